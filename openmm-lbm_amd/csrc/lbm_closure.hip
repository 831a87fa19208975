/*
 * Coupling kernels of the learned closure (include/lbm_amd_closure.h).
 *
 *  closureFeatures  one group of 32 lanes per particle: lanes 0-26 pull the 19 populations of the 27 nodes around the
 *                   nearest node and form u = j/rho (FP64, as the stencil coupling does); lane 27 writes the offset,
 *                   lane 28 the particle velocity, lanes 29-31 the zero padding, all in the particle's frame (below).
 *                   The particle is out of distribution if a normalized velocity exceeds its bound, a neighbour is
 *                   solid, or the neighbourhood leaves the lattice at an open face.
 *  closureTargets   u_Keys(x) - u(nearest) with the stencil interpolation of the coupling, in the particle's frame
 *                   (training pairs).
 *  closureCouple    one lane per particle: the nearest-node drag with u(nearest) + du (du turned back from the
 *                   particle's frame) and its reaction on the nearest node, exactly as the nearest-node stencil
 *                   coupling computes them (ReferenceLBMKernels.cpp:503-553); with du = 0 the two are identical bit for
 *                   bit.
 *
 * The particle's frame: the Keys stencil and the lattice are symmetric under the 48 reflections and axis permutations
 * of the cube, so the network sees every particle after the one that makes its sub-node offset o non-negative and
 * sorted, o'_0 >= o'_1 >= o'_2 >= 0: frame axis i is lattice axis p_i, reflected where o_(p_i) < 0 (sign s_i = -1).
 * A vector w of the lattice frame has w'_i = s_i w_(p_i) there, and the node nearest + d sits at d'_i = s_i d_(p_i) of
 * the 3x3x3 block.  The network then learns on a 48th of the offsets' cube.
 *  closureFallback  the Keys coupling of the out-of-distribution particles (ReferenceLBMKernels.cpp:643-697): fixed
 *                   grid, each wave scans the flags of 64 particles at a time with a fixed stride, so the wall partial
 *                   sums have a fixed order.
 */
#include "lbm_coupling.h"
#include "lbm_mlp.h"

namespace lbm {

namespace {

constexpr int FEATURE_LANES = 32;
constexpr int FALLBACK_MAX_BLOCKS = 512;

/* The frame of a particle whose sub-node offset is o (FP64): bit k set where lattice axis k is reflected (o_k < 0);
   bits 3-4 and 5-6 hold p_0 and p_1, the lattice axes of frame axes 0 and 1 (p_2 = 3 - p_0 - p_1), ordered by |o|
   from the largest, equal values in axis order. */
__device__ __forceinline__ uint8_t frameCode(const double o[3]) {
    double m[3];
    unsigned code = 0;
    for (int k = 0; k < 3; k++) {
        m[k] = fabs(o[k]);
        if (o[k] < 0)
            code |= 1u << k;
    }
    // Rank of each axis: the axes that precede it.
    int r0 = (m[1] > m[0]) + (m[2] > m[0]);
    int r1 = (m[0] >= m[1]) + (m[2] > m[1]);
    int p0 = r0 == 0 ? 0 : (r1 == 0 ? 1 : 2);
    int p1 = r0 == 1 ? 0 : (r1 == 1 ? 1 : 2);
    return (uint8_t) (code | (unsigned) p0 << 3 | (unsigned) p1 << 5);
}

struct Frame {
    int axis[3];      // p_i: the lattice axis of frame axis i
    double sign[3];   // s_i: -1 where that lattice axis is reflected
};

__device__ __forceinline__ Frame frameOf(unsigned code) {
    Frame f;
    f.axis[0] = (int) (code >> 3) & 3;
    f.axis[1] = (int) (code >> 5) & 3;
    f.axis[2] = 3 - f.axis[0] - f.axis[1];
    for (int i = 0; i < 3; i++)
        f.sign[i] = (code >> f.axis[i]) & 1u ? -1.0 : 1.0;
    return f;
}

/* v[k] for a lane-varying k, by selects (no indexed private array). */
template <typename T>
__device__ __forceinline__ T pick(const T v[3], int k) {
    return k == 0 ? v[0] : (k == 1 ? v[1] : v[2]);
}

template <typename S>
__global__ void __launch_bounds__(BLOCK) closureFeatures(const S* __restrict__ src, const uint32_t* __restrict__ mask,
        const double* __restrict__ positions, const double* __restrict__ velocities, const int* __restrict__ particles,
        const double* __restrict__ ucm, ParticleArgs a, ClosureNorm norm, const double* __restrict__ field,
        _Float16* __restrict__ X, float* __restrict__ raw, uint32_t* __restrict__ centreNode,
        double* __restrict__ centreVelocity, uint8_t* __restrict__ ood, uint8_t* __restrict__ frame) {
    constexpr int GROUPS = BLOCK/FEATURE_LANES;
    int lane = threadIdx.x % FEATURE_LANES;
    int i = blockIdx.x*GROUPS + threadIdx.x/FEATURE_LANES;
    bool active = i < a.numParticles;
    bool out = false;
    int part = active ? particles[i] : 0;
    const double* x = positions + 3*(size_t) part;
    int size[3] = {(int) a.g.nx, (int) a.g.ny, (int) a.g.nz};
    // The wrapped coordinate and the nearest node of axis k (FP64, as the coupling computes them) by the group's lane k,
    // shared with the others; every lane then forms the particle's frame from the three offsets.  The axis length is
    // selected, not indexed by the lane (a lane-varying index into the kernel argument's fields would be read once for
    // the whole wavefront).
    double swLane = 0.0;
    int nearLane = 0;
    if (active && lane < 3) {
        const int axisSize = lane == 0 ? size[0] : (lane == 1 ? size[1] : size[2]);
        swLane = wrapCoordinate(x[lane], a.dx, axisSize);
        nearLane = nearestIndex(swLane, axisSize);
    }
    double sw[3], o[3];
    int near[3];
    for (int k = 0; k < 3; k++) {
        sw[k] = __shfl(swLane, k, FEATURE_LANES);
        near[k] = __shfl(nearLane, k, FEATURE_LANES);
        o[k] = sw[k] - floor(sw[k] + 0.5);
    }
    const int code = frameCode(o);
    if (active && lane == 0)
        frame[i] = (uint8_t) code;
    const Frame fr = frameOf((unsigned) code);
    if (active) {
        int openAxis[3] = {a.g.openX, a.g.openY, a.g.openZ};
        _Float16* xi = X + (size_t) i*MLP_IN_PAD;
        float* ri = raw ? raw + (size_t) i*MLP_IN : nullptr;
        if (lane < 27) {
            int d[3] = {lane % 3 - 1, (lane/3) % 3 - 1, lane/9 - 1};
            int idx[3];
            for (int k = 0; k < 3; k++) {
                int t = near[k] + d[k];
                if (t < 0 || t >= size[k]) {
                    if (openAxis[k])
                        out = true;
                    t = (t + size[k]) % size[k];
                }
                idx[k] = t;
            }
            uint32_t node = (uint32_t) (idx[0] + (int) a.g.nx*(idx[1] + (int) a.g.ny*idx[2]));
            double u[3] = {0, 0, 0};
            if (solidAt(mask, node))
                out = true;
            else if (field != nullptr)
                for (int c = 0; c < 3; c++)
                    u[c] = field[3*(size_t) node + c];
            else
                nodeVelocity(src, a.g, node, ucm, a.removal, 1.0, u);
            if (lane == 13) {
                centreNode[i] = node;
                for (int c = 0; c < 3; c++)
                    centreVelocity[3*(size_t) i + c] = u[c];
            }
            // This node's place in the particle's frame, and its velocity there.
            int slot = 13;
            for (int k = 0, stride = 1; k < 3; k++, stride *= 3)
                slot += stride*(fr.sign[k] < 0 ? -pick(d, fr.axis[k]) : pick(d, fr.axis[k]));
            for (int c = 0; c < 3; c++) {
                float r = (float) (fr.sign[c]*pick(u, fr.axis[c]));
                float f = r*norm.uScale;
                if (!(fabsf(f) <= norm.uBound))
                    out = true;
                xi[3 + 3*slot + c] = (_Float16) f;
                if (ri)
                    ri[3 + 3*slot + c] = r;
            }
        }
        else if (lane == 27) {
            float r[3];
            for (int k = 0; k < 3; k++)
                r[k] = (float) (2.0*(sw[k] - floor(sw[k] + 0.5)));
            for (int c = 0; c < 3; c++) {
                float q = fabsf(pick(r, fr.axis[c]));
                xi[c] = (_Float16) q;
                if (ri)
                    ri[c] = q;
            }
        }
        else if (lane == 28) {
            const double* v = velocities + 3*(size_t) part;
            for (int c = 0; c < 3; c++) {
                float r = (float) (fr.sign[c]*pick(v, fr.axis[c])/a.velocityScale);
                float f = r*norm.vScale;
                if (!(fabsf(f) <= norm.vBound))
                    out = true;
                xi[84 + c] = (_Float16) f;
                if (ri)
                    ri[84 + c] = r;
            }
        }
        else {
            for (int c = 0; c < 3; c++)
                xi[MLP_IN + 3*(lane - 29) + c] = (_Float16) 0.0f;
        }
    }
    // The 32 lanes of this particle are one half of the wave's ballot.
    unsigned long long votes = __ballot(out);
    unsigned mine = (unsigned) (votes >> (threadIdx.x & 32));
    if (active && lane == 0)
        ood[i] = mine != 0u ? 1 : 0;
}

/* u = j/rho of every fluid node (zero at solid nodes), FP64, [node][3]: one coalesced pass over the populations that
   replaces the 27 scattered pulls per particle when the particles are dense. */
template <typename S>
__global__ void __launch_bounds__(BLOCK) closureField(const S* __restrict__ src, const uint32_t* __restrict__ mask,
        const double* __restrict__ ucm, ParticleArgs a, double* __restrict__ field) {
    uint32_t node = blockIdx.x*BLOCK + threadIdx.x;
    if (node >= a.g.n)
        return;
    double u[3] = {0, 0, 0};
    if (!solidAt(mask, node))
        nodeVelocity(src, a.g, node, ucm, a.removal, 1.0, u);
    for (int c = 0; c < 3; c++)
        field[3*(size_t) node + c] = u[c];
}

template <typename S>
__global__ void __launch_bounds__(BLOCK) closureTargets(const S* __restrict__ src, const uint32_t* __restrict__ mask,
        const double* __restrict__ positions, const int* __restrict__ particles, const double* __restrict__ ucm,
        ParticleArgs a, const double* __restrict__ centreVelocity, const uint8_t* __restrict__ frame,
        float* __restrict__ targets) {
    constexpr int LANES = 64, GROUPS = BLOCK/LANES;
    int lane = threadIdx.x % LANES;
    int i = blockIdx.x*GROUPS + threadIdx.x/LANES;
    bool active = i < a.numParticles;
    int part = active ? particles[i] : 0;
    double u[3];
    interpolateGroup<S, LANES>(src, mask, positions + 3*(size_t) part, active, lane, ucm, a, u);
    if (active && lane == 0) {
        const Frame fr = frameOf(frame[i]);
        double du[3];
        for (int c = 0; c < 3; c++)
            du[c] = u[c] - centreVelocity[3*(size_t) i + c];
        for (int c = 0; c < 3; c++)
            targets[3*(size_t) i + c] = (float) (fr.sign[c]*pick(du, fr.axis[c]));
    }
}

__global__ void __launch_bounds__(BLOCK) closureCouple(const double* __restrict__ velocities,
        const int* __restrict__ particles, const double* __restrict__ mass, ParticleArgs a,
        const uint32_t* __restrict__ centreNode, const double* __restrict__ centreVelocity, const float* __restrict__ du,
        const uint8_t* __restrict__ ood, const uint8_t* __restrict__ frame, double* __restrict__ particleForce,
        unsigned long long* __restrict__ reactionFixed, uint32_t* __restrict__ touched,
        unsigned long long* __restrict__ counters) {
    int i = blockIdx.x*BLOCK + threadIdx.x;
    bool active = i < a.numParticles;
    bool fallback = active && ood[i] != 0;
    bool network = active && !fallback;
    // The work-group's counts, added to its own pair of counters (one writer each: no atomics on a shared word).
    __shared__ unsigned waveCounts[BLOCK/64][2];
    unsigned long long netVotes = __ballot(network), fallVotes = __ballot(fallback);
    if ((threadIdx.x & 63) == 0) {
        waveCounts[threadIdx.x >> 6][0] = (unsigned) __popcll(netVotes);
        waveCounts[threadIdx.x >> 6][1] = (unsigned) __popcll(fallVotes);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned net = 0, fall = 0;
        for (int w = 0; w < BLOCK/64; w++) {
            net += waveCounts[w][0];
            fall += waveCounts[w][1];
        }
        counters[2*blockIdx.x] += net;
        counters[2*blockIdx.x + 1] += fall;
    }
    if (!network)
        return;
    // du is in the particle's frame: lattice component p_k is s_k du_k.
    const Frame fr = frameOf(frame[i]);
    double dq[3];
    for (int k = 0; k < 3; k++)
        dq[k] = fr.sign[k]*(double) du[3*(size_t) i + k];
    double u[3];
    for (int c = 0; c < 3; c++) {
        double back = fr.axis[0] == c ? dq[0] : (fr.axis[1] == c ? dq[1] : dq[2]);
        u[c] = centreVelocity[3*(size_t) i + c] + back;
    }
    double f[3];
    dragForce(velocities + 3*(size_t) particles[i], u, (uint32_t) i, mass[i], a, f);
    for (int c = 0; c < 3; c++)
        particleForce[3*(size_t) i + c] = f[c]*a.forceScale;
    if (a.isStep) {
        StencilSlot sl;
        sl.node = centreNode[i];
        sl.weight = 1.0;
        sl.inStencil = true;
        sl.solid = false;
        double wallAcc[3] = {0, 0, 0};
        markTouched<1>(touched, sl.node, spreadReaction(sl, f, a, reactionFixed, wallAcc), 0);
    }
}

template <typename S>
__global__ void __launch_bounds__(BLOCK) closureFallback(const S* __restrict__ src, const uint32_t* __restrict__ mask,
        const double* __restrict__ positions, const double* __restrict__ velocities, const int* __restrict__ particles,
        const double* __restrict__ mass, const double* __restrict__ ucm, ParticleArgs a, const uint8_t* __restrict__ ood,
        double* __restrict__ particleForce, unsigned long long* __restrict__ reactionFixed, uint32_t* __restrict__ touched,
        double* __restrict__ partials) {
    constexpr int LANES = 64;
    int lane = threadIdx.x % LANES;
    int wave = blockIdx.x*(BLOCK/LANES) + threadIdx.x/LANES;
    int waves = gridDim.x*(BLOCK/LANES);
    double acc[3] = {0, 0, 0};
    // Each wave reads the flags of 64 particles at a time and couples the flagged ones one after another, all 64 lanes
    // on the 64 slots of the Keys stencil: a fixed assignment and order, so the wall partial sums are reproducible.
    for (int base = wave*LANES; base < a.numParticles; base += waves*LANES) {
        int j = base + lane;
        unsigned long long votes = __ballot(j < a.numParticles && ood[j] != 0);
        while (votes != 0ull) {
            int i = base + __ffsll((long long) votes) - 1;
            votes &= votes - 1ull;
            int part = particles[i];
            double wallAcc[3] = {0, 0, 0};
            coupleGroup<S, LANES>(src, mask, positions + 3*(size_t) part, velocities + 3*(size_t) part, (uint32_t) i,
                    mass[i], true, lane, ucm, a, particleForce + 3*(size_t) i, reactionFixed, touched, wallAcc);
            for (int k = 0; k < 3; k++)
                acc[k] += wallAcc[k];
        }
    }
    if (partials != nullptr)
        storeWallPartial<BLOCK>(acc, partials);
}

unsigned gridOf(int64_t items, int perBlock) {
    return (unsigned) ((items + perBlock - 1)/perBlock);
}

}  // namespace

#define LBM_CLOSURE_STORAGE(storage, CALL) \
    switch (storage) { \
    case LBM_STORAGE_FP16: { using S = __half; CALL; break; } \
    case LBM_STORAGE_BF16: { using S = __hip_bfloat16; CALL; break; } \
    case LBM_STORAGE_FP64: { using S = double; CALL; break; } \
    default: { using S = float; CALL; break; } \
    }

void launchClosureFeatures(int storage, const void* src, const uint32_t* mask, const double* positions,
        const double* velocities, const int* particles, const double* ucm, const ParticleArgs& a, const ClosureNorm& norm,
        const double* field, _Float16* X, float* raw, uint32_t* centreNode, double* centreVelocity, uint8_t* ood,
        uint8_t* frame, hipStream_t stream) {
    if (a.numParticles == 0)
        return;
    unsigned grid = gridOf(a.numParticles, BLOCK/FEATURE_LANES);
    LBM_CLOSURE_STORAGE(storage, (closureFeatures<S><<<grid, BLOCK, 0, stream>>>((const S*) src, mask, positions,
            velocities, particles, ucm, a, norm, field, X, raw, centreNode, centreVelocity, ood, frame)))
}

void launchClosureField(int storage, const void* src, const uint32_t* mask, const double* ucm, const ParticleArgs& a,
        double* field, hipStream_t stream) {
    unsigned grid = gridOf(a.g.n, BLOCK);
    LBM_CLOSURE_STORAGE(storage, (closureField<S><<<grid, BLOCK, 0, stream>>>((const S*) src, mask, ucm, a, field)))
}

void launchClosureTargets(int storage, const void* src, const uint32_t* mask, const double* positions,
        const int* particles, const double* ucm, const ParticleArgs& a, const double* centreVelocity,
        const uint8_t* frame, float* targets, hipStream_t stream) {
    if (a.numParticles == 0)
        return;
    unsigned grid = gridOf(a.numParticles, BLOCK/64);
    LBM_CLOSURE_STORAGE(storage, (closureTargets<S><<<grid, BLOCK, 0, stream>>>((const S*) src, mask, positions,
            particles, ucm, a, centreVelocity, frame, targets)))
}

void launchClosureCouple(const double* velocities, const int* particles, const double* mass, const ParticleArgs& a,
        const uint32_t* centreNode, const double* centreVelocity, const float* du, const uint8_t* ood,
        const uint8_t* frame, double* particleForce, unsigned long long* reactionFixed, uint32_t* touched,
        unsigned long long* counters, hipStream_t stream) {
    if (a.numParticles == 0)
        return;
    closureCouple<<<gridOf(a.numParticles, BLOCK), BLOCK, 0, stream>>>(velocities, particles, mass, a, centreNode,
            centreVelocity, du, ood, frame, particleForce, reactionFixed, touched, counters);
}

int closureCoupleBlocks(int numParticles) {
    return (int) gridOf(std::max(numParticles, 1), BLOCK);
}

int closureFallbackBlocks(int numParticles) {
    return std::max(1, std::min((int) gridOf(numParticles, BLOCK), FALLBACK_MAX_BLOCKS));
}

void launchClosureFallback(int storage, const void* src, const uint32_t* mask, const double* positions,
        const double* velocities, const int* particles, const double* mass, const double* ucm, const ParticleArgs& a,
        const uint8_t* ood, double* particleForce, unsigned long long* reactionFixed, uint32_t* touched, double* partials,
        hipStream_t stream) {
    if (a.numParticles == 0)
        return;
    unsigned grid = (unsigned) closureFallbackBlocks(a.numParticles);
    LBM_CLOSURE_STORAGE(storage, (closureFallback<S><<<grid, BLOCK, 0, stream>>>((const S*) src, mask, positions,
            velocities, particles, mass, ucm, a, ood, particleForce, reactionFixed, touched, partials)))
}

}  // namespace lbm
