/*
 * Particle coupling kernels of the gfx950 LBMForce platform.
 *
 *  - reflectParticles: a coupled particle whose nearest node is solid and that moves into the wall
 *    has its velocity reversed; the wall receives 2 m v (ReferenceLBMKernels.cpp:308-339, 1020-1050;
 *    lbmCoupling.cc:116-171).
 *  - coupleParticles: explicit Euler-Maruyama drag against the fluid velocity interpolated with the
 *    stencil (nearest, trilinear, three-point, Keys): F = -gamma m (v - u(X)) + sqrt(2 gamma m kT) xi
 *    (ReferenceLBMKernels.cpp:503-553, 643-697; lbmCoupling.cc:184-245, 1132-1238).  One group of
 *    lanes per particle, one lane per stencil node: each lane pulls the 19 populations of its node
 *    and forms rho and j, the group reduces u.  The reaction -w_j F is spread onto the nodes with
 *    64-bit fixed-point atomics (2^-50 lattice units): integer addition is associative, so the
 *    per-node sums are bit-identical from run to run whatever the order of the atomics, which is
 *    the reference's determinism guarantee (docs/architecture.md:223-227) without its sort.
 *    Each touched node is flagged in a one-bit-per-node mask; the fused pass converts the accumulators
 *    of the flagged nodes and clears them and the mask (the reference's sumCellReactions and
 *    clearCellReactions, lbmCoupling.cc:251-272, 1852-1877, folded into the lattice pass).
 */
#include "lbm_coupling.h"

#include <hipcub/hipcub.hpp>
#include <stdexcept>

namespace lbm {

/* ------------------------------------------------------------------------------------------ */
/* Also clears the other wall partial sums of the step, allPartials[0, 3*count) outside this kernel's own slots (the
   kernels that write them run later on the stream). */
__global__ void __launch_bounds__(BLOCK) reflectParticles(const double* __restrict__ positions, double* __restrict__ velocities,
        const int* __restrict__ particles, const double* __restrict__ mass, const uint32_t* __restrict__ mask,
        ParticleArgs a, double* __restrict__ allPartials, int count, int ownOffset) {
    double* partials = allPartials + 3*(size_t) ownOffset;
    for (int j = blockIdx.x*BLOCK + threadIdx.x; j < 3*count; j += gridDim.x*BLOCK)
        if (j < 3*ownOffset || j >= 3*(ownOffset + (int) gridDim.x))
            allPartials[j] = 0.0;
    double acc[3] = {0, 0, 0};
    int i = blockIdx.x*BLOCK + threadIdx.x;
    if (i < a.numParticles) {
        int part = particles[i];
        const double* x = positions + 3*(size_t) part;
        double* v = velocities + 3*(size_t) part;
        int size[3] = {(int) a.g.nx, (int) a.g.ny, (int) a.g.nz};
        int near[3], idx[3][2];
        double wt[3][2];
        for (int k = 0; k < 3; k++) {
            double s = wrapCoordinate(x[k], a.dx, size[k]);
            near[k] = nearestIndex(s, size[k]);
            double lower = floor(s);
            idx[k][0] = ((int) lower) % size[k];
            idx[k][1] = (idx[k][0] + 1) % size[k];
            wt[k][1] = s - lower;
            wt[k][0] = 1.0 - wt[k][1];
        }
        uint32_t nearNode = (uint32_t) (near[0] + (int) a.g.nx*(near[1] + (int) a.g.ny*near[2]));
        if (isSolid(mask, nearNode)) {
            double solid[2][2][2];
            for (int p = 0; p < 2; p++)
                for (int q = 0; q < 2; q++)
                    for (int r = 0; r < 2; r++)
                        solid[p][q][r] = isSolid(mask, (uint32_t) (idx[0][p] + (int) a.g.nx*(idx[1][q] + (int) a.g.ny*idx[2][r]))) ? 1.0 : 0.0;
            double n[3] = {0, 0, 0};
            for (int p = 0; p < 2; p++)
                for (int q = 0; q < 2; q++) {
                    n[0] += wt[1][p]*wt[2][q]*(solid[1][p][q] - solid[0][p][q]);
                    n[1] += wt[0][p]*wt[2][q]*(solid[p][1][q] - solid[p][0][q]);
                    n[2] += wt[0][p]*wt[1][q]*(solid[p][q][1] - solid[p][q][0]);
                }
            if (v[0]*n[0] + v[1]*n[1] + v[2]*n[2] > 0) {
                for (int k = 0; k < 3; k++) {
                    acc[k] = v[k]*(2.0*mass[i]*a.cellMass);
                    v[k] = -v[k];
                }
            }
        }
    }
    storeWallPartial<BLOCK>(acc, partials);
}

/* ------------------------------------------------------------------------------------------ */
template <typename S, int LANES>
__global__ void __launch_bounds__(BLOCK) coupleParticles(const S* __restrict__ src, const uint32_t* __restrict__ mask,
        const double* __restrict__ positions, const double* __restrict__ velocities, const int* __restrict__ particles,
        const int* __restrict__ order, const double* __restrict__ mass, const double* __restrict__ ucm, ParticleArgs a,
        double* __restrict__ particleForce, unsigned long long* __restrict__ reactionFixed, uint32_t* __restrict__ touched,
        double* __restrict__ partials, double* __restrict__ systemForces) {
    constexpr int GROUPS = BLOCK/LANES;
    int lane = threadIdx.x % LANES;
    int rank = blockIdx.x*GROUPS + threadIdx.x/LANES;    // position in the processing order
    bool active = rank < a.numParticles;
    int i = active && order != nullptr ? order[rank] : rank;
    int part = active ? particles[i] : 0;
    double m = active ? mass[i] : 0.0;
    double u[3];
    StencilSlot sl = interpolateGroup<S, LANES>(src, mask, positions + 3*(size_t) part, active, lane, ucm, a, u);
    // The normals of the work-group's particles, drawn once by its first GROUPS lanes instead of by every group: a
    // wavefront pays for the FP64 Box-Muller once however many of its lanes need it.  Drawn after the interpolation,
    // so that the first wavefront's loads are in flight meanwhile.
    __shared__ double noise[LANES > 1 ? GROUPS : 1][4];
    const double* myNoise = nullptr;
    if (LANES > 1 && a.randomForce) {
        if (threadIdx.x < GROUPS) {
            double r[4] = {0, 0, 0, 0};
            int k = blockIdx.x*GROUPS + threadIdx.x;
            if (k < a.numParticles)
                particleNoise(order != nullptr ? order[k] : k, a, r);
            for (int k = 0; k < 4; k++)
                noise[threadIdx.x][k] = r[k];
        }
        __syncthreads();
        myNoise = noise[threadIdx.x/LANES];
    }
    double wallAcc[3] = {0, 0, 0};
    finishGroup<LANES>(sl, u, velocities + 3*(size_t) part, (uint32_t) i, m, active, lane, a,
            particleForce + 3*(size_t) (active ? i : 0), reactionFixed, touched, wallAcc, myNoise,
            systemForces != nullptr ? systemForces + 3*(size_t) part : nullptr);
    if (partials != nullptr)
        storeWallPartial<BLOCK>(wallAcc, partials);
}

/* Batched replicas: the coupling of every particle of every replica in one launch.  Each work-group couples up to
   BLOCK/LANES particles of one replica (particleTiles[blockIdx.x] names it), so the replica's descriptor is uniform
   across the work-group and read with scalar loads. */
template <int LANES>
__global__ void __launch_bounds__(BLOCK) coupleBatch(const ReplicaDesc* __restrict__ table,
        const uint16_t* __restrict__ particleTiles, const double* __restrict__ positions,
        const double* __restrict__ velocities, const double* __restrict__ mass, double* __restrict__ forces,
        const int64_t* __restrict__ stepCounter) {
    constexpr int GROUPS = BLOCK/LANES;
    int lane = threadIdx.x % LANES;
    const ReplicaDesc* d = table + particleTiles[blockIdx.x];
    int local = (int) (blockIdx.x - d->particleTileBegin)*GROUPS + threadIdx.x/LANES;
    bool active = local < d->part.numParticles;
    int pp = d->particleBegin + (active ? local : 0);
    ParticleArgs a = d->part;
    a.step = *stepCounter;
    a.isStep = 1;
    const float* src = (a.step & 1) ? d->popB : d->popA;
    double u[3];
    StencilSlot sl = interpolateGroup<float, LANES>(src, nullptr, positions + 3*(size_t) pp, active, lane, nullptr, a, u);
    // The work-group's normals, drawn once by its first GROUPS lanes (as in coupleParticles).
    __shared__ double noise[LANES > 1 ? GROUPS : 1][4];
    const double* myNoise = nullptr;
    if (LANES > 1 && a.randomForce) {
        if (threadIdx.x < GROUPS) {
            double r[4] = {0, 0, 0, 0};
            int k = (int) (blockIdx.x - d->particleTileBegin)*GROUPS + threadIdx.x;
            if (k < a.numParticles)
                particleNoise((uint32_t) k, a, r);
            for (int c = 0; c < 4; c++)
                noise[threadIdx.x][c] = r[c];
        }
        __syncthreads();
        myNoise = noise[threadIdx.x/LANES];
    }
    double wallAcc[3] = {0, 0, 0};
    finishGroup<LANES>(sl, u, velocities + 3*(size_t) pp, (uint32_t) (pp - d->particleBegin), active ? mass[pp] : 0.0, active,
            lane, a, forces + 3*(size_t) pp, d->reactionFixed, d->touched, wallAcc, myNoise);
}

/* Leapfrog Verlet of the batch (the coupling force is the only force).  The step counter is advanced by advanceStep,
   launched after this kernel on the same stream. */
__global__ void __launch_bounds__(BLOCK) verletBatch(const ReplicaDesc* __restrict__ table,
        const uint16_t* __restrict__ particleReplica, int numParticles, double* __restrict__ positions,
        double* __restrict__ velocities, const double* __restrict__ forces, const double* __restrict__ invMass) {
    int p = blockIdx.x*BLOCK + threadIdx.x;
    if (p >= numParticles)
        return;
    double dt = table[particleReplica[p]].dt;
    for (int k = 0; k < 3; k++) {
        size_t idx = 3*(size_t) p + k;
        double v = velocities[idx] + dt*forces[idx]*invMass[p];
        velocities[idx] = v;
        positions[idx] += dt*v;
    }
}

__global__ void advanceStep(int64_t* stepCounter) {
    if (threadIdx.x == 0)
        *stepCounter += 1;
}

__global__ void __launch_bounds__(BLOCK) addForces(const double* __restrict__ particleForce, const int* __restrict__ particles,
        int count, double* __restrict__ forces) {
    int i = blockIdx.x*BLOCK + threadIdx.x;
    if (i >= count)
        return;
    int part = particles[i];
    for (int k = 0; k < 3; k++)
        forces[3*(size_t) part + k] += particleForce[3*(size_t) i + k];
}

/* ------------------------------------------------------------------------------------------ */
static inline unsigned gridFor(uint64_t n) {
    return (unsigned) ((n + BLOCK - 1)/BLOCK);
}

void launchReflect(const double* positions, double* velocities, const int* particles, const double* mass,
        const uint32_t* mask, const ParticleArgs& a, double* allPartials, int count, int ownOffset, hipStream_t stream) {
    if (a.numParticles == 0)
        return;
    reflectParticles<<<gridFor(a.numParticles), BLOCK, 0, stream>>>(positions, velocities, particles, mass, mask, a,
            allPartials, count, ownOffset);
}

int coupleLanes(int stencil) {
    switch (stencil) {
    case LBM_STENCIL_TRILINEAR: return 8;
    case LBM_STENCIL_THREE_POINT: return 32;
    case LBM_STENCIL_KEYS: return 64;
    default: return 1;
    }
}

int coupleBlocks(int stencil, int numParticles) {
    int groups = BLOCK/coupleLanes(stencil);
    return (numParticles + groups - 1)/groups;
}

template <typename S>
static void launchCoupleT(const void* src, const uint32_t* mask, const double* positions, const double* velocities,
        const int* particles, const int* order, const double* mass, const double* ucm, const ParticleArgs& a,
        double* particleForce, unsigned long long* reactionFixed, uint32_t* touched, double* partials,
        double* systemForces, hipStream_t stream) {
    unsigned grid = (unsigned) coupleBlocks(a.stencil, a.numParticles);
    const S* s = (const S*) src;
    switch (coupleLanes(a.stencil)) {
    case 8:
        coupleParticles<S, 8><<<grid, BLOCK, 0, stream>>>(s, mask, positions, velocities, particles, order, mass, ucm, a,
                particleForce, reactionFixed, touched, partials, systemForces);
        break;
    case 32:
        coupleParticles<S, 32><<<grid, BLOCK, 0, stream>>>(s, mask, positions, velocities, particles, order, mass, ucm, a,
                particleForce, reactionFixed, touched, partials, systemForces);
        break;
    case 64:
        coupleParticles<S, 64><<<grid, BLOCK, 0, stream>>>(s, mask, positions, velocities, particles, order, mass, ucm, a,
                particleForce, reactionFixed, touched, partials, systemForces);
        break;
    default:
        coupleParticles<S, 1><<<grid, BLOCK, 0, stream>>>(s, mask, positions, velocities, particles, order, mass, ucm, a,
                particleForce, reactionFixed, touched, partials, systemForces);
        break;
    }
}

void launchCouple(int storage, const void* src, const uint32_t* mask, const double* positions, const double* velocities,
        const int* particles, const int* order, const double* mass, const double* ucm, const ParticleArgs& a,
        double* particleForce, unsigned long long* reactionFixed, uint32_t* touched, double* partials,
        double* systemForces, hipStream_t stream) {
    if (a.numParticles == 0)
        return;
    switch (storage) {
    case LBM_STORAGE_FP16:
        launchCoupleT<__half>(src, mask, positions, velocities, particles, order, mass, ucm, a, particleForce, reactionFixed,
                touched, partials, systemForces, stream);
        break;
    case LBM_STORAGE_BF16:
        launchCoupleT<__hip_bfloat16>(src, mask, positions, velocities, particles, order, mass, ucm, a, particleForce,
                reactionFixed, touched, partials, systemForces, stream);
        break;
    case LBM_STORAGE_FP64:
        launchCoupleT<double>(src, mask, positions, velocities, particles, order, mass, ucm, a, particleForce, reactionFixed,
                touched, partials, systemForces, stream);
        break;
    default:
        launchCoupleT<float>(src, mask, positions, velocities, particles, order, mass, ucm, a, particleForce, reactionFixed,
                touched, partials, systemForces, stream);
        break;
    }
}

void launchCoupleBatch(int stencil, const ReplicaDesc* table, const uint16_t* particleTiles, int numParticleTiles,
        const double* positions, const double* velocities, const double* mass, double* forces,
        const int64_t* stepCounter, hipStream_t stream) {
    if (numParticleTiles == 0)
        return;
    unsigned grid = (unsigned) numParticleTiles;
    switch (coupleLanes(stencil)) {
    case 8:
        coupleBatch<8><<<grid, BLOCK, 0, stream>>>(table, particleTiles, positions, velocities, mass, forces, stepCounter);
        break;
    case 32:
        coupleBatch<32><<<grid, BLOCK, 0, stream>>>(table, particleTiles, positions, velocities, mass, forces, stepCounter);
        break;
    case 64:
        coupleBatch<64><<<grid, BLOCK, 0, stream>>>(table, particleTiles, positions, velocities, mass, forces, stepCounter);
        break;
    default:
        coupleBatch<1><<<grid, BLOCK, 0, stream>>>(table, particleTiles, positions, velocities, mass, forces, stepCounter);
        break;
    }
}

void launchVerletBatch(const ReplicaDesc* table, const uint16_t* particleReplica, int numParticles, double* positions,
        double* velocities, const double* forces, const double* invMass, int64_t* stepCounter, hipStream_t stream) {
    if (numParticles > 0)
        verletBatch<<<gridFor(numParticles), BLOCK, 0, stream>>>(table, particleReplica, numParticles, positions,
                velocities, forces, invMass);
    advanceStep<<<1, 64, 0, stream>>>(stepCounter);
}

void launchAddForces(const double* particleForce, const int* particles, int count, double* forces, hipStream_t stream) {
    if (count == 0)
        return;
    addForces<<<gridFor(count), BLOCK, 0, stream>>>(particleForce, particles, count, forces);
}

/* The nearest node of every coupled particle, as the coupling kernels compute it (nearestNodeOf). */
__global__ void __launch_bounds__(BLOCK) particleNodes(const double* __restrict__ positions,
        const int* __restrict__ particles, ParticleArgs a, int* __restrict__ nodes) {
    int i = blockIdx.x*BLOCK + threadIdx.x;
    if (i < a.numParticles)
        nodes[i] = (int) nearestNodeOf(positions + 3*(size_t) particles[i], a);
}

void launchParticleNodes(const double* positions, const int* particles, const ParticleArgs& a, int* nodes,
        hipStream_t stream) {
    if (a.numParticles > 0)
        particleNodes<<<gridFor(a.numParticles), BLOCK, 0, stream>>>(positions, particles, a, nodes);
}


/* ------------------------------------------------------------------------------------------ */
/* Processing order of the coupling.                                                           */
/* ------------------------------------------------------------------------------------------ */
struct CoupleOrder {
    int count = 0;
    int sortBits = 1;
    uint32_t* keys = nullptr;
    uint32_t* sortedKeys = nullptr;
    int* values = nullptr;
    int* order = nullptr;
    void* temp = nullptr;
    size_t tempBytes = 0;
    int64_t sortedAt = 0;
    bool valid = false;
};

__global__ void __launch_bounds__(BLOCK) orderKeys(const double* __restrict__ positions, const int* __restrict__ particles,
        ParticleArgs a, uint32_t* __restrict__ keys, int* __restrict__ values) {
    int i = blockIdx.x*BLOCK + threadIdx.x;
    if (i >= a.numParticles)
        return;
    // Descending node index: the coupling ends on the low nodes, where the lattice pass that follows (in increasing
    // node order) starts, so the accumulator lines that it consumes first are still cached (measured: 7 us faster).
    keys[i] = a.g.n - 1u - nearestNodeOf(positions + 3*(size_t) particles[i], a);
    values[i] = i;
}

static void orderCheck(hipError_t e, const char* what) {
    if (e != hipSuccess)
        throw std::runtime_error(std::string("HIP error in the coupling order (") + what + "): " + hipGetErrorString(e));
}

void coupleOrderDestroy(CoupleOrder* o) {
    if (o == nullptr)
        return;
    void* buffers[] = {o->keys, o->sortedKeys, o->values, o->order, o->temp};
    for (void* b : buffers)
        if (b)
            (void) hipFree(b);
    delete o;
}

CoupleOrder* coupleOrderCreate(int numParticles, uint32_t numNodes) {
    CoupleOrder* o = new CoupleOrder;
    try {
        o->count = numParticles;
        while (o->sortBits < 32 && (1ull << o->sortBits) < numNodes)
            o->sortBits++;
        orderCheck(hipMalloc((void**) &o->keys, sizeof(uint32_t)*numParticles), "allocation");
        orderCheck(hipMalloc((void**) &o->sortedKeys, sizeof(uint32_t)*numParticles), "allocation");
        orderCheck(hipMalloc((void**) &o->values, sizeof(int)*numParticles), "allocation");
        orderCheck(hipMalloc((void**) &o->order, sizeof(int)*numParticles), "allocation");
        orderCheck(hipcub::DeviceRadixSort::SortPairs(nullptr, o->tempBytes, o->keys, o->sortedKeys, o->values, o->order,
                numParticles, 0, o->sortBits, (hipStream_t) 0), "sort size");
        orderCheck(hipMalloc(&o->temp, std::max<size_t>(o->tempBytes, 1)), "allocation");
    }
    catch (...) {
        coupleOrderDestroy(o);
        throw;
    }
    return o;
}

const int* coupleOrderAt(CoupleOrder* o, int64_t step, const double* positions, const int* particles,
        const ParticleArgs& a, hipStream_t stream) {
    if (o->valid && step >= o->sortedAt && step - o->sortedAt < COUPLE_ORDER_STEPS)
        return o->order;
    orderKeys<<<gridFor(o->count), BLOCK, 0, stream>>>(positions, particles, a, o->keys, o->values);
    size_t bytes = o->tempBytes;
    orderCheck(hipcub::DeviceRadixSort::SortPairs(o->temp, bytes, o->keys, o->sortedKeys, o->values, o->order, o->count,
            0, o->sortBits, stream), "sort");
    orderCheck(hipGetLastError(), "launch");
    o->sortedAt = step;
    o->valid = true;
    return o->order;
}

}  // namespace lbm
