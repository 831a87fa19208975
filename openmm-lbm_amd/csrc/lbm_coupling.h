/*
 * Device helpers of the particle coupling shared by the stencil coupling (lbm_particles.hip) and the learned
 * closure (lbm_closure.hip): the lattice coordinate and nearest node of a particle, the stencils, and the coupling of
 * one particle by a group of lanes.
 */
#pragma once

#include "lbm_internal.h"

namespace lbm {

/* Position wrapped into [0, n) in lattice spacings (ReferenceLBMKernels.cpp:978-979), without FMA
   contraction so that the node of a particle is the same bit for bit as on the CPU. */
__device__ __forceinline__ double wrapCoordinate(double x, double dx, int n) {
#pragma clang fp contract(off)
    double s = x/dx;
    s -= floor(s/n)*n;
    return s;
}

__device__ __forceinline__ int nearestIndex(double s, int n) {
#pragma clang fp contract(off)
    return ((int) floor(s + 0.5)) % n;
}

/* The node nearest to a particle at x (nm). */
__device__ __forceinline__ uint32_t nearestNodeOf(const double* x, const ParticleArgs& a) {
    int size[3] = {(int) a.g.nx, (int) a.g.ny, (int) a.g.nz};
    int idx[3];
    for (int k = 0; k < 3; k++)
        idx[k] = nearestIndex(wrapCoordinate(x[k], a.dx, size[k]), size[k]);
    return (uint32_t) (idx[0] + (int) a.g.nx*(idx[1] + (int) a.g.ny*idx[2]));
}

/* Kernel phi(r) of the stencils (LBMStencils.h:42-62). */
__device__ __forceinline__ double stencilKernel(int stencil, double r) {
#pragma clang fp contract(off)
    double s = fabs(r);
    switch (stencil) {
    case LBM_STENCIL_TRILINEAR:
        return s < 1.0 ? 1.0 - s : 0.0;
    case LBM_STENCIL_THREE_POINT:
        if (s <= 0.5)
            return (1.0 + sqrt(1.0 - 3.0*s*s))/3.0;
        if (s < 1.5)
            return (5.0 - 3.0*s - sqrt(fmax(0.0, 1.0 - 3.0*(1.0 - s)*(1.0 - s))))/6.0;
        return 0.0;
    case LBM_STENCIL_KEYS:
        if (s <= 1.0)
            return 1.0 - 2.5*s*s + 1.5*s*s*s;
        if (s < 2.0)
            return 2.0 - 4.0*s + 2.5*s*s - 0.5*s*s*s;
        return 0.0;
    default:
        return (r >= -0.5 && r < 0.5) ? 1.0 : 0.0;
    }
}

/* Nodes and weights of the stencil along one axis (LBMStencils.h:73-95): width w nodes from floor(s),
   round(s)-1 or floor(s)-1, wrapped or clamped at open faces, repeated nodes merged into the first slot. */
__device__ __forceinline__ void axisWeights(int stencil, int w, double s, int n, bool open, int index[4], double weight[4]) {
#pragma clang fp contract(off)
    int first;
    if (stencil == LBM_STENCIL_THREE_POINT)
        first = (int) floor(s + 0.5) - 1;
    else if (stencil == LBM_STENCIL_KEYS)
        first = (int) floor(s) - 1;
    else if (stencil == LBM_STENCIL_TRILINEAR)
        first = (int) floor(s);
    else
        first = (int) floor(s + 0.5);
    for (int a = 0; a < 4; a++) {
        if (a >= w)
            break;
        int i = first + a;
        weight[a] = stencilKernel(stencil, s - (double) i);
        index[a] = open ? min(max(i, 0), n - 1) : ((i % n) + n) % n;
    }
    for (int a = 1; a < 4; a++) {
        if (a >= w)
            break;
        for (int b = 0; b < a; b++)
            if (index[b] == index[a]) {
                weight[b] += weight[a];
                weight[a] = 0.0;
                break;
            }
    }
}

/* The stencil that a group of LANES lanes couples with (one lane per slot): 1 nearest, 8 trilinear, 32 three-point (27
   slots), 64 Keys; the launchers pick LANES from the stencil. */
template <int LANES>
constexpr int groupStencil() {
    return LANES == 64 ? LBM_STENCIL_KEYS : LANES == 32 ? LBM_STENCIL_THREE_POINT :
           LANES == 8 ? LBM_STENCIL_TRILINEAR : LBM_STENCIL_NEAREST;
}

/*
 * Node and weight of slot `sub` of the stencil along one axis: axisWeights for one slot, one kernel evaluation.  The
 * slots are w consecutive indices from `first`, within [-1, n + 1]; they fall on distinct nodes unless the lattice is
 * narrower than the stencil (periodic) or the stencil reaches past an open face (clamped): there the full rule, with
 * repeated nodes merged into the first slot, applies.
 */
template <int STENCIL>
__device__ __forceinline__ void axisSlot(double s, int n, bool open, int sub, int& index, double& weight) {
#pragma clang fp contract(off)
    if constexpr (STENCIL == LBM_STENCIL_NEAREST) {
        (void) open;
        (void) sub;
        index = nearestIndex(s, n);
        weight = 1.0;
    }
    else {
        constexpr int w = STENCIL == LBM_STENCIL_KEYS ? 4 : STENCIL == LBM_STENCIL_THREE_POINT ? 3 : 2;
        int first = STENCIL == LBM_STENCIL_THREE_POINT ? (int) floor(s + 0.5) - 1 :
                    STENCIL == LBM_STENCIL_KEYS ? (int) floor(s) - 1 : (int) floor(s);
        bool merged = open ? (first < 0 || first + w > n) : n < w;
        if (!merged) {
            int i = first + sub;
            weight = stencilKernel(STENCIL, s - (double) i);
            index = open ? i : (i < 0 ? i + n : (i >= n ? i - n : i));
        }
        else {
            int ix[4];
            double wx[4];
            axisWeights(STENCIL, w, s, n, open, ix, wx);
            index = ix[sub];
            weight = wx[sub];
        }
    }
}

__device__ __forceinline__ bool solidAt(const uint32_t* mask, uint32_t node) {
    return mask != nullptr && isSolid(mask, node);
}

/* One lane's slot of a particle's stencil. */
struct StencilSlot {
    uint32_t node;
    double weight;
    bool inStencil;
    bool solid;
};

/* rho and j (with j -= rho u_cm when the removal is due) of one node from its pulled populations, FP64. */
template <typename S>
__device__ __forceinline__ double nodeMoments(const S* __restrict__ src, const Geometry& g, uint32_t node,
        const double* __restrict__ ucm, int removal, double j[3]) {
    uint32_t ni, nj, nk;
    nodeCoords(g, node, ni, nj, nk);
    typename StorageTraits<S>::Real f[NQ];
    pullNode(src, g, node, ni, nj, nk, f);
    double dr = 0, jx = 0, jy = 0, jz = 0;
    staticFor<NQ>([&](auto I) {
        constexpr int q = decltype(I)::value;
        double v = (double) f[q];
        dr += v;
        jx = addTerm<CX[q]>(jx, v);
        jy = addTerm<CY[q]>(jy, v);
        jz = addTerm<CZ[q]>(jz, v);
    });
    double rho = 1.0 + dr;
    if (removal) {
        jx -= rho*ucm[0];
        jy -= rho*ucm[1];
        jz -= rho*ucm[2];
    }
    j[0] = jx;
    j[1] = jy;
    j[2] = jz;
    return rho;
}

/* The fluid velocity of one node as the coupling sees it (computeMoments): u = j/rho of the pulled populations, with
   j -= rho u_cm when the momentum removal is due, times weight (ReferenceLBMKernels.cpp:503-553, 643-697).  Zero where
   rho <= 0. */
template <typename S>
__device__ __forceinline__ void nodeVelocity(const S* __restrict__ src, const Geometry& g, uint32_t node,
        const double* __restrict__ ucm, int removal, double weight, double u[3]) {
    double j[3];
    double rho = nodeMoments(src, g, node, ucm, removal, j);
    u[0] = u[1] = u[2] = 0.0;
    if (rho > 0) {
        u[0] = j[0]*(weight/rho);
        u[1] = j[1]*(weight/rho);
        u[2] = j[2]*(weight/rho);
    }
}

/*
 * The moments of every node, (jx, jy, jz, rho) with the removal applied, as nodeMoments forms them: a lattice-wide pass
 * that lets dense particles interpolate from 32 bytes per node instead of pulling 19 populations per stencil slot.
 */
template <typename S>
__device__ __forceinline__ void momentFieldNode(const S* __restrict__ src, const Geometry& g, uint32_t node,
        const double* __restrict__ ucm, int removal, double4* __restrict__ field) {
    double j[3];
    double rho = nodeMoments(src, g, node, ucm, removal, j);
    field[node] = make_double4(j[0], j[1], j[2], rho);
}

/* nodeVelocity from the moment field: the same arithmetic on the same moments, so the same bits. */
__device__ __forceinline__ void fieldVelocity(const double4* __restrict__ field, uint32_t node, double weight,
        double u[3]) {
    const double4 m = field[node];
    u[0] = u[1] = u[2] = 0.0;
    if (m.w > 0) {
        u[0] = m.x*(weight/m.w);
        u[1] = m.y*(weight/m.w);
        u[2] = m.z*(weight/m.w);
    }
}

/*
 * The interpolated fluid velocity u(X) of one particle by a group of LANES lanes, one lane per slot of the stencil
 * a.stencil (every lane of the group calls it, inactive ones with active = false): each lane forms w_j u_j of its node
 * (zero at solid nodes, whose rho is 0 in computeMoments) and the group sums them; every lane returns the sum in u and
 * its own slot.
 */
template <typename S, int LANES>
__device__ __forceinline__ StencilSlot interpolateGroup(const S* __restrict__ src, const uint32_t* __restrict__ mask,
        const double* x, bool active, int lane, const double* __restrict__ ucm, const ParticleArgs& a, double u[3],
        const double4* __restrict__ field = nullptr, const AxisStencil* __restrict__ pre = nullptr) {
    constexpr int STENCIL = groupStencil<LANES>();
    constexpr int w = STENCIL == LBM_STENCIL_KEYS ? 4 : STENCIL == LBM_STENCIL_THREE_POINT ? 3 :
                      STENCIL == LBM_STENCIL_TRILINEAR ? 2 : 1;
    int slot = lane;
    StencilSlot sl;
    sl.inStencil = active && slot < w*w*w;
    sl.weight = 0.0;
    sl.node = 0;
    int size[3] = {(int) a.g.nx, (int) a.g.ny, (int) a.g.nz};
    int openAxis[3] = {a.g.openX, a.g.openY, a.g.openZ};
    if (sl.inStencil) {
        int sub[3] = {slot % w, (slot/w) % w, slot/(w*w)};
        int nodeIdx[3];
        sl.weight = 1.0;
        for (int k = 0; k < 3; k++) {
            double wk;
            if (pre != nullptr) {
                // The particle's stencil as stencilSetup formed it: the same axisSlot values.
                nodeIdx[k] = pre->index[k][sub[k]];
                wk = pre->weight[k][sub[k]];
            }
            else
                axisSlot<STENCIL>(wrapCoordinate(x[k], a.dx, size[k]), size[k], openAxis[k] != 0, sub[k], nodeIdx[k],
                        wk);
            sl.weight *= wk;
        }
        sl.node = (uint32_t) (nodeIdx[0] + (int) a.g.nx*(nodeIdx[1] + (int) a.g.ny*nodeIdx[2]));
    }
    sl.solid = sl.inStencil && solidAt(mask, sl.node);
    double ux = 0, uy = 0, uz = 0;
    if (sl.inStencil && !sl.solid) {
        double un[3];
        if (field != nullptr)
            fieldVelocity(field, sl.node, sl.weight, un);
        else
            nodeVelocity(src, a.g, sl.node, ucm, a.removal, sl.weight, un);
        ux = un[0];
        uy = un[1];
        uz = un[2];
    }
#pragma unroll
    for (int off = 1; off < LANES; off <<= 1) {
        ux += __shfl_xor(ux, off, LANES);
        uy += __shfl_xor(uy, off, LANES);
        uz += __shfl_xor(uz, off, LANES);
    }
    u[0] = ux;
    u[1] = uy;
    u[2] = uz;
    return sl;
}

/* The explicit Euler-Maruyama drag in lattice units, F = -gamma m (v - u) + sqrt(2 gamma m kT) xi, with the particle's
   random numbers of the step (keyed by index and step). */
__device__ __forceinline__ void particleNoise(uint32_t index, const ParticleArgs& a, double noise[4]) {
    // The precision of a random number: FP64 with FP64 storage, else FP32 as the fluid's noise (boxMullerPair).
    if (a.noiseFP64)
        drawNormals<double, 4>(a.seed, TAG_PARTICLE, index, a.step, noise);
    else {
        float r[4];
        drawNormals<float, 4>(a.seed, TAG_PARTICLE, index, a.step, r);
        for (int k = 0; k < 4; k++)
            noise[k] = (double) r[k];
    }
}

/*
 * Normal k (0-3) of particleNoise's four, from the words drawNormals gives its pair p = k/2: the radius is word p of the
 * first Philox block, the angle half p of the first word of the third; even k take the cosine, odd k the sine, as in
 * boxMullerPair, in the precision particleNoise uses.  Four lanes drawing one normal each give the particle's normals
 * in the time of one logarithm, square root and cosine or sine.
 */
__device__ __forceinline__ double particleNormal(uint32_t index, int k, const ParticleArgs& a) {
    const uint32_t lo = (uint32_t) (uint64_t) a.step, hi = (uint32_t) ((uint64_t) a.step >> 32);
    const uint4x radii = philox(index, lo, hi, 0u, a.seed, TAG_PARTICLE);
    const uint4x angles = philox(index, lo, hi, 2u, a.seed, TAG_PARTICLE);
    const int p = k >> 1;
    const uint32_t word = p == 0 ? radii.x : radii.y, half = (angles.x >> (16*p)) & 0xFFFFu;
    double pair[2];
    if (a.noiseFP64)
        boxMullerPair<double>(word, half, pair);
    else {
        float f[2];
        boxMullerPair<float>(word, half, f);
        pair[0] = (double) f[0];
        pair[1] = (double) f[1];
    }
    return pair[k & 1];
}

/* preset: the particle's normals when already drawn (particleNoise), else nullptr. */
__device__ __forceinline__ void dragForce(const double* vel, const double u[3], uint32_t index, double m,
        const ParticleArgs& a, double f[3], const double* preset = nullptr) {
    double noise[4] = {0, 0, 0, 0};
    if (a.randomForce) {
        if (preset != nullptr)
            for (int k = 0; k < 3; k++)
                noise[k] = preset[k];
        else
            particleNoise(index, a, noise);
    }
    double amp = sqrt(2.0*a.gamma*m*a.kT);
    for (int k = 0; k < 3; k++) {
        double v = vel[k]/a.velocityScale;
        f[k] = (v - u[k])*(-a.gamma*m);
        if (a.randomForce)
            f[k] += noise[k]*amp;
    }
}

/* The reaction -w F of one slot: on a fluid node into its fixed-point accumulators; on a solid node it leaves the
   fluid and goes to the wall (ReferenceLBMKernels.cpp:959-971).  Returns whether an accumulator received a nonzero
   value (the node is then flagged in the touched mask, which the fused pass reads and clears). */
__device__ __forceinline__ bool spreadReaction(const StencilSlot& sl, const double f[3], const ParticleArgs& a,
        unsigned long long* __restrict__ reactionFixed, double wallAcc[3]) {
    bool any = false;
    for (int k = 0; k < 3; k++) {
        double r = -sl.weight*f[k];
        if (sl.solid)
            wallAcc[k] = r;
        else if (reactionInRange(r, a.overflow)) {
            long long fixed = llrint(r*FIXED_SCALE);
            if (fixed != 0) {
                __hip_atomic_fetch_add(asGlobal(reactionFixed + reactionIndex(sl.node, k, a.g.n, a.reactionPlanes)),
                        (unsigned long long) fixed,
                        __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                any = true;
            }
        }
    }
    return any;
}

/*
 * Marks the nodes that several slots reach: each live slot's node bit goes into seen, and into shared where it was set
 * already.  The W slots of a stencil row are consecutive lanes on x-consecutive nodes, which mostly share a word: the
 * first lane of each run of lanes on one word sets the run's bits with one atomic OR and, from the bits it returns and
 * any bit two lanes of the run hold, the run's shared bits with another.  A word the run visits twice (a lattice
 * narrower than the stencil) may flag a node shared that is not, which only costs an atomic.  Every lane of the
 * wavefront calls it.
 */
template <int W>
__device__ __forceinline__ void markSharedRun(uint32_t* __restrict__ seen, uint32_t* __restrict__ shared, uint32_t node,
        bool live, int lane) {
    const uint32_t word = node >> 5;
    const uint32_t bit = live ? 1u << (node & 31) : 0u;
    uint32_t bits = bit, twice = 0u;
    bool first = true;
    if constexpr (W > 1) {
        const int x = lane % W;
#pragma unroll
        for (int d = 1; d < W; d++) {
            uint32_t w2 = __shfl_down(word, d, 64);
            uint32_t b2 = __shfl_down(bit, d, 64);
            if (x + d < W && w2 == word) {
                twice |= bits & b2;
                bits |= b2;
            }
        }
        uint32_t before = __shfl_up(word, 1, 64);
        first = x == 0 || before != word;
    }
    if (first && bits != 0u) {
        uint32_t old = atomicOr(seen + word, bits);
        uint32_t both = (old & bits) | twice;
        if (both != 0u)
            atomicOr(shared + word, both);
    }
}

/*
 * Flags the nodes of the slots with `flagged` in the touched mask (bit node & 31 of word node >> 5).  The W slots of
 * a stencil row are consecutive lanes on x-consecutive nodes, which mostly share a mask word: the first lane of each
 * run of lanes on one word ORs in the bits of the row's lanes on that word, one atomic instead of up to W.  A bit may
 * be set twice where a row's words are not in runs (a lattice narrower than the stencil); OR makes that harmless.
 * Every lane of the wavefront calls it.
 */
template <int W>
__device__ __forceinline__ void markTouched(uint32_t* __restrict__ touched, uint32_t node, bool flagged, int lane) {
    const uint32_t word = node >> 5;
    const uint32_t bit = flagged ? 1u << (node & 31) : 0u;
    uint32_t bits = bit;
    bool first = true;
    if constexpr (W > 1) {
        const int x = lane % W;                     // the slot's place in its stencil row
#pragma unroll
        for (int d = 1; d < W; d++) {
            uint32_t w2 = __shfl_down(word, d, 64);
            uint32_t b2 = __shfl_down(bit, d, 64);
            if (x + d < W && w2 == word)
                bits |= b2;
        }
        uint32_t before = __shfl_up(word, 1, 64);
        first = x == 0 || before != word;
    }
    if (first && bits != 0u)
        __hip_atomic_fetch_or(asGlobal(touched + word), bits, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

/* The wall momentum acc of the work-group's threads, summed into partials[3*blockIdx.x + k] in a fixed order (the
   butterfly within each wavefront, then the wavefronts in order), so that runs are reproducible.  Every thread of the
   work-group calls it. */
template <int THREADS>
__device__ __forceinline__ void storeWallPartial(const double acc[3], double* __restrict__ partials) {
    __shared__ double waveAcc[THREADS/64][3];
    for (int k = 0; k < 3; k++) {
        double v = acc[k];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1)
            v += __shfl_xor(v, off, 64);
        if ((threadIdx.x & 63) == 0)
            waveAcc[threadIdx.x >> 6][k] = v;
    }
    __syncthreads();
    if (threadIdx.x == 0)
        for (int k = 0; k < 3; k++) {
            double s = 0.0;
            for (int w = 0; w < THREADS/64; w++)
                s += waveAcc[w][k];
            partials[3*blockIdx.x + k] = s;
        }
}

/* The second half of coupleGroup, after the interpolation: the drag force (written by the group's first lane) and,
   in a lattice step, the spread of the reaction. */
template <int LANES>
__device__ __forceinline__ void finishGroup(const StencilSlot& sl, const double u[3], const double* vel, uint32_t index,
        double m, bool active, int lane, const ParticleArgs& a, double* forceOut,
        unsigned long long* __restrict__ reactionFixed, uint32_t* __restrict__ touched, double wallAcc[3],
        const double* noise, double* systemForce = nullptr) {
    // Every lane of the group holds u; each computes the force (identical values).
    double f[3] = {0, 0, 0};
    if (active) {
        dragForce(vel, u, index, m, a, f, noise);
        if (lane == 0)
            for (int k = 0; k < 3; k++) {
                forceOut[k] = f[k]*a.forceScale;
                // The caller's force array, when given: one addition per coupled particle (atomic, in case a particle
                // were coupled twice).
                if (systemForce != nullptr)
                    atomicAdd(systemForce + k, f[k]*a.forceScale);
            }
    }
    if (a.isStep) {
        bool flagged = sl.inStencil && spreadReaction(sl, f, a, reactionFixed, wallAcc);
        constexpr int STENCIL = groupStencil<LANES>();
        constexpr int W = STENCIL == LBM_STENCIL_KEYS ? 4 : STENCIL == LBM_STENCIL_THREE_POINT ? 3 :
                          STENCIL == LBM_STENCIL_TRILINEAR ? 2 : 1;
        markTouched<W>(touched, sl.node, flagged, lane);
    }
}

/*
 * The coupling of one particle by a group of LANES lanes (one lane per stencil slot; every lane of the group calls
 * it, inactive ones with active = false).  x, v: the particle's position (nm) and velocity (nm/ps); index: its
 * coupled-particle index (the counter of its random numbers); m: its mass in lattice units; noise: its normals of the
 * step when drawn beforehand (else nullptr: drawn here).  The force (kJ/mol/nm) is
 * written to forceOut by the first lane; in a lattice step the reaction goes to the accumulators and, on solid nodes,
 * into wallAcc (lattice momentum; the host converts the wall partial sums with m_c dx/dt).
 */
template <typename S, int LANES>
__device__ __forceinline__ void coupleGroup(const S* __restrict__ src, const uint32_t* __restrict__ mask,
        const double* x, const double* vel, uint32_t index, double m, bool active, int lane,
        const double* __restrict__ ucm, const ParticleArgs& a, double* forceOut,
        unsigned long long* __restrict__ reactionFixed, uint32_t* __restrict__ touched, double wallAcc[3],
        const double* noise = nullptr) {
    double u[3];
    StencilSlot sl = interpolateGroup<S, LANES>(src, mask, x, active, lane, ucm, a, u);
    finishGroup<LANES>(sl, u, vel, index, m, active, lane, a, forceOut, reactionFixed, touched, wallAcc, noise);
}

}  // namespace lbm
