/*
 * The centred drag (LBMForce::Centered) on gfx950.
 *
 * With h = 1/2, a = gamma h, the velocity of a particle at the time of the force v~ = v(t - h)/(dx/dt) + h Fc/m (Fc:
 * the other forces of the System, read from the force array before the coupling adds to it) and the fluid velocity
 * u~ = (j + h rho g)/rho of a node:
 *
 *  - nearest node (ReferenceLBMKernels.cpp:555-641): the particles of one node are coupled to each other only.  With M,
 *    P~ = sum m v~ and R the sums of their masses, momenta and random forces,
 *        S = [-gamma (P~ - M u~) + R]/(1 + a + a M/rho),  u_c = u~ - h S/rho,  F_k = [-gamma m_k (v~_k - u_c) + R_k]/(1 + a),
 *    and the node receives -S (a solid node is a wall at rest: u_c = 0).  The particles are sorted by node (a stable
 *    radix sort of the node keys, so a node's particles stay in particle order) and the first particle of each run
 *    sums them in that order and solves its node: the sums have the reference's order.
 *  - stencil (ReferenceLBMKernels.cpp:726-901): (1 + a)/m_k F_k + a sum_l K_kl F_l = b_k/m_k with
 *    K_kl = sum_j xi_jk xi_jl/rho_j and b_k = -gamma m_k (v~_k - sum_j xi_jk u~_j) + R_k, solved per component by
 *    conjugate gradients preconditioned with the diagonal, from the diagonal solution, to a residual of 1e-13 of the
 *    right-hand side.  A product spreads the vector onto the nodes and interpolates it back; the spread accumulates in
 *    64-bit fixed point with a scale set from the largest entry of the vector (2^52 / that entry), so the sums are
 *    exact to 2^-52 of it and independent of the order of the atomics.  With sparse particles only the slots at nodes
 *    that several slots share are spread; a slot alone at its node takes its own particle's w x directly.  Two arrays
 *    of node sums serve the products in turn, each gather clearing the previous product's.  Dot products reduce over a fixed grid in a
 *    fixed order.  The device tests the residual after every iteration and the kernels of later iterations then skip
 *    their work; the host enqueues as many iterations as the previous solve took before it reads the residual, then
 *    one at a time, so the iterations and results are those of a check after every iteration without a round trip
 *    to the host for each.
 *
 * The reaction goes into the same fixed-point accumulators as the explicit drag (consumed by the fused pass; with the
 * split, stored where no other slot reaches the node, since those accumulators are zero, and added atomically where
 * several do), and into the wall partial sums at solid nodes.
 */
#include "lbm_coupling.h"
#include <hipcub/hipcub.hpp>
#include <cmath>
#include <stdexcept>

namespace lbm {

namespace {

constexpr int RED_BLOCKS = 256;          // fixed grid of the CG reductions
constexpr double CG_TOLERANCE = 1e-13;
constexpr int CG_MAX_ITERATIONS = 1000;

/* Device scalars of the conjugate gradients. */
enum Scalar {
    RZ = 0, RZ_NEW = 3, RR = 6, NORM = 9, DAD = 12, ALPHA = 15, BETA = 18, SCALE = 21, INV_SCALE = 22,
    DONE = 23,        // 1 once the residual is within tolerance: the remaining iterations' kernels return at once
    ITERS = 24,       // iterations run in this solve
    NUM_SCALARS = 25
};

void check(hipError_t e, const char* what) {
    if (e != hipSuccess)
        throw std::runtime_error(std::string("centred drag: ") + what + ": " + hipGetErrorString(e));
}

template <typename T>
T* alloc(size_t count) {
    void* p = nullptr;
    check(hipMalloc(&p, std::max<size_t>(count, 1)*sizeof(T)), "hipMalloc");
    return (T*) p;
}

/* v~ and the random force R of coupled particle i, lattice units (ReferenceLBMKernels.cpp:577-585). */
__device__ __forceinline__ void particleKnowns(const double* vel, const double* fc, double m, uint32_t i,
        const ParticleArgs& a, double known[3], double random[3]) {
    double noise[4] = {0, 0, 0, 0};
    if (a.randomForce)
        particleNoise(i, a, noise);
    double amp = sqrt(2.0*a.gamma*m*a.kT);
    for (int k = 0; k < 3; k++) {
        known[k] = vel[k]*(1.0/a.velocityScale) + (fc ? fc[k] : 0.0)*(0.5/(m*a.forceScale));
        random[k] = a.randomForce ? noise[k]*amp : 0.0;
    }
}

/* The reaction r on a node: into its fixed-point accumulators, or to the wall on a solid node.  Returns whether an
   accumulator changed (the node is then flagged in the touched mask by markTouched). */
__device__ __forceinline__ bool addReaction(uint32_t node, bool solid, const double r[3], const ParticleArgs& a,
        unsigned long long* reactionFixed, double wallAcc[3]) {
    if (solid) {
        for (int k = 0; k < 3; k++)
            wallAcc[k] += r[k];          // lattice momentum; the host converts the wall partials
        return false;
    }
    bool any = false;
    for (int k = 0; k < 3; k++) {
        if (!reactionInRange(r[k], a.overflow))
            continue;
        long long fixed = llrint(r[k]*FIXED_SCALE);
        if (fixed != 0) {
            atomicAdd(reactionFixed + reactionIndex(node, k, a.g.n, a.reactionPlanes), (unsigned long long) fixed);
            any = true;
        }
    }
    return any;
}

/* addReaction for a fluid node that no other slot of the step reaches (the CG split's unshared nodes): its
   accumulators are zero (the lattice pass consumed and cleared them), so a store puts the same bits as the atomic add. */
__device__ __forceinline__ bool storeReaction(uint32_t node, const double r[3], const ParticleArgs& a,
        unsigned long long* reactionFixed) {
    bool any = false;
    for (int k = 0; k < 3; k++) {
        if (!reactionInRange(r[k], a.overflow))
            continue;
        long long fixed = llrint(r[k]*FIXED_SCALE);
        if (fixed != 0) {
            reactionFixed[reactionIndex(node, k, a.g.n, a.reactionPlanes)] = (unsigned long long) fixed;
            any = true;
        }
    }
    return any;
}

__device__ __forceinline__ void blockPartials(const double acc[3], double* partials) {
    if (partials != nullptr)
        storeWallPartial<BLOCK>(acc, partials);
}

/* ---------------------------------------------------------------------------------------------------------------- */
/* Nearest node.                                                                                                     */
/* ---------------------------------------------------------------------------------------------------------------- */
__global__ void __launch_bounds__(BLOCK) centredKeys(const double* __restrict__ positions,
        const int* __restrict__ particles, ParticleArgs a, uint32_t* __restrict__ keys, int* __restrict__ values) {
    int i = blockIdx.x*BLOCK + threadIdx.x;
    if (i >= a.numParticles)
        return;
    keys[i] = nearestNodeOf(positions + 3*(size_t) particles[i], a);
    values[i] = i;
}

template <typename S>
__global__ void __launch_bounds__(BLOCK) centredNearest(const S* __restrict__ src, const uint32_t* __restrict__ mask,
        const double* __restrict__ velocities, const double* __restrict__ otherForces, const int* __restrict__ particles,
        const double* __restrict__ mass, const double* __restrict__ ucm, ParticleArgs a, double gx, double gy, double gz,
        const uint32_t* __restrict__ sortedKeys, const int* __restrict__ sortedIdx, double* __restrict__ particleForce,
        unsigned long long* __restrict__ reactionFixed, uint32_t* __restrict__ touched, double* __restrict__ partials) {
    const double h = 0.5, gamma = a.gamma, aa = a.gamma*h;
    double wallAcc[3] = {0, 0, 0};
    int s = blockIdx.x*BLOCK + threadIdx.x;
    if (s < a.numParticles && (s == 0 || sortedKeys[s] != sortedKeys[s - 1])) {
        uint32_t node = sortedKeys[s];
        double M = 0, P[3] = {0, 0, 0}, R[3] = {0, 0, 0};
        int e = s;
        for (; e < a.numParticles && sortedKeys[e] == node; e++) {
            int i = sortedIdx[e];
            int part = particles[i];
            double known[3], random[3];
            particleKnowns(velocities + 3*(size_t) part, otherForces ? otherForces + 3*(size_t) part : nullptr, mass[i],
                    (uint32_t) i, a, known, random);
            M += mass[i];
            for (int k = 0; k < 3; k++) {
                P[k] += known[k]*mass[i];
                R[k] += random[k];
            }
        }
        bool solid = solidAt(mask, node);
        double rho = 0, j[3] = {0, 0, 0};
        if (!solid)
            rho = nodeMoments(src, a.g, node, ucm, a.removal, j);
        double g[3] = {gx, gy, gz};
        double u[3] = {0, 0, 0}, sum[3];
        if (rho > 0) {
            for (int k = 0; k < 3; k++) {
                double knownFluid = (j[k] + h*(rho*g[k]))*(1.0/rho);
                sum[k] = ((P[k] - knownFluid*M)*(-gamma) + R[k])*(1.0/(1.0 + aa + aa*M/rho));
                u[k] = knownFluid - sum[k]*(h/rho);
            }
        }
        else
            for (int k = 0; k < 3; k++)
                sum[k] = ((P[k] - u[k]*M)*(-gamma) + R[k])*(1.0/(1.0 + aa));
        for (int t = s; t < e; t++) {
            int i = sortedIdx[t];
            int part = particles[i];
            double known[3], random[3];
            particleKnowns(velocities + 3*(size_t) part, otherForces ? otherForces + 3*(size_t) part : nullptr, mass[i],
                    (uint32_t) i, a, known, random);
            for (int k = 0; k < 3; k++)
                particleForce[3*(size_t) i + k] = ((known[k] - u[k])*(-gamma*mass[i]) + random[k])*(1.0/(1.0 + aa))*
                        a.forceScale;
        }
        if (a.isStep) {
            double r[3] = {-sum[0], -sum[1], -sum[2]};
            markTouched<1>(touched, node, addReaction(node, solid, r, a, reactionFixed, wallAcc), 0);
        }
    }
    blockPartials(wallAcc, partials);
}

/* ---------------------------------------------------------------------------------------------------------------- */
/* Stencils: conjugate gradients.                                                                                    */
/* ---------------------------------------------------------------------------------------------------------------- */
struct CgArgs {
    double* rhs;
    double* diag;
    double* F;
    double* r;
    double* z;
    double* d;
    double* Ad;
    uint32_t* slotNode;
    double* slotW;
    double* slotWInvRho;     // w/rho, 0 at solid nodes
    uint8_t* slotSolid;
    uint8_t* slotSpreads;    // 1 where the slot takes part in the spread of a product (set by the solve's first one)
    unsigned long long* spread;        // [node][3] fixed point: the node sums of the current product
    unsigned long long* spreadOther;   // the other of the two arrays: the previous product's, cleared by the gather
    double* partials;        // [RED_BLOCKS][8]
    double* scalars;
    int slots;               // w^3
    // Sparse particles (split = 1): a product spreads only the slots at nodes that several slots share (spreadsAt); a
    // slot alone at its node contributes (w/rho) w x of its own particle directly.  seen/shared: node bits set by
    // cgMarkShared and cleared by cgForces.
    int split;
    uint32_t* seen;
    uint32_t* shared;
};

/* Whether a slot's node takes part in the spread of a product: every node without the split, else the shared ones. */
__device__ __forceinline__ bool spreadsAt(const CgArgs& c, uint32_t node) {
    return !c.split || ((c.shared[node >> 5] >> (node & 31)) & 1u);
}

/* One group of LANES lanes per particle, one lane per stencil slot: the slot's node, weight and w/rho, the particle's
   right-hand side, diagonal and F = rhs/diagonal. */
template <typename S, int LANES>
__global__ void __launch_bounds__(BLOCK) cgSetup(const S* __restrict__ src, const uint32_t* __restrict__ mask,
        const double* __restrict__ positions, const double* __restrict__ velocities,
        const double* __restrict__ otherForces, const int* __restrict__ particles, const double* __restrict__ mass,
        const double* __restrict__ ucm, ParticleArgs a, double gx, double gy, double gz, CgArgs c) {
    constexpr int GROUPS = BLOCK/LANES;
    const double h = 0.5, aa = a.gamma*h;
    int lane = threadIdx.x % LANES;
    int i = blockIdx.x*GROUPS + threadIdx.x/LANES;
    bool active = i < a.numParticles;
    int part = active ? particles[i] : 0;
    const double* x = positions + 3*(size_t) part;
    // The stencil of the lanes (groupStencil): each lane evaluates its own slot's kernel weight per axis.
    constexpr int STENCIL = groupStencil<LANES>();
    constexpr int w = STENCIL == LBM_STENCIL_KEYS ? 4 : STENCIL == LBM_STENCIL_THREE_POINT ? 3 : 2;
    bool inStencil = active && lane < c.slots;
    double ux = 0, uy = 0, uz = 0, self = 0;
    if (inStencil) {
        int size[3] = {(int) a.g.nx, (int) a.g.ny, (int) a.g.nz};
        int openAxis[3] = {a.g.openX, a.g.openY, a.g.openZ};
        int sub[3] = {lane % w, (lane/w) % w, lane/(w*w)};
        int nodeIdx[3];
        double weight = 1.0;
        for (int k = 0; k < 3; k++) {
            double wk;
            axisSlot<STENCIL>(wrapCoordinate(x[k], a.dx, size[k]), size[k], openAxis[k] != 0, sub[k], nodeIdx[k], wk);
            weight *= wk;
        }
        uint32_t node = (uint32_t) (nodeIdx[0] + (int) a.g.nx*(nodeIdx[1] + (int) a.g.ny*nodeIdx[2]));
        bool solid = solidAt(mask, node);
        double winv = 0.0;
        if (!solid) {
            double j[3];
            double rho = nodeMoments(src, a.g, node, ucm, a.removal, j);
            if (rho > 0) {
                ux = (j[0] + h*(rho*gx))*(weight/rho);
                uy = (j[1] + h*(rho*gy))*(weight/rho);
                uz = (j[2] + h*(rho*gz))*(weight/rho);
                self = weight*weight/rho;
                winv = weight/rho;
            }
        }
        size_t sl = (size_t) i*c.slots + lane;
        c.slotNode[sl] = node;
        c.slotW[sl] = weight;
        c.slotWInvRho[sl] = winv;
        c.slotSolid[sl] = solid ? 1 : 0;
    }
#pragma unroll
    for (int off = 1; off < LANES; off <<= 1) {
        ux += __shfl_xor(ux, off, LANES);
        uy += __shfl_xor(uy, off, LANES);
        uz += __shfl_xor(uz, off, LANES);
        self += __shfl_xor(self, off, LANES);
    }
    if (active && lane == 0) {
        double m = mass[i];
        double known[3], random[3];
        particleKnowns(velocities + 3*(size_t) part, otherForces ? otherForces + 3*(size_t) part : nullptr, m, (uint32_t) i,
                a, known, random);
        double u[3] = {ux, uy, uz};
        double diag = (1.0 + aa)/m + aa*self;
        c.diag[i] = diag;
        for (int k = 0; k < 3; k++) {
            double b = ((known[k] - u[k])*(-a.gamma*m) + random[k])*(1.0/m);
            c.rhs[3*(size_t) i + k] = b;
            c.F[3*(size_t) i + k] = b*(1.0/diag);
        }
    }
}

/* The nodes that several fluid slots share (with the split), one group of LANES lanes per particle, one lane per slot
   (markSharedRun: one atomic OR per run of a stencil row's lanes on one word). */
template <int LANES>
__global__ void __launch_bounds__(BLOCK) cgMarkShared(int n, CgArgs c) {
    constexpr int GROUPS = BLOCK/LANES;
    constexpr int W = LANES == 64 ? 4 : LANES == 32 ? 3 : 2;    // slots per stencil row
    const int lane = threadIdx.x % LANES;
    const int i = blockIdx.x*GROUPS + threadIdx.x/LANES;
    uint32_t node = 0;
    bool live = false;
    if (i < n && lane < c.slots) {
        size_t sl = (size_t) i*c.slots + lane;
        live = c.slotWInvRho[sl] != 0.0;
        node = c.slotNode[sl];
    }
    markSharedRun<W>(c.seen, c.shared, node, live, lane);
}

/* Every component's residual within CG_TOLERANCE of its right-hand side (the host's test and the device's). */
__host__ __device__ inline bool residualConverged(const double* s) {
    for (int k = 0; k < 3; k++)
        if (s[RR + k] > CG_TOLERANCE*CG_TOLERANCE*s[NORM + k])
            return false;
    return true;
}

/* Largest |x| over the particles and components -> partials[block][0]. */
__global__ void __launch_bounds__(BLOCK) cgMaxAbs(const double* __restrict__ x, int n, double* __restrict__ partials,
        const double* done) {
    if (done != nullptr && *done != 0.0)
        return;
    double m = 0.0;
    for (int i = blockIdx.x*BLOCK + threadIdx.x; i < 3*n; i += gridDim.x*BLOCK)
        m = fmax(m, fabs(x[i]));
    __shared__ double lds[BLOCK];
    lds[threadIdx.x] = m;
    __syncthreads();
    for (int s = BLOCK/2; s > 0; s >>= 1) {
        if ((int) threadIdx.x < s)
            lds[threadIdx.x] = fmax(lds[threadIdx.x], lds[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0)
        partials[8*blockIdx.x] = lds[0];
}

/* The partial rows of the RED_BLOCKS blocks of a reduction, combined by one wavefront (launched with 64 threads):
   lane l takes the rows l, l + 64, ... in order, then the lanes combine in a fixed butterfly, so the result is the same
   in every run.  MAX takes the largest value instead of the sum.  Every lane returns the result. */
template <int COUNT, bool MAX = false>
__device__ __forceinline__ void combinePartials(const double* __restrict__ partials, double v[COUNT]) {
    static_assert(RED_BLOCKS % 64 == 0, "whole rows per lane");
    for (int q = 0; q < COUNT; q++)
        v[q] = 0.0;
    for (int b = threadIdx.x; b < RED_BLOCKS; b += 64)
        for (int q = 0; q < COUNT; q++)
            v[q] = MAX ? fmax(v[q], partials[8*b + q]) : v[q] + partials[8*b + q];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
        for (int q = 0; q < COUNT; q++) {
            double o = __shfl_xor(v[q], off, 64);
            v[q] = MAX ? fmax(v[q], o) : v[q] + o;
        }
}

/* spread[node] += w x (fixed point), one thread per particle slot.  The fixed-point scale, 2^(52 - e) with
   2^e >= max |x| (1 if x is zero), is formed by the first wavefront of every work-group from the largest |x| of the
   blocks in c.partials (one fixed combination, so every work-group has the same value); work-group 0 keeps it and its
   inverse in the scalars for the gather.  The solve's first product (first = 1) records which slots spread. */
__global__ void __launch_bounds__(BLOCK) cgSpread(const double* __restrict__ x, int n, CgArgs c,
        const double* done, int first) {
    if (done != nullptr && *done != 0.0)
        return;
    __shared__ double sharedScale;
    if (threadIdx.x < 64) {
        double m[1];
        combinePartials<1, true>(c.partials, m);
        if (threadIdx.x == 0) {
            int e = m[0] > 0 ? ilogb(m[0]) + 1 : 0;
            sharedScale = ldexp(1.0, 52 - e);
            if (blockIdx.x == 0) {
                c.scalars[SCALE] = sharedScale;
                c.scalars[INV_SCALE] = ldexp(1.0, e - 52);
            }
        }
    }
    __syncthreads();
    size_t t = (size_t) blockIdx.x*BLOCK + threadIdx.x;
    if (t >= (size_t) n*c.slots)
        return;
    double w = c.slotW[t];
    uint32_t node = c.slotNode[t];
    bool spreads;
    if (first) {
        spreads = w != 0.0 && spreadsAt(c, node);
        c.slotSpreads[t] = spreads ? 1 : 0;
    }
    else
        spreads = c.slotSpreads[t] != 0;
    if (!spreads)
        return;
    size_t i = t/c.slots;
    const double scale = sharedScale;
    for (int k = 0; k < 3; k++) {
        long long v = llrint(w*x[3*i + k]*scale);
        if (v != 0)
            atomicAdd(c.spread + 3*(size_t) node + k, (unsigned long long) v);
    }
}

/* Ax = x (1 + a)/m + a sum_slots (w/rho) spread[node], one group of LANES lanes per particle. */
template <int LANES>
__global__ void __launch_bounds__(BLOCK) cgGather(const double* __restrict__ x, double* __restrict__ Ax,
        const double* __restrict__ mass, int n, double aa, CgArgs c, const double* done) {
    if (done != nullptr && *done != 0.0)
        return;
    constexpr int GROUPS = BLOCK/LANES;
    int lane = threadIdx.x % LANES;
    int i = blockIdx.x*GROUPS + threadIdx.x/LANES;
    bool active = i < n;
    double s[3] = {0, 0, 0};
    if (active && lane < c.slots) {
        size_t sl = (size_t) i*c.slots + lane;
        double winv = c.slotWInvRho[sl];
        const bool spreads = c.slotSpreads[sl] != 0;
        const uint32_t node = c.slotNode[sl];
        // The previous product's node sums, which every reader has read, cleared for the next product.
        if (spreads && c.spreadOther != nullptr)
            for (int k = 0; k < 3; k++)
                c.spreadOther[3*(size_t) node + k] = 0ull;
        if (winv != 0.0) {
            if (spreads) {
                double inv = c.scalars[INV_SCALE];
                for (int k = 0; k < 3; k++)
                    s[k] = (double) (long long) c.spread[3*(size_t) node + k]*inv*winv;
            }
            else {
                // Alone at its node: the node's sum is this slot's own w x.
                double ww = c.slotW[sl]*winv;
                for (int k = 0; k < 3; k++)
                    s[k] = ww*x[3*(size_t) i + k];
            }
        }
    }
#pragma unroll
    for (int off = 1; off < LANES; off <<= 1)
        for (int k = 0; k < 3; k++)
            s[k] += __shfl_xor(s[k], off, LANES);
    if (active && lane == 0)
        for (int k = 0; k < 3; k++)
            Ax[3*(size_t) i + k] = x[3*(size_t) i + k]*((1.0 + aa)/mass[i]) + s[k]*aa;
}

/* Block sums of up to 8 per-particle quantities into partials[block][0..7], fixed order. */
__device__ __forceinline__ void reduceBlock(double v[8], int count, double* partials) {
    __shared__ double lds[8][BLOCK];
    for (int q = 0; q < count; q++)
        lds[q][threadIdx.x] = v[q];
    __syncthreads();
    for (int s = BLOCK/2; s > 0; s >>= 1) {
        if ((int) threadIdx.x < s)
            for (int q = 0; q < count; q++)
                lds[q][threadIdx.x] += lds[q][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int q = 0; q < count; q++)
            partials[8*blockIdx.x + q] = lds[q][0];
}

/* r = rhs - Ax, z = r/diag, d = z; partials: r.z, rhs.rhs, r.r (3 each -> 9 values in two reductions). */
__global__ void __launch_bounds__(BLOCK) cgInit(int n, CgArgs c, double* partialsB) {
    double v[8] = {0, 0, 0, 0, 0, 0, 0, 0}, w[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int i = blockIdx.x*BLOCK + threadIdx.x; i < n; i += gridDim.x*BLOCK)
        for (int k = 0; k < 3; k++) {
            size_t t = 3*(size_t) i + k;
            double r = c.rhs[t] - c.Ad[t];
            double z = r*(1.0/c.diag[i]);
            c.r[t] = r;
            c.z[t] = z;
            c.d[t] = z;
            v[k] += r*z;
            v[3 + k] += c.rhs[t]*c.rhs[t];
            w[k] += r*r;
        }
    reduceBlock(v, 6, c.partials);
    __syncthreads();
    reduceBlock(w, 3, partialsB);
}

__global__ void __launch_bounds__(64) cgFinishInit(const double* __restrict__ partials,
        const double* __restrict__ partialsB, double* __restrict__ scalars) {
    double v[6], w[3];
    combinePartials<6>(partials, v);
    combinePartials<3>(partialsB, w);
    if (threadIdx.x != 0)
        return;
    for (int k = 0; k < 3; k++) {
        scalars[RZ + k] = v[k];
        scalars[NORM + k] = v[3 + k];
        scalars[RR + k] = w[k];
    }
    scalars[DONE] = residualConverged(scalars) ? 1.0 : 0.0;
    scalars[ITERS] = 0.0;
}

/* d . Ad into c.partials (alpha = r.z/(d.Ad) follows in cgUpdate). */
__global__ void __launch_bounds__(BLOCK) cgDot(int n, CgArgs c, const double* done) {
    if (done != nullptr && *done != 0.0)
        return;
    double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int i = blockIdx.x*BLOCK + threadIdx.x; i < n; i += gridDim.x*BLOCK)
        for (int k = 0; k < 3; k++)
            v[k] += c.d[3*(size_t) i + k]*c.Ad[3*(size_t) i + k];
    reduceBlock(v, 3, c.partials);
}

/* alpha = r.z/(d.Ad), from the blocks' d.Ad in c.partials (formed by the first wavefront of every work-group, one
   fixed combination; work-group 0 keeps it in the scalars); F += alpha d, r -= alpha Ad, z = r/diag; partialsOut:
   r.z (new), r.r (a buffer of its own: other work-groups may still be reading c.partials). */
__global__ void __launch_bounds__(BLOCK) cgUpdate(int n, CgArgs c, double* __restrict__ partialsOut,
        const double* done) {
    if (done != nullptr && *done != 0.0)
        return;
    __shared__ double sharedAlpha[3];
    if (threadIdx.x < 64) {
        double dad[3];
        combinePartials<3>(c.partials, dad);
        if (threadIdx.x == 0)
            for (int k = 0; k < 3; k++) {
                sharedAlpha[k] = dad[k] > 0 ? c.scalars[RZ + k]/dad[k] : 0.0;
                if (blockIdx.x == 0) {
                    c.scalars[DAD + k] = dad[k];
                    c.scalars[ALPHA + k] = sharedAlpha[k];
                }
            }
    }
    __syncthreads();
    double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    double alpha[3] = {sharedAlpha[0], sharedAlpha[1], sharedAlpha[2]};
    for (int i = blockIdx.x*BLOCK + threadIdx.x; i < n; i += gridDim.x*BLOCK)
        for (int k = 0; k < 3; k++) {
            size_t t = 3*(size_t) i + k;
            c.F[t] += alpha[k]*c.d[t];
            double r = c.r[t] - alpha[k]*c.Ad[t];
            double z = r*(1.0/c.diag[i]);
            c.r[t] = r;
            c.z[t] = z;
            v[k] += r*z;
            v[3 + k] += r*r;
        }
    reduceBlock(v, 6, partialsOut);
}

/* beta = r.z(new)/r.z(old), r.z <- r.z(new), r.r. */
__global__ void __launch_bounds__(64) cgFinishBeta(const double* __restrict__ partials, double* __restrict__ scalars,
        const double* done) {
    if (done != nullptr && *done != 0.0)
        return;
    double v[6];
    combinePartials<6>(partials, v);
    if (threadIdx.x != 0)
        return;
    for (int k = 0; k < 3; k++) {
        double rz = scalars[RZ + k];
        scalars[BETA + k] = rz > 0 ? v[k]/rz : 0.0;
        scalars[RZ + k] = v[k];
        scalars[RZ_NEW + k] = v[k];
        scalars[RR + k] = v[3 + k];
    }
    scalars[DONE] = residualConverged(scalars) ? 1.0 : 0.0;
    scalars[ITERS] += 1.0;
}

/* d = z + beta d, and the largest |d| of each block -> partials[block][0] (for the scale of the next spread). */
__global__ void __launch_bounds__(BLOCK) cgDirection(int n, CgArgs c, const double* done) {
    if (done != nullptr && *done != 0.0)
        return;
    double beta[3] = {c.scalars[BETA], c.scalars[BETA + 1], c.scalars[BETA + 2]};
    double m = 0.0;
    for (int i = blockIdx.x*BLOCK + threadIdx.x; i < n; i += gridDim.x*BLOCK)
        for (int k = 0; k < 3; k++) {
            size_t t = 3*(size_t) i + k;
            double d = c.z[t] + beta[k]*c.d[t];
            c.d[t] = d;
            m = fmax(m, fabs(d));
        }
    __shared__ double lds[BLOCK];
    lds[threadIdx.x] = m;
    __syncthreads();
    for (int s = BLOCK/2; s > 0; s >>= 1) {
        if ((int) threadIdx.x < s)
            lds[threadIdx.x] = fmax(lds[threadIdx.x], lds[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0)
        c.partials[8*blockIdx.x] = lds[0];
}

/* The forces and the reactions -w F of the solution, one group of LANES lanes per particle (block layout of the explicit
   coupling, so the wall partials fit the same slots). */
template <int LANES>
__global__ void __launch_bounds__(BLOCK) cgForces(int n, ParticleArgs a, CgArgs c, double* __restrict__ particleForce,
        unsigned long long* __restrict__ reactionFixed, uint32_t* __restrict__ touched, double* __restrict__ partials) {
    constexpr int GROUPS = BLOCK/LANES;
    int lane = threadIdx.x % LANES;
    int i = blockIdx.x*GROUPS + threadIdx.x/LANES;
    double wallAcc[3] = {0, 0, 0};
    uint32_t node = 0;
    bool flagged = false;
    if (i < n) {
        double F[3] = {c.F[3*(size_t) i], c.F[3*(size_t) i + 1], c.F[3*(size_t) i + 2]};
        if (lane == 0)
            for (int k = 0; k < 3; k++)
                particleForce[3*(size_t) i + k] = F[k]*a.forceScale;
        if (a.isStep && lane < c.slots) {
            size_t sl = (size_t) i*c.slots + lane;
            double w = c.slotW[sl];
            node = c.slotNode[sl];
            if (w != 0.0) {
                double r[3] = {-w*F[0], -w*F[1], -w*F[2]};
                bool solid = c.slotSolid[sl] != 0;
                // With the split, a fluid slot that spreads nothing is alone at its node: a store, not an atomic.
                if (c.split && !solid && c.slotSpreads[sl] == 0 && c.slotWInvRho[sl] != 0.0)
                    flagged = storeReaction(node, r, a, reactionFixed);
                else
                    flagged = addReaction(node, solid, r, a, reactionFixed, wallAcc);
            }
        }
        // Both arrays of node sums and the node bits of the split, cleared for the next solve (every word that a
        // spread or cgMarkShared set belongs to one of these slots; nothing reads them after the solve).
        if (lane < c.slots) {
            size_t sl = (size_t) i*c.slots + lane;
            uint32_t at = c.slotNode[sl];
            if (c.slotSpreads[sl] != 0)
                for (int k = 0; k < 3; k++) {
                    c.spread[3*(size_t) at + k] = 0ull;
                    c.spreadOther[3*(size_t) at + k] = 0ull;
                }
            if (c.split && c.slotWInvRho[sl] != 0.0) {
                c.seen[at >> 5] = 0u;
                c.shared[at >> 5] = 0u;
            }
        }
    }
    // The slots of a particle are its lanes, x fastest, as in the explicit coupling: one OR per run of a row's lanes.
    if (a.isStep)
        markTouched<LANES == 64 ? 4 : LANES == 32 ? 3 : 2>(touched, node, flagged, lane);
    blockPartials(wallAcc, partials);
}

unsigned blocksOf(size_t items, int perBlock) {
    return (unsigned) ((items + perBlock - 1)/perBlock);
}

}  // namespace

struct CentredState {
    int n = 0, stencil = 0, slots = 1, lanes = 1;
    uint32_t numNodes = 0;
    // Nearest node: the radix sort of the node keys.
    uint32_t* keys = nullptr;
    uint32_t* sortedKeys = nullptr;
    int* values = nullptr;
    int* sortedValues = nullptr;
    void* sortTemp = nullptr;
    size_t sortBytes = 0;
    int sortBits = 1;
    // Stencils: conjugate gradients.
    CgArgs cg{};
    unsigned long long* sums[2] = {nullptr, nullptr};   // the two arrays of node sums, used in turn by the products
    double* partialsB = nullptr;
    double* hostScalars = nullptr;        // pinned: the convergence check reads the scalars
    int lastIterations = 1;               // iterations of the previous solve: enqueued before the first check
    bool planned = true;                  // false (LBM_AMD_CG_PLANNED=0): the host reads the residual after each
    std::vector<void*> buffers;
};

CentredState* centredCreate(int numParticles, int stencil, uint32_t numNodes) {
    CentredState* st = new CentredState();
    try {
        int n = std::max(numParticles, 1);
        st->n = numParticles;
        st->stencil = stencil;
        st->numNodes = numNodes;
        if (const char* plannedEnv = getenv("LBM_AMD_CG_PLANNED"))
            st->planned = plannedEnv[0] != '0';
        int w = stencilWidth(stencil);
        st->slots = w*w*w;
        st->lanes = coupleLanes(stencil);
        if (stencil == LBM_STENCIL_NEAREST) {
            st->keys = alloc<uint32_t>(n);
            st->sortedKeys = alloc<uint32_t>(n);
            st->values = alloc<int>(n);
            st->sortedValues = alloc<int>(n);
            while ((1ull << st->sortBits) < numNodes)
                st->sortBits++;
            check(hipcub::DeviceRadixSort::SortPairs(nullptr, st->sortBytes, st->keys, st->sortedKeys, st->values,
                    st->sortedValues, n, 0, st->sortBits, (hipStream_t) 0), "radix sort size");
            st->sortTemp = alloc<char>(st->sortBytes);
            st->buffers = {st->keys, st->sortedKeys, st->values, st->sortedValues, st->sortTemp};
        }
        else {
            CgArgs& c = st->cg;
            c.slots = st->slots;
            c.rhs = alloc<double>(3*(size_t) n);
            c.diag = alloc<double>(n);
            c.F = alloc<double>(3*(size_t) n);
            c.r = alloc<double>(3*(size_t) n);
            c.z = alloc<double>(3*(size_t) n);
            c.d = alloc<double>(3*(size_t) n);
            c.Ad = alloc<double>(3*(size_t) n);
            c.slotNode = alloc<uint32_t>((size_t) n*st->slots);
            c.slotW = alloc<double>((size_t) n*st->slots);
            c.slotWInvRho = alloc<double>((size_t) n*st->slots);
            c.slotSolid = alloc<uint8_t>((size_t) n*st->slots);
            c.slotSpreads = alloc<uint8_t>((size_t) n*st->slots);
            st->sums[0] = alloc<unsigned long long>(3*(size_t) numNodes);
            st->sums[1] = alloc<unsigned long long>(3*(size_t) numNodes);
            c.spread = st->sums[0];
            c.spreadOther = st->sums[1];
            // The split for sparse particles (stencil slots under a quarter of the nodes: few share a node);
            // LBM_AMD_CG_SPLIT=0 or 1 forces the choice.
            bool sparse = (uint64_t) numParticles*(uint64_t) st->slots*4 <= (uint64_t) numNodes;
            const char* splitEnv = getenv("LBM_AMD_CG_SPLIT");
            c.split = (splitEnv ? splitEnv[0] == '1' : sparse) ? 1 : 0;
            c.seen = nullptr;
            c.shared = nullptr;
            if (c.split) {
                size_t words = ((size_t) numNodes + 31)/32;
                c.seen = alloc<uint32_t>(words);
                c.shared = alloc<uint32_t>(words);
                check(hipMemset(c.seen, 0, words*sizeof(uint32_t)), "memset");
                check(hipMemset(c.shared, 0, words*sizeof(uint32_t)), "memset");
            }
            c.partials = alloc<double>(8*RED_BLOCKS);
            c.scalars = alloc<double>(NUM_SCALARS);
            st->partialsB = alloc<double>(8*RED_BLOCKS);
            check(hipMemset(st->sums[0], 0, 3*(size_t) numNodes*sizeof(unsigned long long)), "memset");
            check(hipMemset(st->sums[1], 0, 3*(size_t) numNodes*sizeof(unsigned long long)), "memset");
            st->buffers = {c.rhs, c.diag, c.F, c.r, c.z, c.d, c.Ad, c.slotNode, c.slotW, c.slotWInvRho, c.slotSolid,
                           c.slotSpreads, st->sums[0], st->sums[1], c.partials, c.scalars, st->partialsB, c.seen,
                           c.shared};
            check(hipHostMalloc((void**) &st->hostScalars, NUM_SCALARS*sizeof(double), hipHostMallocDefault),
                    "pinned scalars");
        }
    }
    catch (...) {
        centredDestroy(st);
        throw;
    }
    return st;
}

void centredDestroy(CentredState* st) {
    if (!st)
        return;
    for (void* p : st->buffers)
        if (p)
            (void) hipFree(p);
    if (st->hostScalars)
        (void) hipHostFree(st->hostScalars);
    delete st;
}

namespace {

/* Ax = A x, with the node sums in array `sums` (the gather clears the other, the previous product's).  The largest |x|
   of each block must be in c.partials (cgMaxAbs, or cgDirection for x = d).  done: the solve's convergence flag inside
   the iteration loop (nullptr outside it); first: the solve's first product. */
template <typename S>
void multiply(CentredState* st, const double* x, double* Ax, const double* mass, double aa, const double* done,
        int sums, bool first, hipStream_t stream) {
    CgArgs c = st->cg;
    c.spread = st->sums[sums];
    c.spreadOther = st->sums[1 - sums];
    int n = st->n;
    size_t slotItems = (size_t) n*c.slots;
    cgSpread<<<blocksOf(slotItems, BLOCK), BLOCK, 0, stream>>>(x, n, c, done, first ? 1 : 0);
    switch (st->lanes) {
    case 8: cgGather<8><<<blocksOf(n, BLOCK/8), BLOCK, 0, stream>>>(x, Ax, mass, n, aa, c, done); break;
    case 32: cgGather<32><<<blocksOf(n, BLOCK/32), BLOCK, 0, stream>>>(x, Ax, mass, n, aa, c, done); break;
    default: cgGather<64><<<blocksOf(n, BLOCK/64), BLOCK, 0, stream>>>(x, Ax, mass, n, aa, c, done); break;
    }
}

bool converged(CentredState* st, hipStream_t stream) {
    double* s = st->hostScalars;
    check(hipMemcpyAsync(s, st->cg.scalars, NUM_SCALARS*sizeof(double), hipMemcpyDeviceToHost, stream), "download");
    check(hipStreamSynchronize(stream), "sync");
    return residualConverged(s);
}

template <typename S>
void centredCoupleT(CentredState* st, const S* src, const uint32_t* mask, const double* positions,
        const double* velocities, const double* otherForces, const int* particles, const double* mass,
        const double* ucm, const ParticleArgs& a, const double g[3], double* particleForce,
        unsigned long long* reactionFixed, uint32_t* touched, double* partials, hipStream_t stream) {
    int n = a.numParticles;
    if (n == 0)
        return;
    if (st->stencil == LBM_STENCIL_NEAREST) {
        centredKeys<<<blocksOf(n, BLOCK), BLOCK, 0, stream>>>(positions, particles, a, st->keys, st->values);
        size_t bytes = st->sortBytes;
        check(hipcub::DeviceRadixSort::SortPairs(st->sortTemp, bytes, st->keys, st->sortedKeys, st->values,
                st->sortedValues, n, 0, st->sortBits, stream), "radix sort");
        centredNearest<S><<<blocksOf(n, BLOCK), BLOCK, 0, stream>>>(src, mask, velocities, otherForces, particles, mass,
                ucm, a, g[0], g[1], g[2], st->sortedKeys, st->sortedValues, particleForce, reactionFixed, touched, partials);
        check(hipGetLastError(), "launch");
        return;
    }
    CgArgs& c = st->cg;
    double aa = 0.5*a.gamma;
    unsigned groups = blocksOf(n, BLOCK/st->lanes);
    switch (st->lanes) {
    case 8:
        cgSetup<S, 8><<<groups, BLOCK, 0, stream>>>(src, mask, positions, velocities, otherForces, particles, mass, ucm, a,
                g[0], g[1], g[2], c);
        break;
    case 32:
        cgSetup<S, 32><<<groups, BLOCK, 0, stream>>>(src, mask, positions, velocities, otherForces, particles, mass, ucm, a,
                g[0], g[1], g[2], c);
        break;
    default:
        cgSetup<S, 64><<<groups, BLOCK, 0, stream>>>(src, mask, positions, velocities, otherForces, particles, mass, ucm, a,
                g[0], g[1], g[2], c);
        break;
    }
    if (c.split)
        switch (st->lanes) {
        case 8: cgMarkShared<8><<<groups, BLOCK, 0, stream>>>(n, c); break;
        case 32: cgMarkShared<32><<<groups, BLOCK, 0, stream>>>(n, c); break;
        default: cgMarkShared<64><<<groups, BLOCK, 0, stream>>>(n, c); break;
        }
    cgMaxAbs<<<RED_BLOCKS, BLOCK, 0, stream>>>(c.F, n, c.partials, nullptr);
    multiply<S>(st, c.F, c.Ad, mass, aa, nullptr, 1, true, stream);
    cgInit<<<RED_BLOCKS, BLOCK, 0, stream>>>(n, c, st->partialsB);
    cgFinishInit<<<1, 64, 0, stream>>>(c.partials, st->partialsB, c.scalars);
    check(hipGetLastError(), "launch");
    // Each iteration starts with the new direction (d = z after cgInit), so that the last one computes none.  The
    // previous solve's count of iterations is enqueued at once (those after convergence do nothing), then the host
    // reads the residual and adds one iteration at a time until it has converged.
    const double* done = c.scalars + DONE;
    int iteration = 0;
    for (int batch = st->planned ? std::max(st->lastIterations, 1) : 1;; batch = 1) {
        for (int b = 0; b < batch; b++, iteration++) {
            if (iteration == 0)
                cgMaxAbs<<<RED_BLOCKS, BLOCK, 0, stream>>>(c.d, n, c.partials, done);
            else
                cgDirection<<<RED_BLOCKS, BLOCK, 0, stream>>>(n, c, done);
            multiply<S>(st, c.d, c.Ad, mass, aa, done, iteration & 1, false, stream);
            cgDot<<<RED_BLOCKS, BLOCK, 0, stream>>>(n, c, done);
            cgUpdate<<<RED_BLOCKS, BLOCK, 0, stream>>>(n, c, st->partialsB, done);
            cgFinishBeta<<<1, 64, 0, stream>>>(st->partialsB, c.scalars, done);
            check(hipGetLastError(), "launch");
        }
        if (converged(st, stream))
            break;
        if (iteration >= CG_MAX_ITERATIONS) {
            // cgForces, which clears the node sums and bits after a solve, does not run: clear them here, so that a
            // caller who goes on after the error starts the next solve from zero.
            const size_t sumBytes = 3*(size_t) st->numNodes*sizeof(unsigned long long);
            check(hipMemsetAsync(st->sums[0], 0, sumBytes, stream), "memset");
            check(hipMemsetAsync(st->sums[1], 0, sumBytes, stream), "memset");
            if (c.split) {
                const size_t bitBytes = (((size_t) st->numNodes + 31)/32)*sizeof(uint32_t);
                check(hipMemsetAsync(c.seen, 0, bitBytes, stream), "memset");
                check(hipMemsetAsync(c.shared, 0, bitBytes, stream), "memset");
            }
            throw std::runtime_error("LBMForce: the centred drag with the interpolation stencil did not converge in 1000 "
                    "iterations of the conjugate gradients");
        }
    }
    st->lastIterations = (int) st->hostScalars[ITERS];
    switch (st->lanes) {
    case 8: cgForces<8><<<groups, BLOCK, 0, stream>>>(n, a, c, particleForce, reactionFixed, touched, partials); break;
    case 32: cgForces<32><<<groups, BLOCK, 0, stream>>>(n, a, c, particleForce, reactionFixed, touched, partials); break;
    default: cgForces<64><<<groups, BLOCK, 0, stream>>>(n, a, c, particleForce, reactionFixed, touched, partials); break;
    }
    check(hipGetLastError(), "launch");
}

}  // namespace

void centredCouple(CentredState* st, int storage, const void* src, const uint32_t* mask, const double* positions,
        const double* velocities, const double* otherForces, const int* particles, const double* mass,
        const double* ucm, const ParticleArgs& a, const double g[3], double* particleForce,
        unsigned long long* reactionFixed, uint32_t* touched, double* partials, hipStream_t stream) {
    switch (storage) {
    case LBM_STORAGE_FP16:
        centredCoupleT<__half>(st, (const __half*) src, mask, positions, velocities, otherForces, particles, mass, ucm, a,
                g, particleForce, reactionFixed, touched, partials, stream);
        break;
    case LBM_STORAGE_BF16:
        centredCoupleT<__hip_bfloat16>(st, (const __hip_bfloat16*) src, mask, positions, velocities, otherForces,
                particles, mass, ucm, a, g, particleForce, reactionFixed, touched, partials, stream);
        break;
    case LBM_STORAGE_FP64:
        centredCoupleT<double>(st, (const double*) src, mask, positions, velocities, otherForces, particles, mass, ucm, a,
                g, particleForce, reactionFixed, touched, partials, stream);
        break;
    default:
        centredCoupleT<float>(st, (const float*) src, mask, positions, velocities, otherForces, particles, mass, ucm, a,
                g, particleForce, reactionFixed, touched, partials, stream);
        break;
    }
}

}  // namespace lbm
