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
#include "lbm_internal.h"

namespace lbm {

__host__ __device__ int stencilWidth(int stencil) {
    switch (stencil) {
    case LBM_STENCIL_TRILINEAR: return 2;
    case LBM_STENCIL_THREE_POINT: return 3;
    case LBM_STENCIL_KEYS: return 4;
    default: return 1;
    }
}

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

__device__ __forceinline__ bool solidAt(const uint32_t* mask, uint32_t node) {
    return mask != nullptr && isSolid(mask, node);
}

/* ------------------------------------------------------------------------------------------ */
__global__ void __launch_bounds__(BLOCK) reflectParticles(const double* __restrict__ positions, double* __restrict__ velocities,
        const int* __restrict__ particles, const double* __restrict__ mass, const uint32_t* __restrict__ mask,
        ParticleArgs a, double* __restrict__ partials) {
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
    __shared__ double lds[BLOCK][3];
    for (int k = 0; k < 3; k++)
        lds[threadIdx.x][k] = acc[k];
    __syncthreads();
    if (threadIdx.x == 0) {
        double s[3] = {0, 0, 0};
        for (int t = 0; t < BLOCK; t++)
            for (int k = 0; k < 3; k++)
                s[k] += lds[t][k];
        for (int k = 0; k < 3; k++)
            partials[3*blockIdx.x + k] = s[k];
    }
}

/* ------------------------------------------------------------------------------------------ */
/*
 * The coupling of one particle by a group of LANES lanes (one lane per stencil slot; every lane of the group calls
 * it, inactive ones with active = false).  x, v: the particle's position (nm) and velocity (nm/ps); index: its
 * coupled-particle index (the counter of its random numbers); m: its mass in lattice units.  The force (kJ/mol/nm) is
 * written to forceOut by the first lane; in a lattice step the reaction goes to the accumulators and, on solid nodes,
 * into wallAcc (lattice momentum x momentumScale).
 */
template <typename S, int LANES>
__device__ __forceinline__ void coupleGroup(const S* __restrict__ src, const uint32_t* __restrict__ mask,
        const double* x, const double* vel, uint32_t index, double m, bool active, int lane,
        const double* __restrict__ ucm, const ParticleArgs& a, double* forceOut,
        unsigned long long* __restrict__ reactionFixed, uint32_t* __restrict__ touched, double wallAcc[3]) {
    int w = stencilWidth(a.stencil);
    int numSlots = w*w*w;
    int slot = lane;
    bool inStencil = active && slot < numSlots;
    int size[3] = {(int) a.g.nx, (int) a.g.ny, (int) a.g.nz};
    int openAxis[3] = {a.g.openX, a.g.openY, a.g.openZ};
    double weight = 0.0;
    uint32_t node = 0;
    if (inStencil) {
        int sub[3] = {slot % w, (slot/w) % w, slot/(w*w)};
        int nodeIdx[3];
        weight = 1.0;
        for (int k = 0; k < 3; k++) {
            double s = wrapCoordinate(x[k], a.dx, size[k]);
            int ix[4];
            double wx[4];
            if (a.stencil == LBM_STENCIL_NEAREST) {
                ix[0] = nearestIndex(s, size[k]);
                wx[0] = 1.0;
            }
            else
                axisWeights(a.stencil, w, s, size[k], openAxis[k] != 0, ix, wx);
            nodeIdx[k] = ix[sub[k]];
            weight *= wx[sub[k]];
        }
        node = (uint32_t) (nodeIdx[0] + (int) a.g.nx*(nodeIdx[1] + (int) a.g.ny*nodeIdx[2]));
    }
    // Moments of the stencil node, as computeMoments gives them: solid nodes have rho = 0 and are skipped.
    double ux = 0, uy = 0, uz = 0;
    bool solid = inStencil && solidAt(mask, node);
    if (inStencil && !solid) {
        uint32_t ni, nj, nk;
        nodeCoords(a.g, node, ni, nj, nk);
        Neighbours nb = neighbours(a.g, ni, nj, nk);
        typename StorageTraits<S>::Real f[NQ];
        pullNode(src, a.g.n, node, nb, f);
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
        if (a.removal) {
            jx -= rho*ucm[0];
            jy -= rho*ucm[1];
            jz -= rho*ucm[2];
        }
        if (rho > 0) {
            ux = jx*(weight/rho);
            uy = jy*(weight/rho);
            uz = jz*(weight/rho);
        }
    }
#pragma unroll
    for (int off = 1; off < LANES; off <<= 1) {
        ux += __shfl_xor(ux, off, LANES);
        uy += __shfl_xor(uy, off, LANES);
        uz += __shfl_xor(uz, off, LANES);
    }
    // Every lane of the group now holds u; each computes the force (identical values).
    double f[3] = {0, 0, 0};
    if (active) {
        double noise[4] = {0, 0, 0, 0};
        if (a.randomForce)
            drawNormals<double, 4>(a.seed, TAG_PARTICLE, index, a.step, noise);
        double u[3] = {ux, uy, uz};
        double amp = sqrt(2.0*a.gamma*m*a.kT);
        for (int k = 0; k < 3; k++) {
            double v = vel[k]/a.velocityScale;
            f[k] = (v - u[k])*(-a.gamma*m);
            if (a.randomForce)
                f[k] += noise[k]*amp;
        }
        if (lane == 0)
            for (int k = 0; k < 3; k++)
                forceOut[k] = f[k]*a.forceScale;
    }
    if (a.isStep && inStencil) {
        // The reaction -w F: on a fluid node into its fixed-point accumulators (flagged in the touched mask, which
        // the fused pass reads and clears); on a solid node it leaves the fluid and goes to the wall
        // (ReferenceLBMKernels.cpp:959-971).
        bool any = false;
        for (int k = 0; k < 3; k++) {
            double r = -weight*f[k];
            if (solid)
                wallAcc[k] = r*a.momentumScale;
            else {
                long long fixed = llrint(r*FIXED_SCALE);
                if (fixed != 0) {
                    atomicAdd(reactionFixed + (size_t) k*a.g.n + node, (unsigned long long) fixed);
                    any = true;
                }
            }
        }
        if (any)
            atomicOr(touched + (node >> 5), 1u << (node & 31));
    }
}

template <typename S, int LANES>
__global__ void __launch_bounds__(BLOCK) coupleParticles(const S* __restrict__ src, const uint32_t* __restrict__ mask,
        const double* __restrict__ positions, const double* __restrict__ velocities, const int* __restrict__ particles,
        const double* __restrict__ mass, const double* __restrict__ ucm, ParticleArgs a, double* __restrict__ particleForce,
        unsigned long long* __restrict__ reactionFixed, uint32_t* __restrict__ touched, double* __restrict__ partials) {
    constexpr int GROUPS = BLOCK/LANES;
    int lane = threadIdx.x % LANES;
    int i = blockIdx.x*GROUPS + threadIdx.x/LANES;
    bool active = i < a.numParticles;
    int part = active ? particles[i] : 0;
    double m = active ? mass[i] : 0.0;
    double wallAcc[3] = {0, 0, 0};
    coupleGroup<S, LANES>(src, mask, positions + 3*(size_t) part, velocities + 3*(size_t) part, (uint32_t) i, m, active,
            lane, ucm, a, particleForce + 3*(size_t) (active ? i : 0), reactionFixed, touched, wallAcc);
    if (partials != nullptr) {
        __shared__ double lds[BLOCK][3];
        for (int k = 0; k < 3; k++)
            lds[threadIdx.x][k] = wallAcc[k];
        __syncthreads();
        if (threadIdx.x == 0) {
            double s[3] = {0, 0, 0};
            for (int t = 0; t < BLOCK; t++)
                for (int k = 0; k < 3; k++)
                    s[k] += lds[t][k];
            for (int k = 0; k < 3; k++)
                partials[3*blockIdx.x + k] = s[k];
        }
    }
}

/* Batched replicas: the coupling of every particle of every replica in one launch (particles of one replica are
   contiguous; particleReplica gives the replica of each). */
template <int LANES>
__global__ void __launch_bounds__(BLOCK) coupleBatch(const ReplicaDesc* __restrict__ table,
        const uint16_t* __restrict__ particleReplica, int numParticles, const double* __restrict__ positions,
        const double* __restrict__ velocities, const double* __restrict__ mass, double* __restrict__ forces,
        const int64_t* __restrict__ stepCounter) {
    constexpr int GROUPS = BLOCK/LANES;
    int lane = threadIdx.x % LANES;
    int p = blockIdx.x*GROUPS + threadIdx.x/LANES;
    bool active = p < numParticles;
    int pp = active ? p : 0;
    const ReplicaDesc* d = table + particleReplica[pp];
    ParticleArgs a = d->part;
    a.step = *stepCounter;
    a.isStep = 1;
    const float* src = (a.step & 1) ? d->popB : d->popA;
    double wallAcc[3] = {0, 0, 0};
    coupleGroup<float, LANES>(src, nullptr, positions + 3*(size_t) pp, velocities + 3*(size_t) pp,
            (uint32_t) (pp - d->particleBegin), active ? mass[pp] : 0.0, active, lane, nullptr, a, forces + 3*(size_t) pp,
            d->reactionFixed, d->touched, wallAcc);
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
        const uint32_t* mask, const ParticleArgs& a, double* partials, hipStream_t stream) {
    if (a.numParticles == 0)
        return;
    reflectParticles<<<gridFor(a.numParticles), BLOCK, 0, stream>>>(positions, velocities, particles, mass, mask, a, partials);
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
        const int* particles, const double* mass, const double* ucm, const ParticleArgs& a, double* particleForce,
        unsigned long long* reactionFixed, uint32_t* touched, double* partials, hipStream_t stream) {
    unsigned grid = (unsigned) coupleBlocks(a.stencil, a.numParticles);
    const S* s = (const S*) src;
    switch (coupleLanes(a.stencil)) {
    case 8:
        coupleParticles<S, 8><<<grid, BLOCK, 0, stream>>>(s, mask, positions, velocities, particles, mass, ucm, a,
                particleForce, reactionFixed, touched, partials);
        break;
    case 32:
        coupleParticles<S, 32><<<grid, BLOCK, 0, stream>>>(s, mask, positions, velocities, particles, mass, ucm, a,
                particleForce, reactionFixed, touched, partials);
        break;
    case 64:
        coupleParticles<S, 64><<<grid, BLOCK, 0, stream>>>(s, mask, positions, velocities, particles, mass, ucm, a,
                particleForce, reactionFixed, touched, partials);
        break;
    default:
        coupleParticles<S, 1><<<grid, BLOCK, 0, stream>>>(s, mask, positions, velocities, particles, mass, ucm, a,
                particleForce, reactionFixed, touched, partials);
        break;
    }
}

void launchCouple(int storage, const void* src, const uint32_t* mask, const double* positions, const double* velocities,
        const int* particles, const double* mass, const double* ucm, const ParticleArgs& a, double* particleForce,
        unsigned long long* reactionFixed, uint32_t* touched, double* partials, hipStream_t stream) {
    if (a.numParticles == 0)
        return;
    switch (storage) {
    case LBM_STORAGE_FP16:
        launchCoupleT<__half>(src, mask, positions, velocities, particles, mass, ucm, a, particleForce, reactionFixed,
                touched, partials, stream);
        break;
    case LBM_STORAGE_BF16:
        launchCoupleT<__hip_bfloat16>(src, mask, positions, velocities, particles, mass, ucm, a, particleForce,
                reactionFixed, touched, partials, stream);
        break;
    case LBM_STORAGE_FP64:
        launchCoupleT<double>(src, mask, positions, velocities, particles, mass, ucm, a, particleForce, reactionFixed,
                touched, partials, stream);
        break;
    default:
        launchCoupleT<float>(src, mask, positions, velocities, particles, mass, ucm, a, particleForce, reactionFixed,
                touched, partials, stream);
        break;
    }
}

void launchCoupleBatch(int stencil, const ReplicaDesc* table, const uint16_t* particleReplica, int numParticles,
        const double* positions, const double* velocities, const double* mass, double* forces,
        const int64_t* stepCounter, hipStream_t stream) {
    if (numParticles == 0)
        return;
    unsigned grid = (unsigned) coupleBlocks(stencil, numParticles);
    switch (coupleLanes(stencil)) {
    case 8:
        coupleBatch<8><<<grid, BLOCK, 0, stream>>>(table, particleReplica, numParticles, positions, velocities, mass,
                forces, stepCounter);
        break;
    case 32:
        coupleBatch<32><<<grid, BLOCK, 0, stream>>>(table, particleReplica, numParticles, positions, velocities, mass,
                forces, stepCounter);
        break;
    case 64:
        coupleBatch<64><<<grid, BLOCK, 0, stream>>>(table, particleReplica, numParticles, positions, velocities, mass,
                forces, stepCounter);
        break;
    default:
        coupleBatch<1><<<grid, BLOCK, 0, stream>>>(table, particleReplica, numParticles, positions, velocities, mass,
                forces, stepCounter);
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

}  // namespace lbm
