/*
 * Fluid kernels of the gfx950 LBMForce platform.
 *
 * Layout in HBM: two population arrays A and B of 19*N values (N = nx*ny*nz), [z][y][q][x]: each
 * lattice row holds its 19 populations one after the other, x contiguous (lbm_device.h), holding
 * the POST-collision deviations C_q - w_q of every node at the node itself.  One lattice step is one pass of fusedStep over all nodes: pull the 19
 * values C_q(x - c_q) from A (the pre-collision populations of x), form the moments, collide with
 * Guo forcing (and the fluctuating part), write the 19 results to B at x.  Every population is
 * read once and written once: 2*19*sizeof(storage) bytes per node per step.
 *
 * Walls and faces are folded into the array instead of into the hot kernel: after the pass, the
 * slots that a fluid node will pull from a solid node (bounce-back) or from across an open face
 * (regularized rebuild) are overwritten in B by two small sparse kernels, so the next pass pulls
 * the right values with no test at all.  Reference counterparts: computeFluidMoments and
 * collideAndStream (lbmFluid.cc:89-139, 230-310; ReferenceLBMKernels.cpp:430-469, 1066-1120),
 * bounceBack and computeWallExchange (lbmFluid.cc:358-403; ReferenceLBMKernels.cpp:1156-1205),
 * applyBoundaries (lbmFluid.cc:437-602; ReferenceLBMKernels.cpp:1207-1343).
 */
#include "lbm_fused.h"

namespace lbm {

/* ------------------------------------------------------------------------------------------ */
/* The fused pull-stream + regularized collision + Guo forcing (+ Philox noise) pass.          */
/* ------------------------------------------------------------------------------------------ */
template <typename S, bool SOLIDS, bool REACTION, bool FLUCT, int THREADS, bool WIDE>
__global__ void __launch_bounds__(THREADS) fusedStep(const S* __restrict__ src, S* __restrict__ dst,
        const uint32_t* __restrict__ mask, unsigned long long* __restrict__ reactionFixed, uint32_t* __restrict__ touched,
        const double* __restrict__ ucm, StepArgs a) {
    uint32_t node = blockIdx.x*THREADS + threadIdx.x;
    if (node >= a.g.n)
        return;
    fusedNode<S, SOLIDS, REACTION, FLUCT, WIDE>(src, dst, mask, reactionFixed, touched, ucm, a, node);
}

/* The storage value v as the load type (FP32 for FP16/BF16). */
template <typename S>
__device__ __forceinline__ typename StorageTraits<S>::Real asReal(S v) {
    if constexpr (std::is_same<S, __half>::value)
        return (float) __builtin_bit_cast(_Float16, v);
    else if constexpr (std::is_same<S, __hip_bfloat16>::value)
        return __uint_as_float(((uint32_t) __builtin_bit_cast(uint16_t, v)) << 16);
    else
        return v;
}

/* c_qy + 1 and c_qz + 1 of population q, two bits each, for a q known only at run time. */
constexpr uint64_t packedDirections(const int* c) {
    uint64_t bits = 0;
    for (int q = 0; q < NQ; q++)
        bits |= (uint64_t) (c[q] + 1) << (2*q);
    return bits;
}
constexpr uint64_t PACKED_CY = packedDirections(CY);
constexpr uint64_t PACKED_CZ = packedDirections(CZ);

__device__ __forceinline__ int directionOf(uint64_t packed, int q) {
    return (int) ((packed >> (2*q)) & 3u) - 1;
}

/*
 * The fused pass with the pulls staged through LDS.  A work-group of ROW_THREADS threads covers as many consecutive
 * nodes, in pieces of SEG nodes along x: one piece of a row when SEG = ROW_THREADS (nx a multiple of it), else
 * ROW_THREADS/SEG whole rows (SEG = nx).  For each population q and piece, the source run [x0 - V, x0 + SEG + V) of the
 * row (j - c_qy, k - c_qz) is loaded in aligned 16-byte chunks of V values (wrapped along x and across the periodic
 * faces) into LDS, and each thread takes its value at x - c_qx from there: the global loads are 16 bytes per lane and
 * aligned, NQ*(SEG/V + 2) chunks per piece instead of NQ single-value loads per node.  Same arithmetic as fusedStep, so
 * bit-identical results.  Requires arrays of at most 4 GiB (32-bit offsets) and nx a multiple of V.
 */
template <typename S, bool SOLIDS, bool REACTION, bool FLUCT, int SEG>
__global__ void __launch_bounds__(ROW_THREADS) __attribute__((amdgpu_waves_per_eu(sizeof(S) == 8 ? 1 : 8)))
fusedRows(const S* __restrict__ src, S* __restrict__ dst,
        const uint32_t* __restrict__ mask, unsigned long long* __restrict__ reactionFixed, uint32_t* __restrict__ touched,
        const double* __restrict__ ucm, StepArgs a) {
    typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
    constexpr int T = ROW_THREADS;
    constexpr int V = 16/(int) sizeof(S);
    constexpr int PIECES = T/SEG;
    constexpr int CH = SEG/V + 2;
    constexpr int TOTAL = NQ*PIECES*CH;
    constexpr int ITER = (TOTAL + T - 1)/T;
    static_assert(SEG % V == 0 && T % SEG == 0, "pieces of whole chunks");
    __shared__ u32x4 stage[TOTAL];
    const uint32_t base = blockIdx.x*T;
    const int nx = (int) a.g.nx, ny = (int) a.g.ny, nz = (int) a.g.nz;
    const auto* chunks = asGlobal(reinterpret_cast<const u32x4*>(src));
    u32x4 v[ITER];
#pragma unroll
    for (int it = 0; it < ITER; it++) {
        v[it] = u32x4{0u, 0u, 0u, 0u};
        const int c = (int) threadIdx.x + it*T;
        if (c < TOTAL) {
            const int q = c/(PIECES*CH);
            const int rem = c - q*(PIECES*CH);
            const int p = rem/CH;
            const int ch = rem - p*CH;
            const uint32_t first = base + (uint32_t) (p*SEG);
            if (first < a.g.n) {
                uint32_t i0, j, k;
                nodeCoords(a.g, first, i0, j, k);
                int js = (int) j - directionOf(PACKED_CY, q);
                int ks = (int) k - directionOf(PACKED_CZ, q);
                js = js < 0 ? js + ny : (js >= ny ? js - ny : js);
                ks = ks < 0 ? ks + nz : (ks >= nz ? ks - nz : ks);
                int xs = (int) i0 - V + ch*V;
                xs = xs < 0 ? xs + nx : (xs >= nx ? xs - nx : xs);
                const uint32_t elem = (((uint32_t) ks*(uint32_t) ny + (uint32_t) js)*NQ + (uint32_t) q)*(uint32_t) nx +
                                      (uint32_t) xs;
                v[it] = __builtin_nontemporal_load(chunks + elem/V);
            }
        }
    }
#pragma unroll
    for (int it = 0; it < ITER; it++) {
        const int c = (int) threadIdx.x + it*T;
        if (c < TOTAL)
            stage[c] = v[it];
    }
    __syncthreads();
    const uint32_t node = base + threadIdx.x;
    if (node >= a.g.n)
        return;
    const uint32_t touchedWord = takeTouchedWord<REACTION>(touched, node);
    if constexpr (SOLIDS)
        if (isSolid(mask, node))
            return;
    const int p = (int) threadIdx.x/SEG, t = (int) threadIdx.x - p*SEG;
    const S* staged = reinterpret_cast<const S*>(stage);
    typename StorageTraits<S>::Real fr[NQ];
    staticFor<NQ>([&](auto I) {
        constexpr int q = decltype(I)::value;
        fr[q] = asReal(staged[(q*PIECES + p)*CH*V + t + V - CX[q]]);
    });
    uint32_t i, j, k;
    nodeCoords(a.g, node, i, j, k);
    collideNode<S, REACTION, FLUCT, false>(fr, dst, reactionFixed, touchedWord, ucm, a, node,
            elementOf<uint32_t>(a.g, i, j, k));
}

/*
 * The same pass over every replica of a batch in one launch: work-group b covers THREADS nodes of replica
 * tileReplica[b].  The step index comes from device memory, so one captured launch serves every step.
 */
template <bool FLUCT, int THREADS>
__global__ void __launch_bounds__(THREADS) fusedBatch(const ReplicaDesc* __restrict__ table,
        const uint16_t* __restrict__ tileReplica, const int64_t* __restrict__ stepCounter, int xcdOrder) {
    // Runs of 64 neighbouring tiles per XCD: the batched config4 pass (17.6 M nodes in 64 replicas) 0.476 -> 0.465 ms,
    // its reads' overfetch at tile borders shared in one L2 (runs of 128: 0.468 ms; of 256 or one run per XCD: none).
    const uint32_t tile = xcdOrder ? xcdChunkedTile(blockIdx.x, gridDim.x) : blockIdx.x;
    const ReplicaDesc* d = table + tileReplica[tile];
    uint32_t node = (tile - d->tileBegin)*THREADS + threadIdx.x;
    StepArgs a = d->step;
    if (node >= a.g.n)
        return;
    a.step = *stepCounter;
    a.roundSeed = 0;
    const float* src = (a.step & 1) ? d->popB : d->popA;
    float* dst = (a.step & 1) ? d->popA : d->popB;
    fusedNode<float, false, true, FLUCT>(src, dst, nullptr, d->reactionFixed, d->touched, nullptr, a, node);
}

void launchFusedBatch(int threads, bool fluct, const ReplicaDesc* table, const uint16_t* tileReplica, uint32_t numTiles,
        const int64_t* stepCounter, bool xcdOrder, hipStream_t stream) {
#define LBM_BATCH_CASE(T) \
    case T: \
        if (fluct) fusedBatch<true, T><<<numTiles, T, 0, stream>>>(table, tileReplica, stepCounter, xcdOrder); \
        else fusedBatch<false, T><<<numTiles, T, 0, stream>>>(table, tileReplica, stepCounter, xcdOrder); \
        break;
    switch (threads) {
    LBM_BATCH_CASE(256)
    LBM_BATCH_CASE(1024)
    default:
    LBM_BATCH_CASE(512)
    }
#undef LBM_BATCH_CASE
}

/* ------------------------------------------------------------------------------------------ */
/* FP64 reductions: drho and j summed over the fluid nodes (removal, drift checks), max |j/rho|^2 */
/* (Mach).  Fixed grid, fixed order: reproducible.                                              */
/* ------------------------------------------------------------------------------------------ */
__device__ __forceinline__ double waveSum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
        v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ double waveMax(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
        v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}

/* Sum (or max) of K values over a block of BLOCK threads, result in thread 0. */
template <int K>
__device__ __forceinline__ void blockReduce(double v[K], const bool isMax[K]) {
    __shared__ double lds[BLOCK/64][K];
    int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int c = 0; c < K; c++)
        v[c] = isMax[c] ? waveMax(v[c]) : waveSum(v[c]);
    if (lane == 0)
#pragma unroll
        for (int c = 0; c < K; c++)
            lds[wave][c] = v[c];
    __syncthreads();
    if (threadIdx.x == 0)
#pragma unroll
        for (int c = 0; c < K; c++) {
            double s = lds[0][c];
            for (int w = 1; w < BLOCK/64; w++)
                s = isMax[c] ? fmax(s, lds[w][c]) : s + lds[w][c];
            v[c] = s;
        }
}

template <typename S, bool SOLIDS>
__global__ void __launch_bounds__(BLOCK) momentReduce(const S* __restrict__ src, const uint32_t* __restrict__ mask,
        Geometry g, double* __restrict__ partials) {
    using R = typename StorageTraits<S>::Real;
    double acc[5] = {0, 0, 0, 0, 0};
    for (uint32_t node = blockIdx.x*BLOCK + threadIdx.x; node < g.n; node += gridDim.x*BLOCK) {
        if constexpr (SOLIDS)
            if (isSolid(mask, node))
                continue;
        uint32_t i, j, k;
        nodeCoords(g, node, i, j, k);
        R f[NQ];
        pullNode(src, g, node, i, j, k, f);
        double dr = 0, jx = 0, jy = 0, jz = 0;
        staticFor<NQ>([&](auto I) {
            constexpr int q = decltype(I)::value;
            double v = (double) f[q];
            dr += v;
            jx = addTerm<CX[q]>(jx, v);
            jy = addTerm<CY[q]>(jy, v);
            jz = addTerm<CZ[q]>(jz, v);
        });
        acc[0] += dr;
        acc[1] += jx;
        acc[2] += jy;
        acc[3] += jz;
        double rho = 1.0 + dr;
        acc[4] = fmax(acc[4], (jx*jx + jy*jy + jz*jz)/(rho*rho));
    }
    const bool isMax[5] = {false, false, false, false, true};
    blockReduce<5>(acc, isMax);
    if (threadIdx.x == 0)
        for (int c = 0; c < 5; c++)
            partials[5*blockIdx.x + c] = acc[c];
}

/* totals = {sum drho, px, py, pz, max speed^2}; ucm = p/(numFluid + sum drho). */
__global__ void __launch_bounds__(BLOCK) finishReduce(const double* __restrict__ partials, double numFluid,
        double* __restrict__ totals, double* __restrict__ ucm) {
    double acc[5] = {0, 0, 0, 0, 0};
    for (int b = threadIdx.x; b < REDUCE_BLOCKS; b += BLOCK) {
        for (int c = 0; c < 4; c++)
            acc[c] += partials[5*b + c];
        acc[4] = fmax(acc[4], partials[5*b + 4]);
    }
    const bool isMax[5] = {false, false, false, false, true};
    blockReduce<5>(acc, isMax);
    if (threadIdx.x == 0) {
        for (int c = 0; c < 5; c++)
            totals[c] = acc[c];
        double mass = numFluid + acc[0];
        ucm[0] = acc[1]/mass;
        ucm[1] = acc[2]/mass;
        ucm[2] = acc[3]/mass;
    }
}

/* Sum of count xyz partials (wall momentum), one block. */
__global__ void __launch_bounds__(BLOCK) reduceWall(const double* __restrict__ partials, int count, double* __restrict__ out) {
    double acc[3] = {0, 0, 0};
    for (int b = threadIdx.x; b < count; b += BLOCK)
        for (int c = 0; c < 3; c++)
            acc[c] += partials[3*b + c];
    const bool isMax[3] = {false, false, false};
    blockReduce<3>(acc, isMax);
    if (threadIdx.x == 0)
        for (int c = 0; c < 3; c++)
            out[c] = acc[c];
}

/* ------------------------------------------------------------------------------------------ */
/* Halfway bounce-back (lbmFluid.cc:358-403; ReferenceLBMKernels.cpp:1156-1205).               */
/* For a fluid node x and a link q with x + c_q solid: the population C_q(x) that x sent into   */
/* the solid node comes back as its population along -c_q, which x pulls from the slot         */
/* B[opp q][x + c_q] of the solid node: write it there.  The wall receives 2 c_q (C_q + w_q);  */
/* the part of w_q is constant and added on the host.                                           */
/* ------------------------------------------------------------------------------------------ */
template <typename S>
__global__ void __launch_bounds__(BLOCK) bounceBack(S* __restrict__ dst, const int* __restrict__ wallNodes,
        const int* __restrict__ wallLinks, int count, Geometry g, double* __restrict__ partials) {
    double acc[3] = {0, 0, 0};
    int b = blockIdx.x*BLOCK + threadIdx.x;
    if (b < count) {
        uint32_t node = (uint32_t) wallNodes[b];
        int links = wallLinks[b];
        uint32_t i, j, k;
        nodeCoords(g, node, i, j, k);
        const NodeSlots at = nodeSlots(g, node, i, j, k);
        staticFor<NQ - 1>([&](auto I) {
            constexpr int q = decltype(I)::value + 1;
            if (links & (1 << q)) {
                S v = dst[at.own(q)];
                dst[at.template pushed<q, OPP[q]>()] = v;      // population opp(q) of the solid node x + c_q
                double dv = (double) loadPop(&v);
                acc[0] = addTerm<CX[q]>(acc[0], 2.0*dv);
                acc[1] = addTerm<CY[q]>(acc[1], 2.0*dv);
                acc[2] = addTerm<CZ[q]>(acc[2], 2.0*dv);
            }
        });
    }
    const bool isMax[3] = {false, false, false};
    blockReduce<3>(acc, isMax);
    if (threadIdx.x == 0)
        for (int c = 0; c < 3; c++)
            partials[3*blockIdx.x + c] = acc[c];
}

/* ------------------------------------------------------------------------------------------ */
/* Regularized walls and open faces (ReferenceLBMKernels.cpp:1207-1343; lbmFluid.cc:437-602),   */
/* one thread per boundary node, in FP64.  The node recomputes its pre-collision moments from   */
/* A (and the random numbers of its collision from their counter), reads the values that       */
/* arrived in B, rebuilds its unknown populations and writes each into the slot it will pull   */
/* it from next step: B[q][x - c_q], a slot of a solid node or across an open face, which no    */
/* other node pulls.  All reads of a thread precede its writes, and no two threads share a slot.*/
/* ------------------------------------------------------------------------------------------ */
__device__ __forceinline__ void equilibriumDeviationD(double dr, double ux, double uy, double uz, double out[NQ]) {
    double rho = 1.0 + dr, uu = ux*ux + uy*uy + uz*uz;
#pragma unroll
    for (int q = 0; q < NQ; q++) {
        double cu = CX[q]*ux + CY[q]*uy + CZ[q]*uz;
        out[q] = WQ[q]*(dr + rho*(3.0*cu + 4.5*cu*cu - 1.5*uu));
    }
}

__device__ __forceinline__ void guoD(double ux, double uy, double uz, double fx, double fy, double fz, double out[NQ]) {
    double uf = ux*fx + uy*fy + uz*fz;
#pragma unroll
    for (int q = 0; q < NQ; q++) {
        double cu = CX[q]*ux + CY[q]*uy + CZ[q]*uz;
        double cf = CX[q]*fx + CY[q]*fy + CZ[q]*fz;
        out[q] = WQ[q]*(3.0*(cf - uf) + 9.0*cu*cf);
    }
}

__device__ __forceinline__ void regularizedD(const double pi[6], double out[NQ]) {
#pragma unroll
    for (int q = 0; q < NQ; q++) {
        double hxx = CX[q]*CX[q] - 1.0/3.0, hyy = CY[q]*CY[q] - 1.0/3.0, hzz = CZ[q]*CZ[q] - 1.0/3.0;
        out[q] = 4.5*WQ[q]*(hxx*pi[0] + hyy*pi[1] + hzz*pi[2] + 2.0*(CX[q]*CY[q]*pi[3] + CX[q]*CZ[q]*pi[4] + CY[q]*CZ[q]*pi[5]));
    }
}

template <typename S>
__global__ void __launch_bounds__(BLOCK) boundaryKernel(const S* __restrict__ src, S* __restrict__ dst,
        const BoundaryNode* __restrict__ list, const double* __restrict__ ucm, BoundaryArgs a,
        double* __restrict__ partials) {
    double acc[3] = {0, 0, 0};
    int b = blockIdx.x*BLOCK + threadIdx.x;
    if (b < a.count) {
        BoundaryNode bn = list[b];
        uint32_t node = (uint32_t) bn.node;
        int unknown = bn.unknown, solid = bn.solid, kind = bn.kindFace & 7, face = (bn.kindFace >> 3) - 1;
        uint32_t i, j, k;
        nodeCoords(a.g, node, i, j, k);
        const NodeSlots at = nodeSlots(a.g, node, i, j, k);
        // Pre-collision moments of the step (as the fused pass computed them).
        double f[NQ], post[NQ], own[NQ];
        staticFor<NQ>([&](auto I) {
            constexpr int q = decltype(I)::value;
            f[q] = (double) loadPop(src + at.template pulled<q>());
            post[q] = (double) loadPop(dst + at.template pulled<q>());
            own[q] = (double) loadPop(dst + at.own(q));
        });
        Moments<double> m = nodeMoments(f);
        double pi[6];
        piNeq(m, pi);
        double rho0 = 1.0 + m.dr;
        double jm[3] = {m.jx, m.jy, m.jz};
        if (a.removal)
            for (int c = 0; c < 3; c++)
                jm[c] -= rho0*ucm[c];
        const double g[3] = {a.gx, a.gy, a.gz};
        double rb = rho0, dr = m.dr;
        double ub[3];
        for (int c = 0; c < 3; c++)
            ub[c] = (jm[c] + 0.5*(rb*g[c]))/rb;
        if (kind == 0 || kind == 3)
            ub[0] = ub[1] = ub[2] = 0.0;
        if (kind == 1 || kind == 4)
            for (int c = 0; c < 3; c++)
                ub[c] = a.faceVelocity[face][c];
        if (kind == 2 || kind == 3 || kind == 4) {
            dr = a.faceDensity[face] - 1.0;
            rb = a.faceDensity[face];
        }
        if (kind == 2) {
            int axis = face/2;
            double inward = face % 2 == 0 ? 1.0 : -1.0, sum = 0.0;
#pragma unroll
            for (int q = 0; q < NQ; q++)
                sum += (unknown & (1 << q)) ? post[OPP[q]] : post[q];
            ub[axis] = 0.5*(inward*(dr - sum)/rb + ub[axis]);
        }
        double Fb[3] = {g[0]*rb, g[1]*rb, g[2]*rb};
        double eq[NQ], neq[NQ], srcTerm[NQ], xi[NQ];
        equilibriumDeviationD(dr, ub[0], ub[1], ub[2], eq);
        regularizedD(pi, neq);
        guoD(ub[0], ub[1], ub[2], Fb[0], Fb[1], Fb[2], srcTerm);
        double omega = a.omega;
        if (a.fluct) {
            double r[20];
            double cW = (1.0 - omega)*(1.0 - omega)/(omega*(2.0 - omega));
            if (kind == 0 || kind == 4) {
                // Equilibrium noise minus c times the stress part of the node's own collision (D3Q19.h:194-208).
                drawNormals<double, 19>(a.seed, TAG_BOUNDARY, node, a.step, r);
                double camp = sqrt(a.mu*rb), samp = sqrt(a.mu*rb*fmax(0.0, 1.0 - cW));
                double s[6], gq[NQ];
                stressNoise(r, samp, s);
                ghostNoise(r, camp, gq);
                double rf[16], own[NQ];
                drawNormals<double, 8>(a.seed, TAG_FLUID, node, a.step, rf);
                stressNoisePopulations(rf, sqrt(a.mu*rho0*omega*(2.0 - omega)), own);
#pragma unroll
                for (int q = 0; q < NQ; q++) {
                    double h = (CX[q]*CX[q] - 1.0/3.0)*s[0] + (CY[q]*CY[q] - 1.0/3.0)*s[1] + (CZ[q]*CZ[q] - 1.0/3.0)*s[2]
                            + 2.0*(CX[q]*CY[q]*s[3] + CX[q]*CZ[q]*s[4] + CY[q]*CZ[q]*s[5]);
                    xi[q] = WQ[q]*(4.5*h + gq[q] + camp*(r[15] + 1.7320508075688772*(CX[q]*r[16] + CY[q]*r[17] + CZ[q]*r[18])))
                            - cW*own[q];
                }
            }
            else {
                drawNormals<double, 15>(a.seed, TAG_BOUNDARY, node, a.step, r);
                double s[6], gq[NQ];
                stressNoise(r, sqrt(a.mu*rb*omega*(2.0 - omega)), s);
                ghostNoise(r, sqrt(a.mu*rb), gq);
#pragma unroll
                for (int q = 0; q < NQ; q++) {
                    double h = (CX[q]*CX[q] - 1.0/3.0)*s[0] + (CY[q]*CY[q] - 1.0/3.0)*s[1] + (CZ[q]*CZ[q] - 1.0/3.0)*s[2]
                            + 2.0*(CX[q]*CY[q]*s[3] + CX[q]*CZ[q]*s[4] + CY[q]*CZ[q]*s[5]);
                    xi[q] = WQ[q]*(4.5*h + gq[q]);
                }
            }
        }
        else
            for (int q = 0; q < NQ; q++)
                xi[q] = 0.0;
        if (kind == 0 || kind == 1) {
            // Mass balance (ReferenceLBMKernels.cpp:1290-1325), linear in drho.
            double e0[NQ], sg[NQ];
            equilibriumDeviationD(0.0, ub[0], ub[1], ub[2], e0);
            guoD(ub[0], ub[1], ub[2], g[0], g[1], g[2], sg);
            double num = 0.0, den = 0.0;
#pragma unroll
            for (int q = 1; q < NQ; q++) {
                if (!(unknown & (1 << q)))
                    continue;
                double arriving, flux = 0.0;
                if (solid & (1 << q))
                    arriving = own[OPP[q]];
                else if (!(unknown & (1 << OPP[q]))) {
                    arriving = post[OPP[q]];
                    flux = 6.0*WQ[q]*(CX[q]*ub[0] + CY[q]*ub[1] + CZ[q]*ub[2]);
                }
                else
                    continue;
                double rest = (1.0 - omega)*neq[q] + xi[q];
                num += arriving - rest - e0[q] - 0.5*sg[q] + flux;
                den += WQ[q] + e0[q] + 0.5*sg[q] - flux;
            }
            if (den != 0) {
                dr = num/den;
                rb = 1.0 + dr;
                for (int c = 0; c < 3; c++)
                    Fb[c] = g[c]*rb;
                equilibriumDeviationD(dr, ub[0], ub[1], ub[2], eq);
                guoD(ub[0], ub[1], ub[2], Fb[0], Fb[1], Fb[2], srcTerm);
            }
        }
        uint32_t rnd = mixBits(node ^ a.roundSeed);
        staticFor<NQ - 1>([&](auto I) {
            constexpr int q = decltype(I)::value + 1;
            if (unknown & (1 << q)) {
                double value = eq[q] + (1.0 - omega)*neq[q] + 0.5*srcTerm[q] + xi[q];
                if (solid & (1 << q)) {
                    double t = own[OPP[q]] + value + 2.0*WQ[q];
                    acc[0] = addTerm<-CX[q]>(acc[0], t);
                    acc[1] = addTerm<-CY[q]>(acc[1], t);
                    acc[2] = addTerm<-CZ[q]>(acc[2], t);
                }
                uint32_t r = (rnd ^ (uint32_t) (q*0x9E3779B9u))*0x85EBCA6Bu;
                storePop(dst + at.template pulled<q>(), (typename StorageTraits<S>::Real) value, r ^ (r >> 13));
            }
        });
    }
    const bool isMax[3] = {false, false, false};
    blockReduce<3>(acc, isMax);
    if (threadIdx.x == 0)
        for (int c = 0; c < 3; c++)
            partials[3*blockIdx.x + c] = acc[c];
}

/* ------------------------------------------------------------------------------------------ */
/* State transfer: gather (pull into [q*N + node] doubles), scatter (the inverse), fields.      */
/* ------------------------------------------------------------------------------------------ */
template <typename S>
__global__ void __launch_bounds__(BLOCK) gatherState(const S* __restrict__ src, double* __restrict__ out, Geometry g) {
    uint32_t node = blockIdx.x*BLOCK + threadIdx.x;
    if (node >= g.n)
        return;
    uint32_t i, j, k;
    nodeCoords(g, node, i, j, k);
    const NodeSlots at = nodeSlots(g, node, i, j, k);
    staticFor<NQ>([&](auto I) {
        constexpr int q = decltype(I)::value;
        out[(size_t) q*g.n + node] = (double) loadPop(src + at.template pulled<q>());
    });
}

template <typename S>
__device__ __forceinline__ void storeNearest(S* p, double v) {
    if constexpr (std::is_same<S, __half>::value)
        *p = __float2half_rn((float) v);
    else if constexpr (std::is_same<S, __hip_bfloat16>::value)
        *p = __float2bfloat16((float) v);
    else
        *p = (S) v;
}

template <typename S>
__global__ void __launch_bounds__(BLOCK) scatterState(S* __restrict__ dst, const double* __restrict__ in, Geometry g) {
    uint32_t node = blockIdx.x*BLOCK + threadIdx.x;
    if (node >= g.n)
        return;
    uint32_t i, j, k;
    nodeCoords(g, node, i, j, k);
    const NodeSlots at = nodeSlots(g, node, i, j, k);
    staticFor<NQ>([&](auto I) {
        constexpr int q = decltype(I)::value;
        storeNearest(dst + at.template pulled<q>(), in[(size_t) q*g.n + node]);
    });
}

/*
 * Density and velocity at the entries of the output layout (ReferenceLBMKernels.cpp:1349-1368, 1403-1436): entry l of
 * the lattice padded by `pad` nodes on every side, i fastest.  A padded entry holds the node it wraps onto across a
 * periodic boundary if the halo of that field is exchanged, NaN otherwise and beyond an open face
 * (LBMDecomposition.cpp:183-204 with one domain).
 */
template <typename S>
__global__ void __launch_bounds__(BLOCK) fieldsKernel(const S* __restrict__ src, const uint32_t* __restrict__ mask,
        Geometry g, FieldsLayout out, uint32_t count, double density, double velocityScale, double gx, double gy,
        double gz, double* __restrict__ rhoOut, double* __restrict__ uOut) {
    uint32_t l = blockIdx.x*BLOCK + threadIdx.x;
    if (l >= count)
        return;
    uint32_t node = l;
    bool wantRho = rhoOut != nullptr, wantU = uOut != nullptr;
    if (out.pad > 0) {
        uint32_t wx = g.nx + 2*out.pad, wy = g.ny + 2*out.pad;
        int li = (int) (l % wx) - out.pad, lj = (int) ((l/wx) % wy) - out.pad, lk = (int) (l/(wx*wy)) - out.pad;
        int c[3] = {li, lj, lk}, n[3] = {(int) g.nx, (int) g.ny, (int) g.nz}, open[3] = {g.openX, g.openY, g.openZ};
        bool inside = true, beyondFace = false;
        for (int a = 0; a < 3; a++) {
            if (c[a] < 0 || c[a] >= n[a]) {
                inside = false;
                beyondFace = beyondFace || open[a] != 0;
                c[a] = (c[a] + n[a]) % n[a];
            }
        }
        bool rhoValid = !beyondFace && (inside || out.densityHalo != 0);
        bool uValid = !beyondFace && (inside || out.velocityHalo != 0);
        const double nan = __builtin_nan("");
        if (wantRho && !rhoValid)
            rhoOut[l] = nan;
        if (wantU && !uValid)
            uOut[3*(size_t) l] = uOut[3*(size_t) l + 1] = uOut[3*(size_t) l + 2] = nan;
        wantRho = wantRho && rhoValid;
        wantU = wantU && uValid;
        if (!wantRho && !wantU)
            return;
        node = (uint32_t) c[0] + g.nx*((uint32_t) c[1] + g.ny*(uint32_t) c[2]);
    }
    if (mask && isSolid(mask, node)) {
        if (wantRho)
            rhoOut[l] = 0.0;
        if (wantU)
            uOut[3*(size_t) l] = uOut[3*(size_t) l + 1] = uOut[3*(size_t) l + 2] = 0.0;
        return;
    }
    uint32_t i, j, k;
    nodeCoords(g, node, i, j, k);
    const NodeSlots at = nodeSlots(g, node, i, j, k);
    double r = 1.0, jx = 0, jy = 0, jz = 0;
    staticFor<NQ>([&](auto I) {
        constexpr int q = decltype(I)::value;
        double v = (double) loadPop(src + at.template pulled<q>());
        r += v;
        jx = addTerm<CX[q]>(jx, v);
        jy = addTerm<CY[q]>(jy, v);
        jz = addTerm<CZ[q]>(jz, v);
    });
    if (wantRho)
        rhoOut[l] = r*density;
    if (wantU) {
        uOut[3*(size_t) l] = (jx*(1.0/r) + gx*0.5)*velocityScale;
        uOut[3*(size_t) l + 1] = (jy*(1.0/r) + gy*0.5)*velocityScale;
        uOut[3*(size_t) l + 2] = (jz*(1.0/r) + gz*0.5)*velocityScale;
    }
}

/* Equilibrium at density 1 and velocity u everywhere, -w_q (no fluid) at solid nodes, written in the pull layout. */
template <typename S>
__global__ void __launch_bounds__(BLOCK) initEquilibrium(S* __restrict__ dst, const uint32_t* __restrict__ mask, Geometry g,
        double ux, double uy, double uz) {
    uint32_t node = blockIdx.x*BLOCK + threadIdx.x;
    if (node >= g.n)
        return;
    bool solid = mask && isSolid(mask, node);
    uint32_t i, j, k;
    nodeCoords(g, node, i, j, k);
    const NodeSlots at = nodeSlots(g, node, i, j, k);
    double eq[NQ];
    equilibriumDeviationD(0.0, ux, uy, uz, eq);
    staticFor<NQ>([&](auto I) {
        constexpr int q = decltype(I)::value;
        storeNearest(dst + at.template pulled<q>(), solid ? -WQ[q] : eq[q]);
    });
}

template <typename S>
__global__ void __launch_bounds__(BLOCK) fillZero(S* __restrict__ dst, size_t count) {
    size_t idx = (size_t) blockIdx.x*BLOCK + threadIdx.x;
    for (; idx < count; idx += (size_t) gridDim.x*BLOCK)
        dst[idx] = S(0);
}

/* ------------------------------------------------------------------------------------------ */
/* Launchers.                                                                                 */
/* ------------------------------------------------------------------------------------------ */
static inline unsigned gridFor(uint64_t n) {
    return (unsigned) ((n + BLOCK - 1)/BLOCK);
}

#define LBM_DISPATCH_STORAGE(storage, S, ...) \
    switch (storage) { \
    case LBM_STORAGE_FP16: { using S = __half; __VA_ARGS__; break; } \
    case LBM_STORAGE_BF16: { using S = __hip_bfloat16; __VA_ARGS__; break; } \
    case LBM_STORAGE_FP64: { using S = double; __VA_ARGS__; break; } \
    default: { using S = float; __VA_ARGS__; break; } \
    }

template <typename S, bool SOL, bool REA, bool WIDE>
static void launchFusedT(bool fluct, const void* src, void* dst, const uint32_t* mask, unsigned long long* reactionFixed,
        uint32_t* touched, const double* ucm, const StepArgs& a, hipStream_t stream) {
    // FP16/BF16 storage: 256-thread work-groups (2% faster on config3 than 512; 512 is best for FP32).
    constexpr int T = sizeof(S) == 2 ? FUSED_THREADS_REDUCED : FUSED_THREADS;
    unsigned grid = (unsigned) ((a.g.n + T - 1)/T);
    if (fluct)
        fusedStep<S, SOL, REA, true, T, WIDE><<<grid, T, 0, stream>>>((const S*) src, (S*) dst, mask, reactionFixed,
                touched, ucm, a);
    else
        fusedStep<S, SOL, REA, false, T, WIDE><<<grid, T, 0, stream>>>((const S*) src, (S*) dst, mask, reactionFixed,
                touched, ucm, a);
}

template <typename S, bool SOL, bool REA, int SEG>
static void launchRowsT(bool fluct, const void* src, void* dst, const uint32_t* mask, unsigned long long* reactionFixed,
        uint32_t* touched, const double* ucm, const StepArgs& a, hipStream_t stream) {
    unsigned grid = (unsigned) ((a.g.n + ROW_THREADS - 1)/ROW_THREADS);
    if (fluct)
        fusedRows<S, SOL, REA, true, SEG><<<grid, ROW_THREADS, 0, stream>>>((const S*) src, (S*) dst, mask,
                reactionFixed, touched, ucm, a);
    else
        fusedRows<S, SOL, REA, false, SEG><<<grid, ROW_THREADS, 0, stream>>>((const S*) src, (S*) dst, mask,
                reactionFixed, touched, ucm, a);
}

template <typename S, bool SOL, bool REA>
static void launchPathT(int path, bool fluct, const void* src, void* dst, const uint32_t* mask,
        unsigned long long* reactionFixed, uint32_t* touched, const double* ucm, const StepArgs& a, hipStream_t stream) {
    switch (path) {
    case FUSED_WIDE:
        launchFusedT<S, SOL, REA, true>(fluct, src, dst, mask, reactionFixed, touched, ucm, a, stream);
        break;
    case FUSED_ROWS:
        if (a.g.nx == 128)
            launchRowsT<S, SOL, REA, 128>(fluct, src, dst, mask, reactionFixed, touched, ucm, a, stream);
        else
            launchRowsT<S, SOL, REA, ROW_THREADS>(fluct, src, dst, mask, reactionFixed, touched, ucm, a, stream);
        break;
    default:
        launchFusedT<S, SOL, REA, false>(fluct, src, dst, mask, reactionFixed, touched, ucm, a, stream);
    }
}

bool fusedRowsFit(uint32_t nx, bool wide, uint32_t rowShift) {
    return !wide && rowShift == 0 && (nx % ROW_THREADS == 0 || nx == 128);
}

void launchFused(int storage, int path, bool solids, bool reaction, bool fluct, const void* src, void* dst,
        const uint32_t* mask, unsigned long long* reactionFixed, uint32_t* touched, const double* ucm, const StepArgs& a,
        hipStream_t stream) {
    LBM_DISPATCH_STORAGE(storage, S, {
        if (solids) {
            if (reaction) launchPathT<S, true, true>(path, fluct, src, dst, mask, reactionFixed, touched, ucm, a, stream);
            else launchPathT<S, true, false>(path, fluct, src, dst, mask, reactionFixed, touched, ucm, a, stream);
        }
        else {
            if (reaction) launchPathT<S, false, true>(path, fluct, src, dst, mask, reactionFixed, touched, ucm, a, stream);
            else launchPathT<S, false, false>(path, fluct, src, dst, mask, reactionFixed, touched, ucm, a, stream);
        }
    })
}

void launchMomentReduce(int storage, bool solids, const void* src, const uint32_t* mask, const Geometry& g,
        double* partials, hipStream_t stream) {
    LBM_DISPATCH_STORAGE(storage, S, {
        if (solids)
            momentReduce<S, true><<<REDUCE_BLOCKS, BLOCK, 0, stream>>>((const S*) src, mask, g, partials);
        else
            momentReduce<S, false><<<REDUCE_BLOCKS, BLOCK, 0, stream>>>((const S*) src, mask, g, partials);
    })
}

void launchFinishReduce(const double* partials, double numFluid, double* totals, double* ucm, hipStream_t stream) {
    finishReduce<<<1, BLOCK, 0, stream>>>(partials, numFluid, totals, ucm);
}

void launchReduceWall(const double* partials, int count, double* out, hipStream_t stream) {
    reduceWall<<<1, BLOCK, 0, stream>>>(partials, count, out);
}

void launchBounceBack(int storage, void* dst, const int* wallNodes, const int* wallLinks, int count, const Geometry& g,
        double* partials, hipStream_t stream) {
    if (count == 0)
        return;
    LBM_DISPATCH_STORAGE(storage, S, {
        bounceBack<S><<<gridFor(count), BLOCK, 0, stream>>>((S*) dst, wallNodes, wallLinks, count, g, partials);
    })
}

void launchBoundaries(int storage, const void* src, void* dst, const BoundaryNode* nodes, const double* ucm,
        const BoundaryArgs& a, double* partials, hipStream_t stream) {
    if (a.count == 0)
        return;
    LBM_DISPATCH_STORAGE(storage, S, {
        boundaryKernel<S><<<gridFor(a.count), BLOCK, 0, stream>>>((const S*) src, (S*) dst, nodes, ucm, a, partials);
    })
}

void launchGather(int storage, const void* src, double* out, const Geometry& g, hipStream_t stream) {
    LBM_DISPATCH_STORAGE(storage, S, { gatherState<S><<<gridFor(g.n), BLOCK, 0, stream>>>((const S*) src, out, g); })
}

void launchScatter(int storage, void* dst, const double* in, const Geometry& g, hipStream_t stream) {
    LBM_DISPATCH_STORAGE(storage, S, { scatterState<S><<<gridFor(g.n), BLOCK, 0, stream>>>((S*) dst, in, g); })
}

uint64_t fieldsCount(const Geometry& g, const FieldsLayout& out) {
    return (uint64_t) (g.nx + 2*out.pad)*(g.ny + 2*out.pad)*(g.nz + 2*out.pad);
}

void launchFields(int storage, const void* src, const uint32_t* mask, const Geometry& g, const FieldsLayout& out,
        double density, double velocityScale, double gx, double gy, double gz, double* rhoOut, double* uOut,
        hipStream_t stream) {
    uint64_t count = fieldsCount(g, out);
    if (count > 0xFFFFFFFFull)
        throw std::runtime_error("getFluidFields: the lattice with its halo exceeds 2^32 nodes");
    LBM_DISPATCH_STORAGE(storage, S, {
        fieldsKernel<S><<<gridFor(count), BLOCK, 0, stream>>>((const S*) src, mask, g, out, (uint32_t) count, density,
                velocityScale, gx, gy, gz, rhoOut, uOut);
    })
}

void launchInitEquilibrium(int storage, void* dst, const uint32_t* mask, const Geometry& g, double ux, double uy,
        double uz, hipStream_t stream) {
    LBM_DISPATCH_STORAGE(storage, S, {
        initEquilibrium<S><<<gridFor(g.n), BLOCK, 0, stream>>>((S*) dst, mask, g, ux, uy, uz);
    })
}

void launchFill(int storage, void* dst, size_t count, hipStream_t stream) {
    unsigned grid = (unsigned) std::min<uint64_t>(gridFor(count), 65536);
    if (grid == 0)
        return;
    LBM_DISPATCH_STORAGE(storage, S, { fillZero<S><<<grid, BLOCK, 0, stream>>>((S*) dst, count); })
}

}  // namespace lbm
