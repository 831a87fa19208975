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
/* The stencil width along one axis of the stencil that a group of LANES lanes couples with. */
template <int LANES>
constexpr int groupWidth() {
    return LANES == 64 ? 4 : LANES == 32 ? 3 : LANES == 8 ? 2 : 1;
}

/* The first node of a stencil along one axis (the rule of axisSlot), wrapped into [0, n). */
template <int STENCIL>
__device__ __forceinline__ int anchorIndex(double s, int n) {
    int first = STENCIL == LBM_STENCIL_NEAREST ? nearestIndex(s, n) :
                STENCIL == LBM_STENCIL_THREE_POINT ? (int) floor(s + 0.5) - 1 :
                STENCIL == LBM_STENCIL_KEYS ? (int) floor(s) - 1 : (int) floor(s);
    return first < 0 ? first + n : (first >= n ? first - n : first);
}

/* The bin of a particle at x: the tile of its stencil's first node; *offset receives that node's place in the tile
   (x | y << 5 | z << 7). */
template <int STENCIL>
__device__ __forceinline__ uint32_t anchorTile(const double* x, const ParticleArgs& a, const SpreadBins& bins,
        uint16_t* offset, const AxisStencil* pre) {
    const int size[3] = {(int) a.g.nx, (int) a.g.ny, (int) a.g.nz};
    const int tileSize[3] = {SPREAD_TX, SPREAD_TY, SPREAD_TZ};
    uint32_t t[3], o[3];
    for (int k = 0; k < 3; k++) {
        // The first slot's node: on the lattices of whole tiles with periodic faces that the bins serve, no stencil is
        // merged and it is the first node wrapped.
        int first = pre != nullptr ? pre->index[k][0] : anchorIndex<STENCIL>(wrapCoordinate(x[k], a.dx, size[k]), size[k]);
        t[k] = (uint32_t) (first/tileSize[k]);
        o[k] = (uint32_t) (first - (int) t[k]*tileSize[k]);
    }
    *offset = (uint16_t) (o[0] | (o[1] << 5) | (o[2] << 7));
    return t[0] + bins.ntx*(t[1] + bins.nty*t[2]);
}

/*
 * finishGroup for the one-writer spread: the drag force as in finishGroup; in a lattice step the reactions on solid
 * nodes go to the wall as there, and the particle goes into the bin of its anchor tile with its force, for
 * spreadTiles.  A particle whose bin is full spreads its reactions here with the fixed-point atomics instead (the sums
 * are integers, so both ways give the same bits).
 */
template <int LANES>
__device__ __forceinline__ void finishGroupBinned(const StencilSlot& sl, const double u[3], const double* vel,
        const double* x, uint32_t index, double m, bool active, int lane, const ParticleArgs& a, double* forceOut,
        unsigned long long* __restrict__ reactionFixed, uint32_t* __restrict__ touched, double wallAcc[3],
        const double* noise, double* systemForce, const SpreadBins& bins, const AxisStencil* pre) {
    double f[3] = {0, 0, 0};
    if (active) {
        dragForce(vel, u, index, m, a, f, noise);
        if (lane == 0)
            for (int k = 0; k < 3; k++) {
                forceOut[k] = f[k]*a.forceScale;
                if (systemForce != nullptr)
                    atomicAdd(systemForce + k, f[k]*a.forceScale);
            }
    }
    if (!a.isStep)
        return;
    int full = 0;
    if (active && lane == 0) {
        uint16_t offset;
        uint32_t tile = anchorTile<groupStencil<LANES>()>(x, a, bins, &offset, pre);
        int slot = atomicAdd(bins.count + tile, 1);
        if (slot < bins.cap) {
            const size_t entry = (size_t) tile*bins.cap + slot;
            bins.beads[entry] = (int) index;
            bins.anchor[entry] = offset;
            for (int k = 0; k < 3; k++)
                bins.force[3*entry + k] = f[k];
        }
        else {
            full = 1;
            atomicOr(bins.spilled, 1);
        }
    }
    full = __shfl(full, 0, LANES);
    if (active && !full && sl.inStencil)
        bins.slotWeight[(size_t) index*LANES + lane] = sl.weight;     // spreadTiles forms the slot's node itself
    bool flagged = false;
    if (sl.inStencil) {
        if (full)
            flagged = spreadReaction(sl, f, a, reactionFixed, wallAcc);
        else if (sl.solid)
            for (int k = 0; k < 3; k++)
                wallAcc[k] = -sl.weight*f[k];
    }
    if (__any(full))
        markTouched<groupWidth<LANES>()>(touched, sl.node, flagged, lane);
}

/*
 * The stencil of every coupled particle along each axis, four lanes per particle (lane k < 3 takes axis k): the wrapped
 * coordinate and axisSlot of each slot, which every lane of a coupling group would otherwise form again (the FP64
 * divisions of the wrap and the kernel polynomials run once per particle and axis instead of once per group); with a
 * random force, lane k also draws the particle's normal k, off the coupling's path.  t: the thread's place among
 * 4*numParticles.
 */
template <int STENCIL>
__device__ __forceinline__ void stencilSetupLane(int t, const double* __restrict__ positions,
        const int* __restrict__ particles, const ParticleArgs& a, AxisStencil* __restrict__ out) {
    constexpr int w = STENCIL == LBM_STENCIL_KEYS ? 4 : STENCIL == LBM_STENCIL_THREE_POINT ? 3 :
                      STENCIL == LBM_STENCIL_TRILINEAR ? 2 : 1;
    const int i = t >> 2, k = t & 3;
    if (i >= a.numParticles)
        return;
    if (a.randomForce)
        out[i].normal[k] = particleNormal((uint32_t) i, k, a);
    if (k == 3)
        return;
    // The size and openness of axis k, selected between values held in scalar registers: k differs between the lanes,
    // and an array of the argument's fields indexed by it was compiled into one scalar load at the first lane's address
    // (every lane then took that lane's axis).
    const int nx = __builtin_amdgcn_readfirstlane((int) a.g.nx), ny = __builtin_amdgcn_readfirstlane((int) a.g.ny),
              nz = __builtin_amdgcn_readfirstlane((int) a.g.nz);
    const int ox = __builtin_amdgcn_readfirstlane(a.g.openX), oy = __builtin_amdgcn_readfirstlane(a.g.openY),
              oz = __builtin_amdgcn_readfirstlane(a.g.openZ);
    const int n = k == 0 ? nx : (k == 1 ? ny : nz);
    const bool open = (k == 0 ? ox : (k == 1 ? oy : oz)) != 0;
    const double s = wrapCoordinate(positions[3*(size_t) particles[i] + k], a.dx, n);
    AxisStencil& st = out[i];
#pragma unroll
    for (int sub = 0; sub < 4; sub++) {
        int index = 0;
        double weight = 0.0;
        if (sub < w)
            axisSlot<STENCIL>(s, n, open, sub, index, weight);
        st.index[k][sub] = index;
        st.weight[k][sub] = weight;
    }
}

template <int STENCIL>
__global__ void __launch_bounds__(BLOCK) stencilSetup(const double* __restrict__ positions,
        const int* __restrict__ particles, ParticleArgs a, AxisStencil* __restrict__ out) {
    stencilSetupLane<STENCIL>(blockIdx.x*BLOCK + threadIdx.x, positions, particles, a, out);
}

static unsigned stencilSetupBlocks(int numParticles) {
    return (unsigned) ((4*(size_t) numParticles + BLOCK - 1)/BLOCK);
}

void launchStencilSetup(const double* positions, const int* particles, const ParticleArgs& a, AxisStencil* stencils,
        hipStream_t stream) {
    if (a.numParticles == 0)
        return;
    const unsigned grid = stencilSetupBlocks(a.numParticles);
    switch (a.stencil) {
    case LBM_STENCIL_KEYS:
        stencilSetup<LBM_STENCIL_KEYS><<<grid, BLOCK, 0, stream>>>(positions, particles, a, stencils);
        break;
    case LBM_STENCIL_THREE_POINT:
        stencilSetup<LBM_STENCIL_THREE_POINT><<<grid, BLOCK, 0, stream>>>(positions, particles, a, stencils);
        break;
    case LBM_STENCIL_TRILINEAR:
        stencilSetup<LBM_STENCIL_TRILINEAR><<<grid, BLOCK, 0, stream>>>(positions, particles, a, stencils);
        break;
    default:
        break;
    }
}

template <typename S, int LANES, bool BINNED>
__global__ void __launch_bounds__(BLOCK) coupleParticles(const S* __restrict__ src, const uint32_t* __restrict__ mask,
        const double* __restrict__ positions, const double* __restrict__ velocities, const int* __restrict__ particles,
        const int* __restrict__ order, const double* __restrict__ mass, const double* __restrict__ ucm, ParticleArgs a,
        double* __restrict__ particleForce, unsigned long long* __restrict__ reactionFixed, uint32_t* __restrict__ touched,
        double* __restrict__ partials, double* __restrict__ systemForces, SpreadBins bins,
        const double4* __restrict__ field, const AxisStencil* __restrict__ stencils) {
    constexpr int GROUPS = BLOCK/LANES;
    int lane = threadIdx.x % LANES;
    // Sparse beads with FP16/BF16 populations: each XCD takes a contiguous run of the work-groups, so that beads near
    // in the processing order (sorted by node) share an L2 for the 128 B lines, 64 nodes wide, that they read
    // (coupled step of 256^3 with 20,000 Keys beads 0.401 -> 0.397 ms; no gain with FP32's 32-node lines).
    const uint32_t group = isReduced<S>() && !BINNED && a.xcdOrder ? xcdContiguousTile(blockIdx.x, gridDim.x) : blockIdx.x;
    int rank = (int) group*GROUPS + threadIdx.x/LANES;    // position in the processing order
    bool active = rank < a.numParticles;
    int i = active && order != nullptr ? order[rank] : rank;
    int part = active ? particles[i] : 0;
    double m = active ? mass[i] : 0.0;
    double u[3];
    const AxisStencil* pre = stencils != nullptr && active ? stencils + i : nullptr;
    StencilSlot sl = interpolateGroup<S, LANES>(src, mask, positions + 3*(size_t) part, active, lane, ucm, a, u, field,
            pre);
    // The normals of the work-group's particles: formed with the stencils where those were formed beforehand (no
    // barrier here), else drawn by the work-group's first 4*GROUPS lanes, one normal each, instead of by every group:
    // a wavefront pays for the logarithm and cosine or sine once however many of its lanes need them.  Drawn after the
    // interpolation, so that the first wavefront's loads are in flight meanwhile.
    __shared__ double noise[LANES > 1 ? GROUPS : 1][4];
    const double* myNoise = nullptr;
    if (LANES > 1 && a.randomForce && stencils == nullptr) {
        if (threadIdx.x < 4*GROUPS) {
            int k = (int) group*GROUPS + threadIdx.x/4;
            noise[threadIdx.x/4][threadIdx.x & 3] =
                    k < a.numParticles ? particleNormal(order != nullptr ? order[k] : k, threadIdx.x & 3, a) : 0.0;
        }
        __syncthreads();
        myNoise = noise[threadIdx.x/LANES];
    }
    else if (a.randomForce && pre != nullptr)
        myNoise = pre->normal;
    double wallAcc[3] = {0, 0, 0};
    if constexpr (BINNED)
        finishGroupBinned<LANES>(sl, u, velocities + 3*(size_t) part, positions + 3*(size_t) part, (uint32_t) i, m, active,
                lane, a, particleForce + 3*(size_t) (active ? i : 0), reactionFixed, touched, wallAcc, myNoise,
                systemForces != nullptr ? systemForces + 3*(size_t) part : nullptr, bins, pre);
    else
        finishGroup<LANES>(sl, u, velocities + 3*(size_t) part, (uint32_t) i, m, active, lane, a,
                particleForce + 3*(size_t) (active ? i : 0), reactionFixed, touched, wallAcc, myNoise,
                systemForces != nullptr ? systemForces + 3*(size_t) part : nullptr);
    if (partials != nullptr)
        storeWallPartial<BLOCK>(wallAcc, partials);
}

/*
 * The one-writer spread: one work-group per tile of 32 x 4 x 4 nodes, one thread per node.  The beads that can reach
 * the tile are those binned in it and in its neighbours one tile back along each axis (a stencil is at most 4 nodes
 * wide, no wider than a tile).  The work-group takes their stencil rows along x one per thread, forms the row's nodes
 * from the bead's first stencil node (its tile and anchor offset) and, where the row lies in the tile, loads its slots'
 * weights and adds the fixed-point reactions -w F of the slots inside the tile to the tile's accumulators in LDS.
 * Then each fluid node's sums are written by its own thread (a solid node's part went to the wall in the coupling),
 * and each tile row's touched-mask word by the row's first lane: no global atomics.  Weights and fixed-point values are those of the atomic spread
 * (spreadReaction), and integer sums do not depend on their order, so both give the same bits.  Lattices of whole tiles
 * with periodic faces only (spreadTilesFit): no stencil there is merged, so slot (sx, sy, sz) lies at the first node +
 * (sx, sy, sz), wrapped.  The reference builds its per-node sums from a sort of the stencil slots by node
 * (lbmCoupling.cc:251-272, 1245-1266).
 */
template <int LANES>
__global__ void __launch_bounds__(SPREAD_THREADS) spreadTiles(ParticleArgs a, SpreadBins bins,
        const uint32_t* __restrict__ mask, unsigned long long* __restrict__ reactionFixed, uint32_t* __restrict__ touched) {
    constexpr int W = groupWidth<LANES>();
    // The bin entries are taken SPREAD_CHUNK at a time, so that the candidates' LDS stays small and four work-groups
    // share a CU; a tile's eight bins hold about 8x the mean beads per tile, one chunk in the usual case.
    constexpr int SPREAD_CHUNK = 256;
    __shared__ unsigned long long acc[3*SPREAD_THREADS];
    __shared__ int candidates[SPREAD_CHUNK];
    __shared__ short candidateLocal[SPREAD_CHUNK][3];
    __shared__ double candidateForce[SPREAD_CHUNK][3];
    __shared__ int counts[8], firsts[8];
    __shared__ int numCandidates;
    const uint32_t tile = blockIdx.x;
    const uint32_t tx = tile % bins.ntx, ty = (tile/bins.ntx) % bins.nty, tz = tile/(bins.ntx*bins.nty);
    const int t = (int) threadIdx.x;
    if (t < 8) {
        uint32_t fx = (t & 1) ? (tx == 0 ? bins.ntx - 1 : tx - 1) : tx;
        uint32_t fy = (t & 2) ? (ty == 0 ? bins.nty - 1 : ty - 1) : ty;
        uint32_t fz = (t & 4) ? (tz == 0 ? bins.ntz - 1 : tz - 1) : tz;
        uint32_t from = fx + bins.ntx*(fy + bins.nty*fz);
        counts[t] = min(bins.count[from], bins.cap);
        firsts[t] = (int) from;
    }
    if (t == 0) {
        bins.nextCount[tile] = 0;
        if (tile == 0)
            *bins.nextSpilled = 0;
    }
    __syncthreads();
    int total = 0;
    for (int d = 0; d < 8; d++)
        total += counts[d];
    if (total == 0)
        return;
    for (int k = 0; k < 3; k++)
        acc[k*SPREAD_THREADS + t] = 0ull;
    const int lane = t & 63;
    for (int chunk = 0; chunk < total; chunk += SPREAD_CHUNK) {
        // The candidates of this chunk of bin entries that reach into this tile, gathered by every thread: a bead
        // binned one tile back along an axis reaches it when its first stencil node lies within W - 1 nodes of that
        // tile's end.  Each keeps its first node relative to this tile's origin (negative for the tiles behind).
        if (t == 0)
            numCandidates = 0;
        __syncthreads();
        for (int c = chunk + t; c < min(total, chunk + SPREAD_CHUNK); c += SPREAD_THREADS) {
            int r = c, d = 0;
            while (r >= counts[d]) {
                r -= counts[d];
                d++;
            }
            const size_t entry = (size_t) firsts[d]*bins.cap + r;
            const uint32_t off = bins.anchor[entry];
            const int ox = (int) (off & 31u), oy = (int) ((off >> 5) & 3u), oz = (int) ((off >> 7) & 3u);
            const bool reaches = (!(d & 1) || ox >= SPREAD_TX - W + 1) && (!(d & 2) || oy >= SPREAD_TY - W + 1) &&
                                 (!(d & 4) || oz >= SPREAD_TZ - W + 1);
            if (reaches) {
                const int at = atomicAdd(&numCandidates, 1);
                candidates[at] = bins.beads[entry];
                candidateLocal[at][0] = (short) ((d & 1) ? ox - SPREAD_TX : ox);
                candidateLocal[at][1] = (short) ((d & 2) ? oy - SPREAD_TY : oy);
                candidateLocal[at][2] = (short) ((d & 4) ? oz - SPREAD_TZ : oz);
                for (int k = 0; k < 3; k++)
                    candidateForce[at][k] = bins.force[3*entry + k];
            }
        }
        __syncthreads();
        // One item per (candidate, stencil row along x); each thread loads the weights of UNROLL rows before it adds
        // any, so that their loads overlap.
        constexpr int ROWS = W*W;
        const int items = numCandidates*ROWS;
        constexpr int UNROLL = 2;
        for (int base = t; base < items; base += SPREAD_THREADS*UNROLL) {
            int rowLocal[UNROLL], x0[UNROLL], cand[UNROLL];
            double weight[UNROLL][W];
#pragma unroll
            for (int u = 0; u < UNROLL; u++) {
                const int item = base + u*SPREAD_THREADS;
                rowLocal[u] = -1;
                if (item < items) {
                    const int c = item/ROWS, row = item % ROWS;
                    const int ly = candidateLocal[c][1] + row % W, lz = candidateLocal[c][2] + row/W;
                    if (ly >= 0 && ly < SPREAD_TY && lz >= 0 && lz < SPREAD_TZ) {
                        const double* w = bins.slotWeight + (size_t) candidates[c]*LANES + W*row;
#pragma unroll
                        for (int sx = 0; sx < W; sx++)
                            weight[u][sx] = w[sx];
                        rowLocal[u] = SPREAD_TX*(ly + SPREAD_TY*lz);
                        x0[u] = candidateLocal[c][0];
                        cand[u] = c;
                    }
                }
            }
#pragma unroll
            for (int u = 0; u < UNROLL; u++) {
                if (rowLocal[u] < 0)
                    continue;
#pragma unroll
                for (int sx = 0; sx < W; sx++) {
                    const int lx = x0[u] + sx;
                    if (lx < 0 || lx >= SPREAD_TX)
                        continue;
                    for (int k = 0; k < 3; k++) {
                        double r = -weight[u][sx]*candidateForce[cand[u]][k];
                        if (!reactionInRange(r, a.overflow))
                            continue;
                        long long fixed = llrint(r*FIXED_SCALE);
                        if (fixed != 0)
                            atomicAdd(&acc[k*SPREAD_THREADS + rowLocal[u] + lx], (unsigned long long) fixed);
                    }
                }
            }
        }
        __syncthreads();             // before the next chunk's candidates replace these
    }
    const int lo[3] = {(int) tx*SPREAD_TX, (int) ty*SPREAD_TY, (int) tz*SPREAD_TZ};
    const int lx = t % SPREAD_TX, ly = (t/SPREAD_TX) % SPREAD_TY, lz = t/(SPREAD_TX*SPREAD_TY);
    const uint32_t node = (uint32_t) (lo[0] + lx) + a.g.nx*((uint32_t) (lo[1] + ly) + a.g.ny*(uint32_t) (lo[2] + lz));
    const unsigned long long v[3] = {acc[t], acc[SPREAD_THREADS + t], acc[2*SPREAD_THREADS + t]};
    const bool any = (v[0] | v[1] | v[2]) != 0ull && !solidAt(mask, node);
    const uint64_t ballot = __ballot(any);
    const uint32_t rowBits = (uint32_t) (ballot >> (lane & 32));
    // Bits of beads that spread themselves (full bins); without any, the fused pass left the word clear.
    const bool spilled = *bins.spilled != 0;
    uint32_t old = 0;
    if (spilled && lx == 0 && rowBits != 0u)
        old = touched[node >> 5];
    old = __shfl(old, lane & 32, 64);
    if (any) {
        const bool add = (old >> lx) & 1u;
        for (int k = 0; k < 3; k++) {
            size_t at = reactionIndex(node, k, a.g.n, a.reactionPlanes);
            reactionFixed[at] = add ? reactionFixed[at] + v[k] : v[k];
        }
    }
    if (lx == 0 && rowBits != 0u)
        touched[node >> 5] = old | rowBits;
}

/* The moment field, after SETUP work-groups that form the particles' stencils (stencilSetupLane) when STENCIL is a
   stencil wider than one node: the two are independent, and one launch overlaps the setup with the field's stream. */
template <typename S, int STENCIL>
__global__ void __launch_bounds__(BLOCK) momentField(const S* __restrict__ src, const uint32_t* __restrict__ mask,
        Geometry g, const double* __restrict__ ucm, int removal, double4* __restrict__ field,
        const double* __restrict__ positions, const int* __restrict__ particles, ParticleArgs a,
        AxisStencil* __restrict__ stencils, unsigned setupBlocks) {
    if constexpr (STENCIL != LBM_STENCIL_NEAREST) {
        if (blockIdx.x < setupBlocks) {
            stencilSetupLane<STENCIL>(blockIdx.x*BLOCK + threadIdx.x, positions, particles, a, stencils);
            return;
        }
    }
    const uint32_t block = blockIdx.x - setupBlocks, blocks = gridDim.x - setupBlocks;
    for (uint32_t node = block*BLOCK + threadIdx.x; node < g.n; node += blocks*BLOCK) {
        if (solidAt(mask, node))
            continue;          // solid slots do not read the field
        momentFieldNode(src, g, node, ucm, removal, field);
    }
}

template <typename S>
static void launchMomentFieldT(const S* src, const uint32_t* mask, const ParticleArgs& a, const double* ucm,
        double4* field, const double* positions, const int* particles, AxisStencil* stencils, unsigned grid,
        hipStream_t stream) {
    const unsigned setup = stencils != nullptr ? stencilSetupBlocks(a.numParticles) : 0u;
    switch (stencils != nullptr ? a.stencil : LBM_STENCIL_NEAREST) {
    case LBM_STENCIL_KEYS:
        momentField<S, LBM_STENCIL_KEYS><<<grid + setup, BLOCK, 0, stream>>>(src, mask, a.g, ucm, a.removal, field,
                positions, particles, a, stencils, setup);
        break;
    case LBM_STENCIL_THREE_POINT:
        momentField<S, LBM_STENCIL_THREE_POINT><<<grid + setup, BLOCK, 0, stream>>>(src, mask, a.g, ucm, a.removal,
                field, positions, particles, a, stencils, setup);
        break;
    case LBM_STENCIL_TRILINEAR:
        momentField<S, LBM_STENCIL_TRILINEAR><<<grid + setup, BLOCK, 0, stream>>>(src, mask, a.g, ucm, a.removal,
                field, positions, particles, a, stencils, setup);
        break;
    default:
        momentField<S, LBM_STENCIL_NEAREST><<<grid, BLOCK, 0, stream>>>(src, mask, a.g, ucm, a.removal, field,
                positions, particles, a, nullptr, 0u);
        break;
    }
}

void launchMomentField(int storage, const void* src, const uint32_t* mask, const ParticleArgs& a, const double* ucm,
        double4* field, const double* positions, const int* particles, AxisStencil* stencils, hipStream_t stream) {
    unsigned grid = (unsigned) std::min<uint64_t>((a.g.n + BLOCK - 1)/BLOCK, 65535u*4u);
    switch (storage) {
    case LBM_STORAGE_FP16:
        launchMomentFieldT((const __half*) src, mask, a, ucm, field, positions, particles, stencils, grid, stream);
        break;
    case LBM_STORAGE_BF16:
        launchMomentFieldT((const __hip_bfloat16*) src, mask, a, ucm, field, positions, particles, stencils, grid,
                stream);
        break;
    case LBM_STORAGE_FP64:
        launchMomentFieldT((const double*) src, mask, a, ucm, field, positions, particles, stencils, grid, stream);
        break;
    default:
        launchMomentFieldT((const float*) src, mask, a, ucm, field, positions, particles, stencils, grid, stream);
        break;
    }
}

bool spreadTilesFit(int nx, int ny, int nz, bool anyOpen) {
    return !anyOpen && nx % SPREAD_TX == 0 && ny % SPREAD_TY == 0 && nz % SPREAD_TZ == 0 && nx >= 2*SPREAD_TX &&
           ny >= 2*SPREAD_TY && nz >= 2*SPREAD_TZ;
}

void launchSpreadTiles(const double* positions, const int* particles, const uint32_t* mask, const ParticleArgs& a,
        const SpreadBins& bins, unsigned long long* reactionFixed, uint32_t* touched, hipStream_t stream) {
    (void) positions;
    (void) particles;
    unsigned grid = bins.ntx*bins.nty*bins.ntz;
    switch (coupleLanes(a.stencil)) {
    case 64:
        spreadTiles<64><<<grid, SPREAD_THREADS, 0, stream>>>(a, bins, mask, reactionFixed, touched);
        break;
    case 32:
        spreadTiles<32><<<grid, SPREAD_THREADS, 0, stream>>>(a, bins, mask, reactionFixed, touched);
        break;
    case 8:
        spreadTiles<8><<<grid, SPREAD_THREADS, 0, stream>>>(a, bins, mask, reactionFixed, touched);
        break;
    default:
        spreadTiles<1><<<grid, SPREAD_THREADS, 0, stream>>>(a, bins, mask, reactionFixed, touched);
        break;
    }
}

/* Batched replicas: the coupling of every particle of every replica in one launch.  Each work-group of THREADS
   threads couples up to THREADS/LANES particles of one replica (particleTiles[blockIdx.x] names it), so the replica's
   descriptor is uniform across the work-group and read with scalar loads. */
template <int LANES, int THREADS>
__global__ void __launch_bounds__(THREADS) coupleBatch(const ReplicaDesc* __restrict__ table,
        const uint16_t* __restrict__ particleTiles, const double* __restrict__ positions,
        const double* __restrict__ velocities, const double* __restrict__ mass, double* __restrict__ forces,
        const int64_t* __restrict__ stepCounter, int xcdOrder) {
    constexpr int GROUPS = THREADS/LANES;
    int lane = threadIdx.x % LANES;
    const uint32_t tile = xcdOrder ? xcdContiguousTile(blockIdx.x, gridDim.x) : blockIdx.x;
    const ReplicaDesc* d = table + particleTiles[tile];
    int local = (int) (tile - d->particleTileBegin)*GROUPS + threadIdx.x/LANES;
    bool active = local < d->part.numParticles;
    int pp = d->particleBegin + (active ? local : 0);
    ParticleArgs a = d->part;
    a.step = *stepCounter;
    a.isStep = 1;
    const float* src = (a.step & 1) ? d->popB : d->popA;
    double u[3];
    StencilSlot sl = interpolateGroup<float, LANES>(src, nullptr, positions + 3*(size_t) pp, active, lane, nullptr, a, u);
    // The work-group's normals, drawn by its first 4*GROUPS lanes (as in coupleParticles).
    __shared__ double noise[LANES > 1 ? GROUPS : 1][4];
    const double* myNoise = nullptr;
    if (LANES > 1 && a.randomForce) {
        if (threadIdx.x < 4*GROUPS) {
            int k = (int) (tile - d->particleTileBegin)*GROUPS + threadIdx.x/4;
            noise[threadIdx.x/4][threadIdx.x & 3] = k < a.numParticles ? particleNormal((uint32_t) k, threadIdx.x & 3, a) : 0.0;
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

template <typename S, bool BINNED>
static void launchCoupleB(const void* src, const uint32_t* mask, const double* positions, const double* velocities,
        const int* particles, const int* order, const double* mass, const double* ucm, const ParticleArgs& a,
        double* particleForce, unsigned long long* reactionFixed, uint32_t* touched, double* partials,
        double* systemForces, const SpreadBins& bins, const double4* field, AxisStencil* stencils, hipStream_t stream) {
    unsigned grid = (unsigned) coupleBlocks(a.stencil, a.numParticles);
    const S* s = (const S*) src;
    if (a.stencil == LBM_STENCIL_NEAREST)
        stencils = nullptr;              // one lane per particle: nothing to share
    switch (coupleLanes(a.stencil)) {
    case 8:
        coupleParticles<S, 8, BINNED><<<grid, BLOCK, 0, stream>>>(s, mask, positions, velocities, particles, order, mass,
                ucm, a, particleForce, reactionFixed, touched, partials, systemForces, bins, field, stencils);
        break;
    case 32:
        coupleParticles<S, 32, BINNED><<<grid, BLOCK, 0, stream>>>(s, mask, positions, velocities, particles, order, mass,
                ucm, a, particleForce, reactionFixed, touched, partials, systemForces, bins, field, stencils);
        break;
    case 64:
        coupleParticles<S, 64, BINNED><<<grid, BLOCK, 0, stream>>>(s, mask, positions, velocities, particles, order, mass,
                ucm, a, particleForce, reactionFixed, touched, partials, systemForces, bins, field, stencils);
        break;
    default:
        coupleParticles<S, 1, BINNED><<<grid, BLOCK, 0, stream>>>(s, mask, positions, velocities, particles, order, mass,
                ucm, a, particleForce, reactionFixed, touched, partials, systemForces, bins, field, stencils);
        break;
    }
}

template <typename S>
static void launchCoupleT(const void* src, const uint32_t* mask, const double* positions, const double* velocities,
        const int* particles, const int* order, const double* mass, const double* ucm, const ParticleArgs& a,
        double* particleForce, unsigned long long* reactionFixed, uint32_t* touched, double* partials,
        double* systemForces, const SpreadBins* bins, const double4* field, AxisStencil* stencils, hipStream_t stream) {
    if (bins != nullptr && a.isStep)
        launchCoupleB<S, true>(src, mask, positions, velocities, particles, order, mass, ucm, a, particleForce,
                reactionFixed, touched, partials, systemForces, *bins, field, stencils, stream);
    else
        launchCoupleB<S, false>(src, mask, positions, velocities, particles, order, mass, ucm, a, particleForce,
                reactionFixed, touched, partials, systemForces, SpreadBins(), field, stencils, stream);
}

void launchCouple(int storage, const void* src, const uint32_t* mask, const double* positions, const double* velocities,
        const int* particles, const int* order, const double* mass, const double* ucm, const ParticleArgs& a,
        double* particleForce, unsigned long long* reactionFixed, uint32_t* touched, double* partials,
        double* systemForces, const SpreadBins* bins, const double4* field, AxisStencil* stencils, hipStream_t stream) {
    if (a.numParticles == 0)
        return;
    if (field != nullptr) {
        // The moment field replaces the pulls: the storage type no longer matters.
        launchCoupleT<float>(nullptr, mask, positions, velocities, particles, order, mass, ucm, a, particleForce,
                reactionFixed, touched, partials, systemForces, bins, field, stencils, stream);
        return;
    }
    switch (storage) {
    case LBM_STORAGE_FP16:
        launchCoupleT<__half>(src, mask, positions, velocities, particles, order, mass, ucm, a, particleForce, reactionFixed,
                touched, partials, systemForces, bins, nullptr, stencils, stream);
        break;
    case LBM_STORAGE_BF16:
        launchCoupleT<__hip_bfloat16>(src, mask, positions, velocities, particles, order, mass, ucm, a, particleForce,
                reactionFixed, touched, partials, systemForces, bins, nullptr, stencils, stream);
        break;
    case LBM_STORAGE_FP64:
        launchCoupleT<double>(src, mask, positions, velocities, particles, order, mass, ucm, a, particleForce, reactionFixed,
                touched, partials, systemForces, bins, nullptr, stencils, stream);
        break;
    default:
        launchCoupleT<float>(src, mask, positions, velocities, particles, order, mass, ucm, a, particleForce, reactionFixed,
                touched, partials, systemForces, bins, nullptr, stencils, stream);
        break;
    }
}

template <int LANES>
static void launchCoupleBatchT(int threads, unsigned grid, const ReplicaDesc* table, const uint16_t* particleTiles,
        const double* positions, const double* velocities, const double* mass, double* forces,
        const int64_t* stepCounter, bool xcdOrder, hipStream_t stream) {
    switch (threads) {
    case 64:
        coupleBatch<LANES, 64><<<grid, 64, 0, stream>>>(table, particleTiles, positions, velocities, mass, forces,
                stepCounter, xcdOrder);
        break;
    case 128:
        coupleBatch<LANES, 128><<<grid, 128, 0, stream>>>(table, particleTiles, positions, velocities, mass, forces,
                stepCounter, xcdOrder);
        break;
    default:
        coupleBatch<LANES, BLOCK><<<grid, BLOCK, 0, stream>>>(table, particleTiles, positions, velocities, mass, forces,
                stepCounter, xcdOrder);
        break;
    }
}

void launchCoupleBatch(int stencil, int threads, const ReplicaDesc* table, const uint16_t* particleTiles,
        int numParticleTiles, const double* positions, const double* velocities, const double* mass, double* forces,
        const int64_t* stepCounter, bool xcdOrder, hipStream_t stream) {
    if (numParticleTiles == 0)
        return;
    unsigned grid = (unsigned) numParticleTiles;
    switch (coupleLanes(stencil)) {
    case 8:
        launchCoupleBatchT<8>(threads, grid, table, particleTiles, positions, velocities, mass, forces, stepCounter,
                xcdOrder, stream);
        break;
    case 32:
        launchCoupleBatchT<32>(threads, grid, table, particleTiles, positions, velocities, mass, forces, stepCounter,
                xcdOrder, stream);
        break;
    case 64:
        launchCoupleBatchT<64>(threads, grid, table, particleTiles, positions, velocities, mass, forces, stepCounter,
                xcdOrder, stream);
        break;
    default:
        launchCoupleBatchT<1>(threads, grid, table, particleTiles, positions, velocities, mass, forces, stepCounter,
                xcdOrder, stream);
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
