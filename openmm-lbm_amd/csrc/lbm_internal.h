/*
 * Internal declarations shared by the host orchestration (lbm_context.hip) and the kernel
 * launchers (lbm_fluid.hip, lbm_particles.hip).  Not part of the public C-ABI.
 */
#pragma once

#include "lbm_device.h"
#include "../../include/lbm_amd.h"
#include <string>
#include <vector>

namespace lbm {

constexpr int BLOCK = 256;
constexpr int FUSED_THREADS = 512;         // work-group size of the fused pass (measured best on MI355X)
constexpr int FUSED_THREADS_REDUCED = 256; // the same with FP16/BF16 storage
constexpr int ROW_THREADS = 256;           // work-group size of the LDS-staged fused pass (fusedRows)
constexpr int REDUCE_BLOCKS = 1024;       // fixed grid of the FP64 reductions: fixed summation order
// Fixed-point scale of the particle reactions, 2^50: each contribution is rounded to 8.9e-16 lattice units, and a node
// holds up to +-8192 lattice units per step (a bead of 6,000 cell masses at gamma dt = 2 and a velocity difference of
// 0.1 on one node is 1,200).
constexpr double FIXED_SCALE = 1125899906842624.0;

/* Arguments of one fused pull-stream/collide step. */
struct StepArgs {
    Geometry g;
    double omega;
    double gx, gy, gz;            // body acceleration, lattice units
    int removal;                  // 1: subtract rho*ucm from j
    uint32_t seed;
    int64_t step;
    double mu;                    // 3 kT in lattice units (fluctuations)
    uint32_t roundSeed;           // stochastic-rounding seed of the step
    int reactionPlanes;           // layout of the reaction accumulators (reactionIndex)
};

/* One boundary node of a regularized wall or an open face (LBMBoundaries.h:46-109). */
struct BoundaryNode {
    int node;
    int unknown;     // bits 1<<q of the unknown directions
    int solid;       // bits 1<<q of the unknown directions whose source is solid
    int kindFace;    // kind + 8*(face + 1)
};

struct BoundaryArgs {
    Geometry g;
    int count;
    double omega, gx, gy, gz;
    int removal;
    int fluct;
    uint32_t seed;
    int64_t step;
    double mu;
    double faceVelocity[6][3];
    double faceDensity[6];
    uint32_t roundSeed;
};

struct ParticleArgs {
    Geometry g;
    int numParticles;
    int stencil;            // LBM_STENCIL_*
    double dx;              // nm
    double gamma;           // friction*dt
    double kT;              // lattice units
    double velocityScale;   // dx/dt
    double forceScale;      // m_c dx/dt^2
    double cellMass;
    int removal;
    int isStep;
    int randomForce;
    int noiseFP64 = 0;            // particle normals in FP64 (FP64 storage); else FP32, from the hardware functions
    uint32_t seed;
    int64_t step;
    int hasSolids;
    int reactionPlanes;     // layout of the reaction accumulators (reactionIndex)
    int xcdOrder = 1;       // FP16/BF16 stencil coupling: each XCD takes a contiguous run of the particles (0: plain)
    unsigned int* overflow = nullptr;   // set to 1 by a reaction beyond REACTION_LIMIT (reactionInRange)
};

/*
 * The largest particle reaction one contribution may carry, in lattice units.  The fixed-point accumulators of a node
 * hold +-8192 (2^63 of 2^-50); a single contribution of 64 corresponds to a force of about 4e8 kJ/mol/nm at dx = 1 nm,
 * dt = 0.01 ps and the density of water, so one beyond it means a run that has blown up (or a NaN): it is dropped and
 * the flag set, and the host reports LBM_ERROR_RUNTIME at its next check.
 */
constexpr double REACTION_LIMIT = 64.0;

/*
 * The bins of the one-writer spread of the particle reactions (spreadTiles): the lattice in tiles of 32 x 4 x 4 nodes
 * (x fastest; a tile row of 32 nodes is one word of the touched mask); per tile the indices of up to cap beads whose
 * stencil starts in it (their anchor node), and their count.  Two count arrays alternate: the spread of one step
 * clears the one the next step fills.  force: each binned bead's coupling force, lattice units.
 */
constexpr int SPREAD_TX = 32, SPREAD_TY = 4, SPREAD_TZ = 4;
constexpr int SPREAD_THREADS = SPREAD_TX*SPREAD_TY*SPREAD_TZ;
constexpr int SPREAD_MAX_CAP = 128;

/* A particle's stencil along each axis (axisSlot of every slot): the node index and weight of slot sub along axis k,
   and the particle's four normals of the step when the coupling has a random force (particleNormal).  Formed once per
   particle and step, four lanes per particle, for the groups that couple it. */
struct AxisStencil {
    double weight[3][4];
    double normal[4];
    int index[3][4];
};

struct SpreadBins {
    int* count = nullptr;
    int* nextCount = nullptr;
    int* beads = nullptr;
    uint16_t* anchor = nullptr;      // [tile][cap]: the bead's first stencil node within its tile, x | y << 5 | z << 7
    double* force = nullptr;         // [tile][cap][3]: the bead's drag force
    double* slotWeight = nullptr;    // [bead][slot]: the weight of each stencil slot
    int* spilled = nullptr;          // set when a bead of the step found its bin full (spread with the atomics)
    int* nextSpilled = nullptr;      // the flag of the next step, cleared by spreadTiles
    int cap = 0;
    uint32_t ntx = 0, nty = 0, ntz = 0;
};

__device__ __forceinline__ bool reactionInRange(double r, unsigned int* overflow) {
    if (fabs(r) < REACTION_LIMIT)
        return true;
    if (overflow != nullptr)
        atomicOr(overflow, 1u);
    return false;
}

/*
 * Index of component c of a node's three reaction accumulators.  Sparse particles use node-major [node][3]: a stencil
 * row's reactions share cache lines, which the lattice pass reads and clears (13 us less per step on 20,000 Keys beads
 * on 256^3).  Dense particles (the stencils covering more than a quarter of the lattice) use component planes [3][node]:
 * neighbouring stencils then spread their concurrent atomics over three times as many lines (node-major costs 100,000
 * chain beads on 128^3 23% more).
 */
__host__ __device__ inline size_t reactionIndex(uint32_t node, int c, uint32_t n, int planes) {
    return planes ? (size_t) c*n + node : 3*(size_t) node + c;
}

/* Particles are dense when their stencil slots outnumber a quarter of the nodes. */
inline bool denseCoupling(uint64_t numParticles, int slotsPerParticle, uint64_t numNodes) {
    return numParticles*(uint64_t) slotsPerParticle*4 > numNodes;
}

/*
 * One replica of a batch (lbm_batch.hip): a periodic lattice with its own parameters, FP32 populations, particles
 * and reaction accumulators.  The populations of step s are in popA for even s and popB for odd s, so that every
 * step of a batch is the same sequence of launches (the step index is read from device memory).
 */
struct ReplicaDesc {
    StepArgs step;               // geometry, omega, body force, seed, mu (step field set in the kernels)
    ParticleArgs part;           // coupling parameters (step field set in the kernels)
    float* popA;
    float* popB;
    unsigned long long* reactionFixed;
    uint32_t* touched;
    double dt;                   // ps
    uint32_t tileBegin;          // first work-group of this replica in the fused grid
    int particleBegin;           // first particle of this replica in the batch arrays
    uint32_t particleTileBegin;  // first work-group of this replica in the coupling grid
};

/*
 * The MI355X hands the work-groups of a launch to its 8 XCDs in turn (work-group b to XCD b mod 8), and each XCD has
 * its own L2.  The tile of work-group b of n when every XCD takes a contiguous run of the tiles instead, so that the
 * tiles of one replica share an L2: XCD x takes tiles [x q + min(x, r), ...), q = n / 8, r = n mod 8, a bijection of
 * [0, n).
 */
__device__ __forceinline__ uint32_t xcdContiguousTile(uint32_t b, uint32_t n) {
    const uint32_t x = b & 7u, i = b >> 3, q = n >> 3, r = n & 7u;
    return x*q + (x < r ? x : r) + i;
}

/*
 * The tile of work-group b of n when each XCD takes runs of XCD_CHUNK consecutive tiles in turn: XCD x's i-th
 * work-group gets tile ((i / C) 8 + x) C + i mod C (C = XCD_CHUNK) over the first n - n mod 8C work-groups, the rest
 * keep their own index; a bijection of [0, n).  Neighbouring tiles then run on one XCD at about the same time and
 * share its L2 for the lines at their borders, while every XCD still streams through the whole array.
 */
constexpr uint32_t XCD_CHUNK = 64;
__device__ __forceinline__ uint32_t xcdChunkedTile(uint32_t b, uint32_t n) {
    const uint32_t full = n/(8*XCD_CHUNK)*(8*XCD_CHUNK);
    if (b >= full)
        return b;
    const uint32_t x = b & 7u, i = b >> 3;
    return ((i/XCD_CHUNK)*8 + x)*XCD_CHUNK + i%XCD_CHUNK;
}

/* Host-side launchers (defined in lbm_fluid.hip / lbm_particles.hip); xcdOrder: coupling tiles by xcdContiguousTile,
   fused tiles by xcdChunkedTile. */
void launchFusedBatch(int threads, bool fluct, const ReplicaDesc* table, const uint16_t* tileReplica, uint32_t numTiles,
        const int64_t* stepCounter, bool xcdOrder, hipStream_t stream);
void launchCoupleBatch(int stencil, int threads, const ReplicaDesc* table, const uint16_t* particleTiles,
        int numParticleTiles, const double* positions, const double* velocities, const double* mass, double* forces,
        const int64_t* stepCounter, bool xcdOrder, hipStream_t stream);
int coupleLanes(int stencil);
void launchVerletBatch(const ReplicaDesc* table, const uint16_t* particleReplica, int numParticles, double* positions,
        double* velocities, const double* forces, const double* invMass, int64_t* stepCounter, hipStream_t stream);
/* Which fused-pass kernel a context runs: fusedStep with 32-bit or 64-bit offsets, or fusedRows (LDS-staged pulls). */
enum FusedPath { FUSED_DIRECT = 0, FUSED_WIDE = 1, FUSED_ROWS = 2 };
bool fusedRowsFit(uint32_t nx, bool wide, uint32_t rowShift);
void launchFused(int storage, int path, bool solids, bool reaction, bool fluct, const void* src, void* dst,
        const uint32_t* mask, unsigned long long* reactionFixed, uint32_t* touched, const double* ucm, const StepArgs& a,
        hipStream_t stream);
void launchMomentReduce(int storage, bool solids, const void* src, const uint32_t* mask, const Geometry& g,
        double* partials, hipStream_t stream);
void launchFinishReduce(const double* partials, double numFluid, double* totals, double* ucm, hipStream_t stream);
void launchBounceBack(int storage, void* dst, const int* wallNodes, const int* wallLinks, int count, const Geometry& g,
        double* partials, hipStream_t stream);
void launchBoundaries(int storage, const void* src, void* dst, const BoundaryNode* nodes, const double* ucm,
        const BoundaryArgs& a, double* partials, hipStream_t stream);
void launchGather(int storage, const void* src, double* out, const Geometry& g, hipStream_t stream);
void launchScatter(int storage, void* dst, const double* in, const Geometry& g, hipStream_t stream);
/* The entries that launchFields writes: pad = 0, the lattice in node order; pad = 1, the lattice with a layer one node
   thick around it (getFluidFields with halo), whose entries hold the node across the periodic boundary for the fields
   whose halo is exchanged, NaN for the others and beyond the open faces. */
struct FieldsLayout {
    int pad = 0;
    int densityHalo = 0, velocityHalo = 0;
};
uint64_t fieldsCount(const Geometry& g, const FieldsLayout& out);
void launchFields(int storage, const void* src, const uint32_t* mask, const Geometry& g, const FieldsLayout& out,
        double density, double velocityScale, double gx, double gy, double gz, double* rhoOut, double* uOut,
        hipStream_t stream);
void launchInitEquilibrium(int storage, void* dst, const uint32_t* mask, const Geometry& g, double ux, double uy,
        double uz, hipStream_t stream);
void launchFill(int storage, void* dst, size_t count, hipStream_t stream);
void launchReduceWall(const double* partials, int count, double* out, hipStream_t stream);

/* The reflection writes its wall momentum to allPartials[3*ownOffset...] and clears the rest of allPartials[0, 3*count). */
void launchReflect(const double* positions, double* velocities, const int* particles, const double* mass,
        const uint32_t* mask, const ParticleArgs& a, double* allPartials, int count, int ownOffset, hipStream_t stream);
/* order: the coupled particles in the order to process them (a permutation of 0..numParticles-1), or nullptr;
   systemForces: the caller's force array [particle][3] (kJ/mol/nm) to add the coupling forces to, or nullptr. */
void launchCouple(int storage, const void* src, const uint32_t* mask, const double* positions, const double* velocities,
        const int* particles, const int* order, const double* mass, const double* ucm, const ParticleArgs& a,
        double* particleForce, unsigned long long* reactionFixed, uint32_t* touched, double* partials,
        double* systemForces, const SpreadBins* bins, const double4* field, AxisStencil* stencils, hipStream_t stream);
/* The moments (jx, jy, jz, rho) of every fluid node into field, for launchCouple's dense path; with stencils, the
   particles' stencils (launchStencilSetup) in the same launch. */
void launchMomentField(int storage, const void* src, const uint32_t* mask, const ParticleArgs& a, const double* ucm,
        double4* field, const double* positions, const int* particles, AxisStencil* stencils, hipStream_t stream);
/* The particles' stencils (AxisStencil) for launchCouple, in a launch of their own (launchMomentField forms them
   alongside the moment field when given them). */
void launchStencilSetup(const double* positions, const int* particles, const ParticleArgs& a, AxisStencil* stencils,
        hipStream_t stream);
/* The one-writer spread of the binned beads' reactions, one work-group per tile (after launchCouple with bins). */
void launchSpreadTiles(const double* positions, const int* particles, const uint32_t* mask, const ParticleArgs& a,
        const SpreadBins& bins, unsigned long long* reactionFixed, uint32_t* touched, hipStream_t stream);
/* Whether the lattice suits the tile bins: every axis periodic, nx a multiple of 32 and ny, nz of 4, at least two tiles
   along each axis. */
bool spreadTilesFit(int nx, int ny, int nz, bool anyOpen);

/*
 * The coupled particles sorted by nearest node (z slowest, descending): processed in this order, the particles of one
 * work-group and of neighbouring work-groups read populations from nearby rows of the lattice.  The order changes no result:
 * each particle's coupling is independent of the others and the reactions add as integers.  It is refreshed every
 * COUPLE_ORDER_STEPS steps (particles move a small fraction of a node per step).  Throws std::runtime_error.
 */
constexpr int COUPLE_ORDER_STEPS = 128;
struct CoupleOrder;
CoupleOrder* coupleOrderCreate(int numParticles, uint32_t numNodes);
void coupleOrderDestroy(CoupleOrder* order);
/* The order for a coupling at `step`, sorted anew from the positions when stale. */
const int* coupleOrderAt(CoupleOrder* order, int64_t step, const double* positions, const int* particles,
        const ParticleArgs& a, hipStream_t stream);
void launchAddForces(const double* particleForce, const int* particles, int count, double* forces, hipStream_t stream);
void launchParticleNodes(const double* positions, const int* particles, const ParticleArgs& a, int* nodes,
        hipStream_t stream);

/* The centred drag (lbm_centred.hip): its buffers for numParticles coupled particles, and one coupling.  otherForces
   (kJ/mol/nm, all System particles, may be null) are the forces on the particles before the coupling; g is the body
   acceleration in lattice units.  Throws std::runtime_error. */
struct CentredState;
CentredState* centredCreate(int numParticles, int stencil, uint32_t numNodes);
void centredDestroy(CentredState* state);
void centredCouple(CentredState* state, int storage, const void* src, const uint32_t* mask, const double* positions,
        const double* velocities, const double* otherForces, const int* particles, const double* mass,
        const double* ucm, const ParticleArgs& a, const double g[3], double* particleForce,
        unsigned long long* reactionFixed, uint32_t* touched, double* partials, hipStream_t stream);

/* Nodes per axis of an interpolation stencil (LBMStencils.h:42-104). */
__host__ __device__ inline int stencilWidth(int stencil) {
    switch (stencil) {
    case LBM_STENCIL_TRILINEAR: return 2;
    case LBM_STENCIL_THREE_POINT: return 3;
    case LBM_STENCIL_KEYS: return 4;
    default: return 1;
    }
}

}  // namespace lbm
