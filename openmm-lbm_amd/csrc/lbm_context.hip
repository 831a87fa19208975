/*
 * Host orchestration of the gfx950 LBMForce platform and its C-ABI (include/lbm_amd.h).
 *
 * Restates, for one GPU and one domain, the host side of the reference's GPU platforms
 * (CommonCalcLBMForceKernel, platforms/common/src/CommonLBMKernels.cpp:160-2097) behind the
 * kernel interface CalcLBMForceKernel (openmmapi/include/LBMKernels.h:97-189), with the step
 * semantics of the Reference platform (ReferenceLBMKernels.cpp:308-401):
 *
 *   beginStep:  step index <- step count; reflect the coupled particles at solid nodes.
 *   execute:    first call after beginStep -> advanceFluid; later calls between steps -> the
 *               coupling of the next step without its reaction (same random numbers, since they
 *               are keyed by the step index); a repeated evaluation reuses the forces.
 *   advanceFluid (one lattice step):
 *      [removal due]  momentReduce(A) -> FP64 u_cm
 *      [particles]    coupleParticles(A) -> forces, fixed-point reactions, touched mask
 *      fusedStep(A -> B)   (pull-stream + collision + Guo + reaction (consumed) + noise)
 *      [bounce-back]  bounceBack(B) + wall momentum
 *      [boundaries]   boundaryKernel(A, B) + wall momentum
 *      swap A, B; step index++; [Mach check due] momentReduce(A) -> host
 *
 * Everything runs in order on one HIP stream owned by the context.
 */
#include "lbm_mlp.h"
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <sstream>
#include <stdexcept>

namespace lbm {

static thread_local std::string g_lastError;

void setLastError(const std::string& message) {
    g_lastError = message;
}

struct Error : public std::runtime_error {
    int status;
    Error(int s, const std::string& m) : std::runtime_error(m), status(s) {}
};

#define HIP_CHECK(call) do { \
        hipError_t err_ = (call); \
        if (err_ != hipSuccess) { \
            std::stringstream msg_; \
            msg_ << "HIP error " << hipGetErrorString(err_) << " at " << __FILE__ << ":" << __LINE__; \
            throw Error(LBM_ERROR_RUNTIME, msg_.str()); \
        } \
    } while (0)

static FastDiv makeFastDiv(uint32_t d) {
    FastDiv f;
    f.d = d;
    f.l = 0;
    while ((1ull << f.l) < d)
        f.l++;
    f.m = d <= 1 ? 0 : (uint32_t) (((1ull << 32)*((1ull << f.l) - d))/d + 1);
    return f;
}

template <typename T>
static T* deviceAlloc(size_t count) {
    if (count == 0)
        return nullptr;
    void* ptr = nullptr;
    HIP_CHECK(hipMalloc(&ptr, count*sizeof(T)));
    return (T*) ptr;
}

static size_t storageSize(int storage) {
    switch (storage) {
    case LBM_STORAGE_FP16:
    case LBM_STORAGE_BF16:
        return 2;
    case LBM_STORAGE_FP64:
        return 8;
    default:
        return 4;
    }
}

static const int HCX[NQ] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1};
static const int HCY[NQ] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 1, -1, 1, -1, 0, 0, 0, 0};
static const int HCZ[NQ] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, 1, -1};

/* The kernel's part of a checkpoint: this header, the FP64 state [q*N + node], the wall momentum of the last step (3
   doubles).  Version 2 adds the seed of the random numbers, which a seed of 0 chooses per Context: with it a restart
   draws the same numbers.  Version 1 (without the seed) is still read. */
struct CheckpointHeader {
    char magic[8];
    int32_t version, nx, ny, nz, numParticles, dragScheme;
    uint32_t seed;
};
constexpr size_t CHECKPOINT_V1_HEADER = 8 + 6*sizeof(int32_t);

}  // namespace lbm

using namespace lbm;

struct lbm_context {
    lbm_params p;
    Geometry g;
    size_t elem = 4;
    hipStream_t stream = nullptr;
    void* popA = nullptr;
    void* popB = nullptr;
    uint32_t* mask = nullptr;
    std::vector<char> hostFluid;          // empty: all fluid
    double numFluid = 0;
    uint32_t seed = 0;
    // Walls and boundaries.
    int* wallNodes = nullptr;
    int* wallLinks = nullptr;
    int numWall = 0;
    double staticWall[3] = {0, 0, 0};     // sum over the links of 2 c_q w_q (lattice units)
    BoundaryNode* boundary = nullptr;
    int numBoundary = 0;
    // Reductions.
    double* reducePartials = nullptr;
    double* totals = nullptr;
    double* ucm = nullptr;
    unsigned int* overflow = nullptr;     // set by a particle reaction beyond REACTION_LIMIT
    double* wallPartials = nullptr;
    int wallBounceOffset = 0, wallBoundaryOffset = 0, wallCoupleOffset = 0, wallReflectOffset = 0, wallPartialCount = 0;
    bool wallStepDone = false;
    bool wallOverride = false;
    double wallOverrideValue[3] = {0, 0, 0};
    // Particles.
    int numParticles = 0, numSystem = 0, numSlots = 0;
    bool dense = false;                   // denseCoupling(): index order, component-plane reactions
    bool wide = false;                    // 64-bit element indices in the fused pass (populations above 4 GiB)
    bool coupleXcd = true;                // LBM_AMD_COUPLE_XCD: the coupling's particles in XCD-contiguous runs
    int fusedPath = FUSED_DIRECT;         // the fused-pass kernel (FusedPath)
    int* particles = nullptr;
    double* mass = nullptr;               // lattice units m/m_c
    std::vector<int> hostParticles;
    double* particleForce = nullptr;
    std::vector<double> hostForce;
    unsigned long long* reactionFixed = nullptr;   // [node][3] fixed-point accumulators of the particle reactions
    uint32_t* touched = nullptr;                   // one bit per node: accumulators holding a reaction
    double* stagePos = nullptr;
    double* stageVel = nullptr;
    // Learned closure (include/lbm_amd_closure.h): the network and its per-particle buffers.
    lbm_mlp* closure = nullptr;
    _Float16* closureX = nullptr;             // normalized inputs [n][96]
    uint32_t* closureNode = nullptr;          // nearest node
    double* closureU = nullptr;               // u(nearest), lattice units
    float* closureDu = nullptr;               // network output
    uint8_t* closureOod = nullptr;            // out-of-distribution flags
    uint8_t* closureFrame = nullptr;          // each particle's frame (reflections and axis order of its inputs)
    unsigned long long* closureCounters = nullptr;   // couplings through the network / the fallback
    double* closureField = nullptr;           // node velocities [node][3], when the particles are dense
    // Centred drag: its solver state and the other forces of the host path.
    CentredState* centred = nullptr;
    double* stageForce = nullptr;
    // The processing order of the stencil coupling (explicit drag).
    CoupleOrder* order = nullptr;
    // The one-writer spread of the explicit stencil coupling (spreadTiles), where the lattice allows it: its bins, and
    // which of the two count arrays the next lattice step fills.
    bool binned = false;
    SpreadBins bins;
    // Dense particles interpolate from the moments of every node (momentField) rather than pulling per slot.
    double4* momentField = nullptr;
    // The particles' stencils along each axis, formed once per particle before the stencil coupling (stencilSetup).
    AxisStencil* axisStencils = nullptr;
    int* binCounts[2] = {nullptr, nullptr};
    int* binSpilled = nullptr;               // [2]: the spilled flags of the two count arrays
    int binParity = 0;
    // Set by lbm_execute_device for the duration of one call: the caller's force array, into which the stencil coupling
    // adds the forces of the coupled particles itself (forcesAdded then says so, and no separate addForces runs).
    double* forceTarget = nullptr;
    bool forcesAdded = false;
    // Timing of the fused pass with HIP events on the context's stream (diagnostics).
    bool profile = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> events;
    // Step state (ReferenceLBMKernels.h:158-176).
    bool stepPending = false, stepForcesCurrent = false;
    int64_t stepIndex = 0;

    bool fluct() const {
        return p.fluid_fluctuations && p.fluid_kT > 0;
    }
    double cellMass() const {
        return p.density*p.dx*p.dx*p.dx;
    }
    bool openAxis(int a) const {
        return p.face_boundary[2*a] != LBM_BOUNDARY_PERIODIC;
    }
    bool removalDue(int64_t step) const {
        return p.momentum_removal_frequency > 0 && step % p.momentum_removal_frequency == 0;
    }
    int nodeIndex(int i, int j, int k) const {
        return i + p.nx*(j + p.ny*k);
    }
};

namespace {

void findWalls(lbm_context* c, std::vector<int>& nodes, std::vector<int>& links) {
    // Bounce-back links (LBMBoundaries.h:118-149): fluid node x, direction q with x + c_q solid,
    // without the links that cross an open face.
    const lbm_params& p = c->p;
    if (c->hostFluid.empty() || p.wall_scheme != LBM_WALL_BOUNCE_BACK)
        return;
    int size[3] = {p.nx, p.ny, p.nz};
    int n = p.nx*p.ny*p.nz;
    // sum 2 c_q w_q over the links, in units of 1/18 (2 w_q = 2/18 along the axes, 1/18 on the diagonals): summed as
    // integers it is exact, and exactly zero for a wall with fluid on both sides.
    long long units[3] = {0, 0, 0};
    for (int node = 0; node < n; node++) {
        if (!c->hostFluid[node])
            continue;
        int idx[3] = {node % p.nx, (node/p.nx) % p.ny, node/(p.nx*p.ny)};
        int bits = 0;
        for (int q = 1; q < NQ; q++) {
            int cq[3] = {HCX[q], HCY[q], HCZ[q]}, t[3];
            bool crosses = false;
            for (int a = 0; a < 3; a++) {
                t[a] = idx[a] + cq[a];
                crosses = crosses || (c->openAxis(a) && (t[a] < 0 || t[a] >= size[a]));
                t[a] = (t[a] + size[a]) % size[a];
            }
            if (!crosses && !c->hostFluid[c->nodeIndex(t[0], t[1], t[2])]) {
                bits |= 1 << q;
                for (int a = 0; a < 3; a++)
                    units[a] += cq[a]*(q < 7 ? 2 : 1);
            }
        }
        if (bits) {
            nodes.push_back(node);
            links.push_back(bits);
        }
    }
    for (int a = 0; a < 3; a++)
        c->staticWall[a] = (double) units[a]/18.0;
}

void findBoundaries(lbm_context* c, std::vector<BoundaryNode>& list) {
    // Boundary nodes of regularized walls and open faces (LBMBoundaries.h:46-109).
    const lbm_params& p = c->p;
    bool regularized = !c->hostFluid.empty() && p.wall_scheme == LBM_WALL_REGULARIZED;
    bool open = c->openAxis(0) || c->openAxis(1) || c->openAxis(2);
    if (!regularized && !open)
        return;
    int size[3] = {p.nx, p.ny, p.nz};
    int n = p.nx*p.ny*p.nz;
    for (int node = 0; node < n; node++) {
        if (!c->hostFluid.empty() && !c->hostFluid[node])
            continue;
        int idx[3] = {node % p.nx, (node/p.nx) % p.ny, node/(p.nx*p.ny)};
        int solidBits = 0, outsideBits = 0;
        for (int q = 1; q < NQ; q++) {
            int cq[3] = {HCX[q], HCY[q], HCZ[q]}, s[3];
            bool beyond = false;
            for (int a = 0; a < 3; a++) {
                s[a] = idx[a] - cq[a];
                if (s[a] < 0 || s[a] >= size[a]) {
                    beyond = beyond || c->openAxis(a);
                    s[a] = (s[a] + size[a]) % size[a];
                }
            }
            if (beyond)
                outsideBits |= 1 << q;
            else if (regularized && !c->hostFluid[c->nodeIndex(s[0], s[1], s[2])])
                solidBits |= 1 << q;
        }
        if (solidBits == 0 && outsideBits == 0)
            continue;
        int kind = 0, face = -1;
        if (solidBits == 0) {
            int velocityFace = -1, densityFace = -1, numDensity = 0;
            for (int f = 0; f < 6; f++) {
                int a = f/2;
                if (!c->openAxis(a) || idx[a] != (f % 2 == 0 ? 0 : size[a] - 1))
                    continue;
                int type = p.face_boundary[f];
                if ((type == LBM_BOUNDARY_VELOCITY || type == LBM_BOUNDARY_DENSITY_VELOCITY) && velocityFace < 0)
                    velocityFace = f;
                if (type == LBM_BOUNDARY_DENSITY && numDensity++ == 0)
                    densityFace = f;
            }
            if (velocityFace >= 0) {
                kind = p.face_boundary[velocityFace] == LBM_BOUNDARY_DENSITY_VELOCITY ? 4 : 1;
                face = velocityFace;
            }
            else {
                kind = numDensity == 1 ? 2 : 3;
                face = densityFace;
            }
        }
        list.push_back(BoundaryNode{node, solidBits | outsideBits, solidBits, kind + 8*(face + 1)});
    }
}

StepArgs stepArgs(lbm_context* c, int64_t step) {
    StepArgs a;
    a.g = c->g;
    a.omega = c->p.omega;
    a.gx = c->p.body_acceleration[0];
    a.gy = c->p.body_acceleration[1];
    a.gz = c->p.body_acceleration[2];
    a.removal = c->removalDue(step) ? 1 : 0;
    a.seed = c->seed;
    a.step = step;
    a.mu = 3.0*c->p.fluid_kT*c->p.dt*c->p.dt/(c->cellMass()*c->p.dx*c->p.dx);
    a.roundSeed = (uint32_t) (c->seed*0x9E3779B9u) ^ (uint32_t) (step*0x85EBCA6Bll) ^ (uint32_t) ((uint64_t) step >> 32);
    a.reactionPlanes = c->dense ? 1 : 0;
    return a;
}

ParticleArgs particleArgs(lbm_context* c, int64_t step, bool isStep) {
    ParticleArgs a;
    a.g = c->g;
    a.numParticles = c->numParticles;
    a.stencil = c->p.interpolation_stencil;
    a.dx = c->p.dx;
    a.gamma = c->p.friction*c->p.dt;
    a.kT = c->p.kT*c->p.dt*c->p.dt/(c->cellMass()*c->p.dx*c->p.dx);
    a.velocityScale = c->p.dx/c->p.dt;
    a.forceScale = c->cellMass()*c->p.dx/(c->p.dt*c->p.dt);
    a.cellMass = c->cellMass();
    a.removal = c->removalDue(step) ? 1 : 0;
    a.isStep = isStep ? 1 : 0;
    a.randomForce = (a.kT > 0 && a.gamma > 0) ? 1 : 0;
    a.noiseFP64 = c->p.storage == LBM_STORAGE_FP64 ? 1 : 0;
    a.seed = c->seed;
    a.step = step;
    a.hasSolids = c->mask != nullptr;
    a.reactionPlanes = c->dense ? 1 : 0;
    a.xcdOrder = c->coupleXcd ? 1 : 0;
    a.overflow = c->overflow;
    return a;
}

void reduceMoments(lbm_context* c) {
    launchMomentReduce(c->p.storage, c->mask != nullptr, c->popA, c->mask, c->g, c->reducePartials, c->stream);
    launchFinishReduce(c->reducePartials, c->numFluid, c->totals, c->ucm, c->stream);
}

void couple(lbm_context* c, const double* dPos, const double* dVel, const double* dForces, int64_t step, bool isStep) {
    ParticleArgs a = particleArgs(c, step, isStep);
    if (a.removal)
        reduceMoments(c);
    double* partials = isStep && c->mask != nullptr ? c->wallPartials + 3*(size_t) c->wallCoupleOffset : nullptr;
    if (c->centred != nullptr) {
        // Centred drag: the other forces (dForces, before the coupling adds to them) enter the particle velocities.
        centredCouple(c->centred, c->p.storage, c->popA, c->mask, dPos, dVel, dForces, c->particles, c->mass, c->ucm, a,
                c->p.body_acceleration, c->particleForce, c->reactionFixed, c->touched, partials, c->stream);
        return;
    }
    if (c->closure != nullptr) {
        // Nearest node + network for the particles in distribution, Keys for the others.
        ClosureNorm norm = mlpClosureNorm(c->closure);
        if (c->closureField != nullptr)
            launchClosureField(c->p.storage, c->popA, c->mask, c->ucm, a, c->closureField, c->stream);
        launchClosureFeatures(c->p.storage, c->popA, c->mask, dPos, dVel, c->particles, c->ucm, a, norm,
                c->closureField, c->closureX, nullptr, c->closureNode, c->closureU, c->closureOod, c->closureFrame,
                c->stream);
        launchMlpInfer(mlpDeviceWeights(c->closure), c->closureX, c->numParticles, norm.outScale, c->closureDu,
                c->stream);
        launchClosureCouple(dVel, c->particles, c->mass, a, c->closureNode, c->closureU, c->closureDu, c->closureOod,
                c->closureFrame, c->particleForce, c->reactionFixed, c->touched, c->closureCounters, c->stream);
        launchClosureFallback(c->p.storage, c->popA, c->mask, dPos, dVel, c->particles, c->mass, c->ucm, a,
                c->closureOod, c->particleForce, c->reactionFixed, c->touched, partials, c->stream);
        return;
    }
    const int* order = nullptr;
    try {
        if (c->order != nullptr)
            order = coupleOrderAt(c->order, step, dPos, c->particles, a, c->stream);
    }
    catch (const std::runtime_error& e) {
        throw Error(LBM_ERROR_RUNTIME, e.what());
    }
    const SpreadBins* bins = nullptr;
    if (c->binned && isStep) {
        c->bins.count = c->binCounts[c->binParity];
        c->bins.nextCount = c->binCounts[c->binParity ^ 1];
        c->bins.spilled = c->binSpilled + c->binParity;
        c->bins.nextSpilled = c->binSpilled + (c->binParity ^ 1);
        c->binParity ^= 1;
        bins = &c->bins;
    }
    if (c->momentField != nullptr)
        launchMomentField(c->p.storage, c->popA, c->mask, a, c->ucm, c->momentField, dPos, c->particles, c->axisStencils,
                c->stream);
    else if (c->axisStencils != nullptr)
        launchStencilSetup(dPos, c->particles, a, c->axisStencils, c->stream);
    launchCouple(c->p.storage, c->popA, c->mask, dPos, dVel, c->particles, order, c->mass, c->ucm, a, c->particleForce,
            c->reactionFixed, c->touched, partials, c->forceTarget, bins, c->momentField, c->axisStencils, c->stream);
    if (bins != nullptr)
        launchSpreadTiles(dPos, c->particles, c->mask, a, *bins, c->reactionFixed, c->touched, c->stream);
    c->forcesAdded = c->forceTarget != nullptr;
}

void advanceFluid(lbm_context* c, const double* dPos, const double* dVel, const double* dForces) {
    StepArgs a = stepArgs(c, c->stepIndex);
    bool particles = c->numParticles > 0;
    if (particles)
        couple(c, dPos, dVel, dForces, c->stepIndex, true);   // runs the removal reduction when due
    else if (a.removal)
        reduceMoments(c);
    std::pair<hipEvent_t, hipEvent_t> ev{nullptr, nullptr};
    if (c->profile) {
        HIP_CHECK(hipEventCreate(&ev.first));
        HIP_CHECK(hipEventCreate(&ev.second));
        HIP_CHECK(hipEventRecord(ev.first, c->stream));
    }
    launchFused(c->p.storage, c->fusedPath, c->mask != nullptr, particles, c->fluct(), c->popA, c->popB, c->mask,
            c->reactionFixed, c->touched, c->ucm, a, c->stream);
    if (c->profile) {
        HIP_CHECK(hipEventRecord(ev.second, c->stream));
        c->events.push_back(ev);
    }
    if (c->numWall > 0)
        launchBounceBack(c->p.storage, c->popB, c->wallNodes, c->wallLinks, c->numWall, c->g,
                c->wallPartials + 3*(size_t) c->wallBounceOffset, c->stream);
    if (c->numBoundary > 0) {
        BoundaryArgs b;
        b.g = c->g;
        b.count = c->numBoundary;
        b.omega = c->p.omega;
        b.gx = a.gx;
        b.gy = a.gy;
        b.gz = a.gz;
        b.removal = a.removal;
        b.fluct = c->fluct() ? 1 : 0;
        b.seed = c->seed;
        b.step = c->stepIndex;
        b.mu = a.mu;
        for (int f = 0; f < 6; f++) {
            for (int k = 0; k < 3; k++)
                b.faceVelocity[f][k] = c->p.face_velocity[f][k];
            b.faceDensity[f] = c->p.face_density[f];
        }
        b.roundSeed = a.roundSeed ^ 0x5BD1E995u;
        launchBoundaries(c->p.storage, c->popA, c->popB, c->boundary, c->ucm, b,
                c->wallPartials + 3*(size_t) c->wallBoundaryOffset, c->stream);
    }
    HIP_CHECK(hipGetLastError());
    std::swap(c->popA, c->popB);
    c->wallStepDone = true;
    c->stepForcesCurrent = true;
    c->stepIndex++;
    bool machDue = c->p.mach_check_frequency > 0 && c->stepIndex % c->p.mach_check_frequency == 0;
    if (particles && (machDue || c->stepIndex % 100 == 0)) {
        // A particle reaction beyond the range of the fixed-point accumulators was dropped: the run has blown up.
        unsigned int flag = 0;
        HIP_CHECK(hipMemcpyAsync(&flag, c->overflow, sizeof(flag), hipMemcpyDeviceToHost, c->stream));
        HIP_CHECK(hipStreamSynchronize(c->stream));
        if (flag != 0) {
            std::stringstream msg;
            msg << "LBMForce: a particle reaction beyond " << REACTION_LIMIT << " lattice units (or not a number) by "
                << "lattice step " << c->stepIndex << ": the coupling forces have diverged.  Reduce the time step, the "
                << "friction or the forces on the particles.";
            throw Error(LBM_ERROR_RUNTIME, msg.str());
        }
    }
    if (machDue) {
        reduceMoments(c);
        double totals[5];
        HIP_CHECK(hipMemcpyAsync(totals, c->totals, sizeof(totals), hipMemcpyDeviceToHost, c->stream));
        HIP_CHECK(hipStreamSynchronize(c->stream));
        double mach = std::sqrt(3.0*totals[4]);
        if (!(mach <= c->p.mach_number_limit)) {
            std::stringstream msg;
            msg << "LBMForce: the Mach number of the fluid is " << mach << " after " << c->stepIndex << " lattice steps, "
                << "above the limit " << c->p.mach_number_limit << " (docs/theory.md, section 6). Reduce the forces on the "
                << "fluid, the time step or the friction.";
            throw Error(LBM_ERROR_MACH, msg.str());
        }
    }
}

void beginStepDevice(lbm_context* c, int64_t step, const double* dPos, double* dVel) {
    c->stepIndex = step;
    c->stepPending = true;
    c->wallStepDone = false;
    c->wallOverride = false;
    if (c->mask != nullptr && c->numParticles > 0) {
        // The reflection also clears the step's other wall partial sums.
        ParticleArgs a = particleArgs(c, step, true);
        launchReflect(dPos, dVel, c->particles, c->mass, c->mask, a, c->wallPartials, c->wallPartialCount,
                c->wallReflectOffset, c->stream);
    }
    else if (c->wallPartialCount > 0)
        launchFill(LBM_STORAGE_FP64, c->wallPartials, 3*(size_t) c->wallPartialCount, c->stream);
    HIP_CHECK(hipGetLastError());
}

/* Returns true if the forces of the coupled particles in c->particleForce are to be added. */
void executeDevice(lbm_context* c, int64_t stepCount, bool includeForces, const double* dPos, const double* dVel,
        const double* dForces) {
    bool repeated = c->stepForcesCurrent && stepCount == c->stepIndex - 1;
    if (c->stepPending) {
        c->stepPending = false;
        advanceFluid(c, dPos, dVel, dForces);
    }
    else if (includeForces && c->numParticles > 0 && stepCount > 0 && !repeated) {
        c->stepForcesCurrent = false;
        couple(c, dPos, dVel, dForces, stepCount, false);
    }
    HIP_CHECK(hipGetLastError());
}

void clearEvents(lbm_context* c) {
    for (auto& e : c->events) {
        (void) hipEventDestroy(e.first);
        (void) hipEventDestroy(e.second);
    }
    c->events.clear();
}

void destroy(lbm_context* c) {
    if (c->stream)
        (void) hipStreamSynchronize(c->stream);
    clearEvents(c);
    centredDestroy(c->centred);
    coupleOrderDestroy(c->order);
    void* buffers[] = {c->popA, c->popB, c->mask, c->wallNodes, c->wallLinks, c->boundary, c->reducePartials, c->totals,
                       c->overflow, c->binCounts[0], c->binCounts[1], c->binSpilled, c->bins.beads, c->bins.anchor, c->bins.force,
                       c->bins.slotWeight, c->momentField, c->axisStencils,
                       c->ucm, c->wallPartials, c->particles, c->mass, c->particleForce, c->reactionFixed,
                       c->touched, c->stagePos, c->stageVel, c->closureX, c->closureNode, c->closureU, c->closureDu,
                       c->closureOod, c->closureFrame, c->closureCounters, c->closureField, c->stageForce};
    for (void* b : buffers)
        if (b)
            (void) hipFree(b);
    if (c->stream)
        (void) hipStreamDestroy(c->stream);
    delete c;
}

template <typename T>
void upload(T* dst, const std::vector<T>& src, hipStream_t stream) {
    if (!src.empty())
        HIP_CHECK(hipMemcpyAsync(dst, src.data(), src.size()*sizeof(T), hipMemcpyHostToDevice, stream));
}

void create(const lbm_params* params, const int* solid, int numSolid, const int* parts, const double* masses,
        int numParts, int numSys, lbm_context* c) {
    const lbm_params& p = *params;
    c->p = p;
    if (p.nx <= 0 || p.ny <= 0 || p.nz <= 0)
        throw Error(LBM_ERROR_INVALID, "LBMForce: the grid size must be set to positive values with setGridSize()");
    if (p.drag_scheme != LBM_DRAG_EXPLICIT && p.drag_scheme != LBM_DRAG_CENTERED)
        throw Error(LBM_ERROR_INVALID, "LBMForce: unknown drag scheme");
    if (p.storage < LBM_STORAGE_FP32 || p.storage > LBM_STORAGE_FP64)
        throw Error(LBM_ERROR_INVALID, "LBMForce: unknown storage type");
    if (numParts < 0 || numSys < 0 || (numParts > 0 && (!parts || !masses)))
        throw Error(LBM_ERROR_INVALID, "lbm_create: invalid particle arrays");
    uint64_t n = (uint64_t) p.nx*p.ny*p.nz;
    if (n > 0x7FFFFFFFull)
        throw Error(LBM_ERROR_INVALID, "LBMForce: the lattice has more than 2147483647 nodes");
    c->g.nx = p.nx;
    c->g.ny = p.ny;
    c->g.nz = p.nz;
    c->g.nxy = (uint32_t) (p.nx*p.ny);
    c->g.n = (uint32_t) n;
    c->g.divx = makeFastDiv(p.nx);
    c->g.divxy = makeFastDiv(c->g.nxy);
    c->g.openX = c->openAxis(0);
    c->g.openY = c->openAxis(1);
    c->g.openZ = c->openAxis(2);
    c->elem = storageSize(p.storage);
    // Rows per block of the population layout (lbm_device.h): one by default; LBM_AMD_ROW_BLOCK=2 or 4 groups rows
    // (where they divide ny).  Measured on 256^3 with FP16 storage the lattice pass takes 0.213, 0.232 and 0.221 ms
    // with 1, 2 and 4 rows per block, with FP32 0.435, 0.432 and 0.434 ms.
    const char* blockEnv = getenv("LBM_AMD_ROW_BLOCK");
    int rowBlock = blockEnv ? std::atoi(blockEnv) : 1;
    c->g.rowShift = rowBlock == 4 ? 2u : rowBlock == 2 ? 1u : 0u;
    while (c->g.rowShift > 0 && p.ny % (1 << c->g.rowShift) != 0)
        c->g.rowShift--;
    c->g.qstride = (uint32_t) p.nx << c->g.rowShift;
    if (n > (1ull << 30))
        throw Error(LBM_ERROR_INVALID, "LBMForce: the lattice is too large for one GPU: at most 2^30 nodes");
    // The fused pass addresses populations by 32-bit byte offsets while an array of them fits in 4 GiB (up to 56.5 M
    // nodes with FP32 storage), and by 64-bit indices beyond.  LBM_AMD_WIDE_ADDRESSING=1 takes the 64-bit path always.
    const char* wideEnv = getenv("LBM_AMD_WIDE_ADDRESSING");
    c->wide = NQ*n*c->elem > (1ull << 32) || (wideEnv && wideEnv[0] == '1');
    // The LDS-staged pass (fusedRows) where the row length allows it; LBM_AMD_FUSED_ROWS=0 or 1 forces the choice
    // (where it fits).
    const char* rowsEnv = getenv("LBM_AMD_FUSED_ROWS");
    bool rows = rowsEnv ? rowsEnv[0] == '1' : false;
    c->fusedPath = c->wide ? FUSED_WIDE : (rows && fusedRowsFit((uint32_t) p.nx, false, c->g.rowShift) ? FUSED_ROWS : FUSED_DIRECT);
    c->seed = (uint32_t) p.random_number_seed;
    if (c->seed == 0)
        c->seed = (uint32_t) std::chrono::high_resolution_clock::now().time_since_epoch().count() | 1u;
    HIP_CHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    c->popA = deviceAlloc<char>(NQ*n*c->elem);
    c->popB = deviceAlloc<char>(NQ*n*c->elem);
    // Solid nodes.
    c->numFluid = (double) n;
    if (numSolid > 0) {
        c->hostFluid.assign(n, 1);
        std::vector<uint32_t> bits((n + 31)/32, 0);
        for (int s = 0; s < numSolid; s++) {
            int node = solid[s];
            if (node < 0 || (uint64_t) node >= n)
                throw Error(LBM_ERROR_INVALID, "LBMForce: a solid node index is out of range");
            if (!c->hostFluid[node])
                throw Error(LBM_ERROR_INVALID, "LBMForce: a solid node is listed more than once");
            c->hostFluid[node] = 0;
            bits[node >> 5] |= 1u << (node & 31);
        }
        c->numFluid = (double) (n - numSolid);
        c->mask = deviceAlloc<uint32_t>(bits.size());
        upload(c->mask, bits, c->stream);
    }
    std::vector<int> wallNodes, wallLinks;
    findWalls(c, wallNodes, wallLinks);
    c->numWall = (int) wallNodes.size();
    c->wallNodes = deviceAlloc<int>(wallNodes.size());
    c->wallLinks = deviceAlloc<int>(wallLinks.size());
    upload(c->wallNodes, wallNodes, c->stream);
    upload(c->wallLinks, wallLinks, c->stream);
    std::vector<BoundaryNode> list;
    findBoundaries(c, list);
    c->numBoundary = (int) list.size();
    c->boundary = deviceAlloc<BoundaryNode>(list.size());
    upload(c->boundary, list, c->stream);
    // Reductions.
    c->reducePartials = deviceAlloc<double>(5*REDUCE_BLOCKS);
    c->totals = deviceAlloc<double>(8);
    c->overflow = deviceAlloc<unsigned int>(1);
    HIP_CHECK(hipMemsetAsync(c->overflow, 0, sizeof(unsigned int), c->stream));
    c->ucm = deviceAlloc<double>(4);
    HIP_CHECK(hipMemsetAsync(c->ucm, 0, 4*sizeof(double), c->stream));
    // Particles.
    c->numParticles = numParts;
    c->numSystem = numSys;
    int w = stencilWidth(p.interpolation_stencil);
    c->numSlots = w*w*w;
    c->dense = denseCoupling((uint64_t) numParts, c->numSlots, n);
    auto blocks = [](int count) { return (count + BLOCK - 1)/BLOCK; };
    int coupleGroups = BLOCK/(p.interpolation_stencil == LBM_STENCIL_KEYS ? 64 : p.interpolation_stencil ==
            LBM_STENCIL_THREE_POINT ? 32 : p.interpolation_stencil == LBM_STENCIL_TRILINEAR ? 8 : 1);
    c->wallBounceOffset = 0;
    c->wallBoundaryOffset = c->wallBounceOffset + blocks(c->numWall);
    // The coupling and the reflection give momentum to walls only at solid nodes: without any, no partials.
    int wallParts = c->mask != nullptr ? numParts : 0;
    c->wallCoupleOffset = c->wallBoundaryOffset + blocks(c->numBoundary);
    c->wallReflectOffset = c->wallCoupleOffset + (wallParts + coupleGroups - 1)/coupleGroups;
    c->wallPartialCount = c->wallReflectOffset + blocks(wallParts);
    c->wallPartials = deviceAlloc<double>(3*(size_t) std::max(c->wallPartialCount, 1));
    HIP_CHECK(hipMemsetAsync(c->wallPartials, 0, 3*sizeof(double)*std::max(c->wallPartialCount, 1), c->stream));
    if (numParts > 0) {
        c->hostParticles.assign(parts, parts + numParts);
        std::vector<double> m(numParts);
        for (int i = 0; i < numParts; i++) {
            if (parts[i] < 0 || parts[i] >= numSys)
                throw Error(LBM_ERROR_INVALID, "LBMForce: a coupled particle index is out of range");
            m[i] = masses[i]/c->cellMass();
        }
        c->particles = deviceAlloc<int>(numParts);
        c->mass = deviceAlloc<double>(numParts);
        upload(c->particles, c->hostParticles, c->stream);
        upload(c->mass, m, c->stream);
        c->particleForce = deviceAlloc<double>(3*(size_t) numParts);
        HIP_CHECK(hipMemsetAsync(c->particleForce, 0, 3*sizeof(double)*numParts, c->stream));
        c->hostForce.assign(3*(size_t) numParts, 0.0);
        c->reactionFixed = deviceAlloc<unsigned long long>(3*n);
        HIP_CHECK(hipMemsetAsync(c->reactionFixed, 0, 3*n*sizeof(unsigned long long), c->stream));
        c->touched = deviceAlloc<uint32_t>((n + 31)/32);
        HIP_CHECK(hipMemsetAsync(c->touched, 0, ((n + 31)/32)*sizeof(uint32_t), c->stream));
        c->stagePos = deviceAlloc<double>(3*(size_t) std::max(numSys, 1));
        c->stageVel = deviceAlloc<double>(3*(size_t) std::max(numSys, 1));
        if (p.drag_scheme == LBM_DRAG_CENTERED)
            c->stageForce = deviceAlloc<double>(3*(size_t) std::max(numSys, 1));
        try {
            if (p.drag_scheme == LBM_DRAG_CENTERED)
                c->centred = centredCreate(numParts, p.interpolation_stencil, (uint32_t) n);
            else if (!c->dense)
                // Sparse particles gain DRAM locality from the node order; dense ones would contend for the same
                // accumulators when neighbours are processed together (measured: 100,000 chain beads with Keys on
                // 128^3 couple 6% slower in node order), so they keep the index order.
                c->order = coupleOrderCreate(numParts, (uint32_t) n);
        }
        catch (const std::runtime_error& e) {
            throw Error(LBM_ERROR_RUNTIME, e.what());
        }
        // The one-writer spread (spreadTiles) for the explicit drag on lattices of whole tiles with periodic faces.
        // LBM_AMD_SPREAD=atomic or =tiles forces the choice where it fits; by default the tiles serve dense particles
        // (the stencils cover more than a quarter of the lattice), whose concurrent atomics contend.
        const char* spreadEnv = getenv("LBM_AMD_SPREAD");
        if (const char* xcdEnv = getenv("LBM_AMD_COUPLE_XCD"))
            c->coupleXcd = xcdEnv[0] != '0';
        bool anyOpen = c->openAxis(0) || c->openAxis(1) || c->openAxis(2);
        bool fit = p.drag_scheme == LBM_DRAG_EXPLICIT && spreadTilesFit(p.nx, p.ny, p.nz, anyOpen);
        bool want = spreadEnv ? std::string(spreadEnv) == "tiles" : c->dense;
        // The moment field: by default for dense particles with the explicit drag; LBM_AMD_FIELD=0 or 1 forces it.
        const char* fieldEnv = getenv("LBM_AMD_FIELD");
        bool field = fieldEnv ? fieldEnv[0] == '1' : c->dense;
        if (p.drag_scheme == LBM_DRAG_EXPLICIT && field)
            c->momentField = deviceAlloc<double4>(n);
        // The stencils formed once per particle for the groups of the explicit stencil coupling, by default for dense
        // particles (100,000 Keys beads on 128^3: the coupling kernel 148 -> 110 us, the setup 11 us; for 20,000 sparse
        // beads on 256^3 the extra launch costs what it saves); LBM_AMD_STENCIL_SETUP=0 or 1 forces the choice.
        const char* setupEnv = getenv("LBM_AMD_STENCIL_SETUP");
        bool setup = setupEnv ? setupEnv[0] == '1' : c->dense;
        if (p.drag_scheme == LBM_DRAG_EXPLICIT && p.interpolation_stencil != LBM_STENCIL_NEAREST && setup)
            c->axisStencils = deviceAlloc<AxisStencil>((size_t) numParts);
        if (fit && want) {
            SpreadBins& b = c->bins;
            b.ntx = (uint32_t) (p.nx/SPREAD_TX);
            b.nty = (uint32_t) (p.ny/SPREAD_TY);
            b.ntz = (uint32_t) (p.nz/SPREAD_TZ);
            size_t tiles = (size_t) b.ntx*b.nty*b.ntz;
            // Room for four times the mean number of beads per tile and 16 more; a bead beyond it spreads itself with
            // the atomics.
            double mean = (double) numParts/(double) tiles;
            b.cap = std::min(SPREAD_MAX_CAP, ((int) std::ceil(4.0*mean) + 16 + 15)/16*16);
            c->binCounts[0] = deviceAlloc<int>(tiles);
            c->binCounts[1] = deviceAlloc<int>(tiles);
            HIP_CHECK(hipMemsetAsync(c->binCounts[0], 0, tiles*sizeof(int), c->stream));
            HIP_CHECK(hipMemsetAsync(c->binCounts[1], 0, tiles*sizeof(int), c->stream));
            b.beads = deviceAlloc<int>(tiles*(size_t) b.cap);
            b.anchor = deviceAlloc<uint16_t>(tiles*(size_t) b.cap);
            b.force = deviceAlloc<double>(3*tiles*(size_t) b.cap);
            c->binSpilled = deviceAlloc<int>(2);
            HIP_CHECK(hipMemsetAsync(c->binSpilled, 0, 2*sizeof(int), c->stream));
            int lanes = coupleLanes(p.interpolation_stencil);
            b.slotWeight = deviceAlloc<double>((size_t) numParts*lanes);
            c->binned = true;
        }
    }
    // The fluid at equilibrium with the initial velocity; solid nodes hold no fluid.
    launchInitEquilibrium(p.storage, c->popA, c->mask, c->g, p.initial_velocity[0], p.initial_velocity[1],
            p.initial_velocity[2], c->stream);
    HIP_CHECK(hipMemcpyAsync(c->popB, c->popA, NQ*n*c->elem, hipMemcpyDeviceToDevice, c->stream));
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipStreamSynchronize(c->stream));
}

std::vector<double> getState(lbm_context* c) {
    size_t count = NQ*(size_t) c->g.n;
    double* tmp = deviceAlloc<double>(count);
    std::vector<double> out(count);
    try {
        launchGather(c->p.storage, c->popA, tmp, c->g, c->stream);
        HIP_CHECK(hipMemcpyAsync(out.data(), tmp, count*sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIP_CHECK(hipStreamSynchronize(c->stream));
    }
    catch (...) {
        (void) hipFree(tmp);
        throw;
    }
    HIP_CHECK(hipFree(tmp));
    return out;
}

void setState(lbm_context* c, const double* state) {
    size_t count = NQ*(size_t) c->g.n;
    double* tmp = deviceAlloc<double>(count);
    try {
        HIP_CHECK(hipMemcpyAsync(tmp, state, count*sizeof(double), hipMemcpyHostToDevice, c->stream));
        launchScatter(c->p.storage, c->popA, tmp, c->g, c->stream);
        HIP_CHECK(hipStreamSynchronize(c->stream));
    }
    catch (...) {
        (void) hipFree(tmp);
        throw;
    }
    HIP_CHECK(hipFree(tmp));
    c->stepForcesCurrent = false;
}

void wallMomentum(lbm_context* c, double out[3]) {
    out[0] = out[1] = out[2] = 0.0;
    if (c->wallOverride) {
        for (int k = 0; k < 3; k++)
            out[k] = c->wallOverrideValue[k];
        return;
    }
    if (c->wallPartialCount == 0)
        return;
    std::vector<double> partials(3*(size_t) c->wallPartialCount);
    HIP_CHECK(hipMemcpyAsync(partials.data(), c->wallPartials, partials.size()*sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
    double fluid[3] = {0, 0, 0}, reflect[3] = {0, 0, 0};
    for (int b = 0; b < c->wallReflectOffset; b++)
        for (int k = 0; k < 3; k++)
            fluid[k] += partials[3*(size_t) b + k];
    for (int b = c->wallReflectOffset; b < c->wallPartialCount; b++)
        for (int k = 0; k < 3; k++)
            reflect[k] += partials[3*(size_t) b + k];
    double scale = c->cellMass()*c->p.dx/c->p.dt;
    for (int k = 0; k < 3; k++) {
        double lattice = fluid[k] + (c->wallStepDone && c->numWall > 0 ? c->staticWall[k] : 0.0);
        out[k] = lattice*scale + reflect[k];
    }
}

int fail(const Error& e) {
    g_lastError = e.what();
    return e.status;
}

}  // namespace

#define LBM_TRY(...) \
    try { __VA_ARGS__; g_lastError.clear(); return LBM_OK; } \
    catch (const Error& e) { return fail(e); } \
    catch (const std::bad_alloc&) { g_lastError = "out of host memory"; return LBM_ERROR_RUNTIME; } \
    catch (const std::exception& e) { g_lastError = e.what(); return LBM_ERROR_RUNTIME; }

#define LBM_CHECK_CTX(ctx) if (!(ctx)) { g_lastError = "null context"; return LBM_ERROR_INVALID; }

extern "C" {

const char* lbm_last_error(void) {
    return g_lastError.c_str();
}

int lbm_create(const lbm_params* params, const int* solid, int numSolid, const int* parts, const double* masses,
        int numParts, int numSys, lbm_context** out) {
    if (!params || !out) {
        g_lastError = "lbm_create: null argument";
        return LBM_ERROR_INVALID;
    }
    lbm_context* c = new lbm_context();
    try {
        create(params, solid, numSolid, parts, masses, numParts, numSys, c);
    }
    catch (const Error& e) {
        destroy(c);
        return fail(e);
    }
    catch (const std::exception& e) {
        destroy(c);
        g_lastError = e.what();
        return LBM_ERROR_RUNTIME;
    }
    *out = c;
    g_lastError.clear();
    return LBM_OK;
}

int lbm_destroy(lbm_context* c) {
    if (c)
        destroy(c);
    return LBM_OK;
}

int lbm_begin_step_device(lbm_context* c, int64_t step, const double* positions, double* velocities) {
    LBM_CHECK_CTX(c);
    LBM_TRY(beginStepDevice(c, step, positions, velocities))
}

int lbm_execute_device(lbm_context* c, int64_t step, int includeForces, const double* positions, const double* velocities,
        double* forces) {
    LBM_CHECK_CTX(c);
    LBM_TRY({
        struct TargetScope {
            lbm_context* c;
            ~TargetScope() { c->forceTarget = nullptr; }
        } scope{c};
        c->forceTarget = includeForces ? forces : nullptr;
        c->forcesAdded = false;
        executeDevice(c, step, includeForces != 0, positions, velocities, forces);
        if (includeForces && forces && !c->forcesAdded)
            launchAddForces(c->particleForce, c->particles, c->numParticles, forces, c->stream);
        HIP_CHECK(hipGetLastError());
    })
}

int lbm_begin_step(lbm_context* c, int64_t step, const double* positions, double* velocities) {
    LBM_CHECK_CTX(c);
    LBM_TRY({
        if (c->mask != nullptr && c->numParticles > 0) {
            size_t bytes = 3*sizeof(double)*(size_t) c->numSystem;
            HIP_CHECK(hipMemcpyAsync(c->stagePos, positions, bytes, hipMemcpyHostToDevice, c->stream));
            HIP_CHECK(hipMemcpyAsync(c->stageVel, velocities, bytes, hipMemcpyHostToDevice, c->stream));
            beginStepDevice(c, step, c->stagePos, c->stageVel);
            HIP_CHECK(hipMemcpyAsync(velocities, c->stageVel, bytes, hipMemcpyDeviceToHost, c->stream));
            HIP_CHECK(hipStreamSynchronize(c->stream));
        }
        else
            beginStepDevice(c, step, nullptr, nullptr);
    })
}

int lbm_execute(lbm_context* c, int64_t step, int includeForces, const double* positions, const double* velocities,
        double* forces) {
    LBM_CHECK_CTX(c);
    LBM_TRY({
        if (c->numParticles > 0) {
            size_t bytes = 3*sizeof(double)*(size_t) c->numSystem;
            HIP_CHECK(hipMemcpyAsync(c->stagePos, positions, bytes, hipMemcpyHostToDevice, c->stream));
            HIP_CHECK(hipMemcpyAsync(c->stageVel, velocities, bytes, hipMemcpyHostToDevice, c->stream));
            if (c->stageForce != nullptr && forces)
                HIP_CHECK(hipMemcpyAsync(c->stageForce, forces, bytes, hipMemcpyHostToDevice, c->stream));
        }
        executeDevice(c, step, includeForces != 0, c->stagePos, c->stageVel, forces ? c->stageForce : nullptr);
        if (includeForces && c->numParticles > 0 && forces) {
            HIP_CHECK(hipMemcpyAsync(c->hostForce.data(), c->particleForce, c->hostForce.size()*sizeof(double),
                    hipMemcpyDeviceToHost, c->stream));
            HIP_CHECK(hipStreamSynchronize(c->stream));
            for (int i = 0; i < c->numParticles; i++)
                for (int k = 0; k < 3; k++)
                    forces[3*(size_t) c->hostParticles[i] + k] += c->hostForce[3*(size_t) i + k];
        }
        else
            HIP_CHECK(hipStreamSynchronize(c->stream));
    })
}

int lbm_copy_parameters(lbm_context* c, const lbm_params* params) {
    LBM_CHECK_CTX(c);
    LBM_TRY({
        if (!params)
            throw Error(LBM_ERROR_INVALID, "lbm_copy_parameters: null argument");
        // The context's device layout, boundary lists, coupling buffers and random streams follow from these
        // settings at creation; LBMForceImpl::updateParametersInContext refuses to change them
        // (LBMForceImpl.cpp:388-414), and so does this entry point, which a caller may use without that check.
        const lbm_params& o = c->p;
        const lbm_params& u = *params;
        auto refuse = [](const char* what) {
            throw Error(LBM_ERROR_INVALID, std::string("updateParametersInContext: ") + what);
        };
        if (u.nx != o.nx || u.ny != o.ny || u.nz != o.nz)
            refuse("the grid size cannot be changed");
        if (u.density != o.density || u.tau != o.tau)
            refuse("the fluid density and viscosity cannot be changed");
        if (u.wall_scheme != o.wall_scheme)
            refuse("the wall scheme cannot be changed");
        if (u.density_halo_exchange != o.density_halo_exchange || u.velocity_halo_exchange != o.velocity_halo_exchange)
            refuse("the exchange of the halo cannot be switched on or off");
        for (int f = 0; f < 6; f++)
            if (u.face_boundary[f] != o.face_boundary[f])
                refuse("the boundary types of the faces cannot be changed");
        if (u.drag_scheme != o.drag_scheme)
            refuse("the drag scheme cannot be changed");
        if (u.interpolation_stencil != o.interpolation_stencil)
            refuse("the interpolation stencil cannot be changed");
        if ((u.fluid_fluctuations != 0) != (o.fluid_fluctuations != 0))
            refuse("the fluid fluctuations cannot be switched on or off");
        int storage = c->p.storage;
        c->p = *params;
        c->p.storage = storage;
    })
}

int lbm_get_fluid_fields(lbm_context* c, double* density, double* velocity, int halo) {
    LBM_CHECK_CTX(c);
    LBM_TRY({
        FieldsLayout layout;
        layout.pad = halo ? 1 : 0;
        layout.densityHalo = c->p.density_halo_exchange;
        layout.velocityHalo = c->p.velocity_halo_exchange;
        size_t n = fieldsCount(c->g, layout);
        if (!density && !velocity) {
            g_lastError.clear();
            return LBM_OK;
        }
        // One staging buffer for both fields: n densities, then 3n velocities.
        double* stage = deviceAlloc<double>(4*n);
        double* dRho = density ? stage : nullptr;
        double* dU = velocity ? stage + n : nullptr;
        try {
            launchFields(c->p.storage, c->popA, c->mask, c->g, layout, c->p.density, c->p.dx/c->p.dt,
                    c->p.body_acceleration[0], c->p.body_acceleration[1], c->p.body_acceleration[2], dRho, dU,
                    c->stream);
            HIP_CHECK(hipGetLastError());
            if (density)
                HIP_CHECK(hipMemcpyAsync(density, dRho, n*sizeof(double), hipMemcpyDeviceToHost, c->stream));
            if (velocity)
                HIP_CHECK(hipMemcpyAsync(velocity, dU, 3*n*sizeof(double), hipMemcpyDeviceToHost, c->stream));
            HIP_CHECK(hipStreamSynchronize(c->stream));
        }
        catch (...) {
            (void) hipStreamSynchronize(c->stream);
            (void) hipFree(stage);
            throw;
        }
        HIP_CHECK(hipFree(stage));
    })
}

int lbm_get_particle_nodes(lbm_context* c, const double* positions, int* nodes) {
    LBM_CHECK_CTX(c);
    LBM_TRY({
        if (c->numParticles == 0)
            return LBM_OK;
        size_t bytes = 3*sizeof(double)*(size_t) c->numSystem;
        HIP_CHECK(hipMemcpyAsync(c->stagePos, positions, bytes, hipMemcpyHostToDevice, c->stream));
        int* dNodes = deviceAlloc<int>(c->numParticles);
        ParticleArgs a = particleArgs(c, c->stepIndex, false);
        launchParticleNodes(c->stagePos, c->particles, a, dNodes, c->stream);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess)
            e = hipMemcpyAsync(nodes, dNodes, sizeof(int)*(size_t) c->numParticles, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess)
            e = hipStreamSynchronize(c->stream);
        (void) hipStreamSynchronize(c->stream);
        (void) hipFree(dNodes);
        HIP_CHECK(e);
    })
}

int lbm_get_fluid_mach_number(lbm_context* c, double* mach) {
    LBM_CHECK_CTX(c);
    LBM_TRY({
        reduceMoments(c);
        double totals[5];
        HIP_CHECK(hipMemcpyAsync(totals, c->totals, sizeof(totals), hipMemcpyDeviceToHost, c->stream));
        HIP_CHECK(hipStreamSynchronize(c->stream));
        *mach = std::sqrt(3.0*totals[4]);
    })
}

int lbm_get_fluid_totals(lbm_context* c, double* mass, double momentum[3]) {
    LBM_CHECK_CTX(c);
    LBM_TRY({
        reduceMoments(c);
        double totals[5];
        HIP_CHECK(hipMemcpyAsync(totals, c->totals, sizeof(totals), hipMemcpyDeviceToHost, c->stream));
        HIP_CHECK(hipStreamSynchronize(c->stream));
        *mass = c->numFluid + totals[0];
        for (int k = 0; k < 3; k++)
            momentum[k] = totals[1 + k];
    })
}

int lbm_get_wall_force(lbm_context* c, double force[3]) {
    LBM_CHECK_CTX(c);
    LBM_TRY({
        double m[3];
        wallMomentum(c, m);
        for (int k = 0; k < 3; k++)
            force[k] = m[k]/c->p.dt;
    })
}

int lbm_get_fluid_state(lbm_context* c, double* state, size_t size) {
    LBM_CHECK_CTX(c);
    LBM_TRY({
        if (size != NQ*(size_t) c->g.n)
            throw Error(LBM_ERROR_INVALID, "LBMForce: getFluidState() was called with a state of the wrong size");
        std::vector<double> s = getState(c);
        std::memcpy(state, s.data(), size*sizeof(double));
    })
}

int lbm_set_fluid_state(lbm_context* c, const double* state, size_t size) {
    LBM_CHECK_CTX(c);
    LBM_TRY({
        if (size != NQ*(size_t) c->g.n)
            throw Error(LBM_ERROR_INVALID, "LBMForce: setFluidState() was called with a state of the wrong size");
        setState(c, state);
    })
}

int lbm_checkpoint_size(lbm_context* c, size_t* size) {
    LBM_CHECK_CTX(c);
    LBM_TRY(*size = sizeof(CheckpointHeader) + NQ*(size_t) c->g.n*sizeof(double) + 3*sizeof(double))
}

int lbm_create_checkpoint(lbm_context* c, void* buffer, size_t size) {
    LBM_CHECK_CTX(c);
    LBM_TRY({
        size_t stateBytes = NQ*(size_t) c->g.n*sizeof(double);
        if (size < sizeof(CheckpointHeader) + stateBytes + 3*sizeof(double))
            throw Error(LBM_ERROR_INVALID, "lbm_create_checkpoint: buffer too small");
        CheckpointHeader h;
        std::memcpy(h.magic, "LBMAMDC1", 8);
        h.version = 2;
        h.seed = c->seed;
        h.nx = c->p.nx;
        h.ny = c->p.ny;
        h.nz = c->p.nz;
        h.numParticles = c->numParticles;
        h.dragScheme = c->p.drag_scheme;
        std::vector<double> s = getState(c);
        double m[3];
        wallMomentum(c, m);
        char* out = (char*) buffer;
        std::memcpy(out, &h, sizeof(h));
        std::memcpy(out + sizeof(h), s.data(), stateBytes);
        std::memcpy(out + sizeof(h) + stateBytes, m, 3*sizeof(double));
    })
}

int lbm_load_checkpoint(lbm_context* c, const void* buffer, size_t size) {
    LBM_CHECK_CTX(c);
    LBM_TRY({
        size_t stateBytes = NQ*(size_t) c->g.n*sizeof(double);
        CheckpointHeader h;
        if (size < CHECKPOINT_V1_HEADER)
            throw Error(LBM_ERROR_INVALID, "LBMForce: the checkpoint is truncated");
        std::memcpy(&h, buffer, CHECKPOINT_V1_HEADER);
        if (std::memcmp(h.magic, "LBMAMDC1", 8) != 0 || (h.version != 1 && h.version != 2))
            throw Error(LBM_ERROR_INVALID, "LBMForce: the data is not a checkpoint of this LBMForce platform");
        size_t headerBytes = h.version == 1 ? CHECKPOINT_V1_HEADER : sizeof(h);
        if (size < headerBytes)
            throw Error(LBM_ERROR_INVALID, "LBMForce: the checkpoint is truncated");
        std::memcpy(&h, buffer, headerBytes);
        if (h.nx != c->p.nx || h.ny != c->p.ny || h.nz != c->p.nz || h.numParticles != c->numParticles)
            throw Error(LBM_ERROR_INVALID, "LBMForce: the checkpoint was written for another lattice or System");
        if (h.dragScheme != c->p.drag_scheme)
            throw Error(LBM_ERROR_INVALID, "LBMForce: the checkpoint was written with a different drag scheme");
        if (size < headerBytes + stateBytes + 3*sizeof(double))
            throw Error(LBM_ERROR_INVALID, "LBMForce: the checkpoint is truncated");
        const char* in = (const char*) buffer;
        std::vector<double> state(NQ*(size_t) c->g.n);
        std::memcpy(state.data(), in + headerBytes, stateBytes);
        setState(c, state.data());
        std::memcpy(c->wallOverrideValue, in + headerBytes + stateBytes, 3*sizeof(double));
        c->wallOverride = true;
        if (h.version >= 2)
            c->seed = h.seed;
    })
}

int lbm_synchronize(lbm_context* c) {
    LBM_CHECK_CTX(c);
    LBM_TRY(HIP_CHECK(hipStreamSynchronize(c->stream)))
}

int lbm_profile_fused(lbm_context* c, int enable, double* totalMs, int* count) {
    LBM_CHECK_CTX(c);
    LBM_TRY({
        HIP_CHECK(hipStreamSynchronize(c->stream));
        double total = 0.0;
        for (auto& e : c->events) {
            float ms = 0.0f;
            HIP_CHECK(hipEventElapsedTime(&ms, e.first, e.second));
            total += ms;
        }
        if (totalMs)
            *totalMs = total;
        if (count)
            *count = (int) c->events.size();
        clearEvents(c);
        c->profile = enable != 0;
    })
}

int lbm_profile_fused_launches(lbm_context* c, float* ms, int capacity, int* count) {
    LBM_CHECK_CTX(c);
    LBM_TRY({
        HIP_CHECK(hipStreamSynchronize(c->stream));
        int n = (int) c->events.size();
        for (int i = 0; i < n && i < capacity; i++)
            HIP_CHECK(hipEventElapsedTime(&ms[i], c->events[i].first, c->events[i].second));
        if (count)
            *count = n;
    })
}

int lbm_closure_pairs(lbm_context* c, const double* positions, const double* velocities, float* features,
        float* targets) {
    LBM_CHECK_CTX(c);
    LBM_TRY({
        if (c->p.interpolation_stencil != LBM_STENCIL_KEYS)
            throw Error(LBM_ERROR_INVALID, "lbm_closure_pairs: the context must use the Keys stencil");
        int n = c->numParticles;
        if (n == 0)
            return LBM_OK;
        size_t bytes = 3*sizeof(double)*(size_t) c->numSystem;
        HIP_CHECK(hipMemcpyAsync(c->stagePos, positions, bytes, hipMemcpyHostToDevice, c->stream));
        HIP_CHECK(hipMemcpyAsync(c->stageVel, velocities, bytes, hipMemcpyHostToDevice, c->stream));
        // The coupling of the next step would see this state: the same removal shift.
        ParticleArgs a = particleArgs(c, c->stepIndex, false);
        if (a.removal)
            reduceMoments(c);
        ClosureNorm norm = {1.0f, 1.0f, 1.0f, 3.0e38f, 3.0e38f};
        _Float16* X = deviceAlloc<_Float16>((size_t) n*MLP_IN_PAD);
        float* raw = deviceAlloc<float>((size_t) n*MLP_IN);
        uint32_t* node = deviceAlloc<uint32_t>(n);
        double* u = deviceAlloc<double>(3*(size_t) n);
        uint8_t* ood = deviceAlloc<uint8_t>(n);
        uint8_t* frame = deviceAlloc<uint8_t>(n);
        float* t = deviceAlloc<float>(3*(size_t) n);
        launchClosureFeatures(c->p.storage, c->popA, c->mask, c->stagePos, c->stageVel, c->particles, c->ucm, a, norm,
                nullptr, X, raw, node, u, ood, frame, c->stream);
        launchClosureTargets(c->p.storage, c->popA, c->mask, c->stagePos, c->particles, c->ucm, a, u, frame, t,
                c->stream);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess)
            e = hipMemcpyAsync(features, raw, (size_t) n*MLP_IN*sizeof(float), hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(targets, t, 3*(size_t) n*sizeof(float), hipMemcpyDeviceToHost, c->stream);
        (void) hipStreamSynchronize(c->stream);
        void* scratch[] = {X, raw, node, u, ood, frame, t};
        for (void* b : scratch)
            (void) hipFree(b);
        HIP_CHECK(e);
    })
}

int lbm_set_closure(lbm_context* c, lbm_mlp* mlp) {
    LBM_CHECK_CTX(c);
    LBM_TRY({
        if (mlp != nullptr) {
            if (c->p.interpolation_stencil != LBM_STENCIL_KEYS)
                throw Error(LBM_ERROR_INVALID, "lbm_set_closure: the context must use the Keys stencil (the fallback)");
            if (c->p.drag_scheme != LBM_DRAG_EXPLICIT)
                throw Error(LBM_ERROR_UNSUPPORTED, "lbm_set_closure: the closure corrects the explicit drag only");
            size_t n = (size_t) std::max(c->numParticles, 1);
            if (c->closureX == nullptr) {
                c->closureX = deviceAlloc<_Float16>(n*MLP_IN_PAD);
                c->closureNode = deviceAlloc<uint32_t>(n);
                c->closureU = deviceAlloc<double>(3*n);
                c->closureDu = deviceAlloc<float>(3*n);
                c->closureOod = deviceAlloc<uint8_t>(n);
                c->closureFrame = deviceAlloc<uint8_t>(n);
                const size_t counters = 2*(size_t) closureCoupleBlocks(c->numParticles);
                c->closureCounters = deviceAlloc<unsigned long long>(counters);
                HIP_CHECK(hipMemsetAsync(c->closureCounters, 0, counters*sizeof(unsigned long long), c->stream));
            }
            // Dense particles (more neighbourhood nodes than lattice nodes): one velocity pass over the lattice.
            if (c->closureField == nullptr && 27*(int64_t) c->numParticles >= (int64_t) c->g.n)
                c->closureField = deviceAlloc<double>(3*(size_t) c->g.n);
        }
        HIP_CHECK(hipStreamSynchronize(c->stream));
        c->closure = mlp;
        c->stepForcesCurrent = false;
    })
}

int lbm_closure_stats(lbm_context* c, int64_t* network, int64_t* fallback) {
    LBM_CHECK_CTX(c);
    LBM_TRY({
        unsigned long long counts[2] = {0, 0};
        if (c->closureCounters != nullptr) {
            // Two counts per work-group of the closure coupling, summed here.
            std::vector<unsigned long long> blocks(2*(size_t) closureCoupleBlocks(c->numParticles));
            const size_t bytes = blocks.size()*sizeof(unsigned long long);
            HIP_CHECK(hipMemcpyAsync(blocks.data(), c->closureCounters, bytes, hipMemcpyDeviceToHost, c->stream));
            HIP_CHECK(hipMemsetAsync(c->closureCounters, 0, bytes, c->stream));
            HIP_CHECK(hipStreamSynchronize(c->stream));
            for (size_t b = 0; b < blocks.size(); b += 2) {
                counts[0] += blocks[b];
                counts[1] += blocks[b + 1];
            }
        }
        if (network)
            *network = (int64_t) counts[0];
        if (fallback)
            *fallback = (int64_t) counts[1];
    })
}

/* Internal: the device pointers and the stream of a context, for the harness (lbm_harness.hip). */
hipStream_t lbm_internal_stream(lbm_context* c) {
    return c->stream;
}

}  // extern "C"
