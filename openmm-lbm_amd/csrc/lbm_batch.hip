/*
 * Batched replicas (include/lbm_amd_batch.h): R independent periodic lattices advanced together.
 *
 * Per step, three launches cover every replica: coupleBatch (all particles), fusedBatch (all nodes, one work-group
 * per tile of one replica, looked up in tileReplica) and verletBatch (+ the step counter).  The step index lives in
 * device memory and the populations alternate between the two arrays by its parity, so the launches of a step are the
 * same at every step: the batch captures STEPS_PER_GRAPH steps into a hipGraph once and replays it.
 *
 * lbm_batch_autotune times the work-group size of the fused pass (256, 512, 1024 threads) on the batch itself and
 * keeps the fastest, rebuilding the tile table and the graph.
 */
#include "lbm_internal.h"
#include "../../include/lbm_amd_batch.h"
#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <sstream>
#include <stdexcept>

namespace lbm {
void setLastError(const std::string& message);
}

using namespace lbm;

namespace {

constexpr int STEPS_PER_GRAPH = 10;

struct BatchError : public std::runtime_error {
    int status;
    BatchError(int s, const std::string& m) : std::runtime_error(m), status(s) {}
};

void hipCheck(hipError_t e, const char* what) {
    if (e != hipSuccess)
        throw BatchError(LBM_ERROR_RUNTIME, std::string(what) + ": " + hipGetErrorString(e));
}

template <typename T>
T* alloc(size_t count, const char* what) {
    void* p = nullptr;
    hipCheck(hipMalloc(&p, std::max<size_t>(count, 1)*sizeof(T)), what);
    return (T*) p;
}

FastDiv fastDivOf(uint32_t d) {
    FastDiv f;
    f.d = d;
    f.l = 0;
    while ((1ull << f.l) < d)
        f.l++;
    f.m = d <= 1 ? 0 : (uint32_t) (((1ull << 32)*((1ull << f.l) - d))/d + 1);
    return f;
}

}  // namespace

struct lbm_batch {
    int numReplicas = 0;
    int stencil = 0;
    bool fluct = false;
    hipStream_t stream = nullptr;
    std::vector<lbm_params> params;
    std::vector<ReplicaDesc> desc;
    ReplicaDesc* dDesc = nullptr;
    uint16_t* tileReplica = nullptr;
    uint32_t numTiles = 0;
    int threads = FUSED_THREADS;
    int coupleThreads = BLOCK;              // work-group size of the coupling (particles per tile: coupleThreads/lanes)
    int xcdOrder = 3;                       // bit 0: coupling tiles in XCD-contiguous runs, bit 1: fused tiles in
                                            // runs of XCD_CHUNK per XCD
    int numParticles = 0;
    uint16_t* particleReplica = nullptr;
    uint16_t* particleTiles = nullptr;      // replica of each work-group of the coupling
    int numParticleTiles = 0;
    double* pos = nullptr;
    double* vel = nullptr;
    double* force = nullptr;
    double* mass = nullptr;        // lattice units (per replica cell mass)
    double* invMass = nullptr;     // 1/Da
    int64_t* stepCounter = nullptr;
    int64_t step = 0;
    int64_t totalNodes = 0;
    hipGraphExec_t graph = nullptr;
    std::vector<void*> buffers;
    bool profiling = false;                                // eager steps with events around the fused pass
    std::vector<std::pair<hipEvent_t, hipEvent_t>> events;
};

namespace {

/* The coupling's tile table for work-groups of `threads` threads: threads/lanes particles of one replica each. */
void buildParticleTiles(lbm_batch* b, int threads) {
    std::vector<uint16_t> tiles;
    int perTile = threads/coupleLanes(b->stencil);
    for (int r = 0; r < b->numReplicas; r++) {
        b->desc[r].particleTileBegin = (uint32_t) tiles.size();
        int count = (b->desc[r].part.numParticles + perTile - 1)/perTile;
        tiles.insert(tiles.end(), count, (uint16_t) r);
    }
    if (b->particleTiles)
        hipCheck(hipFree(b->particleTiles), "hipFree");
    b->particleTiles = nullptr;
    b->numParticleTiles = (int) tiles.size();
    if (!tiles.empty()) {
        b->particleTiles = alloc<uint16_t>(tiles.size(), "particle tiles");
        hipCheck(hipMemcpy(b->particleTiles, tiles.data(), tiles.size()*sizeof(uint16_t), hipMemcpyHostToDevice),
                "upload particle tiles");
    }
    b->coupleThreads = threads;
}

void buildTiles(lbm_batch* b, int threads) {
    std::vector<uint16_t> tiles;
    for (int r = 0; r < b->numReplicas; r++) {
        b->desc[r].tileBegin = (uint32_t) tiles.size();
        uint32_t n = b->desc[r].step.g.n;
        uint32_t count = (n + threads - 1)/threads;
        tiles.insert(tiles.end(), count, (uint16_t) r);
    }
    if (b->tileReplica)
        hipCheck(hipFree(b->tileReplica), "hipFree");
    b->tileReplica = alloc<uint16_t>(tiles.size(), "tile table");
    hipCheck(hipMemcpy(b->tileReplica, tiles.data(), tiles.size()*sizeof(uint16_t), hipMemcpyHostToDevice), "upload tiles");
    hipCheck(hipMemcpy(b->dDesc, b->desc.data(), b->desc.size()*sizeof(ReplicaDesc), hipMemcpyHostToDevice), "upload table");
    b->numTiles = (uint32_t) tiles.size();
    b->threads = threads;
}

void launchStep(lbm_batch* b, bool timed = false) {
    launchCoupleBatch(b->stencil, b->coupleThreads, b->dDesc, b->particleTiles, b->numParticleTiles, b->pos, b->vel,
            b->mass, b->force, b->stepCounter, (b->xcdOrder & 1) != 0, b->stream);
    std::pair<hipEvent_t, hipEvent_t> ev{nullptr, nullptr};
    if (timed) {
        hipCheck(hipEventCreate(&ev.first), "event");
        hipCheck(hipEventCreate(&ev.second), "event");
        b->events.push_back(ev);
        hipCheck(hipEventRecord(ev.first, b->stream), "event");
    }
    launchFusedBatch(b->threads, b->fluct, b->dDesc, b->tileReplica, b->numTiles, b->stepCounter,
            (b->xcdOrder & 2) != 0, b->stream);
    if (timed)
        hipCheck(hipEventRecord(ev.second, b->stream), "event");
    launchVerletBatch(b->dDesc, b->particleReplica, b->numParticles, b->pos, b->vel, b->force, b->invMass,
            b->stepCounter, b->stream);
}

void buildGraph(lbm_batch* b) {
    if (b->graph) {
        hipCheck(hipGraphExecDestroy(b->graph), "hipGraphExecDestroy");
        b->graph = nullptr;
    }
    hipGraph_t g;
    hipCheck(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal), "begin capture");
    for (int s = 0; s < STEPS_PER_GRAPH; s++)
        launchStep(b);
    hipCheck(hipStreamEndCapture(b->stream, &g), "end capture");
    hipCheck(hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0), "instantiate graph");
    hipCheck(hipGraphDestroy(g), "hipGraphDestroy");
}

void clearEvents(lbm_batch* b) {
    for (auto& e : b->events) {
        (void) hipEventDestroy(e.first);
        (void) hipEventDestroy(e.second);
    }
    b->events.clear();
}

void stepBatch(lbm_batch* b, int numSteps) {
    if (b->profiling) {
        for (int i = 0; i < numSteps; i++)
            launchStep(b, true);
        hipCheck(hipGetLastError(), "launch");
        b->step += numSteps;
        return;
    }
    int full = numSteps/STEPS_PER_GRAPH;
    for (int i = 0; i < full; i++)
        hipCheck(hipGraphLaunch(b->graph, b->stream), "graph launch");
    for (int i = 0; i < numSteps - full*STEPS_PER_GRAPH; i++)
        launchStep(b);
    hipCheck(hipGetLastError(), "launch");
    b->step += numSteps;
}

void destroyBatch(lbm_batch* b) {
    if (b->stream)
        (void) hipStreamSynchronize(b->stream);
    clearEvents(b);
    if (b->graph)
        (void) hipGraphExecDestroy(b->graph);
    for (void* p : b->buffers)
        (void) hipFree(p);
    void* others[] = {b->dDesc, b->tileReplica, b->particleReplica, b->particleTiles, b->pos, b->vel, b->force, b->mass,
                      b->invMass,
                      b->stepCounter};
    for (void* p : others)
        if (p)
            (void) hipFree(p);
    if (b->stream)
        (void) hipStreamDestroy(b->stream);
    delete b;
}

void createBatch(lbm_batch* b, int R, const lbm_params* params, const int* numParticles, const double* masses,
        const double* positions, const double* velocities) {
    if (R <= 0 || R > 65535)
        throw BatchError(LBM_ERROR_INVALID, "lbm_batch_create: the number of replicas must be in [1, 65535]");
    b->numReplicas = R;
    b->params.assign(params, params + R);
    b->stencil = params[0].interpolation_stencil;
    b->fluct = params[0].fluid_fluctuations && params[0].fluid_kT > 0;
    for (int r = 0; r < R; r++) {
        const lbm_params& p = params[r];
        if (p.interpolation_stencil != b->stencil)
            throw BatchError(LBM_ERROR_INVALID, "lbm_batch_create: every replica must use the same interpolation stencil");
        if ((p.fluid_fluctuations && p.fluid_kT > 0) != b->fluct)
            throw BatchError(LBM_ERROR_INVALID, "lbm_batch_create: every replica must have fluctuations on, or every one off");
        for (int f = 0; f < 6; f++)
            if (p.face_boundary[f] != LBM_BOUNDARY_PERIODIC)
                throw BatchError(LBM_ERROR_UNSUPPORTED, "lbm_batch_create: replicas must be periodic");
        if (p.drag_scheme != LBM_DRAG_EXPLICIT)
            throw BatchError(LBM_ERROR_UNSUPPORTED, "lbm_batch_create: only the explicit drag is supported");
        if (p.storage != LBM_STORAGE_FP32)
            throw BatchError(LBM_ERROR_UNSUPPORTED, "lbm_batch_create: replicas are stored in FP32");
        if (p.momentum_removal_frequency != 0)
            throw BatchError(LBM_ERROR_UNSUPPORTED, "lbm_batch_create: momentum removal is not supported in a batch");
        if (p.nx <= 0 || p.ny <= 0 || p.nz <= 0)
            throw BatchError(LBM_ERROR_INVALID, "lbm_batch_create: invalid grid size");
        if ((int64_t) NQ*p.nx*p.ny*p.nz*(int64_t) sizeof(float) > (1ll << 32))
            throw BatchError(LBM_ERROR_INVALID, "lbm_batch_create: a replica holds at most 56.5 M nodes (its "
                    "populations within 4 GiB)");
        if (numParticles[r] < 0)
            throw BatchError(LBM_ERROR_INVALID, "lbm_batch_create: negative particle count");
    }
    hipCheck(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking), "stream");
    if (const char* xcdEnv = getenv("LBM_AMD_BATCH_XCD"))
        b->xcdOrder = atoi(xcdEnv) & 3;
    b->desc.resize(R);
    int particleBegin = 0;
    std::vector<uint16_t> particleReplica;
    std::vector<double> latticeMass, invMass;
    for (int r = 0; r < R; r++) {
        const lbm_params& p = params[r];
        ReplicaDesc& d = b->desc[r];
        d = ReplicaDesc();
        Geometry& g = d.step.g;
        g.nx = p.nx;
        g.ny = p.ny;
        g.nz = p.nz;
        g.nxy = (uint32_t) (p.nx*p.ny);
        g.n = (uint32_t) ((int64_t) p.nx*p.ny*p.nz);
        g.divx = fastDivOf(p.nx);
        g.divxy = fastDivOf(g.nxy);
        g.rowShift = 0;
        g.qstride = g.nx;
        g.openX = g.openY = g.openZ = 0;
        double cellMass = p.density*p.dx*p.dx*p.dx;
        uint32_t seed = (uint32_t) p.random_number_seed;
        if (seed == 0)
            seed = (uint32_t) std::chrono::high_resolution_clock::now().time_since_epoch().count() + 2654435761u*(uint32_t) r;
        d.step.omega = p.omega;
        d.step.gx = p.body_acceleration[0];
        d.step.gy = p.body_acceleration[1];
        d.step.gz = p.body_acceleration[2];
        d.step.removal = 0;
        d.step.seed = seed;
        d.step.step = 0;
        d.step.mu = 3.0*p.fluid_kT*p.dt*p.dt/(cellMass*p.dx*p.dx);
        d.step.roundSeed = 0;
        int w = stencilWidth(p.interpolation_stencil);
        d.step.reactionPlanes = denseCoupling((uint64_t) numParticles[r], w*w*w, (uint64_t) g.n) ? 1 : 0;
        ParticleArgs& a = d.part;
        a.g = g;
        a.numParticles = numParticles[r];
        a.stencil = p.interpolation_stencil;
        a.dx = p.dx;
        a.gamma = p.friction*p.dt;
        a.kT = p.kT*p.dt*p.dt/(cellMass*p.dx*p.dx);
        a.velocityScale = p.dx/p.dt;
        a.forceScale = cellMass*p.dx/(p.dt*p.dt);
        a.cellMass = cellMass;
        a.removal = 0;
        a.isStep = 1;
        a.randomForce = (a.kT > 0 && a.gamma > 0) ? 1 : 0;
        a.seed = seed;
        a.step = 0;
        a.hasSolids = 0;
        a.reactionPlanes = d.step.reactionPlanes;
        d.dt = p.dt;
        d.particleBegin = particleBegin;
        size_t n = g.n;
        d.popA = alloc<float>(NQ*n, "populations");
        d.popB = alloc<float>(NQ*n, "populations");
        d.reactionFixed = alloc<unsigned long long>(3*n, "reactions");
        d.touched = alloc<uint32_t>((n + 31)/32, "touched mask");
        b->buffers.insert(b->buffers.end(), {d.popA, d.popB, d.reactionFixed, d.touched});
        hipCheck(hipMemsetAsync(d.reactionFixed, 0, 3*n*sizeof(unsigned long long), b->stream), "memset");
        hipCheck(hipMemsetAsync(d.touched, 0, ((n + 31)/32)*sizeof(uint32_t), b->stream), "memset");
        launchInitEquilibrium(LBM_STORAGE_FP32, d.popA, nullptr, g, p.initial_velocity[0], p.initial_velocity[1],
                p.initial_velocity[2], b->stream);
        hipCheck(hipMemcpyAsync(d.popB, d.popA, NQ*n*sizeof(float), hipMemcpyDeviceToDevice, b->stream), "copy");
        for (int i = 0; i < numParticles[r]; i++) {
            double m = masses[particleBegin + i];
            if (!(m > 0))
                throw BatchError(LBM_ERROR_INVALID, "LBMForce: coupled particles must have a positive mass");
            particleReplica.push_back((uint16_t) r);
            latticeMass.push_back(m/cellMass);
            invMass.push_back(1.0/m);
        }
        particleBegin += numParticles[r];
        b->totalNodes += (int64_t) n;
    }
    b->numParticles = particleBegin;
    size_t np = (size_t) b->numParticles;
    b->dDesc = alloc<ReplicaDesc>(R, "replica table");
    b->particleReplica = alloc<uint16_t>(np, "particle table");
    b->pos = alloc<double>(3*np, "positions");
    b->vel = alloc<double>(3*np, "velocities");
    b->force = alloc<double>(3*np, "forces");
    b->mass = alloc<double>(np, "masses");
    b->invMass = alloc<double>(np, "masses");
    b->stepCounter = alloc<int64_t>(1, "step counter");
    if (np > 0) {
        hipCheck(hipMemcpyAsync(b->particleReplica, particleReplica.data(), np*sizeof(uint16_t), hipMemcpyHostToDevice, b->stream), "upload");
        hipCheck(hipMemcpyAsync(b->pos, positions, 3*np*sizeof(double), hipMemcpyHostToDevice, b->stream), "upload");
        hipCheck(hipMemcpyAsync(b->vel, velocities, 3*np*sizeof(double), hipMemcpyHostToDevice, b->stream), "upload");
        hipCheck(hipMemcpyAsync(b->mass, latticeMass.data(), np*sizeof(double), hipMemcpyHostToDevice, b->stream), "upload");
        hipCheck(hipMemcpyAsync(b->invMass, invMass.data(), np*sizeof(double), hipMemcpyHostToDevice, b->stream), "upload");
    }
    hipCheck(hipMemsetAsync(b->stepCounter, 0, sizeof(int64_t), b->stream), "memset");
    hipCheck(hipStreamSynchronize(b->stream), "sync");
    buildParticleTiles(b, BLOCK);
    buildTiles(b, FUSED_THREADS);
    buildGraph(b);
}

int failBatch(const std::exception& e, int status) {
    setLastError(e.what());
    return status;
}

}  // namespace

#define BATCH_TRY(...) \
    try { __VA_ARGS__; setLastError(""); return LBM_OK; } \
    catch (const BatchError& e) { return failBatch(e, e.status); } \
    catch (const std::exception& e) { return failBatch(e, LBM_ERROR_RUNTIME); }

extern "C" {

int lbm_batch_create(int R, const lbm_params* params, const int* numParticles, const double* masses,
        const double* positions, const double* velocities, lbm_batch** out) {
    if (!params || !numParticles || !out) {
        setLastError("lbm_batch_create: null argument");
        return LBM_ERROR_INVALID;
    }
    lbm_batch* b = new lbm_batch();
    try {
        createBatch(b, R, params, numParticles, masses, positions, velocities);
    }
    catch (const BatchError& e) {
        destroyBatch(b);
        return failBatch(e, e.status);
    }
    catch (const std::exception& e) {
        destroyBatch(b);
        return failBatch(e, LBM_ERROR_RUNTIME);
    }
    *out = b;
    setLastError("");
    return LBM_OK;
}

int lbm_batch_autotune(lbm_batch* b, int stepsPerTrial, double* timesMs, int* bestFused, int* bestCouple) {
    BATCH_TRY({
        // Tile shapes: nodes per work-group of the lattice pass (256, 512, 1024 threads, one node each) times
        // particles per work-group of the coupling (64, 128, 256 threads: a particle per 1, 8, 27 or 64 lanes by the
        // stencil).  The candidates are timed in interleaved rounds and each keeps its best round, so that drift of
        // the clocks or of the memory system during the trials does not favour one of them.
        const int fused[3] = {256, 512, 1024};
        const int couple[3] = {64, 128, BLOCK};
        const int rounds = 3;
        double times[9];
        for (double& t : times)
            t = 1e300;
        hipEvent_t e0, e1;
        hipCheck(hipEventCreate(&e0), "event");
        hipCheck(hipEventCreate(&e1), "event");
        int trials = std::max(stepsPerTrial, STEPS_PER_GRAPH);
        for (int r = 0; r < rounds; r++)
            for (int c = 0; c < 9; c++) {
                buildParticleTiles(b, couple[c % 3]);
                buildTiles(b, fused[c/3]);
                buildGraph(b);
                stepBatch(b, STEPS_PER_GRAPH);             // warm-up
                hipCheck(hipEventRecord(e0, b->stream), "event");
                stepBatch(b, trials);
                hipCheck(hipEventRecord(e1, b->stream), "event");
                hipCheck(hipEventSynchronize(e1), "event");
                float ms = 0;
                hipCheck(hipEventElapsedTime(&ms, e0, e1), "event");
                times[c] = std::min(times[c], (double) ms/trials);
            }
        int best = 0;
        for (int c = 0; c < 9; c++) {
            if (timesMs)
                timesMs[c] = times[c];
            if (times[c] < times[best])
                best = c;
        }
        (void) hipEventDestroy(e0);
        (void) hipEventDestroy(e1);
        buildParticleTiles(b, couple[best % 3]);
        buildTiles(b, fused[best/3]);
        buildGraph(b);
        if (bestFused)
            *bestFused = fused[best/3];
        if (bestCouple)
            *bestCouple = couple[best % 3];
    })
}

int lbm_batch_profile_fused(lbm_batch* b, int enable, double* totalMs, int* count) {
    BATCH_TRY({
        if (enable) {
            clearEvents(b);
            b->profiling = true;
        }
        else {
            hipCheck(hipStreamSynchronize(b->stream), "sync");
            double total = 0;
            for (auto& e : b->events) {
                float ms = 0;
                hipCheck(hipEventElapsedTime(&ms, e.first, e.second), "event");
                total += ms;
            }
            if (totalMs)
                *totalMs = total;
            if (count)
                *count = (int) b->events.size();
            clearEvents(b);
            b->profiling = false;
        }
    })
}

int lbm_batch_step(lbm_batch* b, int numSteps) {
    BATCH_TRY(stepBatch(b, numSteps))
}

int lbm_batch_synchronize(lbm_batch* b) {
    BATCH_TRY(hipCheck(hipStreamSynchronize(b->stream), "sync"))
}

int lbm_batch_get_fluid_fields(lbm_batch* b, int r, double* density, double* velocity) {
    BATCH_TRY({
        if (r < 0 || r >= b->numReplicas)
            throw BatchError(LBM_ERROR_INVALID, "lbm_batch_get_fluid_fields: replica out of range");
        const ReplicaDesc& d = b->desc[r];
        const lbm_params& p = b->params[r];
        size_t n = d.step.g.n;
        double* dRho = alloc<double>(n, "fields");
        double* dU = alloc<double>(3*n, "fields");
        const float* cur = (b->step & 1) ? d.popB : d.popA;
        launchFields(LBM_STORAGE_FP32, cur, nullptr, d.step.g, FieldsLayout(), p.density, p.dx/p.dt, p.body_acceleration[0],
                p.body_acceleration[1], p.body_acceleration[2], dRho, dU, b->stream);
        hipError_t e1 = density ? hipMemcpyAsync(density, dRho, n*sizeof(double), hipMemcpyDeviceToHost, b->stream) : hipSuccess;
        hipError_t e2 = velocity ? hipMemcpyAsync(velocity, dU, 3*n*sizeof(double), hipMemcpyDeviceToHost, b->stream) : hipSuccess;
        hipError_t e3 = hipStreamSynchronize(b->stream);
        (void) hipFree(dRho);
        (void) hipFree(dU);
        hipCheck(e1, "download");
        hipCheck(e2, "download");
        hipCheck(e3, "sync");
    })
}

int lbm_batch_get_particles(lbm_batch* b, double* positions, double* velocities) {
    BATCH_TRY({
        size_t bytes = 3*sizeof(double)*(size_t) b->numParticles;
        hipCheck(hipStreamSynchronize(b->stream), "sync");
        if (positions && bytes)
            hipCheck(hipMemcpy(positions, b->pos, bytes, hipMemcpyDeviceToHost), "download");
        if (velocities && bytes)
            hipCheck(hipMemcpy(velocities, b->vel, bytes, hipMemcpyDeviceToHost), "download");
    })
}

int lbm_batch_info(lbm_batch* b, int64_t* totalNodes, int64_t* step) {
    BATCH_TRY({
        if (totalNodes)
            *totalNodes = b->totalNodes;
        if (step)
            *step = b->step;
    })
}

int lbm_batch_destroy(lbm_batch* b) {
    if (b)
        destroyBatch(b);
    return LBM_OK;
}

}  // extern "C"
