/*
 * lbm_amd_batch.h -- many independent replicas stepped together on one GPU.
 *
 * No counterpart in the reference's kernel interface: the reference runs the replicas of a
 * seed-averaged study as separate runs (examples/cocomo/kick.py:30,105 --repeats).  Here R periodic
 * lattices of different sizes, relaxation times, temperatures, seeds and particle counts advance in
 * one fused launch per step over a table of work-groups, with the particles coupled and integrated
 * (leapfrog Verlet, coupling force only) on the device; a step is captured once in a hipGraph and
 * replayed.  FP32 storage; periodic faces, no solid nodes, explicit drag.
 */
#ifndef LBM_AMD_BATCH_H_
#define LBM_AMD_BATCH_H_

#include "lbm_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct lbm_batch lbm_batch;

/*
 * params[r]: the lattice parameters of replica r (lbm_compute_lattice_parameters), all with the same
 * interpolation stencil and the same fluid_fluctuations setting; num_particles[r]: its coupled particles;
 * masses (Da), positions (nm), velocities (nm/ps): all particles, replica after replica (xyz per particle).
 * At most 65,535 replicas of at most 56.5 M nodes each (LBM_ERROR_INVALID beyond); periodic faces, no solid
 * nodes, the explicit drag, FP32 storage and no momentum removal (LBM_ERROR_UNSUPPORTED otherwise).
 * The environment variable LBM_AMD_BATCH_XCD (read here; default 3) sets which launches group their tiles by XCD, so
 * that neighbouring tiles share one of the 8 XCDs' L2s: bit 0 the coupling (each XCD takes one contiguous run of the
 * particle tiles), bit 1 the fused pass (each XCD takes runs of 64 consecutive node tiles in turn); 0 hands work-group
 * b its own tile.  It changes no result.
 */
int lbm_batch_create(int num_replicas, const lbm_params* params, const int* num_particles, const double* masses,
        const double* positions, const double* velocities, lbm_batch** out);

/* Time each tile shape, the nodes per work-group of the lattice pass (256, 512, 1024) times the threads per
   work-group of the coupling (64, 128, 256: particles per tile = threads/lanes, 1, 8, 27 or 64 lanes per particle by
   the stencil), in three interleaved rounds of steps_per_trial steps (the state advances), and keep the fastest.
   times_ms (9 entries, may be NULL) receives the best round's time per step, lattice-pass size major; *best_fused and
   *best_couple (may be NULL) the choice. */
int lbm_batch_autotune(lbm_batch* batch, int steps_per_trial, double* times_ms, int* best_fused, int* best_couple);

/* Advance every replica by num_steps steps (asynchronous; replays the captured graph). */
int lbm_batch_step(lbm_batch* batch, int num_steps);

int lbm_batch_synchronize(lbm_batch* batch);

/* As lbm_profile_fused: while enabled, steps are launched one by one (not from the graph) with events around each
   fused pass; disabling synchronizes and returns their total time and count. */
int lbm_batch_profile_fused(lbm_batch* batch, int enable, double* total_ms, int* count);

/* Density (Da/nm^3) and velocity (nm/ps, xyz) of replica r, as lbm_get_fluid_fields. */
int lbm_batch_get_fluid_fields(lbm_batch* batch, int replica, double* density, double* velocity);

/* The particles of all replicas (either pointer may be NULL). */
int lbm_batch_get_particles(lbm_batch* batch, double* positions, double* velocities);

/* Total nodes of all replicas, and the step count. */
int lbm_batch_info(lbm_batch* batch, int64_t* total_nodes, int64_t* step);

int lbm_batch_destroy(lbm_batch* batch);

#ifdef __cplusplus
}
#endif

#endif /* LBM_AMD_BATCH_H_ */
