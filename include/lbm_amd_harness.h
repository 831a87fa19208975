/*
 * lbm_amd_harness.h -- device-resident particle host for benchmarks and long runs.
 *
 * No counterpart in the reference's plugin: these entry points stand in for what OpenMM's
 * Context and VerletIntegrator do around CalcLBMForceKernel (LBMForceImpl.cpp:366-380), kept on
 * the GPU so that a coupled step crosses no PCIe: per step lbm_begin_step_device(), the extra forces
 * of the benchmark configurations (wall potential, harmonic chain bonds), lbm_execute_device(),
 * and the leapfrog Verlet update.  They are exported by the same library as lbm_amd.h.
 */
#ifndef LBM_AMD_HARNESS_H_
#define LBM_AMD_HARNESS_H_

#include "lbm_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct lbm_md_system lbm_md_system;

const char* lbm_md_last_error(void);

/* n particles with masses (Da), positions (nm) and velocities (nm/ps), xyz per particle; dt in ps. */
int lbm_md_create(int n, const double* masses, const double* positions, const double* velocities, double dt,
        lbm_md_system** out);

/* Repulsive walls U = eps[(sigma/(z - zlo))^12 + (sigma/(zhi - z))^12] (kJ/mol, nm). */
int lbm_md_set_wall_potential(lbm_md_system* md, double eps, double sigma, double zlo, double zhi);

/* Harmonic bonds k/2 (r - r0)^2 between the pairs (2 ints per bond), minimum image in the box;
   at most two bonds per particle (chains). */
int lbm_md_set_bonds(lbm_md_system* md, const int* pairs, int num_bonds, double k, double r0, const double box[3]);

/* num_steps integration steps, the first with step count step0, on the stream of ctx (asynchronous). */
int lbm_md_step(lbm_md_system* md, lbm_context* ctx, int64_t step0, int num_steps);

/* Synchronize and copy the particles back (either pointer may be NULL). */
int lbm_md_get(lbm_md_system* md, lbm_context* ctx, double* positions, double* velocities);

int lbm_md_destroy(lbm_md_system* md);

#ifdef __cplusplus
}
#endif

#endif /* LBM_AMD_HARNESS_H_ */
