/*
 * lbm_amd.h -- C-ABI of the MI355X (gfx950) LBMForce platform.
 *
 * The drop-in boundary is the reference's kernel interface CalcLBMForceKernel
 * (openmmapi/include/LBMKernels.h:97-189): one context per OpenMM Context, created
 * from the converted lattice parameters (LBMLatticeParameters, LBMKernels.h:30-92),
 * entered once per integration step through lbm_begin_step() and lbm_execute().
 * Every entry point returns LBM_OK (0) or a negative status; the message of the last
 * failure of the calling thread is lbm_last_error().  C++ exceptions never cross this
 * boundary.  Plain pointers and sizes only; callers own every host array, the context
 * owns every device allocation.  A context is not thread-safe; it runs its device work
 * in order on one HIP stream of its own.
 *
 * Units at this boundary are those of the reference's kernel interface: lattice
 * parameters in lattice units (computed by lbm_compute_lattice_parameters(), the
 * restatement of LBMForceImpl::computeLatticeParameters), particle positions in nm,
 * velocities in nm/ps, forces in kJ/mol/nm, masses in Da.
 *
 * The same functions are exported by the FP64 CPU oracle library used by the tests
 * (oracle/liblbm_oracle.so), so that every test can run on both.
 */
#ifndef LBM_AMD_H_
#define LBM_AMD_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. */
#define LBM_OK 0
#define LBM_ERROR_INVALID (-1)      /* invalid argument or setup (OpenMMException at setup)        */
#define LBM_ERROR_RUNTIME (-2)      /* HIP failure, allocation failure                           */
#define LBM_ERROR_MACH (-3)         /* Mach number above the limit (ReferenceLBMKernels.cpp:412-420) */
#define LBM_ERROR_UNSUPPORTED (-4)  /* a feature this build does not implement                    */

/* Enumerations, with the values of the reference's enums (LBMForce.h:46-158). */
#define LBM_COUPLING_EULER_MARUYAMA 0
#define LBM_COUPLING_NVE 1
#define LBM_DRAG_EXPLICIT 0
#define LBM_DRAG_CENTERED 1
#define LBM_STENCIL_NEAREST 0
#define LBM_STENCIL_TRILINEAR 1
#define LBM_STENCIL_THREE_POINT 2
#define LBM_STENCIL_KEYS 3
#define LBM_WALL_BOUNCE_BACK 0
#define LBM_WALL_REGULARIZED 1
#define LBM_FACE_XMIN 0
#define LBM_FACE_XMAX 1
#define LBM_FACE_YMIN 2
#define LBM_FACE_YMAX 3
#define LBM_FACE_ZMIN 4
#define LBM_FACE_ZMAX 5
#define LBM_BOUNDARY_PERIODIC 0
#define LBM_BOUNDARY_VELOCITY 1
#define LBM_BOUNDARY_DENSITY 2
#define LBM_BOUNDARY_DENSITY_VELOCITY 3

/* Storage of the populations on the device (no counterpart in the reference, whose
   GPU platforms store them in the OpenMM "mixed" type).  FP16/BF16 store the deviations
   f_q - w_q with stochastic rounding; collision arithmetic is FP32 for these and FP32
   storage, FP64 for FP64 storage. */
#define LBM_STORAGE_FP32 0
#define LBM_STORAGE_FP16 1
#define LBM_STORAGE_BF16 2
#define LBM_STORAGE_FP64 3

/*
 * The converted lattice parameters: LBMLatticeParameters (LBMKernels.h:30-92) without
 * the MPI fields (procs, particleCopiesCheck), plus the storage type.  The coupled
 * particles and solid nodes are passed to lbm_create() as arrays.  The halo-exchange
 * flags (LBMKernels.h:40-42) only select which fields lbm_get_fluid_fields() fills
 * in its halo layer: with the whole lattice in one domain there is nothing to exchange.
 */
typedef struct lbm_params {
    int nx, ny, nz;
    double dx, dt;                   /* nm, ps */
    double density;                  /* Da/nm^3 of the fluid at rest */
    double tau, omega;
    double initial_velocity[3];      /* lattice units */
    double body_acceleration[3];     /* lattice units */
    double friction;                 /* 1/ps */
    double kT;                       /* kJ/mol; 0 with the NVE scheme */
    int drag_scheme;
    int interpolation_stencil;
    int fluid_fluctuations;
    double fluid_kT;                 /* kJ/mol; 0 without fluctuations */
    int random_number_seed;          /* 0: chosen at creation */
    int momentum_removal_frequency;
    int mach_check_frequency;
    double mach_number_limit;
    int wall_scheme;
    int face_boundary[6];
    double face_velocity[6][3];      /* lattice units */
    double face_density[6];          /* lattice units, 1 = fluid at rest */
    int storage;                     /* LBM_STORAGE_* */
    int density_halo_exchange;       /* densityHaloExchange */
    int velocity_halo_exchange;      /* velocityHaloExchange */
} lbm_params;

/*
 * The user-level settings of an LBMForce (LBMForce.h:41-691, defaults LBMForce.cpp:21-32),
 * in OpenMM units, without the MPI settings.  lbm_force_defaults() fills the reference's
 * defaults.
 */
typedef struct lbm_force_settings {
    int nx, ny, nz;
    double fluid_density;            /* Da/nm^3 */
    double kinematic_viscosity;      /* nm^2/ps */
    double friction;                 /* 1/ps */
    double temperature;              /* K */
    int coupling_scheme, drag_scheme, interpolation_stencil, wall_scheme;
    int fluid_fluctuations;
    int random_number_seed, momentum_removal_frequency, mach_check_frequency;
    double mach_number_limit;
    double body_acceleration[3];     /* nm/ps^2 */
    double initial_velocity[3];      /* nm/ps */
    int face_boundary[6];
    double face_velocity[6][3];      /* nm/ps */
    double face_density[6];          /* Da/nm^3, 0 = fluid at rest */
    int storage;
    int density_halo_exchange, velocity_halo_exchange;   /* LBMForce.h:242-263, off by default */
} lbm_force_settings;

typedef struct lbm_context lbm_context;

/* Message of the last failure on the calling thread ("" if none). */
const char* lbm_last_error(void);

/* Fill the defaults of LBMForce::LBMForce() (LBMForce.cpp:21-32). */
void lbm_force_defaults(lbm_force_settings* settings);

/*
 * Replaces LBMForceImpl::computeLatticeParameters (openmmapi/src/LBMForceImpl.cpp:44-214):
 * unit conversion and validation with the reference's error messages.  box: the lengths
 * of the rectangular periodic box (nm); step_size: the integrator step (ps); the coupled
 * particles (system indices) and the masses of all num_system_particles particles.
 */
int lbm_compute_lattice_parameters(const lbm_force_settings* settings, const double box[3], double step_size,
        const int* solid_nodes, int num_solid, const int* particles, int num_particles,
        const double* system_masses, int num_system_particles, lbm_params* out);

/*
 * Replaces CalcLBMForceKernel::initialize (LBMKernels.h:111; GPU CommonLBMKernels.cpp:160-1197):
 * allocate the lattice, put the fluid at equilibrium with the initial velocity.
 * particles: system indices of the coupled particles; particle_masses: their masses (Da);
 * num_system_particles: length of the position/velocity/force arrays of the other calls.
 * solid_nodes: node indices i + nx*(j + ny*k), sorted, without repetitions.
 * At most 2^30 nodes (LBM_ERROR_INVALID beyond).  LBM_AMD_WIDE_ADDRESSING=1 in the environment makes the lattice
 * pass use 64-bit indices at any size (it does by itself where an array of populations exceeds 4 GiB).
 * LBM_AMD_FUSED_ROWS=1 makes it stage its pulls through LDS where rows hold 128 or a multiple of 256 nodes (the same
 * results; slower on MI355X, so off by default).  Other switches that change the work, never the results (each run
 * is bit-identical across them): LBM_AMD_ROW_BLOCK=1|2|4 (lattice rows per block of the population layout, default
 * 1); and, for the explicit drag, LBM_AMD_SPREAD=atomic|tiles (the reaction spread), LBM_AMD_FIELD=0|1 (stencils read a
 * lattice-wide moment field) and LBM_AMD_STENCIL_SETUP=0|1 (each particle's stencil formed once), by default tiles,
 * field and setup for dense particles (stencil slots above a quarter of the nodes) and none of them otherwise, and
 * LBM_AMD_COUPLE_XCD=0|1 (with FP16/BF16 storage, each of the 8 XCDs couples a contiguous run of the particles in
 * their processing order rather than every 8th work-group's; default 1); for the
 * centred drag with a stencil, LBM_AMD_CG_PLANNED=0 makes the host read the conjugate gradients' residual after every
 * iteration instead of after the previous solve's count of them (default 1), and LBM_AMD_CG_SPLIT=0|1 spreads every
 * slot in their products or only those at nodes several slots share (default: the latter for sparse particles; the
 * results agree to the solver's tolerance, not bit for bit).
 */
int lbm_create(const lbm_params* params, const int* solid_nodes, int num_solid, const int* particles,
        const double* particle_masses, int num_particles, int num_system_particles, lbm_context** out);

/* Release the context and its device memory. */
int lbm_destroy(lbm_context* ctx);

/*
 * Replaces CalcLBMForceKernel::beginStep (LBMKernels.h:119; ReferenceLBMKernels.cpp:308-339):
 * record the step count of the Context and reflect the coupled particles moving into a
 * solid node (velocities are changed in place, xyz per system particle).
 */
int lbm_begin_step(lbm_context* ctx, int64_t step_count, const double* positions, double* velocities);

/*
 * Replaces CalcLBMForceKernel::execute (LBMKernels.h:128; ReferenceLBMKernels.cpp:341-365):
 * the first call after lbm_begin_step() advances the fluid by one lattice step; other
 * calls return the coupling force of the next step without advancing the fluid.  With
 * include_forces the coupling forces are ADDED to forces (xyz per system particle).
 */
int lbm_execute(lbm_context* ctx, int64_t step_count, int include_forces, const double* positions,
        const double* velocities, double* forces);

/* The same two calls with device pointers (positions, velocities, forces in device memory,
   FP64 xyz per system particle), for an integrator that keeps the particles on the GPU.
   The oracle library returns LBM_ERROR_UNSUPPORTED. */
int lbm_begin_step_device(lbm_context* ctx, int64_t step_count, const double* positions, double* velocities);
int lbm_execute_device(lbm_context* ctx, int64_t step_count, int include_forces, const double* positions,
        const double* velocities, double* forces);

/*
 * Replaces CalcLBMForceKernel::copyParametersToContext (LBMKernels.h:135), with the checks
 * that LBMForceImpl::updateParametersInContext makes before it (LBMForceImpl.cpp:388-414):
 * a change of the grid, the fluid density or viscosity, the wall scheme, the halo exchange,
 * the face boundary types, the drag scheme, the interpolation stencil or the fluctuations is
 * refused with LBM_ERROR_INVALID and the reference's message, and the context is unchanged.
 * (The coupled particles and the solid nodes are not part of lbm_params: the caller keeps them.)
 */
int lbm_copy_parameters(lbm_context* ctx, const lbm_params* params);

/*
 * Replaces CalcLBMForceKernel::getFluidFields (LBMKernels.h:136-145; ReferenceLBMKernels.cpp:1349-1368,
 * 1403-1436): density (Da/nm^3) and velocity (nm/ps, xyz per node); zero at solid nodes.  Either may be
 * NULL.  halo = 0: N values in node order i + nx*(j + ny*k).  halo = 1: the lattice with a layer one node
 * thick around it, (nx+2)(ny+2)(nz+2) values, node (i, j, k) for i from -1 to nx and so on at index
 * (i+1) + (nx+2)*((j+1) + (ny+2)*(k+1)) (LBMDecomposition.cpp:183-204 with one domain): the layer holds the
 * fields of the node across the periodic boundary for the fields whose halo is exchanged
 * (density_halo_exchange, velocity_halo_exchange), and NaN for the other fields and beyond open faces.
 */
int lbm_get_fluid_fields(lbm_context* ctx, double* density, double* velocity, int halo);

/*
 * Diagnostics with no counterpart in the reference: the nearest node (i + nx (j + ny k)) of every coupled particle,
 * in coupling order, for the positions (nm) of all particles of the System, by the rule of nearestNode
 * (ReferenceLBMKernels.cpp:973-983) as the coupling kernels apply it.
 */
int lbm_get_particle_nodes(lbm_context* ctx, const double* positions, int* nodes);

/* Replaces getFluidMachNumber (LBMKernels.h:149; ReferenceLBMKernels.cpp:1449-1468). */
int lbm_get_fluid_mach_number(lbm_context* ctx, double* mach);

/* Replaces getWallForce (LBMKernels.h:154; ReferenceLBMKernels.cpp:1438-1443), kJ/mol/nm. */
int lbm_get_wall_force(lbm_context* ctx, double force[3]);

/* Replace getFluidState/setFluidState (LBMKernels.h:159-164): FP64 deviations f_q - w_q
   in the layout [q*numNodes + node], size 19*numNodes. */
int lbm_get_fluid_state(lbm_context* ctx, double* state, size_t size);
int lbm_set_fluid_state(lbm_context* ctx, const double* state, size_t size);

/* Replace createCheckpoint/loadCheckpoint (LBMKernels.h:170-174) with byte buffers:
   lbm_checkpoint_size() gives the size, lbm_create_checkpoint() writes it. */
int lbm_checkpoint_size(lbm_context* ctx, size_t* size);
int lbm_create_checkpoint(lbm_context* ctx, void* buffer, size_t size);
int lbm_load_checkpoint(lbm_context* ctx, const void* buffer, size_t size);

/*
 * Diagnostics with no counterpart in the reference: the FP64 total mass (lattice units,
 * numFluidNodes + sum of the deviations) and momentum of the current fluid state, reduced
 * on the device in FP64 (drift checks for the reduced-precision storage).
 */
int lbm_get_fluid_totals(lbm_context* ctx, double* mass, double momentum[3]);

/* Synchronize the context's stream (for timing). */
int lbm_synchronize(lbm_context* ctx);

/*
 * Diagnostics with no counterpart in the reference: HIP-event timing of the fused lattice pass on the
 * context's stream.  Returns the total time (ms) and the number of passes recorded since the last call,
 * then switches recording on (enable = 1) or off (0).  The oracle library returns LBM_ERROR_UNSUPPORTED.
 */
int lbm_profile_fused(lbm_context* ctx, int enable, double* total_ms, int* count);

/*
 * Diagnostics with no counterpart in the reference: the duration (ms) of each fused pass recorded since recording
 * was switched on, oldest first, without clearing them.  Writes at most capacity values and the number recorded to
 * *count.  The oracle library does not export it.
 */
int lbm_profile_fused_launches(lbm_context* ctx, float* ms, int capacity, int* count);

#ifdef __cplusplus
}
#endif

#endif /* LBM_AMD_H_ */
