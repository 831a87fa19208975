/*
 * The LBMForce settings and their conversion into lattice parameters.
 *
 * lbm_force_defaults() gives the defaults of LBMForce::LBMForce() (openmmapi/src/LBMForce.cpp:21-32);
 * lbm_compute_lattice_parameters() restates LBMForceImpl::computeLatticeParameters
 * (openmmapi/src/LBMForceImpl.cpp:44-214) for one domain: the same validation, in the same order,
 * with the same messages, and the same unit conversions.  The MPI checks of the reference
 * (domain sizes, decomposition) have no counterpart on one GPU; the 32-bit index limit is
 * lifted to 2^31 nodes per population array with 64-bit offsets between the arrays.
 */
#include "../../include/lbm_amd.h"
#include <algorithm>
#include <cmath>
#include <set>
#include <sstream>
#include <string>
#include <vector>

namespace lbm {
void setLastError(const std::string& message);
}

namespace {

/* kJ/(mol K): the gas constant of OpenMM (SimTKOpenMMRealType.h: BOLTZ = RGAS/1000). */
const double BOLTZ = 1.380649e-23*6.02214076e23/1000.0;

struct SetupError {
    std::string message;
};

void require(bool ok, const std::string& message) {
    if (!ok)
        throw SetupError{message};
}

}  // namespace

extern "C" void lbm_force_defaults(lbm_force_settings* s) {
    *s = lbm_force_settings();
    s->nx = s->ny = s->nz = 0;
    s->fluid_density = 602.214;
    s->kinematic_viscosity = 1.0035;
    s->friction = 1.0;
    s->temperature = 300.0;
    s->coupling_scheme = LBM_COUPLING_EULER_MARUYAMA;
    s->drag_scheme = LBM_DRAG_EXPLICIT;
    s->interpolation_stencil = LBM_STENCIL_NEAREST;
    s->wall_scheme = LBM_WALL_BOUNCE_BACK;
    s->fluid_fluctuations = 0;
    s->random_number_seed = 0;
    s->momentum_removal_frequency = 1;
    s->mach_check_frequency = 100;
    s->mach_number_limit = 0.3;
    for (int k = 0; k < 3; k++)
        s->body_acceleration[k] = s->initial_velocity[k] = 0.0;
    for (int f = 0; f < 6; f++) {
        s->face_boundary[f] = LBM_BOUNDARY_PERIODIC;
        s->face_velocity[f][0] = s->face_velocity[f][1] = s->face_velocity[f][2] = 0.0;
        s->face_density[f] = 0.0;
    }
    s->storage = LBM_STORAGE_FP32;
    s->density_halo_exchange = s->velocity_halo_exchange = 0;
}

extern "C" int lbm_compute_lattice_parameters(const lbm_force_settings* s, const double box[3], double stepSize,
        const int* solidNodes, int numSolid, const int* particles, int numParticles, const double* systemMasses,
        int numSystemParticles, lbm_params* out) {
    try {
        require(s != nullptr && box != nullptr && out != nullptr, "lbm_compute_lattice_parameters: null argument");
        lbm_params p = lbm_params();
        p.nx = s->nx;
        p.ny = s->ny;
        p.nz = s->nz;
        require(p.nx > 0 && p.ny > 0 && p.nz > 0, "LBMForce: the grid size must be set to positive values with setGridSize()");
        long long numNodes = (long long) p.nx*p.ny*p.nz;
        if (numNodes > 2147483647LL) {
            std::stringstream msg;
            msg << "LBMForce: the lattice has " << numNodes << " nodes, more than the 2147483647 that the plugin can index";
            throw SetupError{msg.str()};
        }
        // The lattice spans the periodic box, with cubic cells (LBMForceImpl.cpp:80-92).
        p.dx = box[0]/p.nx;
        double dy = box[1]/p.ny, dz = box[2]/p.nz;
        if (std::fabs(dy - p.dx) > 1e-6*p.dx || std::fabs(dz - p.dx) > 1e-6*p.dx) {
            std::stringstream msg;
            msg << "LBMForce: the lattice cells must be cubic, but the box and grid sizes give spacings " << p.dx << ", "
                << dy << ", " << dz << " nm";
            throw SetupError{msg.str()};
        }
        require(stepSize > 0, "LBMForce: the integrator step size must be positive");
        p.dt = stepSize;
        require(s->fluid_density > 0, "LBMForce: the fluid density must be positive");
        require(s->kinematic_viscosity > 0, "LBMForce: the kinematic viscosity must be positive");
        p.density = s->fluid_density;
        p.tau = 3.0*s->kinematic_viscosity*p.dt/(p.dx*p.dx) + 0.5;
        p.omega = 1.0/p.tau;
        for (int k = 0; k < 3; k++) {
            p.initial_velocity[k] = s->initial_velocity[k]*(p.dt/p.dx);
            p.body_acceleration[k] = s->body_acceleration[k]*(p.dt*p.dt/p.dx);
        }
        require(s->friction >= 0, "LBMForce: the friction must not be negative");
        require(s->temperature >= 0, "LBMForce: the temperature must not be negative");
        require(s->momentum_removal_frequency >= 0, "LBMForce: the fluid momentum removal frequency must not be negative");
        require(s->mach_check_frequency >= 0, "LBMForce: the Mach number check frequency must not be negative");
        require(s->mach_number_limit > 0, "LBMForce: the Mach number limit must be positive");
        p.mach_check_frequency = s->mach_check_frequency;
        p.mach_number_limit = s->mach_number_limit;
        // Solid nodes (LBMForceImpl.cpp:130-144): sorted, in range, not repeated, not all.
        std::vector<int> solid(solidNodes, solidNodes + (solidNodes ? numSolid : 0));
        std::sort(solid.begin(), solid.end());
        for (size_t i = 0; i < solid.size(); i++) {
            require(solid[i] >= 0 && solid[i] < numNodes, "LBMForce: a solid node index is out of range");
            require(i == 0 || solid[i] != solid[i-1], "LBMForce: a solid node is listed more than once");
        }
        require((long long) solid.size() != numNodes, "LBMForce: all lattice nodes are solid");
        require(s->wall_scheme == LBM_WALL_BOUNCE_BACK || s->wall_scheme == LBM_WALL_REGULARIZED, "LBMForce: unknown wall scheme");
        p.wall_scheme = s->wall_scheme;
        // Faces (LBMForceImpl.cpp:146-172).
        const char* axisName[3] = {"x", "y", "z"};
        int size[3] = {p.nx, p.ny, p.nz};
        for (int f = 0; f < 6; f++) {
            int type = s->face_boundary[f];
            require(type == LBM_BOUNDARY_PERIODIC || type == LBM_BOUNDARY_VELOCITY || type == LBM_BOUNDARY_DENSITY ||
                    type == LBM_BOUNDARY_DENSITY_VELOCITY, "LBMForce: unknown boundary type of a face");
            p.face_boundary[f] = type;
            for (int k = 0; k < 3; k++)
                p.face_velocity[f][k] = s->face_velocity[f][k]*(p.dt/p.dx);
            require(s->face_density[f] >= 0, "LBMForce: the density of a face must not be negative");
            p.face_density[f] = s->face_density[f] == 0 ? 1.0 : s->face_density[f]/p.density;
        }
        bool open = false;
        for (int a = 0; a < 3; a++) {
            bool periodicLow = p.face_boundary[2*a] == LBM_BOUNDARY_PERIODIC;
            bool periodicHigh = p.face_boundary[2*a+1] == LBM_BOUNDARY_PERIODIC;
            require(periodicLow == periodicHigh, std::string("LBMForce: the two faces perpendicular to ") + axisName[a] +
                    " must be both periodic or both open (Velocity, Density or DensityVelocity)");
            require(periodicLow || size[a] >= 3, std::string("LBMForce: with open faces perpendicular to ") + axisName[a] +
                    " the grid needs at least 3 nodes along " + axisName[a]);
            open = open || !periodicLow;
        }
        require(!(open && s->momentum_removal_frequency > 0), "LBMForce: with open faces the fluid exchanges momentum with "
                "the outside, and its momentum cannot be removed: call setFluidMomentumRemovalFrequency(0)");
        p.friction = s->friction;
        require(s->coupling_scheme == LBM_COUPLING_EULER_MARUYAMA || s->coupling_scheme == LBM_COUPLING_NVE,
                "LBMForce: unknown coupling scheme");
        p.kT = s->coupling_scheme == LBM_COUPLING_NVE ? 0.0 : BOLTZ*s->temperature;
        require(s->drag_scheme == LBM_DRAG_EXPLICIT || s->drag_scheme == LBM_DRAG_CENTERED, "LBMForce: unknown drag scheme");
        p.drag_scheme = s->drag_scheme;
        require(s->interpolation_stencil >= LBM_STENCIL_NEAREST && s->interpolation_stencil <= LBM_STENCIL_KEYS,
                "LBMForce: unknown interpolation stencil");
        p.interpolation_stencil = s->interpolation_stencil;
        p.fluid_fluctuations = s->fluid_fluctuations ? 1 : 0;
        p.fluid_kT = p.fluid_fluctuations ? BOLTZ*s->temperature : 0.0;
        if (p.fluid_fluctuations)
            for (int f = 0; f < 6; f++)
                require(p.face_boundary[f] != LBM_BOUNDARY_DENSITY, "LBMForce: Density faces cannot be used with fluid "
                        "fluctuations: they do not keep the fluctuating fluid next to the face at equilibrium (its momentum "
                        "fluctuates more than at the temperature of the force, and coupled particles near the face are "
                        "hotter). Use DensityVelocity faces, which impose the density and the velocity, or Velocity faces");
        p.random_number_seed = s->random_number_seed;
        p.momentum_removal_frequency = s->momentum_removal_frequency;
        std::set<int> seen;
        for (int i = 0; i < numParticles; i++) {
            int part = particles[i];
            require(part >= 0 && part < numSystemParticles, "LBMForce: a coupled particle index is out of range");
            require(seen.find(part) == seen.end(), "LBMForce: a particle is coupled more than once");
            require(systemMasses != nullptr && systemMasses[part] > 0, "LBMForce: coupled particles must have a positive mass");
            seen.insert(part);
        }
        require(s->storage >= LBM_STORAGE_FP32 && s->storage <= LBM_STORAGE_FP64, "LBMForce: unknown storage type");
        p.storage = s->storage;
        p.density_halo_exchange = s->density_halo_exchange ? 1 : 0;     // LBMForceImpl.cpp:126-127
        p.velocity_halo_exchange = s->velocity_halo_exchange ? 1 : 0;
        *out = p;
        lbm::setLastError("");
        return LBM_OK;
    }
    catch (const SetupError& e) {
        lbm::setLastError(e.message);
        return LBM_ERROR_INVALID;
    }
}
