/*
 * Device-resident particle host for benchmarks and long runs: what OpenMM's Context and
 * VerletIntegrator do around the LBMForce kernel in the reference, kept on the GPU so that
 * a coupled step crosses no PCIe.  Per step, on the context's stream:
 *   lbm_begin_step_device -> zero forces -> extra forces (wall potential, harmonic bonds)
 *   -> lbm_execute_device (adds the coupling forces) -> leapfrog Verlet update
 *      v(t+dt/2) = v(t-dt/2) + dt F(t)/m,  x(t+dt) = x(t) + dt v(t+dt/2).
 * The extra forces are those of the benchmark configurations: a repulsive wall potential
 * eps[(sigma/d1)^12 + (sigma/d2)^12] between two planes z = zlo and z = zhi (the pattern of
 * the reference's examples/cocomo/shear.py:291-310), and harmonic bonds k/2 (r - r0)^2 with the
 * minimum image (bead chains, at most two bonds per particle), summed per particle in a fixed order.
 */
#include "lbm_internal.h"
#include <stdexcept>
#include <vector>

extern "C" hipStream_t lbm_internal_stream(lbm_context* c);

namespace {

struct MD {
    int n = 0;
    double dt = 0;
    double* pos = nullptr;
    double* vel = nullptr;
    double* force = nullptr;
    double* invMass = nullptr;
    bool wall = false;
    double eps = 0, sigma = 0, zlo = 0, zhi = 0;
    int numBonds = 0;
    int* bonds = nullptr;
    double bondK = 0, bondR0 = 0;
    double box[3] = {0, 0, 0};
};

thread_local std::string g_err;

/* The forces of the System other than the LBMForce, written (not added) for every particle: the repulsive walls
   U = eps[(s/d1)^12 + (s/d2)^12] along z (when set) plus the harmonic bonds to the (at most two) chain partners, in
   that order. */
__global__ void otherForces(const double* pos, double* f, int n, int wall, double eps, double sigma, double zlo,
        double zhi, const int* partners, double k, double r0, double bx, double by, double bz) {
    int i = blockIdx.x*256 + threadIdx.x;
    if (i >= n)
        return;
    double acc[3] = {0, 0, 0};
    if (wall) {
        double z = pos[3*(size_t) i + 2];
        double d1 = z - zlo, d2 = zhi - z;
        // -dU/dz = 12 eps s^12/d1^13 - 12 eps s^12/d2^13.
        double s1 = sigma/d1, s2 = sigma/d2;
        double s1_12 = s1*s1*s1*s1*s1*s1*s1*s1*s1*s1*s1*s1, s2_12 = s2*s2*s2*s2*s2*s2*s2*s2*s2*s2*s2*s2;
        acc[2] = 12.0*eps*(s1_12/d1 - s2_12/d2);
    }
    if (partners != nullptr) {
        double box[3] = {bx, by, bz}, b[3] = {0, 0, 0};
        for (int p = 0; p < 2; p++) {
            int j = partners[2*i + p];
            if (j < 0)
                continue;
            double d[3];
            for (int a = 0; a < 3; a++) {
                d[a] = pos[3*(size_t) j + a] - pos[3*(size_t) i + a];
                d[a] -= box[a]*rint(d[a]/box[a]);
            }
            double r = sqrt(d[0]*d[0] + d[1]*d[1] + d[2]*d[2]);
            double s = r > 0 ? k*(r - r0)/r : 0.0;
            for (int a = 0; a < 3; a++)
                b[a] += s*d[a];
        }
        for (int a = 0; a < 3; a++)
            acc[a] += b[a];
    }
    for (int a = 0; a < 3; a++)
        f[3*(size_t) i + a] = acc[a];
}

__global__ void verlet(double* pos, double* vel, const double* f, const double* invMass, int n, double dt) {
    int i = blockIdx.x*256 + threadIdx.x;
    if (i >= n)
        return;
    for (int a = 0; a < 3; a++) {
        size_t k = 3*(size_t) i + a;
        double v = vel[k] + dt*f[k]*invMass[i];
        vel[k] = v;
        pos[k] += dt*v;
    }
}

void hipCheck(hipError_t e) {
    if (e != hipSuccess)
        throw std::runtime_error(std::string("HIP error ") + hipGetErrorString(e));
}

}  // namespace

extern "C" {

typedef struct lbm_md_system lbm_md_system;

const char* lbm_md_last_error(void) {
    return g_err.c_str();
}

int lbm_md_create(int n, const double* masses, const double* positions, const double* velocities, double dt,
        lbm_md_system** out) {
    MD* md = new MD();
    try {
        md->n = n;
        md->dt = dt;
        size_t bytes = 3*sizeof(double)*(size_t) n;
        hipCheck(hipMalloc(&md->pos, bytes));
        hipCheck(hipMalloc(&md->vel, bytes));
        hipCheck(hipMalloc(&md->force, bytes));
        hipCheck(hipMalloc(&md->invMass, sizeof(double)*(size_t) n));
        std::vector<double> inv(n);
        for (int i = 0; i < n; i++)
            inv[i] = masses[i] > 0 ? 1.0/masses[i] : 0.0;
        hipCheck(hipMemcpy(md->pos, positions, bytes, hipMemcpyHostToDevice));
        hipCheck(hipMemcpy(md->vel, velocities, bytes, hipMemcpyHostToDevice));
        hipCheck(hipMemcpy(md->invMass, inv.data(), sizeof(double)*(size_t) n, hipMemcpyHostToDevice));
    }
    catch (const std::exception& e) {
        g_err = e.what();
        delete md;
        return LBM_ERROR_RUNTIME;
    }
    *out = (lbm_md_system*) md;
    return LBM_OK;
}

int lbm_md_set_wall_potential(lbm_md_system* s, double eps, double sigma, double zlo, double zhi) {
    MD* md = (MD*) s;
    md->wall = true;
    md->eps = eps;
    md->sigma = sigma;
    md->zlo = zlo;
    md->zhi = zhi;
    return LBM_OK;
}

int lbm_md_set_bonds(lbm_md_system* s, const int* pairs, int numBonds, double k, double r0, const double box[3]) {
    MD* md = (MD*) s;
    try {
        if (md->bonds)
            hipCheck(hipFree(md->bonds));
        md->bonds = nullptr;
        md->numBonds = numBonds;
        if (numBonds > 0) {
            std::vector<int> partners(2*(size_t) md->n, -1);
            for (int b = 0; b < numBonds; b++) {
                int i = pairs[2*b], j = pairs[2*b + 1];
                if (i < 0 || j < 0 || i >= md->n || j >= md->n)
                    throw std::runtime_error("lbm_md_set_bonds: particle index out of range");
                for (int end = 0; end < 2; end++) {
                    int a = end == 0 ? i : j, other = end == 0 ? j : i;
                    int slot = partners[2*(size_t) a] < 0 ? 0 : partners[2*(size_t) a + 1] < 0 ? 1 : -1;
                    if (slot < 0)
                        throw std::runtime_error("lbm_md_set_bonds: a particle has more than two bonds (chains only)");
                    partners[2*(size_t) a + slot] = other;
                }
            }
            hipCheck(hipMalloc(&md->bonds, partners.size()*sizeof(int)));
            hipCheck(hipMemcpy(md->bonds, partners.data(), partners.size()*sizeof(int), hipMemcpyHostToDevice));
        }
        md->bondK = k;
        md->bondR0 = r0;
        for (int a = 0; a < 3; a++)
            md->box[a] = box[a];
    }
    catch (const std::exception& e) {
        g_err = e.what();
        return LBM_ERROR_RUNTIME;
    }
    return LBM_OK;
}

/* Run numSteps integration steps starting at step count step0. */
int lbm_md_step(lbm_md_system* s, lbm_context* ctx, int64_t step0, int numSteps) {
    MD* md = (MD*) s;
    hipStream_t stream = lbm_internal_stream(ctx);
    int grid = (md->n + 255)/256;
    for (int t = 0; t < numSteps; t++) {
        int64_t step = step0 + t;
        int status = lbm_begin_step_device(ctx, step, md->pos, md->vel);
        if (status != LBM_OK) {
            g_err = lbm_last_error();
            return status;
        }
        otherForces<<<grid, 256, 0, stream>>>(md->pos, md->force, md->n, md->wall ? 1 : 0, md->eps, md->sigma, md->zlo,
                md->zhi, md->numBonds > 0 ? md->bonds : nullptr, md->bondK, md->bondR0, md->box[0], md->box[1],
                md->box[2]);
        status = lbm_execute_device(ctx, step, 1, md->pos, md->vel, md->force);
        if (status != LBM_OK) {
            g_err = lbm_last_error();
            return status;
        }
        verlet<<<grid, 256, 0, stream>>>(md->pos, md->vel, md->force, md->invMass, md->n, md->dt);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) {
            g_err = hipGetErrorString(e);
            return LBM_ERROR_RUNTIME;
        }
    }
    return LBM_OK;
}

int lbm_md_get(lbm_md_system* s, lbm_context* ctx, double* positions, double* velocities) {
    MD* md = (MD*) s;
    hipStream_t stream = lbm_internal_stream(ctx);
    size_t bytes = 3*sizeof(double)*(size_t) md->n;
    if (hipStreamSynchronize(stream) != hipSuccess)
        return LBM_ERROR_RUNTIME;
    if (positions && hipMemcpy(positions, md->pos, bytes, hipMemcpyDeviceToHost) != hipSuccess)
        return LBM_ERROR_RUNTIME;
    if (velocities && hipMemcpy(velocities, md->vel, bytes, hipMemcpyDeviceToHost) != hipSuccess)
        return LBM_ERROR_RUNTIME;
    return LBM_OK;
}

int lbm_md_destroy(lbm_md_system* s) {
    MD* md = (MD*) s;
    if (!md)
        return LBM_OK;
    void* bufs[] = {md->pos, md->vel, md->force, md->invMass, md->bonds};
    for (void* b : bufs)
        if (b)
            (void) hipFree(b);
    delete md;
    return LBM_OK;
}

}  // extern "C"
