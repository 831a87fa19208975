/*
 * FP64 CPU ORACLE -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain double-precision restatement of the reference's Reference platform for the
 * one-GPU LBMForce path, used by tests/, the smoke check at the repository root and bench.py's CPU
 * baseline as the reference results.  Nothing in the product library links or calls it.
 *
 * It follows, step for step, platforms/reference/src/ReferenceLBMKernels.cpp (push
 * streaming into a single population array, moments, removal, explicit coupling at the
 * nearest node and with the interpolation stencils, bounce-back, momentum exchange,
 * regularized walls and open faces, Mach check, fields, state), with the algebra of
 * openmmapi/include/internal/D3Q19.h, the stencils of internal/LBMStencils.h and the
 * boundary discovery of internal/LBMBoundaries.h.  It is deliberately a different
 * scheme (push, one array, stored moments) from the GPU library (pull, two arrays,
 * fused), so that agreement between the two checks the GPU's restructuring.
 *
 * The one intended difference from the reference: the random numbers.  The reference
 * draws them from an SFMT generator owned by the kernel, in a sequence order
 * (ReferenceLBMKernels.cpp:1052-1064, 945-957, 1080-1087, 1279-1288).  Here, as on the
 * GPU, every normal number is a function of (seed, purpose, index, step): Philox4x32-10
 * (Salmon et al., SC'11) followed by the Box-Muller transform.  Exact values of the
 * reference's SFMT stream are therefore not reproduced (parity unpinned for that stream;
 * the reference's own GPU platforms differ from it too, TestLBMFluctuations.h:8-15).
 *
 * Pinned by: tests/test_closed_forms.py (the closed forms of the reference's fluid, wall,
 * face, coupling, stencil, checkpoint and parameter tests: uniform flow, conservation, body
 * force, removal timing, shear-wave viscosity, exact bounce-back and regularized Poiseuille,
 * wall balance, Couette, reservoir, inlet, duct, first-step drag, momentum balances, stencil
 * exactness, restarts), tests/test_centered.py (the centred drag against its closed forms and a
 * Gaussian-elimination solution), tests/test_fluctuations.py (mode variances, spectra,
 * equipartition), tests/test_oracle.py (Philox known-answer vectors, algebra identities, stencil
 * moments) and, for the D3Q19 algebra, the reference's own D3Q19.h compiled into oracle/_ref
 * (tests/test_oracle.py::test_algebra_matches_reference_header).
 */
#include "../include/lbm_amd.h"
#include <algorithm>
#include <chrono>
#include <cmath>
#ifdef _OPENMP
#include <omp.h>
#endif
#include <cstring>
#include <limits>
#include <random>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

/* Threads of the node loops (lbm_oracle_threads): one by default, the serial loops of the reference. */
static int oracleThreads = 1;

namespace {

thread_local std::string g_lastError;

struct OracleError : public std::runtime_error {
    int status;
    OracleError(int s, const std::string& m) : std::runtime_error(m), status(s) {}
};

// ---------------------------------------------------------------------------------------------
// D3Q19 algebra (internal/D3Q19.h:26-208).
// ---------------------------------------------------------------------------------------------
const int NQ = 19;
const int CX[NQ] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1};
const int CY[NQ] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 1, -1, 1, -1, 0, 0, 0, 0};
const int CZ[NQ] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, 1, -1};
const int OPP[NQ] = {0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17};
const double W[NQ] = {1.0/3.0, 1.0/18.0, 1.0/18.0, 1.0/18.0, 1.0/18.0, 1.0/18.0, 1.0/18.0,
        1.0/36.0, 1.0/36.0, 1.0/36.0, 1.0/36.0, 1.0/36.0, 1.0/36.0, 1.0/36.0, 1.0/36.0, 1.0/36.0,
        1.0/36.0, 1.0/36.0, 1.0/36.0};
const double MODE_NORM[NQ] = {1.0, 1.0/3.0, 1.0/3.0, 1.0/3.0, 2.0/9.0, 2.0/9.0, 2.0/9.0, 1.0/9.0,
        1.0/9.0, 1.0/9.0, 2.0/27.0, 2.0/27.0, 2.0/27.0, 1.0/18.0, 1.0/18.0, 1.0/18.0, 7.0/162.0,
        5.0/126.0, 1.0/30.0};

// Deviation of the equilibrium from the rest state (D3Q19.h:55-62).
void equilibriumDeviation(double dr, double ux, double uy, double uz, double out[NQ]) {
    double rho = 1.0 + dr;
    double uu = ux*ux + uy*uy + uz*uz;
    for (int q = 0; q < NQ; q++) {
        double cu = CX[q]*ux + CY[q]*uy + CZ[q]*uz;
        out[q] = W[q]*(dr + rho*(3.0*cu + 4.5*cu*cu - 1.5*uu));
    }
}

void hermite2(int q, double h[6]) {
    h[0] = CX[q]*CX[q] - 1.0/3.0;
    h[1] = CY[q]*CY[q] - 1.0/3.0;
    h[2] = CZ[q]*CZ[q] - 1.0/3.0;
    h[3] = CX[q]*CY[q];
    h[4] = CX[q]*CZ[q];
    h[5] = CY[q]*CZ[q];
}

// Pi_neq from the deviations, equilibrium at j/rho (D3Q19.h:98-110).
void piNeqFromDeviation(const double df[NQ], double dr, double jx, double jy, double jz, double pi[6]) {
    double rho = 1.0 + dr;
    double eq[NQ], h[6];
    equilibriumDeviation(dr, jx/rho, jy/rho, jz/rho, eq);
    for (int k = 0; k < 6; k++)
        pi[k] = 0.0;
    for (int q = 0; q < NQ; q++) {
        hermite2(q, h);
        double neq = df[q] - eq[q];
        for (int k = 0; k < 6; k++)
            pi[k] += h[k]*neq;
    }
}

// Regularized non-equilibrium part (D3Q19.h:116-122).
void regularizedNeq(const double pi[6], double out[NQ]) {
    double h[6];
    for (int q = 0; q < NQ; q++) {
        hermite2(q, h);
        out[q] = 4.5*W[q]*(h[0]*pi[0] + h[1]*pi[1] + h[2]*pi[2] + 2.0*(h[3]*pi[3] + h[4]*pi[4] + h[5]*pi[5]));
    }
}

// Guo source term (D3Q19.h:127-134).
void guoSource(double ux, double uy, double uz, double fx, double fy, double fz, double out[NQ]) {
    double uf = ux*fx + uy*fy + uz*fz;
    for (int q = 0; q < NQ; q++) {
        double cu = CX[q]*ux + CY[q]*uy + CZ[q]*uz;
        double cf = CX[q]*fx + CY[q]*fy + CZ[q]*fz;
        out[q] = W[q]*(3.0*(cf - uf) + 9.0*cu*cf);
    }
}

// Orthogonal basis of Lulli et al. (D3Q19.h:145-169).
double basisMode(int k, int q) {
    double x = CX[q], y = CY[q], z = CZ[q];
    double x2 = x*x, y2 = y*y, z2 = z*z;
    switch (k) {
    case 0: return 1.0;
    case 1: return x;
    case 2: return y;
    case 3: return z;
    case 4: return x2 - 1.0/3.0;
    case 5: return y2 - 1.0/3.0;
    case 6: return z2 - 1.0/3.0;
    case 7: return x*y;
    case 8: return x*z;
    case 9: return y*z;
    case 10: return y*(x2 - 1.0/3.0);
    case 11: return z*(x2 - 1.0/3.0);
    case 12: return x*(y2 - 1.0/3.0);
    case 13: return 0.5*x*(y2 + 2.0*z2 - 1.0);
    case 14: return 0.5*y*(x2 + 2.0*z2 - 1.0);
    case 15: return 0.5*z*(x2 + 2.0*y2 - 1.0);
    case 16: return x2*y2 - x2/3.0 - y2/3.0 + z2/6.0 + 1.0/18.0;
    case 17: return (2.0/7.0)*x2*y2 + x2*z2 - (3.0/7.0)*x2 + y2/14.0 - (2.0/7.0)*z2 + 1.0/14.0;
    default: return 0.4*x2*y2 + 0.4*x2*z2 + y2*z2 - 0.1*x2 - 0.4*y2 - 0.4*z2 + 0.1;
    }
}

// Random part of a collision, modes 4..18 (D3Q19.h:179-186).
void collisionNoise(const double r[15], double stressAmp, double ghostAmp, double xi[NQ]) {
    for (int q = 0; q < NQ; q++) {
        double sum = 0.0;
        for (int k = 4; k < NQ; k++)
            sum += basisMode(k, q)/std::sqrt(MODE_NORM[k])*(k < 10 ? stressAmp : ghostAmp)*r[k-4];
        xi[q] = W[q]*sum;
    }
}

// Random part of a node at equilibrium, all modes (D3Q19.h:194-199).
void equilibriumNoise(const double r[NQ], double conservedAmp, double stressAmp, double ghostAmp, double xi[NQ]) {
    collisionNoise(r, stressAmp, ghostAmp, xi);
    for (int q = 0; q < NQ; q++)
        xi[q] += W[q]*conservedAmp*(r[15] + std::sqrt(3.0)*(CX[q]*r[16] + CY[q]*r[17] + CZ[q]*r[18]));
}

double boundaryStressWeight(double omega) {
    return (1.0 - omega)*(1.0 - omega)/(omega*(2.0 - omega));
}

// ---------------------------------------------------------------------------------------------
// Interpolation stencils (internal/LBMStencils.h:29-104).
// ---------------------------------------------------------------------------------------------
int stencilWidth(int stencil) {
    switch (stencil) {
    case LBM_STENCIL_TRILINEAR: return 2;
    case LBM_STENCIL_THREE_POINT: return 3;
    case LBM_STENCIL_KEYS: return 4;
    default: return 1;
    }
}

double stencilKernel(int stencil, double r) {
    double s = std::fabs(r);
    switch (stencil) {
    case LBM_STENCIL_TRILINEAR:
        return s < 1.0 ? 1.0 - s : 0.0;
    case LBM_STENCIL_THREE_POINT:
        if (s <= 0.5)
            return (1.0 + std::sqrt(1.0 - 3.0*s*s))/3.0;
        if (s < 1.5)
            return (5.0 - 3.0*s - std::sqrt(std::max(0.0, 1.0 - 3.0*(1.0 - s)*(1.0 - s))))/6.0;
        return 0.0;
    case LBM_STENCIL_KEYS:
        if (s <= 1.0)
            return 1.0 - 2.5*s*s + 1.5*s*s*s;
        if (s < 2.0)
            return 2.0 - 4.0*s + 2.5*s*s - 0.5*s*s*s;
        return 0.0;
    default:
        return (r >= -0.5 && r < 0.5) ? 1.0 : 0.0;
    }
}

void axisWeights(int stencil, double s, int n, bool open, int* index, double* weight) {
    int w = stencilWidth(stencil);
    int first;
    if (stencil == LBM_STENCIL_THREE_POINT)
        first = (int) std::floor(s + 0.5) - 1;
    else if (stencil == LBM_STENCIL_KEYS)
        first = (int) std::floor(s) - 1;
    else if (stencil == LBM_STENCIL_TRILINEAR)
        first = (int) std::floor(s);
    else
        first = (int) std::floor(s + 0.5);
    for (int a = 0; a < w; a++) {
        weight[a] = stencilKernel(stencil, s - (first + a));
        int i = first + a;
        index[a] = open ? std::min(std::max(i, 0), n - 1) : ((i % n) + n) % n;
    }
    // Repeated nodes: the sum goes into the first slot, in slot order.
    for (int a = 1; a < w; a++)
        for (int b = 0; b < a; b++)
            if (index[b] == index[a]) {
                weight[b] += weight[a];
                weight[a] = 0.0;
                break;
            }
}

// ---------------------------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) and the Box-Muller transform.
// ---------------------------------------------------------------------------------------------
const uint32_t PHILOX_M0 = 0xD2511F53u, PHILOX_M1 = 0xCD9E8D57u;
const uint32_t PHILOX_W0 = 0x9E3779B9u, PHILOX_W1 = 0xBB67AE85u;
const uint32_t TAG_FLUID = 0x464C5549u, TAG_BOUNDARY = 0x424F554Eu, TAG_PARTICLE = 0x50415254u;

void philox4x32(const uint32_t ctrIn[4], const uint32_t keyIn[2], uint32_t out[4]) {
    uint32_t c0 = ctrIn[0], c1 = ctrIn[1], c2 = ctrIn[2], c3 = ctrIn[3];
    uint32_t k0 = keyIn[0], k1 = keyIn[1];
    for (int round = 0; round < 10; round++) {
        uint64_t p0 = (uint64_t) PHILOX_M0*c0;
        uint64_t p1 = (uint64_t) PHILOX_M1*c2;
        uint32_t hi0 = (uint32_t) (p0 >> 32), lo0 = (uint32_t) p0;
        uint32_t hi1 = (uint32_t) (p1 >> 32), lo1 = (uint32_t) p1;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n1 = lo1;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        uint32_t n3 = lo0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += PHILOX_W0;
        k1 += PHILOX_W1;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

// Normal numbers for (seed, tag, index, step), in blocks of 16: three Philox blocks with counters
// (index, step_lo, step_hi, 3*block + c) give 12 words; normals 2p and 2p+1 of the block are the Box-Muller
// pair of the radius word w[p] and the 16-bit angle word (w[8 + p/2] >> 16*(p%2)) & 0xFFFF:
// u1 = (a + 1) 2^-32 in (0, 1], u2 = b 2^-16 in [0, 1), r = sqrt(-2 ln u1), (r cos 2 pi u2, r sin 2 pi u2).
void drawNormals(uint32_t seed, uint32_t tag, uint32_t index, int64_t step, int count, double* out) {
    const double twoPi = 6.283185307179586476925287;
    for (int block = 0; 16*block < count; block++) {
        uint32_t w[12];
        for (int c = 0; c < 3; c++) {
            uint32_t ctr[4] = {index, (uint32_t) (uint64_t) step, (uint32_t) ((uint64_t) step >> 32), (uint32_t) (3*block + c)};
            uint32_t key[2] = {seed, tag};
            philox4x32(ctr, key, w + 4*c);
        }
        for (int p = 0; p < 8 && 16*block + 2*p < count; p++) {
            uint32_t angle = (w[8 + p/2] >> (16*(p & 1))) & 0xFFFFu;
            double r = std::sqrt(-2.0*std::log(((double) w[p] + 1.0)*2.3283064365386963e-10));
            double a = twoPi*((double) angle*1.52587890625e-05);
            out[16*block + 2*p] = r*std::cos(a);
            if (16*block + 2*p + 1 < count)
                out[16*block + 2*p + 1] = r*std::sin(a);
        }
    }
}

// ---------------------------------------------------------------------------------------------
// Boundary kinds (internal/LBMBoundaries.h:30).
// ---------------------------------------------------------------------------------------------
enum Kind { KIND_WALL = 0, KIND_VELOCITY = 1, KIND_DENSITY = 2, KIND_DENSITY_AT_REST = 3, KIND_DENSITY_VELOCITY = 4 };

bool hasEquilibriumNoise(int kind) {
    return kind == KIND_WALL || kind == KIND_DENSITY_VELOCITY;
}

struct Vec3d {
    double x[3];
    Vec3d() { x[0] = x[1] = x[2] = 0.0; }
    Vec3d(double a, double b, double c) { x[0] = a; x[1] = b; x[2] = c; }
    double& operator[](int i) { return x[i]; }
    double operator[](int i) const { return x[i]; }
};

// ---------------------------------------------------------------------------------------------
// The kernel (ReferenceCalcLBMForceKernel without the domain decomposition).
// ---------------------------------------------------------------------------------------------
class OracleKernel {
public:
    lbm_params p;
    int numNodes = 0;
    int numSystemParticles = 0;
    uint32_t seed = 0;
    std::vector<int> particles;
    std::vector<double> particleMass;           // lattice units m/m_c
    std::vector<int> solidNodes;
    std::vector<double> pop;                    // [q*N + node], deviations f - w
    std::vector<double> rho, dev, mom, piNeq, force;   // moments of the current step
    std::vector<char> isFluid;                  // empty: all fluid
    std::vector<int> wallNodes, wallLinks;
    std::vector<int> bNodes, bUnknown, bSolid, bKind, bFace;
    std::vector<int> boundaryOf;
    std::vector<double> boundaryStress;
    std::vector<double> reaction;
    std::vector<Vec3d> particleForces;
    bool stepPending = false, stepForcesCurrent = false;
    int64_t stepIndex = 0;
    Vec3d wallMomentum;

    bool openAxis(int a) const {
        return p.face_boundary[2*a] != LBM_BOUNDARY_PERIODIC;
    }
    bool hasOpenFaces() const {
        return openAxis(0) || openAxis(1) || openAxis(2);
    }
    double cellMass() const {
        return p.density*p.dx*p.dx*p.dx;
    }
    int nodeIndex(int i, int j, int k) const {
        return i + p.nx*(j + p.ny*k);
    }
    bool fluid(int node) const {
        return isFluid.empty() || isFluid[node];
    }

    void initialize(const lbm_params& params, const int* solid, int numSolid, const int* parts, const double* masses,
            int numParts, int numSys) {
        p = params;
        numNodes = p.nx*p.ny*p.nz;
        numSystemParticles = numSys;
        seed = (uint32_t) p.random_number_seed;
        if (seed == 0)
            seed = (uint32_t) std::chrono::high_resolution_clock::now().time_since_epoch().count() | 1u;
        // Equilibrium at density 1 and the initial velocity; solid nodes hold no fluid (ReferenceLBMKernels.cpp:53-79).
        pop.assign((size_t) NQ*numNodes, 0.0);
        double eq[NQ];
        equilibriumDeviation(0.0, p.initial_velocity[0], p.initial_velocity[1], p.initial_velocity[2], eq);
        for (int q = 0; q < NQ; q++)
            std::fill(pop.begin() + (size_t) q*numNodes, pop.begin() + (size_t) (q+1)*numNodes, eq[q]);
        rho.assign(numNodes, 0.0);
        dev.assign(numNodes, 0.0);
        mom.assign(3*(size_t) numNodes, 0.0);
        piNeq.assign(6*(size_t) numNodes, 0.0);
        force.assign(3*(size_t) numNodes, 0.0);
        solidNodes.assign(solid, solid + numSolid);
        isFluid.clear();
        if (numSolid > 0) {
            isFluid.assign(numNodes, 1);
            for (int node : solidNodes) {
                isFluid[node] = 0;
                for (int q = 0; q < NQ; q++)
                    pop[(size_t) q*numNodes + node] = -W[q];
            }
        }
        findWallLinks();
        findBoundaries();
        boundaryOf.clear();
        boundaryStress.clear();
        bool eqNoise = false;
        for (int kind : bKind)
            eqNoise = eqNoise || hasEquilibriumNoise(kind);
        if (p.fluid_fluctuations && eqNoise) {
            boundaryOf.assign(numNodes, -1);
            for (size_t b = 0; b < bNodes.size(); b++)
                boundaryOf[bNodes[b]] = (int) b;
            boundaryStress.assign((size_t) NQ*bNodes.size(), 0.0);
        }
        particles.assign(parts, parts + numParts);
        particleMass.resize(numParts);
        for (int i = 0; i < numParts; i++)
            particleMass[i] = masses[i]/cellMass();
        particleForces.assign(numParts, Vec3d());
        reaction.assign(3*(size_t) numNodes, 0.0);
        wallMomentum = Vec3d();
    }

    // Bounce-back links (LBMBoundaries.h:118-149).
    void findWallLinks() {
        wallNodes.clear();
        wallLinks.clear();
        if (solidNodes.empty() || p.wall_scheme != LBM_WALL_BOUNCE_BACK)
            return;
        int size[3] = {p.nx, p.ny, p.nz};
        for (int node = 0; node < numNodes; node++) {
            if (!isFluid[node])
                continue;
            int idx[3] = {node % p.nx, (node/p.nx) % p.ny, node/(p.nx*p.ny)};
            int links = 0;
            for (int q = 1; q < NQ; q++) {
                int c[3] = {CX[q], CY[q], CZ[q]}, t[3];
                bool crosses = false;
                for (int a = 0; a < 3; a++) {
                    t[a] = idx[a] + c[a];
                    crosses = crosses || (openAxis(a) && (t[a] < 0 || t[a] >= size[a]));
                    t[a] = (t[a] + size[a]) % size[a];
                }
                if (!crosses && !isFluid[nodeIndex(t[0], t[1], t[2])])
                    links |= 1 << q;
            }
            if (links) {
                wallNodes.push_back(node);
                wallLinks.push_back(links);
            }
        }
    }

    // Boundary nodes of regularized walls and open faces (LBMBoundaries.h:46-109).
    void findBoundaries() {
        bNodes.clear(); bUnknown.clear(); bSolid.clear(); bKind.clear(); bFace.clear();
        bool regularized = !solidNodes.empty() && p.wall_scheme == LBM_WALL_REGULARIZED;
        if (!regularized && !hasOpenFaces())
            return;
        int size[3] = {p.nx, p.ny, p.nz};
        for (int node = 0; node < numNodes; node++) {
            if (!fluid(node))
                continue;
            int idx[3] = {node % p.nx, (node/p.nx) % p.ny, node/(p.nx*p.ny)};
            int solidBits = 0, outsideBits = 0;
            for (int q = 1; q < NQ; q++) {
                int c[3] = {CX[q], CY[q], CZ[q]}, s[3];
                bool beyond = false;
                for (int a = 0; a < 3; a++) {
                    s[a] = idx[a] - c[a];
                    if (s[a] < 0 || s[a] >= size[a]) {
                        beyond = beyond || openAxis(a);
                        s[a] = (s[a] + size[a]) % size[a];
                    }
                }
                if (beyond)
                    outsideBits |= 1 << q;
                else if (regularized && !isFluid[nodeIndex(s[0], s[1], s[2])])
                    solidBits |= 1 << q;
            }
            if (solidBits == 0 && outsideBits == 0)
                continue;
            int kind = KIND_WALL, face = -1;
            if (solidBits == 0) {
                int velocityFace = -1, densityFace = -1, numDensity = 0;
                for (int f = 0; f < 6; f++) {
                    int a = f/2;
                    if (!openAxis(a) || idx[a] != (f % 2 == 0 ? 0 : size[a] - 1))
                        continue;
                    int type = p.face_boundary[f];
                    if ((type == LBM_BOUNDARY_VELOCITY || type == LBM_BOUNDARY_DENSITY_VELOCITY) && velocityFace < 0)
                        velocityFace = f;
                    if (type == LBM_BOUNDARY_DENSITY && numDensity++ == 0)
                        densityFace = f;
                }
                if (velocityFace >= 0) {
                    kind = p.face_boundary[velocityFace] == LBM_BOUNDARY_DENSITY_VELOCITY ? KIND_DENSITY_VELOCITY : KIND_VELOCITY;
                    face = velocityFace;
                }
                else {
                    kind = numDensity == 1 ? KIND_DENSITY : KIND_DENSITY_AT_REST;
                    face = densityFace;
                }
            }
            bNodes.push_back(node);
            bUnknown.push_back(solidBits | outsideBits);
            bSolid.push_back(solidBits);
            bKind.push_back(kind);
            bFace.push_back(face);
        }
    }

    void beginStep(int64_t stepCount, const double* positions, double* velocities) {
        // ReferenceLBMKernels.cpp:308-339.
        stepIndex = stepCount;
        stepPending = true;
        wallMomentum = Vec3d();
        if (isFluid.empty())
            return;
        for (size_t i = 0; i < particles.size(); i++) {
            int part = particles[i];
            const double* x = positions + 3*(size_t) part;
            double* v = velocities + 3*(size_t) part;
            Vec3d n = wallNormal(x);
            if (!isFluid[nearestNode(x)] && v[0]*n[0] + v[1]*n[1] + v[2]*n[2] > 0) {
                for (int k = 0; k < 3; k++) {
                    wallMomentum[k] += v[k]*(2.0*particleMass[i]*cellMass());
                    v[k] = -v[k];
                }
            }
        }
    }

    void execute(int64_t stepCount, bool includeForces, const double* positions, const double* velocities, double* forces) {
        // ReferenceLBMKernels.cpp:341-365.
        bool repeated = stepForcesCurrent && stepCount == stepIndex - 1;
        if (stepPending) {
            stepPending = false;
            advanceFluid(positions, velocities, forces);
        }
        else if (includeForces && !particles.empty() && stepCount > 0 && !repeated)
            nextStepForces(stepCount, positions, velocities, forces);
        if (includeForces)
            for (size_t i = 0; i < particles.size(); i++)
                for (int k = 0; k < 3; k++)
                    forces[3*(size_t) particles[i] + k] += particleForces[i][k];
    }

    void advanceFluid(const double* positions, const double* velocities, const double* forces) {
        // ReferenceLBMKernels.cpp:367-389.
        computeMoments();
        if (p.momentum_removal_frequency > 0 && stepIndex % p.momentum_removal_frequency == 0)
            removeMomentum();
        if (!particles.empty())
            couple(positions, velocities, forces, true, stepIndex);
        collideAndStream();
        if (!wallNodes.empty()) {
            bounceBack();
            wallExchange();
        }
        if (!bNodes.empty())
            applyBoundaries();
        stepForcesCurrent = true;
        stepIndex++;
        if (p.mach_check_frequency > 0 && stepIndex % p.mach_check_frequency == 0) {
            double mach = machNumber();
            if (mach > p.mach_number_limit) {
                std::stringstream msg;
                msg << "LBMForce: the Mach number of the fluid is " << mach << " after " << stepIndex
                    << " lattice steps, above the limit " << p.mach_number_limit;
                throw OracleError(LBM_ERROR_MACH, msg.str());
            }
        }
    }

    void nextStepForces(int64_t stepCount, const double* positions, const double* velocities, const double* forces) {
        // ReferenceLBMKernels.cpp:391-401.
        stepForcesCurrent = false;
        computeMoments();
        if (p.momentum_removal_frequency > 0 && stepCount % p.momentum_removal_frequency == 0)
            removeMomentum();
        couple(positions, velocities, forces, false, stepCount);
    }

    void computeMoments() {
        // ReferenceLBMKernels.cpp:430-469.  Nodes are independent: with OpenMP the loop is shared among threads,
        // with results independent of their number.
#pragma omp parallel for schedule(static) num_threads(oracleThreads)
        for (int node = 0; node < numNodes; node++) {
            double f[NQ];
            if (!fluid(node)) {
                rho[node] = dev[node] = 0.0;
                for (int k = 0; k < 3; k++)
                    mom[3*(size_t) node + k] = force[3*(size_t) node + k] = 0.0;
                for (int k = 0; k < 6; k++)
                    piNeq[6*(size_t) node + k] = 0.0;
                continue;
            }
            double dr = 0, jx = 0, jy = 0, jz = 0;
            for (int q = 0; q < NQ; q++) {
                f[q] = pop[(size_t) q*numNodes + node];
                dr += f[q];
                jx += CX[q]*f[q];
                jy += CY[q]*f[q];
                jz += CZ[q]*f[q];
            }
            double r = 1.0 + dr;
            rho[node] = r;
            dev[node] = dr;
            mom[3*(size_t) node] = jx;
            mom[3*(size_t) node + 1] = jy;
            mom[3*(size_t) node + 2] = jz;
            piNeqFromDeviation(f, dr, jx, jy, jz, &piNeq[6*(size_t) node]);
            for (int k = 0; k < 3; k++)
                force[3*(size_t) node + k] = r*p.body_acceleration[k];
        }
    }

    void removeMomentum() {
        // ReferenceLBMKernels.cpp:471-501.
        double mass = 0, px = 0, py = 0, pz = 0;
        for (int node = 0; node < numNodes; node++) {
            mass += rho[node];
            px += mom[3*(size_t) node];
            py += mom[3*(size_t) node + 1];
            pz += mom[3*(size_t) node + 2];
        }
        double ux = px/mass, uy = py/mass, uz = pz/mass;
        for (int node = 0; node < numNodes; node++) {
            mom[3*(size_t) node] -= rho[node]*ux;
            mom[3*(size_t) node + 1] -= rho[node]*uy;
            mom[3*(size_t) node + 2] -= rho[node]*uz;
        }
    }

    int nearestNode(const double* x) const {
        // ReferenceLBMKernels.cpp:973-983.
        int size[3] = {p.nx, p.ny, p.nz}, idx[3];
        for (int k = 0; k < 3; k++) {
            double s = x[k]/p.dx;
            s -= std::floor(s/size[k])*size[k];
            idx[k] = ((int) std::floor(s + 0.5)) % size[k];
        }
        return nodeIndex(idx[0], idx[1], idx[2]);
    }

    void stencilNodes(const double* x, std::vector<int>& nodes, std::vector<double>& weights) const {
        // ReferenceLBMKernels.cpp:998-1018.
        int size[3] = {p.nx, p.ny, p.nz};
        int idx[3][4];
        double wt[3][4];
        for (int k = 0; k < 3; k++) {
            double s = x[k]/p.dx;
            s -= std::floor(s/size[k])*size[k];
            axisWeights(p.interpolation_stencil, s, size[k], openAxis(k), idx[k], wt[k]);
        }
        int w = stencilWidth(p.interpolation_stencil);
        nodes.clear();
        weights.clear();
        for (int c = 0; c < w; c++)
            for (int b = 0; b < w; b++)
                for (int a = 0; a < w; a++) {
                    nodes.push_back(nodeIndex(idx[0][a], idx[1][b], idx[2][c]));
                    weights.push_back(wt[0][a]*wt[1][b]*wt[2][c]);
                }
    }

    Vec3d wallNormal(const double* x) const {
        // ReferenceLBMKernels.cpp:1020-1050.
        int size[3] = {p.nx, p.ny, p.nz};
        int idx[3][2];
        double wt[3][2];
        for (int k = 0; k < 3; k++) {
            double s = x[k]/p.dx;
            s -= std::floor(s/size[k])*size[k];
            double lower = std::floor(s);
            idx[k][0] = ((int) lower) % size[k];
            idx[k][1] = (idx[k][0] + 1) % size[k];
            wt[k][1] = s - lower;
            wt[k][0] = 1.0 - wt[k][1];
        }
        double solid[2][2][2];
        for (int a = 0; a < 2; a++)
            for (int b = 0; b < 2; b++)
                for (int c = 0; c < 2; c++)
                    solid[a][b][c] = isFluid[nodeIndex(idx[0][a], idx[1][b], idx[2][c])] ? 0.0 : 1.0;
        Vec3d g;
        for (int a = 0; a < 2; a++)
            for (int b = 0; b < 2; b++) {
                g[0] += wt[1][a]*wt[2][b]*(solid[1][a][b] - solid[0][a][b]);
                g[1] += wt[0][a]*wt[2][b]*(solid[a][1][b] - solid[a][0][b]);
                g[2] += wt[0][a]*wt[1][b]*(solid[a][b][1] - solid[a][b][0]);
            }
        return g;
    }

    void couple(const double* positions, const double* velocities, const double* forces, bool isStep,
            int64_t noiseStep) {
        if (p.drag_scheme == LBM_DRAG_CENTERED) {
            if (isStep)
                std::fill(reaction.begin(), reaction.end(), 0.0);
            if (p.interpolation_stencil == LBM_STENCIL_NEAREST)
                coupleCentered(positions, velocities, forces, isStep, noiseStep);
            else
                coupleCenteredStencil(positions, velocities, forces, isStep, noiseStep);
            if (isStep)
                applyReaction();
            return;
        }
        // Explicit drag (ReferenceLBMKernels.cpp:503-553 nearest, 643-697 stencils).  The random numbers of a step
        // are keyed by the step index, so every evaluation of the same step sees the same ones.
        double gamma = p.friction*p.dt;
        double kT = p.kT*p.dt*p.dt/(cellMass()*p.dx*p.dx);
        double velocityScale = p.dx/p.dt;
        double forceScale = cellMass()*p.dx/(p.dt*p.dt);
        bool randomForce = kT > 0 && gamma > 0;
        if (isStep)
            std::fill(reaction.begin(), reaction.end(), 0.0);
        std::vector<int> nodes;
        std::vector<double> weights;
        for (size_t i = 0; i < particles.size(); i++) {
            int part = particles[i];
            const double* x = positions + 3*(size_t) part;
            if (p.interpolation_stencil == LBM_STENCIL_NEAREST) {
                nodes.assign(1, nearestNode(x));
                weights.assign(1, 1.0);
            }
            else
                stencilNodes(x, nodes, weights);
            double u[3] = {0, 0, 0};
            for (size_t n = 0; n < nodes.size(); n++) {
                int node = nodes[n];
                if (rho[node] > 0)
                    for (int k = 0; k < 3; k++)
                        u[k] += mom[3*(size_t) node + k]*(weights[n]/rho[node]);
            }
            double m = particleMass[i];
            double noise[3] = {0, 0, 0};
            if (randomForce)
                drawNormals(seed, TAG_PARTICLE, (uint32_t) i, noiseStep, 3, noise);
            double f[3];
            for (int k = 0; k < 3; k++) {
                double v = velocities[3*(size_t) part + k]/velocityScale;
                f[k] = (v - u[k])*(-gamma*m);
                if (randomForce)
                    f[k] += noise[k]*std::sqrt(2.0*gamma*m*kT);
                particleForces[i][k] = f[k]*forceScale;
            }
            if (isStep)
                for (size_t n = 0; n < nodes.size(); n++)
                    for (int k = 0; k < 3; k++)
                        reaction[3*(size_t) nodes[n] + k] -= weights[n]*f[k];
        }
        if (isStep)
            applyReaction();
    }

    void applyReaction() {
        // ReferenceLBMKernels.cpp:959-971.
        for (size_t k = 0; k < reaction.size(); k++)
            force[k] += reaction[k];
        double scale = cellMass()*p.dx/p.dt;
        for (int node : solidNodes)
            for (int k = 0; k < 3; k++)
                wallMomentum[k] += reaction[3*(size_t) node + k]*scale;
    }

    /* The particle's velocity at the time of the force before its drag, v~ = v(t - h) + h Fc/m (lattice units), and its
       random force, as both centred schemes use them. */
    void centredKnowns(const double* velocities, const double* forces, int64_t noiseStep, std::vector<Vec3d>& known,
            std::vector<Vec3d>& random) const {
        double gamma = p.friction*p.dt;
        double kT = p.kT*p.dt*p.dt/(cellMass()*p.dx*p.dx);
        double velocityScale = p.dx/p.dt;
        double forceScale = cellMass()*p.dx/(p.dt*p.dt);
        double h = 0.5;
        bool randomForce = kT > 0 && gamma > 0;
        size_t n = particles.size();
        known.assign(n, Vec3d());
        random.assign(n, Vec3d());
        for (size_t i = 0; i < n; i++) {
            int part = particles[i];
            double m = particleMass[i];
            for (int k = 0; k < 3; k++) {
                double fc = forces ? forces[3*(size_t) part + k] : 0.0;
                known[i][k] = velocities[3*(size_t) part + k]*(1.0/velocityScale) + fc*(h/(m*forceScale));
            }
            if (randomForce) {
                double noise[3];
                drawNormals(seed, TAG_PARTICLE, (uint32_t) i, noiseStep, 3, noise);
                for (int k = 0; k < 3; k++)
                    random[i][k] = noise[k]*std::sqrt(2.0*gamma*m*kT);
            }
        }
    }

    void coupleCentered(const double* positions, const double* velocities, const double* forces, bool isStep,
            int64_t noiseStep) {
        // Centred drag at the nearest node (ReferenceLBMKernels.cpp:555-641): with a = gamma h, h = 1/2, for the
        // particles of one node (sums M, P~ = sum m v~, R in particle order) and u~ = (j + h Fbody)/rho,
        //   S = [-gamma (P~ - M u~) + R]/(1 + a + a M/rho),  u_c = u~ - h S/rho,
        //   F_k = [-gamma m_k (v~_k - u_c) + R_k]/(1 + a),
        // the node receiving -S; at a solid node u_c = 0 and S = [-gamma P~ + R]/(1 + a).
        double gamma = p.friction*p.dt, h = 0.5, a = gamma*h;
        double forceScale = cellMass()*p.dx/(p.dt*p.dt);
        std::vector<Vec3d> known, random;
        centredKnowns(velocities, forces, noiseStep, known, random);
        size_t n = particles.size();
        std::vector<long long> keys(n);
        for (size_t i = 0; i < n; i++)
            keys[i] = (long long) nearestNode(positions + 3*(size_t) particles[i])*(long long) n + (long long) i;
        std::sort(keys.begin(), keys.end());
        for (size_t first = 0; first < n; ) {
            int node = (int) (keys[first]/(long long) n);
            size_t end = first;
            double mass = 0;
            Vec3d pm, rsum;
            for (; end < n && keys[end]/(long long) n == node; end++) {
                size_t i = (size_t) (keys[end] % (long long) n);
                mass += particleMass[i];
                for (int k = 0; k < 3; k++) {
                    pm[k] += known[i][k]*particleMass[i];
                    rsum[k] += random[i][k];
                }
            }
            Vec3d u, sum;
            if (rho[node] > 0) {
                for (int k = 0; k < 3; k++) {
                    double knownFluid = (mom[3*(size_t) node + k] + h*force[3*(size_t) node + k])*(1.0/rho[node]);
                    sum[k] = ((pm[k] - knownFluid*mass)*(-gamma) + rsum[k])*(1.0/(1.0 + a + a*mass/rho[node]));
                    u[k] = knownFluid - sum[k]*(h/rho[node]);
                }
            }
            else
                for (int k = 0; k < 3; k++)
                    sum[k] = ((pm[k] - u[k]*mass)*(-gamma) + rsum[k])*(1.0/(1.0 + a));
            for (size_t e = first; e < end; e++) {
                size_t i = (size_t) (keys[e] % (long long) n);
                for (int k = 0; k < 3; k++)
                    particleForces[i][k] = ((known[i][k] - u[k])*(-gamma*particleMass[i]) + random[i][k])*(1.0/(1.0 + a))*
                            forceScale;
            }
            if (isStep)
                for (int k = 0; k < 3; k++)
                    reaction[3*(size_t) node + k] = -sum[k];
            first = end;
        }
    }

    void coupleCenteredStencil(const double* positions, const double* velocities, const double* forces, bool isStep,
            int64_t noiseStep) {
        // Centred drag with a stencil (ReferenceLBMKernels.cpp:726-901): (1 + a)/m_k F_k + a sum_l K_kl F_l = b_k/m_k,
        // K_kl = sum_j xi_jk xi_jl/rho_j, b_k = -gamma m_k (v~_k - sum_j xi_jk u~_j) + R_k, solved for each component by
        // conjugate gradients preconditioned with the diagonal, from the diagonal solution, until the residual is
        // 1e-13 of the right-hand side; the products spread on the nodes and interpolate back.
        double gamma = p.friction*p.dt, h = 0.5, a = gamma*h;
        double forceScale = cellMass()*p.dx/(p.dt*p.dt);
        std::vector<Vec3d> known, random;
        centredKnowns(velocities, forces, noiseStep, known, random);
        size_t n = particles.size();
        std::vector<std::vector<int>> nodes(n);
        std::vector<std::vector<double>> weights(n);
        std::vector<Vec3d> rhs(n);
        std::vector<double> diagonal(n, 0.0);
        for (size_t i = 0; i < n; i++) {
            stencilNodes(positions + 3*(size_t) particles[i], nodes[i], weights[i]);
            double m = particleMass[i];
            Vec3d u;
            double self = 0;
            for (size_t e = 0; e < nodes[i].size(); e++) {
                int node = nodes[i][e];
                if (rho[node] > 0) {
                    for (int k = 0; k < 3; k++)
                        u[k] += (mom[3*(size_t) node + k] + h*force[3*(size_t) node + k])*(weights[i][e]/rho[node]);
                    self += weights[i][e]*weights[i][e]/rho[node];
                }
            }
            for (int k = 0; k < 3; k++)
                rhs[i][k] = ((known[i][k] - u[k])*(-gamma*m) + random[i][k])*(1.0/m);
            diagonal[i] = (1.0 + a)/m + a*self;
        }
        std::vector<double> spread(3*(size_t) numNodes, 0.0);
        auto multiply = [&](const std::vector<Vec3d>& x, std::vector<Vec3d>& result) {
            for (size_t i = 0; i < n; i++)
                for (size_t e = 0; e < nodes[i].size(); e++)
                    for (int k = 0; k < 3; k++)
                        spread[3*(size_t) nodes[i][e] + k] += weights[i][e]*x[i][k];
            for (size_t i = 0; i < n; i++) {
                Vec3d sum;
                for (size_t e = 0; e < nodes[i].size(); e++) {
                    int node = nodes[i][e];
                    if (rho[node] > 0)
                        for (int k = 0; k < 3; k++)
                            sum[k] += spread[3*(size_t) node + k]*(weights[i][e]/rho[node]);
                }
                for (int k = 0; k < 3; k++)
                    result[i][k] = x[i][k]*((1.0 + a)/particleMass[i]) + sum[k]*a;
            }
            for (size_t i = 0; i < n; i++)
                for (size_t e = 0; e < nodes[i].size(); e++)
                    for (int k = 0; k < 3; k++)
                        spread[3*(size_t) nodes[i][e] + k] = 0.0;
        };
        std::vector<Vec3d> F(n), r(n), z(n), d(n), Ad(n);
        Vec3d rz, norm;
        for (size_t i = 0; i < n; i++)
            for (int k = 0; k < 3; k++)
                F[i][k] = rhs[i][k]*(1.0/diagonal[i]);
        multiply(F, Ad);
        for (size_t i = 0; i < n; i++)
            for (int k = 0; k < 3; k++) {
                r[i][k] = rhs[i][k] - Ad[i][k];
                z[i][k] = r[i][k]*(1.0/diagonal[i]);
                d[i][k] = z[i][k];
                rz[k] += r[i][k]*z[i][k];
                norm[k] += rhs[i][k]*rhs[i][k];
            }
        const double tolerance = 1e-13;
        for (int iteration = 0; ; iteration++) {
            Vec3d residual;
            for (size_t i = 0; i < n; i++)
                for (int k = 0; k < 3; k++)
                    residual[k] += r[i][k]*r[i][k];
            bool converged = true;
            for (int k = 0; k < 3; k++)
                if (residual[k] > tolerance*tolerance*norm[k])
                    converged = false;
            if (converged)
                break;
            if (iteration == 1000)
                throw OracleError(LBM_ERROR_RUNTIME, "LBMForce: the centred drag with the interpolation stencil did not "
                        "converge in 1000 iterations of the conjugate gradients");
            multiply(d, Ad);
            Vec3d dAd, rzNew, alpha;
            for (size_t i = 0; i < n; i++)
                for (int k = 0; k < 3; k++)
                    dAd[k] += d[i][k]*Ad[i][k];
            for (int k = 0; k < 3; k++)
                alpha[k] = dAd[k] > 0 ? rz[k]/dAd[k] : 0.0;
            for (size_t i = 0; i < n; i++)
                for (int k = 0; k < 3; k++) {
                    F[i][k] += alpha[k]*d[i][k];
                    r[i][k] -= alpha[k]*Ad[i][k];
                    z[i][k] = r[i][k]*(1.0/diagonal[i]);
                    rzNew[k] += r[i][k]*z[i][k];
                }
            for (size_t i = 0; i < n; i++)
                for (int k = 0; k < 3; k++)
                    d[i][k] = z[i][k] + (rz[k] > 0 ? rzNew[k]/rz[k] : 0.0)*d[i][k];
            rz = rzNew;
        }
        for (size_t i = 0; i < n; i++) {
            for (int k = 0; k < 3; k++)
                particleForces[i][k] = F[i][k]*forceScale;
            if (isStep)
                for (size_t e = 0; e < nodes[i].size(); e++)
                    for (int k = 0; k < 3; k++)
                        reaction[3*(size_t) nodes[i][e] + k] -= weights[i][e]*F[i][k];
        }
    }

    void collideAndStream() {
        // ReferenceLBMKernels.cpp:1066-1120.
        bool fluct = p.fluid_fluctuations && p.fluid_kT > 0;
        double mu = 3.0*p.fluid_kT*p.dt*p.dt/(cellMass()*p.dx*p.dx);
        double omega = p.omega;
        int nx = p.nx, ny = p.ny, nz = p.nz;
        // Each population slot is written by exactly one node and none is read (the moments are stored): the nodes
        // are shared among OpenMP threads, with results independent of their number.
#pragma omp parallel for schedule(static) num_threads(oracleThreads)
        for (int node = 0; node < numNodes; node++) {
            double eq[NQ], neq[NQ], src[NQ], xi[NQ], normal[16];
            if (!fluid(node))
                continue;
            int i = node % nx, j = (node/nx) % ny, k = node/(nx*ny);
            double r = rho[node];
            const double* F = &force[3*(size_t) node];
            double ux = (mom[3*(size_t) node] + 0.5*F[0])/r;
            double uy = (mom[3*(size_t) node + 1] + 0.5*F[1])/r;
            double uz = (mom[3*(size_t) node + 2] + 0.5*F[2])/r;
            equilibriumDeviation(dev[node], ux, uy, uz, eq);
            regularizedNeq(&piNeq[6*(size_t) node], neq);
            guoSource(ux, uy, uz, F[0], F[1], F[2], src);
            if (fluct) {
                drawNormals(seed, TAG_FLUID, (uint32_t) node, stepIndex, 16, normal);
                collisionNoise(normal, std::sqrt(mu*r*omega*(2.0 - omega)), std::sqrt(mu*r), xi);
                if (!boundaryOf.empty() && boundaryOf[node] >= 0)
                    collisionNoise(normal, std::sqrt(mu*r*omega*(2.0 - omega)), 0.0,
                            &boundaryStress[(size_t) NQ*boundaryOf[node]]);
            }
            for (int q = 0; q < NQ; q++) {
                int di = (i + CX[q] + nx) % nx, dj = (j + CY[q] + ny) % ny, dk = (k + CZ[q] + nz) % nz;
                double value = eq[q] + (1.0 - omega)*neq[q] + 0.5*src[q];
                if (fluct)
                    value += xi[q];
                pop[(size_t) q*numNodes + nodeIndex(di, dj, dk)] = value;
            }
        }
    }

    void bounceBack() {
        // ReferenceLBMKernels.cpp:1156-1174.
        int nx = p.nx, ny = p.ny, nz = p.nz;
        for (size_t b = 0; b < wallNodes.size(); b++) {
            int node = wallNodes[b];
            int i = node % nx, j = (node/nx) % ny, k = node/(nx*ny);
            for (int q = 1; q < NQ; q++) {
                if (!(wallLinks[b] & (1 << q)))
                    continue;
                int solid = nodeIndex((i + CX[q] + nx) % nx, (j + CY[q] + ny) % ny, (k + CZ[q] + nz) % nz);
                pop[(size_t) OPP[q]*numNodes + node] = pop[(size_t) q*numNodes + solid];
            }
        }
    }

    void wallExchange() {
        // ReferenceLBMKernels.cpp:1176-1205.
        int nx = p.nx, ny = p.ny, nz = p.nz;
        Vec3d ex;
        for (int node : solidNodes) {
            int i = node % nx, j = (node/nx) % ny, k = node/(nx*ny);
            for (int q = 1; q < NQ; q++) {
                int di = i + CX[q], dj = j + CY[q], dk = k + CZ[q];
                if ((openAxis(0) && (di < 0 || di >= nx)) || (openAxis(1) && (dj < 0 || dj >= ny)) ||
                        (openAxis(2) && (dk < 0 || dk >= nz)))
                    continue;
                int target = nodeIndex((di + nx) % nx, (dj + ny) % ny, (dk + nz) % nz);
                if (!isFluid[target])
                    continue;
                double df = pop[(size_t) OPP[q]*numNodes + node];
                ex[0] -= CX[q]*2.0*(df + W[q]);
                ex[1] -= CY[q]*2.0*(df + W[q]);
                ex[2] -= CZ[q]*2.0*(df + W[q]);
            }
        }
        double scale = cellMass()*p.dx/p.dt;
        for (int k = 0; k < 3; k++)
            wallMomentum[k] += ex[k]*scale;
    }

    void applyBoundaries() {
        // ReferenceLBMKernels.cpp:1207-1343.
        int nx = p.nx, ny = p.ny, nz = p.nz;
        const double* g = p.body_acceleration;
        double omega = p.omega;
        bool fluct = p.fluid_fluctuations && p.fluid_kT > 0;
        double mu = 3.0*p.fluid_kT*p.dt*p.dt/(cellMass()*p.dx*p.dx);
        double eq[NQ], neq[NQ], src[NQ], xi[NQ], normal[20];
        double c = boundaryStressWeight(omega);
        Vec3d ex;
        for (size_t b = 0; b < bNodes.size(); b++) {
            int node = bNodes[b], unknown = bUnknown[b], solid = bSolid[b], kind = bKind[b], face = bFace[b];
            int i = node % nx, j = (node/nx) % ny, k = node/(nx*ny);
            double rb = rho[node], dr = dev[node];
            double ub[3];
            for (int a = 0; a < 3; a++)
                ub[a] = (mom[3*(size_t) node + a] + 0.5*(rb*g[a]))/rb;
            if (kind == KIND_WALL || kind == KIND_DENSITY_AT_REST)
                ub[0] = ub[1] = ub[2] = 0.0;
            if (kind == KIND_VELOCITY || kind == KIND_DENSITY_VELOCITY)
                for (int a = 0; a < 3; a++)
                    ub[a] = p.face_velocity[face][a];
            if (kind == KIND_DENSITY || kind == KIND_DENSITY_AT_REST || kind == KIND_DENSITY_VELOCITY) {
                dr = p.face_density[face] - 1.0;
                rb = p.face_density[face];
            }
            if (kind == KIND_DENSITY) {
                int axis = face/2;
                double inward = face % 2 == 0 ? 1.0 : -1.0, sum = 0.0;
                for (int q = 0; q < NQ; q++)
                    sum += pop[(size_t) ((unknown & (1 << q)) ? OPP[q] : q)*numNodes + node];
                ub[axis] = 0.5*(inward*(dr - sum)/rb + ub[axis]);
            }
            double Fb[3] = {g[0]*rb, g[1]*rb, g[2]*rb};
            equilibriumDeviation(dr, ub[0], ub[1], ub[2], eq);
            regularizedNeq(&piNeq[6*(size_t) node], neq);
            guoSource(ub[0], ub[1], ub[2], Fb[0], Fb[1], Fb[2], src);
            if (fluct && hasEquilibriumNoise(kind)) {
                drawNormals(seed, TAG_BOUNDARY, (uint32_t) node, stepIndex, 19, normal);
                equilibriumNoise(normal, std::sqrt(mu*rb), std::sqrt(mu*rb*std::max(0.0, 1.0 - c)), std::sqrt(mu*rb), xi);
                for (int q = 0; q < NQ; q++)
                    xi[q] -= c*boundaryStress[(size_t) NQ*b + q];
            }
            else if (fluct) {
                drawNormals(seed, TAG_BOUNDARY, (uint32_t) node, stepIndex, 15, normal);
                collisionNoise(normal, std::sqrt(mu*rb*omega*(2.0 - omega)), std::sqrt(mu*rb), xi);
            }
            if (kind == KIND_WALL || kind == KIND_VELOCITY) {
                double e0[NQ], sg[NQ];
                equilibriumDeviation(0.0, ub[0], ub[1], ub[2], e0);
                guoSource(ub[0], ub[1], ub[2], g[0], g[1], g[2], sg);
                double num = 0.0, den = 0.0;
                for (int q = 1; q < NQ; q++) {
                    if (!(unknown & (1 << q)))
                        continue;
                    double arriving, flux = 0.0;
                    if (solid & (1 << q)) {
                        int source = nodeIndex((i - CX[q] + nx) % nx, (j - CY[q] + ny) % ny, (k - CZ[q] + nz) % nz);
                        arriving = pop[(size_t) OPP[q]*numNodes + source];
                    }
                    else if (!(unknown & (1 << OPP[q]))) {
                        arriving = pop[(size_t) OPP[q]*numNodes + node];
                        flux = 6.0*W[q]*(CX[q]*ub[0] + CY[q]*ub[1] + CZ[q]*ub[2]);
                    }
                    else
                        continue;
                    double rest = (1.0 - omega)*neq[q] + (fluct ? xi[q] : 0.0);
                    num += arriving - rest - e0[q] - 0.5*sg[q] + flux;
                    den += W[q] + e0[q] + 0.5*sg[q] - flux;
                }
                if (den != 0) {
                    dr = num/den;
                    rb = 1.0 + dr;
                    for (int a = 0; a < 3; a++)
                        Fb[a] = g[a]*rb;
                    equilibriumDeviation(dr, ub[0], ub[1], ub[2], eq);
                    guoSource(ub[0], ub[1], ub[2], Fb[0], Fb[1], Fb[2], src);
                }
            }
            for (int q = 1; q < NQ; q++) {
                if (!(unknown & (1 << q)))
                    continue;
                double value = eq[q] + (1.0 - omega)*neq[q] + 0.5*src[q];
                if (fluct)
                    value += xi[q];
                if (solid & (1 << q)) {
                    int source = nodeIndex((i - CX[q] + nx) % nx, (j - CY[q] + ny) % ny, (k - CZ[q] + nz) % nz);
                    double sent = pop[(size_t) OPP[q]*numNodes + source];
                    double t = sent + value + 2.0*W[q];
                    ex[0] -= CX[q]*t;
                    ex[1] -= CY[q]*t;
                    ex[2] -= CZ[q]*t;
                }
                pop[(size_t) q*numNodes + node] = value;
            }
        }
        double scale = cellMass()*p.dx/p.dt;
        for (int k = 0; k < 3; k++)
            wallMomentum[k] += ex[k]*scale;
    }

    void nodeFields(int node, double& density, double u[3]) const {
        // ReferenceLBMKernels.cpp:1349-1368.
        if (!fluid(node)) {
            density = 0;
            u[0] = u[1] = u[2] = 0;
            return;
        }
        double r = 1.0, jx = 0, jy = 0, jz = 0;
        for (int q = 0; q < NQ; q++) {
            double f = pop[(size_t) q*numNodes + node];
            r += f;
            jx += CX[q]*f;
            jy += CY[q]*f;
            jz += CZ[q]*f;
        }
        density = r;
        u[0] = jx*(1.0/r) + p.body_acceleration[0]*0.5;
        u[1] = jy*(1.0/r) + p.body_acceleration[1]*0.5;
        u[2] = jz*(1.0/r) + p.body_acceleration[2]*0.5;
    }

    double machNumber() const {
        // ReferenceLBMKernels.cpp:1449-1468.
        double maxSpeed2 = 0.0;
        for (int node = 0; node < numNodes; node++) {
            if (!fluid(node))
                continue;
            double r = 1.0, jx = 0, jy = 0, jz = 0;
            for (int q = 0; q < NQ; q++) {
                double f = pop[(size_t) q*numNodes + node];
                r += f;
                jx += CX[q]*f;
                jy += CY[q]*f;
                jz += CZ[q]*f;
            }
            maxSpeed2 = std::max(maxSpeed2, (jx*jx + jy*jy + jz*jz)/(r*r));
        }
        return std::sqrt(3.0*maxSpeed2);
    }
};

// The kernel's checkpoint of the GPU library (version 2: with the seed of the random numbers; version 1 without).
struct CheckpointHeader {
    char magic[8];
    int32_t version, nx, ny, nz, numParticles, dragScheme;
    uint32_t seed;
};
constexpr size_t CHECKPOINT_V1_HEADER = 8 + 6*sizeof(int32_t);

int fail(const OracleError& e) {
    g_lastError = e.what();
    return e.status;
}

}  // namespace

struct lbm_context {
    OracleKernel kernel;
};

#define ORACLE_TRY(...) \
    try { __VA_ARGS__; g_lastError.clear(); return LBM_OK; } \
    catch (const OracleError& e) { return fail(e); } \
    catch (const std::exception& e) { g_lastError = e.what(); return LBM_ERROR_RUNTIME; }

extern "C" {

const char* lbm_last_error(void) {
    return g_lastError.c_str();
}

int lbm_create(const lbm_params* params, const int* solid, int numSolid, const int* parts, const double* masses,
        int numParts, int numSys, lbm_context** out) {
    ORACLE_TRY({
        if (!params || !out)
            throw OracleError(LBM_ERROR_INVALID, "lbm_create: null argument");
        lbm_context* ctx = new lbm_context();
        try {
            ctx->kernel.initialize(*params, solid, numSolid, parts, masses, numParts, numSys);
        }
        catch (...) {
            delete ctx;
            throw;
        }
        *out = ctx;
    })
}

int lbm_destroy(lbm_context* ctx) {
    delete ctx;
    return LBM_OK;
}

int lbm_begin_step(lbm_context* ctx, int64_t step, const double* positions, double* velocities) {
    ORACLE_TRY(ctx->kernel.beginStep(step, positions, velocities))
}

int lbm_execute(lbm_context* ctx, int64_t step, int includeForces, const double* positions, const double* velocities,
        double* forces) {
    ORACLE_TRY(ctx->kernel.execute(step, includeForces != 0, positions, velocities, forces))
}

int lbm_begin_step_device(lbm_context*, int64_t, const double*, double*) {
    g_lastError = "the oracle has no device path";
    return LBM_ERROR_UNSUPPORTED;
}

int lbm_execute_device(lbm_context*, int64_t, int, const double*, const double*, double*) {
    g_lastError = "the oracle has no device path";
    return LBM_ERROR_UNSUPPORTED;
}

int lbm_copy_parameters(lbm_context* ctx, const lbm_params* params) {
    // LBMForceImpl::updateParametersInContext (LBMForceImpl.cpp:388-414): the checks, then copyParametersToContext
    // (ReferenceLBMKernels.cpp:1345-1347).
    ORACLE_TRY({
        const lbm_params& o = ctx->kernel.p;
        const lbm_params& u = *params;
        auto refuse = [](const char* what) {
            throw OracleError(LBM_ERROR_INVALID, std::string("updateParametersInContext: ") + what);
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
        int storage = ctx->kernel.p.storage;
        ctx->kernel.p = *params;
        ctx->kernel.p.storage = storage;
    })
}

int lbm_get_fluid_fields(lbm_context* ctx, double* density, double* velocity, int halo) {
    // ReferenceLBMKernels.cpp:1403-1436 with the nodes of LBMDecomposition::getDomainNodes (LBMDecomposition.cpp:
    // 183-204) for one domain: with halo, the lattice padded by one node, a padded node wrapped onto the lattice, or
    // -1 beyond an open face; outside the lattice a field is valid only if its halo is exchanged.
    ORACLE_TRY({
        OracleKernel& k = ctx->kernel;
        const double nan = std::numeric_limits<double>::quiet_NaN();
        int pad = halo ? 1 : 0;
        int n[3] = {k.p.nx, k.p.ny, k.p.nz};
        size_t l = 0;
        for (int kk = -pad; kk < n[2] + pad; kk++)
            for (int jj = -pad; jj < n[1] + pad; jj++)
                for (int ii = -pad; ii < n[0] + pad; ii++, l++) {
                    int c[3] = {ii, jj, kk};
                    bool inside = true, beyondFace = false;
                    for (int a = 0; a < 3; a++) {
                        inside = inside && c[a] >= 0 && c[a] < n[a];
                        if (c[a] < 0 || c[a] >= n[a]) {
                            beyondFace = beyondFace || k.openAxis(a);
                            c[a] = (c[a] + n[a])%n[a];
                        }
                    }
                    int node = beyondFace ? -1 : c[0] + n[0]*(c[1] + n[1]*c[2]);
                    bool densityValid = node >= 0 && (inside || k.p.density_halo_exchange);
                    bool velocityValid = node >= 0 && (inside || k.p.velocity_halo_exchange);
                    double r = nan, u[3] = {nan, nan, nan};
                    if (densityValid || velocityValid)
                        k.nodeFields(node, r, u);
                    if (density)
                        density[l] = densityValid ? r*k.p.density : nan;
                    if (velocity)
                        for (int a = 0; a < 3; a++)
                            velocity[3*l + a] = velocityValid ? u[a]*(k.p.dx/k.p.dt) : nan;
                }
    })
}

/* Threads of the oracle's node loops (OpenMP), 1 unless set; n <= 0 leaves them.  Returns the number in use. */
int lbm_oracle_threads(int n) {
#ifdef _OPENMP
    if (n > 0)
        oracleThreads = n;
    return oracleThreads;
#else
    (void) n;
    return 1;
#endif
}

int lbm_get_particle_nodes(lbm_context* ctx, const double* positions, int* nodes) {
    ORACLE_TRY({
        for (size_t i = 0; i < ctx->kernel.particles.size(); i++)
            nodes[i] = ctx->kernel.nearestNode(positions + 3*(size_t) ctx->kernel.particles[i]);
    })
}

int lbm_get_fluid_mach_number(lbm_context* ctx, double* mach) {
    ORACLE_TRY(*mach = ctx->kernel.machNumber())
}

int lbm_get_wall_force(lbm_context* ctx, double f[3]) {
    ORACLE_TRY({ for (int k = 0; k < 3; k++) f[k] = ctx->kernel.wallMomentum[k]/ctx->kernel.p.dt; })
}

int lbm_get_fluid_state(lbm_context* ctx, double* state, size_t size) {
    ORACLE_TRY({
        if (size != ctx->kernel.pop.size())
            throw OracleError(LBM_ERROR_INVALID, "LBMForce: getFluidState() was called with a state of the wrong size");
        std::memcpy(state, ctx->kernel.pop.data(), size*sizeof(double));
    })
}

int lbm_set_fluid_state(lbm_context* ctx, const double* state, size_t size) {
    ORACLE_TRY({
        if (size != ctx->kernel.pop.size())
            throw OracleError(LBM_ERROR_INVALID, "LBMForce: setFluidState() was called with a state of the wrong size");
        std::memcpy(ctx->kernel.pop.data(), state, size*sizeof(double));
        ctx->kernel.stepForcesCurrent = false;
    })
}

int lbm_checkpoint_size(lbm_context* ctx, size_t* size) {
    ORACLE_TRY(*size = sizeof(CheckpointHeader) + ctx->kernel.pop.size()*sizeof(double) + 3*sizeof(double))
}

int lbm_create_checkpoint(lbm_context* ctx, void* buffer, size_t size) {
    ORACLE_TRY({
        OracleKernel& k = ctx->kernel;
        size_t need = sizeof(CheckpointHeader) + k.pop.size()*sizeof(double) + 3*sizeof(double);
        if (size < need)
            throw OracleError(LBM_ERROR_INVALID, "lbm_create_checkpoint: buffer too small");
        CheckpointHeader h;
        std::memcpy(h.magic, "LBMAMDC1", 8);
        h.version = 2; h.nx = k.p.nx; h.ny = k.p.ny; h.nz = k.p.nz; h.numParticles = (int) k.particles.size(); h.dragScheme = k.p.drag_scheme;
        h.seed = k.seed;
        char* out = (char*) buffer;
        std::memcpy(out, &h, sizeof(h));
        std::memcpy(out + sizeof(h), k.pop.data(), k.pop.size()*sizeof(double));
        std::memcpy(out + sizeof(h) + k.pop.size()*sizeof(double), k.wallMomentum.x, 3*sizeof(double));
    })
}

int lbm_load_checkpoint(lbm_context* ctx, const void* buffer, size_t size) {
    ORACLE_TRY({
        OracleKernel& k = ctx->kernel;
        CheckpointHeader h;
        if (size < CHECKPOINT_V1_HEADER)
            throw OracleError(LBM_ERROR_INVALID, "LBMForce: the checkpoint is truncated");
        std::memcpy(&h, buffer, CHECKPOINT_V1_HEADER);
        if (std::memcmp(h.magic, "LBMAMDC1", 8) != 0 || (h.version != 1 && h.version != 2))
            throw OracleError(LBM_ERROR_INVALID, "LBMForce: the data is not a checkpoint of this LBMForce platform");
        size_t headerBytes = h.version == 1 ? CHECKPOINT_V1_HEADER : sizeof(h);
        if (size < headerBytes)
            throw OracleError(LBM_ERROR_INVALID, "LBMForce: the checkpoint is truncated");
        std::memcpy(&h, buffer, headerBytes);
        size_t need = headerBytes + k.pop.size()*sizeof(double) + 3*sizeof(double);
        if (h.nx != k.p.nx || h.ny != k.p.ny || h.nz != k.p.nz || h.numParticles != (int) k.particles.size())
            throw OracleError(LBM_ERROR_INVALID, "LBMForce: the checkpoint was written for another lattice or System");
        if (h.dragScheme != k.p.drag_scheme)
            throw OracleError(LBM_ERROR_INVALID, "LBMForce: the checkpoint was written with a different drag scheme");
        if (size < need)
            throw OracleError(LBM_ERROR_INVALID, "LBMForce: the checkpoint is truncated");
        const char* in = (const char*) buffer;
        std::memcpy(k.pop.data(), in + headerBytes, k.pop.size()*sizeof(double));
        std::memcpy(k.wallMomentum.x, in + headerBytes + k.pop.size()*sizeof(double), 3*sizeof(double));
        if (h.version >= 2)
            k.seed = h.seed;
        k.stepForcesCurrent = false;
    })
}

int lbm_get_fluid_totals(lbm_context* ctx, double* mass, double momentum[3]) {
    ORACLE_TRY({
        OracleKernel& k = ctx->kernel;
        double m = 0, px = 0, py = 0, pz = 0;
        for (int node = 0; node < k.numNodes; node++) {
            if (!k.fluid(node))
                continue;
            m += 1.0;
            for (int q = 0; q < NQ; q++) {
                double f = k.pop[(size_t) q*k.numNodes + node];
                m += f;
                px += CX[q]*f;
                py += CY[q]*f;
                pz += CZ[q]*f;
            }
        }
        *mass = m;
        momentum[0] = px; momentum[1] = py; momentum[2] = pz;
    })
}

int lbm_synchronize(lbm_context*) {
    return LBM_OK;
}

int lbm_profile_fused(lbm_context*, int, double*, int*) {
    g_lastError = "the oracle has no device path";
    return LBM_ERROR_UNSUPPORTED;
}

/* Test hooks of the oracle (not part of the C-ABI of the GPU library). */
void oracle_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    philox4x32(ctr, key, out);
}

void oracle_normals(uint32_t seed, uint32_t tag, uint32_t index, int64_t step, int count, double* out) {
    drawNormals(seed, tag, index, step, count, out);
}

void oracle_axis_weights(int stencil, double s, int n, int open, int* index, double* weight) {
    axisWeights(stencil, s, n, open != 0, index, weight);
}

void oracle_equilibrium_deviation(double dr, double ux, double uy, double uz, double* out) {
    equilibriumDeviation(dr, ux, uy, uz, out);
}

void oracle_pi_neq(const double* df, double dr, double jx, double jy, double jz, double* pi) {
    piNeqFromDeviation(df, dr, jx, jy, jz, pi);
}

void oracle_regularized_neq(const double* pi, double* out) {
    regularizedNeq(pi, out);
}

void oracle_guo(double ux, double uy, double uz, double fx, double fy, double fz, double* out) {
    guoSource(ux, uy, uz, fx, fy, fz, out);
}

double oracle_mode(int k, int q) {
    return basisMode(k, q);
}

void oracle_collision_noise(const double* r, double stressAmp, double ghostAmp, double* xi) {
    collisionNoise(r, stressAmp, ghostAmp, xi);
}

void oracle_equilibrium_noise(const double* r, double c, double s, double g, double* xi) {
    equilibriumNoise(r, c, s, g, xi);
}

int oracle_num_boundary_nodes(lbm_context* ctx) {
    return (int) ctx->kernel.bNodes.size();
}

/* The kind (internal/LBMBoundaries.h:30) and the face of a boundary node; 0 if the node is not one. */
int oracle_boundary_node(lbm_context* ctx, int node, int* kind, int* face) {
    const auto& k = ctx->kernel;
    for (size_t b = 0; b < k.bNodes.size(); b++)
        if (k.bNodes[b] == node) {
            *kind = k.bKind[b];
            *face = k.bFace[b];
            return 1;
        }
    return 0;
}

uint32_t oracle_seed(lbm_context* ctx) {
    return ctx->kernel.seed;
}

}  // extern "C"
