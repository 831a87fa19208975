/*
 * TEST INFRASTRUCTURE: a C entry-point harness around the reference's own D3Q19 algebra
 * (openmmapi/include/internal/D3Q19.h, which depends only on <cmath>), compiled from the
 * reference tree by oracle/Makefile into oracle/_ref/libref_d3q19.so.  tests/test_oracle.py
 * (test_algebra_matches_reference_header) compares the oracle's restatement of each function with
 * it on random node states, and tests/golden/make_d3q19_vectors.py writes fixtures from it.
 * Nothing else loads it; it is built only where the reference tree is present.
 */
#include "internal/D3Q19.h"

using LBMPlugin::D3Q19;

extern "C" {

void ref_equilibrium_deviation(double dr, double ux, double uy, double uz, double* out) {
    D3Q19::equilibriumDeviation(dr, ux, uy, uz, out);
}

void ref_pi_neq(const double* df, double dr, double jx, double jy, double jz, double* pi) {
    D3Q19::nonEquilibriumMomentFromDeviation(df, dr, jx, jy, jz, pi);
}

void ref_regularized_neq(const double* pi, double* out) {
    D3Q19::regularizedNonEquilibrium(pi, out);
}

void ref_guo(double ux, double uy, double uz, double fx, double fy, double fz, double* out) {
    D3Q19::guoForcing(ux, uy, uz, fx, fy, fz, out);
}

double ref_mode(int k, int q) {
    return D3Q19::mode(k, q);
}

void ref_collision_noise(const double* r, double stressAmp, double ghostAmp, double* xi) {
    D3Q19::fluctuation(r, stressAmp, ghostAmp, xi);
}

void ref_equilibrium_noise(const double* r, double c, double s, double g, double* xi) {
    D3Q19::equilibriumFluctuation(r, c, s, g, xi);
}

double ref_boundary_stress_weight(double omega) {
    return D3Q19::boundaryStressWeight(omega);
}

}
