/*
 * Device-side building blocks of the gfx950 LBMForce platform: the D3Q19 velocity set, the
 * moments of a node, the regularized collision with Guo forcing written in its closed
 * 10-moment Hermite form, the fluctuating part, Philox4x32-10 normal numbers, and the
 * conversions of the population storage types.
 *
 * The collision is the one of ReferenceLBMKernels.cpp:1090-1120 (D3Q19.h:55-134): the
 * post-collision population
 *   C_q = feq_q(rho, u) + (1 - omega) fneq_q(Pi_neq) + S_q(u, F)/2,   u = (j + F/2)/rho,
 * is a second-order Hermite expansion in c_q, so it is computed exactly as
 *   C_q - w_q = w_q [drho + 3 c_q.(j + F) + 4.5 H2(c_q) : P],
 *   P = rho u u + (1 - omega) Pi_neq + (u F + F u)/2,
 * and Pi_neq = sum_q c c df_q - I drho/3 - j j/rho (the same moment as D3Q19.h:98-110, since
 * sum_q H2 dfeq_q = rho u u on D3Q19).  Only the rounding differs from the reference.
 */
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <hip/hip_bf16.h>
#include <stdint.h>
#include <utility>
#include <type_traits>

namespace lbm {

constexpr int NQ = 19;
constexpr int CX[NQ] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1};
constexpr int CY[NQ] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 1, -1, 1, -1, 0, 0, 0, 0};
constexpr int CZ[NQ] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, 1, -1};
constexpr int OPP[NQ] = {0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17};
constexpr double WQ[NQ] = {1.0/3.0, 1.0/18.0, 1.0/18.0, 1.0/18.0, 1.0/18.0, 1.0/18.0, 1.0/18.0,
        1.0/36.0, 1.0/36.0, 1.0/36.0, 1.0/36.0, 1.0/36.0, 1.0/36.0, 1.0/36.0, 1.0/36.0, 1.0/36.0,
        1.0/36.0, 1.0/36.0, 1.0/36.0};

/* 1/sqrt(b_k) of the orthogonal basis of Lulli et al. (norms in D3Q19.h:136-137). */
constexpr double INV_SQRT_NORM[NQ] = {1.0, 1.7320508075688772, 1.7320508075688772, 1.7320508075688772,
        2.1213203435596424, 2.1213203435596424, 2.1213203435596424, 3.0, 3.0, 3.0,
        3.6742346141747673, 3.6742346141747673, 3.6742346141747673, 4.242640687119285, 4.242640687119285,
        4.242640687119285, 4.810702354423639, 5.019960159204453, 5.477225575051661};

/* The basis polynomial e_k evaluated at c_q (D3Q19.h:145-169), as a constant expression. */
constexpr double basisMode(int k, int q) {
    double x = CX[q], y = CY[q], z = CZ[q];
    double x2 = x*x, y2 = y*y, z2 = z*z;
    return k == 0 ? 1.0 : k == 1 ? x : k == 2 ? y : k == 3 ? z :
           k == 4 ? x2 - 1.0/3.0 : k == 5 ? y2 - 1.0/3.0 : k == 6 ? z2 - 1.0/3.0 :
           k == 7 ? x*y : k == 8 ? x*z : k == 9 ? y*z :
           k == 10 ? y*(x2 - 1.0/3.0) : k == 11 ? z*(x2 - 1.0/3.0) : k == 12 ? x*(y2 - 1.0/3.0) :
           k == 13 ? 0.5*x*(y2 + 2.0*z2 - 1.0) : k == 14 ? 0.5*y*(x2 + 2.0*z2 - 1.0) :
           k == 15 ? 0.5*z*(x2 + 2.0*y2 - 1.0) :
           k == 16 ? x2*y2 - x2/3.0 - y2/3.0 + z2/6.0 + 1.0/18.0 :
           k == 17 ? (2.0/7.0)*x2*y2 + x2*z2 - (3.0/7.0)*x2 + y2/14.0 - (2.0/7.0)*z2 + 1.0/14.0 :
           0.4*x2*y2 + 0.4*x2*z2 + y2*z2 - 0.1*x2 - 0.4*y2 - 0.4*z2 + 0.1;
}

/* Compile-time loop: f(std::integral_constant<int, i>) for i = 0..N-1. */
template <typename F, int... Is>
__device__ __forceinline__ void staticForImpl(F&& f, std::integer_sequence<int, Is...>) {
    (f(std::integral_constant<int, Is>{}), ...);
}
template <int N, typename F>
__device__ __forceinline__ void staticFor(F&& f) {
    staticForImpl(f, std::make_integer_sequence<int, N>{});
}

/* c*x for a constant integer c in {-1, 0, 1}, added to acc without ever multiplying by 0. */
template <int C, typename R>
__device__ __forceinline__ R addTerm(R acc, R x) {
    if constexpr (C == 1)
        return acc + x;
    else if constexpr (C == -1)
        return acc - x;
    else
        return acc;
}

/*
 * w_q*v.  In FP32 the weights 1/3, 1/18, 1/36 round up by 3.0e-8 and 7.4e-9 relative: multiplied in
 * every collision, that bias would scale the momentum by (1 + 7.4e-9) per step, 7e-6 after 1000
 * steps.  The weight is split into w_hi = float(w) and w_lo = float(w - w_hi), and w*v is formed
 * as fma(v, w_hi, v*w_lo), exact to about 1e-15 relative before the final rounding.
 */
template <typename R, bool SPLIT>
__device__ __forceinline__ R weighted(int q, R v) {
    if constexpr (std::is_same<R, float>::value && SPLIT) {
        constexpr float hi0 = (float) (1.0/3.0), hi1 = (float) (1.0/18.0), hi2 = (float) (1.0/36.0);
        constexpr float lo0 = (float) (1.0/3.0 - (double) hi0), lo1 = (float) (1.0/18.0 - (double) hi1),
                lo2 = (float) (1.0/36.0 - (double) hi2);
        float hi = q == 0 ? hi0 : (q < 7 ? hi1 : hi2), lo = q == 0 ? lo0 : (q < 7 ? lo1 : lo2);
        return fmaf(v, hi, v*lo);
    }
    else
        return R(WQ[q])*v;
}

/* Moments of one node from its 19 deviations: drho, j, and the raw second moments sum c c df. */
template <typename R>
struct Moments {
    R dr, jx, jy, jz, sxx, syy, szz, sxy, sxz, syz;
};

/* The directions come in opposite pairs (2p+1, 2p+2).  With s_p = f_q + f_opp and d_p = f_q - f_opp, the odd
   moments take d_p times c_q and the even ones s_p times c_q c_q: 51 additions instead of 90. */
constexpr bool oppositePairs() {
    for (int p = 0; p < 9; p++)
        if (OPP[2*p + 1] != 2*p + 2 || CX[2*p + 1] != -CX[2*p + 2] || CY[2*p + 1] != -CY[2*p + 2] ||
            CZ[2*p + 1] != -CZ[2*p + 2])
            return false;
    return true;
}
static_assert(oppositePairs(), "the velocity set is not ordered in opposite pairs");

template <typename R>
__device__ __forceinline__ Moments<R> nodeMoments(const R f[NQ]) {
    R s[9], d[9];
#pragma unroll
    for (int p = 0; p < 9; p++) {
        s[p] = f[2*p + 1] + f[2*p + 2];
        d[p] = f[2*p + 1] - f[2*p + 2];
    }
    // Pairs: 0 x, 1 y, 2 z, 3 (1,1,0), 4 (1,-1,0), 5 (0,1,1), 6 (0,1,-1), 7 (1,0,1), 8 (-1,0,1).
    static_assert(CX[7] == 1 && CY[7] == 1 && CX[9] == 1 && CY[9] == -1 && CY[11] == 1 && CZ[11] == 1 &&
                  CY[13] == 1 && CZ[13] == -1 && CX[15] == 1 && CZ[15] == 1 && CX[17] == -1 && CZ[17] == 1,
                  "pair table of nodeMoments");
    Moments<R> m;
    m.dr = f[0] + (((s[0] + s[1]) + (s[2] + s[3])) + ((s[4] + s[5]) + (s[6] + s[7])) + s[8]);
    m.jx = d[0] + ((d[3] + d[4]) + (d[7] - d[8]));
    m.jy = d[1] + ((d[3] - d[4]) + (d[5] + d[6]));
    m.jz = d[2] + ((d[5] - d[6]) + (d[7] + d[8]));
    R xy = s[3] + s[4], yz = s[5] + s[6], xz = s[7] + s[8];   // the diagonal pairs of each coordinate plane
    m.sxx = s[0] + (xy + xz);
    m.syy = s[1] + (xy + yz);
    m.szz = s[2] + (yz + xz);
    m.sxy = s[3] - s[4];
    m.syz = s[5] - s[6];
    m.sxz = s[7] - s[8];
    return m;
}

/* Pi_neq (xx, yy, zz, xy, xz, yz) with the equilibrium at j/rho (D3Q19.h:98-110). */
template <typename R>
__device__ __forceinline__ void piNeq(const Moments<R>& m, R pi[6]) {
    R rho = R(1) + m.dr;
    R inv = R(1)/rho;
    R third = m.dr*R(1.0/3.0);
    pi[0] = m.sxx - third - m.jx*m.jx*inv;
    pi[1] = m.syy - third - m.jy*m.jy*inv;
    pi[2] = m.szz - third - m.jz*m.jz*inv;
    pi[3] = m.sxy - m.jx*m.jy*inv;
    pi[4] = m.sxz - m.jx*m.jz*inv;
    pi[5] = m.syz - m.jy*m.jz*inv;
}

/*
 * Post-collision deviations out[q] = C_q - w_q from drho, the momentum j (after the removal),
 * Pi_neq and the force density F; stress[6] is an extra symmetric tensor added to P (the stress
 * modes of the fluctuating part, already divided by 4.5, or zeros), ghost[q] the ghost part of
 * the fluctuating part divided by w_q (or nullptr).
 */
template <typename R, bool GHOST, bool SPLIT = true>
__device__ __forceinline__ void collide(R dr, R jx, R jy, R jz, const R pi[6], R fx, R fy, R fz, R omega,
        const R stress[6], const R* ghost, R out[NQ]) {
    R rho = R(1) + dr;
    R inv = R(1)/rho;
    R ux = (jx + R(0.5)*fx)*inv, uy = (jy + R(0.5)*fy)*inv, uz = (jz + R(0.5)*fz)*inv;
    R relax = R(1) - omega;
    R pxx = rho*ux*ux + relax*pi[0] + ux*fx + stress[0];
    R pyy = rho*uy*uy + relax*pi[1] + uy*fy + stress[1];
    R pzz = rho*uz*uz + relax*pi[2] + uz*fz + stress[2];
    R pxy = rho*ux*uy + relax*pi[3] + R(0.5)*(ux*fy + uy*fx) + stress[3];
    R pxz = rho*ux*uz + relax*pi[4] + R(0.5)*(ux*fz + uz*fx) + stress[4];
    R pyz = rho*uy*uz + relax*pi[5] + R(0.5)*(uy*fz + uz*fy) + stress[5];
    R mx = jx + fx, my = jy + fy, mz = jz + fz;
    // Shared parts of the diagonal terms: dr + 4.5 (P_aa + P_bb - T), T = tr(P)/3, for the three planes
    // (4.5 T = 1.5 tr(P): no rounded 1/3).
    R base0 = dr - R(1.5)*(pxx + pyy + pzz);
    R ax = base0 + R(4.5)*pxx, ay = base0 + R(4.5)*pyy, az = base0 + R(4.5)*pzz;
    R dxy = ax + R(4.5)*pyy, dxz = ax + R(4.5)*pzz, dyz = ay + R(4.5)*pzz;
    R sxy = R(9)*pxy, sxz = R(9)*pxz, syz = R(9)*pyz;
    R m3x = R(3)*mx, m3y = R(3)*my, m3z = R(3)*mz;
    R e[NQ];
    e[0] = base0;
    e[1] = ax + m3x;  e[2] = ax - m3x;
    e[3] = ay + m3y;  e[4] = ay - m3y;
    e[5] = az + m3z;  e[6] = az - m3z;
    e[7] = dxy + sxy + (m3x + m3y);   e[8] = dxy + sxy - (m3x + m3y);
    e[9] = dxy - sxy + (m3x - m3y);   e[10] = dxy - sxy - (m3x - m3y);
    e[11] = dyz + syz + (m3y + m3z);  e[12] = dyz + syz - (m3y + m3z);
    e[13] = dyz - syz + (m3y - m3z);  e[14] = dyz - syz - (m3y - m3z);
    e[15] = dxz + sxz + (m3x + m3z);  e[16] = dxz + sxz - (m3x + m3z);
    e[17] = dxz - sxz - (m3x - m3z);  e[18] = dxz - sxz + (m3x - m3z);
#pragma unroll
    for (int q = 0; q < NQ; q++) {
        R v = e[q];
        if constexpr (GHOST)
            v += ghost[q];
        out[q] = weighted<R, SPLIT>(q, v);
    }
}

/*
 * Ghost part of the fluctuating collision divided by w_q: g_q = amp * sum_{k=10..18} e_k(c_q)/sqrt(b_k) r[k-4]
 * (D3Q19.h:179-186), with the constant basis folded at compile time.
 */
template <typename R>
__device__ __forceinline__ void ghostNoise(const R r[16], R amp, R g[NQ]) {
    staticFor<NQ>([&](auto I) {
        constexpr int q = decltype(I)::value;
        R sum = R(0);
        staticFor<9>([&](auto K) {
            constexpr int k = decltype(K)::value + 10;
            constexpr double coef = basisMode(k, q)*INV_SQRT_NORM[k];
            if constexpr (coef != 0.0)
                sum += R(coef)*r[k - 4];
        });
        g[q] = amp*sum;
    });
}

/* Ghost modes 10-15 are odd in c and 16-18 even (D3Q19.h:145-169), which ghostNoiseWeighted relies on. */
constexpr bool ghostParity() {
    for (int q = 1; q < NQ; q += 2)
        for (int k = 10; k < NQ; k++)
            if (basisMode(k, q + 1) != (k < 16 ? -1.0 : 1.0)*basisMode(k, q))
                return false;
    for (int k = 10; k < 16; k++)
        if (basisMode(k, 0) != 0.0)
            return false;
    return true;
}
static_assert(ghostParity(), "parity of the ghost modes");

/* c*x for a compile-time coefficient c, added to acc; a zero coefficient adds nothing. */
template <typename R>
__device__ __forceinline__ R addScaled(R acc, double c, R x) {
    return c == 0.0 ? acc : acc + R(c)*x;
}

/*
 * w_q g_q for the ghost part g_q of ghostNoise (the same normals r[6..14] and amplitude), formed per opposite pair:
 * with the odd modes' sum O_q and the even modes' sum E_q at c_q, w g_q = E_q + O_q and w g_opp = E_q - O_q.
 */
template <typename R>
__device__ __forceinline__ void ghostNoiseWeighted(const R r[16], R amp, R out[NQ]) {
    R a[9];
#pragma unroll
    for (int k = 0; k < 9; k++)
        a[k] = amp*r[k + 6];
    out[0] = R(0);
    staticFor<3>([&](auto K) {
        constexpr int k = decltype(K)::value + 16;
        out[0] = addScaled(out[0], WQ[0]*basisMode(k, 0)*INV_SQRT_NORM[k], a[k - 10]);
    });
    staticFor<9>([&](auto P) {
        constexpr int q = 2*decltype(P)::value + 1;
        R odd = R(0), even = R(0);
        staticFor<9>([&](auto K) {
            constexpr int k = decltype(K)::value + 10;
            constexpr double c = WQ[q]*basisMode(k, q)*INV_SQRT_NORM[k];
            if constexpr (k < 16)
                odd = addScaled(odd, c, a[k - 10]);
            else
                even = addScaled(even, c, a[k - 10]);
        });
        out[q] = even + odd;
        out[q + 1] = even - odd;
    });
}

/* The stress modes (k = 4..9) of the fluctuating part as a tensor added to P: the contribution
   w_q sum_k e_k/sqrt(b_k) amp r_k equals w_q 4.5 H2:S with S = diag(a4, a5, a6)/4.5 and
   off-diagonal a7/9, a8/9, a9/9 (a_k = amp r_k/sqrt(b_k)). */
template <typename R>
__device__ __forceinline__ void stressNoise(const R r[16], R amp, R s[6]) {
    R d = amp*R(INV_SQRT_NORM[4]/4.5), o = amp*R(INV_SQRT_NORM[7]/9.0);
    s[0] = d*r[0];
    s[1] = d*r[1];
    s[2] = d*r[2];
    s[3] = o*r[3];
    s[4] = o*r[4];
    s[5] = o*r[5];
}

/* Stress part of the fluctuating collision, per population (w_q included): D3Q19::fluctuation() with
   a zero ghost amplitude. */
template <typename R>
__device__ __forceinline__ void stressNoisePopulations(const R r[16], R amp, R out[NQ]) {
    staticFor<NQ>([&](auto I) {
        constexpr int q = decltype(I)::value;
        R sum = R(0);
        staticFor<6>([&](auto K) {
            constexpr int k = decltype(K)::value + 4;
            constexpr double coef = basisMode(k, q)*INV_SQRT_NORM[k];
            if constexpr (coef != 0.0)
                sum += R(coef)*r[k - 4];
        });
        out[q] = R(WQ[q])*amp*sum;
    });
}

/* ------------------------------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon et al., SC'11) and Box-Muller normals.                               */
/* ------------------------------------------------------------------------------------------ */
constexpr uint32_t PHILOX_M0 = 0xD2511F53u, PHILOX_M1 = 0xCD9E8D57u;
constexpr uint32_t PHILOX_W0 = 0x9E3779B9u, PHILOX_W1 = 0xBB67AE85u;
constexpr uint32_t TAG_FLUID = 0x464C5549u, TAG_BOUNDARY = 0x424F554Eu, TAG_PARTICLE = 0x50415254u;

struct uint4x {
    uint32_t x, y, z, w;
};

__device__ __forceinline__ uint4x philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int round = 0; round < 10; round++) {
        uint64_t p0 = (uint64_t) PHILOX_M0*c0;
        uint64_t p1 = (uint64_t) PHILOX_M1*c2;
        uint32_t n0 = (uint32_t) (p1 >> 32) ^ c1 ^ k0;
        uint32_t n2 = (uint32_t) (p0 >> 32) ^ c3 ^ k1;
        c1 = (uint32_t) p1;
        c3 = (uint32_t) p0;
        c0 = n0;
        c2 = n2;
        k0 += PHILOX_W0;
        k1 += PHILOX_W1;
    }
    return {c0, c1, c2, c3};
}

/*
 * A pair of normals by the Box-Muller transform of a 32-bit word a (radius) and a 16-bit word b
 * (angle): u1 = (a + 1) 2^-32 in (0, 1], u2 = b 2^-16 in [0, 1), r = sqrt(-2 ln u1),
 * (r cos 2 pi u2, r sin 2 pi u2).  The angle needs no more than 16 bits, so three Philox blocks
 * (12 words: 8 radii, 4 words of two angles each) give 16 normals.
 */
template <typename R>
__device__ __forceinline__ void boxMullerPair(uint32_t a, uint32_t b, R* out) {
    if constexpr (std::is_same<R, double>::value) {
        const double twoPi = 6.283185307179586476925287;
        double r = sqrt(-2.0*log(((double) a + 1.0)*2.3283064365386963e-10));
        double angle = twoPi*((double) b*1.52587890625e-05);
        out[0] = r*cos(angle);
        out[1] = r*sin(angle);
    }
    else {
        // Hardware transcendentals: v_log_f32 is log2 (so -2 ln u1 = -2 ln 2 log2 u1), v_sin_f32/v_cos_f32
        // take the angle in revolutions (sin 2 pi u2), v_sqrt_f32 is within 1 ulp.
        float u1 = fmaf((float) a, 2.3283064365386963e-10f, 2.3283064365386963e-10f);
        float r = __builtin_amdgcn_sqrtf(-1.3862943611198906f*__builtin_amdgcn_logf(u1));
        float turns = (float) b*1.52587890625e-05f;
        out[0] = r*__builtin_amdgcn_cosf(turns);
        out[1] = r*__builtin_amdgcn_sinf(turns);
    }
}

/* COUNT (<= 20) normals for (seed, tag, index, step): blocks of 16 normals from three Philox calls with
   counters (index, step_lo, step_hi, 3*block + c). */
template <typename R, int COUNT>
__device__ __forceinline__ void drawNormals(uint32_t seed, uint32_t tag, uint32_t index, int64_t step, R* out) {
    uint32_t lo = (uint32_t) (uint64_t) step, hi = (uint32_t) ((uint64_t) step >> 32);
#pragma unroll
    for (int block = 0; block < (COUNT + 15)/16; block++) {
        uint32_t w[12];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            uint4x b = philox(index, lo, hi, (uint32_t) (3*block + c), seed, tag);
            w[4*c] = b.x;
            w[4*c + 1] = b.y;
            w[4*c + 2] = b.z;
            w[4*c + 3] = b.w;
        }
#pragma unroll
        for (int p = 0; p < 8; p++) {
            if (16*block + 2*p >= COUNT)
                break;
            uint32_t angle = (w[8 + p/2] >> (16*(p & 1))) & 0xFFFFu;
            boxMullerPair<R>(w[p], angle, out + 16*block + 2*p);
        }
    }
}

/* ------------------------------------------------------------------------------------------ */
/* Storage types.                                                                             */
/* ------------------------------------------------------------------------------------------ */
/* Real: the type of values loaded from and stored to the storage (and of the reaction field).
   Compute: the type of the moments and of the collision.  FP32 storage collides in FP64: summing
   the moments of ~1e-3 deviations in FP32 injects ~1e-9 of rounding into j at every step, which
   over 1000 steps drifts the velocity by ~5e-6 relative; in FP64 only the storage rounds (~1e-7).
   On CDNA the non-packed FP64 vector rate equals the FP32 one and the pass is memory-bound.  FP16
   and BF16 collide in FP32: their storage rounding dominates. */
template <typename S> struct StorageTraits;
template <> struct StorageTraits<float> { using Real = float; using Compute = double; };
template <> struct StorageTraits<double> { using Real = double; using Compute = double; };
template <> struct StorageTraits<__half> { using Real = float; using Compute = float; };
template <> struct StorageTraits<__hip_bfloat16> { using Real = float; using Compute = float; };

/* p as a pointer to global memory: every population, accumulator and mask lives there.  Accessed through it, a
   pointer that was read from memory (a table of replicas) still gives global loads and stores (with a scalar base
   where uniform) instead of generic flat ones. */
template <typename T>
__device__ __forceinline__ __attribute__((address_space(1))) T* asGlobal(T* p) {
    return (__attribute__((address_space(1))) T*) p;
}

template <typename S>
__device__ __forceinline__ typename StorageTraits<S>::Real loadPop(const S* p) {
    if constexpr (std::is_same<S, __half>::value)
        return (float) __builtin_bit_cast(_Float16, *asGlobal(reinterpret_cast<const uint16_t*>(p)));
    else if constexpr (std::is_same<S, __hip_bfloat16>::value)
        return __uint_as_float(((uint32_t) *asGlobal(reinterpret_cast<const uint16_t*>(p))) << 16);
    else
        return *asGlobal(p);
}

/* Integer hash for the stochastic rounding of reduced-precision stores (a murmur3-style finalizer). */
__device__ __forceinline__ uint32_t mixBits(uint32_t h) {
    h ^= h >> 16;
    h *= 0x7FEB352Du;
    h ^= h >> 15;
    h *= 0x846CA68Bu;
    h ^= h >> 16;
    return h;
}

/* Store with stochastic rounding for FP16/BF16 (gfx950 v_cvt_sr_*), plain stores otherwise. */
template <typename S>
__device__ __forceinline__ void storePop(S* p, typename StorageTraits<S>::Real v, uint32_t rnd) {
    if constexpr (std::is_same<S, __half>::value) {
        typedef _Float16 half2v __attribute__((ext_vector_type(2)));
        half2v h = {0, 0};
        h = __builtin_amdgcn_cvt_sr_f16_f32(h, v, rnd, false);
        *reinterpret_cast<_Float16*>(p) = h[0];
    }
    else if constexpr (std::is_same<S, __hip_bfloat16>::value) {
        typedef __bf16 bf2v __attribute__((ext_vector_type(2)));
        bf2v h = {0, 0};
        h = __builtin_amdgcn_cvt_sr_bf16_f32(h, v, rnd, false);
        *reinterpret_cast<__bf16*>(p) = h[0];
    }
    else {
        (void) rnd;
        *p = v;
    }
}

template <typename S>
constexpr bool isReduced() {
    return std::is_same<S, __half>::value || std::is_same<S, __hip_bfloat16>::value;
}

/* Streaming (nontemporal) variants for the fused pass: every population is touched once per step,
   so its lines should not displace others in the caches. */
template <typename S>
__device__ __forceinline__ typename StorageTraits<S>::Real loadPopStream(const S* p) {
    if constexpr (std::is_same<S, __half>::value) {
        uint16_t bits = __builtin_nontemporal_load(asGlobal(reinterpret_cast<const uint16_t*>(p)));
        return (float) __builtin_bit_cast(_Float16, bits);
    }
    else if constexpr (std::is_same<S, __hip_bfloat16>::value) {
        uint16_t bits = __builtin_nontemporal_load(asGlobal(reinterpret_cast<const uint16_t*>(p)));
        return __uint_as_float(((uint32_t) bits) << 16);
    }
    else
        return __builtin_nontemporal_load(asGlobal(p));
}

template <typename S>
__device__ __forceinline__ void storePopStream(S* p, typename StorageTraits<S>::Real v, uint32_t rnd) {
    if constexpr (std::is_same<S, __half>::value) {
        typedef _Float16 half2v __attribute__((ext_vector_type(2)));
        half2v h = {0, 0};
        h = __builtin_amdgcn_cvt_sr_f16_f32(h, v, rnd, false);
        __builtin_nontemporal_store(__builtin_bit_cast(uint16_t, h[0]), asGlobal(reinterpret_cast<uint16_t*>(p)));
    }
    else if constexpr (std::is_same<S, __hip_bfloat16>::value) {
        typedef __bf16 bf2v __attribute__((ext_vector_type(2)));
        bf2v h = {0, 0};
        h = __builtin_amdgcn_cvt_sr_bf16_f32(h, v, rnd, false);
        __builtin_nontemporal_store(__builtin_bit_cast(uint16_t, h[0]), asGlobal(reinterpret_cast<uint16_t*>(p)));
    }
    else {
        (void) rnd;
        __builtin_nontemporal_store(v, asGlobal(p));
    }
}

/* ------------------------------------------------------------------------------------------ */
/* Lattice geometry.                                                                          */
/* ------------------------------------------------------------------------------------------ */

/* Unsigned division by an invariant integer (Granlund and Montgomery 1994, the round-up
   variant with a 33-bit multiplier): q = (t + ((n - t) >> 1)) >> (l - 1), t = mulhi(m, n). */
struct FastDiv {
    uint32_t d, m, l;
};

__device__ __forceinline__ uint32_t fastDiv(uint32_t n, FastDiv f) {
    if (f.d == 1)
        return n;
    uint32_t t = __umulhi(f.m, n);
    return (t + ((n - t) >> 1)) >> (f.l - 1);
}

struct Geometry {
    uint32_t nx, ny, nz, nxy, n;
    FastDiv divx, divxy;
    int openX, openY, openZ;
    uint32_t rowShift = 0;      // log2 of the rows per block of the population layout (elementOf)
    uint32_t qstride = 0;       // elements between populations of one node: nx << rowShift
};

/*
 * Population layout.  The rows (j, k) of the lattice are grouped in blocks of R = 2^rowShift consecutive rows (r = j +
 * ny k; R divides ny); a block holds its 19 populations one after the other, each as R rows of nx values.  Population
 * q of node (i, j, k) is the element ((r >> rowShift)*19 + q)*R*nx + (r & (R - 1))*nx + i.  With R = 1 each row holds
 * its 19 populations in turn, q*nx + 19*node - 18*i.  A work-group of consecutive nodes then reads and writes a few
 * contiguous runs of R*nx values per population instead of 19 planes N values apart, which the HBM serves faster (on
 * 256^3 a bare pull measured 73% of the peak instead of 69%); larger R lengthens the runs for 2-byte storage.
 */

/* Element of population 0 of the rows' block base for row r (add q*qstride for population q, then i). */
template <typename I>
__device__ __forceinline__ I rowElement(const Geometry& g, uint32_t r) {
    return (I) (r >> g.rowShift)*(I) NQ*(I) g.qstride + (I) (r & ((1u << g.rowShift) - 1u))*(I) g.nx;
}

/* Element of population 0 of node (i, j, k). */
template <typename I>
__device__ __forceinline__ I elementOf(const Geometry& g, uint32_t i, uint32_t j, uint32_t k) {
    return rowElement<I>(g, j + g.ny*k) + (I) i;
}

/* The offsets, in population elements, that take a node to its neighbours one step along -c (the pull source) on a
   periodic lattice.  Moving by ny rows keeps a row's place in its block (R divides ny), so the offsets add up for the
   diagonal directions. */
template <typename I>
struct NeighboursOf {
    I xm, xp, ym, yp, zm, zp;   // offsets to (i-1), (i+1), (j-1), (j+1), (k-1), (k+1), wrapped
};

__device__ __forceinline__ void nodeCoords(const Geometry& g, uint32_t node, uint32_t& i, uint32_t& j, uint32_t& k) {
    k = fastDiv(node, g.divxy);
    uint32_t rem = node - k*g.nxy;
    j = fastDiv(rem, g.divx);
    i = rem - j*g.nx;
}

template <typename I>
__device__ __forceinline__ NeighboursOf<I> elementNeighbours(const Geometry& g, uint32_t i, uint32_t j, uint32_t k) {
    const I slab = (I) NQ*(I) g.nxy, all = (I) NQ*(I) g.n;
    const uint32_t r = j + g.ny*k;
    const I here = rowElement<I>(g, r);
    NeighboursOf<I> nb;
    nb.xm = i > 0 ? (I) -1 : (I) g.nx - 1;
    nb.xp = i + 1 < g.nx ? (I) 1 : (I) 1 - (I) g.nx;
    nb.ym = rowElement<I>(g, j > 0 ? r - 1 : r + g.ny - 1) - here;
    nb.yp = rowElement<I>(g, j + 1 < g.ny ? r + 1 : r + 1 - g.ny) - here;
    nb.zm = k > 0 ? -slab : all - slab;
    nb.zp = k + 1 < g.nz ? slab : slab - all;
    return nb;
}

/* Offset of the node x - c_q (the pull source of direction q) from x. */
template <int Q, typename I>
__device__ __forceinline__ I pullOffset(const NeighboursOf<I>& nb) {
    I off = 0;
    if constexpr (CX[Q] == 1) off += nb.xm;
    if constexpr (CX[Q] == -1) off += nb.xp;
    if constexpr (CY[Q] == 1) off += nb.ym;
    if constexpr (CY[Q] == -1) off += nb.yp;
    if constexpr (CZ[Q] == 1) off += nb.zm;
    if constexpr (CZ[Q] == -1) off += nb.zp;
    return off;
}

/* Offset of the node x + c_q from x. */
template <int Q, typename I>
__device__ __forceinline__ I pushOffset(const NeighboursOf<I>& nb) {
    return pullOffset<OPP[Q]>(nb);
}

/* Where a node's populations are, for the kernels off the hot path: 64-bit element indices. */
struct NodeSlots {
    uint64_t e;                      // element of population 0 of the node
    NeighboursOf<long long> nb;      // element offsets to its neighbours
    uint64_t nx;                     // element stride between populations (Geometry::qstride)

    /* Population q of the node itself, of its pull source x - c_q, of the node x + c_q. */
    __device__ __forceinline__ uint64_t own(int q) const { return (uint64_t) q*nx + e; }
    template <int Q> __device__ __forceinline__ uint64_t pulled() const {
        return (uint64_t) Q*nx + (uint64_t) ((long long) e + pullOffset<Q>(nb));
    }
    template <int Q, int P> __device__ __forceinline__ uint64_t pushed() const {   // population P at x + c_Q
        return (uint64_t) P*nx + (uint64_t) ((long long) e + pushOffset<Q>(nb));
    }
};

__device__ __forceinline__ NodeSlots nodeSlots(const Geometry& g, uint32_t node, uint32_t i, uint32_t j, uint32_t k) {
    (void) node;
    return NodeSlots{elementOf<uint64_t>(g, i, j, k), elementNeighbours<long long>(g, i, j, k), g.qstride};
}

/* The element at a 32-bit byte offset from a base (the array plus q*nx for population q): the base stays in scalar
   registers and the access takes the offset from one vector register.  Used where the whole array is at most 4 GiB (checked at context
   creation; larger lattices take the 64-bit path). */
template <typename T>
__device__ __forceinline__ T* atByte(T* base, uint32_t offset) {
    using Byte = typename std::conditional<std::is_const<T>::value, const char, char>::type;
    return reinterpret_cast<T*>(reinterpret_cast<Byte*>(base) + offset);
}

/* Pull the 19 pre-collision deviations of a node: f_q(x) = post-collision f_q stored at x - c_q (64-bit indices). */
template <typename S>
__device__ __forceinline__ void pullNode(const S* __restrict__ src, const Geometry& g, uint32_t node, uint32_t i, uint32_t j,
        uint32_t k, typename StorageTraits<S>::Real f[NQ]) {
    const NodeSlots at = nodeSlots(g, node, i, j, k);
    staticFor<NQ>([&](auto I) {
        constexpr int q = decltype(I)::value;
        f[q] = loadPop(src + at.template pulled<q>());
    });
}

/*
 * The fused pass's pull, with streaming loads.  WIDE = false: 32-bit byte offsets from the base of each population
 * row block (whole array at most 4 GiB); WIDE = true: 64-bit element indices.
 */
template <bool WIDE>
using ElementIndex = typename std::conditional<WIDE, uint64_t, uint32_t>::type;

/* Whether the fused pass pulls population Q with a nontemporal (streaming) load or a plain one.  Measured as the
   coupled step of the 256^3 lattice with 20,000 Keys beads: FP32 is fastest with every pull nontemporal (0.627 ms;
   0.630 ms with the x-shifted pulls plain, 0.635 ms with all plain); FP16 with the x-shifted pulls plain and the others
   nontemporal (0.402 ms; 0.404 ms all plain, 0.409 ms all nontemporal): a wavefront's 2-byte pulls span one 128 B
   line, and an x-shifted one straddles the line of the neighbouring wavefront, which a plain load leaves in L2 for
   it. */
template <typename S, int Q>
constexpr bool pullStreams() {
    return !isReduced<S>() || CX[Q] == 0;
}

template <typename S, bool WIDE>
__device__ __forceinline__ void pullNodeStream(const S* __restrict__ src, uint32_t nx, ElementIndex<WIDE> e,
        const NeighboursOf<typename std::conditional<WIDE, long long, int>::type>& nb,
        typename StorageTraits<S>::Real f[NQ]) {
    if constexpr (WIDE) {
        staticFor<NQ>([&](auto I) {
            constexpr int q = decltype(I)::value;
            const S* p = src + (uint64_t) q*nx + (uint64_t) ((long long) e + pullOffset<q>(nb));
            f[q] = pullStreams<S, q>() ? loadPopStream(p) : loadPop(p);
        });
    }
    else {
        // Byte offsets modulo 2^32: the sums wrap to the right offset, which lies within the 4 GiB array.
        constexpr uint32_t B = (uint32_t) sizeof(S);
        const NeighboursOf<uint32_t> nbBytes = {(uint32_t) nb.xm*B, (uint32_t) nb.xp*B, (uint32_t) nb.ym*B,
                                                (uint32_t) nb.yp*B, (uint32_t) nb.zm*B, (uint32_t) nb.zp*B};
        const uint32_t at = e*B;
        staticFor<NQ>([&](auto I) {
            constexpr int q = decltype(I)::value;
            const S* p = atByte(src + (size_t) q*nx, at + pullOffset<q>(nbBytes));
            f[q] = pullStreams<S, q>() ? loadPopStream(p) : loadPop(p);
        });
    }
}

/* Solid-node bitmask (bit node&31 of word node>>5). */
__device__ __forceinline__ bool isSolid(const uint32_t* __restrict__ mask, uint32_t node) {
    return (mask[node >> 5] >> (node & 31)) & 1u;
}

}  // namespace lbm
