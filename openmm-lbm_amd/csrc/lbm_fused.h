/*
 * The per-node body of the fused pull-stream + regularized collision + Guo forcing (+ Philox noise) pass,
 * shared by the single-lattice kernel (lbm_fluid.hip) and the batched-replica kernel (lbm_batch.hip).
 * Reference counterparts: computeFluidMoments + collideAndStream (lbmFluid.cc:89-139, 230-310;
 * ReferenceLBMKernels.cpp:430-469, 1066-1120).
 */
#pragma once

#include "lbm_internal.h"

namespace lbm {

/*
 * The touched-mask word of a node's wavefront (32 nodes): every lane reads it, then its first lane (always inside the
 * lattice, since words are 32-aligned) clears it for the next step.
 */
template <bool REACTION>
__device__ __forceinline__ uint32_t takeTouchedWord(uint32_t* __restrict__ touched, uint32_t node) {
    uint32_t touchedWord = 0;
    if constexpr (REACTION) {
        touchedWord = *asGlobal(touched + (node >> 5));
        if ((node & 31) == 0 && touchedWord != 0)
            *asGlobal(touched + (node >> 5)) = 0;
    }
    return touchedWord;
}

/*
 * One fluid node from its pulled pre-collision deviations fr[q] = C_q(x - c_q): form the moments, subtract the removal
 * velocity, add the body force and the particle reaction (consumed from the fixed-point accumulators where the
 * touched bit is set), collide, add the fluctuating part, store C_q(x) to dst at element e of x.
 */
template <typename S, bool REACTION, bool FLUCT, bool WIDE = false>
__device__ __forceinline__ void collideNode(const typename StorageTraits<S>::Real fr[NQ], S* __restrict__ dst,
        unsigned long long* __restrict__ reactionFixed, uint32_t touchedWord, const double* __restrict__ ucm,
        const StepArgs& a, uint32_t node, ElementIndex<WIDE> e) {
    using R = typename StorageTraits<S>::Real;
    using C = typename StorageTraits<S>::Compute;
    C f[NQ];
#pragma unroll
    for (int q = 0; q < NQ; q++)
        f[q] = (C) fr[q];
    Moments<C> m = nodeMoments(f);
    C pi[6];
    piNeq(m, pi);
    C rho = C(1) + m.dr;
    C jx = m.jx, jy = m.jy, jz = m.jz;
    if (a.removal) {
        jx -= rho*C(ucm[0]);
        jy -= rho*C(ucm[1]);
        jz -= rho*C(ucm[2]);
    }
    C fx = rho*C(a.gx), fy = rho*C(a.gy), fz = rho*C(a.gz);
    if constexpr (REACTION) {
        if ((touchedWord >> (node & 31)) & 1u) {
            // Reaction of the coupled particles on this node (fixed point), consumed and cleared: with FP16/BF16
            // populations by one atomic exchange per word (coupled step of 256^3 with 20,000 Keys beads 0.403 ->
            // 0.401 ms), with FP32 or FP64 by a load and a store (the exchange costs 0.627 -> 0.634 ms there).
            C r[3];
#pragma unroll
            for (int c = 0; c < 3; c++) {
                unsigned long long* p = reactionFixed + reactionIndex(node, c, a.g.n, a.reactionPlanes);
                unsigned long long v;
                if constexpr (isReduced<S>())
                    v = __hip_atomic_exchange(p, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                else {
                    v = *asGlobal(p);
                    *asGlobal(p) = 0ull;
                }
                r[c] = C((double) (long long) v*(1.0/FIXED_SCALE));
            }
            fx += r[0];
            fy += r[1];
            fz += r[2];
        }
    }
    C stress[6] = {C(0), C(0), C(0), C(0), C(0), C(0)};
    C omega = C(a.omega);
    // The random part is formed in the storage's load type R (FP32, or FP64 with FP64 storage): its precision is
    // that of a random number.  Its stress modes join P; its ghost modes, weighted, are added after the collision.
    R noise[NQ];
    if constexpr (FLUCT) {
        R r[16], s[6];
        drawNormals<R, 16>(a.seed, TAG_FLUID, node, a.step, r);
        R mu = R(a.mu), rhoR = R(rho), om = R(a.omega);
        R stressAmp, ghostAmp;
        if constexpr (std::is_same<R, float>::value) {
            stressAmp = __builtin_amdgcn_sqrtf(mu*rhoR*om*(2.0f - om));
            ghostAmp = __builtin_amdgcn_sqrtf(mu*rhoR);
        }
        else {
            stressAmp = sqrt(mu*rhoR*om*(R(2) - om));
            ghostAmp = sqrt(mu*rhoR);
        }
        stressNoise(r, stressAmp, s);
        ghostNoiseWeighted(r, ghostAmp, noise);
#pragma unroll
        for (int c = 0; c < 6; c++)
            stress[c] = C(s[c]);
    }
    C out[NQ];
    // The exact weights matter where the storage keeps FP32 or FP64 precision; FP16/BF16 rounding is far coarser.
    collide<C, false, !isReduced<S>()>(m.dr, jx, jy, jz, pi, fx, fy, fz, omega, stress, nullptr, out);
    uint32_t rnd = 0;
    if constexpr (isReduced<S>())
        rnd = mixBits(node ^ a.roundSeed);
#pragma unroll
    for (int q = 0; q < NQ; q++) {
        uint32_t r = 0;
        if constexpr (isReduced<S>()) {
            r = (rnd ^ (uint32_t) (q*0x9E3779B9u))*0x85EBCA6Bu;
            r ^= r >> 13;
        }
        R v = (R) out[q];
        if constexpr (FLUCT)
            v += noise[q];
        if constexpr (WIDE)
            storePopStream(dst + (uint64_t) q*a.g.qstride + e, v, r);
        else
            storePopStream(atByte(dst + (size_t) q*a.g.qstride, e*(uint32_t) sizeof(S)), v, r);
    }
}

/*
 * One node: pull C_q(x - c_q) from src and collide (collideNode).  Returns without writing for solid nodes.
 */
template <typename S, bool SOLIDS, bool REACTION, bool FLUCT, bool WIDE = false>
__device__ __forceinline__ void fusedNode(const S* __restrict__ src, S* __restrict__ dst, const uint32_t* __restrict__ mask,
        unsigned long long* __restrict__ reactionFixed, uint32_t* __restrict__ touched, const double* __restrict__ ucm,
        const StepArgs& a, uint32_t node) {
    uint32_t touchedWord = takeTouchedWord<REACTION>(touched, node);
    if constexpr (SOLIDS)
        if (isSolid(mask, node))
            return;
    uint32_t i, j, k;
    nodeCoords(a.g, node, i, j, k);
    using Offset = typename std::conditional<WIDE, long long, int>::type;
    const ElementIndex<WIDE> e = elementOf<ElementIndex<WIDE>>(a.g, i, j, k);
    typename StorageTraits<S>::Real fr[NQ];
    pullNodeStream<S, WIDE>(src, a.g.qstride, e, elementNeighbours<Offset>(a.g, i, j, k), fr);
    collideNode<S, REACTION, FLUCT, WIDE>(fr, dst, reactionFixed, touchedWord, ucm, a, node, e);
}

}  // namespace lbm
