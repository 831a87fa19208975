/*
 * Internal declarations of the learned coupling closure (include/lbm_amd_closure.h): the device layout of the
 * network and the launchers shared by the network kernels (lbm_mlp.hip), the closure's coupling kernels
 * (lbm_closure.hip) and the context (lbm_context.hip).
 */
#pragma once

#include "lbm_internal.h"
#include "../../include/lbm_amd_closure.h"

namespace lbm {

constexpr int MLP_IN = 87, MLP_IN_PAD = 96, MLP_H = 64, MLP_OUT = 3, MLP_OUT_PAD = 16;
constexpr int MLP_OFF_W1 = 0;
constexpr int MLP_OFF_B1 = MLP_OFF_W1 + MLP_IN_PAD*MLP_H;
constexpr int MLP_OFF_W2 = MLP_OFF_B1 + MLP_H;
constexpr int MLP_OFF_B2 = MLP_OFF_W2 + MLP_H*MLP_H;
constexpr int MLP_OFF_W3 = MLP_OFF_B2 + MLP_H;
constexpr int MLP_OFF_B3 = MLP_OFF_W3 + MLP_H*MLP_OUT_PAD;
constexpr int MLP_PARAMS = MLP_OFF_B3 + MLP_OUT_PAD;
static_assert(MLP_PARAMS == LBM_MLP_NUM_PARAMS, "parameter layout");

/*
 * The FP16 operands of the MFMA kernels, converted from the FP32 master parameters after every update: each matrix in
 * the orientation whose 4-element K runs are contiguous for the A operand of v_mfma_f32_16x16x16_f16 (output-major for
 * the forward pass, input-major for the backward pass), and the FP32 biases.
 */
struct MlpDevice {
    _Float16 w1t[MLP_H*MLP_IN_PAD];     // [o][i]
    _Float16 w2t[MLP_H*MLP_H];          // [o][i]
    _Float16 w3t[MLP_OUT_PAD*MLP_H];    // [o][i]
    _Float16 w2[MLP_H*MLP_H];           // [i][o]
    _Float16 w3[MLP_H*MLP_OUT_PAD];     // [i][o]
    float b1[MLP_H], b2[MLP_H], b3[MLP_OUT_PAD];
};

/* Normalization and the in-distribution bounds, as the kernels use them. */
struct ClosureNorm {
    float uScale, vScale, outScale;
    float uBound, vBound;
};

/* The network's device weights and normalization (valid while the lbm_mlp lives). */
const MlpDevice* mlpDeviceWeights(const lbm_mlp* mlp);
ClosureNorm mlpClosureNorm(const lbm_mlp* mlp);

/* du[b][0..2] = network(X[b]) / outScale for b < n (X: normalized FP16 inputs [n][96]). */
void launchMlpInfer(const MlpDevice* weights, const _Float16* X, int n, float outScale, float* du, hipStream_t stream);

/*
 * The closure's coupling kernels (lbm_closure.hip) for the coupled particles of one context.
 *  features: normalized FP16 inputs X [n][96] (and raw FP32 inputs [n][87] if raw != nullptr) in each particle's
 *            frame (the reflection and axis order that make its sub-node offset non-negative and sorted), the frame's
 *            code, the nearest node and its FP64 velocity, and the out-of-distribution flag of every particle;
 *  targets:  u_Keys - u(nearest) of every particle in its frame (training pairs; a.stencil must be Keys);
 *  couple:   nearest-node drag with u(nearest) + du (du turned back from the particle's frame) for in-distribution
 *            particles, reaction on the nearest node;
 *            counters[2b] and counters[2b + 1] count the particles of work-group b coupled through the network and
 *            left to the fallback (closureCoupleBlocks(n) work-groups);
 *  fallback: the stencil coupling (a.stencil) of the out-of-distribution particles, wall momentum into
 *            closureFallbackBlocks(n) partials.
 */
void launchClosureFeatures(int storage, const void* src, const uint32_t* mask, const double* positions,
        const double* velocities, const int* particles, const double* ucm, const ParticleArgs& a, const ClosureNorm& norm,
        const double* field, _Float16* X, float* raw, uint32_t* centreNode, double* centreVelocity, uint8_t* ood,
        uint8_t* frame, hipStream_t stream);
/* The node velocities u = j/rho of the whole lattice, FP64 [node][3] (zero at solid nodes), for
   launchClosureFeatures(field = ...) when 27 x particles exceed the nodes: one coalesced pass instead of 27 scattered
   pulls per particle.  The values are those the per-particle path forms, bit for bit. */
void launchClosureField(int storage, const void* src, const uint32_t* mask, const double* ucm, const ParticleArgs& a,
        double* field, hipStream_t stream);
void launchClosureTargets(int storage, const void* src, const uint32_t* mask, const double* positions,
        const int* particles, const double* ucm, const ParticleArgs& a, const double* centreVelocity,
        const uint8_t* frame, float* targets, hipStream_t stream);
/* Work-groups of launchClosureCouple for numParticles particles: its counters hold two counts per work-group. */
int closureCoupleBlocks(int numParticles);
void launchClosureCouple(const double* velocities, const int* particles, const double* mass, const ParticleArgs& a,
        const uint32_t* centreNode, const double* centreVelocity, const float* du, const uint8_t* ood,
        const uint8_t* frame, double* particleForce, unsigned long long* reactionFixed, uint32_t* touched,
        unsigned long long* counters, hipStream_t stream);
int closureFallbackBlocks(int numParticles);
void launchClosureFallback(int storage, const void* src, const uint32_t* mask, const double* positions,
        const double* velocities, const int* particles, const double* mass, const double* ucm, const ParticleArgs& a,
        const uint8_t* ood, double* particleForce, unsigned long long* reactionFixed, uint32_t* touched, double* partials,
        hipStream_t stream);

}  // namespace lbm
