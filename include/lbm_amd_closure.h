/*
 * lbm_amd_closure.h -- a learned coupling closure: nearest-node coupling corrected by a small MLP toward the Keys
 * stencil.
 *
 * No counterpart in the reference.  It replaces, for one context, the stencil interpolation of the explicit drag
 * (ReferenceLBMKernels.cpp:643-697; the Keys stencil of LBMStencils.h:42-104): every coupled particle is coupled at
 * its nearest node (ReferenceLBMKernels.cpp:503-553) with the fluid velocity u(nearest) + Du, where Du is the output
 * of an 87 -> 64 -> 64 -> 3 network (SiLU) evaluated by FP16 MFMA kernels with FP32 accumulation.  Its inputs:
 *
 *   [0:3]    2 (s - round(s)): the sub-node offset of the particle, s = x/dx wrapped into the lattice
 *   [3:84]   u at the 27 nodes nearest + (a, b, c), a, b, c in {-1, 0, 1} (a fastest), xyz each, lattice units
 *   [84:87]  the particle velocity v dt/dx
 *
 * and its target Du = u_Keys(x) - u(nearest), that is (F_Keys - F_nearest)/(gamma m) of the explicit drag at the same
 * state, all in the particle's frame: the Keys stencil and the lattice are symmetric under the 48 reflections and axis
 * permutations of the cube, and the particle's frame is the one that makes its offset o non-negative and sorted.
 * Frame axis i is lattice axis p_i (|o| in decreasing order, equal values in axis order), reflected where
 * o_(p_i) < 0 (s_i = -1): a vector w is w'_i = s_i w_(p_i) there, the node nearest + d sits at d'_i = s_i d_(p_i) of
 * the block, and the offset inputs are |o_(p_i)|.  The network's output is turned back before the coupling uses it.
 * The reaction goes to the nearest node.  A particle whose inputs fall outside the training range (a
 * normalized velocity beyond the largest seen in training), or whose 3x3x3 neighbourhood holds a solid node or leaves
 * the lattice at an open face, is coupled with the Keys stencil instead: Keys is both the fallback and the yardstick of
 * the network's error.  The context must use the Keys stencil and the explicit drag.
 *
 * Parameters: a flat FP32 vector of LBM_MLP_NUM_PARAMS values, W1[i][o] (96 x 64, inputs padded to 96 with zero
 * rows), b1[64], W2[i][o] (64 x 64), b2[64], W3[i][o] (64 x 16, outputs padded to 16), b3[16]; 9,987 of them are real.
 * Normalization: network input = raw input times (1, u_scale, v_scale) by group, network output = Du times out_scale.
 */
#ifndef LBM_AMD_CLOSURE_H_
#define LBM_AMD_CLOSURE_H_

#include "lbm_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

#define LBM_MLP_NUM_INPUTS 87
#define LBM_MLP_NUM_OUTPUTS 3
#define LBM_MLP_NUM_PARAMS 11408

typedef struct lbm_mlp lbm_mlp;

typedef struct lbm_mlp_normalization {
    double u_scale, v_scale, out_scale;   /* multiply the raw velocity inputs / bead velocity / target */
    double u_bound, v_bound;              /* largest |normalized input| in distribution, per group */
} lbm_mlp_normalization;

/* A network with seeded random weights (normal, variance 1/fan-in; zero biases) on the current device. */
int lbm_mlp_create(uint64_t seed, lbm_mlp** out);
int lbm_mlp_destroy(lbm_mlp* mlp);

int lbm_mlp_get_parameters(lbm_mlp* mlp, float* params, size_t count);
int lbm_mlp_set_parameters(lbm_mlp* mlp, const float* params, size_t count);
int lbm_mlp_get_normalization(lbm_mlp* mlp, lbm_mlp_normalization* norm);
int lbm_mlp_set_normalization(lbm_mlp* mlp, const lbm_mlp_normalization* norm);

/*
 * Train on n pairs (features: raw inputs [n][87]; targets: raw Du [n][3]; host arrays) with Adam (beta 0.9, 0.999),
 * minibatches of batch pairs in a seeded random order, epochs passes, learning rate lr * decay^epoch.  Sets the
 * normalization from the data first (scales: inverse standard deviations; bounds: largest normalized |input|).  The
 * forward and backward passes run in the library's FP16 MFMA kernels.  epoch_loss (epochs values, may be NULL)
 * receives the mean normalized loss of each epoch.
 */
int lbm_mlp_train(lbm_mlp* mlp, const float* features, const float* targets, int64_t n, int epochs, int batch,
        double lr, double decay, uint64_t seed, double* epoch_loss);

/* Du for n raw inputs (host arrays), through the inference kernel. */
int lbm_mlp_forward(lbm_mlp* mlp, const float* features, int64_t n, float* out);

/*
 * No counterpart in the reference: lbm_mlp_forward for callers that cannot allocate device memory (serving loops and
 * harnesses that forbid hipMalloc).  The same kernels and results, with no hipMalloc, hipFree or allocating library
 * call: the device working memory of up to LBM_MLP_NOALLOC_MAX_ROWS rows is a fixed array that each GPU holds until
 * the process ends (9.0 MB).  n beyond that maximum is refused with LBM_ERROR_INVALID.  Every call on one GPU shares
 * that GPU's array: calls that overlap on one GPU (networks on different streams, or threads) race.
 */
#define LBM_MLP_NOALLOC_MAX_ROWS 16384
int lbm_mlp_forward_noalloc(lbm_mlp* mlp, const float* features, int64_t n, float* out);

/*
 * The gradient of the loss mean_b (1/3)|y_b - t_b|^2 (normalized outputs and targets) with respect to the
 * parameters, for one batch of n raw pairs, computed by the training kernels (no update); loss may be NULL.
 */
int lbm_mlp_gradients(lbm_mlp* mlp, const float* features, const float* targets, int64_t n, float* grad,
        double* loss);

/*
 * Training pairs at the context's current state for its coupled particles (positions in nm, velocities in nm/ps of
 * all num_particles particles of the System, host arrays): features [n][87] and targets [n][3] in each particle's
 * frame, n = coupled count.
 */
int lbm_closure_pairs(lbm_context* ctx, const double* positions, const double* velocities, float* features,
        float* targets);

/* Couple this context's particles through the closure (NULL restores the stencil).  The network is read at every
   coupling: it must outlive the attachment and must not be trained meanwhile. */
int lbm_set_closure(lbm_context* ctx, lbm_mlp* mlp);

/* Particle couplings through the network and through the Keys fallback since the last call (then reset). */
int lbm_closure_stats(lbm_context* ctx, int64_t* network, int64_t* fallback);

#ifdef __cplusplus
}
#endif

#endif /* LBM_AMD_CLOSURE_H_ */
