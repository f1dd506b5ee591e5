/* Learned SGS closure for the LES residual: a small eddy-viscosity network trained
 * on the device from filtered high-order solutions, and the sampling of those
 * solutions.
 *
 * The reference's LES models are algebraic eddy viscosities (Smagorinsky and WALE,
 * calc_sgsf, src/eles.cpp:2456-2778 and SGS_flux_kernel, src/cuda_kernels.cu:
 * 1466-1700).  These entry points have no counterpart in the reference: they add
 * a learned eddy viscosity with the WALE model as its fallback.
 *
 * The network: 10 inputs, three hidden layers of 64 SiLU units, a softplus output;
 * BF16 weights and activations as the operands of every product (MFMA), FP32
 * accumulation, FP32 master weights trained with Adam.  Parameter vector (9,089
 * FP32 values): W1 (64x10, [out][in]), b1 (64), W2 (64x64), b2, W3 (64x64), b3,
 * w4 (64), b4 (1).
 *
 * Inputs of the closure at a point with velocity gradient g_ij = du_i/dx_j,
 * G = |g| (Frobenius), filter width D and kinematic viscosity nu:
 *   x_1..9 = g_ij / G (row-major i, j),   x_10 = log(D^2 G / nu)
 * standardised by the model's mean and scale; the output y >= 0 is the
 * eddy-viscosity coefficient C = nu_t / (D^2 G) divided by the model's output
 * scale.  All calls are synchronous on the caller's stream (or the default one).
 */
#ifndef HIFR_SGS_H
#define HIFR_SGS_H

#include "hifr.h"

#ifdef __cplusplus
extern "C" {
#endif

#define HIFR_MLP_INPUTS 10
#define HIFR_MLP_HIDDEN 64
#define HIFR_MLP_PARAMS 9089

typedef struct hifr_mlp hifr_mlp;

/* Network with He-uniform weights (64-bit seed), zero biases, Adam state zeroed. */
hifr_status hifr_mlp_create(hifr_mlp** out, unsigned long long seed);
hifr_status hifr_mlp_destroy(hifr_mlp* m);
hifr_status hifr_mlp_set_stream(hifr_mlp* m, void* hip_stream);

/* FP32 master parameters (HIFR_MLP_PARAMS values); setting them resets Adam. */
hifr_status hifr_mlp_set_params(hifr_mlp* m, const float* params);
hifr_status hifr_mlp_get_params(const hifr_mlp* m, float* params);

/* Input standardisation (mean, scale: 10 each), training range of the raw inputs
 * (lo, hi: 10 each; outside it the closure falls back to WALE) and output scale. */
hifr_status hifr_mlp_set_normalization(hifr_mlp* m, const float* mean, const float* scale, const float* lo,
                                       const float* hi, float out_scale);

/* Loss mean((y - t)^2) over n standardised samples x (n x 10, row-major) with
 * targets t (n), and its gradient (HIFR_MLP_PARAMS values): the device's forward
 * and backward kernels without an update (for checking the gradients). */
hifr_status hifr_mlp_loss_grad(hifr_mlp* m, long n, const float* x, const float* t, float* loss, float* grad);

/* Training set, copied to the device (n standardised samples). */
hifr_status hifr_mlp_set_data(hifr_mlp* m, long n, const float* x, const float* t);

/* epochs passes of Adam (lr, beta1 0.9, beta2 0.999, eps 1e-8) over the training
 * set in mini-batches of `batch` samples, in order; loss_per_epoch (optional,
 * epochs values) receives each epoch's mean batch loss. */
hifr_status hifr_mlp_train(hifr_mlp* m, int epochs, int batch, float lr, float* loss_per_epoch);

/* Network output y (n) for n standardised samples x (n x 10). */
hifr_status hifr_mlp_predict(hifr_mlp* m, long n, const float* x, float* y);

/* Training samples from the state of a FP64 hexahedral context with per-element
 * metrics (affine hexes).  Per element, the state is filtered onto degree
 * order_lo: filter_1d (n_lo x n_hi, row-major, n = order + 1) maps nodal values at
 * the context's 1-D solution points onto the truncated Legendre expansion
 * evaluated at the degree-order_lo Gauss points, applied along r, s, t;
 * deriv_1d (n_lo x n_lo, row-major) differentiates the Lagrange interpolant on
 * those points.  From the filtered rho, momenta and products m_i m_j / rho: the
 * Favre velocity v~, its gradient g, the SGS stress tau and
 *   c_exact = -(tau^d : S~d) / (2 rho-bar S~d : S~d) / (delta^2 |g|),
 *   c_wale  = nu_WALE(g) / (delta^2 |g|)   (the fallback model, for comparison),
 *   x       = the closure's raw inputs (g_ij / |g|, log(delta^2 |g| / nu)),
 * with delta = filter_ratio (8 detJ)^(1/3) / (order_lo + 1).  Outputs are host
 * arrays of N n_lo^3 samples (sample = low-order point + n_lo^3 element; x is
 * (samples x 10)); u_lo (optional, NULL to skip) receives the filtered
 * conservative state (n_lo^3, N, 5). */
hifr_status hifr_sgs_samples(hifr_ctx* c, int order_lo, const double* filter_1d, const double* deriv_1d,
                             double filter_ratio, float* x, float* c_exact, float* c_wale, double* u_lo);

/* Attaches a network to a context created with sgs_model HIFR_SGS_LEARNED: the
 * context copies its current weights and normalisation into buffers of its own
 * (after the network's stream has finished its work), so the network may be
 * trained further or destroyed afterwards; attach again to take new weights.
 * m = NULL detaches (the context's copy is zeroed): the WALE fallback everywhere,
 * as before any network is attached. */
hifr_status hifr_set_sgs_network(hifr_ctx* c, const hifr_mlp* m);

#ifdef __cplusplus
}
#endif

#endif /* HIFR_SGS_H */
