/*
 * acemlp.h — C ABI of libacemlp.so, the network of the learned surrogate:
 * an ensemble of MLPs (IN -> 128 -> 128 -> 128 -> 1, SiLU) with BF16
 * weights/activations and FP32 accumulation on MFMA (gfx950).
 *
 * No counterpart in the reference: ACE.jl has no learned part.  The
 * surrogate predicts the correlation-order-3/4 share of the LinearACEModel
 * site energy from the order <= 2 invariants (B functions of the same
 * SymmetricBasis, symmbasis.jl:297-315); plain ACE (libacehip.so) is the
 * fallback and the reference its error is measured against.
 *
 * Inputs X are BF16 [nrows][160] (normalized features, zero padded), device
 * memory.  Parameters are FP32 masters [nmemb][acemlp_nparam()] with BF16
 * copies used by the kernels.  Return 0 on success, < 0 on error
 * (acemlp_last_error()).
 */
#ifndef ACEMLP_H
#define ACEMLP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct acemlp_ctx acemlp_ctx;

int acemlp_nparam(void);
int acemlp_input_width(void);
const char* acemlp_last_error(void);

/* nmemb members, nin valid input features (<= 160), He-uniform init from seed. */
int acemlp_create(int nmemb, int nin, int device, uint64_t seed, acemlp_ctx** out);
void acemlp_destroy(acemlp_ctx* ctx);

/* host FP32 parameter transfer, n = nmemb * acemlp_nparam() */
int acemlp_get_params(acemlp_ctx* ctx, float* host, int64_t n);
int acemlp_set_params(acemlp_ctx* ctx, const float* host, int64_t n);

/* y[nmemb][nrows] (device FP32) */
int acemlp_predict(acemlp_ctx* ctx, int64_t nrows, const void* X, float* y, void* stream);

/* y[nmemb][nrows] and dX[nmemb][nrows][160] = d y_m / d X (device FP32) */
int acemlp_predict_grad(acemlp_ctx* ctx, int64_t nrows, const void* X, float* y, float* dX, void* stream);

/* one Adam step on the mean-squared error to targets t[nrows] (device FP32);
 * every member sees the same batch.  loss_host[nmemb] (optional, syncs). */
int acemlp_train_step(acemlp_ctx* ctx, int64_t nrows, const void* X, const float* t, float lr,
                      float weight_decay, float* loss_host, void* stream);

/* gradient of the MSE w.r.t. the parameters (host FP32 [nmemb][nparam]), no update */
int acemlp_loss_grad(acemlp_ctx* ctx, int64_t nrows, const void* X, const float* t, float* grad_host,
                     float* loss_host);

/* 16x16 FP32 result of one MFMA on known operands (lane-layout check) */
int acemlp_layout_probe(float* host16x16);

#ifdef __cplusplus
}
#endif
#endif
