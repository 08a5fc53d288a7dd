/* C ABI of the surrogate network of the MI355X light-curve engine
 * (liblcirs.so).  No counterpart in the reference: the network predicts,
 * from fourteen features of a source position, mu = log(A_finite / A_point) and
 * s = log(sigma) of the extended-source magnification, so that positions far
 * from the caustics can skip the ray shooting (see
 * triplelens_amd/surrogate.py for the features, the gate and the audit).
 *
 * Network: 14 -> 256 -> 256 -> 256 -> 256 -> 2, GELU, BF16 weights and
 * activations on the matrix cores with FP32 accumulation, FP32 master
 * weights, Adam.  Loss: Gaussian negative log-likelihood
 * 0.5 ((y - mu) e^-s)^2 + s, mean over the batch.
 *
 * Functions return 0 (or a handle) on success; on failure -1 (or NULL) with
 * the reason in tl_mlp_last_error().  Calls on one handle are serialised. */
#ifndef TL_SURROGATE_H
#define TL_SURROGATE_H

#ifdef __cplusplus
extern "C" {
#endif

/* A new network with He-initialised weights (deterministic in seed). */
void* tl_mlp_new(unsigned long long seed);
void tl_mlp_delete(void* net);

/* Number of parameters in the logical order used by get/set_params:
 * W1 (14 x 256), b1 (256), W2 (256 x 256), b2, W3, b3, W4, b4, W5 (256 x 2),
 * b5 (2), each W row-major (input index first): 201730 values. */
long tl_mlp_param_count(void);
int tl_mlp_get_params(void* net, float* out);
int tl_mlp_set_params(void* net, const float* in);

/* Feature standardisation applied on the GPU: (x - mean) / scale, 14 each. */
int tl_mlp_set_scaling(void* net, const float* mean, const float* scale);

/* Inference: x is n rows of 14 features, out n rows of (mu, s). */
int tl_mlp_predict(void* net, int n, const float* x, float* out);

/* Inference that allocates nothing, for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid allocation).  No counterpart in the reference.  x (n x 14 features)
 * and out (n x (mu, s)) are DEVICE pointers the caller owns; the
 * activations live in a fixed array in the library, one per GPU, of
 * 4.4 MiB, held until the process ends.  n is at most 4096 (larger n is
 * rejected with an error).  Calls are serialised within a process, and the
 * array keeps the last call's contents.  The result equals tl_mlp_predict's
 * for the same rows.  Returns 0, or -1 with tl_mlp_last_error(). */
int tl_mlp_predict_noalloc(void* net, int n, const float* x, float* out);

/* One forward and backward pass over n rows (a multiple of 64): the gradient
 * of the mean loss in the logical parameter order, and the loss. */
int tl_mlp_grads(void* net, int n, const float* x, const float* y, float* grads, float* loss);

/* Adam training (beta 0.9, 0.999, eps 1e-8) for `steps` minibatches of
 * `batch` rows (a multiple of 64) drawn without replacement from the n rows,
 * reshuffled every pass (deterministic in seed).  losses (may be NULL)
 * receives each step's batch loss. */
int tl_mlp_train(void* net, int n, const float* x, const float* y, int steps, int batch, float lr,
                 unsigned long long seed, float* losses);

const char* tl_mlp_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
