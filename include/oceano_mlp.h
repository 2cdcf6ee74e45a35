/* oceano_mlp.h -- C ABI of the learned sub-grid closure network (liboceano_dg.so).
 *
 * No counterpart in the reference: oceano has no closure.  The network corrects the
 * momentum source of coarse cells (the source slot of ShallowWater.h:425-429, used at
 * OceanDG.h:734-736) and is coupled to a DG operator with oc_dg_set_closure (oceano_dg.h).
 *
 * Shape: n_in (<= 32) inputs -> 3 hidden layers of 64 (GELU, erf form) -> n_out (<= 32)
 * outputs.  Products run on BF16 matrix cores with FP32 accumulation; biases, activations,
 * the loss and Adam are FP32.  Inputs are normalised inside the kernels,
 * x_n = (x - x_mean) / x_std, and outputs de-normalised, y = y_n * y_scale + y_shift;
 * training minimises the mean squared error of the normalised outputs.
 *
 * Parameters (oc_mlp_n_params() FP32 values), each weight block row-major [fan_in][fan_out]
 * at its padded size: W1 [32][64], b1 [64], W2 [64][64], b2, W3 [64][64], b3, W4 [64][32],
 * b4 [32].  Pads of unused inputs / outputs are zero.
 *
 * All arrays passed to forward / gradients / train_step are device pointers, row-major
 * [n][n_in] (x) and [n][n_out] (t, y); calls are ordered on `stream`; status codes as in
 * oceano_dg.h.  A handle is not thread-safe. */
#ifndef OCEANO_MLP_H
#define OCEANO_MLP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oc_mlp oc_mlp;

int oc_mlp_create(int32_t n_in, int32_t n_out, uint64_t seed, oc_mlp **out);
void oc_mlp_destroy(oc_mlp *mlp);
int32_t oc_mlp_n_params(void);
int oc_mlp_dims(const oc_mlp *mlp, int32_t *n_in, int32_t *n_out);
/* host arrays of oc_mlp_n_params() floats; set_params also resets the Adam moments */
int oc_mlp_get_params(oc_mlp *mlp, float *params);
int oc_mlp_set_params(oc_mlp *mlp, const float *params);
/* host arrays: x_mean, x_std [n_in]; y_shift, y_scale [n_out] (std and scale > 0) */
int oc_mlp_set_normalization(oc_mlp *mlp, const float *x_mean, const float *x_std,
                             const float *y_shift, const float *y_scale);
/* host arrays lo, hi [n_out] (either may be null = unbounded): inference saturates the
 * normalised output to [lo, hi] before de-normalising (the trained range); training is
 * unaffected.  Default unbounded. */
int oc_mlp_set_output_range(oc_mlp *mlp, const float *lo, const float *hi);
/* y = network(x) */
int oc_mlp_forward(oc_mlp *mlp, int64_t n, const float *x, float *y, void *stream);
/* gradient of the loss over the batch; grad (host, n_params floats) and loss (host) may be
 * null; either one makes the call synchronise the stream */
int oc_mlp_gradients(oc_mlp *mlp, int64_t n, const float *x, const float *t, float *grad,
                     double *loss, void *stream);
/* one Adam step (bias-corrected) with the gradient of the last oc_mlp_gradients call */
int oc_mlp_adam(oc_mlp *mlp, float lr, float beta1, float beta2, float eps, void *stream);
/* gradients + Adam (beta1 0.9, beta2 0.999, eps 1e-8); loss as in oc_mlp_gradients */
int oc_mlp_train_step(oc_mlp *mlp, int64_t n, const float *x, const float *t, float lr,
                      double *loss, void *stream);

#ifdef __cplusplus
}
#endif

#endif
