/* A law net's trainer on the card, C ABI: the gradient of a regression loss over rows and a clipped Adam step, for any
 * net a law declares (sizes, activation).
 *
 * Replaces the reference's generated trainer NAME_train_rows / NAME_train_grad / NAME_adam
 * (src/field/cuda/CudaNet.cpp:279-369), which its world model drives to fit its dynamics net to observed transitions
 * (src/field/mind/WorldModel.cpp:395-417): for rows x (inputs) with targets t and row weights m, the net's output
 * y = out_scale * f(normalised x) and the loss sum over rows of 0.5 m |y - t|^2; the gradient of each weight and bias
 * is the mean over the rows; Adam (b1 0.9, b2 0.999, eps 1e-8) with each gradient clipped to +-clip first. The
 * normalisation (in_shift, in_scale) and out_scale are the caller's, not trained.
 *
 * This build's form: 16 rows a wave, the three products of every layer (forward, the deltas back, the weight
 * gradients) on the matrix cores in fp32 (v_mfma_f32_16x16x4f32), the activations and deltas of a tile in LDS, the
 * tiles' gradients summed in a fixed order (the same rows give the same bits). fp32 throughout, as the reference's.
 *
 * The weights travel in the reference's GPU net image (net_layout / net_image, src/field/cuda/CudaNet.cpp:97-136):
 * in_shift[in], in_scale[in], out_scale, zeros up to the next multiple of 4 floats, then per layer W[out][in] row-major
 * and b[out]. Entry points return 0, or -1 with the reason in fe_nettrain_error(); they wait for the card. */
#ifndef FE_NETTRAIN_H
#define FE_NETTRAIN_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fe_nettrain fe_nettrain;

/* sizes[0..count-1]: inputs, hidden widths, outputs (2..9 entries, each 1..128); act: the hidden layers' activation
 * (0 tanh, 1 relu, 2 silu, 3 linear, 4 gelu, tanh form; the last layer is linear), the Activation order of a law net.
 * The weights start at zero, the normalisation at (0, 1), out_scale at 1. */
fe_nettrain* fe_nettrain_create(const int* sizes, int count, int act, char* why, size_t why_bytes);
void fe_nettrain_free(fe_nettrain* t);
const char* fe_nettrain_error(fe_nettrain* t);
/* Floats of the image (2 in + 1 padded to a multiple of 4, then the weights and biases). */
int fe_nettrain_image_size(fe_nettrain* t);
int fe_nettrain_set(fe_nettrain* t, const float* image);
int fe_nettrain_get(fe_nettrain* t, float* image);
/* The loss (mean over rows) and, when grad is not NULL, the gradient of the weights and biases (the image's floats
 * from the first weight on, in its order) for rows x (rows x in), t (rows x out), m (rows, or NULL for 1 each). No
 * update. */
int fe_nettrain_gradient(fe_nettrain* t, const float* x, const float* target, const float* m, int rows, float* loss, float* grad);
/* One Adam step on those rows (the step count carries over calls); the loss before the step into *loss (may be NULL). */
int fe_nettrain_step(fe_nettrain* t, const float* x, const float* target, const float* m, int rows, float lr, float clip, float* loss);

#ifdef __cplusplus
}
#endif

#endif
