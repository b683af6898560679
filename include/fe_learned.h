/* The learned flux correction for a coarse 2-D two-field lattice, on the card: C ABI.
 *
 * A small net (18 inputs: the 3x3 patch of both fields around a cell in the frame of one of its faces; 64, 64 hidden,
 * GELU; 2 outputs, one flux a field) gives each face a flux F = s (g(P) - g(mirror P')) / 2, antisymmetric under the
 * mirror through the face; a cell's correction is (inflow - outflow) / h. bf16 weights and activations on MFMA, fp32
 * accumulation. The plain law it corrects: du = Du lap u + u - u^3 - k - v, dv = Dv lap v + u - v (5-point Laplacian,
 * zero flux through the lattice's edges). Training: the loss is the mean squared gap between the correction and what
 * a coarsened fine run says the coarse law missed, ((s1 - s0) / dt - plain(s0)); gradients by the net's own backward
 * kernels, Adam.
 *
 * The reference's nets are inference-only here: `flux`/`mlp`/`ldiffuse` evaluate fp16 mma fragments on the card
 * (src/field/cuda/CudaNet.cpp:189-277) and law nets train in JAX (bench/model2_front.py:136-151); its GPU training
 * exists only for the world model (src/field/WorldModel.cpp:180-252). No entry point below has a direct counterpart.
 *
 * Lattices: `count` lattices of ny rows of nx cells, field u then field v each count * nx * ny device floats.
 * Parameters (fe_lnet_param_count() = 5506 floats): W1 [18][64], b1[64], W2 [64][64], b2[64], W3 [64][2], b3[2].
 * Entry points return 0, or -1 with the reason in fe_lnet_error(). Work is queued on the library's stream. */
#ifndef FE_LEARNED_H
#define FE_LEARNED_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fe_lnet fe_lnet;

fe_lnet* fe_lnet_create(const float* params, float scale, char* why, size_t why_bytes);
void fe_lnet_free(fe_lnet* n);
const char* fe_lnet_error(fe_lnet* n);
int fe_lnet_param_count(void);
int fe_lnet_get_params(fe_lnet* n, float* out);
int fe_lnet_set_params(fe_lnet* n, const float* in);
/* Face fluxes of every face: faces_out[(2 * cell + dir) * 2 + field], dir 0 the +x face, 1 the +y face (0 across the
 * lattice's edge). */
int fe_lnet_faces(fe_lnet* n, const float* u, const float* v, int nx, int ny, int count, double h, float* faces_out);
/* One explicit step of dt out of place (u2, v2 distinct from u, v): the plain law, plus the correction when `learned`.
 * A learned step that leaves any value non-finite or beyond `bound` is replaced by the plain law's step on the card
 * (counted: fe_lnet_fallbacks). */
int fe_lnet_step(fe_lnet* n, const float* u, const float* v, int nx, int ny, int count, double h, double du, double dv, double k,
    double dt, int learned, double bound, float* u2, float* v2);
/* fe_lnet_step with learned = 1, allocating nothing: no counterpart in the reference; for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops). The face fluxes and the breach counter live in a
 * static device array of the library (16 MiB, one per device, held until the process ends) instead of the net's
 * lazily grown buffers, so count * nx * ny may be at most 2^20 cells: larger batches are refused with an error (use
 * fe_lnet_step). Every call on one device shares that array: calls that overlap on one device (other streams, other
 * threads) race. Same kernels, same bits as fe_lnet_step; fallbacks are counted in the net as there. */
int fe_lnet_step_noalloc(fe_lnet* n, const float* u, const float* v, int nx, int ny, int count, double h, double du, double dv,
    double k, double dt, double bound, float* u2, float* v2);
int fe_lnet_fallbacks(fe_lnet* n, unsigned* out);
/* The loss's gradient for the pair (s0 -> s1 over dt) into the net's gradient buffer (and grad_out, loss_out when not
 * NULL: then it waits for the card). The loss is `weight` times the mean squared miss. Fixed-order sums: the same bits
 * every run. */
int fe_lnet_gradient(fe_lnet* n, const float* u0, const float* v0, const float* u1, const float* v1, int nx, int ny, int count, double h,
    double du, double dv, double k, double dt, double weight, float* grad_out, float* loss_out);
/* Adam on the last gradient (beta 0.9, 0.999, eps 1e-8), each update clipped to +-clip; repacks the fragments. */
int fe_lnet_adam(fe_lnet* n, double lr, double clip);
/* Block averages of `factor` x `factor` cells, lattice by lattice (coarse: count * (fnx/factor) * (fny/factor)). */
int fe_lnet_coarsen(const float* fine, int fnx, int fny, int count, int factor, float* coarse);

#ifdef __cplusplus
}
#endif

#endif
