/* A world model's planner on the card: model-predictive path integral control (MPPI) over a learned ODE, C ABI.
 *
 * Replaces the reference's mppi_rollout / mppi_update kernels (src/field/mind/WorldModel.cpp:192-240, driven by
 * WorldModel::plan, :521-595): for every agent at once, `samples` futures of `horizon` steps each, from its
 * observation s (F features), under actions a (2 a step, |a_i| <= 1) drawn around its current plan with Gaussian noise
 * that is a hash of (agent id, tick, future, step, axis) -- the same world plans the same way; future 0 is the plan
 * itself. Each step integrates the learned dynamics ds/dt = f(s, a) (an MLP, F + 2 -> hidden -> hidden -> F, tanh) as
 * s += dt * f(s, a) * scale, and adds the step's cost to the future's; the new plan is the futures' actions weighted by
 * exp(-(J - J_min) / lambda), lambda a quarter of the cost spread (the reference's rule), summed in a fixed order.
 *
 * This build's differences, by design: the dynamics run on the matrix cores (fp16 operands, fp32 accumulation,
 * v_mfma_f32_16x16x16_f16, 16 futures a wave), not one thread a future; the noise is a hash of this build's own
 * (splitmix64's finaliser over its own key, then Box-Muller), so plans are reproducible here but not the reference's
 * bit for bit; the cost is a weighted quadratic about a
 * target, sum_f w_f (s_f - target_f)^2 + action_weight |a|^2 a step, set by the caller (the reference generates the
 * cost from a creature's `plan` line, a part of the law language this build's front end leaves to the reference).
 *
 * Weights (fe_mppi_set_dynamics) in the .net layout: per layer W[out][in] row-major then b[out]; the inputs are
 * normalised as (x - in_shift) * in_scale before the first layer. Entry points return 0, or -1 with the reason in
 * fe_mppi_error(). They wait for the card. */
#ifndef FE_MPPI_H
#define FE_MPPI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fe_mppi fe_mppi;

/* features 1..62, hidden 16..128 (a multiple of 16), horizon 1..64, samples 16..1024 (a multiple of 16). */
fe_mppi* fe_mppi_create(int features, int hidden, int horizon, int samples, char* why, size_t why_bytes);
void fe_mppi_free(fe_mppi* p);
const char* fe_mppi_error(fe_mppi* p);
/* weights: (F + 2) * hidden + hidden + hidden * hidden + hidden + hidden * F + F floats; in_shift, in_scale: F + 2;
 * scale: F (what a predicted rate is multiplied by before it moves the state). */
int fe_mppi_set_dynamics(fe_mppi* p, const float* weights, const float* in_shift, const float* in_scale, const float* scale);
int fe_mppi_set_cost(fe_mppi* p, const float* target, const float* weight, float action_weight);
/* One planning pass for `agents` agents: obs agents x F, ids agents, mean agents x horizon x 2 (in: the plans to start
 * from; out: the new plans). cost_out (agents x samples, may be NULL): every future's cost. */
int fe_mppi_plan(fe_mppi* p, int agents, const float* obs, const uint32_t* ids, uint64_t tick, float sigma, float dt, float* mean,
    float* cost_out);

#ifdef __cplusplus
}
#endif

#endif
