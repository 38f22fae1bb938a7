/*
 * bastr_closure.h -- learned subgrid closure for coarse 2-D runs.  No
 * counterpart in the reference, which resolves every scale it keeps; the
 * filter is its own sharp spherical dealiasing (dealiasing(field, factor),
 * src/tool.F90:149-173) at the coarse grid's cutoff k_c = floor(n/3), and
 * the closure force enters where compute_ut2D forms the tendency
 * (src/solution.F90:841-846).
 *
 * A coarse run of the filtered field u_bar = P_c u (P_c keeps |k| < k_c) misses
 * the subgrid force
 *     tau = P_c N(u) - D_c N(u_bar),   N(u) = -u_j d_j u_i,
 * D_c the coarse run's dealiasing (the viscous terms commute with P_c and
 * cancel; a projection, if on, applies to both).  A per-point MLP
 * (36 -> 64 -> 64 -> 2, GELU, BF16 weights and MFMA operands, FP32
 * accumulation) models tau from the 3x3 stencil of the coarse u1, u2,
 * theta = d1 u1 + d2 u2 and omega = d1 u2 - d2 u1.  A handle with a network
 * attached adds its output to N(u_bar) at every RK3 stage (evaluated on each
 * stage's field, or once per step: bastr_closure_per_step).
 *
 * Conventions as in bastr.h (status codes, bastr_last_error, Fortran-order
 * fields; calls are synchronous).  Grids: square, a power of two in
 * [32, 4096]; the fine and coarse handles of one call share their precision.
 */
#ifndef BASTR_CLOSURE_H
#define BASTR_CLOSURE_H

#include <stdint.h>
#include "bastr.h"

#ifdef __cplusplus
extern "C" {
#endif

#define BASTR_CLOSURE_NPARAM 6658  /* W1[64][36] b1[64] W2[64][64] b2[64] W3[2][64] b3[2] */
#define BASTR_CLOSURE_NSCALE 6     /* scales of u1, u2, theta, omega, tau1, tau2 */

typedef struct bastr_pairs bastr_pairs;
typedef struct bastr_closure bastr_closure;

/*
 * The coarse state of a one-field handle `coarse` (n x n) becomes the sharp
 * filter of the one-field handle `fine` (N x N, N >= n): the modes
 * |k| < floor(n/3) of fine's normalised spectrum, zero elsewhere; its nstep
 * and time become fine's.  No entry dealias is pending afterwards.
 */
int bastr_filter_to(bastr_handle* fine, bastr_handle* coarse);

/*
 * Training pairs of one coarse grid n x n, on the device, room for `capacity`
 * snapshots.  bastr_pairs_add filters the field of the one-field handle
 * `fine` (as bastr_filter_to) and stores, in FP32, the coarse fields
 * (u1, u2, theta, omega) and the subgrid force (tau1, tau2) in physical space.
 * bastr_pairs_get copies snapshot i out: fields 4*n*n, target 2*n*n floats
 * (Fortran order, field-major); either may be NULL.
 */
int bastr_pairs_create(int32_t n, int32_t capacity, bastr_pairs** out);
int bastr_pairs_add(bastr_pairs* ps, bastr_handle* fine);
int bastr_pairs_count(const bastr_pairs* ps);
int bastr_pairs_get(const bastr_pairs* ps, int32_t i, float* fields, float* target);
int bastr_pairs_destroy(bastr_pairs* ps);

/*
 * A network with weights drawn from `seed` (uniform, +-sqrt(6/(fan_in+fan_out)),
 * zero biases) and unit scales.  Parameters and scales are host arrays of
 * BASTR_CLOSURE_NPARAM and BASTR_CLOSURE_NSCALE floats (the FP32 master copy;
 * the device keeps BF16 copies of W1, W2, W3).  bastr_closure_fit_scales sets
 * each scale to the root mean square of its field over the pairs; of the pairs
 * (u1, u2), (theta, omega), (tau1, tau2), a field below 1e-6 of its partner
 * (omega of a potential flow, which Burgers keeps potential) takes the
 * partner's scale.
 */
int bastr_closure_create(uint64_t seed, bastr_closure** out);
int bastr_closure_destroy(bastr_closure* net);
int bastr_closure_set_params(bastr_closure* net, const float* params, const float* scales);
int bastr_closure_get_params(const bastr_closure* net, float* params, float* scales);
int bastr_closure_fit_scales(bastr_closure* net, const bastr_pairs* ps);

/*
 * Loss over pairs first..first+count-1, every point:
 *   L = sum_points sum_c (out_c - tau_c / scale_c)^2 / (2 * points),
 * and its gradient with respect to the parameters (grad: NPARAM floats or
 * NULL), from the training kernels' forward and backward passes.
 */
int bastr_closure_loss_grad(bastr_closure* net, const bastr_pairs* ps, int32_t first, int32_t count, double* loss,
                            float* grad);
/*
 * `iters` Adam steps (beta 0.9, 0.999, eps 1e-8, learning rate lr), each on
 * `batch` random 32-point tiles of all pairs (a counter-based draw of `seed`);
 * losses: NULL or iters entries (the loss of each step's batch).
 */
int bastr_closure_train(bastr_closure* net, const bastr_pairs* ps, int32_t iters, int32_t batch, float lr,
                        uint64_t seed, double* losses);
/* The network's force (scale_c * out_c) on snapshot i of the pairs: 2*n*n floats. */
int bastr_closure_apply(bastr_closure* net, const bastr_pairs* ps, int32_t i, float* force);

/*
 * Attach a network to a one-field 2-D RK3 handle (square, n >= 32): every
 * later bastr_step adds its force, evaluated on each stage's field, to the
 * nonlinear term before that stage's forward transform (the dealiasing and
 * projection of the tendency then apply to it).  net = NULL detaches.  The
 * network must outlive the attachment; its parameters may change between
 * steps.
 */
int bastr_closure_attach(bastr_handle* h, bastr_closure* net);
/*
 * Cadence of the attached force.  per_step = 0 (the default): evaluated on
 * each RK3 stage's field.  per_step = 1: evaluated once per time step, on the
 * step's input field (stage A), and the same force added at the three stages
 * (the closure term held over the step: a third of the network evaluations).
 */
int bastr_closure_per_step(bastr_handle* h, int32_t per_step);

#ifdef __cplusplus
}
#endif

#endif /* BASTR_CLOSURE_H */
