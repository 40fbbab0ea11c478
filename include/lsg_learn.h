/*
 * lsg_learn.h - learned preconditioner of the MI355X Lippmann-Schwinger solver.
 *
 * No reference counterpart: the reference's only preconditioner is HODLR
 * (Hybrid/, out of scope). A small convolutional operator on a uniform G x G
 * grid maps (q, Re b^, Im b^) to the correction (Re, Im) of x^ - b^,
 * x^ ~ A^-1 b^ for a right-hand side b^ normalised to unit RMS; the
 * preconditioner is M(v) = v + sigma * correction(v / sigma), sigma = RMS(v).
 * It supplies the initial guess and the right preconditioner of flexible
 * GMRES. The last layer starts at zero, so an untrained network is M = I. Training and inference run in
 * hand-written HIP kernels (BF16 weights and activations, FP32 accumulation
 * and FP32 master weights with Adam).
 *
 * Layout: activations NHWC (sample, row y, column x, channel), weights
 * [cout][ky][kx][cin] with kernel offsets (ky-1, kx-1) * dilation, zero padding.
 * All layers but the last apply ReLU.
 */
#ifndef LSG_LEARN_H_
#define LSG_LEARN_H_

#include <stdint.h>

#include "lsg.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct lsgn_net lsgn_net;

#define LSGN_MAX_LAYERS 8

/* Network of n_layers 3x3 convolutions: 3 input channels, `width` hidden
 * channels, 2 outputs; dilations[l] per layer; weights drawn from a seeded
 * He-normal stream, biases zero. Grid side G (multiple of 16, <= 512) and the
 * largest training batch are fixed at creation. */
int lsgn_create(int n_layers, int width, const int* dilations, int grid, int max_batch, uint64_t seed,
                lsgn_net** out);
void lsgn_destroy(lsgn_net* net);
const char* lsgn_last_error(const lsgn_net* net);

/* FP32 master parameters, layer by layer: weights then biases. */
int64_t lsgn_num_params(const lsgn_net* net);
int lsgn_get_params(const lsgn_net* net, float* out);
int lsgn_set_params(lsgn_net* net, const float* in);

/* Host-buffer forward: in B x G x G x 3, out B x G x G x 2 (FP32 in host
 * memory; BF16 inside). */
int lsgn_forward(lsgn_net* net, const float* in, int batch, float* out);
/* Host-buffer loss and gradient: loss = mean over B G G 2 of (out - target)^2;
 * grad receives d loss / d params in the lsgn_get_params order. */
int lsgn_loss_grad(lsgn_net* net, const float* in, const float* target, int batch, float* loss, float* grad);

/* Training set resident on the device: n samples of inputs (G G 3) and
 * targets (G G 2), FP32 host arrays copied once. */
int lsgn_set_training_data(lsgn_net* net, const float* in, const float* target, int n);
/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8) on samples idx[0..batch)
 * of the resident set; loss (before the step) returned in *loss. */
int lsgn_train_step(lsgn_net* net, const int* idx, int batch, float lr, float* loss);

/* Attach the network to a solver context as its learned preconditioner: the
 * grid maps (Chebyshev interpolation leaf nodes -> grid, bilinear grid ->
 * nodes) are built from the context's tree; qk2/k^2 of the context (or of
 * column 0 of lsg_set_contrast) is the q channel. net = NULL detaches. The
 * network must outlive the attachment. M(v) = v + sigma P_nodes(net(q,
 * P_grid(v)/sigma)). */
int lsg_attach_preconditioner(lsg_ctx* ctx, lsgn_net* net);
/* z = M(v) for host vectors of N complex128 (test hook). */
int lsg_apply_preconditioner(lsg_ctx* ctx, const double* v, double* z);
/* Node <-> grid maps of the context for host arrays (test hook):
 * to_grid: v[N] complex -> g[G G] complex; to_nodes: g -> v. */
int lsg_grid_resample(lsg_ctx* ctx, const double* in, double* out, int to_grid);

#define LSG_PREC_NONE 0
#define LSG_PREC_GUESS 1          /* x0 = M(b) when x is zero on entry */
#define LSG_PREC_RIGHT 2          /* flexible GMRES with right preconditioner M */
/* Flexible GMRES (FP64 operator, FP64 Gram-Schmidt) with the attached
 * network: flags = LSG_PREC_GUESS | LSG_PREC_RIGHT. With LSG_PREC_RIGHT the
 * network preconditions the first restart cycle; the remaining cycles are
 * plain GMRES from the current iterate. The true FP64 residual decides
 * convergence as in lsg_gmres_ex. */
int lsg_gmres_learned(lsg_ctx* ctx, const double* b, double* x, const lsg_solve_opts* opts, int flags,
                      lsg_stats* stats);

#ifdef __cplusplus
}
#endif

#endif /* LSG_LEARN_H_ */
