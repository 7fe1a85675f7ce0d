/* vpfc.h - learned coarse-velocity closure for the VPFP step on MI355X.
 *
 * No counterpart in the reference: VlaPy has no learned component.  The
 * closure corrects an under-resolved coarse velocity grid (nv = 256) after
 * every step of the plain solver:
 *
 *     f(x, .) <- f(x, .) + out_scale * net(in_scale * f(x, .))   for every x row
 *
 * net: four 1-D convolutions over v, kernel 5, zero padding 2, 64 channels,
 *     1 -> 64 (tanh) -> 64 (tanh) -> 64 (tanh) -> 1 (linear).
 * Arithmetic: every weight, activation and gradient that enters a product is
 * rounded to BF16 (round to nearest even); products accumulate in FP32; the
 * forward pass runs all four layers on MFMA (v_mfma_f32_16x16x32_bf16), the
 * backward pass the input and weight gradients of the two 64 -> 64 layers;
 * tanh is evaluated to within 4e-7 relative
 * before its BF16 rounding; post-tanh activations are stored in BF16 and
 * tanh' = 1 - a^2 uses the stored value.
 *
 * Flat FP32 parameter vector (VPFC_NPARAMS = 41793 floats):
 *   W1 [64][5]   (c, k)                 offset 0
 *   b1 [64]                             320
 *   W2 [64][320] (co, k*64 + ci)        384
 *   b2 [64]                             20864
 *   W3 [64][320]                        20928
 *   b3 [64]                             41408
 *   W4 [320]     (k*64 + ci)            41472
 *   b4 [1]                              41792
 * so that layer l computes out[co][v] = b[co] + sum_{k, ci} W[co][k*64+ci] in[ci][v+k-2].
 *
 * All arrays passed to the entry points are device pointers (rows of 256
 * FP32 values, row-major) unless stated otherwise; calls enqueue on `stream`.
 * Return codes are those of vpfp.h; vpfc_last_error() holds the message.
 */
#ifndef VPFC_H
#define VPFC_H

#include "vpfp.h"

#ifdef __cplusplus
extern "C" {
#endif

#define VPFC_NV 256
#define VPFC_CHANNELS 64
#define VPFC_KERNEL 5
#define VPFC_NPARAMS 41793

typedef struct vpfc_net_s* vpfc_net_t;

/* A network with workspace for training batches of up to max_rows rows
 * (inference takes any number of rows).  The compute calls take a stream
 * each and the network does not track them, so destroy and the setters
 * below wait for the whole device (setup-time calls, not for a step loop). */
int vpfc_create(int max_rows, vpfc_net_t* out);
int vpfc_destroy(vpfc_net_t net);

/* Parameters from / to host memory (FP32, the flat layout above);
 * synchronous.  Setting them resets the Adam moments and step count. */
int vpfc_set_params(vpfc_net_t net, const float* host_params);
int vpfc_get_params(vpfc_net_t net, float* host_params);
int vpfc_set_scales(vpfc_net_t net, float in_scale, float out_scale);

/* Background profile bg [256] (host, FP32; NULL clears it).  With a
 * background every row is seen relative to it and to its own size:
 *     d = f - bg,  s = sqrt(mean_v d^2 + 1e-30),  input in_scale * d / s,
 * the correction applied is out_scale * s * net(...) and the training target
 * is compared in units of s (target / s).  Without one, d = f and s = 1. */
int vpfc_set_background(vpfc_net_t net, const float* host_bg);

/* Conservation projection of the correction (host arrays v [256], bg
 * [256], FP64; v == NULL clears it): the network output y is replaced by
 *     y - bg(v) (c0 + c1 v + c2 v^2)
 * with c chosen so that the trapz moments sum_v w_v v^l y (l = 0, 1, 2;
 * w = 1, ends 1/2) vanish: the closure then adds no density, momentum or
 * kinetic energy to a row (to FP32 rounding).  Applied in vpfc_forward and
 * vpfc_apply; vpfc_gradients and vpfc_train_step see the unprojected
 * output (train on projected targets). */
int vpfc_set_conservation(vpfc_net_t net, const double* v, const double* bg);

/* y = net(in_scale * d / s) for `rows` rows (d, s as above) */
int vpfc_forward(vpfc_net_t net, const float* x, int rows, float* y, void* stream);

/* f <- f + out_scale * s * net(in_scale * d / s), in place, `rows` rows */
int vpfc_apply(vpfc_net_t net, float* f, int rows, void* stream);

/* Loss L = mean over rows and v of (net(in_scale * d / s) - target / s)^2 and its
 * gradient with respect to the flat parameters: grad [VPFC_NPARAMS] and
 * loss [1] are device arrays.  rows <= max_rows. */
int vpfc_gradients(vpfc_net_t net, const float* x, const float* target, int rows, float* grad, float* loss,
                   void* stream);

/* One Adam step (beta1, beta2, eps as in Kingma & Ba, bias-corrected) on
 * the gradient of vpfc_gradients over x, target; loss [1] may be NULL. */
int vpfc_train_step(vpfc_net_t net, const float* x, const float* target, int rows, float lr, float beta1,
                    float beta2, float eps, float* loss, void* stream);

/* Attach a closure to a batch of the parameter scan (vpfp.h): every step
 * of vpfp_batch_step_n is then followed by vpfc_apply on all n_inst * nx
 * rows, and the next step's first half kick runs as a launch of its own
 * instead of being fused after the collision.  The batch must be FP32 with
 * nv == VPFC_NV; net == NULL detaches.  The net must outlive its use. */
int vpfp_batch_set_closure(vpfp_batch_t batch, vpfc_net_t net);

const char* vpfc_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* VPFC_H */
