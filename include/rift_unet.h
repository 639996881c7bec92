/*
 * rift_unet.h -- C ABI of the U-Net flood surrogate on MI355X (gfx950).
 *
 * The reference has no learned component: none of these entry points
 * replaces a reference function.  They serve ensemble screening: a network
 * trained on maps the solver generates (maximum depth, arrival time) predicts
 * them for new scenarios, and the scenarios whose prediction lies near a
 * decision threshold are re-run with the solver (rift_kp.h).
 *
 * Network: `levels` levels of `base_channels << l` channels; per encoder
 * level two 3x3 convolutions with ReLU, 2x2 average pooling between levels;
 * per decoder level a 2x2 stride-2 transposed convolution, concatenation
 * with the encoder output of that level, two 3x3 convolutions with ReLU;
 * a 1x1 convolution head.  Weights and activations are bfloat16 on the
 * matrix cores with FP32 accumulation; the optimizer keeps FP32 master
 * weights (Adam).
 *
 * Conventions as in rift_kp.h: RIFT_OK or a negative RIFT_E* code, message
 * in rift_last_error().  Inputs and outputs of the compute calls are DEVICE
 * pointers to planar FP32 arrays [batch][channels][height][width], row 0 of
 * an image = its first row in memory; work runs on the handle's stream and
 * the calls return after it completes.
 *
 * Canonical parameter vector (get/set_params, gradient): layer by layer --
 * enc0a, enc0b, ..., enc{L-1}a, enc{L-1}b, then for l = L-2..0 up{l},
 * dec{l}a, dec{l}b, then head -- each weight followed by its bias.
 * conv3x3 weight [cout][9][cin] (tap 3*ky+kx reads pixel (i+ky-1, j+kx-1),
 * zero padded); transposed conv [4][cout][cin] (sub-pixel 2a+b writes
 * (2i+a, 2j+b)); head [out_channels][base_channels].  A decoder's first
 * convolution reads [up, skip] concatenated in that channel order.
 */
#ifndef RIFT_UNET_H
#define RIFT_UNET_H

#include "rift_kp.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rift_unet rift_unet;

typedef struct rift_unet_desc {
    int height, width;        /* multiples of 2^(levels-1)                     */
    int in_channels;          /* 1..32                                         */
    int out_channels;         /* 1..32                                         */
    int base_channels;        /* multiple of 32, base << (levels-1) <= 512     */
    int levels;               /* 2..5                                          */
    int max_batch;            /* largest batch of any later call               */
    unsigned long long seed;  /* He-normal initial weights, zero biases        */
    int device;
} rift_unet_desc;

int rift_unet_create(const rift_unet_desc* desc, rift_unet** out);
int rift_unet_destroy(rift_unet* net);

/* Number of canonical parameters. */
int rift_unet_param_count(const rift_unet* net, long* n);
int rift_unet_get_params(rift_unet* net, float* host);
int rift_unet_set_params(rift_unet* net, const float* host);

/* Weights of the output channels in the loss (default 1 each):
 * loss = sum_c w_c * mean over pixels of (y_c - target_c)^2. */
int rift_unet_set_loss_weights(rift_unet* net, const float* w);
/* Adam hyper-parameters (defaults 0.9, 0.999, 1e-8); resets the moments. */
int rift_unet_set_adam(rift_unet* net, float beta1, float beta2, float eps);

/* y = net(x).  x [batch][in_channels][H][W], y [batch][out_channels][H][W]. */
int rift_unet_forward(rift_unet* net, const float* x, int batch, float* y);

/* Forward, loss and backward without an update: the loss and the gradient
 * of the loss with respect to the canonical parameters (host array). */
int rift_unet_gradient(rift_unet* net, const float* x, const float* target, int batch, double* loss,
                       float* grad_host);

/* One Adam step on the batch; *loss (may be NULL) = loss before the step. */
int rift_unet_train_step(rift_unet* net, const float* x, const float* target, int batch, float lr,
                         double* loss);

/* Algorithmic FLOPs of one forward pass and one training step at `batch`
 * (multiply-add = 2). */
int rift_unet_flops(const rift_unet* net, int batch, double* forward, double* train);

#ifdef __cplusplus
}
#endif

#endif /* RIFT_UNET_H */
