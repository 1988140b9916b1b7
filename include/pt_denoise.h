/*
 * pt_denoise.h -- learned denoiser for the progressive render (no
 * counterpart in the reference; the plain progressive render stays the
 * fallback and the measure of its error).
 *
 * A 5-level convolutional encoder-decoder with skip connections (3x3
 * kernels, 32-48-64-80-96 channels, 2x2 average pooling, nearest
 * upsampling), over 9 input channels per pixel: log1p of the linear
 * radiance estimate, first-hit albedo and shading normal (the renderer's
 * feature buffers, pt_renderer_enable_aovs).  It predicts a correction to
 * the input's log radiance; the output is expm1 of the sum.  Weights and
 * activations are bf16 with FP32 accumulation on MFMA; master weights and
 * the Adam state are FP32.  Loss: mean |y - log1p(t)| + mean
 * (expm1(y) - t)^2 / (t^2 + 0.01) over pixels and channels.
 *
 * Feature arrays are [n][h][w][9] float (radiance rgb, albedo rgb, normal
 * xyz); radiance and targets are [n][h][w][3].  Training and test shapes
 * must be multiples of 16; ptd_denoise_renderer pads the frame itself.
 * Errors: a pt_status, with the message in ptd_last_error().
 */
#ifndef PT_DENOISE_H
#define PT_DENOISE_H

#include "pt_render.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ptd_network ptd_network;

const char *ptd_last_error(void);
/* Number of FP32 master parameters (weights [co][tap][ci] then biases, layer by layer). */
int64_t ptd_parameter_count(void);
/* A network with seeded He-normal weights (output layer scaled down, so it starts near the identity). */
pt_status ptd_create(uint64_t seed, ptd_network **out);
void ptd_destroy(ptd_network *net);
pt_status ptd_get_parameters(ptd_network *net, float *host, int64_t count);
/* Replaces the weights and clears the optimizer state. */
pt_status ptd_set_parameters(ptd_network *net, const float *host, int64_t count);

/* Denoised radiance of host features. */
pt_status ptd_forward(ptd_network *net, const float *features, int32_t n, int32_t h, int32_t w, float *radiance);
/* Loss and its gradient with respect to every parameter (no update). */
pt_status ptd_loss_and_gradients(ptd_network *net, const float *features, const float *targets, int32_t n, int32_t h,
                                 int32_t w, float *loss, float *gradients, int64_t count);

/* Training crops resident on the device: features [crops][h][w][9], targets [crops][h][w][3]. */
pt_status ptd_set_training_data(ptd_network *net, const float *features, const float *targets, int32_t crops,
                                int32_t h, int32_t w);
/* `steps` Adam steps; step s trains on crops crop_ids[s*batch .. s*batch+batch-1].
 * losses[s] receives step s's loss (may be NULL); *milliseconds the device time. */
pt_status ptd_train(ptd_network *net, const int32_t *crop_ids, int32_t batch, int32_t steps, float learning_rate,
                    float *losses, float *milliseconds);

/* Denoise a renderer's current accumulation (its feature buffers enabled),
 * on the device: radiance receives width*height*3 floats in the layout of
 * pt_renderer_read_image; *milliseconds the device time of the denoiser. */
pt_status ptd_denoise_renderer(ptd_network *net, const pt_renderer *renderer, float *radiance, float *milliseconds);

#ifdef __cplusplus
}
#endif
#endif
