/* bugdef_fno.h -- C-ABI of the absorbed-power surrogate network (gfx950).
 *
 * The reference has no counterpart: these entry points serve the coupled epoch
 * loop's learned replacement of the photon-transport step.  The network maps
 * per-voxel optical properties of a 64^3 tile (plus beam features) to the
 * log-scaled fluence of the tile; the engine multiplies by mu_a and the
 * incident irradiance to obtain absorbed power.
 *
 * Architecture: a 3-D Fourier neural operator, width 32, 4 Fourier layers,
 * 8 input channels, projection 32 -> 64 -> 1, kept modes (kx, ky, kz) with kx
 * and ky signed (half non-negative, half negative) and kz non-negative, over a
 * zero-padded transform of length np >= 64.  Activations are bf16, channel
 * mixes run on BF16 MFMA with FP32 accumulation, spectral transforms and the
 * per-mode mixing in FP32, parameters and optimiser state in FP32.
 *
 * Tensors passed as *_dev are DEVICE pointers (HBM, e.g. from hipMalloc or a
 * framework's device tensors) on the context's device; the context queues its
 * work on its own stream, so the caller makes the inputs ready before the call
 * and calls bd_fno_sync before reading outputs.  Layouts, channels last:
 *   x  : float32 [B][64][64][64][8]
 *   y  : float32 [B][64][64][64]
 *   t,w,gy : float32 [B][64][64][64]
 * Parameters are one flat float32 vector (order: lift W[8][32], b[32]; per
 * layer: spectral W[kx][ky][kz][32][32][re,im], W[32][32], b[32]; projection
 * P1[32][64], c1[64], p2[64], c2[1]).
 *
 * Status codes are those of bugdef.h (BD_OK, BD_EINVAL, BD_ENODEV, BD_EHIP).
 */
#ifndef BUGDEF_FNO_H
#define BUGDEF_FNO_H

#include <stdint.h>

#include "bugdef.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct bd_fno bd_fno;

typedef struct {
    int32_t kx, ky, kz; /* kept modes: kx, ky even (<= 64), 1 <= kz <= 8 */
    int32_t np;         /* transform length (>= 64; > 64 zero-pads the tile) */
} bd_fno_spec;

/* Create a network context on `device` for batches of up to max_batch tiles.
 * train != 0 also allocates gradients, AdamW state and backward buffers. */
int bd_fno_create(int device, const bd_fno_spec *spec, int max_batch, int train, bd_fno **out);
int bd_fno_destroy(bd_fno *f);
int bd_fno_num_params(const bd_fno *f, int64_t *n);

/* Host copies of the flat parameter vector (set also clears the AdamW state)
 * and of the last computed gradient. */
int bd_fno_set_params(bd_fno *f, const float *host, int64_t n);
int bd_fno_get_params(bd_fno *f, float *host, int64_t n);
int bd_fno_get_grads(bd_fno *f, float *host, int64_t n);

/* y = network(x) for B tiles (B <= max_batch). */
int bd_fno_forward(bd_fno *f, const float *x_dev, int B, float *y_dev);

/* Gradient of sum(gy * y) with respect to the parameters, for the inputs of
 * the last bd_fno_forward (same x, same B); training contexts only. */
int bd_fno_backward(bd_fno *f, const float *x_dev, const float *gy_dev, int B);

/* Weighted squared error L = sum(w (y - t)^2) / B (w sums to one per tile)
 * and its gradient gy = 2 w (y - t) / B; returns L (FP64, fixed-order sum). */
int bd_fno_loss(bd_fno *f, const float *y_dev, const float *t_dev, const float *w_dev, int B, float *gy_dev,
                double *loss);

/* One AdamW update of the parameters from the current gradient (bias-corrected
 * moments, decoupled weight decay). */
int bd_fno_adamw(bd_fno *f, double lr, double beta1, double beta2, double eps, double weight_decay);

/* Wait for the context's queued work. */
int bd_fno_sync(bd_fno *f);

/* ---- coupling to an engine context (bugdef.h) -------------------------------
 * The network sees a 64^3 window of the domain.  Supported beams: launched
 * along +z with azimuth 0 or pi (the drift of the collimated beam stays in
 * the x-z plane); others are rejected with BD_EINVAL.
 *
 * Window origin (voxel indices, written to origin[0..2]) laterally centred on
 * the beam axis with its top `margin` voxels above the first non-void voxel
 * on the axis; origin[3] receives that voxel's z index (Nz if none). */
int bd_sur_window(bd_ctx *ctx, const bd_beam *beam, int margin, int32_t *origin);

/* The tiled window set: the window of bd_sur_window, then windows stacked
 * above it along the beam axis, each overlapping the one below by 8 voxels,
 * until one reaches 8 voxels above the tissue surface around the crater (the
 * first non-void z over nine columns within +-28 voxels of the axis) or the
 * domain top; at most max_windows (1..4).  origins[4k..4k+3] as in
 * bd_sur_window, bottom window first; the count goes to *n_windows. */
int bd_sur_windows(bd_ctx *ctx, const bd_beam *beam, int margin, int max_windows, int32_t *origins,
                   int *n_windows);

/* Input features of the window (x_dev: float32 [64][64][64][8], channels
 * log1p(mu_a), log1p(mu_s/10), g, n-1, ballistic beam log-intensity,
 * sin(theta), spot/0.5mm - 1, ln(1 + optical depth along the beam)/ln(101))
 * from the packed transport optics.  The ballistic march starts at the
 * domain top, so channels 4 and 7 include what lies above the window. */
int bd_sur_features(bd_ctx *ctx, const bd_beam *beam, const int32_t *origin, float *x_dev);

/* The same for n windows (origins as from bd_sur_windows) into
 * x_dev [n][64][64][64][8], in one launch. */
int bd_sur_features_n(bd_ctx *ctx, const bd_beam *beam, int n, const int32_t *origins, float *x_dev);

/* bd_sur_features_n, but with mc_photons > 0 channel 5 carries the fluence of
 * the context's last MC run instead of sin(theta): ln(Phi/E0 + 1e-4)/ln(1e4) + 1,
 * Phi = L / (mc_photons V), E0 the beam's peak irradiance per unit power (the
 * scaling of bd_sur_targets), 0 outside the domain.  That run must have had
 * mc_photons photons and score_fluence set.  A network trained on these
 * features refines a low-photon MC field into the high-photon one. */
int bd_sur_features_mc(bd_ctx *ctx, const bd_beam *beam, int n, const int32_t *origins, float *x_dev,
                       int64_t mc_photons);

/* Training target t = ln(Phi/E0 + 1e-4) and loss weights (sum one) of the
 * window from the context's last MC tallies (bd_mc_run with score_fluence). */
int bd_sur_targets(bd_ctx *ctx, const bd_beam *beam, const int32_t *origin, float *t_dev, float *w_dev);

/* Compose the absorbed power from a prediction y_dev ([64][64][64]):
 * E0 mu_a (exp(y) - 1e-4)+ inside the window, blended over 8 voxels into the
 * field kept by bd_sur_keep_stale at faces inside the domain, that field
 * elsewhere.  into_q != 0 writes the context's Q (then run bd_epoch_q);
 * into_q == 0 writes a separate field that bd_sur_error compares with Q. */
int bd_sur_apply(bd_ctx *ctx, const bd_beam *beam, const int32_t *origin, const float *y_dev, int into_q);

/* The same over n windows (y_dev [n][64][64][64]): per voxel the ramp-weighted
 * mean of the windows holding it, sum_k r_k q_k / max(W, 1), plus
 * (1 - min(W, 1)) of the kept field, W = sum_k r_k (r_k: the 8-voxel face
 * ramps of window k at its faces inside the domain; overlapping ramps sum
 * to one).  The network's absorbed power is multiplied by `scale` (> 0): the
 * coupled loop passes the MC / network ratio of the total absorbed power from
 * its last judge epoch. */
int bd_sur_apply_n(bd_ctx *ctx, const bd_beam *beam, int n, const int32_t *origins, const float *y_dev,
                   int into_q, double scale);

/* Keep the current Q (an MC epoch's field) as the field outside the window. */
int bd_sur_keep_stale(bd_ctx *ctx);

/* Relative L2 error of the composed field against Q over non-ablated voxels:
 * out[0] per voxel, out[1] over 4x4x4 block sums (MC shot noise averaged). */
int bd_sur_error(bd_ctx *ctx, double *out);

/* The same comparison with the totals: out[0] per-voxel and out[1] block-4
 * relative L2 as bd_sur_error, out[2] the sum of Q and out[3] the sum of the
 * composed field, both over non-ablated voxels. */
int bd_sur_judge(bd_ctx *ctx, double *out);

/* A second kept field: copy the context's Q (typically just written by an MC
 * run of its own) into it. */
int bd_sur_keep_aux(bd_ctx *ctx);

/* Coarse amplitudes from the kept MC field: for each block of block^3 voxels
 * (block 1..8), the non-ablated voxels of the composed field (Q with
 * into_q != 0, else the judged field) are scaled by sum kept / sum composed
 * over the block's non-ablated voxels, when both sums are positive.  The
 * network then supplies the shape inside each block and a low-photon MC the
 * block totals. */
int bd_sur_correct(bd_ctx *ctx, int block, int into_q);

/* bd_sur_judge for the kept MC field against Q: with the kept field from an
 * independent MC of the same state this is the MC-vs-MC distance, the shot
 * noise any comparison with one MC run carries. */
int bd_sur_judge_aux(bd_ctx *ctx, double *out);

/* Plain device buffers for the tensors above, so a host program needs no other
 * GPU runtime: allocation on `device`, synchronous host<->device and
 * device<->device copies (bytes), and a float32 fill of n values. */
int bd_dev_alloc(int device, int64_t bytes, void **ptr);
int bd_dev_free(void *ptr);
int bd_dev_upload(void *dst_dev, const void *src_host, int64_t bytes);
int bd_dev_download(void *dst_host, const void *src_dev, int64_t bytes);
int bd_dev_copy(void *dst_dev, const void *src_dev, int64_t bytes);
int bd_dev_fill_f32(void *dst_dev, float value, int64_t n);

#ifdef __cplusplus
}
#endif
#endif /* BUGDEF_FNO_H */
