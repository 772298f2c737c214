/* Test infrastructure, not product code. The C interface of the two CPU
 * oracles, loaded by the tests (and by the CPU-baseline leg of bench.py)
 * through ctypes:
 *
 *   oa_*  Oracle A: a double-precision restatement of the reference's CPU
 *         reference, src/train/reference.cpp (every function cites the lines
 *         it follows).
 *   ob_*  Oracle B: a single-precision model of this project's GPU pipeline,
 *         operation for operation, which the GPU must match bit for bit on
 *         everything integer (sort, binning, tile ranges, stops, fixed-point
 *         sums) and on the floats that feed them.
 *
 * Every function returns 0 on success and a negative value on bad input.
 */
#pragma once
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* GaussianCloud (src/splat/gaussian_cloud.hpp:21-38): raw parameters. */
typedef struct {
  const float *positions; /* 3 per Gaussian */
  const float *scales;    /* 3, log standard deviations */
  const float *rotations; /* 4, quaternion w x y z, not normalised */
  const float *opacities; /* 1, logit */
  const float *sh;        /* 3 (degree+1)^2, [coefficient][rgb] */
  uint32_t count;
  int32_t sh_degree;
} oc_cloud;

/* Camera (src/render/camera.hpp:9-18); view is column-major like glm. */
typedef struct {
  float view[16];
  float position[3];
  float fx, fy, cx, cy;
  int32_t width, height;
} oc_camera;

/* ReferenceOptions (src/train/reference.hpp:18-33). */
typedef struct {
  float background[3];
  float near_plane;
  int32_t no_cutoffs;
  float opacity_reg, scale_reg;
  int32_t mip_filter_2d;
  const float *filter3d; /* NULL or one per Gaussian */
  uint32_t filter3d_count;
} oa_options;

int oa_render(const oc_cloud *cloud, const oc_camera *cam, const oa_options *opt, int threads, double *rgb, double *log_t,
              uint64_t *stop);
int oa_contributions(const oc_cloud *cloud, const oc_camera *cam, const oa_options *opt, int threads, double *out);
/* reference.cpp:248-272: per Gaussian, the composited pixels whose mean error is
 * at least share of the view's worst (target RGBA8, width*height*4 bytes). */
int oa_marked_pixels(const oc_cloud *cloud, const oc_camera *cam, const uint8_t *target, double share, const oa_options *opt,
                     int threads, double *out);
int oa_loss(const double *rgb, const uint8_t *target, int width, int height, double lambda, int threads, double *adjoint,
            double *loss_l1_ssim);
int oa_gradients(const oc_cloud *cloud, const oc_camera *cam, const double *adjoint, const oa_options *opt, int threads,
                 double *positions, double *scales, double *rotations, double *opacities, double *sh,
                 double *absolute_centre);

/* reference.cpp:33-38 and 40-67, exposed for the closed-form tests. */
int oa_quat_to_mat(const double *q /* w x y z, unit */, double *R /* row-major 3x3 */);
int oa_sh_basis(const double *dir, int degree, double *out /* 16 */);

/* ---- Oracle B: the FP32 tile model ---- */

/* The per-view constants the GPU kernels read, as the product's camera block. */
typedef struct {
  float view[16];
  float position[3];
  float near_plane;
  float fx, fy, cx, cy;
  float tan_half_x, tan_half_y; /* 0.5 W / fx, 0.5 H / fy, computed in float */
  uint32_t width, height;
  uint32_t count, sh_degree, sh_words, sh_coefficients;
  uint32_t sh_format; /* 0: DC float + rest halves (the reference's layout); 1: all bf16 */
} ob_camera;

typedef struct {
  int32_t mip_filter_2d;
  float sh_dropout;
  uint32_t sh_retained, sh_seed;
} ob_preprocess_options;

int ob_preprocess(const float *positions, const float *scales, const float *rotations, const float *opacities,
                  const uint32_t *sh_words, const float *filter3d, const ob_camera *cam, const ob_preprocess_options *opt,
                  float *splats /* 12 per Gaussian */, uint32_t *keys, uint32_t *values);

/* Stable sort of (key, value) by key, as std::stable_sort. */
int ob_stable_sort(uint32_t *keys, uint32_t *values, uint32_t count);

/* Tile coverage per depth-sorted splat, and the entries in depth order. With
 * entry_keys == NULL only counts; returns totals in counts[0] visible,
 * counts[1] entries. */
int ob_bin(const uint32_t *keys, const uint32_t *values, const float *splats, uint32_t count, uint32_t grid_w,
           uint32_t grid_h, uint32_t *coverage, uint32_t *entry_keys, uint32_t *entry_values, uint64_t entry_capacity,
           uint64_t *counts);

int ob_tile_ranges(const uint32_t *entry_keys, uint64_t entries, uint32_t tiles, uint32_t *ranges /* 2 per tile */);

int ob_forward(const float *splats, const uint32_t *entry_values, const uint32_t *ranges, uint32_t width, uint32_t height,
               const float *background, float *image /* 4 per pixel */, uint32_t *stops);

/* L1 + SSIM: the SSIM fields pass and the adjoint pass. tile_totals gets 2
 * floats per tile (SSIM total, L1 total). */
int ob_loss(const float *image, const uint8_t *target, uint32_t width, uint32_t height, float lambda, float *fields,
            float *adjoint /* 4 per pixel */, float *tile_totals);

/* Backward with fixed-point sums: 14 int64 per Gaussian, accumulated from the
 * quantised wave totals. fixed_scale as the product computes it. */
int ob_backward_fixed(const float *splats, const uint32_t *entry_values, const uint32_t *ranges, uint32_t width,
                      uint32_t height, const float *background, const float *image, const float *adjoint,
                      const uint32_t *stops, float fixed_scale, uint32_t count, int64_t *raw,
                      const float *edge /* a scored view's edge weights, else NULL */);
/* The same sums in float per wave total, added in a deterministic order; the
 * GPU's float atomics add them in some order, so this is for tolerance checks. */
int ob_backward_float(const float *splats, const uint32_t *entry_values, const uint32_t *ranges, uint32_t width,
                      uint32_t height, const float *background, const float *image, const float *adjoint,
                      const uint32_t *stops, uint32_t count, float *raw, const float *edge);
/* Fixed-point sums to floats, as the chain-rule kernel reads them. */
int ob_fixed_to_float(const int64_t *raw, uint64_t n, float fixed_scale, float *out);

typedef struct {
  float opacity_reg, scale_reg; /* already divided by the Gaussian count */
  int32_t mip_filter_2d;
  float sh_dropout;
  uint32_t sh_retained, sh_seed;
  uint32_t mode;  /* 0 statistics, 1 gradients, 2 step, 3 curvature */
  uint32_t flags; /* 1 statistics, 4 selective, 16 second order */
  float rates[5];
  float rest_scale, beta1, beta2, epsilon, correction1, correction2;
  uint32_t seed, sh_floats;
  /* Second order (train_step.glsl:44-67): the trust region's epsilon, the
   * curvature's bias correction 1 / (1 - beta^k), beta, and (curvature mode)
   * 1 / the values in the image. */
  float trust_region, curvature_scale, curvature_beta, per_value;
  /* When set (gradients and step modes): 6 floats per Gaussian, t x v then
   * v = Rv dL/dp with t the camera-space position (train_gradient.comp:276-279);
   * left as they are for Gaussians no pixel touched. */
  float *pose_terms;
} ob_chain_options;

/* Chain rule from the 14 raw float sums per Gaussian. In gradients mode
 * writes the five gradient arrays; in step mode updates values and state in
 * place (SH through the packed words) and raw is rewritten for the SH pass.
 * stats: 8 floats per Gaussian (DensifyStat), updated when flags has 1. */
int ob_chain(float *positions, float *scales, float *rotations, float *opacities, uint32_t *sh_words,
             const float *filter3d, const ob_camera *cam, float *raw, const float *splats, float *stats,
             uint32_t *state_pos, uint32_t *state_scale, uint32_t *state_rot, uint32_t *state_opacity,
             uint32_t *state_sh, float *g_pos, float *g_scale, float *g_rot, float *g_opacity, float *g_sh,
             const ob_chain_options *opt);
/* SH's part of Adam, per packed word, after ob_chain in step mode. */
int ob_sh_step(uint32_t *sh_words, uint32_t *state_sh, const float *raw, const ob_camera *cam,
               const ob_chain_options *opt);

/* One Adam step of every parameter from summed gradients (the batched step). */
int ob_adam(float *positions, float *scales, float *rotations, float *opacities, uint32_t *sh_words, uint32_t *state_pos,
            uint32_t *state_scale, uint32_t *state_rot, uint32_t *state_opacity, uint32_t *state_sh, const float *g_pos,
            const float *g_scale, const float *g_rot, const float *g_opacity, const float *g_sh, uint32_t count,
            uint32_t sh_words_per, uint32_t sh_format, const ob_chain_options *opt);

/* The exp and log the model uses, for the math tests. */
void ob_exp_array(const float *x, float *y, uint64_t n);
void ob_log_array(const float *x, float *y, uint64_t n);
/* The dropout and rounding hash (32 bits, and as a float in [0, 1)) and the
 * two stochastic roundings, elementwise. */
void ob_hash_array(const uint32_t *index, uint32_t seed, uint64_t n, uint32_t *out);
void ob_uniform_array(const uint32_t *index, uint32_t seed, uint64_t n, float *out);
void ob_bf16_stochastic_array(const float *x, const uint32_t *index, uint32_t seed, uint64_t n, uint32_t *out);
void ob_half_stochastic_array(const float *x, const uint32_t *index, uint32_t seed, uint64_t n, uint32_t *out);

#ifdef __cplusplus
}
#endif
