/* vsplat: an MI355X-native differentiable Gaussian-splat tile rasterizer.
 *
 * The C interface of the trainer's rasterizer path. It replaces the reference's
 * `class Trainer` (src/train/trainer.hpp:89-270) behind plain pointers and
 * sizes; include/vsplat_trainer.hpp wraps it back into that class shape with
 * the reference's exception types.
 *
 * Conventions: every function returning int returns VSPLAT_OK (0) or a
 * negative status; vsplat_last_error() then holds a message for the calling
 * thread. Host arrays are borrowed for the duration of the call unless noted.
 * A trainer is used by one thread at a time; each call that returns results is
 * synchronous, and the vsplat_enqueue_* calls only queue work on the trainer's
 * stream (vsplat_synchronize waits for it).
 */
#ifndef VSPLAT_H
#define VSPLAT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Statuses, and the reference exception each one stands for. */
#define VSPLAT_OK 0
#define VSPLAT_RUNTIME_ERROR (-1)    /* std::runtime_error: empty cloud, bad image size, target/camera mismatch (trainer.cpp:174, 276, 571) */
#define VSPLAT_BUDGET_EXCEEDED (-2)  /* splat::BudgetExceeded, recoverable (trainer.hpp:46-49, trainer.cpp:359) */
#define VSPLAT_INVALID_ARGUMENT (-3) /* std::invalid_argument (trainer.cpp:802), and null or malformed arguments */
#define VSPLAT_LOGIC_ERROR (-4)      /* std::logic_error: gradients before compute_gradients (trainer.cpp:853) */
#define VSPLAT_DEVICE_ERROR (-5)     /* a HIP call failed, or no GPU */
#define VSPLAT_UNSUPPORTED (-6)      /* a call this build does not take in that state (a batch with second order or an exposure) */

/* Trainer::kBytesPerTileEntry (trainer.hpp:194): entry keys, values and their sort copies. */
#define VSPLAT_BYTES_PER_TILE_ENTRY 16u

typedef struct vsplat_trainer vsplat_trainer;

/* GaussianCloud (src/splat/gaussian_cloud.hpp:21-38): raw parameters, arrays of
 * count*3, count*3, count*4, count and count*3*(degree+1)^2 floats. */
typedef struct {
  const float *positions;
  const float *scales;    /* log standard deviations */
  const float *rotations; /* quaternion w x y z, not necessarily normalised */
  const float *opacities; /* logit */
  const float *sh;        /* [Gaussian][coefficient][rgb] */
  uint32_t count;
  int32_t sh_degree; /* 0..3 */
} vsplat_cloud;

/* Camera (src/render/camera.hpp:9-18): view is world->camera, column-major as glm stores it. */
typedef struct {
  float view[16];
  float position[3];
  float fx, fy, cx, cy;
  int32_t width, height;
} vsplat_camera;

/* TrainerOptions (src/train/trainer.hpp:24-42). */
typedef struct {
  float lambda_ssim;
  float background[3];
  float near_plane;
  int32_t force_fixed_point; /* 64-bit fixed-point sums, deterministic; else float atomics */
  int32_t selective_adam;
  float opacity_reg, scale_reg;
  float position_noise;   /* 3DGS-MCMC's noise_lr: after each step, position noise scaled by it times rates->position; 0 for none */
  int32_t mip_filter_2d;
  int32_t pose_gradients; /* fill vsplat_step_stats::pose_gradient in vsplat_compute_gradients and vsplat_step */
  uint64_t memory_budget; /* bytes; UINT64_MAX for none */
  /* How the SH coefficients are stored on the GPU (no reference counterpart:
   * the reference always uses VSPLAT_SH_HALF). VSPLAT_SH_BF16 stores every
   * coefficient as bf16 (96 bytes per Gaussian at degree 3 instead of 104);
   * arithmetic stays fp32, and a step rounds the stored values at random. */
  int32_t sh_storage;
  /* Nonzero: the backward pass recomputes each pixel's transmittance and stop
   * with a front-to-back sweep of its tile instead of reading what the forward
   * pass stored (no reference counterpart; the results are identical). */
  int32_t recompute_transmittance;
  /* The raster tile shape (no reference counterpart; the reference's is 16x16):
   * VSPLAT_TILES_16X16, _8X8, _32X8, _32X32, or _AUTO, which times the four
   * on the first synchronous call for each view size (vsplat_step,
   * compute_gradients, step_views, capture_views, or vsplat_autotune_tiles) and
   * keeps the fastest. Large tiles list far fewer entries for scenes of large
   * splats, small ones make shorter lists per pixel block.
   * Every shape gives the same images and sums; only the tile lists differ. */
  int32_t tile_shape;
  /* How a batch of views runs (no reference counterpart; the results are the
   * same bits either way): VSPLAT_BATCH_VIEW_STREAMS, each view's passes on a
   * stream of its own, the views overlapping; or VSPLAT_BATCH_ONE_LAUNCH, each
   * view's lists on its own stream, then one forward launch and one backward
   * launch over every view's 8x8 pixel blocks (per tile shape). */
  int32_t batch_raster;
} vsplat_trainer_options;

enum { VSPLAT_BATCH_VIEW_STREAMS = 0, VSPLAT_BATCH_ONE_LAUNCH = 1 };

enum { VSPLAT_TILES_16X16 = 0, VSPLAT_TILES_8X8 = 1, VSPLAT_TILES_32X8 = 2, VSPLAT_TILES_AUTO = 3, VSPLAT_TILES_32X32 = 4 };

enum { VSPLAT_SH_HALF = 0, VSPLAT_SH_BF16 = 1 };

/* LearningRates (src/train/trainer.hpp:51-61). trust_region: once
 * vsplat_enable_second_order was called, the epsilon of the trust-region steps
 * (position, scale, opacity and SH; rotation keeps Adam). */
typedef struct {
  float position, scale, rotation, opacity, sh, sh_rest_scale, trust_region;
} vsplat_learning_rates;

/* StepStats (src/train/trainer.hpp:63-73). The exposure gradient is filled
 * when an exposure is set (vsplat_set_exposure); the pose gradient, dL/d(w, v)
 * of the camera in camera space (see vsplat_refined_pose), with
 * options.pose_gradients, by the synchronous calls (the queued and batched
 * calls leave it zero). */
typedef struct {
  float loss, l1, ssim;
  uint32_t visible_splats;
  uint32_t tile_entries;
  float exposure_gradient[12];
  float pose_gradient[6];
} vsplat_step_stats;

/* One stage of Trainer::timings() (trainer.hpp:186): average ms per timed iteration. */
typedef struct {
  char name[32];
  double milliseconds;
} vsplat_timing;

/* The defaults of TrainerOptions and LearningRates. */
void vsplat_default_options(vsplat_trainer_options *options);
void vsplat_default_rates(vsplat_learning_rates *rates);

/* Trainer::Trainer (trainer.hpp:91): copies the cloud to the GPU. options may be NULL. */
int vsplat_trainer_create(const vsplat_cloud *cloud, const vsplat_trainer_options *options, vsplat_trainer **out);
/* Trainer::~Trainer (trainer.hpp:92). NULL is ignored. */
void vsplat_trainer_destroy(vsplat_trainer *trainer);

/* Trainer::render (trainer.hpp:98): width*height*4 floats, rgb and log transmittance. */
int vsplat_render(vsplat_trainer *trainer, const vsplat_camera *camera, float *out_rgba);
/* No counterpart in the reference. vsplat_render without allocating, for
 * callers that cannot allocate device memory (tight serving or viewer loops,
 * harnesses that forbid hipMalloc). The image (width*height*4 floats, rgb and
 * log transmittance, as vsplat_render) is written to device_rgba, device memory
 * of capacity_bytes. The lists and per-pixel state live in one static device
 * array per GPU of VSPLAT_NOALLOC_SCRATCH_BYTES (held until the process ends),
 * so calls that overlap on one GPU race: make them one after another. Clouds,
 * views and tile grids past the limits below are refused
 * (VSPLAT_INVALID_ARGUMENT) before anything runs; a view that needs more tile
 * entries returns VSPLAT_BUDGET_EXCEEDED. Synchronous, on the trainer's stream. */
#define VSPLAT_NOALLOC_MAX_GAUSSIANS (1u << 17)
#define VSPLAT_NOALLOC_MAX_ENTRIES (1u << 21)
#define VSPLAT_NOALLOC_MAX_PIXELS (1u << 21)
#define VSPLAT_NOALLOC_MAX_TILES (1u << 16)
#define VSPLAT_NOALLOC_SCRATCH_BYTES (56ull << 20)
int vsplat_render_noalloc(vsplat_trainer *trainer, const vsplat_camera *camera, float *device_rgba, uint64_t capacity_bytes);
/* Trainer::compute_gradients (trainer.hpp:104): forward, loss, backward; the
 * gradients stay on the GPU for vsplat_get_gradients and vsplat_apply_gradients.
 * target is RGBA8, width*height*4 bytes. stats may be NULL. */
int vsplat_compute_gradients(vsplat_trainer *trainer, const vsplat_camera *camera, const uint8_t *target, size_t target_bytes,
                             vsplat_step_stats *stats);
/* Trainer::collect_statistics (trainer.hpp:106): densification statistics only. */
int vsplat_collect_statistics(vsplat_trainer *trainer, const vsplat_camera *camera, const uint8_t *target, size_t target_bytes,
                              vsplat_step_stats *stats);
/* Trainer::score_view (trainer.hpp:111, trainer.cpp:519-532): statistics of the
 * view with its high-error pixels marked (error at least error_share of the
 * view's worst), then each Gaussian's marked pixels, those times the view's
 * loss, and its edge score (the target's edge weight times its blending
 * weight) added to its densification statistics, in pixels times
 * pixel_weight. VSPLAT_INVALID_ARGUMENT for error_share <= 0. */
int vsplat_score_view(vsplat_trainer *trainer, const vsplat_camera *camera, const uint8_t *target, size_t target_bytes,
                      float error_share, float pixel_weight, vsplat_step_stats *stats);
/* Trainer::enable_second_order and Trainer::update_curvature (trainer.hpp:
 * 113-118, trainer.cpp:642-667): from enable_second_order on, the states'
 * second halves hold a running average of diag(J^T J) (rotation keeps Adam),
 * and vsplat_step / vsplat_enqueue step within the trust region that
 * rates->trust_region sets (3DGS^2-TR). update_curvature adds one Hutchinson
 * estimate from the last view's lists and image (VSPLAT_LOGIC_ERROR before any
 * view). The batched steps and vsplat_apply_gradients then refuse to step
 * (VSPLAT_UNSUPPORTED). */
int vsplat_enable_second_order(vsplat_trainer *trainer);
int vsplat_update_curvature(vsplat_trainer *trainer);
/* Trainer::reset_opacity (trainer.hpp:162-163, trainer.cpp:810-819): every
 * opacity capped at `maximum` (its logit caps the stored value) and the
 * opacity's optimiser state cleared. VSPLAT_INVALID_ARGUMENT unless
 * 0 < maximum < 1. */
int vsplat_reset_opacity(vsplat_trainer *trainer, float maximum);
/* Trainer::add_position_noise (trainer.hpp:164-165, trainer.cpp:806-808,
 * train_noise.comp): one round of 3DGS-MCMC's position noise, each position
 * moved by scaler * gate(1 - opacity) * Sigma n with n three normals drawn
 * from a hash of the Gaussian's index and `iteration`. A step with
 * options.position_noise > 0 adds the same round itself, scaler
 * position_noise * rates->position, iteration the steps taken so far. */
int vsplat_add_position_noise(vsplat_trainer *trainer, float scaler, uint32_t iteration);
/* Trainer::set_exposure (trainer.hpp:154): the view's affine colour transform,
 * 12 floats, rows of (r, g, b, 1): the loss compares T (rgb, 1) with the
 * target, and the synchronous calls report dL/dT in stats->exposure_gradient.
 * The neutral transform (1 0 0 0, 0 1 0 0, 0 0 1 0) skips the passes. The
 * queued and batched calls refuse a non-neutral one (VSPLAT_UNSUPPORTED). */
int vsplat_set_exposure(vsplat_trainer *trainer, const float *transform);
/* refined_pose (src/train/trainer.hpp:78-80, trainer.cpp:120-137): the camera
 * moved by move = (w, v), a rotation w (axis times angle) and then a
 * translation v in camera space, the move a pose gradient is taken for. A CPU
 * function: no trainer, no device. */
int vsplat_refined_pose(const vsplat_camera *camera, const float *move, vsplat_camera *out);
/* load_ply (src/splat/gaussian_cloud.hpp:60-63): a 3DGS .ply, binary
 * little-endian or ASCII; unknown properties such as normals are ignored and
 * the SH degree follows from the number of f_rest_* properties (0, 9, 24, 45).
 * vsplat_ply_info reads the header: the vertex count and SH degree, for sizing
 * the arrays vsplat_read_ply fills in GaussianCloud's layout (count must be the
 * file's). A file that is not a Gaussian splat file, truncated or malformed
 * gives VSPLAT_RUNTIME_ERROR with the reference's message. No GPU is used. */
int vsplat_ply_info(const char *path, uint64_t *count, int32_t *sh_degree);
int vsplat_read_ply(const char *path, float *positions, float *scales, float *rotations, float *opacities, float *sh, uint64_t count);
/* save_ply (src/splat/gaussian_cloud.hpp:65-67): binary little-endian in the
 * standard 3DGS layout (zero normals, f_rest channel-major). No GPU is used. */
int vsplat_write_ply(const char *path, const vsplat_cloud *cloud);
/* The trainer's cloud (Trainer::cloud) written as by vsplat_write_ply. */
int vsplat_save_ply(vsplat_trainer *trainer, const char *path);
/* Trainer::apply_gradients (trainer.hpp:121): one Adam step from the last compute_gradients. */
int vsplat_apply_gradients(vsplat_trainer *trainer, const vsplat_learning_rates *rates);
/* Trainer::step (trainer.hpp:125-126): compute_gradients and the Adam step in one
 * pass. next_target (may be NULL) is the following step's image: it is copied
 * to staging memory on a worker thread and must stay valid until the next
 * vsplat_step or vsplat_compute_gradients. */
int vsplat_step(vsplat_trainer *trainer, const vsplat_camera *camera, const uint8_t *target, size_t target_bytes,
                const vsplat_learning_rates *rates, const uint8_t *next_target, size_t next_target_bytes, vsplat_step_stats *stats);
/* Trainer::prefetch_target (trainer.hpp:129). The same borrowing rule as next_target. */
int vsplat_prefetch_target(vsplat_trainer *trainer, const uint8_t *target, size_t target_bytes);

/* Trainer::gradients (trainer.hpp:141-144): fp32 per parameter, laid out as the
 * cloud. VSPLAT_LOGIC_ERROR before any compute_gradients. Any pointer may be NULL. */
int vsplat_get_gradients(vsplat_trainer *trainer, float *positions, float *scales, float *rotations, float *opacities, float *sh);
/* Trainer::cloud (trainer.hpp:131), into caller arrays sized as the cloud. */
int vsplat_get_cloud(vsplat_trainer *trainer, float *positions, float *scales, float *rotations, float *opacities, float *sh);
/* Trainer::densify_stats (trainer.hpp:176): 8 floats per Gaussian, the
 * DensifyStat of train_types.glsl:18-21 (gradient sum, visits, largest radius,
 * contribution, absolute gradient sum, marked pixels, marked loss, edge score). */
int vsplat_get_densify_stats(vsplat_trainer *trainer, float *out);
/* Trainer::reset_densify_stats (trainer.hpp:178). */
int vsplat_reset_densify_stats(vsplat_trainer *trainer);
/* Trainer::apply_changes (trainer.hpp:204, trainer.cpp:749-787), the device
 * side of densify, relocate and the remove_* calls (trainer.hpp:156-183): the
 * Gaussians become `count` new ones, DensifyResult's cloud and source
 * (densify.hpp:52-63). New Gaussian g keeps the optimiser state of old Gaussian
 * source[g], gathered on the GPU, or starts with zero moments when source[g]
 * is negative. Its values come from `cloud` (count Gaussians of the trainer's
 * SH degree), or, when cloud is NULL, from old Gaussian source[g] on the GPU,
 * which needs every source >= 0 (pruning and copies, with no upload). The
 * densification statistics, the 3D filter, stored gradients and a captured
 * batch are dropped; Adam's step count is kept. VSPLAT_INVALID_ARGUMENT for a
 * source past the old count or a mismatched cloud, VSPLAT_RUNTIME_ERROR for
 * count 0 (the reference: "densification removed every Gaussian"). */
int vsplat_apply_changes(vsplat_trainer *trainer, const vsplat_cloud *cloud, const int64_t *source, uint32_t count);

/* Trainer::set_active_sh_degree (trainer.hpp:151), clamped to [0, degree]. */
int vsplat_set_active_sh_degree(vsplat_trainer *trainer, int32_t degree);
/* Trainer::set_filter3d (trainer.hpp:168): one value per Gaussian, else VSPLAT_INVALID_ARGUMENT. */
int vsplat_set_filter3d(vsplat_trainer *trainer, const float *filter, size_t count);
/* Trainer::set_sh_dropout (trainer.hpp:171). */
int vsplat_set_sh_dropout(vsplat_trainer *trainer, float probability, int32_t retained_degree, uint32_t seed);

/* Trainer::parameters (trainer.hpp:137): the device buffer of group 0 positions,
 * 1 scales, 2 rotations, 3 opacities, 4 SH (packed as half_sh_words describes).
 * Valid until the trainer is destroyed. */
int vsplat_parameters(vsplat_trainer *trainer, uint32_t group, void **device_pointer, uint64_t *bytes);

/* Trainer accessors (trainer.hpp:138-146, 187-193). */
uint32_t vsplat_gaussian_count(const vsplat_trainer *trainer);
int32_t vsplat_sh_degree(const vsplat_trainer *trainer);
int32_t vsplat_active_sh_degree(const vsplat_trainer *trainer);
int32_t vsplat_uses_fixed_point(const vsplat_trainer *trainer);
uint64_t vsplat_gpu_bytes(const vsplat_trainer *trainer);
uint64_t vsplat_bytes_per_gaussian(const vsplat_trainer *trainer);
uint64_t vsplat_staging_bytes(const vsplat_trainer *trainer);
/* Trainer::timings (trainer.hpp:186): up to capacity stages; *count gets how many exist. */
int vsplat_timings(const vsplat_trainer *trainer, vsplat_timing *out, uint32_t capacity, uint32_t *count);

/* Forgets the stage times gathered so far (waits for queued work first). */
int vsplat_reset_timings(vsplat_trainer *trainer);
/* Whether steps queued with vsplat_enqueue record per-stage events (on by
 * default); the synchronous calls always do. No reference counterpart. */
int vsplat_set_stage_timing(vsplat_trainer *trainer, int32_t enabled);

/* The message of the calling thread's last failed call. */
const char *vsplat_last_error(void);

/* ---- Device-resident path: no reference counterpart; for callers that keep
 * targets in GPU memory and do not want a host round trip per step. ---- */

/* Queues compute_gradients (mode 1) or step (mode 2, rates required) against a
 * target already in device memory (width*height*4 bytes, on this trainer's GPU). */
int vsplat_enqueue(vsplat_trainer *trainer, const vsplat_camera *camera, const uint8_t *device_target, int32_t mode,
                   const vsplat_learning_rates *rates);
/* Waits for queued work; stats (may be NULL) of the last queued step. Returns
 * VSPLAT_RUNTIME_ERROR if a queued view needed more tile entries than were
 * reserved: that step and every step queued after it computed nothing and
 * changed nothing (Adam's step count included), while the steps before it
 * completed. vsplat_compute_gradients with the view, or vsplat_reserve_entries,
 * grows the reservation. */
int vsplat_synchronize(vsplat_trainer *trainer, vsplat_step_stats *stats);
/* The trainer's HIP stream (a hipStream_t). */
void *vsplat_stream(vsplat_trainer *trainer);
/* Reserves room for this many tile entries now (subject to the memory budget). */
int vsplat_reserve_entries(vsplat_trainer *trainer, uint64_t entries);

/* ---- The batched step: no reference counterpart (the reference trains one
 * view per iteration). For callers that train on several views at once. ---- */

/* One training iteration over n views: each view's forward, loss and backward,
 * the gradients summed over the views in view order (each view's objective
 * includes the penalties, and adds to the densification statistics), then for
 * mode 2 one Adam step from the sum (one step count, one bias correction); mode 1
 * keeps the sum for vsplat_get_gradients. targets[k] is view k's RGBA8 image in
 * host memory, bytes[k] its size. stats (may be NULL) gets n entries. A view that
 * does not fit the tile-entry reservation grows it and the batch runs again, as
 * vsplat_step does. */
int vsplat_step_views(vsplat_trainer *trainer, uint32_t n, const vsplat_camera *cameras, const uint8_t *const *targets,
                      const size_t *bytes, int32_t mode, const vsplat_learning_rates *rates, vsplat_step_stats *stats);
/* The same queued against images already in GPU memory (device_targets[k],
 * width*height*4 bytes each). */
int vsplat_enqueue_views(vsplat_trainer *trainer, uint32_t n, const vsplat_camera *cameras, const uint8_t *const *device_targets,
                         int32_t mode, const vsplat_learning_rates *rates);
/* Waits for a batch queued with vsplat_enqueue_views: up to capacity views'
 * stats into stats (may be NULL), *count (may be NULL) the batch's view count.
 * VSPLAT_RUNTIME_ERROR, as vsplat_synchronize, if a view did not fit: the batch
 * changed nothing. */
int vsplat_synchronize_views(vsplat_trainer *trainer, vsplat_step_stats *stats, uint32_t capacity, uint32_t *count);
/* Captures the batch vsplat_enqueue_views would queue as a HIP graph (replacing
 * any captured before), without running it. The trainer's buffers are sized
 * for it first; the tile-entry reservation is the one of the moment (a
 * vsplat_step_views over the same views sizes it). Each replay is one training
 * iteration with the captured cameras, device targets (which must stay valid)
 * and rates; Adam's step count and bias corrections advance on the device, so
 * replays continue the optimisation. */
int vsplat_capture_views(vsplat_trainer *trainer, uint32_t n, const vsplat_camera *cameras, const uint8_t *const *device_targets,
                         int32_t mode, const vsplat_learning_rates *rates);
/* Queues the captured batch `replays` times on the trainer's stream; wait with
 * vsplat_synchronize_views (the last replay's stats; VSPLAT_RUNTIME_ERROR if a
 * view did not fit, in which case that replay and every later one changed
 * nothing). VSPLAT_LOGIC_ERROR if nothing is captured, or if a buffer the graph
 * uses was reallocated or a setting baked into it (active SH degree, SH
 * dropout, 3D filter) changed since the capture. */
int vsplat_launch_captured(vsplat_trainer *trainer, uint32_t replays);
/* Frees the captured graph. */
int vsplat_release_captured(vsplat_trainer *trainer);
/* Times the three tile shapes on one view (forward, loss and backward, no
 * parameter change; device_target as for vsplat_enqueue), keeps the fastest
 * for the view's size (used when tile_shape is VSPLAT_TILES_AUTO) and returns
 * it in *tile_width, *tile_height (either may be NULL). */
int vsplat_autotune_tiles(vsplat_trainer *trainer, const vsplat_camera *camera, const uint8_t *device_target, uint32_t repeats,
                          uint32_t *tile_width, uint32_t *tile_height);
/* Tunes the tile shapes of a batch's view sizes together (no reference
 * counterpart): the batch's passes up to the backward are timed as a whole
 * (the views overlap on their streams), first with each size's current shape,
 * then each size in turn with the other shapes, keeping those that make the
 * batch faster. Used by vsplat_step_views and vsplat_capture_views with
 * VSPLAT_TILES_AUTO the first time they meet a batch's list of sizes; the
 * shapes kept apply to those sizes from then on. Each view's shape goes to
 * tile_widths[k], tile_heights[k] (either may be NULL). Device targets as for
 * vsplat_enqueue_views. No parameter, gradient or statistic changes. */
int vsplat_autotune_views(vsplat_trainer *trainer, uint32_t n, const vsplat_camera *cameras, const uint8_t *const *device_targets,
                          uint32_t repeats, uint32_t *tile_widths, uint32_t *tile_heights);

/* ---- Introspection, for parity tests: no reference counterpart. ---- */

/* Copies one of the trainer's intermediate device buffers, as the last call
 * left it, to host memory: "splats" (48 bytes per Gaussian), "keys", "values"
 * (depth-sorted, 4 bytes per Gaussian), "coverage", "entry_keys" (each
 * entry's tile), "entry_blocks" (the 8x8 pixel blocks of the tile its splat
 * reaches, bit q for block q, row-major; every block but for 32x32 tiles), "entry_values" (4 bytes per tile
 * entry each), "ranges" (8 bytes per tile),
 * "image" (16 bytes per pixel: rgb, log T), "stops", "adjoint" (16 per pixel),
 * "fields" (36 per pixel), "tile_totals" (8 per tile), "block_work" (per
 * raster launch index, 4 bytes: the entries the backward walks for that 8x8
 * block), "block_order" and "list_order" (the backward's and the forward's
 * launch orders: permutations of those indices, heaviest first on each XCD,
 * by block work and by tile list length), "raw" (14 int64 per
 * Gaussian with fixed point, else 14 float), "counters" (8 words: visible splats, entries, entries listed, big splats, big-splat bit words, overflow, skipped steps, unused),
 * "state_positions", "state_scales", "state_rotations", "state_opacities",
 * "state_sh" (Adam's state, one word per parameter: bf16 m low, bf16 v high),
 * "stats" (8 floats per Gaussian). *bytes
 * gets the size in use; at most `capacity` bytes are copied. */
int vsplat_inspect(vsplat_trainer *trainer, const char *buffer, void *out, uint64_t capacity, uint64_t *bytes);
/* The same for view `view` of the last batch (its k-th camera; view 0 is also
 * the single-view calls'); the shared buffers (state_*, stats) are the same
 * for every view. VSPLAT_INVALID_ARGUMENT for a view no batch has made. */
int vsplat_inspect_view(vsplat_trainer *trainer, uint32_t view, const char *buffer, void *out, uint64_t capacity,
                        uint64_t *bytes);

/* Device memory for callers without a HIP runtime of their own (no reference
 * counterpart): targets for vsplat_enqueue and the batched calls. Allocates
 * bytes of GPU memory, copying them from host when it is not NULL. */
int vsplat_device_alloc(uint64_t bytes, const void *host, void **out_device);
/* Copies bytes from GPU memory to host memory. */
int vsplat_device_read(const void *device, void *host, uint64_t bytes);
/* Copies bytes from one GPU buffer to another `repeats` times, back to back on
 * a stream of its own, and sets *ms_per_copy (may be NULL) to the mean time of
 * one copy by HIP events: the device's copy rate, for measurements. */
int vsplat_device_copy(void *dst, const void *src, uint64_t bytes, uint32_t repeats, float *ms_per_copy);
/* Frees memory from vsplat_device_alloc. NULL is ignored. */
void vsplat_device_free(void *device);

/* Runs one of the device's elementwise functions over n floats (device
 * arithmetic check): 0 exp, 1 log, 2 sqrt, 3 reciprocal (1/x), 4 float to half
 * and back, 5 the half in the low 16 bits of the input's bits to float, 6 the
 * backward's short division of pairs: out[2k] = in[2k] / in[2k + 1] for
 * 2^-100 <= in[2k] <= 2^60 and 2^-20 <= in[2k + 1] <= 1 (outside that range
 * the result is undefined), out[2k + 1] = 0. in and out are host arrays. */
int vsplat_math_probe(int32_t function, const float *in, float *out, uint64_t n);

#ifdef __cplusplus
}
#endif

#endif /* VSPLAT_H */
