// Host-side launchers of the rasterizer's kernels, one translation unit per
// stage. Every launcher only enqueues on the given stream.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "device.hpp"

namespace vsplat {

// What the chain-rule pass does with the gradients (train_step.glsl:346-353).
constexpr uint32_t kModeStatistics = 0, kModeGradients = 1, kModeStep = 2;
constexpr uint32_t kModeCurvature = 3; // the curvature's running average (Trainer::update_curvature)
constexpr uint32_t kModeRender = 0xFFFFFFFFu; // forward only (Trainer::render)
constexpr uint32_t kFlagStatistics = 1, kFlagSelective = 4;
constexpr uint32_t kFlagAccumulate = 8; // gradients mode: add to the gradient buffers instead of writing them
constexpr uint32_t kFlagSecondOrder = 16; // step mode: the trust-region step, curvature in the states' second halves
constexpr uint32_t kFlagPose = 32;        // gradients and step modes: each Gaussian's pose terms into poseTerms
constexpr uint32_t kShPassFloats = 10; // per Gaussian, what the SH step needs from the chain rule

struct ChainParams {
  float opacityReg, scaleReg; // per Gaussian: the option over the count
  uint32_t mipFilter2d;
  float shDropout;
  uint32_t shRetained, shSeed;
  uint32_t mode, flags;
  float rates[5]; // position, scale, rotation, opacity, SH
  float restScale, beta1, beta2, epsilon, correction1, correction2;
  uint32_t seed, shFloats;
  uint32_t fixedPoint;
  float fixedScale;
  // The optimiser's step count on the device: [steps taken, beta1^taken,
  // beta2^taken] (the powers as float bits, kept by repeated multiplication).
  // When set, a step takes its bias corrections and rounding seed from it
  // rather than from correction1, correction2 and seed, so that a captured
  // graph can replay steps.
  uint32_t *adamState;
  // Second order (kFlagSecondOrder, kModeCurvature): the trust region's
  // epsilon, the curvature's bias correction 1 / (1 - beta^k), beta, and in
  // curvature mode 1 / the values in the image (seed: the update count).
  float trustRegion, curvatureScale, curvatureBeta, perValue;
};

struct ChainBuffers {
  float *positions, *scales, *rotations, *opacities;
  uint32_t *shWords;
  const float *filter3d;
  const Splat *splats;
  const int64_t *rawFixed;
  const float *rawFloat;
  float *shPass;
  float *stats; // 8 per Gaussian, or null
  uint32_t *statePos, *stateScale, *stateRot, *stateOpacity, *stateSh;
  float *gradPos, *gradScale, *gradRot, *gradOpacity, *gradSh;
  float4 *dirTerms; // per Gaussian: the view direction's term of the position gradient (between the chain rule's two kernels)
  float *poseTerms; // 6 per Gaussian with kFlagPose: t x v, then v = Rv dL/dp; untouched Gaussians' left as they are
};

// One view of a batched chain rule: its camera, its screen-space sums (fixed
// point with its scale, or float when rawFixed is null) and its splats.
struct ChainView {
  ViewParams c;
  const int64_t *rawFixed;
  const float *rawFloat;
  const Splat *splats;
  float fixedScale;
};
constexpr uint32_t kChainViewsPerLaunch = 8;
// Up to kChainViewsPerLaunch views, passed by value as the kernel's argument.
struct ChainViews {
  ChainView view[kChainViewsPerLaunch];
  uint32_t count;
};

struct TileRecord;

// Workgroups (of four waves) of the kernels that count and list the big splats
// (boxes of more than kMaskTiles tiles), whose waves take the listed splats in turn.
constexpr uint32_t kBigSplatBlocks = 2048;

// Projection and tile coverage: splats, depth keys and values (the identity),
// and per Gaussian the tile count and TileRecord; adds the visible count. The
// big splats are listed in bigList (count[kBigSplats], cleared before) and
// counted by their own kernel. bigList holds one word per Gaussian; bigBits
// (bitCapacity words) receives their reached-tile bits, one word per 64
// candidates, for the list pass (a splat that finds no room keeps none).
void launch_project(hipStream_t stream, const float *positions, const float *scales, const float *rotations, const float *opacities,
                    const uint32_t *shWords, const float *filter3d, const ViewParams &view, bool mipFilter2d, float shDropout,
                    uint32_t shRetained, uint32_t shSeed, uint32_t gridW, uint32_t gridH, Splat *splats, uint32_t *keys,
                    uint32_t *values, uint32_t *tileCount, TileRecord *records, uint32_t *counters, uint32_t *bigList,
                    uint64_t *bigBits, uint32_t bitCapacity);

// Scratch for launch_radix_sort over up to capacity keys.
size_t radix_sort_scratch_words(uint64_t capacity);
void launch_radix_sort(hipStream_t stream, uint32_t *keys, uint32_t *values, uint32_t *tmpKeys, uint32_t *tmpValues,
                       const uint32_t *countDev, uint32_t capacity, uint32_t bits, uint32_t *scratch);
void launch_copy_u32(hipStream_t stream, const uint32_t *src, uint32_t *dst, const uint32_t *countDev, uint32_t capacity);
void launch_fill_u32(hipStream_t stream, uint32_t *dst, uint64_t count, uint32_t value);

// One per-Gaussian array for launch_gather: `stride` 32-bit words per Gaussian.
struct GatherArray {
  const uint32_t *from;
  uint32_t *to;
  uint32_t stride;
};
constexpr uint32_t kMaxGatherArrays = 10;
// to[g * stride + k] = source[g] < 0 ? 0 : from[source[g] * stride + k] for the
// new Gaussians g < gaussians, for each of `count` (at most kMaxGatherArrays) arrays.
void launch_gather(hipStream_t stream, const int32_t *source, uint32_t gaussians, const GatherArray *arrays, uint32_t count);

// Scratch for launch_coverage_scan over count splats (per chunk of splats, its
// entries and its visible splats; written before it is read).
size_t coverage_scan_words(uint32_t count);
// Offsets of each depth-sorted splat's entries, the visible and entry counts
// and the capacity check.
void launch_coverage_scan(hipStream_t stream, const uint32_t *keys, const uint32_t *values, const uint32_t *tileCount, uint32_t count,
                          uint32_t *offsets, uint32_t *counters, uint32_t capacity, bool isStep, uint32_t *totals);
// A batch's views' status, up to kStatusViewsPerLaunch views (views first ..
// first + count - 1) a launch: each view's counters and loss sums copied to its
// kViewStatusBytes slot of `slots`, then status[overflow] (the first view's
// counters) |= every other view's.
constexpr uint32_t kStatusViewsPerLaunch = 8, kViewStatusBytes = 48;
struct ViewStatus {
  const uint32_t *counters[kStatusViewsPerLaunch];
  const double *sums[kStatusViewsPerLaunch];
  uint32_t first, count;
};
void launch_gather_views(hipStream_t stream, const ViewStatus &views, uint8_t *slots, uint32_t *status);
// The tile lists in depth order: the small splats' entries from their bits,
// then the big splats' (bigList, with the bits launch_project kept). records is written: a
// big splat's first entry is kept in its maskLo for the second kernel.
// The tile entries (keys: the tile in the low tileBits bits, above them the
// tile's 8x8 blocks the splat reaches, marked for 32x32 tiles and all set for
// the others; values: the Gaussian), for a view of width x height pixels.
void launch_bin_emit(hipStream_t stream, const uint32_t *keys, const uint32_t *values, const Splat *splats, TileRecord *records,
                     uint32_t count, uint32_t gridW, uint32_t tileW, uint32_t tileH, uint32_t width, uint32_t height, uint32_t tileBits,
                     const uint32_t *offsets, const uint32_t *counters, const uint32_t *bigList, const uint64_t *bigBits,
                     uint32_t *entryKeys, uint32_t *entryValues);
void launch_tile_ranges(hipStream_t stream, const uint32_t *entryKeys, const uint32_t *counters, uint32_t capacity, uint32_t tileBits,
                        uint2 *ranges);

// The raster launches' workgroups for a view of `tiles` list tiles of TW x TH
// pixels: one per 8x8 block, the tile count rounded up to a multiple of the
// chip's eight XCDs so that all of a tile's blocks run on one XCD.
constexpr uint32_t kRasterXcds = 8;
inline uint32_t raster_blocks(uint32_t tiles, uint32_t tileW, uint32_t tileH) {
  return (tiles + kRasterXcds - 1) / kRasterXcds * kRasterXcds * (tileW / 8u) * (tileH / 8u);
}

// One view of a batched raster launch (its buffers and sizes), and up to
// kRasterViewsPerLaunch of them, passed by value as the kernel's argument.
// The views' workgroups follow each other: view v's are firstBlock ..
// firstBlock + blocks - 1 of the grid, blocks = raster_blocks(its tiles), so
// every firstBlock is a multiple of kRasterXcds.
struct RasterView {
  const Splat *splats;
  const uint32_t *entryKeys, *entryValues;
  const uint2 *ranges;
  float4 *image;
  uint32_t *stops;
  const float4 *adjoint;
  int64_t *rawFixed;
  float *rawFloat;
  float fixedScale;
  uint32_t width, height, gridW, firstBlock, blocks;
  uint32_t tileBits;
};
constexpr uint32_t kRasterViewsPerLaunch = 8;
struct RasterViews {
  RasterView view[kRasterViewsPerLaunch];
  uint32_t count;
};
// The batched forward and backward over views whose lists share one tile shape.
void launch_forward_views(hipStream_t stream, const RasterViews &views, uint32_t tileW, uint32_t tileH, const float background[3]);
void launch_backward_views(hipStream_t stream, bool fixedPoint, bool recompute, const RasterViews &views, uint32_t tileW, uint32_t tileH,
                           const float background[3]);

// The raster kernels for one tile shape: 8x8, 16x16, 32x8 or 32x32 (tileW x tileH).
// The lists are read as launch_bin_emit left them: entryKeys carry each
// entry's tile (low tileBits bits) and blocks, entryValues its Gaussian.
void launch_forward(hipStream_t stream, const Splat *splats, const uint32_t *entryKeys, const uint32_t *entryValues, uint32_t tileBits,
                    const uint2 *ranges, uint32_t width, uint32_t height, uint32_t tileW, uint32_t tileH, const float background[3],
                    float4 *image, uint32_t *stops, uint32_t *work = nullptr, const uint32_t *order = nullptr);
// (stops may be null: not kept when the backward recomputes them. work, when
// given, receives per launch index the entries the backward will walk for
// that 8x8 block, raster_blocks(tiles) of them. order: the blocks' launch
// order from launch_list_order, or null for raster_blocks' own.)
// recompute: the transmittance and stops are recomputed per tile rather than
// read from the forward pass (image and stops may then be null). edge: a
// scored view's edge weights (the adjoint's fourth value its marks), which
// fill the last two raw sums; null otherwise. order: the blocks' launch order
// from launch_work_order, or null for raster_blocks' own.
void launch_backward(hipStream_t stream, bool fixedPoint, bool recompute, const Splat *splats, const uint32_t *entryKeys,
                     const uint32_t *entryValues, uint32_t tileBits, const uint2 *ranges, uint32_t width, uint32_t height, uint32_t tileW,
                     uint32_t tileH, const float background[3], const float4 *image, const float4 *adjoint, const uint32_t *stops,
                     float fixedScale, int64_t *rawFixed, float *rawFloat, const float *edge = nullptr, const uint32_t *order = nullptr);
// The backward's launch order, heaviest first on each XCD by the forward's work
// (each block keeps its XCD): a permutation of the items = raster_blocks(tiles)
// launch indices.
void launch_work_order(hipStream_t stream, const uint32_t *work, uint32_t items, uint32_t *order);
// The forward's launch order, heaviest first on each XCD by the length of each
// block's tile list.
void launch_list_order(hipStream_t stream, const uint2 *ranges, uint32_t tiles, uint32_t tileW, uint32_t tileH, uint32_t *order);

// errorMax and edge: when scoring, the view's worst error (cleared before)
// and the target's edge weights, and the adjoint's fourth value the pixel's
// error; both null otherwise.
void launch_loss(hipStream_t stream, const float4 *image, const uint8_t *target, uint32_t width, uint32_t height, float lambda,
                 float *fields, float4 *adjoint, float2 *tileTotals, double *sums, uint32_t *errorMax, float *edge);

// A view's affine colour transform (Trainer::set_exposure): rows of (r, g, b, 1).
struct Exposure {
  float transform[12];
};
// Workgroups of launch_exposure_backward: twelve totals each.
uint32_t exposure_groups(uint32_t pixels);
void launch_exposure(hipStream_t stream, const float4 *image, float4 *exposed, uint32_t pixels, const Exposure &t);
void launch_exposure_backward(hipStream_t stream, const float4 *image, float4 *adjoint, uint32_t pixels, const Exposure &t,
                              float *groupTotals);
void launch_error_mask(hipStream_t stream, float4 *adjoint, const uint32_t *errorMax, uint32_t pixels, float share);

// The chain rule over the views of a batch chunk, summed per Gaussian in view
// order (kFlagAccumulate: continuing from the sums an earlier chunk stored).
// Gradients mode stores the sums; step mode takes the geometric parameters'
// Adam step from them and stores the SH sums for launch_sh_adam.
void launch_chain_views(hipStream_t stream, const ChainBuffers &buffers, const ChainViews &views, const ChainParams &params,
                        const uint32_t *counters);
// Adam for the SH coefficients from the stored SH gradients, then the step count advanced.
void launch_sh_adam(hipStream_t stream, const ChainBuffers &buffers, const ViewParams &view, const ChainParams &params,
                    const uint32_t *counters);
// One Adam step of every parameter from the stored gradients in buffers.grad*
// (Trainer::apply_gradients); params.mode is kModeStep.
void launch_adam(hipStream_t stream, const ChainBuffers &buffers, const ViewParams &view, const ChainParams &params,
                 uint32_t *counters);
void launch_chain(hipStream_t stream, const ChainBuffers &buffers, const ViewParams &view, const ChainParams &params,
                  const uint32_t *counters);
// Position noise after a step (the iteration from *iterationDev when set, the
// device's step count; nothing when counters say the step was skipped), and
// opacities clamped to a logit (reset_opacity).
void launch_noise(hipStream_t stream, const ChainBuffers &buffers, uint32_t count, float scaler, const uint32_t *iterationDev,
                  uint32_t iterationHost, const uint32_t *counters);
void launch_clamp_opacity(hipStream_t stream, float *opacities, uint32_t count, float limit);
// The pose terms' totals per group of 256 Gaussians (6 floats per group).
void launch_pose_totals(hipStream_t stream, const float *terms, uint32_t count, float *totals);
inline uint32_t pose_groups(uint32_t count) { return (count + 255) / 256; }
// Second order: the Hutchinson adjoint of random signs (update number seed),
// and the states' second halves cleared (enable_second_order).
void launch_hutchinson(hipStream_t stream, float4 *adjoint, uint32_t pixels, uint32_t seed);
void launch_keep_first_moment(hipStream_t stream, uint32_t *state, uint64_t count);
// A scored view's marked pixels, marked loss and edge score added to each
// Gaussian's statistics (stats: two float4 per Gaussian) from raw sums 12 and 13.
void launch_score(hipStream_t stream, const int64_t *rawFixed, const float *rawFloat, float fixedScale, float4 *stats, uint32_t count,
                  float toPixels, float loss);

} // namespace vsplat
