// The trainer's host side (src/train/trainer.cpp:157-863 for what it does; the
// layout is this project's): buffers sized for the view, the passes queued on
// one HIP stream, one wait per step for the counts and loss totals.
//
// The reference reads the tile-entry count back between two submissions to
// size the entry buffers (trainer.cpp:328-375). Here the whole step is queued
// at once against the entry capacity already reserved; the binning stages
// leave the count on the device, every later kernel reads it from there, and
// a view that does not fit sets a flag that makes the rest of the step a no-op
// (nothing is written to the parameters). The host then grows the reservation,
// within the memory budget, and runs the step again. In steady state that is
// one synchronisation per step and no host work between kernels.
#include "trainer.hpp"
#include "tiles.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace vsplat {

// The fixed-point scale of the loss's sums: per-pixel derivatives carry
// 1 / (3 W H), which 3 W H undoes, then kFixedSteps steps per unit
// (trainer.cpp:586-591 has 64 steps; the sums here are 64-bit).
static float fixed_scale(const ViewParams &v) { return 3.0f * static_cast<float>(v.width) * static_cast<float>(v.height) * kFixedSteps; }
// The curvature's Hutchinson adjoint is +-1 per value, without the 1 / (3 W H).
static float curvature_fixed_scale() { return kFixedSteps; }

namespace {

void check(hipError_t e, const char *what) {
  if (e != hipSuccess) throw Error(VSPLAT_DEVICE_ERROR, std::string(what) + ": " + hipGetErrorString(e));
}

uint32_t half_words(uint32_t coefficients) { return 3u + (3u * (coefficients - 1u) + 1u) / 2u; }
uint32_t bf16_words(uint32_t coefficients) { return (3u * coefficients + 1u) / 2u; }

// Float to bf16, round to nearest even (as the device's bf16_round), and back.
uint32_t float_to_bf16(float f) {
  uint32_t bits;
  std::memcpy(&bits, &f, 4);
  return (bits + 0x7FFFu + ((bits >> 16) & 1u)) >> 16;
}
float bf16_to_float(uint32_t h) {
  const uint32_t bits = h << 16;
  float f;
  std::memcpy(&f, &bits, 4);
  return f;
}

// Float to IEEE binary16, round to nearest even (as the device's float_to_half).
uint16_t float_to_half(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  const uint32_t sign = (u >> 16) & 0x8000u, a = u & 0x7FFFFFFFu;
  if (a >= 0x7F800000u) return static_cast<uint16_t>(sign | 0x7C00u | (a > 0x7F800000u ? 0x200u | ((a >> 13) & 0x3FFu) : 0u));
  if (a >= 0x477FF000u) return static_cast<uint16_t>(sign | 0x7C00u);
  if (a < 0x38800000u) {
    float v;
    std::memcpy(&v, &a, 4);
    return static_cast<uint16_t>(sign | static_cast<uint32_t>(std::nearbyint(v * 16777216.0f)));
  }
  return static_cast<uint16_t>(sign | ((a + 0x0FFFu + ((a >> 13) & 1u) - 0x38000000u) >> 13));
}

float half_to_float(uint32_t h) {
  const uint32_t sign = (h & 0x8000u) << 16, e = (h >> 10) & 0x1Fu, m = h & 0x3FFu;
  uint32_t bits;
  if (e == 0) {
    const float v = static_cast<float>(m) * 5.9604644775390625e-08f;
    return sign ? -v : v;
  }
  bits = e == 31 ? (sign | 0x7F800000u | (m << 13)) : (sign | ((e + 112u) << 23) | (m << 13));
  float f;
  std::memcpy(&f, &bits, 4);
  return f;
}

uint64_t div_up(uint64_t a, uint64_t b) { return (a + b - 1) / b; }

// The big splats' bit words a view keeps room for: one per Gaussian and one per
// 16 tile entries (boxes with up to four candidates per listed entry keep
// their bits; the rest are tested again when listed).
uint64_t big_bit_words(const ViewBuffers &b) { return std::min<uint64_t>(b.bigBits.bytes() / 8, 0xFFFFFFFFull); }

} // namespace

uint64_t Trainer::big_bit_words_for(uint64_t entries) const { return count_ + entries / 16; }

void DeviceArray::ensure(size_t bytes) {
  if (bytes <= bytes_ && ptr_ != nullptr) return;
  if (!owned_) throw Error(VSPLAT_RUNTIME_ERROR, "a borrowed device array is too small");
  release();
  check(hipMalloc(&ptr_, std::max<size_t>(bytes, 256)), "hipMalloc");
  bytes_ = std::max<size_t>(bytes, 256);
}
void DeviceArray::release() {
  if (ptr_ != nullptr && owned_) (void)hipFree(ptr_);
  ptr_ = nullptr;
  bytes_ = 0;
  owned_ = true;
}
void PinnedArray::ensure(size_t bytes) {
  if (bytes <= bytes_ && ptr_ != nullptr) return;
  release();
  check(hipHostMalloc(&ptr_, std::max<size_t>(bytes, 256), hipHostMallocDefault), "hipHostMalloc");
  bytes_ = std::max<size_t>(bytes, 256);
}
void PinnedArray::release() {
  if (ptr_ != nullptr) (void)hipHostFree(ptr_);
  ptr_ = nullptr;
  bytes_ = 0;
}

Trainer::Trainer(const vsplat_cloud &c, const vsplat_trainer_options &options) : options_(options) {
  if (c.count == 0) throw Error(VSPLAT_RUNTIME_ERROR, "nothing to train: the cloud is empty");
  if (c.sh_degree < 0 || c.sh_degree > 3) throw Error(VSPLAT_INVALID_ARGUMENT, "the SH degree must be 0 to 3");
  if (c.positions == nullptr || c.scales == nullptr || c.rotations == nullptr || c.opacities == nullptr || c.sh == nullptr) {
    throw Error(VSPLAT_INVALID_ARGUMENT, "the cloud has a null array");
  }
  if (!(options.position_noise >= 0.0f)) throw Error(VSPLAT_INVALID_ARGUMENT, "position noise must be 0 or above");
  int devices = 0;
  if (hipGetDeviceCount(&devices) != hipSuccess || devices == 0) throw Error(VSPLAT_DEVICE_ERROR, "no GPU is visible");
  count_ = c.count;
  shDegree_ = activeShDegree_ = c.sh_degree;
  const uint32_t coefficients = static_cast<uint32_t>((c.sh_degree + 1) * (c.sh_degree + 1));
  shFloats_ = 3u * coefficients;
  if (options.sh_storage != VSPLAT_SH_HALF && options.sh_storage != VSPLAT_SH_BF16) {
    throw Error(VSPLAT_INVALID_ARGUMENT, "sh_storage must be VSPLAT_SH_HALF or VSPLAT_SH_BF16");
  }
  shFormat_ = options.sh_storage == VSPLAT_SH_BF16 ? kShBf16 : kShHalf;
  if (options.batch_raster != VSPLAT_BATCH_VIEW_STREAMS && options.batch_raster != VSPLAT_BATCH_ONE_LAUNCH) {
    throw Error(VSPLAT_INVALID_ARGUMENT, "batch_raster must be VSPLAT_BATCH_VIEW_STREAMS or VSPLAT_BATCH_ONE_LAUNCH");
  }
  if (options.tile_shape < VSPLAT_TILES_16X16 || options.tile_shape > VSPLAT_TILES_32X32) {
    throw Error(VSPLAT_INVALID_ARGUMENT, "tile_shape must be one of VSPLAT_TILES_16X16, _8X8, _32X8, _32X32 or _AUTO");
  }
  shWords_ = shFormat_ == kShBf16 ? bf16_words(coefficients) : half_words(coefficients);
  // Float atomics are the default on this GPU, as on any with them (trainer.cpp:181).
  fixedPoint_ = options.force_fixed_point != 0;
  check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
  slots_.resize(kTimingSlots);
  for (auto &slot : slots_)
    for (auto &e : slot.events) check(hipEventCreate(&e), "hipEventCreate");

  const size_t n = count_;
  DeviceArray *values[5] = {&positions_, &scales_, &rotations_, &opacities_, &shWordsBuf_};
  upload_cloud(c, values);
  const size_t stateFloats[5] = {n * 3, n * 3, n * 4, n, n * shFloats_};
  for (int g = 0; g < 5; ++g) {
    state_[g].ensure(stateFloats[g] * 4);
    launch_fill_u32(stream_, state_[g].as<uint32_t>(), stateFloats[g], 0u);
  }
  allocate_per_gaussian();
  readback_.ensure(kCounterWords * 4 + 2 * sizeof(double));
  adamState_.ensure(4 * 4);
  const uint32_t adamStart[4] = {0u, 0x3F800000u, 0x3F800000u, 0u}; // no steps; beta^0 = 1
  check(hipMemcpy(adamState_.as<uint32_t>(), adamStart, sizeof(adamStart), hipMemcpyHostToDevice), "adam state");
  check(hipEventCreateWithFlags(&batchStart_, hipEventDisableTiming), "hipEventCreate");
  check(hipEventCreateWithFlags(&batchRaster_, hipEventDisableTiming), "hipEventCreate");
  view_buffers(0);
  check(hipStreamSynchronize(stream_), "trainer set-up");
}

// The cloud's values in the trainer's layout: positions, scales, rotations and
// opacities as floats; the SH packed as pack_half_sh does
// (src/splat/gaussian_cloud.cpp:537-550: DC as floats, the rest as halves), or
// every coefficient as bf16.
void Trainer::upload_cloud(const vsplat_cloud &c, DeviceArray *const values[5]) const {
  const size_t n = c.count;
  const float *from[4] = {c.positions, c.scales, c.rotations, c.opacities};
  const size_t floats[4] = {3, 3, 4, 1};
  for (int g = 0; g < 4; ++g) {
    values[g]->ensure(n * floats[g] * 4);
    check(hipMemcpy(values[g]->as<float>(), from[g], n * floats[g] * 4, hipMemcpyHostToDevice), "upload parameters");
  }
  std::vector<uint32_t> words(n * shWords_, 0u);
  for (size_t i = 0; i < n; ++i) {
    const float *in = c.sh + i * shFloats_;
    uint32_t *to = words.data() + i * shWords_;
    if (shFormat_ == kShBf16) {
      for (uint32_t j = 0; j < shFloats_; ++j) to[j / 2] |= float_to_bf16(in[j]) << (16 * (j & 1));
    } else {
      for (uint32_t j = 0; j < 3; ++j) std::memcpy(&to[j], &in[j], 4);
      for (uint32_t j = 3; j < shFloats_; ++j) to[3 + (j - 3) / 2] |= static_cast<uint32_t>(float_to_half(in[j])) << (16 * ((j - 3) & 1));
    }
  }
  values[4]->ensure(n * shWords_ * 4);
  check(hipMemcpy(values[4]->as<uint32_t>(), words.data(), words.size() * 4, hipMemcpyHostToDevice), "upload SH");
}

// The per-Gaussian buffers the views share (trainer.cpp:242-262), cleared:
// densification statistics, 3D filter values, the SH pass's scratch.
void Trainer::allocate_per_gaussian() {
  const size_t n = count_;
  shPass_.ensure(n * kShPassFloats * 4);
  launch_fill_u32(stream_, shPass_.as<uint32_t>(), n * kShPassFloats, 0u);
  dirTerms_.ensure(n * 16);
  stats_.ensure(n * 32);
  launch_fill_u32(stream_, stats_.as<uint32_t>(), n * 8, 0u);
  filter3d_.ensure(n * 4);
  launch_fill_u32(stream_, filter3d_.as<uint32_t>(), n, 0u);
  filterSet_ = false;
}

// Waits for the views' streams and frees their buffers and streams (the first
// view's stream is the trainer's and stays).
void Trainer::release_views() {
  if (stream_ != nullptr) (void)hipStreamSynchronize(stream_);
  for (auto &v : views_) {
    if (v->ownsStream) {
      (void)hipStreamSynchronize(v->stream);
      (void)hipStreamDestroy(v->stream);
    }
    if (v->done != nullptr) (void)hipEventDestroy(v->done);
  }
  views_.clear();
}

Trainer::~Trainer() {
  if (prefetchCopy_.valid()) prefetchCopy_.wait();
  release_captured();
  release_views();
  if (batchStart_ != nullptr) (void)hipEventDestroy(batchStart_);
  if (batchRaster_ != nullptr) (void)hipEventDestroy(batchRaster_);
  for (auto &slot : slots_)
    for (auto &e : slot.events)
      if (e != nullptr) (void)hipEventDestroy(e);
  if (stream_ != nullptr) (void)hipStreamDestroy(stream_);
}

// The k-th view's buffers, sized for the cloud on first use; the first runs on
// the trainer's stream, the others each on a stream of their own.
ViewBuffers &Trainer::view_buffers(size_t k) {
  while (views_.size() <= k) {
    auto v = std::make_unique<ViewBuffers>();
    if (views_.empty()) {
      v->stream = stream_;
    } else {
      check(hipStreamCreateWithFlags(&v->stream, hipStreamNonBlocking), "hipStreamCreate");
      v->ownsStream = true;
    }
    check(hipEventCreateWithFlags(&v->done, hipEventDisableTiming), "hipEventCreate");
    const size_t n = count_;
    v->splats.ensure(n * sizeof(Splat));
    v->keys.ensure(n * 4);
    v->values.ensure(n * 4);
    v->tmpKeys.ensure(n * 4);
    v->tmpValues.ensure(n * 4);
    v->coverage.ensure(n * 4);
    v->tileCount.ensure(n * 4);
    v->tileRecords.ensure(n * sizeof(TileRecord));
    v->bigList.ensure(n * 4);
    v->bigBits.ensure(n * 8); // grows with the entries (grow_entries)
    if (fixedPoint_) {
      v->rawFixed.ensure(n * kRawPerSplat * 8);
      launch_fill_u32(v->stream, v->rawFixed.as<uint32_t>(), n * kRawPerSplat * 2, 0u);
    } else {
      v->rawFloat.ensure(n * kRawPerSplat * 4);
      launch_fill_u32(v->stream, v->rawFloat.as<uint32_t>(), n * kRawPerSplat, 0u);
    }
    v->scanScratch.ensure(coverage_scan_words(static_cast<uint32_t>(n)) * 4);
    v->sortScratch.ensure(radix_sort_scratch_words(n) * 4);
    v->counters.ensure(kCounterWords * 4);
    launch_fill_u32(v->stream, v->counters.as<uint32_t>(), kCounterWords, 0u);
    v->sums.ensure(2 * sizeof(double));
    check(hipStreamSynchronize(v->stream), "view buffers");
    views_.push_back(std::move(v));
    if (reservedEntries_ > 0) grow_entries(*views_.back(), reservedEntries_);
  }
  return *views_[k];
}

ViewParams Trainer::view_params(const vsplat_camera &cam) const {
  ViewParams v{};
  std::memcpy(v.view, cam.view, sizeof(v.view));
  v.position[0] = cam.position[0];
  v.position[1] = cam.position[1];
  v.position[2] = cam.position[2];
  v.nearPlane = options_.near_plane;
  v.fx = cam.fx;
  v.fy = cam.fy;
  v.cx = cam.cx;
  v.cy = cam.cy;
  // upload_camera (trainer.cpp:316-317), in float.
  v.tanHalfX = 0.5f * static_cast<float>(cam.width) / cam.fx;
  v.tanHalfY = 0.5f * static_cast<float>(cam.height) / cam.fy;
  v.width = static_cast<uint32_t>(cam.width);
  v.height = static_cast<uint32_t>(cam.height);
  v.count = count_;
  v.shDegree = static_cast<uint32_t>(activeShDegree_);
  v.shWords = shWords_;
  v.shCoefficients = static_cast<uint32_t>((activeShDegree_ + 1) * (activeShDegree_ + 1));
  v.shFormat = shFormat_;
  return v;
}

// The raster tile shape for a view size: the option's, or in auto mode the
// one the autotuner chose for that size (16x16 until it has run).
std::pair<uint32_t, uint32_t> Trainer::tile_shape_for(int width, int height) const {
  switch (options_.tile_shape) {
  case VSPLAT_TILES_8X8: return {8u, 8u};
  case VSPLAT_TILES_32X8: return {32u, 8u};
  case VSPLAT_TILES_32X32: return {32u, 32u};
  case VSPLAT_TILES_AUTO: {
    const auto found = tuned_.find({width, height});
    return found != tuned_.end() ? found->second : std::make_pair(16u, 16u);
  }
  default: return {16u, 16u};
  }
}

bool Trainer::needs_tuning(int width, int height) const {
  return options_.tile_shape == VSPLAT_TILES_AUTO && tuned_.find({width, height}) == tuned_.end();
}

void Trainer::ensure_size(ViewBuffers &b, int width, int height) {
  const auto shape = tile_shape_for(width, height);
  ensure_layout(b, width, height, shape.first, shape.second);
}

void Trainer::ensure_layout(ViewBuffers &b, int width, int height, uint32_t tileW, uint32_t tileH) {
  if (width <= 0 || height <= 0) throw Error(VSPLAT_RUNTIME_ERROR, "bad image size");
  if (width == b.width && height == b.height && tileW == b.tileW && tileH == b.tileH) return;
  set_geometry(b, width, height, tileW, tileH);
  const size_t pixels = static_cast<size_t>(width) * static_cast<size_t>(height);
  b.image.ensure(pixels * 16);
  // A backward that recomputes the transmittance reads no stops: none are kept.
  if (options_.recompute_transmittance == 0) b.stops.ensure(pixels * 4);
  b.adjoint.ensure(pixels * 16);
  b.fields.ensure(pixels * 36);
  b.ranges.ensure(static_cast<size_t>(b.tiles) * 8);
  b.tileTotals.ensure(static_cast<size_t>(b.lossTiles) * 8);
  const size_t blocks = raster_blocks(b.tiles, tileW, tileH);
  b.listOrder.ensure(blocks * 4);
  b.blockWork.ensure(blocks * 4);
  b.blockOrder.ensure(blocks * 4);
}

// The tile grid of a view size and tile shape (no buffers).
void Trainer::set_geometry(ViewBuffers &b, int width, int height, uint32_t tileW, uint32_t tileH) {
  if (width <= 0 || height <= 0) throw Error(VSPLAT_RUNTIME_ERROR, "bad image size");
  const uint64_t gridW = div_up(static_cast<uint64_t>(width), tileW), gridH = div_up(static_cast<uint64_t>(height), tileH);
  if (gridW * gridH > (1u << 24)) throw Error(VSPLAT_RUNTIME_ERROR, "the image is too large for the tile index");
  // Tile coordinates are kept as 16-bit pairs per splat.
  if (gridW > 0xFFFFu || gridH > 0xFFFFu) throw Error(VSPLAT_RUNTIME_ERROR, "the image is too wide or too tall");
  b.width = width;
  b.height = height;
  b.tileW = tileW;
  b.tileH = tileH;
  b.gridW = static_cast<uint32_t>(gridW);
  b.gridH = static_cast<uint32_t>(gridH);
  b.tiles = b.gridW * b.gridH;
  b.tileBits = 1;
  while ((uint64_t{1} << b.tileBits) < b.tiles) ++b.tileBits;
  // An entry's key holds its tile and, above it, one bit per 8x8 block of the
  // tile; the list pass packs pixel coordinates in 16 bits.
  if (b.tileBits + (tileW / 8u) * (tileH / 8u) > 32u) {
    throw Error(VSPLAT_RUNTIME_ERROR, "the image has too many tiles of this shape for the tile keys");
  }
  if (width > 0xFFFF || height > 0xFFFF) throw Error(VSPLAT_RUNTIME_ERROR, "the image is too wide or too tall");
  // The loss works on 16x16 tiles of its own, whatever the raster's shape.
  b.lossTiles = static_cast<uint32_t>(div_up(static_cast<uint64_t>(width), 16) * div_up(static_cast<uint64_t>(height), 16));
}

// Times the view's passes up to the backward (forward, loss, backward: no
// parameter changes) with each tile shape and keeps the fastest for the view's
// size. Every shape gives the same images and sums (each wave holds the same
// 8x8 block of pixels), so the choice changes only the time.
std::pair<uint32_t, uint32_t> Trainer::autotune_tiles(const vsplat_camera &camera, const uint8_t *deviceTarget, uint32_t repeats) {
  if (deviceTarget == nullptr) throw Error(VSPLAT_INVALID_ARGUMENT, "no target");
  ViewBuffers &b = *views_[0];
  const ViewParams v = view_params(camera);
  const float savedDropout = dropoutNow_;
  dropoutNow_ = 0.0f;
  hipEvent_t begin = nullptr, end = nullptr;
  check(hipEventCreate(&begin), "hipEventCreate");
  check(hipEventCreate(&end), "hipEventCreate");
  const std::pair<uint32_t, uint32_t> shapes[] = {{16u, 16u}, {8u, 8u}, {32u, 8u}, {32u, 32u}};
  std::pair<uint32_t, uint32_t> best = shapes[0];
  float bestMs = 0.0f;
  uint64_t bestEntries = 0;
  try {
    for (const auto &shape : shapes) {
      ensure_layout(b, camera.width, camera.height, shape.first, shape.second);
      // One run sizes the tile-entry buffers for this shape (also down, after
      // a shape that needed more).
      uint64_t entries = 0;
      for (int attempt = 0;; ++attempt) {
        enqueue_passes(b, v, deviceTarget, kModeGradients, nullptr, false, false);
        bool overflow = false;
        const vsplat_step_stats stats = read_stats(overflow);
        entries = stats.tile_entries;
        if (!overflow) break;
        launch_fill_u32(stream_, b.counters.as<uint32_t>() + kOverflow, 2, 0u);
        if (attempt > 0 && stats.tile_entries <= b.entryCapacity) throw Error(VSPLAT_DEVICE_ERROR, "tile entries overflowed twice");
        grow_entries(b, stats.tile_entries);
      }
      fit_entries(b, entries);
      float least = 0.0f;
      for (uint32_t r = 0; r < (repeats > 0 ? repeats : 1); ++r) {
        check(hipEventRecord(begin, stream_), "hipEventRecord");
        enqueue_passes(b, v, deviceTarget, kModeGradients, nullptr, false, false);
        check(hipEventRecord(end, stream_), "hipEventRecord");
        check(hipEventSynchronize(end), "hipEventSynchronize");
        float ms = 0.0f;
        check(hipEventElapsedTime(&ms, begin, end), "hipEventElapsedTime");
        least = r == 0 ? ms : std::min(least, ms);
      }
      if (bestMs == 0.0f || least < bestMs) {
        bestMs = least;
        best = shape;
        bestEntries = entries;
      }
    }
  } catch (...) {
    (void)hipEventDestroy(begin);
    (void)hipEventDestroy(end);
    dropoutNow_ = savedDropout;
    throw;
  }
  (void)hipEventDestroy(begin);
  (void)hipEventDestroy(end);
  dropoutNow_ = savedDropout;
  tuned_[{camera.width, camera.height}] = best;
  ensure_size(b, camera.width, camera.height);
  fit_entries(b, bestEntries);
  lastView_.width = 0; // the buffers hold the timed passes of some shape, not a step
  return best;
}

// The tile shapes of a batch's view sizes, tuned together: the batch's passes
// up to the backward (all a shape changes; the chain rule does not depend on
// it) are timed with events around the whole batch, first with each size's
// current shape (16x16 for a size not tuned yet), then for each size in turn
// with the other two shapes, keeping a shape that makes the batch faster. A
// batch's views overlap on their streams, so the fastest shape for a view on
// its own is not always the fastest inside the batch. Nothing of the trainer's
// state changes but the chosen shapes and the buffers sized for them.
void Trainer::autotune_views(uint32_t n, const vsplat_camera *cameras, const uint8_t *const *targets, uint32_t repeats,
                             uint32_t *tileW, uint32_t *tileH) {
  prepare_views(n, cameras, targets, kModeGradients, nullptr);
  std::vector<std::pair<int, int>> sizes;
  for (uint32_t k = 0; k < n; ++k) {
    const std::pair<int, int> size{cameras[k].width, cameras[k].height};
    if (std::find(sizes.begin(), sizes.end(), size) == sizes.end()) sizes.push_back(size);
    if (tuned_.find(size) == tuned_.end()) tuned_[size] = tile_shape_for(size.first, size.second);
  }
  const float savedDropout = dropoutNow_;
  dropoutNow_ = 0.0f;
  hipEvent_t begin = nullptr, end = nullptr;
  std::vector<ViewParams> params(n);
  std::vector<vsplat_step_stats> stats(n);
  // Each view's tile entries under each shape it was run with.
  std::vector<std::map<std::pair<uint32_t, uint32_t>, uint64_t>> entriesOf(n);
  const auto time_batch = [&]() -> float {
    for (uint32_t k = 0; k < n; ++k) ensure_size(*views_[k], cameras[k].width, cameras[k].height);
    // One run sizes the tile-entry buffers for these shapes (also down).
    for (int attempt = 0;; ++attempt) {
      enqueue_batch_passes(n, cameras, targets, params.data());
      bool overflow = false;
      synchronize_views(stats.data(), n, overflow);
      if (!overflow) break;
      bool grew = false;
      for (uint32_t k = 0; k < n; ++k) {
        grew = grew || stats[k].tile_entries > views_[k]->entryCapacity;
        grow_entries(*views_[k], stats[k].tile_entries);
      }
      if (attempt > 0 && !grew) throw Error(VSPLAT_DEVICE_ERROR, "tile entries overflowed twice");
    }
    for (uint32_t k = 0; k < n; ++k) {
      entriesOf[k][tuned_[{cameras[k].width, cameras[k].height}]] = stats[k].tile_entries;
      fit_entries(*views_[k], stats[k].tile_entries);
    }
    float least = 0.0f;
    for (uint32_t r = 0; r < (repeats > 0 ? repeats : 1); ++r) {
      check(hipEventRecord(begin, stream_), "hipEventRecord");
      enqueue_batch_passes(n, cameras, targets, params.data());
      check(hipEventRecord(end, stream_), "hipEventRecord");
      check(hipEventSynchronize(end), "hipEventSynchronize");
      float ms = 0.0f;
      check(hipEventElapsedTime(&ms, begin, end), "hipEventElapsedTime");
      least = r == 0 ? ms : std::min(least, ms);
    }
    queuedViewSizes_.clear();
    return least;
  };
  const std::pair<uint32_t, uint32_t> shapes[] = {{16u, 16u}, {8u, 8u}, {32u, 8u}, {32u, 32u}};
  try {
    check(hipEventCreate(&begin), "hipEventCreate");
    check(hipEventCreate(&end), "hipEventCreate");
    float best = time_batch();
    for (const auto &size : sizes) {
      auto keep = tuned_[size];
      for (const auto &shape : shapes) {
        if (shape == keep) continue;
        tuned_[size] = shape;
        const float ms = time_batch();
        if (ms < best) {
          best = ms;
          keep = shape;
        }
      }
      tuned_[size] = keep;
    }
  } catch (...) {
    if (begin != nullptr) (void)hipEventDestroy(begin);
    if (end != nullptr) (void)hipEventDestroy(end);
    dropoutNow_ = savedDropout;
    throw;
  }
  (void)hipEventDestroy(begin);
  (void)hipEventDestroy(end);
  dropoutNow_ = savedDropout;
  for (uint32_t k = 0; k < n; ++k) {
    ensure_size(*views_[k], cameras[k].width, cameras[k].height);
    const auto shape = tuned_[{cameras[k].width, cameras[k].height}];
    fit_entries(*views_[k], entriesOf[k][shape]);
    if (tileW != nullptr) tileW[k] = shape.first;
    if (tileH != nullptr) tileH[k] = shape.second;
  }
  batchTunedKey_ = sizes_key(n, cameras);
  lastView_.width = 0; // the buffers hold the timed passes, not a step
}

// The view sizes of a batch, in order, as one number.
uint64_t Trainer::sizes_key(uint32_t n, const vsplat_camera *cameras) {
  uint64_t h = 1469598103934665603ull ^ n;
  for (uint32_t k = 0; k < n; ++k) {
    h = (h ^ static_cast<uint32_t>(cameras[k].width)) * 1099511628211ull;
    h = (h ^ static_cast<uint32_t>(cameras[k].height)) * 1099511628211ull;
  }
  return h;
}

void Trainer::grow_entries(ViewBuffers &b, uint64_t entries) {
  if (entries <= b.entryCapacity) return;
  if (entries > 0xFFFFFFFFull) throw Error(VSPLAT_BUDGET_EXCEEDED, std::to_string(entries) + " tile entries exceed the 32-bit index");
  const uint64_t budget = options_.memory_budget;
  const uint64_t current = b.entryCapacity * VSPLAT_BYTES_PER_TILE_ENTRY;
  const uint64_t others = gpu_bytes() - current;
  // Checked before allocating: running out of memory is not recoverable, this is (trainer.cpp:354-360).
  if (others + entries * VSPLAT_BYTES_PER_TILE_ENTRY > budget) {
    throw Error(VSPLAT_BUDGET_EXCEEDED, std::to_string(entries) + " tile entries do not fit the memory budget");
  }
  const uint64_t wanted = std::min<uint64_t>(entries + entries / 2 + 1024, 0xFFFFFFFFull);
  const uint64_t affordable = budget == ~uint64_t{0} ? wanted : (budget - others) / VSPLAT_BYTES_PER_TILE_ENTRY;
  const uint64_t capacity = std::max(std::min(wanted, affordable), entries);
  check(hipStreamSynchronize(b.stream), "grow entries");
  for (DeviceArray *a : {&b.entryKeys, &b.entryValues, &b.tmpEntryKeys, &b.tmpEntryValues}) {
    a->release();
    a->ensure(capacity * 4);
  }
  b.sortScratch.ensure(std::max(radix_sort_scratch_words(capacity), radix_sort_scratch_words(count_)) * 4);
  b.bigBits.ensure(big_bit_words_for(capacity) * 8);
  b.entryCapacity = capacity;
}

// A view's tile-entry buffers sized for `entries` by grow_entries' rule (half
// as many again), never below the reserved count: grown when too small, and
// reallocated smaller when more than twice that, as after the tile-shape
// tuner has tried a shape that lists many more entries (the sort's grids and
// histograms are sized by the capacity, so an oversized buffer costs time on
// every pass, not only memory).
void Trainer::fit_entries(ViewBuffers &b, uint64_t entries) {
  entries = std::max(entries, reservedEntries_);
  if (entries > b.entryCapacity) {
    grow_entries(b, entries);
    return;
  }
  const uint64_t wanted = std::min<uint64_t>(entries + entries / 2 + 1024, 0xFFFFFFFFull);
  if (b.entryCapacity <= 2 * wanted) return;
  check(hipStreamSynchronize(b.stream), "fit entries");
  for (DeviceArray *a : {&b.entryKeys, &b.entryValues, &b.tmpEntryKeys, &b.tmpEntryValues}) {
    a->release();
    a->ensure(wanted * 4);
  }
  b.sortScratch.release();
  b.sortScratch.ensure(std::max(radix_sort_scratch_words(wanted), radix_sort_scratch_words(count_)) * 4);
  b.bigBits.release();
  b.bigBits.ensure(big_bit_words_for(wanted) * 8);
  b.entryCapacity = wanted;
}

void Trainer::reserve_entries(uint64_t entries) {
  reservedEntries_ = std::max(reservedEntries_, entries);
  for (auto &v : views_) grow_entries(*v, entries);
}

void Trainer::begin_timing(bool iteration) {
  if (slotPending_ == slots_.size()) {
    check(hipStreamSynchronize(stream_), "timing");
    drain_timing();
  }
  current_ = &slots_[slotHead_ % slots_.size()];
  current_->marks.clear();
  current_->iteration = iteration;
  check(hipEventRecord(current_->events[0], stream_), "hipEventRecord");
  ++slotHead_;
  ++slotPending_;
}

void Trainer::mark(const char *stage) {
  if (current_ == nullptr || current_->marks.size() + 1 > static_cast<size_t>(kMaxMarks)) return;
  current_->marks.push_back(stage);
  check(hipEventRecord(current_->events[current_->marks.size()], stream_), "hipEventRecord");
}

// Forgets the last step's events (a step that ran again).
void Trainer::cancel_timing() {
  if (current_ != nullptr && slotPending_ > 0) {
    --slotHead_;
    --slotPending_;
  }
  current_ = nullptr;
}

// After the stream has passed every pending slot: add their stage times up.
void Trainer::drain_timing() {
  for (size_t k = slotHead_ - slotPending_; k < slotHead_; ++k) {
    TimingSlot &slot = slots_[k % slots_.size()];
    for (size_t m = 0; m < slot.marks.size(); ++m) {
      float ms = 0.0f;
      check(hipEventElapsedTime(&ms, slot.events[m], slot.events[m + 1]), "hipEventElapsedTime");
      auto it = std::find_if(stageTotals_.begin(), stageTotals_.end(), [&](const auto &s) { return s.first == slot.marks[m]; });
      if (it == stageTotals_.end()) {
        stageTotals_.emplace_back(slot.marks[m], static_cast<double>(ms));
      } else {
        it->second += static_cast<double>(ms);
      }
    }
    if (slot.iteration) ++timedIterations_;
  }
  slotPending_ = 0;
  current_ = nullptr;
}

void Trainer::reset_timings() {
  check(hipStreamSynchronize(stream_), "reset timings");
  slotPending_ = 0;
  current_ = nullptr;
  stageTotals_.clear();
  timedIterations_ = 0;
}

std::vector<std::pair<std::string, double>> Trainer::timings() const {
  auto out = stageTotals_;
  for (auto &s : out) s.second /= static_cast<double>(std::max<uint64_t>(timedIterations_, 1));
  return out;
}

// A view's passes in order, on its stream, with a stage event after each: the
// tile lists, the forward pass, the loss and its adjoint, the backward pass,
// and (chain) the chain rule.
void Trainer::enqueue_passes(ViewBuffers &b, const ViewParams &view, const uint8_t *deviceTarget, uint32_t mode,
                             const vsplat_learning_rates *rates, bool timed, bool chain) {
  current_ = nullptr;
  if (timed) begin_timing(true);
  const ViewParams v = with_tiles(b, view);
  enqueue_lists(b, v, mode, true);
  // Each pass's 8x8 blocks launch heaviest first on their XCD: the forward's by
  // their tiles' list lengths, the backward's by the work the forward found.
  const bool backward = mode != kModeRender;
  uint32_t *listOrder = b.listOrder.as<uint32_t>(); // none in the allocation-free render
  if (listOrder != nullptr) launch_list_order(b.stream, b.ranges.as<uint2>(), b.tiles, b.tileW, b.tileH, listOrder);
  launch_forward(b.stream, b.splats.as<Splat>(), b.entryKeys.as<uint32_t>(), b.entryValues.as<uint32_t>(), b.tileBits, b.ranges.as<uint2>(),
                 v.width, v.height, b.tileW, b.tileH, options_.background, b.image.as<float4>(),
                 options_.recompute_transmittance != 0 ? nullptr : b.stops.as<uint32_t>(), backward ? b.blockWork.as<uint32_t>() : nullptr,
                 listOrder);
  if (backward) launch_work_order(b.stream, b.blockWork.as<uint32_t>(), raster_blocks(b.tiles, b.tileW, b.tileH), b.blockOrder.as<uint32_t>());
  mark("composite");
  if (!backward) return;
  enqueue_loss(b, v, deviceTarget);
  mark("loss");
  // Per-pixel derivatives carry 1 / (3 pixels); the scale undoes it and leaves
  // a resolution of 2^-20 (trainer.cpp:586-591 keeps 1/64 in 32 bits).
  const bool scoring = scoreShare_ > 0.0f;
  launch_backward(b.stream, fixedPoint_, options_.recompute_transmittance != 0, b.splats.as<Splat>(), b.entryKeys.as<uint32_t>(),
                  b.entryValues.as<uint32_t>(), b.tileBits, b.ranges.as<uint2>(), v.width, v.height, b.tileW, b.tileH, options_.background,
                  b.image.as<float4>(),
                  b.adjoint.as<float4>(), b.stops.as<uint32_t>(), fixed_scale(v), b.rawFixed.as<int64_t>(), b.rawFloat.as<float>(),
                  scoring ? b.edge.as<float>() : nullptr, b.blockOrder.as<uint32_t>());
  mark("backward");
  if (!chain) return;
  enqueue_chain(b, v, mode, rates, b.stream, b.counters.as<uint32_t>());
  mark(mode == kModeStep ? "chain rule and step" : "chain rule");
}

ViewParams Trainer::with_tiles(const ViewBuffers &b, const ViewParams &view) const {
  ViewParams v = view;
  v.tileW = b.tileW;
  v.tileH = b.tileH;
  return v;
}

// The projection, the depth sort, the tile counts and the sorted tile lists
// with their ranges (marks: with the stage events of a timed step).
void Trainer::enqueue_lists(ViewBuffers &b, const ViewParams &v, uint32_t mode, bool marks) {
  hipStream_t s = b.stream;
  uint32_t *counters = b.counters.as<uint32_t>();
  launch_fill_u32(s, counters, kClearedPerPass, 0u); // not the sticky overflow word
  launch_project(s, positions_.as<float>(), scales_.as<float>(), rotations_.as<float>(), opacities_.as<float>(),
                 shWordsBuf_.as<uint32_t>(), filterSet_ ? filter3d_.as<float>() : nullptr, v, options_.mip_filter_2d != 0, dropoutNow_,
                 shRetained_, shSeed_, b.gridW, b.gridH, b.splats.as<Splat>(), b.keys.as<uint32_t>(), b.values.as<uint32_t>(),
                 b.tileCount.as<uint32_t>(), b.tileRecords.as<TileRecord>(), counters, b.bigList.as<uint32_t>(),
                 b.bigBits.as<uint64_t>(), big_bit_words(b));
  if (marks) mark("project");
  launch_radix_sort(s, b.keys.as<uint32_t>(), b.values.as<uint32_t>(), b.tmpKeys.as<uint32_t>(), b.tmpValues.as<uint32_t>(), nullptr, count_,
                    32, b.sortScratch.as<uint32_t>());
  if (marks) mark("depth sort");
  launch_coverage_scan(s, b.keys.as<uint32_t>(), b.values.as<uint32_t>(), b.tileCount.as<uint32_t>(), count_, b.coverage.as<uint32_t>(),
                       counters, static_cast<uint32_t>(b.entryCapacity), mode == kModeStep, b.scanScratch.as<uint32_t>());
  if (marks) mark("count tiles");
  launch_fill_u32(s, b.ranges.as<uint32_t>(), static_cast<uint64_t>(b.tiles) * 2, 0u);
  if (b.entryCapacity > 0) {
    launch_bin_emit(s, b.keys.as<uint32_t>(), b.values.as<uint32_t>(), b.splats.as<Splat>(), b.tileRecords.as<TileRecord>(), count_, b.gridW,
                    b.tileW, b.tileH, static_cast<uint32_t>(b.width), static_cast<uint32_t>(b.height), b.tileBits,
                    b.coverage.as<uint32_t>(), counters, b.bigList.as<uint32_t>(), b.bigBits.as<uint64_t>(), b.entryKeys.as<uint32_t>(),
                    b.entryValues.as<uint32_t>());
    launch_radix_sort(s, b.entryKeys.as<uint32_t>(), b.entryValues.as<uint32_t>(), b.tmpEntryKeys.as<uint32_t>(), b.tmpEntryValues.as<uint32_t>(),
                      counters + kEntriesUsed, static_cast<uint32_t>(b.entryCapacity), b.tileBits, b.sortScratch.as<uint32_t>());
    launch_tile_ranges(s, b.entryKeys.as<uint32_t>(), counters, static_cast<uint32_t>(b.entryCapacity), b.tileBits, b.ranges.as<uint2>());
  }
  if (marks) mark("tile lists");
}

// The loss and its adjoint from the view's image (as its exposure makes it),
// with the scoring outputs when a view is scored; then the raw sums cleared
// for the backward pass.
void Trainer::enqueue_loss(ViewBuffers &b, const ViewParams &v, const uint8_t *deviceTarget) {
  hipStream_t s = b.stream;
  // The loss sees the image as this view's exposure makes it (trainer.cpp:397-425).
  const uint32_t pixels = v.width * v.height;
  const float4 *seen = b.image.as<float4>();
  // Their buffers are made on first use; the queued, batched and captured
  // paths, which must not allocate, refuse exposure and never score.
  if (exposing_) {
    b.exposed.ensure(static_cast<size_t>(pixels) * 16);
    b.exposureTotals.ensure(static_cast<size_t>(exposure_groups(pixels)) * 12 * 4);
  }
  if (scoreShare_ > 0.0f) {
    b.edge.ensure(static_cast<size_t>(pixels) * 4);
    b.errorMax.ensure(4);
  }
  if (exposing_) {
    launch_exposure(s, b.image.as<float4>(), b.exposed.as<float4>(), pixels, exposure_);
    seen = b.exposed.as<float4>();
  }
  const bool scoring = scoreShare_ > 0.0f;
  if (scoring) launch_fill_u32(s, b.errorMax.as<uint32_t>(), 1, 0u);
  launch_loss(s, seen, deviceTarget, v.width, v.height, options_.lambda_ssim, b.fields.as<float>(), b.adjoint.as<float4>(),
              b.tileTotals.as<float2>(), b.sums.as<double>(), scoring ? b.errorMax.as<uint32_t>() : nullptr,
              scoring ? b.edge.as<float>() : nullptr);
  if (exposing_) {
    launch_exposure_backward(s, b.image.as<float4>(), b.adjoint.as<float4>(), pixels, exposure_, b.exposureTotals.as<float>());
  }
  if (scoring) launch_error_mask(s, b.adjoint.as<float4>(), b.errorMax.as<uint32_t>(), pixels, scoreShare_);
  if (fixedPoint_) {
    launch_fill_u32(s, b.rawFixed.as<uint32_t>(), static_cast<uint64_t>(count_) * kRawPerSplat * 2, 0u);
  } else {
    launch_fill_u32(s, b.rawFloat.as<uint32_t>(), static_cast<uint64_t>(count_) * kRawPerSplat, 0u);
  }
}

ChainParams Trainer::chain_params(const ViewParams &v, uint32_t mode, const vsplat_learning_rates *rates) const {
  ChainParams p{};
  p.opacityReg = options_.opacity_reg / static_cast<float>(count_);
  p.scaleReg = options_.scale_reg / static_cast<float>(count_);
  p.mipFilter2d = options_.mip_filter_2d ? 1u : 0u;
  p.shDropout = dropoutNow_;
  p.shRetained = shRetained_;
  p.shSeed = shSeed_;
  p.mode = mode;
  p.flags = (options_.selective_adam ? kFlagSelective : 0u) | (statisticsNext_ ? kFlagStatistics : 0u) |
            (accumulateNext_ ? kFlagAccumulate : 0u) | (secondOrder_ ? kFlagSecondOrder : 0u) |
            (poseNext_ && (mode == kModeGradients || mode == kModeStep) ? kFlagPose : 0u);
  // Second order (trainer.cpp:455-474): the trust region, and the curvature's
  // bias correction 1 / (1 - beta^k), beta^k kept in float by repeated
  // multiplication over k = max(updates, 1).
  p.curvatureBeta = 0.999f;
  float betaPower = 1.0f;
  for (uint32_t k = 0; k < std::max(curvatureUpdates_, 1u); ++k) betaPower = betaPower * p.curvatureBeta;
  p.curvatureScale = 1.0f / (1.0f - betaPower);
  p.trustRegion = rates != nullptr ? rates->trust_region : 0.0f;
  if (mode == kModeCurvature) {
    p.seed = curvatureUpdates_;
    p.perValue = 1.0f / (3.0f * static_cast<float>(v.width) * static_cast<float>(v.height));
  }
  p.beta1 = 0.9f;
  p.beta2 = 0.999f;
  p.epsilon = 1e-15f; // trainer.cpp:23
  p.adamState = adamState_.as<uint32_t>();
  if (mode == kModeStep) {
    p.rates[0] = rates->position;
    p.rates[1] = rates->scale;
    p.rates[2] = rates->rotation;
    p.rates[3] = rates->opacity;
    p.rates[4] = rates->sh;
    p.restScale = rates->sh_rest_scale;
    // The bias corrections 1 - beta^t (trainer.cpp:467-468) come from the
    // device's step count, beta^t kept by repeated multiplication.
  }
  p.shFloats = shFloats_;
  p.fixedPoint = fixedPoint_ ? 1u : 0u;
  p.fixedScale = mode == kModeCurvature ? curvature_fixed_scale() : fixed_scale(v);
  return p;
}

void Trainer::ensure_gradients() {
  const size_t n = count_;
  const size_t floats[5] = {n * 3, n * 3, n * 4, n, n * shFloats_};
  for (int g = 0; g < 5; ++g) grad_[g].ensure(floats[g] * 4);
}

ChainBuffers Trainer::chain_buffers(const ViewBuffers &v) const {
  ChainBuffers b{};
  b.positions = positions_.as<float>();
  b.scales = scales_.as<float>();
  b.rotations = rotations_.as<float>();
  b.opacities = opacities_.as<float>();
  b.shWords = shWordsBuf_.as<uint32_t>();
  b.poseTerms = poseNext_ ? poseTerms_.as<float>() : nullptr;
  b.filter3d = filterSet_ ? filter3d_.as<float>() : nullptr;
  b.splats = v.splats.as<Splat>();
  b.rawFixed = v.rawFixed.as<int64_t>();
  b.rawFloat = v.rawFloat.as<float>();
  b.shPass = shPass_.as<float>();
  b.stats = stats_.as<float>();
  b.statePos = state_[0].as<uint32_t>();
  b.stateScale = state_[1].as<uint32_t>();
  b.stateRot = state_[2].as<uint32_t>();
  b.stateOpacity = state_[3].as<uint32_t>();
  b.stateSh = state_[4].as<uint32_t>();
  b.gradPos = grad_[0].as<float>();
  b.gradScale = grad_[1].as<float>();
  b.gradRot = grad_[2].as<float>();
  b.gradOpacity = grad_[3].as<float>();
  b.gradSh = grad_[4].as<float>();
  b.dirTerms = dirTerms_.as<float4>();
  return b;
}

void Trainer::enqueue_chain(const ViewBuffers &b, const ViewParams &v, uint32_t mode, const vsplat_learning_rates *rates,
                            hipStream_t stream, uint32_t *counters) {
  if (mode == kModeGradients) ensure_gradients();
  const bool pose = poseNext_ && (mode == kModeGradients || mode == kModeStep);
  if (pose) launch_fill_u32(stream, poseTerms_.as<uint32_t>(), static_cast<uint64_t>(count_) * 6, 0u);
  launch_chain(stream, chain_buffers(b), v, chain_params(v, mode, rates), counters);
  if (pose) launch_pose_totals(stream, poseTerms_.as<float>(), count_, poseTotals_.as<float>());
  if (mode == kModeStep) enqueue_noise(b, rates, stream, counters);
}

// After a step with TrainerOptions::position_noise set (trainer.cpp:601): the
// noise scaled by the position rate, its iteration the device's step count
// (the step just taken), nothing when the step was skipped.
void Trainer::enqueue_noise(const ViewBuffers &b, const vsplat_learning_rates *rates, hipStream_t stream, const uint32_t *counters) {
  if (!(options_.position_noise > 0.0f) || rates == nullptr) return;
  launch_noise(stream, chain_buffers(b), count_, options_.position_noise * rates->position, adamState_.as<uint32_t>(), 0u, counters);
}

// Trainer::add_position_noise (trainer.cpp:806-808): one round of the noise
// with the caller's scaler and iteration.
void Trainer::add_position_noise(float scaler, uint32_t iteration) {
  launch_noise(stream_, chain_buffers(*views_[0]), count_, scaler, nullptr, iteration, nullptr);
  check(hipStreamSynchronize(stream_), "position noise");
}

// Trainer::reset_opacity (trainer.cpp:810-819): opacities capped at the logit
// of `maximum`, and the opacity's optimiser state cleared.
void Trainer::reset_opacity(float maximum) {
  if (!(maximum > 0.0f && maximum < 1.0f)) throw Error(VSPLAT_INVALID_ARGUMENT, "the maximum opacity must lie in (0, 1)");
  const float limit = std::log(maximum / (1.0f - maximum));
  launch_clamp_opacity(stream_, opacities_.as<float>(), count_, limit);
  launch_fill_u32(stream_, state_[3].as<uint32_t>(), count_, 0u); // one word per opacity: both moments
  check(hipStreamSynchronize(stream_), "reset opacity");
}

vsplat_step_stats Trainer::read_stats(bool &overflow) {
  ViewBuffers &b = *views_[0];
  uint8_t *host = readback_.as<uint8_t>();
  check(hipMemcpyAsync(host, b.counters.as<uint32_t>(), kCounterWords * 4, hipMemcpyDeviceToHost, stream_), "read counters");
  check(hipMemcpyAsync(host + kCounterWords * 4, b.sums.as<double>(), 2 * sizeof(double), hipMemcpyDeviceToHost, stream_), "read sums");
  check(hipStreamSynchronize(stream_), "step");
  uint32_t counters[kCounterWords];
  double sums[2];
  std::memcpy(counters, host, sizeof(counters));
  std::memcpy(sums, host + kCounterWords * 4, sizeof(sums));
  overflow = counters[kOverflow] != 0;
  return make_stats(counters, sums, b.width, b.height);
}

vsplat_step_stats Trainer::make_stats(const uint32_t *counters, const double *sums, int width, int height) const {
  vsplat_step_stats st{};
  // trainer.cpp:606-620: the per-tile totals summed in double, then in float as the reference.
  const double perValue = 1.0 / (3.0 * static_cast<double>(width) * static_cast<double>(height));
  st.l1 = static_cast<float>(sums[1] * perValue);
  st.ssim = static_cast<float>(sums[0] * perValue);
  st.loss = (1.0f - options_.lambda_ssim) * st.l1 + options_.lambda_ssim * (1.0f - st.ssim);
  st.visible_splats = counters[kVisible];
  st.tile_entries = counters[kEntries];
  return st;
}

void Trainer::wait_prefetch() {
  if (prefetchCopy_.valid()) prefetchCopy_.get();
}

// Starts copying the next step's target into pinned memory on another host
// thread; the step that takes this image waits for the copy and swaps the
// pinned buffers instead of copying (stage_target).
void Trainer::prefetch_target(const uint8_t *target, size_t bytes) {
  if (target == nullptr || bytes == 0) return;
  wait_prefetch(); // the pinned buffer may be resized below
  next_.ensure(bytes);
  uint8_t *into = next_.as<uint8_t>();
  prefetched_ = target;
  prefetchedBytes_ = bytes;
  prefetchCopy_ = std::async(std::launch::async, [into, target, bytes] { std::memcpy(into, target, bytes); });
}

void Trainer::stage_target(const uint8_t *target, size_t bytes) {
  if (prefetched_ == target && prefetchedBytes_ == bytes) {
    // Started during the previous step (trainer.cpp:572-576).
    wait_prefetch();
    staging_.swap(next_);
    prefetched_ = nullptr;
  } else {
    staging_.ensure(bytes);
    std::memcpy(staging_.as<uint8_t>(), target, bytes);
  }
  target_.ensure(bytes);
  check(hipMemcpyAsync(target_.as<uint8_t>(), staging_.as<uint8_t>(), bytes, hipMemcpyHostToDevice, stream_), "upload target");
}

vsplat_step_stats Trainer::run(const vsplat_camera &camera, const uint8_t *target, size_t bytes, uint32_t mode,
                               const vsplat_learning_rates *rates, const uint8_t *next, size_t nextBytes) {
  ViewBuffers &b = *views_[0];
  ensure_size(b, camera.width, camera.height);
  const size_t expected = static_cast<size_t>(b.width) * static_cast<size_t>(b.height) * 4;
  if (target == nullptr || bytes != expected) throw Error(VSPLAT_RUNTIME_ERROR, "the target image does not match the camera");
  if (mode == kModeStep && rates == nullptr) throw Error(VSPLAT_INVALID_ARGUMENT, "a step needs learning rates");
  // Dropout trains; scoring and statistics see the whole model (trainer.cpp:569).
  dropoutNow_ = rates != nullptr ? shDropout_ : 0.0f;
  stage_target(target, bytes);
  if (next != nullptr && nextBytes > 0) prefetch_target(next, nextBytes);
  if (needs_tuning(camera.width, camera.height)) autotune_tiles(camera, target_.as<uint8_t>(), 2);
  const ViewParams v = view_params(camera);
  statisticsNext_ = true; // run() always adds to the densification statistics (trainer.cpp:600)
  // The pose gradient (trainer.cpp:455, 630-637), from the chain rule's pass.
  poseNext_ = options_.pose_gradients != 0 && (mode == kModeGradients || mode == kModeStep);
  struct ClearOnExit {
    bool &flag;
    ~ClearOnExit() { flag = false; } // also when a pass throws: later calls write no pose terms
  } clearPose{poseNext_};
  if (poseNext_) {
    poseTerms_.ensure(static_cast<size_t>(count_) * 6 * 4);
    poseTotals_.ensure(static_cast<size_t>(pose_groups(count_)) * 6 * 4);
  }
  vsplat_step_stats stats{};
  for (int attempt = 0;; ++attempt) {
    enqueue_passes(b, v, target_.as<uint8_t>(), mode, rates, true);
    bool overflow = false;
    stats = read_stats(overflow);
    if (!overflow) break;
    cancel_timing();
    launch_fill_u32(stream_, b.counters.as<uint32_t>() + kOverflow, 2, 0u); // overflow and skipped
    if (attempt > 0 && stats.tile_entries <= b.entryCapacity) throw Error(VSPLAT_DEVICE_ERROR, "tile entries overflowed twice");
    grow_entries(b, stats.tile_entries);
  }
  drain_timing();
  if (exposing_) {
    // dL/d(transform): the workgroups' totals summed in double in order (trainer.cpp:620-628).
    const uint32_t groups = exposure_groups(static_cast<uint32_t>(camera.width) * static_cast<uint32_t>(camera.height));
    std::vector<float> totals(static_cast<size_t>(groups) * 12);
    check(hipMemcpy(totals.data(), b.exposureTotals.as<float>(), totals.size() * 4, hipMemcpyDeviceToHost), "read exposure totals");
    double sums[12] = {};
    for (uint32_t g = 0; g < groups; ++g)
      for (int k = 0; k < 12; ++k) sums[k] += static_cast<double>(totals[static_cast<size_t>(g) * 12 + k]);
    for (int k = 0; k < 12; ++k) stats.exposure_gradient[k] = static_cast<float>(sums[k]);
  }
  if (poseNext_) {
    // dL/d(w, v): the groups' totals summed in double in order (trainer.cpp:630-637).
    const uint32_t groups = pose_groups(count_);
    std::vector<float> totals(static_cast<size_t>(groups) * 6);
    check(hipMemcpy(totals.data(), poseTotals_.as<float>(), totals.size() * 4, hipMemcpyDeviceToHost), "read pose totals");
    double sums[6] = {};
    for (uint32_t g = 0; g < groups; ++g)
      for (int k = 0; k < 6; ++k) sums[k] += static_cast<double>(totals[static_cast<size_t>(g) * 6 + k]);
    for (int k = 0; k < 6; ++k) stats.pose_gradient[k] = static_cast<float>(sums[k]);
  }
  if (mode == kModeGradients) haveGradients_ = true;
  lastView_ = v;
  lastViewIndex_ = 0;
  return stats;
}

// Trainer::score_view (trainer.cpp:519-532): statistics of the view with the
// high-error pixels marked (errorShare of the view's worst error), then each
// Gaussian's marked pixels, that count times the view's loss, and its edge
// score added to its statistics, scaled to pixels times pixelWeight.
vsplat_step_stats Trainer::score_view(const vsplat_camera &camera, const uint8_t *target, size_t bytes, float errorShare,
                                      float pixelWeight) {
  if (!(errorShare > 0.0f)) throw Error(VSPLAT_INVALID_ARGUMENT, "the error share must be above 0");
  scoreShare_ = errorShare;
  vsplat_step_stats stats{};
  try {
    stats = run(camera, target, bytes, kModeStatistics, nullptr, nullptr, 0);
  } catch (...) {
    scoreShare_ = 0.0f;
    throw;
  }
  scoreShare_ = 0.0f;
  const ViewBuffers &b = *views_[0];
  const float toPixels = 3.0f * static_cast<float>(camera.width) * static_cast<float>(camera.height) * pixelWeight;
  launch_score(stream_, fixedPoint_ ? b.rawFixed.as<int64_t>() : nullptr, b.rawFloat.as<float>(), fixed_scale(lastView_),
               stats_.as<float4>(), count_, toPixels, stats.loss);
  check(hipStreamSynchronize(stream_), "score");
  return stats;
}

void Trainer::set_exposure(const float transform[12]) {
  static const float neutral[12] = {1.0f, 0.0f, 0.0f, 0.0f, 0.0f, 1.0f, 0.0f, 0.0f, 0.0f, 0.0f, 1.0f, 0.0f};
  std::memcpy(exposure_.transform, transform, sizeof(exposure_.transform));
  // A neutral exposure changes nothing, so its passes are left out (trainer.cpp:399-400).
  exposing_ = std::memcmp(transform, neutral, sizeof(neutral)) != 0;
}

void Trainer::render(const vsplat_camera &camera, float *out) {
  ViewBuffers &b = *views_[0];
  ensure_size(b, camera.width, camera.height);
  dropoutNow_ = 0.0f; // reports and evaluation see the whole model
  const ViewParams v = view_params(camera);
  for (int attempt = 0;; ++attempt) {
    enqueue_passes(b, v, nullptr, kModeRender, nullptr, false);
    bool overflow = false;
    const vsplat_step_stats stats = read_stats(overflow);
    if (!overflow) break;
    launch_fill_u32(stream_, b.counters.as<uint32_t>() + kOverflow, 2, 0u); // overflow and skipped
    if (attempt > 0 && stats.tile_entries <= b.entryCapacity) throw Error(VSPLAT_DEVICE_ERROR, "tile entries overflowed twice");
    grow_entries(b, stats.tile_entries);
  }
  // As the reference's render uploads its camera, a curvature update now takes this view.
  lastView_ = v;
  lastViewIndex_ = 0;
  const size_t pixels = static_cast<size_t>(b.width) * static_cast<size_t>(b.height);
  check(hipMemcpy(out, b.image.as<float4>(), pixels * 16, hipMemcpyDeviceToHost), "download image");
}

// render() without allocating (vsplat_render_noalloc). The per-Gaussian,
// per-entry and per-pixel buffers are carved out of `scratch` at the sizes of
// the VSPLAT_NOALLOC_MAX_* limits, the same layout on every call; the image is
// written straight into the caller's buffer. The counters are cleared first
// (the scratch keeps the last call's contents); every other carved buffer is
// written before it is read. The
// trainer's own view buffers, its last view and its statistics are not touched.
void Trainer::render_into(const vsplat_camera &camera, float *deviceOut, uint64_t capacity, void *scratch, uint64_t scratchBytes) {
  if (count_ > VSPLAT_NOALLOC_MAX_GAUSSIANS) {
    throw Error(VSPLAT_INVALID_ARGUMENT, std::to_string(count_) + " Gaussians: the allocation-free render takes at most " +
                                             std::to_string(VSPLAT_NOALLOC_MAX_GAUSSIANS));
  }
  if (camera.width <= 0 || camera.height <= 0) throw Error(VSPLAT_RUNTIME_ERROR, "bad image size");
  const uint64_t pixels = static_cast<uint64_t>(camera.width) * static_cast<uint64_t>(camera.height);
  if (pixels > VSPLAT_NOALLOC_MAX_PIXELS) {
    throw Error(VSPLAT_INVALID_ARGUMENT, std::to_string(pixels) + " pixels: the allocation-free render takes at most " +
                                             std::to_string(VSPLAT_NOALLOC_MAX_PIXELS));
  }
  if (capacity < pixels * 16) throw Error(VSPLAT_INVALID_ARGUMENT, "the output holds fewer than width*height*16 bytes");
  ViewBuffers b;
  b.stream = stream_;
  const auto shape = tile_shape_for(camera.width, camera.height);
  set_geometry(b, camera.width, camera.height, shape.first, shape.second);
  if (b.tiles > VSPLAT_NOALLOC_MAX_TILES) {
    throw Error(VSPLAT_INVALID_ARGUMENT, std::to_string(b.tiles) + " tiles: the allocation-free render takes at most " +
                                             std::to_string(VSPLAT_NOALLOC_MAX_TILES));
  }
  uint8_t *base = static_cast<uint8_t *>(scratch);
  uint64_t used = 0;
  const auto carve = [&](DeviceArray &a, uint64_t bytes) {
    bytes = div_up(bytes, 256) * 256;
    if (used + bytes > scratchBytes) throw Error(VSPLAT_RUNTIME_ERROR, "the allocation-free render's scratch is smaller than its layout");
    a.borrow(base + used, bytes);
    used += bytes;
  };
  const uint64_t n = VSPLAT_NOALLOC_MAX_GAUSSIANS, entries = VSPLAT_NOALLOC_MAX_ENTRIES;
  carve(b.splats, n * sizeof(Splat));
  for (DeviceArray *a : {&b.keys, &b.values, &b.tmpKeys, &b.tmpValues, &b.coverage, &b.tileCount}) carve(*a, n * 4);
  carve(b.tileRecords, n * sizeof(TileRecord));
  carve(b.bigList, n * 4);
  carve(b.bigBits, (n + entries / 16) * 8);
  for (DeviceArray *a : {&b.entryKeys, &b.entryValues, &b.tmpEntryKeys, &b.tmpEntryValues}) carve(*a, entries * 4);
  carve(b.stops, uint64_t{VSPLAT_NOALLOC_MAX_PIXELS} * 4);
  carve(b.ranges, uint64_t{VSPLAT_NOALLOC_MAX_TILES} * 8);
  carve(b.sortScratch, radix_sort_scratch_words(std::max(entries, n)) * 4);
  carve(b.scanScratch, coverage_scan_words(static_cast<uint32_t>(n)) * 4);
  carve(b.counters, kCounterWords * 4);
  b.image.borrow(deviceOut, capacity);
  b.entryCapacity = entries;
  // Fill kernels rather than memsets, as everywhere in a step.
  launch_fill_u32(stream_, b.counters.as<uint32_t>(), kCounterWords, 0u);
  dropoutNow_ = 0.0f; // as render(): the whole model
  enqueue_passes(b, view_params(camera), nullptr, kModeRender, nullptr, false);
  uint32_t *host = readback_.as<uint32_t>(); // pinned at construction
  check(hipMemcpyAsync(host, b.counters.as<uint32_t>(), kCounterWords * 4, hipMemcpyDeviceToHost, stream_), "read counters");
  check(hipStreamSynchronize(stream_), "render");
  if (host[kOverflow] != 0) {
    throw Error(VSPLAT_BUDGET_EXCEEDED, std::to_string(host[kEntries]) + " tile entries: the allocation-free render lists at most " +
                                            std::to_string(VSPLAT_NOALLOC_MAX_ENTRIES));
  }
}

void Trainer::apply_gradients(const vsplat_learning_rates &rates) {
  if (!haveGradients_) throw Error(VSPLAT_LOGIC_ERROR, "no gradients: compute_gradients() comes first");
  if (secondOrder_) throw Error(VSPLAT_UNSUPPORTED, "second-order steps are taken by step(), not from stored gradients");
  dropoutNow_ = 0.0f;
  statisticsNext_ = false;
  begin_timing(false);
  // Adam from the stored gradients: the same rule and rounding as the chain
  // rule's own step, and it also serves a batch's summed gradients.
  launch_adam(stream_, chain_buffers(*views_[0]), lastView_, chain_params(lastView_, kModeStep, &rates), views_[0]->counters.as<uint32_t>());
  mark("adam step");
  check(hipStreamSynchronize(stream_), "apply gradients");
  drain_timing();
}

void Trainer::enqueue(const vsplat_camera &camera, const uint8_t *deviceTarget, uint32_t mode, const vsplat_learning_rates *rates) {
  if (mode != kModeGradients && mode != kModeStep) throw Error(VSPLAT_INVALID_ARGUMENT, "mode must be 1 (gradients) or 2 (step)");
  if (mode == kModeStep && rates == nullptr) throw Error(VSPLAT_INVALID_ARGUMENT, "a step needs learning rates");
  if (deviceTarget == nullptr) throw Error(VSPLAT_INVALID_ARGUMENT, "no target");
  if (exposing_) throw Error(VSPLAT_UNSUPPORTED, "queued steps take the neutral exposure: its gradient is read by the synchronous calls");
  ViewBuffers &b = *views_[0];
  ensure_size(b, camera.width, camera.height);
  dropoutNow_ = rates != nullptr ? shDropout_ : 0.0f;
  statisticsNext_ = true;
  const ViewParams v = view_params(camera);
  enqueue_passes(b, v, deviceTarget, mode, rates, timeQueued_);
  if (mode == kModeGradients) haveGradients_ = true;
  lastView_ = v;
  lastViewIndex_ = 0;
}

vsplat_step_stats Trainer::synchronize() {
  bool overflow = false;
  const vsplat_step_stats stats = read_stats(overflow);
  drain_timing();
  if (overflow) {
    // Adam's step count on the device (and so its bias corrections) covers
    // only the steps that ran: a skipped step does not advance it.
    clear_overflow();
    throw Error(VSPLAT_RUNTIME_ERROR, "a queued view needed more tile entries than were reserved; it and the views queued after it were skipped");
  }
  return stats;
}

// Clears every view's sticky overflow word and skipped count (after a failed
// queue has been reported).
void Trainer::clear_overflow() {
  for (auto &v : views_) launch_fill_u32(stream_, v->counters.as<uint32_t>() + kOverflow, 2, 0u);
  check(hipStreamSynchronize(stream_), "clear overflow");
}

// The batched step (no reference counterpart): every view's passes up to its
// backward pass, each on its own buffers and stream, side by side; then on the
// trainer's stream, once all are through, each view's chain rule in view order
// adding its gradients to the sums (the first view writes them), and one Adam
// step from the sums. A batch is one training iteration: one Adam step count,
// one bias correction. A view that does not fit skips the whole batch: its
// overflow word is or-ed into the first view's, which the chain rule and the
// Adam step check.
// Checks a batch's arguments and makes every buffer it needs (so that queueing
// it allocates nothing).
void Trainer::prepare_views(uint32_t n, const vsplat_camera *cameras, const uint8_t *const *targets, uint32_t mode,
                            const vsplat_learning_rates *rates) {
  if (n == 0 || cameras == nullptr || targets == nullptr) throw Error(VSPLAT_INVALID_ARGUMENT, "a batch needs views and targets");
  if (mode != kModeGradients && mode != kModeStep) throw Error(VSPLAT_INVALID_ARGUMENT, "mode must be 1 (gradients) or 2 (step)");
  if (mode == kModeStep && rates == nullptr) throw Error(VSPLAT_INVALID_ARGUMENT, "a step needs learning rates");
  if (secondOrder_ && mode == kModeStep) throw Error(VSPLAT_UNSUPPORTED, "a batch of views takes Adam's step, not the second-order one");
  if (exposing_) throw Error(VSPLAT_UNSUPPORTED, "a batch of views takes the neutral exposure (one exposure is one photo's)");
  for (uint32_t k = 0; k < n; ++k) {
    if (targets[k] == nullptr) throw Error(VSPLAT_INVALID_ARGUMENT, "no target for a view");
    ensure_size(view_buffers(k), cameras[k].width, cameras[k].height);
  }
  ensure_gradients();
  viewStats_.ensure(static_cast<size_t>(n) * kViewStatsBytes);
}

// Each view's passes up to its backward pass, on the view's stream, after
// everything already queued on the trainer's stream (the previous step's
// update of the parameters); each view's counters and loss sums copied to its
// stats slot; the trainer's stream then waits for every view and gathers their
// sticky flags into the first view's counters, the batch's status: a view that
// did not fit skips the batch's whole chain rule and step.
void Trainer::enqueue_view_passes(uint32_t n, const vsplat_camera *cameras, const uint8_t *const *targets, ViewParams *params) {
  queuedViewSizes_.clear();
  current_ = nullptr; // no stage events in a batch
  check(hipEventRecord(batchStart_, stream_), "hipEventRecord");
  for (uint32_t k = 0; k < n; ++k) {
    ViewBuffers &b = *views_[k];
    if (b.stream != stream_) check(hipStreamWaitEvent(b.stream, batchStart_, 0), "hipStreamWaitEvent");
    params[k] = view_params(cameras[k]);
    enqueue_passes(b, params[k], targets[k], kModeGradients, nullptr, false, false);
    check(hipEventRecord(b.done, b.stream), "hipEventRecord");
    queuedViewSizes_.emplace_back(cameras[k].width, cameras[k].height);
  }
  for (uint32_t k = 1; k < n; ++k) check(hipStreamWaitEvent(stream_, views_[k]->done, 0), "hipStreamWaitEvent");
  gather_view_status(n);
}

// Once every view's passes are through (on the trainer's stream): each view's
// counters and loss sums into its stats slot, and the views' sticky flags into
// the first view's counters, the batch's status.
void Trainer::gather_view_status(uint32_t n) {
  uint32_t *status = views_[0]->counters.as<uint32_t>();
  for (uint32_t first = 0; first < n; first += kStatusViewsPerLaunch) {
    ViewStatus v{};
    v.first = first;
    v.count = std::min(n - first, kStatusViewsPerLaunch);
    for (uint32_t k = 0; k < v.count; ++k) {
      v.counters[k] = views_[first + k]->counters.as<uint32_t>();
      v.sums[k] = views_[first + k]->sums.as<double>();
    }
    launch_gather_views(stream_, v, viewStats_.as<uint8_t>(), status);
  }
}

// A batch's passes up to each view's backward pass, by the schedule the options ask for.
void Trainer::enqueue_batch_passes(uint32_t n, const vsplat_camera *cameras, const uint8_t *const *targets, ViewParams *params) {
  if (options_.batch_raster == VSPLAT_BATCH_ONE_LAUNCH) {
    enqueue_view_passes_one_launch(n, cameras, targets, params);
  } else {
    enqueue_view_passes(n, cameras, targets, params);
  }
}

// A batch with one forward launch and one backward launch over every view's
// 8x8 pixel blocks (one per tile shape among the views, eight views a launch):
// each view's lists on its own stream; once every view's lists are made, the
// forward launch on the trainer's stream; each view's loss on its own stream;
// then the backward launch. Each (view, block) pair runs the single-view
// kernels' code, so the images, sums and gradients are the same bits.
void Trainer::enqueue_view_passes_one_launch(uint32_t n, const vsplat_camera *cameras, const uint8_t *const *targets, ViewParams *params) {
  queuedViewSizes_.clear();
  current_ = nullptr; // no stage events in a batch
  check(hipEventRecord(batchStart_, stream_), "hipEventRecord");
  for (uint32_t k = 0; k < n; ++k) {
    ViewBuffers &b = *views_[k];
    if (b.stream != stream_) check(hipStreamWaitEvent(b.stream, batchStart_, 0), "hipStreamWaitEvent");
    params[k] = view_params(cameras[k]);
    enqueue_lists(b, with_tiles(b, params[k]), kModeGradients, false);
    check(hipEventRecord(b.done, b.stream), "hipEventRecord");
    queuedViewSizes_.emplace_back(cameras[k].width, cameras[k].height);
  }
  for (uint32_t k = 1; k < n; ++k) check(hipStreamWaitEvent(stream_, views_[k]->done, 0), "hipStreamWaitEvent");
  launch_batch_raster(n, params, false);
  check(hipEventRecord(batchRaster_, stream_), "hipEventRecord");
  for (uint32_t k = 0; k < n; ++k) {
    ViewBuffers &b = *views_[k];
    if (b.stream != stream_) check(hipStreamWaitEvent(b.stream, batchRaster_, 0), "hipStreamWaitEvent");
    enqueue_loss(b, with_tiles(b, params[k]), targets[k]);
    check(hipEventRecord(b.done, b.stream), "hipEventRecord");
  }
  for (uint32_t k = 1; k < n; ++k) check(hipStreamWaitEvent(stream_, views_[k]->done, 0), "hipStreamWaitEvent");
  gather_view_status(n);
  launch_batch_raster(n, params, true);
}

// The batch's forward (or backward) passes on the trainer's stream: the views
// grouped by tile shape, up to kRasterViewsPerLaunch per launch.
void Trainer::launch_batch_raster(uint32_t n, const ViewParams *params, bool backward) {
  std::vector<bool> done(n, false);
  for (uint32_t first = 0; first < n; ++first) {
    if (done[first]) continue;
    const uint32_t tileW = views_[first]->tileW, tileH = views_[first]->tileH;
    RasterViews group{};
    uint32_t blocks = 0;
    const auto flush = [&]() {
      if (group.count == 0) return;
      if (backward) {
        launch_backward_views(stream_, fixedPoint_, options_.recompute_transmittance != 0, group, tileW, tileH, options_.background);
      } else {
        launch_forward_views(stream_, group, tileW, tileH, options_.background);
      }
      group = RasterViews{};
      blocks = 0;
    };
    for (uint32_t k = first; k < n; ++k) {
      ViewBuffers &b = *views_[k];
      if (done[k] || b.tileW != tileW || b.tileH != tileH) continue;
      done[k] = true;
      if (group.count == kRasterViewsPerLaunch) flush();
      RasterView &w = group.view[group.count++];
      w.splats = b.splats.as<Splat>();
      w.entryKeys = b.entryKeys.as<uint32_t>();
      w.entryValues = b.entryValues.as<uint32_t>();
      w.tileBits = b.tileBits;
      w.ranges = b.ranges.as<uint2>();
      w.image = b.image.as<float4>();
      w.stops = options_.recompute_transmittance != 0 ? nullptr : b.stops.as<uint32_t>();
      w.adjoint = b.adjoint.as<float4>();
      w.rawFixed = fixedPoint_ ? b.rawFixed.as<int64_t>() : nullptr;
      w.rawFloat = b.rawFloat.as<float>();
      w.fixedScale = fixed_scale(params[k]);
      w.width = b.width;
      w.height = b.height;
      w.gridW = b.gridW;
      w.firstBlock = blocks;
      w.blocks = raster_blocks(b.tiles, tileW, tileH);
      blocks += w.blocks;
    }
    flush();
  }
}

// The batch's chain rule on the trainer's stream, after every view's passes
// (enqueue_view_passes): chain_views_kernel over the views in chunks of
// kChainViewsPerLaunch, each chunk continuing the sums of the one before; the
// last takes the step. (Each view's chain rule on its own stream as soon as
// its backward had run, in view order, measured slower: the views' streams
// share the device's four hardware queues, so a view's chain pass waited
// behind a later view's passes queued before it.)
void Trainer::enqueue_view_chains(uint32_t n, const ViewParams *params, uint32_t mode, const vsplat_learning_rates *rates) {
  uint32_t *status = views_[0]->counters.as<uint32_t>();
  const ChainBuffers buffers = chain_buffers(*views_[0]);
  for (uint32_t first = 0; first < n; first += kChainViewsPerLaunch) {
    const uint32_t count = std::min(n - first, kChainViewsPerLaunch);
    const bool last = first + count == n;
    ChainViews chunk{};
    chunk.count = count;
    for (uint32_t k = 0; k < count; ++k) {
      ChainView &w = chunk.view[k];
      const ViewBuffers &b = *views_[first + k];
      w.c = params[first + k];
      w.rawFixed = fixedPoint_ ? b.rawFixed.as<int64_t>() : nullptr;
      w.rawFloat = b.rawFloat.as<float>();
      w.splats = b.splats.as<Splat>();
      w.fixedScale = fixed_scale(w.c);
    }
    accumulateNext_ = first > 0;
    launch_chain_views(stream_, buffers, chunk, chain_params(params[first], last ? mode : kModeGradients, rates), status);
  }
  accumulateNext_ = false;
}

void Trainer::enqueue_views(uint32_t n, const vsplat_camera *cameras, const uint8_t *const *targets, uint32_t mode,
                            const vsplat_learning_rates *rates) {
  prepare_views(n, cameras, targets, mode, rates);
  dropoutNow_ = rates != nullptr ? shDropout_ : 0.0f;
  statisticsNext_ = true;
  std::vector<ViewParams> params(n);
  enqueue_batch_passes(n, cameras, targets, params.data());
  uint32_t *status = views_[0]->counters.as<uint32_t>();
  enqueue_view_chains(n, params.data(), mode, rates);
  const ChainBuffers buffers = chain_buffers(*views_[0]);
  lastView_ = params[n - 1];
  lastViewIndex_ = n - 1;
  if (mode == kModeStep) {
    launch_sh_adam(stream_, buffers, lastView_, chain_params(lastView_, kModeStep, rates), status);
    enqueue_noise(*views_[0], rates, stream_, status);
  } else {
    haveGradients_ = true;
  }
}

// What a captured batch depends on: every buffer it names (address and size)
// and the settings baked into its kernels' arguments.
uint64_t Trainer::capture_key(uint32_t n) const {
  uint64_t h = 1469598103934665603ull;
  auto mix = [&h](uint64_t x) {
    for (int k = 0; k < 8; ++k) {
      h ^= (x >> (8 * k)) & 0xFFu;
      h *= 1099511628211ull;
    }
  };
  auto array = [&mix](const DeviceArray &a) {
    mix(reinterpret_cast<uintptr_t>(a.as<void>()));
    mix(a.bytes());
  };
  for (const DeviceArray *a : {&positions_, &scales_, &rotations_, &opacities_, &shWordsBuf_, &adamState_, &shPass_, &stats_, &filter3d_,
                               &viewStats_})
    array(*a);
  for (int g = 0; g < 5; ++g) {
    array(state_[g]);
    array(grad_[g]);
  }
  mix(n);
  for (uint32_t k = 0; k < n && k < views_.size(); ++k) {
    const ViewBuffers &b = *views_[k];
    for (const DeviceArray *a : {&b.splats, &b.keys, &b.values, &b.tmpKeys, &b.tmpValues, &b.coverage, &b.tileCount, &b.tileRecords,
                                 &b.bigList, &b.bigBits, &b.rawFixed, &b.rawFloat, &b.entryKeys, &b.entryValues, &b.tmpEntryKeys, &b.tmpEntryValues, &b.image,
                                 &b.stops, &b.adjoint, &b.fields, &b.ranges, &b.tileTotals, &b.listOrder, &b.blockWork, &b.blockOrder, &b.sortScratch,
                                 &b.scanScratch, &b.counters, &b.sums})
      array(*a);
    mix(b.entryCapacity);
    mix(static_cast<uint64_t>(b.width) << 32 | static_cast<uint32_t>(b.height));
    mix(static_cast<uint64_t>(b.tileW) << 32 | b.tileH);
  }
  mix(count_);
  mix(static_cast<uint64_t>(activeShDegree_));
  mix(filterSet_ ? 1u : 0u);
  uint32_t dropout;
  std::memcpy(&dropout, &shDropout_, 4);
  mix(dropout);
  mix(static_cast<uint64_t>(shRetained_) << 32 | shSeed_);
  // Settings the batch refuses or bakes in: an exposure set since the capture
  // would be ignored by a replay, and second order changes the step.
  mix((exposing_ ? 1u : 0u) | (secondOrder_ ? 2u : 0u));
  return h;
}

void Trainer::release_captured() {
  if (graph_ != nullptr) {
    (void)hipStreamSynchronize(stream_);
    (void)hipGraphExecDestroy(graph_);
  }
  graph_ = nullptr;
  graphViewSizes_.clear();
}

// Captures one batch (enqueue_views) as a HIP graph on the trainer's stream.
// Everything is allocated first: nothing may be allocated while a stream is
// being captured. Replays run the batch exactly as captured: the same cameras,
// targets and rates, the tile-entry reservation of the moment; Adam's step
// count and bias corrections come from the device, so each replay is the next
// step.
void Trainer::capture_views(uint32_t n, const vsplat_camera *cameras, const uint8_t *const *targets, uint32_t mode,
                            const vsplat_learning_rates *rates) {
  release_captured();
  if (options_.tile_shape == VSPLAT_TILES_AUTO && n > 0 && cameras != nullptr && targets != nullptr &&
      sizes_key(n, cameras) != batchTunedKey_) {
    autotune_views(n, cameras, targets, 2, nullptr, nullptr);
  }
  prepare_views(n, cameras, targets, mode, rates);
  check(hipStreamSynchronize(stream_), "before capture");
  check(hipStreamBeginCapture(stream_, hipStreamCaptureModeRelaxed), "hipStreamBeginCapture");
  hipGraph_t graph = nullptr;
  try {
    enqueue_views(n, cameras, targets, mode, rates);
  } catch (...) {
    (void)hipStreamEndCapture(stream_, &graph);
    if (graph != nullptr) (void)hipGraphDestroy(graph);
    queuedViewSizes_.clear();
    throw;
  }
  check(hipStreamEndCapture(stream_, &graph), "hipStreamEndCapture");
  hipGraphExec_t exec = nullptr;
  const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  check(e, "hipGraphInstantiate");
  graph_ = exec;
  graphViews_ = n;
  graphViewSizes_ = queuedViewSizes_;
  graphKey_ = capture_key(n);
  graphLastView_ = lastView_;
  queuedViewSizes_.clear();
  lastView_.width = 0; // captured, not run
}

void Trainer::launch_captured(uint32_t replays) {
  if (graph_ == nullptr) throw Error(VSPLAT_LOGIC_ERROR, "no batch is captured");
  if (capture_key(graphViews_) != graphKey_) {
    throw Error(VSPLAT_LOGIC_ERROR, "the trainer's buffers or settings changed since the batch was captured: capture it again");
  }
  for (uint32_t r = 0; r < replays; ++r) check(hipGraphLaunch(graph_, stream_), "hipGraphLaunch");
  if (replays > 0) {
    queuedViewSizes_ = graphViewSizes_;
    lastView_ = graphLastView_;
    lastViewIndex_ = graphViews_ - 1;
  }
}

uint32_t Trainer::synchronize_views(vsplat_step_stats *stats, uint32_t capacity, bool &overflow) {
  (void)read_stats(overflow);
  drain_timing();
  const uint32_t n = static_cast<uint32_t>(queuedViewSizes_.size());
  std::vector<uint8_t> host(static_cast<size_t>(n) * kViewStatsBytes);
  if (n > 0) check(hipMemcpy(host.data(), viewStats_.as<uint8_t>(), host.size(), hipMemcpyDeviceToHost), "read view stats");
  for (uint32_t k = 0; k < n && k < capacity && stats != nullptr; ++k) {
    uint32_t counters[kCounterWords];
    double sums[2];
    std::memcpy(counters, host.data() + static_cast<size_t>(k) * kViewStatsBytes, sizeof(counters));
    std::memcpy(sums, host.data() + static_cast<size_t>(k) * kViewStatsBytes + kCounterWords * 4, sizeof(sums));
    stats[k] = make_stats(counters, sums, queuedViewSizes_[k].first, queuedViewSizes_[k].second);
  }
  if (overflow) clear_overflow();
  return n;
}

uint32_t Trainer::step_views(uint32_t n, const vsplat_camera *cameras, const uint8_t *const *targets, const size_t *bytes, uint32_t mode,
                             const vsplat_learning_rates *rates, vsplat_step_stats *stats) {
  if (n == 0 || cameras == nullptr || targets == nullptr || bytes == nullptr) {
    throw Error(VSPLAT_INVALID_ARGUMENT, "a batch needs views and targets");
  }
  while (viewTargets_.size() < n) viewTargets_.push_back(std::make_unique<DeviceArray>());
  std::vector<const uint8_t *> device(n);
  for (uint32_t k = 0; k < n; ++k) {
    const size_t expected = static_cast<size_t>(cameras[k].width) * static_cast<size_t>(cameras[k].height) * 4;
    if (cameras[k].width <= 0 || cameras[k].height <= 0 || targets[k] == nullptr || bytes[k] != expected) {
      throw Error(VSPLAT_RUNTIME_ERROR, "a target image does not match its camera");
    }
    viewTargets_[k]->ensure(bytes[k]);
    check(hipMemcpyAsync(viewTargets_[k]->as<uint8_t>(), targets[k], bytes[k], hipMemcpyHostToDevice, stream_), "upload target");
    device[k] = viewTargets_[k]->as<uint8_t>();
  }
  if (options_.tile_shape == VSPLAT_TILES_AUTO && sizes_key(n, cameras) != batchTunedKey_) {
    autotune_views(n, cameras, device.data(), 2, nullptr, nullptr);
  }
  std::vector<vsplat_step_stats> all(n);
  for (int attempt = 0;; ++attempt) {
    enqueue_views(n, cameras, device.data(), mode, rates);
    bool overflow = false;
    synchronize_views(all.data(), n, overflow);
    if (!overflow) break;
    bool grew = false;
    for (uint32_t k = 0; k < n; ++k) {
      grew = grew || all[k].tile_entries > views_[k]->entryCapacity;
      grow_entries(*views_[k], all[k].tile_entries);
    }
    if (attempt > 0 && !grew) throw Error(VSPLAT_DEVICE_ERROR, "tile entries overflowed twice");
  }
  if (stats != nullptr) std::copy(all.begin(), all.end(), stats);
  return n;
}

void Trainer::gradients(float *positions, float *scales, float *rotations, float *opacities, float *sh) const {
  if (!haveGradients_) throw Error(VSPLAT_LOGIC_ERROR, "no gradients: compute_gradients() stores them");
  const size_t n = count_;
  float *out[5] = {positions, scales, rotations, opacities, sh};
  const size_t floats[5] = {n * 3, n * 3, n * 4, n, n * shFloats_};
  for (int g = 0; g < 5; ++g) {
    if (out[g] != nullptr) check(hipMemcpy(out[g], grad_[g].as<float>(), floats[g] * 4, hipMemcpyDeviceToHost), "download gradients");
  }
}

void Trainer::cloud(float *positions, float *scales, float *rotations, float *opacities, float *sh) const {
  const size_t n = count_;
  check(hipStreamSynchronize(stream_), "cloud");
  if (positions) check(hipMemcpy(positions, positions_.as<float>(), n * 12, hipMemcpyDeviceToHost), "download positions");
  if (scales) check(hipMemcpy(scales, scales_.as<float>(), n * 12, hipMemcpyDeviceToHost), "download scales");
  if (rotations) check(hipMemcpy(rotations, rotations_.as<float>(), n * 16, hipMemcpyDeviceToHost), "download rotations");
  if (opacities) check(hipMemcpy(opacities, opacities_.as<float>(), n * 4, hipMemcpyDeviceToHost), "download opacities");
  if (sh) {
    std::vector<uint32_t> words(n * shWords_);
    check(hipMemcpy(words.data(), shWordsBuf_.as<uint32_t>(), words.size() * 4, hipMemcpyDeviceToHost), "download SH");
    for (size_t i = 0; i < n; ++i) {
      const uint32_t *in = words.data() + i * shWords_;
      float *to = sh + i * shFloats_;
      if (shFormat_ == kShBf16) {
        for (uint32_t j = 0; j < shFloats_; ++j) to[j] = bf16_to_float((in[j / 2] >> (16 * (j & 1))) & 0xFFFFu);
      } else {
        for (uint32_t j = 0; j < 3; ++j) std::memcpy(&to[j], &in[j], 4);
        for (uint32_t j = 3; j < shFloats_; ++j) to[j] = half_to_float((in[3 + (j - 3) / 2] >> (16 * ((j - 3) & 1))) & 0xFFFFu);
      }
    }
  }
}

void Trainer::densify_stats(float *out) const {
  check(hipStreamSynchronize(stream_), "densify stats");
  check(hipMemcpy(out, stats_.as<float>(), static_cast<size_t>(count_) * 32, hipMemcpyDeviceToHost), "download stats");
}

// Trainer::apply_changes (trainer.cpp:749-787): the Gaussians replaced by a
// changed cloud of `count`. New Gaussian g keeps the optimiser state of old
// Gaussian source[g] (gathered on the GPU), or starts fresh when source[g] < 0.
// Its values come from `cloud`, or, when cloud is null, from old Gaussian
// source[g] as well (every source >= 0 then), so that pruning and copying need
// no upload. Everything is allocated before anything changes, so a failure
// leaves the trainer as it was. The densification statistics, the 3D filter,
// the stored gradients and a captured batch are dropped, as the reference's
// allocate_per_gaussian clears them; Adam's step count is kept.
void Trainer::apply_changes(const vsplat_cloud *cloud, const int64_t *source, uint32_t count) {
  if (count == 0) throw Error(VSPLAT_RUNTIME_ERROR, "the changes remove every Gaussian");
  if (source == nullptr) throw Error(VSPLAT_INVALID_ARGUMENT, "no source table");
  if (cloud != nullptr) {
    if (cloud->count != count) throw Error(VSPLAT_INVALID_ARGUMENT, "the cloud's count differs from the source table's");
    if (cloud->sh_degree != shDegree_) throw Error(VSPLAT_INVALID_ARGUMENT, "the cloud's SH degree differs from the trainer's");
    if (cloud->positions == nullptr || cloud->scales == nullptr || cloud->rotations == nullptr || cloud->opacities == nullptr ||
        cloud->sh == nullptr) {
      throw Error(VSPLAT_INVALID_ARGUMENT, "the cloud has a null array");
    }
  }
  std::vector<int32_t> table(count);
  for (uint32_t g = 0; g < count; ++g) {
    const int64_t from = source[g];
    if (from >= static_cast<int64_t>(count_)) throw Error(VSPLAT_INVALID_ARGUMENT, "a source index is past the old cloud");
    if (from < 0 && cloud == nullptr) throw Error(VSPLAT_INVALID_ARGUMENT, "a new Gaussian (source < 0) needs values: pass the cloud");
    table[g] = from < 0 ? -1 : static_cast<int32_t>(from);
  }
  release_captured();
  for (auto &v : views_) check(hipStreamSynchronize(v->stream), "queued work");

  const size_t n = count;
  const uint32_t valueWords[5] = {3, 3, 4, 1, shWords_}, stateWords[5] = {3, 3, 4, 1, shFloats_};
  DeviceArray sourceTable, values[5], state[5];
  sourceTable.ensure(n * 4);
  for (int g = 0; g < 5; ++g) {
    state[g].ensure(n * stateWords[g] * 4);
    if (cloud == nullptr) values[g].ensure(n * valueWords[g] * 4);
  }
  if (cloud != nullptr) {
    DeviceArray *to[5] = {&values[0], &values[1], &values[2], &values[3], &values[4]};
    upload_cloud(*cloud, to);
  }
  check(hipMemcpy(sourceTable.as<int32_t>(), table.data(), n * 4, hipMemcpyHostToDevice), "upload source table");
  DeviceArray *old[5] = {&positions_, &scales_, &rotations_, &opacities_, &shWordsBuf_};
  GatherArray arrays[kMaxGatherArrays];
  uint32_t k = 0;
  for (int g = 0; g < 5; ++g) arrays[k++] = {state_[g].as<uint32_t>(), state[g].as<uint32_t>(), stateWords[g]};
  if (cloud == nullptr)
    for (int g = 0; g < 5; ++g) arrays[k++] = {old[g]->as<uint32_t>(), values[g].as<uint32_t>(), valueWords[g]};
  launch_gather(stream_, sourceTable.as<int32_t>(), count, arrays, k);
  check(hipGetLastError(), "gather launch");
  check(hipStreamSynchronize(stream_), "gather");

  for (int g = 0; g < 5; ++g) {
    old[g]->swap(values[g]);
    state_[g].swap(state[g]);
    grad_[g].release();
  }
  count_ = count;
  haveGradients_ = false;
  accumulateNext_ = false;
  lastView_.width = 0;
  release_views();
  allocate_per_gaussian();
  view_buffers(0);
  check(hipStreamSynchronize(stream_), "per-Gaussian buffers");
}

// Trainer::enable_second_order (trainer.cpp:642-654): from now on the states'
// second halves hold the curvature, from zero (rotations keep Adam's moments).
void Trainer::enable_second_order() {
  if (secondOrder_) return;
  release_captured();
  const size_t n = count_;
  const size_t floats[5] = {n * 3, n * 3, n * 4, n, n * shFloats_};
  for (int g = 0; g < 5; ++g) {
    if (g == 2) continue; // rotation stays with Adam
    launch_keep_first_moment(stream_, state_[g].as<uint32_t>(), floats[g]);
  }
  check(hipStreamSynchronize(stream_), "enable second order");
  secondOrder_ = true;
  curvatureUpdates_ = 0;
}

// Trainer::update_curvature (trainer.cpp:656-667): a Hutchinson adjoint through
// the backward pass of the last view's lists and image, then the chain rule
// adding each parameter's squared gradient to its curvature.
void Trainer::update_curvature() {
  // The last view that ran, on the buffers it ran on (a batch's last view has
  // its own); nothing after an autotune or a capture, which leave the buffers
  // with passes of other shapes or none.
  if (lastView_.width == 0 || lastViewIndex_ >= views_.size()) {
    throw Error(VSPLAT_LOGIC_ERROR, "update_curvature() needs a view: step() or compute_gradients() first");
  }
  ViewBuffers &b = *views_[lastViewIndex_];
  if (b.width != static_cast<int>(lastView_.width) || b.height != static_cast<int>(lastView_.height)) {
    throw Error(VSPLAT_LOGIC_ERROR, "update_curvature(): the last view's buffers were remade; run a step first");
  }
  ++curvatureUpdates_;
  const ViewParams &v = lastView_;
  const uint32_t pixels = v.width * v.height;
  uint32_t *counters = b.counters.as<uint32_t>();
  launch_hutchinson(stream_, b.adjoint.as<float4>(), pixels, curvatureUpdates_);
  if (fixedPoint_) {
    launch_fill_u32(stream_, b.rawFixed.as<uint32_t>(), static_cast<uint64_t>(count_) * kRawPerSplat * 2, 0u);
  } else {
    launch_fill_u32(stream_, b.rawFloat.as<uint32_t>(), static_cast<uint64_t>(count_) * kRawPerSplat, 0u);
  }
  launch_backward(stream_, fixedPoint_, options_.recompute_transmittance != 0, b.splats.as<Splat>(), b.entryKeys.as<uint32_t>(),
                  b.entryValues.as<uint32_t>(), b.tileBits, b.ranges.as<uint2>(), v.width, v.height, b.tileW, b.tileH, options_.background,
                  b.image.as<float4>(),
                  b.adjoint.as<float4>(), b.stops.as<uint32_t>(), curvature_fixed_scale(), b.rawFixed.as<int64_t>(),
                  b.rawFloat.as<float>());
  const bool statistics = statisticsNext_;
  statisticsNext_ = false; // the curvature adds nothing to the densification statistics
  enqueue_chain(b, v, kModeCurvature, nullptr, stream_, counters);
  statisticsNext_ = statistics;
  check(hipStreamSynchronize(stream_), "update curvature");
}

void Trainer::reset_densify_stats() {
  launch_fill_u32(stream_, stats_.as<uint32_t>(), static_cast<uint64_t>(count_) * 8, 0u);
  check(hipStreamSynchronize(stream_), "reset stats");
}

void Trainer::set_active_sh_degree(int degree) { activeShDegree_ = std::clamp(degree, 0, shDegree_); }

void Trainer::set_filter3d(const float *filter, size_t count) {
  if (filter == nullptr || count != count_) throw Error(VSPLAT_INVALID_ARGUMENT, "the filter must have one value per Gaussian");
  check(hipMemcpy(filter3d_.as<float>(), filter, count * 4, hipMemcpyHostToDevice), "upload filter");
  filterSet_ = true;
}

void Trainer::set_sh_dropout(float probability, int retained, uint32_t seed) {
  shDropout_ = std::clamp(probability, 0.0f, 1.0f);
  shRetained_ = static_cast<uint32_t>(std::clamp(retained, 0, 3));
  shSeed_ = seed;
}

void Trainer::parameters(uint32_t group, void **pointer, uint64_t *bytes) const {
  const DeviceArray *g[5] = {&positions_, &scales_, &rotations_, &opacities_, &shWordsBuf_};
  if (group >= 5) throw Error(VSPLAT_INVALID_ARGUMENT, "parameter groups are 0 to 4");
  const size_t n = count_;
  const size_t size[5] = {n * 12, n * 12, n * 16, n * 4, n * shWords_ * 4};
  if (pointer) *pointer = g[group]->as<void>();
  if (bytes) *bytes = size[group];
}

uint64_t Trainer::bytes_per_gaussian() const {
  const uint64_t floats = 3 + 3 + 4 + 1 + shFloats_;
  // Values (SH as packed words) and one optimiser word per parameter.
  const uint64_t parameters = (floats - shFloats_ + shWords_) * 4 + floats * 4;
  // Splat, sort keys and values with copies, coverage, tile count, filter, big
  // splat list, tile record, sums, SH-pass data, the direction term between the
  // chain kernels, statistics.
  const uint64_t raw = fixedPoint_ ? kRawPerSplat * 8 : kRawPerSplat * 4;
  return parameters + sizeof(Splat) + 4 * 8 + sizeof(TileRecord) + raw + kShPassFloats * 4 + 16 + 32;
}

uint64_t Trainer::gpu_bytes() const {
  // One view's per-Gaussian share is in bytes_per_gaussian; the others' are added as allocated.
  uint64_t total = bytes_per_gaussian() * count_ + target_.bytes() + viewStats_.bytes();
  for (size_t k = 0; k < views_.size(); ++k) {
    const ViewBuffers &b = *views_[k];
    for (const DeviceArray *a : {&b.image, &b.stops, &b.adjoint, &b.fields, &b.ranges, &b.tileTotals, &b.listOrder, &b.blockWork,
                                 &b.blockOrder, &b.entryKeys, &b.entryValues, &b.tmpEntryKeys, &b.tmpEntryValues, &b.bigBits, &b.exposed,
                                 &b.edge, &b.exposureTotals, &b.errorMax}) {
      total += a->bytes();
    }
    if (k == 0) continue;
    for (const DeviceArray *a : {&b.splats, &b.keys, &b.values, &b.tmpKeys, &b.tmpValues, &b.coverage, &b.tileCount, &b.tileRecords,
                                 &b.bigList, &b.rawFixed, &b.rawFloat, &b.sortScratch, &b.scanScratch}) {
      total += a->bytes();
    }
  }
  for (const DeviceArray &a : grad_) total += a.bytes();
  total += poseTerms_.bytes() + poseTotals_.bytes();
  return total;
}

} // namespace vsplat

namespace vsplat {

void Trainer::inspect(const std::string &name, void *out, uint64_t capacity, uint64_t *bytes, uint32_t view) const {
  if (view >= views_.size()) throw Error(VSPLAT_INVALID_ARGUMENT, "no such view: a batch makes one per view");
  const ViewBuffers &b = *views_[view];
  const uint64_t n = count_, pixels = static_cast<uint64_t>(b.width) * static_cast<uint64_t>(b.height);
  uint64_t entries = 0;
  if (name == "entry_keys" || name == "entry_values" || name == "entry_blocks") {
    uint32_t counters[kCounterWords];
    check(hipStreamSynchronize(stream_), "inspect");
    check(hipMemcpy(counters, b.counters.as<uint32_t>(), sizeof(counters), hipMemcpyDeviceToHost), "inspect counters");
    entries = counters[kEntriesUsed];
  }
  struct Item {
    const char *name;
    const DeviceArray *array;
    uint64_t size;
  };
  const Item items[] = {
      {"splats", &b.splats, n * sizeof(Splat)},
      {"keys", &b.keys, n * 4},
      {"values", &b.values, n * 4},
      {"coverage", &b.coverage, n * 4},
      {"entry_keys", &b.entryKeys, entries * 4},   // each entry's tile
      {"entry_blocks", &b.entryKeys, entries * 4}, // the tile's 8x8 blocks (bit q: block q) the entry's splat reaches
      {"entry_values", &b.entryValues, entries * 4},
      {"ranges", &b.ranges, static_cast<uint64_t>(b.tiles) * 8},
      {"image", &b.image, pixels * 16},
      {"stops", &b.stops, pixels * 4},
      {"adjoint", &b.adjoint, pixels * 16},
      {"fields", &b.fields, pixels * 36},
      {"tile_totals", &b.tileTotals, static_cast<uint64_t>(b.lossTiles) * 8},
      {"block_work", &b.blockWork, static_cast<uint64_t>(raster_blocks(b.tiles, b.tileW, b.tileH)) * 4},
      {"block_order", &b.blockOrder, static_cast<uint64_t>(raster_blocks(b.tiles, b.tileW, b.tileH)) * 4},
      {"list_order", &b.listOrder, static_cast<uint64_t>(raster_blocks(b.tiles, b.tileW, b.tileH)) * 4},
      {"raw", fixedPoint_ ? &b.rawFixed : &b.rawFloat, n * kRawPerSplat * (fixedPoint_ ? 8u : 4u)},
      {"counters", &b.counters, kCounterWords * 4},
      {"state_positions", &state_[0], n * 12},
      {"state_scales", &state_[1], n * 12},
      {"state_rotations", &state_[2], n * 16},
      {"state_opacities", &state_[3], n * 4},
      {"state_sh", &state_[4], n * shFloats_ * 4},
      {"stats", &stats_, n * 32},
  };
  if (name == "stops" && options_.recompute_transmittance != 0) {
    throw Error(VSPLAT_LOGIC_ERROR, "no stops are kept: the backward recomputes them");
  }
  for (const Item &item : items) {
    if (name != item.name) continue;
    if (bytes != nullptr) *bytes = item.size;
    const uint64_t copy = std::min(item.size, capacity);
    if (out != nullptr && copy > 0) {
      check(hipStreamSynchronize(stream_), "inspect");
      check(hipMemcpy(out, item.array->as<void>(), copy, hipMemcpyDeviceToHost), "inspect copy");
      // The keys hold both: the tile in the low tileBits bits, the blocks above.
      uint32_t *words = static_cast<uint32_t *>(out);
      if (name == "entry_keys") {
        for (uint64_t k = 0; k < copy / 4; ++k) words[k] &= (1u << b.tileBits) - 1u;
      } else if (name == "entry_blocks") {
        for (uint64_t k = 0; k < copy / 4; ++k) words[k] >>= b.tileBits;
      }
    }
    return;
  }
  throw Error(VSPLAT_INVALID_ARGUMENT, "no buffer named " + name);
}

} // namespace vsplat
