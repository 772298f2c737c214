// The trainer's host side: device buffers, the order of passes, and the
// reads back of counts and loss totals (the reference's Trainer,
// src/train/trainer.cpp:157-863, re-laid out for HIP).
#pragma once

#include <hip/hip_runtime.h>

#include <future>
#include <map>
#include <memory>
#include <cstdint>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <utility>
#include <vector>

#include "../../include/vsplat.h"
#include "kernels.hpp"

namespace vsplat {

struct Error : std::runtime_error {
  int status;
  Error(int s, const std::string &what) : std::runtime_error(what), status(s) {}
};

// The allocation-free render's scratch on the current GPU (noalloc.hip): its
// device address, looked up once per device, and its size in bytes.
void *noalloc_scratch(uint64_t *bytes);

// One hipMalloc'd array; grows by reallocation, contents are not kept. A
// borrowed array points into memory it does not own (never freed, never grown:
// ensure() beyond its size throws).
class DeviceArray {
public:
  DeviceArray() = default;
  DeviceArray(const DeviceArray &) = delete;
  DeviceArray &operator=(const DeviceArray &) = delete;
  ~DeviceArray() { release(); }
  void ensure(size_t bytes);
  void release();
  void borrow(void *ptr, size_t bytes) {
    release();
    ptr_ = ptr;
    bytes_ = bytes;
    owned_ = false;
  }
  template <class T> T *as() const { return static_cast<T *>(ptr_); }
  size_t bytes() const { return bytes_; }
  void swap(DeviceArray &other) {
    std::swap(ptr_, other.ptr_);
    std::swap(bytes_, other.bytes_);
    std::swap(owned_, other.owned_);
  }

private:
  void *ptr_ = nullptr;
  size_t bytes_ = 0;
  bool owned_ = true;
};

class PinnedArray {
public:
  PinnedArray() = default;
  PinnedArray(const PinnedArray &) = delete;
  PinnedArray &operator=(const PinnedArray &) = delete;
  ~PinnedArray() { release(); }
  void ensure(size_t bytes);
  void release();
  template <class T> T *as() const { return static_cast<T *>(ptr_); }
  size_t bytes() const { return bytes_; }
  void swap(PinnedArray &other) {
    std::swap(ptr_, other.ptr_);
    std::swap(bytes_, other.bytes_);
  }

private:
  void *ptr_ = nullptr;
  size_t bytes_ = 0;
};

// Everything one view's passes write: the per-Gaussian, per-pixel and
// per-entry buffers, the counters, and the stream the passes run on. The
// single-view calls use the first; a batch uses one per view, so that the
// views' passes can run side by side.
struct ViewBuffers {
  ViewBuffers() = default;
  ViewBuffers(const ViewBuffers &) = delete;
  ViewBuffers &operator=(const ViewBuffers &) = delete;
  hipStream_t stream = nullptr;
  bool ownsStream = false;
  hipEvent_t done = nullptr; // recorded on `stream` after the view's backward pass
  int width = 0, height = 0;
  uint32_t tileW = 16, tileH = 16; // the raster tile shape
  uint32_t gridW = 0, gridH = 0, tiles = 0, tileBits = 0;
  uint32_t lossTiles = 0;          // the loss's 16x16 tiles
  uint64_t entryCapacity = 0;
  // Per Gaussian.
  DeviceArray splats, keys, values, tmpKeys, tmpValues, coverage, tileCount, tileRecords, bigList, bigBits, rawFixed, rawFloat;
  // Per tile entry.
  DeviceArray entryKeys, entryValues, tmpEntryKeys, tmpEntryValues;
  // Per pixel and per tile; per 8x8 block (raster launch index), the backward's
  // work found by the forward and the backward's launch order.
  DeviceArray image, stops, adjoint, fields, ranges, tileTotals, listOrder, blockWork, blockOrder;
  // Exposure and scoring (used only when set): the exposed image, the target's
  // edge weights, the exposure backward's workgroup totals, the worst error.
  DeviceArray exposed, edge, exposureTotals, errorMax;
  // Scratch and counters.
  DeviceArray sortScratch, scanScratch, counters, sums;
};

class Trainer {
public:
  Trainer(const vsplat_cloud &cloud, const vsplat_trainer_options &options);
  ~Trainer();
  Trainer(const Trainer &) = delete;
  Trainer &operator=(const Trainer &) = delete;

  void render(const vsplat_camera &camera, float *out);
  // render() without allocating: the view's lists and per-pixel state in
  // `scratch` (scratchBytes, 256-byte aligned device memory), the image written
  // to deviceOut (capacity bytes). Sizes past the limits in include/vsplat.h
  // are refused before anything runs.
  void render_into(const vsplat_camera &camera, float *deviceOut, uint64_t capacity, void *scratch, uint64_t scratchBytes);
  vsplat_step_stats run(const vsplat_camera &camera, const uint8_t *target, size_t bytes, uint32_t mode,
                        const vsplat_learning_rates *rates, const uint8_t *next, size_t nextBytes);
  void apply_gradients(const vsplat_learning_rates &rates);
  vsplat_step_stats score_view(const vsplat_camera &camera, const uint8_t *target, size_t bytes, float errorShare, float pixelWeight);
  void set_exposure(const float transform[12]);
  void prefetch_target(const uint8_t *target, size_t bytes);

  void enqueue(const vsplat_camera &camera, const uint8_t *deviceTarget, uint32_t mode, const vsplat_learning_rates *rates);
  vsplat_step_stats synchronize();
  // The batched step: n views' gradients summed, then (mode step) one Adam step.
  void enqueue_views(uint32_t n, const vsplat_camera *cameras, const uint8_t *const *targets, uint32_t mode,
                     const vsplat_learning_rates *rates);
  // Waits; each queued view's stats into stats[0 .. min(n, capacity)); returns n.
  uint32_t synchronize_views(vsplat_step_stats *stats, uint32_t capacity, bool &overflow);
  uint32_t step_views(uint32_t n, const vsplat_camera *cameras, const uint8_t *const *targets, const size_t *bytes, uint32_t mode,
                      const vsplat_learning_rates *rates, vsplat_step_stats *stats);
  void reserve_entries(uint64_t entries);
  // Times each tile shape on this view and keeps the fastest for its size.
  std::pair<uint32_t, uint32_t> autotune_tiles(const vsplat_camera &camera, const uint8_t *deviceTarget, uint32_t repeats);
  // Tunes the shapes of a batch's view sizes together, timing the whole batch;
  // each view's chosen shape into tileW[k], tileH[k] (either may be null).
  void autotune_views(uint32_t n, const vsplat_camera *cameras, const uint8_t *const *targets, uint32_t repeats, uint32_t *tileW,
                      uint32_t *tileH);
  // A batch captured as a HIP graph, replayed with launch_captured and waited
  // for with synchronize_views.
  void capture_views(uint32_t n, const vsplat_camera *cameras, const uint8_t *const *targets, uint32_t mode,
                     const vsplat_learning_rates *rates);
  void launch_captured(uint32_t replays);
  void release_captured();

  void gradients(float *positions, float *scales, float *rotations, float *opacities, float *sh) const;
  void cloud(float *positions, float *scales, float *rotations, float *opacities, float *sh) const;
  void densify_stats(float *out) const;
  void reset_densify_stats();
  void apply_changes(const vsplat_cloud *cloud, const int64_t *source, uint32_t count);
  void enable_second_order();
  void update_curvature();
  void reset_opacity(float maximum);
  void add_position_noise(float scaler, uint32_t iteration);
  bool second_order() const { return secondOrder_; }
  void set_active_sh_degree(int degree);
  void set_filter3d(const float *filter, size_t count);
  void set_sh_dropout(float probability, int retained, uint32_t seed);
  void parameters(uint32_t group, void **pointer, uint64_t *bytes) const;
  std::vector<std::pair<std::string, double>> timings() const;
  void reset_timings();
  void set_queued_timing(bool on) { timeQueued_ = on; }
  // One of view `view`'s buffers (a batch's k-th view; 0 for the single-view calls), or a shared one.
  void inspect(const std::string &buffer, void *out, uint64_t capacity, uint64_t *bytes, uint32_t view = 0) const;

  uint32_t count() const { return count_; }
  int sh_degree() const { return shDegree_; }
  int active_sh_degree() const { return activeShDegree_; }
  bool fixed_point() const { return fixedPoint_; }
  uint64_t gpu_bytes() const;
  uint64_t bytes_per_gaussian() const;
  uint64_t staging_bytes() const { return staging_.bytes(); }
  hipStream_t stream() const { return stream_; }

private:
  ViewParams view_params(const vsplat_camera &camera) const;
  ViewBuffers &view_buffers(size_t k); // the k-th view's buffers, made on first use
  void release_views();
  void enqueue_view_passes(uint32_t n, const vsplat_camera *cameras, const uint8_t *const *targets, ViewParams *params);
  void enqueue_view_chains(uint32_t n, const ViewParams *params, uint32_t mode, const vsplat_learning_rates *rates);
  static uint64_t sizes_key(uint32_t n, const vsplat_camera *cameras);
  void upload_cloud(const vsplat_cloud &cloud, DeviceArray *const values[5]) const;
  void allocate_per_gaussian();
  void ensure_size(ViewBuffers &b, int width, int height);
  void gather_view_status(uint32_t n);
  void ensure_layout(ViewBuffers &b, int width, int height, uint32_t tileW, uint32_t tileH);
  static void set_geometry(ViewBuffers &b, int width, int height, uint32_t tileW, uint32_t tileH);
  std::pair<uint32_t, uint32_t> tile_shape_for(int width, int height) const;
  bool needs_tuning(int width, int height) const;
  void grow_entries(ViewBuffers &b, uint64_t entries);
  void fit_entries(ViewBuffers &b, uint64_t entries);
  uint64_t big_bit_words_for(uint64_t entries) const;
  void stage_target(const uint8_t *target, size_t bytes);
  // The passes up to the backward pass on b's stream, then (chain) the chain rule.
  void enqueue_passes(ViewBuffers &b, const ViewParams &view, const uint8_t *deviceTarget, uint32_t mode,
                      const vsplat_learning_rates *rates, bool timed, bool chain = true);
  ViewParams with_tiles(const ViewBuffers &b, const ViewParams &view) const;
  void enqueue_lists(ViewBuffers &b, const ViewParams &view, uint32_t mode, bool marks);
  void enqueue_loss(ViewBuffers &b, const ViewParams &view, const uint8_t *deviceTarget);
  // A batch's passes with one forward and one backward launch over all views (VSPLAT_BATCH_ONE_LAUNCH).
  void enqueue_view_passes_one_launch(uint32_t n, const vsplat_camera *cameras, const uint8_t *const *targets, ViewParams *params);
  void enqueue_batch_passes(uint32_t n, const vsplat_camera *cameras, const uint8_t *const *targets, ViewParams *params);
  void launch_batch_raster(uint32_t n, const ViewParams *params, bool backward);
  void enqueue_chain(const ViewBuffers &b, const ViewParams &view, uint32_t mode, const vsplat_learning_rates *rates,
                     hipStream_t stream, uint32_t *counters);
  void enqueue_noise(const ViewBuffers &b, const vsplat_learning_rates *rates, hipStream_t stream, const uint32_t *counters);
  ChainParams chain_params(const ViewParams &view, uint32_t mode, const vsplat_learning_rates *rates) const;
  ChainBuffers chain_buffers(const ViewBuffers &b) const;
  void ensure_gradients();
  void prepare_views(uint32_t n, const vsplat_camera *cameras, const uint8_t *const *targets, uint32_t mode,
                     const vsplat_learning_rates *rates);
  uint64_t capture_key(uint32_t n) const;
  void clear_overflow();
  vsplat_step_stats make_stats(const uint32_t *counters, const double *sums, int width, int height) const;
  void begin_timing(bool iteration);
  void mark(const char *stage);
  void cancel_timing();
  void drain_timing();
  void wait_prefetch();
  vsplat_step_stats read_stats(bool &overflow);

  vsplat_trainer_options options_;
  uint32_t count_ = 0;
  int shDegree_ = 0, activeShDegree_ = 0;
  uint32_t shFloats_ = 3, shWords_ = 3;
  uint32_t shFormat_ = 0; // kShHalf or kShBf16
  bool fixedPoint_ = true;
  float shDropout_ = 0.0f, dropoutNow_ = 0.0f;
  Exposure exposure_{{1.0f, 0.0f, 0.0f, 0.0f, 0.0f, 1.0f, 0.0f, 0.0f, 0.0f, 0.0f, 1.0f, 0.0f}};
  bool exposing_ = false;     // exposure_ is not the neutral one
  float scoreShare_ = 0.0f;   // while score_view runs: the error share that marks a pixel
  bool secondOrder_ = false;  // enable_second_order() was called
  uint32_t curvatureUpdates_ = 0;
  uint32_t shRetained_ = 0, shSeed_ = 0;
  bool haveGradients_ = false;
  bool statisticsNext_ = true; // the next chain pass adds to the densification statistics
  bool poseNext_ = false;      // the next chain pass writes the pose terms (the synchronous calls, with pose_gradients)
  DeviceArray poseTerms_, poseTotals_; // 6 floats per Gaussian, and per group of 256
  bool accumulateNext_ = false; // the next chain pass (gradients mode) adds to the gradient sums
  // Per queued view of a batch: its counters and loss sums, copied on the device.
  static constexpr size_t kViewStatsBytes = kViewStatusBytes;
  std::vector<std::pair<int, int>> queuedViewSizes_;
  ViewParams lastView_{};   // the camera of the last view that ran (width 0: none)
  size_t lastViewIndex_ = 0; // whose buffers hold its lists, image and sums

  hipStream_t stream_ = nullptr;
  // Parameters and optimiser state, in GaussianCloud's layout; SH packed as halves.
  DeviceArray positions_, scales_, rotations_, opacities_, shWordsBuf_;
  DeviceArray state_[5], grad_[5];
  DeviceArray adamState_; // steps taken, beta1^steps, beta2^steps (ChainParams::adamState)
  // Per Gaussian, shared by the views.
  DeviceArray shPass_, stats_, filter3d_, dirTerms_;
  bool filterSet_ = false;
  // The views' buffers; the first is the single-view calls' and runs on stream_.
  std::vector<std::unique_ptr<ViewBuffers>> views_;
  hipEvent_t batchStart_ = nullptr, batchRaster_ = nullptr;
  // The captured batch: its executable graph, view count and sizes, and the
  // capture_key it was made under.
  hipGraphExec_t graph_ = nullptr;
  uint32_t graphViews_ = 0;
  std::vector<std::pair<int, int>> graphViewSizes_;
  uint64_t graphKey_ = 0;
  ViewParams graphLastView_{}; // the captured batch's last view
  uint64_t reservedEntries_ = 0; // what reserve_entries asked for, for views made later
  std::map<std::pair<int, int>, std::pair<uint32_t, uint32_t>> tuned_; // view size -> tile shape the autotuner chose
  uint64_t batchTunedKey_ = 0; // sizes_key of the batch autotune_views last tuned
  DeviceArray target_, viewStats_;
  std::vector<std::unique_ptr<DeviceArray>> viewTargets_; // step_views' targets, one per view
  PinnedArray staging_, next_, readback_;

  // The next step's target (the reference's prefetch, trainer.cpp:539-564):
  // one asynchronous host copy into pinned memory (next_) per prefetch, of the
  // caller's image at prefetched_, which stays borrowed until the next step.
  std::future<void> prefetchCopy_;
  const uint8_t *prefetched_ = nullptr;
  size_t prefetchedBytes_ = 0;

  // Stage timing with events (trainer.cpp:201-240): one set of events per
  // queued step, in a ring, read once the stream has passed them.
  static constexpr int kMaxMarks = 12;
  static constexpr size_t kTimingSlots = 64;
  struct TimingSlot {
    hipEvent_t events[kMaxMarks + 1] = {};
    std::vector<const char *> marks;
    bool iteration = true; // counts as a training iteration in the averages
  };
  std::vector<TimingSlot> slots_;
  size_t slotHead_ = 0, slotPending_ = 0;
  TimingSlot *current_ = nullptr;
  std::vector<std::pair<std::string, double>> stageTotals_;
  uint64_t timedIterations_ = 0;
  bool timeQueued_ = true; // the enqueue path records stage events too
};

} // namespace vsplat
