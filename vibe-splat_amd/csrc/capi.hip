// The C entry points of include/vsplat.h: argument checks, then the Trainer,
// with every exception turned into a status and a per-thread message.
#include "trainer.hpp"
#include "ply.hpp"

#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

struct vsplat_trainer {
  vsplat::Trainer impl;
  vsplat_trainer(const vsplat_cloud &c, const vsplat_trainer_options &o) : impl(c, o) {}
};

namespace {

thread_local std::string lastError;

template <class F> int guarded(F &&f) {
  try {
    f();
    return VSPLAT_OK;
  } catch (const vsplat::Error &e) {
    lastError = e.what();
    return e.status;
  } catch (const std::bad_alloc &) {
    lastError = "out of host memory";
    return VSPLAT_RUNTIME_ERROR;
  } catch (const std::exception &e) {
    lastError = e.what();
    return VSPLAT_RUNTIME_ERROR;
  } catch (...) {
    lastError = "unknown error";
    return VSPLAT_RUNTIME_ERROR;
  }
}

int invalid(const char *what) {
  lastError = what;
  return VSPLAT_INVALID_ARGUMENT;
}

} // namespace

extern "C" {

void vsplat_default_options(vsplat_trainer_options *o) {
  if (o == nullptr) return;
  std::memset(o, 0, sizeof(*o));
  o->lambda_ssim = 0.2f; // trainer.hpp:25-41
  o->near_plane = 0.01f;
  o->selective_adam = 1;
  o->memory_budget = ~uint64_t{0};
}

void vsplat_default_rates(vsplat_learning_rates *r) {
  if (r == nullptr) return;
  *r = vsplat_learning_rates{1.6e-4f, 0.005f, 0.001f, 0.05f, 0.025f, 0.05f, 0.0f}; // trainer.hpp:51-61
}

int vsplat_trainer_create(const vsplat_cloud *cloud, const vsplat_trainer_options *options, vsplat_trainer **out) {
  if (cloud == nullptr || out == nullptr) return invalid("null cloud or output");
  *out = nullptr;
  vsplat_trainer_options defaults;
  vsplat_default_options(&defaults);
  return guarded([&] { *out = new vsplat_trainer(*cloud, options != nullptr ? *options : defaults); });
}

void vsplat_trainer_destroy(vsplat_trainer *t) { delete t; }

int vsplat_render(vsplat_trainer *t, const vsplat_camera *camera, float *out) {
  if (t == nullptr || camera == nullptr || out == nullptr) return invalid("null argument");
  return guarded([&] { t->impl.render(*camera, out); });
}

int vsplat_render_noalloc(vsplat_trainer *t, const vsplat_camera *camera, float *device_rgba, uint64_t capacity_bytes) {
  if (t == nullptr || camera == nullptr || device_rgba == nullptr) return invalid("null argument");
  return guarded([&] {
    uint64_t bytes = 0;
    void *scratch = vsplat::noalloc_scratch(&bytes);
    t->impl.render_into(*camera, device_rgba, capacity_bytes, scratch, bytes);
  });
}

int vsplat_compute_gradients(vsplat_trainer *t, const vsplat_camera *camera, const uint8_t *target, size_t bytes, vsplat_step_stats *stats) {
  if (t == nullptr || camera == nullptr) return invalid("null argument");
  return guarded([&] {
    const vsplat_step_stats s = t->impl.run(*camera, target, bytes, vsplat::kModeGradients, nullptr, nullptr, 0);
    if (stats) *stats = s;
  });
}

int vsplat_collect_statistics(vsplat_trainer *t, const vsplat_camera *camera, const uint8_t *target, size_t bytes, vsplat_step_stats *stats) {
  if (t == nullptr || camera == nullptr) return invalid("null argument");
  return guarded([&] {
    const vsplat_step_stats s = t->impl.run(*camera, target, bytes, vsplat::kModeStatistics, nullptr, nullptr, 0);
    if (stats) *stats = s;
  });
}

int vsplat_score_view(vsplat_trainer *t, const vsplat_camera *camera, const uint8_t *target, size_t bytes, float error_share,
                      float pixel_weight, vsplat_step_stats *stats) {
  if (t == nullptr || camera == nullptr) return invalid("null argument");
  return guarded([&] {
    const vsplat_step_stats s = t->impl.score_view(*camera, target, bytes, error_share, pixel_weight);
    if (stats) *stats = s;
  });
}

// refined_pose (trainer.cpp:120-137): the rotation w by Rodrigues' formula
// (with the limits of its factors at small angles) turns the view, then v
// shifts it, in double; the view and position are rounded to float.
int vsplat_refined_pose(const vsplat_camera *camera, const float *move, vsplat_camera *out) {
  if (camera == nullptr || move == nullptr || out == nullptr) return invalid("null argument");
  const double w[3] = {move[0], move[1], move[2]}, v[3] = {move[3], move[4], move[5]};
  const double angle = std::sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
  const double a = angle > 1e-8 ? std::sin(angle) / angle : 1.0;
  const double b = angle > 1e-8 ? (1.0 - std::cos(angle)) / (angle * angle) : 0.5;
  // K is the cross-product matrix of w: K x = w x x.
  const double K[3][3] = {{0.0, -w[2], w[1]}, {w[2], 0.0, -w[0]}, {-w[1], w[0], 0.0}};
  double R[3][3];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      double kk = 0.0;
      for (int j = 0; j < 3; ++j) kk += K[r][j] * K[j][c];
      R[r][c] = (r == c ? 1.0 : 0.0) + a * K[r][c] + b * kk;
    }
  const float *V = camera->view; // column-major: V[c * 4 + r]
  double rotation[3][3], translation[3];
  for (int r = 0; r < 3; ++r) {
    for (int c = 0; c < 3; ++c) {
      double s = 0.0;
      for (int j = 0; j < 3; ++j) s += R[r][j] * static_cast<double>(V[c * 4 + j]);
      rotation[r][c] = s;
    }
    double s = 0.0;
    for (int j = 0; j < 3; ++j) s += R[r][j] * static_cast<double>(V[12 + j]);
    translation[r] = s + v[r];
  }
  vsplat_camera result = *camera;
  for (int c = 0; c < 4; ++c)
    for (int r = 0; r < 4; ++r) {
      double x = r == 3 ? (c == 3 ? 1.0 : 0.0) : (c == 3 ? translation[r] : rotation[r][c]);
      result.view[c * 4 + r] = static_cast<float>(x);
    }
  for (int c = 0; c < 3; ++c) {
    double s = 0.0;
    for (int r = 0; r < 3; ++r) s += rotation[r][c] * translation[r];
    result.position[c] = static_cast<float>(-s);
  }
  *out = result;
  return VSPLAT_OK;
}

int vsplat_ply_info(const char *path, uint64_t *count, int32_t *sh_degree) {
  if (path == nullptr) return invalid("null argument");
  return guarded([&] {
    const vsplat::PlyInfo info = vsplat::ply_info(path);
    if (count) *count = info.count;
    if (sh_degree) *sh_degree = info.shDegree;
  });
}

int vsplat_read_ply(const char *path, float *positions, float *scales, float *rotations, float *opacities, float *sh, uint64_t count) {
  if (path == nullptr || positions == nullptr || scales == nullptr || rotations == nullptr || opacities == nullptr || sh == nullptr) {
    return invalid("null argument");
  }
  return guarded([&] {
    if (vsplat::ply_info(path).count != count) throw vsplat::Error(VSPLAT_INVALID_ARGUMENT, "the arrays are not sized for the file's vertices");
    vsplat::read_ply(path, vsplat::PlyCloud{positions, scales, rotations, opacities, sh});
  });
}

int vsplat_write_ply(const char *path, const vsplat_cloud *cloud) {
  if (path == nullptr || cloud == nullptr) return invalid("null argument");
  return guarded([&] {
    vsplat::write_ply(path, vsplat::ConstPlyCloud{cloud->positions, cloud->scales, cloud->rotations, cloud->opacities, cloud->sh,
                                                  cloud->count, cloud->sh_degree});
  });
}

int vsplat_save_ply(vsplat_trainer *t, const char *path) {
  if (t == nullptr || path == nullptr) return invalid("null argument");
  return guarded([&] {
    const size_t n = t->impl.count(), coeffs = static_cast<size_t>((t->impl.sh_degree() + 1) * (t->impl.sh_degree() + 1));
    std::vector<float> positions(n * 3), scales(n * 3), rotations(n * 4), opacities(n), sh(n * coeffs * 3);
    t->impl.cloud(positions.data(), scales.data(), rotations.data(), opacities.data(), sh.data());
    vsplat::write_ply(path, vsplat::ConstPlyCloud{positions.data(), scales.data(), rotations.data(), opacities.data(), sh.data(), n,
                                                  t->impl.sh_degree()});
  });
}

int vsplat_enable_second_order(vsplat_trainer *t) {
  if (t == nullptr) return invalid("null trainer");
  return guarded([&] { t->impl.enable_second_order(); });
}

int vsplat_update_curvature(vsplat_trainer *t) {
  if (t == nullptr) return invalid("null trainer");
  return guarded([&] { t->impl.update_curvature(); });
}

int vsplat_reset_opacity(vsplat_trainer *t, float maximum) {
  if (t == nullptr) return invalid("null trainer");
  return guarded([&] { t->impl.reset_opacity(maximum); });
}

int vsplat_add_position_noise(vsplat_trainer *t, float scaler, uint32_t iteration) {
  if (t == nullptr) return invalid("null trainer");
  return guarded([&] { t->impl.add_position_noise(scaler, iteration); });
}

int vsplat_set_exposure(vsplat_trainer *t, const float *transform) {
  if (t == nullptr || transform == nullptr) return invalid("null argument");
  return guarded([&] { t->impl.set_exposure(transform); });
}

int vsplat_apply_gradients(vsplat_trainer *t, const vsplat_learning_rates *rates) {
  if (t == nullptr || rates == nullptr) return invalid("null argument");
  return guarded([&] { t->impl.apply_gradients(*rates); });
}

int vsplat_step(vsplat_trainer *t, const vsplat_camera *camera, const uint8_t *target, size_t bytes, const vsplat_learning_rates *rates,
                const uint8_t *next, size_t nextBytes, vsplat_step_stats *stats) {
  if (t == nullptr || camera == nullptr || rates == nullptr) return invalid("null argument");
  return guarded([&] {
    const vsplat_step_stats s = t->impl.run(*camera, target, bytes, vsplat::kModeStep, rates, next, nextBytes);
    if (stats) *stats = s;
  });
}

int vsplat_prefetch_target(vsplat_trainer *t, const uint8_t *target, size_t bytes) {
  if (t == nullptr) return invalid("null trainer");
  return guarded([&] { t->impl.prefetch_target(target, bytes); });
}

int vsplat_get_gradients(vsplat_trainer *t, float *positions, float *scales, float *rotations, float *opacities, float *sh) {
  if (t == nullptr) return invalid("null trainer");
  return guarded([&] { t->impl.gradients(positions, scales, rotations, opacities, sh); });
}

int vsplat_get_cloud(vsplat_trainer *t, float *positions, float *scales, float *rotations, float *opacities, float *sh) {
  if (t == nullptr) return invalid("null trainer");
  return guarded([&] { t->impl.cloud(positions, scales, rotations, opacities, sh); });
}

int vsplat_get_densify_stats(vsplat_trainer *t, float *out) {
  if (t == nullptr || out == nullptr) return invalid("null argument");
  return guarded([&] { t->impl.densify_stats(out); });
}

int vsplat_reset_densify_stats(vsplat_trainer *t) {
  if (t == nullptr) return invalid("null trainer");
  return guarded([&] { t->impl.reset_densify_stats(); });
}

int vsplat_apply_changes(vsplat_trainer *t, const vsplat_cloud *cloud, const int64_t *source, uint32_t count) {
  if (t == nullptr || (source == nullptr && count > 0)) return invalid("null argument");
  return guarded([&] { t->impl.apply_changes(cloud, source, count); });
}

int vsplat_set_active_sh_degree(vsplat_trainer *t, int32_t degree) {
  if (t == nullptr) return invalid("null trainer");
  return guarded([&] { t->impl.set_active_sh_degree(degree); });
}

int vsplat_set_filter3d(vsplat_trainer *t, const float *filter, size_t count) {
  if (t == nullptr) return invalid("null trainer");
  return guarded([&] { t->impl.set_filter3d(filter, count); });
}

int vsplat_set_sh_dropout(vsplat_trainer *t, float probability, int32_t retained, uint32_t seed) {
  if (t == nullptr) return invalid("null trainer");
  return guarded([&] { t->impl.set_sh_dropout(probability, retained, seed); });
}

int vsplat_parameters(vsplat_trainer *t, uint32_t group, void **pointer, uint64_t *bytes) {
  if (t == nullptr) return invalid("null trainer");
  return guarded([&] { t->impl.parameters(group, pointer, bytes); });
}

uint32_t vsplat_gaussian_count(const vsplat_trainer *t) { return t != nullptr ? t->impl.count() : 0u; }
int32_t vsplat_sh_degree(const vsplat_trainer *t) { return t != nullptr ? t->impl.sh_degree() : -1; }
int32_t vsplat_active_sh_degree(const vsplat_trainer *t) { return t != nullptr ? t->impl.active_sh_degree() : -1; }
int32_t vsplat_uses_fixed_point(const vsplat_trainer *t) { return t != nullptr && t->impl.fixed_point() ? 1 : 0; }
uint64_t vsplat_gpu_bytes(const vsplat_trainer *t) { return t != nullptr ? t->impl.gpu_bytes() : 0u; }
uint64_t vsplat_bytes_per_gaussian(const vsplat_trainer *t) { return t != nullptr ? t->impl.bytes_per_gaussian() : 0u; }
uint64_t vsplat_staging_bytes(const vsplat_trainer *t) { return t != nullptr ? t->impl.staging_bytes() : 0u; }

int vsplat_timings(const vsplat_trainer *t, vsplat_timing *out, uint32_t capacity, uint32_t *count) {
  if (t == nullptr) return invalid("null trainer");
  return guarded([&] {
    const auto stages = t->impl.timings();
    if (count) *count = static_cast<uint32_t>(stages.size());
    for (uint32_t i = 0; i < capacity && i < stages.size() && out != nullptr; ++i) {
      std::memset(out[i].name, 0, sizeof(out[i].name));
      std::strncpy(out[i].name, stages[i].first.c_str(), sizeof(out[i].name) - 1);
      out[i].milliseconds = stages[i].second;
    }
  });
}

const char *vsplat_last_error(void) { return lastError.c_str(); }

int vsplat_enqueue(vsplat_trainer *t, const vsplat_camera *camera, const uint8_t *deviceTarget, int32_t mode,
                   const vsplat_learning_rates *rates) {
  if (t == nullptr || camera == nullptr) return invalid("null argument");
  return guarded([&] { t->impl.enqueue(*camera, deviceTarget, static_cast<uint32_t>(mode), rates); });
}

int vsplat_synchronize(vsplat_trainer *t, vsplat_step_stats *stats) {
  if (t == nullptr) return invalid("null trainer");
  return guarded([&] {
    const vsplat_step_stats s = t->impl.synchronize();
    if (stats) *stats = s;
  });
}

int vsplat_step_views(vsplat_trainer *t, uint32_t n, const vsplat_camera *cameras, const uint8_t *const *targets, const size_t *bytes,
                      int32_t mode, const vsplat_learning_rates *rates, vsplat_step_stats *stats) {
  if (t == nullptr) return invalid("null trainer");
  return guarded([&] { t->impl.step_views(n, cameras, targets, bytes, static_cast<uint32_t>(mode), rates, stats); });
}

int vsplat_enqueue_views(vsplat_trainer *t, uint32_t n, const vsplat_camera *cameras, const uint8_t *const *device_targets, int32_t mode,
                         const vsplat_learning_rates *rates) {
  if (t == nullptr) return invalid("null trainer");
  return guarded([&] { t->impl.enqueue_views(n, cameras, device_targets, static_cast<uint32_t>(mode), rates); });
}

int vsplat_synchronize_views(vsplat_trainer *t, vsplat_step_stats *stats, uint32_t capacity, uint32_t *count) {
  if (t == nullptr) return invalid("null trainer");
  return guarded([&] {
    bool overflow = false;
    const uint32_t n = t->impl.synchronize_views(stats, capacity, overflow);
    if (count) *count = n;
    if (overflow) {
      throw vsplat::Error(VSPLAT_RUNTIME_ERROR,
                          "a queued view needed more tile entries than were reserved; its batch and the steps queued after it were skipped");
    }
  });
}

int vsplat_capture_views(vsplat_trainer *t, uint32_t n, const vsplat_camera *cameras, const uint8_t *const *device_targets, int32_t mode,
                         const vsplat_learning_rates *rates) {
  if (t == nullptr) return invalid("null trainer");
  return guarded([&] { t->impl.capture_views(n, cameras, device_targets, static_cast<uint32_t>(mode), rates); });
}

int vsplat_launch_captured(vsplat_trainer *t, uint32_t replays) {
  if (t == nullptr) return invalid("null trainer");
  return guarded([&] { t->impl.launch_captured(replays); });
}

int vsplat_autotune_tiles(vsplat_trainer *t, const vsplat_camera *camera, const uint8_t *device_target, uint32_t repeats,
                          uint32_t *tile_width, uint32_t *tile_height) {
  if (t == nullptr || camera == nullptr) return invalid("null argument");
  return guarded([&] {
    const auto shape = t->impl.autotune_tiles(*camera, device_target, repeats);
    if (tile_width) *tile_width = shape.first;
    if (tile_height) *tile_height = shape.second;
  });
}

int vsplat_autotune_views(vsplat_trainer *t, uint32_t n, const vsplat_camera *cameras, const uint8_t *const *device_targets,
                          uint32_t repeats, uint32_t *tile_widths, uint32_t *tile_heights) {
  if (t == nullptr || cameras == nullptr || device_targets == nullptr) return invalid("null argument");
  return guarded([&] { t->impl.autotune_views(n, cameras, device_targets, repeats, tile_widths, tile_heights); });
}

int vsplat_release_captured(vsplat_trainer *t) {
  if (t == nullptr) return invalid("null trainer");
  return guarded([&] { t->impl.release_captured(); });
}

void *vsplat_stream(vsplat_trainer *t) { return t != nullptr ? static_cast<void *>(t->impl.stream()) : nullptr; }

int vsplat_reserve_entries(vsplat_trainer *t, uint64_t entries) {
  if (t == nullptr) return invalid("null trainer");
  return guarded([&] { t->impl.reserve_entries(entries); });
}

} // extern "C"

namespace vsplat {
void math_probe(int function, const float *in, float *out, uint64_t n);
}

extern "C" int vsplat_inspect(vsplat_trainer *t, const char *buffer, void *out, uint64_t capacity, uint64_t *bytes) {
  if (t == nullptr || buffer == nullptr) return invalid("null argument");
  return guarded([&] { t->impl.inspect(buffer, out, capacity, bytes); });
}

extern "C" int vsplat_inspect_view(vsplat_trainer *t, uint32_t view, const char *buffer, void *out, uint64_t capacity,
                                   uint64_t *bytes) {
  if (t == nullptr || buffer == nullptr) return invalid("null argument");
  return guarded([&] { t->impl.inspect(buffer, out, capacity, bytes, view); });
}

extern "C" int vsplat_math_probe(int32_t function, const float *in, float *out, uint64_t n) {
  if ((in == nullptr || out == nullptr) && n > 0) return invalid("null argument");
  return guarded([&] { vsplat::math_probe(function, in, out, n); });
}

extern "C" int vsplat_device_alloc(uint64_t bytes, const void *host, void **out) {
  if (out == nullptr) return invalid("null argument");
  *out = nullptr;
  return guarded([&] {
    void *p = nullptr;
    if (hipMalloc(&p, bytes > 0 ? bytes : 1) != hipSuccess) throw vsplat::Error(VSPLAT_DEVICE_ERROR, "hipMalloc failed");
    if (host != nullptr && bytes > 0 && hipMemcpy(p, host, bytes, hipMemcpyHostToDevice) != hipSuccess) {
      (void)hipFree(p);
      throw vsplat::Error(VSPLAT_DEVICE_ERROR, "hipMemcpy to the device failed");
    }
    *out = p;
  });
}

extern "C" int vsplat_device_read(const void *device, void *host, uint64_t bytes) {
  if ((device == nullptr || host == nullptr) && bytes > 0) return invalid("null argument");
  return guarded([&] {
    if (bytes > 0 && hipMemcpy(host, device, bytes, hipMemcpyDeviceToHost) != hipSuccess) {
      throw vsplat::Error(VSPLAT_DEVICE_ERROR, "hipMemcpy to the host failed");
    }
  });
}

extern "C" int vsplat_device_copy(void *dst, const void *src, uint64_t bytes, uint32_t repeats, float *ms_per_copy) {
  if ((dst == nullptr || src == nullptr) && bytes > 0) return invalid("null argument");
  if (repeats == 0) return invalid("no copies asked for");
  return guarded([&] {
    const auto check = [](hipError_t e, const char *what) {
      if (e != hipSuccess) throw vsplat::Error(VSPLAT_DEVICE_ERROR, std::string(what) + ": " + hipGetErrorString(e));
    };
    hipStream_t stream = nullptr;
    hipEvent_t begin = nullptr, end = nullptr;
    auto release = [&] {
      if (begin != nullptr) (void)hipEventDestroy(begin);
      if (end != nullptr) (void)hipEventDestroy(end);
      if (stream != nullptr) (void)hipStreamDestroy(stream);
    };
    try {
      check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreateWithFlags");
      check(hipEventCreate(&begin), "hipEventCreate");
      check(hipEventCreate(&end), "hipEventCreate");
      check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, stream), "hipMemcpyAsync"); // warm-up
      check(hipEventRecord(begin, stream), "hipEventRecord");
      for (uint32_t r = 0; r < repeats; ++r) check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, stream), "hipMemcpyAsync");
      check(hipEventRecord(end, stream), "hipEventRecord");
      check(hipEventSynchronize(end), "hipEventSynchronize");
      float ms = 0.0f;
      check(hipEventElapsedTime(&ms, begin, end), "hipEventElapsedTime");
      if (ms_per_copy != nullptr) *ms_per_copy = ms / static_cast<float>(repeats);
    } catch (...) {
      release();
      throw;
    }
    release();
  });
}

extern "C" void vsplat_device_free(void *device) {
  if (device != nullptr) (void)hipFree(device);
}

extern "C" int vsplat_set_stage_timing(vsplat_trainer *t, int32_t enabled) {
  if (t == nullptr) return invalid("null trainer");
  return guarded([&] { t->impl.set_queued_timing(enabled != 0); });
}

extern "C" int vsplat_reset_timings(vsplat_trainer *t) {
  if (t == nullptr) return invalid("null trainer");
  return guarded([&] { t->impl.reset_timings(); });
}
