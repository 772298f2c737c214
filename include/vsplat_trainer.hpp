// vsplat_trainer.hpp: the trainer's rasterizer path as a C++ class, over the C
// interface of vsplat.h. It has the shape of the reference's `class Trainer`
// (src/train/trainer.hpp:89-270) and its option, rate and statistics structs
// (:24-73), and throws the exceptions the reference throws:
//
//   VSPLAT_RUNTIME_ERROR     std::runtime_error       (trainer.cpp:174, 276, 281, 571)
//   VSPLAT_BUDGET_EXCEEDED   vsplat::BudgetExceeded   (a std::runtime_error; trainer.hpp:46-49)
//   VSPLAT_INVALID_ARGUMENT  std::invalid_argument    (trainer.cpp:802)
//   VSPLAT_LOGIC_ERROR       std::logic_error         (trainer.cpp:853)
//   VSPLAT_DEVICE_ERROR      std::runtime_error
//   VSPLAT_UNSUPPORTED       vsplat::Unsupported      (a std::logic_error)
//
// Header-only, C++17; with C++20 the image arguments also take std::span as the
// reference's do. Link against libvsplat.so.
#pragma once

#include "vsplat.h"

#include <array>
#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#if __cplusplus >= 202002L
#include <span>
#endif

namespace vsplat {

class BudgetExceeded : public std::runtime_error {
public:
  using std::runtime_error::runtime_error;
};

class Unsupported : public std::logic_error {
public:
  using std::logic_error::logic_error;
};

namespace detail {
inline void check(int status) {
  if (status == VSPLAT_OK) return;
  const char *message = vsplat_last_error();
  const std::string what = message != nullptr ? message : "vsplat error";
  switch (status) {
  case VSPLAT_BUDGET_EXCEEDED:
    throw BudgetExceeded(what);
  case VSPLAT_INVALID_ARGUMENT:
    throw std::invalid_argument(what);
  case VSPLAT_LOGIC_ERROR:
    throw std::logic_error(what);
  case VSPLAT_UNSUPPORTED:
    throw Unsupported(what);
  default:
    throw std::runtime_error(what);
  }
}
} // namespace detail

// GaussianCloud (src/splat/gaussian_cloud.hpp:21-38).
struct GaussianCloud {
  int shDegree = 0;
  std::vector<float> positions, scales, rotations, opacities, sh;
  std::size_t size() const noexcept { return opacities.size(); }
  bool empty() const noexcept { return opacities.empty(); }
  int shCoeffs() const noexcept { return (shDegree + 1) * (shDegree + 1); }
};

// Camera (src/render/camera.hpp:9-18); view is column-major, view[c * 4 + r].
struct Camera {
  std::array<float, 16> view{1.f, 0.f, 0.f, 0.f, 0.f, 1.f, 0.f, 0.f, 0.f, 0.f, 1.f, 0.f, 0.f, 0.f, 0.f, 1.f};
  std::array<float, 3> position{0.f, 0.f, 0.f};
  float fx = 1.f, fy = 1.f, cx = 0.f, cy = 0.f;
  int width = 1, height = 1;
};

// TrainerOptions (trainer.hpp:24-42), with the same defaults.
struct TrainerOptions {
  float lambdaSsim = 0.2f;
  std::array<float, 3> background{0.f, 0.f, 0.f};
  float nearPlane = 0.01f;
  bool forceFixedPoint = false;
  bool selectiveAdam = true;
  float opacityReg = 0.f;
  float scaleReg = 0.f;
  float positionNoise = 0.f;
  bool mipFilter2d = false;
  bool poseGradients = false;
  std::uint64_t memoryBudget = ~std::uint64_t{0};
  bool bf16Sh = false; // this build only: SH stored as bf16 (vsplat_trainer_options::sh_storage)
  bool recomputeTransmittance = false; // this build only (vsplat_trainer_options::recompute_transmittance)
  int tileShape = VSPLAT_TILES_16X16;  // this build only (vsplat_trainer_options::tile_shape)
};

// LearningRates (trainer.hpp:51-61).
struct LearningRates {
  float position = 1.6e-4f;
  float scale = 0.005f;
  float rotation = 0.001f;
  float opacity = 0.05f;
  float sh = 0.025f;
  float shRestScale = 0.05f;
  float trustRegion = 0.f;
};

// Exposure and kNeutralExposure (trainer.hpp:76-77).
using Exposure = std::array<float, 12>;
inline constexpr Exposure kNeutralExposure{1.f, 0.f, 0.f, 0.f, 0.f, 1.f, 0.f, 0.f, 0.f, 0.f, 1.f, 0.f};

// StepStats (trainer.hpp:63-73).
struct StepStats {
  float loss = 0.f;
  float l1 = 0.f;
  float ssim = 0.f;
  std::uint32_t visibleSplats = 0;
  std::uint32_t tileEntries = 0;
  std::array<float, 12> exposureGradient{};
  std::array<float, 6> poseGradient{};
};

// A borrowed byte range: the reference's std::span<const std::uint8_t>.
struct Bytes {
  const std::uint8_t *data = nullptr;
  std::size_t size = 0;
  Bytes() = default;
  Bytes(const std::uint8_t *d, std::size_t n) : data(d), size(n) {}
  Bytes(const std::vector<std::uint8_t> &v) : data(v.data()), size(v.size()) {}
#if __cplusplus >= 202002L
  Bytes(std::span<const std::uint8_t> s) : data(s.data()), size(s.size()) {}
#endif
};

class Trainer {
public:
  struct Gradients {
    std::vector<float> positions, scales, rotations, opacities, sh;
  };

  explicit Trainer(const GaussianCloud &initial, const TrainerOptions &options = {}) {
    if (initial.empty()) throw std::runtime_error("the cloud is empty");
    const std::size_t n = initial.size();
    if (initial.positions.size() != 3 * n || initial.scales.size() != 3 * n || initial.rotations.size() != 4 * n ||
        initial.sh.size() != 3 * n * static_cast<std::size_t>(initial.shCoeffs()))
      throw std::runtime_error("the cloud's arrays do not match its size");
    vsplat_cloud c{initial.positions.data(), initial.scales.data(), initial.rotations.data(), initial.opacities.data(),
                   initial.sh.data(),        static_cast<std::uint32_t>(n), initial.shDegree};
    vsplat_trainer_options o;
    vsplat_default_options(&o);
    o.lambda_ssim = options.lambdaSsim;
    for (int i = 0; i < 3; ++i) o.background[i] = options.background[i];
    o.near_plane = options.nearPlane;
    o.force_fixed_point = options.forceFixedPoint ? 1 : 0;
    o.selective_adam = options.selectiveAdam ? 1 : 0;
    o.opacity_reg = options.opacityReg;
    o.scale_reg = options.scaleReg;
    o.position_noise = options.positionNoise;
    o.mip_filter_2d = options.mipFilter2d ? 1 : 0;
    o.pose_gradients = options.poseGradients ? 1 : 0;
    o.memory_budget = options.memoryBudget;
    o.sh_storage = options.bf16Sh ? VSPLAT_SH_BF16 : VSPLAT_SH_HALF;
    o.recompute_transmittance = options.recomputeTransmittance ? 1 : 0;
    o.tile_shape = options.tileShape;
    detail::check(vsplat_trainer_create(&c, &o, &trainer_));
  }
  ~Trainer() { vsplat_trainer_destroy(trainer_); }
  Trainer(const Trainer &) = delete;
  Trainer &operator=(const Trainer &) = delete;

  // width * height * 4 floats: rgb and the log of the transmittance.
  std::vector<float> render(const Camera &camera) {
    const vsplat_camera c = to_c(camera);
    std::vector<float> out(static_cast<std::size_t>(camera.width) * static_cast<std::size_t>(camera.height) * 4);
    detail::check(vsplat_render(trainer_, &c, out.data()));
    return out;
  }

  StepStats compute_gradients(const Camera &camera, Bytes target) {
    const vsplat_camera c = to_c(camera);
    vsplat_step_stats s{};
    detail::check(vsplat_compute_gradients(trainer_, &c, target.data, target.size, &s));
    return from_c(s);
  }

  StepStats collect_statistics(const Camera &camera, Bytes target) {
    const vsplat_camera c = to_c(camera);
    vsplat_step_stats s{};
    detail::check(vsplat_collect_statistics(trainer_, &c, target.data, target.size, &s));
    return from_c(s);
  }

  // Trainer::score_view (trainer.hpp:111).
  StepStats score_view(const Camera &camera, Bytes target, float errorShare, float pixelWeight = 1.f) {
    const vsplat_camera c = to_c(camera);
    vsplat_step_stats s{};
    detail::check(vsplat_score_view(trainer_, &c, target.data, target.size, errorShare, pixelWeight, &s));
    return from_c(s);
  }

  // Trainer::enable_second_order and Trainer::update_curvature (trainer.hpp:117-118).
  void enable_second_order() { detail::check(vsplat_enable_second_order(trainer_)); }
  void update_curvature() { detail::check(vsplat_update_curvature(trainer_)); }

  // Trainer::reset_opacity and Trainer::add_position_noise (trainer.hpp:162-165).
  void reset_opacity(float maximum) { detail::check(vsplat_reset_opacity(trainer_, maximum)); }
  void add_position_noise(float scaler, std::uint32_t iteration) {
    detail::check(vsplat_add_position_noise(trainer_, scaler, iteration));
  }

  // Trainer::set_exposure (trainer.hpp:154): rows of (r, g, b, 1).
  void set_exposure(const Exposure &exposure) { detail::check(vsplat_set_exposure(trainer_, exposure.data())); }

  void apply_gradients(const LearningRates &rates) {
    const vsplat_learning_rates r = to_c(rates);
    detail::check(vsplat_apply_gradients(trainer_, &r));
  }

  // nextTarget stays borrowed until the next step() or compute_gradients().
  StepStats step(const Camera &camera, Bytes target, const LearningRates &rates, Bytes nextTarget = {}) {
    const vsplat_camera c = to_c(camera);
    const vsplat_learning_rates r = to_c(rates);
    vsplat_step_stats s{};
    detail::check(vsplat_step(trainer_, &c, target.data, target.size, &r, nextTarget.data, nextTarget.size, &s));
    return from_c(s);
  }

  void prefetch_target(Bytes target) { detail::check(vsplat_prefetch_target(trainer_, target.data, target.size)); }

  GaussianCloud cloud() const {
    GaussianCloud c;
    c.shDegree = sh_degree();
    const std::size_t n = gaussian_count();
    c.positions.resize(3 * n);
    c.scales.resize(3 * n);
    c.rotations.resize(4 * n);
    c.opacities.resize(n);
    c.sh.resize(3 * n * static_cast<std::size_t>(c.shCoeffs()));
    detail::check(vsplat_get_cloud(trainer_, c.positions.data(), c.scales.data(), c.rotations.data(), c.opacities.data(), c.sh.data()));
    return c;
  }

  Gradients gradients() const {
    Gradients g;
    const std::size_t n = gaussian_count();
    const std::size_t coeffs = static_cast<std::size_t>((sh_degree() + 1) * (sh_degree() + 1));
    g.positions.resize(3 * n);
    g.scales.resize(3 * n);
    g.rotations.resize(4 * n);
    g.opacities.resize(n);
    g.sh.resize(3 * n * coeffs);
    detail::check(vsplat_get_gradients(trainer_, g.positions.data(), g.scales.data(), g.rotations.data(), g.opacities.data(), g.sh.data()));
    return g;
  }

  // The device buffer of a parameter group (0 positions, 1 scales, 2 rotations,
  // 3 opacities, 4 SH as half_sh_words describes) and its size in bytes.
  std::pair<void *, std::uint64_t> parameters(std::size_t group) const {
    void *pointer = nullptr;
    std::uint64_t bytes = 0;
    detail::check(vsplat_parameters(trainer_, static_cast<std::uint32_t>(group), &pointer, &bytes));
    return {pointer, bytes};
  }

  int sh_degree() const noexcept { return vsplat_sh_degree(trainer_); }
  int active_sh_degree() const noexcept { return vsplat_active_sh_degree(trainer_); }
  bool uses_fixed_point() const noexcept { return vsplat_uses_fixed_point(trainer_) != 0; }
  void set_active_sh_degree(int degree) { detail::check(vsplat_set_active_sh_degree(trainer_, degree)); }
  void set_filter3d(const std::vector<float> &filter) { detail::check(vsplat_set_filter3d(trainer_, filter.data(), filter.size())); }
  void set_sh_dropout(float probability, int retainedDegree, std::uint32_t seed) {
    detail::check(vsplat_set_sh_dropout(trainer_, probability, retainedDegree, seed));
  }

  // 8 floats per Gaussian: the DensifyStat of shaders/train_types.glsl:18-21.
  std::vector<float> densify_stats() const {
    std::vector<float> out(8 * gaussian_count());
    detail::check(vsplat_get_densify_stats(trainer_, out.data()));
    return out;
  }
  void reset_densify_stats() { detail::check(vsplat_reset_densify_stats(trainer_)); }

  // Trainer::apply_changes (trainer.cpp:749-787) with DensifyResult's cloud and
  // source (densify.hpp:52-63): new Gaussian g keeps old Gaussian source[g]'s
  // optimiser state, or starts fresh when source[g] < 0.
  void apply_changes(const GaussianCloud &cloud, const std::vector<std::int64_t> &source) {
    if (cloud.empty()) throw std::runtime_error("densification removed every Gaussian");
    const std::size_t n = cloud.size();
    if (source.size() != n || cloud.positions.size() != 3 * n || cloud.scales.size() != 3 * n || cloud.rotations.size() != 4 * n ||
        cloud.sh.size() != 3 * n * static_cast<std::size_t>(cloud.shCoeffs()))
      throw std::invalid_argument("the cloud's arrays or the source table do not match its size");
    vsplat_cloud c{cloud.positions.data(), cloud.scales.data(), cloud.rotations.data(), cloud.opacities.data(),
                   cloud.sh.data(),        static_cast<std::uint32_t>(n), cloud.shDegree};
    detail::check(vsplat_apply_changes(trainer_, &c, source.data(), static_cast<std::uint32_t>(n)));
  }
  // The same with the values gathered on the device as well (every source >= 0):
  // prunes and plain copies, with nothing transferred but the table.
  void apply_changes(const std::vector<std::int64_t> &source) {
    detail::check(vsplat_apply_changes(trainer_, nullptr, source.data(), static_cast<std::uint32_t>(source.size())));
  }

  std::vector<std::pair<std::string, double>> timings() const {
    std::uint32_t count = 0;
    detail::check(vsplat_timings(trainer_, nullptr, 0, &count));
    std::vector<vsplat_timing> raw(count);
    detail::check(vsplat_timings(trainer_, raw.data(), count, &count));
    std::vector<std::pair<std::string, double>> out;
    for (std::uint32_t i = 0; i < count && i < raw.size(); ++i) out.emplace_back(raw[i].name, raw[i].milliseconds);
    return out;
  }

  std::size_t gaussian_count() const noexcept { return vsplat_gaussian_count(trainer_); }
  std::uint64_t gpu_bytes() const { return vsplat_gpu_bytes(trainer_); }
  std::uint64_t staging_bytes() const noexcept { return vsplat_staging_bytes(trainer_); }
  std::uint64_t bytes_per_gaussian() const noexcept { return vsplat_bytes_per_gaussian(trainer_); }
  static constexpr std::uint64_t kBytesPerTileEntry = VSPLAT_BYTES_PER_TILE_ENTRY;

  // The C handle, for the device-resident calls of vsplat.h.
  vsplat_trainer *handle() const noexcept { return trainer_; }

private:
  static vsplat_camera to_c(const Camera &camera) {
    vsplat_camera c{};
    for (int i = 0; i < 16; ++i) c.view[i] = camera.view[i];
    for (int i = 0; i < 3; ++i) c.position[i] = camera.position[i];
    c.fx = camera.fx;
    c.fy = camera.fy;
    c.cx = camera.cx;
    c.cy = camera.cy;
    c.width = camera.width;
    c.height = camera.height;
    return c;
  }
  static vsplat_learning_rates to_c(const LearningRates &r) {
    return vsplat_learning_rates{r.position, r.scale, r.rotation, r.opacity, r.sh, r.shRestScale, r.trustRegion};
  }
  static StepStats from_c(const vsplat_step_stats &s) {
    StepStats out;
    out.loss = s.loss;
    out.l1 = s.l1;
    out.ssim = s.ssim;
    out.visibleSplats = s.visible_splats;
    out.tileEntries = s.tile_entries;
    for (int i = 0; i < 12; ++i) out.exposureGradient[i] = s.exposure_gradient[i];
    for (int i = 0; i < 6; ++i) out.poseGradient[i] = s.pose_gradient[i];
    return out;
  }

  vsplat_trainer *trainer_ = nullptr;
};

// refined_pose (trainer.hpp:78-80): the camera moved by (rotation w, translation v).
[[nodiscard]] inline Camera refined_pose(const Camera &camera, const std::array<float, 6> &move) {
  vsplat_camera c{};
  for (int i = 0; i < 16; ++i) c.view[i] = camera.view[i];
  for (int i = 0; i < 3; ++i) c.position[i] = camera.position[i];
  c.fx = camera.fx, c.fy = camera.fy, c.cx = camera.cx, c.cy = camera.cy, c.width = camera.width, c.height = camera.height;
  vsplat_camera moved{};
  detail::check(vsplat_refined_pose(&c, move.data(), &moved));
  Camera out = camera;
  for (int i = 0; i < 16; ++i) out.view[i] = moved.view[i];
  for (int i = 0; i < 3; ++i) out.position[i] = moved.position[i];
  return out;
}

} // namespace vsplat
