// Drives the reference-shaped C++ wrapper (include/vsplat_trainer.hpp) through
// real training on the GPU, as a caller of the reference's Trainer would:
//
//   overfit   tests/test_train.cpp:493-520: 400 steps on one view from a
//             40 x 40 grid of grey Gaussians; prints the first and last loss.
//   prefetch  tests/test_train.cpp:786-810: a prefetched image smaller than
//             the next step's does not spoil that step; prints the losses.
//
// Usage: wrapper_train overfit <target.rgba> | prefetch. Prints key=value lines.
#include "vsplat_trainer.hpp"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iterator>
#include <random>
#include <vector>

namespace {

vsplat::Camera camera(int width, int height, float focal) {
  vsplat::Camera c; // at the origin looking along +z (identity view)
  for (int i = 0; i < 16; ++i) c.view[i] = (i % 5 == 0) ? 1.f : 0.f;
  c.fx = c.fy = focal;
  c.cx = 0.5f * static_cast<float>(width);
  c.cy = 0.5f * static_cast<float>(height);
  c.width = width;
  c.height = height;
  return c;
}

// A side x side grid of grey Gaussians 2 units ahead of a camera at the origin.
vsplat::GaussianCloud grid(int side) {
  vsplat::GaussianCloud g;
  g.shDegree = 0;
  const std::size_t n = static_cast<std::size_t>(side) * side;
  const float spacing = 2.6f / static_cast<float>(side);
  for (std::size_t i = 0; i < n; ++i) {
    const float x = (static_cast<float>(i % side) + 0.5f - 0.5f * static_cast<float>(side)) * spacing;
    const float y = (static_cast<float>(i / side) + 0.5f - 0.5f * static_cast<float>(side)) * spacing;
    g.positions.insert(g.positions.end(), {x, y, 2.f});
    const float s = std::log(spacing * 0.75f);
    g.scales.insert(g.scales.end(), {s, s, s});
    g.rotations.insert(g.rotations.end(), {1.f, 0.f, 0.f, 0.f});
    g.opacities.push_back(0.f);
    g.sh.insert(g.sh.end(), {0.f, 0.f, 0.f});
  }
  return g;
}

std::vector<std::uint8_t> noise(int width, int height, unsigned seed) {
  std::mt19937 rng(seed);
  std::vector<std::uint8_t> out(static_cast<std::size_t>(width) * height * 4);
  for (std::size_t i = 0; i < out.size(); ++i) out[i] = (i % 4 == 3) ? 255 : static_cast<std::uint8_t>(rng() >> 24);
  return out;
}

vsplat::GaussianCloud random_cloud(int count, unsigned seed) {
  std::mt19937 rng(seed);
  std::uniform_real_distribution<float> u(-1.f, 1.f);
  vsplat::GaussianCloud c;
  c.shDegree = 0;
  for (int i = 0; i < count; ++i) {
    c.positions.insert(c.positions.end(), {u(rng), u(rng) * 0.75f, 3.f + u(rng)});
    c.scales.insert(c.scales.end(), {-2.5f + 0.5f * u(rng), -2.5f + 0.5f * u(rng), -2.5f + 0.5f * u(rng)});
    c.rotations.insert(c.rotations.end(), {1.f, 0.2f * u(rng), 0.2f * u(rng), 0.2f * u(rng)});
    c.opacities.push_back(u(rng));
    c.sh.insert(c.sh.end(), {u(rng), u(rng), u(rng)});
  }
  return c;
}

int overfit(const char *path) {
  std::ifstream in(path, std::ios::binary);
  const std::vector<std::uint8_t> target((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
  const vsplat::Camera view = camera(128, 96, 100.f);
  if (target.size() != 128u * 96u * 4u) {
    std::printf("error=target size %zu\n", target.size());
    return 2;
  }
  const vsplat::GaussianCloud initial = grid(40);
  vsplat::Trainer trainer(initial);
  vsplat::LearningRates rates;
  rates.position = 1e-3f;
  float first = 0.f, last = 0.f;
  for (int iteration = 0; iteration < 400; ++iteration) {
    const vsplat::StepStats stats = trainer.step(view, target, rates);
    if (!std::isfinite(stats.loss)) {
      std::printf("error=loss not finite at %d\n", iteration);
      return 3;
    }
    if (iteration == 0) first = stats.loss;
    last = stats.loss;
  }
  const vsplat::GaussianCloud trained = trainer.cloud();
  std::printf("first=%.6f\nlast=%.6f\n", first, last);
  std::printf("count=%zu\n", trained.size());
  std::printf("positions_moved=%d\nsh_moved=%d\n", trained.positions != initial.positions ? 1 : 0, trained.sh != initial.sh ? 1 : 0);
  return 0;
}

int prefetch() {
  const vsplat::Camera small = camera(64, 48, 60.f), large = camera(1024, 768, 60.f * 16.f);
  const auto smallTarget = noise(small.width, small.height, 71);
  const auto largeTarget = noise(large.width, large.height, 73), otherTarget = noise(large.width, large.height, 79);
  const vsplat::GaussianCloud cloud = random_cloud(60, 83);
  vsplat::Trainer reference(cloud);
  const float expected = reference.compute_gradients(large, largeTarget).loss;
  vsplat::LearningRates still;
  still.position = still.scale = still.rotation = still.opacity = still.sh = 0.f;
  vsplat::Trainer trainer(cloud);
  (void)trainer.render(large);            // large image buffers first
  trainer.prefetch_target(smallTarget);   // becomes the staging buffer
  (void)trainer.compute_gradients(small, smallTarget);
  const float afterSmall = trainer.step(large, largeTarget, still, otherTarget).loss;
  const std::uint64_t staging = trainer.staging_bytes();
  (void)trainer.step(large, largeTarget, still, largeTarget); // the prefetch is used by the next step
  const float prefetched = trainer.step(large, largeTarget, still).loss;
  std::printf("expected=%.7f\nafter_small=%.7f\nprefetched=%.7f\nstaging=%llu\nlarge_bytes=%zu\n", expected, afterSmall, prefetched,
              static_cast<unsigned long long>(staging), largeTarget.size());
  return 0;
}

} // namespace

int main(int argc, char **argv) {
  try {
    if (argc >= 3 && std::strcmp(argv[1], "overfit") == 0) return overfit(argv[2]);
    if (argc >= 2 && std::strcmp(argv[1], "prefetch") == 0) return prefetch();
  } catch (const std::exception &e) {
    std::printf("error=%s\n", e.what());
    return 1;
  }
  std::printf("usage: wrapper_train overfit <target.rgba> | prefetch\n");
  return 2;
}
