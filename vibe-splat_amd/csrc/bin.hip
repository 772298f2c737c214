// Tile binning: how many 16x16 tiles each depth-sorted splat reaches, the
// (tile, splat) entries in depth order, and each tile's run of entries after
// the stable sort by tile (train_bin_count.comp, train_bin_emit.comp,
// train_tile_ranges.comp).
#include "device.hpp"
#include "kernels.hpp"

namespace vsplat {

namespace {

// Inclusive tile box of a splat on a gridW x gridH grid (train_types.glsl:34-39).
__device__ __forceinline__ void splat_tiles(const Splat &s, uint32_t gridW, uint32_t gridH, uint32_t t[4]) {
  const float lastX = static_cast<float>(gridW) - 1.0f, lastY = static_cast<float>(gridH) - 1.0f;
  t[0] = static_cast<uint32_t>(clampf(floorf((s.cx - s.ex - 1.0f) * 0.0625f), 0.0f, lastX));
  t[1] = static_cast<uint32_t>(clampf(floorf((s.cy - s.ey - 1.0f) * 0.0625f), 0.0f, lastY));
  t[2] = static_cast<uint32_t>(clampf(floorf((s.cx + s.ex + 1.0f) * 0.0625f), 0.0f, lastX));
  t[3] = static_cast<uint32_t>(clampf(floorf((s.cy + s.ey + 1.0f) * 0.0625f), 0.0f, lastY));
}

// The least of q = qx x^2 + 2 qy x y + qz y^2 along one edge of the tile's
// square of pixel centres, at the clamped stationary point.
__device__ __forceinline__ float edge_min_vertical(const Splat &s, float iqz, float x, float ylo, float yhi) {
  const float y = clampf((-s.qy * x) * iqz, ylo, yhi);
  return fma32(s.qz * y, y, fma32(2.0f * s.qy * x, y, s.qx * x * x));
}
__device__ __forceinline__ float edge_min_horizontal(const Splat &s, float iqx, float v, float xlo, float xhi) {
  const float u = clampf((-s.qy * v) * iqx, xlo, xhi);
  return fma32(s.qz * v, v, fma32(2.0f * s.qy * u, v, s.qx * u * u));
}

// Whether the splat's 1/255 ellipse reaches a pixel centre of tile (tx, ty)
// (train_types.glsl:56-77), with the same slack so rounding never drops a pixel.
// iqx and iqz are 1/qx and 1/qz, computed once per splat.
__device__ __forceinline__ bool reaches_tile(const Splat &s, float reach, float iqx, float iqz, uint32_t tx, uint32_t ty) {
  const float lowx = (static_cast<float>(tx) * 16.0f + 0.5f) - s.cx, lowy = (static_cast<float>(ty) * 16.0f + 0.5f) - s.cy;
  const float highx = lowx + 15.0f, highy = lowy + 15.0f;
  if (lowx <= 0.0f && lowy <= 0.0f && highx >= 0.0f && highy >= 0.0f) return true;
  float least = 1e30f;
  least = minf(least, edge_min_vertical(s, iqz, lowx, lowy, highy));
  least = minf(least, edge_min_horizontal(s, iqx, lowy, lowx, highx));
  least = minf(least, edge_min_vertical(s, iqz, highx, lowy, highy));
  least = minf(least, edge_min_horizontal(s, iqx, highy, lowx, highx));
  return least <= fma32(reach, 1.001f, 1e-3f);
}

__global__ __launch_bounds__(256) void bin_count_kernel(const uint32_t *__restrict__ keys, const uint32_t *__restrict__ values,
                                                        const Splat *__restrict__ splats, uint32_t count, uint32_t gridW,
                                                        uint32_t gridH, uint32_t *__restrict__ coverage, uint32_t *counters) {
  const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  bool visible = false;
  if (r < count) {
    uint32_t n = 0;
    if (keys[r] != kInvisible) {
      visible = true;
      const Splat s = load_splat(splats, values[r]);
      const float reach = reach_of(s.o), iqx = 1.0f / s.qx, iqz = 1.0f / s.qz;
      uint32_t t[4];
      splat_tiles(s, gridW, gridH, t);
      for (uint32_t y = t[1]; y <= t[3]; ++y)
        for (uint32_t x = t[0]; x <= t[2]; ++x) n += reaches_tile(s, reach, iqx, iqz, x, y) ? 1u : 0u;
    }
    coverage[r] = n;
  }
  // One atomic per wave for the visible count.
  const uint64_t mask = __ballot(visible);
  if ((threadIdx.x & 63u) == 0 && mask != 0) atomicAdd(&counters[kVisible], static_cast<uint32_t>(__popcll(mask)));
}

// Whether the entries fit. A view that does not fit sets the sticky overflow
// word; from then on every queued pass lists nothing and its chain rule does
// nothing, so the passes that ran are exactly those before the first overflow.
__global__ void bin_check_kernel(uint32_t *counters, uint32_t capacity, uint32_t isStep) {
  const uint32_t entries = counters[kEntries];
  if (counters[kOverflow] != 0 || entries > capacity) {
    counters[kEntriesUsed] = 0;
    counters[kOverflow] = 1;
    counters[kSkipped] += isStep;
  } else {
    counters[kEntriesUsed] = entries;
  }
}

__global__ __launch_bounds__(256) void bin_emit_kernel(const uint32_t *__restrict__ keys, const uint32_t *__restrict__ values,
                                                       const Splat *__restrict__ splats, uint32_t count, uint32_t gridW,
                                                       uint32_t gridH, const uint32_t *__restrict__ offsets, const uint32_t *counters,
                                                       uint32_t *__restrict__ entryKeys, uint32_t *__restrict__ entryValues) {
  const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= count || counters[kOverflow] != 0 || keys[r] == kInvisible) return;
  const uint32_t splat = values[r];
  const Splat s = load_splat(splats, splat);
  const float reach = reach_of(s.o), iqx = 1.0f / s.qx, iqz = 1.0f / s.qz;
  uint32_t t[4];
  splat_tiles(s, gridW, gridH, t);
  uint32_t slot = offsets[r];
  for (uint32_t y = t[1]; y <= t[3]; ++y) {
    for (uint32_t x = t[0]; x <= t[2]; ++x) {
      if (!reaches_tile(s, reach, iqx, iqz, x, y)) continue;
      entryKeys[slot] = y * gridW + x;
      entryValues[slot] = splat;
      ++slot;
    }
  }
}

// Start and end of each tile's run; every write belongs to one thread (train_tile_ranges.comp:11-22).
__global__ __launch_bounds__(256) void tile_ranges_kernel(const uint32_t *__restrict__ entryKeys, const uint32_t *counters,
                                                          uint2 *__restrict__ ranges) {
  const uint32_t count = counters[kEntriesUsed];
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < count; k += gridDim.x * blockDim.x) {
    const uint32_t tile = entryKeys[k];
    if (k == 0 || entryKeys[k - 1] != tile) ranges[tile].x = k;
    if (k + 1 == count || entryKeys[k + 1] != tile) ranges[tile].y = k + 1;
  }
}

} // namespace

void launch_bin_count(hipStream_t stream, const uint32_t *keys, const uint32_t *values, const Splat *splats, uint32_t count,
                      uint32_t gridW, uint32_t gridH, uint32_t *coverage, uint32_t *counters) {
  if (count == 0) return;
  hipLaunchKernelGGL(bin_count_kernel, dim3((count + 255) / 256), dim3(256), 0, stream, keys, values, splats, count, gridW, gridH,
                     coverage, counters);
}

void launch_bin_check(hipStream_t stream, uint32_t *counters, uint32_t capacity, bool isStep) {
  hipLaunchKernelGGL(bin_check_kernel, dim3(1), dim3(1), 0, stream, counters, capacity, isStep ? 1u : 0u);
}

void launch_bin_emit(hipStream_t stream, const uint32_t *keys, const uint32_t *values, const Splat *splats, uint32_t count,
                     uint32_t gridW, uint32_t gridH, const uint32_t *offsets, const uint32_t *counters, uint32_t *entryKeys,
                     uint32_t *entryValues) {
  if (count == 0) return;
  hipLaunchKernelGGL(bin_emit_kernel, dim3((count + 255) / 256), dim3(256), 0, stream, keys, values, splats, count, gridW, gridH,
                     offsets, counters, entryKeys, entryValues);
}

void launch_tile_ranges(hipStream_t stream, const uint32_t *entryKeys, const uint32_t *counters, uint32_t capacity, uint2 *ranges) {
  if (capacity == 0) return;
  const uint32_t blocks = (capacity + 255) / 256 < 16384 ? (capacity + 255) / 256 : 16384;
  hipLaunchKernelGGL(tile_ranges_kernel, dim3(blocks), dim3(256), 0, stream, entryKeys, counters, ranges);
}

} // namespace vsplat
