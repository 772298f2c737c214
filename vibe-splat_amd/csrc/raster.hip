// Per-tile compositing and its derivative (train_forward.comp:38-100,
// train_backward.glsl:78-158). One workgroup of 256 threads per 16x16 tile,
// one thread per pixel; the tile's splats are staged through LDS in batches
// of 256 with their pixel box already cut to the tile, and every thread works
// on its own pixel in registers.
//
// Backward: wave w of a tile holds pixel rows 4w..4w+3. For each splat the 14
// per-pixel sums are reduced over the wave by a transposed butterfly (each
// xor step halves the values a lane carries, so 16 slots cost 15 shuffles),
// after which lane 4j holds total j; the totals go to global memory with one
// atomic each, as 64-bit fixed point (quantised per wave total, so the integer
// sums are the same in any order) or as float atomics.
#include "device.hpp"
#include "kernels.hpp"

namespace vsplat {

namespace {

// A splat as the pixel loops use it: the extent is replaced by the pixel box
// within this tile.
struct Staged {
  float qx, qy, qz, o;
  float r, g, b, radius;
  float cx, cy;
  uint32_t box, pad;
};
static_assert(sizeof(Staged) == 48, "48-byte staged splat");

__device__ __forceinline__ Staged stage(const Splat &s, uint32_t x0, uint32_t y0, uint32_t x1, uint32_t y1) {
  return Staged{s.qx, s.qy, s.qz, s.o, s.r, s.g, s.b, s.radius, s.cx, s.cy, splat_box(s, x0, y0, x1, y1), 0u};
}

// splat_at (train_types.glsl:81-94); e = exp(power) is kept for the backward pass.
__device__ __forceinline__ bool splat_at(const Staged &s, uint32_t x, uint32_t y, float &dx, float &dy, float &e, float &alpha) {
  dx = (static_cast<float>(x) + 0.5f) - s.cx;
  dy = (static_cast<float>(y) + 0.5f) - s.cy;
  if (fabsf(dx) > s.radius || fabsf(dy) > s.radius) return false;
  const float power = fma32(-0.5f, fma32(s.qx * dx, dx, (s.qz * dy) * dy), -((s.qy * dx) * dy));
  if (power > 0.0f) return false;
  e = vs_exp(power);
  const float a = s.o * e;
  alpha = a < 0.99f ? a : 0.99f;
  return alpha >= kMinAlpha;
}

struct TileGeometry {
  uint32_t x0, y0, x1, y1;
};
__device__ __forceinline__ TileGeometry tile_geometry(uint32_t tile, uint32_t gridW, uint32_t width, uint32_t height) {
  const uint32_t tx = tile % gridW, ty = tile / gridW;
  const uint32_t x0 = tx * kTileSize, y0 = ty * kTileSize;
  return {x0, y0, x0 + kTileSize < width ? x0 + kTileSize : width, y0 + kTileSize < height ? y0 + kTileSize : height};
}

__global__ __launch_bounds__(256) void forward_kernel(const Splat *__restrict__ splats, const uint32_t *__restrict__ entryValues,
                                                      const uint2 *__restrict__ ranges, uint32_t width, uint32_t height, uint32_t gridW,
                                                      float bgR, float bgG, float bgB, float4 *__restrict__ image,
                                                      uint32_t *__restrict__ stops) {
  __shared__ Staged batch[kTilePixels];
  const uint32_t tile = blockIdx.x;
  const TileGeometry geo = tile_geometry(tile, gridW, width, height);
  const uint32_t lx = threadIdx.x & 15u, ly = threadIdx.x >> 4;
  const uint32_t x = geo.x0 + lx, y = geo.y0 + ly;
  const bool inside = x < width && y < height;
  const uint2 range = ranges[tile];
  float r = 0.0f, g = 0.0f, b = 0.0f, T = 1.0f, logT = 0.0f;
  uint32_t stop = range.x;
  bool done = !inside;
  for (uint32_t end = range.y; end > range.x;) {
    const uint32_t count = end - range.x < kTilePixels ? end - range.x : kTilePixels;
    end -= count;
    // Also the barrier that keeps the batch from being overwritten while read.
    if (__syncthreads_count(!done) == 0) break;
    if (threadIdx.x < count) batch[threadIdx.x] = stage(load_splat(splats, entryValues[end + threadIdx.x]), geo.x0, geo.y0, geo.x1, geo.y1);
    __syncthreads();
    if (done) continue;
    for (uint32_t k = count; k-- > 0;) {
      const Staged s = batch[k];
      if (!in_box(s.box, lx, ly)) continue;
      float dx, dy, e, alpha;
      if (!splat_at(s, x, y, dx, dy, e, alpha)) continue;
      const float logRest = vs_log(1.0f - alpha);
      if (logT + logRest < kLogMinTransmittance) {
        stop = end + k + 1;
        done = true;
        break;
      }
      const float w = T * alpha;
      r = fma32(w, s.r, r);
      g = fma32(w, s.g, g);
      b = fma32(w, s.b, b);
      T = T * (1.0f - alpha);
      logT = logT + logRest;
    }
  }
  if (inside) {
    const uint32_t at = y * width + x;
    image[at] = make_float4(fma32(T, bgR, r), fma32(T, bgG, g), fma32(T, bgB, b), logT);
    stops[at] = stop;
  }
}

// The fixed-point quantisation of a wave total (train_backward_fixed.comp:19-21), NaN as 0.
__device__ __forceinline__ int32_t quantise(float v, float scale) {
  float s = v * scale;
  if (s != s) s = 0.0f;
  s = s > 1e9f ? 1e9f : (s < -1e9f ? -1e9f : s);
  return static_cast<int32_t>(__builtin_rintf(s));
}

template <int Count>
__device__ __forceinline__ void halve(float (&v)[16], uint32_t lane, int offset) {
  const bool upper = (lane & static_cast<uint32_t>(offset)) != 0;
#pragma unroll
  for (int i = 0; i < Count / 2; ++i) {
    const float send = upper ? v[i] : v[i + Count / 2];
    const float keep = upper ? v[i + Count / 2] : v[i];
    v[i] = keep + __shfl_xor(send, offset, 64);
  }
}

template <bool Fixed>
__global__ __launch_bounds__(256) void backward_kernel(const Splat *__restrict__ splats, const uint32_t *__restrict__ entryValues,
                                                       const uint2 *__restrict__ ranges, uint32_t width, uint32_t height, uint32_t gridW,
                                                       float bgR, float bgG, float bgB, const float4 *__restrict__ image,
                                                       const float4 *__restrict__ adjoint, const uint32_t *__restrict__ stops,
                                                       float fixedScale, unsigned long long *__restrict__ rawFixed,
                                                       float *__restrict__ rawFloat) {
  __shared__ Staged batch[kTilePixels];
  __shared__ uint32_t batchIndex[kTilePixels];
  __shared__ uint32_t tileStop;
  const uint32_t tile = blockIdx.x;
  const TileGeometry geo = tile_geometry(tile, gridW, width, height);
  const uint32_t lx = threadIdx.x & 15u, ly = threadIdx.x >> 4, lane = threadIdx.x & 63u;
  const uint32_t x = geo.x0 + lx, y = geo.y0 + ly;
  const bool inside = x < width && y < height;
  const uint32_t at = inside ? y * width + x : 0u;
  const float logTotal = inside ? image[at].w : 0.0f;
  float4 adj = inside ? adjoint[at] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  const float perPixel = 1.0f / (3.0f * static_cast<float>(width) * static_cast<float>(height));
  const uint2 range = ranges[tile];
  const uint32_t stop = inside ? stops[at] : range.y;
  if (threadIdx.x == 0) tileStop = range.y;
  __syncthreads();
  atomicMin(&tileStop, stop);
  __syncthreads();
  const uint32_t first = tileStop;
  float Br = bgR, Bg = bgG, Bb = bgB, logBehind = 0.0f;

  for (uint32_t start = first; start < range.y; start += kTilePixels) {
    const uint32_t count = range.y - start < kTilePixels ? range.y - start : kTilePixels;
    __syncthreads();
    if (threadIdx.x < count) {
      const uint32_t index = entryValues[start + threadIdx.x];
      batchIndex[threadIdx.x] = index;
      batch[threadIdx.x] = stage(load_splat(splats, index), geo.x0, geo.y0, geo.x1, geo.y1);
    }
    __syncthreads();
    for (uint32_t k = 0; k < count; ++k) {
      const Staged s = batch[k];
      float v[16];
#pragma unroll
      for (int i = 0; i < 16; ++i) v[i] = 0.0f;
      bool active = false;
      float dx, dy, e, alpha;
      if (inside && start + k >= stop && in_box(s.box, lx, ly) && splat_at(s, x, y, dx, dy, e, alpha)) {
        active = true;
        const float logRest = vs_log(1.0f - alpha);
        const float T = vs_exp((logTotal - logBehind) - logRest);
        const float Ta = T * alpha;
        v[6] = Ta * adj.x;
        v[7] = Ta * adj.y;
        v[8] = Ta * adj.z;
        v[9] = Ta * perPixel;
        // At the 0.99 clamp alpha no longer depends on the parameters.
        if (s.o * e < 0.99f) {
          const float dAlpha = T * fma32(adj.z, s.b - Bb, fma32(adj.y, s.g - Bg, adj.x * (s.r - Br)));
          v[0] = dAlpha * e;
          const float dPower = dAlpha * alpha;
          const float vx = fma32(s.qy, dy, s.qx * dx), vy = fma32(s.qz, dy, s.qy * dx);
          v[1] = dPower * vx;
          v[2] = dPower * vy;
          v[10] = fabsf(v[1]);
          v[11] = fabsf(v[2]);
          const float hp = 0.5f * dPower;
          v[3] = (hp * vx) * vx;
          v[4] = (hp * vx) * vy;
          v[5] = (hp * vy) * vy;
        }
        Br = fma32(alpha, s.r - Br, Br);
        Bg = fma32(alpha, s.g - Bg, Bg);
        Bb = fma32(alpha, s.b - Bb, Bb);
        logBehind = logBehind + logRest;
      }
      if (!__any(active)) continue;
      halve<16>(v, lane, 32);
      halve<8>(v, lane, 16);
      halve<4>(v, lane, 8);
      halve<2>(v, lane, 4);
      v[0] += __shfl_xor(v[0], 2, 64);
      v[0] += __shfl_xor(v[0], 1, 64);
      const uint32_t value = lane >> 2;
      if ((lane & 3u) == 0 && value < kRawPerSplat) {
        const size_t slot = static_cast<size_t>(batchIndex[k]) * kRawPerSplat + value;
        if constexpr (Fixed) {
          const int32_t q = quantise(v[0], fixedScale);
          if (q != 0) atomicAdd(&rawFixed[slot], static_cast<unsigned long long>(static_cast<long long>(q)));
        } else {
          if (v[0] != 0.0f) atomicAdd(&rawFloat[slot], v[0]);
        }
      }
    }
  }
}

} // namespace

void launch_forward(hipStream_t stream, const Splat *splats, const uint32_t *entryValues, const uint2 *ranges, uint32_t width,
                    uint32_t height, const float background[3], float4 *image, uint32_t *stops) {
  const uint32_t gridW = (width + kTileSize - 1) / kTileSize, gridH = (height + kTileSize - 1) / kTileSize;
  hipLaunchKernelGGL(forward_kernel, dim3(gridW * gridH), dim3(kTilePixels), 0, stream, splats, entryValues, ranges, width, height,
                     gridW, background[0], background[1], background[2], image, stops);
}

void launch_backward(hipStream_t stream, bool fixedPoint, const Splat *splats, const uint32_t *entryValues, const uint2 *ranges,
                     uint32_t width, uint32_t height, const float background[3], const float4 *image, const float4 *adjoint,
                     const uint32_t *stops, float fixedScale, int64_t *rawFixed, float *rawFloat) {
  const uint32_t gridW = (width + kTileSize - 1) / kTileSize, gridH = (height + kTileSize - 1) / kTileSize;
  if (fixedPoint) {
    hipLaunchKernelGGL(backward_kernel<true>, dim3(gridW * gridH), dim3(kTilePixels), 0, stream, splats, entryValues, ranges, width,
                       height, gridW, background[0], background[1], background[2], image, adjoint, stops, fixedScale,
                       reinterpret_cast<unsigned long long *>(rawFixed), rawFloat);
  } else {
    hipLaunchKernelGGL(backward_kernel<false>, dim3(gridW * gridH), dim3(kTilePixels), 0, stream, splats, entryValues, ranges, width,
                       height, gridW, background[0], background[1], background[2], image, adjoint, stops, fixedScale,
                       reinterpret_cast<unsigned long long *>(rawFixed), rawFloat);
  }
}

} // namespace vsplat
