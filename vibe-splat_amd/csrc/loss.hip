// Loss and its derivative per pixel, L = (1 - lambda) L1 + lambda (1 - SSIM)
// (train_ssim.comp:42-123, train_loss.comp:52-113). The 11x11 Gaussian window
// (sigma 1.5) is applied separably: a workgroup takes a strip 16 columns wide
// and TH rows tall, a horizontal pass over its 16 columns and TH + 10 rows (5
// rows of halo above and below) into LDS, then a vertical pass out of it; the
// taller the strip, the fewer halo rows are worked twice (26 per 16 output rows
// at TH = 16, 74 per 64 at TH = 64). The window is cut at the image border and
// renormalised there, as in the reference.
//
// Pass 1 writes per pixel the three fields P = dS/dmu - 2 mu dS/dvar - mu_y dS/dcov,
// Q = dS/dvar, R = dS/dcov (each over the window's weight), 9 floats; pass 2
// applies the same window to them to get dSSIM/dimage, and adds the L1 term.
// Each pass also leaves the total (SSIM, L1) of every 16x16 tile of its strip,
// summed in a fixed tree over the tile's 256 pixels whatever the strip height.
#include "device.hpp"
#include "kernels.hpp"

namespace vsplat {

namespace {

__constant__ float kWeights[11] = {0.00102838f, 0.00759876f, 0.03600077f, 0.10936069f, 0.21300554f, 0.26601172f,
                                   0.21300554f, 0.10936069f, 0.03600077f, 0.00759876f, 0.00102838f};
// The same weights as compile-time constants, for the windows that lie inside
// the image: their loops unroll with the weights as immediates and the taps'
// loads at fixed offsets, adding in the same order as the general loops.
constexpr float kTaps[11] = {0.00102838f, 0.00759876f, 0.03600077f, 0.10936069f, 0.21300554f, 0.26601172f,
                             0.21300554f, 0.10936069f, 0.03600077f, 0.00759876f, 0.00102838f};
constexpr int kHalo = 5;

// The sum of the 11 weights, added in tap order (window_norm of a window inside the image).
__device__ __forceinline__ float full_norm() {
  float norm = 0.0f;
#pragma unroll
  for (int k = 0; k < 11; ++k) norm = norm + kTaps[k];
  return norm;
}
// Strip heights: the SSIM pass's 73 registers allow six waves per SIMD, which
// 32 rows keep (8 waves and 40 KB of LDS a workgroup); the loss pass's 44 allow
// eight at 64 rows (16 waves, 43 KB).
constexpr int kSsimRows = 32, kLossRows = 64;
template <int TH> constexpr int staged_pixels() { return (TH + 2 * kHalo) * 16; }

__device__ __forceinline__ float window_norm(int at, int size) {
  float norm = 0.0f;
  const int lo = at - 5 > 0 ? at - 5 : 0, hi = at + 5 < size - 1 ? at + 5 : size - 1;
  for (int i = lo; i <= hi; ++i) norm = norm + kWeights[i - at + 5];
  return norm;
}

__device__ __forceinline__ float unorm8(uint32_t v) { return static_cast<float>(v) * (1.0f / 255.0f); }

// The total of one value per thread over each group of 256 threads (a 16x16
// tile): xor butterfly within each wave (lane l plus lane l ^ m, m = 32 ... 1),
// then the group's four waves in order.
__device__ __forceinline__ float tile_sum(float v, float *waveTotals) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v = v + __shfl_xor(v, m, 64);
  if ((threadIdx.x & 63u) == 0) waveTotals[threadIdx.x >> 6] = v;
  __syncthreads();
  const float *w = waveTotals + (threadIdx.x >> 8) * 4u;
  const float total = ((w[0] + w[1]) + w[2]) + w[3];
  __syncthreads();
  return total;
}

// The 16x16 tile that thread t of the strip at (x0, y0) sums into, when that
// tile lies in the image (its thread t & 255 == 0 stores it); else -1. Thread t
// works pixel (x0 + (t & 15), y0 + (t >> 4)).
__device__ __forceinline__ int strip_tile(int x0, int y0, int H, uint32_t gridW) {
  const int ty = y0 / 16 + static_cast<int>(threadIdx.x >> 8);
  return ty * 16 < H ? ty * static_cast<int>(gridW) + x0 / 16 : -1;
}

template <int TH>
__global__ __launch_bounds__(16 * TH) void ssim_kernel(const float4 *__restrict__ image, const uchar4 *__restrict__ target,
                                                       uint32_t width, uint32_t height, uint32_t gridW, float *__restrict__ fields,
                                                       float2 *__restrict__ tileTotals) {
  constexpr int kThreads = 16 * TH, kStaged = staged_pixels<TH>();
  __shared__ float h[15][kStaged];
  __shared__ float waveTotals[kThreads / 64];
  const int W = static_cast<int>(width), H = static_cast<int>(height);
  const int x0 = static_cast<int>(blockIdx.x % gridW) * 16, y0 = static_cast<int>(blockIdx.x / gridW) * TH;
  for (int i = static_cast<int>(threadIdx.x); i < kStaged; i += kThreads) {
    const int row = i / 16, col = i % 16;
    const int y = y0 + row - kHalo, x = x0 + col;
    float acc[15];
#pragma unroll
    for (int q = 0; q < 15; ++q) acc[q] = 0.0f;
    if (y >= 0 && y < H && x < W) {
      // One tap's five sums per channel: mean, squares and product.
      const auto tap = [&acc](float w, const float4 a4, const uchar4 b4) {
        const float a[3] = {a4.x, a4.y, a4.z};
        const float b[3] = {unorm8(b4.x), unorm8(b4.y), unorm8(b4.z)};
#pragma unroll
        for (int ch = 0; ch < 3; ++ch) {
          const float wa = w * a[ch], wb = w * b[ch];
          acc[ch] = acc[ch] + wa;
          acc[3 + ch] = acc[3 + ch] + wb;
          acc[6 + ch] = fma32(wa, a[ch], acc[6 + ch]);
          acc[9 + ch] = fma32(wb, b[ch], acc[9 + ch]);
          acc[12 + ch] = fma32(wa, b[ch], acc[12 + ch]);
        }
      };
      float norm = 0.0f;
      if (x >= 5 && x + 5 < W) {
        const size_t row = static_cast<size_t>(y) * W + (x - 5);
#pragma unroll
        for (int k = 0; k < 11; ++k) tap(kTaps[k], image[row + k], target[row + k]);
        norm = full_norm();
      } else {
        const int u0 = x - 5 > 0 ? x - 5 : 0, u1 = x + 5 < W - 1 ? x + 5 : W - 1;
        for (int u = u0; u <= u1; ++u) {
          const float w = kWeights[u - x + 5];
          const size_t at = static_cast<size_t>(y) * W + u;
          tap(w, image[at], target[at]);
          norm = norm + w;
        }
      }
      const float inorm = 1.0f / norm;
#pragma unroll
      for (int q = 0; q < 15; ++q) acc[q] = acc[q] * inorm;
    }
#pragma unroll
    for (int q = 0; q < 15; ++q) h[q][i] = acc[q];
  }
  __syncthreads();

  const int lx = static_cast<int>(threadIdx.x & 15u), px = x0 + lx, py = y0 + static_cast<int>(threadIdx.x >> 4);
  float own = 0.0f;
  if (px < W && py < H) {
    const float rowNorm = window_norm(py, H);
    const float irow = 1.0f / rowNorm;
    float acc[15];
#pragma unroll
    for (int q = 0; q < 15; ++q) acc[q] = 0.0f;
    if (py >= 5 && py + 5 < H) {
      const int idx = (py - y0) * 16 + lx; // the window's top row (v = py - 5) in the stage
#pragma unroll
      for (int k = 0; k < 11; ++k)
#pragma unroll
        for (int q = 0; q < 15; ++q) acc[q] = fma32(kTaps[k], h[q][idx + k * 16], acc[q]);
    } else {
      const int v0 = py - 5 > 0 ? py - 5 : 0, v1 = py + 5 < H - 1 ? py + 5 : H - 1;
      for (int v = v0; v <= v1; ++v) {
        const float w = kWeights[v - py + 5];
        const int idx = (v - y0 + kHalo) * 16 + lx;
#pragma unroll
        for (int q = 0; q < 15; ++q) acc[q] = fma32(w, h[q][idx], acc[q]);
      }
    }
    const float inorm = 1.0f / (rowNorm * window_norm(px, W));
    const size_t at = static_cast<size_t>(py) * W + px;
    const float C1 = 0.0001f, C2 = 0.0009f;
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) {
      const float mx = acc[ch] * irow, my = acc[3 + ch] * irow;
      const float vx = fma32(-mx, mx, acc[6 + ch] * irow), vy = fma32(-my, my, acc[9 + ch] * irow),
                  cxy = fma32(-mx, my, acc[12 + ch] * irow);
      const float n1 = fma32(2.0f * mx, my, C1), n2 = fma32(2.0f, cxy, C2);
      const float d1 = fma32(mx, mx, fma32(my, my, C1)), d2 = (vx + vy) + C2;
      const float dd = d1 * d2;
      // One reciprocal per denominator instead of a division per use.
      const float idd = 1.0f / dd, id1 = 1.0f / d1, id2 = 1.0f / d2;
      const float s = (n1 * n2) * idd;
      own = ch == 0 ? s : own + s;
      // n1 and n2 are cancelled by hand: n2 can be zero where the derivative is not.
      const float dMu = ((2.0f * my) * n2) * idd - ((2.0f * s) * mx) * id1;
      const float dVar = -s * id2;
      const float dCov = (2.0f * n1) * idd;
      fields[at * 9 + ch] = fma32(-my, dCov, fma32(-2.0f * mx, dVar, dMu)) * inorm;
      fields[at * 9 + 3 + ch] = dVar * inorm;
      fields[at * 9 + 6 + ch] = dCov * inorm;
    }
  }
  const float total = tile_sum(own, waveTotals);
  const int tile = strip_tile(x0, y0, H, gridW);
  if ((threadIdx.x & 255u) == 0 && tile >= 0) tileTotals[tile].x = total;
}

// The target's luminance at a pixel, clamped to the image (train_loss.comp:34-40):
// 0.299 r + 0.587 g + 0.114 b, added left to right.
__device__ __forceinline__ float target_luminance(const uchar4 *__restrict__ target, int x, int y, int W, int H) {
  x = x < 0 ? 0 : (x > W - 1 ? W - 1 : x);
  y = y < 0 ? 0 : (y > H - 1 ? H - 1 : y);
  const uchar4 t = target[static_cast<size_t>(y) * W + x];
  return (unorm8(t.x) * 0.299f + unorm8(t.y) * 0.587f) + unorm8(t.z) * 0.114f;
}

// The edge weight of the edge-aware score (train_loss.comp:42-51): |Laplacian
// of the target's luminance| / 4 in [0, 0.999], as differences from the centre
// so that a flat patch gives exactly zero.
__device__ __forceinline__ float target_edge(const uchar4 *__restrict__ target, int x, int y, int W, int H) {
  const float c = target_luminance(target, x, y, W, H);
  const float laplacian = (((target_luminance(target, x - 1, y, W, H) - c) + (target_luminance(target, x + 1, y, W, H) - c)) +
                           (target_luminance(target, x, y - 1, W, H) - c)) +
                          (target_luminance(target, x, y + 1, W, H) - c);
  const float e = fabsf(laplacian) / 4.0f;
  return e < 0.0f ? 0.0f : (e > 0.999f ? 0.999f : e);
}

// When scoring (errorMax not null): the adjoint's fourth value is the pixel's
// mean error, errorMax the worst over the view in 1/65535 (one atomic per
// wave; a maximum does not depend on the order), and edge the target's edge
// weight (train_loss.comp:106-110). Otherwise the fourth value is 0.
template <int TH>
__global__ __launch_bounds__(16 * TH) void loss_kernel(const float4 *__restrict__ image, const uchar4 *__restrict__ target,
                                                       uint32_t width, uint32_t height, uint32_t gridW, float lambda,
                                                       const float *__restrict__ fields, float4 *__restrict__ adjoint,
                                                       float2 *__restrict__ tileTotals, uint32_t *__restrict__ errorMax,
                                                       float *__restrict__ edge) {
  constexpr int kThreads = 16 * TH, kStaged = staged_pixels<TH>();
  __shared__ float h[9][kStaged];
  __shared__ float waveTotals[kThreads / 64];
  const int W = static_cast<int>(width), H = static_cast<int>(height);
  const int x0 = static_cast<int>(blockIdx.x % gridW) * 16, y0 = static_cast<int>(blockIdx.x / gridW) * TH;
  for (int i = static_cast<int>(threadIdx.x); i < kStaged; i += kThreads) {
    const int row = i / 16, col = i % 16;
    const int y = y0 + row - kHalo, x = x0 + col;
    float acc[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) acc[q] = 0.0f;
    if (y >= 0 && y < H && x < W) {
      if (x >= 5 && x + 5 < W) {
        const float *f = fields + (static_cast<size_t>(y) * W + (x - 5)) * 9;
#pragma unroll
        for (int k = 0; k < 11; ++k)
#pragma unroll
          for (int q = 0; q < 9; ++q) acc[q] = fma32(kTaps[k], f[k * 9 + q], acc[q]);
      } else {
        const int u0 = x - 5 > 0 ? x - 5 : 0, u1 = x + 5 < W - 1 ? x + 5 : W - 1;
        for (int u = u0; u <= u1; ++u) {
          const float w = kWeights[u - x + 5];
          const float *f = fields + (static_cast<size_t>(y) * W + u) * 9;
#pragma unroll
          for (int q = 0; q < 9; ++q) acc[q] = fma32(w, f[q], acc[q]);
        }
      }
    }
#pragma unroll
    for (int q = 0; q < 9; ++q) h[q][i] = acc[q];
  }
  __syncthreads();

  const int lx = static_cast<int>(threadIdx.x & 15u), px = x0 + lx, py = y0 + static_cast<int>(threadIdx.x >> 4);
  float own = 0.0f;
  uint32_t worst = 0;
  if (px < W && py < H) {
    float acc[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) acc[q] = 0.0f;
    if (py >= 5 && py + 5 < H) {
      const int idx = (py - y0) * 16 + lx; // the window's top row (v = py - 5) in the stage
#pragma unroll
      for (int k = 0; k < 11; ++k)
#pragma unroll
        for (int q = 0; q < 9; ++q) acc[q] = fma32(kTaps[k], h[q][idx + k * 16], acc[q]);
    } else {
      const int v0 = py - 5 > 0 ? py - 5 : 0, v1 = py + 5 < H - 1 ? py + 5 : H - 1;
      for (int v = v0; v <= v1; ++v) {
        const float w = kWeights[v - py + 5];
        const int idx = (v - y0 + kHalo) * 16 + lx;
#pragma unroll
        for (int q = 0; q < 9; ++q) acc[q] = fma32(w, h[q][idx], acc[q]);
      }
    }
    const size_t at = static_cast<size_t>(py) * W + px;
    const float perValue = 1.0f / (3.0f * static_cast<float>(W) * static_cast<float>(H));
    const float keep = 1.0f - lambda;
    const float4 x4 = image[at];
    const uchar4 y4 = target[at];
    const float xs[3] = {x4.x, x4.y, x4.z};
    const float ys[3] = {unorm8(y4.x), unorm8(y4.y), unorm8(y4.z)};
    float out[3];
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) {
      const float d = xs[ch] - ys[ch];
      own = ch == 0 ? fabsf(d) : own + fabsf(d);
      const float sgn = d > 0.0f ? 1.0f : (d < 0.0f ? -1.0f : 0.0f);
      const float inner = fma32(ys[ch], acc[6 + ch], fma32(2.0f * xs[ch], acc[3 + ch], acc[ch]));
      out[ch] = fma32(-lambda, inner, keep * sgn) * perValue;
    }
    float error = 0.0f;
    if (errorMax != nullptr) {
      error = own / 3.0f;
      worst = static_cast<uint32_t>((error < 0.0f ? 0.0f : (error > 1.0f ? 1.0f : error)) * 65535.0f);
      edge[at] = target_edge(target, px, py, W, H);
    }
    adjoint[at] = make_float4(out[0], out[1], out[2], error);
  }
  if (errorMax != nullptr) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
      const uint32_t other = static_cast<uint32_t>(__shfl_xor(static_cast<int>(worst), m, 64));
      worst = other > worst ? other : worst;
    }
    if ((threadIdx.x & 63u) == 0 && worst != 0) atomicMax(errorMax, worst);
  }
  const float total = tile_sum(own, waveTotals);
  const int tile = strip_tile(x0, y0, H, gridW);
  if ((threadIdx.x & 255u) == 0 && tile >= 0) tileTotals[tile].y = total;
}

// The exposed image (train_exposure.comp:17-24): rgb' = T (rgb, 1), each row's
// dot product added left to right; the fourth value (log T) copied.
__global__ __launch_bounds__(256) void exposure_kernel(const float4 *__restrict__ image, float4 *__restrict__ exposed, uint32_t pixels,
                                                       Exposure t) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= pixels) return;
  const float4 c = image[i];
  float out[3];
#pragma unroll
  for (int row = 0; row < 3; ++row) {
    const float *a = t.transform + 4 * row;
    out[row] = ((a[0] * c.x + a[1] * c.y) + a[2] * c.z) + a[3];
  }
  exposed[i] = make_float4(out[0], out[1], out[2], c.w);
}

// The adjoint taken back through the exposure, in place, and per workgroup of
// 256 pixels the twelve sums dL/dT[row][column] = g[row] (rgb, 1)[column]
// (train_exposure_backward.comp:22-40), each summed in the tile_sum tree.
__global__ __launch_bounds__(256) void exposure_backward_kernel(const float4 *__restrict__ image, float4 *__restrict__ adjoint,
                                                                uint32_t pixels, Exposure t, float *__restrict__ groupTotals) {
  __shared__ float waveTotals[4];
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  float g[3] = {0.0f, 0.0f, 0.0f}, rgb1[4] = {0.0f, 0.0f, 0.0f, 0.0f};
  if (i < pixels) {
    const float4 a = adjoint[i];
    const float4 c = image[i];
    g[0] = a.x;
    g[1] = a.y;
    g[2] = a.z;
    rgb1[0] = c.x;
    rgb1[1] = c.y;
    rgb1[2] = c.z;
    rgb1[3] = 1.0f;
    float back[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) back[k] = (g[0] * t.transform[k] + g[1] * t.transform[4 + k]) + g[2] * t.transform[8 + k];
    adjoint[i] = make_float4(back[0], back[1], back[2], a.w);
  }
#pragma unroll 1
  for (int row = 0; row < 3; ++row) {
#pragma unroll 1
    for (int column = 0; column < 4; ++column) {
      const float total = tile_sum(g[row] * rgb1[column], waveTotals);
      if (threadIdx.x == 0) groupTotals[blockIdx.x * 12u + row * 4u + column] = total;
    }
  }
}

// When scoring, the adjoint's fourth value becomes the mark of a high-error
// pixel (train_error_mask.comp:14-19): 1 where its error is at least `share`
// of the view's worst, else 0.
__global__ __launch_bounds__(256) void error_mask_kernel(float4 *__restrict__ adjoint, const uint32_t *__restrict__ errorMax,
                                                         uint32_t pixels, float share) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= pixels) return;
  const uint32_t worst = *errorMax;
  const float limit = (share * static_cast<float>(worst)) / 65535.0f;
  adjoint[i].w = worst > 0u && adjoint[i].w >= limit ? 1.0f : 0.0f;
}

// Sums the tiles' totals in double, in a fixed order: thread t adds tiles t,
// t + 256, ..., then a tree over the threads. sums[0] SSIM, sums[1] L1.
__global__ __launch_bounds__(256) void total_kernel(const float2 *__restrict__ tileTotals, uint32_t tiles, double *__restrict__ sums) {
  __shared__ double s[2][256];
  double ssim = 0.0, l1 = 0.0;
  for (uint32_t t = threadIdx.x; t < tiles; t += 256) {
    ssim += static_cast<double>(tileTotals[t].x);
    l1 += static_cast<double>(tileTotals[t].y);
  }
  s[0][threadIdx.x] = ssim;
  s[1][threadIdx.x] = l1;
  __syncthreads();
  for (uint32_t half = 128; half > 0; half >>= 1) {
    if (threadIdx.x < half) {
      s[0][threadIdx.x] += s[0][threadIdx.x + half];
      s[1][threadIdx.x] += s[1][threadIdx.x + half];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    sums[0] = s[0][0];
    sums[1] = s[1][0];
  }
}

} // namespace

void launch_loss(hipStream_t stream, const float4 *image, const uint8_t *target, uint32_t width, uint32_t height, float lambda,
                 float *fields, float4 *adjoint, float2 *tileTotals, double *sums, uint32_t *errorMax, float *edge) {
  const uint32_t gridW = (width + 15) / 16, tiles = gridW * ((height + 15) / 16);
  const uint32_t ssimStrips = gridW * ((height + kSsimRows - 1) / kSsimRows), lossStrips = gridW * ((height + kLossRows - 1) / kLossRows);
  const uchar4 *t = reinterpret_cast<const uchar4 *>(target);
  hipLaunchKernelGGL(ssim_kernel<kSsimRows>, dim3(ssimStrips), dim3(16 * kSsimRows), 0, stream, image, t, width, height, gridW, fields,
                     tileTotals);
  hipLaunchKernelGGL(loss_kernel<kLossRows>, dim3(lossStrips), dim3(16 * kLossRows), 0, stream, image, t, width, height, gridW, lambda,
                     fields, adjoint, tileTotals, errorMax, edge);
  hipLaunchKernelGGL(total_kernel, dim3(1), dim3(256), 0, stream, tileTotals, tiles, sums);
}

uint32_t exposure_groups(uint32_t pixels) { return (pixels + 255) / 256; }

void launch_exposure(hipStream_t stream, const float4 *image, float4 *exposed, uint32_t pixels, const Exposure &t) {
  if (pixels == 0) return;
  hipLaunchKernelGGL(exposure_kernel, dim3(exposure_groups(pixels)), dim3(256), 0, stream, image, exposed, pixels, t);
}

void launch_exposure_backward(hipStream_t stream, const float4 *image, float4 *adjoint, uint32_t pixels, const Exposure &t,
                              float *groupTotals) {
  if (pixels == 0) return;
  hipLaunchKernelGGL(exposure_backward_kernel, dim3(exposure_groups(pixels)), dim3(256), 0, stream, image, adjoint, pixels, t,
                     groupTotals);
}

void launch_error_mask(hipStream_t stream, float4 *adjoint, const uint32_t *errorMax, uint32_t pixels, float share) {
  if (pixels == 0) return;
  hipLaunchKernelGGL(error_mask_kernel, dim3(exposure_groups(pixels)), dim3(256), 0, stream, adjoint, errorMax, pixels, share);
}

} // namespace vsplat
