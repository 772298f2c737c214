// Device-side definitions shared by the rasterizer's kernels: the per-view
// constant block, the 48-byte projected splat, and the single-precision exp
// and log every kernel uses.
//
// Arithmetic rules (they make the GPU's results reproducible on any IEEE host):
// the code is compiled with -ffp-contract=off, so each + - * / sqrt rounds on
// its own; fused multiply-adds appear only where written as __builtin_fmaf;
// division and sqrt are the correctly rounded ones; exp and log below use only
// those operations. Integer results (sort keys, tile lists, stops, fixed-point
// sums) therefore come out the same on every run and on the CPU model the
// tests hold them against.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace vsplat {

constexpr uint32_t kTileSize = 16;       // train_types.glsl:6
constexpr uint32_t kTilePixels = 256;    // one workgroup of 256 threads per tile
constexpr uint32_t kInvisible = 0xFFFFFFFFu;
constexpr uint32_t kRawPerSplat = 14;    // train_types.glsl:15
// Fixed-point sums (train_backward_fixed.comp:19-21, trainer.cpp:586-591, widened
// to 64 bits): each wave total is quantised at kFixedSteps steps per unit and
// clamped to kFixedClamp steps, the reference's 64 steps and 1e9 times 2^14.
// A unit is 1 / (3 W H) for the loss's adjoint and 1 for the curvature's.
// At most kFixedClamp per wave keeps a Gaussian's sum inside int64 for up to
// 2^19 waves (over 33 million pixels).
constexpr float kFixedSteps = 1048576.0f; // 2^20
constexpr float kFixedClamp = 1.6384e13f; // 1e9 * 2^14, exact in float

// The per-view constants, passed by value as a kernel argument (camera block of
// splat_types.glsl:313-319, with the near plane and the frustum half-tangents
// split out).
struct ViewParams {
  float view[16]; // world to camera, column-major: view[c * 4 + r]
  float position[3];
  float nearPlane;
  float fx, fy, cx, cy;
  float tanHalfX, tanHalfY; // 0.5 W / fx, 0.5 H / fy
  uint32_t width, height;
  uint32_t count, shDegree, shWords, shCoefficients; // active degree and coefficients
  uint32_t shFormat;                                 // kShHalf or kShBf16
  uint32_t tileW, tileH;                             // the raster tile shape (powers of two: 8, 16 or 32 by 8 or 16)
};

// How a Gaussian's SH coefficients are stored, in words of 32 bits:
// kShHalf, the reference trainer's layout (the DC term as three floats, the rest
// as halves two to a word, low half first; splat_types.glsl:299-310), or
// kShBf16, every coefficient as bf16, two to a word, low half first.
constexpr uint32_t kShHalf = 0, kShBf16 = 1;

// The projected splat (train_types.glsl:26-31): conic xx xy yy, opacity, rgb,
// radius, centre, extent. 48 bytes, read as three 16-byte vectors.
struct alignas(16) Splat {
  float qx, qy, qz, o;
  float r, g, b, radius;
  float cx, cy, ex, ey;
};
static_assert(sizeof(Splat) == 48, "48-byte splat");

// Counters the binning stages leave on the device, read by later kernels and
// once per step by the host.
enum Counter : uint32_t {
  kVisible = 0,      // splats in view
  kEntries = 1,      // tile entries this view needs
  kEntriesUsed = 2,  // entries actually listed: kEntries, or 0 if they did not fit
  kBigSplats = 3,    // splats whose tile box is larger than a TileRecord's bits (listed for their own kernels)
  kBigBitWords = 4,  // 64-bit words of the big splats' reached-tile bits handed out so far
  kOverflow = 5,     // sticky: a view needed more entries than the buffers hold; every later pass is skipped until the host clears it
  kSkipped = 6,      // step-mode passes skipped that way since the host last cleared it
  kCounterWords = 8,  // word 7 unused (the readback keeps 8 words, so the sums after them stay 8-byte aligned)
  kClearedPerPass = 5, // the first five are cleared at the start of every pass
};

__device__ __forceinline__ float fma32(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ float as_float(uint32_t u) { return __uint_as_float(u); }
__device__ __forceinline__ uint32_t as_uint(float f) { return __float_as_uint(f); }
__device__ __forceinline__ float minf(float a, float b) { return b < a ? b : a; }
__device__ __forceinline__ float maxf(float a, float b) { return a < b ? b : a; }
__device__ __forceinline__ float clampf(float v, float lo, float hi) { return v < lo ? lo : (v > hi ? hi : v); }

// a / b correctly rounded, for 2^-100 <= a <= 2^60 and 2^-20 <= b <= 1: the
// hardware's IEEE division sequence (reciprocal, one Newton step on it, the
// quotient and two residual corrections by fma) without V_DIV_SCALE and
// V_DIV_FMAS's range scaling and V_DIV_FIXUP's special cases. V_DIV_SCALE
// scales only a denominator that is denormal or has a denormal reciprocal, a
// numerator at or below 2^-103, or a quotient at or above 2^96 or denormal;
// V_DIV_FIXUP changes only NaN, infinite, zero or out-of-range results. In the
// range above none of these occur, so the result is `a / b` bit for bit.
__device__ __forceinline__ float div_in_range(float a, float b) {
  const float r0 = __builtin_amdgcn_rcpf(b);
  const float r = fma32(fma32(-b, r0, 1.0f), r0, r0);
  const float q0 = a * r;
  const float q1 = fma32(fma32(-b, q0, a), r, q0);
  return fma32(fma32(-b, q1, a), r, q1);
}

// e^x: x = n ln2 + r, |r| <= ln2 / 2 (ln2 split so that n ln2_hi is exact),
// e^r = 1 + r + r^2 q(r) with q a degree-4 near-minimax fit of (e^r - 1 - r) / r^2,
// then 2^n applied as two multiplications so that overflow to +inf and gradual
// underflow both round once. Cody-Waite reduction (Muller, Elementary Functions, ch. 11).
__device__ __forceinline__ float vs_exp(float x) {
  if (x != x) return x + x;
  x = x > 89.0f ? 89.0f : x;
  if (x < -104.0f) return 0.0f;
  const float n = __builtin_rintf(x * 0x1.715476p+0f);
  float r = fma32(n, -0x1.62e4p-1f, x);
  r = fma32(n, -0x1.7f7d1cp-20f, r);
  float q = fma32(0x1.6d752p-10f, r, 0x1.120b66p-7f);
  q = fma32(q, r, 0x1.5554b8p-5f);
  q = fma32(q, r, 0x1.5554dcp-3f);
  q = fma32(q, r, 0x1p-1f);
  const float p = fma32(r * r, q, r) + 1.0f;
  const int ni = static_cast<int>(n);
  const int n1 = ni >> 1, n2 = ni - n1;
  return (p * as_float(static_cast<uint32_t>(n1 + 127) << 23)) * as_float(static_cast<uint32_t>(n2 + 127) << 23);
}

// log(x): x = m 2^e with m in [sqrt(1/2), sqrt(2)), f = m - 1 exactly,
// log1p(f) = f + f^2 Q(f) with Q a degree-8 near-minimax fit of (log1p(f) - f) / f^2.
__device__ __forceinline__ float vs_log(float x) {
  if (!(x > 0.0f)) return x == 0.0f ? -__builtin_inff() : (x != x ? x + x : __builtin_nanf(""));
  if (x == __builtin_inff()) return x;
  int eadj = 0;
  if (x < 0x1p-126f) {
    x *= 0x1p23f;
    eadj = -23;
  }
  const uint32_t u = as_uint(x) - 0x3f3504f3u;
  const int e = static_cast<int32_t>(u) >> 23;
  const float m = as_float((u & 0x007fffffu) + 0x3f3504f3u);
  const float f = m - 1.0f;
  float q = fma32(-0x1.4a4f98p-4f, f, 0x1.088dd0p-3f);
  q = fma32(q, f, -0x1.0bba2ep-3f);
  q = fma32(q, f, 0x1.2245e4p-3f);
  q = fma32(q, f, -0x1.548f6ap-3f);
  q = fma32(q, f, 0x1.99a7bap-3f);
  q = fma32(q, f, -0x1.0001f8p-2f);
  q = fma32(q, f, 0x1.555548p-2f);
  q = fma32(q, f, -0x1.fffffep-2f);
  const float lp = fma32(f * f, q, f);
  const float ef = static_cast<float>(e + eadj);
  return fma32(ef, 0x1.62e4p-1f, fma32(ef, 0x1.7f7d1cp-20f, lp));
}

// vs_exp on [-87, 87] without its special cases, for arguments that are never
// NaN: the argument is clamped into that range (one med3), where the result
// p 2^n is a normal float (n >= -126 with p > 1 at the low end), so ldexp
// applies 2^n exactly, as vs_exp's two multiplications do. Equal to vs_exp(x)
// for every x in [-87, 87].
__device__ __forceinline__ float exp_finite(float x) {
  x = __builtin_amdgcn_fmed3f(x, -87.0f, 87.0f);
  const float n = __builtin_rintf(x * 0x1.715476p+0f);
  float r = fma32(n, -0x1.62e4p-1f, x);
  r = fma32(n, -0x1.7f7d1cp-20f, r);
  float q = fma32(0x1.6d752p-10f, r, 0x1.120b66p-7f);
  q = fma32(q, r, 0x1.5554b8p-5f);
  q = fma32(q, r, 0x1.5554dcp-3f);
  q = fma32(q, r, 0x1p-1f);
  const float p = fma32(r * r, q, r) + 1.0f;
  return __builtin_ldexpf(p, static_cast<int>(n));
}

// vs_log for a positive normal finite x, without its special cases; equal to
// vs_log(x) there. The pixel loops take it of 1 - alpha, in [0.01, 1].
__device__ __forceinline__ float log_normal(float x) {
  const uint32_t u = as_uint(x) - 0x3f3504f3u;
  const int e = static_cast<int32_t>(u) >> 23;
  const float m = as_float((u & 0x007fffffu) + 0x3f3504f3u);
  const float f = m - 1.0f;
  float q = fma32(-0x1.4a4f98p-4f, f, 0x1.088dd0p-3f);
  q = fma32(q, f, -0x1.0bba2ep-3f);
  q = fma32(q, f, 0x1.2245e4p-3f);
  q = fma32(q, f, -0x1.548f6ap-3f);
  q = fma32(q, f, 0x1.99a7bap-3f);
  q = fma32(q, f, -0x1.0001f8p-2f);
  q = fma32(q, f, 0x1.555548p-2f);
  q = fma32(q, f, -0x1.fffffep-2f);
  const float lp = fma32(f * f, q, f);
  const float ef = static_cast<float>(e);
  return fma32(ef, 0x1.62e4p-1f, fma32(ef, 0x1.7f7d1cp-20f, lp));
}

// IEEE binary16 to float: the hardware conversion, which is exact (every half,
// denormals included, is a float).
__device__ __forceinline__ float half_to_float(uint32_t h) {
  return static_cast<float>(__builtin_bit_cast(_Float16, static_cast<uint16_t>(h)));
}

// Float to binary16, round to nearest even; overflow to infinity.
__device__ __forceinline__ uint32_t float_to_half(float f) {
  const uint32_t u = as_uint(f), sign = (u >> 16) & 0x8000u;
  const uint32_t a = u & 0x7FFFFFFFu;
  if (a >= 0x7F800000u) return sign | 0x7C00u | (a > 0x7F800000u ? 0x200u | ((a >> 13) & 0x3FFu) : 0u);
  if (a >= 0x477FF000u) return sign | 0x7C00u;
  if (a < 0x38800000u) return sign | static_cast<uint32_t>(__builtin_rintf(as_float(a) * 0x1p24f));
  const uint32_t rounded = a + 0x0FFFu + ((a >> 13) & 1u);
  return sign | ((rounded - 0x38000000u) >> 13);
}

// Element j of a Gaussian's SH (trainer_sh_value, splat_types.glsl:299-310, for
// kShHalf). The format is a template argument: the kernels that read every
// coefficient are compiled once per format.
template <uint32_t Format>
__device__ __forceinline__ float sh_value(const uint32_t *words, uint32_t base, uint32_t j) {
  if constexpr (Format == kShBf16) {
    const uint32_t w = words[base + j / 2];
    return as_float((j & 1u) ? (w & 0xFFFF0000u) : (w << 16));
  } else {
    if (j < 3) return as_float(words[base + j]);
    const uint32_t w = words[base + 3 + (j - 3) / 2];
    return half_to_float(((j - 3) & 1u) ? (w >> 16) : (w & 0xFFFFu));
  }
}

// Copies `words` consecutive 32-bit words from global memory to LDS with
// 16-byte loads, the block's threads striding over them (both ends 16-byte
// aligned); a tail of fewer than four words goes word by word. Neighbouring
// threads read neighbouring vectors, which is what keeps the address and data
// units from touching a cache line once per lane when each thread's data is a
// record of its own.
template <uint32_t Threads>
__device__ __forceinline__ void stage_words(uint32_t *lds, const uint32_t *src, uint32_t words) {
  const uint32_t vectors = words / 4;
  const uint4 *from = reinterpret_cast<const uint4 *>(src);
  uint4 *to = reinterpret_cast<uint4 *>(lds);
  for (uint32_t v = threadIdx.x; v < vectors; v += Threads) to[v] = from[v];
  const uint32_t tail = vectors * 4 + threadIdx.x;
  if (tail < words) lds[tail] = src[tail];
}

// World to camera, fused in the order x, y, z then the translation last.
__device__ __forceinline__ void to_camera(const ViewParams &c, float px, float py, float pz, float &tx, float &ty, float &tz) {
  const float *V = c.view;
  tx = fma32(V[8], pz, fma32(V[4], py, fma32(V[0], px, V[12])));
  ty = fma32(V[9], pz, fma32(V[5], py, fma32(V[1], px, V[13])));
  tz = fma32(V[10], pz, fma32(V[6], py, fma32(V[2], px, V[14])));
}

// Rotation matrix of a unit quaternion (w, x, y, z) as R[row][col] (train_types.glsl:172-177).
__device__ __forceinline__ void quat_to_mat(float w, float x, float y, float z, float R[3][3]) {
  const float xx = x * x, yy = y * y, zz = z * z, xy = x * y, xz = x * z, yz = y * z, wx = w * x, wy = w * y, wz = w * z;
  R[0][0] = 1.0f - 2.0f * (yy + zz);
  R[0][1] = 2.0f * (xy - wz);
  R[0][2] = 2.0f * (xz + wy);
  R[1][0] = 2.0f * (xy + wz);
  R[1][1] = 1.0f - 2.0f * (xx + zz);
  R[1][2] = 2.0f * (yz - wx);
  R[2][0] = 2.0f * (xz - wy);
  R[2][1] = 2.0f * (yz + wx);
  R[2][2] = 1.0f - 2.0f * (xx + yy);
}

// a . b over three terms, fused in the order 0, 1, 2.
__device__ __forceinline__ float dot3(float a0, float a1, float a2, float b0, float b1, float b2) {
  return fma32(a2, b2, fma32(a1, b1, a0 * b0));
}

// The 3D smoothing filter's opacity factor sqrt(prod s2 / (s2 + f)) (train_preprocess.comp:62-63).
// With no filter each ratio s2 / (s2 + 0) is exactly 1 for a finite positive s2
// and NaN for 0 or infinity, which the shortcut returns without dividing.
__device__ __forceinline__ float filter_factor(float s2x, float s2y, float s2z, float f) {
  if (f == 0.0f) {
    const float inf = __builtin_inff();
    const bool finite = s2x > 0.0f && s2x < inf && s2y > 0.0f && s2y < inf && s2z > 0.0f && s2z < inf;
    return finite ? 1.0f : __builtin_nanf("");
  }
  return sqrtf((s2x / (s2x + f)) * (s2y / (s2y + f)) * (s2z / (s2z + f)));
}

// The real SH basis up to degree 3 in single precision (train_types.glsl:97-123).
__device__ __forceinline__ void sh_basis(float x, float y, float z, uint32_t degree, float b[16]) {
  const float C1 = 0.4886025119029199f;
  const float xx = x * x, yy = y * y, zz = z * z;
#pragma unroll
  for (int k = 0; k < 16; ++k) b[k] = 0.0f;
  b[0] = 0.28209479177387814f;
  if (degree > 0) {
    b[1] = -C1 * y;
    b[2] = C1 * z;
    b[3] = -C1 * x;
  }
  if (degree > 1) {
    b[4] = 1.0925484305920792f * x * y;
    b[5] = -1.0925484305920792f * y * z;
    b[6] = 0.31539156525252005f * (2.0f * zz - xx - yy);
    b[7] = -1.0925484305920792f * x * z;
    b[8] = 0.5462742152960396f * (xx - yy);
  }
  if (degree > 2) {
    b[9] = -0.5900435899266435f * y * (3.0f * xx - yy);
    b[10] = 2.890611442640554f * x * y * z;
    b[11] = -0.4570457994644658f * y * (4.0f * zz - xx - yy);
    b[12] = 0.3731763325901154f * z * (2.0f * zz - 3.0f * xx - 3.0f * yy);
    b[13] = -0.4570457994644658f * x * (4.0f * zz - xx - yy);
    b[14] = 1.445305721320277f * z * (xx - yy);
    b[15] = -0.5900435899266435f * x * (xx - 3.0f * yy);
  }
}

// PCG's output permutation as a hash to [0, 1) (train_types.glsl:156-160).
__device__ __forceinline__ float train_uniform(uint32_t index, uint32_t seed) {
  const uint32_t state = (index ^ (seed * 0x9E3779B9u)) * 747796405u + 2891336453u;
  const uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
  return static_cast<float>((word >> 22u) ^ word) * 0x1p-32f;
}

// SH dropout (train_types.glsl:165-170): how many coefficients Gaussian `index` uses.
__device__ __forceinline__ uint32_t sh_coefficients_used(uint32_t index, uint32_t coefficients, float dropout, uint32_t retained,
                                                         uint32_t seed) {
  if (dropout > 0.0f && train_uniform(index, seed) < dropout) {
    const uint32_t keep = (retained + 1u) * (retained + 1u);
    return coefficients < keep ? coefficients : keep;
  }
  return coefficients;
}

// 2 log(255 o), the squared reach of splat_at's 1/255 cut-off (train_types.glsl:51-58).
__device__ __forceinline__ float reach_of(float opacity) {
  const float r = 2.0f * vs_log(255.0f * opacity);
  return r > 0.0f ? r : 0.0f;
}

__device__ __forceinline__ Splat load_splat(const Splat *splats, uint32_t i) {
  const float4 *p = reinterpret_cast<const float4 *>(splats + i);
  const float4 a = p[0], b = p[1], c = p[2];
  return Splat{a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w, c.x, c.y, c.z, c.w};
}

// The pixels a splat can reach before any clipping, half-open: columns
// [ax, bx), rows [ay, by), from the extent box grown by half a pixel
// (train_types.glsl:43-48).
struct PixelBox {
  uint32_t ax, ay, bx, by;
};
__device__ __forceinline__ PixelBox splat_pixel_box(float cx, float cy, float ex, float ey) {
  const float lx = clampf(cx - ex - 0.5f, 0.0f, 16777216.0f), ly = clampf(cy - ey - 0.5f, 0.0f, 16777216.0f);
  const float hx = clampf(cx + ex + 0.5f, 0.0f, 16777216.0f), hy = clampf(cy + ey + 0.5f, 0.0f, 16777216.0f);
  return PixelBox{static_cast<uint32_t>(lx), static_cast<uint32_t>(ly), static_cast<uint32_t>(hx) + 1u, static_cast<uint32_t>(hy) + 1u};
}

// The pixels of the tile at (x0, y0, x1, y1) that a splat can reach, half-open,
// tile-local, packed as four bytes x0 | y0 << 8 | x1 << 16 | y1 << 24 (train_types.glsl:43-48).
__device__ __forceinline__ uint32_t splat_box(const Splat &s, uint32_t x0, uint32_t y0, uint32_t x1, uint32_t y1) {
  const PixelBox p = splat_pixel_box(s.cx, s.cy, s.ex, s.ey);
  uint32_t ax = p.ax, ay = p.ay, bx = p.bx, by = p.by;
  ax = ax > x0 ? ax : x0;
  ay = ay > y0 ? ay : y0;
  bx = bx < x1 ? bx : x1;
  by = by < y1 ? by : y1;
  // Empty boxes become 0..0; everything else is within the tile, so 0..16 fits a byte.
  if (ax >= bx || ay >= by) return 0u;
  return (ax - x0) | ((ay - y0) << 8) | ((bx - x0) << 16) | ((by - y0) << 24);
}

__device__ __forceinline__ bool in_box(uint32_t box, uint32_t lx, uint32_t ly) {
  return lx >= (box & 0xFFu) && ly >= ((box >> 8) & 0xFFu) && lx < ((box >> 16) & 0xFFu) && ly < (box >> 24);
}

constexpr float kMinAlpha = 1.0f / 255.0f;
constexpr float kMinTransmittance = 1e-4f; // train_types.glsl:11 holds its log, LOG_MIN_TRANSMITTANCE

} // namespace vsplat
