// Test infrastructure, not product code. Oracle B: a single-precision model of
// this project's GPU pipeline, written as plain sequential C++ and kept
// operation for operation in step with the HIP kernels, so that everything the
// GPU computes can be compared bit for bit. Where the GPU's result depends on
// how work is split (the fixed-point backward quantises one total per wave of
// 64 pixels), the model reproduces that split exactly.
//
// The algorithm follows the reference's training shaders; each function names
// the shader lines it restates. Its arithmetic (operation order, fused
// multiply-adds, the exp and log of fp32_math.hpp) is this project's own
// choice, and is tied to the FP64 restatement of the reference (oracle_a.cpp)
// by the tolerance tests in tests/test_oracle_b.py.
//
// The hash-driven stochastic roundings (the bf16 second moment, bf16.glsl:16-23,
// and the fp16 SH after a step, train_step.glsl:70-80) are pinned by
// tests/test_rounding.py: the hash exactly, by inverting it through the
// published PCG constants, and each rounding rule by its unbiasedness over
// 2^18 draws per value (a truncating or a round-to-nearest copy fails). The
// reference holds no vectors of its own for them: the reference GPU's bit
// patterns themselves are not available to compare with.
#include "fp32_math.hpp"
#include "oracle_api.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <vector>

namespace oracle {
namespace {

using std::fma;

constexpr std::uint32_t kTile = 16;
constexpr std::uint32_t kInvisible = 0xFFFFFFFFu;
constexpr float kMinAlpha = 1.0f / 255.0f;            // train_types.glsl:92-93
constexpr float kMinTransmittance = 1e-4f; // train_types.glsl:11 holds its log
constexpr int kRaw = 14;                              // train_types.glsl:15

// The splat record, 12 floats: conic xx xy yy, opacity, rgb, radius, centre xy, extent xy
// (train_types.glsl:26-31).
struct Splat {
  float qx, qy, qz, o, r, g, b, radius, cx, cy, ex, ey;
};
static_assert(sizeof(Splat) == 48, "48 byte splat");

Splat load_splat(const float *s, std::uint32_t i) {
  Splat out;
  std::memcpy(&out, s + static_cast<std::size_t>(i) * 12, sizeof(Splat));
  return out;
}

float clampf(float v, float lo, float hi) { return v < lo ? lo : (v > hi ? hi : v); }
float minf(float a, float b) { return b < a ? b : a; }
float maxf(float a, float b) { return a < b ? b : a; }

// Half float <-> float, round to nearest even (IEEE binary16).
float half_to_float(std::uint32_t h) {
  const std::uint32_t sign = (h & 0x8000u) << 16, e = (h >> 10) & 0x1Fu, m = h & 0x3FFu;
  if (e == 0) {
    const float v = static_cast<float>(m) * 0x1p-24f;
    return sign ? -v : v;
  }
  if (e == 31) return bits_to_float(sign | 0x7F800000u | (m << 13));
  return bits_to_float(sign | ((e + 112u) << 23) | (m << 13));
}
std::uint32_t float_to_half(float f) {
  const std::uint32_t u = float_to_bits(f), sign = (u >> 16) & 0x8000u;
  const std::uint32_t a = u & 0x7FFFFFFFu;
  if (a >= 0x7F800000u) return sign | 0x7C00u | (a > 0x7F800000u ? 0x200u | ((a >> 13) & 0x3FFu) : 0u);
  if (a >= 0x477FF000u) return sign | 0x7C00u; // rounds to beyond the largest half
  if (a < 0x38800000u) {                      // below the smallest normal half: a denormal
    const float scaled = bits_to_float(a) * 0x1p24f; // exact: a is at least 2^-150 apart
    const std::uint32_t m = static_cast<std::uint32_t>(std::rint(scaled));
    return sign | m;
  }
  const std::uint32_t rounded = a + 0x0FFFu + ((a >> 13) & 1u);
  return sign | ((rounded - 0x38000000u) >> 13);
}

// The SH element j of Gaussian `base`'s words: DC as floats, the rest as halves
// (splat_types.glsl:305-310); or, format 1, every coefficient a bf16, two to a word.
float sh_value(const std::uint32_t *words, std::uint32_t base, std::uint32_t j, std::uint32_t format) {
  if (format == 1) {
    const std::uint32_t w = words[base + j / 2];
    return bits_to_float((j & 1u) ? (w & 0xFFFF0000u) : (w << 16));
  }
  if (j < 3) return bits_to_float(words[base + j]);
  const std::uint32_t w = words[base + 3 + (j - 3) / 2];
  return half_to_float(((j - 3) & 1u) ? (w >> 16) : (w & 0xFFFFu));
}

// train_types.glsl:97-123, in single precision.
void sh_basis(float x, float y, float z, std::uint32_t degree, float b[16]) {
  const float C1 = 0.4886025119029199f;
  const float xx = x * x, yy = y * y, zz = z * z;
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

// train_types.glsl:126-153: d basis / d direction, 3 per function.
void sh_basis_gradient(float x, float y, float z, std::uint32_t degree, float g[16][3]) {
  const float C1 = 0.4886025119029199f, A = 1.0925484305920792f, K = 0.31539156525252005f, L = 0.5462742152960396f;
  const float C = -0.5900435899266435f, E = 2.890611442640554f, F = -0.4570457994644658f, G = 0.3731763325901154f,
              H = 1.445305721320277f;
  const float xx = x * x, yy = y * y, zz = z * z;
  for (int k = 0; k < 16; ++k) g[k][0] = g[k][1] = g[k][2] = 0.0f;
  if (degree > 0) {
    g[1][1] = -C1;
    g[2][2] = C1;
    g[3][0] = -C1;
  }
  if (degree > 1) {
    g[4][0] = A * y;
    g[4][1] = A * x;
    g[5][1] = -A * z;
    g[5][2] = -A * y;
    g[6][0] = K * (-2.0f * x);
    g[6][1] = K * (-2.0f * y);
    g[6][2] = K * (4.0f * z);
    g[7][0] = -A * z;
    g[7][2] = -A * x;
    g[8][0] = L * (2.0f * x);
    g[8][1] = L * (-2.0f * y);
  }
  if (degree > 2) {
    g[9][0] = C * (6.0f * x * y);
    g[9][1] = C * (3.0f * xx - 3.0f * yy);
    g[10][0] = E * (y * z);
    g[10][1] = E * (x * z);
    g[10][2] = E * (x * y);
    g[11][0] = F * (-2.0f * x * y);
    g[11][1] = F * (4.0f * zz - xx - 3.0f * yy);
    g[11][2] = F * (8.0f * y * z);
    g[12][0] = G * (-6.0f * x * z);
    g[12][1] = G * (-6.0f * y * z);
    g[12][2] = G * (6.0f * zz - 3.0f * xx - 3.0f * yy);
    g[13][0] = F * (4.0f * zz - 3.0f * xx - yy);
    g[13][1] = F * (-2.0f * x * y);
    g[13][2] = F * (8.0f * x * z);
    g[14][0] = H * (2.0f * x * z);
    g[14][1] = H * (-2.0f * y * z);
    g[14][2] = H * (xx - yy);
    g[15][0] = C * (3.0f * xx - 3.0f * yy);
    g[15][1] = C * (-6.0f * x * y);
  }
}

// PCG's output permutation (RXS M XS, O'Neill 2014; the GPU hash of Jarzynski
// and Olano 2020) of index ^ (seed * 0x9E3779B9) after one LCG step
// (train_types.glsl:156-160, bf16.glsl:19-22).
std::uint32_t train_hash(std::uint32_t index, std::uint32_t seed) {
  const std::uint32_t state = (index ^ (seed * 0x9E3779B9u)) * 747796405u + 2891336453u;
  const std::uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
  return (word >> 22u) ^ word;
}

// The hash as a float in [0, 1) (train_types.glsl:156-160).
float train_uniform(std::uint32_t index, std::uint32_t seed) { return static_cast<float>(train_hash(index, seed)) * 0x1p-32f; }

// train_types.glsl:165-170
std::uint32_t sh_coefficients_used(std::uint32_t index, std::uint32_t coefficients, float dropout, std::uint32_t retained,
                                   std::uint32_t seed) {
  if (dropout > 0.0f && train_uniform(index, seed) < dropout) {
    return std::min(coefficients, (retained + 1u) * (retained + 1u));
  }
  return coefficients;
}

// World to camera, the same fused order everywhere.
void to_camera(const ob_camera &c, float px, float py, float pz, float &tx, float &ty, float &tz) {
  const float *V = c.view;
  tx = fma(V[8], pz, fma(V[4], py, fma(V[0], px, V[12])));
  ty = fma(V[9], pz, fma(V[5], py, fma(V[1], px, V[13])));
  tz = fma(V[10], pz, fma(V[6], py, fma(V[2], px, V[14])));
}

// Rotation matrix of a unit quaternion (w, x, y, z), R[row][col] (train_types.glsl:172-177).
void quat_to_mat(float w, float x, float y, float z, float R[3][3]) {
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

// sum_k a[i][k] b[j][k] over three terms, fused in order k = 0, 1, 2.
float dot3(float a0, float a1, float a2, float b0, float b1, float b2) { return fma(a2, b2, fma(a1, b1, a0 * b0)); }

// The opacity factor of the 3D filter, sqrt(prod s2 / (s2 + f)) (train_preprocess.comp:62-63).
// With no filter each ratio s2 / (s2 + 0) is exactly 1 for a finite positive s2
// and NaN for 0 or infinity, which is what the shortcut returns.
float filter_factor(float s2x, float s2y, float s2z, float f) {
  if (f == 0.0f) {
    const bool finite = s2x > 0.0f && s2x < INFINITY && s2y > 0.0f && s2y < INFINITY && s2z > 0.0f && s2z < INFINITY;
    return finite ? 1.0f : NAN;
  }
  return std::sqrt((s2x / (s2x + f)) * (s2y / (s2y + f)) * (s2z / (s2z + f)));
}

struct Projection {
  float J00, J02, J11, J12; // the clamped EWA Jacobian's non-zero entries
  float T[2][3];            // J times the view rotation
};
Projection project(const ob_camera &c, float tx, float ty, float iz) {
  const float limX = 1.3f * c.tan_half_x, limY = 1.3f * c.tan_half_y;
  const float ux = clampf(tx * iz, -limX, limX), uy = clampf(ty * iz, -limY, limY);
  Projection p;
  p.J00 = c.fx * iz;
  p.J02 = -(c.fx * ux) * iz;
  p.J11 = c.fy * iz;
  p.J12 = -(c.fy * uy) * iz;
  const float *V = c.view; // W[r][k] = V[k*4 + r]
  for (int k = 0; k < 3; ++k) {
    p.T[0][k] = fma(p.J02, V[k * 4 + 2], p.J00 * V[k * 4 + 0]);
    p.T[1][k] = fma(p.J12, V[k * 4 + 2], p.J11 * V[k * 4 + 1]);
  }
  return p;
}

struct Reach {
  float reach;
};
// 2 log(255 o): the ellipse of splat_at's cut-off, q <= reach (train_types.glsl:51-58).
float reach_of(float opacity) {
  const float r = 2.0f * flog(255.0f * opacity);
  return r > 0.0f ? r : 0.0f;
}

// train_types.glsl:34-39: inclusive tile box.
void splat_tiles(const Splat &s, std::uint32_t gridW, std::uint32_t gridH, std::uint32_t t[4]) {
  const float lastX = static_cast<float>(gridW) - 1.0f, lastY = static_cast<float>(gridH) - 1.0f;
  t[0] = static_cast<std::uint32_t>(clampf(std::floor((s.cx - s.ex - 1.0f) * 0.0625f), 0.0f, lastX));
  t[1] = static_cast<std::uint32_t>(clampf(std::floor((s.cy - s.ey - 1.0f) * 0.0625f), 0.0f, lastY));
  t[2] = static_cast<std::uint32_t>(clampf(std::floor((s.cx + s.ex + 1.0f) * 0.0625f), 0.0f, lastX));
  t[3] = static_cast<std::uint32_t>(clampf(std::floor((s.cy + s.ey + 1.0f) * 0.0625f), 0.0f, lastY));
}

// q(x, y) = qx x^2 + 2 qy x y + qz y^2 along a line, least at the clamped
// stationary point; iqx, iqz are 1/qx and 1/qz.
float edge_min(const Splat &s, float iqx, float iqz, float x, float ylo, float yhi, bool vertical) {
  if (vertical) {
    const float y = clampf((-s.qy * x) * iqz, ylo, yhi);
    return fma(s.qz * y, y, fma(2.0f * s.qy * x, y, s.qx * x * x));
  }
  const float v = x; // the line y = v, parameter u along x
  const float u = clampf((-s.qy * v) * iqx, ylo, yhi);
  return fma(s.qz * v, v, fma(2.0f * s.qy * u, v, s.qx * u * u));
}

// train_types.glsl:56-77
bool reaches_tile(const Splat &s, float reach, float iqx, float iqz, std::uint32_t tx, std::uint32_t ty) {
  const float lowx = (static_cast<float>(tx) * 16.0f + 0.5f) - s.cx, lowy = (static_cast<float>(ty) * 16.0f + 0.5f) - s.cy;
  const float highx = lowx + 15.0f, highy = lowy + 15.0f;
  if (lowx <= 0.0f && lowy <= 0.0f && highx >= 0.0f && highy >= 0.0f) return true;
  float least = 1e30f;
  least = minf(least, edge_min(s, iqx, iqz, lowx, lowy, highy, true));
  least = minf(least, edge_min(s, iqx, iqz, lowy, lowx, highx, false));
  least = minf(least, edge_min(s, iqx, iqz, highx, lowy, highy, true));
  least = minf(least, edge_min(s, iqx, iqz, highy, lowx, highx, false));
  return least <= fma(reach, 1.001f, 1e-3f);
}

// The pixels of the tile at (x0, y0) that a splat can reach, tile-local and
// half-open: x in [box[0], box[2]), y in [box[1], box[3]) (train_types.glsl:43-48).
void splat_box(const Splat &s, std::uint32_t x0, std::uint32_t y0, std::uint32_t x1, std::uint32_t y1, std::uint32_t box[4]) {
  // Clamped to [0, 2^24] so that the conversions to integers are defined.
  const float lx = clampf(s.cx - s.ex - 0.5f, 0.0f, 16777216.0f), ly = clampf(s.cy - s.ey - 0.5f, 0.0f, 16777216.0f);
  const float hx = clampf(s.cx + s.ex + 0.5f, 0.0f, 16777216.0f), hy = clampf(s.cy + s.ey + 0.5f, 0.0f, 16777216.0f);
  const std::uint32_t ax = std::max(static_cast<std::uint32_t>(lx), x0), ay = std::max(static_cast<std::uint32_t>(ly), y0);
  const std::uint32_t bx = std::min(static_cast<std::uint32_t>(hx) + 1u, x1), by = std::min(static_cast<std::uint32_t>(hy) + 1u, y1);
  box[0] = ax;
  box[1] = ay;
  box[2] = bx;
  box[3] = by;
}

// splat_at (train_types.glsl:81-94). e = exp(power) is kept for the backward pass.
bool splat_at(const Splat &s, std::uint32_t x, std::uint32_t y, float &dx, float &dy, float &power, float &e, float &alpha) {
  dx = (static_cast<float>(x) + 0.5f) - s.cx;
  dy = (static_cast<float>(y) + 0.5f) - s.cy;
  if (std::fabs(dx) > s.radius || std::fabs(dy) > s.radius) return false;
  power = fma(-0.5f, fma(s.qx * dx, dx, (s.qz * dy) * dy), -((s.qy * dx) * dy));
  if (power > 0.0f) return false;
  e = fexp(power);
  const float a = s.o * e;
  alpha = a < 0.99f ? a : 0.99f;
  return alpha >= kMinAlpha;
}

struct TileGeometry {
  std::uint32_t x0, y0, x1, y1; // pixels, upper bounds exclusive
};
TileGeometry tile_geometry(std::uint32_t tile, std::uint32_t gridW, std::uint32_t width, std::uint32_t height) {
  const std::uint32_t tx = tile % gridW, ty = tile / gridW;
  return {tx * kTile, ty * kTile, std::min(tx * kTile + kTile, width), std::min(ty * kTile + kTile, height)};
}

// The fixed-point quantisation of one wave total (train_backward_fixed.comp:19-21), NaN as 0,
// in 64 bits: the clamp is the reference's 1e9 times 2^14, for a scale of 2^20
// steps per unit where the reference's has 64 (the caller's fixedScale).
constexpr float kFixedClamp = 1.6384e13f;
std::int64_t quantise(float v, float scale) {
  float s = v * scale;
  if (s != s) s = 0.0f;
  s = s > kFixedClamp ? kFixedClamp : (s < -kFixedClamp ? -kFixedClamp : s);
  return static_cast<std::int64_t>(std::rint(s));
}

// The backward pass's reduction tree over a wave: lane l and lane l ^ m are
// added, for m = 32, 16, 15, 7, 2, 1 in turn (the partners of the GPU's
// permlane swaps, row mirror, half-row mirror and quad permutations).
float wave_sum(float lanes[64]) {
  float a[64];
  std::copy(lanes, lanes + 64, a);
  for (int m : {32, 16, 15, 7, 2, 1}) {
    float b[64];
    for (int l = 0; l < 64; ++l) b[l] = a[l] + a[l ^ m];
    std::copy(b, b + 64, a);
  }
  return a[0];
}

// bf16 (bf16.glsl:10-29)
float bf16_load(std::uint32_t word, std::uint32_t slot) { return bits_to_float(slot == 0 ? word << 16 : word & 0xFFFF0000u); }
std::uint32_t bf16_round(float x) {
  const std::uint32_t bits = float_to_bits(x);
  return (bits + 0x7FFFu + ((bits >> 16) & 1u)) >> 16;
}
// Up or down at random: the low 16 bits of the hash added to the float's bits
// carry into the kept half with probability (low 16 bits of x) / 2^16.
std::uint32_t bf16_stochastic(float x, std::uint32_t index, std::uint32_t seed) {
  return (float_to_bits(x) + (train_hash(index, seed) & 0xFFFFu)) >> 16;
}
std::uint32_t bf16_store(std::uint32_t word, std::uint32_t slot, std::uint32_t bits) {
  return slot == 0 ? (word & 0xFFFF0000u) | bits : (word & 0xFFFFu) | (bits << 16);
}

// Adam with bf16 moments (train_step.glsl:44-61), or with radius >= 0 the
// second-order step: the averaged gradient over the curvature (the state's
// second half), clipped to the radius; only the first moment is stored.
void step_parameter(const ob_chain_options &o, float &value, std::uint32_t &state, float g, float rate, float radius,
                    std::uint32_t index) {
  const float m = fma(o.beta1, bf16_load(state, 0), (1.0f - o.beta1) * g);
  if (radius >= 0.0f) {
    const float change = (m / o.correction1) / std::max(bf16_load(state, 1) * o.curvature_scale, 1e-30f);
    value = value - std::min(std::max(change, -radius), radius);
    state = bf16_store(state, 0, bf16_round(m));
    return;
  }
  const float v = fma(o.beta2, bf16_load(state, 1), ((1.0f - o.beta2) * g) * g);
  value = value - (rate * (m / o.correction1)) / (std::sqrt(v / o.correction2) + o.epsilon);
  state = bf16_store(bf16_store(state, 0, bf16_round(m)), 1, bf16_stochastic(v, index, o.seed));
}

// One estimate of diag(J^T J) into its running average, the state's second
// half rounded at random (train_step.glsl:63-67).
std::uint32_t with_curvature(const ob_chain_options &o, std::uint32_t state, float g, std::uint32_t index) {
  const float c = o.curvature_beta * bf16_load(state, 1) + (((1.0f - o.curvature_beta) * g) * g) * o.per_value;
  return bf16_store(state, 1, bf16_stochastic(c, index, o.seed));
}

// Half floats rounded at random (train_step.glsl:409-419).
std::uint32_t half_stochastic(float x, std::uint32_t index, std::uint32_t seed) {
  const std::uint32_t nearest = float_to_half(x);
  const float rounded = half_to_float(nearest);
  if (rounded == x || std::isinf(rounded)) return nearest;
  const std::uint32_t other = std::fabs(x) > std::fabs(rounded) ? nearest + 1u : nearest - 1u;
  const float gap = half_to_float(other) - rounded;
  return train_uniform(index, seed) < (x - rounded) / gap ? other : nearest;
}

constexpr std::uint32_t kModeStatistics = 0, kModeGradients = 1, kModeStep = 2, kModeCurvature = 3;
constexpr std::uint32_t kFlagStatistics = 1, kFlagSelective = 4, kFlagSecondOrder = 16;

// The trust-region radii of 3DGS^2-TR (train_gradient.comp:131-143): how far
// a parameter may move before the Gaussian's contribution changes by more than
// e in squared Hellinger distance; -1 (Adam) without second order. Rotation
// always takes Adam's step.
constexpr float kShC0 = 0.28209479177387814f;
struct Radii {
  float position, scale, opacity, sh[3];
};
Radii radii_of(const ob_chain_options &o, float opacity, const float s2[3], const float dc[3]) {
  Radii r{-1.0f, -1.0f, -1.0f, {-1.0f, -1.0f, -1.0f}};
  if ((o.flags & kFlagSecondOrder) == 0) return r;
  const float alpha = std::max(opacity, 1e-3f), e = o.trust_region;
  r.position = std::sqrt((((8.0f * ((s2[0] + s2[1]) + s2[2])) / 3.0f) * e) / alpha);
  r.scale = std::sqrt((2.0f * e) / alpha);
  r.opacity = std::sqrt((4.0f * alpha) * e) / std::max(alpha * (1.0f - alpha), 1e-3f);
  for (int c = 0; c < 3; ++c) {
    const float colour = std::min(std::max(0.5f + kShC0 * dc[c], 0.02f), 1.0f);
    r.sh[c] = std::sqrt(((4.0f * colour) * e) / alpha) / kShC0;
  }
  return r;
}

} // namespace
} // namespace oracle

using namespace oracle;

// train_preprocess.comp:31-134
extern "C" int ob_preprocess(const float *positions, const float *scales, const float *rotations, const float *opacities,
                             const uint32_t *shWords, const float *filter3d, const ob_camera *cam, const ob_preprocess_options *opt,
                             float *splats, uint32_t *keys, uint32_t *values) {
  if (cam == nullptr || opt == nullptr) return -1;
  const ob_camera &c = *cam;
  for (std::uint32_t i = 0; i < c.count; ++i) {
    keys[i] = kInvisible;
    values[i] = i;
    const float px = positions[i * 3], py = positions[i * 3 + 1], pz = positions[i * 3 + 2];
    float tx, ty, tz;
    to_camera(c, px, py, pz, tx, ty, tz);
    if (!(tz >= c.near_plane)) continue;
    const float opacity = 1.0f / (1.0f + fexp(-opacities[i]));
    if (!(opacity >= kMinAlpha)) continue;
    float qw = rotations[i * 4], qx = rotations[i * 4 + 1], qy = rotations[i * 4 + 2], qz = rotations[i * 4 + 3];
    const float qlen = std::sqrt(fma(qz, qz, fma(qy, qy, fma(qx, qx, qw * qw))));
    if (!(qlen >= 1e-4f)) continue;
    const float iq = 1.0f / qlen;
    qw = qw * iq;
    qx = qx * iq;
    qy = qy * iq;
    qz = qz * iq;
    const float sx = fexp(scales[i * 3]), sy = fexp(scales[i * 3 + 1]), sz = fexp(scales[i * 3 + 2]);
    float R[3][3];
    quat_to_mat(qw, qx, qy, qz, R);
    const float s[3] = {sx, sy, sz};
    float M[3][3];
    for (int r = 0; r < 3; ++r)
      for (int k = 0; k < 3; ++k) M[r][k] = R[r][k] * s[k];
    const float f = filter3d != nullptr ? filter3d[i] : 0.0f;
    float S[3][3];
    for (int r = 0; r < 3; ++r)
      for (int k = 0; k < 3; ++k) S[r][k] = dot3(M[r][0], M[r][1], M[r][2], M[k][0], M[k][1], M[k][2]);
    S[0][0] = S[0][0] + f;
    S[1][1] = S[1][1] + f;
    S[2][2] = S[2][2] + f;
    const float opacity3 = opacity * filter_factor(sx * sx, sy * sy, sz * sz, f);
    if (!(opacity3 >= kMinAlpha)) continue;

    const float iz = 1.0f / tz;
    const Projection p = project(c, tx, ty, iz);
    float U[2][3];
    for (int r = 0; r < 2; ++r)
      for (int k = 0; k < 3; ++k) U[r][k] = dot3(p.T[r][0], p.T[r][1], p.T[r][2], S[0][k], S[1][k], S[2][k]);
    const float a0 = dot3(U[0][0], U[0][1], U[0][2], p.T[0][0], p.T[0][1], p.T[0][2]);
    const float b = dot3(U[0][0], U[0][1], U[0][2], p.T[1][0], p.T[1][1], p.T[1][2]);
    const float c0 = dot3(U[1][0], U[1][1], U[1][2], p.T[1][0], p.T[1][1], p.T[1][2]);
    const float dilation = opt->mip_filter_2d ? 0.1f : 0.3f;
    float a = a0 + dilation, cc = c0 + dilation;
    float det = fma(a, cc, -(b * b));
    float rawDet = opt->mip_filter_2d ? fma(a0, c0, -(b * b)) : 0.0f;
    // Near-degenerate covariance: the 2D covariance built again in double from
    // the float M and T, and inverted there.
    const bool degenerate = !(det > 1e-6f) || det < 0x1p-12f * (a * cc);
    double da = 0.0, db = 0.0, dc = 0.0, ddet = 0.0;
    if (degenerate) {
      double SD[3][3];
      for (int r = 0; r < 3; ++r)
        for (int k = 0; k < 3; ++k) {
          double acc = 0.0;
          for (int j = 0; j < 3; ++j) acc = acc + static_cast<double>(M[r][j]) * static_cast<double>(M[k][j]);
          SD[r][k] = r == k ? acc + static_cast<double>(f) : acc;
        }
      double UD[2][3];
      for (int r = 0; r < 2; ++r)
        for (int k = 0; k < 3; ++k) {
          UD[r][k] = static_cast<double>(p.T[r][0]) * SD[0][k] + static_cast<double>(p.T[r][1]) * SD[1][k] +
                     static_cast<double>(p.T[r][2]) * SD[2][k];
        }
      const double da0 = UD[0][0] * p.T[0][0] + UD[0][1] * p.T[0][1] + UD[0][2] * p.T[0][2];
      db = UD[0][0] * p.T[1][0] + UD[0][1] * p.T[1][1] + UD[0][2] * p.T[1][2];
      const double dc0 = UD[1][0] * p.T[1][0] + UD[1][1] * p.T[1][1] + UD[1][2] * p.T[1][2];
      da = da0 + dilation;
      dc = dc0 + dilation;
      ddet = da * dc - db * db;
      a = static_cast<float>(da);
      cc = static_cast<float>(dc);
      det = static_cast<float>(ddet);
      if (opt->mip_filter_2d) rawDet = static_cast<float>(da0 * dc0 - db * db);
    }
    if (!(det > 1e-6f)) continue;
    float opacityUsed = opacity3;
    if (opt->mip_filter_2d) {
      opacityUsed = rawDet > 0.0f ? opacity3 * std::sqrt(rawDet / det) : 0.0f;
      if (!(opacityUsed >= kMinAlpha)) continue;
    }
    const float mid = 0.5f * (a + cc);
    const float disc = fma(mid, mid, -det);
    const float lambda = mid + std::sqrt(disc > 0.1f ? disc : 0.1f);
    const float radius = std::ceil(3.0f * std::sqrt(lambda));
    const float centerX = fma(c.fx, tx * iz, c.cx), centerY = fma(c.fy, ty * iz, c.cy);
    if (centerX + radius < 0.0f || centerY + radius < 0.0f || centerX - radius > static_cast<float>(c.width) ||
        centerY - radius > static_cast<float>(c.height)) {
      continue;
    }
    const float dx = px - c.position[0], dy = py - c.position[1], dz = pz - c.position[2];
    const float dl = std::sqrt(fma(dz, dz, fma(dy, dy, dx * dx)));
    float basis[16];
    const float idl = 1.0f / dl;
    sh_basis(dx * idl, dy * idl, dz * idl, c.sh_degree, basis);
    const std::uint32_t used = sh_coefficients_used(i, c.sh_coefficients, opt->sh_dropout, opt->sh_retained, opt->sh_seed);
    const std::uint32_t base = i * c.sh_words;
    float col[3] = {0.5f, 0.5f, 0.5f};
    for (std::uint32_t k = 0; k < used; ++k)
      for (std::uint32_t ch = 0; ch < 3; ++ch) col[ch] = fma(basis[k], sh_value(shWords, base, k * 3 + ch, c.sh_format), col[ch]);
    if (!(std::fabs(col[0]) < 1e30f && std::fabs(col[1]) < 1e30f && std::fabs(col[2]) < 1e30f)) continue;
    const float reach = reach_of(opacityUsed);
    const float ex = std::sqrt(reach * a), ey = std::sqrt(reach * cc);
    Splat out;
    const float id = 1.0f / det;
    out.qx = degenerate ? static_cast<float>(dc / ddet) : cc * id;
    out.qy = degenerate ? static_cast<float>(-db / ddet) : -b * id;
    out.qz = degenerate ? static_cast<float>(da / ddet) : a * id;
    out.o = opacityUsed;
    out.r = maxf(col[0], 0.0f);
    out.g = maxf(col[1], 0.0f);
    out.b = maxf(col[2], 0.0f);
    out.radius = radius;
    out.cx = centerX;
    out.cy = centerY;
    out.ex = minf(ex, radius);
    out.ey = minf(ey, radius);
    std::memcpy(splats + static_cast<std::size_t>(i) * 12, &out, sizeof(Splat));
    keys[i] = 0xFFFFFFFFu - float_to_bits(tz);
  }
  return 0;
}

extern "C" int ob_stable_sort(uint32_t *keys, uint32_t *values, uint32_t count) {
  std::vector<std::uint32_t> order(count);
  std::iota(order.begin(), order.end(), 0u);
  std::stable_sort(order.begin(), order.end(), [&](std::uint32_t a, std::uint32_t b) { return keys[a] < keys[b]; });
  std::vector<std::uint32_t> k(count), v(count);
  for (std::uint32_t i = 0; i < count; ++i) {
    k[i] = keys[order[i]];
    v[i] = values[order[i]];
  }
  std::copy(k.begin(), k.end(), keys);
  std::copy(v.begin(), v.end(), values);
  return 0;
}

// train_bin_count.comp:18-37 and train_bin_emit.comp:18-37, with the exclusive
// scan between them (the entries of splat r start at the sum of the counts before it).
extern "C" int ob_bin(const uint32_t *keys, const uint32_t *values, const float *splats, uint32_t count, uint32_t gridW,
                      uint32_t gridH, uint32_t *coverage, uint32_t *entryKeys, uint32_t *entryValues, uint64_t capacity,
                      uint64_t *counts) {
  std::uint64_t visible = 0, total = 0;
  for (std::uint32_t r = 0; r < count; ++r) {
    std::uint32_t n = 0;
    if (keys[r] != kInvisible) {
      const Splat s = load_splat(splats, values[r]);
      const float reach = reach_of(s.o), iqx = 1.0f / s.qx, iqz = 1.0f / s.qz;
      std::uint32_t t[4];
      splat_tiles(s, gridW, gridH, t);
      for (std::uint32_t y = t[1]; y <= t[3]; ++y) {
        for (std::uint32_t x = t[0]; x <= t[2]; ++x) {
          if (!reaches_tile(s, reach, iqx, iqz, x, y)) continue;
          if (entryKeys != nullptr) {
            if (total + n >= capacity) return -2;
            entryKeys[total + n] = y * gridW + x;
            entryValues[total + n] = values[r];
          }
          ++n;
        }
      }
      ++visible;
    }
    if (coverage != nullptr) coverage[r] = n;
    total += n;
  }
  counts[0] = visible;
  counts[1] = total;
  return 0;
}

// train_tile_ranges.comp:11-22
extern "C" int ob_tile_ranges(const uint32_t *entryKeys, uint64_t entries, uint32_t tiles, uint32_t *ranges) {
  std::fill(ranges, ranges + 2 * static_cast<std::size_t>(tiles), 0u);
  for (std::uint64_t k = 0; k < entries; ++k) {
    const std::uint32_t tile = entryKeys[k];
    if (tile >= tiles) return -2;
    if (k == 0 || entryKeys[k - 1] != tile) ranges[tile * 2] = static_cast<std::uint32_t>(k);
    if (k + 1 == entries || entryKeys[k + 1] != tile) ranges[tile * 2 + 1] = static_cast<std::uint32_t>(k + 1);
  }
  return 0;
}

// train_forward.comp:38-100, pixel by pixel (a pixel's result does not depend on the batching).
// A pixel stops where the product of 1 - alpha would fall under 1e-4 (the
// shader tests the sum of the logs against log 1e-4, the same bound), and
// its log transmittance is the log of that product.
extern "C" int ob_forward(const float *splats, const uint32_t *entryValues, const uint32_t *ranges, uint32_t width, uint32_t height,
                          const float *background, float *image, uint32_t *stops) {
  const std::uint32_t gridW = (width + kTile - 1) / kTile, gridH = (height + kTile - 1) / kTile;
  for (std::uint32_t tile = 0; tile < gridW * gridH; ++tile) {
    const TileGeometry g = tile_geometry(tile, gridW, width, height);
    const std::uint32_t begin = ranges[tile * 2], end = ranges[tile * 2 + 1];
    for (std::uint32_t y = g.y0; y < g.y1; ++y) {
      for (std::uint32_t x = g.x0; x < g.x1; ++x) {
        float rgb[3] = {0.0f, 0.0f, 0.0f}, T = 1.0f;
        std::uint32_t stop = begin;
        for (std::uint32_t e = end; e-- > begin;) {
          const Splat s = load_splat(splats, entryValues[e]);
          std::uint32_t box[4];
          splat_box(s, g.x0, g.y0, g.x1, g.y1, box);
          if (x < box[0] || x >= box[2] || y < box[1] || y >= box[3]) continue;
          float dx, dy, power, ep, alpha;
          if (!splat_at(s, x, y, dx, dy, power, ep, alpha)) continue;
          const float Tnext = T * (1.0f - alpha);
          if (Tnext < kMinTransmittance) {
            stop = e + 1;
            break;
          }
          const float w = T * alpha;
          rgb[0] = fma(w, s.r, rgb[0]);
          rgb[1] = fma(w, s.g, rgb[1]);
          rgb[2] = fma(w, s.b, rgb[2]);
          T = Tnext;
        }
        const float logT = flog(T);
        const std::size_t at = static_cast<std::size_t>(y) * width + x;
        image[at * 4] = fma(T, background[0], rgb[0]);
        image[at * 4 + 1] = fma(T, background[1], rgb[1]);
        image[at * 4 + 2] = fma(T, background[2], rgb[2]);
        image[at * 4 + 3] = logT;
        stops[at] = stop;
      }
    }
  }
  return 0;
}

namespace {

// reference.cpp:12-13 / train_ssim.comp:25-26
const float kWeights[11] = {0.00102838f, 0.00759876f, 0.03600077f, 0.10936069f, 0.21300554f, 0.26601172f,
                            0.21300554f, 0.10936069f, 0.03600077f, 0.00759876f, 0.00102838f};

// train_ssim.comp:34-40: the window's weight inside the image.
float window_norm(int at, int size) {
  float norm = 0.0f;
  for (int i = std::max(at - 5, 0); i <= std::min(at + 5, size - 1); ++i) norm = norm + kWeights[i - at + 5];
  return norm;
}

float unorm8(std::uint8_t v) { return static_cast<float>(v) * (1.0f / 255.0f); }

// The sum over a tile's 256 pixels (zero outside the image): per wave of 64,
// lane l plus lane l ^ m for m = 32, 16, 8, 4, 2, 1, then the four waves in order.
float tile_sum(const float values[256]) {
  float total = 0.0f;
  for (int w = 0; w < 4; ++w) {
    float a[64];
    std::copy(values + w * 64, values + w * 64 + 64, a);
    for (int m = 32; m >= 1; m >>= 1) {
      float b[64];
      for (int l = 0; l < 64; ++l) b[l] = a[l] + a[l ^ m];
      std::copy(b, b + 64, a);
    }
    total = w == 0 ? a[0] : total + a[0];
  }
  return total;
}

} // namespace

// train_ssim.comp:42-123 then train_loss.comp:52-113. fields: 9 floats per
// pixel (P rgb, Q rgb, R rgb).
extern "C" int ob_loss(const float *image, const uint8_t *target, uint32_t width, uint32_t height, float lambda, float *fields,
                       float *adjoint, float *tileTotals) {
  const int W = static_cast<int>(width), H = static_cast<int>(height);
  const std::size_t pixels = static_cast<std::size_t>(W) * H;
  const float C1 = 0.0001f, C2 = 0.0009f;
  // Horizontal window means: 15 per pixel (mx, my, xx, yy, xy, each rgb).
  std::vector<float> hsum(pixels * 15);
  for (int y = 0; y < H; ++y) {
    for (int x = 0; x < W; ++x) {
      float acc[15] = {};
      float norm = 0.0f;
      for (int u = std::max(x - 5, 0); u <= std::min(x + 5, W - 1); ++u) {
        const float w = kWeights[u - x + 5];
        const std::size_t at = static_cast<std::size_t>(y) * W + u;
        for (int ch = 0; ch < 3; ++ch) {
          const float a = image[at * 4 + ch], b = unorm8(target[at * 4 + ch]);
          const float wa = w * a, wb = w * b;
          acc[ch] = acc[ch] + wa;
          acc[3 + ch] = acc[3 + ch] + wb;
          acc[6 + ch] = fma(wa, a, acc[6 + ch]);
          acc[9 + ch] = fma(wb, b, acc[9 + ch]);
          acc[12 + ch] = fma(wa, b, acc[12 + ch]);
        }
        norm = norm + w;
      }
      const float inorm = 1.0f / norm;
      for (int q = 0; q < 15; ++q) hsum[(static_cast<std::size_t>(y) * W + x) * 15 + q] = acc[q] * inorm;
    }
  }
  std::vector<float> ownSsim(pixels);
  for (int py = 0; py < H; ++py) {
    const float rowNorm = window_norm(py, H);
    const float irow = 1.0f / rowNorm;
    for (int px = 0; px < W; ++px) {
      float acc[15] = {};
      for (int v = std::max(py - 5, 0); v <= std::min(py + 5, H - 1); ++v) {
        const float w = kWeights[v - py + 5];
        for (int q = 0; q < 15; ++q) acc[q] = fma(w, hsum[(static_cast<std::size_t>(v) * W + px) * 15 + q], acc[q]);
      }
      const float inorm = 1.0f / (rowNorm * window_norm(px, W));
      const std::size_t at = static_cast<std::size_t>(py) * W + px;
      float own = 0.0f;
      for (int ch = 0; ch < 3; ++ch) {
        const float mx = acc[ch] * irow, my = acc[3 + ch] * irow;
        const float vx = fma(-mx, mx, acc[6 + ch] * irow), vy = fma(-my, my, acc[9 + ch] * irow),
                    cxy = fma(-mx, my, acc[12 + ch] * irow);
        const float n1 = fma(2.0f * mx, my, C1), n2 = fma(2.0f, cxy, C2);
        const float d1 = fma(mx, mx, fma(my, my, C1)), d2 = (vx + vy) + C2;
        const float dd = d1 * d2;
        const float idd = 1.0f / dd, id1 = 1.0f / d1, id2 = 1.0f / d2;
        const float s = (n1 * n2) * idd;
        own = ch == 0 ? s : own + s;
        const float dMu = ((2.0f * my) * n2) * idd - ((2.0f * s) * mx) * id1;
        const float dVar = -s * id2;
        const float dCov = (2.0f * n1) * idd;
        fields[at * 9 + ch] = fma(-my, dCov, fma(-2.0f * mx, dVar, dMu)) * inorm;
        fields[at * 9 + 3 + ch] = dVar * inorm;
        fields[at * 9 + 6 + ch] = dCov * inorm;
      }
      ownSsim[at] = own;
    }
  }
  // Loss pass: the window applied to the fields, horizontally then vertically.
  std::vector<float> hf(pixels * 9);
  for (int y = 0; y < H; ++y) {
    for (int x = 0; x < W; ++x) {
      float acc[9] = {};
      for (int u = std::max(x - 5, 0); u <= std::min(x + 5, W - 1); ++u) {
        const float w = kWeights[u - x + 5];
        const std::size_t at = static_cast<std::size_t>(y) * W + u;
        for (int q = 0; q < 9; ++q) acc[q] = fma(w, fields[at * 9 + q], acc[q]);
      }
      std::copy(acc, acc + 9, hf.begin() + static_cast<std::ptrdiff_t>((static_cast<std::size_t>(y) * W + x) * 9));
    }
  }
  const float perValue = 1.0f / (3.0f * static_cast<float>(W) * static_cast<float>(H));
  const float keep = 1.0f - lambda;
  std::vector<float> ownL1(pixels);
  for (int py = 0; py < H; ++py) {
    for (int px = 0; px < W; ++px) {
      float acc[9] = {};
      for (int v = std::max(py - 5, 0); v <= std::min(py + 5, H - 1); ++v) {
        const float w = kWeights[v - py + 5];
        for (int q = 0; q < 9; ++q) acc[q] = fma(w, hf[(static_cast<std::size_t>(v) * W + px) * 9 + q], acc[q]);
      }
      const std::size_t at = static_cast<std::size_t>(py) * W + px;
      float own = 0.0f;
      for (int ch = 0; ch < 3; ++ch) {
        const float x = image[at * 4 + ch], y = unorm8(target[at * 4 + ch]);
        const float d = x - y;
        own = ch == 0 ? std::fabs(d) : own + std::fabs(d);
        const float sgn = d > 0.0f ? 1.0f : (d < 0.0f ? -1.0f : 0.0f);
        const float inner = fma(y, acc[6 + ch], fma(2.0f * x, acc[3 + ch], acc[ch]));
        adjoint[at * 4 + ch] = fma(-lambda, inner, keep * sgn) * perValue;
      }
      adjoint[at * 4 + 3] = 0.0f;
      ownL1[at] = own;
    }
  }
  // Per-tile totals in the tile's reduction order.
  const std::uint32_t gridW = (width + kTile - 1) / kTile, gridH = (height + kTile - 1) / kTile;
  for (std::uint32_t tile = 0; tile < gridW * gridH; ++tile) {
    const TileGeometry g = tile_geometry(tile, gridW, width, height);
    float s[256] = {}, l[256] = {};
    for (std::uint32_t y = g.y0; y < g.y1; ++y) {
      for (std::uint32_t x = g.x0; x < g.x1; ++x) {
        const std::uint32_t lane = (y - g.y0) * 16 + (x - g.x0);
        s[lane] = ownSsim[static_cast<std::size_t>(y) * W + x];
        l[lane] = ownL1[static_cast<std::size_t>(y) * W + x];
      }
    }
    tileTotals[tile * 2] = tile_sum(s);
    tileTotals[tile * 2 + 1] = tile_sum(l);
  }
  return 0;
}

namespace {

// train_backward.glsl:78-158 for one pixel and one entry: the 14 values, or
// false when the pixel takes nothing from the splat. B and behind advance:
// behind is the transmittance in front of the entries visited so far (at the
// start exp(logTotal)), and an entry's own transmittance is behind / (1 - alpha),
// where train_backward.glsl:122-123 takes exp(logTotal - logBehind - logRest).
struct PixelState {
  float logTotal, A[3], B[3], behind;
  float marked, edge; // a scored view: the adjoint's fourth value and the target's edge weight
  std::uint32_t stop;
  bool inside;
};

bool backward_pair(const Splat &s, const std::uint32_t box[4], std::uint32_t x, std::uint32_t y, std::uint32_t e, float perPixel,
                   PixelState &p, float g[kRaw]) {
  for (int k = 0; k < kRaw; ++k) g[k] = 0.0f;
  if (!p.inside || e < p.stop) return false;
  if (x < box[0] || x >= box[2] || y < box[1] || y >= box[3]) return false;
  float dx, dy, power, ep, alpha;
  if (!splat_at(s, x, y, dx, dy, power, ep, alpha)) return false;
  const float T = p.behind / (1.0f - alpha);
  const float Ta = T * alpha;
  g[6] = Ta * p.A[0];
  g[7] = Ta * p.A[1];
  g[8] = Ta * p.A[2];
  g[9] = Ta * perPixel;
  // A scored view's two more sums (train_backward.glsl:130-131): zero otherwise.
  g[12] = p.marked * perPixel;
  g[13] = (p.edge * Ta) * perPixel;
  if (s.o * ep < 0.99f) {
    const float dAlpha = T * fma(p.A[2], s.b - p.B[2], fma(p.A[1], s.g - p.B[1], p.A[0] * (s.r - p.B[0])));
    g[0] = dAlpha * ep;
    const float dPower = dAlpha * alpha;
    const float vx = fma(s.qy, dy, s.qx * dx), vy = fma(s.qz, dy, s.qy * dx);
    g[1] = dPower * vx;
    g[2] = dPower * vy;
    g[10] = std::fabs(g[1]);
    g[11] = std::fabs(g[2]);
    const float hp = 0.5f * dPower;
    g[3] = (hp * vx) * vx;
    g[4] = (hp * vx) * vy;
    g[5] = (hp * vy) * vy;
  }
  p.B[0] = fma(alpha, s.r - p.B[0], p.B[0]);
  p.B[1] = fma(alpha, s.g - p.B[1], p.B[1]);
  p.B[2] = fma(alpha, s.b - p.B[2], p.B[2]);
  p.behind = T;
  return true;
}

// The pixel of lane `lane` of wave `wave`: the wave holds the 8x8 quadrant
// (wave & 1, wave >> 1) of the tile, the lane its pixel (lane & 7, lane >> 3).
std::uint32_t wave_x(const TileGeometry &g, std::uint32_t wave, std::uint32_t lane) { return g.x0 + (wave & 1u) * 8u + (lane & 7u); }
std::uint32_t wave_y(const TileGeometry &g, std::uint32_t wave, std::uint32_t lane) { return g.y0 + (wave >> 1) * 8u + (lane >> 3); }

// Visits every (tile, wave, entry) in order with the 64 lanes' values; `emit`
// receives the wave's 14 totals.
template <class Emit>
void backward_waves(const float *splats, const uint32_t *entryValues, const uint32_t *ranges, uint32_t width, uint32_t height,
                    const float *background, const float *image, const float *adjoint, const uint32_t *stops, const float *edge,
                    Emit &&emit) {
  const std::uint32_t gridW = (width + kTile - 1) / kTile, gridH = (height + kTile - 1) / kTile;
  const float perPixel = 1.0f / (3.0f * static_cast<float>(width) * static_cast<float>(height));
  for (std::uint32_t tile = 0; tile < gridW * gridH; ++tile) {
    const TileGeometry geo = tile_geometry(tile, gridW, width, height);
    const std::uint32_t begin = ranges[tile * 2], end = ranges[tile * 2 + 1];
    std::uint32_t tileStop = end;
    for (std::uint32_t y = geo.y0; y < geo.y1; ++y)
      for (std::uint32_t x = geo.x0; x < geo.x1; ++x) tileStop = std::min(tileStop, stops[static_cast<std::size_t>(y) * width + x]);
    (void)begin;
    for (std::uint32_t wave = 0; wave < 4; ++wave) {
      PixelState px[64];
      for (std::uint32_t lane = 0; lane < 64; ++lane) {
        const std::uint32_t x = wave_x(geo, wave, lane), y = wave_y(geo, wave, lane);
        PixelState &p = px[lane];
        p.inside = x < width && y < height;
        const std::size_t at = p.inside ? static_cast<std::size_t>(y) * width + x : 0;
        p.logTotal = p.inside ? image[at * 4 + 3] : 0.0f;
        for (int ch = 0; ch < 3; ++ch) {
          p.A[ch] = p.inside ? adjoint[at * 4 + ch] : 0.0f;
          p.B[ch] = background[ch];
        }
        p.behind = fexp(p.logTotal);
        p.marked = p.inside && edge != nullptr ? adjoint[at * 4 + 3] : 0.0f;
        p.edge = p.inside && edge != nullptr ? edge[at] : 0.0f;
        p.stop = p.inside ? stops[at] : end;
      }
      for (std::uint32_t e = tileStop; e < end; ++e) {
        const std::uint32_t index = entryValues[e];
        const Splat s = load_splat(splats, index);
        std::uint32_t box[4];
        splat_box(s, geo.x0, geo.y0, geo.x1, geo.y1, box);
        float lanes[kRaw][64];
        bool any = false;
        for (std::uint32_t lane = 0; lane < 64; ++lane) {
          const std::uint32_t x = wave_x(geo, wave, lane), y = wave_y(geo, wave, lane);
          float g[kRaw];
          any = backward_pair(s, box, x, y, e, perPixel, px[lane], g) || any;
          for (int k = 0; k < kRaw; ++k) lanes[k][lane] = g[k];
        }
        if (!any) continue;
        float totals[kRaw];
        for (int k = 0; k < kRaw; ++k) totals[k] = wave_sum(lanes[k]);
        emit(index, totals);
      }
    }
  }
}

} // namespace

extern "C" int ob_backward_fixed(const float *splats, const uint32_t *entryValues, const uint32_t *ranges, uint32_t width,
                                 uint32_t height, const float *background, const float *image, const float *adjoint,
                                 const uint32_t *stops, float fixedScale, uint32_t count, int64_t *raw, const float *edge) {
  std::fill(raw, raw + static_cast<std::size_t>(count) * kRaw, 0);
  backward_waves(splats, entryValues, ranges, width, height, background, image, adjoint, stops, edge,
                 [&](std::uint32_t index, const float totals[kRaw]) {
                   for (int k = 0; k < kRaw; ++k) raw[static_cast<std::size_t>(index) * kRaw + k] += quantise(totals[k], fixedScale);
                 });
  return 0;
}

extern "C" int ob_backward_float(const float *splats, const uint32_t *entryValues, const uint32_t *ranges, uint32_t width,
                                 uint32_t height, const float *background, const float *image, const float *adjoint,
                                 const uint32_t *stops, uint32_t count, float *raw, const float *edge) {
  std::fill(raw, raw + static_cast<std::size_t>(count) * kRaw, 0.0f);
  backward_waves(splats, entryValues, ranges, width, height, background, image, adjoint, stops, edge,
                 [&](std::uint32_t index, const float totals[kRaw]) {
                   for (int k = 0; k < kRaw; ++k) raw[static_cast<std::size_t>(index) * kRaw + k] += totals[k];
                 });
  return 0;
}

// train_fixed_to_float.comp:9-14, through double: the int64 converts exactly, is scaled by the
// reciprocal of the scale in double, and is rounded to float once.
extern "C" int ob_fixed_to_float(const int64_t *raw, uint64_t n, float fixedScale, float *out) {
  const double inverse = 1.0 / static_cast<double>(fixedScale);
  for (std::uint64_t i = 0; i < n; ++i) out[i] = static_cast<float>(static_cast<double>(raw[i]) * inverse);
  return 0;
}

// train_gradient.comp:105-315 (statistics, gradients and Adam modes).
extern "C" int ob_chain(float *positions, float *scales, float *rotations, float *opacities, uint32_t *shWords, const float *filter3d,
                        const ob_camera *cam, float *raw, const float *splats, float *stats, uint32_t *statePos,
                        uint32_t *stateScale, uint32_t *stateRot, uint32_t *stateOpacity, uint32_t * /*stateSh: stepped by ob_sh_step*/, float *gPos,
                        float *gScale, float *gRot, float *gOpacity, float *gSh, const ob_chain_options *opt) {
  if (cam == nullptr || opt == nullptr) return -1;
  const ob_camera &c = *cam;
  const ob_chain_options &o = *opt;
  const float W = static_cast<float>(c.width), H = static_cast<float>(c.height);
  for (std::uint32_t i = 0; i < c.count; ++i) {
    float *g = raw + static_cast<std::size_t>(i) * kRaw;
    bool touched = false;
    for (int k = 0; k < kRaw; ++k) touched = touched || g[k] != 0.0f;
    const bool zeros = o.mode == kModeGradients || (o.mode == kModeStep && (o.flags & kFlagSelective) == 0);
    const bool penalties = o.opacity_reg > 0.0f || o.scale_reg > 0.0f;
    const auto add_statistics = [&]() {
      float *st = stats + static_cast<std::size_t>(i) * 8;
      const float ax = (g[1] * 0.5f) * W, ay = (g[2] * 0.5f) * H;
      st[0] = st[0] + std::sqrt(fma(ay, ay, ax * ax));
      st[1] = st[1] + 1.0f;
      const float radius = splats[static_cast<std::size_t>(i) * 12 + 7];
      st[2] = maxf(st[2], radius);
      st[3] = fma(((g[9] * 3.0f) * W), H, st[3]);
      const float bx = (g[10] * 0.5f) * W, by = (g[11] * 0.5f) * H;
      st[4] = st[4] + std::sqrt(fma(by, by, bx * bx));
    };
    if (o.mode == kModeStatistics || (!touched && !zeros && !penalties)) {
      if (touched && (o.flags & kFlagStatistics) != 0) add_statistics();
      continue;
    }
    const float logScale[3] = {scales[i * 3], scales[i * 3 + 1], scales[i * 3 + 2]};
    const float opacity = 1.0f / (1.0f + fexp(-opacities[i]));
    const float s2[3] = {fexp(2.0f * logScale[0]), fexp(2.0f * logScale[1]), fexp(2.0f * logScale[2])};
    const std::uint32_t base = i * c.sh_words, stored = o.sh_floats / 3;
    const float dc[3] = {sh_value(shWords, base, 0, c.sh_format), sh_value(shWords, base, 1, c.sh_format),
                         sh_value(shWords, base, 2, c.sh_format)};
    const Radii radii = radii_of(o, opacity, s2, dc);
    // Emitting: gradients stored, Adam's or the second-order step, or the
    // curvature (train_gradient.comp:57-72).
    const auto emit = [&](float *values, std::uint32_t *state, float *grad, std::uint32_t at, float gv, float rate, float radius,
                          std::uint32_t group) {
      if (o.mode == kModeGradients) {
        grad[at] = gv;
      } else if (gv != 0.0f || (o.mode == kModeStep && (o.flags & kFlagSelective) == 0 && radius < 0.0f)) {
        if (o.mode == kModeCurvature) {
          state[at] = with_curvature(o, state[at], gv, at * 5u + group);
          return;
        }
        float v = values[at];
        std::uint32_t st = state[at];
        step_parameter(o, v, st, gv, rate, radius, at * 5u + group);
        values[at] = v;
        state[at] = st;
      }
    };
    if (!touched) {
      const float op = penalties ? o.opacity_reg * opacity * (1.0f - opacity) : 0.0f;
      for (std::uint32_t a = 0; a < 3; ++a) {
        emit(positions, statePos, gPos, i * 3 + a, 0.0f, o.rates[0], radii.position, 0);
        emit(scales, stateScale, gScale, i * 3 + a, penalties ? o.scale_reg * fexp(logScale[a]) : 0.0f, o.rates[1], radii.scale, 1);
      }
      if (o.mode != kModeCurvature)
        for (std::uint32_t a = 0; a < 4; ++a) emit(rotations, stateRot, gRot, i * 4 + a, 0.0f, o.rates[2], -1.0f, 2);
      emit(opacities, stateOpacity, gOpacity, i, op, o.rates[3], radii.opacity, 3);
      if (o.mode == kModeGradients)
        for (std::uint32_t j = 0; j < o.sh_floats; ++j) gSh[static_cast<std::size_t>(i) * o.sh_floats + j] = 0.0f;
      continue;
    }
    if ((o.flags & kFlagStatistics) != 0) add_statistics();

    const float px = positions[i * 3], py = positions[i * 3 + 1], pz = positions[i * 3 + 2];
    float qw = rotations[i * 4], qx = rotations[i * 4 + 1], qy = rotations[i * 4 + 2], qz = rotations[i * 4 + 3];
    const float qlen = std::sqrt(fma(qz, qz, fma(qy, qy, fma(qx, qx, qw * qw))));
    if (qlen > 0.0f) {
      const float iq = 1.0f / qlen;
      qw = qw * iq;
      qx = qx * iq;
      qy = qy * iq;
      qz = qz * iq;
    } else {
      qw = 1.0f;
      qx = qy = qz = 0.0f;
    }
    float tx, ty, tz;
    to_camera(c, px, py, pz, tx, ty, tz);
    const float iz = 1.0f / tz, iz2 = iz * iz;
    const float gc0 = g[1], gc1 = g[2];
    float G00 = g[3], G01 = g[4], G11 = g[5];
    float k2 = 1.0f;
    float R[3][3];
    quat_to_mat(qw, qx, qy, qz, R);
    const float f = filter3d != nullptr ? filter3d[i] : 0.0f;
    // Sigma = R D R^T + f I, then X = Wv Sigma Wv^T.
    float S[3][3];
    for (int r = 0; r < 3; ++r)
      for (int k = 0; k < 3; ++k) S[r][k] = dot3(R[r][0] * s2[0], R[r][1] * s2[1], R[r][2] * s2[2], R[k][0], R[k][1], R[k][2]);
    S[0][0] = S[0][0] + f;
    S[1][1] = S[1][1] + f;
    S[2][2] = S[2][2] + f;
    const float *V = c.view;
    float Y[3][3], X[3][3];
    for (int r = 0; r < 3; ++r)
      for (int k = 0; k < 3; ++k) Y[r][k] = dot3(V[0 * 4 + r], V[1 * 4 + r], V[2 * 4 + r], S[0][k], S[1][k], S[2][k]);
    for (int r = 0; r < 3; ++r)
      for (int k = 0; k < 3; ++k) X[r][k] = dot3(Y[r][0], Y[r][1], Y[r][2], V[0 * 4 + k], V[1 * 4 + k], V[2 * 4 + k]);
    const float k3 = filter_factor(s2[0], s2[1], s2[2], f);
    const float opacity3 = opacity * k3;
    const float limX = 1.3f * c.tan_half_x, limY = 1.3f * c.tan_half_y;
    const float rx = tx * iz, ry = ty * iz;
    const bool clampedX = std::fabs(rx) > limX, clampedY = std::fabs(ry) > limY;
    const Projection pr = project(c, tx, ty, iz);
    const float ux = clampf(rx, -limX, limX), uy = clampf(ry, -limY, limY);
    // X J^T, 3x2: the product every trace term shares.
    float XJ[3][2];
    for (int r = 0; r < 3; ++r) {
      XJ[r][0] = fma(X[r][2], pr.J02, X[r][0] * pr.J00);
      XJ[r][1] = fma(X[r][2], pr.J12, X[r][1] * pr.J11);
    }
    if (opt->mip_filter_2d) {
      // C = J X J^T, and the opacity factor sqrt(det C / det(C + 0.1 I)) (train_gradient.comp:197-211).
      const float C00 = fma(pr.J02, XJ[2][0], pr.J00 * XJ[0][0]);
      const float C01 = fma(pr.J02, XJ[2][1], pr.J00 * XJ[0][1]);
      const float C10 = fma(pr.J12, XJ[2][0], pr.J11 * XJ[1][0]);
      const float C11 = fma(pr.J12, XJ[2][1], pr.J11 * XJ[1][1]);
      const float F00 = C00 + 0.1f, F11 = C11 + 0.1f;
      const float detC = fma(C00, C11, -(C01 * C10)), detF = fma(F00, F11, -(C01 * C10));
      if (detC > 0.0f && detF > 0.0f) {
        const float k = std::sqrt(detC / detF);
        const float m = ((g[0] * opacity3) * 0.5f) * k;
        G00 = fma(m, C11 / detC - F11 / detF, G00);
        G01 = fma(m, -C01 / detC + C01 / detF, G01);
        G11 = fma(m, C00 / detC - F00 / detF, G11);
        k2 = k;
      } else {
        k2 = 0.0f;
      }
    }
    // Position: the centre's Jacobian (unclamped) and the trace terms of dJ.
    const float Jc02 = -(c.fx * tx) * iz2, Jc12 = -(c.fy * ty) * iz2;
    float gT0 = pr.J00 * gc0, gT1 = pr.J11 * gc1, gT2 = fma(Jc12, gc1, Jc02 * gc0);
    const float ax = clampedX ? 0.0f : -c.fx * iz2, ay = clampedY ? 0.0f : -c.fy * iz2;
    const float zx = (c.fx * ux) * iz2, zy = (c.fy * uy) * iz2;
    const float az0 = -c.fx * iz2, az1 = -c.fy * iz2, az2 = clampedX ? zx : 2.0f * zx, az3 = clampedY ? zy : 2.0f * zy;
    const float trX = ax * fma(G01, XJ[2][1], G00 * XJ[2][0]);
    const float trY = ay * fma(G11, XJ[2][1], G01 * XJ[2][0]);
    const float N00 = fma(az2, XJ[2][0], az0 * XJ[0][0]), N01 = fma(az2, XJ[2][1], az0 * XJ[0][1]);
    const float N10 = fma(az3, XJ[2][0], az1 * XJ[1][0]), N11 = fma(az3, XJ[2][1], az1 * XJ[1][1]);
    const float trZ = fma(G11, N11, fma(G01, N10, fma(G01, N01, G00 * N00)));
    gT0 = fma(2.0f, trX, gT0);
    gT1 = fma(2.0f, trY, gT1);
    gT2 = fma(2.0f, trZ, gT2);
    float gP[3];
    for (int k = 0; k < 3; ++k) gP[k] = dot3(V[k * 4 + 0], V[k * 4 + 1], V[k * 4 + 2], gT0, gT1, gT2);

    // The 3D covariance: G3 = T^T G T.
    float GT[2][3];
    for (int k = 0; k < 3; ++k) {
      GT[0][k] = fma(G01, pr.T[1][k], G00 * pr.T[0][k]);
      GT[1][k] = fma(G11, pr.T[1][k], G01 * pr.T[0][k]);
    }
    float G3[3][3];
    for (int r = 0; r < 3; ++r)
      for (int k = 0; k < 3; ++k) G3[r][k] = fma(pr.T[1][r], GT[1][k], pr.T[0][r] * GT[0][k]);
    float dScale[3];
    for (int a = 0; a < 3; ++a) {
      float v[3];
      for (int r = 0; r < 3; ++r) v[r] = dot3(G3[r][0], G3[r][1], G3[r][2], R[0][a], R[1][a], R[2][a]);
      const float quad = dot3(R[0][a], R[1][a], R[2][a], v[0], v[1], v[2]);
      const float filterTerm = f != 0.0f ? ((((g[0] * k2) * opacity) * k3) * f) / (s2[a] + f) : 0.0f;
      dScale[a] = fma(o.scale_reg, fexp(logScale[a]), fma(2.0f * s2[a], quad, filterTerm));
    }
    // Rotation: P = G3 R D (half of the reference's P), dL/dq = 2 <P, dR/dq>.
    float P[3][3];
    for (int r = 0; r < 3; ++r)
      for (int k = 0; k < 3; ++k) P[r][k] = dot3(G3[r][0], G3[r][1], G3[r][2], R[0][k], R[1][k], R[2][k]) * s2[k];
    const float w = qw, x = qx, y = qy, z = qz;
    const float sw = fma(x, P[2][1] - P[1][2], fma(y, P[0][2] - P[2][0], z * (P[1][0] - P[0][1])));
    const float sx = fma(w, P[2][1] - P[1][2], fma(z, P[0][2] + P[2][0], y * (P[0][1] + P[1][0]))) - 2.0f * x * (P[1][1] + P[2][2]);
    const float sy = fma(w, P[0][2] - P[2][0], fma(z, P[1][2] + P[2][1], x * (P[0][1] + P[1][0]))) - 2.0f * y * (P[0][0] + P[2][2]);
    const float sz = fma(w, P[1][0] - P[0][1], fma(y, P[1][2] + P[2][1], x * (P[0][2] + P[2][0]))) - 2.0f * z * (P[0][0] + P[1][1]);
    const float gq[4] = {4.0f * sw, 4.0f * sx, 4.0f * sy, 4.0f * sz};
    const float qd = fma(z, gq[3], fma(y, gq[2], fma(x, gq[1], w * gq[0])));
    const float qv[4] = {w, x, y, z};
    const float denom = qlen > 1e-12f ? qlen : 1e-12f;
    float tangent[4];
    const float idenom = 1.0f / denom;
    for (int k = 0; k < 4; ++k) tangent[k] = fma(-qv[k], qd, gq[k]) * idenom;

    const float dOpacity = ((fma(g[0] * k2, k3, o.opacity_reg) * opacity) * (1.0f - opacity));

    // Colour and the view direction's part of the position gradient.
    const float gColour[3] = {g[6], g[7], g[8]};
    const float ddx = px - c.position[0], ddy = py - c.position[1], ddz = pz - c.position[2];
    const float dist = std::sqrt(fma(ddz, ddz, fma(ddy, ddy, ddx * ddx)));
    const float idist = 1.0f / dist;
    const float dir[3] = {ddx * idist, ddy * idist, ddz * idist};
    float basis[16];
    sh_basis(dir[0], dir[1], dir[2], c.sh_degree, basis);
    const std::uint32_t used = sh_coefficients_used(i, c.sh_coefficients, o.sh_dropout, o.sh_retained, o.sh_seed);
    float colour[3] = {0.5f, 0.5f, 0.5f};
    for (std::uint32_t k = 0; k < used; ++k)
      for (std::uint32_t ch = 0; ch < 3; ++ch) colour[ch] = fma(basis[k], sh_value(shWords, base, k * 3 + ch, c.sh_format), colour[ch]);
    const float unclamped[3] = {colour[0] > 0.0f ? 1.0f : 0.0f, colour[1] > 0.0f ? 1.0f : 0.0f, colour[2] > 0.0f ? 1.0f : 0.0f};
    const float gcu[3] = {gColour[0] * unclamped[0], gColour[1] * unclamped[1], gColour[2] * unclamped[2]};
    if (c.sh_degree > 0) {
      float bg[16][3];
      sh_basis_gradient(dir[0], dir[1], dir[2], c.sh_degree, bg);
      float gDir[3] = {0.0f, 0.0f, 0.0f};
      for (std::uint32_t k = 1; k < used; ++k) {
        const float d = dot3(gcu[0], gcu[1], gcu[2], sh_value(shWords, base, k * 3, c.sh_format),
                             sh_value(shWords, base, k * 3 + 1, c.sh_format), sh_value(shWords, base, k * 3 + 2, c.sh_format));
        for (int a = 0; a < 3; ++a) gDir[a] = fma(d, bg[k][a], gDir[a]);
      }
      const float dd = dot3(dir[0], dir[1], dir[2], gDir[0], gDir[1], gDir[2]);
      for (int a = 0; a < 3; ++a) gP[a] = gP[a] + fma(-dir[a], dd, gDir[a]) * idist;
    }

    if (o.pose_terms != nullptr && o.mode != kModeCurvature && (gP[0] != 0.0f || gP[1] != 0.0f || gP[2] != 0.0f)) {
      // train_gradient.comp:276-279: v = Rv gP, w = t x v.
      const float *V = c.view;
      float v[3];
      for (int r = 0; r < 3; ++r) v[r] = dot3(V[r], V[4 + r], V[8 + r], gP[0], gP[1], gP[2]);
      float *out = o.pose_terms + static_cast<std::size_t>(i) * 6;
      out[0] = fma(ty, v[2], -(tz * v[1]));
      out[1] = fma(tz, v[0], -(tx * v[2]));
      out[2] = fma(tx, v[1], -(ty * v[0]));
      out[3] = v[0];
      out[4] = v[1];
      out[5] = v[2];
    }
    for (std::uint32_t a = 0; a < 3; ++a) {
      emit(positions, statePos, gPos, i * 3 + a, gP[a], o.rates[0], radii.position, 0);
      emit(scales, stateScale, gScale, i * 3 + a, dScale[a], o.rates[1], radii.scale, 1);
    }
    if (o.mode != kModeCurvature)
      for (std::uint32_t a = 0; a < 4; ++a) emit(rotations, stateRot, gRot, i * 4 + a, tangent[a], o.rates[2], -1.0f, 2);
    emit(opacities, stateOpacity, gOpacity, i, dOpacity, o.rates[3], radii.opacity, 3);
    if (o.mode == kModeGradients) {
      for (std::uint32_t k = 0; k < stored; ++k)
        for (std::uint32_t ch = 0; ch < 3; ++ch)
          gSh[static_cast<std::size_t>(i) * o.sh_floats + k * 3 + ch] = k < used ? basis[k] * gcu[ch] : 0.0f;
    } else {
      // For the SH pass (train_gradient.comp:300-313): colour gradient, direction, coefficients used.
      g[0] = gcu[0];
      g[1] = gcu[1];
      g[2] = gcu[2];
      g[3] = dir[0];
      g[4] = dir[1];
      g[5] = dir[2];
      g[6] = radii.sh[0];
      g[7] = radii.sh[1];
      g[8] = radii.sh[2];
      g[9] = bits_to_float(used);
    }
  }
  return 0;
}

// train_sh_step.comp:19-70 (first-order Adam).
// The SH floats packed word q holds, and one of them read from and stored
// back into the word: halves (and the DC floats as they are) in the
// reference's layout, bf16 in format 1; stored values are rounded at random.
void sh_word_floats(std::uint32_t format, std::uint32_t q, std::uint32_t shFloats, std::uint32_t &first, std::uint32_t &last) {
  if (format == 1) {
    first = 2 * q;
  } else {
    first = q < 3 ? q : 3 + 2 * (q - 3);
    if (q < 3) {
      last = q;
      return;
    }
  }
  last = std::min(first + 1, shFloats - 1);
}
float sh_word_get(std::uint32_t format, std::uint32_t word, std::uint32_t j, std::uint32_t first) {
  if (format == 1) return bf16_load(word, j - first);
  if (j < 3) return bits_to_float(word);
  return half_to_float(((j - 3) & 1u) ? (word >> 16) : (word & 0xFFFFu));
}
std::uint32_t sh_word_put(std::uint32_t format, std::uint32_t word, std::uint32_t j, std::uint32_t first, float value, std::uint32_t s,
                          std::uint32_t seed) {
  if (format == 1) return bf16_store(word, j - first, bf16_stochastic(value, s, seed ^ 0x5DEECE66u));
  if (j < 3) return float_to_bits(value);
  const std::uint32_t shift = 16u * ((j - 3) & 1u);
  return (word & ~(0xFFFFu << shift)) | (half_stochastic(value, s, seed ^ 0x5DEECE66u) << shift);
}

extern "C" int ob_sh_step(uint32_t *shWords, uint32_t *stateSh, const float *raw, const ob_camera *cam, const ob_chain_options *opt) {
  const ob_camera &c = *cam;
  const ob_chain_options &o = *opt;
  const std::uint32_t words = c.sh_words;
  for (std::uint32_t w = 0; w < c.count * words; ++w) {
    const std::uint32_t i = w / words, q = w - i * words;
    const float *g = raw + static_cast<std::size_t>(i) * kRaw;
    const float gColour[3] = {g[0], g[1], g[2]};
    const std::uint32_t used = float_to_bits(g[9]);
    const bool secondOrder = (o.flags & kFlagSecondOrder) != 0;
    const bool zeros = o.mode == kModeStep && (o.flags & kFlagSelective) == 0 && !secondOrder;
    if (used == 0 && !zeros) continue;
    float basis[16];
    sh_basis(g[3], g[4], g[5], c.sh_degree, basis);
    std::uint32_t first, last;
    sh_word_floats(c.sh_format, q, o.sh_floats, first, last);
    std::uint32_t word = shWords[w];
    bool changed = false;
    for (std::uint32_t j = first; j <= last; ++j) {
      const std::uint32_t k = j / 3, ch = j - 3 * k;
      const float gv = k < used ? basis[k] * gColour[ch] : 0.0f;
      if (gv == 0.0f && !zeros) continue;
      const std::uint32_t s = i * o.sh_floats + j;
      std::uint32_t state = stateSh[s];
      if (o.mode == kModeCurvature) {
        stateSh[s] = with_curvature(o, state, gv, s * 5u + 4u);
        continue;
      }
      float value = sh_word_get(c.sh_format, word, j, first);
      step_parameter(o, value, state, gv, k == 0 ? o.rates[4] : o.rates[4] * o.rest_scale, secondOrder ? g[6 + ch] : -1.0f,
                     s * 5u + 4u);
      stateSh[s] = state;
      word = sh_word_put(c.sh_format, word, j, first, value, s, o.seed);
      changed = true;
    }
    if (changed) shWords[w] = word;
  }
  return 0;
}

// The batched step's Adam from summed gradients: every non-SH parameter with
// the chain rule's rule and rounding stream (index at * 5 + group), then SH
// per packed word as ob_sh_step does, with the gradient read from g_sh.
extern "C" int ob_adam(float *positions, float *scales, float *rotations, float *opacities, uint32_t *shWords, uint32_t *statePos,
                       uint32_t *stateScale, uint32_t *stateRot, uint32_t *stateOpacity, uint32_t *stateSh, const float *gPos,
                       const float *gScale, const float *gRot, const float *gOpacity, const float *gSh, uint32_t count,
                       uint32_t shWordsPer, uint32_t shFormat, const ob_chain_options *opt) {
  const ob_chain_options &o = *opt;
  const bool zeros = (o.flags & kFlagSelective) == 0;
  const auto step = [&](float *values, std::uint32_t *state, const float *grad, std::uint32_t n, float rate, std::uint32_t group) {
    for (std::uint32_t at = 0; at < n; ++at) {
      const float gv = grad[at];
      if (gv == 0.0f && !zeros) continue;
      float v = values[at];
      std::uint32_t st = state[at];
      step_parameter(o, v, st, gv, rate, -1.0f, at * 5u + group);
      values[at] = v;
      state[at] = st;
    }
  };
  step(positions, statePos, gPos, count * 3, o.rates[0], 0);
  step(scales, stateScale, gScale, count * 3, o.rates[1], 1);
  step(rotations, stateRot, gRot, count * 4, o.rates[2], 2);
  step(opacities, stateOpacity, gOpacity, count, o.rates[3], 3);
  for (std::uint32_t w = 0; w < count * shWordsPer; ++w) {
    const std::uint32_t i = w / shWordsPer, q = w - i * shWordsPer;
    std::uint32_t first, last;
    sh_word_floats(shFormat, q, o.sh_floats, first, last);
    std::uint32_t word = shWords[w];
    bool changed = false;
    for (std::uint32_t j = first; j <= last; ++j) {
      const std::uint32_t s = i * o.sh_floats + j;
      const float gv = gSh[s];
      if (gv == 0.0f && !zeros) continue;
      std::uint32_t state = stateSh[s];
      float value = sh_word_get(shFormat, word, j, first);
      step_parameter(o, value, state, gv, j < 3 ? o.rates[4] : o.rates[4] * o.rest_scale, -1.0f, s * 5u + 4u);
      stateSh[s] = state;
      word = sh_word_put(shFormat, word, j, first, value, s, o.seed);
      changed = true;
    }
    if (changed) shWords[w] = word;
  }
  return 0;
}

extern "C" void ob_exp_array(const float *x, float *y, uint64_t n) {
  for (std::uint64_t i = 0; i < n; ++i) y[i] = fexp(x[i]);
}
extern "C" void ob_log_array(const float *x, float *y, uint64_t n) {
  for (std::uint64_t i = 0; i < n; ++i) y[i] = flog(x[i]);
}

// The hash and the two stochastic roundings elementwise, for the tests that pin
// them: train_hash(index[i], seed), train_uniform(index[i], seed), bf16_stochastic(x[i], index[i], seed)
// and half_stochastic(x[i], index[i], seed) (16 bits per output word).
extern "C" void ob_hash_array(const uint32_t *index, uint32_t seed, uint64_t n, uint32_t *out) {
  for (std::uint64_t i = 0; i < n; ++i) out[i] = oracle::train_hash(index[i], seed);
}
extern "C" void ob_uniform_array(const uint32_t *index, uint32_t seed, uint64_t n, float *out) {
  for (std::uint64_t i = 0; i < n; ++i) out[i] = oracle::train_uniform(index[i], seed);
}
extern "C" void ob_bf16_stochastic_array(const float *x, const uint32_t *index, uint32_t seed, uint64_t n, uint32_t *out) {
  for (std::uint64_t i = 0; i < n; ++i) out[i] = oracle::bf16_stochastic(x[i], index[i], seed);
}
extern "C" void ob_half_stochastic_array(const float *x, const uint32_t *index, uint32_t seed, uint64_t n, uint32_t *out) {
  for (std::uint64_t i = 0; i < n; ++i) out[i] = oracle::half_stochastic(x[i], index[i], seed);
}
