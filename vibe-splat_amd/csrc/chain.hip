// The chain rule from the backward pass's 14 screen-space sums per Gaussian to
// the raw parameters, and what becomes of the result (train_gradient.comp:105-315):
// the densification statistics only, the gradients stored (one float per
// parameter), or Adam's step taken at once with bf16 moments. The SH
// coefficients step in a pass of their own, one thread per packed word
// (train_sh_step.comp:19-70), so that neighbouring threads touch neighbouring memory.
//
// The fixed-point sums are converted here (train_fixed_to_float.comp:9-14):
// int64 to double (exact), times the scale's reciprocal in double, rounded to float once.
#include "device.hpp"
#include "kernels.hpp"

namespace vsplat {

namespace {

// bf16 moments (bf16.glsl:10-29).
__device__ __forceinline__ float bf16_load(uint32_t word, uint32_t slot) { return as_float(slot == 0 ? word << 16 : word & 0xFFFF0000u); }
__device__ __forceinline__ uint32_t bf16_round(float x) {
  const uint32_t bits = as_uint(x);
  return (bits + 0x7FFFu + ((bits >> 16) & 1u)) >> 16;
}
__device__ __forceinline__ uint32_t bf16_stochastic(float x, uint32_t index, uint32_t seed) {
  const uint32_t state = (index ^ (seed * 0x9E3779B9u)) * 747796405u + 2891336453u;
  const uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
  return (as_uint(x) + (((word >> 22u) ^ word) & 0xFFFFu)) >> 16;
}
__device__ __forceinline__ uint32_t bf16_store(uint32_t word, uint32_t slot, uint32_t bits) {
  return slot == 0 ? (word & 0xFFFF0000u) | bits : (word & 0xFFFFu) | (bits << 16);
}

// Adam (train_step.glsl:44-61): the step uses both moments before they are
// rounded; m is rounded to nearest, v at random. With radius >= 0 the
// second-order step of 3DGS^2-TR instead: the averaged gradient over the
// curvature (the state's second half), clipped to the radius; only m is stored.
__device__ __forceinline__ void step_parameter(const ChainParams &o, float &value, uint32_t &state, float g, float rate, float radius,
                                               uint32_t index) {
  const float m = fma32(o.beta1, bf16_load(state, 0), (1.0f - o.beta1) * g);
  if (radius >= 0.0f) {
    const float curvature = bf16_load(state, 1) * o.curvatureScale;
    const float change = (m / o.correction1) / (curvature > 1e-30f ? curvature : 1e-30f);
    value = value - (change < -radius ? -radius : (change > radius ? radius : change));
    state = bf16_store(state, 0, bf16_round(m));
    return;
  }
  const float v = fma32(o.beta2, bf16_load(state, 1), ((1.0f - o.beta2) * g) * g);
  value = value - (rate * (m / o.correction1)) / (sqrtf(v / o.correction2) + o.epsilon);
  state = bf16_store(bf16_store(state, 0, bf16_round(m)), 1, bf16_stochastic(v, index, o.seed));
}

// One estimate of diag(J^T J) into its running average in the state's second
// half, rounded at random (train_step.glsl:63-67).
__device__ __forceinline__ uint32_t with_curvature(const ChainParams &o, uint32_t state, float g, uint32_t index) {
  const float c = o.curvatureBeta * bf16_load(state, 1) + (((1.0f - o.curvatureBeta) * g) * g) * o.perValue;
  return bf16_store(state, 1, bf16_stochastic(c, index, o.seed));
}

// The trust-region radii of one Gaussian (train_gradient.comp:131-143), from
// its parameters before the step: how far each may move before its
// contribution changes by more than epsilon in squared Hellinger distance;
// -1 (Adam's step) without second order. Rotation always takes Adam's step.
struct Radii {
  float position, scale, opacity;
};
__device__ __forceinline__ Radii radii_of(const ChainParams &o, const float *opacities, const float *scales, uint32_t i) {
  if ((o.flags & kFlagSecondOrder) == 0) return Radii{-1.0f, -1.0f, -1.0f};
  const float opacity = 1.0f / (1.0f + vs_exp(-opacities[i]));
  const float alpha = opacity > 1e-3f ? opacity : 1e-3f, e = o.trustRegion;
  const float s2[3] = {vs_exp(2.0f * scales[i * 3]), vs_exp(2.0f * scales[i * 3 + 1]), vs_exp(2.0f * scales[i * 3 + 2])};
  const float spread = alpha * (1.0f - alpha);
  return Radii{sqrtf((((8.0f * ((s2[0] + s2[1]) + s2[2])) / 3.0f) * e) / alpha), sqrtf((2.0f * e) / alpha),
               sqrtf((4.0f * alpha) * e) / (spread > 1e-3f ? spread : 1e-3f)};
}
// The SH coefficients' radius for colour channel ch, from its DC coefficient.
__device__ __forceinline__ float sh_radius(const ChainParams &o, float opacity, float dc) {
  const float alpha = opacity > 1e-3f ? opacity : 1e-3f;
  const float shC0 = 0.28209479177387814f;
  const float raw = 0.5f + shC0 * dc;
  const float colour = raw < 0.02f ? 0.02f : (raw > 1.0f ? 1.0f : raw);
  return sqrtf(((4.0f * colour) * o.trustRegion) / alpha) / shC0;
}

// The step about to be taken, from the device state when there is one.
__device__ __forceinline__ void adam_step_now(ChainParams &o) {
  if (o.adamState == nullptr) return;
  const uint32_t taken = o.adamState[0];
  o.correction1 = 1.0f - as_float(o.adamState[1]) * o.beta1;
  o.correction2 = 1.0f - as_float(o.adamState[2]) * o.beta2;
  o.seed = taken + 1u;
}

// Half floats rounded at random in proportion to the distance (train_step.glsl:409-419).
__device__ __forceinline__ uint32_t half_stochastic(float x, uint32_t index, uint32_t seed) {
  const uint32_t nearest = float_to_half(x);
  const float rounded = half_to_float(nearest);
  if (rounded == x || fabsf(rounded) == __builtin_inff()) return nearest;
  const uint32_t other = fabsf(x) > fabsf(rounded) ? nearest + 1u : nearest - 1u;
  const float gap = half_to_float(other) - rounded;
  return train_uniform(index, seed) < (x - rounded) / gap ? other : nearest;
}

// What becomes of gradient g of parameter `at` (train_gradient.comp:57-72):
// stored, a step taken (Adam's, or within `radius` when >= 0), or (curvature
// mode) its square added to the curvature.
__device__ __forceinline__ void emit(const ChainParams &o, float *values, uint32_t *state, float *grad, uint32_t at, float g, float rate,
                                     float radius, uint32_t group) {
  if (o.mode == kModeGradients) {
    grad[at] = (o.flags & kFlagAccumulate) != 0 ? grad[at] + g : g;
  } else if (g != 0.0f || (o.mode == kModeStep && (o.flags & kFlagSelective) == 0 && radius < 0.0f)) {
    if (o.mode == kModeCurvature) {
      state[at] = with_curvature(o, state[at], g, at * 5u + group);
      return;
    }
    float v = values[at];
    uint32_t s = state[at];
    step_parameter(o, v, s, g, rate, radius, at * 5u + group);
    values[at] = v;
    state[at] = s;
  }
}

// d basis_k / d direction for one basis function k >= 1 (train_types.glsl:126-153).
__device__ __forceinline__ void sh_basis_gradient_term(uint32_t k, float x, float y, float z, float g[3]) {
  const float C1 = 0.4886025119029199f, A = 1.0925484305920792f, K = 0.31539156525252005f, L = 0.5462742152960396f;
  const float C = -0.5900435899266435f, E = 2.890611442640554f, F = -0.4570457994644658f, G = 0.3731763325901154f,
              H = 1.445305721320277f;
  const float xx = x * x, yy = y * y, zz = z * z;
  g[0] = g[1] = g[2] = 0.0f;
  switch (k) {
  case 1: g[1] = -C1; break;
  case 2: g[2] = C1; break;
  case 3: g[0] = -C1; break;
  case 4: g[0] = A * y; g[1] = A * x; break;
  case 5: g[1] = -A * z; g[2] = -A * y; break;
  case 6: g[0] = K * (-2.0f * x); g[1] = K * (-2.0f * y); g[2] = K * (4.0f * z); break;
  case 7: g[0] = -A * z; g[2] = -A * x; break;
  case 8: g[0] = L * (2.0f * x); g[1] = L * (-2.0f * y); break;
  case 9: g[0] = C * (6.0f * x * y); g[1] = C * (3.0f * xx - 3.0f * yy); break;
  case 10: g[0] = E * (y * z); g[1] = E * (x * z); g[2] = E * (x * y); break;
  case 11: g[0] = F * (-2.0f * x * y); g[1] = F * (4.0f * zz - xx - 3.0f * yy); g[2] = F * (8.0f * y * z); break;
  case 12: g[0] = G * (-6.0f * x * z); g[1] = G * (-6.0f * y * z); g[2] = G * (6.0f * zz - 3.0f * xx - 3.0f * yy); break;
  case 13: g[0] = F * (4.0f * zz - 3.0f * xx - yy); g[1] = F * (-2.0f * x * y); g[2] = F * (8.0f * x * z); break;
  case 14: g[0] = H * (2.0f * x * z); g[1] = H * (-2.0f * y * z); g[2] = H * (xx - yy); break;
  default: g[0] = C * (3.0f * xx - 3.0f * yy); g[1] = C * (-6.0f * x * y); break;
  }
}

// The fixed-point or float sums of one Gaussian in one view, as floats, from
// `raw` (the Gaussian's kRawPerSplat int64 or float values, in global memory or LDS).
__device__ __forceinline__ bool convert_raw(const uint32_t *raw, bool fixed, float fixedScale, float g[kRawPerSplat]) {
  if (fixed) {
    const double inverse = 1.0 / static_cast<double>(fixedScale);
#pragma unroll
    for (uint32_t k = 0; k < kRawPerSplat; ++k) {
      const long long v = static_cast<long long>(static_cast<uint64_t>(raw[2 * k]) | (static_cast<uint64_t>(raw[2 * k + 1]) << 32));
      g[k] = static_cast<float>(static_cast<double>(v) * inverse);
    }
  } else {
#pragma unroll
    for (uint32_t k = 0; k < kRawPerSplat; ++k) g[k] = as_float(raw[k]);
  }
  bool touched = false;
#pragma unroll
  for (uint32_t k = 0; k < kRawPerSplat; ++k) touched = touched || g[k] != 0.0f;
  return touched;
}
__device__ __forceinline__ bool load_raw(const int64_t *rawFixed, const float *rawFloat, float fixedScale, uint32_t i,
                                         float g[kRawPerSplat]) {
  if (rawFixed != nullptr) {
    return convert_raw(reinterpret_cast<const uint32_t *>(rawFixed + static_cast<size_t>(i) * kRawPerSplat), true, fixedScale, g);
  }
  return convert_raw(reinterpret_cast<const uint32_t *>(rawFloat + static_cast<size_t>(i) * kRawPerSplat), false, fixedScale, g);
}

// One view's raw sums of Gaussian i, fetched into registers ahead of their
// use (the loads are waited for where the words are first read), and with
// `statistics` the splat's radius. Sums 2k and 2k + 1 sit in raw[4k .. 4k + 3]
// as two int64 (low word first), or as two floats in raw[4k] and raw[4k + 1]:
// the same places for both kinds, so that the registers stay registers. The
// last two sums are a scored view's; a view that is not scored leaves them zero
// and they are not loaded.
constexpr uint32_t kUnscoredSums = 12;
__device__ __forceinline__ void fetch_raw(const ChainView &w, uint32_t i, bool statistics, uint32_t raw[2 * kRawPerSplat], float &radius) {
  if (w.rawFixed != nullptr) {
    const uint4 *src = reinterpret_cast<const uint4 *>(w.rawFixed + static_cast<size_t>(i) * kRawPerSplat);
#pragma unroll
    for (uint32_t k = 0; k < kRawPerSplat / 2; ++k) {
      const uint4 v = 2 * k < kUnscoredSums ? src[k] : make_uint4(0u, 0u, 0u, 0u);
      raw[4 * k] = v.x;
      raw[4 * k + 1] = v.y;
      raw[4 * k + 2] = v.z;
      raw[4 * k + 3] = v.w;
    }
  } else {
    const uint2 *src = reinterpret_cast<const uint2 *>(w.rawFloat + static_cast<size_t>(i) * kRawPerSplat);
#pragma unroll
    for (uint32_t k = 0; k < kRawPerSplat / 2; ++k) {
      const uint2 v = 2 * k < kUnscoredSums ? src[k] : make_uint2(0u, 0u);
      raw[4 * k] = v.x;
      raw[4 * k + 1] = v.y;
      raw[4 * k + 2] = 0u;
      raw[4 * k + 3] = 0u;
    }
  }
  if (statistics) radius = w.splats[i].radius;
}

// convert_raw for the words fetch_raw placed.
__device__ __forceinline__ bool convert_fetched(const uint32_t raw[2 * kRawPerSplat], bool fixed, float fixedScale, float g[kRawPerSplat]) {
  if (fixed) return convert_raw(raw, true, fixedScale, g);
  uint32_t words[kRawPerSplat];
#pragma unroll
  for (uint32_t k = 0; k < kRawPerSplat; ++k) words[k] = raw[4 * (k / 2) + (k & 1u)];
  return convert_raw(words, false, fixedScale, g);
}

// DensifyStat (train_gradient.comp:94-101), per half image rather than per pixel.
__device__ __forceinline__ void add_statistics(float4 &a, float4 &b, const float g[kRawPerSplat], float W, float H, float radius) {
  const float ax = (g[1] * 0.5f) * W, ay = (g[2] * 0.5f) * H;
  a.x = a.x + sqrtf(fma32(ay, ay, ax * ax));
  a.y = a.y + 1.0f;
  a.z = maxf(a.z, radius);
  a.w = fma32((g[9] * 3.0f) * W, H, a.w);
  const float bx = (g[10] * 0.5f) * W, by = (g[11] * 0.5f) * H;
  b.x = b.x + sqrtf(fma32(by, by, bx * bx));
}

// What one view gives the geometric parameters.
struct ViewTerms {
  float position[3], scale[3], rotation[4], opacity;
};

// The penalties' part alone, for a Gaussian no pixel of the view saw.
__device__ __forceinline__ void untouched_terms(const ChainBuffers &B, const ChainParams &o, uint32_t i, bool penalties, ViewTerms &t) {
  const float opacity = 1.0f / (1.0f + vs_exp(-B.opacities[i]));
  const float op = penalties ? o.opacityReg * opacity * (1.0f - opacity) : 0.0f;
#pragma unroll
  for (uint32_t a = 0; a < 3; ++a) {
    t.position[a] = 0.0f;
    t.scale[a] = penalties ? o.scaleReg * vs_exp(B.scales[i * 3 + a]) : 0.0f;
  }
#pragma unroll
  for (uint32_t a = 0; a < 4; ++a) t.rotation[a] = 0.0f;
  t.opacity = op;
}

// The chain rule for a Gaussian the view saw (train_gradient.comp:105-315):
// from the screen-space sums g to the geometric terms, with the SH part handed
// to sh(basis, colour gradient, direction, coefficients used) as soon as it is
// known. Reads the parameters only.
// Its SH part: the colour, clamped at zero, the SH terms handed to sh(...),
// and dirTerm, the view direction's part of the position gradient.
template <int Degree, uint32_t Format, typename ShSink>
__device__ __forceinline__ void view_sh_terms(const ViewParams &c, const ChainParams &o, uint32_t i, const float position[3],
                                              const uint32_t *words, const float g[kRawPerSplat], float dirTerm[3], ShSink &&sh) {
  const float ddx = position[0] - c.position[0], ddy = position[1] - c.position[1], ddz = position[2] - c.position[2];
  const float dist = sqrtf(fma32(ddz, ddz, fma32(ddy, ddy, ddx * ddx)));
  const float idist = 1.0f / dist;
  const float dir[3] = {ddx * idist, ddy * idist, ddz * idist};
  constexpr uint32_t kCoefficients = (Degree + 1) * (Degree + 1);
  float basis[16];
  sh_basis(dir[0], dir[1], dir[2], Degree, basis);
  const uint32_t used = sh_coefficients_used(i, c.shCoefficients, o.shDropout, o.shRetained, o.shSeed);
  // The coefficients (`words`: this Gaussian's, staged in LDS) are read where
  // they are used (colour, then the direction term) rather than held, which
  // keeps the kernel's register count down.
  const uint32_t base = 0;
  float colour[3] = {0.5f, 0.5f, 0.5f};
#pragma unroll
  for (uint32_t k = 0; k < kCoefficients; ++k) {
#pragma unroll
    for (uint32_t ch = 0; ch < 3; ++ch) {
      const float v = sh_value<Format>(words, base, k * 3 + ch);
      colour[ch] = k < used ? fma32(basis[k], v, colour[ch]) : colour[ch];
    }
  }
  const float gcu[3] = {g[6] * (colour[0] > 0.0f ? 1.0f : 0.0f), g[7] * (colour[1] > 0.0f ? 1.0f : 0.0f),
                        g[8] * (colour[2] > 0.0f ? 1.0f : 0.0f)};
  dirTerm[0] = dirTerm[1] = dirTerm[2] = 0.0f;
  if (Degree > 0) {
    float gDir[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
    for (uint32_t k = 1; k < kCoefficients; ++k) {
      const float d = dot3(gcu[0], gcu[1], gcu[2], sh_value<Format>(words, base, k * 3),
                           sh_value<Format>(words, base, k * 3 + 1), sh_value<Format>(words, base, k * 3 + 2));
      float bg[3];
      sh_basis_gradient_term(k, dir[0], dir[1], dir[2], bg);
#pragma unroll
      for (int a = 0; a < 3; ++a) gDir[a] = k < used ? fma32(d, bg[a], gDir[a]) : gDir[a];
    }
    const float dd = dot3(dir[0], dir[1], dir[2], gDir[0], gDir[1], gDir[2]);
#pragma unroll
    for (int a = 0; a < 3; ++a) dirTerm[a] = fma32(-dir[a], dd, gDir[a]) * idist;
  }
  sh(basis, gcu, dir, used);
}

// What the chain rule needs of a Gaussian whatever the view: its position,
// exp(log-scale) and its square, the rotation normalised (the identity for a
// zero quaternion) with the quaternion's length, the rotation matrix R, the
// 3D covariance Sigma = R diag(s2) R^T + f I, the opacity, the 3D filter f
// and its opacity factor. Computed once per Gaussian however many views use it.
struct GaussianFrame {
  float p[3], e[3], s2[3];
  float q[4], qlen;
  float R[3][3], S[3][3];
  float opacity, f, k3;
};

__device__ __forceinline__ void gaussian_frame(const ChainBuffers &B, uint32_t i, GaussianFrame &G) {
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    G.p[a] = B.positions[i * 3 + a];
    const float logScale = B.scales[i * 3 + a];
    G.e[a] = vs_exp(logScale);
    G.s2[a] = vs_exp(2.0f * logScale);
  }
  G.opacity = 1.0f / (1.0f + vs_exp(-B.opacities[i]));
  float qw = B.rotations[i * 4], qx = B.rotations[i * 4 + 1], qy = B.rotations[i * 4 + 2], qz = B.rotations[i * 4 + 3];
  G.qlen = sqrtf(fma32(qz, qz, fma32(qy, qy, fma32(qx, qx, qw * qw))));
  if (G.qlen > 0.0f) {
    const float iq = 1.0f / G.qlen;
    qw = qw * iq;
    qx = qx * iq;
    qy = qy * iq;
    qz = qz * iq;
  } else {
    qw = 1.0f;
    qx = qy = qz = 0.0f;
  }
  G.q[0] = qw;
  G.q[1] = qx;
  G.q[2] = qy;
  G.q[3] = qz;
  quat_to_mat(qw, qx, qy, qz, G.R);
  G.f = B.filter3d != nullptr ? B.filter3d[i] : 0.0f;
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int k = 0; k < 3; ++k)
      G.S[r][k] = dot3(G.R[r][0] * G.s2[0], G.R[r][1] * G.s2[1], G.R[r][2] * G.s2[2], G.R[k][0], G.R[k][1], G.R[k][2]);
  G.S[0][0] = G.S[0][0] + G.f;
  G.S[1][1] = G.S[1][1] + G.f;
  G.S[2][2] = G.S[2][2] + G.f;
  G.k3 = filter_factor(G.s2[0], G.s2[1], G.s2[2], G.f);
}

// Its geometric part: position (with dirTerm added when the colour depends on
// the view), log-scale, rotation and opacity terms.
__device__ __forceinline__ void view_geometry_terms(const GaussianFrame &G, const ViewParams &c, const ChainParams &o,
                                                    const float g[kRawPerSplat], const float dirTerm[3], bool withDir, ViewTerms &t) {
  const float opacity = G.opacity, f = G.f;
  const float *s2 = G.s2;
  const float qw = G.q[0], qx = G.q[1], qy = G.q[2], qz = G.q[3], qlen = G.qlen;
  float tx, ty, tz;
  to_camera(c, G.p[0], G.p[1], G.p[2], tx, ty, tz);
  const float iz = 1.0f / tz, iz2 = iz * iz;
  const float gc0 = g[1], gc1 = g[2];
  float G00 = g[3], G01 = g[4], G11 = g[5];
  float k2 = 1.0f;
  const float (&R)[3][3] = G.R;
  // X = Wv Sigma Wv^T with Wv the view rotation.
  const float (&S)[3][3] = G.S;
  const float *V = c.view;
  float Y[3][3], X[3][3];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int k = 0; k < 3; ++k) Y[r][k] = dot3(V[0 * 4 + r], V[1 * 4 + r], V[2 * 4 + r], S[0][k], S[1][k], S[2][k]);
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int k = 0; k < 3; ++k) X[r][k] = dot3(Y[r][0], Y[r][1], Y[r][2], V[0 * 4 + k], V[1 * 4 + k], V[2 * 4 + k]);
  const float k3 = G.k3;
  const float opacity3 = opacity * k3;
  const float limX = 1.3f * c.tanHalfX, limY = 1.3f * c.tanHalfY;
  const float rx = tx * iz, ry = ty * iz;
  const bool clampedX = fabsf(rx) > limX, clampedY = fabsf(ry) > limY;
  const float ux = clampf(rx, -limX, limX), uy = clampf(ry, -limY, limY);
  const float J00 = c.fx * iz, J02 = -(c.fx * ux) * iz, J11 = c.fy * iz, J12 = -(c.fy * uy) * iz;
  float T[2][3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    T[0][k] = fma32(J02, V[k * 4 + 2], J00 * V[k * 4 + 0]);
    T[1][k] = fma32(J12, V[k * 4 + 2], J11 * V[k * 4 + 1]);
  }
  float XJ[3][2];
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    XJ[r][0] = fma32(X[r][2], J02, X[r][0] * J00);
    XJ[r][1] = fma32(X[r][2], J12, X[r][1] * J11);
  }
  if (o.mipFilter2d) {
    // The 2D filter's opacity factor k = sqrt(det C / det(C + 0.1 I)) moves with C too.
    const float C00 = fma32(J02, XJ[2][0], J00 * XJ[0][0]);
    const float C01 = fma32(J02, XJ[2][1], J00 * XJ[0][1]);
    const float C10 = fma32(J12, XJ[2][0], J11 * XJ[1][0]);
    const float C11 = fma32(J12, XJ[2][1], J11 * XJ[1][1]);
    const float F00 = C00 + 0.1f, F11 = C11 + 0.1f;
    const float detC = fma32(C00, C11, -(C01 * C10)), detF = fma32(F00, F11, -(C01 * C10));
    if (detC > 0.0f && detF > 0.0f) {
      const float k = sqrtf(detC / detF);
      const float m = ((g[0] * opacity3) * 0.5f) * k;
      G00 = fma32(m, C11 / detC - F11 / detF, G00);
      G01 = fma32(m, -C01 / detC + C01 / detF, G01);
      G11 = fma32(m, C00 / detC - F00 / detF, G11);
      k2 = k;
    } else {
      k2 = 0.0f;
    }
  }
  // Position: the centre's (unclamped) Jacobian, and tr(G dJ X J^T) for dJ along x, y, z.
  const float Jc02 = -(c.fx * tx) * iz2, Jc12 = -(c.fy * ty) * iz2;
  float gT0 = J00 * gc0, gT1 = J11 * gc1, gT2 = fma32(Jc12, gc1, Jc02 * gc0);
  const float ax = clampedX ? 0.0f : -c.fx * iz2, ay = clampedY ? 0.0f : -c.fy * iz2;
  const float zx = (c.fx * ux) * iz2, zy = (c.fy * uy) * iz2;
  const float az0 = -c.fx * iz2, az1 = -c.fy * iz2, az2 = clampedX ? zx : 2.0f * zx, az3 = clampedY ? zy : 2.0f * zy;
  const float trX = ax * fma32(G01, XJ[2][1], G00 * XJ[2][0]);
  const float trY = ay * fma32(G11, XJ[2][1], G01 * XJ[2][0]);
  const float N00 = fma32(az2, XJ[2][0], az0 * XJ[0][0]), N01 = fma32(az2, XJ[2][1], az0 * XJ[0][1]);
  const float N10 = fma32(az3, XJ[2][0], az1 * XJ[1][0]), N11 = fma32(az3, XJ[2][1], az1 * XJ[1][1]);
  const float trZ = fma32(G11, N11, fma32(G01, N10, fma32(G01, N01, G00 * N00)));
  gT0 = fma32(2.0f, trX, gT0);
  gT1 = fma32(2.0f, trY, gT1);
  gT2 = fma32(2.0f, trZ, gT2);
#pragma unroll
  for (int k = 0; k < 3; ++k) t.position[k] = dot3(V[k * 4 + 0], V[k * 4 + 1], V[k * 4 + 2], gT0, gT1, gT2);

  // The 3D covariance: G3 = T^T G T, dL = tr(G3 dSigma).
  float GT[2][3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    GT[0][k] = fma32(G01, T[1][k], G00 * T[0][k]);
    GT[1][k] = fma32(G11, T[1][k], G01 * T[0][k]);
  }
  float G3[3][3];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int k = 0; k < 3; ++k) G3[r][k] = fma32(T[1][r], GT[1][k], T[0][r] * GT[0][k]);
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    float v[3];
#pragma unroll
    for (int r = 0; r < 3; ++r) v[r] = dot3(G3[r][0], G3[r][1], G3[r][2], R[0][a], R[1][a], R[2][a]);
    const float quad = dot3(R[0][a], R[1][a], R[2][a], v[0], v[1], v[2]);
    const float filterTerm = f != 0.0f ? ((((g[0] * k2) * opacity) * k3) * f) / (s2[a] + f) : 0.0f;
    t.scale[a] = fma32(o.scaleReg, G.e[a], fma32(2.0f * s2[a], quad, filterTerm));
  }
  // Rotation: P = G3 R D, dL/dq = 2 <2P, dR/dq> written out per component.
  float P[3][3];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int k = 0; k < 3; ++k) P[r][k] = dot3(G3[r][0], G3[r][1], G3[r][2], R[0][k], R[1][k], R[2][k]) * s2[k];
  const float w = qw, x = qx, y = qy, z = qz;
  const float sw = fma32(x, P[2][1] - P[1][2], fma32(y, P[0][2] - P[2][0], z * (P[1][0] - P[0][1])));
  const float sx = fma32(w, P[2][1] - P[1][2], fma32(z, P[0][2] + P[2][0], y * (P[0][1] + P[1][0]))) - 2.0f * x * (P[1][1] + P[2][2]);
  const float sy = fma32(w, P[0][2] - P[2][0], fma32(z, P[1][2] + P[2][1], x * (P[0][1] + P[1][0]))) - 2.0f * y * (P[0][0] + P[2][2]);
  const float sz = fma32(w, P[1][0] - P[0][1], fma32(y, P[1][2] + P[2][1], x * (P[0][2] + P[2][0]))) - 2.0f * z * (P[0][0] + P[1][1]);
  const float gq[4] = {4.0f * sw, 4.0f * sx, 4.0f * sy, 4.0f * sz};
  const float qd = fma32(z, gq[3], fma32(y, gq[2], fma32(x, gq[1], w * gq[0])));
  const float qv[4] = {w, x, y, z};
  const float denom = qlen > 1e-12f ? qlen : 1e-12f;
  const float idenom = 1.0f / denom;
#pragma unroll
  for (int k = 0; k < 4; ++k) t.rotation[k] = fma32(-qv[k], qd, gq[k]) * idenom;

  t.opacity = (fma32(g[0] * k2, k3, o.opacityReg) * opacity) * (1.0f - opacity);

  if (withDir) {
#pragma unroll
    for (int a = 0; a < 3; ++a) t.position[a] = t.position[a] + dirTerm[a];
  }
}

// Both parts in one thread (the batched chain rule).
template <int Degree, uint32_t Format, typename ShSink>
__device__ __forceinline__ void view_terms(const GaussianFrame &G, const ViewParams &c, const ChainParams &o, uint32_t i,
                                           const uint32_t *words, const float g[kRawPerSplat], ViewTerms &t, ShSink &&sh) {
  float dirTerm[3];
  view_sh_terms<Degree, Format>(c, o, i, G.p, words, g, dirTerm, sh);
  // Keep the scheduler from hoisting the geometry's work into the SH section.
  __builtin_amdgcn_sched_barrier(0);
  view_geometry_terms(G, c, o, g, dirTerm, Degree > 0, t);
}

// untouched_terms from a Gaussian's frame (the same values).
__device__ __forceinline__ void untouched_terms(const GaussianFrame &G, const ChainParams &o, bool penalties, ViewTerms &t) {
  const float op = penalties ? o.opacityReg * G.opacity * (1.0f - G.opacity) : 0.0f;
#pragma unroll
  for (uint32_t a = 0; a < 3; ++a) {
    t.position[a] = 0.0f;
    t.scale[a] = penalties ? o.scaleReg * G.e[a] : 0.0f;
  }
#pragma unroll
  for (uint32_t a = 0; a < 4; ++a) t.rotation[a] = 0.0f;
  t.opacity = op;
}

__device__ __forceinline__ void emit_terms(const ChainBuffers &B, const ChainParams &o, uint32_t i, const ViewTerms &t, const Radii &r) {
#pragma unroll
  for (uint32_t a = 0; a < 3; ++a) {
    emit(o, B.positions, B.statePos, B.gradPos, i * 3 + a, t.position[a], o.rates[0], r.position, 0);
    emit(o, B.scales, B.stateScale, B.gradScale, i * 3 + a, t.scale[a], o.rates[1], r.scale, 1);
  }
  if (o.mode != kModeCurvature) {
#pragma unroll
    for (uint32_t a = 0; a < 4; ++a) emit(o, B.rotations, B.stateRot, B.gradRot, i * 4 + a, t.rotation[a], o.rates[2], -1.0f, 2);
  }
  emit(o, B.opacities, B.stateOpacity, B.gradOpacity, i, t.opacity, o.rates[3], r.opacity, 3);
}

constexpr uint32_t kChainThreads = 256;

// The SH gradients of one Gaussian row: zero (or unchanged, accumulating) where
// value(j) is zero; rows of a multiple of four floats go out four at a time.
template <typename Value>
__device__ __forceinline__ void store_sh_row(float *out, uint32_t floats, bool add, Value &&value) {
  if ((floats & 3u) == 0u) {
    float4 *out4 = reinterpret_cast<float4 *>(out);
#pragma unroll
    for (uint32_t q = 0; q < 12; ++q) {
      if (4 * q >= floats) break;
      float4 v = make_float4(value(4 * q), value(4 * q + 1), value(4 * q + 2), value(4 * q + 3));
      if (add) {
        const float4 p = out4[q];
        v = make_float4(p.x + v.x, p.y + v.y, p.z + v.z, p.w + v.w);
      }
      out4[q] = v;
    }
  } else {
    for (uint32_t j = 0; j < floats; ++j) out[j] = add ? out[j] + value(j) : value(j);
  }
}

// One view, in two kernels over the Gaussians: the SH part (colour, SH
// gradients or the SH step's inputs, and the view direction's term of the
// position gradient, kept in dirTerms), then the geometric part with the
// statistics and the emits. Split, each needs far fewer registers than the
// whole: the chain rule's long dependent chains (divisions, square roots, the
// fixed-point conversion) are hidden by more waves per SIMD. The SH part runs
// first because a step changes the positions it reads.
template <int Degree, uint32_t Format>
__device__ __forceinline__ void chain_sh_gradients(const ChainBuffers &B, const ViewParams &c, const ChainParams &o, uint32_t i,
                                                   float *shRows);

// In gradients mode the SH gradient rows go through LDS (dynamic shared
// memory, kShRowStride floats per row, 64 rows per wave): each lane writes its
// row there, then the block stores its rows, which lie consecutively in
// gradSh, with consecutive lanes on consecutive floats. Stored one row per
// lane, every store instruction would touch 64 lines 192 bytes apart.
constexpr uint32_t kShRowStride = 49; // odd, so that a row's lanes hit distinct banks
constexpr uint32_t kShRowsLds = kChainThreads * kShRowStride * 4;

template <int Degree, uint32_t Format>
__global__ __launch_bounds__(kChainThreads) void chain_sh_kernel(ChainBuffers B, ViewParams c, ChainParams o, const uint32_t *counters) {
  extern __shared__ float shRows[];
  if (counters[kOverflow] != 0) return; // a view did not fit: this pass computed nothing (uniform)
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (o.mode == kModeGradients) {
    chain_sh_gradients<Degree, Format>(B, c, o, i, shRows);
    return;
  }
  if (i >= c.count) return;
  float g[kRawPerSplat];
  const bool touched = load_raw(o.fixedPoint ? B.rawFixed : nullptr, B.rawFloat, o.fixedScale, i, g);
  const bool zeros = o.mode == kModeGradients || (o.mode == kModeStep && (o.flags & kFlagSelective) == 0);
  const bool penalties = o.opacityReg > 0.0f || o.scaleReg > 0.0f;
  float *shPass = B.shPass + static_cast<size_t>(i) * kShPassFloats;
  if (o.mode == kModeStatistics || !touched) {
    // Nothing from the pixels: no SH step, and zero SH gradients where the
    // single kernel wrote them (gradients mode, or penalties in step mode).
    if (o.mode == kModeStep || o.mode == kModeCurvature) shPass[9] = as_float(0u);
    if (o.mode == kModeGradients && (zeros || penalties)) {
      store_sh_row(B.gradSh + static_cast<size_t>(i) * o.shFloats, o.shFloats, (o.flags & kFlagAccumulate) != 0,
                   [](uint32_t) { return 0.0f; });
    }
    return;
  }
  constexpr uint32_t kCoefficients = (Degree + 1) * (Degree + 1);
  float dirTerm[3];
  const float position[3] = {B.positions[i * 3], B.positions[i * 3 + 1], B.positions[i * 3 + 2]};
  view_sh_terms<Degree, Format>(c, o, i, position, B.shWords + static_cast<size_t>(i) * c.shWords, g, dirTerm,
                                [&](const float basis[16], const float gcu[3], const float dir[3], uint32_t used) {
    if (o.mode == kModeGradients) {
      // Coefficients above the active degree get zero (they are stored, not trained).
      store_sh_row(B.gradSh + static_cast<size_t>(i) * o.shFloats, o.shFloats, (o.flags & kFlagAccumulate) != 0, [&](uint32_t j) {
        const uint32_t k = j / 3, ch = j - 3 * k;
        return j < kCoefficients * 3 && k < used ? basis[k] * gcu[ch] : 0.0f;
      });
    } else {
      // For the SH pass: the colour gradient, the direction and the coefficients used.
      shPass[0] = gcu[0];
      shPass[1] = gcu[1];
      shPass[2] = gcu[2];
      shPass[3] = dir[0];
      shPass[4] = dir[1];
      shPass[5] = dir[2];
      if ((o.flags & kFlagSecondOrder) != 0) {
        const float opacity = 1.0f / (1.0f + vs_exp(-B.opacities[i]));
        const uint32_t *words = B.shWords + static_cast<size_t>(i) * c.shWords;
#pragma unroll
        for (uint32_t ch = 0; ch < 3; ++ch) shPass[6 + ch] = sh_radius(o, opacity, sh_value<Format>(words, 0, ch));
      }
      shPass[9] = as_float(used);
    }
  });
  B.dirTerms[i] = make_float4(dirTerm[0], dirTerm[1], dirTerm[2], 0.0f);
}

// chain_sh_kernel's gradients mode: the same terms, each lane's row (zero for
// a Gaussian no pixel touched, or above the active degree) written to LDS, then
// the block's rows stored (or added, kFlagAccumulate) in order.
template <int Degree, uint32_t Format>
__device__ __forceinline__ void chain_sh_gradients(const ChainBuffers &B, const ViewParams &c, const ChainParams &o, uint32_t i,
                                                   float *shRows) {
  const uint32_t floats = o.shFloats;
  float *row = shRows + threadIdx.x * kShRowStride;
  if (i < c.count) {
    float g[kRawPerSplat];
    const bool touched = load_raw(o.fixedPoint ? B.rawFixed : nullptr, B.rawFloat, o.fixedScale, i, g);
    for (uint32_t j = 0; j < floats; ++j) row[j] = 0.0f;
    if (touched) {
      constexpr uint32_t kCoefficients = (Degree + 1) * (Degree + 1);
      float dirTerm[3];
      const float position[3] = {B.positions[i * 3], B.positions[i * 3 + 1], B.positions[i * 3 + 2]};
      view_sh_terms<Degree, Format>(c, o, i, position, B.shWords + static_cast<size_t>(i) * c.shWords, g, dirTerm,
                                    [&](const float basis[16], const float gcu[3], const float *, uint32_t used) {
        for (uint32_t k = 0; k < kCoefficients && k < used; ++k) {
#pragma unroll
          for (uint32_t ch = 0; ch < 3; ++ch) row[3 * k + ch] = basis[k] * gcu[ch];
        }
      });
      B.dirTerms[i] = make_float4(dirTerm[0], dirTerm[1], dirTerm[2], 0.0f);
    }
  }
  __syncthreads();
  const uint32_t first = blockIdx.x * blockDim.x;
  const uint32_t rows = c.count - first < blockDim.x ? c.count - first : blockDim.x;
  const uint32_t total = rows * floats;
  float *out = B.gradSh + static_cast<size_t>(first) * floats;
  const bool add = (o.flags & kFlagAccumulate) != 0;
  // Element e of the block's rows is row e / floats, column e % floats; kept
  // incrementally as e advances by the block size.
  uint32_t r = threadIdx.x / floats, col = threadIdx.x - r * floats;
  const uint32_t stepRows = blockDim.x / floats, stepCols = blockDim.x - stepRows * floats;
  for (uint32_t e = threadIdx.x; e < total; e += blockDim.x) {
    const float v = shRows[r * kShRowStride + col];
    out[e] = add ? out[e] + v : v;
    r += stepRows;
    col += stepCols;
    if (col >= floats) {
      col -= floats;
      ++r;
    }
  }
}

__global__ __launch_bounds__(kChainThreads) void chain_geometry_kernel(ChainBuffers B, ViewParams c, ChainParams o,
                                                                      const uint32_t *counters) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= c.count) return;
  if (counters[kOverflow] != 0) return;
  if (o.mode == kModeStep) adam_step_now(o);
  float g[kRawPerSplat];
  const bool touched = load_raw(o.fixedPoint ? B.rawFixed : nullptr, B.rawFloat, o.fixedScale, i, g);
  const bool zeros = o.mode == kModeGradients || (o.mode == kModeStep && (o.flags & kFlagSelective) == 0);
  const bool penalties = o.opacityReg > 0.0f || o.scaleReg > 0.0f;
  if ((touched && (o.flags & kFlagStatistics) != 0) && B.stats != nullptr) {
    float4 *st = reinterpret_cast<float4 *>(B.stats + static_cast<size_t>(i) * 8);
    float4 a = st[0], b = st[1];
    add_statistics(a, b, g, static_cast<float>(c.width), static_cast<float>(c.height), B.splats[i].radius);
    st[0] = a;
    st[1] = b;
  }
  if (o.mode == kModeStatistics || (!touched && !zeros && !penalties)) return;
  ViewTerms t;
  if (!touched) {
    // Nothing from the pixels; the penalties still apply.
    untouched_terms(B, o, i, penalties, t);
  } else {
    const float4 d = B.dirTerms[i];
    const float dirTerm[3] = {d.x, d.y, d.z};
    GaussianFrame G;
    gaussian_frame(B, i, G);
    view_geometry_terms(G, c, o, g, dirTerm, c.shDegree > 0, t);
    const float *gP = t.position;
    if ((o.flags & kFlagPose) != 0 && o.mode != kModeCurvature && (gP[0] != 0.0f || gP[1] != 0.0f || gP[2] != 0.0f)) {
      // How the loss moves with the camera through this Gaussian's position
      // (train_gradient.comp:276-279): v = Rv gP, w = t x v.
      float tx, ty, tz;
      to_camera(c, B.positions[i * 3], B.positions[i * 3 + 1], B.positions[i * 3 + 2], tx, ty, tz);
      const float *V = c.view;
      float v[3];
#pragma unroll
      for (int r = 0; r < 3; ++r) v[r] = dot3(V[r], V[4 + r], V[8 + r], gP[0], gP[1], gP[2]);
      float *out = B.poseTerms + static_cast<size_t>(i) * 6;
      out[0] = fma32(ty, v[2], -(tz * v[1]));
      out[1] = fma32(tz, v[0], -(tx * v[2]));
      out[2] = fma32(tx, v[1], -(ty * v[0]));
      out[3] = v[0];
      out[4] = v[1];
      out[5] = v[2];
    }
  }
  // Everything above read the parameters before any of them changes.
  emit_terms(B, o, i, t, radii_of(o, B.opacities, B.scales, i));
}

// The batched chain rule (no reference counterpart): for each Gaussian, every
// view of the batch in turn, each view's terms added to the sums in view order
// in float32 (the first view's are the sums' start), so the result is the
// per-view chain rule's gradients added up in that order. Unlike the
// single-view chain rule it stays one kernel: split, the second half would
// read every view's raw sums and direction terms again (about 1 GB for eight
// views at 1M Gaussians), which measured slower. What does not depend on the
// view (the Gaussian's frame) is computed once, and each view's raw sums are
// loaded while the view before is worked, which the kernel's low occupancy
// (its LDS) needs to keep loads in flight. The geometric sums stay in
// registers and the SH sums in LDS, one row per coefficient so that a wave's
// lanes hit consecutive banks. In gradients mode the sums are stored; in step
// mode the geometric parameters take Adam's step from them here, and the SH
// sums are stored for the SH step.
template <int Degree, uint32_t Format>
__global__ __launch_bounds__(kChainThreads) void chain_views_kernel(ChainBuffers B, ChainViews views, ChainParams o,
                                                                   const uint32_t *counters) {
  constexpr uint32_t kCoefficients = (Degree + 1) * (Degree + 1);
  __shared__ float shSum[kCoefficients * 3][kChainThreads];
  __shared__ uint32_t stage[kChainThreads * 26];
  const uint32_t lane = threadIdx.x;
  const uint32_t i = blockIdx.x * blockDim.x + lane;
  const uint32_t count = views.view[0].c.count, shWords = views.view[0].c.shWords;
  if (counters[kOverflow] != 0) return;
  {
    // The block's SH words, read by every view: staged once through LDS with
    // 16-byte loads (measured faster here than each view reading them again).
    const uint32_t first = blockIdx.x * kChainThreads;
    const uint32_t inBlock = count - first < kChainThreads ? count - first : kChainThreads;
    stage_words<kChainThreads>(stage, B.shWords + static_cast<size_t>(first) * shWords, inBlock * shWords);
    __syncthreads();
  }
  if (i >= count) return;
  const uint32_t *words = stage + lane * shWords;
  if (o.mode == kModeStep) adam_step_now(o);
  const bool penalties = o.opacityReg > 0.0f || o.scaleReg > 0.0f;
  const bool statistics = (o.flags & kFlagStatistics) != 0 && B.stats != nullptr;
  // A later chunk of a batch continues from the sums the earlier ones stored.
  ViewTerms sum;
  bool started = (o.flags & kFlagAccumulate) != 0;
  if (started) {
#pragma unroll
    for (uint32_t k = 0; k < 3; ++k) {
      sum.position[k] = B.gradPos[i * 3 + k];
      sum.scale[k] = B.gradScale[i * 3 + k];
    }
#pragma unroll
    for (uint32_t k = 0; k < 4; ++k) sum.rotation[k] = B.gradRot[i * 4 + k];
    sum.opacity = B.gradOpacity[i];
    const float *prev = B.gradSh + static_cast<size_t>(i) * o.shFloats;
#pragma unroll
    for (uint32_t j = 0; j < kCoefficients * 3; ++j) shSum[j][lane] = prev[j];
  }
  o.flags &= ~kFlagAccumulate; // the sums below are complete: they are stored, not added
  float4 a = make_float4(0.0f, 0.0f, 0.0f, 0.0f), b = a;
  bool statsChanged = false;
  GaussianFrame G;
  gaussian_frame(B, i, G);
  // Each view's raw sums are loaded kAhead views ahead of the view being worked,
  // in kAhead register sets that the unrolled loop uses in turn: two where the
  // kernel stays within 256 VGPRs with them (two waves per SIMD), one for
  // degree 3 with 16-bit SH, whose larger SH terms leave no room.
  constexpr uint32_t kAhead = Degree == 3 && Format != kShBf16 ? 1u : 2u;
  uint32_t raw[kAhead][2 * kRawPerSplat];
  float radius[kAhead];
#pragma unroll
  for (uint32_t k = 0; k < kAhead; ++k) {
    radius[k] = 0.0f;
    if (k < views.count) fetch_raw(views.view[k], i, statistics, raw[k], radius[k]);
  }
#pragma unroll
  for (uint32_t v = 0; v < kChainViewsPerLaunch; ++v) {
    if (v >= views.count) break;
    const ChainView &w = views.view[v];
    const uint32_t set = v % kAhead;
    float g[kRawPerSplat];
    const bool touched = convert_fetched(raw[set], w.rawFixed != nullptr, w.fixedScale, g);
    const float viewRadius = radius[set];
    if (v + kAhead < views.count) fetch_raw(views.view[v + kAhead], i, statistics, raw[set], radius[set]);
    if (touched && statistics) {
      if (!statsChanged) {
        const float4 *st = reinterpret_cast<const float4 *>(B.stats + static_cast<size_t>(i) * 8);
        a = st[0];
        b = st[1];
        statsChanged = true;
      }
      add_statistics(a, b, g, static_cast<float>(w.c.width), static_cast<float>(w.c.height), viewRadius);
    }
    ViewTerms t;
    if (touched) {
      view_terms<Degree, Format>(G, w.c, o, i, words, g, t, [&](const float basis[16], const float gcu[3], const float[3], uint32_t used) {
#pragma unroll
        for (uint32_t k = 0; k < kCoefficients; ++k)
#pragma unroll
          for (uint32_t ch = 0; ch < 3; ++ch) {
            const float x = k < used ? basis[k] * gcu[ch] : 0.0f;
            shSum[k * 3 + ch][lane] = started ? shSum[k * 3 + ch][lane] + x : x;
          }
      });
    } else {
      untouched_terms(G, o, penalties, t);
#pragma unroll
      for (uint32_t j = 0; j < kCoefficients * 3; ++j) shSum[j][lane] = started ? shSum[j][lane] + 0.0f : 0.0f;
    }
#pragma unroll
    for (uint32_t k = 0; k < 3; ++k) {
      sum.position[k] = started ? sum.position[k] + t.position[k] : t.position[k];
      sum.scale[k] = started ? sum.scale[k] + t.scale[k] : t.scale[k];
    }
#pragma unroll
    for (uint32_t k = 0; k < 4; ++k) sum.rotation[k] = started ? sum.rotation[k] + t.rotation[k] : t.rotation[k];
    sum.opacity = started ? sum.opacity + t.opacity : t.opacity;
    started = true;
  }
  if (statsChanged) {
    float4 *st = reinterpret_cast<float4 *>(B.stats + static_cast<size_t>(i) * 8);
    st[0] = a;
    st[1] = b;
  }
  // Coefficients above the active degree get zero (they are stored, not trained).
  float *out = B.gradSh + static_cast<size_t>(i) * o.shFloats;
#pragma unroll
  for (uint32_t j = 0; j < kCoefficients * 3; ++j) out[j] = shSum[j][lane];
  for (uint32_t j = kCoefficients * 3; j < o.shFloats; ++j) out[j] = 0.0f;
  emit_terms(B, o, i, sum, Radii{-1.0f, -1.0f, -1.0f}); // a batch takes Adam's step
}

// The SH floats packed word q holds: [first, last].
__device__ __forceinline__ void sh_word_floats(uint32_t format, uint32_t q, uint32_t shFloats, uint32_t &first, uint32_t &last) {
  if (format == kShBf16) {
    first = 2 * q;
  } else {
    first = q < 3 ? q : 3 + 2 * (q - 3);
    if (q < 3) {
      last = q;
      return;
    }
  }
  last = first + 1 < shFloats - 1 ? first + 1 : shFloats - 1;
}

// SH float j (of the word whose first float is `first`) read from, and stored
// back into, its packed word. Stored values are rounded at random to the
// storage type (train_step.glsl:409-419 for the halves); the DC floats of
// kShHalf are stored as they are.
__device__ __forceinline__ float sh_word_get(uint32_t format, uint32_t word, uint32_t j, uint32_t first) {
  if (format == kShBf16) return bf16_load(word, j - first);
  if (j < 3) return as_float(word);
  return half_to_float(((j - 3) & 1u) ? (word >> 16) : (word & 0xFFFFu));
}
__device__ __forceinline__ uint32_t sh_word_put(uint32_t format, uint32_t word, uint32_t j, uint32_t first, float value, uint32_t s,
                                                uint32_t seed) {
  if (format == kShBf16) return bf16_store(word, j - first, bf16_stochastic(value, s, seed ^ 0x5DEECE66u));
  if (j < 3) return as_uint(value);
  const uint32_t shift = 16u * ((j - 3) & 1u);
  return (word & ~(0xFFFFu << shift)) | (half_stochastic(value, s, seed ^ 0x5DEECE66u) << shift);
}

__global__ __launch_bounds__(256) void sh_step_kernel(uint32_t *__restrict__ shWords, uint32_t *__restrict__ stateSh,
                                                      const float *__restrict__ shPass, ViewParams c, ChainParams o,
                                                      const uint32_t *counters) {
  const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t words = c.shWords;
  if (w >= c.count * words) return;
  if (counters[kOverflow] != 0) return;
  if (o.mode == kModeStep) adam_step_now(o);
  const uint32_t i = w / words, q = w - i * words;
  const float *g = shPass + static_cast<size_t>(i) * kShPassFloats;
  const uint32_t used = as_uint(g[9]);
  const bool secondOrder = (o.flags & kFlagSecondOrder) != 0;
  const bool zeros = o.mode == kModeStep && (o.flags & kFlagSelective) == 0 && !secondOrder;
  if (used == 0 && !zeros) return;
  float basis[16];
  sh_basis(g[3], g[4], g[5], c.shDegree, basis);
  uint32_t first, last;
  sh_word_floats(c.shFormat, q, o.shFloats, first, last);
  uint32_t word = shWords[w];
  bool changed = false;
  for (uint32_t j = first; j <= last; ++j) {
    const uint32_t k = j / 3, ch = j - 3 * k;
    const float gv = k < used ? basis[k] * g[ch] : 0.0f;
    if (gv == 0.0f && !zeros) continue;
    const uint32_t s = i * o.shFloats + j;
    uint32_t state = stateSh[s];
    if (o.mode == kModeCurvature) {
      stateSh[s] = with_curvature(o, state, gv, s * 5u + 4u);
      continue;
    }
    float value = sh_word_get(c.shFormat, word, j, first);
    step_parameter(o, value, state, gv, k == 0 ? o.rates[4] : o.rates[4] * o.restScale, secondOrder ? g[6 + ch] : -1.0f, s * 5u + 4u);
    stateSh[s] = state;
    word = sh_word_put(c.shFormat, word, j, first, value, s, o.seed);
    changed = true;
  }
  if (changed) shWords[w] = word;
}

// One Adam step from summed gradients (the batched step): thread t takes
// parameter t of positions, scales, rotations and opacities in turn, with the
// same rule and the same rounding stream (index at * 5 + group) as the chain
// rule's own step, so a batch of one view steps exactly as the chain rule does.
__global__ __launch_bounds__(256) void adam_kernel(ChainBuffers B, ViewParams c, ChainParams o, uint32_t *counters) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (counters[kOverflow] != 0) {
    if (t == 0) counters[kSkipped] += 1; // this step did not happen (the device step count stays)
    return;
  }
  adam_step_now(o);
  const uint32_t n = c.count;
  if (t < 3 * n) {
    emit(o, B.positions, B.statePos, nullptr, t, B.gradPos[t], o.rates[0], -1.0f, 0);
  } else if (t < 6 * n) {
    emit(o, B.scales, B.stateScale, nullptr, t - 3 * n, B.gradScale[t - 3 * n], o.rates[1], -1.0f, 1);
  } else if (t < 10 * n) {
    emit(o, B.rotations, B.stateRot, nullptr, t - 6 * n, B.gradRot[t - 6 * n], o.rates[2], -1.0f, 2);
  } else if (t < 11 * n) {
    emit(o, B.opacities, B.stateOpacity, nullptr, t - 10 * n, B.gradOpacity[t - 10 * n], o.rates[3], -1.0f, 3);
  }
}

// The SH part of that step, per packed word (as sh_step_kernel, with the
// gradient read from the summed SH gradients).
__global__ __launch_bounds__(256) void sh_adam_kernel(uint32_t *__restrict__ shWords, uint32_t *__restrict__ stateSh,
                                                      const float *__restrict__ gradSh, ViewParams c, ChainParams o,
                                                      const uint32_t *counters) {
  const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t words = c.shWords;
  if (w >= c.count * words) return;
  if (counters[kOverflow] != 0) return;
  adam_step_now(o);
  const uint32_t i = w / words, q = w - i * words;
  const bool zeros = (o.flags & kFlagSelective) == 0;
  uint32_t first, last;
  sh_word_floats(c.shFormat, q, o.shFloats, first, last);
  uint32_t word = shWords[w];
  bool changed = false;
  for (uint32_t j = first; j <= last; ++j) {
    const uint32_t s = i * o.shFloats + j;
    const float gv = gradSh[s];
    if (gv == 0.0f && !zeros) continue;
    uint32_t state = stateSh[s];
    float value = sh_word_get(c.shFormat, word, j, first);
    step_parameter(o, value, state, gv, j < 3 ? o.rates[4] : o.rates[4] * o.restScale, -1.0f, s * 5u + 4u);
    stateSh[s] = state;
    word = sh_word_put(c.shFormat, word, j, first, value, s, o.seed);
    changed = true;
  }
  if (changed) shWords[w] = word;
}

// After a step's last kernel: one more step taken, unless the step was skipped.
__global__ void advance_adam_kernel(uint32_t *state, float beta1, float beta2, const uint32_t *counters) {
  if (counters[kOverflow] != 0) return;
  state[0] += 1u;
  state[1] = as_uint(as_float(state[1]) * beta1);
  state[2] = as_uint(as_float(state[2]) * beta2);
}

// After a scored view (train_score.comp:17-26): each Gaussian's marked pixels
// (raw sum 12, over 3 pixels of the image like the derivatives, times
// toPixels), that count times the view's loss, and its edge score (raw sum 13
// likewise) added to the second half of its statistics.
__global__ __launch_bounds__(256) void score_kernel(const int64_t *__restrict__ rawFixed, const float *__restrict__ rawFloat,
                                                    float fixedScale, float4 *__restrict__ stats, uint32_t count, float toPixels,
                                                    float loss) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= count) return;
  float g[kRawPerSplat];
  (void)load_raw(rawFixed, rawFloat, fixedScale, i, g);
  float4 more = stats[static_cast<size_t>(i) * 2 + 1];
  const float pixels = g[12] * toPixels;
  more.y = more.y + pixels;
  more.z = more.z + pixels * loss;
  more.w = more.w + g[13] * toPixels;
  stats[static_cast<size_t>(i) * 2 + 1] = more;
}

// 3DGS-MCMC's exploration (train_noise.comp:30-70): each Gaussian's position
// moves by scaler * gate(1 - opacity) * Sigma * n, with n three standard
// normals (Box-Muller from a hash of the index and the iteration), Sigma its
// covariance and the gate a steep logistic, so that only the nearly transparent
// wander. iterationDev, when set, holds the step count on the device (after a
// step: the step just taken), and a step that was skipped moves nothing.
__device__ __forceinline__ uint32_t pcg_hash(uint32_t v) {
  const uint32_t state = v * 747796405u + 2891336453u;
  const uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
  return (word >> 22u) ^ word;
}
__global__ __launch_bounds__(256) void noise_kernel(float *__restrict__ positions, const float *__restrict__ scales,
                                                   const float *__restrict__ rotations, const float *__restrict__ opacities, uint32_t count,
                                                   float scaler, const uint32_t *iterationDev, uint32_t iterationHost,
                                                   const uint32_t *counters) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= count) return;
  if (counters != nullptr && counters[kOverflow] != 0) return;
  const uint32_t iteration = iterationDev != nullptr ? iterationDev[0] : iterationHost;
  const float opacity = 1.0f / (1.0f + vs_exp(-opacities[i]));
  const float gate = 1.0f / (1.0f + vs_exp(-100.0f * ((1.0f - opacity) - 0.995f)));
  if (gate < 1e-6f) return;
  float u[4];
#pragma unroll
  for (uint32_t k = 0; k < 4; ++k) {
    // Uniform in (0, 1], never 0, so that the logarithm is finite.
    u[k] = (static_cast<float>(pcg_hash(pcg_hash(i * 4u + k) ^ (iteration * 0x9E3779B9u))) + 1.0f) / 4294967297.0f;
  }
  const float r0 = sqrtf(-2.0f * vs_log(u[0])), r1 = sqrtf(-2.0f * vs_log(u[2]));
  const float twoPi = 6.283185307f;
  const float n[3] = {r0 * cosf(twoPi * u[1]), r0 * sinf(twoPi * u[1]), r1 * cosf(twoPi * u[3])};
  float q[4] = {rotations[i * 4], rotations[i * 4 + 1], rotations[i * 4 + 2], rotations[i * 4 + 3]};
  const float length = sqrtf(fma32(q[3], q[3], fma32(q[2], q[2], fma32(q[1], q[1], q[0] * q[0]))));
  if (length > 0.0f) {
    for (int k = 0; k < 4; ++k) q[k] = q[k] / length;
  } else {
    q[0] = 1.0f;
    q[1] = q[2] = q[3] = 0.0f;
  }
  float R[3][3];
  quat_to_mat(q[0], q[1], q[2], q[3], R);
  const float s2[3] = {vs_exp(2.0f * scales[i * 3]), vs_exp(2.0f * scales[i * 3 + 1]), vs_exp(2.0f * scales[i * 3 + 2])};
  // Sigma n = R (s2 * (R^T n)).
  float m[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) m[k] = s2[k] * dot3(R[0][k], R[1][k], R[2][k], n[0], n[1], n[2]);
  float move[3];
  bool finite = true;
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    move[r] = (scaler * gate) * dot3(R[r][0], R[r][1], R[r][2], m[0], m[1], m[2]);
    finite = finite && __builtin_isfinite(move[r]); // NaN or infinity skips the Gaussian (train_noise.comp)
  }
  if (!finite) return;
#pragma unroll
  for (int r = 0; r < 3; ++r) positions[i * 3 + r] = positions[i * 3 + r] + move[r];
}

// The pose gradient's per-group totals (train_gradient.comp:323-337): the
// terms of 256 Gaussians summed by a tree in LDS (element l plus element
// l + s, s = 128, 64, ..., 1), one float per component, so the result does not
// depend on scheduling. The host sums the groups in double, in order.
__global__ __launch_bounds__(256) void pose_totals_kernel(const float *__restrict__ terms, uint32_t count, float *__restrict__ totals) {
  __shared__ float lds[6][256];
  const uint32_t l = threadIdx.x, i = blockIdx.x * 256u + l;
#pragma unroll
  for (int k = 0; k < 6; ++k) lds[k][l] = i < count ? terms[static_cast<size_t>(i) * 6 + k] : 0.0f;
  __syncthreads();
  for (uint32_t s = 128; s >= 1; s >>= 1) {
    if (l < s) {
#pragma unroll
      for (int k = 0; k < 6; ++k) lds[k][l] = lds[k][l] + lds[k][l + s];
    }
    __syncthreads();
  }
  if (l < 6) totals[blockIdx.x * 6u + l] = lds[l][0];
}

// Trainer::reset_opacity (trainer.cpp:810-819): opacities above the limit (a
// logit) brought down to it.
__global__ __launch_bounds__(256) void clamp_opacity_kernel(float *__restrict__ opacities, uint32_t count, float limit) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i < count) opacities[i] = opacities[i] > limit ? limit : opacities[i];
}

// Second-order training: an adjoint of random signs, one per pixel and
// channel (train_hutchinson.comp:15-29). Taken through the backward pass it
// gives J^T u, whose square has the diagonal of J^T J as its expectation
// (Hutchinson's estimator in image space).
__device__ __forceinline__ uint32_t mix_bits(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7FEB352Du;
  x ^= x >> 15;
  x *= 0x846CA68Bu;
  x ^= x >> 16;
  return x;
}
__global__ __launch_bounds__(256) void hutchinson_kernel(float4 *__restrict__ adjoint, uint32_t pixels, uint32_t seed) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= pixels) return;
  const uint32_t bits = mix_bits(i ^ mix_bits(seed));
  adjoint[i] = make_float4((bits & 1u) ? 1.0f : -1.0f, (bits & 2u) ? 1.0f : -1.0f, (bits & 4u) ? 1.0f : -1.0f, 0.0f);
}

// Trainer::enable_second_order (trainer.cpp:642-654): the curvature takes the
// place of the second moment, from zero; the first moment is kept.
__global__ __launch_bounds__(256) void keep_first_moment_kernel(uint32_t *__restrict__ state, uint64_t count) {
  for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < count; i += static_cast<uint64_t>(gridDim.x) * 256u) state[i] &= 0xFFFFu;
}

} // namespace

// Calls LAUNCH(degree, format) with both as compile-time constants.
#define VSPLAT_SH_DISPATCH(degree, format, LAUNCH)                                                                                 \
  switch ((degree) * 2 + ((format) == kShBf16 ? 1 : 0)) {                                                                          \
  case 0: LAUNCH(0, kShHalf); break;                                                                                               \
  case 1: LAUNCH(0, kShBf16); break;                                                                                               \
  case 2: LAUNCH(1, kShHalf); break;                                                                                               \
  case 3: LAUNCH(1, kShBf16); break;                                                                                               \
  case 4: LAUNCH(2, kShHalf); break;                                                                                               \
  case 5: LAUNCH(2, kShBf16); break;                                                                                               \
  case 6: LAUNCH(3, kShHalf); break;                                                                                               \
  default: LAUNCH(3, kShBf16); break;                                                                                              \
  }

void launch_chain(hipStream_t stream, const ChainBuffers &buffers, const ViewParams &view, const ChainParams &params,
                  const uint32_t *counters) {
  if (view.count == 0) return;
  const dim3 grid((view.count + 255) / 256), block(256);
  const uint32_t rowsLds = params.mode == kModeGradients ? kShRowsLds : 0u;
#define VSPLAT_CHAIN(D, F) hipLaunchKernelGGL((chain_sh_kernel<D, F>), grid, block, rowsLds, stream, buffers, view, params, counters)
  VSPLAT_SH_DISPATCH(view.shDegree, view.shFormat, VSPLAT_CHAIN)
#undef VSPLAT_CHAIN
  hipLaunchKernelGGL(chain_geometry_kernel, grid, block, 0, stream, buffers, view, params, counters);
  if (params.mode == kModeStep || params.mode == kModeCurvature) {
    const uint64_t words = static_cast<uint64_t>(view.count) * view.shWords;
    hipLaunchKernelGGL(sh_step_kernel, dim3(static_cast<uint32_t>((words + 255) / 256)), dim3(256), 0, stream, buffers.shWords,
                       buffers.stateSh, buffers.shPass, view, params, counters);
    if (params.mode == kModeStep && params.adamState != nullptr) {
      hipLaunchKernelGGL(advance_adam_kernel, dim3(1), dim3(1), 0, stream, params.adamState, params.beta1, params.beta2, counters);
    }
  }
}

void launch_chain_views(hipStream_t stream, const ChainBuffers &buffers, const ChainViews &views, const ChainParams &params,
                        const uint32_t *counters) {
  const uint32_t count = views.view[0].c.count;
  if (count == 0 || views.count == 0) return;
  const dim3 grid((count + 255) / 256), block(256);
#define VSPLAT_CHAIN(D, F) hipLaunchKernelGGL((chain_views_kernel<D, F>), grid, block, 0, stream, buffers, views, params, counters)
  VSPLAT_SH_DISPATCH(views.view[0].c.shDegree, views.view[0].c.shFormat, VSPLAT_CHAIN)
#undef VSPLAT_CHAIN
}

void launch_sh_adam(hipStream_t stream, const ChainBuffers &buffers, const ViewParams &view, const ChainParams &params,
                    const uint32_t *counters) {
  if (view.count == 0) return;
  const uint64_t words = static_cast<uint64_t>(view.count) * view.shWords;
  hipLaunchKernelGGL(sh_adam_kernel, dim3(static_cast<uint32_t>((words + 255) / 256)), dim3(256), 0, stream, buffers.shWords,
                     buffers.stateSh, buffers.gradSh, view, params, counters);
  if (params.adamState != nullptr) {
    hipLaunchKernelGGL(advance_adam_kernel, dim3(1), dim3(1), 0, stream, params.adamState, params.beta1, params.beta2, counters);
  }
}

void launch_noise(hipStream_t stream, const ChainBuffers &buffers, uint32_t count, float scaler, const uint32_t *iterationDev,
                  uint32_t iterationHost, const uint32_t *counters) {
  if (count == 0) return;
  hipLaunchKernelGGL(noise_kernel, dim3((count + 255) / 256), dim3(256), 0, stream, buffers.positions, buffers.scales, buffers.rotations,
                     buffers.opacities, count, scaler, iterationDev, iterationHost, counters);
}

void launch_pose_totals(hipStream_t stream, const float *terms, uint32_t count, float *totals) {
  if (count == 0) return;
  hipLaunchKernelGGL(pose_totals_kernel, dim3(pose_groups(count)), dim3(256), 0, stream, terms, count, totals);
}

void launch_clamp_opacity(hipStream_t stream, float *opacities, uint32_t count, float limit) {
  if (count == 0) return;
  hipLaunchKernelGGL(clamp_opacity_kernel, dim3((count + 255) / 256), dim3(256), 0, stream, opacities, count, limit);
}

void launch_hutchinson(hipStream_t stream, float4 *adjoint, uint32_t pixels, uint32_t seed) {
  if (pixels == 0) return;
  hipLaunchKernelGGL(hutchinson_kernel, dim3((pixels + 255) / 256), dim3(256), 0, stream, adjoint, pixels, seed);
}

void launch_keep_first_moment(hipStream_t stream, uint32_t *state, uint64_t count) {
  if (count == 0) return;
  const uint64_t blocks64 = (count + 255) / 256;
  hipLaunchKernelGGL(keep_first_moment_kernel, dim3(blocks64 < 8192 ? static_cast<uint32_t>(blocks64) : 8192u), dim3(256), 0, stream, state,
                     count);
}

void launch_score(hipStream_t stream, const int64_t *rawFixed, const float *rawFloat, float fixedScale, float4 *stats, uint32_t count,
                  float toPixels, float loss) {
  if (count == 0) return;
  hipLaunchKernelGGL(score_kernel, dim3((count + 255) / 256), dim3(256), 0, stream, rawFixed, rawFloat, fixedScale, stats, count,
                     toPixels, loss);
}

void launch_adam(hipStream_t stream, const ChainBuffers &buffers, const ViewParams &view, const ChainParams &params,
                 uint32_t *counters) {
  if (view.count == 0) return;
  const uint64_t parameters = static_cast<uint64_t>(view.count) * 11u;
  hipLaunchKernelGGL(adam_kernel, dim3(static_cast<uint32_t>((parameters + 255) / 256)), dim3(256), 0, stream, buffers, view, params,
                     counters);
  const uint64_t words = static_cast<uint64_t>(view.count) * view.shWords;
  hipLaunchKernelGGL(sh_adam_kernel, dim3(static_cast<uint32_t>((words + 255) / 256)), dim3(256), 0, stream, buffers.shWords,
                     buffers.stateSh, buffers.gradSh, view, params, counters);
  if (params.adamState != nullptr) {
    hipLaunchKernelGGL(advance_adam_kernel, dim3(1), dim3(1), 0, stream, params.adamState, params.beta1, params.beta2, counters);
  }
}

} // namespace vsplat
