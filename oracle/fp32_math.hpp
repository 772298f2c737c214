// Test infrastructure, not product code: the oracle's own single-precision
// exp and log. The GPU kernels carry an independent copy of the same published
// method (Cody-Waite range reduction + a fitted polynomial, Muller,
// "Elementary Functions", ch. 11), and the parity tests require the two to
// agree bit for bit; the oracle's copy is in turn pinned against the C
// library's expf/logf by an ulp-error sweep (tests/test_oracle_math.py).
//
// Only IEEE-754 basic operations (+ - * / sqrt, fused multiply-add,
// round-to-nearest-even) and integer bit manipulation are used, so the results
// are the same on any IEEE host and on gfx950 with contraction off.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>

namespace oracle {

inline float bits_to_float(std::uint32_t u) {
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}
inline std::uint32_t float_to_bits(float f) {
  std::uint32_t u;
  std::memcpy(&u, &f, 4);
  return u;
}

// e^x. Range reduction x = n ln2 + r with |r| <= ln2/2, ln2 split so that n*ln2_hi
// is exact; e^r = 1 + r + r^2 q(r) with q a degree-4 near-minimax fit of
// (e^r - 1 - r) / r^2; the scaling by 2^n is two exact-or-once-rounded
// multiplications so that overflow and gradual underflow come out right.
inline float fexp(float x) {
  if (x != x) return x + x;
  if (x > 89.0f) x = 89.0f;  // e^89 overflows to +inf in the scaling below
  if (x < -104.0f) return 0.0f;
  const float n = std::rint(x * 0x1.715476p+0f);
  float r = std::fma(n, -0x1.62e4p-1f, x);
  r = std::fma(n, -0x1.7f7d1cp-20f, r);
  float q = std::fma(0x1.6d752p-10f, r, 0x1.120b66p-7f);
  q = std::fma(q, r, 0x1.5554b8p-5f);
  q = std::fma(q, r, 0x1.5554dcp-3f);
  q = std::fma(q, r, 0x1p-1f);
  const float p = std::fma(r * r, q, r) + 1.0f;
  const int ni = static_cast<int>(n);
  const int n1 = ni >> 1, n2 = ni - n1;
  return (p * bits_to_float(static_cast<std::uint32_t>(n1 + 127) << 23)) * bits_to_float(static_cast<std::uint32_t>(n2 + 127) << 23);
}

// log(x). x = m 2^e with m in [sqrt(1/2), sqrt(2)), f = m - 1 (exact),
// log1p(f) = f + f^2 Q(f) with Q a degree-8 near-minimax fit of
// (log1p(f) - f) / f^2, then e ln2 added with ln2 split as in fexp.
inline float flog(float x) {
  if (!(x > 0.0f)) return x == 0.0f ? -INFINITY : (x != x ? x + x : NAN);
  if (x == INFINITY) return x;
  int eadj = 0;
  if (x < 0x1p-126f) {
    x *= 0x1p23f;
    eadj = -23;
  }
  const std::uint32_t u = float_to_bits(x) - 0x3f3504f3u;
  const int e = static_cast<std::int32_t>(u) >> 23;
  const float m = bits_to_float((u & 0x007fffffu) + 0x3f3504f3u);
  const float f = m - 1.0f;
  float q = std::fma(-0x1.4a4f98p-4f, f, 0x1.088dd0p-3f);
  q = std::fma(q, f, -0x1.0bba2ep-3f);
  q = std::fma(q, f, 0x1.2245e4p-3f);
  q = std::fma(q, f, -0x1.548f6ap-3f);
  q = std::fma(q, f, 0x1.99a7bap-3f);
  q = std::fma(q, f, -0x1.0001f8p-2f);
  q = std::fma(q, f, 0x1.555548p-2f);
  q = std::fma(q, f, -0x1.fffffep-2f);
  const float lp = std::fma(f * f, q, f);
  const float ef = static_cast<float>(e + eadj);
  return std::fma(ef, 0x1.62e4p-1f, std::fma(ef, 0x1.7f7d1cp-20f, lp));
}

} // namespace oracle
