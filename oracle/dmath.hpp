// Test infrastructure: small double-precision vector and matrix helpers for the
// FP64 oracle. Matrices are indexed m[row][col] in mathematical order; where
// the reference builds a glm matrix column by column, the oracle transcribes
// the same entries into this row/column order.
#pragma once

#include <cmath>

namespace oracle {

struct D2 {
  double x = 0, y = 0;
};
struct D3 {
  double x = 0, y = 0, z = 0;
  double &operator[](int i) { return i == 0 ? x : i == 1 ? y : z; }
  double operator[](int i) const { return i == 0 ? x : i == 1 ? y : z; }
};
inline D3 operator+(D3 a, D3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline D3 operator-(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline D3 operator*(double s, D3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline D3 operator*(D3 a, D3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
inline double dot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline double length(D3 a) { return std::sqrt(dot(a, a)); }

struct M3 {
  double m[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  static M3 diag(double a, double b, double c) {
    M3 r;
    r.m[0][0] = a;
    r.m[1][1] = b;
    r.m[2][2] = c;
    return r;
  }
  D3 col(int c) const { return {m[0][c], m[1][c], m[2][c]}; }
};
inline M3 operator*(const M3 &a, const M3 &b) {
  M3 r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0;
      for (int k = 0; k < 3; ++k) s += a.m[i][k] * b.m[k][j];
      r.m[i][j] = s;
    }
  return r;
}
inline M3 operator+(const M3 &a, const M3 &b) {
  M3 r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r.m[i][j] = a.m[i][j] + b.m[i][j];
  return r;
}
inline M3 transpose(const M3 &a) {
  M3 r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r.m[i][j] = a.m[j][i];
  return r;
}
inline D3 operator*(const M3 &a, D3 v) {
  return {a.m[0][0] * v.x + a.m[0][1] * v.y + a.m[0][2] * v.z, a.m[1][0] * v.x + a.m[1][1] * v.y + a.m[1][2] * v.z,
          a.m[2][0] * v.x + a.m[2][1] * v.y + a.m[2][2] * v.z};
}

// 2x2 and 2x3 (two rows, three columns) matrices.
struct M2 {
  double m[2][2] = {{0, 0}, {0, 0}};
};
struct M23 {
  double m[2][3] = {{0, 0, 0}, {0, 0, 0}};
};
inline M23 operator*(const M23 &a, const M3 &b) {
  M23 r;
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0;
      for (int k = 0; k < 3; ++k) s += a.m[i][k] * b.m[k][j];
      r.m[i][j] = s;
    }
  return r;
}
// a * b^T for two 2x3 matrices: 2x2.
inline M2 mul_bt(const M23 &a, const M23 &b) {
  M2 r;
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) {
      double s = 0;
      for (int k = 0; k < 3; ++k) s += a.m[i][k] * b.m[j][k];
      r.m[i][j] = s;
    }
  return r;
}
inline M2 operator*(const M2 &a, const M2 &b) {
  M2 r;
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) r.m[i][j] = a.m[i][0] * b.m[0][j] + a.m[i][1] * b.m[1][j];
  return r;
}
// a^T g a for a 2x3 and a 2x2: 3x3.
inline M3 at_g_a(const M23 &a, const M2 &g) {
  M3 r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0;
      for (int p = 0; p < 2; ++p)
        for (int q = 0; q < 2; ++q) s += a.m[p][i] * g.m[p][q] * a.m[q][j];
      r.m[i][j] = s;
    }
  return r;
}

} // namespace oracle
