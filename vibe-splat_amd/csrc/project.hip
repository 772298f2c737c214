// Projection: raw parameters to screen-space splats and depth keys
// (train_preprocess.comp:31-134). One thread per Gaussian; the parameter
// arrays are read as the GaussianCloud layout keeps them (xyz, xyz, wxyz, o),
// contiguous across a wave.
#include "device.hpp"
#include "kernels.hpp"

namespace vsplat {

struct ProjectOptions {
  uint32_t mipFilter2d;
  float shDropout;
  uint32_t shRetained, shSeed;
};

template <int Degree>
__global__ __launch_bounds__(256) void project_kernel(const float *__restrict__ positions, const float *__restrict__ scales,
                                                      const float *__restrict__ rotations, const float *__restrict__ opacities,
                                                      const uint32_t *__restrict__ shWords, const float *__restrict__ filter3d,
                                                      ViewParams c, ProjectOptions opt, Splat *__restrict__ splats,
                                                      uint32_t *__restrict__ keys, uint32_t *__restrict__ values) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= c.count) return;
  keys[i] = kInvisible;
  values[i] = i;
  const float px = positions[i * 3], py = positions[i * 3 + 1], pz = positions[i * 3 + 2];
  float tx, ty, tz;
  to_camera(c, px, py, pz, tx, ty, tz);
  // Written so that NaN fails every test, as the reference does.
  if (!(tz >= c.nearPlane)) return;
  const float opacity = 1.0f / (1.0f + vs_exp(-opacities[i]));
  if (!(opacity >= kMinAlpha)) return;
  float qw = rotations[i * 4], qx = rotations[i * 4 + 1], qy = rotations[i * 4 + 2], qz = rotations[i * 4 + 3];
  const float qlen = sqrtf(fma32(qz, qz, fma32(qy, qy, fma32(qx, qx, qw * qw))));
  if (!(qlen >= 1e-4f)) return;
  const float iq = 1.0f / qlen;
  qw = qw * iq;
  qx = qx * iq;
  qy = qy * iq;
  qz = qz * iq;
  const float sx = vs_exp(scales[i * 3]), sy = vs_exp(scales[i * 3 + 1]), sz = vs_exp(scales[i * 3 + 2]);
  float R[3][3];
  quat_to_mat(qw, qx, qy, qz, R);
  const float s[3] = {sx, sy, sz};
  float M[3][3];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int k = 0; k < 3; ++k) M[r][k] = R[r][k] * s[k];
  const float f = filter3d != nullptr ? filter3d[i] : 0.0f;
  float S[3][3];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int k = 0; k < 3; ++k) S[r][k] = dot3(M[r][0], M[r][1], M[r][2], M[k][0], M[k][1], M[k][2]);
  S[0][0] = S[0][0] + f;
  S[1][1] = S[1][1] + f;
  S[2][2] = S[2][2] + f;
  const float opacity3 = opacity * filter_factor(sx * sx, sy * sy, sz * sz, f);
  if (!(opacity3 >= kMinAlpha)) return;

  // EWA: J, clamped to 1.3x the frustum, times the view rotation.
  const float iz = 1.0f / tz;
  const float limX = 1.3f * c.tanHalfX, limY = 1.3f * c.tanHalfY;
  const float ux = clampf(tx * iz, -limX, limX), uy = clampf(ty * iz, -limY, limY);
  const float J00 = c.fx * iz, J02 = -(c.fx * ux) * iz, J11 = c.fy * iz, J12 = -(c.fy * uy) * iz;
  float T[2][3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    T[0][k] = fma32(J02, c.view[k * 4 + 2], J00 * c.view[k * 4 + 0]);
    T[1][k] = fma32(J12, c.view[k * 4 + 2], J11 * c.view[k * 4 + 1]);
  }
  float U[2][3];
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int k = 0; k < 3; ++k) U[r][k] = dot3(T[r][0], T[r][1], T[r][2], S[0][k], S[1][k], S[2][k]);
  const float a0 = dot3(U[0][0], U[0][1], U[0][2], T[0][0], T[0][1], T[0][2]);
  const float b = dot3(U[0][0], U[0][1], U[0][2], T[1][0], T[1][1], T[1][2]);
  const float c0 = dot3(U[1][0], U[1][1], U[1][2], T[1][0], T[1][1], T[1][2]);
  const float dilation = opt.mipFilter2d ? 0.1f : 0.3f;
  const float a = a0 + dilation, cc = c0 + dilation;
  const float det = fma32(a, cc, -(b * b));
  if (!(det > 1e-6f)) return;
  float opacityUsed = opacity3;
  if (opt.mipFilter2d) {
    const float rawDet = fma32(a0, c0, -(b * b));
    opacityUsed = rawDet > 0.0f ? opacity3 * sqrtf(rawDet / det) : 0.0f;
    if (!(opacityUsed >= kMinAlpha)) return;
  }
  const float mid = 0.5f * (a + cc);
  const float disc = fma32(mid, mid, -det);
  const float lambda = mid + sqrtf(disc > 0.1f ? disc : 0.1f);
  const float radius = ceilf(3.0f * sqrtf(lambda));
  const float centerX = fma32(c.fx, tx * iz, c.cx), centerY = fma32(c.fy, ty * iz, c.cy);
  if (centerX + radius < 0.0f || centerY + radius < 0.0f || centerX - radius > static_cast<float>(c.width) ||
      centerY - radius > static_cast<float>(c.height)) {
    return;
  }
  // Colour from SH along the view direction; the active degree is a template
  // argument so that the coefficient loop unrolls. Dropped coefficients are
  // skipped by selection, which leaves the sum exactly as a shorter loop would.
  const float dx = px - c.position[0], dy = py - c.position[1], dz = pz - c.position[2];
  const float dl = sqrtf(fma32(dz, dz, fma32(dy, dy, dx * dx)));
  const float idl = 1.0f / dl;
  float basis[16];
  sh_basis(dx * idl, dy * idl, dz * idl, Degree, basis);
  const uint32_t used = sh_coefficients_used(i, c.shCoefficients, opt.shDropout, opt.shRetained, opt.shSeed);
  const uint32_t base = i * c.shWords;
  float col[3] = {0.5f, 0.5f, 0.5f};
#pragma unroll
  for (uint32_t k = 0; k < (Degree + 1) * (Degree + 1); ++k) {
#pragma unroll
    for (uint32_t ch = 0; ch < 3; ++ch) {
      const float v = sh_value(shWords, base, k * 3 + ch);
      col[ch] = k < used ? fma32(basis[k], v, col[ch]) : col[ch];
    }
  }
  if (!(fabsf(col[0]) < 1e30f && fabsf(col[1]) < 1e30f && fabsf(col[2]) < 1e30f)) return;
  // The extent of the 1/255 ellipse, which is what binning and the pixel loops use.
  const float reach = reach_of(opacityUsed);
  const float ex = sqrtf(reach * a), ey = sqrtf(reach * cc);
  float4 *out = reinterpret_cast<float4 *>(splats + i);
  const float id = 1.0f / det;
  out[0] = make_float4(cc * id, -b * id, a * id, opacityUsed);
  out[1] = make_float4(maxf(col[0], 0.0f), maxf(col[1], 0.0f), maxf(col[2], 0.0f), radius);
  out[2] = make_float4(centerX, centerY, minf(ex, radius), minf(ey, radius));
  keys[i] = 0xFFFFFFFFu - as_uint(tz);
}

void launch_project(hipStream_t stream, const float *positions, const float *scales, const float *rotations, const float *opacities,
                    const uint32_t *shWords, const float *filter3d, const ViewParams &view, bool mipFilter2d, float shDropout,
                    uint32_t shRetained, uint32_t shSeed, Splat *splats, uint32_t *keys, uint32_t *values) {
  if (view.count == 0) return;
  const ProjectOptions opt{mipFilter2d ? 1u : 0u, shDropout, shRetained, shSeed};
  const dim3 grid((view.count + 255) / 256), block(256);
  switch (view.shDegree) {
  case 0: hipLaunchKernelGGL(project_kernel<0>, grid, block, 0, stream, positions, scales, rotations, opacities, shWords, filter3d, view, opt, splats, keys, values); break;
  case 1: hipLaunchKernelGGL(project_kernel<1>, grid, block, 0, stream, positions, scales, rotations, opacities, shWords, filter3d, view, opt, splats, keys, values); break;
  case 2: hipLaunchKernelGGL(project_kernel<2>, grid, block, 0, stream, positions, scales, rotations, opacities, shWords, filter3d, view, opt, splats, keys, values); break;
  default: hipLaunchKernelGGL(project_kernel<3>, grid, block, 0, stream, positions, scales, rotations, opacities, shWords, filter3d, view, opt, splats, keys, values); break;
  }
}

} // namespace vsplat
