// Test infrastructure, not product code. Oracle A: a double-precision
// restatement of the reference's CPU reference, src/train/reference.cpp, in this
// project's own code (no glm: the matrices are written out in dmath.hpp).
// Each function cites the lines it follows. The finite-difference tests in
// tests/test_oracle_a.py pin it the same way the reference pins its own
// (tests/test_train.cpp:86-182).
//
// Work is split over horizontal bands of pixel rows, one thread per band, with
// per-thread accumulators reduced in thread order, so results do not depend
// on the thread count beyond the order of those final sums.
#include "dmath.hpp"
#include "oracle_api.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <thread>
#include <vector>

namespace oracle {
namespace {

// reference.cpp:12-15
constexpr double kWeights[11] = {0.00102838, 0.00759876, 0.03600077, 0.10936069, 0.21300554, 0.26601172,
                                 0.21300554, 0.10936069, 0.03600077, 0.00759876, 0.00102838};
constexpr double kC1 = 0.0001, kC2 = 0.0009;
constexpr double kLogMinTransmittance = -9.210340372;

struct Splat { // reference.cpp:24-31
  std::size_t index = 0;
  D2 center;
  double radius = 0.0, depth = 0.0;
  D3 conic; // xx, xy, yy
  double opacity = 0.0;
  D3 color;
};

struct Options {
  D3 background;
  double nearPlane = 0.01;
  bool noCutoffs = false;
  double opacityReg = 0.0, scaleReg = 0.0;
  bool mipFilter2d = false;
  const float *filter3d = nullptr;
  std::size_t filter3dCount = 0;
  double filter(std::size_t i) const { return i < filter3dCount ? static_cast<double>(filter3d[i]) : 0.0; }
};

Options options_of(const oa_options *o) {
  Options r;
  if (o == nullptr) return r;
  r.background = {o->background[0], o->background[1], o->background[2]};
  r.nearPlane = o->near_plane;
  r.noCutoffs = o->no_cutoffs != 0;
  r.opacityReg = o->opacity_reg;
  r.scaleReg = o->scale_reg;
  r.mipFilter2d = o->mip_filter_2d != 0;
  r.filter3d = o->filter3d;
  r.filter3dCount = o->filter3d != nullptr ? o->filter3d_count : 0;
  return r;
}

// The camera's view as a 4x4 in row/column order: view[c*4+r] is glm's view[c][r].
struct View {
  double v[4][4];
  D3 position;
  double fx, fy, cx, cy;
  int width, height;
};
View view_of(const oc_camera *c) {
  View r;
  for (int col = 0; col < 4; ++col)
    for (int row = 0; row < 4; ++row) r.v[row][col] = c->view[col * 4 + row];
  r.position = {c->position[0], c->position[1], c->position[2]};
  r.fx = c->fx;
  r.fy = c->fy;
  r.cx = c->cx;
  r.cy = c->cy;
  r.width = c->width;
  r.height = c->height;
  return r;
}
M3 rotation_of(const View &v) {
  M3 r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r.m[i][j] = v.v[i][j];
  return r;
}
D3 transform(const View &v, D3 p) {
  return {v.v[0][0] * p.x + v.v[0][1] * p.y + v.v[0][2] * p.z + v.v[0][3], v.v[1][0] * p.x + v.v[1][1] * p.y + v.v[1][2] * p.z + v.v[1][3],
          v.v[2][0] * p.x + v.v[2][1] * p.y + v.v[2][2] * p.z + v.v[2][3]};
}

// reference.cpp:33-38. glm's columns (1-2(y2+z2), 2(xy+wz), 2(xz-wy)), ... as rows/columns.
M3 quat_to_mat(double w, double x, double y, double z) {
  M3 r;
  r.m[0][0] = 1.0 - 2.0 * (y * y + z * z);
  r.m[1][0] = 2.0 * (x * y + w * z);
  r.m[2][0] = 2.0 * (x * z - w * y);
  r.m[0][1] = 2.0 * (x * y - w * z);
  r.m[1][1] = 1.0 - 2.0 * (x * x + z * z);
  r.m[2][1] = 2.0 * (y * z + w * x);
  r.m[0][2] = 2.0 * (x * z + w * y);
  r.m[1][2] = 2.0 * (y * z - w * x);
  r.m[2][2] = 1.0 - 2.0 * (x * x + y * y);
  return r;
}

// reference.cpp:40-67
std::array<double, 16> sh_basis(D3 dir, int degree) {
  const double x = dir.x, y = dir.y, z = dir.z, xx = x * x, yy = y * y, zz = z * z;
  constexpr double C1 = 0.4886025119029199;
  std::array<double, 16> b{};
  b[0] = 0.28209479177387814;
  if (degree > 0) {
    b[1] = -C1 * y;
    b[2] = C1 * z;
    b[3] = -C1 * x;
  }
  if (degree > 1) {
    b[4] = 1.0925484305920792 * x * y;
    b[5] = -1.0925484305920792 * y * z;
    b[6] = 0.31539156525252005 * (2.0 * zz - xx - yy);
    b[7] = -1.0925484305920792 * x * z;
    b[8] = 0.5462742152960396 * (xx - yy);
  }
  if (degree > 2) {
    b[9] = -0.5900435899266435 * y * (3.0 * xx - yy);
    b[10] = 2.890611442640554 * x * y * z;
    b[11] = -0.4570457994644658 * y * (4.0 * zz - xx - yy);
    b[12] = 0.3731763325901154 * z * (2.0 * zz - 3.0 * xx - 3.0 * yy);
    b[13] = -0.4570457994644658 * x * (4.0 * zz - xx - yy);
    b[14] = 1.445305721320277 * z * (xx - yy);
    b[15] = -0.5900435899266435 * x * (xx - 3.0 * yy);
  }
  return b;
}

// reference.cpp:69-96
std::array<D3, 16> sh_basis_gradient(D3 dir, int degree) {
  constexpr double C1 = 0.4886025119029199, A = 1.0925484305920792, K = 0.31539156525252005, L = 0.5462742152960396;
  constexpr double C = -0.5900435899266435, E = 2.890611442640554, F = -0.4570457994644658, G = 0.3731763325901154,
                   H = 1.445305721320277;
  const double x = dir.x, y = dir.y, z = dir.z, xx = x * x, yy = y * y, zz = z * z;
  std::array<D3, 16> g{};
  if (degree > 0) {
    g[1] = {0.0, -C1, 0.0};
    g[2] = {0.0, 0.0, C1};
    g[3] = {-C1, 0.0, 0.0};
  }
  if (degree > 1) {
    g[4] = A * D3{y, x, 0.0};
    g[5] = -A * D3{0.0, z, y};
    g[6] = K * D3{-2.0 * x, -2.0 * y, 4.0 * z};
    g[7] = -A * D3{z, 0.0, x};
    g[8] = L * D3{2.0 * x, -2.0 * y, 0.0};
  }
  if (degree > 2) {
    g[9] = C * D3{6.0 * x * y, 3.0 * xx - 3.0 * yy, 0.0};
    g[10] = E * D3{y * z, x * z, x * y};
    g[11] = F * D3{-2.0 * x * y, 4.0 * zz - xx - 3.0 * yy, 8.0 * y * z};
    g[12] = G * D3{-6.0 * x * z, -6.0 * y * z, 6.0 * zz - 3.0 * xx - 3.0 * yy};
    g[13] = F * D3{4.0 * zz - 3.0 * xx - yy, -2.0 * x * y, 8.0 * x * z};
    g[14] = H * D3{2.0 * x * z, -2.0 * y * z, xx - yy};
    g[15] = C * D3{3.0 * xx - 3.0 * yy, -6.0 * x * y, 0.0};
  }
  return g;
}

struct Cloud {
  const oc_cloud *c;
  std::size_t size() const { return c->count; }
  int coefficients() const { return (c->sh_degree + 1) * (c->sh_degree + 1); }
  D3 position(std::size_t i) const { return {c->positions[i * 3], c->positions[i * 3 + 1], c->positions[i * 3 + 2]}; }
  // reference.cpp:102-107
  void quat(std::size_t i, double q[4], double &len) const {
    for (int k = 0; k < 4; ++k) q[k] = c->rotations[i * 4 + k];
    len = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    if (len > 0.0) {
      for (int k = 0; k < 4; ++k) q[k] /= len;
    } else {
      q[0] = 1.0;
      q[1] = q[2] = q[3] = 0.0;
    }
  }
  D3 sh(std::size_t i, int k) const {
    const std::size_t at = (i * static_cast<std::size_t>(coefficients()) + static_cast<std::size_t>(k)) * 3;
    return {c->sh[at], c->sh[at + 1], c->sh[at + 2]};
  }
};

D3 normalize(D3 v) {
  const double l = length(v);
  return {v.x / l, v.y / l, v.z / l};
}

// reference.cpp:109-176: splats, stably sorted back to front.
std::vector<Splat> prepare(const Cloud &cloud, const View &cam, const Options &options, int threads) {
  const M3 Rv = rotation_of(cam);
  const double fx = cam.fx, fy = cam.fy, cx = cam.cx, cy = cam.cy;
  const double limX = 1.3 * 0.5 * cam.width / fx, limY = 1.3 * 0.5 * cam.height / fy;
  const std::size_t n = cloud.size();
  std::vector<Splat> all(n);
  std::vector<unsigned char> keep(n, 0);
  const auto work = [&](std::size_t begin, std::size_t end) {
    for (std::size_t i = begin; i < end; ++i) {
      const D3 p = cloud.position(i);
      const D3 t = transform(cam, p);
      if (!(t.z >= options.nearPlane)) continue;
      const double opacity = 1.0 / (1.0 + std::exp(-static_cast<double>(cloud.c->opacities[i])));
      if (!(opacity >= 1.0 / 255.0)) continue;
      double q[4], quatLength = 0.0;
      cloud.quat(i, q, quatLength);
      if (!(quatLength >= 1e-4)) continue;
      const D3 s{std::exp(static_cast<double>(cloud.c->scales[i * 3])), std::exp(static_cast<double>(cloud.c->scales[i * 3 + 1])),
                 std::exp(static_cast<double>(cloud.c->scales[i * 3 + 2]))};
      const M3 M = quat_to_mat(q[0], q[1], q[2], q[3]) * M3::diag(s.x, s.y, s.z);
      const double filter = options.filter(i);
      const M3 sigma = M * transpose(M) + M3::diag(filter, filter, filter);
      double opacity3 = opacity;
      for (int a = 0; a < 3; ++a) opacity3 *= std::sqrt(s[a] * s[a] / (s[a] * s[a] + filter));
      if (!(opacity3 >= 1.0 / 255.0)) continue;

      const double tx = std::clamp(t.x / t.z, -limX, limX) * t.z, ty = std::clamp(t.y / t.z, -limY, limY) * t.z;
      // reference.cpp:136: J with rows (fx/z, 0, -fx tx/z^2) and (0, fy/z, -fy ty/z^2), third row zero.
      M3 J;
      J.m[0][0] = fx / t.z;
      J.m[1][1] = fy / t.z;
      J.m[0][2] = -fx * tx / (t.z * t.z);
      J.m[1][2] = -fy * ty / (t.z * t.z);
      const M3 T = J * Rv;
      const M3 cov = T * sigma * transpose(T);
      const double dilation = options.mipFilter2d ? 0.1 : 0.3;
      const double a = cov.m[0][0] + dilation, b = cov.m[1][0], c = cov.m[1][1] + dilation;
      const double det = a * c - b * b;
      if (!(det > 1e-6)) continue;
      double opacityUsed = opacity3;
      if (options.mipFilter2d) {
        const double rawDet = cov.m[0][0] * cov.m[1][1] - b * b;
        opacityUsed = rawDet > 0.0 ? opacity3 * std::sqrt(rawDet / det) : 0.0;
        if (!(opacityUsed >= 1.0 / 255.0)) continue;
      }
      const double mid = 0.5 * (a + c);
      const double lambda = mid + std::sqrt(std::max(0.1, mid * mid - det));

      Splat out;
      out.index = i;
      out.radius = std::ceil(3.0 * std::sqrt(lambda));
      out.center = {fx * t.x / t.z + cx, fy * t.y / t.z + cy};
      if (out.center.x + out.radius < 0.0 || out.center.y + out.radius < 0.0 || out.center.x - out.radius > cam.width ||
          out.center.y - out.radius > cam.height) {
        continue;
      }
      out.depth = t.z;
      out.conic = {c / det, -b / det, a / det};
      out.opacity = opacityUsed;
      const auto basis = sh_basis(normalize(p - cam.position), cloud.c->sh_degree);
      D3 color{0.5, 0.5, 0.5};
      for (int k = 0; k < cloud.coefficients(); ++k) color = color + basis[k] * cloud.sh(i, k);
      if (!(std::abs(color.x) < 1e30 && std::abs(color.y) < 1e30 && std::abs(color.z) < 1e30)) continue;
      out.color = {std::max(color.x, 0.0), std::max(color.y, 0.0), std::max(color.z, 0.0)};
      all[i] = out;
      keep[i] = 1;
    }
  };
  const std::size_t parts = static_cast<std::size_t>(std::max(1, threads));
  std::vector<std::thread> pool;
  for (std::size_t t = 0; t < parts; ++t) pool.emplace_back(work, n * t / parts, n * (t + 1) / parts);
  for (auto &t : pool) t.join();
  std::vector<Splat> splats;
  for (std::size_t i = 0; i < n; ++i)
    if (keep[i]) splats.push_back(all[i]);
  std::stable_sort(splats.begin(), splats.end(), [](const Splat &a, const Splat &b) { return a.depth > b.depth; });
  return splats;
}

// reference.cpp:178-186
bool splat_at(const Splat &s, int x, int y, D2 &d, double &power, double &alpha, bool noCutoffs) {
  d = {x + 0.5 - s.center.x, y + 0.5 - s.center.y};
  if (!noCutoffs && (std::abs(d.x) > s.radius || std::abs(d.y) > s.radius)) return false;
  power = -0.5 * (s.conic.x * d.x * d.x + s.conic.z * d.y * d.y) - s.conic.y * d.x * d.y;
  if (power > 0.0) return false;
  alpha = std::min(0.99, s.opacity * std::exp(power));
  return noCutoffs || alpha >= 1.0 / 255.0;
}

// The pixel box of a splat, reference.cpp:207-211.
void bounds(const Splat &s, const View &cam, bool noCutoffs, int &x0, int &x1, int &y0, int &y1) {
  const double reach = noCutoffs ? 1e6 : s.radius;
  x0 = std::max(0, static_cast<int>(std::floor(s.center.x - reach)) - 1);
  x1 = std::min(cam.width - 1, static_cast<int>(std::ceil(s.center.x + reach)) + 1);
  y0 = std::max(0, static_cast<int>(std::floor(s.center.y - reach)) - 1);
  y1 = std::min(cam.height - 1, static_cast<int>(std::ceil(s.center.y + reach)) + 1);
}

struct Image {
  int width = 0, height = 0;
  std::vector<D3> rgb;
  std::vector<double> logT;
  std::vector<std::size_t> stop;
};

template <class F> void for_bands(int height, int threads, F &&f) {
  const int parts = std::max(1, std::min(threads, height));
  std::vector<std::thread> pool;
  for (int t = 0; t < parts; ++t) pool.emplace_back(f, t, height * t / parts, height * (t + 1) / parts);
  for (auto &t : pool) t.join();
}

// reference.cpp:193-234. contributions (per thread, then summed in thread order) when given.
Image composite(const Cloud &cloud, const View &cam, const Options &options, int threads, std::vector<double> *contributions) {
  Image image;
  image.width = cam.width;
  image.height = cam.height;
  const std::size_t pixels = static_cast<std::size_t>(cam.width) * cam.height;
  image.rgb.assign(pixels, D3{});
  image.logT.assign(pixels, 0.0);
  image.stop.assign(pixels, 0);
  std::vector<unsigned char> done(pixels, 0);
  const std::vector<Splat> splats = prepare(cloud, cam, options, threads);
  const int parts = std::max(1, std::min(threads, cam.height));
  std::vector<std::vector<double>> partial(contributions != nullptr ? static_cast<std::size_t>(parts) : 0);
  for_bands(cam.height, parts, [&](int part, int rowBegin, int rowEnd) {
    std::vector<double> *mine = nullptr;
    if (contributions != nullptr) {
      partial[static_cast<std::size_t>(part)].assign(cloud.size(), 0.0);
      mine = &partial[static_cast<std::size_t>(part)];
    }
    for (std::size_t j = splats.size(); j-- > 0;) {
      const Splat &s = splats[j];
      int x0, x1, y0, y1;
      bounds(s, cam, options.noCutoffs, x0, x1, y0, y1);
      y0 = std::max(y0, rowBegin);
      y1 = std::min(y1, rowEnd - 1);
      for (int y = y0; y <= y1; ++y) {
        for (int x = x0; x <= x1; ++x) {
          D2 d;
          double power = 0.0, alpha = 0.0;
          if (!splat_at(s, x, y, d, power, alpha, options.noCutoffs)) continue;
          const std::size_t at = static_cast<std::size_t>(y) * cam.width + x;
          if (done[at]) continue;
          const double logRest = std::log(1.0 - alpha);
          if (!options.noCutoffs && image.logT[at] + logRest < kLogMinTransmittance) {
            image.stop[at] = j + 1;
            done[at] = 1;
            continue;
          }
          image.rgb[at] = image.rgb[at] + (std::exp(image.logT[at]) * alpha) * s.color;
          if (mine != nullptr) (*mine)[s.index] += std::exp(image.logT[at]) * alpha;
          image.logT[at] += logRest;
        }
      }
    }
  });
  for (std::size_t at = 0; at < pixels; ++at) image.rgb[at] = image.rgb[at] + std::exp(image.logT[at]) * options.background;
  if (contributions != nullptr) {
    contributions->assign(cloud.size(), 0.0);
    for (const auto &p : partial)
      for (std::size_t i = 0; i < p.size(); ++i) (*contributions)[i] += p[i];
  }
  return image;
}

double sign(double v) { return v > 0.0 ? 1.0 : (v < 0.0 ? -1.0 : 0.0); }

// tr(G A X J^T), reference.cpp:188-191.
double trace_term(const M2 &G, const M23 &A, const M3 &X, const M23 &J) {
  const M2 product = G * mul_bt(A * X, J);
  return product.m[0][0] + product.m[1][1];
}

} // namespace
} // namespace oracle

using namespace oracle;

extern "C" int oa_render(const oc_cloud *c, const oc_camera *camera, const oa_options *opt, int threads, double *rgb, double *logT,
                         uint64_t *stop) {
  if (c == nullptr || camera == nullptr || camera->width <= 0 || camera->height <= 0) return -1;
  const Image image = composite(Cloud{c}, view_of(camera), options_of(opt), threads, nullptr);
  for (std::size_t i = 0; i < image.rgb.size(); ++i) {
    if (rgb) {
      rgb[i * 3] = image.rgb[i].x;
      rgb[i * 3 + 1] = image.rgb[i].y;
      rgb[i * 3 + 2] = image.rgb[i].z;
    }
    if (logT) logT[i] = image.logT[i];
    if (stop) stop[i] = image.stop[i];
  }
  return 0;
}

// reference.cpp:242-246
extern "C" int oa_contributions(const oc_cloud *c, const oc_camera *camera, const oa_options *opt, int threads, double *out) {
  if (c == nullptr || camera == nullptr || out == nullptr) return -1;
  std::vector<double> contributions;
  (void)composite(Cloud{c}, view_of(camera), options_of(opt), threads, &contributions);
  std::copy(contributions.begin(), contributions.end(), out);
  return 0;
}

// reference.cpp:248-272: per Gaussian, the pixels it is composited into whose
// mean error (over the channels, against the target / 255) is at least
// `share` of the view's worst.
extern "C" int oa_marked_pixels(const oc_cloud *c, const oc_camera *camera, const uint8_t *target, double share,
                                const oa_options *opt, int threads, double *out) {
  if (c == nullptr || camera == nullptr || target == nullptr || out == nullptr) return -1;
  const Cloud cloud{c};
  const View cam = view_of(camera);
  const Options options = options_of(opt);
  const Image image = composite(cloud, cam, options, threads, nullptr);
  std::vector<double> errors(image.rgb.size(), 0.0);
  double worst = 0.0;
  for (std::size_t at = 0; at < errors.size(); ++at) {
    const double rgb[3] = {image.rgb[at].x, image.rgb[at].y, image.rgb[at].z};
    for (std::size_t k = 0; k < 3; ++k) errors[at] += std::abs(rgb[k] - target[at * 4 + k] / 255.0) / 3.0;
    worst = std::max(worst, errors[at]);
  }
  std::fill(out, out + cloud.size(), 0.0);
  const std::vector<Splat> splats = prepare(cloud, cam, options, threads);
  for (std::size_t j = 0; j < splats.size(); ++j) {
    const Splat &s = splats[j];
    int x0, x1, y0, y1;
    bounds(s, cam, options.noCutoffs, x0, x1, y0, y1);
    for (int y = y0; y <= y1; ++y) {
      for (int x = x0; x <= x1; ++x) {
        D2 d;
        double power = 0.0, alpha = 0.0;
        if (!splat_at(s, x, y, d, power, alpha, options.noCutoffs)) continue;
        const std::size_t at = static_cast<std::size_t>(y) * cam.width + x;
        if (j >= image.stop[at] && errors[at] >= share * worst) out[s.index] += 1.0;
      }
    }
  }
  return 0;
}

// reference.cpp:276-341
extern "C" int oa_loss(const double *rgb, const uint8_t *target, int w, int h, double lambda, int threads, double *adjoint,
                       double *stats) {
  if (rgb == nullptr || target == nullptr || w <= 0 || h <= 0) return -1;
  const std::size_t pixels = static_cast<std::size_t>(w) * h;
  const auto image_at = [&](std::size_t at) { return D3{rgb[at * 3], rgb[at * 3 + 1], rgb[at * 3 + 2]}; };
  const auto target_at = [&](std::size_t at) {
    const uint8_t *t = target + at * 4;
    return D3{t[0] / 255.0, t[1] / 255.0, t[2] / 255.0};
  };
  std::vector<D3> P(pixels), Q(pixels), R(pixels);
  const int parts = std::max(1, std::min(threads, h));
  std::vector<double> ssimParts(static_cast<std::size_t>(parts), 0.0), l1Parts(static_cast<std::size_t>(parts), 0.0);
  for_bands(h, parts, [&](int part, int rowBegin, int rowEnd) {
    double ssimSum = 0.0;
    for (int row = rowBegin; row < rowEnd; ++row) {
      for (int px = 0; px < w; ++px) {
        D3 mx, my, xx, yy, xy;
        double norm = 0.0;
        for (int v = std::max(row - 5, 0); v <= std::min(row + 5, h - 1); ++v) {
          for (int u = std::max(px - 5, 0); u <= std::min(px + 5, w - 1); ++u) {
            const double weight = kWeights[v - row + 5] * kWeights[u - px + 5];
            const std::size_t at = static_cast<std::size_t>(v) * w + u;
            const D3 a = image_at(at), b = target_at(at);
            mx = mx + weight * a;
            my = my + weight * b;
            xx = xx + weight * (a * a);
            yy = yy + weight * (b * b);
            xy = xy + weight * (a * b);
            norm += weight;
          }
        }
        mx = {mx.x / norm, mx.y / norm, mx.z / norm};
        my = {my.x / norm, my.y / norm, my.z / norm};
        // reference.cpp:302-312, channel by channel.
        D3 p, q, r;
        for (int k = 0; k < 3; ++k) {
          const double vx = xx[k] / norm - mx[k] * mx[k], vy = yy[k] / norm - my[k] * my[k], cxy = xy[k] / norm - mx[k] * my[k];
          const double n1 = 2.0 * mx[k] * my[k] + kC1, n2 = 2.0 * cxy + kC2, d1 = mx[k] * mx[k] + my[k] * my[k] + kC1,
                       d2 = vx + vy + kC2;
          const double s = (n1 * n2) / (d1 * d2);
          ssimSum += s;
          const double dMu = 2.0 * my[k] * n2 / (d1 * d2) - s * 2.0 * mx[k] / d1, dVar = -s / d2, dCov = 2.0 * n1 / (d1 * d2);
          p[k] = (dMu - 2.0 * mx[k] * dVar - my[k] * dCov) / norm;
          q[k] = dVar / norm;
          r[k] = dCov / norm;
        }
        const std::size_t at = static_cast<std::size_t>(row) * w + px;
        P[at] = p;
        Q[at] = q;
        R[at] = r;
      }
    }
    ssimParts[static_cast<std::size_t>(part)] = ssimSum;
  });
  const double perValue = 1.0 / (3.0 * static_cast<double>(pixels));
  for_bands(h, parts, [&](int part, int rowBegin, int rowEnd) {
    double l1Sum = 0.0;
    for (int row = rowBegin; row < rowEnd; ++row) {
      for (int px = 0; px < w; ++px) {
        const std::size_t at = static_cast<std::size_t>(row) * w + px;
        const D3 x = image_at(at), y = target_at(at), difference = x - y;
        l1Sum += std::abs(difference.x) + std::abs(difference.y) + std::abs(difference.z);
        D3 p, q, r;
        for (int v = std::max(row - 5, 0); v <= std::min(row + 5, h - 1); ++v) {
          for (int u = std::max(px - 5, 0); u <= std::min(px + 5, w - 1); ++u) {
            const double weight = kWeights[v - row + 5] * kWeights[u - px + 5];
            const std::size_t other = static_cast<std::size_t>(v) * w + u;
            p = p + weight * P[other];
            q = q + weight * Q[other];
            r = r + weight * R[other];
          }
        }
        if (adjoint != nullptr) {
          for (int k = 0; k < 3; ++k) {
            adjoint[at * 3 + static_cast<std::size_t>(k)] =
                ((1.0 - lambda) * sign(difference[k]) - lambda * (p[k] + 2.0 * x[k] * q[k] + y[k] * r[k])) * perValue;
          }
        }
      }
    }
    l1Parts[static_cast<std::size_t>(part)] = l1Sum;
  });
  double ssimSum = 0.0, l1Sum = 0.0;
  for (int t = 0; t < parts; ++t) {
    ssimSum += ssimParts[static_cast<std::size_t>(t)];
    l1Sum += l1Parts[static_cast<std::size_t>(t)];
  }
  if (stats != nullptr) {
    stats[1] = l1Sum * perValue;
    stats[2] = ssimSum * perValue;
    stats[0] = (1.0 - lambda) * stats[1] + lambda * (1.0 - stats[2]);
  }
  return 0;
}

// reference.cpp:343-519
extern "C" int oa_gradients(const oc_cloud *c, const oc_camera *camera, const double *adjointIn, const oa_options *opt, int threads,
                            double *gPositions, double *gScales, double *gRotations, double *gOpacities, double *gSh,
                            double *gAbsolute) {
  if (c == nullptr || camera == nullptr || adjointIn == nullptr) return -1;
  const Cloud cloud{c};
  const View cam = view_of(camera);
  const Options options = options_of(opt);
  const Image total = composite(cloud, cam, options, threads, nullptr);
  const std::size_t pixels = static_cast<std::size_t>(cam.width) * cam.height;
  std::vector<D3> behind(pixels, options.background);
  std::vector<double> logBehind(pixels, 0.0);
  const std::size_t n = cloud.size();
  const std::size_t floats = static_cast<std::size_t>(3 * cloud.coefficients());
  std::vector<double> positions(n * 3, 0.0), scales(n * 3, 0.0), rotations(n * 4, 0.0), opacities(n, 0.0), sh(n * floats, 0.0),
      absolute(n * 2, 0.0);

  // reference.cpp:359-401, per band of rows.
  const std::vector<Splat> splats = prepare(cloud, cam, options, threads);
  const int parts = std::max(1, std::min(threads, cam.height));
  std::vector<std::vector<double>> rawParts(static_cast<std::size_t>(parts)), absParts(static_cast<std::size_t>(parts));
  for_bands(cam.height, parts, [&](int part, int rowBegin, int rowEnd) {
    std::vector<double> &raw = rawParts[static_cast<std::size_t>(part)];
    std::vector<double> &abs = absParts[static_cast<std::size_t>(part)];
    raw.assign(n * 9, 0.0);
    abs.assign(n * 2, 0.0);
    for (std::size_t j = 0; j < splats.size(); ++j) {
      const Splat &s = splats[j];
      double *g = raw.data() + s.index * 9;
      int x0, x1, y0, y1;
      bounds(s, cam, options.noCutoffs, x0, x1, y0, y1);
      y0 = std::max(y0, rowBegin);
      y1 = std::min(y1, rowEnd - 1);
      for (int y = y0; y <= y1; ++y) {
        for (int x = x0; x <= x1; ++x) {
          D2 d;
          double power = 0.0, alpha = 0.0;
          if (!splat_at(s, x, y, d, power, alpha, options.noCutoffs)) continue;
          const std::size_t at = static_cast<std::size_t>(y) * cam.width + x;
          if (j < total.stop[at]) continue;
          const double logRest = std::log(1.0 - alpha);
          const double T = std::exp(total.logT[at] - logBehind[at] - logRest);
          const D3 A{adjointIn[at * 3], adjointIn[at * 3 + 1], adjointIn[at * 3 + 2]}, B = behind[at];
          g[6] += T * alpha * A.x;
          g[7] += T * alpha * A.y;
          g[8] += T * alpha * A.z;
          if (s.opacity * std::exp(power) < 0.99) {
            const double dAlpha = T * dot(A, s.color - B);
            g[0] += dAlpha * std::exp(power);
            const double dPower = dAlpha * alpha;
            const D2 v{s.conic.x * d.x + s.conic.y * d.y, s.conic.y * d.x + s.conic.z * d.y};
            g[1] += dPower * v.x;
            g[2] += dPower * v.y;
            abs[s.index * 2] += std::abs(dPower * v.x);
            abs[s.index * 2 + 1] += std::abs(dPower * v.y);
            g[3] += dPower * 0.5 * v.x * v.x;
            g[4] += dPower * 0.5 * v.x * v.y;
            g[5] += dPower * 0.5 * v.y * v.y;
          }
          // glm::mix(B, colour, alpha) = B (1 - alpha) + colour alpha
          behind[at] = (1.0 - alpha) * B + alpha * s.color;
          logBehind[at] += logRest;
        }
      }
    }
  });
  std::vector<double> raw(n * 9, 0.0);
  for (int t = 0; t < parts; ++t) {
    for (std::size_t i = 0; i < raw.size(); ++i) raw[i] += rawParts[static_cast<std::size_t>(t)][i];
    for (std::size_t i = 0; i < absolute.size(); ++i) absolute[i] += absParts[static_cast<std::size_t>(t)][i];
  }
  rawParts.clear();
  absParts.clear();

  // reference.cpp:403-508, the chain rule per Gaussian.
  const M3 Rv = rotation_of(cam);
  const double fx = cam.fx, fy = cam.fy;
  const double limX = 1.3 * 0.5 * cam.width / fx, limY = 1.3 * 0.5 * cam.height / fy;
  const int coefficients = cloud.coefficients();
  const auto chain = [&](std::size_t begin, std::size_t end) {
    for (std::size_t i = begin; i < end; ++i) {
      const double *g = raw.data() + i * 9;
      bool touched = false;
      for (int k = 0; k < 9; ++k) touched = touched || g[k] != 0.0;
      if (!touched) continue;
      const D3 p = cloud.position(i);
      double quat[4], quatLength = 1.0;
      cloud.quat(i, quat, quatLength);
      const D3 t = transform(cam, p);
      const double z = t.z, iz = 1.0 / z, iz2 = iz * iz;
      const double gCentre[2] = {g[1], g[2]};
      M2 G;
      G.m[0][0] = g[3];
      G.m[0][1] = G.m[1][0] = g[4];
      G.m[1][1] = g[5];
      const double opacity = 1.0 / (1.0 + std::exp(-static_cast<double>(c->opacities[i])));
      double k2 = 1.0;
      const M3 R = quat_to_mat(quat[0], quat[1], quat[2], quat[3]);
      const D3 s2{std::exp(2.0 * static_cast<double>(c->scales[i * 3])), std::exp(2.0 * static_cast<double>(c->scales[i * 3 + 1])),
                  std::exp(2.0 * static_cast<double>(c->scales[i * 3 + 2]))};
      const M3 D = M3::diag(s2.x, s2.y, s2.z);
      const double filter = options.filter(i);
      const M3 X = Rv * (R * D * transpose(R) + M3::diag(filter, filter, filter)) * transpose(Rv);
      double k3 = 1.0;
      for (int a = 0; a < 3; ++a) k3 *= std::sqrt(s2[a] / (s2[a] + filter));
      const double opacity3 = opacity * k3;

      const double rx = t.x * iz, ry = t.y * iz;
      const bool clampedX = std::abs(rx) > limX, clampedY = std::abs(ry) > limY;
      const double tx = std::clamp(rx, -limX, limX) * z, ty = std::clamp(ry, -limY, limY) * z;
      M23 J;
      J.m[0][0] = fx * iz;
      J.m[1][1] = fy * iz;
      J.m[0][2] = -fx * tx * iz2;
      J.m[1][2] = -fy * ty * iz2;
      if (options.mipFilter2d) {
        const M2 C = mul_bt(J * X, J);
        M2 F = C;
        F.m[0][0] += 0.1;
        F.m[1][1] += 0.1;
        const double detC = C.m[0][0] * C.m[1][1] - C.m[0][1] * C.m[1][0], detF = F.m[0][0] * F.m[1][1] - F.m[0][1] * F.m[1][0];
        if (detC > 0.0 && detF > 0.0) {
          const double k = std::sqrt(detC / detF);
          const double cofC[2][2] = {{C.m[1][1], -C.m[0][1]}, {-C.m[1][0], C.m[0][0]}};
          const double cofF[2][2] = {{F.m[1][1], -F.m[0][1]}, {-F.m[1][0], F.m[0][0]}};
          for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b) G.m[a][b] += g[0] * opacity3 * 0.5 * k * (cofC[a][b] / detC - cofF[a][b] / detF);
          k2 = k;
        } else {
          k2 = 0.0;
        }
      }
      M23 Jc;
      Jc.m[0][0] = fx * iz;
      Jc.m[1][1] = fy * iz;
      Jc.m[0][2] = -fx * t.x * iz2;
      Jc.m[1][2] = -fy * t.y * iz2;
      D3 gT{Jc.m[0][0] * gCentre[0] + Jc.m[1][0] * gCentre[1], Jc.m[0][1] * gCentre[0] + Jc.m[1][1] * gCentre[1],
            Jc.m[0][2] * gCentre[0] + Jc.m[1][2] * gCentre[1]};
      M23 dJx, dJy, dJz;
      dJx.m[0][2] = clampedX ? 0.0 : -fx * iz2;
      dJy.m[1][2] = clampedY ? 0.0 : -fy * iz2;
      dJz.m[0][0] = -fx * iz2;
      dJz.m[1][1] = -fy * iz2;
      dJz.m[0][2] = clampedX ? fx * tx * iz2 * iz : 2.0 * fx * tx * iz2 * iz;
      dJz.m[1][2] = clampedY ? fy * ty * iz2 * iz : 2.0 * fy * ty * iz2 * iz;
      gT = gT + 2.0 * D3{trace_term(G, dJx, X, J), trace_term(G, dJy, X, J), trace_term(G, dJz, X, J)};
      D3 gP = transpose(Rv) * gT;

      const M23 T = J * Rv;
      const M3 G3 = at_g_a(T, G);
      for (int a = 0; a < 3; ++a) {
        const D3 axis = R.col(a);
        scales[i * 3 + static_cast<std::size_t>(a)] =
            2.0 * s2[a] * dot(axis, G3 * axis) + g[0] * k2 * opacity * k3 * filter / (s2[a] + filter);
      }

      const M3 P = G3 * R * D; // P / 2 of reference.cpp:467
      const double w = quat[0], x = quat[1], y = quat[2], zq = quat[3];
      // reference.cpp:469-472, dR/dw, dR/dx, dR/dy, dR/dz given by glm columns.
      const double cols[4][3][3] = {
          {{0.0, 2.0 * zq, -2.0 * y}, {-2.0 * zq, 0.0, 2.0 * x}, {2.0 * y, -2.0 * x, 0.0}},
          {{0.0, 2.0 * y, 2.0 * zq}, {2.0 * y, -4.0 * x, 2.0 * w}, {2.0 * zq, -2.0 * w, -4.0 * x}},
          {{-4.0 * y, 2.0 * x, -2.0 * w}, {2.0 * x, 0.0, 2.0 * zq}, {2.0 * w, 2.0 * zq, -4.0 * y}},
          {{-4.0 * zq, 2.0 * w, 2.0 * x}, {-2.0 * w, -4.0 * zq, 2.0 * y}, {2.0 * x, 2.0 * y, 0.0}}};
      double gq[4];
      for (int k = 0; k < 4; ++k) {
        double sum = 0.0;
        for (int col = 0; col < 3; ++col)
          for (int row = 0; row < 3; ++row) sum += 2.0 * P.m[row][col] * cols[k][col][row];
        gq[k] = sum;
      }
      const double qdot = quat[0] * gq[0] + quat[1] * gq[1] + quat[2] * gq[2] + quat[3] * gq[3];
      for (int k = 0; k < 4; ++k) rotations[i * 4 + static_cast<std::size_t>(k)] = (gq[k] - quat[k] * qdot) / std::max(quatLength, 1e-12);

      opacities[i] = g[0] * k2 * k3 * opacity * (1.0 - opacity);

      const D3 gColour{g[6], g[7], g[8]};
      const auto basis = sh_basis(normalize(p - cam.position), c->sh_degree);
      D3 colour{0.5, 0.5, 0.5};
      for (int k = 0; k < coefficients; ++k) colour = colour + basis[k] * cloud.sh(i, k);
      if (c->sh_degree > 0) {
        const D3 toSplat = p - cam.position;
        const double distance = length(toSplat);
        const D3 dir{toSplat.x / distance, toSplat.y / distance, toSplat.z / distance};
        const auto basisGradient = sh_basis_gradient(dir, c->sh_degree);
        const D3 unclamped{colour.x > 0.0 ? 1.0 : 0.0, colour.y > 0.0 ? 1.0 : 0.0, colour.z > 0.0 ? 1.0 : 0.0};
        D3 gDir;
        for (int k = 1; k < coefficients; ++k) gDir = gDir + dot(gColour * unclamped, cloud.sh(i, k)) * basisGradient[k];
        gP = gP + (1.0 / distance) * (gDir - dot(dir, gDir) * dir);
      }
      for (int a = 0; a < 3; ++a) positions[i * 3 + static_cast<std::size_t>(a)] = gP[a];
      for (int k = 0; k < coefficients; ++k) {
        for (int ch = 0; ch < 3; ++ch) {
          sh[(i * static_cast<std::size_t>(coefficients) + static_cast<std::size_t>(k)) * 3 + static_cast<std::size_t>(ch)] =
              colour[ch] > 0.0 ? basis[k] * gColour[ch] : 0.0;
        }
      }
    }
  };
  {
    const std::size_t threadCount = static_cast<std::size_t>(std::max(1, threads));
    std::vector<std::thread> pool;
    for (std::size_t t = 0; t < threadCount; ++t) pool.emplace_back(chain, n * t / threadCount, n * (t + 1) / threadCount);
    for (auto &t : pool) t.join();
  }
  // reference.cpp:509-517
  if (options.opacityReg > 0.0 || options.scaleReg > 0.0) {
    const double count = static_cast<double>(n);
    for (std::size_t i = 0; i < n; ++i) {
      const double o = 1.0 / (1.0 + std::exp(-static_cast<double>(c->opacities[i])));
      opacities[i] += options.opacityReg / count * o * (1.0 - o);
      for (std::size_t a = 0; a < 3; ++a) scales[i * 3 + a] += options.scaleReg / count * std::exp(static_cast<double>(c->scales[i * 3 + a]));
    }
  }
  const auto put = [](double *to, const std::vector<double> &from) {
    if (to != nullptr) std::copy(from.begin(), from.end(), to);
  };
  put(gPositions, positions);
  put(gScales, scales);
  put(gRotations, rotations);
  put(gOpacities, opacities);
  put(gSh, sh);
  put(gAbsolute, absolute);
  return 0;
}

extern "C" int oa_quat_to_mat(const double *q, double *R) {
  const M3 m = quat_to_mat(q[0], q[1], q[2], q[3]);
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) R[r * 3 + c] = m.m[r][c];
  return 0;
}

extern "C" int oa_sh_basis(const double *dir, int degree, double *out) {
  const auto b = sh_basis(D3{dir[0], dir[1], dir[2]}, degree);
  for (int k = 0; k < 16; ++k) out[k] = b[k];
  return 0;
}
