// Projection: raw parameters to screen-space splats and depth keys
// (train_preprocess.comp:31-134), and each splat's tile coverage
// (train_bin_count.comp:18-37), in one pass. One thread per Gaussian. A
// workgroup first issues every load it needs: each thread its Gaussian's
// position, scales, rotation and opacity (contiguous across a wave), and the
// workgroup its 256 Gaussians' packed SH words as one contiguous run of
// 16-byte loads into LDS, so that the 104-byte SH records are not read lane by
// lane. The coverage is then counted in Gaussian order while the splat is
// still in registers, by the wave's shared walk over candidate tiles
// (tiles.hpp); the bits it finds are kept for the list pass.
#include "device.hpp"
#include "kernels.hpp"
#include "tiles.hpp"

namespace vsplat {

constexpr uint32_t kProjectThreads = 256;
constexpr uint32_t kMaxShWords = 26; // degree 3: three floats, then 45 halves in 23 words
constexpr uint32_t kStageVectors = (kProjectThreads * kMaxShWords / 4 + kProjectThreads - 1) / kProjectThreads;
static_assert(kProjectThreads * kMaxShWords * 4 >= kProjectThreads * 64, "the stage holds the walk's owners");

struct ProjectOptions {
  uint32_t mipFilter2d;
  float shDropout;
  uint32_t shRetained, shSeed;
};

// A Gaussian's raw parameters as read from the cloud's arrays.
struct GaussianIn {
  float px, py, pz;        // position
  float ls[3];             // log standard deviations
  float qw, qx, qy, qz;    // rotation, not normalised
  float logit;             // opacity
  float filter;            // Mip-Splatting 3D filter, or 0
};

// One Gaussian: false if it is culled, else its splat and depth key. `sh` is
// this Gaussian's packed SH words.
template <int Degree, uint32_t Format>
__device__ __forceinline__ bool project_one(uint32_t i, const GaussianIn &g, const uint32_t *sh, const ViewParams &c,
                                            const ProjectOptions &opt, Splat &out, uint32_t &key) {
  const float px = g.px, py = g.py, pz = g.pz;
  float tx, ty, tz;
  to_camera(c, px, py, pz, tx, ty, tz);
  // Written so that NaN fails every test, as the reference does.
  if (!(tz >= c.nearPlane)) return false;
  const float opacity = 1.0f / (1.0f + vs_exp(-g.logit));
  if (!(opacity >= kMinAlpha)) return false;
  float qw = g.qw, qx = g.qx, qy = g.qy, qz = g.qz;
  const float qlen = sqrtf(fma32(qz, qz, fma32(qy, qy, fma32(qx, qx, qw * qw))));
  if (!(qlen >= 1e-4f)) return false;
  const float iq = 1.0f / qlen;
  qw = qw * iq;
  qx = qx * iq;
  qy = qy * iq;
  qz = qz * iq;
  const float sx = vs_exp(g.ls[0]), sy = vs_exp(g.ls[1]), sz = vs_exp(g.ls[2]);
  float R[3][3];
  quat_to_mat(qw, qx, qy, qz, R);
  const float s[3] = {sx, sy, sz};
  float M[3][3];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int k = 0; k < 3; ++k) M[r][k] = R[r][k] * s[k];
  const float f = g.filter;
  float S[3][3];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int k = 0; k < 3; ++k) S[r][k] = dot3(M[r][0], M[r][1], M[r][2], M[k][0], M[k][1], M[k][2]);
  S[0][0] = S[0][0] + f;
  S[1][1] = S[1][1] + f;
  S[2][2] = S[2][2] + f;
  const float opacity3 = opacity * filter_factor(sx * sx, sy * sy, sz * sz, f);
  if (!(opacity3 >= kMinAlpha)) return false;

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
  float a = a0 + dilation, cc = c0 + dilation;
  float det = fma32(a, cc, -(b * b));
  float rawDet = opt.mipFilter2d ? fma32(a0, c0, -(b * b)) : 0.0f;
  // A near-degenerate covariance (a long, thin splat) loses its small axis in
  // float: in the sums of Sigma and T Sigma T^T, and then in a c - b b. When
  // the float determinant is under the cull threshold or under 2^-12 of a c,
  // the 2D covariance is built again in double from the float M and T, and
  // inverted there. (With the dilation the true determinant is at least 0.09.)
  const bool degenerate = !(det > 1e-6f) || det < 0x1p-12f * (a * cc);
  double da = 0.0, db = 0.0, dc = 0.0, ddet = 0.0;
  if (degenerate) {
    double SD[3][3];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        double acc = 0.0;
#pragma unroll
        for (int j = 0; j < 3; ++j) acc = acc + static_cast<double>(M[r][j]) * static_cast<double>(M[k][j]);
        SD[r][k] = r == k ? acc + static_cast<double>(f) : acc;
      }
    double UD[2][3];
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        UD[r][k] = static_cast<double>(T[r][0]) * SD[0][k] + static_cast<double>(T[r][1]) * SD[1][k] +
                   static_cast<double>(T[r][2]) * SD[2][k];
      }
    const double da0 = UD[0][0] * T[0][0] + UD[0][1] * T[0][1] + UD[0][2] * T[0][2];
    db = UD[0][0] * T[1][0] + UD[0][1] * T[1][1] + UD[0][2] * T[1][2];
    const double dc0 = UD[1][0] * T[1][0] + UD[1][1] * T[1][1] + UD[1][2] * T[1][2];
    da = da0 + dilation;
    dc = dc0 + dilation;
    ddet = da * dc - db * db;
    a = static_cast<float>(da);
    cc = static_cast<float>(dc);
    det = static_cast<float>(ddet);
    if (opt.mipFilter2d) rawDet = static_cast<float>(da0 * dc0 - db * db);
  }
  if (!(det > 1e-6f)) return false;
  float opacityUsed = opacity3;
  if (opt.mipFilter2d) {
    opacityUsed = rawDet > 0.0f ? opacity3 * sqrtf(rawDet / det) : 0.0f;
    if (!(opacityUsed >= kMinAlpha)) return false;
  }
  const float mid = 0.5f * (a + cc);
  const float disc = fma32(mid, mid, -det);
  const float lambda = mid + sqrtf(disc > 0.1f ? disc : 0.1f);
  const float radius = ceilf(3.0f * sqrtf(lambda));
  const float centerX = fma32(c.fx, tx * iz, c.cx), centerY = fma32(c.fy, ty * iz, c.cy);
  if (centerX + radius < 0.0f || centerY + radius < 0.0f || centerX - radius > static_cast<float>(c.width) ||
      centerY - radius > static_cast<float>(c.height)) {
    return false;
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
  float col[3] = {0.5f, 0.5f, 0.5f};
#pragma unroll
  for (uint32_t k = 0; k < (Degree + 1) * (Degree + 1); ++k) {
#pragma unroll
    for (uint32_t ch = 0; ch < 3; ++ch) {
      const float v = sh_value<Format>(sh, 0, k * 3 + ch);
      col[ch] = k < used ? fma32(basis[k], v, col[ch]) : col[ch];
    }
  }
  if (!(fabsf(col[0]) < 1e30f && fabsf(col[1]) < 1e30f && fabsf(col[2]) < 1e30f)) return false;
  // The extent of the 1/255 ellipse, which is what binning and the pixel loops use.
  const float reach = reach_of(opacityUsed);
  const float ex = sqrtf(reach * a), ey = sqrtf(reach * cc);
  const float id = 1.0f / det;
  const float conicX = degenerate ? static_cast<float>(dc / ddet) : cc * id;
  const float conicY = degenerate ? static_cast<float>(-db / ddet) : -b * id;
  const float conicZ = degenerate ? static_cast<float>(da / ddet) : a * id;
  out = Splat{conicX, conicY, conicZ, opacityUsed, maxf(col[0], 0.0f), maxf(col[1], 0.0f), maxf(col[2], 0.0f), radius,
              centerX, centerY, minf(ex, radius), minf(ey, radius)};
  key = 0xFFFFFFFFu - as_uint(tz);
  return true;
}

template <int Degree, uint32_t Format>
__global__ __launch_bounds__(256) void project_kernel(const float *__restrict__ positions, const float *__restrict__ scales,
                                                      const float *__restrict__ rotations, const float *__restrict__ opacities,
                                                      const uint32_t *__restrict__ shWords, const float *__restrict__ filter3d,
                                                      ViewParams c, ProjectOptions opt, uint32_t gridW, uint32_t gridH,
                                                      Splat *__restrict__ splats, uint32_t *__restrict__ keys,
                                                      uint32_t *__restrict__ values, uint32_t *__restrict__ tileCount,
                                                      TileRecord *__restrict__ records, uint32_t *counters, uint32_t *__restrict__ bigList) {
  // The block's SH words, staged with coalesced loads; afterwards the walk's owners.
  __shared__ uint32_t stage[kProjectThreads * kMaxShWords];
  __shared__ uint32_t flags[kProjectThreads]; // per wave: the walk's round marks
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x & ~63u;
  const uint32_t first = blockIdx.x * kProjectThreads;
  const uint32_t inBlock = c.count - first < kProjectThreads ? c.count - first : kProjectThreads;

  // Every load is issued before any of them is needed: the parameters of this
  // thread's Gaussian, and this thread's share of the block's SH words.
  GaussianIn g{};
  if (i < c.count) {
    g.px = positions[i * 3];
    g.py = positions[i * 3 + 1];
    g.pz = positions[i * 3 + 2];
    g.ls[0] = scales[i * 3];
    g.ls[1] = scales[i * 3 + 1];
    g.ls[2] = scales[i * 3 + 2];
    const float4 q = reinterpret_cast<const float4 *>(rotations)[i];
    g.qw = q.x;
    g.qx = q.y;
    g.qy = q.z;
    g.qz = q.w;
    g.logit = opacities[i];
    g.filter = filter3d != nullptr ? filter3d[i] : 0.0f;
  }
  {
    // 256 Gaussians' words start 16-byte aligned (256 * 4 bytes is a multiple of 16).
    const uint32_t words = inBlock * c.shWords;
    const uint4 *src = reinterpret_cast<const uint4 *>(shWords + static_cast<size_t>(first) * c.shWords);
    uint4 part[kStageVectors];
#pragma unroll
    for (uint32_t k = 0; k < kStageVectors; ++k) {
      const uint32_t v = k * kProjectThreads + threadIdx.x;
      part[k] = v < words / 4 ? src[v] : make_uint4(0, 0, 0, 0);
    }
    const uint32_t tail = (words / 4) * 4 + threadIdx.x; // the last words, fewer than four
    const uint32_t last = tail < words ? shWords[static_cast<size_t>(first) * c.shWords + tail] : 0u;
#pragma unroll
    for (uint32_t k = 0; k < kStageVectors; ++k) {
      const uint32_t v = k * kProjectThreads + threadIdx.x;
      if (v < words / 4) reinterpret_cast<uint4 *>(stage)[v] = part[k];
    }
    if (tail < words) stage[tail] = last;
  }
  __syncthreads();

  Splat s{};
  uint32_t key = kInvisible;
  const bool visible = i < c.count && project_one<Degree, Format>(i, g, stage + threadIdx.x * c.shWords, c, opt, s, key);
  if (i < c.count) {
    keys[i] = key;
    values[i] = i;
    if (visible) {
      float4 *out = reinterpret_cast<float4 *>(splats + i);
      out[0] = make_float4(s.qx, s.qy, s.qz, s.o);
      out[1] = make_float4(s.r, s.g, s.b, s.radius);
      out[2] = make_float4(s.cx, s.cy, s.ex, s.ey);
    }
  }
  __syncthreads(); // the SH words are read; the stage now holds the walk's owners
  WalkOwner *owners = reinterpret_cast<WalkOwner *>(stage);

  // Candidate tiles: the bounding box of each visible splat. A box of more
  // than kMaskTiles tiles is a big splat's: it is listed for big_count_kernel
  // (one slot per wave from one atomic) and left out of the walk.
  uint32_t t[4] = {0, 0, 0, 0};
  uint32_t box = 0;
  if (visible) {
    splat_tiles(s.cx, s.cy, s.ex, s.ey, gridW, gridH, TileShape{c.tileW, c.tileH}, t);
    box = (t[2] - t[0] + 1) * (t[3] - t[1] + 1);
  }
  const bool big = box > kMaskTiles;
  const uint64_t bigLanes = __ballot(big);
  if (bigLanes != 0) {
    uint32_t base = 0;
    if (lane == static_cast<uint32_t>(__builtin_ctzll(bigLanes))) {
      base = atomicAdd(&counters[kBigSplats], static_cast<uint32_t>(__popcll(bigLanes)));
    }
    base = __shfl(base, static_cast<int>(__builtin_ctzll(bigLanes)), 64);
    if (big) bigList[base + static_cast<uint32_t>(__popcll(bigLanes & ((1ull << lane) - 1ull)))] = i;
  }
  const bool walks = visible && !big;
  const uint32_t candidates = walks ? box : 0u;
  const uint64_t walkers = __ballot(walks);
  const uint32_t width = t[2] - t[0] + 1;
  const uint32_t end = wave_inclusive_sum(candidates, lane);
  const uint32_t start = end - candidates;
  const uint32_t total = __shfl(end, 63, 64);
  if (walks) {
    const uint32_t k = static_cast<uint32_t>(__popcll(walkers & ((1ull << lane) - 1ull)));
    owners[wave + k] = WalkOwner{tile_test(s.qx, s.qy, s.qz, s.o, s.cx, s.cy), t[0], t[1], width, start,
                                 1.0f / static_cast<float>(width), i, 0u, 0u};
  }
  __builtin_amdgcn_wave_barrier();

  uint32_t reached = 0, before = 0;
  uint64_t bits = 0;
  for (uint32_t base = 0; base < total; base += 64) {
    const uint32_t slot = walk_round(flags + wave, lane, base, walks, start, before);
    const uint32_t j = base + lane;
    bool hit = false;
    if (j < total) {
      const WalkOwner &o = owners[wave + slot];
      uint32_t tx, ty;
      box_tile(o, j - o.start, tx, ty);
      hit = reaches_tile(o.test, tx, ty, TileShape{c.tileW, c.tileH});
    }
    const uint64_t round = __ballot(hit);
    // This lane's own candidates within the round: lanes start - base .. start + candidates - base.
    if (candidates != 0 && start < base + 64 && start + candidates > base) {
      const uint32_t lo = start > base ? start - base : 0u;
      const uint32_t hi = start + candidates - base < 64u ? start + candidates - base : 64u;
      const uint64_t mine = round & bit_range(lo, hi);
      reached += static_cast<uint32_t>(__popcll(mine));
      if (candidates <= kMaskTiles) bits |= (mine >> lo) << (base + lo - start);
    }
  }
  if (i < c.count) {
    if (!big) tileCount[i] = reached; // a big splat's count comes from big_count_kernel
    const uint32_t height = t[3] - t[1] + 1;
    records[i] = TileRecord{static_cast<uint32_t>(bits), static_cast<uint32_t>(bits >> 32), t[0] | (t[1] << 16),
                            visible ? (width | (height << 16)) : 0u};
  }
}

// Big splats' reached tiles, counted by the waves: a wave loads 64 listed
// splats at once (one per lane), then walks each one's box in turn, 64
// candidates a round. The listed splats are dealt out over all the waves.
__global__ __launch_bounds__(256) void big_count_kernel(const Splat *__restrict__ splats, TileRecord *__restrict__ records,
                                                        const uint32_t *__restrict__ bigList, uint32_t *counters, TileShape shape,
                                                        uint64_t *__restrict__ bigBits, uint32_t bitCapacity,
                                                        uint32_t *__restrict__ tileCount) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t waves = gridDim.x * (blockDim.x / 64u), wave = (blockIdx.x * blockDim.x + threadIdx.x) / 64u;
  const uint32_t listed = counters[kBigSplats];
  // Wave w takes listed splats w, w + waves, w + 2 waves, ..., 64 at a time (one
  // per lane), so that a short list still spreads over every wave.
  for (uint32_t base = 0; wave + base * waves < listed; base += 64u) {
    const uint32_t mineAt = wave + (base + lane) * waves;
    BigSplat mine{};
    uint32_t rounds = 0;
    if (mineAt < listed) {
      const uint32_t i = bigList[mineAt];
      const Splat s = load_splat(splats, i);
      const TileRecord rec = records[i];
      mine = BigSplat{tile_test(s.qx, s.qy, s.qz, s.o, s.cx, s.cy), rec.origin & 0xFFFFu, rec.origin >> 16, rec.size & 0xFFFFu,
                      (rec.size & 0xFFFFu) * (rec.size >> 16), i, 0u};
      rounds = (mine.candidates + 63u) / 64u;
    }
    // Room for each splat's bits, one word a round, from one atomic per wave.
    const uint32_t end = wave_inclusive_sum(rounds, lane);
    uint32_t wordBase = 0;
    if (lane == 63u && end != 0) wordBase = atomicAdd(&counters[kBigBitWords], end);
    wordBase = __shfl(wordBase, 63, 64);
    const uint32_t bitStart = wordBase + (end - rounds);
    // Past the capacity a splat keeps no bits, and its list pass tests its tiles again.
    mine.first = bitStart + rounds <= bitCapacity ? bitStart : kNoBits;
    const uint32_t inChunk = static_cast<uint32_t>(__popcll(__ballot(mineAt < listed))); // a prefix of the lanes
    uint32_t myHits = 0;
    for (uint32_t j = 0; j < inChunk; ++j) {
      const BigSplat b = big_splat_of_lane(mine, j);
      BoxCursor at(lane, b.width);
      uint32_t hits = 0, round = 0;
      for (uint32_t k = lane; k - lane < b.candidates; k += 64u, ++round) {
        const bool hit = k < b.candidates && reaches_tile(b.test, b.tx0 + at.col, b.ty0 + at.row, shape);
        const uint64_t mask = __ballot(hit);
        if (lane == 0 && b.first != kNoBits) bigBits[b.first + round] = mask;
        hits += static_cast<uint32_t>(__popcll(mask));
        at.advance();
      }
      if (lane == j) myHits = hits;
    }
    if (mineAt < listed) {
      tileCount[mine.gaussian] = myHits;
      records[mine.gaussian].maskHi = mine.first; // where its bits are, for the list pass
    }
  }
}

void launch_project(hipStream_t stream, const float *positions, const float *scales, const float *rotations, const float *opacities,
                    const uint32_t *shWords, const float *filter3d, const ViewParams &view, bool mipFilter2d, float shDropout,
                    uint32_t shRetained, uint32_t shSeed, uint32_t gridW, uint32_t gridH, Splat *splats, uint32_t *keys,
                    uint32_t *values, uint32_t *tileCount, TileRecord *records, uint32_t *counters, uint32_t *bigList,
                    uint64_t *bigBits, uint32_t bitCapacity) {
  if (view.count == 0) return;
  const ProjectOptions opt{mipFilter2d ? 1u : 0u, shDropout, shRetained, shSeed};
  const dim3 grid((view.count + 255) / 256), block(256);
#define VSPLAT_PROJECT(D, F)                                                                                                       \
  hipLaunchKernelGGL((project_kernel<D, F>), grid, block, 0, stream, positions, scales, rotations, opacities, shWords, filter3d,    \
                     view, opt, gridW, gridH, splats, keys, values, tileCount, records, counters, bigList)
#define VSPLAT_PROJECT_DEGREE(F)                                                                                                   \
  switch (view.shDegree) {                                                                                                         \
  case 0: VSPLAT_PROJECT(0, F); break;                                                                                             \
  case 1: VSPLAT_PROJECT(1, F); break;                                                                                             \
  case 2: VSPLAT_PROJECT(2, F); break;                                                                                             \
  default: VSPLAT_PROJECT(3, F); break;                                                                                            \
  }
  if (view.shFormat == kShBf16) {
    VSPLAT_PROJECT_DEGREE(kShBf16)
  } else {
    VSPLAT_PROJECT_DEGREE(kShHalf)
  }
#undef VSPLAT_PROJECT_DEGREE
#undef VSPLAT_PROJECT
  // The big splats' counts; the grid takes them in turn (how many there are is only known on the device).
  hipLaunchKernelGGL(big_count_kernel, dim3(kBigSplatBlocks), dim3(256), 0, stream, splats, records, bigList, counters,
                     TileShape{view.tileW, view.tileH}, bigBits, bitCapacity, tileCount);
}

} // namespace vsplat
