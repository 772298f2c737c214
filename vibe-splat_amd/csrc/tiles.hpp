// Which tiles a splat reaches (train_types.glsl:34-77, for the reference's
// 16x16; here any power-of-two width and height, 8 to 32), and the
// wave-cooperative walk over candidate tiles that projection and list
// emission share.
//
// A splat's candidates are the tiles of its bounding box (its centre plus and
// minus its extent and one pixel). A wave walks the candidates of its 64
// splats together, 64 candidates per round: lane j of a round takes candidate
// base + j of the wave's concatenated list, finds which splat owns it by a
// binary search over the running totals, and tests it. Every lane does useful
// work in every round but the last, whatever the spread of bounding-box sizes
// in the wave.
#pragma once

#include "device.hpp"

namespace vsplat {

// A tile shape: width and height in pixels, powers of two (the reference's is
// 16 x 16). Dividing by them is multiplying by an exact power of two.
struct TileShape {
  uint32_t w, h;
};

// Inclusive tile box of a splat on a gridW x gridH grid (train_types.glsl:34-39).
__device__ __forceinline__ void splat_tiles(float cx, float cy, float ex, float ey, uint32_t gridW, uint32_t gridH, TileShape shape,
                                            uint32_t t[4]) {
  const float lastX = static_cast<float>(gridW) - 1.0f, lastY = static_cast<float>(gridH) - 1.0f;
  const float ix = 1.0f / static_cast<float>(shape.w), iy = 1.0f / static_cast<float>(shape.h);
  t[0] = static_cast<uint32_t>(clampf(floorf((cx - ex - 1.0f) * ix), 0.0f, lastX));
  t[1] = static_cast<uint32_t>(clampf(floorf((cy - ey - 1.0f) * iy), 0.0f, lastY));
  t[2] = static_cast<uint32_t>(clampf(floorf((cx + ex + 1.0f) * ix), 0.0f, lastX));
  t[3] = static_cast<uint32_t>(clampf(floorf((cy + ey + 1.0f) * iy), 0.0f, lastY));
}

// What the reaches-tile test needs of a splat: the conic, the centre, the
// threshold 2 ln(255 o) * 1.001 + 1e-3, and 1/qx, 1/qz.
struct TileTest {
  float qx, qy, qz, cx, cy, limit, iqx, iqz;
};

__device__ __forceinline__ TileTest tile_test(float qx, float qy, float qz, float o, float cx, float cy) {
  return TileTest{qx, qy, qz, cx, cy, fma32(reach_of(o), 1.001f, 1e-3f), 1.0f / qx, 1.0f / qz};
}

// The least of q = qx x^2 + 2 qy x y + qz y^2 along one edge of the tile's
// square of pixel centres, at the clamped stationary point.
__device__ __forceinline__ float edge_min_vertical(const TileTest &s, float x, float ylo, float yhi) {
  const float y = clampf((-s.qy * x) * s.iqz, ylo, yhi);
  return fma32(s.qz * y, y, fma32(2.0f * s.qy * x, y, s.qx * x * x));
}
__device__ __forceinline__ float edge_min_horizontal(const TileTest &s, float v, float xlo, float xhi) {
  const float u = clampf((-s.qy * v) * s.iqx, xlo, xhi);
  return fma32(s.qz * v, v, fma32(2.0f * s.qy * u, v, s.qx * u * u));
}

// Whether the splat's 1/255 ellipse reaches a pixel centre of tile (tx, ty)
// (train_types.glsl:56-77), with the reference's slack so that rounding never
// drops a pixel the pixel loops would composite.
__device__ __forceinline__ bool reaches_tile(const TileTest &s, uint32_t tx, uint32_t ty, TileShape shape) {
  const float w = static_cast<float>(shape.w), h = static_cast<float>(shape.h);
  const float lowx = (static_cast<float>(tx) * w + 0.5f) - s.cx, lowy = (static_cast<float>(ty) * h + 0.5f) - s.cy;
  const float highx = lowx + (w - 1.0f), highy = lowy + (h - 1.0f);
  if (lowx <= 0.0f && lowy <= 0.0f && highx >= 0.0f && highy >= 0.0f) return true;
  float least = 1e30f;
  least = minf(least, edge_min_vertical(s, lowx, lowy, highy));
  least = minf(least, edge_min_horizontal(s, lowy, lowx, highx));
  least = minf(least, edge_min_vertical(s, highx, lowy, highy));
  least = minf(least, edge_min_horizontal(s, highy, lowx, highx));
  return least <= s.limit;
}

// Inclusive prefix sum over the 64 lanes of a wave.
__device__ __forceinline__ uint32_t wave_inclusive_sum(uint32_t v, uint32_t lane) {
#pragma unroll
  for (uint32_t d = 1; d < 64; d <<= 1) {
    const uint32_t t = __shfl_up(v, d, 64);
    v += lane >= d ? t : 0u;
  }
  return v;
}

// A splat as the lanes testing its candidates read it from LDS. The splats
// with candidates are packed in lane order (slot k is the k-th such lane), so
// their candidate ranges follow each other without gaps.
struct WalkOwner {
  TileTest test;
  uint32_t tx0, ty0, width, start; // the box, row-major, and its first candidate in the wave's list
  float invWidth;                  // 1 / width, rounded
  uint32_t gaussian;
  uint32_t maskLo, maskHi;
};
static_assert(sizeof(WalkOwner) == 64, "64-byte walk owner");

// Candidate `local` of a box `width` tiles wide: its tile. The quotient from
// the rounded reciprocal is at most a few units off and is corrected exactly.
__device__ __forceinline__ void box_tile(const WalkOwner &o, uint32_t local, uint32_t &tx, uint32_t &ty) {
  int q = static_cast<int>(static_cast<float>(local) * o.invWidth);
  int r = static_cast<int>(local) - static_cast<int>(__umul24(static_cast<uint32_t>(q), o.width));
  while (r < 0) {
    --q;
    r += static_cast<int>(o.width);
  }
  while (r >= static_cast<int>(o.width)) {
    ++q;
    r -= static_cast<int>(o.width);
  }
  tx = o.tx0 + static_cast<uint32_t>(r);
  ty = o.ty0 + static_cast<uint32_t>(q);
}

// One round of a wave's walk: candidates base .. base + 63 of the wave's list.
// `flags` is the wave's 64-word scratch, `before` the number of owners whose
// candidates start before base (advanced here). Returns the slot of the owner
// of candidate base + lane (meaningful when that candidate exists).
//
// An owner starting in the round marks its position; a ballot of the marks
// gives, by a popcount below each lane, how many owners start at or before it.
__device__ __forceinline__ uint32_t walk_round(uint32_t *flags, uint32_t lane, uint32_t base, bool owner, uint32_t start,
                                               uint32_t &before) {
  __builtin_amdgcn_wave_barrier();
  flags[lane] = 0u;
  if (owner && start >= base && start - base < 64u) flags[start - base] = 1u;
  __builtin_amdgcn_wave_barrier();
  const uint64_t starts = __ballot(flags[lane] != 0u);
  const uint64_t upTo = lane == 63u ? ~0ull : ((2ull << lane) - 1ull);
  const uint32_t slot = before + static_cast<uint32_t>(__popcll(starts & upTo)) - 1u;
  before += static_cast<uint32_t>(__popcll(starts));
  return slot;
}

// Bits lo..hi-1 of a 64-bit mask (0 <= lo <= hi <= 64).
__device__ __forceinline__ uint64_t bit_range(uint32_t lo, uint32_t hi) {
  const uint64_t below_hi = hi >= 64 ? ~0ull : ((1ull << hi) - 1ull);
  const uint64_t below_lo = lo >= 64 ? ~0ull : ((1ull << lo) - 1ull);
  return below_hi & ~below_lo;
}

// Splats whose bounding box has at most this many tiles keep the outcome of
// every candidate as one bit (row-major over the box); larger ones are tested
// again when the lists are written.
constexpr uint32_t kMaskTiles = 64;

// Per Gaussian, in Gaussian order: the reached-tile bits and the box they refer to.
struct TileRecord {
  uint32_t maskLo, maskHi;
  uint32_t origin; // tx0 | ty0 << 16
  uint32_t size;   // width | height << 16, 0 when the splat is not visible
};

} // namespace vsplat
