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

// A small splat as the list pass reads it from LDS: its box, its reached-tile
// bits and where its entries start in the list.
struct EmitOwner {
  uint32_t tx0, ty0, width, start; // the box, row-major, and its first candidate in the wave's list
  float invWidth;                  // 1 / width, rounded
  uint32_t gaussian;
  uint32_t maskLo, maskHi;
  uint32_t offset; // its first entry
  uint32_t last;   // the box's last tile column and row, tx1 | ty1 << 16
  uint32_t edges;  // box_block_edges
  uint32_t pad;
};
static_assert(sizeof(EmitOwner) == 48, "48-byte emit owner");

// Candidate `local` of a box `width` tiles wide: its tile. The quotient from
// the rounded reciprocal is at most a few units off and is corrected exactly.
template <class Owner>
__device__ __forceinline__ void box_tile(const Owner &o, uint32_t local, uint32_t &tx, uint32_t &ty) {
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

// A wave's walk over one big splat's box, 64 candidates a round, row-major:
// lane l starts at candidate l and every round moves 64 on (a carry of at most
// one row past the wrap, since 64 mod width < width).
struct BoxCursor {
  uint32_t col, row, stepCols, stepRows, width;
  __device__ __forceinline__ BoxCursor(uint32_t lane, uint32_t w) : width(w) {
    // lane / w and 64 / w from a rounded reciprocal, corrected exactly (w >= 1).
    const float inv = 1.0f / static_cast<float>(w);
    row = static_cast<uint32_t>(static_cast<float>(lane) * inv);
    while (row * w > lane) --row;
    while ((row + 1u) * w <= lane) ++row;
    col = lane - row * w;
    stepRows = static_cast<uint32_t>(64.0f * inv);
    while (stepRows * w > 64u) --stepRows;
    while ((stepRows + 1u) * w <= 64u) ++stepRows;
    stepCols = 64u - stepRows * w;
  }
  __device__ __forceinline__ void advance() {
    col += stepCols;
    row += stepRows;
    if (col >= width) {
      col -= width;
      row += 1u;
    }
  }
};

// One of 64 listed big splats, as every lane of a wave reads it from the lane
// that loaded it (a wave loads 64 at once so that their loads overlap).
struct BigSplat {
  TileTest test;
  uint32_t tx0, ty0, width, candidates, gaussian, first; // first: its first list entry, or (counting) its first bit word
};
// A big splat whose reached-tile bits did not fit the bit buffer (its list pass tests its tiles again).
constexpr uint32_t kNoBits = 0xFFFFFFFFu;
__device__ __forceinline__ uint32_t lane_u32(uint32_t v, uint32_t j) {
  return static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v), static_cast<int>(j)));
}
__device__ __forceinline__ float lane_f32(float v, uint32_t j) { return __uint_as_float(lane_u32(__float_as_uint(v), j)); }
__device__ __forceinline__ BigSplat big_splat_of_lane(const BigSplat &b, uint32_t j) {
  return BigSplat{TileTest{lane_f32(b.test.qx, j), lane_f32(b.test.qy, j), lane_f32(b.test.qz, j), lane_f32(b.test.cx, j),
                           lane_f32(b.test.cy, j), lane_f32(b.test.limit, j), lane_f32(b.test.iqx, j), lane_f32(b.test.iqz, j)},
                  lane_u32(b.tx0, j), lane_u32(b.ty0, j), lane_u32(b.width, j), lane_u32(b.candidates, j), lane_u32(b.gaussian, j),
                  lane_u32(b.first, j)};
}

// Where a splat's pixel box starts and ends in the 8x8 pixel blocks of the
// first and last tiles of its tile box (tx0..tx1, ty0..ty1), packed as
// c0 | c1 << 8 | r0 << 16 | r1 << 24: in the first tile column the box's
// block columns are c0 and up, in the last those below c1 (and the same for
// rows); a column or row the box misses has c0 = blocks across (r0 = blocks
// down), c1 = 0 (r1 = 0). Tiles clip at the image's edge.
__device__ __forceinline__ uint32_t box_block_edges(const PixelBox &p, uint32_t tx0, uint32_t ty0, uint32_t tx1, uint32_t ty1,
                                                    TileShape shape, uint32_t width, uint32_t height) {
  const uint32_t across = shape.w >> 3, down = shape.h >> 3;
  const uint32_t x0 = tx0 * shape.w, y0 = ty0 * shape.h, x1 = tx1 * shape.w, y1 = ty1 * shape.h;
  const uint32_t lx = p.ax > x0 ? p.ax : x0, ly = p.ay > y0 ? p.ay : y0;
  const uint32_t endX0 = x0 + shape.w < width ? x0 + shape.w : width, endY0 = y0 + shape.h < height ? y0 + shape.h : height;
  const uint32_t c0 = lx >= endX0 ? across : (lx - x0) >> 3, r0 = ly >= endY0 ? down : (ly - y0) >> 3;
  const uint32_t endX1 = x1 + shape.w < width ? x1 + shape.w : width, endY1 = y1 + shape.h < height ? y1 + shape.h : height;
  const uint32_t hx = p.bx < endX1 ? p.bx : endX1, hy = p.by < endY1 ? p.by : endY1;
  const uint32_t c1 = hx > x1 ? (((hx - 1u - x1) >> 3) + 1u < across ? ((hx - 1u - x1) >> 3) + 1u : across) : 0u;
  const uint32_t r1 = hy > y1 ? (((hy - 1u - y1) >> 3) + 1u < down ? ((hy - 1u - y1) >> 3) + 1u : down) : 0u;
  return c0 | (c1 << 8) | (r0 << 16) | (r1 << 24);
}

// The 8x8 pixel blocks of tile (tx, ty) of a splat's tile box that its pixel
// box reaches, bit q for the tile's block q (row-major, shape.w / 8 across,
// as the raster numbers them): exactly the blocks for which splat_box, given
// the block, is not empty, so the raster pass stages the entry there only.
__device__ __forceinline__ uint32_t block_mask(uint32_t edges, uint32_t tx, uint32_t ty, uint32_t tx0, uint32_t ty0, uint32_t tx1,
                                               uint32_t ty1, TileShape shape) {
  const uint32_t across = shape.w >> 3, down = shape.h >> 3;
  const uint32_t c0 = tx == tx0 ? edges & 0xFFu : 0u, c1 = tx == tx1 ? (edges >> 8) & 0xFFu : across;
  const uint32_t r0 = ty == ty0 ? (edges >> 16) & 0xFFu : 0u, r1 = ty == ty1 ? edges >> 24 : down;
  const uint32_t row = (1u << across) - 1u;                  // one row of blocks
  const uint32_t column = 0xFFFFFFFFu / row & ((1u << (across * down)) - 1u); // bit 0 of every row
  const uint32_t cols = (row << c0) & (row >> (across - c1)) & row;
  const uint32_t rows = (column << (r0 * across)) & (column >> ((down - r1) * across));
  return cols * rows;
}

// Splats whose bounding box has at most this many tiles keep the outcome of
// every candidate as one bit (row-major over the box) and are counted and
// listed by the waves' shared walks; larger ones ("big" splats, a few per
// scene with ordinary sizes, very many with heavy-tailed scales) are counted
// and listed by their own kernels, a wave per splat, so that one large box
// does not hold up a wave's other 63 splats.
constexpr uint32_t kMaskTiles = 64;

// Per Gaussian, in Gaussian order: the reached-tile bits and the box they refer to.
struct TileRecord {
  uint32_t maskLo, maskHi;
  uint32_t origin; // tx0 | ty0 << 16
  uint32_t size;   // width | height << 16, 0 when the splat is not visible
};

} // namespace vsplat
