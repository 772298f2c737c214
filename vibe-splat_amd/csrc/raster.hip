// Per-tile compositing and its derivative (train_forward.comp:38-100,
// train_backward.glsl:78-158). The lists are per tile (16x16 by default, or
// 8x8, 32x8 or 32x32); the pixels are worked in 8x8 blocks, one workgroup of one
// wave64 per block, lane l holding pixel (l & 7, l >> 3). A block reads its
// tile's list in batches of 64 entries: each lane stages one splat with the
// block's pixels it can touch already worked out, and only the splats that
// can touch a pixel of the block are kept, packed in list order. A block
// stops when its own pixels are done, with no barrier across the tile.
//
// Staging: for each splat, an 8-bit column mask and an 8-bit row mask of the
// pixels inside both its extent box (splat_pixels, train_types.glsl:43-48) and
// its radius square (the first test of splat_at): a pixel is a candidate iff
// its column and row bits are set. The masks are computed with the same float
// operations as the per-pixel tests they replace, so the set is the same.
// A wave skips a splat none of its lanes is a candidate for; otherwise every
// lane evaluates the Gaussian without branches and the result is selected.
// Every list shape has the same 8x8 blocks, so images and sums do not depend
// on the shape.
//
// Backward: for each splat the 12 non-zero of the 14 per-pixel sums are
// reduced over the wave by a transposed butterfly in which the steps halve the
// values a lane carries where they can: partner lanes l ^ 32 and l ^ 16
// through the permlane swaps, l ^ 15 and l ^ 7 through DPP row and half-row
// mirrors, then l ^ 2 and l ^ 1 through quad permutations. Twelve lanes end
// with one total each, which goes to global memory with one atomic: 64-bit
// fixed point (each wave total quantised, so the integer sums are the same in
// any order) or float.
#include "device.hpp"
#include "kernels.hpp"
#include "blockscan.hpp"
#include "tiles.hpp"

namespace vsplat {

namespace {

struct Staged {
  float qx, qy, qz, o;
  float r, g, b;
  uint32_t entry; // the splat's index in the tile's list
  float cx, cy;
  uint32_t cols, rows; // the block's columns (bits 0-7) and rows (bits 16-23) the splat can touch; rows unused
};
static_assert(sizeof(Staged) == 48, "48-byte staged splat");

constexpr uint32_t kBlock = 64; // pixels of an 8x8 block, threads of its workgroup, entries per batch

struct BlockGeometry {
  uint32_t x0, y0, x1, y1;
  uint32_t block; // the block's number in its tile (its bit in the entries' block masks)
  bool reach;     // lists of tiles larger than the block: stage only the splats whose ellipse reaches it
};

// Workgroups are dealt to the chip's eight XCDs in turn (workgroup h runs on
// XCD h % 8), and each XCD has its own L2. Launch index h works 8x8 block q of
// list tile t such that a tile's blocks all run on one XCD, one after another,
// and read its list through one L2, while the tiles are still dealt over the
// XCDs in turn (XCD x works tiles x, x + 8, x + 16, ...): with u = h / 8,
// t = (u / blocks) * 8 + h % 8 and q = u % blocks, the tile's blocks row-major
// (TW / 8 across). A grid of raster_blocks(tiles) covers every (t, q) once;
// the blocks of the tiles past the last (t >= tiles) start below the image.
template <uint32_t TW, uint32_t TH>
__device__ __forceinline__ BlockGeometry block_geometry(uint32_t h, uint32_t gridW, uint32_t width, uint32_t height,
                                                        uint32_t &tile) {
  constexpr uint32_t kAcross = TW / 8u, kPerTile = (TW / 8u) * (TH / 8u);
  const uint32_t u = h / kRasterXcds;
  tile = (u / kPerTile) * kRasterXcds + h % kRasterXcds;
  const uint32_t q = u % kPerTile;
  const uint32_t x0 = (tile % gridW) * TW + (q % kAcross) * 8u, y0 = (tile / gridW) * TH + (q / kAcross) * 8u;
  return {x0, y0, x0 + 8u < width ? x0 + 8u : width, y0 + 8u < height ? y0 + 8u : height, q, TW * TH > 64u};
}

// Heaviest first: a block's work (entries) as one of kOrderKeys keys, 8 per
// octave (the exponent and the next 3 bits of work + 1), the largest work the
// smallest key.
constexpr uint32_t kOrderKeys = 256;
__device__ __forceinline__ uint32_t heavy_first(uint32_t work) {
  const uint32_t v = work < 0xFFFFFFFFu ? work + 1u : work;
  const uint32_t e = 31u - static_cast<uint32_t>(__builtin_clz(v));
  const uint32_t fraction = (e >= 3u ? v >> (e - 3u) : v << (3u - e)) & 7u;
  return kOrderKeys - 1u - (e * 8u + fraction);
}

// The block's pixels a splat can touch: inside the extent box and within the
// radius (|d| <= r on both axes, d from the pixel centre). A larger tile's list
// holds every splat whose extent box meets the block (the block marks), and
// about a quarter of those (elongated, turned splats) have boxes that meet the
// block where their 1/255 ellipse does not: those are dropped here by the test
// that makes the lists of 8x8 tiles (reaches_tile, with its slack against
// rounding), on the same values, so that every tile shape stages the same splats
// and none that could take a pixel of the block.
__device__ __forceinline__ Staged stage(const Splat &s, const BlockGeometry &geo, uint32_t entry) {
  uint32_t box = splat_box(s, geo.x0, geo.y0, geo.x1, geo.y1);
  if (box != 0u && geo.reach &&
      !reaches_tile(tile_test(s.qx, s.qy, s.qz, s.o, s.cx, s.cy), geo.x0 >> 3, geo.y0 >> 3, TileShape{8u, 8u})) {
    box = 0u;
  }
  uint32_t mask = 0;
  if (box != 0u) {
    const uint32_t bx0 = box & 0xFFu, by0 = (box >> 8) & 0xFFu, bx1 = (box >> 16) & 0xFFu, by1 = box >> 24;
    // Kept as a loop: unrolled, its temporaries would set the kernel's register count.
#pragma unroll 1
    for (uint32_t k = 0; k < 8u; ++k) {
      const float dx = (static_cast<float>(geo.x0 + k) + 0.5f) - s.cx;
      const float dy = (static_cast<float>(geo.y0 + k) + 0.5f) - s.cy;
      const bool col = k >= bx0 && k < bx1 && !(fabsf(dx) > s.radius);
      const bool row = k >= by0 && k < by1 && !(fabsf(dy) > s.radius);
      mask |= ((col ? 1u : 0u) << k) | ((row ? 1u : 0u) << (16u + k));
    }
  }
  return Staged{s.qx, s.qy, s.qz, s.o, s.r, s.g, s.b, entry, s.cx, s.cy, mask, 0u};
}

// Lane l's pixel in its block and its two mask bits.
__device__ __forceinline__ uint32_t lane_bits(uint32_t lx, uint32_t ly) { return (1u << lx) | (1u << (16u + ly)); }
__device__ __forceinline__ bool covers(const Staged &s, uint32_t bits) { return (s.cols & bits) == bits; }

// The wave votes as scalar masks: a ballot of one comparison is that
// comparison's lane mask, so the loops vote with the AND of the ballots of
// single comparisons (scalar work) rather than the ballot of a combined
// predicate (which costs vector instructions to form).
__device__ __forceinline__ uint64_t vote(bool p) { return __builtin_amdgcn_ballot_w64(p); }

// The position of set bit number k (from 0) of m, which has more than k.
__device__ __forceinline__ uint32_t kth_set_bit(uint64_t m, uint32_t k) {
  uint32_t at = 0;
#pragma unroll
  for (uint32_t w = 32; w >= 1; w >>= 1) {
    const uint32_t low = static_cast<uint32_t>(__popcll(m & ((1ull << w) - 1ull)));
    if (k >= low) {
      k -= low;
      m >>= w;
      at += w;
    }
  }
  return at;
}

// The next entries of the block's walk over its tile's list, from `pos` on
// towards `limit` (up: pos, pos + 1, ...; down: pos - 1, pos - 2, ...), that
// have the block's bit (blockBit) in their key: up to kBlock of them, in walk
// order, into `queue`; `pos` is left at the first entry not yet looked at (up)
// or one past it (down). The keys are read a window of kBlock at a time; a
// window is taken up to the first wanted entry that does not fit. Returns how
// many were queued. The caller has made sure the previous queue is read.
template <bool Down>
__device__ __forceinline__ uint32_t gather_entries(const uint32_t *__restrict__ entryKeys, uint32_t blockBit, uint32_t &pos,
                                                   uint32_t limit, uint32_t *queue) {
  const uint32_t lane = threadIdx.x;
  uint32_t queued = 0;
  while (queued < kBlock && pos != limit) {
    const uint32_t left = Down ? pos - limit : limit - pos;
    const uint32_t n = left < kBlock ? left : kBlock;
    const uint32_t e = Down ? pos - 1u - lane : pos + lane;
    const bool wanted = lane < n && (entryKeys[e] & blockBit) != 0u;
    const uint64_t m = __ballot(wanted);
    const uint32_t found = static_cast<uint32_t>(__popcll(m)), room = kBlock - queued;
    const uint32_t taken = found > room ? kth_set_bit(m, room) : n; // window entries used up
    if (wanted && lane < taken) queue[queued + static_cast<uint32_t>(__popcll(m & ((1ull << lane) - 1ull)))] = e;
    queued += found < room ? found : room;
    pos = Down ? pos - taken : pos + taken;
  }
  return queued;
}

// The queued entries: each lane stages one, and the splats that can touch a
// pixel of the block are packed at the front of `batch` in queue order (with
// their Gaussians in `index`, when given). Returns how many were kept.
__device__ __forceinline__ uint32_t stage_batch(const Splat *__restrict__ splats, const uint32_t *__restrict__ entryValues,
                                                const uint32_t *queue, uint32_t count, const BlockGeometry &geo, Staged *batch,
                                                uint32_t *index) {
  const uint32_t lane = threadIdx.x;
  Staged st{};
  uint32_t gaussian = 0;
  bool keep = false;
  if (lane < count) {
    const uint32_t entry = queue[lane];
    gaussian = entryValues[entry];
    st = stage(load_splat(splats, gaussian), geo, entry);
    keep = (st.cols & 0xFFu) != 0u && (st.cols >> 16) != 0u;
  }
  const uint64_t kept = __ballot(keep);
  if (keep) {
    const uint32_t slot = static_cast<uint32_t>(__popcll(kept & ((1ull << lane) - 1ull)));
    batch[slot] = st;
    if (index != nullptr) index[slot] = gaussian;
  }
  __syncthreads();
  return static_cast<uint32_t>(__popcll(kept));
}

// The Gaussian's exponent at offset (dx, dy) (splat_at, train_types.glsl:88).
__device__ __forceinline__ float power_at(const Staged &s, float dx, float dy) {
  return fma32(-0.5f, fma32(s.qx * dx, dx, (s.qz * dy) * dy), -((s.qy * dx) * dy));
}

// Front to back over a tile's list (train_forward.comp:38-100): the colour
// (with Colour), the transmittance and its log, and the stop index of the
// calling lane's pixel. The forward pass and the backward pass that
// recomputes the transmittance both run this, so they agree bit for bit.
// A pixel stops at the splat that would take its transmittance, the product
// of 1 - alpha, under 1e-4; train_forward.comp:84-88 asks the same of the sum
// of the logs (log T < log 1e-4), which costs a log per pair, and rounds
// further from the exact product than the product does. The log of the
// final transmittance is taken once per pixel.
template <bool Colour>
__device__ __forceinline__ void composite_block(const Splat *__restrict__ splats, const uint32_t *__restrict__ entryKeys,
                                                const uint32_t *__restrict__ entryValues, uint32_t blockBit, uint2 range,
                                                const BlockGeometry &geo, Staged *batch, uint32_t *queue, bool inside, float px, float py,
                                                uint32_t bits, float &r, float &g, float &b, float &T, float &logT, uint32_t &stop) {
  stop = range.x;
  bool done = !inside;
  uint64_t doneLanes = ~vote(inside); // lanes whose pixel is finished (or outside the image)
  for (uint32_t pos = range.y; pos > range.x;) {
    if (doneLanes == ~0ull) break;
    __syncthreads(); // the previous batch and queue are read
    const uint32_t queued = gather_entries<true>(entryKeys, blockBit, pos, range.x, queue);
    __syncthreads();
    const uint32_t kept = stage_batch(splats, entryValues, queue, queued, geo, batch, nullptr);
    // The batch runs front to back already (the queue was filled walking down).
    for (uint32_t k = 0; k < kept; ++k) {
      const Staged s = batch[k];
      const bool covered = covers(s, bits);
      const uint64_t candidates = vote(covered) & ~doneLanes;
      if (candidates == 0) continue;
      const bool candidate = !done && covered;
      const float dx = px - s.cx, dy = py - s.cy;
      const float power = power_at(s, dx, dy);
      // Below -87 alpha is far under 1/255 either way, so the clamp in exp_finite decides nothing.
      const float e = exp_finite(power);
      const float a = s.o * e;
      const float alpha = a < 0.99f ? a : 0.99f;
      const bool front = !(power > 0.0f), visible = alpha >= kMinAlpha;
      const uint64_t oks = candidates & vote(front) & vote(visible);
      if (oks == 0) continue;
      const bool ok = candidate && front && visible;
      // Without branches: a lane that takes nothing adds a zero weight
      // (fma(0, c, r) is r) and keeps its transmittance by selection.
      const float Tnext = T * (1.0f - alpha);
      const bool below = Tnext < kMinTransmittance;
      const bool stopHere = ok && below;
      const bool take = ok && !stopHere;
      if constexpr (Colour) {
        const float w = take ? T * alpha : 0.0f;
        r = fma32(w, s.r, r);
        g = fma32(w, s.g, g);
        b = fma32(w, s.b, b);
      }
      T = take ? Tnext : T;
      stop = stopHere ? s.entry + 1u : stop;
      done = done || stopHere;
      doneLanes |= oks & vote(below);
      if (doneLanes == ~0ull) break;
    }
  }
  logT = log_normal(T); // T in [1e-4, 1]: a normal float
}

// The forward pass of one 8x8 block (item) of a view.
template <uint32_t TW, uint32_t TH>
__device__ __forceinline__ void forward_block(uint32_t item, const Splat *__restrict__ splats, const uint32_t *__restrict__ entryKeys,
                                              const uint32_t *__restrict__ entryValues, uint32_t tileBits,
                                              const uint2 *__restrict__ ranges, uint32_t width, uint32_t height, uint32_t gridW,
                                              float bgR, float bgG, float bgB, float4 *__restrict__ image, uint32_t *__restrict__ stops,
                                              uint32_t *__restrict__ work, Staged *batch, uint32_t *queue) {
  uint32_t tile;
  const BlockGeometry geo = block_geometry<TW, TH>(item, gridW, width, height, tile);
  if (geo.x0 >= width || geo.y0 >= height) { // a block past the image's edge (uniform)
    if (work != nullptr && threadIdx.x == 0) work[item] = 0;
    return;
  }
  const uint32_t lx = threadIdx.x & 7u, ly = threadIdx.x >> 3;
  const uint32_t x = geo.x0 + lx, y = geo.y0 + ly;
  const bool inside = x < width && y < height;
  const float px = static_cast<float>(x) + 0.5f, py = static_cast<float>(y) + 0.5f;
  float r = 0.0f, g = 0.0f, b = 0.0f, T = 1.0f, logT = 0.0f;
  uint32_t stop;
  const uint2 range = ranges[tile];
  composite_block<true>(splats, entryKeys, entryValues, 1u << (tileBits + geo.block), range, geo, batch, queue, inside, px, py,
                        lane_bits(lx, ly), r, g, b, T, logT, stop);
  if (inside) {
    const uint32_t at = y * width + x;
    image[at] = make_float4(fma32(T, bgR, r), fma32(T, bgG, g), fma32(T, bgB, b), logT);
    if (stops != nullptr) stops[at] = stop; // none when the backward recomputes them
  }
  if (work != nullptr) {
    // The backward's walk over this block: the entries from its lowest stop to the list's end.
    uint32_t low = inside ? stop : range.y;
#pragma unroll
    for (uint32_t d = 32; d >= 1; d >>= 1) {
      const uint32_t other = static_cast<uint32_t>(__shfl_xor(static_cast<int>(low), static_cast<int>(d), 64));
      low = other < low ? other : low;
    }
    if (threadIdx.x == 0) work[item] = range.y - low;
  }
}

template <uint32_t TW, uint32_t TH>
__global__ __launch_bounds__(kBlock) void forward_kernel(const Splat *__restrict__ splats, const uint32_t *__restrict__ entryKeys,
                                                         const uint32_t *__restrict__ entryValues, uint32_t tileBits,
                                                         const uint2 *__restrict__ ranges, uint32_t width, uint32_t height,
                                                         uint32_t gridW, float bgR, float bgG, float bgB, float4 *__restrict__ image,
                                                         uint32_t *__restrict__ stops, uint32_t *__restrict__ work,
                                                         const uint32_t *__restrict__ order) {
  __shared__ Staged batch[kBlock];
  __shared__ uint32_t queue[kBlock];
  const uint32_t item = order != nullptr ? order[blockIdx.x] : blockIdx.x;
  forward_block<TW, TH>(item, splats, entryKeys, entryValues, tileBits, ranges, width, height, gridW, bgR, bgG, bgB, image, stops, work,
                        batch, queue);
}

// The view of a batched launch that block `item` belongs to, and the block's
// index within it (the views' blocks follow each other in view order).
__device__ __forceinline__ uint32_t view_of(const RasterViews &views, uint32_t &item) {
  uint32_t v = 0;
  while (v + 1 < views.count && item >= views.view[v + 1].firstBlock) ++v;
  item -= views.view[v].firstBlock;
  return v;
}

// The forward pass of a batch's views in one launch: one workgroup per
// (view, 8x8 block) pair, the views' tile lists all of shape TW x TH.
template <uint32_t TW, uint32_t TH>
__global__ __launch_bounds__(kBlock) void forward_views_kernel(RasterViews views, float bgR, float bgG, float bgB) {
  __shared__ Staged batch[kBlock];
  __shared__ uint32_t queue[kBlock];
  uint32_t item = blockIdx.x;
  const RasterView &w = views.view[view_of(views, item)];
  forward_block<TW, TH>(item, w.splats, w.entryKeys, w.entryValues, w.tileBits, w.ranges, w.width, w.height, w.gridW, bgR, bgG, bgB,
                        w.image, w.stops, nullptr, batch, queue);
}

// The fixed-point quantisation of a wave total (train_backward_fixed.comp:19-21),
// NaN as 0, widened to 64 bits: the scale carries 2^20 steps per unit where the
// reference's carries 64, and the clamp is the reference's 1e9 times the same
// 2^14, so a total spans the same range in units at 16384 times the resolution.
__device__ __forceinline__ int64_t quantise(float v, float scale) {
  float s = v * scale;
  if (s != s) s = 0.0f;
  s = s > kFixedClamp ? kFixedClamp : (s < -kFixedClamp ? -kFixedClamp : s);
  return static_cast<int64_t>(__builtin_rintf(s));
}

__device__ __forceinline__ float dpp_row_mirror(float v) { return as_float(__builtin_amdgcn_update_dpp(0, static_cast<int>(as_uint(v)), 0x140, 0xF, 0xF, false)); }
__device__ __forceinline__ float dpp_half_mirror(float v) { return as_float(__builtin_amdgcn_update_dpp(0, static_cast<int>(as_uint(v)), 0x141, 0xF, 0xF, false)); }
__device__ __forceinline__ float dpp_xor2(float v) { return as_float(__builtin_amdgcn_update_dpp(0, static_cast<int>(as_uint(v)), 0x4E, 0xF, 0xF, false)); }
__device__ __forceinline__ float dpp_xor1(float v) { return as_float(__builtin_amdgcn_update_dpp(0, static_cast<int>(as_uint(v)), 0xB1, 0xF, 0xF, false)); }

// Twelve values per lane to the wave's twelve totals, every value through the
// same tree (partners l ^ 32, 16, 15, 7, 2, 1 in turn; a + b and b + a are the
// same float). Afterwards lane l holds in v[0] the total of value
// reduced_value(l); wave_emits(l) picks one lane per value.
__device__ __forceinline__ void reduce_wave(float (&v)[12], uint32_t lane) {
  // Partner l ^ 32: lanes 0-31 keep values 0-5, lanes 32-63 values 6-11.
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    const auto p = __builtin_amdgcn_permlane32_swap(as_uint(v[i]), as_uint(v[i + 6]), false, false);
    v[i] = as_float(p[0]) + as_float(p[1]);
  }
  // Partner l ^ 16: rows 0 and 2 keep their first three, rows 1 and 3 the others.
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const auto p = __builtin_amdgcn_permlane16_swap(as_uint(v[i]), as_uint(v[i + 3]), false, false);
    v[i] = as_float(p[0]) + as_float(p[1]);
  }
  // Partner l ^ 15 (row mirror), all three kept.
#pragma unroll
  for (int i = 0; i < 3; ++i) v[i] = v[i] + dpp_row_mirror(v[i]);
  // Partner l ^ 7 (half-row mirror): lanes with bit 2 clear keep the first of
  // the first two, the others the second; all keep the third.
  {
    const bool upper = (lane & 4u) != 0u;
    const float send = upper ? v[0] : v[1];
    const float keep = upper ? v[1] : v[0];
    v[0] = keep + dpp_half_mirror(send);
    v[2] = v[2] + dpp_half_mirror(v[2]);
  }
  // Partner l ^ 2: lanes with bit 1 clear keep the one they kept, the others the third.
  {
    const bool upper = (lane & 2u) != 0u;
    const float send = upper ? v[0] : v[2];
    const float keep = upper ? v[2] : v[0];
    v[0] = keep + dpp_xor2(send);
  }
  v[0] = v[0] + dpp_xor1(v[0]);
}
// Sixteen values per lane to one total per lane group, for the scored pass's
// fourteen sums: afterwards lane l holds the wave's total of value (l >> 2) & 15
// in v[0]. The same partner tree as reduce_wave.
__device__ __forceinline__ void reduce_wave16(float (&v)[16], uint32_t lane) {
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const auto p = __builtin_amdgcn_permlane32_swap(as_uint(v[i]), as_uint(v[i + 8]), false, false);
    v[i] = as_float(p[0]) + as_float(p[1]);
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const auto p = __builtin_amdgcn_permlane16_swap(as_uint(v[i]), as_uint(v[i + 4]), false, false);
    v[i] = as_float(p[0]) + as_float(p[1]);
  }
  {
    const bool upper = (lane & 8u) != 0u;
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const float send = upper ? v[i] : v[i + 2];
      const float keep = upper ? v[i + 2] : v[i];
      v[i] = keep + dpp_row_mirror(send);
    }
  }
  {
    const bool upper = (lane & 4u) != 0u;
    const float send = upper ? v[0] : v[1];
    const float keep = upper ? v[1] : v[0];
    v[0] = keep + dpp_half_mirror(send);
  }
  v[0] = v[0] + dpp_xor2(v[0]);
  v[0] = v[0] + dpp_xor1(v[0]);
}

// The value whose total lane l holds after reduce_wave: 6 b5 + 3 b4 + (b1 ? 2 : b2),
// with bk bit k of l.
__device__ __forceinline__ uint32_t reduced_value(uint32_t lane) {
  return 6u * ((lane >> 5) & 1u) + 3u * ((lane >> 4) & 1u) + (((lane >> 1) & 1u) != 0u ? 2u : (lane >> 2) & 1u);
}
// One lane per value: bits 0 and 3 clear, and not both bits 1 and 2 (those
// lanes hold the third value twice).
__device__ __forceinline__ bool wave_emits(uint32_t lane) { return (lane & 9u) == 0u && (lane & 6u) != 6u; }

// Back to front over the tile's list (train_backward.glsl:78-158), for the
// block's pixels, from the block's lowest stop. With Recompute the pixel's log
// transmittance and stop come from a front-to-back sweep of its own instead
// of the forward pass's image and stops, which are then not read. With Score
// (a scored view) the last two sums are filled too: the pixel's high-error
// mark (the adjoint's fourth value) and its edge weight times the blending
// weight, each over 3 pixels of the image.
template <bool Fixed, bool Recompute, bool Score, uint32_t TW, uint32_t TH>
__device__ __forceinline__ void backward_block(uint32_t item, const Splat *__restrict__ splats, const uint32_t *__restrict__ entryKeys,
                                               const uint32_t *__restrict__ entryValues, uint32_t tileBits,
                                               const uint2 *__restrict__ ranges, uint32_t width, uint32_t height, uint32_t gridW,
                                               float bgR, float bgG, float bgB, const float4 *__restrict__ image,
                                               const float4 *__restrict__ adjoint, const uint32_t *__restrict__ stops,
                                               float fixedScale, unsigned long long *__restrict__ rawFixed,
                                               float *__restrict__ rawFloat, const float *__restrict__ edge, Staged *batch,
                                               uint32_t *batchIndex, uint32_t *queue) {
  uint32_t tile;
  const BlockGeometry geo = block_geometry<TW, TH>(item, gridW, width, height, tile);
  if (geo.x0 >= width || geo.y0 >= height) return; // a block past the image's edge (uniform)
  const uint32_t lane = threadIdx.x, lx = lane & 7u, ly = lane >> 3;
  const uint32_t bits = lane_bits(lx, ly);
  const uint32_t x = geo.x0 + lx, y = geo.y0 + ly;
  const bool inside = x < width && y < height;
  const float px = static_cast<float>(x) + 0.5f, py = static_cast<float>(y) + 0.5f;
  const uint32_t at = inside ? y * width + x : 0u;
  const float4 adj = inside ? adjoint[at] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  const float edgeWeight = Score && inside ? edge[at] : 0.0f;
  const float perPixel = 1.0f / (3.0f * static_cast<float>(width) * static_cast<float>(height));
  const uint2 range = ranges[tile];
  float logTotal;
  uint32_t stop;
  if constexpr (Recompute) {
    float r = 0.0f, g = 0.0f, b = 0.0f, T = 1.0f;
    logTotal = 0.0f;
    composite_block<false>(splats, entryKeys, entryValues, 1u << (tileBits + geo.block), range, geo, batch, queue, inside, px, py, bits, r,
                           g, b, T, logTotal, stop);
    if (!inside) stop = range.y;
  } else {
    logTotal = inside ? image[at].w : 0.0f;
    stop = inside ? stops[at] : range.y;
  }
  // The block's lowest stop: entries before it have no candidate lane here.
  uint32_t first = stop;
#pragma unroll
  for (uint32_t d = 32; d >= 1; d >>= 1) {
    const uint32_t other = static_cast<uint32_t>(__shfl_xor(static_cast<int>(first), static_cast<int>(d), 64));
    first = other < first ? other : first;
  }
  // The colour behind the splats visited so far, and the transmittance in
  // front of them: before any, the background and the pixel's final
  // transmittance. Walking back, a splat's own transmittance is the one in
  // front of those behind it over 1 - alpha, one rounding a step, where
  // train_backward.glsl:119 takes the exp of the difference of two log sums.
  float Br = bgR, Bg = bgG, Bb = bgB;
  float Tbehind = exp_finite(logTotal);

  const uint32_t blockBit = 1u << (tileBits + geo.block);
  const uint64_t insideLanes = vote(inside);
  for (uint32_t pos = first < range.y ? first : range.y; pos < range.y;) {
    __syncthreads(); // the previous batch and queue are read
    const uint32_t queued = gather_entries<false>(entryKeys, blockBit, pos, range.y, queue);
    __syncthreads();
    const uint32_t kept = stage_batch(splats, entryValues, queue, queued, geo, batch, batchIndex);
    for (uint32_t k = 0; k < kept; ++k) {
      const Staged s = batch[k];
      const bool live = s.entry >= stop, covered = covers(s, bits);
      const uint64_t candidates = insideLanes & vote(live) & vote(covered);
      if (candidates == 0) continue;
      const bool candidate = inside && live && covered;
      const float dx = px - s.cx, dy = py - s.cy;
      const float power = power_at(s, dx, dy);
      const float e = exp_finite(power);
      const float a = s.o * e;
      const float alpha = a < 0.99f ? a : 0.99f;
      const bool front = !(power > 0.0f), visible = alpha >= kMinAlpha;
      if ((candidates & vote(front) & vote(visible)) == 0) continue;
      const bool ok = candidate && front && visible;
      // Used where ok only: there 0.01 <= 1 - alpha <= 1 and the transmittance
      // is at least the pixel's final one, which the forward keeps at or above
      // 1e-4, so the short division is exact `Tbehind / (1 - alpha)`.
      const float T = div_in_range(Tbehind, 1.0f - alpha);
      // A lane that takes nothing from the splat carries zero factors, so its
      // values below are all zero (every other factor is finite); the sign of
      // a zero cannot change a total that is not zero.
      const float Ta = ok ? T * alpha : 0.0f;
      // At the 0.99 clamp alpha no longer depends on the parameters.
      const bool free = ok && a < 0.99f;
      const float dAlpha = free ? T * fma32(adj.z, s.b - Bb, fma32(adj.y, s.g - Bg, adj.x * (s.r - Br))) : 0.0f;
      const float dPower = dAlpha * alpha;
      const float vx = fma32(s.qy, dy, s.qx * dx), vy = fma32(s.qz, dy, s.qy * dx);
      const float g1 = dPower * vx, g2 = dPower * vy;
      const float hp = 0.5f * dPower;
      constexpr int kValues = Score ? 16 : 12;
      float v[kValues];
      v[0] = dAlpha * e;
      v[1] = g1;
      v[2] = g2;
      v[3] = (hp * vx) * vx;
      v[4] = (hp * vx) * vy;
      v[5] = (hp * vy) * vy;
      v[6] = Ta * adj.x;
      v[7] = Ta * adj.y;
      v[8] = Ta * adj.z;
      v[9] = Ta * perPixel;
      v[10] = fabsf(g1);
      v[11] = fabsf(g2);
      if constexpr (Score) {
        v[12] = ok ? adj.w * perPixel : 0.0f;
        v[13] = (edgeWeight * Ta) * perPixel;
        v[14] = v[15] = 0.0f;
      }
      // The colour behind and the transmittance in front: unchanged where alpha is taken as 0.
      const float alphaBehind = ok ? alpha : 0.0f;
      Br = fma32(alphaBehind, s.r - Br, Br);
      Bg = fma32(alphaBehind, s.g - Bg, Bg);
      Bb = fma32(alphaBehind, s.b - Bb, Bb);
      Tbehind = ok ? T : Tbehind;
      uint32_t value;
      bool emits;
      if constexpr (Score) {
        reduce_wave16(v, lane);
        value = lane >> 2;
        emits = (lane & 3u) == 0 && value < kRawPerSplat;
      } else {
        // Values 12 and 13 of the kRawPerSplat slots are zero without scoring.
        reduce_wave(v, lane);
        value = reduced_value(lane);
        emits = wave_emits(lane);
      }
      if (emits) {
        const size_t slot = static_cast<size_t>(batchIndex[k]) * kRawPerSplat + value;
        if constexpr (Fixed) {
          // Totals under 2^31 steps (nearly all) round and convert in 32 bits,
          // exactly as quantise() would; the wave takes quantise() if any is larger.
          const float scaled = v[0] * fixedScale;
          int64_t q;
          if (__builtin_amdgcn_ballot_w64(!(fabsf(scaled) < 2147483648.0f)) == 0u) {
            q = static_cast<int32_t>(__builtin_rintf(scaled));
          } else {
            q = quantise(v[0], fixedScale);
          }
          if (q != 0) atomicAdd(&rawFixed[slot], static_cast<unsigned long long>(q));
        } else {
          if (v[0] != 0.0f) atomicAdd(&rawFloat[slot], v[0]);
        }
      }
    }
  }
}


template <bool Fixed, bool Recompute, bool Score, uint32_t TW, uint32_t TH>
__global__ __launch_bounds__(kBlock) void backward_kernel(const Splat *__restrict__ splats, const uint32_t *__restrict__ entryKeys,
                                                          const uint32_t *__restrict__ entryValues, uint32_t tileBits,
                                                          const uint2 *__restrict__ ranges, uint32_t width, uint32_t height, uint32_t gridW,
                                                          float bgR, float bgG, float bgB, const float4 *__restrict__ image,
                                                          const float4 *__restrict__ adjoint, const uint32_t *__restrict__ stops,
                                                          float fixedScale, unsigned long long *__restrict__ rawFixed,
                                                          float *__restrict__ rawFloat, const float *__restrict__ edge,
                                                          const uint32_t *__restrict__ order) {
  __shared__ Staged batch[kBlock];
  __shared__ uint32_t batchIndex[kBlock];
  __shared__ uint32_t queue[kBlock];
  const uint32_t item = order != nullptr ? order[blockIdx.x] : blockIdx.x;
  backward_block<Fixed, Recompute, Score, TW, TH>(item, splats, entryKeys, entryValues, tileBits, ranges, width, height, gridW, bgR,
                                                  bgG, bgB, image, adjoint, stops, fixedScale, rawFixed, rawFloat, edge, batch, batchIndex,
                                                  queue);
}

// The backward pass of a batch's views in one launch, one workgroup per
// (view, 8x8 block) pair; each view's sums go to its own raw buffer.
template <bool Fixed, bool Recompute, uint32_t TW, uint32_t TH>
__global__ __launch_bounds__(kBlock) void backward_views_kernel(RasterViews views, float bgR, float bgG, float bgB) {
  __shared__ Staged batch[kBlock];
  __shared__ uint32_t batchIndex[kBlock];
  __shared__ uint32_t queue[kBlock];
  uint32_t item = blockIdx.x;
  const RasterView &w = views.view[view_of(views, item)];
  backward_block<Fixed, Recompute, false, TW, TH>(item, w.splats, w.entryKeys, w.entryValues, w.tileBits, w.ranges, w.width, w.height,
                                                  w.gridW, bgR, bgG, bgB, w.image, w.adjoint, w.stops, w.fixedScale,
                                                  reinterpret_cast<unsigned long long *>(w.rawFixed), w.rawFloat, nullptr, batch,
                                                  batchIndex, queue);
}

constexpr uint32_t kOrderThreads = 1024;
// Keys kept in LDS between the two passes, up to this many items per XCD
// (a 2880x2880 view of 8x8 blocks); past it the second pass reads the work again.
constexpr uint32_t kOrderStaged = 16384;

// Adds each lane's item to its key's count, one atomic per distinct key of the
// wave (most blocks share a few small keys), and returns the lane's place
// among its key's items.
__device__ __forceinline__ uint32_t key_place(uint32_t *counts, uint32_t key, bool valid) {
  const uint32_t lane = threadIdx.x & 63u;
  uint64_t same = __ballot(valid);
#pragma unroll
  for (uint32_t b = 0; b < 8; ++b) { // key < kOrderKeys
    const uint64_t ones = __ballot((key >> b) & 1u);
    same &= ((key >> b) & 1u) ? ones : ~ones;
  }
  const uint32_t lower = static_cast<uint32_t>(__popcll(same & ((uint64_t{1} << lane) - 1u)));
  uint32_t base = 0;
  if (valid && lower == 0) base = atomicAdd(&counts[key], static_cast<uint32_t>(__popcll(same)));
  const int leader = valid ? __builtin_ctzll(same) : static_cast<int>(lane);
  return static_cast<uint32_t>(__shfl(static_cast<int>(base), leader, 64)) + lower;
}

// A raster launch order, heaviest first: one workgroup per XCD x sorts the
// launch indices h = 8 i + x of XCD x by key(h) (a counting sort over
// kOrderKeys keys in LDS) and gives the j-th to launch index 8 j + x, which
// runs on XCD x too (so a tile's blocks still share one L2). Items of one key
// land in any order, which changes nothing: each block's pixels are its own
// and its sums are added with atomics. `items` is a multiple of 8 (raster_blocks).
template <class Key>
__device__ __forceinline__ void xcd_order(const Key &key, uint32_t items, uint32_t *__restrict__ order) {
  __shared__ uint32_t counts[kOrderKeys];
  __shared__ uint32_t waveTotals[kOrderThreads / 64];
  __shared__ uint8_t staged[kOrderStaged];
  const uint32_t x = blockIdx.x, n = items / kRasterXcds;
  for (uint32_t k = threadIdx.x; k < kOrderKeys; k += kOrderThreads) counts[k] = 0;
  __syncthreads();
  // Every lane runs every trip, so that the ballots see whole waves.
  for (uint32_t start = 0; start < n; start += kOrderThreads) {
    const uint32_t i = start + threadIdx.x;
    const bool valid = i < n;
    const uint32_t k = valid ? key(i * kRasterXcds + x) : 0u;
    if (valid && i < kOrderStaged) staged[i] = static_cast<uint8_t>(k);
    (void)key_place(counts, k, valid);
  }
  __syncthreads();
  const uint32_t c = threadIdx.x < kOrderKeys ? counts[threadIdx.x] : 0u;
  uint32_t before;
  block_exclusive_sum(c, before, waveTotals);
  if (threadIdx.x < kOrderKeys) counts[threadIdx.x] = before;
  __syncthreads();
  for (uint32_t start = 0; start < n; start += kOrderThreads) {
    const uint32_t i = start + threadIdx.x;
    const bool valid = i < n;
    const uint32_t k = !valid ? 0u : i < kOrderStaged ? staged[i] : key(i * kRasterXcds + x);
    const uint32_t j = key_place(counts, k, valid);
    if (valid) order[j * kRasterXcds + x] = i * kRasterXcds + x;
  }
}

// The backward's key: the entries the forward found that block walks.
struct WorkKey {
  const uint32_t *work;
  __device__ uint32_t operator()(uint32_t h) const { return heavy_first(work[h]); }
};

// The forward's key: the length of the block's tile list (block_geometry's
// mapping from launch index to tile), which the forward's walk follows.
struct ListKey {
  const uint2 *ranges;
  uint32_t tiles, perTile;
  __device__ uint32_t operator()(uint32_t h) const {
    const uint32_t tile = (h / kRasterXcds / perTile) * kRasterXcds + h % kRasterXcds;
    if (tile >= tiles) return kOrderKeys - 1u;
    const uint2 r = ranges[tile];
    return heavy_first(r.y - r.x);
  }
};

__global__ __launch_bounds__(kOrderThreads) void work_order_kernel(const uint32_t *__restrict__ work, uint32_t items,
                                                                    uint32_t *__restrict__ order) {
  xcd_order(WorkKey{work}, items, order);
}

__global__ __launch_bounds__(kOrderThreads) void list_order_kernel(const uint2 *__restrict__ ranges, uint32_t tiles, uint32_t perTile,
                                                                    uint32_t items, uint32_t *__restrict__ order) {
  xcd_order(ListKey{ranges, tiles, perTile}, items, order);
}
} // namespace

void launch_work_order(hipStream_t stream, const uint32_t *work, uint32_t items, uint32_t *order) {
  if (items == 0) return;
  hipLaunchKernelGGL(work_order_kernel, dim3(kRasterXcds), dim3(kOrderThreads), 0, stream, work, items, order);
}

void launch_list_order(hipStream_t stream, const uint2 *ranges, uint32_t tiles, uint32_t tileW, uint32_t tileH, uint32_t *order) {
  const uint32_t items = raster_blocks(tiles, tileW, tileH);
  if (items == 0) return;
  hipLaunchKernelGGL(list_order_kernel, dim3(kRasterXcds), dim3(kOrderThreads), 0, stream, ranges, tiles, (tileW / 8u) * (tileH / 8u), items,
                     order);
}


// Calls LAUNCH(TW, TH) with the tile shape as compile-time constants.
#define VSPLAT_TILE_DISPATCH(tileW, tileH, LAUNCH)                                                                                 \
  if ((tileW) == 8 && (tileH) == 8) {                                                                                              \
    LAUNCH(8, 8);                                                                                                                  \
  } else if ((tileW) == 32 && (tileH) == 8) {                                                                                      \
    LAUNCH(32, 8);                                                                                                                 \
  } else if ((tileW) == 32 && (tileH) == 32) {                                                                                     \
    LAUNCH(32, 32);                                                                                                                \
  } else {                                                                                                                         \
    LAUNCH(16, 16);                                                                                                                \
  }

void launch_forward(hipStream_t stream, const Splat *splats, const uint32_t *entryKeys, const uint32_t *entryValues, uint32_t tileBits,
                    const uint2 *ranges, uint32_t width, uint32_t height, uint32_t tileW, uint32_t tileH, const float background[3],
                    float4 *image, uint32_t *stops, uint32_t *work, const uint32_t *order) {
  const uint32_t gridW = (width + tileW - 1) / tileW, gridH = (height + tileH - 1) / tileH;
#define VSPLAT_FORWARD(W, H)                                                                                                       \
  hipLaunchKernelGGL((forward_kernel<W, H>), dim3(raster_blocks(gridW * gridH, W, H)), dim3(kBlock), 0, stream, splats, entryKeys,   \
                     entryValues, tileBits, ranges, width, height, gridW, background[0], background[1], background[2], image, stops, work,  \
                     order)
  VSPLAT_TILE_DISPATCH(tileW, tileH, VSPLAT_FORWARD)
#undef VSPLAT_FORWARD
}

void launch_backward(hipStream_t stream, bool fixedPoint, bool recompute, const Splat *splats, const uint32_t *entryKeys,
                     const uint32_t *entryValues, uint32_t tileBits, const uint2 *ranges, uint32_t width, uint32_t height, uint32_t tileW,
                     uint32_t tileH, const float background[3], const float4 *image, const float4 *adjoint, const uint32_t *stops,
                     float fixedScale, int64_t *rawFixed, float *rawFloat, const float *edge, const uint32_t *order) {
  const uint32_t gridW = (width + tileW - 1) / tileW, gridH = (height + tileH - 1) / tileH;
#define VSPLAT_BACKWARD_SHAPE(F, R, S, W, H)                                                                                       \
  hipLaunchKernelGGL((backward_kernel<F, R, S, W, H>), dim3(raster_blocks(gridW * gridH, W, H)), dim3(kBlock), 0, stream, splats,   \
                     entryKeys, entryValues, tileBits, ranges, width, height, gridW, background[0], background[1], background[2], image, \
                     adjoint, stops, fixedScale, reinterpret_cast<unsigned long long *>(rawFixed), rawFloat, edge, order)
#define VSPLAT_BACKWARD_TTT(W, H) VSPLAT_BACKWARD_SHAPE(true, true, true, W, H)
#define VSPLAT_BACKWARD_TTF(W, H) VSPLAT_BACKWARD_SHAPE(true, true, false, W, H)
#define VSPLAT_BACKWARD_TFT(W, H) VSPLAT_BACKWARD_SHAPE(true, false, true, W, H)
#define VSPLAT_BACKWARD_TFF(W, H) VSPLAT_BACKWARD_SHAPE(true, false, false, W, H)
#define VSPLAT_BACKWARD_FTT(W, H) VSPLAT_BACKWARD_SHAPE(false, true, true, W, H)
#define VSPLAT_BACKWARD_FTF(W, H) VSPLAT_BACKWARD_SHAPE(false, true, false, W, H)
#define VSPLAT_BACKWARD_FFT(W, H) VSPLAT_BACKWARD_SHAPE(false, false, true, W, H)
#define VSPLAT_BACKWARD_FFF(W, H) VSPLAT_BACKWARD_SHAPE(false, false, false, W, H)
  const bool score = edge != nullptr;
  const int variant = (fixedPoint ? 4 : 0) | (recompute ? 2 : 0) | (score ? 1 : 0);
  switch (variant) {
  case 7: VSPLAT_TILE_DISPATCH(tileW, tileH, VSPLAT_BACKWARD_TTT) break;
  case 6: VSPLAT_TILE_DISPATCH(tileW, tileH, VSPLAT_BACKWARD_TTF) break;
  case 5: VSPLAT_TILE_DISPATCH(tileW, tileH, VSPLAT_BACKWARD_TFT) break;
  case 4: VSPLAT_TILE_DISPATCH(tileW, tileH, VSPLAT_BACKWARD_TFF) break;
  case 3: VSPLAT_TILE_DISPATCH(tileW, tileH, VSPLAT_BACKWARD_FTT) break;
  case 2: VSPLAT_TILE_DISPATCH(tileW, tileH, VSPLAT_BACKWARD_FTF) break;
  case 1: VSPLAT_TILE_DISPATCH(tileW, tileH, VSPLAT_BACKWARD_FFT) break;
  default: VSPLAT_TILE_DISPATCH(tileW, tileH, VSPLAT_BACKWARD_FFF) break;
  }
#undef VSPLAT_BACKWARD_FFF
#undef VSPLAT_BACKWARD_FFT
#undef VSPLAT_BACKWARD_FTF
#undef VSPLAT_BACKWARD_FTT
#undef VSPLAT_BACKWARD_TFF
#undef VSPLAT_BACKWARD_TFT
#undef VSPLAT_BACKWARD_TTF
#undef VSPLAT_BACKWARD_TTT
#undef VSPLAT_BACKWARD_SHAPE
}

// The batched raster launches: one grid over every view's 8x8 blocks.
void launch_forward_views(hipStream_t stream, const RasterViews &views, uint32_t tileW, uint32_t tileH, const float background[3]) {
  if (views.count == 0) return;
  const uint32_t blocks = views.view[views.count - 1].firstBlock + views.view[views.count - 1].blocks;
#define VSPLAT_FORWARD_VIEWS(W, H)                                                                                                 \
  hipLaunchKernelGGL((forward_views_kernel<W, H>), dim3(blocks), dim3(kBlock), 0, stream, views, background[0], background[1],      \
                     background[2])
  VSPLAT_TILE_DISPATCH(tileW, tileH, VSPLAT_FORWARD_VIEWS)
#undef VSPLAT_FORWARD_VIEWS
}

void launch_backward_views(hipStream_t stream, bool fixedPoint, bool recompute, const RasterViews &views, uint32_t tileW, uint32_t tileH,
                           const float background[3]) {
  if (views.count == 0) return;
  const uint32_t blocks = views.view[views.count - 1].firstBlock + views.view[views.count - 1].blocks;
#define VSPLAT_BACKWARD_VIEWS_SHAPE(F, R, W, H)                                                                                     \
  hipLaunchKernelGGL((backward_views_kernel<F, R, W, H>), dim3(blocks), dim3(kBlock), 0, stream, views, background[0], background[1], \
                     background[2])
#define VSPLAT_BACKWARD_VIEWS_TT(W, H) VSPLAT_BACKWARD_VIEWS_SHAPE(true, true, W, H)
#define VSPLAT_BACKWARD_VIEWS_TF(W, H) VSPLAT_BACKWARD_VIEWS_SHAPE(true, false, W, H)
#define VSPLAT_BACKWARD_VIEWS_FT(W, H) VSPLAT_BACKWARD_VIEWS_SHAPE(false, true, W, H)
#define VSPLAT_BACKWARD_VIEWS_FF(W, H) VSPLAT_BACKWARD_VIEWS_SHAPE(false, false, W, H)
  if (fixedPoint && recompute) {
    VSPLAT_TILE_DISPATCH(tileW, tileH, VSPLAT_BACKWARD_VIEWS_TT)
  } else if (fixedPoint) {
    VSPLAT_TILE_DISPATCH(tileW, tileH, VSPLAT_BACKWARD_VIEWS_TF)
  } else if (recompute) {
    VSPLAT_TILE_DISPATCH(tileW, tileH, VSPLAT_BACKWARD_VIEWS_FT)
  } else {
    VSPLAT_TILE_DISPATCH(tileW, tileH, VSPLAT_BACKWARD_VIEWS_FF)
  }
#undef VSPLAT_BACKWARD_VIEWS_FF
#undef VSPLAT_BACKWARD_VIEWS_FT
#undef VSPLAT_BACKWARD_VIEWS_TF
#undef VSPLAT_BACKWARD_VIEWS_TT
#undef VSPLAT_BACKWARD_VIEWS_SHAPE
}

} // namespace vsplat
