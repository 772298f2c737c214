// Tile lists: the (tile, splat) entries in depth order at the offsets the scan
// of the tile counts gave (train_bin_emit.comp:18-37), and each tile's run of
// entries after the stable sort by tile (train_tile_ranges.comp:11-22). The
// counts themselves come from the projection (project.hip).
#include "device.hpp"
#include "kernels.hpp"
#include "blockscan.hpp"
#include "tiles.hpp"

namespace vsplat {

namespace {

constexpr uint32_t kScanThreads = 256;
constexpr uint32_t kScanPerThread = 4;
constexpr uint32_t kScanTile = kScanThreads * kScanPerThread; // 1024 splats per workgroup

// The list offsets: the exclusive sum, in depth order, of the tile counts the
// projection left in Gaussian order (train_bin_count.comp's counts, scanned as
// scan_blocks.comp does), in two launches over chunks of kScanTile splats.
// coverage_count_kernel gathers each chunk's counts in depth order into the
// offsets buffer (an invisible splat counts 0) and stores the chunk's total and
// its visible splats; coverage_offsets_kernel adds up the totals of the chunks
// before its own and turns its counts into offsets in place. The last chunk
// also stores the visible count and the entries and decides whether they fit:
// the later kernels only list and sort what fits. A view that does not fit sets
// the sticky overflow word; from then on every queued pass lists nothing and
// its chain rule does nothing, so the passes that ran are exactly those before
// the first overflow.
//
// totals: per chunk, the entries (totals[c]) and the visible splats
// (totals[chunks + c]), each written before it is read.
__global__ __launch_bounds__(kScanThreads) void coverage_count_kernel(const uint32_t *__restrict__ keys, const uint32_t *__restrict__ values,
                                                                      const uint32_t *__restrict__ tileCount, uint32_t count,
                                                                      uint32_t *__restrict__ offsets, uint32_t *__restrict__ totals) {
  __shared__ uint32_t waveTotals[kScanThreads / 64];
  const uint32_t chunk = blockIdx.x;
  uint32_t sum = 0, seen = 0;
#pragma unroll
  for (uint32_t k = 0; k < kScanPerThread; ++k) {
    const uint32_t r = chunk * kScanTile + k * kScanThreads + threadIdx.x;
    const bool visible = r < count && keys[r] != kInvisible;
    const uint32_t v = visible ? tileCount[values[r]] : 0u;
    if (r < count) offsets[r] = v;
    sum += v;
    seen += visible ? 1u : 0u;
  }
  uint32_t unused;
  const uint32_t chunkTotal = block_exclusive_sum(sum, unused, waveTotals);
  const uint32_t chunkSeen = block_exclusive_sum(seen, unused, waveTotals);
  if (threadIdx.x == 0) {
    totals[chunk] = chunkTotal;
    totals[gridDim.x + chunk] = chunkSeen;
  }
}

__global__ __launch_bounds__(kScanThreads) void coverage_offsets_kernel(uint32_t count, uint32_t *__restrict__ offsets,
                                                                        const uint32_t *__restrict__ totals, uint32_t *counters,
                                                                        uint32_t capacity, uint32_t isStep) {
  __shared__ uint32_t waveTotals[kScanThreads / 64];
  const uint32_t chunk = blockIdx.x;
  const bool last = chunk == gridDim.x - 1;
  // The entries of the chunks before this one (and, for the last, every
  // chunk's visible splats).
  uint32_t earlier = 0, seen = 0;
  for (uint32_t c = threadIdx.x; c < chunk; c += kScanThreads) earlier += totals[c];
  if (last) {
    for (uint32_t c = threadIdx.x; c < gridDim.x; c += kScanThreads) seen += totals[gridDim.x + c];
  }
  uint32_t unused;
  earlier = block_exclusive_sum(earlier, unused, waveTotals);
  // Each thread its kScanPerThread consecutive splats: their counts, their sum,
  // the exclusive sum over the chunk's threads, then the running offsets.
  const uint32_t first = chunk * kScanTile + threadIdx.x * kScanPerThread;
  uint32_t v[kScanPerThread];
  uint32_t sum = 0;
#pragma unroll
  for (uint32_t k = 0; k < kScanPerThread; ++k) {
    v[k] = first + k < count ? offsets[first + k] : 0u;
    sum += v[k];
  }
  uint32_t exclusive;
  const uint32_t chunkTotal = block_exclusive_sum(sum, exclusive, waveTotals);
  uint32_t running = earlier + exclusive;
#pragma unroll
  for (uint32_t k = 0; k < kScanPerThread; ++k) {
    if (first + k < count) offsets[first + k] = running;
    running += v[k];
  }
  if (last) {
    const uint32_t visible = block_exclusive_sum(seen, unused, waveTotals);
    if (threadIdx.x == 0) {
      const uint32_t entries = earlier + chunkTotal;
      counters[kVisible] = visible;
      counters[kEntries] = entries;
      if (counters[kOverflow] != 0 || entries > capacity) {
        counters[kEntriesUsed] = 0;
        counters[kOverflow] = 1;
        counters[kSkipped] += isStep;
      } else {
        counters[kEntriesUsed] = entries;
      }
    }
  }
}

// A batch's views each have their own counters and loss sums: each view's go
// to its slot as they are, and the first view's counters gather the sticky
// overflow word of all (a view that did not fit). One workgroup; the slots are
// written before the first view's counters change.
static_assert(kViewStatusBytes == kCounterWords * 4 + 2 * sizeof(double), "a view's slot: its counters, then its two sums");
__global__ __launch_bounds__(128) void gather_views_kernel(ViewStatus v, uint8_t *__restrict__ slots, uint32_t *status) {
  constexpr uint32_t kWords = kViewStatusBytes / 4;
  for (uint32_t k = threadIdx.x; k < v.count * kWords; k += blockDim.x) {
    const uint32_t view = k / kWords, w = k - view * kWords;
    uint32_t *slot = reinterpret_cast<uint32_t *>(slots + static_cast<size_t>(v.first + view) * kViewStatusBytes);
    slot[w] = w < kCounterWords ? v.counters[view][w] : reinterpret_cast<const uint32_t *>(v.sums[view])[w - kCounterWords];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t overflow = 0;
    for (uint32_t view = 0; view < v.count; ++view) {
      if (v.first + view != 0) overflow |= v.counters[view][kOverflow];
    }
    status[kOverflow] |= overflow;
  }
}

// A wave takes 64 consecutive depth-sorted splats and walks their candidate
// tiles 64 at a time (tiles.hpp), reading each small splat's outcome from the
// bits the projection kept. An entry goes to its splat's first entry (the
// scan's offset) plus the number of its splat's reached tiles before it, so a
// round's stores are contiguous except where a big splat's run lies between
// two small ones. A big splat's first entry is left in its record for
// big_emit_kernel. An entry's key is its tile in the low tileBits bits and,
// above them, the tile's 8x8 pixel blocks its splat can reach (block_mask;
// with blockMasks off, all of them): the sort orders by the tile alone, and
// the raster pass stages an entry only in the blocks whose bit is set.
__global__ __launch_bounds__(256) void bin_emit_kernel(const uint32_t *__restrict__ keys, const uint32_t *__restrict__ values,
                                                       const Splat *__restrict__ splats, TileRecord *__restrict__ records,
                                                       uint32_t count, uint32_t gridW, TileShape shape, uint32_t width, uint32_t height,
                                                       uint32_t tileBits, bool blockMasks, const uint32_t *__restrict__ offsets,
                                                       const uint32_t *counters, uint32_t *__restrict__ entryKeys,
                                                       uint32_t *__restrict__ entryValues) {
  __shared__ EmitOwner owners[256]; // per wave: its small splats with candidates, packed
  __shared__ uint32_t flags[256];   // per wave: the walk's round marks
  if (counters[kOverflow] != 0) return; // the same for every thread of the launch
  const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x & ~63u;
  uint32_t candidates = 0, i = 0;
  TileRecord rec{};
  if (r < count && keys[r] != kInvisible) {
    i = values[r];
    rec = records[i];
    candidates = (rec.size & 0xFFFFu) * (rec.size >> 16);
  }
  if (candidates > kMaskTiles) {
    records[i].maskLo = offsets[r]; // a big splat: listed by big_emit_kernel
    candidates = 0;
  }
  const bool owner = candidates != 0;
  const uint64_t owning = __ballot(owner);
  const uint32_t end = wave_inclusive_sum(candidates, lane);
  const uint32_t total = __shfl(end, 63, 64);
  if (total == 0) return; // nothing in this wave; no barrier follows
  const uint32_t start = end - candidates;
  if (owner) {
    const uint32_t boxWidth = rec.size & 0xFFFFu, tx0 = rec.origin & 0xFFFFu, ty0 = rec.origin >> 16;
    const uint32_t tx1 = tx0 + boxWidth - 1u, ty1 = ty0 + (rec.size >> 16) - 1u;
    const uint32_t k = static_cast<uint32_t>(__popcll(owning & ((1ull << lane) - 1ull)));
    uint32_t edges = 0;
    if (blockMasks) {
      const float4 c = reinterpret_cast<const float4 *>(splats + i)[2]; // cx, cy, ex, ey
      edges = box_block_edges(splat_pixel_box(c.x, c.y, c.z, c.w), tx0, ty0, tx1, ty1, shape, width, height);
    }
    owners[wave + k] = EmitOwner{tx0, ty0, boxWidth, start, 1.0f / static_cast<float>(boxWidth), i, rec.maskLo, rec.maskHi, offsets[r],
                                 tx1 | (ty1 << 16), edges, 0u};
  }
  const uint32_t allBlocks = (1u << ((shape.w >> 3) * (shape.h >> 3))) - 1u;
  __builtin_amdgcn_wave_barrier();
  const uint32_t limit = counters[kEntriesUsed];
  uint32_t before = 0;
  for (uint32_t base = 0; base < total; base += 64) {
    const uint32_t k = walk_round(flags + wave, lane, base, owner, start, before);
    const uint32_t j = base + lane;
    if (j < total) {
      const EmitOwner &o = owners[wave + k];
      const uint32_t local = j - o.start;
      const uint64_t mask = static_cast<uint64_t>(o.maskLo) | (static_cast<uint64_t>(o.maskHi) << 32);
      if (((mask >> local) & 1ull) != 0) {
        uint32_t tx, ty;
        box_tile(o, local, tx, ty);
        const uint32_t at = o.offset + static_cast<uint32_t>(__popcll(mask & ((1ull << local) - 1ull)));
        // The hits are the ones the projection counted, so at < limit; the test only guards the buffer.
        if (at < limit) {
          const uint32_t blocks = blockMasks ? block_mask(o.edges, tx, ty, o.tx0, o.ty0, o.last & 0xFFFFu, o.last >> 16, shape) : allBlocks;
          entryKeys[at] = (ty * gridW + tx) | (blocks << tileBits);
          entryValues[at] = o.gaussian;
        }
      }
    }
  }
}

// Big splats' entries, listed by the waves: a wave loads 64 listed splats at
// once (one per lane), then walks each one's box in turn, 64 candidates a
// round, row-major, each round's outcome read from the bits the counting pass
// kept (or tested again where it kept none), the reached ones numbered by a
// prefix count from the splat's first entry. The listed splats are dealt out
// over all the waves. Their entries mark every 8x8 block of the tile (the
// tiles inside a big box are covered whole).
__global__ __launch_bounds__(256) void big_emit_kernel(const Splat *__restrict__ splats, const TileRecord *__restrict__ records,
                                                       const uint32_t *__restrict__ bigList, const uint64_t *__restrict__ bigBits,
                                                       uint32_t gridW, TileShape shape, uint32_t tileBits, const uint32_t *counters,
                                                       uint32_t *__restrict__ entryKeys, uint32_t *__restrict__ entryValues) {
  if (counters[kOverflow] != 0) return; // the same for every thread of the launch
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t waves = gridDim.x * (blockDim.x / 64u), wave = (blockIdx.x * blockDim.x + threadIdx.x) / 64u;
  const uint32_t listed = counters[kBigSplats], limit = counters[kEntriesUsed];
  const uint32_t allBlocks = ((1u << ((shape.w >> 3) * (shape.h >> 3))) - 1u) << tileBits;
  // Wave w takes listed splats w, w + waves, w + 2 waves, ..., 64 at a time (one
  // per lane), so that a short list still spreads over every wave.
  for (uint32_t base = 0; wave + base * waves < listed; base += 64u) {
    const uint32_t mineAt = wave + (base + lane) * waves;
    BigSplat mine{};
    uint32_t bits = kNoBits;
    if (mineAt < listed) {
      const uint32_t i = bigList[mineAt];
      const TileRecord rec = records[i];
      bits = rec.maskHi;
      TileTest test{};
      if (bits == kNoBits) {
        const Splat s = load_splat(splats, i);
        test = tile_test(s.qx, s.qy, s.qz, s.o, s.cx, s.cy);
      }
      mine = BigSplat{test, rec.origin & 0xFFFFu, rec.origin >> 16, rec.size & 0xFFFFu, (rec.size & 0xFFFFu) * (rec.size >> 16), i,
                      rec.maskLo};
    }
    const uint32_t inChunk = static_cast<uint32_t>(__popcll(__ballot(mineAt < listed))); // a prefix of the lanes
    for (uint32_t j = 0; j < inChunk; ++j) {
      const BigSplat b = big_splat_of_lane(mine, j);
      const uint32_t bitStart = lane_u32(bits, j);
      BoxCursor cursor(lane, b.width);
      uint32_t next = b.first; // the splat's first entry
      for (uint32_t k = lane, round = 0; k - lane < b.candidates; k += 64u, ++round) {
        const uint32_t tx = b.tx0 + cursor.col, ty = b.ty0 + cursor.row;
        // The counting pass's outcome of this round, or the test again where it kept none.
        const uint64_t mask = bitStart != kNoBits ? bigBits[bitStart + round]
                                                  : __ballot(k < b.candidates && reaches_tile(b.test, tx, ty, shape));
        const bool hit = ((mask >> lane) & 1ull) != 0;
        const uint32_t at = next + static_cast<uint32_t>(__popcll(mask & ((1ull << lane) - 1ull)));
        if (hit && at < limit) {
          entryKeys[at] = (ty * gridW + tx) | allBlocks;
          entryValues[at] = b.gaussian;
        }
        next += static_cast<uint32_t>(__popcll(mask));
        cursor.advance();
      }
    }
  }
}

// Start and end of each tile's run; every write belongs to one thread
// (train_tile_ranges.comp:11-22). A thread takes four consecutive keys with one
// 16-byte load, and the keys either side of them; a key's tile is its low bits
// (tileMask), its block bits above them.
__global__ __launch_bounds__(256) void tile_ranges_kernel(const uint32_t *__restrict__ entryKeys, const uint32_t *counters,
                                                          uint32_t tileMask, uint2 *__restrict__ ranges) {
  const uint32_t count = counters[kEntriesUsed];
  const uint32_t quads = (count + 3u) / 4u;
  for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < quads; q += gridDim.x * blockDim.x) {
    const uint32_t k0 = q * 4u;
    uint32_t key[4];
    if (k0 + 4u <= count) {
      const uint4 v = reinterpret_cast<const uint4 *>(entryKeys)[q];
      key[0] = v.x;
      key[1] = v.y;
      key[2] = v.z;
      key[3] = v.w;
    } else {
#pragma unroll
      for (uint32_t j = 0; j < 4u; ++j) key[j] = k0 + j < count ? entryKeys[k0 + j] : ~0u;
    }
#pragma unroll
    for (uint32_t j = 0; j < 4u; ++j) key[j] &= tileMask;
    const uint32_t before = k0 == 0 ? ~0u : entryKeys[k0 - 1u] & tileMask;
    const uint32_t after = k0 + 4u < count ? entryKeys[k0 + 4u] & tileMask : ~0u;
#pragma unroll
    for (uint32_t j = 0; j < 4u; ++j) {
      const uint32_t k = k0 + j;
      if (k >= count) break;
      const uint32_t prev = j == 0 ? before : key[j - 1u];
      const uint32_t next = j == 3u ? after : key[j + 1u];
      if (k == 0 || prev != key[j]) ranges[key[j]].x = k;
      if (k + 1u == count || next != key[j]) ranges[key[j]].y = k + 1u;
    }
  }
}

} // namespace

size_t coverage_scan_words(uint32_t count) { return 2 * static_cast<size_t>(count == 0 ? 1 : (count + kScanTile - 1) / kScanTile); }

void launch_coverage_scan(hipStream_t stream, const uint32_t *keys, const uint32_t *values, const uint32_t *tileCount, uint32_t count,
                          uint32_t *offsets, uint32_t *counters, uint32_t capacity, bool isStep, uint32_t *totals) {
  const uint32_t chunks = count == 0 ? 1 : (count + kScanTile - 1) / kScanTile;
  hipLaunchKernelGGL(coverage_count_kernel, dim3(chunks), dim3(kScanThreads), 0, stream, keys, values, tileCount, count, offsets, totals);
  hipLaunchKernelGGL(coverage_offsets_kernel, dim3(chunks), dim3(kScanThreads), 0, stream, count, offsets, totals, counters, capacity,
                     isStep ? 1u : 0u);
}

void launch_gather_views(hipStream_t stream, const ViewStatus &views, uint8_t *slots, uint32_t *status) {
  if (views.count == 0) return;
  hipLaunchKernelGGL(gather_views_kernel, dim3(1), dim3(128), 0, stream, views, slots, status);
}

void launch_bin_emit(hipStream_t stream, const uint32_t *keys, const uint32_t *values, const Splat *splats, TileRecord *records,
                     uint32_t count, uint32_t gridW, uint32_t tileW, uint32_t tileH, uint32_t width, uint32_t height, uint32_t tileBits,
                     const uint32_t *offsets, const uint32_t *counters, const uint32_t *bigList, const uint64_t *bigBits,
                     uint32_t *entryKeys, uint32_t *entryValues) {
  if (count == 0) return;
  // Blocks are marked per entry for tiles of 16 blocks (32x32), where a block
  // skipping the entries that miss it saves more raster time than marking
  // costs the list pass; smaller tiles mark every block.
  const bool blockMasks = (tileW / 8u) * (tileH / 8u) >= 16u;
  hipLaunchKernelGGL(bin_emit_kernel, dim3((count + 255) / 256), dim3(256), 0, stream, keys, values, splats, records, count, gridW,
                     TileShape{tileW, tileH}, width, height, tileBits, blockMasks, offsets, counters, entryKeys, entryValues);
  hipLaunchKernelGGL(big_emit_kernel, dim3(kBigSplatBlocks), dim3(256), 0, stream, splats, records, bigList, bigBits, gridW,
                     TileShape{tileW, tileH}, tileBits, counters, entryKeys, entryValues);
}

void launch_tile_ranges(hipStream_t stream, const uint32_t *entryKeys, const uint32_t *counters, uint32_t capacity, uint32_t tileBits,
                        uint2 *ranges) {
  if (capacity == 0) return;
  const uint32_t quads = (capacity + 3) / 4;
  const uint32_t blocks = (quads + 255) / 256 < 16384 ? (quads + 255) / 256 : 16384;
  hipLaunchKernelGGL(tile_ranges_kernel, dim3(blocks), dim3(256), 0, stream, entryKeys, counters, (1u << tileBits) - 1u, ranges);
}

} // namespace vsplat
