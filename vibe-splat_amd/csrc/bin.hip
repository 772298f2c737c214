// Tile lists: the (tile, splat) entries in depth order at the offsets the scan
// of the tile counts gave (train_bin_emit.comp:18-37), and each tile's run of
// entries after the stable sort by tile (train_tile_ranges.comp:11-22). The
// counts themselves come from the projection (project.hip).
#include "device.hpp"
#include "kernels.hpp"
#include "tiles.hpp"

namespace vsplat {

namespace {

// The tile count of each depth-sorted splat, gathered from Gaussian order (the
// projection counted them), ready for the exclusive scan that gives the offsets.
__global__ __launch_bounds__(256) void gather_counts_kernel(const uint32_t *__restrict__ keys, const uint32_t *__restrict__ values,
                                                            const uint32_t *__restrict__ tileCount, uint32_t count,
                                                            uint32_t *__restrict__ coverage) {
  const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < count) coverage[r] = keys[r] == kInvisible ? 0u : tileCount[values[r]];
}

// Whether the entries fit. A view that does not fit sets the sticky overflow
// word; from then on every queued pass lists nothing and its chain rule does
// nothing, so the passes that ran are exactly those before the first overflow.
__global__ void bin_check_kernel(uint32_t *counters, uint32_t capacity, uint32_t isStep) {
  const uint32_t entries = counters[kEntries];
  if (counters[kOverflow] != 0 || entries > capacity) {
    counters[kEntriesUsed] = 0;
    counters[kOverflow] = 1;
    counters[kSkipped] += isStep;
  } else {
    counters[kEntriesUsed] = entries;
  }
}

// A wave takes 64 consecutive depth-sorted splats, whose entries fill one
// contiguous run of the list starting at the first one's offset. It walks their
// candidate tiles 64 at a time (tiles.hpp); the reached ones of a round are
// numbered by a ballot prefix count, so each round's stores are contiguous.
// Small splats read the outcome from the bits the projection kept; large ones
// repeat the test on the stored splat.
__global__ __launch_bounds__(256) void bin_emit_kernel(const uint32_t *__restrict__ keys, const uint32_t *__restrict__ values,
                                                       const Splat *__restrict__ splats, const TileRecord *__restrict__ records,
                                                       uint32_t count, uint32_t gridW, const uint32_t *__restrict__ offsets,
                                                       const uint32_t *counters, uint32_t *__restrict__ entryKeys,
                                                       uint32_t *__restrict__ entryValues) {
  __shared__ TileTest tests[256];
  __shared__ uint32_t boxes[256][3]; // tx0, ty0, width
  __shared__ uint64_t masks[256];
  __shared__ uint32_t gaussian[256];
  __shared__ uint32_t ends[256];
  if (counters[kOverflow] != 0) return; // the same for every thread of the launch
  const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x & ~63u;
  uint32_t candidates = 0;
  if (r < count && keys[r] != kInvisible) {
    const uint32_t i = values[r];
    const TileRecord rec = records[i];
    const uint32_t width = rec.size & 0xFFFFu, height = rec.size >> 16;
    candidates = width * height;
    boxes[threadIdx.x][0] = rec.origin & 0xFFFFu;
    boxes[threadIdx.x][1] = rec.origin >> 16;
    boxes[threadIdx.x][2] = width;
    masks[threadIdx.x] = static_cast<uint64_t>(rec.maskLo) | (static_cast<uint64_t>(rec.maskHi) << 32);
    gaussian[threadIdx.x] = i;
    if (candidates > kMaskTiles) {
      const Splat s = load_splat(splats, i);
      tests[threadIdx.x] = tile_test(s.qx, s.qy, s.qz, s.o, s.cx, s.cy);
    }
  }
  const uint32_t end = wave_inclusive_sum(candidates, lane);
  ends[threadIdx.x] = end;
  const uint32_t total = __shfl(end, 63, 64);
  __syncthreads();
  if (total == 0) return; // nothing in this wave; no barrier follows
  uint32_t slot = offsets[r - lane]; // the wave's first splat is in range when it has candidates
  const uint32_t limit = counters[kEntriesUsed];
  for (uint32_t base = 0; base < total; base += 64) {
    const uint32_t j = base + lane;
    bool hit = false;
    uint32_t owner = wave, tx = 0, ty = 0;
    if (j < total) {
      owner = wave + owner_of(ends + wave, j);
      const uint32_t local = j - (owner == wave ? 0u : ends[owner - 1]);
      const uint32_t w = boxes[owner][2];
      const uint32_t row = local / w;
      tx = boxes[owner][0] + (local - row * w);
      ty = boxes[owner][1] + row;
      const uint32_t size = owner == wave ? ends[owner] : ends[owner] - ends[owner - 1];
      hit = size <= kMaskTiles ? ((masks[owner] >> local) & 1ull) != 0 : reaches_tile(tests[owner], tx, ty);
    }
    const uint64_t round = __ballot(hit);
    const uint32_t at = slot + static_cast<uint32_t>(__popcll(round & ((1ull << lane) - 1ull)));
    // The hits are the ones the projection counted, so at < limit; the test only guards the buffer.
    if (hit && at < limit) {
      entryKeys[at] = ty * gridW + tx;
      entryValues[at] = gaussian[owner];
    }
    slot += static_cast<uint32_t>(__popcll(round));
  }
}

// Start and end of each tile's run; every write belongs to one thread (train_tile_ranges.comp:11-22).
__global__ __launch_bounds__(256) void tile_ranges_kernel(const uint32_t *__restrict__ entryKeys, const uint32_t *counters,
                                                          uint2 *__restrict__ ranges) {
  const uint32_t count = counters[kEntriesUsed];
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < count; k += gridDim.x * blockDim.x) {
    const uint32_t tile = entryKeys[k];
    if (k == 0 || entryKeys[k - 1] != tile) ranges[tile].x = k;
    if (k + 1 == count || entryKeys[k + 1] != tile) ranges[tile].y = k + 1;
  }
}

} // namespace

void launch_gather_counts(hipStream_t stream, const uint32_t *keys, const uint32_t *values, const uint32_t *tileCount, uint32_t count,
                          uint32_t *coverage) {
  if (count == 0) return;
  hipLaunchKernelGGL(gather_counts_kernel, dim3((count + 255) / 256), dim3(256), 0, stream, keys, values, tileCount, count, coverage);
}

void launch_bin_check(hipStream_t stream, uint32_t *counters, uint32_t capacity, bool isStep) {
  hipLaunchKernelGGL(bin_check_kernel, dim3(1), dim3(1), 0, stream, counters, capacity, isStep ? 1u : 0u);
}

void launch_bin_emit(hipStream_t stream, const uint32_t *keys, const uint32_t *values, const Splat *splats, const TileRecord *records,
                     uint32_t count, uint32_t gridW, const uint32_t *offsets, const uint32_t *counters, uint32_t *entryKeys,
                     uint32_t *entryValues) {
  if (count == 0) return;
  hipLaunchKernelGGL(bin_emit_kernel, dim3((count + 255) / 256), dim3(256), 0, stream, keys, values, splats, records, count, gridW,
                     offsets, counters, entryKeys, entryValues);
}

void launch_tile_ranges(hipStream_t stream, const uint32_t *entryKeys, const uint32_t *counters, uint32_t capacity, uint2 *ranges) {
  if (capacity == 0) return;
  const uint32_t blocks = (capacity + 255) / 256 < 16384 ? (capacity + 255) / 256 : 16384;
  hipLaunchKernelGGL(tile_ranges_kernel, dim3(blocks), dim3(256), 0, stream, entryKeys, counters, ranges);
}

} // namespace vsplat
