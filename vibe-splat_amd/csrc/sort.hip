// Exclusive scan and stable LSD radix sort, the rasterizer's two integer
// primitives (the reference's radix_histogram.comp, radix_scatter.comp,
// scan_blocks.comp and scan_add.comp, driven by src/core/radix_sort.cpp:58-96).
//
// The sort ranks keys with wave64 ballots instead of the reference's eight
// barrier-separated 1-bit splits per round: for each digit bit a lane's
// __ballot, and-ed together, gives the mask of lanes in its wave holding the
// same digit; its rank is the popcount of that mask below it. Waves, rounds and
// blocks are visited in index order, so the sort is stable, as the tile lists
// need (test_radix_sort.cpp:15-60 checks the reference's against std::stable_sort).
//
// Counts may live in device memory (the tile entries are only known on the
// device): grids are sized for the capacity and blocks past the count exit.
#include "device.hpp"
#include "kernels.hpp"

#include <vector>

namespace vsplat {

namespace {

constexpr uint32_t kScanThreads = 256;
constexpr uint32_t kScanPerThread = 16;
constexpr uint32_t kScanBlock = kScanThreads * kScanPerThread; // 4096 elements per block

constexpr uint32_t kSortThreads = 256;
constexpr uint32_t kSortRounds = 8; // keys per thread
constexpr uint32_t kSortTile = kSortThreads * kSortRounds;

__device__ __forceinline__ uint32_t count_of(const uint32_t *countDev, uint32_t countHost) {
  return countDev != nullptr ? *countDev : countHost;
}

// Block-wide exclusive sum of one value per thread; returns the total.
__device__ __forceinline__ uint32_t block_exclusive_sum(uint32_t v, uint32_t &exclusive, uint32_t *waveTotals) {
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (uint32_t d = 1; d < 64; d <<= 1) {
    const uint32_t y = __shfl_up(x, d, 64);
    if (lane >= d) x += y;
  }
  if (lane == 63) waveTotals[wave] = x;
  __syncthreads();
  uint32_t before = 0, total = 0;
  for (uint32_t w = 0; w < blockDim.x / 64; ++w) {
    const uint32_t t = waveTotals[w];
    if (w < wave) before += t;
    total += t;
  }
  exclusive = before + x - v;
  __syncthreads();
  return total;
}

__global__ __launch_bounds__(kScanThreads) void scan_reduce_kernel(const uint32_t *__restrict__ data, uint32_t count,
                                                                   uint32_t *__restrict__ partials) {
  __shared__ uint32_t waveTotals[kScanThreads / 64];
  const uint32_t base = blockIdx.x * kScanBlock + threadIdx.x * kScanPerThread;
  uint32_t sum = 0;
#pragma unroll
  for (uint32_t k = 0; k < kScanPerThread; ++k) sum += base + k < count ? data[base + k] : 0u;
  uint32_t excl;
  const uint32_t total = block_exclusive_sum(sum, excl, waveTotals);
  if (threadIdx.x == 0) partials[blockIdx.x] = total;
}

// In-place exclusive scan of a block's 4096 elements plus an offset per block.
__global__ __launch_bounds__(kScanThreads) void scan_block_kernel(uint32_t *__restrict__ data, uint32_t count,
                                                                  const uint32_t *__restrict__ offsets, uint32_t *__restrict__ total) {
  __shared__ uint32_t waveTotals[kScanThreads / 64];
  const uint32_t base = blockIdx.x * kScanBlock + threadIdx.x * kScanPerThread;
  uint32_t v[kScanPerThread];
  uint32_t sum = 0;
#pragma unroll
  for (uint32_t k = 0; k < kScanPerThread; ++k) {
    v[k] = base + k < count ? data[base + k] : 0u;
    sum += v[k];
  }
  uint32_t excl;
  const uint32_t blockTotal = block_exclusive_sum(sum, excl, waveTotals);
  uint32_t running = excl + (offsets != nullptr ? offsets[blockIdx.x] : 0u);
#pragma unroll
  for (uint32_t k = 0; k < kScanPerThread; ++k) {
    if (base + k < count) data[base + k] = running;
    running += v[k];
  }
  if (total != nullptr && blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) {
    *total = (offsets != nullptr ? offsets[blockIdx.x] : 0u) + blockTotal;
  }
}

// Per block, how many of its keys fall in each digit; digit-major so that one
// exclusive scan gives every block its scatter base (radix_histogram.comp:15-30).
template <uint32_t Bits>
__global__ __launch_bounds__(kSortThreads) void radix_histogram_kernel(const uint32_t *__restrict__ keys, const uint32_t *countDev,
                                                                       uint32_t countHost, uint32_t shift, uint32_t numBlocks,
                                                                       uint32_t *__restrict__ hist) {
  constexpr uint32_t kRadix = 1u << Bits;
  __shared__ uint32_t counts[kRadix];
  for (uint32_t d = threadIdx.x; d < kRadix; d += kSortThreads) counts[d] = 0;
  __syncthreads();
  const uint32_t count = count_of(countDev, countHost);
  const uint32_t base = blockIdx.x * kSortTile;
#pragma unroll
  for (uint32_t r = 0; r < kSortRounds; ++r) {
    const uint32_t i = base + r * kSortThreads + threadIdx.x;
    if (i < count) atomicAdd(&counts[(keys[i] >> shift) & (kRadix - 1u)], 1u);
  }
  __syncthreads();
  for (uint32_t d = threadIdx.x; d < kRadix; d += kSortThreads) hist[d * numBlocks + blockIdx.x] = counts[d];
}

// Stable scatter (radix_scatter.comp:33-85), ranks from wave ballots.
template <uint32_t Bits>
__global__ __launch_bounds__(kSortThreads) void radix_scatter_kernel(const uint32_t *__restrict__ keysIn, const uint32_t *__restrict__ valsIn,
                                                                     uint32_t *__restrict__ keysOut, uint32_t *__restrict__ valsOut,
                                                                     const uint32_t *countDev, uint32_t countHost, uint32_t shift,
                                                                     uint32_t numBlocks, const uint32_t *__restrict__ offsets) {
  constexpr uint32_t kRadix = 1u << Bits;
  constexpr uint32_t kWaves = kSortThreads / 64;
  __shared__ uint32_t base[kRadix];
  __shared__ uint32_t waveCount[kWaves][kRadix];
  const uint32_t count = count_of(countDev, countHost);
  const uint32_t tileStart = blockIdx.x * kSortTile;
  if (tileStart >= count) return;
  for (uint32_t d = threadIdx.x; d < kRadix; d += kSortThreads) base[d] = offsets[d * numBlocks + blockIdx.x];
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  const uint64_t below = (uint64_t{1} << lane) - 1u;
  for (uint32_t r = 0; r < kSortRounds; ++r) {
    for (uint32_t d = threadIdx.x; d < kRadix * kWaves; d += kSortThreads) (&waveCount[0][0])[d] = 0;
    __syncthreads();
    const uint32_t i = tileStart + r * kSortThreads + threadIdx.x;
    const bool valid = i < count;
    const uint32_t key = valid ? keysIn[i] : 0u;
    const uint32_t val = valid ? valsIn[i] : 0u;
    const uint32_t digit = (key >> shift) & (kRadix - 1u);
    uint64_t same = __ballot(valid);
#pragma unroll
    for (uint32_t b = 0; b < Bits; ++b) {
      const uint64_t ones = __ballot((digit >> b) & 1u);
      same &= ((digit >> b) & 1u) ? ones : ~ones;
    }
    const uint32_t rank = __popcll(same & below);
    if (valid && rank == 0) waveCount[wave][digit] = __popcll(same);
    __syncthreads();
    // Per digit: where each wave's run starts, after the earlier waves and rounds.
    for (uint32_t d = threadIdx.x; d < kRadix; d += kSortThreads) {
      uint32_t s = base[d];
#pragma unroll
      for (uint32_t w = 0; w < kWaves; ++w) {
        const uint32_t c = waveCount[w][d];
        waveCount[w][d] = s;
        s += c;
      }
      base[d] = s;
    }
    __syncthreads();
    if (valid) {
      const uint32_t dst = waveCount[wave][digit] + rank;
      keysOut[dst] = key;
      valsOut[dst] = val;
    }
    __syncthreads();
  }
}

uint32_t div_up(uint64_t a, uint32_t b) { return static_cast<uint32_t>((a + b - 1) / b); }

} // namespace

size_t scan_scratch_words(uint64_t count) {
  size_t words = 0;
  while (count > kScanBlock) {
    count = div_up(count, kScanBlock);
    words += count + 64; // each level's partials, padded
  }
  return words + 64;
}

void launch_exclusive_scan(hipStream_t stream, uint32_t *data, uint32_t count, uint32_t *scratch, uint32_t *total) {
  if (count == 0) return;
  if (count <= kScanBlock) {
    hipLaunchKernelGGL(scan_block_kernel, dim3(1), dim3(kScanThreads), 0, stream, data, count, nullptr, total);
    return;
  }
  const uint32_t blocks = div_up(count, kScanBlock);
  uint32_t *partials = scratch;
  hipLaunchKernelGGL(scan_reduce_kernel, dim3(blocks), dim3(kScanThreads), 0, stream, data, count, partials);
  launch_exclusive_scan(stream, partials, blocks, scratch + blocks + 64, nullptr);
  hipLaunchKernelGGL(scan_block_kernel, dim3(blocks), dim3(kScanThreads), 0, stream, data, count, partials, total);
}

// The histogram (256 digits per block at most) and its scan's partials.
size_t radix_sort_scratch_words(uint64_t capacity) {
  const uint64_t hist = static_cast<uint64_t>(div_up(capacity, kSortTile)) * 256u;
  return hist + 64 + scan_scratch_words(hist);
}

template <uint32_t Bits>
static void sort_pass(hipStream_t stream, const uint32_t *kIn, const uint32_t *vIn, uint32_t *kOut, uint32_t *vOut,
                      const uint32_t *countDev, uint32_t capacity, uint32_t shift, uint32_t *hist, uint32_t *scanScratch) {
  const uint32_t blocks = div_up(capacity, kSortTile);
  hipLaunchKernelGGL((radix_histogram_kernel<Bits>), dim3(blocks), dim3(kSortThreads), 0, stream, kIn, countDev, capacity, shift,
                     blocks, hist);
  launch_exclusive_scan(stream, hist, blocks << Bits, scanScratch, nullptr);
  hipLaunchKernelGGL((radix_scatter_kernel<Bits>), dim3(blocks), dim3(kSortThreads), 0, stream, kIn, vIn, kOut, vOut, countDev,
                     capacity, shift, blocks, hist);
}

// Sorts (keys, values) stably by the low `bits` bits of the keys. The result
// ends in keys/values; tmpKeys/tmpValues are ping-pong buffers. countDev, when
// given, holds the element count (at most capacity).
void launch_radix_sort(hipStream_t stream, uint32_t *keys, uint32_t *values, uint32_t *tmpKeys, uint32_t *tmpValues,
                       const uint32_t *countDev, uint32_t capacity, uint32_t bits, uint32_t *scratch) {
  if (capacity <= 1 || bits == 0) return;
  const uint32_t passes = div_up(bits, 8);
  const uint32_t digit = div_up(bits, passes); // even split, at most 8 bits a pass
  const uint32_t blocks = div_up(capacity, kSortTile);
  uint32_t *hist = scratch;
  uint32_t *scanScratch = scratch + static_cast<size_t>(blocks) * 256u + 64;
  uint32_t *srcK = keys, *srcV = values, *dstK = tmpKeys, *dstV = tmpValues;
  for (uint32_t p = 0; p < passes; ++p) {
    const uint32_t shift = p * digit;
    switch (digit) {
    case 1: sort_pass<1>(stream, srcK, srcV, dstK, dstV, countDev, capacity, shift, hist, scanScratch); break;
    case 2: sort_pass<2>(stream, srcK, srcV, dstK, dstV, countDev, capacity, shift, hist, scanScratch); break;
    case 3: sort_pass<3>(stream, srcK, srcV, dstK, dstV, countDev, capacity, shift, hist, scanScratch); break;
    case 4: sort_pass<4>(stream, srcK, srcV, dstK, dstV, countDev, capacity, shift, hist, scanScratch); break;
    case 5: sort_pass<5>(stream, srcK, srcV, dstK, dstV, countDev, capacity, shift, hist, scanScratch); break;
    case 6: sort_pass<6>(stream, srcK, srcV, dstK, dstV, countDev, capacity, shift, hist, scanScratch); break;
    case 7: sort_pass<7>(stream, srcK, srcV, dstK, dstV, countDev, capacity, shift, hist, scanScratch); break;
    default: sort_pass<8>(stream, srcK, srcV, dstK, dstV, countDev, capacity, shift, hist, scanScratch); break;
    }
    uint32_t *t = srcK;
    srcK = dstK;
    dstK = t;
    t = srcV;
    srcV = dstV;
    dstV = t;
  }
  if (srcK != keys) {
    // An odd number of passes: the result is in the temporaries.
    launch_copy_u32(stream, srcK, keys, countDev, capacity);
    launch_copy_u32(stream, srcV, values, countDev, capacity);
  }
}

__global__ void copy_u32_kernel(const uint32_t *__restrict__ src, uint32_t *__restrict__ dst, const uint32_t *countDev, uint32_t countHost) {
  const uint32_t count = count_of(countDev, countHost);
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) dst[i] = src[i];
}

void launch_copy_u32(hipStream_t stream, const uint32_t *src, uint32_t *dst, const uint32_t *countDev, uint32_t capacity) {
  if (capacity == 0) return;
  const uint32_t blocks = div_up(capacity, 256) < 4096 ? div_up(capacity, 256) : 4096;
  hipLaunchKernelGGL(copy_u32_kernel, dim3(blocks), dim3(256), 0, stream, src, dst, countDev, capacity);
}

__global__ void fill_u32_kernel(uint32_t *__restrict__ dst, uint64_t count, uint32_t value) {
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < count; i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
    dst[i] = value;
}

// Clears with a kernel rather than hipMemsetAsync, so that the whole step can
// be captured into a graph.
void launch_fill_u32(hipStream_t stream, uint32_t *dst, uint64_t count, uint32_t value) {
  if (count == 0) return;
  const uint64_t blocks64 = (count + 255) / 256;
  const uint32_t blocks = blocks64 < 8192 ? static_cast<uint32_t>(blocks64) : 8192;
  hipLaunchKernelGGL(fill_u32_kernel, dim3(blocks), dim3(256), 0, stream, dst, count, value);
}

} // namespace vsplat
