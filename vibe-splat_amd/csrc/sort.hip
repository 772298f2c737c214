// Stable LSD radix sort, the rasterizer's depth sort and tile sort (the
// reference's radix_histogram.comp, radix_scatter.comp and scan_blocks.comp,
// driven by src/core/radix_sort.cpp:58-96). Per digit pass, three launches:
//
//   histogram  per tile of 2048 or 4096 keys, how many hold each digit, stored
//              digit-major (hist[digit * tiles + tile]);
//   scan       one workgroup per digit: the exclusive sum over tiles in place,
//              and the digit's total;
//   scatter    per tile: where each digit starts (the exclusive sum of the
//              digit totals), plus the tile's offset, plus each key's rank
//              among the tile's keys of its digit.
//
// Ranking uses wave64 ballots instead of the reference's barrier-separated
// 1-bit splits: for each digit bit a lane's ballot, and-ed together, gives the
// lanes of its wave holding the same digit, and its rank among them is the
// popcount below it. Each wave ranks a quarter of the tile's keys in
// order, keeping its running digit counts in LDS, so the sort is stable, as
// the tile lists need (test_radix_sort.cpp:15-60 checks the reference's
// against std::stable_sort). The tile's keys then go through LDS in digit
// order, so that each digit's run is written to memory contiguously.
//
// Counts may live in device memory (the tile entries are only known on the
// device): grids are sized for the capacity and tiles past the count exit.
#include "device.hpp"
#include "kernels.hpp"
#include "blockscan.hpp"

namespace vsplat {

namespace {

constexpr uint32_t kSortThreads = 256;
// Keys per thread: 8 (tiles of 2048) below kLargeSort keys, 16 (tiles of 4096)
// from there; the larger tiles halve the histogram and its scan, but at 1M keys
// they would leave the 256 compute units short of work.
constexpr uint32_t kSortPerThreadSmall = 8, kSortPerThreadLarge = 16;
constexpr uint32_t kLargeSort = 4u << 20;
constexpr uint32_t kScanThreads = 1024;

__device__ __forceinline__ uint32_t count_of(const uint32_t *countDev, uint32_t countHost) {
  return countDev != nullptr ? *countDev : countHost;
}

template <uint32_t Bits, uint32_t PerThread>
__global__ __launch_bounds__(kSortThreads) void radix_histogram_kernel(const uint32_t *__restrict__ keys, const uint32_t *countDev,
                                                                       uint32_t countHost, uint32_t shift, uint32_t tiles,
                                                                       uint32_t *__restrict__ hist) {
  constexpr uint32_t kRadix = 1u << Bits;
  constexpr uint32_t kSortTile = kSortThreads * PerThread;
  __shared__ uint32_t counts[kRadix];
  for (uint32_t d = threadIdx.x; d < kRadix; d += kSortThreads) counts[d] = 0;
  __syncthreads();
  const uint32_t count = count_of(countDev, countHost);
  const uint32_t base = blockIdx.x * kSortTile;
  if (base < count) {
#pragma unroll
    for (uint32_t r = 0; r < PerThread; ++r) {
      const uint32_t i = base + r * kSortThreads + threadIdx.x;
      if (i < count) atomicAdd(&counts[(keys[i] >> shift) & (kRadix - 1u)], 1u);
    }
  }
  __syncthreads();
  for (uint32_t d = threadIdx.x; d < kRadix; d += kSortThreads) hist[d * tiles + blockIdx.x] = counts[d];
}

// One workgroup per digit: the exclusive sum of that digit's counts over the
// tiles, in place, and the digit's total.
__global__ __launch_bounds__(kScanThreads) void radix_scan_kernel(uint32_t *__restrict__ hist, uint32_t tiles,
                                                                  uint32_t *__restrict__ totals) {
  __shared__ uint32_t waveTotals[kScanThreads / 64];
  uint32_t *row = hist + static_cast<size_t>(blockIdx.x) * tiles;
  uint32_t carry = 0;
  for (uint32_t start = 0; start < tiles; start += kScanThreads) {
    const uint32_t t = start + threadIdx.x;
    const uint32_t v = t < tiles ? row[t] : 0u;
    uint32_t exclusive;
    const uint32_t sum = block_exclusive_sum(v, exclusive, waveTotals);
    if (t < tiles) row[t] = carry + exclusive;
    carry += sum;
  }
  if (threadIdx.x == 0) totals[blockIdx.x] = carry;
}

template <uint32_t Bits, uint32_t PerThread>
__global__ __launch_bounds__(kSortThreads) void radix_scatter_kernel(const uint32_t *__restrict__ keysIn, const uint32_t *__restrict__ valsIn,
                                                                     uint32_t *__restrict__ keysOut, uint32_t *__restrict__ valsOut,
                                                                     const uint32_t *countDev, uint32_t countHost, uint32_t shift,
                                                                     uint32_t tiles, const uint32_t *__restrict__ hist,
                                                                     const uint32_t *__restrict__ totals) {
  constexpr uint32_t kRadix = 1u << Bits;
  constexpr uint32_t kWaves = kSortThreads / 64;
  constexpr uint32_t kSortTile = kSortThreads * PerThread;
  constexpr uint32_t kWaveKeys = kSortTile / kWaves; // consecutive keys per wave
  __shared__ uint32_t digitBase[kRadix]; // the digit's global start for this tile, less its start within the tile
  __shared__ uint32_t waveCount[kWaves][kRadix];
  __shared__ uint32_t waveTotals[kWaves];
  __shared__ uint32_t sortedKeys[kSortTile], sortedValues[kSortTile];
  const uint32_t count = count_of(countDev, countHost);
  const uint32_t tileStart = blockIdx.x * kSortTile;
  if (tileStart >= count) return; // the same for the whole workgroup
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;

  // Where this tile's keys of each digit go: the digit's start plus the tile's offset.
  const uint32_t total = threadIdx.x < kRadix ? totals[threadIdx.x] : 0u;
  uint32_t start;
  block_exclusive_sum(total, start, waveTotals);
  if (threadIdx.x < kRadix) digitBase[threadIdx.x] = start + hist[threadIdx.x * tiles + blockIdx.x];
  for (uint32_t x = threadIdx.x; x < kWaves * kRadix; x += kSortThreads) (&waveCount[0][0])[x] = 0;
  __syncthreads();

  // Rank within the wave's 512 keys, in order.
  const uint64_t below = (uint64_t{1} << lane) - 1u;
  uint32_t key[PerThread], val[PerThread], rank[PerThread];
  const uint32_t first = tileStart + wave * kWaveKeys + lane;
#pragma unroll
  for (uint32_t r = 0; r < PerThread; ++r) {
    const uint32_t i = first + r * 64;
    key[r] = i < count ? keysIn[i] : 0u;
    val[r] = i < count ? valsIn[i] : 0u;
  }
#pragma unroll
  for (uint32_t r = 0; r < PerThread; ++r) {
    const bool valid = first + r * 64 < count;
    const uint32_t digit = (key[r] >> shift) & (kRadix - 1u);
    uint64_t same = __ballot(valid);
#pragma unroll
    for (uint32_t b = 0; b < Bits; ++b) {
      const uint64_t ones = __ballot((digit >> b) & 1u);
      same &= ((digit >> b) & 1u) ? ones : ~ones;
    }
    const uint32_t lower = static_cast<uint32_t>(__popcll(same & below));
    __builtin_amdgcn_wave_barrier();
    const uint32_t prior = waveCount[wave][digit];
    rank[r] = prior + lower;
    __builtin_amdgcn_wave_barrier();
    if (valid && lower == 0) waveCount[wave][digit] = prior + static_cast<uint32_t>(__popcll(same));
  }
  __syncthreads();
  // Per digit: where its keys start within the tile (in digit order), and each
  // wave's offset from there.
  {
    const uint32_t d = threadIdx.x;
    uint32_t tileCount = 0;
    if (d < kRadix) {
#pragma unroll
      for (uint32_t w = 0; w < kWaves; ++w) tileCount += waveCount[w][d];
    }
    uint32_t localStart;
    block_exclusive_sum(tileCount, localStart, waveTotals);
    if (d < kRadix) {
      digitBase[d] -= localStart;
      uint32_t running = localStart;
#pragma unroll
      for (uint32_t w = 0; w < kWaves; ++w) {
        const uint32_t c = waveCount[w][d];
        waveCount[w][d] = running;
        running += c;
      }
    }
  }
  __syncthreads();
  // The tile's keys in digit order in LDS, then written out run by run: a
  // digit's keys are consecutive in both, so neighbouring threads store to
  // neighbouring addresses.
#pragma unroll
  for (uint32_t r = 0; r < PerThread; ++r) {
    if (first + r * 64 < count) {
      const uint32_t at = waveCount[wave][(key[r] >> shift) & (kRadix - 1u)] + rank[r];
      sortedKeys[at] = key[r];
      sortedValues[at] = val[r];
    }
  }
  __syncthreads();
  const uint32_t inTile = count - tileStart < kSortTile ? count - tileStart : kSortTile;
#pragma unroll
  for (uint32_t r = 0; r < PerThread; ++r) {
    const uint32_t j = r * kSortThreads + threadIdx.x;
    if (j < inTile) {
      const uint32_t k = sortedKeys[j];
      const uint32_t dst = digitBase[(k >> shift) & (kRadix - 1u)] + j;
      keysOut[dst] = k;
      valsOut[dst] = sortedValues[j];
    }
  }
}

uint32_t div_up(uint64_t a, uint32_t b) { return static_cast<uint32_t>((a + b - 1) / b); }

template <uint32_t Bits, uint32_t PerThread>
void sort_pass(hipStream_t stream, const uint32_t *kIn, const uint32_t *vIn, uint32_t *kOut, uint32_t *vOut, const uint32_t *countDev,
               uint32_t capacity, uint32_t shift, uint32_t *scratch) {
  constexpr uint32_t kRadix = 1u << Bits;
  const uint32_t tiles = div_up(capacity, kSortThreads * PerThread);
  uint32_t *totals = scratch;
  uint32_t *hist = scratch + 256;
  hipLaunchKernelGGL((radix_histogram_kernel<Bits, PerThread>), dim3(tiles), dim3(kSortThreads), 0, stream, kIn, countDev, capacity,
                     shift, tiles, hist);
  hipLaunchKernelGGL(radix_scan_kernel, dim3(kRadix), dim3(kScanThreads), 0, stream, hist, tiles, totals);
  hipLaunchKernelGGL((radix_scatter_kernel<Bits, PerThread>), dim3(tiles), dim3(kSortThreads), 0, stream, kIn, vIn, kOut, vOut,
                     countDev, capacity, shift, tiles, hist, totals);
}

template <uint32_t Bits>
void sort_pass(hipStream_t stream, const uint32_t *kIn, const uint32_t *vIn, uint32_t *kOut, uint32_t *vOut, const uint32_t *countDev,
               uint32_t capacity, uint32_t shift, uint32_t *scratch) {
  if (capacity >= kLargeSort) {
    sort_pass<Bits, kSortPerThreadLarge>(stream, kIn, vIn, kOut, vOut, countDev, capacity, shift, scratch);
  } else {
    sort_pass<Bits, kSortPerThreadSmall>(stream, kIn, vIn, kOut, vOut, countDev, capacity, shift, scratch);
  }
}

} // namespace

size_t radix_sort_scratch_words(uint64_t capacity) {
  return 256 + static_cast<size_t>(div_up(capacity, kSortThreads * kSortPerThreadSmall)) * 256u;
}

// Sorts (keys, values) stably by the low `bits` bits of the keys (bits <= 32),
// in ceil(bits / 8) passes of equal width. The result ends in keys/values;
// tmpKeys/tmpValues are the other half of each pass. countDev, when given,
// holds the element count (at most capacity).
void launch_radix_sort(hipStream_t stream, uint32_t *keys, uint32_t *values, uint32_t *tmpKeys, uint32_t *tmpValues,
                       const uint32_t *countDev, uint32_t capacity, uint32_t bits, uint32_t *scratch) {
  if (capacity <= 1 || bits == 0) return;
  const uint32_t passes = div_up(bits, 8);
  const uint32_t digit = div_up(bits, passes); // even split, at most 8 bits a pass
  uint32_t *srcK = keys, *srcV = values, *dstK = tmpKeys, *dstV = tmpValues;
  for (uint32_t p = 0; p < passes; ++p) {
    // The last pass takes the bits that are left, so that no bit above `bits`
    // (which may carry other data) takes part in the order.
    const uint32_t shift = p * digit;
    switch (digit < bits - shift ? digit : bits - shift) {
    case 1: sort_pass<1>(stream, srcK, srcV, dstK, dstV, countDev, capacity, shift, scratch); break;
    case 2: sort_pass<2>(stream, srcK, srcV, dstK, dstV, countDev, capacity, shift, scratch); break;
    case 3: sort_pass<3>(stream, srcK, srcV, dstK, dstV, countDev, capacity, shift, scratch); break;
    case 4: sort_pass<4>(stream, srcK, srcV, dstK, dstV, countDev, capacity, shift, scratch); break;
    case 5: sort_pass<5>(stream, srcK, srcV, dstK, dstV, countDev, capacity, shift, scratch); break;
    case 6: sort_pass<6>(stream, srcK, srcV, dstK, dstV, countDev, capacity, shift, scratch); break;
    case 7: sort_pass<7>(stream, srcK, srcV, dstK, dstV, countDev, capacity, shift, scratch); break;
    default: sort_pass<8>(stream, srcK, srcV, dstK, dstV, countDev, capacity, shift, scratch); break;
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
