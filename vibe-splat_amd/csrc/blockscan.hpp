// Prefix sums within a workgroup, as the tile-count scan, the radix sort and
// the backward's launch order use them.
#pragma once

#include "device.hpp"

namespace vsplat {

// Block-wide exclusive sum of one value per thread (blockDim.x a multiple of
// 64, at most 1024); returns the block total. Contains barriers: every thread
// of the block calls it.
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

} // namespace vsplat
