// Single-pass prefix sums across workgroups by decoupled look-back (Merrill
// and Garland, "Single-pass Parallel Prefix Scan with Decoupled Look-back",
// NVIDIA technical report NVR-2016-002), as the tile-count scan uses it.
//
// Workgroups take tile numbers from an atomic ticket, so a tile's
// predecessors were all started before it and finish without waiting on it.
// Each tile publishes a 64-bit status word: a tag in the high half (which
// launch it belongs to, and whether the low half is the tile's own total or
// the inclusive sum up to it) and the value in the low half. Status words are
// read and written with agent-scope atomics, so they are coherent across the
// GPU's L2 caches; such a read crosses the chip, so one wave reads 64
// predecessors at a time.
//
// A wait that does not end within kSpinLimit polls (about a tenth of a second)
// gives up, sets *fault and returns what it has: the kernel then finishes with
// wrong sums, and the host reports the fault, rather than the GPU hanging.
#pragma once

#include "device.hpp"

namespace vsplat {

constexpr uint32_t kTagAggregate = 1u, kTagInclusive = 2u;
constexpr uint32_t kSpinLimit = 1u << 20;

__device__ __forceinline__ uint64_t status_word(uint32_t tag, uint32_t value) {
  return (static_cast<uint64_t>(tag) << 32) | value;
}

__device__ __forceinline__ void status_publish(uint64_t *slot, uint32_t tag, uint32_t value) {
  __hip_atomic_store(slot, status_word(tag, value), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
  for (uint32_t d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
  return v;
}

// The sum of the values of tiles 0 .. tile-1, called by all 64 lanes of one
// wave. `base` is the tag prefix of this launch. Lane l reads tile end-1-l:
// when the nearest inclusive word in the window has every nearer word ready,
// the sum is done; when the whole window holds totals, it moves 64 back.
__device__ __forceinline__ uint32_t look_back_window(const uint64_t *status, uint32_t tile, uint32_t base, uint32_t lane,
                                                     uint32_t *fault) {
  uint32_t sum = 0, spins = 0;
  uint32_t end = tile;
  while (end > 0) {
    const int t = static_cast<int>(end) - 1 - static_cast<int>(lane);
    // Before tile 0 there is an inclusive sum of nothing.
    const uint64_t w = t >= 0 ? __hip_atomic_load(status + t, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                              : status_word(base | kTagInclusive, 0u);
    const uint32_t tag = static_cast<uint32_t>(w >> 32);
    const bool inclusive = tag == (base | kTagInclusive);
    const bool ready = inclusive || tag == (base | kTagAggregate);
    const uint64_t inclusiveLanes = __ballot(inclusive), readyLanes = __ballot(ready);
    if (inclusiveLanes != 0) {
      const uint32_t nearest = static_cast<uint32_t>(__builtin_ctzll(inclusiveLanes));
      const uint64_t needed = nearest == 63u ? ~0ull : ((2ull << nearest) - 1ull);
      if ((readyLanes & needed) == needed) return sum + wave_sum(lane <= nearest ? static_cast<uint32_t>(w) : 0u);
    } else if (readyLanes == ~0ull) {
      sum += wave_sum(static_cast<uint32_t>(w));
      end -= 64;
      spins = 0;
      continue;
    }
    if (++spins > kSpinLimit) {
      if (lane == 0) atomicOr(fault, 1u);
      return sum;
    }
    __builtin_amdgcn_s_sleep(2);
  }
  return sum;
}

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
