// The per-Gaussian arrays of a changed cloud, gathered from the old ones on the
// GPU (train_gather.comp:16-24, trainer.cpp:749-787): element k of new Gaussian
// g is element k of old Gaussian source[g], or 0 when source[g] is negative (a
// new Gaussian starts with fresh optimiser state). One launch covers every
// array: blockIdx.y picks the array.
#include "kernels.hpp"

namespace vsplat {
namespace {

struct GatherList {
  GatherArray array[kMaxGatherArrays];
};

__global__ __launch_bounds__(256) void gather_kernel(const int32_t *__restrict__ source, uint32_t gaussians, GatherList list) {
  const GatherArray a = list.array[blockIdx.y];
  const uint32_t total = gaussians * a.stride; // below 2^30: the launcher's chunks
  for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < total; i += gridDim.x * 256u) {
    const uint32_t g = i / a.stride, k = i - g * a.stride;
    const int32_t from = source[g];
    a.to[i] = from < 0 ? 0u : a.from[static_cast<uint64_t>(from) * a.stride + k];
  }
}

} // namespace

void launch_gather(hipStream_t stream, const int32_t *source, uint32_t gaussians, const GatherArray *arrays, uint32_t count) {
  if (gaussians == 0 || count == 0) return;
  uint32_t widest = 1;
  for (uint32_t j = 0; j < count; ++j) widest = arrays[j].stride > widest ? arrays[j].stride : widest;
  const uint32_t chunk = (1u << 30) / widest; // Gaussians per launch, so that every element index fits 30 bits
  for (uint32_t first = 0; first < gaussians; first += chunk) {
    const uint32_t n = gaussians - first < chunk ? gaussians - first : chunk;
    GatherList list{};
    for (uint32_t j = 0; j < count; ++j) {
      list.array[j] = {arrays[j].from, arrays[j].to + static_cast<uint64_t>(first) * arrays[j].stride, arrays[j].stride};
    }
    const uint64_t blocks64 = (static_cast<uint64_t>(n) * widest + 255) / 256;
    const uint32_t blocks = blocks64 < 4096 ? static_cast<uint32_t>(blocks64) : 4096;
    hipLaunchKernelGGL(gather_kernel, dim3(blocks, count), dim3(256), 0, stream, source + first, n, list);
  }
}

} // namespace vsplat
