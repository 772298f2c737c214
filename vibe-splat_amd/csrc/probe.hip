// Elementwise runs of the device's single-precision functions, so that tests
// can hold them against the CPU model bit for bit (vsplat_math_probe).
#include "device.hpp"
#include "trainer.hpp"

namespace vsplat {
namespace {

__global__ void math_probe_kernel(int function, const float *__restrict__ in, float *__restrict__ out, uint64_t n) {
  const uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
  if (i >= n) return;
  const float x = in[i];
  float y = 0.0f;
  if (function == 6) {
    // Pairs: the quotient of (in[2k], in[2k + 1]) goes to out[2k], out[2k + 1] is 0.
    out[i] = (i & 1u) != 0 || i + 1 >= n ? 0.0f : div_in_range(x, in[i + 1]);
    return;
  }
  switch (function) {
  case 0: y = vs_exp(x); break;
  case 1: y = vs_log(x); break;
  case 2: y = sqrtf(x); break;
  case 3: y = 1.0f / x; break;
  case 4: y = half_to_float(float_to_half(x)); break;
  default: y = half_to_float(as_uint(x) & 0xFFFFu); break;
  }
  out[i] = y;
}

void check(hipError_t e, const char *what) {
  if (e != hipSuccess) throw Error(VSPLAT_DEVICE_ERROR, std::string(what) + ": " + hipGetErrorString(e));
}

} // namespace

void math_probe(int function, const float *in, float *out, uint64_t n) {
  if (function < 0 || function > 6) throw Error(VSPLAT_INVALID_ARGUMENT, "math probe functions are 0 to 6");
  if (n == 0) return;
  DeviceArray a, b;
  a.ensure(n * 4);
  b.ensure(n * 4);
  check(hipMemcpy(a.as<float>(), in, n * 4, hipMemcpyHostToDevice), "probe upload");
  const uint64_t blocks = (n + 255) / 256;
  hipLaunchKernelGGL(math_probe_kernel, dim3(static_cast<uint32_t>(blocks)), dim3(256), 0, nullptr, function, a.as<float>(), b.as<float>(), n);
  check(hipGetLastError(), "probe launch");
  check(hipMemcpy(out, b.as<float>(), n * 4, hipMemcpyDeviceToHost), "probe download");
}

} // namespace vsplat
