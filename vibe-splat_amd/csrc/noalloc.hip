// The allocation-free render's scratch: one static device array, which every
// GPU the library runs on holds until the process ends (vsplat_render_noalloc,
// include/vsplat.h). In host code the array's name is a host-side stand-in, so
// its device address is looked up once per device, with that device current,
// and only that address is handed to kernels and copies.
#include "trainer.hpp"

#include <mutex>
#include <string>

namespace vsplat {

namespace {

static __device__ __align__(256) unsigned char scratch[VSPLAT_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex scratchMutex;
void *scratchAddress[kMaxDevices] = {};

} // namespace

void *noalloc_scratch(uint64_t *bytes) {
  int device = -1;
  const hipError_t got = hipGetDevice(&device);
  if (got != hipSuccess) throw Error(VSPLAT_DEVICE_ERROR, std::string("hipGetDevice: ") + hipGetErrorString(got));
  if (device < 0 || device >= kMaxDevices) throw Error(VSPLAT_DEVICE_ERROR, "device index out of range");
  std::lock_guard<std::mutex> lock(scratchMutex);
  if (scratchAddress[device] == nullptr) {
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || ptr == nullptr) {
      throw Error(VSPLAT_DEVICE_ERROR, std::string("hipGetSymbolAddress of the render scratch: ") + hipGetErrorString(e));
    }
    scratchAddress[device] = ptr;
  }
  *bytes = VSPLAT_NOALLOC_SCRATCH_BYTES;
  return scratchAddress[device];
}

} // namespace vsplat
