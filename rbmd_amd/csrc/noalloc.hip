// noalloc.hip -- the fixed device scratch of the entry point that allocates
// nothing (rbmd_md_energy_noalloc).
//
// The scratch is one file-scope device array.  Every GPU the library runs on
// holds its own copy of it, at its own address, until the process ends.  In
// host code the array's name is only a host-side stand-in, so its device
// address is taken once per device, with that device current, when the first
// context is created on it (rbmd_ctx_create), and host code uses that address
// wherever it needs the array's memory.  Every no-alloc call on a device
// shares the device's array and leaves its contents behind: calls that
// overlap on one device (two contexts, two threads) race.
#include <hip/hip_runtime.h>

#include <mutex>

#include "dev.hpp"

namespace rbmd {

static __device__ __align__(256) double scratch[kScratchBytes / sizeof(double)];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

hipError_t scratch_init(int device) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[device]) return hipSuccess;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  if (!p) return hipErrorInvalidSymbol;
  g_scratch_ptr[device] = p;
  return hipSuccess;
}

void* scratch_address(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return g_scratch_ptr[device];
}

}  // namespace rbmd
