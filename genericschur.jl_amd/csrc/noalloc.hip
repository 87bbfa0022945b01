// The fixed device scratch array of the no-allocation batched entry point
// (gs_schur_batched_z_noalloc) and the lookup of its device address.
//
// The array is a file-scope __device__ variable: every GPU the library runs
// on holds its own copy until the process ends, at its own address.  Its
// name is a host-side stand-in in host code, never a device address: host
// code only ever uses the pointer hipGetSymbolAddress returns for the
// current device, fetched once per device and cached here.
#include <hip/hip_runtime.h>

#include <mutex>

#include "internal.h"

namespace gs {

namespace {

__device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};

}  // namespace

hipError_t noalloc_scratch(int device, void** ptr) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[device]) {
    int cur = -1;
    hipError_t e = hipGetDevice(&cur);
    if (e != hipSuccess) return e;
    if (cur != device && (e = hipSetDevice(device)) != hipSuccess) return e;
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != device) (void)hipSetDevice(cur);
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    g_ptr[device] = p;
  }
  *ptr = g_ptr[device];
  return hipSuccess;
}

}  // namespace gs
