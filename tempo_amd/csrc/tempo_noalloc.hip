// tempo_hip_run_noalloc: the step for callers that cannot allocate device
// memory.  Its per-point scratch is `scratch`, a static device array of
// kNoAllocBytes per GPU (each GPU that runs this library's code holds its
// own copy until the process ends).
//
// Host code never passes the array's name to a kernel or a copy: the name is
// a host-side stand-in there.  The first call on a device, with that device
// current, looks up the device's copy with hipGetSymbolAddress and keeps the
// pointer; every later call on that device uses the kept pointer.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "tempo_capi_internal.hpp"
#include "tempo_hip.h"

namespace {

constexpr size_t kNoAllocBytes = size_t(256) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};
hipError_t g_err[kMaxDevices] = {};

// this device's copy of `scratch` (current device), looked up once
void* device_scratch(int dev) {
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_ptr[dev]) {
    void* p = nullptr;
    g_err[dev] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (g_err[dev] == hipSuccess) g_ptr[dev] = p;
  }
  return g_ptr[dev];
}

}  // namespace

extern "C" int tempo_hip_run_noalloc(tempo_hip_ctx* ctx, const tempo_hip_fields* f, const tempo_hip_diags* d,
                                     const tempo_hip_dims* dims, double dt, int itimestep, int precision,
                                     void* stream) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices)
    return tempo::set_error(TEMPO_HIP_E_HIP, "tempo_hip_run_noalloc: no current HIP device");
  void* base = device_scratch(dev);
  if (!base) {
    const std::string msg = std::string("tempo_hip_run_noalloc: hipGetSymbolAddress(scratch) failed: ") +
                            hipGetErrorString(g_err[dev]);
    return tempo::set_error(TEMPO_HIP_E_HIP, msg.c_str());
  }
  return tempo::run_with_scratch(ctx, f, d, dims, dt, itimestep, precision, stream, base, kNoAllocBytes);
}
