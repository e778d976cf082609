// Static device scratch for the entry point that allocates nothing
// (solar_compute_rhs_noalloc): the four ghost-strip sets of one right-hand
// side (p' and velocity ghosts for DDX and DDY) live in a file-scope device
// array instead of buffers the handle allocates.  Every GPU the library runs
// on holds its own copy of the array, at its own address, from the moment the
// code object loads until the process ends.
//
// In host code `scratch` names a host-side stand-in, not device memory:
// kernels and copies use the per-device address that hipGetSymbolAddress
// returns, looked up once per device by scratch_address() with that device
// current.
#include <hip/hip_runtime.h>

#include <mutex>

#include "solar_internal.hpp"

namespace {

constexpr size_t kScratchBytes = size_t(128) << 20;  // 128 MiB
constexpr int kMaxDevices = 64;

static __device__ __align__(256) double scratch[kScratchBytes / sizeof(double)];

std::mutex g_mu;
void* g_addr[kMaxDevices] = {};
hipError_t g_status[kMaxDevices] = {};
bool g_done[kMaxDevices] = {};

}  // namespace

namespace solar {

size_t scratch_bytes() { return kScratchBytes; }

// The device address of `scratch` on `device`, which must be the current
// device.  The lookup runs once per device; later calls return the recorded
// address (or the recorded failure).
hipError_t scratch_address(int device, void** addr) {
  *addr = nullptr;
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_done[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    g_status[device] = e;
    g_addr[device] = e == hipSuccess ? p : nullptr;
    g_done[device] = true;
  }
  *addr = g_addr[device];
  return g_status[device];
}

}  // namespace solar
