// The per-device scratch array of set_points without allocation.
#include "noalloc.hpp"

namespace nnf {

// File-scope device array: one copy per GPU, at an address only the runtime
// knows. In host code `scratch` is a host-side stand-in; its device address
// comes from hipGetSymbolAddress and is the only form ever passed to kernels
// or copies.
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_device_lock[kMaxDevices];
std::mutex g_table_lock;
void* g_scratch_ptr[kMaxDevices] = {};

void* scratch_address(int device) {
  std::lock_guard<std::mutex> guard(g_table_lock);
  if (g_scratch_ptr[device]) return g_scratch_ptr[device];
  int current = 0;
  check_hip(hipGetDevice(&current), "hipGetDevice");
  if (current != device) check_hip(hipSetDevice(device), "hipSetDevice");
  void* ptr = nullptr;
  const hipError_t err = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (current != device) check_hip(hipSetDevice(current), "hipSetDevice");
  check_hip(err, "hipGetSymbolAddress(scratch)");
  if (!ptr) throw neonufft::GPUError("Neonufft:: scratch array has no device address");
  g_scratch_ptr[device] = ptr;
  return ptr;
}
}  // namespace

ScratchLease lease_noalloc_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) {
    throw neonufft::InternalError("Neonufft:: device index outside the scratch table");
  }
  void* ptr = scratch_address(device);
  ScratchLease lease{std::unique_lock<std::mutex>(g_device_lock[device]), Workspace{}};
  lease.ws.base = static_cast<char*>(ptr);
  lease.ws.cap = kNoAllocScratchBytes;
  lease.ws.used = 0;
  return lease;
}

}  // namespace nnf
