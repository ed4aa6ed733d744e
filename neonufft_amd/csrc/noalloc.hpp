// Fixed device scratch for nnf_plan_set_points_noalloc: set_points without
// any device allocation, for callers that cannot allocate (serving loops
// that must not call hipMalloc, allocation-free regions of a program).
#pragma once

#include <mutex>

#include "common.hpp"

namespace nnf {

// Size of the per-device scratch array (set_points temporaries: about 4 bytes
// per point plus a few per bin / row, so up to ~16 million points per call).
constexpr std::size_t kNoAllocScratchBytes = std::size_t(64) << 20;

// The calling process holds one such array per GPU, from the library's load
// to the process's end. Every call on a GPU shares its array, so calls are
// serialised per device: the lease holds that device's lock, and its user
// must not return before the device work that reads the array has finished.
struct ScratchLease {
  std::unique_lock<std::mutex> lock;
  Workspace ws;
};

// Leases the scratch array of `device` (its address is looked up once per
// device, with that device current).
ScratchLease lease_noalloc_scratch(int device);

}  // namespace nnf
