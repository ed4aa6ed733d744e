#pragma once

#include <neonufft/allocator.hpp>
#include <neonufft/config.h>

#include <cstddef>

namespace neonufft {
namespace gpu {

/* hipMalloc / hipFree allocator, the default of every GPU plan
 * (include/neonufft/gpu/device_allocator.hpp:14-21). */
class NEONUFFT_EXPORT DeviceAllocator : public Allocator {
public:
  DeviceAllocator() = default;
  void* allocate(std::size_t size) override;
  void deallocate(void* ptr) noexcept override;
};

}  // namespace gpu
}  // namespace neonufft
