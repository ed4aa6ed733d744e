#pragma once

#include <neonufft/config.h>

#include <cstddef>

namespace neonufft {

/* Abstract allocator injected into GPU plans (include/neonufft/allocator.hpp:12-21). */
class Allocator {
public:
  Allocator() = default;
  virtual ~Allocator() = default;
  virtual void* allocate(std::size_t size) = 0;
  virtual void deallocate(void* ptr) noexcept = 0;
};

}  // namespace neonufft
