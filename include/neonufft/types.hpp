#pragma once

#include <cstdint>

namespace neonufft {
/* Index type of the public API (include/neonufft/types.hpp:9 of the reference). */
using IntType = std::int64_t;
}  // namespace neonufft
