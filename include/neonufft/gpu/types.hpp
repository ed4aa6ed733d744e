#pragma once

#include <neonufft/config.h>

#include <hip/hip_complex.h>
#include <hip/hip_runtime_api.h>

#include <type_traits>

namespace neonufft {
namespace gpu {

/* Device complex and stream types (include/neonufft/gpu/types.hpp:25-33). */
using ComplexDoubleType = hipDoubleComplex;
using ComplexFloatType = hipComplex;
using StreamType = hipStream_t;

template <typename T>
using ComplexType = std::conditional_t<std::is_same_v<T, float>, ComplexFloatType, ComplexDoubleType>;

}  // namespace gpu
}  // namespace neonufft
