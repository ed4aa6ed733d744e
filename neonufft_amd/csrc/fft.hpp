// In-place, unnormalised complex-to-complex FFT on rocFFT.
//
// Same transform as the reference's FFTGrid (src/neonufft/gpu/util/fft_grid.cpp:18-70):
// grid stored column-major (x fastest), exp(sign * 2 pi i m g / n) per
// dimension, i.e. forward for sign < 0 and backward (unnormalised) otherwise.
#pragma once

#include <memory>

#include "common.hpp"

struct rocfft_plan_t;
struct rocfft_execution_info_t;

namespace nnf {

class FftPlan {
public:
  FftPlan() = default;
  // n: dims fastest-first; batch: number of contiguous grids transformed per call.
  FftPlan(bool is_double, int dim, const IntType* n, int sign, IntType batch, hipStream_t stream,
          std::shared_ptr<neonufft::Allocator> alloc);
  FftPlan(const FftPlan&) = delete;
  FftPlan& operator=(const FftPlan&) = delete;
  FftPlan(FftPlan&& o) noexcept;
  FftPlan& operator=(FftPlan&& o) noexcept;
  ~FftPlan();

  void execute(void* data) const;
  bool valid() const { return plan_ != nullptr; }
  // Changes the stream later executions are enqueued on.
  void set_stream(hipStream_t stream);

private:
  void destroy() noexcept;
  rocfft_plan_t* plan_ = nullptr;
  rocfft_execution_info_t* info_ = nullptr;
  DeviceBuffer<char> work_;
};

}  // namespace nnf
