#include "fft.hpp"

#include <rocfft/rocfft.h>

#include <mutex>

namespace nnf {

namespace {

void check_fft(rocfft_status st, const char* what) {
  if (st != rocfft_status_success) {
    set_error_detail(std::string(what) + " failed with rocfft_status " + std::to_string(static_cast<int>(st)));
    throw neonufft::GPUFFTError("Neonufft:: rocFFT call failed");
  }
}

void rocfft_init_once() {
  static std::once_flag flag;
  std::call_once(flag, [] { check_fft(rocfft_setup(), "rocfft_setup"); });
}

}  // namespace

FftPlan::FftPlan(bool is_double, int dim, const IntType* n, int sign, IntType batch,
                 hipStream_t stream, std::shared_ptr<neonufft::Allocator> alloc) {
  rocfft_init_once();
  std::size_t lengths[3] = {1, 1, 1};
  for (int d = 0; d < dim; ++d) lengths[d] = static_cast<std::size_t>(n[d]);
  const rocfft_transform_type type =
      sign < 0 ? rocfft_transform_type_complex_forward : rocfft_transform_type_complex_inverse;
  const rocfft_precision prec = is_double ? rocfft_precision_double : rocfft_precision_single;
  try {
    check_fft(rocfft_plan_create(&plan_, rocfft_placement_inplace, type, prec,
                                 static_cast<std::size_t>(dim), lengths,
                                 static_cast<std::size_t>(batch), nullptr),
              "rocfft_plan_create");
    std::size_t work_bytes = 0;
    check_fft(rocfft_plan_get_work_buffer_size(plan_, &work_bytes),
              "rocfft_plan_get_work_buffer_size");
    check_fft(rocfft_execution_info_create(&info_), "rocfft_execution_info_create");
    if (work_bytes) {
      work_.reset(work_bytes, alloc);
      check_fft(rocfft_execution_info_set_work_buffer(info_, work_.data(), work_bytes),
                "rocfft_execution_info_set_work_buffer");
    }
    check_fft(rocfft_execution_info_set_stream(info_, stream), "rocfft_execution_info_set_stream");
  } catch (...) {
    destroy();
    throw;
  }
}

FftPlan::FftPlan(FftPlan&& o) noexcept { *this = std::move(o); }

FftPlan& FftPlan::operator=(FftPlan&& o) noexcept {
  destroy();
  plan_ = o.plan_;
  info_ = o.info_;
  work_ = std::move(o.work_);
  o.plan_ = nullptr;
  o.info_ = nullptr;
  return *this;
}

FftPlan::~FftPlan() { destroy(); }

void FftPlan::destroy() noexcept {
  if (info_) rocfft_execution_info_destroy(info_);
  if (plan_) rocfft_plan_destroy(plan_);
  info_ = nullptr;
  plan_ = nullptr;
  work_.release();
}

void FftPlan::set_stream(hipStream_t stream) {
  if (!info_) throw neonufft::InternalError("Neonufft:: FFT plan not initialised");
  check_fft(rocfft_execution_info_set_stream(info_, stream), "rocfft_execution_info_set_stream");
}

void FftPlan::execute(void* data) const {
  if (!plan_) throw neonufft::InternalError("Neonufft:: FFT plan not initialised");
  void* buffers[1] = {data};
  check_fft(rocfft_execute(plan_, buffers, nullptr, info_), "rocfft_execute");
}

}  // namespace nnf
