// Host-memory plans neonufft::Plan<T,DIM> and PlanT3<T,DIM> (declared in
// include/neonufft/plan.hpp, same interface as the reference's
// include/neonufft/plan.hpp:31-85). The reference runs these on the CPU; here
// they stage host arrays through device buffers and run the GPU plan
// implementation on a private stream of the device current at construction.
// Each call returns once its results are in the caller's host arrays.
#include <neonufft/gpu/device_allocator.hpp>
#include <neonufft/plan.hpp>

#include <algorithm>
#include <vector>

#include "plan_impl.hpp"

namespace nnf {

namespace {

// A non-blocking stream owned by a host plan (default priority).
class OwnedStream {
public:
  OwnedStream() { check_hip(hipStreamCreateWithFlags(&s_, hipStreamNonBlocking), "hipStreamCreate"); }
  ~OwnedStream() {
    if (s_) {
      (void)hipStreamSynchronize(s_);
      (void)hipStreamDestroy(s_);
    }
  }
  OwnedStream(const OwnedStream&) = delete;
  OwnedStream& operator=(const OwnedStream&) = delete;
  hipStream_t get() const { return s_; }

private:
  hipStream_t s_ = nullptr;
};

template <typename E>
void to_device(DeviceBuffer<E>& dst, const E* src, std::size_t n, hipStream_t s) {
  if (n) check_hip(hipMemcpyAsync(dst.data(), src, n * sizeof(E), hipMemcpyHostToDevice, s),
                   "hipMemcpyAsync(host to device)");
}

template <typename E>
void to_host(E* dst, const E* src, std::size_t n, hipStream_t s) {
  if (n) check_hip(hipMemcpyAsync(dst, src, n * sizeof(E), hipMemcpyDeviceToHost, s),
                   "hipMemcpyAsync(device to host)");
}

// Coordinates of DIM dimensions copied to the device (grow-only buffers).
template <typename T, int DIM>
struct DeviceCoords {
  DeviceBuffer<T> d[DIM];
  const T* ptr[DIM] = {};

  void upload(IntType m, const T* const* host, hipStream_t s,
              const std::shared_ptr<neonufft::Allocator>& alloc) {
    if (m < 0) throw neonufft::InputError("Neonufft:: negative number of points");
    const std::size_t n = static_cast<std::size_t>(m);
    for (int k = 0; k < DIM; ++k) {
      if (n && !host[k]) throw neonufft::InputError("Neonufft:: null coordinate pointer");
      d[k].ensure(std::max<std::size_t>(n, 1), alloc);
      to_device(d[k], host[k], n, s);
      ptr[k] = d[k].data();
    }
  }
};

// Element offsets of every mode of an N_0 x ... array (x fastest) under the
// caller's strides.
template <int DIM>
std::vector<std::size_t> mode_offsets(const IntType* modes, const IntType* strides) {
  std::size_t total = 1;
  for (int d = 0; d < DIM; ++d) total *= static_cast<std::size_t>(modes[d]);
  std::vector<std::size_t> off(total);
  std::size_t i = 0;
  const IntType n1 = DIM >= 2 ? modes[1] : 1, n2 = DIM >= 3 ? modes[2] : 1;
  for (IntType k2 = 0; k2 < n2; ++k2) {
    for (IntType k1 = 0; k1 < n1; ++k1) {
      IntType base = 0;
      if constexpr (DIM >= 2) base += k1 * strides[1];
      if constexpr (DIM >= 3) base += k2 * strides[2];
      for (IntType k0 = 0; k0 < modes[0]; ++k0) {
        off[i++] = static_cast<std::size_t>(base + k0 * strides[0]);
      }
    }
  }
  return off;
}

template <typename T, int DIM>
class HostPlanImpl {
public:
  HostPlanImpl(const neonufft::Options& opt, int sign, IntType num_nu, const T* const* loc,
               const IntType* modes)
      : alloc_(std::make_shared<neonufft::gpu::DeviceAllocator>()) {
    nm_ = 1;
    for (int d = 0; d < DIM; ++d) {
      if (modes[d] <= 0) throw neonufft::InputError("Neonufft:: modes must be positive");
      modes_[d] = modes[d];
      nm_ *= static_cast<std::size_t>(modes[d]);
    }
    coords_.upload(num_nu, loc, stream_.get(), alloc_);
    grid_modes_.reset(nm_, alloc_);
    host_modes_.resize(nm_);
    nu_.reset(std::max<std::size_t>(static_cast<std::size_t>(num_nu), 1), alloc_);
    m_ = static_cast<std::size_t>(num_nu);
    for (int d = 0; d < DIM; ++d) contig_[d] = d == 0 ? 1 : contig_[d - 1] * modes_[d - 1];
    gpu_ = std::make_unique<PlanImpl<T, DIM>>(opt, sign, num_nu, coords_.ptr, modes_,
                                              stream_.get(), alloc_);
  }

  void set_points(IntType num_nu, const T* const* loc) {
    coords_.upload(num_nu, loc, stream_.get(), alloc_);
    nu_.ensure(std::max<std::size_t>(static_cast<std::size_t>(num_nu), 1), alloc_);
    m_ = static_cast<std::size_t>(num_nu);
    gpu_->set_points(num_nu, coords_.ptr);
  }

  void transform_type_1(const std::complex<T>* in, std::complex<T>* out, const IntType* strides) {
    if (!out) throw neonufft::InputError("Neonufft:: null output");
    if (m_ && !in) throw neonufft::InputError("Neonufft:: null input");
    const hipStream_t s = stream_.get();
    to_device(nu_, reinterpret_cast<const Cplx<T>*>(in), m_, s);
    gpu_->transform_type_1(nu_.data(), grid_modes_.data(), contig_);
    to_host(host_modes_.data(), grid_modes_.data(), nm_, s);
    check_hip(hipStreamSynchronize(s), "hipStreamSynchronize");
    const std::vector<std::size_t> off = mode_offsets<DIM>(modes_, strides);
    for (std::size_t i = 0; i < nm_; ++i) {
      out[off[i]] = std::complex<T>(host_modes_[i].x, host_modes_[i].y);
    }
  }

  void transform_type_2(const std::complex<T>* in, const IntType* strides, std::complex<T>* out) {
    if (!in) throw neonufft::InputError("Neonufft:: null input");
    if (m_ && !out) throw neonufft::InputError("Neonufft:: null output");
    const std::vector<std::size_t> off = mode_offsets<DIM>(modes_, strides);
    for (std::size_t i = 0; i < nm_; ++i) {
      host_modes_[i] = Cplx<T>{in[off[i]].real(), in[off[i]].imag()};
    }
    const hipStream_t s = stream_.get();
    to_device(grid_modes_, host_modes_.data(), nm_, s);
    gpu_->transform_type_2(grid_modes_.data(), contig_, nu_.data());
    to_host(reinterpret_cast<Cplx<T>*>(out), nu_.data(), m_, s);
    check_hip(hipStreamSynchronize(s), "hipStreamSynchronize");
  }

private:
  // declared first: destroyed last, after the GPU plan that enqueues on it
  OwnedStream stream_;
  std::shared_ptr<neonufft::Allocator> alloc_;
  IntType modes_[3] = {1, 1, 1};
  IntType contig_[3] = {1, 1, 1};
  std::size_t nm_ = 1, m_ = 0;
  DeviceCoords<T, DIM> coords_;
  DeviceBuffer<Cplx<T>> nu_;          // strengths / values at the points
  DeviceBuffer<Cplx<T>> grid_modes_;  // modes, contiguous
  std::vector<Cplx<T>> host_modes_;   // modes, contiguous, before the strided copy
  std::unique_ptr<PlanImpl<T, DIM>> gpu_;
};

template <typename T, int DIM>
class HostPlanT3Impl {
public:
  HostPlanT3Impl(const neonufft::Options& opt, int sign, IntType num_in, const T* const* in_pts,
                 IntType num_out, const T* const* out_pts, IntType batch)
      : alloc_(std::make_shared<neonufft::gpu::DeviceAllocator>()), batch_(batch) {
    if (batch < 1) throw neonufft::InputError("Neonufft:: batch size must be >= 1");
    in_coords_.upload(num_in, in_pts, stream_.get(), alloc_);
    out_coords_.upload(num_out, out_pts, stream_.get(), alloc_);
    size_io(num_in, num_out);
    gpu_ = std::make_unique<PlanT3Impl<T, DIM>>(opt, sign, num_in, in_coords_.ptr, num_out,
                                                out_coords_.ptr, stream_.get(), batch, alloc_);
  }

  HostPlanT3Impl(const neonufft::Options& opt, int sign, const double* in_min,
                 const double* in_max, const double* out_min, const double* out_max,
                 IntType batch)
      : alloc_(std::make_shared<neonufft::gpu::DeviceAllocator>()), batch_(batch) {
    if (batch < 1) throw neonufft::InputError("Neonufft:: batch size must be >= 1");
    gpu_ = std::make_unique<PlanT3Impl<T, DIM>>(opt, sign, in_min, in_max, out_min, out_max,
                                                stream_.get(), batch, alloc_);
  }

  void set_input_points(IntType num_in, const T* const* pts) {
    in_coords_.upload(num_in, pts, stream_.get(), alloc_);
    size_io(num_in, static_cast<IntType>(n_out_));
    gpu_->set_input_points(num_in, in_coords_.ptr);
  }

  void set_output_points(IntType num_out, const T* const* pts) {
    out_coords_.upload(num_out, pts, stream_.get(), alloc_);
    size_io(static_cast<IntType>(n_in_), num_out);
    gpu_->set_output_points(num_out, out_coords_.ptr);
  }

  void add_input(const std::complex<T>* in, IntType bdist) {
    if (bdist <= 0) bdist = static_cast<IntType>(n_in_);
    if (n_in_ && !in) throw neonufft::InputError("Neonufft:: null input");
    const hipStream_t s = stream_.get();
    // batch members at `bdist` in the caller's array -> contiguous on the device
    for (IntType b = 0; b < batch_; ++b) {
      if (n_in_) {
        check_hip(hipMemcpyAsync(in_.data() + b * n_in_, in + b * bdist, n_in_ * sizeof(Cplx<T>),
                                 hipMemcpyHostToDevice, s),
                  "hipMemcpyAsync(host to device)");
      }
    }
    gpu_->add_input(in_.data(), static_cast<IntType>(n_in_));
    check_hip(hipStreamSynchronize(s), "hipStreamSynchronize");
  }

  void transform(std::complex<T>* out, IntType bdist) {
    if (bdist <= 0) bdist = static_cast<IntType>(n_out_);
    if (n_out_ && !out) throw neonufft::InputError("Neonufft:: null output");
    const hipStream_t s = stream_.get();
    gpu_->transform(out_.data(), static_cast<IntType>(n_out_));
    for (IntType b = 0; b < batch_; ++b) {
      if (n_out_) {
        check_hip(hipMemcpyAsync(out + b * bdist, out_.data() + b * n_out_,
                                 n_out_ * sizeof(Cplx<T>), hipMemcpyDeviceToHost, s),
                  "hipMemcpyAsync(device to host)");
      }
    }
    check_hip(hipStreamSynchronize(s), "hipStreamSynchronize");
  }

  void reset() {
    gpu_->reset();
    check_hip(hipStreamSynchronize(stream_.get()), "hipStreamSynchronize");
  }

private:
  void size_io(IntType num_in, IntType num_out) {
    n_in_ = static_cast<std::size_t>(num_in);
    n_out_ = static_cast<std::size_t>(num_out);
    const std::size_t b = static_cast<std::size_t>(batch_);
    in_.ensure(std::max<std::size_t>(b * n_in_, 1), alloc_);
    out_.ensure(std::max<std::size_t>(b * n_out_, 1), alloc_);
  }

  OwnedStream stream_;
  std::shared_ptr<neonufft::Allocator> alloc_;
  IntType batch_;
  std::size_t n_in_ = 0, n_out_ = 0;
  DeviceCoords<T, DIM> in_coords_, out_coords_;
  DeviceBuffer<Cplx<T>> in_, out_;  // batch x points, contiguous
  std::unique_ptr<PlanT3Impl<T, DIM>> gpu_;
};

}  // namespace

}  // namespace nnf

namespace neonufft {

template <typename T, std::size_t DIM>
Plan<T, DIM>::Plan(Options opt, int sign, IntType num_nu, std::array<const T*, DIM> loc,
                   std::array<IntType, DIM> modes)
    : impl_(new nnf::HostPlanImpl<T, DIM>(opt, sign, num_nu, loc.data(), modes.data()),
            [](void* p) { delete static_cast<nnf::HostPlanImpl<T, DIM>*>(p); }) {}

template <typename T, std::size_t DIM>
void Plan<T, DIM>::transform_type_1(const std::complex<T>* in, std::complex<T>* out,
                                    std::array<IntType, DIM> out_strides) {
  static_cast<nnf::HostPlanImpl<T, DIM>*>(impl_.get())
      ->transform_type_1(in, out, out_strides.data());
}

template <typename T, std::size_t DIM>
void Plan<T, DIM>::transform_type_2(const std::complex<T>* in, std::array<IntType, DIM> in_strides,
                                    std::complex<T>* out) {
  static_cast<nnf::HostPlanImpl<T, DIM>*>(impl_.get())
      ->transform_type_2(in, in_strides.data(), out);
}

template <typename T, std::size_t DIM>
void Plan<T, DIM>::set_points(IntType num_nu, std::array<const T*, DIM> loc) {
  static_cast<nnf::HostPlanImpl<T, DIM>*>(impl_.get())->set_points(num_nu, loc.data());
}

// sizeof(complex T) * (padded spread grid + FFT grid): the reference CPU
// plan's formula (src/neonufft/plan_t3_impl.hpp:45-64), whose spread grid
// carries ns / 2 + 2 cells of padding on each side
// (src/neonufft/util/grid_info_t3.hpp:47, src/neonufft/util/spread_padding.hpp:10).
template <typename T, std::size_t DIM>
std::uint64_t PlanT3<T, DIM>::grid_memory_size(const Options& opt, std::array<T, DIM> input_min,
                                               std::array<T, DIM> input_max,
                                               std::array<T, DIM> output_min,
                                               std::array<T, DIM> output_max) {
  double b[4][3] = {};
  for (std::size_t d = 0; d < DIM; ++d) {
    b[0][d] = input_min[d];
    b[1][d] = input_max[d];
    b[2][d] = output_min[d];
    b[3][d] = output_max[d];
  }
  int ns;
  double beta;
  nnf::kernel_width_and_beta(opt.tol, opt.upsampfac, &ns, &beta);
  const nnf::GridInfoT3 gi = nnf::make_grid_info_t3(static_cast<int>(DIM), ns, opt.upsampfac,
                                                    opt.recenter_threshold, b[0], b[1], b[2], b[3]);
  std::uint64_t sg = 1, fg = 1;
  for (std::size_t d = 0; d < DIM; ++d) {
    sg *= static_cast<std::uint64_t>(gi.spread_grid_size[d] + 2 * (ns / 2 + 2));
    fg *= static_cast<std::uint64_t>(gi.fft_grid_size[d]);
  }
  return sizeof(std::complex<T>) * (sg + fg);
}

template <typename T, std::size_t DIM>
PlanT3<T, DIM>::PlanT3(Options opt, int sign, IntType num_in,
                       std::array<const T*, DIM> input_points, IntType num_out,
                       std::array<const T*, DIM> output_points, IntType batch_size)
    : impl_(new nnf::HostPlanT3Impl<T, DIM>(opt, sign, num_in, input_points.data(), num_out,
                                             output_points.data(), batch_size),
            [](void* p) { delete static_cast<nnf::HostPlanT3Impl<T, DIM>*>(p); }) {}

template <typename T, std::size_t DIM>
PlanT3<T, DIM>::PlanT3(Options opt, int sign, std::array<T, DIM> input_min,
                       std::array<T, DIM> input_max, std::array<T, DIM> output_min,
                       std::array<T, DIM> output_max, IntType batch_size)
    : impl_(nullptr, [](void*) {}) {
  double b[4][3] = {};
  for (std::size_t d = 0; d < DIM; ++d) {
    b[0][d] = input_min[d];
    b[1][d] = input_max[d];
    b[2][d] = output_min[d];
    b[3][d] = output_max[d];
  }
  impl_ = decltype(impl_)(
      new nnf::HostPlanT3Impl<T, DIM>(opt, sign, b[0], b[1], b[2], b[3], batch_size),
      [](void* p) { delete static_cast<nnf::HostPlanT3Impl<T, DIM>*>(p); });
}

template <typename T, std::size_t DIM>
void PlanT3<T, DIM>::set_input_points(IntType num_in, std::array<const T*, DIM> input_points) {
  static_cast<nnf::HostPlanT3Impl<T, DIM>*>(impl_.get())
      ->set_input_points(num_in, input_points.data());
}

template <typename T, std::size_t DIM>
void PlanT3<T, DIM>::set_output_points(IntType num_out, std::array<const T*, DIM> output_points) {
  static_cast<nnf::HostPlanT3Impl<T, DIM>*>(impl_.get())
      ->set_output_points(num_out, output_points.data());
}

template <typename T, std::size_t DIM>
void PlanT3<T, DIM>::add_input(const std::complex<T>* in, IntType bdist) {
  static_cast<nnf::HostPlanT3Impl<T, DIM>*>(impl_.get())->add_input(in, bdist);
}

template <typename T, std::size_t DIM>
void PlanT3<T, DIM>::transform(std::complex<T>* out, IntType bdist) {
  static_cast<nnf::HostPlanT3Impl<T, DIM>*>(impl_.get())->transform(out, bdist);
}

template <typename T, std::size_t DIM>
void PlanT3<T, DIM>::reset() {
  static_cast<nnf::HostPlanT3Impl<T, DIM>*>(impl_.get())->reset();
}

template class Plan<float, 1>;
template class Plan<float, 2>;
template class Plan<float, 3>;
template class Plan<double, 1>;
template class Plan<double, 2>;
template class Plan<double, 3>;

template class PlanT3<float, 1>;
template class PlanT3<float, 2>;
template class PlanT3<float, 3>;
template class PlanT3<double, 1>;
template class PlanT3<double, 2>;
template class PlanT3<double, 3>;

}  // namespace neonufft
