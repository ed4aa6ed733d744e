// The drop-in C++ classes neonufft::gpu::Plan<T,DIM> and PlanT3<T,DIM>
// (declared in include/neonufft/gpu/plan.hpp, same interface as the
// reference's include/neonufft/gpu/plan.hpp:25-82) and DeviceAllocator.
// They forward to the shared plan implementation; the explicit instantiations
// are the 12 the reference exports (src/neonufft/gpu/plan.cpp:47-53,122-128).
#include <neonufft/gpu/device_allocator.hpp>
#include <neonufft/gpu/plan.hpp>

#include "plan_impl.hpp"

namespace nnf {

namespace {
thread_local std::string g_error_detail;
}

void set_error_detail(const std::string& msg) { g_error_detail = msg; }
const std::string& error_detail() { return g_error_detail; }

void throw_hip(hipError_t err, const char* what) {
  set_error_detail(std::string(what) + ": " + hipGetErrorName(err) + " (" +
                   hipGetErrorString(err) + ")");
  throw neonufft::GPUError("Neonufft:: HIP runtime error");
}

}  // namespace nnf

namespace neonufft {
namespace gpu {

void* DeviceAllocator::allocate(std::size_t size) {
  void* ptr = nullptr;
  if (size) nnf::check_hip(hipMalloc(&ptr, size), "hipMalloc");
  return ptr;
}

void DeviceAllocator::deallocate(void* ptr) noexcept {
  if (ptr) (void)hipFree(ptr);
}

template <typename T>
using CplxT = nnf::Cplx<T>;

template <typename T, std::size_t DIM>
Plan<T, DIM>::Plan(Options opt, int sign, IntType num_nu, std::array<const T*, DIM> loc,
                   std::array<IntType, DIM> modes, StreamType stream,
                   std::shared_ptr<Allocator> device_alloc)
    : impl_(new nnf::PlanImpl<T, DIM>(opt, sign, num_nu, loc.data(), modes.data(), stream,
                                       std::move(device_alloc)),
            [](void* p) { delete static_cast<nnf::PlanImpl<T, DIM>*>(p); }) {}

template <typename T, std::size_t DIM>
void Plan<T, DIM>::transform_type_1(const ComplexType<T>* in, ComplexType<T>* out,
                                    std::array<IntType, DIM> out_strides) {
  static_cast<nnf::PlanImpl<T, DIM>*>(impl_.get())
      ->transform_type_1(reinterpret_cast<const CplxT<T>*>(in), reinterpret_cast<CplxT<T>*>(out),
                         out_strides.data());
}

template <typename T, std::size_t DIM>
void Plan<T, DIM>::transform_type_2(const ComplexType<T>* in, std::array<IntType, DIM> in_strides,
                                    ComplexType<T>* out) {
  static_cast<nnf::PlanImpl<T, DIM>*>(impl_.get())
      ->transform_type_2(reinterpret_cast<const CplxT<T>*>(in), in_strides.data(),
                         reinterpret_cast<CplxT<T>*>(out));
}

template <typename T, std::size_t DIM>
void Plan<T, DIM>::set_points(IntType num_nu, std::array<const T*, DIM> loc) {
  static_cast<nnf::PlanImpl<T, DIM>*>(impl_.get())->set_points(num_nu, loc.data());
}

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
  return nnf::PlanT3Impl<T, DIM>::grid_memory_size(opt, b[0], b[1], b[2], b[3]);
}

template <typename T, std::size_t DIM>
PlanT3<T, DIM>::PlanT3(Options opt, int sign, IntType num_in,
                       std::array<const T*, DIM> input_points, IntType num_out,
                       std::array<const T*, DIM> output_points, StreamType stream,
                       IntType batch_size, std::shared_ptr<Allocator> device_alloc)
    : impl_(new nnf::PlanT3Impl<T, DIM>(opt, sign, num_in, input_points.data(), num_out,
                                         output_points.data(), stream, batch_size,
                                         std::move(device_alloc)),
            [](void* p) { delete static_cast<nnf::PlanT3Impl<T, DIM>*>(p); }) {}

template <typename T, std::size_t DIM>
PlanT3<T, DIM>::PlanT3(Options opt, int sign, std::array<T, DIM> input_min,
                       std::array<T, DIM> input_max, std::array<T, DIM> output_min,
                       std::array<T, DIM> output_max, StreamType stream, IntType batch_size,
                       std::shared_ptr<Allocator> device_alloc)
    : impl_(nullptr, [](void*) {}) {
  double b[4][3] = {};
  for (std::size_t d = 0; d < DIM; ++d) {
    b[0][d] = input_min[d];
    b[1][d] = input_max[d];
    b[2][d] = output_min[d];
    b[3][d] = output_max[d];
  }
  impl_ = decltype(impl_)(new nnf::PlanT3Impl<T, DIM>(opt, sign, b[0], b[1], b[2], b[3], stream,
                                                       batch_size, std::move(device_alloc)),
                          [](void* p) { delete static_cast<nnf::PlanT3Impl<T, DIM>*>(p); });
}

template <typename T, std::size_t DIM>
void PlanT3<T, DIM>::set_input_points(IntType num_in, std::array<const T*, DIM> input_points) {
  static_cast<nnf::PlanT3Impl<T, DIM>*>(impl_.get())->set_input_points(num_in, input_points.data());
}

template <typename T, std::size_t DIM>
void PlanT3<T, DIM>::set_output_points(IntType num_out, std::array<const T*, DIM> output_points) {
  static_cast<nnf::PlanT3Impl<T, DIM>*>(impl_.get())
      ->set_output_points(num_out, output_points.data());
}

template <typename T, std::size_t DIM>
void PlanT3<T, DIM>::add_input(const ComplexType<T>* in, IntType bdist) {
  static_cast<nnf::PlanT3Impl<T, DIM>*>(impl_.get())
      ->add_input(reinterpret_cast<const CplxT<T>*>(in), bdist);
}

template <typename T, std::size_t DIM>
void PlanT3<T, DIM>::transform(ComplexType<T>* out, IntType bdist) {
  static_cast<nnf::PlanT3Impl<T, DIM>*>(impl_.get())
      ->transform(reinterpret_cast<CplxT<T>*>(out), bdist);
}

template <typename T, std::size_t DIM>
void PlanT3<T, DIM>::reset() {
  static_cast<nnf::PlanT3Impl<T, DIM>*>(impl_.get())->reset();
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

}  // namespace gpu
}  // namespace neonufft
