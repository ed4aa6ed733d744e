#pragma once

#include <neonufft/config.h>
#include <neonufft/enums.h>
#include <neonufft/types.hpp>

#include <array>
#include <complex>
#include <cstddef>
#include <cstdint>
#include <memory>
#include <type_traits>

namespace neonufft {

/* Plan options, same members, order and defaults as the reference
 * (include/neonufft/plan.hpp:19-29). Passed by value, so the layout is ABI.
 * The GPU plans read tol, upsampfac, recenter_threshold and order; the other
 * members are accepted and ignored exactly as the reference GPU plan does. */
struct NEONUFFT_EXPORT Options {
  double tol = 0.001;
  double upsampfac = 2.0;
  double recenter_threshold = 0.1;
  int num_threads = 0;
  bool sort_input = true;
  bool sort_output = true;
  bool kernel_approximation = false;
  NeonufftModeOrder order = NEONUFFT_MODE_ORDER_CMCL;
  NeonufftKernelType kernel_type = NEONUFFT_ES_KERNEL;
};

/* Host-memory plans: the reference's neonufft::Plan / PlanT3
 * (include/neonufft/plan.hpp:31-85), same classes and signatures. Every
 * pointer is a HOST pointer, as in the reference; the work runs on the
 * current GPU through the GPU plans of <neonufft/gpu/plan.hpp>: coordinates
 * are copied to the device at construction / set_points, each call copies its
 * input to the device, transforms there, copies the result back and returns
 * when it is in the caller's array. Strided mode arrays are gathered /
 * scattered on the host, so elements outside the strided positions are never
 * touched. Options::num_threads is accepted and ignored (no CPU worker
 * threads). grid_memory_size follows the reference CPU plan's formula. */
template <typename T, std::size_t DIM>
class NEONUFFT_EXPORT Plan {
public:
  static_assert(std::is_same_v<T, float> || std::is_same_v<T, double>);
  static_assert(DIM == 1 || DIM == 2 || DIM == 3);

  Plan(Options opt, int sign, IntType num_nu, std::array<const T*, DIM> loc,
       std::array<IntType, DIM> modes);

  void transform_type_1(const std::complex<T>* in, std::complex<T>* out,
                        std::array<IntType, DIM> out_strides);

  void transform_type_2(const std::complex<T>* in, std::array<IntType, DIM> in_strides,
                        std::complex<T>* out);

  void set_points(IntType num_nu, std::array<const T*, DIM> loc);

private:
  std::unique_ptr<void, void (*)(void*)> impl_;
};

template <typename T, std::size_t DIM>
class NEONUFFT_EXPORT PlanT3 {
public:
  static_assert(std::is_same_v<T, float> || std::is_same_v<T, double>);
  static_assert(DIM == 1 || DIM == 2 || DIM == 3);

  static std::uint64_t grid_memory_size(const Options& opt, std::array<T, DIM> input_min,
                                        std::array<T, DIM> input_max,
                                        std::array<T, DIM> output_min,
                                        std::array<T, DIM> output_max);

  PlanT3(Options opt, int sign, IntType num_in, std::array<const T*, DIM> input_points,
         IntType num_out, std::array<const T*, DIM> output_points, IntType batch_size = 1);

  PlanT3(Options opt, int sign, std::array<T, DIM> input_min, std::array<T, DIM> input_max,
         std::array<T, DIM> output_min, std::array<T, DIM> output_max, IntType batch_size = 1);

  void set_input_points(IntType num_in, std::array<const T*, DIM> input_points);

  void set_output_points(IntType num_out, std::array<const T*, DIM> output_points);

  void add_input(const std::complex<T>* in, IntType bdist = 0);

  void transform(std::complex<T>* out, IntType bdist = 0);

  void reset();

private:
  std::unique_ptr<void, void (*)(void*)> impl_;
};

}  // namespace neonufft
