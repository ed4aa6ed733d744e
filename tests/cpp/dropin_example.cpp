// A caller written against the reference's public GPU API
// (include/neonufft/gpu/plan.hpp:25-82 of the reference): it must compile and
// link unchanged against this library. Built two ways:
//  - as a program (tests/test_abi.py): without arguments it only prints its
//    usage, so the link check needs no GPU;
//  - as libdropin.so (tests/cpp/Makefile): the extern "C" hooks below drive
//    the C++ classes on caller device buffers, and tests/test_gpu_dropin.py
//    calls them through ctypes and checks the results against the oracle.
#include <neonufft/exceptions.hpp>
#include <neonufft/gpu/plan.hpp>
#include <neonufft/plan.hpp>

#include <hip/hip_runtime_api.h>

#include <complex>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <exception>

namespace {

int fail(const char* what, int code, char* msg, int len) {
  if (msg && len > 0) {
    std::snprintf(msg, static_cast<std::size_t>(len), "%s", what);
  }
  return code;
}

template <typename F>
int guarded(F&& f, char* msg, int len) {
  try {
    f();
    if (hipDeviceSynchronize() != hipSuccess) return fail("hipDeviceSynchronize failed", 6, msg, len);
    return 0;
  } catch (const neonufft::GenericError& e) {
    return fail(e.what(), static_cast<int>(e.error_code()), msg, len);
  } catch (const std::exception& e) {
    return fail(e.what(), -1, msg, len);
  }
}

}  // namespace

extern "C" {

// neonufft::gpu::Plan<float, 3>: construct on points A, type 1 of c1 into
// f1, type 2 of f_in into c_out; then set_points to B and type 1 of c2
// into f2. Mode arrays contiguous (strides 1, n0, n0 n1).
int dropin_plan_f32_3d(int64_t ma, const float* const* xa, int64_t mb, const float* const* xb,
                       const int64_t* modes, double tol, int sign, const void* c1, void* f1,
                       const void* f_in, void* c_out, const void* c2, void* f2, char* msg,
                       int msglen) {
  return guarded(
      [&] {
        neonufft::Options opt;
        opt.tol = tol;
        const std::array<neonufft::IntType, 3> n{modes[0], modes[1], modes[2]};
        const std::array<neonufft::IntType, 3> st{1, modes[0], modes[0] * modes[1]};
        neonufft::gpu::Plan<float, 3> plan(opt, sign, ma, {xa[0], xa[1], xa[2]}, n, nullptr);
        plan.transform_type_1(static_cast<const hipComplex*>(c1), static_cast<hipComplex*>(f1),
                              st);
        plan.transform_type_2(static_cast<const hipComplex*>(f_in), st,
                              static_cast<hipComplex*>(c_out));
        plan.set_points(mb, {xb[0], xb[1], xb[2]});
        plan.transform_type_1(static_cast<const hipComplex*>(c2), static_cast<hipComplex*>(f2),
                              st);
      },
      msg, msglen);
}

// neonufft::gpu::PlanT3<double, 3> from bounds (bounds = in_min[3],
// in_max[3], out_min[3], out_max[3]): sources A and targets S, two add_input
// calls (c1, c2) and transform into out1; then reset, sources B, one
// add_input (c3) and transform into out2. Also reports grid_memory_size.
int dropin_t3_f64_3d(int64_t ma, const double* const* xa, int64_t mb, const double* const* xb,
                     int64_t n, const double* const* s, const double* bounds, double tol,
                     int sign, const void* c1, const void* c2, const void* c3, void* out1,
                     void* out2, uint64_t* grid_bytes, char* msg, int msglen) {
  return guarded(
      [&] {
        neonufft::Options opt;
        opt.tol = tol;
        const std::array<double, 3> in_min{bounds[0], bounds[1], bounds[2]};
        const std::array<double, 3> in_max{bounds[3], bounds[4], bounds[5]};
        const std::array<double, 3> out_min{bounds[6], bounds[7], bounds[8]};
        const std::array<double, 3> out_max{bounds[9], bounds[10], bounds[11]};
        *grid_bytes =
            neonufft::gpu::PlanT3<double, 3>::grid_memory_size(opt, in_min, in_max, out_min, out_max);
        neonufft::gpu::PlanT3<double, 3> plan(opt, sign, in_min, in_max, out_min, out_max, nullptr);
        plan.set_input_points(ma, {xa[0], xa[1], xa[2]});
        plan.set_output_points(n, {s[0], s[1], s[2]});
        plan.add_input(static_cast<const hipDoubleComplex*>(c1));
        plan.add_input(static_cast<const hipDoubleComplex*>(c2));
        plan.transform(static_cast<hipDoubleComplex*>(out1));
        plan.reset();
        plan.set_input_points(mb, {xb[0], xb[1], xb[2]});
        plan.add_input(static_cast<const hipDoubleComplex*>(c3));
        plan.transform(static_cast<hipDoubleComplex*>(out2));
      },
      msg, msglen);
}

// neonufft::gpu::PlanT3<float, 2> from points, batch of 2 with bdist =
// m + 1 (inputs) and n + 1 (outputs).
int dropin_t3_f32_2d(int64_t m, const float* const* x, int64_t n, const float* const* s,
                     double tol, int sign, const void* c, void* out, char* msg, int msglen) {
  return guarded(
      [&] {
        neonufft::Options opt;
        opt.tol = tol;
        neonufft::gpu::PlanT3<float, 2> plan(opt, sign, m, {x[0], x[1]}, n, {s[0], s[1]}, nullptr,
                                             2);
        plan.add_input(static_cast<const hipComplex*>(c), m + 1);
        plan.transform(static_cast<hipComplex*>(out), n + 1);
      },
      msg, msglen);
}

// Host-memory neonufft::Plan<double, 2> (include/neonufft/plan.hpp:31-52 of
// the reference): every pointer a host pointer. Type 1 of c into f1 written
// with out_strides st1, type 2 of f_in read with in_strides st2 into c_out;
// then set_points to B and type 1 of c2 into f2 (contiguous).
int dropin_host_plan_f64_2d(int64_t ma, const double* const* xa, int64_t mb,
                            const double* const* xb, const int64_t* modes, double tol, int sign,
                            const void* c1, void* f1, const int64_t* st1, const void* f_in,
                            const int64_t* st2, void* c_out, const void* c2, void* f2,
                            char* msg, int msglen) {
  return guarded(
      [&] {
        neonufft::Options opt;
        opt.tol = tol;
        using C = std::complex<double>;
        neonufft::Plan<double, 2> plan(opt, sign, ma, {xa[0], xa[1]}, {modes[0], modes[1]});
        plan.transform_type_1(static_cast<const C*>(c1), static_cast<C*>(f1), {st1[0], st1[1]});
        plan.transform_type_2(static_cast<const C*>(f_in), {st2[0], st2[1]},
                              static_cast<C*>(c_out));
        plan.set_points(mb, {xb[0], xb[1]});
        plan.transform_type_1(static_cast<const C*>(c2), static_cast<C*>(f2), {1, modes[0]});
      },
      msg, msglen);
}

// Host-memory neonufft::PlanT3<float, 3> from bounds, batch of 2 with bdist:
// two add_input calls (c1, c2; inputs at in_bdist) then transform into out1
// (at out_bdist); reset, one add_input (c3), transform into out2. Also the
// reference CPU plan's grid_memory_size for the same bounds.
int dropin_host_t3_f32_3d(int64_t m, const float* const* x, int64_t n, const float* const* s,
                          const float* bounds, double tol, int sign, const void* c1,
                          const void* c2, const void* c3, int64_t in_bdist, int64_t out_bdist,
                          void* out1, void* out2, uint64_t* grid_bytes, char* msg, int msglen) {
  return guarded(
      [&] {
        neonufft::Options opt;
        opt.tol = tol;
        using C = std::complex<float>;
        const std::array<float, 3> in_min{bounds[0], bounds[1], bounds[2]};
        const std::array<float, 3> in_max{bounds[3], bounds[4], bounds[5]};
        const std::array<float, 3> out_min{bounds[6], bounds[7], bounds[8]};
        const std::array<float, 3> out_max{bounds[9], bounds[10], bounds[11]};
        *grid_bytes =
            neonufft::PlanT3<float, 3>::grid_memory_size(opt, in_min, in_max, out_min, out_max);
        neonufft::PlanT3<float, 3> plan(opt, sign, in_min, in_max, out_min, out_max, 2);
        plan.set_input_points(m, {x[0], x[1], x[2]});
        plan.set_output_points(n, {s[0], s[1], s[2]});
        plan.add_input(static_cast<const C*>(c1), in_bdist);
        plan.add_input(static_cast<const C*>(c2), in_bdist);
        plan.transform(static_cast<C*>(out1), out_bdist);
        plan.reset();
        plan.add_input(static_cast<const C*>(c3), in_bdist);
        plan.transform(static_cast<C*>(out2), out_bdist);
      },
      msg, msglen);
}

// An invalid call must surface as the reference's exception type: a
// negative point count is an InputError (NEONUFFT_INPUT_ERROR).
int dropin_input_error(char* msg, int msglen) {
  return guarded(
      [&] {
        neonufft::Options opt;
        neonufft::gpu::Plan<double, 1> plan(opt, 1, -5, {nullptr}, {8}, nullptr);
      },
      msg, msglen);
}

}  // extern "C"

int main(int argc, char** argv) {
  if (argc < 2) {
    std::printf("usage: dropin_example run\n");
    return 0;
  }
  neonufft::Options opt;
  opt.tol = 1e-5;
  const float* loc[3] = {nullptr, nullptr, nullptr};
  try {
    neonufft::gpu::Plan<float, 3> plan(opt, -1, 0, {loc[0], loc[1], loc[2]}, {16, 16, 16}, nullptr);
    plan.transform_type_1(nullptr, nullptr, {1, 16, 256});
    neonufft::gpu::PlanT3<double, 2> t3(opt, 1, std::array<double, 2>{0, 0},
                                        std::array<double, 2>{1, 1}, std::array<double, 2>{0, 0},
                                        std::array<double, 2>{1, 1}, nullptr);
    t3.reset();
    std::printf("%llu\n", (unsigned long long)neonufft::gpu::PlanT3<float, 1>::grid_memory_size(
                              opt, {-1.f}, {1.f}, {-2.f}, {2.f}));
  } catch (const neonufft::GenericError& e) {
    std::printf("error %d: %s\n", (int)e.error_code(), e.what());
    return 1;
  }
  return 0;
}
