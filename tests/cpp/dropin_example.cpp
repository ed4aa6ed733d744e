// Written against the reference's public GPU API (include/neonufft/gpu/plan.hpp):
// it must compile and link unchanged against this library.
#include <neonufft/gpu/plan.hpp>
#include <neonufft/exceptions.hpp>

#include <cstdio>

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
