// Autotuner for the spreading / interpolation launch shape (no reference
// counterpart: the reference fixes 8-cell bins, src/neonufft/gpu/util/
// partition_group.hpp:10-15, and one thread per point).
//
// For a plan's precision, dimension, kernel width, fine grid and points it
// times the spreading and/or interpolation kernels for a list of subgrid tile
// shapes and per-workgroup point caps, on the caller's own points, and keeps
// the fastest. 2D plans that spread with row bands time their type 1 for a
// list of band heights instead (the tile shape then only serves
// interpolation), and the two choices are returned together. Results are cached in-process per (precision, dim, ns, fine
// grid, log2 of the point count, kernel selection): the point distribution is
// not part of the key, so a cached answer assumes similar clustering.
#include <hip/hip_runtime.h>

#include <array>
#include <cmath>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include <neonufft/gpu/device_allocator.hpp>
#include <neonufft_amd.h>
#include "plan_impl.hpp"

namespace nnf {

namespace {

using Key = std::tuple<int, int, int, int64_t, int64_t, int64_t, int, int>;

std::mutex g_cache_mutex;
std::map<Key, std::pair<nnf_tuning, double>> g_cache;

struct Candidate {
  int bin[3];
  uint32_t max_points;
};

std::vector<Candidate> candidates(int dim) {
  std::vector<std::array<int, 3>> bins;
  if (dim == 1) {
    bins = {{1024, 1, 1}, {512, 1, 1}, {2048, 1, 1}, {256, 1, 1}};
  } else if (dim == 2) {
    bins = {{64, 32, 1}, {32, 32, 1}, {64, 16, 1}, {128, 16, 1}, {128, 32, 1}, {32, 16, 1}};
  } else {
    bins = {{32, 8, 8}, {16, 16, 8}, {16, 8, 8}, {32, 16, 4}, {24, 8, 8}, {32, 8, 4},
            {64, 8, 4}, {16, 16, 4}};
  }
  std::vector<Candidate> out;
  for (uint32_t mp : {2048u, 1024u}) {
    for (const auto& b : bins) out.push_back(Candidate{{b[0], b[1], b[2]}, mp});
  }
  return out;
}

class EventPair {
public:
  EventPair() {
    check_hip(hipEventCreate(&a_), "hipEventCreate");
    check_hip(hipEventCreate(&b_), "hipEventCreate");
  }
  ~EventPair() {
    (void)hipEventDestroy(a_);
    (void)hipEventDestroy(b_);
  }
  template <typename F>
  double time_ms(hipStream_t s, int reps, F&& f) {
    f();  // warm-up
    check_hip(hipEventRecord(a_, s), "hipEventRecord");
    for (int r = 0; r < reps; ++r) f();
    check_hip(hipEventRecord(b_, s), "hipEventRecord");
    check_hip(hipEventSynchronize(b_), "hipEventSynchronize");
    float ms = 0.0f;
    check_hip(hipEventElapsedTime(&ms, a_, b_), "hipEventElapsedTime");
    return static_cast<double>(ms) / reps;
  }

private:
  hipEvent_t a_ = nullptr, b_ = nullptr;
};

template <typename T, int DIM>
std::pair<nnf_tuning, double> tune(const neonufft::Options& opt, int sign, IntType num_nu,
                                   const T* const* loc, const IntType* modes, hipStream_t stream,
                                   int types, int reps) {
  auto alloc = std::make_shared<neonufft::gpu::DeviceAllocator>();
  const std::size_t m = static_cast<std::size_t>(num_nu);
  DeviceBuffer<Cplx<T>> strengths(m ? m : 1, alloc);
  DeviceBuffer<Cplx<T>> values(m ? m : 1, alloc);
  // any non-zero strengths: the flush skips all-zero tile cells
  check_hip(hipMemsetD32Async(reinterpret_cast<hipDeviceptr_t>(strengths.data()),
                              static_cast<int>(0x3f800000u), strengths.bytes() / 4, stream),
            "hipMemsetD32Async");
  EventPair ev;
  nnf_tuning best{};
  double best_ms = -1.0;
  std::vector<std::array<int64_t, 4>> seen;
  bool bands = false;  // type 1 spreads with row bands: tuned separately below
  for (const Candidate& c : candidates(DIM)) {
    PlanTuning t;
    for (int d = 0; d < 3; ++d) t.bin[d] = d < DIM ? c.bin[d] : 0;
    t.max_points = c.max_points;
    PlanImpl<T, DIM> plan(opt, sign, num_nu, loc, modes, stream, alloc, t);
    nnf_plan_info info;
    plan.info(&info);
    // the LDS budget may shrink a candidate onto one already timed
    const std::array<int64_t, 4> key{info.bin_size[0], info.bin_size[1], info.bin_size[2],
                                     static_cast<int64_t>(c.max_points)};
    bool dup = false;
    for (const auto& s : seen) dup = dup || s == key;
    if (dup) continue;
    seen.push_back(key);
    bands = info.spread_mode != kSpreadTiles;
    double ms = 0.0;
    if ((types & 1) && !bands) {
      plan.spread(strengths.data(), false);
      ms += ev.time_ms(stream, reps, [&] { plan.spread(strengths.data(), true); });
    }
    if (types & 2) ms += ev.time_ms(stream, reps, [&] { plan.interp(values.data()); });
    if (best_ms < 0.0 || ms < best_ms) {
      best_ms = ms;
      best = nnf_tuning{};
      for (int d = 0; d < 3; ++d)
        best.bin_size[d] = d < DIM ? static_cast<int32_t>(info.bin_size[d]) : 0;
      best.max_points = static_cast<int32_t>(c.max_points);
    }
    if (!(types & 2) && bands) break;  // nothing a tile shape changes is timed
  }
  if constexpr (DIM == 2) {
    if ((types & 1) && bands) {
      // band heights: the default, then powers of two up to the LDS budget
      // (heights the plan clamps onto an already timed one are skipped)
      std::size_t nm = 1;
      for (int d = 0; d < DIM; ++d) nm *= static_cast<std::size_t>(modes[d]);
      DeviceBuffer<Cplx<T>> out(nm, alloc);
      const IntType st[3] = {1, modes[0], 1};
      double band_ms = -1.0;
      int band_rows = 0;
      std::vector<int> timed;
      for (int rows : {0, 1, 2, 4, 8, 16, 32}) {
        PlanTuning t;
        t.spread_mode = kSpreadBands;
        t.band_rows = rows;
        try {
          PlanImpl<T, DIM> plan(opt, sign, num_nu, loc, modes, stream, alloc, t);
          nnf_plan_info info;
          plan.info(&info);
          if (info.spread_mode == kSpreadTiles) break;  // the points do not suit bands
          bool dup = false;
          for (int r : timed) dup = dup || r == info.band_rows;
          if (dup) continue;
          timed.push_back(info.band_rows);
          const double ms = ev.time_ms(stream, reps, [&] {
            plan.transform_type_1(strengths.data(), out.data(), st);
          });
          if (band_ms < 0.0 || ms < band_ms) {
            band_ms = ms;
            band_rows = info.band_rows;
          }
        } catch (const neonufft::InputError&) {
          break;
        }
      }
      if (band_ms >= 0.0) {
        best.spread_mode = NNF_SPREAD_AUTO;
        best.band_rows = band_rows;
        best_ms = (best_ms < 0.0 ? 0.0 : best_ms) + band_ms;
      }
    }
  }
  return {best, best_ms};
}

}  // namespace

// Entry used by the C ABI (c_api.cpp).
void autotune(int precision, int dim, const neonufft::Options& opt, int sign, IntType num_nu,
              const void* const* loc, const IntType* modes, hipStream_t stream, int types,
              int reps, bool use_cache, nnf_tuning* best, double* best_ms) {
  if (types < 1 || types > 3) throw neonufft::InputError("Neonufft:: types must be 1, 2 or 3");
  if (reps < 1) throw neonufft::InputError("Neonufft:: reps must be positive");
  if (num_nu < 0) throw neonufft::InputError("Neonufft:: negative number of points");
  if (precision != NNF_F32 && precision != NNF_F64)
    throw neonufft::InputError("Neonufft:: precision must be NNF_F32 or NNF_F64");
  if (dim < 1 || dim > 3) throw neonufft::InputError("Neonufft:: dim must be 1, 2 or 3");
  const KernelSetup ks = make_kernel_setup(opt.tol, opt.upsampfac, precision == NNF_F32);
  int64_t n[3] = {1, 1, 1};
  for (int d = 0; d < dim; ++d) {
    if (modes[d] <= 0) throw neonufft::InputError("Neonufft:: modes must be positive");
    n[d] = std::max<IntType>(2 * ks.ns,
                             static_cast<IntType>(static_cast<double>(modes[d]) * opt.upsampfac));
  }
  const int mbucket = num_nu > 0 ? static_cast<int>(std::log2(static_cast<double>(num_nu))) : -1;
  const Key key{precision, dim, ks.ns, n[0], n[1], n[2], mbucket, types};
  if (use_cache) {
    std::lock_guard<std::mutex> lock(g_cache_mutex);
    auto it = g_cache.find(key);
    if (it != g_cache.end()) {
      *best = it->second.first;
      *best_ms = it->second.second;
      return;
    }
  }
  std::pair<nnf_tuning, double> r;
  if (precision == NNF_F32) {
    const auto* l = reinterpret_cast<const float* const*>(loc);
    if (dim == 1) r = tune<float, 1>(opt, sign, num_nu, l, modes, stream, types, reps);
    if (dim == 2) r = tune<float, 2>(opt, sign, num_nu, l, modes, stream, types, reps);
    if (dim == 3) r = tune<float, 3>(opt, sign, num_nu, l, modes, stream, types, reps);
  } else {
    const auto* l = reinterpret_cast<const double* const*>(loc);
    if (dim == 1) r = tune<double, 1>(opt, sign, num_nu, l, modes, stream, types, reps);
    if (dim == 2) r = tune<double, 2>(opt, sign, num_nu, l, modes, stream, types, reps);
    if (dim == 3) r = tune<double, 3>(opt, sign, num_nu, l, modes, stream, types, reps);
  }
  *best = r.first;
  *best_ms = r.second;
  std::lock_guard<std::mutex> lock(g_cache_mutex);
  g_cache[key] = r;
}

void autotune_clear_cache() {
  std::lock_guard<std::mutex> lock(g_cache_mutex);
  g_cache.clear();
}

}  // namespace nnf
