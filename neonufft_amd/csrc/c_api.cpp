// extern "C" entry points (include/neonufft_amd.h). Each one forwards to the
// same plan implementation as the C++ classes and maps exceptions to the
// reference's NeonufftError codes.
#include <neonufft/gpu/device_allocator.hpp>

#include <cstring>
#include <memory>
#include <new>

#include <neonufft_amd.h>
#include "batch.hpp"

namespace nnf {
void autotune(int precision, int dim, const neonufft::Options& opt, int sign, IntType num_nu,
              const void* const* loc, const IntType* modes, hipStream_t stream, int types,
              int reps, bool use_cache, nnf_tuning* best, double* best_ms);
void autotune_clear_cache();
}  // namespace nnf
#include "plan_impl.hpp"

namespace {

using nnf::Cplx;
using neonufft::IntType;

thread_local std::string g_last_error;

struct PlanIface {
  virtual ~PlanIface() = default;
  virtual void set_points(int64_t n, const void* const* loc) = 0;
  virtual void set_points_noalloc(int64_t n, const void* const* loc) = 0;
  virtual void t1(const void* in, void* out, const int64_t* st) = 0;
  virtual void t2(const void* in, const int64_t* st, void* out) = 0;
  virtual void spread(const void* in, bool accumulate) = 0;
  virtual void spread_tiles(const void* in) = 0;
  virtual void reduce_tiles(bool accumulate) = 0;
  virtual void interp(void* out) = 0;
  virtual void* grid() = 0;
  virtual const void* fseries(int d) = 0;
  virtual void info(nnf_plan_info* i) = 0;
};

template <typename T, int DIM>
struct PlanHolder final : PlanIface {
  nnf::PlanImpl<T, DIM> impl;
  PlanHolder(const neonufft::Options& o, int sign, IntType m, const void* const* loc,
             const int64_t* modes, hipStream_t s, std::shared_ptr<neonufft::Allocator> a,
             const nnf::PlanTuning& t)
      : impl(o, sign, m, reinterpret_cast<const T* const*>(loc), modes, s, std::move(a), t) {}
  void set_points(int64_t n, const void* const* loc) override {
    impl.set_points(n, reinterpret_cast<const T* const*>(loc));
  }
  void set_points_noalloc(int64_t n, const void* const* loc) override {
    impl.set_points_noalloc(n, reinterpret_cast<const T* const*>(loc));
  }
  void t1(const void* in, void* out, const int64_t* st) override {
    impl.transform_type_1(static_cast<const Cplx<T>*>(in), static_cast<Cplx<T>*>(out), st);
  }
  void t2(const void* in, const int64_t* st, void* out) override {
    impl.transform_type_2(static_cast<const Cplx<T>*>(in), st, static_cast<Cplx<T>*>(out));
  }
  void spread(const void* in, bool accumulate) override {
    impl.spread(static_cast<const Cplx<T>*>(in), accumulate);
  }
  void spread_tiles(const void* in) override {
    impl.spread_tiles(static_cast<const Cplx<T>*>(in));
  }
  void reduce_tiles(bool accumulate) override { impl.reduce_tiles(accumulate); }
  void interp(void* out) override { impl.interp(static_cast<Cplx<T>*>(out)); }
  void* grid() override { return impl.fine_grid(); }
  const void* fseries(int d) override { return impl.fseries(d); }
  void info(nnf_plan_info* i) override { impl.info(i); }
};

struct PlanT3Iface {
  virtual ~PlanT3Iface() = default;
  virtual void set_in(int64_t n, const void* const* p) = 0;
  virtual void set_out(int64_t n, const void* const* p) = 0;
  virtual void add_input(const void* in, int64_t bdist) = 0;
  virtual void transform(void* out, int64_t bdist) = 0;
  virtual void reset() = 0;
  virtual void info(nnf_plan_info* i) = 0;
};

// TC: coordinate and I/O precision (differs from T for NNF_MIXED).
template <typename T, int DIM, typename TC = T>
struct PlanT3Holder final : PlanT3Iface {
  nnf::PlanT3Impl<T, DIM, TC> impl;
  PlanT3Holder(const neonufft::Options& o, int sign, IntType num_in, const void* const* in_pts,
               IntType num_out, const void* const* out_pts, hipStream_t s, IntType batch,
               std::shared_ptr<neonufft::Allocator> a)
      : impl(o, sign, num_in, reinterpret_cast<const TC* const*>(in_pts), num_out,
             reinterpret_cast<const TC* const*>(out_pts), s, batch, std::move(a)) {}
  PlanT3Holder(const neonufft::Options& o, int sign, const double* in_min, const double* in_max,
               const double* out_min, const double* out_max, hipStream_t s, IntType batch,
               std::shared_ptr<neonufft::Allocator> a)
      : impl(o, sign, in_min, in_max, out_min, out_max, s, batch, std::move(a)) {}
  void set_in(int64_t n, const void* const* p) override {
    impl.set_input_points(n, reinterpret_cast<const TC* const*>(p));
  }
  void set_out(int64_t n, const void* const* p) override {
    impl.set_output_points(n, reinterpret_cast<const TC* const*>(p));
  }
  void add_input(const void* in, int64_t bdist) override {
    impl.add_input(static_cast<const Cplx<TC>*>(in), bdist);
  }
  void transform(void* out, int64_t bdist) override {
    impl.transform(static_cast<Cplx<TC>*>(out), bdist);
  }
  void reset() override { impl.reset(); }
  void info(nnf_plan_info* i) override { impl.info(i); }
};

struct BatchIface {
  virtual ~BatchIface() = default;
  virtual void set_points(const int64_t* n, const void* const* loc) = 0;
  virtual void t1(const void* const* in, void* const* out) = 0;
  virtual void t2(const void* const* in, void* const* out) = 0;
  virtual void info(nnf_batch_info* i) = 0;
};

template <typename T, int DIM>
struct BatchHolder final : BatchIface {
  nnf::BatchPlanImpl<T, DIM> impl;
  BatchHolder(const neonufft::Options& o, int sign, IntType count, const int64_t* num_nu,
              const void* const* loc, const int64_t* modes, hipStream_t s,
              std::shared_ptr<neonufft::Allocator> a, bool graph, const nnf::PlanTuning& t)
      : impl(o, sign, count, num_nu, reinterpret_cast<const T* const*>(loc), modes, s,
             std::move(a), graph, t) {}
  void set_points(const int64_t* n, const void* const* loc) override {
    impl.set_points(n, reinterpret_cast<const T* const*>(loc));
  }
  void t1(const void* const* in, void* const* out) override {
    impl.transform_type_1(reinterpret_cast<const Cplx<T>* const*>(in),
                          reinterpret_cast<Cplx<T>* const*>(out));
  }
  void t2(const void* const* in, void* const* out) override {
    impl.transform_type_2(reinterpret_cast<const Cplx<T>* const*>(in),
                          reinterpret_cast<Cplx<T>* const*>(out));
  }
  void info(nnf_batch_info* i) override {
    i->precision = std::is_same_v<T, float> ? NNF_F32 : NNF_F64;
    i->dim = DIM;
    i->n_spread = impl.kernel().ns;
    i->count = impl.count();
    i->num_points = static_cast<int64_t>(impl.total_points());
    i->num_subproblems = static_cast<int64_t>(impl.num_subproblems());
    i->num_shapes = impl.num_groups();
    i->grid_cells = static_cast<int64_t>(impl.grid_cells());
    i->graph = impl.graph_active() ? 1 : 0;
    i->spread_mode = !impl.row_bands() ? NNF_SPREAD_TILES
                     : impl.fused_fft() ? NNF_SPREAD_BANDS
                                        : NNF_SPREAD_BANDS_GRID;
  }
};

nnf::PlanTuning to_tuning(const nnf_tuning* tuning) {
  nnf::PlanTuning t;
  if (tuning) {
    for (int d = 0; d < 3; ++d) {
      if (tuning->bin_size[d] < 0) throw neonufft::InputError("Neonufft:: negative bin size");
      t.bin[d] = tuning->bin_size[d];
    }
    if (tuning->max_points < 0) throw neonufft::InputError("Neonufft:: negative max_points");
    if (tuning->max_points > 0) t.max_points = static_cast<uint32_t>(tuning->max_points);
    if (tuning->spread_mode < NNF_SPREAD_AUTO || tuning->spread_mode > NNF_SPREAD_BANDS_GRID)
      throw neonufft::InputError("Neonufft:: unknown spread mode");
    if (tuning->band_rows < 0) throw neonufft::InputError("Neonufft:: negative band rows");
    t.spread_mode = tuning->spread_mode;
    t.band_rows = tuning->band_rows;
  }
  return t;
}

neonufft::Options to_options(const nnf_options* o) {
  neonufft::Options opt;
  if (o) {
    opt.tol = o->tol;
    opt.upsampfac = o->upsampfac;
    opt.recenter_threshold = o->recenter_threshold;
    opt.num_threads = o->num_threads;
    opt.sort_input = o->sort_input != 0;
    opt.sort_output = o->sort_output != 0;
    opt.kernel_approximation = o->kernel_approximation != 0;
    opt.order = o->order == 0 ? NEONUFFT_MODE_ORDER_FFT : NEONUFFT_MODE_ORDER_CMCL;
    opt.kernel_type = NEONUFFT_ES_KERNEL;
  }
  return opt;
}

void check_kind(int precision, int dim) {
  if (precision != NNF_F32 && precision != NNF_F64)
    throw neonufft::InputError("Neonufft:: precision must be NNF_F32 or NNF_F64");
  if (dim < 1 || dim > 3) throw neonufft::InputError("Neonufft:: dim must be 1, 2 or 3");
}

template <template <typename, int> class H, typename Base, typename... A>
Base* make_holder(int precision, int dim, A&&... a) {
  check_kind(precision, dim);
  if (precision == NNF_F32) {
    if (dim == 1) return new H<float, 1>(std::forward<A>(a)...);
    if (dim == 2) return new H<float, 2>(std::forward<A>(a)...);
    return new H<float, 3>(std::forward<A>(a)...);
  }
  if (dim == 1) return new H<double, 1>(std::forward<A>(a)...);
  if (dim == 2) return new H<double, 2>(std::forward<A>(a)...);
  return new H<double, 3>(std::forward<A>(a)...);
}

template <typename... A>
PlanT3Iface* make_t3_holder(int precision, int dim, A&&... a) {
  if (precision == NNF_MIXED) {
    if (dim < 1 || dim > 3) throw neonufft::InputError("Neonufft:: dim must be 1, 2 or 3");
    if (dim == 1) return new PlanT3Holder<float, 1, double>(std::forward<A>(a)...);
    if (dim == 2) return new PlanT3Holder<float, 2, double>(std::forward<A>(a)...);
    return new PlanT3Holder<float, 3, double>(std::forward<A>(a)...);
  }
  return make_holder<PlanT3Holder, PlanT3Iface>(precision, dim, std::forward<A>(a)...);
}

template <typename F>
int guarded(F&& f) {
  try {
    nnf::set_error_detail("");
    f();
    g_last_error.clear();
    return NEONUFFT_SUCCESS;
  } catch (const neonufft::GenericError& e) {
    g_last_error = e.what();
    if (!nnf::error_detail().empty()) g_last_error += std::string(" [") + nnf::error_detail() + "]";
    return e.error_code();
  } catch (const std::bad_alloc&) {
    g_last_error = "Neonufft:: host allocation failed";
    return NEONUFFT_MEMORY_ALLOC_ERROR;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return NEONUFFT_UNKNOWN_ERROR;
  } catch (...) {
    g_last_error = "Neonufft:: unknown error";
    return NEONUFFT_UNKNOWN_ERROR;
  }
}

std::shared_ptr<neonufft::Allocator> default_alloc() {
  return std::make_shared<neonufft::gpu::DeviceAllocator>();
}

}  // namespace

struct nnf_plan_s {
  std::unique_ptr<PlanIface> p;
};
struct nnf_plan_t3_s {
  std::unique_ptr<PlanT3Iface> p;
};

struct nnf_batch_s {
  std::unique_ptr<BatchIface> p;
};

#define NNF_REQUIRE_PLAN(pl)                                                  \
  if (!(pl) || !(pl)->p) throw neonufft::InputError("Neonufft:: null plan handle");

extern "C" {

NEONUFFT_EXPORT void nnf_options_default(nnf_options* opt) {
  if (!opt) return;
  const neonufft::Options d;
  opt->tol = d.tol;
  opt->upsampfac = d.upsampfac;
  opt->recenter_threshold = d.recenter_threshold;
  opt->num_threads = d.num_threads;
  opt->sort_input = d.sort_input;
  opt->sort_output = d.sort_output;
  opt->kernel_approximation = d.kernel_approximation;
  opt->order = d.order == NEONUFFT_MODE_ORDER_FFT ? 0 : 1;
  opt->kernel_type = 0;
}

NEONUFFT_EXPORT const char* nnf_last_error(void) { return g_last_error.c_str(); }

NEONUFFT_EXPORT const char* nnf_version(void) { return NEONUFFT_VERSION; }

NEONUFFT_EXPORT int nnf_plan_create_tuned(nnf_plan* plan, int precision, int dim,
                                          const nnf_options* opt, int sign, int64_t num_nu,
                                          const void* const* loc, const int64_t* modes,
                                          void* stream, const nnf_tuning* tuning) {
  return guarded([&] {
    if (!plan) throw neonufft::InputError("Neonufft:: null plan output pointer");
    *plan = nullptr;
    if (!modes) throw neonufft::InputError("Neonufft:: null modes");
    const nnf::PlanTuning t = to_tuning(tuning);
    auto h = std::make_unique<nnf_plan_s>();
    h->p.reset(make_holder<PlanHolder, PlanIface>(
        precision, dim, to_options(opt), sign, static_cast<IntType>(num_nu), loc, modes,
        static_cast<hipStream_t>(stream), default_alloc(), t));
    *plan = h.release();
  });
}

NEONUFFT_EXPORT int nnf_plan_create(nnf_plan* plan, int precision, int dim, const nnf_options* opt,
                                    int sign, int64_t num_nu, const void* const* loc,
                                    const int64_t* modes, void* stream) {
  return nnf_plan_create_tuned(plan, precision, dim, opt, sign, num_nu, loc, modes, stream,
                               nullptr);
}

NEONUFFT_EXPORT int nnf_plan_set_points(nnf_plan plan, int64_t num_nu, const void* const* loc) {
  return guarded([&] {
    NNF_REQUIRE_PLAN(plan);
    plan->p->set_points(num_nu, loc);
  });
}

NEONUFFT_EXPORT int nnf_plan_set_points_noalloc(nnf_plan plan, int64_t num_nu,
                                                const void* const* loc) {
  return guarded([&] {
    NNF_REQUIRE_PLAN(plan);
    plan->p->set_points_noalloc(num_nu, loc);
  });
}

NEONUFFT_EXPORT int nnf_plan_transform_type_1(nnf_plan plan, const void* in, void* out,
                                              const int64_t* out_strides) {
  return guarded([&] {
    NNF_REQUIRE_PLAN(plan);
    plan->p->t1(in, out, out_strides);
  });
}

NEONUFFT_EXPORT int nnf_plan_transform_type_2(nnf_plan plan, const void* in,
                                              const int64_t* in_strides, void* out) {
  return guarded([&] {
    NNF_REQUIRE_PLAN(plan);
    plan->p->t2(in, in_strides, out);
  });
}

NEONUFFT_EXPORT int nnf_plan_destroy(nnf_plan plan) {
  return guarded([&] { delete plan; });
}

NEONUFFT_EXPORT int nnf_plan_t3_grid_memory_size(int precision, int dim, const nnf_options* opt,
                                                 const double* input_min, const double* input_max,
                                                 const double* output_min,
                                                 const double* output_max, uint64_t* size) {
  return guarded([&] {
    // NNF_MIXED sizes its grids like NNF_F32 (they hold fp32 values)
    if (precision == NNF_MIXED) precision = NNF_F32;
    check_kind(precision, dim);
    if (!size || !input_min || !input_max || !output_min || !output_max)
      throw neonufft::InputError("Neonufft:: null argument");
    const neonufft::Options o = to_options(opt);
    uint64_t s = 0;
    if (precision == NNF_F32) {
      if (dim == 1) s = nnf::PlanT3Impl<float, 1>::grid_memory_size(o, input_min, input_max, output_min, output_max);
      if (dim == 2) s = nnf::PlanT3Impl<float, 2>::grid_memory_size(o, input_min, input_max, output_min, output_max);
      if (dim == 3) s = nnf::PlanT3Impl<float, 3>::grid_memory_size(o, input_min, input_max, output_min, output_max);
    } else {
      if (dim == 1) s = nnf::PlanT3Impl<double, 1>::grid_memory_size(o, input_min, input_max, output_min, output_max);
      if (dim == 2) s = nnf::PlanT3Impl<double, 2>::grid_memory_size(o, input_min, input_max, output_min, output_max);
      if (dim == 3) s = nnf::PlanT3Impl<double, 3>::grid_memory_size(o, input_min, input_max, output_min, output_max);
    }
    *size = s;
  });
}

NEONUFFT_EXPORT int nnf_plan_t3_create(nnf_plan_t3* plan, int precision, int dim,
                                       const nnf_options* opt, int sign, int64_t num_in,
                                       const void* const* input_points, int64_t num_out,
                                       const void* const* output_points, void* stream,
                                       int64_t batch_size) {
  return guarded([&] {
    if (!plan) throw neonufft::InputError("Neonufft:: null plan output pointer");
    *plan = nullptr;
    auto h = std::make_unique<nnf_plan_t3_s>();
    h->p.reset(make_t3_holder(
        precision, dim, to_options(opt), sign, static_cast<IntType>(num_in), input_points,
        static_cast<IntType>(num_out), output_points, static_cast<hipStream_t>(stream),
        static_cast<IntType>(batch_size), default_alloc()));
    *plan = h.release();
  });
}

NEONUFFT_EXPORT int nnf_plan_t3_create_bounds(nnf_plan_t3* plan, int precision, int dim,
                                              const nnf_options* opt, int sign,
                                              const double* input_min, const double* input_max,
                                              const double* output_min, const double* output_max,
                                              void* stream, int64_t batch_size) {
  return guarded([&] {
    if (!plan) throw neonufft::InputError("Neonufft:: null plan output pointer");
    *plan = nullptr;
    if (!input_min || !input_max || !output_min || !output_max)
      throw neonufft::InputError("Neonufft:: null bounds");
    auto h = std::make_unique<nnf_plan_t3_s>();
    h->p.reset(make_t3_holder(
        precision, dim, to_options(opt), sign, input_min, input_max, output_min, output_max,
        static_cast<hipStream_t>(stream), static_cast<IntType>(batch_size), default_alloc()));
    *plan = h.release();
  });
}

NEONUFFT_EXPORT int nnf_plan_t3_set_input_points(nnf_plan_t3 plan, int64_t num_in,
                                                 const void* const* points) {
  return guarded([&] {
    NNF_REQUIRE_PLAN(plan);
    plan->p->set_in(num_in, points);
  });
}

NEONUFFT_EXPORT int nnf_plan_t3_set_output_points(nnf_plan_t3 plan, int64_t num_out,
                                                  const void* const* points) {
  return guarded([&] {
    NNF_REQUIRE_PLAN(plan);
    plan->p->set_out(num_out, points);
  });
}

NEONUFFT_EXPORT int nnf_plan_t3_add_input(nnf_plan_t3 plan, const void* in, int64_t bdist) {
  return guarded([&] {
    NNF_REQUIRE_PLAN(plan);
    plan->p->add_input(in, bdist);
  });
}

NEONUFFT_EXPORT int nnf_plan_t3_transform(nnf_plan_t3 plan, void* out, int64_t bdist) {
  return guarded([&] {
    NNF_REQUIRE_PLAN(plan);
    plan->p->transform(out, bdist);
  });
}

NEONUFFT_EXPORT int nnf_plan_t3_reset(nnf_plan_t3 plan) {
  return guarded([&] {
    NNF_REQUIRE_PLAN(plan);
    plan->p->reset();
  });
}

NEONUFFT_EXPORT int nnf_plan_t3_destroy(nnf_plan_t3 plan) {
  return guarded([&] { delete plan; });
}

NEONUFFT_EXPORT int nnf_plan_get_info(nnf_plan plan, nnf_plan_info* info) {
  return guarded([&] {
    NNF_REQUIRE_PLAN(plan);
    if (!info) throw neonufft::InputError("Neonufft:: null info");
    plan->p->info(info);
  });
}

NEONUFFT_EXPORT int nnf_plan_t3_get_info(nnf_plan_t3 plan, nnf_plan_info* info) {
  return guarded([&] {
    NNF_REQUIRE_PLAN(plan);
    if (!info) throw neonufft::InputError("Neonufft:: null info");
    plan->p->info(info);
  });
}

NEONUFFT_EXPORT int nnf_plan_spread(nnf_plan plan, const void* in, int accumulate) {
  return guarded([&] {
    NNF_REQUIRE_PLAN(plan);
    plan->p->spread(in, accumulate != 0);
  });
}

NEONUFFT_EXPORT int nnf_plan_spread_tiles(nnf_plan plan, const void* in) {
  return guarded([&] {
    NNF_REQUIRE_PLAN(plan);
    plan->p->spread_tiles(in);
  });
}

NEONUFFT_EXPORT int nnf_plan_reduce_tiles(nnf_plan plan, int accumulate) {
  return guarded([&] {
    NNF_REQUIRE_PLAN(plan);
    plan->p->reduce_tiles(accumulate != 0);
  });
}

NEONUFFT_EXPORT int nnf_plan_interp(nnf_plan plan, void* out) {
  return guarded([&] {
    NNF_REQUIRE_PLAN(plan);
    plan->p->interp(out);
  });
}

NEONUFFT_EXPORT int nnf_plan_fine_grid(nnf_plan plan, void** grid) {
  return guarded([&] {
    NNF_REQUIRE_PLAN(plan);
    if (!grid) throw neonufft::InputError("Neonufft:: null output pointer");
    *grid = plan->p->grid();
  });
}

NEONUFFT_EXPORT int nnf_plan_fseries(nnf_plan plan, int d, const void** fs) {
  return guarded([&] {
    NNF_REQUIRE_PLAN(plan);
    if (!fs) throw neonufft::InputError("Neonufft:: null output pointer");
    *fs = plan->p->fseries(d);
    if (!*fs) throw neonufft::InputError("Neonufft:: dimension out of range");
  });
}

NEONUFFT_EXPORT int nnf_autotune(int precision, int dim, const nnf_options* opt, int sign,
                                 int64_t num_nu, const void* const* loc, const int64_t* modes,
                                 void* stream, int types, int reps, int use_cache,
                                 nnf_tuning* best, double* best_ms) {
  return guarded([&] {
    if (!modes || !best || !best_ms) throw neonufft::InputError("Neonufft:: null argument");
    nnf::autotune(precision, dim, to_options(opt), sign, static_cast<IntType>(num_nu), loc, modes,
                  static_cast<hipStream_t>(stream), types, reps, use_cache != 0, best, best_ms);
  });
}

NEONUFFT_EXPORT int nnf_autotune_clear_cache(void) {
  return guarded([&] { nnf::autotune_clear_cache(); });
}

NEONUFFT_EXPORT int nnf_batch_create_tuned(nnf_batch* batch, int precision, int dim,
                                           const nnf_options* opt, int sign, int64_t count,
                                           const int64_t* num_nu, const void* const* loc,
                                           const int64_t* modes, void* stream, int use_graph,
                                           const nnf_tuning* tuning) {
  return guarded([&] {
    if (!batch) throw neonufft::InputError("Neonufft:: null batch output pointer");
    *batch = nullptr;
    if (!modes || !num_nu) throw neonufft::InputError("Neonufft:: null modes or point counts");
    const nnf::PlanTuning t = to_tuning(tuning);
    auto h = std::make_unique<nnf_batch_s>();
    h->p.reset(make_holder<BatchHolder, BatchIface>(
        precision, dim, to_options(opt), sign, static_cast<IntType>(count), num_nu, loc, modes,
        static_cast<hipStream_t>(stream), default_alloc(), use_graph != 0, t));
    *batch = h.release();
  });
}

NEONUFFT_EXPORT int nnf_batch_create(nnf_batch* batch, int precision, int dim,
                                     const nnf_options* opt, int sign, int64_t count,
                                     const int64_t* num_nu, const void* const* loc,
                                     const int64_t* modes, void* stream, int use_graph) {
  return nnf_batch_create_tuned(batch, precision, dim, opt, sign, count, num_nu, loc, modes,
                                stream, use_graph, nullptr);
}

NEONUFFT_EXPORT int nnf_batch_set_points(nnf_batch batch, const int64_t* num_nu,
                                         const void* const* loc) {
  return guarded([&] {
    NNF_REQUIRE_PLAN(batch);
    batch->p->set_points(num_nu, loc);
  });
}

NEONUFFT_EXPORT int nnf_batch_transform_type_1(nnf_batch batch, const void* const* in,
                                               void* const* out) {
  return guarded([&] {
    NNF_REQUIRE_PLAN(batch);
    batch->p->t1(in, out);
  });
}

NEONUFFT_EXPORT int nnf_batch_transform_type_2(nnf_batch batch, const void* const* in,
                                               void* const* out) {
  return guarded([&] {
    NNF_REQUIRE_PLAN(batch);
    batch->p->t2(in, out);
  });
}

NEONUFFT_EXPORT int nnf_batch_get_info(nnf_batch batch, nnf_batch_info* info) {
  return guarded([&] {
    NNF_REQUIRE_PLAN(batch);
    if (!info) throw neonufft::InputError("Neonufft:: null info");
    batch->p->info(info);
  });
}

NEONUFFT_EXPORT int nnf_batch_destroy(nnf_batch batch) {
  return guarded([&] { delete batch; });
}

}  // extern "C"
