// Autotuner for the spreading / interpolation launch shape (no reference
// counterpart: the reference fixes 8-cell bins, src/neonufft/gpu/util/
// partition_group.hpp:10-15, and one thread per point).
//
// For a plan's precision, dimension, kernel width, fine grid and points it
// times the spreading and/or interpolation kernels for a list of subgrid tile
// shapes and per-workgroup point caps, on the caller's own points, and keeps
// the fastest. 2D plans that spread with row bands time their type 1 for a
// list of band heights instead (the tile shape then only serves
// interpolation), and the two choices are returned together.
//
// Results are cached per (precision, dim, ns, fine grid, log2 of the point
// count, kernel selection, point-distribution signature). The signature comes
// from the points binned on the default tiles: the largest bin's count over
// the mean count of the occupied bins (half-octave buckets) and the share of
// empty bins (quarter buckets), so uniform and clustered point sets of equal
// size get separate entries. The cache lives in the process and, unless
// disabled, in a text file that later processes read: NEONUFFT_AMD_TUNE_CACHE
// names the file ("" disables it), else
// $XDG_CACHE_HOME (or $HOME/.cache)/neonufft_amd/autotune-<version>-<gfx arch>.txt.
// Every line of the file also carries the library version and the device's
// architecture name, and only lines matching both are read.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <filesystem>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>
#include <string>
#include <tuple>
#include <vector>

#include <neonufft/gpu/device_allocator.hpp>
#include <neonufft_amd.h>
#include "plan_impl.hpp"

namespace nnf {

namespace {

// precision, dim, ns, n0, n1, n2, log2 M, types, occupancy bucket, empty bucket
using Key = std::tuple<int, int, int, int64_t, int64_t, int64_t, int, int, int, int>;

struct Entry {
  nnf_tuning tuning;
  double ms;
  int source;  // 1: timed in this process, 2: read from the cache file
};

std::mutex g_cache_mutex;
std::map<Key, Entry> g_cache;
std::string g_loaded_file;  // cache file already read into g_cache
bool g_file_read = false;

// The device's architecture name ("gfx950"), file-name safe.
std::string device_arch() {
  int dev = 0;
  check_hip(hipGetDevice(&dev), "hipGetDevice");
  hipDeviceProp_t prop;
  check_hip(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties");
  std::string a(prop.gcnArchName);
  for (char& ch : a) {
    if (!std::isalnum(static_cast<unsigned char>(ch))) ch = '_';
  }
  return a.empty() ? std::string("unknown") : a;
}

std::string cache_file(const std::string& arch) {
  if (const char* env = std::getenv("NEONUFFT_AMD_TUNE_CACHE")) return std::string(env);
  std::string base;
  if (const char* x = std::getenv("XDG_CACHE_HOME"); x && *x) {
    base = x;
  } else if (const char* h = std::getenv("HOME"); h && *h) {
    base = std::string(h) + "/.cache";
  } else {
    return std::string();
  }
  return base + "/neonufft_amd/autotune-" + NEONUFFT_VERSION + "-" + arch + ".txt";
}

// One line per entry: "nnf-tune <version> <arch>" then the key fields, then
// bin_size[3] max_points spread_mode band_rows ms.
std::string format_line(const std::string& arch, const Key& k, const nnf_tuning& t, double ms) {
  std::ostringstream o;
  o << "nnf-tune " << NEONUFFT_VERSION << ' ' << arch << ' ' << std::get<0>(k) << ' '
    << std::get<1>(k) << ' ' << std::get<2>(k) << ' ' << std::get<3>(k) << ' ' << std::get<4>(k)
    << ' ' << std::get<5>(k) << ' ' << std::get<6>(k) << ' ' << std::get<7>(k) << ' '
    << std::get<8>(k) << ' ' << std::get<9>(k) << ' ' << t.bin_size[0] << ' ' << t.bin_size[1]
    << ' ' << t.bin_size[2] << ' ' << t.max_points << ' ' << t.spread_mode << ' ' << t.band_rows
    << ' ' << ms << '\n';
  return o.str();
}

// Reads the file's entries for this version and architecture (once per file
// and process; a missing or unreadable file is an empty cache). Caller holds
// g_cache_mutex.
void load_file(const std::string& path, const std::string& arch) {
  if (path.empty() || (g_file_read && g_loaded_file == path)) return;
  g_file_read = true;
  g_loaded_file = path;
  std::ifstream in(path);
  std::string line;
  while (std::getline(in, line)) {
    std::istringstream is(line);
    std::string tag, ver, a;
    int prec, dim, ns, mb, types, occ, empty;
    int64_t n0, n1, n2;
    nnf_tuning t{};
    double ms;
    if (!(is >> tag >> ver >> a >> prec >> dim >> ns >> n0 >> n1 >> n2 >> mb >> types >> occ >>
          empty >> t.bin_size[0] >> t.bin_size[1] >> t.bin_size[2] >> t.max_points >>
          t.spread_mode >> t.band_rows >> ms)) {
      continue;
    }
    if (tag != "nnf-tune" || ver != NEONUFFT_VERSION || a != arch) continue;
    const Key k{prec, dim, ns, n0, n1, n2, mb, types, occ, empty};
    if (!g_cache.count(k)) g_cache[k] = Entry{t, ms, 2};
  }
}

// Appends one entry (best effort: an unwritable location only loses the
// persistence). Caller holds g_cache_mutex.
void append_file(const std::string& path, const std::string& line) {
  if (path.empty()) return;
  std::error_code ec;
  const std::filesystem::path p(path);
  if (p.has_parent_path()) std::filesystem::create_directories(p.parent_path(), ec);
  std::ofstream out(path, std::ios::app);
  if (out) out << line;
}

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

// Point-distribution signature on the default tiles: (occupancy bucket,
// empty-bin bucket). occupancy = round(2 log2(largest bin count / mean count
// of the occupied bins)), empty = floor(4 x share of empty bins).
template <typename T, int DIM>
std::pair<int, int> point_signature(const neonufft::Options& opt, IntType num_nu,
                                    const T* const* loc, const IntType* modes,
                                    hipStream_t stream) {
  if (num_nu <= 0) return {0, 4};
  auto alloc = std::make_shared<neonufft::gpu::DeviceAllocator>();
  const KernelSetup ks = make_kernel_setup(opt.tol, opt.upsampfac, std::is_same_v<T, float>);
  IntType n[3] = {1, 1, 1};
  for (int d = 0; d < DIM; ++d) {
    n[d] = std::max<IntType>(2 * ks.ns,
                             static_cast<IntType>(static_cast<double>(modes[d]) * opt.upsampfac));
  }
  const int no_bin[3] = {0, 0, 0};
  const BinGeom g = choose_bins(DIM, n, ks.ns, sizeof(Cplx<T>), no_bin);
  Loc<T> l{{nullptr, nullptr, nullptr}};
  for (int d = 0; d < DIM; ++d) l.p[d] = loc[d];
  const FoldParams fp{{0, 0, 0}, {1, 1, 1}};
  PointSet<T, DIM> ps;
  ps.build(l, static_cast<std::size_t>(num_nu), fp, g, ks.ns, kMaxSubproblemPoints, stream, alloc);
  const std::vector<uint32_t> counts = ps.host_bin_counts(stream);
  uint32_t mx = 0;
  std::size_t occupied = 0;
  for (uint32_t c : counts) {
    mx = std::max(mx, c);
    occupied += c > 0;
  }
  const double mean = static_cast<double>(num_nu) / static_cast<double>(std::max<std::size_t>(occupied, 1));
  const int occ = std::clamp(static_cast<int>(std::lround(2.0 * std::log2(std::max(1.0, mx / mean)))), 0, 40);
  const double empty_share =
      1.0 - static_cast<double>(occupied) / static_cast<double>(std::max<std::size_t>(counts.size(), 1));
  const int empty = std::clamp(static_cast<int>(4.0 * empty_share), 0, 4);
  return {occ, empty};
}

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
  std::pair<int, int> sig;
  if (precision == NNF_F32) {
    const auto* l = reinterpret_cast<const float* const*>(loc);
    if (dim == 1) sig = point_signature<float, 1>(opt, num_nu, l, modes, stream);
    if (dim == 2) sig = point_signature<float, 2>(opt, num_nu, l, modes, stream);
    if (dim == 3) sig = point_signature<float, 3>(opt, num_nu, l, modes, stream);
  } else {
    const auto* l = reinterpret_cast<const double* const*>(loc);
    if (dim == 1) sig = point_signature<double, 1>(opt, num_nu, l, modes, stream);
    if (dim == 2) sig = point_signature<double, 2>(opt, num_nu, l, modes, stream);
    if (dim == 3) sig = point_signature<double, 3>(opt, num_nu, l, modes, stream);
  }
  const Key key{precision, dim, ks.ns, n[0], n[1], n[2], mbucket, types, sig.first, sig.second};
  const std::string arch = device_arch();
  const std::string file = cache_file(arch);
  if (use_cache) {
    std::lock_guard<std::mutex> lock(g_cache_mutex);
    load_file(file, arch);
    auto it = g_cache.find(key);
    if (it != g_cache.end()) {
      *best = it->second.tuning;
      best->reserved[0] = it->second.source;
      *best_ms = it->second.ms;
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
  best->reserved[0] = 0;
  *best_ms = r.second;
  std::lock_guard<std::mutex> lock(g_cache_mutex);
  g_cache[key] = Entry{r.first, r.second, 1};
  if (use_cache) append_file(file, format_line(arch, key, r.first, r.second));
}

void autotune_clear_cache() {
  std::lock_guard<std::mutex> lock(g_cache_mutex);
  g_cache.clear();
}

}  // namespace nnf
