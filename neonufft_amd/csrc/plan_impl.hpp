// Plan orchestration: the type 1/2 pipeline and the type-3 pipeline.
//
// Mirrors the stage sequence of the reference's gpu::PlanImpl and
// gpu::PlanT3Impl (src/neonufft/gpu/plan_impl.hpp:36-167,
// src/neonufft/gpu/plan_t3_impl.hpp:48-326): fold + bin on set_points,
// spread -> FFT -> deconvolve for type 1, deconvolve -> FFT -> interpolate
// for type 2, and prephase x spread, deconvolve -> FFT -> interpolate x
// postphase for type 3. All work is enqueued on the plan's stream.
#pragma once

#include <neonufft/gpu/plan.hpp>
#include <neonufft/plan.hpp>

#include <algorithm>
#include <memory>
#include <vector>

#include <neonufft_amd.h>
#include "cover.hpp"
#include "fft.hpp"
#include "noalloc.hpp"
#include "params.hpp"
#include "point_set.hpp"
#include "pruned_fft.hpp"

namespace nnf {

// Spreading decomposition of 2D type 1 (NNF_SPREAD_* of the C ABI).
enum SpreadMode { kSpreadAuto = 0, kSpreadTiles = 1, kSpreadBands = 2, kSpreadBandsGrid = 3 };

struct PlanTuning {
  int bin[3] = {0, 0, 0};  // 0 = default
  uint32_t max_points = kDefaultMaxPoints;
  int spread_mode = kSpreadAuto;
  int band_rows = 0;  // 0 = default
};

// Row-band layout of a 2D plan: rows = 0 when subgrid tiles are used (asked
// for, or the grid does not suit the row view). fft: type 1 runs the x-FFT
// inside the band kernel (n0 a power of two >= 32, rows a power of two). LDS
// rows have pitch n0 + kBandPad (halo cells for runs crossing the row ends,
// and a shift of the banks from row to row for the transposed mode
// write-out).
struct BandLayout {
  int rows = 0;
  int pitch = 0;
  bool fft = false;
};

inline BandLayout plan_band_layout(const IntType* n, int ns, std::size_t acc_bytes,
                                   const PlanTuning& t, int min_bands) {
  BandLayout b;
  if (t.spread_mode == kSpreadTiles) return b;
  const IntType n0 = n[0];
  const bool pow2 = n0 >= 32 && (n0 & (n0 - 1)) == 0;
  b.fft = pow2 && t.spread_mode != kSpreadBandsGrid;
  b.pitch = static_cast<int>(n0 + kBandPad);
  int rows = band_rows(n0, b.pitch, n[1], ns, acc_bytes, kBandBudget, min_bands);
  if (rows > 0 && t.band_rows > 0) {
    rows = std::min(band_rows(n0, b.pitch, n[1], ns, acc_bytes, kBandBudget, 1), t.band_rows);
  }
  if (b.fft && rows > 0) {
    int p = 1;
    while (2 * p <= rows) p *= 2;
    rows = p;
  }
  if (rows <= 0) {
    if (t.spread_mode == kSpreadBands || t.spread_mode == kSpreadBandsGrid)
      throw neonufft::InputError("Neonufft:: grid does not suit row-band spreading");
    return BandLayout{};
  }
  b.rows = rows;
  return b;
}

// Row bands suit a point set when
//  - work: the fullest band's point range holds at most max(65536, 8 x the
//    uniform expectation) points (a band is one workgroup with no further
//    split; subgrid tiles cap a workgroup at 2048 points), and
//  - precision: no cell's fixed-point bound counts more than kBandMaxWindow
//    points (window_max); the rounding of the fixed-point sum grows with
//    that count times the largest strength, so dense grids (many points per
//    cell) keep the tiles' per-subproblem scale.
// row_start: n1 + 1 offsets of the row view.
// An explicit request for bands (check_work false) waives the work rule only.
constexpr uint32_t kBandMaxWindow = 1024;
inline bool bands_suitable(const std::vector<uint32_t>& row_start, int n1, int rows, int ns,
                           uint32_t window_max, bool check_work) {
  if (window_max > kBandMaxWindow) return false;
  if (!check_work) return true;
  const double m = row_start[static_cast<std::size_t>(n1)];
  const double expect = m * static_cast<double>(rows + ns - 1) / static_cast<double>(n1);
  const double limit = std::max(65536.0, 8.0 * expect);
  for (int y0 = 0; y0 < n1; y0 += rows) {
    const int hi = std::min(y0 + rows, n1);
    const int lo = y0 - (ns - 1);
    const double c = lo < 0 ? double(row_start[n1] - row_start[lo + n1]) + row_start[hi]
                            : double(row_start[hi] - row_start[lo]);
    if (c > limit) return false;
  }
  return true;
}

// Pencil height of the row kernels (rows.hip) for a 2D / 3D grid: the largest
// power of two <= 16 whose LDS (two row planes of n0 + 8 cells plus the
// subproblem-range table) fits kBandBudget; 0 when the row kernels do not
// apply (n0 not a power of two >= 32, or not even one row fits).
inline int row_fft_rows(int dim, const BinGeom& g, std::size_t tbytes) {
  const int n0 = g.n[0];
  if (dim < 2 || n0 < 32 || (n0 & (n0 - 1)) != 0) return 0;
  for (int rows = 16; rows >= 1; rows /= 2) {
    if (row_fft_lds_bytes(rows, n0 + 8, tbytes, row_fft_rng_entries(g, rows, dim)) <= kBandBudget) {
      return rows;
    }
  }
  return 0;
}

inline int log2_exact(IntType v) {
  int l = 0;
  while ((IntType(1) << l) < v) ++l;
  return l;
}

// Twiddles of every power-of-two length N <= len, level by level:
// tw[N/2 + q] = exp(sign 2 pi i q / N) for q < N/2 (tw[0] unused), in FP64.
template <typename T>
void upload_twiddles(DeviceBuffer<Cplx<T>>& tw, IntType len, int sign, hipStream_t stream,
                     const std::shared_ptr<neonufft::Allocator>& alloc) {
  std::vector<Cplx<T>> h(static_cast<std::size_t>(std::max<IntType>(len, 2)));
  h[0] = Cplx<T>{T(1), T(0)};
  for (IntType N = 2; N <= len; N *= 2) {
    for (IntType q = 0; q < N / 2; ++q) {
      const double a = 2.0 * kPi * static_cast<double>(q) / static_cast<double>(N);
      h[static_cast<std::size_t>(N / 2 + q)] =
          Cplx<T>{static_cast<T>(std::cos(a)), static_cast<T>(sign * std::sin(a))};
    }
  }
  tw.reset(h.size(), alloc);
  check_hip(hipMemcpyAsync(tw.data(), h.data(), h.size() * sizeof(Cplx<T>), hipMemcpyHostToDevice,
                           stream),
            "hipMemcpyAsync(twiddles)");
  check_hip(hipStreamSynchronize(stream), "hipStreamSynchronize");
}

template <typename T, int DIM>
class PlanImpl {
public:
  PlanImpl(const neonufft::Options& opt, int sign, IntType num_nu, const T* const* loc,
           const IntType* modes, hipStream_t stream, std::shared_ptr<neonufft::Allocator> alloc,
           const PlanTuning& tune = PlanTuning())
      : opt_(opt), sign_(sign), stream_(stream), alloc_(std::move(alloc)), tune_(tune) {
    if (!alloc_) throw neonufft::InputError("Neonufft:: null allocator");
    check_hip(hipGetDevice(&device_), "hipGetDevice");
    ks_ = make_kernel_setup(opt_.tol, opt_.upsampfac, std::is_same_v<T, float>);
    for (int d = 0; d < 3; ++d) {
      modes_[d] = 1;
      n_[d] = 1;
    }
    std::size_t cells = 1;
    for (int d = 0; d < DIM; ++d) {
      if (modes[d] <= 0) throw neonufft::InputError("Neonufft:: modes must be positive");
      modes_[d] = modes[d];
      // n_d = max(2 ns, N_d * sigma) truncated (gpu/plan_impl.hpp:56-60)
      n_[d] = std::max<IntType>(2 * ks_.ns,
                                static_cast<IntType>(static_cast<double>(modes[d]) * opt_.upsampfac));
      cells *= static_cast<std::size_t>(n_[d]);
    }
    grid_.reset(cells, alloc_);
    if constexpr (DIM == 1) {
      fft_ = FftPlan(std::is_same_v<T, double>, DIM, n_, sign_, 1, stream_, alloc_);
    } else {
      pfft_.init(n_, modes_, sign_, 1, stream_, alloc_);
    }
    for (int d = 0; d < DIM; ++d) {
      fs_[d].reset(static_cast<std::size_t>(n_[d] / 2 + 1), alloc_);
      launch_fseries<T>(fs_[d].data(), n_[d], ks_.quad, stream_);
    }
    geom_ = choose_bins(DIM, n_, ks_.ns, sizeof(Cplx<T>), tune_.bin);
    cover_.build(DIM, geom_, ks_.ns, stream_, alloc_);
    if constexpr (DIM == 2) {
      band_ = plan_band_layout(n_, ks_.ns, sizeof(T), tune_, 256);
      band_rows_ = band_.rows;
      if (band_rows_ > 0) bounds_.reset(band_bound_words(static_cast<int>(n_[1]), sizeof(T)), alloc_);
    }
    // row kernels: the x-FFT in LDS joined to the tile reduction (type 1) and
    // to the zero padding (type 2); NNF_SPREAD_TILES keeps the separate
    // reduction, rocFFT x pass and transposes
    if (tune_.spread_mode != kSpreadTiles) row_rows_ = row_fft_rows(DIM, geom_, sizeof(T));
    // 3D with the row kernels: the y and z passes in LDS too, when both
    // extents are powers of two >= 32
    IntType tw_len = n_[0];
    if constexpr (DIM == 3) {
      auto pow2 = [](IntType v) { return v >= 32 && (v & (v - 1)) == 0; };
      if (row_rows_ > 0 && pow2(n_[1]) && pow2(n_[2])) {
        axis_rows_y_ = axis_fft_rows(n_[1], sizeof(T), kBandBudget);
        axis_rows_z_ = axis_fft_rows(n_[2], sizeof(T), kBandBudget);
        if (axis_rows_y_ == 0 || axis_rows_z_ == 0) axis_rows_y_ = axis_rows_z_ = 0;
        if (axis_rows_y_ > 0) tw_len = std::max({n_[0], n_[1], n_[2]});
      }
    }
    // 2D: the y pass of type 1 in LDS (power-of-two n1) uses the same table
    if constexpr (DIM == 2) tw_len = std::max(n_[0], n_[1]);
    if ((DIM == 2 && band_.fft) || row_rows_ > 0) upload_twiddles<T>(tw_, tw_len, sign_, stream_, alloc_);
    set_points(num_nu, loc);
  }

  // ws: null = buffers grow as needed; otherwise nothing is allocated (the
  // temporaries come from *ws, the results must fit the memory an earlier
  // allocating set_points left, else InputError). Until a set_points
  // succeeds, the plan refuses transforms.
  void set_points(IntType num_nu, const T* const* loc, Workspace* ws = nullptr) {
    if (num_nu < 0) throw neonufft::InputError("Neonufft:: negative number of points");
    Loc<T> l{{nullptr, nullptr, nullptr}};
    for (int d = 0; d < DIM; ++d) l.p[d] = loc ? loc[d] : nullptr;
    FoldParams fp{{0, 0, 0}, {1, 1, 1}};
    const std::size_t m = static_cast<std::size_t>(num_nu);
    if (ws) {
      // everything this call takes from ws, checked before any work
      std::size_t need = PointSet<T, DIM>::scratch_bytes(
          m, static_cast<std::size_t>(geom_.nbins[0]) * geom_.nbins[1] * geom_.nbins[2]);
      if constexpr (DIM == 2) {
        if (band_rows_ > 0) {
          need = std::max(need, RowView<T>::scratch_bytes(m, static_cast<std::size_t>(n_[1])));
          if (sorted_.size() < std::max<std::size_t>(m, 1)) {
            throw neonufft::InputError(
                "Neonufft:: more points than the plan holds memory for (set_points without "
                "allocation)");
          }
        }
      }
      if (need > ws->cap) {
        throw neonufft::InputError(
            "Neonufft:: set_points without allocation needs more scratch memory than the "
            "library's fixed maximum");
      }
    }
    points_valid_ = false;
    pts_.build(l, m, fp, geom_, ks_.ns, tune_.max_points, stream_, alloc_, ws);
    bands_on_ = false;
    if constexpr (DIM == 2) {
      if (band_rows_ > 0) {
        if (ws) ws->used = 0;  // the point set's temporaries are no longer needed
        rows_.build(pts_, ks_.ns, stream_, alloc_, ws);
        if (!ws) sorted_.ensure(std::max<std::size_t>(pts_.size(), 1), alloc_);
        row_host_.resize(static_cast<std::size_t>(n_[1]) + 2);
        check_hip(hipMemcpyAsync(row_host_.data(), rows_.row_start(),
                                 (row_host_.size() - 1) * sizeof(uint32_t),
                                 hipMemcpyDeviceToHost, stream_),
                  "hipMemcpyAsync(row offsets)");
        check_hip(hipMemcpyAsync(row_host_.data() + n_[1] + 1, rows_.window_max(),
                                 sizeof(uint32_t), hipMemcpyDeviceToHost, stream_),
                  "hipMemcpyAsync(window max)");
      }
    }
    check_hip(hipStreamSynchronize(stream_), "hipStreamSynchronize");
    if constexpr (DIM == 2) {
      bands_on_ = band_rows_ > 0 &&
                  bands_suitable(row_host_, static_cast<int>(n_[1]), band_rows_, ks_.ns,
                                 row_host_[static_cast<std::size_t>(n_[1]) + 1],
                                 tune_.spread_mode == kSpreadAuto);
    }
    if (!bands_on_) {
      const std::size_t need = std::max<std::size_t>(pts_.nsub(), 1) * tile_cells(geom_);
      if (ws) {
        if (tiles_.size() < need) {
          throw neonufft::InputError(
              "Neonufft:: the points need more spreading tiles than the plan holds memory for "
              "(set_points without allocation)");
        }
      } else if (tiles_.size() < need) {
        // 1/16 headroom, so that point sets of the same size with a few more
        // subproblems fit without a new allocation
        const std::size_t nsub = std::max<std::size_t>(pts_.nsub(), 1);
        tiles_.reset((nsub + nsub / 16 + 1) * tile_cells(geom_), alloc_);
      }
    }
    points_valid_ = true;
  }

  // set_points with no device allocation: temporaries in the device's fixed
  // scratch array (held for the whole call, which ends with the stream
  // synchronised), results in the memory of earlier set_points calls.
  void set_points_noalloc(IntType num_nu, const T* const* loc) {
    ScratchLease lease = lease_noalloc_scratch(device_);
    try {
      set_points(num_nu, loc, &lease.ws);
    } catch (...) {
      // work already enqueued may still read the array: finish it before
      // another call can take the array
      (void)hipStreamSynchronize(stream_);
      throw;
    }
  }

  // Stage 1 of type 1: form the fine grid from the strengths (overwriting
  // it, or adding to it when accumulating). Subgrid tiles: spread into
  // per-subproblem tiles, then reduce them into the grid. Row bands (2D):
  // gather the strengths into row order, then spread band by band.
  void spread(const Cplx<T>* in, bool accumulate = false) {
    spread_tiles(in);
    reduce_tiles(accumulate);
  }

  // The two kernels of spread(), separately (for stage timing).
  void spread_tiles(const Cplx<T>* in) {
    require_points();
    if (pts_.size() && !in) throw neonufft::InputError("Neonufft:: null input");
    if constexpr (DIM == 2) {
      if (bands_on_) {
        launch_clear_bounds(bounds_.data(), bounds_.size(), stream_);
        launch_band_gather<T>(nullptr, band_job(in), 1, pts_.size(), stream_);
        return;
      }
    }
    launch_spread<T, DIM>(ks_.ns, pts_.recs(), pts_.perm(), pts_.subs(), pts_.nsub(), in, nullptr,
                          tiles_.data(), geom_, ks_.es, stream_);
  }
  void reduce_tiles(bool accumulate) {
    require_points();
    if constexpr (DIM == 2) {
      if (bands_on_) {
        const int nb = static_cast<int>((n_[1] + band_rows_ - 1) / band_rows_);
        launch_band_spread<T>(ks_.ns, nullptr, nullptr, band_job(nullptr), 1,
                              static_cast<std::size_t>(nb),
                              band_lds_bytes(band_rows_, band_.pitch, sizeof(T)), accumulate,
                              false, ks_.es, stream_);
        return;
      }
    }
    const ReduceJob<T, DIM> job{tiles_.data(), pts_.sub_starts(), grid_.data(), geom_,
                                cover_.table()};
    launch_tile_reduce<T, DIM>(job, ks_.ns, accumulate, stream_);
  }

  // Stage 3 of type 2: interpolate the current fine grid.
  void interp(Cplx<T>* out) {
    require_points();
    if (pts_.size() && !out) throw neonufft::InputError("Neonufft:: null output");
    launch_interp<T, DIM>(ks_.ns, pts_.recs(), pts_.perm(), pts_.subs(), pts_.nsub(), grid_.data(),
                          nullptr, out, geom_, ks_.es, stream_);
  }

  void transform_type_1(const Cplx<T>* in, Cplx<T>* out, const IntType* out_strides) {
    require_points();
    if (!out || !out_strides) throw neonufft::InputError("Neonufft:: null output");
    T* fs[3] = {fs_[0].data(), fs_[1].data(), fs_[2].data()};
    if constexpr (DIM == 2) {
      if (bands_on_ && band_.fft) {
        // gather, then spread + x-FFT per band straight into the pruned
        // FFT's intermediate; y-FFT and deconvolution follow
        if (pts_.size() && !in) throw neonufft::InputError("Neonufft:: null input");
        launch_clear_bounds(bounds_.data(), bounds_.size(), stream_);
        launch_band_gather<T>(nullptr, band_job(in), 1, pts_.size(), stream_);
        const int nb = static_cast<int>((n_[1] + band_rows_ - 1) / band_rows_);
        launch_band_spread<T>(ks_.ns, nullptr, nullptr, band_job(nullptr), 1,
                              static_cast<std::size_t>(nb),
                              band_lds_bytes(band_rows_, band_.pitch, sizeof(T)), false, true,
                              ks_.es, stream_);
        forward_from_h(fs, out, out_strides);
        return;
      }
    }
    if constexpr (DIM >= 2) {
      if (row_rows_ > 0 && !bands_on_ && bins_regular(DIM, geom_, ks_.ns)) {
        // tiles -> retained x modes in one pass, then the remaining passes
        spread_tiles(in);
        launch_reduce_rows_fft<T, DIM>(row_job(), ks_.ns, stream_);
        if (axis_rows_y_ > 0) {
          launch_axis_fft_type1<T>(axis_job(out, nullptr, out_strides), stream_);
        } else {
          forward_from_h(fs, out, out_strides);
        }
        return;
      }
    }
    spread(in);
    if constexpr (DIM == 1) {
      fft_.execute(grid_.data());
      launch_downsample<T, DIM>(order_fft(), grid_.data(), n_, fs, out, out_strides, modes_,
                                stream_);
    } else {
      pfft_.forward(grid_.data(), fs, order_fft(), out, nullptr, out_strides, stream_);
    }
  }

  void transform_type_2(const Cplx<T>* in, const IntType* in_strides, Cplx<T>* out) {
    require_points();
    if (!in || !in_strides) throw neonufft::InputError("Neonufft:: null input");
    T* fs[3] = {fs_[0].data(), fs_[1].data(), fs_[2].data()};
    if constexpr (DIM == 1) {
      launch_upsample<T, DIM>(order_fft(), in, in_strides, modes_, fs, grid_.data(), n_, stream_);
      fft_.execute(grid_.data());
    } else {
      if (row_rows_ > 0) {
        if (axis_rows_y_ > 0) {
          launch_axis_fft_type2<T>(axis_job(nullptr, in, in_strides), stream_);
        } else if (DIM == 2 && pfft_.y_lds_rows() > 0) {
          pfft_.backward_to_h_lds(tw_.data(), in, nullptr, in_strides, fs, order_fft(), stream_);
        } else {
          pfft_.backward_to_h(in, nullptr, in_strides, fs, order_fft(), stream_);
        }
        launch_expand_rows_ifft<T>(row_job(), stream_);
      } else {
        pfft_.backward(in, nullptr, in_strides, fs, order_fft(), grid_.data(), stream_);
      }
    }
    interp(out);
  }

  void info(nnf_plan_info* info) const {
    *info = nnf_plan_info{};
    info->precision = std::is_same_v<T, float> ? NNF_F32 : NNF_F64;
    info->dim = DIM;
    info->n_spread = ks_.ns;
    info->es_beta = ks_.beta;
    for (int d = 0; d < 3; ++d) {
      info->modes[d] = modes_[d];
      info->grid[d] = n_[d];
      info->spread_grid[d] = n_[d];
      info->bin_size[d] = geom_.bin[d];
    }
    info->num_bins = static_cast<int64_t>(pts_.nbins());
    info->num_subproblems = static_cast<int64_t>(pts_.nsub());
    info->num_points = static_cast<int64_t>(pts_.size());
    info->horner_degree = ks_.horner_deg;
    info->quad_nodes = ks_.quad.nq;
    info->spread_mode = !bands_on_ ? kSpreadTiles : (band_.fft ? kSpreadBands : kSpreadBandsGrid);
    info->band_rows = bands_on_ ? band_rows_ : 0;
  }

  Cplx<T>* fine_grid() const { return grid_.data(); }
  const T* fseries(int d) const { return d >= 0 && d < DIM ? fs_[d].data() : nullptr; }

private:
  int order_fft() const { return opt_.order == NEONUFFT_MODE_ORDER_FFT ? 1 : 0; }

  void require_points() const {
    if (!points_valid_) {
      throw neonufft::InputError("Neonufft:: no valid points (the last set_points failed)");
    }
  }

  // Type 1 after the x pass (h filled): the y pass in LDS for 2D plans with a
  // power-of-two n1, else the rocFFT passes and the deconvolution.
  void forward_from_h(T* const* fs, Cplx<T>* out, const IntType* st) {
    if (DIM == 2 && pfft_.y_lds_rows() > 0) {
      pfft_.forward_from_h_lds(tw_.data(), fs, order_fft(), out, nullptr, st, stream_);
    } else {
      pfft_.forward_from_h(fs, order_fft(), out, nullptr, st, stream_);
    }
  }

  AxisFftJob<T> axis_job(Cplx<T>* out, const Cplx<T>* in, const IntType* st) const {
    AxisFftJob<T> j{};
    j.h = pfft_.h();
    j.k = pfft_.k();
    j.modes_out = out;
    j.modes_in = in;
    for (int d = 0; d < 3; ++d) {
      j.fs[d] = fs_[d].data();
      j.st[d] = static_cast<long long>(st[d]);
    }
    j.order_fft = order_fft();
    j.tw = tw_.data();
    j.N0 = static_cast<int>(modes_[0]);
    j.N1 = static_cast<int>(modes_[1]);
    j.N2 = static_cast<int>(modes_[2]);
    j.n1 = static_cast<int>(n_[1]);
    j.n2 = static_cast<int>(n_[2]);
    j.log2n1 = log2_exact(n_[1]);
    j.log2n2 = log2_exact(n_[2]);
    j.rows_y = axis_rows_y_;
    j.log2rows_y = log2_exact(axis_rows_y_);
    j.rows_z = axis_rows_z_;
    j.log2rows_z = log2_exact(axis_rows_z_);
    return j;
  }

  RowFftJob<T> row_job() const {
    RowFftJob<T> j{};
    j.tiles = tiles_.data();
    j.sub_start = pts_.sub_starts();
    j.g = geom_;
    j.h = pfft_.h();
    j.grid = grid_.data();
    j.tw = tw_.data();
    j.N0 = static_cast<int>(modes_[0]);
    j.log2n0 = log2_exact(n_[0]);
    j.rows = row_rows_;
    j.log2rows = log2_exact(row_rows_);
    j.pitch = static_cast<int>(n_[0] + 8);
    return j;
  }

  BandJob<T> band_job(const Cplx<T>* in) const {
    BandJob<T> j{};
    j.pts = rows_.recs();
    j.perm = rows_.perm();
    j.row_start = rows_.row_start();
    j.vrow = rows_.vrow();
    j.bcount = rows_.block_counts();
    j.rowsum = reinterpret_cast<T*>(bounds_.data());
    j.rowmax = j.rowsum + n_[1];
    j.in = in;
    j.sorted = sorted_.data();
    j.grid = grid_.data();
    j.m = pts_.size();
    j.n0 = static_cast<int>(n_[0]);
    j.n1 = static_cast<int>(n_[1]);
    j.rows = band_rows_;
    j.pitch = band_.pitch;
    if (band_.fft) {
      j.h = pfft_.h();
      j.tw = tw_.data();
      j.tw_len = static_cast<int>(n_[0]);
      j.N0 = static_cast<int>(modes_[0]);
      j.log2n0 = log2_exact(n_[0]);
      j.log2rows = log2_exact(band_rows_);
    }
    return j;
  }

  neonufft::Options opt_;
  int sign_;
  hipStream_t stream_;
  std::shared_ptr<neonufft::Allocator> alloc_;
  PlanTuning tune_;
  int device_ = 0;                // device current at construction
  KernelSetup ks_;
  IntType modes_[3];
  IntType n_[3];
  BinGeom geom_{};
  int band_rows_ = 0;             // 2D row-band spreading (0 = subgrid tiles)
  int row_rows_ = 0;              // pencil height of the row kernels (0 = not used)
  int axis_rows_y_ = 0;           // 3D: lines per workgroup of the LDS y pass (0 = rocFFT)
  int axis_rows_z_ = 0;           // 3D: lines per workgroup of the LDS z pass
  bool bands_on_ = false;         // row bands used for the current points
  bool points_valid_ = false;     // the last set_points succeeded
  std::vector<uint32_t> row_host_;  // row offsets of the view (balance check)
  BandLayout band_;
  DeviceBuffer<Cplx<T>> tw_;      // x-FFT twiddles of the fused type 1
  DeviceBuffer<unsigned long long> bounds_;  // row sums + row maxima of |c| (row bands)
  RowView<T> rows_;               // row-sorted view (row bands)
  DeviceBuffer<Cplx<T>> sorted_;  // strengths in view order (row bands)
  DeviceBuffer<Cplx<T>> grid_;
  DeviceBuffer<Cplx<T>> tiles_;  // spreading tiles, tile_cells(geom_) per subproblem
  CoverBuffers cover_;
  FftPlan fft_;                                // 1D plans
  PrunedFft<T, DIM == 1 ? 2 : DIM> pfft_;      // 2D / 3D plans
  DeviceBuffer<T> fs_[3];
  PointSet<T, DIM> pts_;
};

// T: precision of spreading, FFT and interpolation. TC: precision of the
// coordinates and of the complex inputs / outputs; TC = double with T = float
// is the mixed-precision plan (FP64 fold and phases from double coordinates,
// complex128 I/O converted at the plan's edges, fp32 bulk work).
template <typename T, int DIM, typename TC = T>
class PlanT3Impl {
  static constexpr bool kMixed = !std::is_same_v<T, TC>;

public:
  // sizeof(complex T) * (spread grid + 2 * FFT grid), the reference's formula
  // (gpu/plan_t3_impl.hpp:54-72).
  static uint64_t grid_memory_size(const neonufft::Options& opt, const double* in_min,
                                   const double* in_max, const double* out_min,
                                   const double* out_max) {
    int ns;
    double beta;
    kernel_width_and_beta(opt.tol, opt.upsampfac, &ns, &beta);
    const GridInfoT3 gi = make_grid_info_t3(DIM, ns, opt.upsampfac, opt.recenter_threshold, in_min,
                                            in_max, out_min, out_max);
    uint64_t sg = 1, fg = 1;
    for (int d = 0; d < DIM; ++d) {
      sg *= static_cast<uint64_t>(gi.spread_grid_size[d]);
      fg *= static_cast<uint64_t>(gi.fft_grid_size[d]);
    }
    return sizeof(Cplx<T>) * (sg + 2 * fg);
  }

  PlanT3Impl(const neonufft::Options& opt, int sign, IntType num_in, const TC* const* in_pts,
             IntType num_out, const TC* const* out_pts, hipStream_t stream, IntType batch,
             std::shared_ptr<neonufft::Allocator> alloc, const PlanTuning& tune = PlanTuning())
      : opt_(opt), sign_(sign), stream_(stream), batch_(batch), alloc_(std::move(alloc)), tune_(tune) {
    check_common();
    if (num_in < 0 || num_out < 0) throw neonufft::InputError("Neonufft:: negative point count");
    double bounds[4][3];
    minmax(num_in, in_pts, bounds[0], bounds[1]);
    minmax(num_out, out_pts, bounds[2], bounds[3]);
    init(bounds[0], bounds[1], bounds[2], bounds[3]);
    set_input_points(num_in, in_pts);
    set_output_points(num_out, out_pts);
  }

  PlanT3Impl(const neonufft::Options& opt, int sign, const double* in_min, const double* in_max,
             const double* out_min, const double* out_max, hipStream_t stream, IntType batch,
             std::shared_ptr<neonufft::Allocator> alloc, const PlanTuning& tune = PlanTuning())
      : opt_(opt), sign_(sign), stream_(stream), batch_(batch), alloc_(std::move(alloc)), tune_(tune) {
    check_common();
    init(in_min, in_max, out_min, out_max);
  }

  void set_input_points(IntType num_in, const TC* const* pts) {
    if (num_in < 0) throw neonufft::InputError("Neonufft:: negative point count");
    Loc<TC> l = make_loc(pts);
    FoldParams fp{};
    for (int d = 0; d < 3; ++d) {
      fp.offset[d] = d < DIM ? gi_.input_offsets[d] : 0.0;
      fp.scale[d] = d < DIM ? 1.0 / gi_.gamma[d] : 1.0;
    }
    in_pts_.build(l, static_cast<std::size_t>(num_in), fp, geom_in_, ks_.ns, tune_.max_points,
                  stream_, alloc_);
    tiles_.ensure(std::max<std::size_t>(in_pts_.nsub(), 1) * tile_cells(geom_in_), alloc_);
    has_pre_ = false;
    for (int d = 0; d < DIM; ++d) has_pre_ = has_pre_ || gi_.output_offsets[d] != 0.0;
    if (has_pre_) {
      pre_.ensure(num_in ? static_cast<std::size_t>(num_in) : 1, alloc_);
      launch_prephase<TC, T, DIM>(sign_, l, in_pts_.perm(), static_cast<std::size_t>(num_in),
                                  gi_.output_offsets, pre_.data(), stream_);
    }
    if constexpr (kMixed) io_in_.ensure(num_in ? static_cast<std::size_t>(num_in) : 1, alloc_);
  }

  void set_output_points(IntType num_out, const TC* const* pts) {
    if (num_out < 0) throw neonufft::InputError("Neonufft:: negative point count");
    Loc<TC> l = make_loc(pts);
    FoldParams fp{};
    double scale[3] = {1, 1, 1};
    for (int d = 0; d < 3; ++d) {
      fp.offset[d] = d < DIM ? gi_.output_offsets[d] : 0.0;
      scale[d] = d < DIM ? gi_.gamma[d] * gi_.h[d] : 1.0;
      fp.scale[d] = scale[d];
    }
    out_pts_.build(l, static_cast<std::size_t>(num_out), fp, geom_out_, ks_.ns, tune_.max_points,
                   stream_, alloc_);
    post_.ensure(num_out ? static_cast<std::size_t>(num_out) : 1, alloc_);
    launch_postphase<TC, T, DIM>(sign_, l, out_pts_.perm(), static_cast<std::size_t>(num_out),
                                 gi_.input_offsets, gi_.output_offsets, scale, ks_.quad,
                                 post_.data(), stream_);
    if constexpr (kMixed) io_out_.ensure(num_out ? static_cast<std::size_t>(num_out) : 1, alloc_);
  }

  void add_input(const Cplx<TC>* in, IntType bdist) {
    if (bdist <= 0) bdist = static_cast<IntType>(in_pts_.size());
    if (in_pts_.size() && !in) throw neonufft::InputError("Neonufft:: null input");
    for (IntType b = 0; b < batch_; ++b) {
      const Cplx<T>* src;
      if constexpr (kMixed) {
        launch_convert<TC, T>(in + b * bdist, io_in_.data(), in_pts_.size(), stream_);
        src = io_in_.data();
      } else {
        src = in + b * bdist;
      }
      launch_spread<T, DIM>(ks_.ns, in_pts_.recs(), in_pts_.perm(), in_pts_.subs(), in_pts_.nsub(),
                            src, has_pre_ ? pre_.data() : nullptr, tiles_.data(), geom_in_,
                            ks_.es, stream_);
      const ReduceJob<T, DIM> job{tiles_.data(), in_pts_.sub_starts(),
                                  spread_grid_.data() + b * spread_cells_, geom_in_,
                                  cover_in_.table()};
      launch_tile_reduce<T, DIM>(job, ks_.ns, true, stream_);
    }
  }

  void transform(Cplx<TC>* out, IntType bdist) {
    if (bdist <= 0) bdist = static_cast<IntType>(out_pts_.size());
    if (out_pts_.size() && !out) throw neonufft::InputError("Neonufft:: null output");
    T* fs[3] = {fs_[0].data(), fs_[1].data(), fs_[2].data()};
    IntType strides[3] = {1, 1, 1};
    for (int d = 1; d < DIM; ++d) strides[d] = strides[d - 1] * gi_.spread_grid_size[d - 1];
    for (IntType b = 0; b < batch_; ++b) {
      // The spread grid is read as CMCL-ordered modes (gpu/plan_t3_impl.hpp:227).
      if constexpr (DIM == 1) {
        launch_upsample<T, DIM>(0, spread_grid_.data() + b * spread_cells_, strides,
                                gi_.spread_grid_size, fs, fft_grid_.data(), gi_.fft_grid_size,
                                stream_);
        fft_.execute(fft_grid_.data());
      } else {
        pfft_.backward(spread_grid_.data() + b * spread_cells_, nullptr, strides, fs, 0,
                       fft_grid_.data(), stream_);
      }
      Cplx<T>* dst;
      if constexpr (kMixed) {
        dst = io_out_.data();
      } else {
        dst = out + b * bdist;
      }
      launch_interp<T, DIM>(ks_.ns, out_pts_.recs(), out_pts_.perm(), out_pts_.subs(),
                            out_pts_.nsub(), fft_grid_.data(), post_.data(), dst, geom_out_,
                            ks_.es, stream_);
      if constexpr (kMixed) {
        launch_convert<T, TC>(io_out_.data(), out + b * bdist, out_pts_.size(), stream_);
      }
    }
  }

  void reset() {
    launch_fill_zero<T>(spread_grid_.data(), spread_grid_.size(), stream_);
  }

  void info(nnf_plan_info* info) const {
    *info = nnf_plan_info{};
    info->precision = kMixed ? NNF_MIXED : (std::is_same_v<T, float> ? NNF_F32 : NNF_F64);
    info->dim = DIM;
    info->n_spread = ks_.ns;
    info->es_beta = ks_.beta;
    for (int d = 0; d < 3; ++d) {
      info->modes[d] = gi_.spread_grid_size[d];
      info->grid[d] = gi_.fft_grid_size[d];
      info->spread_grid[d] = gi_.spread_grid_size[d];
      info->bin_size[d] = geom_in_.bin[d];
    }
    info->num_bins = static_cast<int64_t>(in_pts_.nbins());
    info->num_subproblems = static_cast<int64_t>(in_pts_.nsub());
    info->num_points = static_cast<int64_t>(in_pts_.size());
    info->num_out_points = static_cast<int64_t>(out_pts_.size());
    info->horner_degree = ks_.horner_deg;
    info->quad_nodes = ks_.quad.nq;
  }

  const GridInfoT3& grid_info() const { return gi_; }

private:
  void check_common() {
    if (!alloc_) throw neonufft::InputError("Neonufft:: null allocator");
    if (batch_ < 1) throw neonufft::InputError("Neonufft:: batch size must be >= 1");
    ks_ = make_kernel_setup(opt_.tol, opt_.upsampfac, std::is_same_v<T, float>);
  }

  static Loc<TC> make_loc(const TC* const* pts) {
    Loc<TC> l{{nullptr, nullptr, nullptr}};
    for (int d = 0; d < DIM; ++d) l.p[d] = pts ? pts[d] : nullptr;
    return l;
  }

  // Per-dimension min/max of device coordinates, finished on the host
  // (the reference does the same with hipCUB + a blocking copy,
  // gpu/plan_t3_impl.hpp:88-102).
  void minmax(IntType m, const TC* const* pts, double* lo, double* hi) {
    for (int d = 0; d < 3; ++d) {
      lo[d] = 0.0;
      hi[d] = 0.0;
    }
    if (m <= 0) return;
    const int nblk = minmax_blocks(static_cast<std::size_t>(m));
    DeviceBuffer<double> part(static_cast<std::size_t>(2 * nblk * DIM), alloc_);
    for (int d = 0; d < DIM; ++d) {
      if (!pts || !pts[d]) throw neonufft::InputError("Neonufft:: null coordinate pointer");
      launch_minmax<TC>(pts[d], static_cast<std::size_t>(m), part.data() + 2 * nblk * d, stream_);
    }
    std::vector<double> host(static_cast<std::size_t>(2 * nblk * DIM));
    check_hip(hipMemcpyAsync(host.data(), part.data(), host.size() * sizeof(double),
                             hipMemcpyDeviceToHost, stream_),
              "hipMemcpyAsync(min/max)");
    check_hip(hipStreamSynchronize(stream_), "hipStreamSynchronize");
    for (int d = 0; d < DIM; ++d) {
      double a = host[2 * nblk * d], b = host[2 * nblk * d + 1];
      for (int k = 1; k < nblk; ++k) {
        a = std::min(a, host[2 * nblk * d + 2 * k]);
        b = std::max(b, host[2 * nblk * d + 2 * k + 1]);
      }
      lo[d] = a;
      hi[d] = b;
    }
  }

  void init(const double* in_min, const double* in_max, const double* out_min,
            const double* out_max) {
    gi_ = make_grid_info_t3(DIM, ks_.ns, opt_.upsampfac, opt_.recenter_threshold, in_min, in_max,
                            out_min, out_max);
    spread_cells_ = 1;
    std::size_t fft_cells = 1;
    for (int d = 0; d < DIM; ++d) {
      spread_cells_ *= gi_.spread_grid_size[d];
      fft_cells *= static_cast<std::size_t>(gi_.fft_grid_size[d]);
    }
    spread_grid_.reset(static_cast<std::size_t>(spread_cells_ * batch_), alloc_);
    launch_fill_zero<T>(spread_grid_.data(), spread_grid_.size(), stream_);
    fft_grid_.reset(fft_cells, alloc_);
    if constexpr (DIM == 1) {
      fft_ = FftPlan(std::is_same_v<T, double>, DIM, gi_.fft_grid_size, sign_, 1, stream_, alloc_);
    } else {
      pfft_.init(gi_.fft_grid_size, gi_.spread_grid_size, sign_, 1, stream_, alloc_);
    }
    for (int d = 0; d < DIM; ++d) {
      fs_[d].reset(static_cast<std::size_t>(gi_.fft_grid_size[d] / 2 + 1), alloc_);
      launch_fseries<T>(fs_[d].data(), gi_.fft_grid_size[d], ks_.quad, stream_);
    }
    geom_in_ = choose_bins(DIM, gi_.spread_grid_size, ks_.ns, sizeof(Cplx<T>), tune_.bin);
    geom_out_ = choose_bins(DIM, gi_.fft_grid_size, ks_.ns, sizeof(Cplx<T>), tune_.bin);
    cover_in_.build(DIM, geom_in_, ks_.ns, stream_, alloc_);
    check_hip(hipStreamSynchronize(stream_), "hipStreamSynchronize");
  }

  neonufft::Options opt_;
  int sign_;
  hipStream_t stream_;
  IntType batch_;
  std::shared_ptr<neonufft::Allocator> alloc_;
  PlanTuning tune_;
  KernelSetup ks_;
  GridInfoT3 gi_;
  IntType spread_cells_ = 1;
  BinGeom geom_in_{}, geom_out_{};
  DeviceBuffer<Cplx<T>> spread_grid_, fft_grid_, pre_, post_;
  DeviceBuffer<Cplx<T>> io_in_, io_out_;  // mixed precision: converted inputs / outputs
  DeviceBuffer<Cplx<T>> tiles_;           // spreading tiles of the input points
  CoverBuffers cover_in_;
  bool has_pre_ = false;
  FftPlan fft_;                            // 1D plans
  PrunedFft<T, DIM == 1 ? 2 : DIM> pfft_;  // 2D / 3D plans
  DeviceBuffer<T> fs_[3];
  PointSet<T, DIM> in_pts_, out_pts_;
};

}  // namespace nnf
