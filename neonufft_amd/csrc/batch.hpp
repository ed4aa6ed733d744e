// Many independent type 1 / type 2 transforms of uneven sizes in one plan.
//
// No reference counterpart: the reference runs one gpu::Plan per transform
// (src/neonufft/gpu/plan_impl.hpp:36-167). Here every transform keeps its own
// points and fine grid, but the work of all of them is issued together:
//  - one zero fill over all fine grids,
//  - one spreading (or interpolation) launch over the union of all
//    subproblems, each subproblem naming its transform (Subprob::job),
//  - one batched rocFFT plan per distinct grid shape (grids of one shape are
//    contiguous),
//  - one batched deconvolution launch per shape.
// The sequence can be captured once into a HIP graph and replayed; per-call
// input / output pointers live in device-side tables refreshed before each
// replay, so the graph stays valid when the caller's buffers change.
#pragma once

#include <cstring>
#include <map>
#include <memory>
#include <vector>

#include "cover.hpp"
#include "plan_impl.hpp"

namespace nnf {

// Stream-ordered upload of a small host table into a fixed device buffer.
// Unchanged contents are not sent again; otherwise the bytes go through one
// of a few pinned slots, each reused only after its previous copy finished.
class TableUpload {
public:
  TableUpload() = default;
  TableUpload(const TableUpload&) = delete;
  TableUpload& operator=(const TableUpload&) = delete;
  ~TableUpload() {
    for (int k = 0; k < kSlots; ++k) {
      if (ev_[k]) (void)hipEventDestroy(ev_[k]);
      if (slot_[k]) (void)hipHostFree(slot_[k]);
    }
  }

  void init(std::size_t bytes) {
    bytes_ = bytes;
    for (int k = 0; k < kSlots; ++k) {
      check_hip(hipHostMalloc(&slot_[k], bytes ? bytes : 1, hipHostMallocDefault), "hipHostMalloc");
      check_hip(hipEventCreateWithFlags(&ev_[k], hipEventDisableTiming), "hipEventCreate");
    }
    last_.assign(bytes, 0);
    valid_ = false;
  }

  void invalidate() { valid_ = false; }

  void upload(void* dst, const void* src, hipStream_t stream) {
    if (valid_ && std::memcmp(last_.data(), src, bytes_) == 0) return;
    const int k = next_;
    next_ = (next_ + 1) % kSlots;
    if (used_[k]) check_hip(hipEventSynchronize(ev_[k]), "hipEventSynchronize");
    std::memcpy(slot_[k], src, bytes_);
    check_hip(hipMemcpyAsync(dst, slot_[k], bytes_, hipMemcpyHostToDevice, stream),
              "hipMemcpyAsync(table)");
    check_hip(hipEventRecord(ev_[k], stream), "hipEventRecord");
    used_[k] = true;
    std::memcpy(last_.data(), src, bytes_);
    valid_ = true;
  }

private:
  static constexpr int kSlots = 4;
  std::size_t bytes_ = 0;
  void* slot_[kSlots] = {};
  hipEvent_t ev_[kSlots] = {};
  bool used_[kSlots] = {};
  int next_ = 0;
  bool valid_ = false;
  std::vector<unsigned char> last_;
};

template <typename T, int DIM>
class BatchPlanImpl {
public:
  BatchPlanImpl(const neonufft::Options& opt, int sign, IntType count, const IntType* num_nu,
                const T* const* loc, const IntType* modes, hipStream_t stream,
                std::shared_ptr<neonufft::Allocator> alloc, bool use_graph,
                const PlanTuning& tune = PlanTuning())
      : opt_(opt), sign_(sign), stream_(stream), alloc_(std::move(alloc)), tune_(tune),
        use_graph_(use_graph) {
    if (!alloc_) throw neonufft::InputError("Neonufft:: null allocator");
    if (count <= 0) throw neonufft::InputError("Neonufft:: batch count must be positive");
    // one launch dimension (gridDim.y / z <= 65535) indexes the transforms of
    // the batched reductions and up / downsampling
    if (count > 65535) throw neonufft::InputError("Neonufft:: at most 65535 transforms per batch");
    if (!num_nu || !modes) throw neonufft::InputError("Neonufft:: null argument");
    ks_ = make_kernel_setup(opt_.tol, opt_.upsampfac, std::is_same_v<T, float>);
    count_ = static_cast<int>(count);
    items_.resize(count_);
    // group transforms by mode shape (equal modes -> equal fine grid)
    std::map<std::vector<IntType>, int> shape_to_group;
    for (int t = 0; t < count_; ++t) {
      Item& it = items_[t];
      std::vector<IntType> key;
      for (int d = 0; d < 3; ++d) {
        it.modes[d] = 1;
        it.n[d] = 1;
      }
      for (int d = 0; d < DIM; ++d) {
        const IntType N = modes[t * DIM + d];
        if (N <= 0) throw neonufft::InputError("Neonufft:: modes must be positive");
        it.modes[d] = N;
        it.n[d] = std::max<IntType>(2 * ks_.ns,
                                    static_cast<IntType>(static_cast<double>(N) * opt_.upsampfac));
        key.push_back(N);
      }
      auto f = shape_to_group.find(key);
      if (f == shape_to_group.end()) {
        Group g;
        for (int d = 0; d < 3; ++d) {
          g.modes[d] = it.modes[d];
          g.n[d] = it.n[d];
        }
        groups_.push_back(std::move(g));
        f = shape_to_group.emplace(key, static_cast<int>(groups_.size()) - 1).first;
      }
      it.group = f->second;
      groups_[it.group].members.push_back(t);
    }
    // lay the grids out group by group, each group starting on a 256-byte
    // boundary (the up/downsampling kernels store cell pairs as 16 bytes)
    std::size_t off = 0;
    for (Group& g : groups_) {
      g.cells = 1;
      for (int d = 0; d < DIM; ++d) g.cells *= static_cast<std::size_t>(g.n[d]);
      constexpr std::size_t kAlignCells = 256 / sizeof(Cplx<T>);
      off = (off + kAlignCells - 1) / kAlignCells * kAlignCells;
      g.offset = off;
      for (std::size_t k = 0; k < g.members.size(); ++k) {
        items_[g.members[k]].grid_off = off + k * g.cells;
        items_[g.members[k]].member = static_cast<int>(k);
      }
      off += g.cells * g.members.size();
    }
    grids_.reset(off, alloc_);
    for (Group& g : groups_) {
      if constexpr (DIM == 1) {
        g.fft = FftPlan(std::is_same_v<T, double>, DIM, g.n, sign_,
                        static_cast<IntType>(g.members.size()), stream_, alloc_);
      } else {
        g.pfft.init(g.n, g.modes, sign_, static_cast<int>(g.members.size()), stream_, alloc_);
      }
      for (int d = 0; d < DIM; ++d) {
        g.fs[d].reset(static_cast<std::size_t>(g.n[d] / 2 + 1), alloc_);
        launch_fseries<T>(g.fs[d].data(), g.n[d], ks_.quad, stream_);
      }
      g.geom = choose_bins(DIM, g.n, ks_.ns, sizeof(Cplx<T>), tune_.bin);
      g.cover.build(DIM, g.geom, ks_.ns, stream_, alloc_);
      max_lds_ = std::max(max_lds_, tile_bytes<T>(g.geom));
      max_n0_ = std::max(max_n0_, g.geom.n[0]);
      max_rows_ = std::max(max_rows_, g.geom.n[1] * g.geom.n[2]);
      max_bins_ = std::max<std::size_t>(
          max_bins_, static_cast<std::size_t>(g.geom.nbins[0]) * g.geom.nbins[1] * g.geom.nbins[2]);
      regular_ = regular_ && bins_regular(DIM, g.geom, ks_.ns);
      if constexpr (DIM == 2) {
        g.band = plan_band_layout(g.n, ks_.ns, sizeof(T), tune_, 1);
        bands_ = bands_ && g.band.rows > 0;
        fused_ = fused_ && g.band.fft;
        band_lds_ = std::max(band_lds_, band_lds_bytes(g.band.rows, g.band.pitch, sizeof(T)));
        tw_len_ = std::max<IntType>(tw_len_, std::max(g.n[0], g.n[1]));
      }
    }
    if constexpr (DIM != 2) bands_ = false;
    fused_ = fused_ && bands_;
    if (fused_) upload_twiddles<T>(tw_, tw_len_, sign_, stream_, alloc_);
    std::size_t io_off = 0;
    for (Group& g : groups_) {
      g.io_off = io_off;
      io_off += g.members.size();
    }
    pts_.resize(count_);
    sjobs_.reset(count_, alloc_);
    ijobs_.reset(count_, alloc_);
    rjobs_.reset(count_, alloc_);
    host_rjobs_.resize(count_);
    io_.reset(count_, alloc_);
    host_sjobs_.resize(count_);
    host_ijobs_.resize(count_);
    host_io_.resize(count_);
    up_sjobs_.init(sizeof(SpreadJob<T, DIM>) * count_);
    if (bands_) {
      // per transform: n1 row sums and n1 row maxima of |c|
      std::size_t words = 0;
      for (int t = 0; t < count_; ++t) {
        items_[t].bound_off = words;
        words += band_bound_words(static_cast<int>(items_[t].n[1]), sizeof(T));
      }
      bounds_.reset(words, alloc_);
      bjobs_.reset(count_, alloc_);
      host_bjobs_.resize(count_);
      rows_.resize(count_);
      up_bjobs_.init(sizeof(BandJob<T>) * count_);
    }
    up_ijobs_.init(sizeof(InterpJob<T, DIM>) * count_);
    up_io_.init(sizeof(void*) * count_);
    check_hip(hipStreamCreateWithFlags(&capture_stream_, hipStreamNonBlocking), "hipStreamCreate");
    set_points(num_nu, loc);
  }

  ~BatchPlanImpl() {
    drop_graphs();
    if (capture_stream_) (void)hipStreamDestroy(capture_stream_);
  }

  BatchPlanImpl(const BatchPlanImpl&) = delete;
  BatchPlanImpl& operator=(const BatchPlanImpl&) = delete;

  // loc: count * DIM device pointers, transform-major.
  void set_points(const IntType* num_nu, const T* const* loc) {
    if (!num_nu) throw neonufft::InputError("Neonufft:: null point counts");
    drop_graphs();
    std::size_t total_sub = 0, total_tiles = 0;
    total_points_ = 0;
    for (int t = 0; t < count_; ++t) {
      if (num_nu[t] < 0) throw neonufft::InputError("Neonufft:: negative number of points");
      Loc<T> l{{nullptr, nullptr, nullptr}};
      for (int d = 0; d < DIM; ++d) l.p[d] = loc ? loc[t * DIM + d] : nullptr;
      FoldParams fp{{0, 0, 0}, {1, 1, 1}};
      pts_[t].build(l, static_cast<std::size_t>(num_nu[t]), fp, groups_[items_[t].group].geom,
                    ks_.ns, tune_.max_points, stream_, alloc_);
      items_[t].sub_off = total_sub;
      items_[t].pt_off = total_points_;
      total_sub += pts_[t].nsub();
      total_points_ += static_cast<std::size_t>(num_nu[t]);
    }
    bands_on_ = false;
    if constexpr (DIM == 2) {
      if (bands_) {
        total_bands_ = 0;
        row_host_.resize(count_);
        for (int t = 0; t < count_; ++t) {
          const Group& g = groups_[items_[t].group];
          rows_[t].build(pts_[t], ks_.ns, stream_, alloc_);
          items_[t].band_off = total_bands_;
          total_bands_ += static_cast<std::size_t>((g.n[1] + g.band.rows - 1) / g.band.rows);
          row_host_[t].resize(static_cast<std::size_t>(g.n[1]) + 2);
          check_hip(hipMemcpyAsync(row_host_[t].data(), rows_[t].row_start(),
                                   (row_host_[t].size() - 1) * sizeof(uint32_t),
                                   hipMemcpyDeviceToHost, stream_),
                    "hipMemcpyAsync(row offsets)");
          check_hip(hipMemcpyAsync(row_host_[t].data() + g.n[1] + 1, rows_[t].window_max(),
                                   sizeof(uint32_t), hipMemcpyDeviceToHost, stream_),
                    "hipMemcpyAsync(window max)");
        }
        sorted_.ensure(total_points_ ? total_points_ : 1, alloc_);
        up_bjobs_.invalidate();
        check_hip(hipStreamSynchronize(stream_), "hipStreamSynchronize");
        // every transform falls back to tiles when one of them does not
        // suit row bands
        bands_on_ = true;
        for (int t = 0; t < count_ && bands_on_; ++t) {
          const Group& g = groups_[items_[t].group];
          bands_on_ = bands_suitable(row_host_[t], static_cast<int>(g.n[1]), g.band.rows, ks_.ns,
                                     row_host_[t][static_cast<std::size_t>(g.n[1]) + 1],
                                     tune_.spread_mode == kSpreadAuto);
        }
        if (bands_on_) {
          // the transform of every band of the launch
          band_job_host_.resize(total_bands_);
          for (int t = 0; t < count_; ++t) {
            const Group& g = groups_[items_[t].group];
            const std::size_t nb = static_cast<std::size_t>((g.n[1] + g.band.rows - 1) / g.band.rows);
            std::fill_n(band_job_host_.begin() + static_cast<std::ptrdiff_t>(items_[t].band_off), nb,
                        static_cast<uint32_t>(t));
          }
          band_job_.ensure(total_bands_ ? total_bands_ : 1, alloc_);
          if (total_bands_) {
            check_hip(hipMemcpyAsync(band_job_.data(), band_job_host_.data(),
                                     total_bands_ * sizeof(uint32_t), hipMemcpyHostToDevice, stream_),
                      "hipMemcpyAsync(band jobs)");
          }
          build_gather_chunks(num_nu);
        }
      }
    }
    for (int t = 0; t < count_; ++t) {
      items_[t].tiles_off = total_tiles;
      if (!bands_on_) total_tiles += pts_[t].nsub() * tile_cells(groups_[items_[t].group].geom);
    }
    tiles_.ensure(total_tiles ? total_tiles : 1, alloc_);
    for (int t = 0; t < count_; ++t) {
      const Item& it = items_[t];
      const Group& g = groups_[it.group];
      host_rjobs_[t] = ReduceJob<T, DIM>{tiles_.data() + it.tiles_off, pts_[t].sub_starts(),
                                         grids_.data() + it.grid_off, g.geom, g.cover.table()};
    }
    check_hip(hipMemcpyAsync(rjobs_.data(), host_rjobs_.data(),
                             sizeof(ReduceJob<T, DIM>) * count_, hipMemcpyHostToDevice, stream_),
              "hipMemcpyAsync(reduce jobs)");
    subs_.ensure(total_sub ? total_sub : 1, alloc_);
    up_sjobs_.invalidate();
    up_ijobs_.invalidate();
    nsub_ = total_sub;
    for (int t = 0; t < count_; ++t) {
      launch_subprob_retag(pts_[t].subs(), pts_[t].nsub(), static_cast<uint32_t>(t),
                           subs_.data() + items_[t].sub_off, stream_);
    }
    check_hip(hipStreamSynchronize(stream_), "hipStreamSynchronize");
  }

  // in / out: count device pointers each; modes contiguous (x fastest).
  void transform_type_1(const Cplx<T>* const* in, Cplx<T>* const* out) {
    if (!in || !out) throw neonufft::InputError("Neonufft:: null pointer table");
    if constexpr (DIM == 2) {
      if (bands_on_) {
        for (int t = 0; t < count_; ++t) {
          const Item& it = items_[t];
          const Group& g = groups_[it.group];
          BandJob<T> j{};
          j.pts = rows_[t].recs();
          j.perm = rows_[t].perm();
          j.row_start = rows_[t].row_start();
          j.vrow = rows_[t].vrow();
          j.bcount = rows_[t].block_counts();
          j.rowsum = reinterpret_cast<T*>(bounds_.data() + it.bound_off);
          j.rowmax = j.rowsum + g.n[1];
          j.in = in[t];
          j.sorted = sorted_.data() + it.pt_off;
          j.grid = grids_.data() + it.grid_off;
          j.m = pts_[t].size();
          j.pt_off = it.pt_off;
          j.band_off = it.band_off;
          j.n0 = static_cast<int>(g.n[0]);
          j.n1 = static_cast<int>(g.n[1]);
          j.rows = g.band.rows;
          j.pitch = g.band.pitch;
          if (fused_) {
            j.h = g.pfft.h() + static_cast<std::size_t>(it.member) *
                                   static_cast<std::size_t>(g.modes[0] * g.n[1]);
            j.tw = tw_.data();
            j.tw_len = static_cast<int>(tw_len_);
            j.N0 = static_cast<int>(g.modes[0]);
            j.log2n0 = log2_exact(g.n[0]);
            j.log2rows = log2_exact(g.band.rows);
          }
          host_bjobs_[t] = j;
        }
        up_bjobs_.upload(bjobs_.data(), host_bjobs_.data(), stream_);
        upload_io(reinterpret_cast<void* const*>(out));
        run(1);
        return;
      }
    }
    for (int t = 0; t < count_; ++t) {
      const Item& it = items_[t];
      host_sjobs_[t] = SpreadJob<T, DIM>{pts_[t].recs(), pts_[t].perm(), in[t], nullptr,
                                         tiles_.data() + it.tiles_off,
                                         static_cast<uint32_t>(it.sub_off), 0u,
                                         groups_[it.group].geom};
    }
    up_sjobs_.upload(sjobs_.data(), host_sjobs_.data(), stream_);
    upload_io(reinterpret_cast<void* const*>(out));
    run(1);
  }

  void transform_type_2(const Cplx<T>* const* in, Cplx<T>* const* out) {
    if (!in || !out) throw neonufft::InputError("Neonufft:: null pointer table");
    for (int t = 0; t < count_; ++t) {
      const Item& it = items_[t];
      host_ijobs_[t] = InterpJob<T, DIM>{pts_[t].recs(), pts_[t].perm(),
                                         grids_.data() + it.grid_off, nullptr, out[t],
                                         groups_[it.group].geom};
    }
    up_ijobs_.upload(ijobs_.data(), host_ijobs_.data(), stream_);
    upload_io(const_cast<void* const*>(reinterpret_cast<const void* const*>(in)));
    run(2);
  }

  std::size_t total_points() const { return total_points_; }
  std::size_t num_subproblems() const { return nsub_; }
  int num_groups() const { return static_cast<int>(groups_.size()); }
  std::size_t grid_cells() const { return grids_.size(); }
  int count() const { return count_; }
  bool graph_active() const { return graph_ok_[0] || graph_ok_[1]; }
  bool row_bands() const { return bands_on_; }
  bool fused_fft() const { return fused_; }
  const KernelSetup& kernel() const { return ks_; }

private:
  struct Item {
    IntType modes[3];
    IntType n[3];
    std::size_t grid_off = 0;
    std::size_t sub_off = 0;
    std::size_t tiles_off = 0;  // first tile cell of this transform in tiles_
    std::size_t pt_off = 0;     // first point of this transform over all transforms
    std::size_t band_off = 0;   // first row band of this transform (row bands)
    std::size_t bound_off = 0;  // first word of this transform's row sums / maxima (row bands)
    int group = 0;
    int member = 0;             // index of this transform inside its group
  };
  struct Group {
    IntType modes[3];
    IntType n[3];
    std::size_t cells = 0, offset = 0;
    std::vector<int> members;
    FftPlan fft;                            // 1D groups
    PrunedFft<T, DIM == 1 ? 2 : DIM> pfft;  // 2D / 3D groups
    DeviceBuffer<T> fs[3];
    BinGeom geom{};
    CoverBuffers cover;
    std::size_t io_off = 0;  // first entry of this group's caller pointers in io_
    BandLayout band;         // row-band layout (row-band spreading)
  };

  int order_fft() const { return opt_.order == NEONUFFT_MODE_ORDER_FFT ? 1 : 0; }

  // Caller pointers in group-major order: group g reads io_[g.io_off ...].
  void upload_io(void* const* ptrs) {
    std::size_t k = 0;
    for (const Group& g : groups_) {
      for (int t : g.members) host_io_[k++] = ptrs[t];
    }
    up_io_.upload(io_.data(), host_io_.data(), stream_);
  }

  void enqueue(int type, hipStream_t s) {
    if (type == 1) {
      // every fine-grid cell is written, so the grids need no zero fill
      if (bands_on_) {
        if constexpr (DIM == 2) {
          launch_clear_bounds(bounds_.data(), bounds_.size(), s);
          launch_band_gather_chunks<T>(bjobs_.data(), gather_chunks_.data(), n_gather_chunks_, s);
          launch_band_spread<T>(ks_.ns, bjobs_.data(), band_job_.data(), BandJob<T>{}, count_,
                                total_bands_, band_lds_, false, fused_, ks_.es, s);
        }
      } else {
        launch_spread_batch<T, DIM>(ks_.ns, sjobs_.data(), subs_.data(), nsub_, max_lds_, ks_.es,
                                    s);
        launch_tile_reduce_batch<T, DIM>(rjobs_.data(), count_, ks_.ns, regular_, max_bins_,
                                         max_n0_, max_rows_, false, s);
      }
      for (Group& g : groups_) {
        T* fs[3] = {g.fs[0].data(), g.fs[1].data(), g.fs[2].data()};
        IntType st[3] = {1, g.modes[0], g.modes[0] * g.modes[1]};
        Cplx<T>* const* outs = reinterpret_cast<Cplx<T>* const*>(io_.data() + g.io_off);
        if constexpr (DIM == 1) {
          g.fft.set_stream(s);
          g.fft.execute(grids_.data() + g.offset);
          launch_downsample_batch<T, DIM>(order_fft(), grids_.data() + g.offset, g.cells, g.n,
                                          fs, outs, st, g.modes,
                                          static_cast<int>(g.members.size()), s);
        } else {
          g.pfft.set_stream(s);
          if constexpr (DIM == 2) {
            if (fused_ && bands_on_) {
              if (g.pfft.y_lds_rows() > 0) {
                g.pfft.forward_from_h_lds(tw_.data(), fs, order_fft(), nullptr, outs, st, s);
              } else {
                g.pfft.forward_from_h(fs, order_fft(), nullptr, outs, st, s);
              }
              continue;
            }
          }
          g.pfft.forward(grids_.data() + g.offset, fs, order_fft(), nullptr, outs, st, s);
        }
      }
    } else {
      for (Group& g : groups_) {
        T* fs[3] = {g.fs[0].data(), g.fs[1].data(), g.fs[2].data()};
        IntType st[3] = {1, g.modes[0], g.modes[0] * g.modes[1]};
        const Cplx<T>* const* ins =
            reinterpret_cast<const Cplx<T>* const*>(io_.data() + g.io_off);
        if constexpr (DIM == 1) {
          launch_upsample_batch<T, DIM>(order_fft(), ins, st, g.modes, fs,
                                        grids_.data() + g.offset, g.cells, g.n,
                                        static_cast<int>(g.members.size()), s);
          g.fft.set_stream(s);
          g.fft.execute(grids_.data() + g.offset);
        } else {
          g.pfft.set_stream(s);
          if (DIM == 2 && fused_ && g.pfft.y_lds_rows() > 0) {
            // twiddles up to the largest n1 exist for the fused type 1
            g.pfft.backward_lds(tw_.data(), nullptr, ins, st, fs, order_fft(),
                                grids_.data() + g.offset, s);
          } else {
            g.pfft.backward(nullptr, ins, st, fs, order_fft(), grids_.data() + g.offset, s);
          }
        }
      }
      launch_interp_batch<T, DIM>(ks_.ns, ijobs_.data(), subs_.data(), nsub_, max_lds_, ks_.es, s);
    }
  }

  void run(int type) {
    const int k = type - 1;
    if (use_graph_ && !graph_failed_[k] && !graph_ok_[k]) capture(type);
    if (graph_ok_[k]) {
      check_hip(hipGraphLaunch(exec_[k], stream_), "hipGraphLaunch");
      return;
    }
    enqueue(type, stream_);
  }

  void capture(int type) {
    const int k = type - 1;
    hipGraph_t graph = nullptr;
    // the capture runs on a private stream; the replay on the plan's stream
    check_hip(hipStreamBeginCapture(capture_stream_, hipStreamCaptureModeThreadLocal),
              "hipStreamBeginCapture");
    bool ok = true;
    try {
      enqueue(type, capture_stream_);
    } catch (...) {
      ok = false;
    }
    const hipError_t end = hipStreamEndCapture(capture_stream_, &graph);
    (void)hipGetLastError();
    for (Group& g : groups_) {
      if constexpr (DIM == 1) {
        g.fft.set_stream(stream_);
      } else {
        g.pfft.set_stream(stream_);
      }
    }
    if (ok && end == hipSuccess && graph) {
      if (hipGraphInstantiate(&exec_[k], graph, nullptr, nullptr, 0) == hipSuccess) {
        graph_ok_[k] = true;
      } else {
        (void)hipGetLastError();
        exec_[k] = nullptr;
      }
    }
    if (graph) (void)hipGraphDestroy(graph);
    if (!graph_ok_[k]) graph_failed_[k] = true;
  }

  // Gather workgroups: all chunks of a transform at one residue mod 8 of the
  // workgroup index (one XCD), transforms placed largest first on the XCD
  // with the fewest points so far, so each transform's random strength reads
  // stay in one L2 and the XCDs get about equal shares.
  void build_gather_chunks(const IntType* num_nu) {
    constexpr int kXcds = 8;
    std::vector<int> order(static_cast<std::size_t>(count_));
    for (int t = 0; t < count_; ++t) order[static_cast<std::size_t>(t)] = t;
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return num_nu[a] > num_nu[b]; });
    std::vector<std::vector<GatherChunk>> lists(kXcds);
    std::size_t load[kXcds] = {};
    for (int t : order) {
      const int x = static_cast<int>(std::min_element(load, load + kXcds) - load);
      const std::size_t m = static_cast<std::size_t>(num_nu[t]);
      for (std::size_t first = 0; first < m; first += kGatherChunk) {
        lists[static_cast<std::size_t>(x)].push_back(
            GatherChunk{static_cast<uint32_t>(t), static_cast<uint32_t>(first)});
      }
      load[x] += m;
    }
    std::size_t len = 0;
    for (const auto& l : lists) len = std::max(len, l.size());
    gather_chunks_host_.assign(len * kXcds, GatherChunk{0xffffffffu, 0u});
    for (int x = 0; x < kXcds; ++x) {
      const auto& l = lists[static_cast<std::size_t>(x)];
      for (std::size_t k = 0; k < l.size(); ++k) gather_chunks_host_[k * kXcds + x] = l[k];
    }
    n_gather_chunks_ = gather_chunks_host_.size();
    gather_chunks_.ensure(n_gather_chunks_ ? n_gather_chunks_ : 1, alloc_);
    if (n_gather_chunks_) {
      check_hip(hipMemcpyAsync(gather_chunks_.data(), gather_chunks_host_.data(),
                               n_gather_chunks_ * sizeof(GatherChunk), hipMemcpyHostToDevice,
                               stream_),
                "hipMemcpyAsync(gather chunks)");
    }
  }

  void drop_graphs() {
    for (int k = 0; k < 2; ++k) {
      if (exec_[k]) (void)hipGraphExecDestroy(exec_[k]);
      exec_[k] = nullptr;
      graph_ok_[k] = false;
      graph_failed_[k] = false;
    }
  }

  neonufft::Options opt_;
  int sign_;
  hipStream_t stream_;
  std::shared_ptr<neonufft::Allocator> alloc_;
  PlanTuning tune_;
  bool use_graph_;
  KernelSetup ks_;
  int count_ = 0;
  std::vector<Item> items_;
  std::vector<Group> groups_;
  DeviceBuffer<Cplx<T>> grids_;
  std::vector<PointSet<T, DIM>> pts_;
  DeviceBuffer<Subprob> subs_;
  std::size_t nsub_ = 0, total_points_ = 0, max_lds_ = 0;
  int max_n0_ = 0, max_rows_ = 0;
  std::size_t max_bins_ = 0;
  bool regular_ = true;  // every group's bins suit the bin-owned reduction
  bool bands_ = true;    // 2D: row-band spreading for every group
  bool bands_on_ = false;  // ... used for the current points (not too clustered)
  std::vector<std::vector<uint32_t>> row_host_;  // row offsets per transform (balance check)
  bool fused_ = true;    // ... with the x-FFT inside the band kernel
  IntType tw_len_ = 1;
  DeviceBuffer<Cplx<T>> tw_;
  DeviceBuffer<unsigned long long> bounds_;  // per transform: row sums + row maxima of |c|
  std::size_t total_bands_ = 0, band_lds_ = 0;
  std::vector<RowView<T>> rows_;  // row-sorted views (row bands)
  DeviceBuffer<Cplx<T>> sorted_;  // strengths of all transforms in view order
  DeviceBuffer<uint32_t> band_job_;      // transform of every row band (row bands)
  std::vector<uint32_t> band_job_host_;  // its host copy (source of the upload)
  DeviceBuffer<GatherChunk> gather_chunks_;      // gather workgroup ranges (row bands)
  std::vector<GatherChunk> gather_chunks_host_;  // their host copy
  std::size_t n_gather_chunks_ = 0;
  DeviceBuffer<BandJob<T>> bjobs_;
  std::vector<BandJob<T>> host_bjobs_;
  TableUpload up_bjobs_;
  DeviceBuffer<Cplx<T>> tiles_;  // spreading tiles of all transforms
  DeviceBuffer<ReduceJob<T, DIM>> rjobs_;
  std::vector<ReduceJob<T, DIM>> host_rjobs_;
  DeviceBuffer<SpreadJob<T, DIM>> sjobs_;
  DeviceBuffer<InterpJob<T, DIM>> ijobs_;
  DeviceBuffer<void*> io_;  // outputs of type 1, inputs of type 2
  std::vector<SpreadJob<T, DIM>> host_sjobs_;
  std::vector<InterpJob<T, DIM>> host_ijobs_;
  std::vector<void*> host_io_;
  TableUpload up_sjobs_, up_ijobs_, up_io_;
  hipStream_t capture_stream_ = nullptr;
  hipGraphExec_t exec_[2] = {nullptr, nullptr};
  bool graph_ok_[2] = {false, false};
  bool graph_failed_[2] = {false, false};
};

}  // namespace nnf
