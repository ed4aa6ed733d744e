// A set of non-uniform points folded onto a grid and sorted by bin, with the
// subproblem table the spreading / interpolation kernels consume.
#pragma once

#include <memory>
#include <vector>

#include "kernels.hpp"

namespace nnf {

template <typename T, int DIM>
class PointSet {
public:
  // Folds, bins and sorts m points (loc[d][i]); synchronises `stream` once to
  // size the subproblem table.
  //
  // ws == null: every buffer is owned (grown as needed). ws != null: nothing
  // is allocated. The temporaries (bin counts and starts, subproblem counts,
  // scan partials, per-point ranks) come from *ws, and the results go to the
  // buffers an earlier allocating build left; a point count or subproblem
  // count beyond their capacity throws InputError before anything is written
  // to them.
  template <typename TC>
  void build(Loc<TC> loc, std::size_t m, const FoldParams& fp, const BinGeom& g, int ns,
             uint32_t max_points, hipStream_t stream, const std::shared_ptr<neonufft::Allocator>& alloc,
             Workspace* ws = nullptr) {
    if (m > 0xffffffffull) {
      throw neonufft::NotImplementedError("Neonufft:: more than 2^32 - 1 points per plan");
    }
    if (m > 0) {
      for (int d = 0; d < DIM; ++d)
        if (!loc.p[d]) throw neonufft::InputError("Neonufft:: null coordinate pointer");
    }
    if (max_points == 0 || max_points > kMaxSubproblemPoints) max_points = kMaxSubproblemPoints;
    const std::size_t nbins = static_cast<std::size_t>(g.nbins[0]) * g.nbins[1] * g.nbins[2];
    Temps t;
    if (ws) {
      if (recs_.size() < std::max<std::size_t>(m, 1) || perm_.size() < std::max<std::size_t>(m, 1) ||
          sub_start_.size() < nbins + 1) {
        throw neonufft::InputError(
            "Neonufft:: more points than the plan holds memory for (set_points without "
            "allocation)");
      }
      t = carve(*ws, m, nbins);
    } else {
      count_.ensure(nbins + 1, alloc);
      start_.ensure(nbins + 1, alloc);
      nsubb_.ensure(nbins + 1, alloc);
      sub_start_.ensure(nbins + 1, alloc);
      scratch_.ensure(scan_scratch_size(nbins + 1), alloc);
      recs_.ensure(m ? m : 1, alloc);
      perm_.ensure(m ? m : 1, alloc);
      rank_.ensure(m ? m : 1, alloc);
      t = Temps{count_.data(), start_.data(), nsubb_.data(), sub_start_.data(), scratch_.data(),
                rank_.data()};
    }
    g_ = g;
    m_ = m;
    max_points_ = max_points;
    nbins_ = nbins;
    nsub_ = 0;

    launch_fill_u32(t.count, nbins + 1, 0u, stream);
    launch_fill_u32(t.nsubb, nbins + 1, 0u, stream);
    launch_bin_count<TC, DIM>(loc, m, fp, g, ns, t.count, t.rank, stream);
    launch_exclusive_scan(t.count, t.start, nbins + 1, t.scan, stream);
    launch_subprob_counts(t.count, nbins, max_points, t.nsubb, stream);
    launch_exclusive_scan(t.nsubb, t.sub_start, nbins + 1, t.scan, stream);
    uint32_t total = 0;
    check_hip(hipMemcpyAsync(&total, t.sub_start + nbins, sizeof(uint32_t), hipMemcpyDeviceToHost,
                             stream),
              "hipMemcpyAsync(subproblem total)");
    check_hip(hipStreamSynchronize(stream), "hipStreamSynchronize");
    if (ws) {
      if (subs_.size() < std::max<uint32_t>(total, 1)) {
        m_ = 0;
        throw neonufft::InputError(
            "Neonufft:: more subproblems than the plan holds memory for (set_points without "
            "allocation)");
      }
    } else {
      // room for any set of m points (each bin adds at most one partial
      // subproblem), so a later build without allocation fits whenever its
      // point count does
      subs_.ensure(nbins + m / max_points + 1, alloc);
    }
    nsub_ = total;
    launch_subprob_build(t.count, t.start, t.sub_start, nbins, max_points, subs_.data(), stream);
    launch_bin_scatter<TC, T, DIM>(loc, m, fp, g, ns, t.start, t.rank, recs_.data(), perm_.data(),
                                   stream);
    launch_bank_order<T, DIM>(subs_.data(), nsub_, recs_.data(), perm_.data(), g, stream);
    if (ws) {
      check_hip(hipMemcpyAsync(sub_start_.data(), t.sub_start, (nbins + 1) * sizeof(uint32_t),
                               hipMemcpyDeviceToDevice, stream),
                "hipMemcpyAsync(subproblem starts)");
    }
  }

  // Points per bin (nbins values) after an allocating build; synchronises.
  std::vector<uint32_t> host_bin_counts(hipStream_t stream) const {
    std::vector<uint32_t> h(nbins_);
    if (nbins_) {
      check_hip(hipMemcpyAsync(h.data(), count_.data(), nbins_ * sizeof(uint32_t),
                               hipMemcpyDeviceToHost, stream),
                "hipMemcpyAsync(bin counts)");
      check_hip(hipStreamSynchronize(stream), "hipStreamSynchronize");
    }
    return h;
  }

  // Workspace bytes a build without allocation takes for m points over nbins bins.
  static std::size_t scratch_bytes(std::size_t m, std::size_t nbins) {
    Workspace dry;
    carve(dry, m, nbins);
    return dry.used;
  }

  const PointRec<T, DIM>* recs() const { return recs_.data(); }
  const uint32_t* perm() const { return perm_.data(); }
  const Subprob* subs() const { return subs_.data(); }
  // Per bin, nbins + 1 entries: the subproblems of bin b are
  // [sub_starts()[b], sub_starts()[b + 1]) in subs().
  const uint32_t* sub_starts() const { return sub_start_.data(); }
  std::size_t nsub() const { return nsub_; }
  std::size_t size() const { return m_; }
  std::size_t nbins() const { return nbins_; }
  uint32_t max_points() const { return max_points_; }
  const BinGeom& geom() const { return g_; }

private:
  struct Temps {
    uint32_t *count, *start, *nsubb, *sub_start, *scan, *rank;
  };
  static Temps carve(Workspace& ws, std::size_t m, std::size_t nbins) {
    Temps t;
    t.count = ws.take<uint32_t>(nbins + 1);
    t.start = ws.take<uint32_t>(nbins + 1);
    t.nsubb = ws.take<uint32_t>(nbins + 1);
    t.sub_start = ws.take<uint32_t>(nbins + 1);
    t.scan = ws.take<uint32_t>(scan_scratch_size(nbins + 1));
    t.rank = ws.take<uint32_t>(m ? m : 1);
    return t;
  }

  BinGeom g_{};
  std::size_t m_ = 0, nsub_ = 0, nbins_ = 0;
  uint32_t max_points_ = 0;
  DeviceBuffer<PointRec<T, DIM>> recs_;
  DeviceBuffer<uint32_t> perm_, rank_, count_, start_, nsubb_, sub_start_, scratch_;
  DeviceBuffer<Subprob> subs_;
};

// Row-sorted view of a 2D point set for the row-band spreading: the records
// sorted by their first kernel row (counting sort over the n1 rows, built
// from the tile-sorted records on the device, no host synchronisation), then
// bank-ordered inside chunks of at most kMaxSubproblemPoints points of a row
// (consecutive points start on distinct cells mod 32, so the lanes of a wave
// hit distinct LDS banks).
template <typename T>
class RowView {
public:
  // ws as in PointSet::build: null = owned buffers; otherwise the
  // temporaries (row counts, chunk counts and table, scan partials, ranks)
  // come from *ws and the results go to the buffers of an earlier
  // allocating build.
  void build(const PointSet<T, 2>& ps, int ns, hipStream_t stream,
             const std::shared_ptr<neonufft::Allocator>& alloc, Workspace* ws = nullptr) {
    const BinGeom& g = ps.geom();
    const std::size_t n1 = static_cast<std::size_t>(g.n[1]);
    const std::size_t m = ps.size();
    const std::size_t max_chunks = chunk_bound(m, n1);
    Temps t;
    if (ws) {
      if (recs_.size() < std::max<std::size_t>(m, 1) || perm_.size() < std::max<std::size_t>(m, 1) ||
          vrow_.size() < std::max<std::size_t>(m, 1) || start_.size() < n1 + 1 ||
          bcount_.size() < n1 * kBoundBlocks || wmax_.size() < 1) {
        throw neonufft::InputError(
            "Neonufft:: more points than the plan holds memory for (set_points without "
            "allocation)");
      }
      t = carve(*ws, m, n1);
    } else {
      count_.ensure(n1 + 1, alloc);
      start_.ensure(n1 + 1, alloc);
      nchunk_.ensure(n1 + 1, alloc);
      chunk_start_.ensure(n1 + 1, alloc);
      scratch_.ensure(scan_scratch_size(n1 + 1), alloc);
      rank_.ensure(m ? m : 1, alloc);
      recs_.ensure(m ? m : 1, alloc);
      perm_.ensure(m ? m : 1, alloc);
      vrow_.ensure(m ? m : 1, alloc);
      chunks_.ensure(max_chunks, alloc);
      bcount_.ensure(n1 * kBoundBlocks, alloc);
      wmax_.ensure(1, alloc);
      t = Temps{count_.data(), nchunk_.data(), chunk_start_.data(), scratch_.data(), rank_.data(),
                chunks_.data()};
    }
    m_ = m;
    static_assert(sizeof(Subprob) % sizeof(uint32_t) == 0, "Subprob is a whole number of words");
    launch_fill_u32(t.count, n1 + 1, 0u, stream);
    launch_fill_u32(reinterpret_cast<uint32_t*>(t.chunks),
                    max_chunks * (sizeof(Subprob) / sizeof(uint32_t)), 0u, stream);
    launch_row_count<T>(ps.recs(), m, g, ns, t.count, t.rank, stream);
    launch_exclusive_scan(t.count, start_.data(), n1 + 1, t.scan, stream);
    launch_row_scatter<T>(ps.recs(), ps.perm(), m, g, ns, start_.data(), t.rank, recs_.data(),
                          perm_.data(), vrow_.data(), stream);
    launch_subprob_counts(t.count, n1 + 1, kMaxSubproblemPoints, t.nchunk, stream);
    launch_exclusive_scan(t.nchunk, t.chunk_start, n1 + 1, t.scan, stream);
    launch_subprob_build(t.count, start_.data(), t.chunk_start, n1, kMaxSubproblemPoints, t.chunks,
                         stream);
    // residue of a view record = (i0x + ns/2) mod 32 (tile pitch 0 drops the row)
    BinGeom og{};
    og.tile[0] = 0;
    og.tile[1] = 0;
    og.tile[2] = 1;
    launch_bank_order<T, 2>(t.chunks, max_chunks, recs_.data(), perm_.data(), og, stream);
    launch_fill_u32(bcount_.data(), n1 * kBoundBlocks, 0u, stream);
    launch_row_block_count<T>(recs_.data(), m, g.n[0], ns, bcount_.data(), stream);
    launch_fill_u32(wmax_.data(), 1, 0u, stream);
    launch_row_window_max(bcount_.data(), static_cast<int>(n1), ns, wmax_.data(), stream);
  }

  // Workspace bytes a build without allocation takes for m points over n1 rows.
  static std::size_t scratch_bytes(std::size_t m, std::size_t n1) {
    Workspace dry;
    carve(dry, m, n1);
    return dry.used;
  }

  const PointRec<T, 2>* recs() const { return recs_.data(); }
  const uint32_t* perm() const { return perm_.data(); }
  const uint32_t* row_start() const { return start_.data(); }
  // first kernel row of each view point (bank order keeps points in their row)
  const uint16_t* vrow() const { return vrow_.data(); }
  const uint32_t* block_counts() const { return bcount_.data(); }
  // device: the largest point count any cell's fixed-point bound sees
  const uint32_t* window_max() const { return wmax_.data(); }
  std::size_t size() const { return m_; }

private:
  struct Temps {
    uint32_t *count, *nchunk, *chunk_start, *scan, *rank;
    Subprob* chunks;
  };
  // upper bound on the chunks: one partial chunk per row plus the full ones
  static std::size_t chunk_bound(std::size_t m, std::size_t n1) {
    return n1 + m / kMaxSubproblemPoints + 1;
  }
  static Temps carve(Workspace& ws, std::size_t m, std::size_t n1) {
    Temps t;
    t.count = ws.take<uint32_t>(n1 + 1);
    t.nchunk = ws.take<uint32_t>(n1 + 1);
    t.chunk_start = ws.take<uint32_t>(n1 + 1);
    t.scan = ws.take<uint32_t>(scan_scratch_size(n1 + 1));
    t.rank = ws.take<uint32_t>(m ? m : 1);
    t.chunks = ws.take<Subprob>(chunk_bound(m, n1));
    return t;
  }

  std::size_t m_ = 0;
  DeviceBuffer<PointRec<T, 2>> recs_;
  DeviceBuffer<uint32_t> perm_, rank_, count_, start_, nchunk_, chunk_start_, scratch_, bcount_,
      wmax_;
  DeviceBuffer<uint16_t> vrow_;
  DeviceBuffer<Subprob> chunks_;
};

}  // namespace nnf
