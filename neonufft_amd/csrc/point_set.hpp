// A set of non-uniform points folded onto a grid and sorted by bin, with the
// subproblem table the spreading / interpolation kernels consume.
#pragma once

#include <memory>

#include "kernels.hpp"

namespace nnf {

template <typename T, int DIM>
class PointSet {
public:
  // Folds, bins and sorts m points (loc[d][i]); synchronises `stream` once to
  // size the subproblem table.
  template <typename TC>
  void build(Loc<TC> loc, std::size_t m, const FoldParams& fp, const BinGeom& g, int ns,
             uint32_t max_points, hipStream_t stream, const std::shared_ptr<neonufft::Allocator>& alloc) {
    if (m > 0xffffffffull) {
      throw neonufft::NotImplementedError("Neonufft:: more than 2^32 - 1 points per plan");
    }
    if (m > 0) {
      for (int d = 0; d < DIM; ++d)
        if (!loc.p[d]) throw neonufft::InputError("Neonufft:: null coordinate pointer");
    }
    g_ = g;
    m_ = m;
    if (max_points == 0 || max_points > kMaxSubproblemPoints) max_points = kMaxSubproblemPoints;
    max_points_ = max_points;
    const std::size_t nbins = static_cast<std::size_t>(g.nbins[0]) * g.nbins[1] * g.nbins[2];
    nbins_ = nbins;
    count_.ensure(nbins + 1, alloc);
    start_.ensure(nbins + 1, alloc);
    nsubb_.ensure(nbins + 1, alloc);
    sub_start_.ensure(nbins + 1, alloc);
    scratch_.ensure(scan_scratch_size(nbins + 1), alloc);
    recs_.ensure(m ? m : 1, alloc);
    perm_.ensure(m ? m : 1, alloc);
    rank_.ensure(m ? m : 1, alloc);

    check_hip(hipMemsetAsync(count_.data(), 0, (nbins + 1) * sizeof(uint32_t), stream),
              "hipMemsetAsync(bin counts)");
    check_hip(hipMemsetAsync(nsubb_.data(), 0, (nbins + 1) * sizeof(uint32_t), stream),
              "hipMemsetAsync(subproblem counts)");
    launch_bin_count<TC, DIM>(loc, m, fp, g, ns, count_.data(), rank_.data(), stream);
    launch_exclusive_scan(count_.data(), start_.data(), nbins + 1, scratch_.data(), stream);
    launch_subprob_counts(count_.data(), nbins, max_points, nsubb_.data(), stream);
    launch_exclusive_scan(nsubb_.data(), sub_start_.data(), nbins + 1, scratch_.data(), stream);
    uint32_t total = 0;
    check_hip(hipMemcpyAsync(&total, sub_start_.data() + nbins, sizeof(uint32_t),
                             hipMemcpyDeviceToHost, stream),
              "hipMemcpyAsync(subproblem total)");
    check_hip(hipStreamSynchronize(stream), "hipStreamSynchronize");
    nsub_ = total;
    subs_.ensure(nsub_ ? nsub_ : 1, alloc);
    launch_subprob_build(count_.data(), start_.data(), sub_start_.data(), nbins, max_points,
                         subs_.data(), stream);
    launch_bin_scatter<TC, T, DIM>(loc, m, fp, g, ns, start_.data(), rank_.data(), recs_.data(),
                               perm_.data(), stream);
    launch_bank_order<T, DIM>(subs_.data(), nsub_, recs_.data(), perm_.data(), g, stream);
  }

  const PointRec<T, DIM>* recs() const { return recs_.data(); }
  const uint32_t* perm() const { return perm_.data(); }
  const Subprob* subs() const { return subs_.data(); }
  const uint32_t* bin_counts() const { return count_.data(); }
  const uint32_t* bin_starts() const { return start_.data(); }
  // Per bin, nbins + 1 entries: the subproblems of bin b are
  // [sub_starts()[b], sub_starts()[b + 1]) in subs().
  const uint32_t* sub_starts() const { return sub_start_.data(); }
  std::size_t nsub() const { return nsub_; }
  std::size_t size() const { return m_; }
  std::size_t nbins() const { return nbins_; }
  uint32_t max_points() const { return max_points_; }
  const BinGeom& geom() const { return g_; }

private:
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
  void build(const PointSet<T, 2>& ps, int ns, hipStream_t stream,
             const std::shared_ptr<neonufft::Allocator>& alloc) {
    const BinGeom& g = ps.geom();
    const std::size_t n1 = static_cast<std::size_t>(g.n[1]);
    const std::size_t m = ps.size();
    m_ = m;
    count_.ensure(n1 + 1, alloc);
    start_.ensure(n1 + 1, alloc);
    nchunk_.ensure(n1 + 1, alloc);
    chunk_start_.ensure(n1 + 1, alloc);
    scratch_.ensure(scan_scratch_size(n1 + 1), alloc);
    rank_.ensure(m ? m : 1, alloc);
    recs_.ensure(m ? m : 1, alloc);
    perm_.ensure(m ? m : 1, alloc);
    vrow_.ensure(m ? m : 1, alloc);
    // upper bound on the chunks: one partial chunk per row plus the full ones
    const std::size_t max_chunks = n1 + m / kMaxSubproblemPoints + 1;
    chunks_.ensure(max_chunks, alloc);
    check_hip(hipMemsetAsync(count_.data(), 0, (n1 + 1) * sizeof(uint32_t), stream),
              "hipMemsetAsync(row counts)");
    check_hip(hipMemsetAsync(chunks_.data(), 0, max_chunks * sizeof(Subprob), stream),
              "hipMemsetAsync(row chunks)");
    launch_row_count<T>(ps.recs(), m, g, ns, count_.data(), rank_.data(), stream);
    launch_exclusive_scan(count_.data(), start_.data(), n1 + 1, scratch_.data(), stream);
    launch_row_scatter<T>(ps.recs(), ps.perm(), m, g, ns, start_.data(), rank_.data(),
                          recs_.data(), perm_.data(), vrow_.data(), stream);
    launch_subprob_counts(count_.data(), n1 + 1, kMaxSubproblemPoints, nchunk_.data(), stream);
    launch_exclusive_scan(nchunk_.data(), chunk_start_.data(), n1 + 1, scratch_.data(), stream);
    launch_subprob_build(count_.data(), start_.data(), chunk_start_.data(), n1,
                         kMaxSubproblemPoints, chunks_.data(), stream);
    // residue of a view record = (i0x + ns/2) mod 32 (tile pitch 0 drops the row)
    BinGeom og{};
    og.tile[0] = 0;
    og.tile[1] = 0;
    og.tile[2] = 1;
    launch_bank_order<T, 2>(chunks_.data(), max_chunks, recs_.data(), perm_.data(), og, stream);
    bcount_.ensure(n1 * kBoundBlocks, alloc);
    check_hip(hipMemsetAsync(bcount_.data(), 0, n1 * kBoundBlocks * sizeof(uint32_t), stream),
              "hipMemsetAsync(block counts)");
    launch_row_block_count<T>(recs_.data(), m, g.n[0], ns, bcount_.data(), stream);
    wmax_.ensure(1, alloc);
    check_hip(hipMemsetAsync(wmax_.data(), 0, sizeof(uint32_t), stream), "hipMemsetAsync(window max)");
    launch_row_window_max(bcount_.data(), static_cast<int>(n1), ns, wmax_.data(), stream);
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
  std::size_t m_ = 0;
  DeviceBuffer<PointRec<T, 2>> recs_;
  DeviceBuffer<uint32_t> perm_, rank_, count_, start_, nchunk_, chunk_start_, scratch_, bcount_,
      wmax_;
  DeviceBuffer<uint16_t> vrow_;
  DeviceBuffer<Subprob> chunks_;
};

}  // namespace nnf
