// Host-side launchers of the HIP kernels (defined in the *.hip files with
// explicit instantiations for T in {float, double}, DIM in {1, 2, 3}).
#pragma once

#include "common.hpp"

namespace nnf {

template <typename T>
struct Loc {
  const T* p[3];
};

// Quadrature for the Fourier transform of the ES kernel:
// phihat(f) = sum_q weight[q] * cos(f * node[q]), with node = z * ns/2 and
// weight = ns * w * phi(node) for the positive half (z, w) of a 2*nq point
// Gauss-Legendre rule (phi is even). Built on the host in FP64.
constexpr int kMaxQuad = 32;
struct QuadRule {
  double node[kMaxQuad];
  double weight[kMaxQuad];
  int nq;
};

// Upper bound on the points of one subproblem (one workgroup's worth).
constexpr uint32_t kMaxSubproblemPoints = 2048;

// ---- points.hip: fold / rescale, counting sort by bin, subproblems -------
template <typename TC, int DIM>
void launch_bin_count(Loc<TC> loc, std::size_t m, const FoldParams& fp, const BinGeom& g, int ns,
                      uint32_t* count, uint32_t* rank, hipStream_t stream);

// TC: coordinate type read; T: precision of the point records written
// (TC = T, or TC = double with T = float for mixed-precision type 3).
template <typename TC, typename T, int DIM>
void launch_bin_scatter(Loc<TC> loc, std::size_t m, const FoldParams& fp, const BinGeom& g,
                        int ns, const uint32_t* bin_start, const uint32_t* rank,
                        PointRec<T, DIM>* pts, uint32_t* perm, hipStream_t stream);

// Exclusive prefix sum of n uint32 values; `scratch` must hold
// scan_scratch_size(n) values.
std::size_t scan_scratch_size(std::size_t n);
void launch_exclusive_scan(const uint32_t* in, uint32_t* out, std::size_t n, uint32_t* scratch,
                           hipStream_t stream);

// p[0..n) = v (a kernel, not hipMemsetAsync, so graphs may record it).
void launch_fill_u32(uint32_t* p, std::size_t n, uint32_t v, hipStream_t stream);

void launch_subprob_counts(const uint32_t* count, std::size_t nbins, uint32_t max_points,
                           uint32_t* nsub, hipStream_t stream);
void launch_subprob_build(const uint32_t* count, const uint32_t* bin_start,
                          const uint32_t* sub_start, std::size_t nbins, uint32_t max_points,
                          Subprob* subs, hipStream_t stream);

// Reorders the points inside every subproblem for conflict-free LDS access
// (bank-balanced waves); needs the subproblem table.
template <typename T, int DIM>
void launch_bank_order(const Subprob* subs, std::size_t nsub, PointRec<T, DIM>* pts,
                       uint32_t* perm, const BinGeom& g, hipStream_t stream);

// Per-block partial min/max of x (as double): partial[2*b] = min, [2*b+1] = max.
int minmax_blocks(std::size_t m);
template <typename T>
void launch_minmax(const T* x, std::size_t m, double* partial, hipStream_t stream);

// ---- spread_interp.hip: ES spreading / interpolation on binned points -----
// Cells of one subgrid tile (bin + ns halo), the pitch of every stored tile.
__host__ __device__ inline std::size_t tile_cells(const BinGeom& g) {
  return static_cast<std::size_t>(g.tile[0]) * g.tile[1] * g.tile[2];
}

// Spreading reads strength in[perm[p]] for sorted point p (in[p] when perm is
// null, i.e. strengths already in bin order) and writes the whole tile of
// subproblem s (tile_cells(g) cells, x
// fastest) to tiles + s * tile_cells(g); launch_tile_reduce then forms the
// fine grid from the tiles.
template <typename T, int DIM>
void launch_spread(int ns, const PointRec<T, DIM>* pts, const uint32_t* perm, const Subprob* subs,
                   std::size_t nsub, const Cplx<T>* in, const Cplx<T>* pre_sorted, Cplx<T>* tiles,
                   const BinGeom& g, const EsParams& es, hipStream_t stream);

// Fine-grid cell -> covering tiles, one table per dimension: entries
// [start[d][c], start[d][c + 1]) of entry[d] hold (bin coordinate, position
// inside that bin's tile) for cell c of dimension d (built on the host).
struct CoverTable {
  const int* start[3];
  const int2* entry[3];
};

// Everything the tile reduction needs for one fine grid.
template <typename T, int DIM>
struct ReduceJob {
  const Cplx<T>* tiles;       // tile of (local) subproblem s at tiles + s * tile_cells(g)
  const uint32_t* sub_start;  // nbins + 1: subproblems of bin b are [sub_start[b], sub_start[b+1])
  Cplx<T>* grid;
  BinGeom g;
  CoverTable cover;
};

// True when, in every dimension, a cell is covered only by the tiles of its
// own bin and its two neighbours: bin >= ns, at least two bins, and the last
// bin full or at least ns wide. Such geometries use the bin-owned reduction.
__host__ __device__ inline bool bins_regular(int dim, const BinGeom& g, int ns) {
  for (int d = 0; d < dim; ++d) {
    const int rem = g.n[d] % g.bin[d];
    if (g.nbins[d] < 2 || g.bin[d] < ns || (rem != 0 && rem < ns)) return false;
  }
  return true;
}

// grid[c] (+)= sum over the tiles covering cell c; every cell is written, so
// no zero fill is needed before a non-accumulating spread.
template <typename T, int DIM>
void launch_tile_reduce(const ReduceJob<T, DIM>& job, int ns, bool accumulate,
                        hipStream_t stream);
// One reduction per job (device table of `count` jobs) in a single launch.
// regular: every job's geometry satisfies bins_regular; max_bins / max_n0 /
// max_rows: the largest bin count, x extent and row count among the grids.
template <typename T, int DIM>
void launch_tile_reduce_batch(const ReduceJob<T, DIM>* jobs, int count, int ns, bool regular,
                              std::size_t max_bins, int max_n0, int max_rows, bool accumulate,
                              hipStream_t stream);

template <typename T, int DIM>
void launch_interp(int ns, const PointRec<T, DIM>* pts, const uint32_t* perm, const Subprob* subs,
                   std::size_t nsub, const Cplx<T>* grid, const Cplx<T>* post_sorted, Cplx<T>* out,
                   const BinGeom& g, const EsParams& es, hipStream_t stream);

// Per-transform descriptors of the batched kernels (one subproblem table
// across many independent transforms; Subprob::job indexes these).
template <typename T, int DIM>
struct SpreadJob {
  const PointRec<T, DIM>* pts;
  const uint32_t* perm;  // null: `in` is already in bin order
  const Cplx<T>* in;
  const Cplx<T>* pre;  // bin-order prephase or null
  Cplx<T>* tiles;      // this transform's tiles, indexed by subproblem - sub_off
  uint32_t sub_off;    // first subproblem of this transform in the shared table
  uint32_t pad;        // keeps the struct free of indeterminate padding bytes
  BinGeom g;
};

template <typename T, int DIM>
struct InterpJob {
  const PointRec<T, DIM>* pts;
  const uint32_t* perm;
  const Cplx<T>* grid;
  const Cplx<T>* post;  // bin-order postphase or null
  Cplx<T>* out;
  BinGeom g;
};

template <typename T, int DIM>
void launch_spread_batch(int ns, const SpreadJob<T, DIM>* jobs, const Subprob* subs,
                         std::size_t nsub, std::size_t lds_bytes, const EsParams& es,
                         hipStream_t stream);

template <typename T, int DIM>
void launch_interp_batch(int ns, const InterpJob<T, DIM>* jobs, const Subprob* subs,
                         std::size_t nsub, std::size_t lds_bytes, const EsParams& es,
                         hipStream_t stream);

// Dynamic LDS bytes of one spreading / interpolation block for a bin geometry.
template <typename T>
std::size_t tile_bytes(const BinGeom& g);

// ---- band.hip: row-band spreading (2D type 1) ------------------------------
// Row-sorted view of a tile-sorted 2D point set: records sorted by their first
// kernel row i0y (mod n1), with cell = (i0x + ns/2) | (row << 16) and
// bin = row; perm_out maps the view back to the caller's indices and
// row_out[p] is the first kernel row of view point p (n1 <= 65536).
template <typename T>
void launch_row_count(const PointRec<T, 2>* recs, std::size_t m, const BinGeom& g, int ns,
                      uint32_t* count, uint32_t* rank, hipStream_t stream);
template <typename T>
void launch_row_scatter(const PointRec<T, 2>* recs, const uint32_t* perm, std::size_t m,
                        const BinGeom& g, int ns, const uint32_t* row_start, const uint32_t* rank,
                        PointRec<T, 2>* out, uint32_t* perm_out, uint16_t* row_out,
                        hipStream_t stream);

// One 2D transform of a row-band launch. Band b of the launch belongs to the
// job with the largest band_off <= b and covers rows
// [(b - band_off) * rows, + rows) of its n0 x n1 grid.
template <typename T>
struct BandJob {
  const PointRec<T, 2>* pts;   // row-sorted records
  const uint32_t* perm;        // row-sorted position -> caller's point index
  const uint32_t* row_start;   // n1 + 1 offsets into the view
  const uint16_t* vrow;        // first kernel row of each view point
  const uint32_t* bcount;      // n1 x kBoundBlocks: points per (first row, column block)
  T* rowsum;                   // n1: sum of |Re c| + |Im c| over the points of each first row
  T* rowmax;                   // n1: largest |Re c| + |Im c| among them
  const Cplx<T>* in;           // caller's strengths
  Cplx<T>* sorted;             // strengths in view order (gather output)
  Cplx<T>* grid;               // fine grid, x fastest
  std::size_t m;               // points
  std::size_t pt_off;          // first point of this job in a flattened gather
  std::size_t band_off;        // first band of this job in the launch
  int n0, n1, rows, pitch;     // band rows; LDS row pitch (cells)
  // fused x-FFT (n0 and rows powers of two, n0 >= 32): the retained x modes
  // of every row go to h[x' * n1 + y] (not deconvolved)
  Cplx<T>* h;
  const Cplx<T>* tw;           // tw[N/2 + q] = exp(sign 2 pi i q / N), q < N/2, N <= tw_len
  int tw_len, N0, log2n0, log2rows;
};

// Points of each first kernel row whose kernel x-window meets each of the
// kBoundBlocks column blocks of width max(1, ceil(n0 / kBoundBlocks)).
constexpr int kBoundBlocks = 64;
template <typename T>
void launch_row_block_count(const PointRec<T, 2>* view, std::size_t m, int n0, int ns,
                            uint32_t* bcount, hipStream_t stream);
// *out (zeroed by the caller) = the largest sum of bcount over ns consecutive
// rows (mod n1) of one column block: the most points any cell's bound sees.
void launch_row_window_max(const uint32_t* bcount, int n1, int ns, uint32_t* out,
                           hipStream_t stream);
// Zeroes n 8-byte words of strength bounds (a kernel, so graphs may record
// it). Each job owns band_bound_words(n1, sizeof(T)) words: its n1 row sums,
// then its n1 row maxima.
void launch_clear_bounds(unsigned long long* words, std::size_t n, hipStream_t stream);
inline std::size_t band_bound_words(int n1, std::size_t tbytes) {
  return (2 * static_cast<std::size_t>(n1) * tbytes + 7) / 8;
}

// Strengths into view order (sorted[p] = in[perm[p]]) for every job, and per
// first kernel row the sum and the largest of |Re c| + |Im c| over the row's
// points into the job's (cleared) rowsum / rowmax; `jobs` is a device table
// of njobs entries sorted by pt_off, or null to use `single`.
// total_points = sum of m.
template <typename T>
void launch_band_gather(const BandJob<T>* jobs, const BandJob<T>& single, int njobs,
                        std::size_t total_points, hipStream_t stream);
// Points per workgroup of the gather.
constexpr int kGatherChunk = 2048;
// One gather workgroup's range: points [first, first + kGatherChunk) of job
// `job` (job = 0xffffffff: no work).
struct GatherChunk {
  uint32_t job;
  uint32_t first;
};
// The same gather over the device table `jobs`, one workgroup per entry of
// `chunks`. Workgroup i runs on XCD i mod 8, so a table that places all of a
// job's chunks at one residue mod 8 keeps that job's random strength reads in
// one XCD's L2.
template <typename T>
void launch_band_gather_chunks(const BandJob<T>* jobs, const GatherChunk* chunks,
                               std::size_t nchunks, hipStream_t stream);
// Forms (or adds to) every job's fine grid from its gathered strengths, one
// workgroup per band; with `fft`, transforms each band along x in LDS and
// writes the retained modes to job.h instead (no grid, no accumulate).
// `band_job` (optional, with `jobs`): the job of every band of the launch, so
// a workgroup finds its job with one load instead of a search over `jobs`.
template <typename T>
void launch_band_spread(int ns, const BandJob<T>* jobs, const uint32_t* band_job,
                        const BandJob<T>& single, int njobs, std::size_t total_bands,
                        std::size_t lds_bytes, bool accumulate, bool fft, const EsParams& es,
                        hipStream_t stream);
std::size_t band_lds_bytes(int rows, int pitch, std::size_t acc_bytes);
// LDS row of a band: cells x in [-kBandLeft, pitch - kBandLeft) (the cells
// outside [0, n0) catch kernel runs that cross the row's ends and are folded
// back periodically); pitch = n0 + kBandPad keeps every row's cell 0 32-byte
// aligned for the in-LDS FFT and shifts consecutive rows by 24 banks.
constexpr int kBandLeft = 8;
constexpr int kBandPad = 24;

// ---- rows.hip: x-FFT in LDS joined to the tile reduction / zero padding ---
// Pencils of `rows` (a power of two) consecutive y rows of one z plane with
// the whole x extent n0 (a power of two >= 32); LDS rows of `pitch` cells
// (n0 + 8). h: the pruned FFT's intermediate, h[(z N0 + x') n1 + y] for the
// N0 retained x modes (retained order), not deconvolved.
template <typename T>
struct RowFftJob {
  const Cplx<T>* tiles;       // type 1: spreading tiles, tile_cells(g) per subproblem
  const uint32_t* sub_start;  // type 1: nbins + 1 subproblem offsets per bin
  BinGeom g;                  // fine grid and its bins (type 1: bins_regular)
  Cplx<T>* h;
  Cplx<T>* grid;              // type 2: fine grid written (x fastest)
  const Cplx<T>* tw;          // twiddles as for the row bands (tw[N/2 + q])
  int N0, log2n0, rows, log2rows, pitch;
};
// Entries of a reduction pencil's subproblem-range table: every x bin x the y
// bins its rows meet plus one on each side x three z bins (3D).
inline int row_fft_rng_entries(const BinGeom& g, int rows, int dim) {
  return g.nbins[0] * ((rows - 1) / g.bin[1] + 4) * (dim >= 3 ? 3 : 1);
}
inline std::size_t row_fft_lds_bytes(int rows, int pitch, std::size_t tbytes, int rng_entries) {
  return 2 * static_cast<std::size_t>(rows) * static_cast<std::size_t>(pitch) * tbytes +
         8 * static_cast<std::size_t>(rng_entries);
}
// Type 1: tiles -> h for every pencil of the grid (one workgroup each).
template <typename T, int DIM>
void launch_reduce_rows_fft(const RowFftJob<T>& job, int ns, hipStream_t stream);
// Type 2: h -> fine grid.
template <typename T>
void launch_expand_rows_ifft(const RowFftJob<T>& job, hipStream_t stream);

// ---- rows.hip: the y and z passes of a 3D pruned FFT in LDS --------------
// For power-of-two n1, n2 >= 32. Between h[(z N0 + x') n1 + y] and
// k[(y' N0 + x') n2 + z] (y' < N1 retained), the passes the pruned FFT runs
// as rocFFT passes joined by transposes, each in one kernel:
//   type 1: h -> y-FFT -> k (retained y'), then k -> z-FFT -> deconvolved
//           modes in the caller's order and strides;
//   type 2: caller's modes -> deconvolved, zero-padded z rows -> z-FFT -> k,
//           then k -> zero-padded y rows -> y-FFT -> h.
// A workgroup transforms `rows` (a power of two) rows of one axis with LDS
// rows of pitch n + 8.
template <typename T>
struct AxisFftJob {
  Cplx<T>* h;
  Cplx<T>* k;
  Cplx<T>* modes_out;         // type 1: caller's modes written
  const Cplx<T>* modes_in;    // type 2: caller's modes read
  const T* fs[3];        // deconvolution factors per axis, by |mode|
  long long st[3];       // caller's mode strides
  int order_fft;
  const Cplx<T>* tw;     // twiddles (tw[N/2 + q]) up to max(n1, n2)
  int N0, N1, N2, n1, n2, log2n1, log2n2;
  int rows_y, log2rows_y, rows_z, log2rows_z;
};
// Rows per workgroup of an axis pass over lines of n cells (0 = too long).
inline int axis_fft_rows(IntType n, std::size_t tbytes, std::size_t budget) {
  for (int rows = 16; rows >= 1; rows /= 2) {
    if (2 * static_cast<std::size_t>(rows) * static_cast<std::size_t>(n + 8) * tbytes <= budget) {
      return rows;
    }
  }
  return 0;
}
// The y pass of 2D plans in LDS. h holds, per transform b (at b *
// member_stride), N0 retained x' lines of n1 cells, y fastest. One workgroup
// transforms `rows` lines (a power of two) along y.
//   type 1 (launch_y_out2): h -> y-FFT -> the retained y' deconvolved into the
//     transform's modes (`single`, or the device table `outs`);
//   type 2 (launch_y_in2): the transform's modes (`single_in` or `ins`)
//     deconvolved into zero-padded y lines -> y-FFT -> h.
// Modes in the caller's order and strides.
template <typename T>
struct Y2Job {
  Cplx<T>* h;
  long long member_stride;
  Cplx<T>* single;
  Cplx<T>* const* outs;
  const Cplx<T>* single_in;
  const Cplx<T>* const* ins;
  const T* fs0;
  const T* fs1;
  long long st0, st1;
  int order_fft;
  const Cplx<T>* tw;  // twiddles (tw[N/2 + q]) up to n1
  int count, N0, N1, n1, log2n1, rows, log2rows;
};
template <typename T>
void launch_y_out2(const Y2Job<T>& job, hipStream_t stream);
template <typename T>
void launch_y_in2(const Y2Job<T>& job, hipStream_t stream);

template <typename T>
void launch_axis_fft_type1(const AxisFftJob<T>& job, hipStream_t stream);
template <typename T>
void launch_axis_fft_type2(const AxisFftJob<T>& job, hipStream_t stream);

// ---- grid.hip: deconvolution factors, up/down-sampling, phases -----------
template <typename T>
void launch_fill_zero(Cplx<T>* p, std::size_t count, hipStream_t stream);

// fs[k] = (-1)^k / phihat(2 pi k / n), k = 0..n/2.
template <typename T>
void launch_fseries(T* fs, IntType n, const QuadRule& rule, hipStream_t stream);

// order_fft: 1 = NEONUFFT_MODE_ORDER_FFT, 0 = CMCL. Strides in elements.
template <typename T, int DIM>
void launch_upsample(int order_fft, const Cplx<T>* in, const IntType* in_strides,
                     const IntType* modes, T* const* fs, Cplx<T>* grid, const IntType* n,
                     hipStream_t stream);

template <typename T, int DIM>
void launch_downsample(int order_fft, const Cplx<T>* grid, const IntType* n, T* const* fs,
                       Cplx<T>* out, const IntType* out_strides, const IntType* modes,
                       hipStream_t stream);

// Batched up/down-sampling over `count` transforms with equal shapes whose fine
// grids are contiguous with stride `gstride` cells; ins / outs are device
// arrays of `count` pointers.
template <typename T, int DIM>
void launch_upsample_batch(int order_fft, const Cplx<T>* const* ins, const IntType* in_strides,
                           const IntType* modes, T* const* fs, Cplx<T>* grids,
                           std::size_t gstride, const IntType* n, int count, hipStream_t stream);

template <typename T, int DIM>
void launch_downsample_batch(int order_fft, const Cplx<T>* grids, std::size_t gstride,
                             const IntType* n, T* const* fs, Cplx<T>* const* outs,
                             const IntType* out_strides, const IntType* modes, int count,
                             hipStream_t stream);

// Copies a subproblem table into a batched table, tagging every entry with `job`.
void launch_subprob_retag(const Subprob* src, std::size_t nsub, uint32_t job, Subprob* dst,
                          hipStream_t stream);

// Type-3 phase factors, written in bin (sorted) order: value[p] belongs to
// point perm[p].
// TC: coordinate type read (phase arguments are formed in FP64 either way).
template <typename TC, typename T, int DIM>
void launch_prephase(int sign, Loc<TC> loc, const uint32_t* perm, std::size_t m,
                     const double* sbar, Cplx<T>* pre_sorted, hipStream_t stream);

template <typename TC, typename T, int DIM>
void launch_postphase(int sign, Loc<TC> loc, const uint32_t* perm, std::size_t m,
                      const double* xbar, const double* sbar, const double* scale,
                      const QuadRule& rule, Cplx<T>* post_sorted, hipStream_t stream);

// out[i] = in[i] converted between complex precisions.
template <typename TI, typename TO>
void launch_convert(const Cplx<TI>* in, Cplx<TO>* out, std::size_t m, hipStream_t stream);

}  // namespace nnf
