// Point fold / rescale and counting sort by subgrid tile (bin).
//
// Restates the behaviour of the reference's rescale_and_permut and rescale_t3
// (src/neonufft/gpu/kernels/rescale_loc_kernel.cu:23-249) with a different
// design: the fold runs in FP64 for every plan precision, the bin offsets come
// from a parallel scan (bins stay in spatial order), the fine-grid coordinate
// is stored split into an integer kernel start cell and an fp32 Horner
// argument, and the bins are cut into bounded subproblems for the spreading /
// interpolation kernels.
#include "chunk_rank.hpp"
#include "kernels.hpp"

namespace nnf {

namespace {

constexpr int kBlock = 256;

inline unsigned grid_for(std::size_t work, int block, unsigned cap = 65535u * 8u) {
  std::size_t g = (work + block - 1) / block;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return static_cast<unsigned>(g);
}

// Fold one coordinate: v = (x - offset) * scale; u = frac((v + pi) / 2 pi);
// xi = u * n. Returns the containing cell (for the bin), the first kernel
// cell i0 = ceil(xi - ns/2) and the Horner argument i0 - (xi - ns/2) in [0,1).
__device__ inline void fold_coord(double v, int n, int ns, double& hx, int& i0, int& cell) {
  double u = (v + kPi) / (2.0 * kPi);
  u -= floor(u);
  if (u >= 1.0) u = 0.0;
  double xi = u * static_cast<double>(n);
  if (xi >= static_cast<double>(n)) xi -= static_cast<double>(n);
  const double a = xi - 0.5 * ns;
  const double s = ceil(a);
  hx = s - a;
  i0 = static_cast<int>(s);
  cell = min(static_cast<int>(floor(xi)), n - 1);
}

template <typename TC, int DIM>
__device__ inline uint32_t point_bin(const Loc<TC>& loc, std::size_t i, const FoldParams& fp,
                                     const BinGeom& g, int ns, double (&hx)[3], int (&rel)[3]) {
  uint32_t bin = 0, mult = 1;
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    const double v = (static_cast<double>(loc.p[d][i]) - fp.offset[d]) * fp.scale[d];
    int i0, cell;
    fold_coord(v, g.n[d], ns, hx[d], i0, cell);
    const int b = cell / g.bin[d];
    rel[d] = i0 - (b * g.bin[d] - ns / 2);
    bin += static_cast<uint32_t>(b) * mult;
    mult *= static_cast<uint32_t>(g.nbins[d]);
  }
  return bin;
}

// Counting pass of the bin sort: rank[i] = position of point i among the
// points of its bin (hashed per-chunk counts, one global atomic per distinct
// bin of a chunk).
template <typename TC, int DIM>
__global__ void __launch_bounds__(kChunkThreads)
    bin_count_kernel(Loc<TC> loc, std::size_t m, FoldParams fp, BinGeom g, int ns,
                     uint32_t* __restrict__ count, uint32_t* __restrict__ rank) {
  chunk_rank(m, count, rank, [&](std::size_t i) {
    double hx[3];
    int rel[3];
    return point_bin<TC, DIM>(loc, i, fp, g, ns, hx, rel);
  });
}

// TC: coordinate type read; T: precision of the point records written.
template <typename TC, typename T, int DIM>
__global__ void __launch_bounds__(kBlock)
    bin_scatter_kernel(Loc<TC> loc, std::size_t m, FoldParams fp, BinGeom g, int ns,
                       const uint32_t* __restrict__ bin_start, const uint32_t* __restrict__ rank,
                       PointRec<T, DIM>* __restrict__ pts, uint32_t* __restrict__ perm) {
  for (std::size_t i = threadIdx.x + static_cast<std::size_t>(blockIdx.x) * kBlock; i < m;
       i += static_cast<std::size_t>(gridDim.x) * kBlock) {
    double hx[3];
    int rel[3];
    const uint32_t bin = point_bin<TC, DIM>(loc, i, fp, g, ns, hx, rel);
    PointRec<T, DIM> rec{};
    uint32_t cell = 0;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      rec.x[d] = static_cast<T>(hx[d]);
      cell |= (static_cast<uint32_t>(rel[d]) & CellPack<DIM>::mask)
              << (CellPack<DIM>::bits == 32 ? 0 : d * CellPack<DIM>::bits);
    }
    rec.cell = cell;
    if constexpr (DIM == 2) rec.bin = bin;
    const uint32_t pos = bin_start[bin] + rank[i];
    pts[pos] = rec;
    perm[pos] = static_cast<uint32_t>(i);
  }
}

// ---- exclusive scan (three passes: tile sums, scan of sums, tile scans) ----
constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;

// Exclusive scan of one value per thread across the block; returns the prefix
// and writes the block total to *total.
__device__ inline uint32_t block_exclusive_scan(uint32_t v, uint32_t* total) {
  __shared__ uint32_t wave_tot[kScanThreads / 64];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  uint32_t incl = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t t = __shfl_up(incl, off, 64);
    if (lane >= off) incl += t;
  }
  if (lane == 63) wave_tot[wave] = incl;
  __syncthreads();
  uint32_t wave_off = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < kScanThreads / 64; ++w) {
    if (w < wave) wave_off += wave_tot[w];
    tot += wave_tot[w];
  }
  __syncthreads();
  *total = tot;
  return wave_off + incl - v;
}

__global__ void __launch_bounds__(kScanThreads)
    scan_tile_sums(const uint32_t* __restrict__ in, std::size_t n, uint32_t* __restrict__ sums) {
  const std::size_t base = static_cast<std::size_t>(blockIdx.x) * kScanTile;
  uint32_t s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    const std::size_t i = base + static_cast<std::size_t>(k) * kScanThreads + threadIdx.x;
    if (i < n) s += in[i];
  }
  uint32_t tot;
  block_exclusive_scan(s, &tot);
  if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

// Single block: in-place exclusive scan of the tile sums, chunk by chunk.
__global__ void __launch_bounds__(kScanThreads)
    scan_sums_kernel(uint32_t* __restrict__ sums, std::size_t nsums) {
  uint32_t carry = 0;
  for (std::size_t base = 0; base < nsums; base += kScanThreads) {
    const std::size_t i = base + threadIdx.x;
    const uint32_t v = i < nsums ? sums[i] : 0u;
    uint32_t tot;
    const uint32_t ex = block_exclusive_scan(v, &tot);
    if (i < nsums) sums[i] = carry + ex;
    carry += tot;
  }
}

__global__ void __launch_bounds__(kScanThreads)
    scan_tiles_kernel(const uint32_t* __restrict__ in, std::size_t n,
                      const uint32_t* __restrict__ tile_off, uint32_t* __restrict__ out) {
  // each thread owns kScanItems consecutive values
  const std::size_t base =
      static_cast<std::size_t>(blockIdx.x) * kScanTile + static_cast<std::size_t>(threadIdx.x) * kScanItems;
  uint32_t v[kScanItems];
  uint32_t s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    v[k] = (base + k < n) ? in[base + k] : 0u;
    s += v[k];
  }
  uint32_t tot;
  uint32_t run = tile_off[blockIdx.x] + block_exclusive_scan(s, &tot);
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    if (base + k < n) out[base + k] = run;
    run += v[k];
  }
}

__global__ void __launch_bounds__(kBlock)
    fill_u32_kernel(uint32_t* __restrict__ p, std::size_t n, uint32_t v) {
  for (std::size_t i = threadIdx.x + static_cast<std::size_t>(blockIdx.x) * kBlock; i < n;
       i += static_cast<std::size_t>(gridDim.x) * kBlock) {
    p[i] = v;
  }
}

__global__ void __launch_bounds__(kBlock)
    subprob_counts_kernel(const uint32_t* __restrict__ count, std::size_t nbins,
                          uint32_t max_points, uint32_t* __restrict__ nsub) {
  for (std::size_t b = threadIdx.x + static_cast<std::size_t>(blockIdx.x) * kBlock; b < nbins;
       b += static_cast<std::size_t>(gridDim.x) * kBlock) {
    nsub[b] = (count[b] + max_points - 1) / max_points;
  }
}

__global__ void __launch_bounds__(kBlock)
    subprob_build_kernel(const uint32_t* __restrict__ count, const uint32_t* __restrict__ bin_start,
                         const uint32_t* __restrict__ sub_start, std::size_t nbins,
                         uint32_t max_points, Subprob* __restrict__ subs) {
  for (std::size_t b = threadIdx.x + static_cast<std::size_t>(blockIdx.x) * kBlock; b < nbins;
       b += static_cast<std::size_t>(gridDim.x) * kBlock) {
    // split the bin into nsub near-equal chunks of at most max_points
    const uint32_t c = count[b];
    const uint32_t s0 = sub_start[b];
    const uint32_t b0 = bin_start[b];
    const uint32_t ns = (c + max_points - 1) / max_points;
    const uint32_t chunk = ns ? (c + ns - 1) / ns : 0;
    for (uint32_t k = 0; k < ns; ++k) {
      Subprob sp;
      sp.bin = static_cast<uint32_t>(b);
      sp.begin = b0 + k * chunk;
      sp.count = min(chunk, c - k * chunk);
      sp.job = 0;
      subs[s0 + k] = sp;
    }
  }
}

// Reorders the points of each subproblem so that every consecutive run of 32
// points has distinct LDS tile base addresses modulo 32 (as far as the counts
// allow): points are sorted by (rank within residue class, residue), where
// residue = tile-linear index of the point's first kernel cell mod 32. The
// spreading / interpolation kernels give consecutive points to consecutive
// lanes and all lanes step through the same kernel-cell offsets, so a wave
// whose bases are spread over the residues touches every LDS bank equally
// at every step instead of piling onto a few banks.
constexpr int kOrderThreads = 1024;
constexpr int kOrderPerThread = (kMaxSubproblemPoints + kOrderThreads - 1) / kOrderThreads;

template <typename T, int DIM>
__device__ inline uint32_t tile_residue(const PointRec<T, DIM>& rec, int t0, int t1) {
  int base = 0;
  if constexpr (CellPack<DIM>::bits == 32) {
    base = static_cast<int>(rec.cell);
  } else {
    const int r0 = static_cast<int>(rec.cell & CellPack<DIM>::mask);
    const int r1 = static_cast<int>((rec.cell >> CellPack<DIM>::bits) & CellPack<DIM>::mask);
    int r2 = 0;
    if constexpr (DIM == 3)
      r2 = static_cast<int>((rec.cell >> (2 * CellPack<DIM>::bits)) & CellPack<DIM>::mask);
    base = (r2 * t1 + r1) * t0 + r0;
  }
  return static_cast<uint32_t>(base) & 31u;
}

template <typename T, int DIM>
__global__ void __launch_bounds__(kOrderThreads)
    bank_order_kernel(const Subprob* __restrict__ subs, PointRec<T, DIM>* __restrict__ pts,
                      uint32_t* __restrict__ perm, int t0, int t1) {
  __shared__ uint32_t cnt[32];
  if (threadIdx.x < 32) cnt[threadIdx.x] = 0;
  __syncthreads();
  const Subprob sp = subs[blockIdx.x];
  PointRec<T, DIM> rec[kOrderPerThread];
  uint32_t idx[kOrderPerThread], res[kOrderPerThread], rank[kOrderPerThread];
#pragma unroll
  for (int k = 0; k < kOrderPerThread; ++k) {
    const uint32_t i = threadIdx.x + k * kOrderThreads;
    if (i < sp.count) {
      rec[k] = pts[sp.begin + i];
      idx[k] = perm[sp.begin + i];
      res[k] = tile_residue<T, DIM>(rec[k], t0, t1);
      rank[k] = atomicAdd(&cnt[res[k]], 1u);
    }
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kOrderPerThread; ++k) {
    const uint32_t i = threadIdx.x + k * kOrderThreads;
    if (i < sp.count) {
      uint32_t pos = 0;
#pragma unroll
      for (uint32_t r = 0; r < 32; ++r) {
        const uint32_t c = cnt[r];
        pos += min(c, rank[k]);
        if (r < res[k] && c > rank[k]) ++pos;
      }
      pts[sp.begin + pos] = rec[k];
      perm[sp.begin + pos] = idx[k];
    }
  }
}

__global__ void __launch_bounds__(kBlock)
    subprob_retag_kernel(const Subprob* __restrict__ src, std::size_t nsub, uint32_t job,
                         Subprob* __restrict__ dst) {
  for (std::size_t i = threadIdx.x + static_cast<std::size_t>(blockIdx.x) * kBlock; i < nsub;
       i += static_cast<std::size_t>(gridDim.x) * kBlock) {
    Subprob s = src[i];
    s.job = job;
    dst[i] = s;
  }
}

constexpr int kMinMaxBlocks = 512;

template <typename T>
__global__ void __launch_bounds__(kBlock)
    minmax_kernel(const T* __restrict__ x, std::size_t m, double* __restrict__ partial) {
  double lo = HUGE_VAL, hi = -HUGE_VAL;
  for (std::size_t i = threadIdx.x + static_cast<std::size_t>(blockIdx.x) * kBlock; i < m;
       i += static_cast<std::size_t>(gridDim.x) * kBlock) {
    const double v = static_cast<double>(x[i]);
    lo = fmin(lo, v);
    hi = fmax(hi, v);
  }
  __shared__ double slo[kBlock], shi[kBlock];
  slo[threadIdx.x] = lo;
  shi[threadIdx.x] = hi;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      slo[threadIdx.x] = fmin(slo[threadIdx.x], slo[threadIdx.x + s]);
      shi[threadIdx.x] = fmax(shi[threadIdx.x], shi[threadIdx.x + s]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    partial[2 * blockIdx.x] = slo[0];
    partial[2 * blockIdx.x + 1] = shi[0];
  }
}

}  // namespace

template <typename TC, int DIM>
void launch_bin_count(Loc<TC> loc, std::size_t m, const FoldParams& fp, const BinGeom& g, int ns,
                      uint32_t* count, uint32_t* rank, hipStream_t stream) {
  if (m == 0) return;
  hipLaunchKernelGGL((bin_count_kernel<TC, DIM>), dim3(chunk_rank_blocks(m)),
                     dim3(kChunkThreads), 0, stream, loc, m, fp, g, ns, count, rank);
  check_launch("bin_count_kernel");
}

template <typename TC, typename T, int DIM>
void launch_bin_scatter(Loc<TC> loc, std::size_t m, const FoldParams& fp, const BinGeom& g,
                        int ns, const uint32_t* bin_start, const uint32_t* rank,
                        PointRec<T, DIM>* pts, uint32_t* perm, hipStream_t stream) {
  if (m == 0) return;
  hipLaunchKernelGGL((bin_scatter_kernel<TC, T, DIM>), dim3(grid_for(m, kBlock)), dim3(kBlock), 0,
                     stream, loc, m, fp, g, ns, bin_start, rank, pts, perm);
  check_launch("bin_scatter_kernel");
}

template <typename T, int DIM>
void launch_bank_order(const Subprob* subs, std::size_t nsub, PointRec<T, DIM>* pts,
                       uint32_t* perm, const BinGeom& g, hipStream_t stream) {
  if (nsub == 0) return;
  hipLaunchKernelGGL((bank_order_kernel<T, DIM>), dim3(static_cast<unsigned>(nsub)),
                     dim3(kOrderThreads), 0, stream, subs, pts, perm, g.tile[0], g.tile[1]);
  check_launch("bank_order_kernel");
}

std::size_t scan_scratch_size(std::size_t n) { return (n + kScanTile - 1) / kScanTile + 1; }

void launch_exclusive_scan(const uint32_t* in, uint32_t* out, std::size_t n, uint32_t* scratch,
                           hipStream_t stream) {
  if (n == 0) return;
  const std::size_t tiles = (n + kScanTile - 1) / kScanTile;
  hipLaunchKernelGGL(scan_tile_sums, dim3(static_cast<unsigned>(tiles)), dim3(kScanThreads), 0,
                     stream, in, n, scratch);
  check_launch("scan_tile_sums");
  hipLaunchKernelGGL(scan_sums_kernel, dim3(1), dim3(kScanThreads), 0, stream, scratch, tiles);
  check_launch("scan_sums_kernel");
  hipLaunchKernelGGL(scan_tiles_kernel, dim3(static_cast<unsigned>(tiles)), dim3(kScanThreads), 0,
                     stream, in, n, scratch, out);
  check_launch("scan_tiles_kernel");
}

void launch_fill_u32(uint32_t* p, std::size_t n, uint32_t v, hipStream_t stream) {
  if (n == 0) return;
  hipLaunchKernelGGL(fill_u32_kernel, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, stream, p, n, v);
  check_launch("fill_u32_kernel");
}

void launch_subprob_counts(const uint32_t* count, std::size_t nbins, uint32_t max_points,
                           uint32_t* nsub, hipStream_t stream) {
  hipLaunchKernelGGL(subprob_counts_kernel, dim3(grid_for(nbins, kBlock)), dim3(kBlock), 0, stream,
                     count, nbins, max_points, nsub);
  check_launch("subprob_counts_kernel");
}

void launch_subprob_build(const uint32_t* count, const uint32_t* bin_start,
                          const uint32_t* sub_start, std::size_t nbins, uint32_t max_points,
                          Subprob* subs, hipStream_t stream) {
  hipLaunchKernelGGL(subprob_build_kernel, dim3(grid_for(nbins, kBlock)), dim3(kBlock), 0, stream,
                     count, bin_start, sub_start, nbins, max_points, subs);
  check_launch("subprob_build_kernel");
}

void launch_subprob_retag(const Subprob* src, std::size_t nsub, uint32_t job, Subprob* dst,
                          hipStream_t stream) {
  if (nsub == 0) return;
  hipLaunchKernelGGL(subprob_retag_kernel, dim3(grid_for(nsub, kBlock)), dim3(kBlock), 0, stream,
                     src, nsub, job, dst);
  check_launch("subprob_retag_kernel");
}

int minmax_blocks(std::size_t m) {
  std::size_t b = (m + kBlock * 8 - 1) / (kBlock * 8);
  if (b < 1) b = 1;
  if (b > kMinMaxBlocks) b = kMinMaxBlocks;
  return static_cast<int>(b);
}

template <typename T>
void launch_minmax(const T* x, std::size_t m, double* partial, hipStream_t stream) {
  hipLaunchKernelGGL((minmax_kernel<T>), dim3(minmax_blocks(m)), dim3(kBlock), 0, stream, x, m,
                     partial);
  check_launch("minmax_kernel");
}

#define NNF_INST_POINTS(T, DIM)                                                                 \
  template void launch_bin_count<T, DIM>(Loc<T>, std::size_t, const FoldParams&, const BinGeom&, \
                                         int, uint32_t*, uint32_t*, hipStream_t);               \
  template void launch_bin_scatter<T, T, DIM>(Loc<T>, std::size_t, const FoldParams&,           \
                                              const BinGeom&, int, const uint32_t*,             \
                                              const uint32_t*, PointRec<T, DIM>*, uint32_t*,    \
                                              hipStream_t);                                     \
  template void launch_bank_order<T, DIM>(const Subprob*, std::size_t, PointRec<T, DIM>*,       \
                                          uint32_t*, const BinGeom&, hipStream_t);

NNF_INST_POINTS(float, 1)
NNF_INST_POINTS(float, 2)
NNF_INST_POINTS(float, 3)
NNF_INST_POINTS(double, 1)
NNF_INST_POINTS(double, 2)
NNF_INST_POINTS(double, 3)

// double coordinates binned for fp32 records (mixed-precision type 3)
template void launch_bin_scatter<double, float, 1>(Loc<double>, std::size_t, const FoldParams&,
                                                   const BinGeom&, int, const uint32_t*,
                                                   const uint32_t*, PointRec<float, 1>*,
                                                   uint32_t*, hipStream_t);
template void launch_bin_scatter<double, float, 2>(Loc<double>, std::size_t, const FoldParams&,
                                                   const BinGeom&, int, const uint32_t*,
                                                   const uint32_t*, PointRec<float, 2>*,
                                                   uint32_t*, hipStream_t);
template void launch_bin_scatter<double, float, 3>(Loc<double>, std::size_t, const FoldParams&,
                                                   const BinGeom&, int, const uint32_t*,
                                                   const uint32_t*, PointRec<float, 3>*,
                                                   uint32_t*, hipStream_t);

template void launch_minmax<float>(const float*, std::size_t, double*, hipStream_t);
template void launch_minmax<double>(const double*, std::size_t, double*, hipStream_t);

}  // namespace nnf
