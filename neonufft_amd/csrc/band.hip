// Row-band spreading for 2D type 1.
//
// Same sums as the reference's 2D spreading kernel
// (src/neonufft/gpu/kernels/spreading_kernel.cu:527-586): cell
// ((i0x + a) mod n0, (i0y + b) mod n1) receives phi_x(a) phi_y(b) c for every
// point. The decomposition is MI355X-specific: one workgroup owns a band of
// whole fine-grid rows, accumulates every point whose kernel rows reach the
// band in LDS (fixed point, integer atomics) and writes each of its cells
// once. No tiles are stored, there is no tile reduction and no zero fill, and
// the band of rows is the natural unit for a fused x-direction FFT.
//
// Points are read in a row-sorted view (sorted by the first kernel row i0y),
// so the points of a band are one or two contiguous ranges; their strengths
// are first gathered into that order so that the bands read them coalesced.
#include "chunk_rank.hpp"
#include "lds_fft.hpp"

namespace nnf {

namespace {

constexpr int kRowBlock = 256;
constexpr int kGatherItems = 8;
// Band workgroup size and points per thread per pass: every thread issues the
// loads of kBandBatch points before using them. Two workgroups per CU (half
// the LDS each, at most 128 VGPRs) let one accumulate while the other runs its
// FFT or writes out (measured on MI355X, 256 uneven 2D transforms: 5.2 ->
// 4.95 ms against one 1024-thread workgroup per CU).
template <typename T>
constexpr int kBandThreads = sizeof(T) == 4 ? 512 : 256;
// waves per SIMD the register allocation must allow (fp64 needs up to ~210
// VGPRs; two 4-wave workgroups per CU still fit)
template <typename T>
constexpr int kBandWaves = sizeof(T) == 4 ? 4 : 2;
constexpr int kBandBatch = 2;

// First kernel cell (i0x, i0y) of a tile-sorted 2D record, from its bin and
// its bin-relative cell (the packing of bin_scatter_kernel).
template <typename T>
__device__ inline void rec_start(const PointRec<T, 2>& r, const BinGeom& g, int ns, int& i0x,
                                 int& i0y) {
  const int bx = static_cast<int>(r.bin % static_cast<uint32_t>(g.nbins[0]));
  const int by = static_cast<int>(r.bin / static_cast<uint32_t>(g.nbins[0]));
  i0x = bx * g.bin[0] - ns / 2 + static_cast<int>(r.cell & CellPack<2>::mask);
  i0y = by * g.bin[1] - ns / 2 + static_cast<int>(r.cell >> CellPack<2>::bits);
}

template <typename T>
__global__ void __launch_bounds__(kChunkThreads)
    row_count_kernel(const PointRec<T, 2>* __restrict__ recs, std::size_t m, BinGeom g, int ns,
                     uint32_t* __restrict__ count, uint32_t* __restrict__ rank) {
  // the records are bin-sorted: a chunk of them starts on a few dozen rows
  chunk_rank(m, count, rank, [&](std::size_t p) {
    int i0x, i0y;
    rec_start(recs[p], g, ns, i0x, i0y);
    // i0y = ceil(y - ns/2) lies in [-ns/2, n1 - 1]
    return static_cast<uint32_t>(i0y < 0 ? i0y + g.n[1] : i0y);
  });
}

template <typename T>
__global__ void __launch_bounds__(kRowBlock)
    row_scatter_kernel(const PointRec<T, 2>* __restrict__ recs, const uint32_t* __restrict__ perm,
                       std::size_t m, BinGeom g, int ns, const uint32_t* __restrict__ row_start,
                       const uint32_t* __restrict__ rank, PointRec<T, 2>* __restrict__ out,
                       uint32_t* __restrict__ perm_out, uint16_t* __restrict__ row_out) {
  for (std::size_t p = threadIdx.x + static_cast<std::size_t>(blockIdx.x) * kRowBlock; p < m;
       p += static_cast<std::size_t>(gridDim.x) * kRowBlock) {
    PointRec<T, 2> r = recs[p];
    int i0x, i0y;
    rec_start(r, g, ns, i0x, i0y);
    const int row = i0y < 0 ? i0y + g.n[1] : i0y;
    r.cell = static_cast<uint32_t>(i0x + ns / 2) | (static_cast<uint32_t>(row) << CellPack<2>::bits);
    r.bin = static_cast<uint32_t>(row);
    const uint32_t pos = row_start[row] + rank[p];
    out[pos] = r;
    perm_out[pos] = perm[p];
    row_out[pos] = static_cast<uint16_t>(row);
  }
}

// Column blocks of the fixed-point bound: n0 split into at most kBoundBlocks
// blocks of bw cells.
__host__ __device__ inline int bound_block_width(int n0) {
  return (n0 + kBoundBlocks - 1) / kBoundBlocks > 1 ? (n0 + kBoundBlocks - 1) / kBoundBlocks : 1;
}

// bcount[row * kBoundBlocks + k] = view points of first kernel row `row` whose
// kernel x-window meets column block k.
template <typename T>
__global__ void __launch_bounds__(kRowBlock)
    row_block_count_kernel(const PointRec<T, 2>* __restrict__ view, std::size_t m, int n0, int ns,
                           uint32_t* __restrict__ bcount) {
  const int bw = bound_block_width(n0);
  const int nb = (n0 + bw - 1) / bw;
  for (std::size_t p = threadIdx.x + static_cast<std::size_t>(blockIdx.x) * kRowBlock; p < m;
       p += static_cast<std::size_t>(gridDim.x) * kRowBlock) {
    const uint32_t cell = view[p].cell;
    const int row = static_cast<int>(cell >> CellPack<2>::bits);
    const int x0 = static_cast<int>(cell & CellPack<2>::mask) - ns / 2;
    const int xa = x0 < 0 ? x0 + n0 : x0;
    const int xe = x0 + ns - 1;
    const int xb = xe >= n0 ? xe - n0 : xe;
    const int kb = xb / bw;
    uint32_t* c = bcount + static_cast<std::size_t>(row) * kBoundBlocks;
    for (int b = xa / bw;; b = (b + 1 == nb ? 0 : b + 1)) {
      atomicAdd(&c[b], 1u);
      if (b == kb) break;
    }
  }
}

// *out = max over rows r and blocks k of sum_{j < ns} bcount[r - j][k]
// (rows mod n1): the largest point count any cell's bound can see.
__global__ void __launch_bounds__(kRowBlock)
    row_window_max_kernel(const uint32_t* __restrict__ bcount, int n1, int ns,
                          uint32_t* __restrict__ out) {
  const int i = threadIdx.x + blockIdx.x * kRowBlock;
  if (i >= n1 * kBoundBlocks) return;
  const int r = i / kBoundBlocks, k = i - r * kBoundBlocks;
  uint32_t s = 0;
  for (int j = 0; j < ns; ++j) {
    const int row = r - j < 0 ? r - j + n1 : r - j;
    s += bcount[static_cast<std::size_t>(row) * kBoundBlocks + k];
  }
  atomicMax(out, s);
}

__global__ void clear_kernel(unsigned long long* __restrict__ p, std::size_t n) {
  for (std::size_t i = threadIdx.x + static_cast<std::size_t>(blockIdx.x) * blockDim.x; i < n;
       i += static_cast<std::size_t>(gridDim.x) * blockDim.x) {
    p[i] = 0ull;
  }
}

// Running maximum of a non-negative T through the unsigned order of its bits
// (the same as the numeric order for non-negative values).
__device__ inline void max_into(float* slot, float v) {
  atomicMax(reinterpret_cast<unsigned int*>(slot), __float_as_uint(v));
}
__device__ inline void max_into(double* slot, double v) {
  atomicMax(reinterpret_cast<unsigned long long*>(slot),
            static_cast<unsigned long long>(__double_as_longlong(v)));
}

// Index of the job owning flattened item `i` (jobs sorted by `off`).
template <typename T>
__device__ inline int find_job(const BandJob<T>* jobs, int njobs, std::size_t i, bool by_band) {
  int lo = 0, hi = njobs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    const std::size_t o = by_band ? jobs[mid].band_off : jobs[mid].pt_off;
    if (o <= i) {
      lo = mid;
    } else {
      hi = mid - 1;
    }
  }
  return lo;
}

// The fields of a job the gather uses.
template <typename T>
struct GatherJob {
  const uint32_t* perm;
  const uint16_t* vrow;
  const Cplx<T>* in;
  Cplx<T>* sorted;
  T* rowsum;
  T* rowmax;
  std::size_t pt_off, m;
};

template <typename T>
__device__ inline GatherJob<T> gather_fields(const BandJob<T>& b) {
  return GatherJob<T>{b.perm, b.vrow, b.in, b.sorted, b.rowsum, b.rowmax, b.pt_off, b.m};
}

// sorted[p] = in[perm[p]] over the points of every job (flattened), and per
// first kernel row the sum and the largest |Re c| + |Im c| of its points into
// the job's (cleared) rowsum / rowmax. Consecutive lanes take consecutive view
// points (sorted by row), so one row's points are a run of lanes: a segmented
// wave scan combines each run and the run's last lane issues one atomic add
// and one atomic max. A workgroup's loads form two dependent rounds after its
// job's fields: the view indices and rows of all its items, then their
// strengths. With `chunks`, workgroup b gathers the one job range chunks[b]
// names.
template <typename T>
__global__ void __launch_bounds__(kRowBlock)
    band_gather_kernel(const BandJob<T>* __restrict__ jobs, const GatherChunk* __restrict__ chunks,
                       BandJob<T> single, int njobs, std::size_t total) {
  static_assert(kRowBlock * kGatherItems == kGatherChunk, "gather chunk size");
  std::size_t base = static_cast<std::size_t>(blockIdx.x) * kGatherChunk;
  int j = 0;
  GatherJob<T> job = gather_fields(single);
  if (chunks) {
    const GatherChunk e = chunks[blockIdx.x];
    if (e.job == 0xffffffffu) return;
    j = static_cast<int>(e.job);
    job = gather_fields(jobs[j]);
    base = job.pt_off + e.first;
    total = job.pt_off + job.m;  // this job only
  } else if (jobs) {
    if (base >= total) return;
    j = find_job(jobs, njobs, base, false);
    job = gather_fields(jobs[j]);
  }
  if (base >= total) return;
  Cplx<T>* dst[kGatherItems];
  const Cplx<T>* in[kGatherItems];
  uint32_t idx[kGatherItems];
  int lo[kGatherItems], jk[kGatherItems];
#pragma unroll
  for (int k = 0; k < kGatherItems; ++k) {
    const std::size_t i = base + threadIdx.x + static_cast<std::size_t>(k) * kRowBlock;
    dst[k] = nullptr;
    in[k] = nullptr;
    idx[k] = 0u;
    lo[k] = 0;
    jk[k] = j;
    if (i < total) {
      while (i >= job.pt_off + job.m) job = gather_fields(jobs[++j]);
      const std::size_t p = i - job.pt_off;
      dst[k] = job.sorted + p;
      in[k] = job.in;
      idx[k] = job.perm[p];
      lo[k] = static_cast<int>(job.vrow[p]);
      jk[k] = j;
    }
  }
  Cplx<T> v[kGatherItems];
#pragma unroll
  for (int k = 0; k < kGatherItems; ++k) v[k] = dst[k] ? in[k][idx[k]] : Cplx<T>{T(0), T(0)};
#pragma unroll
  for (int k = 0; k < kGatherItems; ++k) {
    if (dst[k]) *dst[k] = v[k];
  }
  const int lane = static_cast<int>(threadIdx.x & 63);
#pragma unroll
  for (int k = 0; k < kGatherItems; ++k) {
    // (job, row) key: rows and jobs are both below 2^16
    const uint32_t key = dst[k] ? (static_cast<uint32_t>(jk[k]) << 16) | static_cast<uint32_t>(lo[k])
                                : 0xffffffffu;
    T s = dst[k] ? fabs(v[k].x) + fabs(v[k].y) : T(0);
    T mx = s;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const T ts = __shfl_up(s, off, 64);
      const T tm = __shfl_up(mx, off, 64);
      const uint32_t kk = __shfl_up(key, off, 64);
      if (lane >= off && kk == key) {
        s += ts;
        mx = max(mx, tm);
      }
    }
    const uint32_t next = __shfl_down(key, 1, 64);
    if (dst[k] && (lane == 63 || next != key)) {
      const GatherJob<T> jb = jk[k] == j ? job : gather_fields(jobs[jk[k]]);
      atomicAdd(jb.rowsum + lo[k], s);
      max_into(jb.rowmax + lo[k], mx);
    }
  }
}

template <typename T>
__device__ inline T wave_max(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off, 64));
  return v;
}


// One band of rows [y0, y0 + rk) of one job.
//  1. bound: every cell value is at most (points of the cell row's NS-row
//     window whose kernel x-window meets the cell's 1/64 column block: counts
//     per row precomputed with the row view) x (the job's largest
//     |Re c| + |Im c|, from the gather); this fixes the fixed-point scale;
//  2. accumulation with integer LDS atomics (rows outside the band skipped;
//     a row no lane of a wave needs is skipped by the whole wave);
//  3. each cell converted back once and written (or added) to the grid, or
//     (FFT) the rows transformed along x in LDS and only the retained modes
//     written, transposed, to job.h[x' * n1 + y].
template <typename T, int NS, bool FFT>
__global__ void __launch_bounds__(kBandThreads<T>)
    __attribute__((amdgpu_waves_per_eu(kBandWaves<T>)))
    band_spread_kernel(const BandJob<T>* __restrict__ jobs, const uint32_t* __restrict__ band_job,
                       BandJob<T> single, int njobs, EsParams es, int accumulate) {
  using Acc = typename FixedAcc<T>::type;
  constexpr int NT = kBandThreads<T>;
  constexpr int KB = kBandBatch;
  extern __shared__ __align__(16) unsigned char smem_raw[];

  const BandJob<T> job =
      jobs ? jobs[band_job ? static_cast<int>(band_job[blockIdx.x])
                           : find_job(jobs, njobs, blockIdx.x, true)]
           : single;
  const int n0 = job.n0, n1 = job.n1, pitch = job.pitch;
  const int y0 = static_cast<int>(blockIdx.x - job.band_off) * job.rows;
  const int rk = min(job.rows, n1 - y0);
  // Row r of the band: cell x in [-kBandLeft, pitch - kBandLeft) at
  // r * pitch + kBandLeft + x of the real plane, `plane` cells later in the
  // imaginary plane (the cells outside [0, n0) are the halo).
  Acc* tre = reinterpret_cast<Acc*>(smem_raw);
  const int plane = rk * pitch;
  Acc* tim = tre + plane;
  for (int i = threadIdx.x; i < 2 * rk * pitch; i += NT) tre[i] = Acc(0);

  // Points whose first kernel row lies in [y0 - NS + 1, y0 + rk - 1] (mod
  // n1): one range of the row-sorted view, or two when it wraps below row 0
  // (rows of the wrapped range sit n1 below their stored row).
  const int lo_row = y0 - (NS - 1);
  uint32_t b0, e0, b1 = 0, e1 = 0;
  int shift0 = 0;
  if (lo_row < 0) {
    b0 = job.row_start[lo_row + n1];
    e0 = job.row_start[n1];
    shift0 = -n1;
    b1 = job.row_start[0];
    e1 = job.row_start[y0 + rk];
  } else {
    b0 = job.row_start[lo_row];
    e0 = job.row_start[y0 + rk];
  }
  const uint32_t cnt0 = e0 - b0;
  const uint32_t total = cnt0 + (e1 - b1);
  const PointRec<T, 2>* __restrict__ pts = job.pts;
  const Cplx<T>* __restrict__ str = job.sorted;

  // Fixed-point bound: cell (r, x) only receives points whose first kernel
  // row lies in [r - NS + 1, r] (its NS-row window), each weighted by at most
  // one. Two bounds hold for it: (points of the window whose kernel x-window
  // meets x's column block) x (the largest |Re c| + |Im c| in the window's
  // rows), and the sum of |Re c| + |Im c| over the window's points (row
  // maxima and sums from the gather). The smaller is taken per (row, block),
  // so a few strengths far above the rest coarsen only the scale of the bands
  // near them, and by at most the window sum. Lane k of wave w takes the band
  // rows w, w + NW, ... and column block k; after the barrier every wave
  // takes the maximum over waves and blocks.
  constexpr int NW = NT / 64;
  __shared__ T part[NW][kBoundBlocks];
  {
    const int k = static_cast<int>(threadIdx.x & 63);
    const int w = static_cast<int>(threadIdx.x >> 6);
    T bm = T(0);
    for (int r = w; r < rk; r += NW) {
      uint32_t bc = 0;
      T ws = T(0), wm = T(0);
#pragma unroll
      for (int j = 0; j < NS; ++j) {
        int row = y0 + r - j;
        row = row < 0 ? row + n1 : row;
        bc += job.bcount[static_cast<std::size_t>(row) * kBoundBlocks + k];
        ws += job.rowsum[row];
        wm = max(wm, job.rowmax[row]);
      }
      bm = max(bm, min(static_cast<T>(bc) * wm, ws));
    }
    part[w][k] = bm;
  }
  __syncthreads();  // zeroed LDS and the partial maxima
  T bpart = T(0);
#pragma unroll
  for (int w = 0; w < NW; ++w) bpart = max(bpart, part[w][threadIdx.x & 63]);
  const T bound = wave_max(bpart);
  const T one = FixedAcc<T>::one;
  const T scale = bound > T(0) ? one / bound : T(0);

  for (uint32_t q0 = 0; q0 < total; q0 += NT * KB) {
    PointRec<T, 2> rec[KB];
    Cplx<T> c[KB];
    int lr0[KB];
#pragma unroll
    for (int k = 0; k < KB; ++k) {
      const uint32_t q = q0 + threadIdx.x + k * NT;
      lr0[k] = NS;  // no row of the band
      rec[k] = PointRec<T, 2>{};
      c[k] = Cplx<T>{T(0), T(0)};
      if (q < total) {
        const bool first = q < cnt0;
        const uint32_t p = first ? b0 + q : b1 + (q - cnt0);
        rec[k] = pts[p];
        c[k] = str[p];
        lr0[k] = first ? shift0 - y0 : -y0;
      }
    }
#pragma unroll
    for (int k = 0; k < KB; ++k) {
      const bool valid = lr0[k] != NS;
      if (__ballot(valid) == 0) break;
      const T cr = c[k].x * scale, ci = c[k].y * scale;
      const int x0 = static_cast<int>(rec[k].cell & CellPack<2>::mask) - NS / 2;
      const int l0 = valid ? static_cast<int>(rec[k].cell >> CellPack<2>::bits) + lr0[k] : rk;
      T kx[NS];
      es_weights<T, NS>(rec[k].x[0], es, kx);
      // cells x0 .. x0 + NS - 1 of the row, unwrapped: a run that crosses
      // either end of the row lands in the row's halo and is folded below,
      // so the NS atomics of a row share one address register
      const int xoff = kBandLeft + x0;
#pragma unroll
      for (int b = 0; b < NS; ++b) {
        const int r = l0 + b;
        const bool act = static_cast<unsigned>(r) < static_cast<unsigned>(rk);
        if (__ballot(act) == 0) continue;
        if (act) {
          const T wy = es_piece<NS>(rec[k].x[1], b, es);
          const T vr = wy * cr, vi = wy * ci;
          Acc* rr = tre + r * pitch + xoff;
          Acc* ri = rr + plane;
#pragma unroll
          for (int a = 0; a < NS; ++a) {
            lds_add_fixed(&rr[a], to_fixed(kx[a] * vr));
            lds_add_fixed(&ri[a], to_fixed(kx[a] * vi));
          }
        }
      }
    }
  }
  __syncthreads();
  // periodic fold of the halos: x in [-NS/2, 0) belongs to x + n0, x in
  // [n0, n0 + NS - NS/2) to x - n0 (sources and targets are disjoint)
  for (int i = threadIdx.x; i < rk * NS; i += NT) {
    const int r = i / NS;
    const int j = i - r * NS;
    constexpr int H = NS / 2;
    const int src = j < H ? j - H : n0 + (j - H);
    const int dst = j < H ? src + n0 : src - n0;
    Acc* row = tre + r * pitch + kBandLeft;
    row[dst] += row[src];
    row[plane + dst] += row[plane + src];
  }
  __syncthreads();

  const T inv = bound > T(0) ? bound / one : T(0);
  if constexpr (FFT) {
    lds_fft_x<T, NT, true>(tre + kBandLeft, tim + kBandLeft, rk, pitch, job.log2n0, inv, job.tw);
    // retained mode xp of row r: X[f] at position bitrev(f). Lanes take
    // (r fastest, then xp from eight interleaved strides of g modes) so that
    // eight lanes write 8 rows of one mode contiguously and the bitrev
    // positions of a wave spread over the LDS banks.
    const T* fre = reinterpret_cast<const T*>(tre + kBandLeft);
    const T* fim = reinterpret_cast<const T*>(tim + kBandLeft);
    const int lr = job.log2rows;
    const int N0 = job.N0;
    const int P = N0 - N0 / 2;
    const int g = (N0 + 7) >> 3;
    const int total_out = (8 * g) << lr;
    const int sh = 32 - job.log2n0;
    Cplx<T>* __restrict__ h = job.h;
    for (int i = threadIdx.x; i < total_out; i += NT) {
      const int r = i & ((1 << lr) - 1);
      const int j = i >> lr;
      const int xp = (j & 7) * g + (j >> 3);
      if (r >= rk || xp >= N0) continue;
      const int f = xp < P ? xp : xp + n0 - N0;
      const int pos = r * pitch + static_cast<int>(__brev(static_cast<unsigned>(f)) >> sh);
      h[static_cast<std::size_t>(xp) * n1 + y0 + r] = Cplx<T>{fre[pos], fim[pos]};
    }
    return;
  }

  // Each cell converted back once; threads cover several rows at a time when
  // a row is shorter than the workgroup.
  Cplx<T>* __restrict__ grid = job.grid + static_cast<std::size_t>(y0) * n0;
  const int rpp = n0 >= NT ? 1 : NT / n0;  // rows per pass
  const int tr = n0 >= NT ? 0 : static_cast<int>(threadIdx.x) / n0;
  const int tx = n0 >= NT ? static_cast<int>(threadIdx.x) : static_cast<int>(threadIdx.x) - tr * n0;
  if (tr < rpp) {
    for (int r = tr; r < rk; r += rpp) {
      Cplx<T>* row = grid + static_cast<std::size_t>(r) * n0;
      for (int x = tx; x < n0; x += (n0 >= NT ? NT : n0)) {
        Cplx<T> v{static_cast<T>(tre[r * pitch + kBandLeft + x]) * inv,
                  static_cast<T>(tim[r * pitch + kBandLeft + x]) * inv};
        if (accumulate) {
          const Cplx<T> o = row[x];
          v.x += o.x;
          v.y += o.y;
        }
        row[x] = v;
      }
    }
  }
}

template <typename T, int NS>
void band_spread_ns(const BandJob<T>* jobs, const uint32_t* band_job, const BandJob<T>& single,
                    int njobs, std::size_t bands, std::size_t lds, bool accumulate, bool fft,
                    const EsParams& es, hipStream_t stream) {
  if (fft) {
    hipLaunchKernelGGL((band_spread_kernel<T, NS, true>), dim3(static_cast<unsigned>(bands)),
                       dim3(kBandThreads<T>), lds, stream, jobs, band_job, single, njobs, es, 0);
  } else {
    hipLaunchKernelGGL((band_spread_kernel<T, NS, false>), dim3(static_cast<unsigned>(bands)),
                       dim3(kBandThreads<T>), lds, stream, jobs, band_job, single, njobs, es,
                       accumulate ? 1 : 0);
  }
  check_launch("band_spread_kernel");
}

unsigned blocks_for(std::size_t work, int per_block) {
  std::size_t g = (work + per_block - 1) / per_block;
  if (g < 1) g = 1;
  if (g > 65535u * 16u) g = 65535u * 16u;
  return static_cast<unsigned>(g);
}

}  // namespace

template <typename T>
void launch_row_count(const PointRec<T, 2>* recs, std::size_t m, const BinGeom& g, int ns,
                      uint32_t* count, uint32_t* rank, hipStream_t stream) {
  if (m == 0) return;
  hipLaunchKernelGGL((row_count_kernel<T>), dim3(chunk_rank_blocks(m)), dim3(kChunkThreads), 0,
                     stream, recs, m, g, ns, count, rank);
  check_launch("row_count_kernel");
}

template <typename T>
void launch_row_scatter(const PointRec<T, 2>* recs, const uint32_t* perm, std::size_t m,
                        const BinGeom& g, int ns, const uint32_t* row_start, const uint32_t* rank,
                        PointRec<T, 2>* out, uint32_t* perm_out, uint16_t* row_out,
                        hipStream_t stream) {
  if (m == 0) return;
  if (g.n[1] > 65536) throw neonufft::InternalError("row view needs at most 65536 rows");
  hipLaunchKernelGGL((row_scatter_kernel<T>), dim3(blocks_for(m, kRowBlock)), dim3(kRowBlock), 0,
                     stream, recs, perm, m, g, ns, row_start, rank, out, perm_out, row_out);
  check_launch("row_scatter_kernel");
}

template <typename T>
void launch_row_block_count(const PointRec<T, 2>* view, std::size_t m, int n0, int ns,
                            uint32_t* bcount, hipStream_t stream) {
  if (m == 0) return;
  hipLaunchKernelGGL((row_block_count_kernel<T>), dim3(blocks_for(m, kRowBlock)),
                     dim3(kRowBlock), 0, stream, view, m, n0, ns, bcount);
  check_launch("row_block_count_kernel");
}

void launch_row_window_max(const uint32_t* bcount, int n1, int ns, uint32_t* out,
                           hipStream_t stream) {
  hipLaunchKernelGGL(row_window_max_kernel, dim3(blocks_for(static_cast<std::size_t>(n1) * kBoundBlocks, kRowBlock)),
                     dim3(kRowBlock), 0, stream, bcount, n1, ns, out);
  check_launch("row_window_max_kernel");
}

void launch_clear_bounds(unsigned long long* words, std::size_t n, hipStream_t stream) {
  if (n == 0) return;
  hipLaunchKernelGGL(clear_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, stream, words, n);
  check_launch("clear_kernel");
}

template <typename T>
void launch_band_gather(const BandJob<T>* jobs, const BandJob<T>& single, int njobs,
                        std::size_t total_points, hipStream_t stream) {
  if (total_points == 0) return;
  const std::size_t blocks = (total_points + kGatherChunk - 1) / kGatherChunk;
  if (blocks > 0x7fffffffu) throw neonufft::InputError("Neonufft:: too many points");
  hipLaunchKernelGGL((band_gather_kernel<T>), dim3(static_cast<unsigned>(blocks)),
                     dim3(kRowBlock), 0, stream, jobs, nullptr, single, njobs, total_points);
  check_launch("band_gather_kernel");
}

template <typename T>
void launch_band_gather_chunks(const BandJob<T>* jobs, const GatherChunk* chunks,
                               std::size_t nchunks, hipStream_t stream) {
  if (nchunks == 0) return;
  if (nchunks > 0x7fffffffu) throw neonufft::InputError("Neonufft:: too many points");
  hipLaunchKernelGGL((band_gather_kernel<T>), dim3(static_cast<unsigned>(nchunks)),
                     dim3(kRowBlock), 0, stream, jobs, chunks, BandJob<T>{}, 0, std::size_t{0});
  check_launch("band_gather_kernel");
}

template <typename T>
void launch_band_spread(int ns, const BandJob<T>* jobs, const uint32_t* band_job,
                        const BandJob<T>& single, int njobs, std::size_t total_bands,
                        std::size_t lds_bytes, bool accumulate, bool fft, const EsParams& es,
                        hipStream_t stream) {
  if (total_bands == 0) return;
  if (fft && accumulate) throw neonufft::InternalError("fused band FFT cannot accumulate");
  if (total_bands > 0x7fffffffu) throw neonufft::InputError("Neonufft:: too many row bands");
#define NNF_BAND_CASE(NSV)                                                                    \
  case NSV:                                                                                  \
    band_spread_ns<T, NSV>(jobs, band_job, single, njobs, total_bands, lds_bytes, accumulate,  \
                           fft, es, stream);                                                 \
    break;
  switch (ns) {
    NNF_BAND_CASE(2)
    NNF_BAND_CASE(3)
    NNF_BAND_CASE(4)
    NNF_BAND_CASE(5)
    NNF_BAND_CASE(6)
    NNF_BAND_CASE(7)
    NNF_BAND_CASE(8)
    NNF_BAND_CASE(9)
    NNF_BAND_CASE(10)
    NNF_BAND_CASE(11)
    NNF_BAND_CASE(12)
    NNF_BAND_CASE(13)
    NNF_BAND_CASE(14)
    NNF_BAND_CASE(15)
    NNF_BAND_CASE(16)
    default:
      throw neonufft::InternalError("n_spread not in [2, 16]");
  }
#undef NNF_BAND_CASE
}

std::size_t band_lds_bytes(int rows, int pitch, std::size_t acc_bytes) {
  return 2 * static_cast<std::size_t>(rows) * static_cast<std::size_t>(pitch) * acc_bytes;
}

#define NNF_INST_BAND(T)                                                                        \
  template void launch_row_count<T>(const PointRec<T, 2>*, std::size_t, const BinGeom&, int,    \
                                    uint32_t*, uint32_t*, hipStream_t);                         \
  template void launch_row_scatter<T>(const PointRec<T, 2>*, const uint32_t*, std::size_t,      \
                                      const BinGeom&, int, const uint32_t*, const uint32_t*,    \
                                      PointRec<T, 2>*, uint32_t*, uint16_t*, hipStream_t);      \
  template void launch_row_block_count<T>(const PointRec<T, 2>*, std::size_t, int, int,          \
                                          uint32_t*, hipStream_t);                               \
  template void launch_band_gather_chunks<T>(const BandJob<T>*, const GatherChunk*, std::size_t, \
                                             hipStream_t);                                        \
  template void launch_band_gather<T>(const BandJob<T>*, const BandJob<T>&, int, std::size_t,   \
                                      hipStream_t);                                             \
  template void launch_band_spread<T>(int, const BandJob<T>*, const uint32_t*, const BandJob<T>&, \
                                      int, std::size_t, std::size_t, bool, bool,                 \
                                      const EsParams&, hipStream_t);

NNF_INST_BAND(float)
NNF_INST_BAND(double)

}  // namespace nnf
