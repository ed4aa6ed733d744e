// ES-kernel spreading (type 1) and interpolation (type 2) on binned points.
//
// Semantics follow the reference's spread_{1,2,3}d_kernel and
// interpolation_{1,2,3}d_kernel (src/neonufft/gpu/kernels/spreading_kernel.cu:477-660,
// interpolation_kernel.cu:27-319): grid[(i0 + j) mod n] gets
// prod_d phi(i0_d + j_d - xi_d) * c (times the type-3 prephase), and the
// interpolated value is the same weighted sum over the grid (times the type-3
// postphase). The design is MI355X-specific:
//  - one workgroup per subproblem (<= max_points points of one bin), with the
//    bin's tile plus its ns-cell halo staged in LDS;
//  - spreading accumulates with LDS atomics, stores the finished tile with
//    plain coalesced stores, and a tile reduction sums overlapping tiles into
//    the grid (no global atomics, fixed summation order); interpolation
//    loads the tile once with coalesced reads and gathers from LDS;
//  - kernel weights come from per-piece Horner polynomials in registers (fp32)
//    or the exact exp/sqrt formula (fp64);
//  - subproblem -> workgroup mapping keeps neighbouring bins on one XCD.
#include "es_device.hpp"

namespace nnf {

namespace {

// Workgroup sizes. Spreading holds its points in registers across the LDS
// accumulation, so 512 threads (4 points each) are best; fp32 interpolation
// keeps nothing across points and gains from 1024 threads hiding LDS latency
// (measured on MI355X: 1.05 vs 1.33 ms spreading, 0.60 vs 0.52 ms
// interpolation, 1e7 points on a 256^3 grid, ns = 6). fp64 interpolation
// needs more than the 128 VGPRs a 1024-thread workgroup leaves per lane.
constexpr int kThreads = 512;
template <typename T>
constexpr int kInterpThreads = sizeof(T) == 4 ? 1024 : 512;
// points per thread in one spreading subproblem
constexpr int kPointsPerThread = kMaxSubproblemPoints / kThreads;
static_assert(kPointsPerThread * kThreads == kMaxSubproblemPoints);

// Tile geometry of one subproblem.
template <int DIM, int NS>
struct TileInfo {
  int org[3];  // grid index of tile cell 0 per dim (bin origin - NS/2)
  int ext[3];  // cells of the tile that can receive contributions
  int t0, t1;  // LDS pitches (x, y)

  __device__ TileInfo(uint32_t bin, const BinGeom& g) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      org[d] = 0;
      ext[d] = 1;
    }
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      const int bc = static_cast<int>(bin % static_cast<uint32_t>(g.nbins[d]));
      bin /= static_cast<uint32_t>(g.nbins[d]);
      const int o = bc * g.bin[d];
      org[d] = o - NS / 2;
      ext[d] = min(g.bin[d], g.n[d] - o) + NS;
    }
    t0 = g.tile[0];
    t1 = DIM >= 2 ? g.tile[1] : 1;
  }
};

// Sum over the workgroup; contains a __syncthreads().
template <typename T>
__device__ inline T block_sum(T v, T* scratch) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  if ((threadIdx.x & 63) == 0) scratch[threadIdx.x >> 6] = v;
  __syncthreads();
  T t = T(0);
#pragma unroll
  for (int w = 0; w < kThreads / 64; ++w) t += scratch[w];
  return t;
}

// One subproblem of spreading. `smem_raw` is the dynamic LDS tile.
template <typename T, int DIM, int NS>
__device__ __forceinline__ void spread_subproblem(const SpreadJob<T, DIM>& job,
                                                  const Subprob& sp, uint32_t sidx,
                                                  const EsParams& es, unsigned char* smem_raw,
                                                  T* wsum) {
  using Acc = typename FixedAcc<T>::type;
  const PointRec<T, DIM>* __restrict__ pts = job.pts;
  const uint32_t* __restrict__ perm = job.perm;
  const Cplx<T>* __restrict__ in = job.in;
  const Cplx<T>* __restrict__ pre = job.pre;
  const BinGeom& g = job.g;
  const TileInfo<DIM, NS> ti(sp.bin, g);
  const int tile_cells = g.tile[0] * (DIM >= 2 ? g.tile[1] : 1) * (DIM >= 3 ? g.tile[2] : 1);
  // fp32: one packed 64-bit word (re + 2^32 im) per cell, one atomic per
  // cell; fp64: real and imaginary parts in separate planes of 64-bit
  // integers. Either way neighbouring cells fall in neighbouring banks.
  constexpr bool kPacked = std::is_same_v<T, float>;
  Acc* tre = reinterpret_cast<Acc*>(smem_raw);
  Acc* tim = tre + tile_cells;
  unsigned long long* tpk = reinterpret_cast<unsigned long long*>(smem_raw);
  for (int i = threadIdx.x; i < 2 * tile_cells; i += kThreads) tre[i] = Acc(0);

  // This thread's points (at most kPointsPerThread) and the tile bound.
  const uint32_t end = sp.begin + sp.count;
  PointRec<T, DIM> rec[kPointsPerThread];
  Cplx<T> cv[kPointsPerThread];
  T bound = T(0);
#pragma unroll
  for (int k = 0; k < kPointsPerThread; ++k) {
    const uint32_t p = sp.begin + threadIdx.x + k * kThreads;
    cv[k] = Cplx<T>{T(0), T(0)};
    if (p < end) {
      rec[k] = pts[p];
      Cplx<T> c = perm ? in[perm[p]] : in[p];
      if (pre) {
        const Cplx<T> q = pre[p];
        c = Cplx<T>{c.x * q.x - c.y * q.y, c.x * q.y + c.y * q.x};
      }
      cv[k] = c;
      bound += fabs(c.x) + fabs(c.y);
    }
  }
  const T total = block_sum(bound, wsum);  // also orders the tile zero fill
  const T one = FixedAcc<T>::one;
  const T scale = total > T(0) ? one / total : T(0);

#pragma unroll
  for (int k = 0; k < kPointsPerThread; ++k) {
    const uint32_t p = sp.begin + threadIdx.x + k * kThreads;
    if (p >= end) continue;
    const Cplx<T> c{cv[k].x * scale, cv[k].y * scale};
    T kx[NS];
    es_weights<T, NS>(rec[k].x[0], es, kx);
    const int r0 = unpack_cell<DIM>(rec[k].cell, 0);
    if constexpr (DIM == 1) {
#pragma unroll
      for (int a = 0; a < NS; ++a) {
        if constexpr (kPacked) {
          lds_add_packed(&tpk[r0 + a], to_fixed(kx[a] * c.x), to_fixed(kx[a] * c.y));
        } else {
          lds_add_fixed(&tre[r0 + a], to_fixed(kx[a] * c.x));
          lds_add_fixed(&tim[r0 + a], to_fixed(kx[a] * c.y));
        }
      }
    } else {
      T ky[NS];
      es_weights<T, NS>(rec[k].x[1], es, ky);
      const int r1 = unpack_cell<DIM>(rec[k].cell, 1);
      int r2 = 0;
      if constexpr (DIM == 3) r2 = unpack_cell<DIM>(rec[k].cell, 2);
      constexpr int NZ = DIM == 3 ? NS : 1;
      constexpr int kUnrollZ = NS <= 8 ? NZ : 1;
#pragma unroll kUnrollZ
      for (int cz = 0; cz < NZ; ++cz) {
        T wz = T(1);
        if constexpr (DIM == 3) wz = es_piece<NS>(rec[k].x[2], cz, es);
        const T czr = wz * c.x, czi = wz * c.y;
#pragma unroll
        for (int b = 0; b < NS; ++b) {
          const T vr = ky[b] * czr, vi = ky[b] * czi;
          const int base = ((r2 + cz) * ti.t1 + r1 + b) * ti.t0 + r0;
          if constexpr (kPacked) {
            // the row's conversions first: a conversion's result is not
            // read by the next instruction (which would need a wait state)
            int fr[NS], fi[NS];
#pragma unroll
            for (int a = 0; a < NS; ++a) fr[a] = to_fixed(kx[a] * vr);
#pragma unroll
            for (int a = 0; a < NS; ++a) fi[a] = to_fixed(kx[a] * vi);
#pragma unroll
            for (int a = 0; a < NS; ++a) lds_add_packed(&tpk[base + a], fr[a], fi[a]);
          } else {
#pragma unroll
            for (int a = 0; a < NS; ++a) {
              lds_add_fixed(&tre[base + a], to_fixed(kx[a] * vr));
              lds_add_fixed(&tim[base + a], to_fixed(kx[a] * vi));
            }
          }
        }
      }
    }
  }
  __syncthreads();

  // Store the whole tile (converted back from fixed point) to this
  // subproblem's slot with coalesced plain stores; the tile reduction adds
  // overlapping tiles into the grid without atomics.
  const T inv = total > T(0) ? total / one : T(0);
  Cplx<T>* __restrict__ dst =
      job.tiles + static_cast<std::size_t>(sidx - job.sub_off) * static_cast<std::size_t>(tile_cells);
  for (int i = threadIdx.x; i < tile_cells; i += kThreads) {
    if constexpr (kPacked) {
      int re, im;
      unpack_packed(tpk[i], &re, &im);
      dst[i] = Cplx<T>{static_cast<T>(re) * inv, static_cast<T>(im) * inv};
    } else {
      dst[i] = Cplx<T>{static_cast<T>(tre[i]) * inv, static_cast<T>(tim[i]) * inv};
    }
  }
}

template <typename T, int DIM, int NS>
__global__ void __launch_bounds__(kThreads)
    spread_kernel(SpreadJob<T, DIM> job, const Subprob* __restrict__ subs, uint32_t nsub,
                  EsParams es) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  __shared__ T wsum[kThreads / 64];
  const uint32_t sidx = xcd_remap(blockIdx.x, nsub);
  spread_subproblem<T, DIM, NS>(job, subs[sidx], sidx, es, smem_raw, wsum);
}

// Many independent transforms in one launch: each subproblem names its job.
template <typename T, int DIM, int NS>
__global__ void __launch_bounds__(kThreads)
    spread_batch_kernel(const SpreadJob<T, DIM>* __restrict__ jobs,
                        const Subprob* __restrict__ subs, uint32_t nsub, EsParams es) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  __shared__ T wsum[kThreads / 64];
  const uint32_t sidx = xcd_remap(blockIdx.x, nsub);
  const Subprob sp = subs[sidx];
  const SpreadJob<T, DIM> job = jobs[sp.job];
  spread_subproblem<T, DIM, NS>(job, sp, sidx, es, smem_raw, wsum);
}

// ---------------------------------------------------------------------------
// Interpolation
// ---------------------------------------------------------------------------
template <typename T, int DIM, int NS>
__device__ __forceinline__ void interp_subproblem(const InterpJob<T, DIM>& job,
                                                  const Subprob& sp, const EsParams& es,
                                                  unsigned char* smem_raw) {
  Cplx<T>* tile = reinterpret_cast<Cplx<T>*>(smem_raw);
  const PointRec<T, DIM>* __restrict__ pts = job.pts;
  const uint32_t* __restrict__ perm = job.perm;
  const Cplx<T>* __restrict__ grid = job.grid;
  const Cplx<T>* __restrict__ post = job.post;
  Cplx<T>* __restrict__ out = job.out;
  const BinGeom& g = job.g;
  const TileInfo<DIM, NS> ti(sp.bin, g);

  // Stage the tile (bin + halo) from the grid. The loop runs over the tile's
  // cells flattened (x fastest, so reads stay coalesced along rows) and each
  // lane issues kStage independent loads before storing any: one 38-cell row
  // per wave and step left 40 % of the lanes idle and exposed the full load
  // latency once per row. A lane's cell (x, y, z) advances by a whole
  // workgroup of cells per step with two carries instead of a division per
  // cell, and tiles that do not cross the grid's edge skip the periodic wrap.
  {
    constexpr int S = kInterpThreads<T>;
    constexpr int kStage = 4;
    const int e0 = ti.ext[0], e1 = ti.ext[1], e2 = ti.ext[2];
    // advance of (x, y, z) per step of S cells: S = drow e0 + dx, drow = dz e1 + dy
    const int dx = S % e0, drow = S / e0;
    const int dy = drow % e1, dz = drow / e1;
    int x, y, z;
    {
      int rest;
      x = div_rem(static_cast<int>(threadIdx.x), e0, 1.0f / static_cast<float>(e0), &rest);
      y = div_rem(rest, e1, 1.0f / static_cast<float>(e1), &z);
    }
    bool interior = true;
#pragma unroll
    for (int d = 0; d < DIM; ++d) interior &= ti.org[d] >= 0 && ti.org[d] + ti.ext[d] <= g.n[d];
    // grids below 2^32 cells whose (y, z) plane count fits 24 bits index in
    // 32 bits with the full-rate 24-bit multiplier
    const bool narrow = static_cast<uint64_t>(g.n[0]) * static_cast<uint64_t>(g.n[1]) *
                                static_cast<uint64_t>(g.n[2]) < (uint64_t{1} << 32) &&
                        g.n[1] * g.n[2] < (1 << 24);
    // one straight-line loop per (interior, narrow) case: a branch inside the
    // step body makes the compiler wait for each load before the next
    auto stage = [&](auto interior_c, auto narrow_c) __attribute__((always_inline)) {
      constexpr bool kInterior = decltype(interior_c)::value;
      constexpr bool kNarrow = decltype(narrow_c)::value;
      while (z < e2) {
        Cplx<T> v[kStage];
        int dst[kStage];
#pragma unroll
        for (int k = 0; k < kStage; ++k) {
          dst[k] = -1;
          if (z < e2) {
            int gx = ti.org[0] + x;
            int gy = DIM >= 2 ? ti.org[1] + y : 0;
            int gz = DIM >= 3 ? ti.org[2] + z : 0;
            if constexpr (!kInterior) {
              gx = wrap_once(gx, g.n[0]);
              if constexpr (DIM >= 2) gy = wrap_once(gy, g.n[1]);
              if constexpr (DIM >= 3) gz = wrap_once(gz, g.n[2]);
            }
            const int row = __mul24(gz, g.n[1]) + gy;
            if constexpr (kNarrow) {
              v[k] = grid[__umul24(static_cast<unsigned>(row), static_cast<unsigned>(g.n[0])) +
                          static_cast<unsigned>(gx)];
            } else {
              v[k] = grid[static_cast<std::size_t>(row) * static_cast<std::size_t>(g.n[0]) +
                          static_cast<std::size_t>(gx)];
            }
            dst[k] = __mul24(__mul24(z, ti.t1) + y, ti.t0) + x;
          }
          x += dx;
          const int cx = x >= e0 ? 1 : 0;
          x -= cx ? e0 : 0;
          y += dy + cx;
          const int cy = y >= e1 ? 1 : 0;
          y -= cy ? e1 : 0;
          z += dz + cy;
        }
#pragma unroll
        for (int k = 0; k < kStage; ++k) {
          if (dst[k] >= 0) tile[dst[k]] = v[k];
        }
      }
    };
    using yes = std::true_type;
    using no = std::false_type;
    if (interior) {
      if (narrow) stage(yes{}, yes{}); else stage(yes{}, no{});
    } else {
      if (narrow) stage(no{}, yes{}); else stage(no{}, no{});
    }
  }
  __syncthreads();

  const uint32_t end = sp.begin + sp.count;
  for (uint32_t p = sp.begin + threadIdx.x; p < end; p += kInterpThreads<T>) {
    const PointRec<T, DIM> rec = pts[p];
    T kx[NS];
    es_weights<T, NS>(rec.x[0], es, kx);
    const int r0 = unpack_cell<DIM>(rec.cell, 0);
    T sr = T(0), si = T(0);
    if constexpr (DIM == 1) {
#pragma unroll
      for (int a = 0; a < NS; ++a) {
        const Cplx<T> v = tile[r0 + a];
        sr += kx[a] * v.x;
        si += kx[a] * v.y;
      }
    } else {
      T ky[NS];
      es_weights<T, NS>(rec.x[1], es, ky);
      const int r1 = unpack_cell<DIM>(rec.cell, 1);
      int r2 = 0;
      if constexpr (DIM == 3) r2 = unpack_cell<DIM>(rec.cell, 2);
      constexpr int NZ = DIM == 3 ? NS : 1;
      // fp64: full unrolling lets the compiler hoist up to ns^3 tile loads
      // ahead of the FMAs and spill them; unroll z in fp32 only, y by 2.
      constexpr int kUnrollZ = (sizeof(T) == 4 && NS <= 8) ? NZ : 1;
      constexpr int kUnrollY = sizeof(T) == 4 ? NS : 2;
#pragma unroll kUnrollZ
      for (int cz = 0; cz < NZ; ++cz) {
        T pr = T(0), pi = T(0);
#pragma unroll kUnrollY
        for (int b = 0; b < NS; ++b) {
          const Cplx<T>* row = tile + ((r2 + cz) * ti.t1 + r1 + b) * ti.t0 + r0;
          T rr = T(0), ri = T(0);
#pragma unroll
          for (int a = 0; a < NS; ++a) {
            const Cplx<T> v = row[a];
            rr += kx[a] * v.x;
            ri += kx[a] * v.y;
          }
          pr += ky[b] * rr;
          pi += ky[b] * ri;
        }
        T wz = T(1);
        if constexpr (DIM == 3) wz = es_piece<NS>(rec.x[2], cz, es);
        sr += wz * pr;
        si += wz * pi;
      }
    }
    if (post) {
      const Cplx<T> q = post[p];
      const T tr = sr * q.x - si * q.y;
      si = sr * q.y + si * q.x;
      sr = tr;
    }
    out[perm[p]] = Cplx<T>{sr, si};
  }
}

template <typename T, int DIM, int NS>
__global__ void __launch_bounds__(kInterpThreads<T>)
    interp_kernel(InterpJob<T, DIM> job, const Subprob* __restrict__ subs, uint32_t nsub,
                  EsParams es) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const Subprob sp = subs[xcd_remap(blockIdx.x, nsub)];
  interp_subproblem<T, DIM, NS>(job, sp, es, smem_raw);
}

template <typename T, int DIM, int NS>
__global__ void __launch_bounds__(kInterpThreads<T>)
    interp_batch_kernel(const InterpJob<T, DIM>* __restrict__ jobs,
                        const Subprob* __restrict__ subs, uint32_t nsub, EsParams es) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const Subprob sp = subs[xcd_remap(blockIdx.x, nsub)];
  const InterpJob<T, DIM> job = jobs[sp.job];
  interp_subproblem<T, DIM, NS>(job, sp, es, smem_raw);
}

template <typename T, int DIM, int NS>
void spread_ns(const SpreadJob<T, DIM>& job, const Subprob* subs, std::size_t nsub,
               const EsParams& es, hipStream_t stream) {
  const std::size_t lds = tile_bytes<T>(job.g);
  hipLaunchKernelGGL((spread_kernel<T, DIM, NS>), dim3(static_cast<unsigned>(nsub)),
                     dim3(kThreads), lds, stream, job, subs, static_cast<uint32_t>(nsub), es);
  check_launch("spread_kernel");
}

template <typename T, int DIM, int NS>
void interp_ns(const InterpJob<T, DIM>& job, const Subprob* subs, std::size_t nsub,
               const EsParams& es, hipStream_t stream) {
  const std::size_t lds = tile_bytes<T>(job.g);
  hipLaunchKernelGGL((interp_kernel<T, DIM, NS>), dim3(static_cast<unsigned>(nsub)),
                     dim3(kInterpThreads<T>), lds, stream, job, subs, static_cast<uint32_t>(nsub), es);
  check_launch("interp_kernel");
}

template <typename T, int DIM, int NS>
void spread_batch_ns(const SpreadJob<T, DIM>* jobs, const Subprob* subs, std::size_t nsub,
                     std::size_t lds, const EsParams& es, hipStream_t stream) {
  hipLaunchKernelGGL((spread_batch_kernel<T, DIM, NS>), dim3(static_cast<unsigned>(nsub)),
                     dim3(kThreads), lds, stream, jobs, subs, static_cast<uint32_t>(nsub), es);
  check_launch("spread_batch_kernel");
}

template <typename T, int DIM, int NS>
void interp_batch_ns(const InterpJob<T, DIM>* jobs, const Subprob* subs, std::size_t nsub,
                     std::size_t lds, const EsParams& es, hipStream_t stream) {
  hipLaunchKernelGGL((interp_batch_kernel<T, DIM, NS>), dim3(static_cast<unsigned>(nsub)),
                     dim3(kInterpThreads<T>), lds, stream, jobs, subs, static_cast<uint32_t>(nsub), es);
  check_launch("interp_batch_kernel");
}

// ---------------------------------------------------------------------------
// Tile reduction: fine-grid cell c receives the sum of every stored tile that
// covers it (each subproblem of each covering bin), read at the cell's
// position inside that tile. Consecutive threads take consecutive x cells, so
// the tile reads and the grid write are coalesced; no atomics, and the
// summation order is fixed.
// ---------------------------------------------------------------------------
constexpr int kReduceThreads = 256;

// Adds tile values of every subproblem of `bin` at position `pos` (general
// path: any number of covering tiles per dimension).
template <typename T, int DIM>
__device__ inline void reduce_cell_loops(const ReduceJob<T, DIM>& job, int x0, int x1, int y0,
                                         int y1, int z0, int z1, T& ar, T& ai) {
  const BinGeom& g = job.g;
  const std::size_t tc = tile_cells(g);
  for (int iz = z0; iz < z1; ++iz) {
    int bz = 0, lz = 0;
    if constexpr (DIM >= 3) {
      const int2 e = job.cover.entry[2][iz];
      bz = e.x;
      lz = e.y;
    }
    for (int iy = y0; iy < y1; ++iy) {
      int by = 0, ly = 0;
      if constexpr (DIM >= 2) {
        const int2 e = job.cover.entry[1][iy];
        by = e.x;
        ly = e.y;
      }
      const int brow = g.nbins[0] * (by + g.nbins[1] * bz);
      const std::size_t lrow = static_cast<std::size_t>(g.tile[0]) * (ly + g.tile[1] * lz);
      for (int ix = x0; ix < x1; ++ix) {
        const int2 e = job.cover.entry[0][ix];
        const int bin = brow + e.x;
        const uint32_t s0 = job.sub_start[bin], s1 = job.sub_start[bin + 1];
        const Cplx<T>* t = job.tiles + lrow + e.y;
        for (uint32_t s = s0; s < s1; ++s) {
          const Cplx<T> v = t[static_cast<std::size_t>(s) * tc];
          ar += v.x;
          ai += v.y;
        }
      }
    }
  }
}

// Sums the tiles covering the x cells of one grid row (y, z fixed: the y / z
// cover lists are the same for every thread of the block). The common case,
// at most two covering tiles per dimension, evaluates all (<= 8) tile
// combinations unrolled so their loads are in flight together; other cells
// take the general loops.
template <typename T, int DIM>
__device__ inline void reduce_row(const ReduceJob<T, DIM>& job, int c1, int c2, int xbegin,
                                  int xstride, bool accumulate) {
  const BinGeom& g = job.g;
  const std::size_t tc = tile_cells(g);
  int y0 = 0, y1 = 1, z0 = 0, z1 = 1;
  if constexpr (DIM >= 2) {
    y0 = job.cover.start[1][c1];
    y1 = job.cover.start[1][c1 + 1];
  }
  if constexpr (DIM >= 3) {
    z0 = job.cover.start[2][c2];
    z1 = job.cover.start[2][c2 + 1];
  }
  constexpr int NY = DIM >= 2 ? 2 : 1;
  constexpr int NZ = DIM >= 3 ? 2 : 1;
  const bool yz_fast = (y1 - y0) <= NY && (z1 - z0) <= NZ;
  // row part of every combination: bin row offset and tile row offset
  int brow[NY * NZ];
  std::size_t lrow[NY * NZ];
  bool vrow[NY * NZ];
#pragma unroll
  for (int jz = 0; jz < NZ; ++jz) {
#pragma unroll
    for (int jy = 0; jy < NY; ++jy) {
      const int k = jy + NY * jz;
      vrow[k] = yz_fast && (y0 + jy < y1) && (z0 + jz < z1);
      int by = 0, ly = 0, bz = 0, lz = 0;
      if (vrow[k]) {
        if constexpr (DIM >= 2) {
          const int2 e = job.cover.entry[1][y0 + jy];
          by = e.x;
          ly = e.y;
        }
        if constexpr (DIM >= 3) {
          const int2 e = job.cover.entry[2][z0 + jz];
          bz = e.x;
          lz = e.y;
        }
      }
      brow[k] = g.nbins[0] * (by + g.nbins[1] * bz);
      lrow[k] = static_cast<std::size_t>(g.tile[0]) * (ly + g.tile[1] * lz);
    }
  }
  Cplx<T>* __restrict__ grow =
      job.grid + static_cast<std::size_t>(g.n[0]) * (c1 + static_cast<std::size_t>(g.n[1]) * c2);
  for (int c0 = xbegin; c0 < g.n[0]; c0 += xstride) {
    const int x0 = job.cover.start[0][c0], x1 = job.cover.start[0][c0 + 1];
    T ar = T(0), ai = T(0);
    if (yz_fast && x1 - x0 <= 2) {
      int2 ex[2];
      ex[0] = job.cover.entry[0][x0];
      ex[1] = x1 - x0 == 2 ? job.cover.entry[0][x0 + 1] : make_int2(0, 0);
      uint32_t s0[2 * NY * NZ], s1[2 * NY * NZ];
#pragma unroll
      for (int k = 0; k < 2 * NY * NZ; ++k) {
        const int jx = k & 1, kr = k >> 1;
        const bool valid = vrow[kr] && x0 + jx < x1;
        const int bin = brow[kr] + ex[jx].x;
        s0[k] = valid ? job.sub_start[bin] : 0u;
        s1[k] = valid ? job.sub_start[bin + 1] : 0u;
      }
      // first subproblem of every combination: independent loads
      Cplx<T> v[2 * NY * NZ];
#pragma unroll
      for (int k = 0; k < 2 * NY * NZ; ++k) {
        v[k] = Cplx<T>{T(0), T(0)};
        if (s1[k] > s0[k]) {
          v[k] = job.tiles[lrow[k >> 1] + ex[k & 1].y + static_cast<std::size_t>(s0[k]) * tc];
        }
      }
#pragma unroll
      for (int k = 0; k < 2 * NY * NZ; ++k) {
        ar += v[k].x;
        ai += v[k].y;
      }
      // bins split into several subproblems
#pragma unroll
      for (int k = 0; k < 2 * NY * NZ; ++k) {
        const Cplx<T>* t = job.tiles + lrow[k >> 1] + ex[k & 1].y;
        for (uint32_t s = s0[k] + 1; s < s1[k]; ++s) {
          const Cplx<T> w = t[static_cast<std::size_t>(s) * tc];
          ar += w.x;
          ai += w.y;
        }
      }
    } else {
      reduce_cell_loops<T, DIM>(job, x0, x1, y0, y1, z0, z1, ar, ai);
    }
    if (accumulate) {
      const Cplx<T> v = grow[c0];
      ar += v.x;
      ai += v.y;
    }
    grow[c0] = Cplx<T>{ar, ai};
  }
}

// ---------------------------------------------------------------------------
// Bin-owned tile reduction for regular geometries (per dimension: bin >= ns,
// at least two bins, last bin full or >= ns wide). Then a cell of bin b is
// covered only by the tiles of b and of its two neighbours: by b at position
// i + ns/2 (i = offset inside the bin), by b - 1 when i < ns - ns/2 (at
// i + w(b - 1) + ns/2) and by b + 1 when i >= w(b) - ns/2 (at
// i - w(b) + ns/2), w = bin width, neighbours periodic. One workgroup takes
// one bin: the subproblem ranges of the 3^DIM neighbour bins go to LDS once,
// and every tile load afterwards is independent of the others.
// ---------------------------------------------------------------------------
constexpr int kBinReduceThreads = 256;

template <int DIM>
constexpr int kNeigh = DIM == 1 ? 3 : (DIM == 2 ? 9 : 27);

template <typename T, int DIM>
__device__ inline void reduce_bin(const ReduceJob<T, DIM>& job, int ns, uint32_t bin,
                                  uint32_t (*rng)[2], bool accumulate) {
  const BinGeom& g = job.g;
  const int h = ns / 2;
  int b[3] = {0, 0, 0}, w[3] = {1, 1, 1}, wprev[3] = {1, 1, 1};
  uint32_t rem = bin;
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    b[d] = static_cast<int>(rem % static_cast<uint32_t>(g.nbins[d]));
    rem /= static_cast<uint32_t>(g.nbins[d]);
    w[d] = min(g.bin[d], g.n[d] - b[d] * g.bin[d]);
    const int bp = b[d] == 0 ? g.nbins[d] - 1 : b[d] - 1;
    wprev[d] = min(g.bin[d], g.n[d] - bp * g.bin[d]);
  }
  // subproblem ranges of the neighbour bins (offset -1, 0, +1 per dimension)
  if (threadIdx.x < kNeigh<DIM>) {
    int k = threadIdx.x, nb = 0, mult = 1;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      const int off = k % 3 - 1;
      k /= 3;
      int bd = b[d] + off;
      bd = bd < 0 ? bd + g.nbins[d] : (bd >= g.nbins[d] ? bd - g.nbins[d] : bd);
      nb += bd * mult;
      mult *= g.nbins[d];
    }
    rng[threadIdx.x][0] = job.sub_start[nb];
    rng[threadIdx.x][1] = job.sub_start[nb + 1];
  }
  __syncthreads();
  const std::size_t tc = tile_cells(g);
  const int cells = w[0] * w[1] * w[2];
  constexpr int kCombos = 1 << DIM;  // own tile or the one covering neighbour, per dimension
  constexpr int kCells = 1;  // cells per thread per pass (more passes beat fewer waves)
  for (int ibase = threadIdx.x; ibase < cells; ibase += kCells * kBinReduceThreads) {
    Cplx<T> v[kCells][kCombos];
    int tpos[kCells][kCombos], nidx[kCells][kCombos];
    bool ok[kCells][kCombos];
#pragma unroll
    for (int u = 0; u < kCells; ++u) {
      const int i = ibase + u * kBinReduceThreads;
      const bool live = i < cells;
      const int li[3] = {i % w[0], DIM >= 2 ? (i / w[0]) % w[1] : 0,
                         DIM >= 3 ? i / (w[0] * w[1]) : 0};
      // per dimension: own position, and the covering neighbour (-1, +1 or
      // none = 0) with its position; regular bins never have both
      int pown[3] = {0, 0, 0}, poth[3] = {0, 0, 0}, ooth[3] = {0, 0, 0};
#pragma unroll
      for (int d = 0; d < DIM; ++d) {
        pown[d] = li[d] + h;
        if (li[d] < ns - h) {
          ooth[d] = -1;
          poth[d] = li[d] + wprev[d] + h;
        } else if (li[d] >= w[d] - h) {
          ooth[d] = 1;
          poth[d] = li[d] - w[d] + h;
        }
      }
#pragma unroll
      for (int k = 0; k < kCombos; ++k) {
        bool good = live;
        int p[3] = {0, 0, 0}, nb = 0, mult = 1;
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
          const bool other = (k >> d) & 1;
          good = good && (!other || ooth[d] != 0);
          p[d] = other ? poth[d] : pown[d];
          nb += ((other ? ooth[d] : 0) + 1) * mult;
          mult *= 3;
        }
        ok[u][k] = good;
        nidx[u][k] = nb;
        tpos[u][k] = p[0] + g.tile[0] * (p[1] + g.tile[1] * p[2]);
      }
    }
#pragma unroll
    for (int u = 0; u < kCells; ++u) {
#pragma unroll
      for (int k = 0; k < kCombos; ++k) {
        v[u][k] = Cplx<T>{T(0), T(0)};
        if (ok[u][k]) {
          const uint32_t s0 = rng[nidx[u][k]][0];
          if (rng[nidx[u][k]][1] > s0) {
            v[u][k] = job.tiles[static_cast<std::size_t>(tpos[u][k]) +
                                static_cast<std::size_t>(s0) * tc];
          }
        }
      }
    }
#pragma unroll
    for (int u = 0; u < kCells; ++u) {
      const int i = ibase + u * kBinReduceThreads;
      if (i >= cells) continue;
      T ar = T(0), ai = T(0);
#pragma unroll
      for (int k = 0; k < kCombos; ++k) {
        ar += v[u][k].x;
        ai += v[u][k].y;
        if (ok[u][k]) {  // bins split into several subproblems
          const uint32_t s1 = rng[nidx[u][k]][1];
          for (uint32_t s = rng[nidx[u][k]][0] + 1; s < s1; ++s) {
            const Cplx<T> x = job.tiles[static_cast<std::size_t>(tpos[u][k]) +
                                        static_cast<std::size_t>(s) * tc];
            ar += x.x;
            ai += x.y;
          }
        }
      }
      const int li[3] = {i % w[0], DIM >= 2 ? (i / w[0]) % w[1] : 0,
                         DIM >= 3 ? i / (w[0] * w[1]) : 0};
      const std::size_t c =
          static_cast<std::size_t>(b[0] * g.bin[0] + li[0]) +
          static_cast<std::size_t>(g.n[0]) *
              (static_cast<std::size_t>(b[1] * g.bin[1] + li[1]) +
               static_cast<std::size_t>(g.n[1]) *
                   static_cast<std::size_t>(b[2] * g.bin[2] + li[2]));
      if (accumulate) {
        const Cplx<T> a = job.grid[c];
        ar += a.x;
        ai += a.y;
      }
      job.grid[c] = Cplx<T>{ar, ai};
    }
  }
  __syncthreads();  // rng is reused by the next bin
}

template <typename T, int DIM>
__global__ void __launch_bounds__(kBinReduceThreads)
    bin_reduce_kernel(ReduceJob<T, DIM> job, int ns, int accumulate) {
  __shared__ uint32_t rng[kNeigh<DIM>][2];
  const uint32_t nbins = static_cast<uint32_t>(job.g.nbins[0]) * job.g.nbins[1] * job.g.nbins[2];
  // neighbouring bins on one XCD so their shared tile lines meet in one L2
  for (uint32_t bin = xcd_remap(blockIdx.x, gridDim.x); bin < nbins; bin += gridDim.x) {
    reduce_bin<T, DIM>(job, ns, bin, rng, accumulate != 0);
  }
}

// blockIdx.y selects the job.
template <typename T, int DIM>
__global__ void __launch_bounds__(kBinReduceThreads)
    bin_reduce_batch_kernel(const ReduceJob<T, DIM>* __restrict__ jobs, int ns, int accumulate) {
  __shared__ uint32_t rng[kNeigh<DIM>][2];
  const ReduceJob<T, DIM> job = jobs[blockIdx.y];
  const uint32_t nbins = static_cast<uint32_t>(job.g.nbins[0]) * job.g.nbins[1] * job.g.nbins[2];
  for (uint32_t bin = blockIdx.x; bin < nbins; bin += gridDim.x) {
    reduce_bin<T, DIM>(job, ns, bin, rng, accumulate != 0);
  }
}

// blockIdx.y strides over the rows (y, z) of the grid, blockIdx.x over x.
template <typename T, int DIM>
__device__ inline void reduce_grid(const ReduceJob<T, DIM>& job, bool accumulate) {
  const int rows = job.g.n[1] * job.g.n[2];
  const int xbegin = threadIdx.x + blockIdx.x * kReduceThreads;
  const int xstride = gridDim.x * kReduceThreads;
  for (int row = blockIdx.y; row < rows; row += gridDim.y) {
    const int c1 = DIM >= 2 ? row % job.g.n[1] : 0;
    const int c2 = DIM >= 3 ? row / job.g.n[1] : 0;
    reduce_row<T, DIM>(job, c1, c2, xbegin, xstride, accumulate);
  }
}

template <typename T, int DIM>
__global__ void __launch_bounds__(kReduceThreads)
    tile_reduce_kernel(ReduceJob<T, DIM> job, int accumulate) {
  reduce_grid<T, DIM>(job, accumulate != 0);
}

// blockIdx.z selects the job.
template <typename T, int DIM>
__global__ void __launch_bounds__(kReduceThreads)
    tile_reduce_batch_kernel(const ReduceJob<T, DIM>* __restrict__ jobs, int accumulate) {
  const ReduceJob<T, DIM> job = jobs[blockIdx.z];
  reduce_grid<T, DIM>(job, accumulate != 0);
}

// Launch shape of a reduction over grids with x extent n0 and `rows` rows:
// enough x blocks to cover a row (at most 8 unless there are few rows, as in
// 1D), rows spread over y.
inline dim3 reduce_shape(int n0, int rows, unsigned zcount, unsigned max_y) {
  unsigned gx = static_cast<unsigned>((n0 + kReduceThreads - 1) / kReduceThreads);
  const unsigned gx_cap = rows < 64 ? 4096u : 8u;
  if (gx > gx_cap) gx = gx_cap;
  unsigned gy = static_cast<unsigned>(rows);
  if (gy > max_y) gy = max_y;
  if (gy < 1u) gy = 1u;
  return dim3(gx, gy, zcount);
}

}  // namespace

template <typename T>
std::size_t tile_bytes(const BinGeom& g) {
  return static_cast<std::size_t>(g.tile[0]) * g.tile[1] * g.tile[2] * sizeof(Cplx<T>);
}

#define NNF_NS_CASES(FN, ...)                                    \
  switch (ns) {                                                  \
    case 2: FN<T, DIM, 2>(__VA_ARGS__); break;                   \
    case 3: FN<T, DIM, 3>(__VA_ARGS__); break;                   \
    case 4: FN<T, DIM, 4>(__VA_ARGS__); break;                   \
    case 5: FN<T, DIM, 5>(__VA_ARGS__); break;                   \
    case 6: FN<T, DIM, 6>(__VA_ARGS__); break;                   \
    case 7: FN<T, DIM, 7>(__VA_ARGS__); break;                   \
    case 8: FN<T, DIM, 8>(__VA_ARGS__); break;                   \
    case 9: FN<T, DIM, 9>(__VA_ARGS__); break;                   \
    case 10: FN<T, DIM, 10>(__VA_ARGS__); break;                 \
    case 11: FN<T, DIM, 11>(__VA_ARGS__); break;                 \
    case 12: FN<T, DIM, 12>(__VA_ARGS__); break;                 \
    case 13: FN<T, DIM, 13>(__VA_ARGS__); break;                 \
    case 14: FN<T, DIM, 14>(__VA_ARGS__); break;                 \
    case 15: FN<T, DIM, 15>(__VA_ARGS__); break;                 \
    case 16: FN<T, DIM, 16>(__VA_ARGS__); break;                 \
    default: throw neonufft::InternalError("n_spread not in [2, 16]"); \
  }

template <typename T, int DIM>
void launch_spread(int ns, const PointRec<T, DIM>* pts, const uint32_t* perm, const Subprob* subs,
                   std::size_t nsub, const Cplx<T>* in, const Cplx<T>* pre_sorted, Cplx<T>* tiles,
                   const BinGeom& g, const EsParams& es, hipStream_t stream) {
  if (nsub == 0) return;
  const SpreadJob<T, DIM> job{pts, perm, in, pre_sorted, tiles, 0u, 0u, g};
  NNF_NS_CASES(spread_ns, job, subs, nsub, es, stream)
}

template <typename T, int DIM>
void launch_interp(int ns, const PointRec<T, DIM>* pts, const uint32_t* perm, const Subprob* subs,
                   std::size_t nsub, const Cplx<T>* grid, const Cplx<T>* post_sorted, Cplx<T>* out,
                   const BinGeom& g, const EsParams& es, hipStream_t stream) {
  if (nsub == 0) return;
  const InterpJob<T, DIM> job{pts, perm, grid, post_sorted, out, g};
  NNF_NS_CASES(interp_ns, job, subs, nsub, es, stream)
}

template <typename T, int DIM>
void launch_spread_batch(int ns, const SpreadJob<T, DIM>* jobs, const Subprob* subs,
                         std::size_t nsub, std::size_t lds_bytes, const EsParams& es,
                         hipStream_t stream) {
  if (nsub == 0) return;
  NNF_NS_CASES(spread_batch_ns, jobs, subs, nsub, lds_bytes, es, stream)
}

template <typename T, int DIM>
void launch_interp_batch(int ns, const InterpJob<T, DIM>* jobs, const Subprob* subs,
                         std::size_t nsub, std::size_t lds_bytes, const EsParams& es,
                         hipStream_t stream) {
  if (nsub == 0) return;
  NNF_NS_CASES(interp_batch_ns, jobs, subs, nsub, lds_bytes, es, stream)
}

template <typename T, int DIM>
void launch_tile_reduce(const ReduceJob<T, DIM>& job, int ns, bool accumulate,
                        hipStream_t stream) {
  if (bins_regular(DIM, job.g, ns)) {
    const std::size_t nbins =
        static_cast<std::size_t>(job.g.nbins[0]) * job.g.nbins[1] * job.g.nbins[2];
    const unsigned blocks = static_cast<unsigned>(std::min<std::size_t>(nbins, 65535u * 16u));
    hipLaunchKernelGGL((bin_reduce_kernel<T, DIM>), dim3(blocks), dim3(kBinReduceThreads), 0,
                       stream, job, ns, accumulate ? 1 : 0);
    check_launch("bin_reduce_kernel");
    return;
  }
  const dim3 shape = reduce_shape(job.g.n[0], job.g.n[1] * job.g.n[2], 1u, 65535u);
  hipLaunchKernelGGL((tile_reduce_kernel<T, DIM>), shape, dim3(kReduceThreads), 0, stream, job,
                     accumulate ? 1 : 0);
  check_launch("tile_reduce_kernel");
}

template <typename T, int DIM>
void launch_tile_reduce_batch(const ReduceJob<T, DIM>* jobs, int count, int ns, bool regular,
                              std::size_t max_bins, int max_n0, int max_rows, bool accumulate,
                              hipStream_t stream) {
  if (count <= 0) return;
  if (count > 65535) throw neonufft::InputError("Neonufft:: too many transforms in one batch");
  if (regular) {
    const unsigned blocks = static_cast<unsigned>(std::min<std::size_t>(max_bins, 4096u));
    hipLaunchKernelGGL((bin_reduce_batch_kernel<T, DIM>),
                       dim3(blocks, static_cast<unsigned>(count)), dim3(kBinReduceThreads), 0,
                       stream, jobs, ns, accumulate ? 1 : 0);
    check_launch("bin_reduce_batch_kernel");
    return;
  }
  const dim3 shape = reduce_shape(max_n0, max_rows, static_cast<unsigned>(count), 1024u);
  hipLaunchKernelGGL((tile_reduce_batch_kernel<T, DIM>), shape, dim3(kReduceThreads), 0, stream,
                     jobs, accumulate ? 1 : 0);
  check_launch("tile_reduce_batch_kernel");
}

#define NNF_INST_SI(T, DIM)                                                                      \
  template void launch_spread<T, DIM>(int, const PointRec<T, DIM>*, const uint32_t*,             \
                                      const Subprob*, std::size_t, const Cplx<T>*,               \
                                      const Cplx<T>*, Cplx<T>*, const BinGeom&, const EsParams&, \
                                      hipStream_t);                                              \
  template void launch_interp<T, DIM>(int, const PointRec<T, DIM>*, const uint32_t*,             \
                                      const Subprob*, std::size_t, const Cplx<T>*,               \
                                      const Cplx<T>*, Cplx<T>*, const BinGeom&, const EsParams&, \
                                      hipStream_t);                                              \
  template void launch_spread_batch<T, DIM>(int, const SpreadJob<T, DIM>*, const Subprob*,       \
                                            std::size_t, std::size_t, const EsParams&,           \
                                            hipStream_t);                                        \
  template void launch_interp_batch<T, DIM>(int, const InterpJob<T, DIM>*, const Subprob*,       \
                                            std::size_t, std::size_t, const EsParams&,           \
                                            hipStream_t);                                        \
  template void launch_tile_reduce<T, DIM>(const ReduceJob<T, DIM>&, int, bool, hipStream_t);   \
  template void launch_tile_reduce_batch<T, DIM>(const ReduceJob<T, DIM>*, int, int, bool,       \
                                                 std::size_t, int, int, bool, hipStream_t);

NNF_INST_SI(float, 1)
NNF_INST_SI(float, 2)
NNF_INST_SI(float, 3)
NNF_INST_SI(double, 1)
NNF_INST_SI(double, 2)
NNF_INST_SI(double, 3)

template std::size_t tile_bytes<float>(const BinGeom&);
template std::size_t tile_bytes<double>(const BinGeom&);

}  // namespace nnf
