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
//  - spreading accumulates with LDS atomics and flushes the tile to HBM once
//    with coalesced float atomics; interpolation loads the tile once with
//    coalesced reads and gathers from LDS;
//  - kernel weights come from per-piece Horner polynomials in registers (fp32)
//    or the exact exp/sqrt formula (fp64);
//  - subproblem -> workgroup mapping keeps neighbouring bins on one XCD.
#include "kernels.hpp"

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

// Map the hardware workgroup id to a subproblem so that consecutive
// subproblems (neighbouring bins) run on the same XCD: hardware dispatches
// workgroup b to XCD b % 8.
__device__ inline uint32_t xcd_remap(uint32_t b, uint32_t n) {
  const uint32_t n8 = n & ~7u;
  if (b >= n8) return b;
  const uint32_t per = n8 >> 3;
  return (b & 7u) * per + (b >> 3);
}

__device__ inline int wrap_once(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

// Weight of kernel piece j (cell i0 + j) for Horner argument x in [0,1):
// phi(x + j - NS/2).
template <int NS>
__device__ inline float es_piece(float x, int j, const EsParams& es) {
  constexpr int DEG = horner_degree(NS);
  const float y = fmaf(2.0f, x, -1.0f);
  const float* c = es.coef + j * (DEG + 1);
  float v = c[DEG];
#pragma unroll
  for (int d = DEG - 1; d >= 0; --d) v = fmaf(v, y, c[d]);
  return v;
}

template <int NS>
__device__ inline double es_piece(double x, int j, const EsParams& es) {
  constexpr double c = 4.0 / (NS * NS);
  const double t = x + static_cast<double>(j) - 0.5 * NS;
  const double a = 1.0 - c * t * t;
  return a > 0.0 ? exp(es.beta * (sqrt(a) - 1.0)) : 0.0;
}

template <typename T, int NS>
__device__ inline void es_weights(T x, const EsParams& es, T (&k)[NS]) {
#pragma unroll
  for (int j = 0; j < NS; ++j) k[j] = es_piece<NS>(x, j, es);
}

template <int DIM>
__device__ inline int unpack_cell(uint32_t cell, int d) {
  if constexpr (CellPack<DIM>::bits == 32) {
    return static_cast<int>(cell);
  } else {
    return static_cast<int>((cell >> (d * CellPack<DIM>::bits)) & CellPack<DIM>::mask);
  }
}

__device__ inline void hbm_add(float* p, float v) {
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void hbm_add(double* p, double v) {
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

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

// ---------------------------------------------------------------------------
// Spreading
//
// The tile accumulates in fixed point with integer LDS atomics: on gfx950,
// ds_add_u32 sustains 5-9e12 lane-ops/s while ds_add_f32 runs at ~2e11.
// Each subproblem picks its own scale from a rigorous bound on any tile value,
// bound = sum over its points of |Re c| + |Im c| (kernel weights are <= 1), so
// fp32 plans keep 30 fractional-range bits in int32 and fp64 plans 62 bits in
// int64. The flush converts back and adds to the grid with float atomics.
// ---------------------------------------------------------------------------
template <typename T>
struct FixedAcc;
template <>
struct FixedAcc<float> {
  using type = int;
  static constexpr int bits = 30;
};
template <>
struct FixedAcc<double> {
  using type = long long;
  static constexpr int bits = 62;
};

// Round to the nearest integer (floor(v + 0.5)) in one VALU instruction; the
// compiler emits rndne + cvt (two) for __float2int_rn. |v| <= 2^30 here.
__device__ inline int to_fixed(float v) {
  int r;
  asm("v_cvt_rpi_i32_f32 %0, %1" : "=v"(r) : "v"(v));
  return r;
}
__device__ inline long long to_fixed(double v) { return __double2ll_rn(v); }

__device__ inline void lds_add_fixed(int* p, int v) {
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ inline void lds_add_fixed(long long* p, long long v) {
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

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
                                                  const Subprob& sp, const EsParams& es,
                                                  unsigned char* smem_raw, T* wsum) {
  using Acc = typename FixedAcc<T>::type;
  const PointRec<T, DIM>* __restrict__ pts = job.pts;
  const uint32_t* __restrict__ perm = job.perm;
  const Cplx<T>* __restrict__ in = job.in;
  const Cplx<T>* __restrict__ pre = job.pre;
  Cplx<T>* __restrict__ grid = job.grid;
  const BinGeom& g = job.g;
  const TileInfo<DIM, NS> ti(sp.bin, g);
  const int tile_cells = g.tile[0] * (DIM >= 2 ? g.tile[1] : 1) * (DIM >= 3 ? g.tile[2] : 1);
  // Real and imaginary parts live in separate planes so that 32-bit atomics
  // of neighbouring cells fall in neighbouring banks.
  Acc* tre = reinterpret_cast<Acc*>(smem_raw);
  Acc* tim = tre + tile_cells;
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
      Cplx<T> c = in[perm[p]];
      if (pre) {
        const Cplx<T> q = pre[p];
        c = Cplx<T>{c.x * q.x - c.y * q.y, c.x * q.y + c.y * q.x};
      }
      cv[k] = c;
      bound += fabs(c.x) + fabs(c.y);
    }
  }
  const T total = block_sum(bound, wsum);  // also orders the tile zero fill
  const T one = static_cast<T>(1ull << FixedAcc<T>::bits);
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
        lds_add_fixed(&tre[r0 + a], to_fixed(kx[a] * c.x));
        lds_add_fixed(&tim[r0 + a], to_fixed(kx[a] * c.y));
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
#pragma unroll
          for (int a = 0; a < NS; ++a) {
            lds_add_fixed(&tre[base + a], to_fixed(kx[a] * vr));
            lds_add_fixed(&tim[base + a], to_fixed(kx[a] * vi));
          }
        }
      }
    }
  }
  __syncthreads();

  // Flush the tile: rows along x are added to the grid with coalesced float
  // atomics (cheap at this access pattern).
  const T inv = total > T(0) ? total / one : T(0);
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int rows = ti.ext[1] * ti.ext[2];
  for (int r = wave; r < rows; r += kThreads / 64) {
    const int i1 = r % ti.ext[1];
    const int i2 = r / ti.ext[1];
    const int gy = DIM >= 2 ? wrap_once(ti.org[1] + i1, g.n[1]) : 0;
    const int gz = DIM >= 3 ? wrap_once(ti.org[2] + i2, g.n[2]) : 0;
    const std::size_t grow =
        (static_cast<std::size_t>(gz) * static_cast<std::size_t>(g.n[1]) + gy) * g.n[0];
    const int trow = (i2 * ti.t1 + i1) * ti.t0;
    for (int i0 = lane; i0 < ti.ext[0]; i0 += 64) {
      const Acc qr = tre[trow + i0];
      const Acc qi = tim[trow + i0];
      if (qr != Acc(0) || qi != Acc(0)) {
        T* dst = reinterpret_cast<T*>(grid + grow + wrap_once(ti.org[0] + i0, g.n[0]));
        hbm_add(dst, static_cast<T>(qr) * inv);
        hbm_add(dst + 1, static_cast<T>(qi) * inv);
      }
    }
  }
}

template <typename T, int DIM, int NS>
__global__ void __launch_bounds__(kThreads)
    spread_kernel(SpreadJob<T, DIM> job, const Subprob* __restrict__ subs, uint32_t nsub,
                  EsParams es) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  __shared__ T wsum[kThreads / 64];
  const Subprob sp = subs[xcd_remap(blockIdx.x, nsub)];
  spread_subproblem<T, DIM, NS>(job, sp, es, smem_raw, wsum);
}

// Many independent transforms in one launch: each subproblem names its job.
template <typename T, int DIM, int NS>
__global__ void __launch_bounds__(kThreads)
    spread_batch_kernel(const SpreadJob<T, DIM>* __restrict__ jobs,
                        const Subprob* __restrict__ subs, uint32_t nsub, EsParams es) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  __shared__ T wsum[kThreads / 64];
  const Subprob sp = subs[xcd_remap(blockIdx.x, nsub)];
  const SpreadJob<T, DIM> job = jobs[sp.job];
  spread_subproblem<T, DIM, NS>(job, sp, es, smem_raw, wsum);
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

  // Stage the tile (bin + halo) from the grid with coalesced row reads.
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int rows = ti.ext[1] * ti.ext[2];
  for (int r = wave; r < rows; r += kInterpThreads<T> / 64) {
    const int i1 = r % ti.ext[1];
    const int i2 = r / ti.ext[1];
    const int gy = DIM >= 2 ? wrap_once(ti.org[1] + i1, g.n[1]) : 0;
    const int gz = DIM >= 3 ? wrap_once(ti.org[2] + i2, g.n[2]) : 0;
    const std::size_t grow =
        (static_cast<std::size_t>(gz) * static_cast<std::size_t>(g.n[1]) + gy) * g.n[0];
    const int trow = (i2 * ti.t1 + i1) * ti.t0;
    for (int i0 = lane; i0 < ti.ext[0]; i0 += 64) {
      tile[trow + i0] = grid[grow + wrap_once(ti.org[0] + i0, g.n[0])];
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
                   std::size_t nsub, const Cplx<T>* in, const Cplx<T>* pre_sorted, Cplx<T>* grid,
                   const BinGeom& g, const EsParams& es, hipStream_t stream) {
  if (nsub == 0) return;
  const SpreadJob<T, DIM> job{pts, perm, in, pre_sorted, grid, g};
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
                                            hipStream_t);

NNF_INST_SI(float, 1)
NNF_INST_SI(float, 2)
NNF_INST_SI(float, 3)
NNF_INST_SI(double, 1)
NNF_INST_SI(double, 2)
NNF_INST_SI(double, 3)

template std::size_t tile_bytes<float>(const BinGeom&);
template std::size_t tile_bytes<double>(const BinGeom&);

}  // namespace nnf
