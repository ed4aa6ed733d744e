// Fine-grid kernels: zero fill, deconvolution factors, up/down-sampling
// (deconvolve + reorder), and the type-3 phase factors.
//
// Semantics: fseries_kernel.cu:24-67, downsample_kernel.cu:133-191,
// upsample_kernel.cu:18-196, prephase_kernel.cu:35-50, postphase_kernel.cu:27-76
// of the reference (src/neonufft/gpu/kernels/). Differences by design:
// deconvolution factors and phases are computed in FP64 (then rounded to T),
// the upsampling kernel writes every fine-grid cell (zero padding included) so
// no separate memset pass is needed, and the type-3 phases are stored in the
// points' bin order so the spreading / interpolation kernels read them
// contiguously.
#include "kernels.hpp"

namespace nnf {

namespace {

constexpr int kBlock = 256;

template <typename T>
struct FsPtrs {
  const T* p[3];
};


// Zeroes n4 16-byte units and, if tail8 is set, one trailing 8-byte unit.
__global__ void __launch_bounds__(kBlock)
    fill_zero_kernel(float4* __restrict__ p, std::size_t n4, float2* __restrict__ tail8) {
  for (std::size_t i = threadIdx.x + static_cast<std::size_t>(blockIdx.x) * kBlock; i < n4;
       i += static_cast<std::size_t>(gridDim.x) * kBlock) {
    p[i] = make_float4(0.f, 0.f, 0.f, 0.f);
  }
  if (tail8 && blockIdx.x == 0 && threadIdx.x == 0) *tail8 = make_float2(0.f, 0.f);
}

__device__ inline double phihat_eval(double f, const QuadRule& r) {
  double s = 0.0;
  for (int q = 0; q < r.nq; ++q) s += r.weight[q] * cos(f * r.node[q]);
  return s;
}

template <typename T>
__global__ void __launch_bounds__(kBlock) fseries_kernel(T* __restrict__ fs, IntType n, QuadRule r) {
  const IntType nout = n / 2 + 1;
  for (IntType k = threadIdx.x + static_cast<IntType>(blockIdx.x) * kBlock; k < nout;
       k += static_cast<IntType>(gridDim.x) * kBlock) {
    const double f = 2.0 * kPi * static_cast<double>(k) / static_cast<double>(n);
    fs[k] = static_cast<T>(((k & 1) ? -1.0 : 1.0) / phihat_eval(f, r));
  }
}

// Up/down-sampling geometry in 32-bit per-axis form: fine cells n, modes N,
// mode-array element strides st (unused axes: n = N = 1).
struct SampleGeom {
  int n[3];
  int N[3];
  long long st[3];
};

// Mode of fine cell g on one axis (k = g for g < N - N/2, k = g - n for
// g >= n - N/2), or false for a padding cell.
__device__ inline bool cell_mode32(int g, int n, int N, int& k) {
  if (g < N - N / 2) {
    k = g;
    return true;
  }
  if (g >= n - N / 2) {
    k = g - n;
    return true;
  }
  return false;
}

// Array index of mode k (CMCL: k + N/2; FFT order: k, negatives after).
__device__ inline int mode_index32(int k, int N, int order_fft) {
  return order_fft ? (k >= 0 ? k : k + N) : k + N / 2;
}

// Value of fine cell (g0, g1, g2) after upsampling: f[k] * prod fs_d[|k_d|] for
// a mode cell, 0 for padding.
template <typename T, int DIM>
__device__ inline Cplx<T> upsample_cell(int order_fft, const Cplx<T>* __restrict__ in,
                                        const SampleGeom& g, const FsPtrs<T>& fs, int g0, int g1,
                                        int g2) {
  int k0, k1 = 0, k2 = 0;
  if (!cell_mode32(g0, g.n[0], g.N[0], k0)) return Cplx<T>{T(0), T(0)};
  if (DIM >= 2 && !cell_mode32(g1, g.n[1], g.N[1], k1)) return Cplx<T>{T(0), T(0)};
  if (DIM >= 3 && !cell_mode32(g2, g.n[2], g.N[2], k2)) return Cplx<T>{T(0), T(0)};
  long long off = static_cast<long long>(mode_index32(k0, g.N[0], order_fft)) * g.st[0];
  T f = fs.p[0][k0 < 0 ? -k0 : k0];
  if (DIM >= 2) {
    off += static_cast<long long>(mode_index32(k1, g.N[1], order_fft)) * g.st[1];
    f *= fs.p[1][k1 < 0 ? -k1 : k1];
  }
  if (DIM >= 3) {
    off += static_cast<long long>(mode_index32(k2, g.N[2], order_fft)) * g.st[2];
    f *= fs.p[2][k2 < 0 ? -k2 : k2];
  }
  const Cplx<T> a = in[off];
  return Cplx<T>{a.x * f, a.y * f};
}

__device__ inline void store_pair(Cplx<float>* p, const Cplx<float>& a, const Cplx<float>& b) {
  *reinterpret_cast<float4*>(p) = make_float4(a.x, a.y, b.x, b.y);
}
__device__ inline void store_pair(Cplx<double>* p, const Cplx<double>& a, const Cplx<double>& b) {
  p[0] = a;
  p[1] = b;
}

// Writes every cell of one fine grid (threads tid, tid + nthr, ...). With an
// even x extent two neighbouring cells go out in one 16-byte (fp32) store.
template <typename T, int DIM>
__device__ inline void upsample_grid(int order_fft, const Cplx<T>* __restrict__ in,
                                     const SampleGeom& g, const FsPtrs<T>& fs,
                                     Cplx<T>* __restrict__ grid, std::size_t tid,
                                     std::size_t nthr) {
  const std::size_t n0 = static_cast<std::size_t>(g.n[0]);
  const std::size_t cells = n0 * g.n[1] * g.n[2];
  if ((n0 & 1) == 0) {
    for (std::size_t q = tid; 2 * q < cells; q += nthr) {
      const std::size_t idx = 2 * q;
      const std::size_t row = idx / n0;
      const int g0 = static_cast<int>(idx - row * n0);
      const int g1 = DIM >= 2 ? static_cast<int>(row % g.n[1]) : 0;
      const int g2 = DIM >= 3 ? static_cast<int>(row / g.n[1]) : 0;
      store_pair(grid + idx, upsample_cell<T, DIM>(order_fft, in, g, fs, g0, g1, g2),
                 upsample_cell<T, DIM>(order_fft, in, g, fs, g0 + 1, g1, g2));
    }
  } else {
    for (std::size_t idx = tid; idx < cells; idx += nthr) {
      const std::size_t row = idx / n0;
      const int g0 = static_cast<int>(idx - row * n0);
      const int g1 = DIM >= 2 ? static_cast<int>(row % g.n[1]) : 0;
      const int g2 = DIM >= 3 ? static_cast<int>(row / g.n[1]) : 0;
      grid[idx] = upsample_cell<T, DIM>(order_fft, in, g, fs, g0, g1, g2);
    }
  }
}

// Type-1 deconvolution of every mode of one transform: out[i] (strided) =
// grid[k mod n] * prod fs_d[|k_d|].
template <typename T, int DIM>
__device__ inline void downsample_grid(int order_fft, const Cplx<T>* __restrict__ grid,
                                       const SampleGeom& g, const FsPtrs<T>& fs,
                                       Cplx<T>* __restrict__ out, std::size_t tid,
                                       std::size_t nthr) {
  const std::size_t N0 = static_cast<std::size_t>(g.N[0]);
  const std::size_t modes = N0 * g.N[1] * g.N[2];
  for (std::size_t idx = tid; idx < modes; idx += nthr) {
    const std::size_t row = idx / N0;
    const int i[3] = {static_cast<int>(idx - row * N0),
                      DIM >= 2 ? static_cast<int>(row % g.N[1]) : 0,
                      DIM >= 3 ? static_cast<int>(row / g.N[1]) : 0};
    long long goff = 0, ooff = 0, gmul = 1;
    T f = T(1);
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      const int N = g.N[d];
      const int k = order_fft ? (i[d] < N - N / 2 ? i[d] : i[d] - N) : i[d] - N / 2;
      f *= fs.p[d][k < 0 ? -k : k];
      goff += static_cast<long long>(k < 0 ? k + g.n[d] : k) * gmul;
      gmul *= g.n[d];
      ooff += static_cast<long long>(i[d]) * g.st[d];
    }
    const Cplx<T> a = grid[goff];
    out[ooff] = Cplx<T>{a.x * f, a.y * f};
  }
}

template <typename T, int DIM>
__global__ void __launch_bounds__(kBlock)
    upsample_kernel(int order_fft, const Cplx<T>* __restrict__ in, SampleGeom g, FsPtrs<T> fs,
                    Cplx<T>* __restrict__ grid) {
  upsample_grid<T, DIM>(order_fft, in, g, fs, grid,
                        threadIdx.x + static_cast<std::size_t>(blockIdx.x) * kBlock,
                        static_cast<std::size_t>(gridDim.x) * kBlock);
}

template <typename T, int DIM>
__global__ void __launch_bounds__(kBlock)
    downsample_kernel(int order_fft, const Cplx<T>* __restrict__ grid, SampleGeom g,
                      FsPtrs<T> fs, Cplx<T>* __restrict__ out) {
  downsample_grid<T, DIM>(order_fft, grid, g, fs, out,
                          threadIdx.x + static_cast<std::size_t>(blockIdx.x) * kBlock,
                          static_cast<std::size_t>(gridDim.x) * kBlock);
}

// Batched variants: blockIdx.y selects one of the transforms of equal mode and
// fine-grid shape whose grids are contiguous with stride `gstride` cells.
template <typename T, int DIM>
__global__ void __launch_bounds__(kBlock)
    upsample_batch_kernel(int order_fft, const Cplx<T>* const* __restrict__ ins, SampleGeom g,
                          FsPtrs<T> fs, Cplx<T>* __restrict__ grids, std::size_t gstride) {
  upsample_grid<T, DIM>(order_fft, ins[blockIdx.y], g, fs, grids + gstride * blockIdx.y,
                        threadIdx.x + static_cast<std::size_t>(blockIdx.x) * kBlock,
                        static_cast<std::size_t>(gridDim.x) * kBlock);
}

template <typename T, int DIM>
__global__ void __launch_bounds__(kBlock)
    downsample_batch_kernel(int order_fft, const Cplx<T>* __restrict__ grids,
                            std::size_t gstride, SampleGeom g, FsPtrs<T> fs,
                            Cplx<T>* const* __restrict__ outs) {
  downsample_grid<T, DIM>(order_fft, grids + gstride * blockIdx.y, g, fs, outs[blockIdx.y],
                          threadIdx.x + static_cast<std::size_t>(blockIdx.x) * kBlock,
                          static_cast<std::size_t>(gridDim.x) * kBlock);
}

struct Vec3d {
  double v[3];
};

template <typename TC, typename T, int DIM>
__global__ void __launch_bounds__(kBlock)
    prephase_kernel(int sign, Loc<TC> loc, const uint32_t* __restrict__ perm, std::size_t m,
                    Vec3d sbar, Cplx<T>* __restrict__ pre) {
  for (std::size_t p = threadIdx.x + static_cast<std::size_t>(blockIdx.x) * kBlock; p < m;
       p += static_cast<std::size_t>(gridDim.x) * kBlock) {
    const uint32_t i = perm[p];
    double ph = 0.0;
#pragma unroll
    for (int d = 0; d < DIM; ++d) ph += static_cast<double>(loc.p[d][i]) * sbar.v[d];
    double s, c;
    sincos(ph, &s, &c);
    pre[p] = Cplx<T>{static_cast<T>(c), static_cast<T>(sign * s)};
  }
}

template <typename TC, typename T, int DIM>
__global__ void __launch_bounds__(kBlock)
    postphase_kernel(int sign, Loc<TC> loc, const uint32_t* __restrict__ perm, std::size_t m,
                     Vec3d xbar, Vec3d sbar, Vec3d scale, QuadRule r, Cplx<T>* __restrict__ post) {
  for (std::size_t p = threadIdx.x + static_cast<std::size_t>(blockIdx.x) * kBlock; p < m;
       p += static_cast<std::size_t>(gridDim.x) * kBlock) {
    const uint32_t i = perm[p];
    double ph = 0.0, phihat = 1.0;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      const double t = static_cast<double>(loc.p[d][i]) - sbar.v[d];
      ph += t * xbar.v[d];
      phihat *= phihat_eval(t * scale.v[d], r);
    }
    double s, c;
    sincos(ph, &s, &c);
    post[p] = Cplx<T>{static_cast<T>(c / phihat), static_cast<T>(sign * s / phihat)};
  }
}

// Complex precision conversion (mixed-precision type-3 I/O), grid-stride.
template <typename TI, typename TO>
__global__ void __launch_bounds__(kBlock)
    convert_kernel(const Cplx<TI>* __restrict__ in, Cplx<TO>* __restrict__ out, std::size_t m) {
  for (std::size_t i = threadIdx.x + static_cast<std::size_t>(blockIdx.x) * kBlock; i < m;
       i += static_cast<std::size_t>(gridDim.x) * kBlock) {
    const Cplx<TI> v = in[i];
    out[i] = Cplx<TO>{static_cast<TO>(v.x), static_cast<TO>(v.y)};
  }
}

inline unsigned blocks_for(std::size_t work, unsigned cap = 65535u * 4u) {
  std::size_t b = (work + kBlock - 1) / kBlock;
  if (b < 1) b = 1;
  if (b > cap) b = cap;
  return static_cast<unsigned>(b);
}

}  // namespace

template <typename T>
void launch_fill_zero(Cplx<T>* p, std::size_t count, hipStream_t stream) {
  // A kernel rather than hipMemsetAsync so that the transforms stay safe to
  // capture in a HIP graph.
  if (count == 0) return;
  const std::size_t bytes = count * sizeof(Cplx<T>);
  const std::size_t n4 = bytes / 16;
  float2* tail = (bytes % 16) ? reinterpret_cast<float2*>(reinterpret_cast<char*>(p) + n4 * 16)
                              : nullptr;
  hipLaunchKernelGGL(fill_zero_kernel, dim3(blocks_for(n4 ? n4 : 1)), dim3(kBlock), 0, stream,
                     reinterpret_cast<float4*>(p), n4, tail);
  check_launch("fill_zero_kernel");
}

template <typename T>
void launch_fseries(T* fs, IntType n, const QuadRule& rule, hipStream_t stream) {
  hipLaunchKernelGGL((fseries_kernel<T>), dim3(blocks_for(static_cast<std::size_t>(n / 2 + 1))),
                     dim3(kBlock), 0, stream, fs, n, rule);
  check_launch("fseries_kernel");
}

template <typename T, int DIM>
SampleGeom sample_geom(const IntType* n, const IntType* modes, const IntType* strides) {
  SampleGeom g{{1, 1, 1}, {1, 1, 1}, {0, 0, 0}};
  for (int d = 0; d < DIM; ++d) {
    g.n[d] = static_cast<int>(n[d]);
    g.N[d] = static_cast<int>(modes[d]);
    g.st[d] = static_cast<long long>(strides[d]);
  }
  return g;
}

template <typename T, int DIM>
FsPtrs<T> fs_ptrs(T* const* fs) {
  FsPtrs<T> f{{nullptr, nullptr, nullptr}};
  for (int d = 0; d < DIM; ++d) f.p[d] = fs[d];
  return f;
}

inline std::size_t up_work(const SampleGeom& g) {
  const std::size_t cells = static_cast<std::size_t>(g.n[0]) * g.n[1] * g.n[2];
  return (g.n[0] & 1) == 0 ? cells / 2 : cells;
}

template <typename T, int DIM>
void launch_upsample(int order_fft, const Cplx<T>* in, const IntType* in_strides,
                     const IntType* modes, T* const* fs, Cplx<T>* grid, const IntType* n,
                     hipStream_t stream) {
  const SampleGeom g = sample_geom<T, DIM>(n, modes, in_strides);
  hipLaunchKernelGGL((upsample_kernel<T, DIM>), dim3(blocks_for(up_work(g))), dim3(kBlock), 0,
                     stream, order_fft, in, g, fs_ptrs<T, DIM>(fs), grid);
  check_launch("upsample_kernel");
}

template <typename T, int DIM>
void launch_downsample(int order_fft, const Cplx<T>* grid, const IntType* n, T* const* fs,
                       Cplx<T>* out, const IntType* out_strides, const IntType* modes,
                       hipStream_t stream) {
  const SampleGeom g = sample_geom<T, DIM>(n, modes, out_strides);
  const std::size_t work = static_cast<std::size_t>(g.N[0]) * g.N[1] * g.N[2];
  hipLaunchKernelGGL((downsample_kernel<T, DIM>), dim3(blocks_for(work)), dim3(kBlock), 0,
                     stream, order_fft, grid, g, fs_ptrs<T, DIM>(fs), out);
  check_launch("downsample_kernel");
}

template <typename T, int DIM>
void launch_upsample_batch(int order_fft, const Cplx<T>* const* ins, const IntType* in_strides,
                           const IntType* modes, T* const* fs, Cplx<T>* grids,
                           std::size_t gstride, const IntType* n, int count, hipStream_t stream) {
  if (count <= 0) return;
  if (count > 65535) throw neonufft::InputError("Neonufft:: too many transforms of one shape");
  const SampleGeom g = sample_geom<T, DIM>(n, modes, in_strides);
  hipLaunchKernelGGL((upsample_batch_kernel<T, DIM>),
                     dim3(blocks_for(up_work(g), 2048u), static_cast<unsigned>(count)),
                     dim3(kBlock), 0, stream, order_fft, ins, g, fs_ptrs<T, DIM>(fs), grids,
                     gstride);
  check_launch("upsample_batch_kernel");
}

template <typename T, int DIM>
void launch_downsample_batch(int order_fft, const Cplx<T>* grids, std::size_t gstride,
                             const IntType* n, T* const* fs, Cplx<T>* const* outs,
                             const IntType* out_strides, const IntType* modes, int count,
                             hipStream_t stream) {
  if (count <= 0) return;
  if (count > 65535) throw neonufft::InputError("Neonufft:: too many transforms of one shape");
  const SampleGeom g = sample_geom<T, DIM>(n, modes, out_strides);
  const std::size_t work = static_cast<std::size_t>(g.N[0]) * g.N[1] * g.N[2];
  hipLaunchKernelGGL((downsample_batch_kernel<T, DIM>),
                     dim3(blocks_for(work, 2048u), static_cast<unsigned>(count)), dim3(kBlock),
                     0, stream, order_fft, grids, gstride, g, fs_ptrs<T, DIM>(fs), outs);
  check_launch("downsample_batch_kernel");
}

template <typename TC, typename T, int DIM>
void launch_prephase(int sign, Loc<TC> loc, const uint32_t* perm, std::size_t m,
                     const double* sbar, Cplx<T>* pre_sorted, hipStream_t stream) {
  if (m == 0) return;
  Vec3d sb{{0, 0, 0}};
  for (int d = 0; d < DIM; ++d) sb.v[d] = sbar[d];
  hipLaunchKernelGGL((prephase_kernel<TC, T, DIM>), dim3(blocks_for(m)), dim3(kBlock), 0, stream, sign,
                     loc, perm, m, sb, pre_sorted);
  check_launch("prephase_kernel");
}

template <typename TC, typename T, int DIM>
void launch_postphase(int sign, Loc<TC> loc, const uint32_t* perm, std::size_t m,
                      const double* xbar, const double* sbar, const double* scale,
                      const QuadRule& rule, Cplx<T>* post_sorted, hipStream_t stream) {
  if (m == 0) return;
  Vec3d xb{{0, 0, 0}}, sb{{0, 0, 0}}, sc{{0, 0, 0}};
  for (int d = 0; d < DIM; ++d) {
    xb.v[d] = xbar[d];
    sb.v[d] = sbar[d];
    sc.v[d] = scale[d];
  }
  hipLaunchKernelGGL((postphase_kernel<TC, T, DIM>), dim3(blocks_for(m)), dim3(kBlock), 0, stream,
                     sign, loc, perm, m, xb, sb, sc, rule, post_sorted);
  check_launch("postphase_kernel");
}

template <typename TI, typename TO>
void launch_convert(const Cplx<TI>* in, Cplx<TO>* out, std::size_t m, hipStream_t stream) {
  if (m == 0) return;
  hipLaunchKernelGGL((convert_kernel<TI, TO>), dim3(blocks_for(m)), dim3(kBlock), 0, stream, in,
                     out, m);
  check_launch("convert_kernel");
}

template void launch_convert<double, float>(const Cplx<double>*, Cplx<float>*, std::size_t,
                                            hipStream_t);
template void launch_convert<float, double>(const Cplx<float>*, Cplx<double>*, std::size_t,
                                            hipStream_t);

#define NNF_INST_GRID_T(T)                                                                   \
  template void launch_fill_zero<T>(Cplx<T>*, std::size_t, hipStream_t);                    \
  template void launch_fseries<T>(T*, IntType, const QuadRule&, hipStream_t);

#define NNF_INST_GRID(T, DIM)                                                                     \
  template void launch_upsample<T, DIM>(int, const Cplx<T>*, const IntType*, const IntType*,      \
                                        T* const*, Cplx<T>*, const IntType*, hipStream_t);        \
  template void launch_downsample<T, DIM>(int, const Cplx<T>*, const IntType*, T* const*,         \
                                          Cplx<T>*, const IntType*, const IntType*, hipStream_t); \
  template void launch_prephase<T, T, DIM>(int, Loc<T>, const uint32_t*, std::size_t,             \
                                           const double*, Cplx<T>*, hipStream_t);                 \
  template void launch_postphase<T, T, DIM>(int, Loc<T>, const uint32_t*, std::size_t,            \
                                            const double*, const double*, const double*,          \
                                            const QuadRule&, Cplx<T>*, hipStream_t);              \
  template void launch_upsample_batch<T, DIM>(int, const Cplx<T>* const*, const IntType*,         \
                                              const IntType*, T* const*, Cplx<T>*, std::size_t,   \
                                              const IntType*, int, hipStream_t);                  \
  template void launch_downsample_batch<T, DIM>(int, const Cplx<T>*, std::size_t, const IntType*, \
                                                T* const*, Cplx<T>* const*, const IntType*,       \
                                                const IntType*, int, hipStream_t);

NNF_INST_GRID_T(float)
NNF_INST_GRID_T(double)
NNF_INST_GRID(float, 1)
NNF_INST_GRID(float, 2)
NNF_INST_GRID(float, 3)
NNF_INST_GRID(double, 1)
NNF_INST_GRID(double, 2)
NNF_INST_GRID(double, 3)

// double coordinates, fp32 phases (mixed-precision type 3)
#define NNF_INST_PHASE_MIXED(DIM)                                                            \
  template void launch_prephase<double, float, DIM>(int, Loc<double>, const uint32_t*,       \
                                                    std::size_t, const double*, Cplx<float>*, \
                                                    hipStream_t);                            \
  template void launch_postphase<double, float, DIM>(                                        \
      int, Loc<double>, const uint32_t*, std::size_t, const double*, const double*,          \
      const double*, const QuadRule&, Cplx<float>*, hipStream_t);
NNF_INST_PHASE_MIXED(1)
NNF_INST_PHASE_MIXED(2)
NNF_INST_PHASE_MIXED(3)

}  // namespace nnf
