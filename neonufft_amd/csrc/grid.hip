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

struct Dims3 {
  IntType v[3];
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

// Mode k of an N-mode axis is in [-N/2, N-1-N/2]. Fine-grid cell g holds mode
// k = g (g < N - N/2) or k = g - n (g >= n - N/2); other cells are padding.
__device__ inline bool cell_mode(IntType g, IntType n, IntType N, IntType& k) {
  const IntType nneg = N / 2;
  if (g < N - nneg) {
    k = g;
    return true;
  }
  if (g >= n - nneg) {
    k = g - n;
    return true;
  }
  return false;
}

// Array index of mode k (CMCL: k + N/2; FFT order: k, negatives after).
__device__ inline IntType mode_index(IntType k, IntType N, int order_fft) {
  if (order_fft) return k >= 0 ? k : k + N;
  return k + N / 2;
}

template <typename T, int DIM>
__global__ void __launch_bounds__(kBlock)
    upsample_kernel(int order_fft, const Cplx<T>* __restrict__ in, Dims3 st, Dims3 N,
                    FsPtrs<T> fs, Cplx<T>* __restrict__ grid, Dims3 n) {
  const IntType rows = (DIM >= 2 ? n.v[1] : 1) * (DIM >= 3 ? n.v[2] : 1);
  for (IntType row = blockIdx.y; row < rows; row += gridDim.y) {
    T frow = T(1);
    bool live = true;
    IntType off = 0;
    if constexpr (DIM >= 2) {
      const IntType g1 = row % n.v[1];
      IntType k1;
      live = cell_mode(g1, n.v[1], N.v[1], k1);
      if (live) {
        frow *= fs.p[1][k1 < 0 ? -k1 : k1];
        off += mode_index(k1, N.v[1], order_fft) * st.v[1];
      }
      if constexpr (DIM >= 3) {
        const IntType g2 = row / n.v[1];
        IntType k2;
        const bool l2 = cell_mode(g2, n.v[2], N.v[2], k2);
        live = live && l2;
        if (live) {
          frow *= fs.p[2][k2 < 0 ? -k2 : k2];
          off += mode_index(k2, N.v[2], order_fft) * st.v[2];
        }
      }
    }
    Cplx<T>* grow = grid + row * n.v[0];
    for (IntType g0 = threadIdx.x + static_cast<IntType>(blockIdx.x) * kBlock; g0 < n.v[0];
         g0 += static_cast<IntType>(gridDim.x) * kBlock) {
      Cplx<T> v{T(0), T(0)};
      IntType k0;
      if (live && cell_mode(g0, n.v[0], N.v[0], k0)) {
        const Cplx<T> a = in[off + mode_index(k0, N.v[0], order_fft) * st.v[0]];
        const T f = frow * fs.p[0][k0 < 0 ? -k0 : k0];
        v = Cplx<T>{a.x * f, a.y * f};
      }
      grow[g0] = v;
    }
  }
}

template <typename T, int DIM>
__global__ void __launch_bounds__(kBlock)
    downsample_kernel(int order_fft, const Cplx<T>* __restrict__ grid, Dims3 n, FsPtrs<T> fs,
                      Cplx<T>* __restrict__ out, Dims3 st, Dims3 N) {
  const IntType rows = (DIM >= 2 ? N.v[1] : 1) * (DIM >= 3 ? N.v[2] : 1);
  for (IntType row = blockIdx.y; row < rows; row += gridDim.y) {
    T frow = T(1);
    IntType goff = 0, ooff = 0;
    if constexpr (DIM >= 2) {
      const IntType i1 = row % N.v[1];
      const IntType k1 = order_fft ? (i1 < N.v[1] - N.v[1] / 2 ? i1 : i1 - N.v[1]) : i1 - N.v[1] / 2;
      frow *= fs.p[1][k1 < 0 ? -k1 : k1];
      goff += (k1 < 0 ? k1 + n.v[1] : k1);
      ooff += i1 * st.v[1];
      if constexpr (DIM >= 3) {
        const IntType i2 = row / N.v[1];
        const IntType k2 =
            order_fft ? (i2 < N.v[2] - N.v[2] / 2 ? i2 : i2 - N.v[2]) : i2 - N.v[2] / 2;
        frow *= fs.p[2][k2 < 0 ? -k2 : k2];
        goff += (k2 < 0 ? k2 + n.v[2] : k2) * n.v[1];
        ooff += i2 * st.v[2];
      }
    }
    const Cplx<T>* grow = grid + goff * n.v[0];
    for (IntType i0 = threadIdx.x + static_cast<IntType>(blockIdx.x) * kBlock; i0 < N.v[0];
         i0 += static_cast<IntType>(gridDim.x) * kBlock) {
      const IntType k0 =
          order_fft ? (i0 < N.v[0] - N.v[0] / 2 ? i0 : i0 - N.v[0]) : i0 - N.v[0] / 2;
      const Cplx<T> a = grow[k0 < 0 ? k0 + n.v[0] : k0];
      const T f = frow * fs.p[0][k0 < 0 ? -k0 : k0];
      out[ooff + i0 * st.v[0]] = Cplx<T>{a.x * f, a.y * f};
    }
  }
}

// Batched variants: blockIdx.z selects one of `count` transforms of equal
// mode and fine-grid shape whose grids are contiguous (stride `gstride`).
template <typename T, int DIM>
__global__ void __launch_bounds__(kBlock)
    upsample_batch_kernel(int order_fft, const Cplx<T>* const* __restrict__ ins, Dims3 st,
                          Dims3 N, FsPtrs<T> fs, Cplx<T>* __restrict__ grids, std::size_t gstride,
                          Dims3 n) {
  const Cplx<T>* __restrict__ in = ins[blockIdx.z];
  Cplx<T>* __restrict__ grid = grids + gstride * blockIdx.z;
  const IntType rows = (DIM >= 2 ? n.v[1] : 1) * (DIM >= 3 ? n.v[2] : 1);
  for (IntType row = blockIdx.y; row < rows; row += gridDim.y) {
    T frow = T(1);
    bool live = true;
    IntType off = 0;
    if constexpr (DIM >= 2) {
      const IntType g1 = row % n.v[1];
      IntType k1;
      live = cell_mode(g1, n.v[1], N.v[1], k1);
      if (live) {
        frow *= fs.p[1][k1 < 0 ? -k1 : k1];
        off += mode_index(k1, N.v[1], order_fft) * st.v[1];
      }
      if constexpr (DIM >= 3) {
        const IntType g2 = row / n.v[1];
        IntType k2;
        const bool l2 = cell_mode(g2, n.v[2], N.v[2], k2);
        live = live && l2;
        if (live) {
          frow *= fs.p[2][k2 < 0 ? -k2 : k2];
          off += mode_index(k2, N.v[2], order_fft) * st.v[2];
        }
      }
    }
    Cplx<T>* grow = grid + row * n.v[0];
    for (IntType g0 = threadIdx.x + static_cast<IntType>(blockIdx.x) * kBlock; g0 < n.v[0];
         g0 += static_cast<IntType>(gridDim.x) * kBlock) {
      Cplx<T> v{T(0), T(0)};
      IntType k0;
      if (live && cell_mode(g0, n.v[0], N.v[0], k0)) {
        const Cplx<T> a = in[off + mode_index(k0, N.v[0], order_fft) * st.v[0]];
        const T f = frow * fs.p[0][k0 < 0 ? -k0 : k0];
        v = Cplx<T>{a.x * f, a.y * f};
      }
      grow[g0] = v;
    }
  }
}

template <typename T, int DIM>
__global__ void __launch_bounds__(kBlock)
    downsample_batch_kernel(int order_fft, const Cplx<T>* __restrict__ grids, std::size_t gstride,
                            Dims3 n, FsPtrs<T> fs, Cplx<T>* const* __restrict__ outs, Dims3 st,
                            Dims3 N) {
  const Cplx<T>* __restrict__ grid = grids + gstride * blockIdx.z;
  Cplx<T>* __restrict__ out = outs[blockIdx.z];
  const IntType rows = (DIM >= 2 ? N.v[1] : 1) * (DIM >= 3 ? N.v[2] : 1);
  for (IntType row = blockIdx.y; row < rows; row += gridDim.y) {
    T frow = T(1);
    IntType goff = 0, ooff = 0;
    if constexpr (DIM >= 2) {
      const IntType i1 = row % N.v[1];
      const IntType k1 = order_fft ? (i1 < N.v[1] - N.v[1] / 2 ? i1 : i1 - N.v[1]) : i1 - N.v[1] / 2;
      frow *= fs.p[1][k1 < 0 ? -k1 : k1];
      goff += (k1 < 0 ? k1 + n.v[1] : k1);
      ooff += i1 * st.v[1];
      if constexpr (DIM >= 3) {
        const IntType i2 = row / N.v[1];
        const IntType k2 =
            order_fft ? (i2 < N.v[2] - N.v[2] / 2 ? i2 : i2 - N.v[2]) : i2 - N.v[2] / 2;
        frow *= fs.p[2][k2 < 0 ? -k2 : k2];
        goff += (k2 < 0 ? k2 + n.v[2] : k2) * n.v[1];
        ooff += i2 * st.v[2];
      }
    }
    const Cplx<T>* grow = grid + goff * n.v[0];
    for (IntType i0 = threadIdx.x + static_cast<IntType>(blockIdx.x) * kBlock; i0 < N.v[0];
         i0 += static_cast<IntType>(gridDim.x) * kBlock) {
      const IntType k0 =
          order_fft ? (i0 < N.v[0] - N.v[0] / 2 ? i0 : i0 - N.v[0]) : i0 - N.v[0] / 2;
      const Cplx<T> a = grow[k0 < 0 ? k0 + n.v[0] : k0];
      const T f = frow * fs.p[0][k0 < 0 ? -k0 : k0];
      out[ooff + i0 * st.v[0]] = Cplx<T>{a.x * f, a.y * f};
    }
  }
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
void launch_upsample(int order_fft, const Cplx<T>* in, const IntType* in_strides,
                     const IntType* modes, T* const* fs, Cplx<T>* grid, const IntType* n,
                     hipStream_t stream) {
  Dims3 st{{0, 0, 0}}, NN{{1, 1, 1}}, nn{{1, 1, 1}};
  FsPtrs<T> f{{nullptr, nullptr, nullptr}};
  for (int d = 0; d < DIM; ++d) {
    st.v[d] = in_strides[d];
    NN.v[d] = modes[d];
    nn.v[d] = n[d];
    f.p[d] = fs[d];
  }
  const IntType rows = nn.v[1] * nn.v[2];
  const unsigned gx = static_cast<unsigned>((nn.v[0] + kBlock - 1) / kBlock);
  const unsigned gy = static_cast<unsigned>(rows < 65535 ? rows : 65535);
  hipLaunchKernelGGL((upsample_kernel<T, DIM>), dim3(gx, gy), dim3(kBlock), 0, stream, order_fft,
                     in, st, NN, f, grid, nn);
  check_launch("upsample_kernel");
}

template <typename T, int DIM>
void launch_downsample(int order_fft, const Cplx<T>* grid, const IntType* n, T* const* fs,
                       Cplx<T>* out, const IntType* out_strides, const IntType* modes,
                       hipStream_t stream) {
  Dims3 st{{0, 0, 0}}, NN{{1, 1, 1}}, nn{{1, 1, 1}};
  FsPtrs<T> f{{nullptr, nullptr, nullptr}};
  for (int d = 0; d < DIM; ++d) {
    st.v[d] = out_strides[d];
    NN.v[d] = modes[d];
    nn.v[d] = n[d];
    f.p[d] = fs[d];
  }
  const IntType rows = NN.v[1] * NN.v[2];
  const unsigned gx = static_cast<unsigned>((NN.v[0] + kBlock - 1) / kBlock);
  const unsigned gy = static_cast<unsigned>(rows < 65535 ? rows : 65535);
  hipLaunchKernelGGL((downsample_kernel<T, DIM>), dim3(gx, gy), dim3(kBlock), 0, stream,
                     order_fft, grid, nn, f, out, st, NN);
  check_launch("downsample_kernel");
}

template <typename T, int DIM>
void launch_upsample_batch(int order_fft, const Cplx<T>* const* ins, const IntType* in_strides,
                           const IntType* modes, T* const* fs, Cplx<T>* grids,
                           std::size_t gstride, const IntType* n, int count, hipStream_t stream) {
  if (count <= 0) return;
  Dims3 st{{0, 0, 0}}, NN{{1, 1, 1}}, nn{{1, 1, 1}};
  FsPtrs<T> f{{nullptr, nullptr, nullptr}};
  for (int d = 0; d < DIM; ++d) {
    st.v[d] = in_strides[d];
    NN.v[d] = modes[d];
    nn.v[d] = n[d];
    f.p[d] = fs[d];
  }
  const IntType rows = nn.v[1] * nn.v[2];
  const unsigned gx = static_cast<unsigned>((nn.v[0] + kBlock - 1) / kBlock);
  const unsigned gy = static_cast<unsigned>(rows < 1024 ? rows : 1024);
  hipLaunchKernelGGL((upsample_batch_kernel<T, DIM>), dim3(gx, gy, static_cast<unsigned>(count)),
                     dim3(kBlock), 0, stream, order_fft, ins, st, NN, f, grids, gstride, nn);
  check_launch("upsample_batch_kernel");
}

template <typename T, int DIM>
void launch_downsample_batch(int order_fft, const Cplx<T>* grids, std::size_t gstride,
                             const IntType* n, T* const* fs, Cplx<T>* const* outs,
                             const IntType* out_strides, const IntType* modes, int count,
                             hipStream_t stream) {
  if (count <= 0) return;
  Dims3 st{{0, 0, 0}}, NN{{1, 1, 1}}, nn{{1, 1, 1}};
  FsPtrs<T> f{{nullptr, nullptr, nullptr}};
  for (int d = 0; d < DIM; ++d) {
    st.v[d] = out_strides[d];
    NN.v[d] = modes[d];
    nn.v[d] = n[d];
    f.p[d] = fs[d];
  }
  const IntType rows = NN.v[1] * NN.v[2];
  const unsigned gx = static_cast<unsigned>((NN.v[0] + kBlock - 1) / kBlock);
  const unsigned gy = static_cast<unsigned>(rows < 1024 ? rows : 1024);
  hipLaunchKernelGGL((downsample_batch_kernel<T, DIM>), dim3(gx, gy, static_cast<unsigned>(count)),
                     dim3(kBlock), 0, stream, order_fft, grids, gstride, nn, f, outs, st, NN);
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
