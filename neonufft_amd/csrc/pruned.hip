// Kernels of the pruned multi-dimensional FFT used by type 1 / type 2 plans
// in 2D and 3D.
//
// Type 1 keeps only N_d of the n_d frequencies per axis after the FFT, type 2
// starts from N_d nonzero entries per axis. Transforming one axis at a time
// with contiguous 1D rocFFT passes, and moving the next axis to the front with
// an LDS-tiled transpose that keeps only the retained entries (or expands
// them back with zeros), does a fraction of the full transform's work:
// 3D: x on n^3, y on n^2 N, z on n N^2 values; the reference instead runs the
// full in-place transform (src/neonufft/gpu/util/fft_grid.cpp:59-70) and then
// downsamples.
//
// Retained index j in [0, N) of an axis with n fine cells: fine index
// j < P ? j : j + n - N, mode k = j < P ? j : j - N, with P = N - N/2
// (mode_index.hpp).
#include "kernels.hpp"
#include "mode_index.hpp"

namespace nnf {

namespace {

constexpr int kTile = 32;   // transpose tile edge
constexpr int kRowsPer = 8; // tile rows handled per pass by a 32 x 8 block

inline unsigned cap_slabs(std::size_t s) {
  return static_cast<unsigned>(s < 65535 ? (s ? s : 1) : 65535);
}

}  // namespace

// ---------------------------------------------------------------------------
// dst[s][c][r] = src[s][r][ret_fine(c)], r < R, c < Nc (retained of nc).
// Slab s = (hi, lo) with lo < lo_count; offsets hi * *_hi + lo * *_lo.
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(kTile* kRowsPer)
    compact_transpose_kernel(const Cplx<T>* __restrict__ src, Cplx<T>* __restrict__ dst, int R,
                             int nc, int Nc, long long src_rs, long long dst_rs, int slabs,
                             int lo_count, long long src_hi, long long src_lo, long long dst_hi,
                             long long dst_lo) {
  __shared__ Cplx<T> tile[kTile][kTile + 1];
  const int c0 = blockIdx.x * kTile, r0 = blockIdx.y * kTile;
  for (int s = blockIdx.z; s < slabs; s += gridDim.z) {
    const long long sb = (s / lo_count) * src_hi + (s % lo_count) * src_lo;
    const long long db = (s / lo_count) * dst_hi + (s % lo_count) * dst_lo;
    const int c = c0 + threadIdx.x;
#pragma unroll
    for (int k = 0; k < kTile; k += kRowsPer) {
      const int r = r0 + threadIdx.y + k;
      if (r < R && c < Nc) tile[threadIdx.y + k][threadIdx.x] = src[sb + r * src_rs + ret_fine(c, nc, Nc)];
    }
    __syncthreads();
    const int r = r0 + threadIdx.x;
#pragma unroll
    for (int k = 0; k < kTile; k += kRowsPer) {
      const int cc = c0 + threadIdx.y + k;
      if (r < R && cc < Nc) dst[db + cc * dst_rs + r] = tile[threadIdx.x][threadIdx.y + k];
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// dst[s][r][f] = src[s][fine_ret(f)][r] for retained fine cells f < nf, else 0.
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(kTile* kRowsPer)
    expand_transpose_kernel(const Cplx<T>* __restrict__ src, Cplx<T>* __restrict__ dst, int R,
                            int nf, int Nf, long long src_rs, long long dst_rs, int slabs,
                            int lo_count, long long src_hi, long long src_lo, long long dst_hi,
                            long long dst_lo) {
  __shared__ Cplx<T> tile[kTile][kTile + 1];
  const int f0 = blockIdx.x * kTile, r0 = blockIdx.y * kTile;
  for (int s = blockIdx.z; s < slabs; s += gridDim.z) {
    const long long sb = (s / lo_count) * src_hi + (s % lo_count) * src_lo;
    const long long db = (s / lo_count) * dst_hi + (s % lo_count) * dst_lo;
    // load src rows (retained index of fine cell f0 + ...) along r
    const int r = r0 + threadIdx.x;
#pragma unroll
    for (int k = 0; k < kTile; k += kRowsPer) {
      const int f = f0 + threadIdx.y + k;
      Cplx<T> v{T(0), T(0)};
      if (f < nf && r < R) {
        const int j = fine_ret(f, nf, Nf);
        if (j >= 0) v = src[sb + j * src_rs + r];
      }
      tile[threadIdx.y + k][threadIdx.x] = v;
    }
    __syncthreads();
    const int f = f0 + threadIdx.x;
#pragma unroll
    for (int k = 0; k < kTile; k += kRowsPer) {
      const int rr = r0 + threadIdx.y + k;
      if (rr < R && f < nf) dst[db + rr * dst_rs + f] = tile[threadIdx.x][threadIdx.y + k];
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Type-1 output: the last FFT axis L is fastest in src, the x-retained axis
// is the row index. Slab = (member b, middle index m) for 3D (m = retained y),
// (b) for 2D. src[b][m][x'][l]: out_b[i0(x') st0 + i1(m) st1 + i2(l') st2] or,
// in 2D, out_b[i0(x') st0 + i1(l') st1], times the deconvolution factors.
// ---------------------------------------------------------------------------
template <typename T, int DIM>
__global__ void __launch_bounds__(kTile* kRowsPer)
    modes_out_kernel(const Cplx<T>* __restrict__ src, long long member_stride, Cplx<T>* single,
                     Cplx<T>* const* __restrict__ outs, int N0, int Nm, int nl, int Nl,
                     long long st0, long long stm, long long stl, const T* __restrict__ fs0,
                     const T* __restrict__ fsm, const T* __restrict__ fsl, int order_fft,
                     int slabs) {
  __shared__ Cplx<T> tile[kTile][kTile + 1];
  const int l0 = blockIdx.x * kTile, x0 = blockIdx.y * kTile;
  for (int s = blockIdx.z; s < slabs; s += gridDim.z) {
    const int b = DIM == 3 ? s / Nm : s;
    const int m = DIM == 3 ? s % Nm : 0;
    const Cplx<T>* sp = src + b * member_stride + (DIM == 3 ? static_cast<long long>(m) * N0 * nl : 0);
    Cplx<T>* out = outs ? outs[b] : single;
    // load [x'][l'] with l' along threadIdx.x (contiguous in src)
    const int l = l0 + threadIdx.x;
#pragma unroll
    for (int k = 0; k < kTile; k += kRowsPer) {
      const int x = x0 + threadIdx.y + k;
      if (x < N0 && l < Nl) tile[threadIdx.y + k][threadIdx.x] = sp[static_cast<long long>(x) * nl + ret_fine(l, nl, Nl)];
    }
    __syncthreads();
    const int x = x0 + threadIdx.x;
    T fm = T(1);
    long long om = 0;
    if (DIM == 3) {
      const int km = ret_mode(m, Nm);
      fm = fsm[km < 0 ? -km : km];
      om = static_cast<long long>(order_index(km, Nm, order_fft)) * stm;
    }
    if (x < N0) {
      const int kx = ret_mode(x, N0);
      const T fx = fs0[kx < 0 ? -kx : kx] * fm;
      const long long ox = static_cast<long long>(order_index(kx, N0, order_fft)) * st0 + om;
#pragma unroll
      for (int k = 0; k < kTile; k += kRowsPer) {
        const int ll = l0 + threadIdx.y + k;
        if (ll < Nl) {
          const int kl = ret_mode(ll, Nl);
          const T f = fx * fsl[kl < 0 ? -kl : kl];
          const Cplx<T> v = tile[threadIdx.x][threadIdx.y + k];
          out[ox + static_cast<long long>(order_index(kl, Nl, order_fft)) * stl] =
              Cplx<T>{v.x * f, v.y * f};
        }
      }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Type-2 input: the reverse of modes_out_kernel, with the last axis expanded
// to all nl fine cells (zeros outside the retained ones):
// dst[b][m][x'][f] = fs * in_b[...] for retained f, else 0.
// ---------------------------------------------------------------------------
template <typename T, int DIM>
__global__ void __launch_bounds__(kTile* kRowsPer)
    modes_in_kernel(const Cplx<T>* single, const Cplx<T>* const* __restrict__ ins,
                    Cplx<T>* __restrict__ dst, long long member_stride, int N0, int Nm, int nl,
                    int Nl, long long st0, long long stm, long long stl,
                    const T* __restrict__ fs0, const T* __restrict__ fsm,
                    const T* __restrict__ fsl, int order_fft, int slabs) {
  __shared__ Cplx<T> tile[kTile][kTile + 1];
  const int f0 = blockIdx.x * kTile, x0 = blockIdx.y * kTile;
  for (int s = blockIdx.z; s < slabs; s += gridDim.z) {
    const int b = DIM == 3 ? s / Nm : s;
    const int m = DIM == 3 ? s % Nm : 0;
    Cplx<T>* dp = dst + b * member_stride + (DIM == 3 ? static_cast<long long>(m) * N0 * nl : 0);
    const Cplx<T>* in = ins ? ins[b] : single;
    T fm = T(1);
    long long om = 0;
    if (DIM == 3) {
      const int km = ret_mode(m, Nm);
      fm = fsm[km < 0 ? -km : km];
      om = static_cast<long long>(order_index(km, Nm, order_fft)) * stm;
    }
    // load along x' (contiguous in the caller's array when st0 = 1)
    const int x = x0 + threadIdx.x;
#pragma unroll
    for (int k = 0; k < kTile; k += kRowsPer) {
      const int f = f0 + threadIdx.y + k;
      Cplx<T> v{T(0), T(0)};
      if (x < N0 && f < nl) {
        const int j = fine_ret(f, nl, Nl);
        if (j >= 0) {
          const int kx = ret_mode(x, N0), kl = ret_mode(j, Nl);
          const T fac = fm * fs0[kx < 0 ? -kx : kx] * fsl[kl < 0 ? -kl : kl];
          const Cplx<T> a = in[static_cast<long long>(order_index(kx, N0, order_fft)) * st0 + om +
                               static_cast<long long>(order_index(kl, Nl, order_fft)) * stl];
          v = Cplx<T>{a.x * fac, a.y * fac};
        }
      }
      tile[threadIdx.y + k][threadIdx.x] = v;
    }
    __syncthreads();
    const int f = f0 + threadIdx.x;
#pragma unroll
    for (int k = 0; k < kTile; k += kRowsPer) {
      const int xx = x0 + threadIdx.y + k;
      if (xx < N0 && f < nl) dp[static_cast<long long>(xx) * nl + f] = tile[threadIdx.x][threadIdx.y + k];
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
template <typename T>
void launch_compact_transpose(const Cplx<T>* src, Cplx<T>* dst, int R, int nc, int Nc,
                              long long src_rs, long long dst_rs, std::size_t slabs, int lo_count,
                              long long src_hi, long long src_lo, long long dst_hi,
                              long long dst_lo, hipStream_t stream) {
  if (!slabs || !R || !Nc) return;
  const dim3 grid((Nc + kTile - 1) / kTile, (R + kTile - 1) / kTile, cap_slabs(slabs));
  hipLaunchKernelGGL((compact_transpose_kernel<T>), grid, dim3(kTile, kRowsPer), 0, stream, src,
                     dst, R, nc, Nc, src_rs, dst_rs, static_cast<int>(slabs), lo_count, src_hi,
                     src_lo, dst_hi, dst_lo);
  check_launch("compact_transpose_kernel");
}

template <typename T>
void launch_expand_transpose(const Cplx<T>* src, Cplx<T>* dst, int R, int nf, int Nf,
                             long long src_rs, long long dst_rs, std::size_t slabs, int lo_count,
                             long long src_hi, long long src_lo, long long dst_hi,
                             long long dst_lo, hipStream_t stream) {
  if (!slabs || !R || !nf) return;
  const dim3 grid((nf + kTile - 1) / kTile, (R + kTile - 1) / kTile, cap_slabs(slabs));
  hipLaunchKernelGGL((expand_transpose_kernel<T>), grid, dim3(kTile, kRowsPer), 0, stream, src,
                     dst, R, nf, Nf, src_rs, dst_rs, static_cast<int>(slabs), lo_count, src_hi,
                     src_lo, dst_hi, dst_lo);
  check_launch("expand_transpose_kernel");
}

template <typename T, int DIM>
void launch_modes_out(const Cplx<T>* src, long long member_stride, Cplx<T>* single,
                      Cplx<T>* const* outs, int count, int N0, int Nm, int nl, int Nl,
                      long long st0, long long stm, long long stl, const T* fs0, const T* fsm,
                      const T* fsl, int order_fft, hipStream_t stream) {
  const std::size_t slabs = static_cast<std::size_t>(count) * (DIM == 3 ? Nm : 1);
  const dim3 grid((Nl + kTile - 1) / kTile, (N0 + kTile - 1) / kTile, cap_slabs(slabs));
  hipLaunchKernelGGL((modes_out_kernel<T, DIM>), grid, dim3(kTile, kRowsPer), 0, stream, src,
                     member_stride, single, outs, N0, Nm, nl, Nl, st0, stm, stl, fs0, fsm, fsl,
                     order_fft, static_cast<int>(slabs));
  check_launch("modes_out_kernel");
}

template <typename T, int DIM>
void launch_modes_in(const Cplx<T>* single, const Cplx<T>* const* ins, int count, Cplx<T>* dst,
                     long long member_stride, int N0, int Nm, int nl, int Nl, long long st0,
                     long long stm, long long stl, const T* fs0, const T* fsm, const T* fsl,
                     int order_fft, hipStream_t stream) {
  const std::size_t slabs = static_cast<std::size_t>(count) * (DIM == 3 ? Nm : 1);
  const dim3 grid((nl + kTile - 1) / kTile, (N0 + kTile - 1) / kTile, cap_slabs(slabs));
  hipLaunchKernelGGL((modes_in_kernel<T, DIM>), grid, dim3(kTile, kRowsPer), 0, stream, single,
                     ins, dst, member_stride, N0, Nm, nl, Nl, st0, stm, stl, fs0, fsm, fsl,
                     order_fft, static_cast<int>(slabs));
  check_launch("modes_in_kernel");
}

#define NNF_INST_PRUNED_T(T)                                                                     \
  template void launch_compact_transpose<T>(const Cplx<T>*, Cplx<T>*, int, int, int, long long,   \
                                            long long, std::size_t, int, long long, long long,    \
                                            long long, long long, hipStream_t);                   \
  template void launch_expand_transpose<T>(const Cplx<T>*, Cplx<T>*, int, int, int, long long,    \
                                           long long, std::size_t, int, long long, long long,     \
                                           long long, long long, hipStream_t);
#define NNF_INST_PRUNED(T, DIM)                                                                  \
  template void launch_modes_out<T, DIM>(const Cplx<T>*, long long, Cplx<T>*, Cplx<T>* const*,    \
                                         int, int, int, int, int, long long, long long,           \
                                         long long, const T*, const T*, const T*, int,            \
                                         hipStream_t);                                            \
  template void launch_modes_in<T, DIM>(const Cplx<T>*, const Cplx<T>* const*, int, Cplx<T>*,     \
                                        long long, int, int, int, int, long long, long long,      \
                                        long long, const T*, const T*, const T*, int,             \
                                        hipStream_t);

NNF_INST_PRUNED_T(float)
NNF_INST_PRUNED_T(double)
NNF_INST_PRUNED(float, 2)
NNF_INST_PRUNED(float, 3)
NNF_INST_PRUNED(double, 2)
NNF_INST_PRUNED(double, 3)

}  // namespace nnf
