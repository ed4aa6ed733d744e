// In-LDS x-direction FFT of whole fine-grid rows (power-of-two length n0 >=
// 32), shared by the row-band spreading (band.hip) and the row kernels that
// join the tile reduction or the zero padding to the x-FFT (rows.hip).
#pragma once

#include "es_device.hpp"

namespace nnf {

namespace {

// ---------------------------------------------------------------------------
// x-direction FFT of rows held in LDS, decimation in frequency: the rows are
// in natural x order (row bands accumulate them there, so bank-ordered points
// stay conflict-free) and the transform leaves X[f] at position bitrev(f):
//   stage B: half-spans n0/2 .. 32 as radix-4 passes through LDS (consecutive
//            lanes take consecutive positions, conflict-free), plus one
//            radix-2 pass at half-span 32 when their count is odd; a thread
//            keeps one position across the rows of the band, so it loads its
//            twiddles once per pass;
//   stage A: half-spans 16 .. 1 inside aligned 32-cell blocks: a quad of
//            lanes owns a block, 8 consecutive positions per lane (16 and 8
//            by DPP quad permutations, 4, 2, 1 in registers).
// With kFixed, the rows hold fixed-point sums (Acc) and the first stage that
// runs converts them to T in place (times inv); else they already hold T.
// Twiddles are stored level by level: W_N^q = exp(sign 2 pi i q / N) at
// tw[N/2 + q] (q < N/2, N = 2, 4, ... up to the largest n0), so the twiddles
// of consecutive positions are consecutive in memory.
// ---------------------------------------------------------------------------
template <typename T>
__device__ inline Cplx<T> twid(const Cplx<T>* __restrict__ tw, int n, int q) {
  return tw[(n >> 1) + q];
}

// cos and sin of 2 pi j / 32, j < 16: the twiddles of the in-register
// levels (W_32^j, W_16^j = W_32^2j, W_8^j = W_32^4j, W_4^j = W_32^8j) as
// literals, so they occupy no registers.
constexpr double kCos32[16] = {
    1.0, 0.9807852804032304, 0.9238795325112867, 0.8314696123025452,
    0.7071067811865476, 0.5555702330196023, 0.38268343236508984, 0.19509032201612833,
    6.123233995736766e-17, -0.1950903220161282, -0.3826834323650897, -0.555570233019602,
    -0.7071067811865475, -0.8314696123025453, -0.9238795325112867, -0.9807852804032304};
constexpr double kSin32[16] = {
    0.0, 0.19509032201612825, 0.3826834323650898, 0.5555702330196022,
    0.7071067811865475, 0.8314696123025452, 0.9238795325112867, 0.9807852804032304,
    1.0, 0.9807852804032304, 0.9238795325112867, 0.8314696123025455,
    0.7071067811865476, 0.5555702330196022, 0.3826834323650899, 0.1950903220161286};

template <typename T>
__device__ inline Cplx<T> w32(int j, T sgn) {
  return Cplx<T>{static_cast<T>(kCos32[j]), sgn * static_cast<T>(kSin32[j])};
}

// (a, b) <- (a + b, (a - b) w)
template <typename T>
__device__ inline void dif(T& ar, T& ai, T& br, T& bi, const Cplx<T>& w) {
  const T dr = ar - br, di = ai - bi;
  ar += br;
  ai += bi;
  br = dr * w.x - di * w.y;
  bi = dr * w.y + di * w.x;
}

// The value of `v` in lane ^ MASK (MASK 1 or 2; quad permutation, no LDS).
template <int MASK>
__device__ inline int quad_xor_bits(int v) {
  constexpr int ctrl = MASK == 1 ? 0xB1 : 0x4E;  // quad_perm [1,0,3,2] / [2,3,0,1]
  return __builtin_amdgcn_update_dpp(0, v, ctrl, 0xF, 0xF, false);
}
template <int MASK>
__device__ inline float quad_xor(float v) {
  return __int_as_float(quad_xor_bits<MASK>(__float_as_int(v)));
}
template <int MASK>
__device__ inline double quad_xor(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = quad_xor_bits<MASK>(static_cast<int>(b));
  const int hi = quad_xor_bits<MASK>(static_cast<int>(b >> 32));
  return __longlong_as_double((static_cast<long long>(hi) << 32) |
                              static_cast<long long>(static_cast<unsigned int>(lo)));
}

// Level of half-span 16 (MASK 2) or 8 (MASK 1) across the lanes of a quad:
// the lane whose chunk bit is clear holds the top element a, its partner b;
// top <- a + b, bottom <- (a - b) w.
template <int MASK, typename T>
__device__ inline void dif_quad(T& xr, T& xi, bool top, const Cplx<T>& w) {
  const T pr = quad_xor<MASK>(xr), pi = quad_xor<MASK>(xi);
  const T dr = pr - xr, di = pi - xi;  // bottom: a - b with a = partner
  xr = top ? xr + pr : dr * w.x - di * w.y;
  xi = top ? xi + pi : dr * w.y + di * w.x;
}

// Element i of a row plane as T: converted from fixed point in the first
// stage, else read back as T.
template <bool CONV, typename T, typename Acc>
__device__ inline T ldv(const Acc* ip, const T* fp, int i, T inv) {
  if constexpr (CONV) {
    return static_cast<T>(ip[i]) * inv;
  } else {
    return fp[i];
  }
}

// Radix-4 DIF over the levels 2u then u at positions base + {0, u, 2u, 3u}.
template <bool CONV, typename T, typename Acc>
__device__ inline void dif4(const Acc* ire, const Acc* iim, T* fre, T* fim, int base, int u,
                            const Cplx<T>& w40, const Cplx<T>& w41, const Cplx<T>& w2, T inv) {
  T x0r = ldv<CONV>(ire, fre, base, inv), x0i = ldv<CONV>(iim, fim, base, inv);
  T x1r = ldv<CONV>(ire, fre, base + u, inv), x1i = ldv<CONV>(iim, fim, base + u, inv);
  T x2r = ldv<CONV>(ire, fre, base + 2 * u, inv), x2i = ldv<CONV>(iim, fim, base + 2 * u, inv);
  T x3r = ldv<CONV>(ire, fre, base + 3 * u, inv), x3i = ldv<CONV>(iim, fim, base + 3 * u, inv);
  dif(x0r, x0i, x2r, x2i, w40);
  dif(x1r, x1i, x3r, x3i, w41);
  dif(x0r, x0i, x1r, x1i, w2);
  dif(x2r, x2i, x3r, x3i, w2);
  fre[base] = x0r;
  fim[base] = x0i;
  fre[base + u] = x1r;
  fim[base + u] = x1i;
  fre[base + 2 * u] = x2r;
  fim[base + 2 * u] = x2i;
  fre[base + 3 * u] = x3r;
  fim[base + 3 * u] = x3i;
}

template <bool CONV, typename T, int NT, typename Acc>
__device__ inline void dif4_pass(const Acc* ire, const Acc* iim, T* fre, T* fim, int rk, int pitch,
                                 int log2n0, int u, T inv, const Cplx<T>* __restrict__ tw) {
  const int lq = log2n0 - 2;  // log2 items per row
  const int items = 1 << lq;
  const int rstep = items >= NT ? 1 : NT >> lq;
  for (int j0 = 0; j0 < items; j0 += NT) {
    const int j = j0 + (items >= NT ? static_cast<int>(threadIdx.x)
                                    : static_cast<int>(threadIdx.x) & (items - 1));
    const int r0 = items >= NT ? 0 : static_cast<int>(threadIdx.x) >> lq;
    const int q = j & (u - 1);
    const int p = (j - q) * 4 + q;
    const Cplx<T> w40 = twid(tw, 4 * u, q), w41 = twid(tw, 4 * u, q + u);
    const Cplx<T> w2 = twid(tw, 2 * u, q);
    // (unrolling two rows here spills registers at 1024 threads)
    for (int r = r0; r < rk; r += rstep) {
      dif4<CONV>(ire, iim, fre, fim, r * pitch + p, u, w40, w41, w2, inv);
    }
  }
}

template <bool CONV, typename T, int NT, typename Acc>
__device__ inline void dif2_pass32(const Acc* ire, const Acc* iim, T* fre, T* fim, int rk,
                                   int pitch, int log2n0, T inv, const Cplx<T>* __restrict__ tw) {
  const int lh = log2n0 - 1;
  const int items = 1 << lh;
  const int rstep = items >= NT ? 1 : NT >> lh;
  for (int j0 = 0; j0 < items; j0 += NT) {
    const int j = j0 + (items >= NT ? static_cast<int>(threadIdx.x)
                                    : static_cast<int>(threadIdx.x) & (items - 1));
    const int r0 = items >= NT ? 0 : static_cast<int>(threadIdx.x) >> lh;
    const int q = j & 31;
    const int p = (j - q) * 2 + q;
    const Cplx<T> w = twid(tw, 64, q);
    for (int r = r0; r < rk; r += rstep) {
      const int a = r * pitch + p;
      T ar = ldv<CONV>(ire, fre, a, inv), ai = ldv<CONV>(iim, fim, a, inv);
      T br = ldv<CONV>(ire, fre, a + 32, inv), bi = ldv<CONV>(iim, fim, a + 32, inv);
      dif(ar, ai, br, bi, w);
      fre[a] = ar;
      fim[a] = ai;
      fre[a + 32] = br;
      fim[a + 32] = bi;
    }
  }
}

// 8 consecutive LDS elements of a 16-byte aligned address as 16-byte accesses
// (lanes 32 bytes apart then meet each bank at most twice per wave; 4-byte
// accesses at that stride would queue 16 lanes on each of 4 banks).
template <typename E>
__device__ inline void lds_load8(const E* p, E (&v)[8]) {
  constexpr int kVec = 16 / sizeof(E);
#pragma unroll
  for (int i = 0; i < 8; i += kVec) {
    const uint4 u = *reinterpret_cast<const uint4*>(p + i);
    __builtin_memcpy(&v[i], &u, 16);
  }
}
template <typename E>
__device__ inline void lds_store8(E* p, const E (&v)[8]) {
  constexpr int kVec = 16 / sizeof(E);
#pragma unroll
  for (int i = 0; i < 8; i += kVec) {
    uint4 u;
    __builtin_memcpy(&u, &v[i], 16);
    *reinterpret_cast<uint4*>(p + i) = u;
  }
}

template <bool CONV, typename T, int NT, typename Acc>
__device__ inline void dif_stage_a(const Acc* ire, const Acc* iim, T* fre, T* fim, int rk,
                                   int pitch, int log2n0, T inv, T sgn) {
  const int lc = log2n0 - 3;  // log2 chunks of 8 per row
  const int chunks = 1 << lc;
  const int rstep = chunks >= NT ? 1 : NT >> lc;
  const int c = threadIdx.x & 3;
  for (int k0 = 0; k0 < chunks; k0 += NT) {
    const int k = k0 + (chunks >= NT ? static_cast<int>(threadIdx.x)
                                     : static_cast<int>(threadIdx.x) & (chunks - 1));
    const int r0 = chunks >= NT ? 0 : static_cast<int>(threadIdx.x) >> lc;
    for (int r = r0; r < rk; r += rstep) {
      const int base = r * pitch + 8 * k;  // pitch is a multiple of 8
      T xr[8], xi[8];
      if constexpr (CONV) {
        Acc ar[8], ai[8];
        lds_load8(ire + base, ar);
        lds_load8(iim + base, ai);
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          xr[e] = static_cast<T>(ar[e]) * inv;
          xi[e] = static_cast<T>(ai[e]) * inv;
        }
      } else {
        lds_load8(fre + base, xr);
        lds_load8(fim + base, xi);
      }
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const Cplx<T> lo = w32<T>(e, sgn), hi = w32<T>(8 + e, sgn);
        dif_quad<2>(xr[e], xi[e], (c & 2) == 0, (c & 1) ? hi : lo);
      }
#pragma unroll
      for (int e = 0; e < 8; ++e) dif_quad<1>(xr[e], xi[e], (c & 1) == 0, w32<T>(2 * e, sgn));
#pragma unroll
      for (int e = 0; e < 4; ++e) dif(xr[e], xi[e], xr[e + 4], xi[e + 4], w32<T>(4 * e, sgn));
#pragma unroll
      for (int e = 0; e < 8; e += 4) {
        dif(xr[e], xi[e], xr[e + 2], xi[e + 2], w32<T>(0, sgn));
        dif(xr[e + 1], xi[e + 1], xr[e + 3], xi[e + 3], w32<T>(8, sgn));
      }
#pragma unroll
      for (int e = 0; e < 8; e += 2) dif(xr[e], xi[e], xr[e + 1], xi[e + 1], w32<T>(0, sgn));
      lds_store8(fre + base, xr);
      lds_store8(fim + base, xi);
    }
  }
}

template <typename T, int NT, bool kFixed, typename Acc>
__device__ inline void lds_fft_x(Acc* ire, Acc* iim, int rk, int pitch, int log2n0, T inv,
                                 const Cplx<T>* __restrict__ tw) {
  // sign of the transform: imaginary part of W_4 = exp(sign i pi / 2)
  const T sgn = twid(tw, 4, 1).y > T(0) ? T(1) : T(-1);
  T* fre = reinterpret_cast<T*>(ire);
  T* fim = reinterpret_cast<T*>(iim);
  const int lb = log2n0 - 5;  // levels of half-span >= 32
  bool conv = kFixed;
  int u = 1 << (log2n0 - 2);
  for (int k = 0; k < lb / 2; ++k, u >>= 2) {
    if (conv) {
      dif4_pass<true, T, NT>(ire, iim, fre, fim, rk, pitch, log2n0, u, inv, tw);
    } else {
      dif4_pass<false, T, NT>(ire, iim, fre, fim, rk, pitch, log2n0, u, inv, tw);
    }
    conv = false;
    __syncthreads();
  }
  if (lb & 1) {
    if (conv) {
      dif2_pass32<true, T, NT>(ire, iim, fre, fim, rk, pitch, log2n0, inv, tw);
    } else {
      dif2_pass32<false, T, NT>(ire, iim, fre, fim, rk, pitch, log2n0, inv, tw);
    }
    conv = false;
    __syncthreads();
  }
  if (conv) {
    dif_stage_a<true, T, NT>(ire, iim, fre, fim, rk, pitch, log2n0, inv, sgn);
  } else {
    dif_stage_a<false, T, NT>(ire, iim, fre, fim, rk, pitch, log2n0, inv, sgn);
  }
  __syncthreads();
}

}  // namespace

}  // namespace nnf
