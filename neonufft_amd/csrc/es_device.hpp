// Device helpers shared by the spreading / interpolation kernels: workgroup
// to XCD mapping, periodic index wrap, ES kernel evaluation (per-piece Horner
// polynomials for fp32, the exact formula for fp64) and the fixed-point LDS
// accumulation used by spreading.
#pragma once

#include "kernels.hpp"

namespace nnf {

namespace {

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

// i = q d + r with 0 <= r < d for 0 <= i < 2^22 and 0 < d < 2^16, from a float
// reciprocal rd = 1/d: the product i * rd carries a relative error below
// 2^-23, i.e. less than (i / d) 2^-23 < 1 / (2 d) absolute, which is less than
// the distance from i / d up to the next integer; so its truncation is q or
// q - 1, fixed by one correction. Returns r, sets *q.
__device__ inline int div_rem(int i, int d, float rd, int* q) {
  int t = static_cast<int>(static_cast<float>(i) * rd);
  int r = i - __mul24(t, d);  // t d <= i < 2^22: exact in the 24-bit multiplier
  if (r >= d) {
    ++t;
    r -= d;
  }
  *q = t;
  return r;
}

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


// ---------------------------------------------------------------------------
// Spreading
//
// The tile accumulates in fixed point with integer LDS atomics: on gfx950,
// ds_add_u32 sustains 5-9e12 lane-ops/s while ds_add_f32 runs at ~2e11.
// Each subproblem picks its own scale from a rigorous bound on any tile value,
// bound = sum over its points of |Re c| + |Im c| (kernel weights are <= 1),
// mapped to FixedAcc<T>::one: the full int32 (fp32 plans) / int64 (fp64
// plans) range less a margin for the rounding of each added term (at most
// one half per add; a cell receives at most a few thousand adds). Finished
// tiles are converted back once per cell.
//
// fp32 tiles pack a cell into one 64-bit word, re + 2^32 im, so that one
// ds_add_u64 replaces two ds_add_u32 (the atomic rate, not the VALU, bounds
// spreading). Each added word is the rounded (re_i, im_i) with the borrow of
// a negative low word taken from the high one (high = im_i + (re_i >> 31)),
// so the 64-bit sum is exactly sum re + 2^32 sum im; |sum re| < 2^31 gives
// re = low word (signed) and im = high word + [re < 0]. The tile holds the
// same integers as separate re / im planes would.
// ---------------------------------------------------------------------------
template <typename T>
struct FixedAcc;
template <>
struct FixedAcc<float> {
  using type = int;
  static constexpr float one = 2147483648.0f - 2097152.0f;  // 2^31 - 2^21
};
template <>
struct FixedAcc<double> {
  using type = long long;
  static constexpr double one = 9223372036854775808.0 - 8796093022208.0;  // 2^63 - 2^43
};

// Round to the nearest integer (floor(v + 0.5)) in one VALU instruction; the
// compiler emits rndne + cvt (two) for __float2int_rn. |v| < 2^31 here.
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
// fp32 packed cell: adds re + 2^32 im (low word re, high word im plus the
// borrow of a negative re).
__device__ inline void lds_add_packed(unsigned long long* p, int re, int im) {
  const unsigned long long w =
      (static_cast<unsigned long long>(static_cast<uint32_t>(im + (re >> 31))) << 32) |
      static_cast<uint32_t>(re);
  __hip_atomic_fetch_add(p, w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
// Sum of packed words -> (re, im) in fixed point.
__device__ inline void unpack_packed(unsigned long long w, int* re, int* im) {
  *re = static_cast<int>(static_cast<uint32_t>(w));
  *im = static_cast<int>(static_cast<uint32_t>(w >> 32)) + (*re < 0 ? 1 : 0);
}

}  // namespace

}  // namespace nnf
