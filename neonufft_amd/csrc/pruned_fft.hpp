// Pruned 2D / 3D FFT between a batch of fine grids and their kept modes
// (type 1: FFT, truncate, deconvolve; type 2: deconvolve, zero-pad, FFT).
//
// Same results as the full in-place transform followed by downsampling (or
// upsampling followed by it); the work is cut by transforming one axis at a
// time with contiguous 1D rocFFT passes and carrying only the retained
// entries of the axes already done:
//   G[b][z][y][x] --x-FFT--> compact-transpose --> H[b][z][x'][y]
//   --y-FFT--> compact-transpose --> K[b][y'][x'][z] --z-FFT--> modes
// (2D stops after H). Type 2 runs the mirror image with zero-padding
// transposes. 3D plans on power-of-two grids replace the passes after H with
// the LDS axis kernels of rows.hip and use only the h() and k() arrays.
#pragma once

#include <memory>

#include "fft.hpp"
#include "kernels.hpp"
#include "params.hpp"

namespace nnf {

template <typename T>
void launch_compact_transpose(const Cplx<T>* src, Cplx<T>* dst, int R, int nc, int Nc,
                              long long src_rs, long long dst_rs, std::size_t slabs, int lo_count,
                              long long src_hi, long long src_lo, long long dst_hi,
                              long long dst_lo, hipStream_t stream);
template <typename T>
void launch_expand_transpose(const Cplx<T>* src, Cplx<T>* dst, int R, int nf, int Nf,
                             long long src_rs, long long dst_rs, std::size_t slabs, int lo_count,
                             long long src_hi, long long src_lo, long long dst_hi,
                             long long dst_lo, hipStream_t stream);
template <typename T, int DIM>
void launch_modes_out(const Cplx<T>* src, long long member_stride, Cplx<T>* single,
                      Cplx<T>* const* outs, int count, int N0, int Nm, int nl, int Nl,
                      long long st0, long long stm, long long stl, const T* fs0, const T* fsm,
                      const T* fsl, int order_fft, hipStream_t stream);
template <typename T, int DIM>
void launch_modes_in(const Cplx<T>* single, const Cplx<T>* const* ins, int count, Cplx<T>* dst,
                     long long member_stride, int N0, int Nm, int nl, int Nl, long long st0,
                     long long stm, long long stl, const T* fs0, const T* fsm, const T* fsl,
                     int order_fft, hipStream_t stream);

// DIM 2 or 3 (1D plans transform their single axis directly).
template <typename T, int DIM>
class PrunedFft {
public:
  PrunedFft() = default;
  PrunedFft(const PrunedFft&) = delete;
  PrunedFft& operator=(const PrunedFft&) = delete;
  PrunedFft(PrunedFft&&) = default;
  PrunedFft& operator=(PrunedFft&&) = default;

  // n: fine grid, N: modes (fastest first); `count` grids, contiguous.
  void init(const IntType* n, const IntType* N, int sign, int count, hipStream_t stream,
            const std::shared_ptr<neonufft::Allocator>& alloc) {
    for (int d = 0; d < 3; ++d) {
      n_[d] = d < DIM ? static_cast<long long>(n[d]) : 1;
      N_[d] = d < DIM ? static_cast<long long>(N[d]) : 1;
    }
    count_ = count;
    y_rows_ = 0;
    if (DIM == 2 && n_[1] >= 32 && (n_[1] & (n_[1] - 1)) == 0) {
      y_rows_ = axis_fft_rows(static_cast<IntType>(n_[1]), sizeof(T), kBandBudget);
    }
    const bool dbl = std::is_same_v<T, double>;
    IntType len = n_[0];
    px_ = FftPlan(dbl, 1, &len, sign, static_cast<IntType>(count * n_[1] * n_[2]), stream, alloc);
    len = n_[1];
    py_ = FftPlan(dbl, 1, &len, sign, static_cast<IntType>(count * n_[2] * N_[0]), stream, alloc);
    h_.reset(static_cast<std::size_t>(count * n_[2] * N_[0] * n_[1]), alloc);
    if (DIM == 3) {
      len = n_[2];
      pz_ = FftPlan(dbl, 1, &len, sign, static_cast<IntType>(count * N_[1] * N_[0]), stream, alloc);
      k_.reset(static_cast<std::size_t>(count * N_[1] * N_[0] * n_[2]), alloc);
    }
  }

  // Intermediate after the x pass: per grid and z plane, N0 x-modes (retained
  // order) of n1 cells each, y fastest. A caller that produces it directly
  // (row-band spreading or the row reduction with their own x-FFT) continues
  // with forward_from_h; type 2 stops there with backward_to_h.
  Cplx<T>* h() const { return h_.data(); }
  // 3D: intermediate after the y pass, per grid N1 retained y' x N0 x' lines
  // of n2 cells, z fastest (k[(y' N0 + x') n2 + z]).
  Cplx<T>* k() const { return k_.data(); }

  // Type 1 from a filled h() (x-FFT done, retained x modes written there by
  // the row-band or row-reduction kernels): the remaining passes and the
  // deconvolution into the outputs.
  void forward_from_h(T* const* fs, int order_fft, Cplx<T>* single, Cplx<T>* const* outs,
                      const IntType* st, hipStream_t s) {
    py_.execute(h_.data());
    finish_forward(fs, order_fft, single, outs, st, s);
  }

  // 2D: workgroup rows of the y pass in LDS (power-of-two n1 >= 32 within
  // the LDS budget), else 0.
  int y_lds_rows() const { return y_rows_; }

  // 2D type 1 from a filled h() with the y pass in LDS (y_lds_rows() > 0;
  // twiddles tw[N/2 + q] of the plan's sign up to n1): the y-FFT and the
  // deconvolution into the outputs read h once and write the modes once,
  // instead of the rocFFT pass plus the transposing deconvolution.
  void forward_from_h_lds(const Cplx<T>* tw, T* const* fs, int order_fft, Cplx<T>* single,
                          Cplx<T>* const* outs, const IntType* st, hipStream_t s) {
    Y2Job<T> j = y_job(tw, fs, order_fft, st);
    j.single = single;
    j.outs = outs;
    launch_y_out2<T>(j, s);
  }

  // 2D type 2 up to h() with the y pass in LDS (same conditions): the
  // deconvolved, zero-padded and y-transformed modes in one read of the modes
  // and one write of h.
  void backward_to_h_lds(const Cplx<T>* tw, const Cplx<T>* single, const Cplx<T>* const* ins,
                         const IntType* st, T* const* fs, int order_fft, hipStream_t s) {
    Y2Job<T> j = y_job(tw, fs, order_fft, st);
    j.single_in = single;
    j.ins = ins;
    launch_y_in2<T>(j, s);
  }

  // 2D type 2 into the fine grids with the y pass in LDS (y_lds_rows() > 0).
  void backward_lds(const Cplx<T>* tw, const Cplx<T>* single, const Cplx<T>* const* ins,
                    const IntType* st, T* const* fs, int order_fft, Cplx<T>* grids,
                    hipStream_t s) {
    backward_to_h_lds(tw, single, ins, st, fs, order_fft, s);
    expand_from_h(grids, s);
  }

  // Type 2 up to h(): deconvolution, zero padding and the passes after x; the
  // caller pads and transforms x (row kernel) into the fine grid.
  void backward_to_h(const Cplx<T>* single, const Cplx<T>* const* ins, const IntType* st,
                     T* const* fs, int order_fft, hipStream_t s) {
    const long long n1 = n_[1], n2 = n_[2], N0 = N_[0], N1 = N_[1], N2 = N_[2];
    if constexpr (DIM == 2) {
      launch_modes_in<T, 2>(single, ins, count_, h_.data(), N0 * n1, static_cast<int>(N0), 1,
                            static_cast<int>(n1), static_cast<int>(N1), st[0], 0, st[1], fs[0],
                            fs[1], fs[1], order_fft, s);
    } else {
      launch_modes_in<T, 3>(single, ins, count_, k_.data(), N1 * N0 * n2, static_cast<int>(N0),
                            static_cast<int>(N1), static_cast<int>(n2), static_cast<int>(N2),
                            st[0], st[1], st[2], fs[0], fs[1], fs[2], order_fft, s);
      pz_.execute(k_.data());
      launch_expand_transpose<T>(k_.data(), h_.data(), static_cast<int>(n2), static_cast<int>(n1),
                                 static_cast<int>(N1), N0 * n2, N0 * n1,
                                 static_cast<std::size_t>(count_ * N0), static_cast<int>(N0),
                                 N1 * N0 * n2, n2, n2 * N0 * n1, n1, s);
    }
    py_.execute(h_.data());
  }

  void set_stream(hipStream_t s) {
    px_.set_stream(s);
    py_.set_stream(s);
    if (DIM == 3) pz_.set_stream(s);
  }

  // Type 1: grids (destroyed) -> deconvolved modes in the caller's order and
  // strides; outputs `single` (count = 1) or the device table `outs`.
  void forward(Cplx<T>* grids, T* const* fs, int order_fft, Cplx<T>* single,
               Cplx<T>* const* outs, const IntType* st, hipStream_t s) {
    const long long n0 = n_[0], n1 = n_[1], n2 = n_[2], N0 = N_[0];
    const long long gstride = n0 * n1 * n2;
    px_.execute(grids);
    launch_compact_transpose<T>(grids, h_.data(), static_cast<int>(n1), static_cast<int>(n0),
                                static_cast<int>(N0), n0, n1,
                                static_cast<std::size_t>(count_ * n2), static_cast<int>(n2),
                                gstride, n1 * n0, n2 * N0 * n1, N0 * n1, s);
    py_.execute(h_.data());
    finish_forward(fs, order_fft, single, outs, st, s);
  }

  // Type 2: modes (caller's order and strides; `single` or the device table
  // `ins`) -> deconvolved, zero-padded and transformed fine grids.
  void backward(const Cplx<T>* single, const Cplx<T>* const* ins, const IntType* st,
                T* const* fs, int order_fft, Cplx<T>* grids, hipStream_t s) {
    backward_to_h(single, ins, st, fs, order_fft, s);
    expand_from_h(grids, s);
  }

private:
  // After the y-FFT of h: 2D deconvolves from h; 3D transposes to K, runs the
  // z-FFT and deconvolves.
  // h -> zero-padded x rows of the fine grids -> x-FFT.
  void expand_from_h(Cplx<T>* grids, hipStream_t s) {
    const long long n0 = n_[0], n1 = n_[1], n2 = n_[2], N0 = N_[0];
    const long long gstride = n0 * n1 * n2;
    launch_expand_transpose<T>(h_.data(), grids, static_cast<int>(n1), static_cast<int>(n0),
                               static_cast<int>(N0), n1, n0,
                               static_cast<std::size_t>(count_ * n2), static_cast<int>(n2),
                               n2 * N0 * n1, N0 * n1, gstride, n1 * n0, s);
    px_.execute(grids);
  }

  Y2Job<T> y_job(const Cplx<T>* tw, T* const* fs, int order_fft, const IntType* st) const {
    Y2Job<T> j{};
    j.h = h_.data();
    j.member_stride = N_[0] * n_[1];
    j.fs0 = fs[0];
    j.fs1 = fs[1];
    j.st0 = st[0];
    j.st1 = st[1];
    j.order_fft = order_fft;
    j.tw = tw;
    j.count = count_;
    j.N0 = static_cast<int>(N_[0]);
    j.N1 = static_cast<int>(N_[1]);
    j.n1 = static_cast<int>(n_[1]);
    j.rows = y_rows_;
    for (j.log2n1 = 0; (1LL << j.log2n1) < n_[1]; ++j.log2n1) {
    }
    for (j.log2rows = 0; (1 << j.log2rows) < y_rows_; ++j.log2rows) {
    }
    return j;
  }

  void finish_forward(T* const* fs, int order_fft, Cplx<T>* single, Cplx<T>* const* outs,
                      const IntType* st, hipStream_t s) {
    const long long n1 = n_[1], n2 = n_[2], N0 = N_[0], N1 = N_[1], N2 = N_[2];
    if constexpr (DIM == 2) {
      launch_modes_out<T, 2>(h_.data(), N0 * n1, single, outs, count_, static_cast<int>(N0), 1,
                             static_cast<int>(n1), static_cast<int>(N1), st[0], 0, st[1], fs[0],
                             fs[1], fs[1], order_fft, s);
    } else {
      launch_compact_transpose<T>(h_.data(), k_.data(), static_cast<int>(n2),
                                  static_cast<int>(n1), static_cast<int>(N1), N0 * n1, N0 * n2,
                                  static_cast<std::size_t>(count_ * N0), static_cast<int>(N0),
                                  n2 * N0 * n1, n1, N1 * N0 * n2, n2, s);
      pz_.execute(k_.data());
      launch_modes_out<T, 3>(k_.data(), N1 * N0 * n2, single, outs, count_, static_cast<int>(N0),
                             static_cast<int>(N1), static_cast<int>(n2), static_cast<int>(N2),
                             st[0], st[1], st[2], fs[0], fs[1], fs[2], order_fft, s);
    }
  }

  long long n_[3] = {1, 1, 1}, N_[3] = {1, 1, 1};
  int count_ = 0;
  int y_rows_ = 0;
  FftPlan px_, py_, pz_;
  DeviceBuffer<Cplx<T>> h_, k_;
};

}  // namespace nnf
