// Index maps between retained modes, fine-grid cells and the caller's mode
// order, shared by the pruned-FFT kernels (pruned.hip) and the fused axis
// passes (rows.hip).
//
// Retained index j in [0, N) of an axis with n fine cells: fine index
// j < P ? j : j + n - N, mode k = j < P ? j : j - N, with P = N - N/2.
#pragma once

namespace nnf {

__device__ inline int ret_fine(int j, int n, int N) {
  const int P = N - N / 2;
  return j < P ? j : j + n - N;
}
__device__ inline int ret_mode(int j, int N) {
  const int P = N - N / 2;
  return j < P ? j : j - N;
}
// Array index of mode k in the caller's order (CMCL: k + N/2; FFT: k, negatives after).
__device__ inline int order_index(int k, int N, int order_fft) {
  return order_fft ? (k >= 0 ? k : k + N) : k + N / 2;
}
// Retained index of fine cell f, or -1 for a padding cell.
__device__ inline int fine_ret(int f, int n, int N) {
  const int P = N - N / 2;
  if (f < P) return f;
  if (f >= n - N / 2) return f - (n - N);
  return -1;
}

}  // namespace nnf
