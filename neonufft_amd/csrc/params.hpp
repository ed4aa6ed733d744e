// Host-side numerical set-up of a plan: kernel width and shape, the fp32
// Horner table of the ES kernel, the quadrature for its Fourier transform,
// the bin (subgrid tile) geometry and the type-3 grid sizing.
#pragma once

#include <array>

#include "kernels.hpp"

namespace nnf {

struct KernelSetup {
  int ns = 0;          // kernel width (n_spread)
  double beta = 0.0;   // ES shape (es_beta)
  int horner_deg = 0;  // fp32 polynomial degree (0 for fp64 plans)
  EsParams es{};       // device-side evaluation parameters
  QuadRule quad{};     // phihat quadrature
};

// n_spread and es_beta from (tol, upsampfac):
// src/neonufft/es_kernel_param.hpp:24-56 of the reference.
void kernel_width_and_beta(double tol, double upsampfac, int* ns, double* beta);

// phi(t) = exp(beta (sqrt(1 - (2t/ns)^2) - 1)) for |t| < ns/2, else 0.
double es_kernel(double t, int ns, double beta);

// Gauss-Legendre rule of order n on [-1, 1] (ascending nodes), FP64 Newton.
void gauss_legendre(int n, double* x, double* w);

// Number of Gauss-Legendre nodes on (0,1) used for phihat: ns + ns%2 + 10,
// the count of the reference's CPU path (src/neonufft/kernels/fseries_kernel.cpp:35).
int quad_nodes(int ns);

// Builds KernelSetup for a plan precision (use_horner: fp32 plans).
KernelSetup make_kernel_setup(double tol, double upsampfac, bool use_horner);

// Chebyshev fit of the ES kernel pieces, converted to monomial coefficients in
// y = 2x - 1; returns the max abs error over a dense check grid.
double fit_horner(int ns, double beta, int deg, double* coef /* ns*(deg+1) */);

// Bin geometry for a grid of size n (DIM entries), element size and ns.
// bin_override: if non-null and positive entries, used instead of the default.
BinGeom choose_bins(int dim, const IntType* n, int ns, std::size_t elem_bytes,
                    const int* bin_override = nullptr);

// Rows per band of the row-band spreading of an n0 x n1 grid with LDS rows of
// `pitch` cells: as many as two accumulator planes (acc_bytes per cell) of
// `budget` LDS bytes hold, at most n1 - ns + 1 (the point range of a band must
// not wrap onto itself) and at most max(2, n1 / min_bands) so that a grid
// yields about min_bands bands; 0 when the grid does not suit the row view
// (n0 + ns/2 or n1 above 65535, or one row over budget).
int band_rows(IntType n0, IntType pitch, IntType n1, int ns, std::size_t acc_bytes,
              std::size_t budget, int min_bands);

// LDS bytes of one row band: two bands per CU (4 rows of a 2048-cell fp32
// grid with the fused FFT's padded pitch fit).
constexpr std::size_t kBandBudget = 68 * 1024;

// Default number of points per subproblem.
constexpr uint32_t kDefaultMaxPoints = 2048;

// Type-3 grid sizing: src/neonufft/util/grid_info_t3.hpp:18-52 and
// src/contrib/es_kernel/util.hpp:39-105 of the reference, in FP64.
struct GridInfoT3 {
  int dim = 0;
  double input_offsets[3] = {0, 0, 0};
  double output_offsets[3] = {0, 0, 0};
  double input_half_extent[3] = {0, 0, 0};
  double output_half_extent[3] = {0, 0, 0};
  IntType spread_grid_size[3] = {1, 1, 1};
  IntType fft_grid_size[3] = {1, 1, 1};
  double h[3] = {0, 0, 0};     // spread grid spacing 2 pi / n
  double gamma[3] = {1, 1, 1}; // input scale: x' = (x - xbar) / gamma
};

IntType next235even(IntType n);
IntType t3_grid_size(double S, double X, double upsampfac, int ns);
GridInfoT3 make_grid_info_t3(int dim, int ns, double upsampfac, double recenter_threshold,
                             const double* in_min, const double* in_max, const double* out_min,
                             const double* out_max);

}  // namespace nnf
