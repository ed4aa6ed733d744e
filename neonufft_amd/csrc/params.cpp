#include "params.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace nnf {

void kernel_width_and_beta(double tol, double upsampfac, int* ns_out, double* beta_out) {
  int ns;
  if (upsampfac == 2.0) {
    ns = static_cast<int>(std::ceil(-std::log10(tol / 10.0)));
  } else {
    ns = static_cast<int>(std::ceil(-std::log(tol) / (kPi * std::sqrt(1.0 - 1.0 / upsampfac))));
  }
  ns = std::max(2, std::min(16, ns));
  double beta_fac = 2.30;
  if (ns == 2) beta_fac = 2.20;
  if (ns == 3) beta_fac = 2.26;
  if (ns == 4) beta_fac = 2.38;
  if (upsampfac != 2.0) beta_fac = 0.97 * kPi * (1.0 - 1.0 / (2.0 * upsampfac));
  *ns_out = ns;
  *beta_out = beta_fac * ns;
}

double es_kernel(double t, int ns, double beta) {
  const double c = 4.0 / (static_cast<double>(ns) * ns);
  const double a = 1.0 - c * t * t;
  return a > 0.0 ? std::exp(beta * (std::sqrt(a) - 1.0)) : 0.0;
}

void gauss_legendre(int n, double* x, double* w) {
  for (int i = 0; i < (n + 1) / 2; ++i) {
    double z = std::cos(kPi * (i + 0.75) / (n + 0.5));
    double dp = 1.0;
    for (int it = 0; it < 100; ++it) {
      double p0 = 1.0, p1 = 0.0;
      for (int j = 1; j <= n; ++j) {
        const double p2 = p1;
        p1 = p0;
        p0 = ((2.0 * j - 1.0) * z * p1 - (j - 1.0) * p2) / j;
      }
      dp = n * (z * p0 - p1) / (z * z - 1.0);
      const double dz = p0 / dp;
      z -= dz;
      if (std::fabs(dz) < 1e-16) break;
    }
    double p0 = 1.0, p1 = 0.0;
    for (int j = 1; j <= n; ++j) {
      const double p2 = p1;
      p1 = p0;
      p0 = ((2.0 * j - 1.0) * z * p1 - (j - 1.0) * p2) / j;
    }
    dp = n * (z * p0 - p1) / (z * z - 1.0);
    const double wi = 2.0 / ((1.0 - z * z) * dp * dp);
    x[i] = -z;
    x[n - 1 - i] = z;
    w[i] = wi;
    w[n - 1 - i] = wi;
  }
  if (n % 2 == 1) x[n / 2] = 0.0;
}

int quad_nodes(int ns) { return ns + ns % 2 + 10; }

double fit_horner(int ns, double beta, int deg, double* coef) {
  const int nc = deg + 1;
  std::vector<double> yk(nc), fk(nc), cheb(nc), mono(nc);
  // Chebyshev polynomial T_m as monomial coefficient rows.
  std::vector<std::vector<double>> Tm(nc, std::vector<double>(nc, 0.0));
  Tm[0][0] = 1.0;
  if (nc > 1) Tm[1][1] = 1.0;
  for (int m = 2; m < nc; ++m) {
    for (int k = 0; k < nc; ++k) {
      double v = -Tm[m - 2][k];
      if (k > 0) v += 2.0 * Tm[m - 1][k - 1];
      Tm[m][k] = v;
    }
  }
  for (int k = 0; k < nc; ++k) yk[k] = std::cos(kPi * (k + 0.5) / nc);
  double max_err = 0.0;
  for (int j = 0; j < ns; ++j) {
    auto f = [&](double y) { return es_kernel(0.5 * (y + 1.0) + j - 0.5 * ns, ns, beta); };
    for (int k = 0; k < nc; ++k) fk[k] = f(yk[k]);
    for (int m = 0; m < nc; ++m) {
      double s = 0.0;
      for (int k = 0; k < nc; ++k) s += fk[k] * std::cos(m * kPi * (k + 0.5) / nc);
      cheb[m] = (m == 0 ? 1.0 : 2.0) * s / nc;
    }
    std::fill(mono.begin(), mono.end(), 0.0);
    for (int m = 0; m < nc; ++m)
      for (int k = 0; k < nc; ++k) mono[k] += cheb[m] * Tm[m][k];
    for (int k = 0; k < nc; ++k) coef[j * nc + k] = mono[k];
    for (int s = 0; s <= 256; ++s) {
      const double y = -1.0 + 2.0 * s / 256.0;
      double v = mono[deg];
      for (int k = deg - 1; k >= 0; --k) v = v * y + mono[k];
      max_err = std::max(max_err, std::fabs(v - f(y)));
    }
  }
  return max_err;
}

KernelSetup make_kernel_setup(double tol, double upsampfac, bool use_horner) {
  if (!(tol > 0.0) || !std::isfinite(tol)) throw neonufft::InputError("Neonufft:: tolerance must be positive");
  if (!(upsampfac > 1.0) || !std::isfinite(upsampfac))
    throw neonufft::InputError("Neonufft:: upsampling factor must be > 1");
  KernelSetup ks;
  kernel_width_and_beta(tol, upsampfac, &ks.ns, &ks.beta);
  std::memset(&ks.es, 0, sizeof(ks.es));
  ks.es.beta = ks.beta;
  if (use_horner) {
    ks.horner_deg = horner_degree(ks.ns);
    std::vector<double> c(static_cast<std::size_t>(ks.ns) * (ks.horner_deg + 1));
    fit_horner(ks.ns, ks.beta, ks.horner_deg, c.data());
    for (std::size_t i = 0; i < c.size(); ++i) ks.es.coef[i] = static_cast<float>(c[i]);
  }
  const int nq = quad_nodes(ks.ns);
  std::vector<double> x(2 * nq), w(2 * nq);
  gauss_legendre(2 * nq, x.data(), w.data());
  ks.quad.nq = nq;
  const double hw = 0.5 * ks.ns;
  for (int q = 0; q < nq; ++q) {
    const double z = x[nq + q];  // positive half
    ks.quad.node[q] = z * hw;
    ks.quad.weight[q] = 2.0 * hw * w[nq + q] * es_kernel(z * hw, ks.ns, ks.beta);
  }
  return ks;
}

int band_rows(IntType n0, IntType pitch, IntType n1, int ns, std::size_t acc_bytes,
              std::size_t budget, int min_bands) {
  if (n0 + ns / 2 > 65535 || n1 > 65535 || n1 < ns) return 0;
  const std::size_t row_bytes = 2 * static_cast<std::size_t>(pitch) * acc_bytes;
  IntType rows = static_cast<IntType>(budget / row_bytes);
  rows = std::min<IntType>(rows, n1 - ns + 1);
  if (min_bands > 1) rows = std::min<IntType>(rows, std::max<IntType>(2, n1 / min_bands));
  return rows >= 1 ? static_cast<int>(rows) : 0;
}

BinGeom choose_bins(int dim, const IntType* n, int ns, std::size_t elem_bytes,
                    const int* bin_override) {
  BinGeom g{};
  int pref[3] = {1, 1, 1};
  // Defaults measured on MI355X for uniform points at ~0.6 points per cell
  // (3D, ns = 6): 32 x 8 x 8 tiles with up to 2048 points per workgroup gave
  // the fastest spreading and interpolation (the LDS budget below shrinks
  // them for wider kernels and fp64).
  if (dim == 1) {
    pref[0] = 1024;
  } else if (dim == 2) {
    pref[0] = 64;
    pref[1] = 32;
  } else {
    pref[0] = 32;
    pref[1] = 8;
    pref[2] = 8;
  }
  if (bin_override) {
    for (int d = 0; d < dim; ++d)
      if (bin_override[d] > 0) pref[d] = bin_override[d];
  }
  // LDS tile bytes per workgroup (160 KB per CU on gfx950). Measured on MI355X
  // with 128^3 modes: fp64 tiles need most of it, since at ns >= 9 a 60 KB
  // tile leaves 4x4x4 bins whose halo is 50x the bin (tol 1e-10: type 1
  // 39.6 -> 10.5 ms, type 2 20.3 -> 6.6 ms); fp32 tiles up to 100 KB keep
  // 32x8x8 bins through ns = 8 (type 2 at ns = 7: 1.13 -> 0.97 ms, ns <= 6
  // unchanged).
  const std::size_t budget = elem_bytes >= 16 ? 144 * 1024 : 100 * 1024;
  auto bytes = [&](const int* b) {
    std::size_t cells = 1;
    for (int d = 0; d < dim; ++d) cells *= static_cast<std::size_t>(b[d] + ns);
    return cells * elem_bytes;
  };
  // shrink the largest bin edge until the LDS tile fits the budget
  while (bytes(pref) > budget) {
    int dmax = 0;
    for (int d = 1; d < dim; ++d)
      if (pref[d] > pref[dmax]) dmax = d;
    if (pref[dmax] <= 1) break;
    pref[dmax] = std::max(1, pref[dmax] / 2);
  }
  // Prefer bin edges that leave no partial last bin narrower than ns (the tile
  // reduction then only needs a bin's direct neighbours): search the nearest
  // edge within max(a quarter of the preferred one, ns) whose tile still fits.
  for (int d = 0; d < dim; ++d) {
    const int nd = static_cast<int>(n[d]);
    const int B = std::min(pref[d], nd);
    const auto good = [&](int b) { const int r = nd % b; return r == 0 || r >= ns; };
    if (B < ns || good(B)) continue;
    const int span = std::max(B / 4, ns);
    for (int delta = 1; delta <= span; ++delta) {
      bool found = false;
      for (int cand : {B - delta, B + delta}) {
        if (cand < ns || cand > nd) continue;
        int trial[3] = {pref[0], pref[1], pref[2]};
        trial[d] = cand;
        if (good(cand) && bytes(trial) <= budget) {
          pref[d] = cand;
          found = true;
          break;
        }
      }
      if (found) break;
    }
  }
  for (int d = 0; d < 3; ++d) {
    if (d < dim) {
      const int nd = static_cast<int>(n[d]);
      const int cap = dim == 3 ? CellPack<3>::max_bin : (dim == 2 ? CellPack<2>::max_bin : CellPack<1>::max_bin);
      g.n[d] = nd;
      g.bin[d] = std::max(1, std::min({pref[d], nd, cap}));
      g.nbins[d] = (nd + g.bin[d] - 1) / g.bin[d];
      g.tile[d] = g.bin[d] + ns;
    } else {
      g.n[d] = 1;
      g.bin[d] = 1;
      g.nbins[d] = 1;
      g.tile[d] = 1;
    }
  }
  return g;
}

IntType next235even(IntType n) {
  if (n <= 2) return 2;
  if (n % 2 == 1) n += 1;
  for (IntType cand = n;; cand += 2) {
    IntType r = cand;
    while (r % 2 == 0) r /= 2;
    while (r % 3 == 0) r /= 3;
    while (r % 5 == 0) r /= 5;
    if (r == 1) return cand;
  }
}

IntType t3_grid_size(double S, double X, double upsampfac, int ns) {
  double Xs = X, Ss = S;
  if (X == 0.0) {
    if (S == 0.0) {
      Xs = 1.0;
      Ss = 1.0;
    } else {
      Xs = std::max(Xs, 1.0 / S);
    }
  } else {
    Ss = std::max(Ss, 1.0 / X);
  }
  IntType size = static_cast<IntType>(2.0 * upsampfac * Ss * Xs / kPi + static_cast<double>(ns + 1));
  if (size < 2 * ns) size = 2 * ns;
  size += size % 2;
  return size;
}

GridInfoT3 make_grid_info_t3(int dim, int ns, double upsampfac, double recenter_threshold,
                             const double* in_min, const double* in_max, const double* out_min,
                             const double* out_max) {
  GridInfoT3 gi;
  gi.dim = dim;
  for (int d = 0; d < dim; ++d) {
    if (!(in_min[d] <= in_max[d])) throw neonufft::InputError("Invalid input bounds.");
    if (!(out_min[d] <= out_max[d])) throw neonufft::InputError("Invalid output bounds.");
  }
  for (int d = 0; d < dim; ++d) {
    double xo = 0.5 * (in_max[d] + in_min[d]);
    double so = 0.5 * (out_max[d] + out_min[d]);
    double X = std::max(std::fabs(in_max[d] - xo), std::fabs(in_min[d] - xo));
    double S = std::max(std::fabs(out_max[d] - so), std::fabs(out_min[d] - so));
    if (std::fabs(xo) < recenter_threshold * X) {
      xo = 0.0;
      X = std::max(std::fabs(in_max[d]), std::fabs(in_min[d]));
    }
    if (std::fabs(so) < recenter_threshold * S) {
      so = 0.0;
      S = std::max(std::fabs(out_max[d]), std::fabs(out_min[d]));
    }
    gi.input_offsets[d] = xo;
    gi.output_offsets[d] = so;
    gi.input_half_extent[d] = X;
    gi.output_half_extent[d] = S;
    const IntType n = t3_grid_size(S, X, upsampfac, ns);
    gi.spread_grid_size[d] = n;
    gi.fft_grid_size[d] =
        next235even(std::max<IntType>(2 * ns, static_cast<IntType>(static_cast<double>(n) * upsampfac)));
    // util.hpp:91-105: h = 2 pi / n; gamma = n / (2 sigma S), S made safe.
    gi.h[d] = 2.0 * kPi / static_cast<double>(n);
    double Ss = S;
    if (X == 0.0 && S == 0.0) {
      Ss = 1.0;
    } else if (S == 0.0) {
      Ss = std::max(Ss, 1.0 / X);
    }
    gi.gamma[d] = static_cast<double>(n) / (2.0 * upsampfac * Ss);
  }
  return gi;
}

}  // namespace nnf
