/*
 * nufft_oracle.c -- CPU restatement of the NeoNUFFT algorithm, in FP64.
 *
 * TEST INFRASTRUCTURE ONLY. This file is the checker for the HIP library in
 * neonufft_amd/. Only tests/, __graft_entry__.smoke() and the cpu_baseline leg
 * of bench.py may load it. The product path never calls it and must fail
 * loudly when its own HIP extension is missing.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the reference repository root). All arithmetic is IEEE double; complex
 * numbers are interleaved (re, im) pairs; grids are column-major (x fastest).
 *
 * Build: see oracle/Makefile (gcc -O2 -fopenmp -shared -fPIC).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef int64_t i64;

static const double ORC_PI = 3.14159265358979323846264338327950288;

/* ------------------------------------------------------------------------ */
/* Kernel parameters: src/neonufft/es_kernel_param.hpp:24-56                 */
/* ------------------------------------------------------------------------ */
int orc_kernel_params(double tol, double upsampfac, int* n_spread, double* es_beta) {
  int ns;
  if (upsampfac == 2.0) {
    ns = (int)ceil(-log10(tol / 10.0));
  } else {
    ns = (int)ceil(-log(tol) / (ORC_PI * sqrt(1.0 - 1.0 / upsampfac)));
  }
  if (ns < 2) ns = 2;
  if (ns > 16) ns = 16;
  double beta_fac = 2.30;
  if (ns == 2) beta_fac = 2.20;
  if (ns == 3) beta_fac = 2.26;
  if (ns == 4) beta_fac = 2.38;
  if (upsampfac != 2.0) beta_fac = 0.97 * ORC_PI * (1.0 - 1.0 / (2.0 * upsampfac));
  *n_spread = ns;
  *es_beta = beta_fac * ns;
  return 0;
}

/* ES kernel phi(t) = exp(beta*(sqrt(1-(2t/ns)^2)-1)) on |t| < ns/2, else 0.
 * src/neonufft/gpu/kernels/es_kernel_eval.cuh:36-46 (double variant). */
double orc_es_kernel(double t, int ns, double beta) {
  const double es_c = 4.0 / ((double)ns * (double)ns);
  const double arg = 1.0 - es_c * t * t;
  if (arg <= 0.0) return 0.0;
  return exp(beta * (sqrt(arg) - 1.0));
}

/* ------------------------------------------------------------------------ */
/* Gauss-Legendre rule on [-1,1], ascending nodes. Newton iteration on the   */
/* three-term Legendre recurrence (textbook method; the reference uses the   */
/* Glaser-Liu-Rokhlin solver, src/contrib/legendre_rule/legendre_rule_fast.* */
/* and gpu/kernels/legendre_rule.cuh:404-490, which yields the same rule).    */
/* ------------------------------------------------------------------------ */
void orc_gauss_legendre(int n, double* x, double* w) {
  for (int i = 0; i < (n + 1) / 2; ++i) {
    double z = cos(ORC_PI * (i + 0.75) / (n + 0.5));
    double dp = 0.0;
    for (int it = 0; it < 100; ++it) {
      double p0 = 1.0, p1 = 0.0;
      for (int j = 1; j <= n; ++j) {
        double p2 = p1;
        p1 = p0;
        p0 = ((2.0 * j - 1.0) * z * p1 - (j - 1.0) * p2) / j;
      }
      dp = n * (z * p0 - p1) / (z * z - 1.0);
      double dz = p0 / dp;
      z -= dz;
      if (fabs(dz) < 1e-16) break;
    }
    /* recompute derivative at the converged root */
    {
      double p0 = 1.0, p1 = 0.0;
      for (int j = 1; j <= n; ++j) {
        double p2 = p1;
        p1 = p0;
        p0 = ((2.0 * j - 1.0) * z * p1 - (j - 1.0) * p2) / j;
      }
      dp = n * (z * p0 - p1) / (z * z - 1.0);
    }
    double wi = 2.0 / ((1.0 - z * z) * dp * dp);
    x[i] = -z;
    x[n - 1 - i] = z;
    w[i] = wi;
    w[n - 1 - i] = wi;
  }
  if (n % 2 == 1) x[n / 2] = 0.0;
}

/* Continuous Fourier transform of phi (support [-ns/2, ns/2]) at angular
 * frequency `freq` in grid units: int phi(t) cos(freq*t) dt, by nq positive
 * nodes of a 2*nq point Gauss-Legendre rule (the kernel is even).
 * Restates the sums in gpu/kernels/fseries_kernel.cu:44-65 and
 * gpu/kernels/postphase_kernel.cu:43-69. */
static double orc_phihat_nodes(double freq, int ns, double beta, int nq, const double* z,
                               const double* wq) {
  const double hw = 0.5 * ns;
  double sum = 0.0;
  for (int q = 0; q < nq; ++q) {
    sum += 2.0 * hw * wq[q] * orc_es_kernel(z[q] * hw, ns, beta) * cos(freq * z[q] * hw);
  }
  return sum;
}

static void orc_half_rule(int nq, double* z, double* wq) {
  double* x = (double*)malloc(sizeof(double) * 2 * nq);
  double* w = (double*)malloc(sizeof(double) * 2 * nq);
  orc_gauss_legendre(2 * nq, x, w);
  for (int q = 0; q < nq; ++q) {
    z[q] = x[nq + q]; /* positive half */
    wq[q] = w[nq + q];
  }
  free(x);
  free(w);
}

double orc_phihat(double freq, int ns, double beta, int nq) {
  double z[64], wq[64];
  if (nq > 64) nq = 64;
  orc_half_rule(nq, z, wq);
  return orc_phihat_nodes(freq, ns, beta, nq, z, wq);
}

/* Deconvolution factors fs[k] = (-1)^k / phihat(2*pi*k/n), k = 0..n/2.
 * gpu/kernels/fseries_kernel.cu:24-67 (GPU), kernels/fseries_kernel.cpp:32-78 (CPU). */
void orc_fseries(i64 n, int ns, double beta, int nq, double* fs) {
  double z[64], wq[64];
  if (nq > 64) nq = 64;
  orc_half_rule(nq, z, wq);
  const i64 n_out = n / 2 + 1;
  for (i64 k = 0; k < n_out; ++k) {
    const double ph = orc_phihat_nodes(2.0 * ORC_PI * (double)k / (double)n, ns, beta, nq, z, wq);
    fs[k] = ((k % 2) ? -1.0 : 1.0) / ph;
  }
}

/* ------------------------------------------------------------------------ */
/* Fold onto [0,1): gpu/kernels/rescale_loc_kernel.cu:104-118                */
/* ------------------------------------------------------------------------ */
double orc_fold(double x) {
  double u = (x + ORC_PI) / (2.0 * ORC_PI);
  u = u - floor(u);
  if (u >= 1.0) u = 0.0;
  return u;
}

void orc_fold_array(i64 m, const double* x, double offset, double scale, double* u) {
#pragma omp parallel for schedule(static)
  for (i64 i = 0; i < m; ++i) u[i] = orc_fold((x[i] - offset) * scale);
}

/* ------------------------------------------------------------------------ */
/* Spreading (type 1): grid[(i0+j) mod n] += prod_d phi(i0_d+j_d-xi_d) * c    */
/* with xi = u*n, i0 = ceil(xi - ns/2), unpadded periodic wrap.               */
/* gpu/kernels/spreading_kernel.cu:588-660 (3D), :527-586 (2D), :477-525 (1D). */
/* ------------------------------------------------------------------------ */
static inline i64 orc_wrap(i64 i, i64 n) {
  i %= n;
  return i < 0 ? i + n : i;
}

static void orc_point_weights(double u, i64 n, int ns, double beta, i64* i0, double* ker) {
  const double xi = u * (double)n;
  const double start = ceil(xi - 0.5 * ns);
  *i0 = (i64)start;
  for (int j = 0; j < ns; ++j) ker[j] = orc_es_kernel(start + j - xi, ns, beta);
}

static void orc_spread_point(int dim, const i64* n, int ns, double beta, const double* u0,
                             const double* u1, const double* u2, i64 j, double cr, double ci,
                             double* grid) {
  double kx[16], ky[16], kz[16];
  i64 ix, iy = 0, iz = 0;
  orc_point_weights(u0[j], n[0], ns, beta, &ix, kx);
  int nsy = 1, nsz = 1;
  ky[0] = 1.0;
  kz[0] = 1.0;
  if (dim >= 2) {
    orc_point_weights(u1[j], n[1], ns, beta, &iy, ky);
    nsy = ns;
  }
  if (dim >= 3) {
    orc_point_weights(u2[j], n[2], ns, beta, &iz, kz);
    nsz = ns;
  }
  for (int c = 0; c < nsz; ++c) {
    const i64 gz = (dim >= 3) ? orc_wrap(iz + c, n[2]) : 0;
    for (int b = 0; b < nsy; ++b) {
      const i64 gy = (dim >= 2) ? orc_wrap(iy + b, n[1]) : 0;
      const double wyz = ky[b] * kz[c];
      const i64 row = (gz * (dim >= 2 ? n[1] : 1) + gy) * n[0];
      for (int a = 0; a < ns; ++a) {
        const i64 g = row + orc_wrap(ix + a, n[0]);
        const double w = kx[a] * wyz;
        grid[2 * g] += w * cr;
        grid[2 * g + 1] += w * ci;
      }
    }
  }
}

/* u[d]: folded coordinates in [0,1); c: M complex strengths; pre: optional M
 * complex factors multiplied into c (type-3 prephase, spreading_kernel.cu:628-632).
 * grid is accumulated into (not zeroed). Parallel over slabs of the last
 * dimension in two colour phases so that concurrent slabs never overlap. */
void orc_spread(int dim, const i64* n, int ns, double beta, i64 m, const double* u0,
                const double* u1, const double* u2, const double* c, const double* pre,
                double* grid) {
  const i64 nlast = n[dim - 1];
  const double* ulast = dim == 1 ? u0 : (dim == 2 ? u1 : u2);
  i64 nslab = nlast / ns;
  if (nslab % 2) nslab -= 1;
  if (nslab < 2) {
    for (i64 j = 0; j < m; ++j) {
      double cr = c[2 * j], ci = c[2 * j + 1];
      if (pre) {
        const double pr = pre[2 * j], pi = pre[2 * j + 1];
        const double tr = cr * pr - ci * pi;
        ci = cr * pi + ci * pr;
        cr = tr;
      }
      orc_spread_point(dim, n, ns, beta, u0, u1, u2, j, cr, ci, grid);
    }
    return;
  }
  /* bucket points by the slab of their first cell along the last dimension */
  i64* count = (i64*)calloc((size_t)nslab + 1, sizeof(i64));
  i64* slab_of = (i64*)malloc(sizeof(i64) * (size_t)(m > 0 ? m : 1));
  const i64 width = nlast / nslab; /* >= ns */
  for (i64 j = 0; j < m; ++j) {
    const double xi = ulast[j] * (double)nlast;
    i64 i0 = orc_wrap((i64)ceil(xi - 0.5 * ns), nlast);
    i64 s = i0 / width;
    if (s >= nslab) s = nslab - 1;
    slab_of[j] = s;
    count[s + 1]++;
  }
  for (i64 s = 0; s < nslab; ++s) count[s + 1] += count[s];
  i64* order = (i64*)malloc(sizeof(i64) * (size_t)(m > 0 ? m : 1));
  i64* fill = (i64*)calloc((size_t)nslab, sizeof(i64));
  for (i64 j = 0; j < m; ++j) {
    const i64 s = slab_of[j];
    order[count[s] + fill[s]++] = j;
  }
  for (int phase = 0; phase < 2; ++phase) {
#pragma omp parallel for schedule(dynamic, 1)
    for (i64 s = phase; s < nslab; s += 2) {
      for (i64 t = count[s]; t < count[s + 1]; ++t) {
        const i64 j = order[t];
        double cr = c[2 * j], ci = c[2 * j + 1];
        if (pre) {
          const double pr = pre[2 * j], pi = pre[2 * j + 1];
          const double tr = cr * pr - ci * pi;
          ci = cr * pi + ci * pr;
          cr = tr;
        }
        orc_spread_point(dim, n, ns, beta, u0, u1, u2, j, cr, ci, grid);
      }
    }
  }
  free(count);
  free(slab_of);
  free(order);
  free(fill);
}

/* Interpolation (type 2): out[j] = sum prod_d phi(...) * grid[(i0+j) mod n],
 * optionally times post[j]. gpu/kernels/interpolation_kernel.cu:193-319
 * (3D), :84-191 (2D), :27-82 (1D). */
void orc_interp(int dim, const i64* n, int ns, double beta, i64 m, const double* u0,
                const double* u1, const double* u2, const double* grid, const double* post,
                double* out) {
#pragma omp parallel for schedule(static)
  for (i64 j = 0; j < m; ++j) {
    double kx[16], ky[16], kz[16];
    i64 ix, iy = 0, iz = 0;
    orc_point_weights(u0[j], n[0], ns, beta, &ix, kx);
    int nsy = 1, nsz = 1;
    ky[0] = 1.0;
    kz[0] = 1.0;
    if (dim >= 2) {
      orc_point_weights(u1[j], n[1], ns, beta, &iy, ky);
      nsy = ns;
    }
    if (dim >= 3) {
      orc_point_weights(u2[j], n[2], ns, beta, &iz, kz);
      nsz = ns;
    }
    double sr = 0.0, si = 0.0;
    for (int c = 0; c < nsz; ++c) {
      const i64 gz = (dim >= 3) ? orc_wrap(iz + c, n[2]) : 0;
      for (int b = 0; b < nsy; ++b) {
        const i64 gy = (dim >= 2) ? orc_wrap(iy + b, n[1]) : 0;
        const double wyz = ky[b] * kz[c];
        const i64 row = (gz * (dim >= 2 ? n[1] : 1) + gy) * n[0];
        for (int a = 0; a < ns; ++a) {
          const i64 g = row + orc_wrap(ix + a, n[0]);
          const double w = kx[a] * wyz;
          sr += w * grid[2 * g];
          si += w * grid[2 * g + 1];
        }
      }
    }
    if (post) {
      const double pr = post[2 * j], pi = post[2 * j + 1];
      const double tr = sr * pr - si * pi;
      si = sr * pi + si * pr;
      sr = tr;
    }
    out[2 * j] = sr;
    out[2 * j + 1] = si;
  }
}

/* ------------------------------------------------------------------------ */
/* Mode <-> fine-grid index maps.                                            */
/* Mode k of an N-mode axis lies in [-N/2, N-1-N/2] (integer division).      */
/* CMCL order: array index i <-> k = i - N/2. FFT order: i < N-N/2 <-> k = i, */
/* otherwise k = i - N. Fine-grid index of k is k mod n.                      */
/* gpu/kernels/downsample_kernel.cu:133-191, upsample_kernel.cu:18-154.       */
/* ------------------------------------------------------------------------ */
static inline i64 orc_mode_of_index(i64 i, i64 N, int order_fft) {
  const i64 nneg = N / 2;
  if (order_fft) return (i < N - nneg) ? i : i - N;
  return i - nneg;
}

/* Type-1 deconvolution: out[i] = F[k mod n] * prod_d fs_d[|k_d|].
 * order_fft: 0 = CMCL (NEONUFFT_MODE_ORDER_CMCL), 1 = FFT order. */
void orc_downsample(int dim, const i64* n, const i64* N, const double* fs0, const double* fs1,
                    const double* fs2, int order_fft, const double* grid, const i64* out_strides,
                    double* out) {
  const i64 N1 = dim >= 2 ? N[1] : 1, N2 = dim >= 3 ? N[2] : 1;
#pragma omp parallel for collapse(2) schedule(static)
  for (i64 i2 = 0; i2 < N2; ++i2) {
    for (i64 i1 = 0; i1 < N1; ++i1) {
      double f21 = 1.0;
      i64 g21 = 0;
      if (dim >= 3) {
        const i64 k2 = orc_mode_of_index(i2, N[2], order_fft);
        f21 *= fs2[k2 < 0 ? -k2 : k2];
        g21 += orc_wrap(k2, n[2]) * n[1];
      }
      if (dim >= 2) {
        const i64 k1 = orc_mode_of_index(i1, N[1], order_fft);
        f21 *= fs1[k1 < 0 ? -k1 : k1];
        g21 += orc_wrap(k1, n[1]);
      }
      for (i64 i0 = 0; i0 < N[0]; ++i0) {
        const i64 k0 = orc_mode_of_index(i0, N[0], order_fft);
        const double f = f21 * fs0[k0 < 0 ? -k0 : k0];
        const i64 g = g21 * n[0] + orc_wrap(k0, n[0]);
        const i64 o = i0 * out_strides[0] + (dim >= 2 ? i1 * out_strides[1] : 0) +
                      (dim >= 3 ? i2 * out_strides[2] : 0);
        out[2 * o] = f * grid[2 * g];
        out[2 * o + 1] = f * grid[2 * g + 1];
      }
    }
  }
}

/* Type-2 (and type-3 inner) pre-correction: grid zeroed, then
 * F[k mod n] = in[i] * prod_d fs_d[|k_d|]. */
void orc_upsample(int dim, const i64* n, const i64* N, const double* fs0, const double* fs1,
                  const double* fs2, int order_fft, const double* in, const i64* in_strides,
                  double* grid) {
  const i64 G = n[0] * (dim >= 2 ? n[1] : 1) * (dim >= 3 ? n[2] : 1);
  memset(grid, 0, sizeof(double) * 2 * (size_t)G);
  const i64 N1 = dim >= 2 ? N[1] : 1, N2 = dim >= 3 ? N[2] : 1;
#pragma omp parallel for collapse(2) schedule(static)
  for (i64 i2 = 0; i2 < N2; ++i2) {
    for (i64 i1 = 0; i1 < N1; ++i1) {
      double f21 = 1.0;
      i64 g21 = 0;
      if (dim >= 3) {
        const i64 k2 = orc_mode_of_index(i2, N[2], order_fft);
        f21 *= fs2[k2 < 0 ? -k2 : k2];
        g21 += orc_wrap(k2, n[2]) * n[1];
      }
      if (dim >= 2) {
        const i64 k1 = orc_mode_of_index(i1, N[1], order_fft);
        f21 *= fs1[k1 < 0 ? -k1 : k1];
        g21 += orc_wrap(k1, n[1]);
      }
      for (i64 i0 = 0; i0 < N[0]; ++i0) {
        const i64 k0 = orc_mode_of_index(i0, N[0], order_fft);
        const double f = f21 * fs0[k0 < 0 ? -k0 : k0];
        const i64 g = g21 * n[0] + orc_wrap(k0, n[0]);
        const i64 o = i0 * in_strides[0] + (dim >= 2 ? i1 * in_strides[1] : 0) +
                      (dim >= 3 ? i2 * in_strides[2] : 0);
        grid[2 * g] = f * in[2 * o];
        grid[2 * g + 1] = f * in[2 * o + 1];
      }
    }
  }
}

/* ------------------------------------------------------------------------ */
/* Type-3 sizing: src/contrib/es_kernel/util.hpp:39-105                     */
/* ------------------------------------------------------------------------ */
i64 orc_next235even(i64 n) {
  if (n <= 2) return 2;
  if (n % 2 == 1) n += 1;
  for (i64 cand = n;; cand += 2) {
    i64 r = cand;
    while (r % 2 == 0) r /= 2;
    while (r % 3 == 0) r /= 3;
    while (r % 5 == 0) r /= 5;
    if (r == 1) return cand;
  }
}

/* S: output half extent, X: input half extent. util.hpp:64-88 */
i64 orc_t3_grid_size(double S, double X, double upsampfac, int ns) {
  double Xs = X, Ss = S;
  if (X == 0.0) {
    if (S == 0.0) {
      Xs = 1.0;
      Ss = 1.0;
    } else {
      Xs = fmax(Xs, 1.0 / S);
    }
  } else {
    Ss = fmax(Ss, 1.0 / X);
  }
  i64 size = (i64)(2.0 * upsampfac * Ss * Xs / ORC_PI + (double)(ns + 1));
  if (size < 2 * ns) size = 2 * ns;
  size += size % 2;
  return size;
}

/* util.hpp:91-105: h = 2 pi / n, gamma = n / (2 sigma S) */
void orc_set_nhg(double S, double X, double upsampfac, i64 n, double* h, double* gam) {
  *h = 2.0 * ORC_PI / (double)n;
  double Ss = S;
  if (X == 0.0 && S == 0.0) {
    Ss = 1.0;
  } else if (S == 0.0) {
    Ss = fmax(Ss, 1.0 / X);
  }
  *gam = (double)n / (2.0 * upsampfac * Ss);
}

/* Prephase e^{i sign sum_d x_d sbar_d}: gpu/kernels/prephase_kernel.cu:35-50 */
void orc_prephase(int dim, int sign, i64 m, const double* x0, const double* x1, const double* x2,
                  const double* sbar, double* out) {
#pragma omp parallel for schedule(static)
  for (i64 j = 0; j < m; ++j) {
    double ph = x0[j] * sbar[0];
    if (dim >= 2) ph += x1[j] * sbar[1];
    if (dim >= 3) ph += x2[j] * sbar[2];
    out[2 * j] = cos(ph);
    out[2 * j + 1] = sign * sin(ph);
  }
}

/* Postphase e^{i sign sum_d (s_d - sbar_d) xbar_d} / prod_d phihat((s_d - sbar_d) scale_d):
 * gpu/kernels/postphase_kernel.cu:27-76. */
void orc_postphase(int dim, int sign, int ns, double beta, int nq, i64 m, const double* s0,
                   const double* s1, const double* s2, const double* xbar, const double* sbar,
                   const double* scale, double* out) {
  double z[64], wq[64];
  if (nq > 64) nq = 64;
  orc_half_rule(nq, z, wq);
#pragma omp parallel for schedule(static)
  for (i64 j = 0; j < m; ++j) {
    const double* s[3] = {s0, s1, s2};
    double phihat = 1.0, ph = 0.0;
    for (int d = 0; d < dim; ++d) {
      const double t = s[d][j] - sbar[d];
      phihat *= orc_phihat_nodes(t * scale[d], ns, beta, nq, z, wq);
      ph += t * xbar[d];
    }
    out[2 * j] = cos(ph) / phihat;
    out[2 * j + 1] = sign * sin(ph) / phihat;
  }
}

/* ------------------------------------------------------------------------ */
/* Direct NUDFT sums (the closed-form ground truth), FP64.                   */
/* tests/nuft_direct_kernels.cpp:126-224 (type 1), :27-124 (type 2),          */
/* :227-293 (type 3).                                                         */
/* ------------------------------------------------------------------------ */

/* e^{i s k x} for k in [k0, k0+len), reseeded by a direct sincos every 32 steps. */
static void orc_power_row(double x, int sign, i64 k0, i64 len, double* re, double* im) {
  const double cs = cos(x), sn = sign * sin(x);
  double cr = 0, ci = 0;
  for (i64 t = 0; t < len; ++t) {
    if (t % 32 == 0) {
      const double a = sign * (double)(k0 + t) * x;
      cr = cos(a);
      ci = sin(a);
    } else {
      const double tr = cr * cs - ci * sn;
      ci = cr * sn + ci * cs;
      cr = tr;
    }
    re[t] = cr;
    im[t] = ci;
  }
}

/* Type 1 at S sampled integer mode vectors kk[3*s + d]:
 * out[s] = sum_j c_j exp(i sign k_s . x_j). */
void orc_direct_t1(int dim, int sign, i64 m, const double* x0, const double* x1, const double* x2,
                   const double* c, i64 S, const i64* kk, double* out) {
  i64 kmin[3] = {0, 0, 0}, kmax[3] = {0, 0, 0};
  for (int d = 0; d < dim; ++d) {
    kmin[d] = kk[d];
    kmax[d] = kk[d];
    for (i64 s = 0; s < S; ++s) {
      if (kk[3 * s + d] < kmin[d]) kmin[d] = kk[3 * s + d];
      if (kk[3 * s + d] > kmax[d]) kmax[d] = kk[3 * s + d];
    }
  }
  memset(out, 0, sizeof(double) * 2 * (size_t)S);
#pragma omp parallel
  {
    double* acc = (double*)calloc(2 * (size_t)S, sizeof(double));
    double* tab[3][2];
    const double* xs[3] = {x0, x1, x2};
    for (int d = 0; d < 3; ++d) {
      const i64 len = kmax[d] - kmin[d] + 1;
      tab[d][0] = (double*)malloc(sizeof(double) * (size_t)len);
      tab[d][1] = (double*)malloc(sizeof(double) * (size_t)len);
    }
#pragma omp for schedule(static)
    for (i64 j = 0; j < m; ++j) {
      for (int d = 0; d < dim; ++d)
        orc_power_row(xs[d][j], sign, kmin[d], kmax[d] - kmin[d] + 1, tab[d][0], tab[d][1]);
      const double cr = c[2 * j], ci = c[2 * j + 1];
      for (i64 s = 0; s < S; ++s) {
        double er = 1.0, ei = 0.0;
        for (int d = 0; d < dim; ++d) {
          const i64 t = kk[3 * s + d] - kmin[d];
          const double ar = tab[d][0][t], ai = tab[d][1][t];
          const double tr = er * ar - ei * ai;
          ei = er * ai + ei * ar;
          er = tr;
        }
        acc[2 * s] += cr * er - ci * ei;
        acc[2 * s + 1] += cr * ei + ci * er;
      }
    }
#pragma omp critical
    {
      for (i64 s = 0; s < 2 * S; ++s) out[s] += acc[s];
    }
    for (int d = 0; d < 3; ++d) {
      free(tab[d][0]);
      free(tab[d][1]);
    }
    free(acc);
  }
}

/* Type 2 at S points: out[s] = sum_k f[k] exp(i sign k . x_s), f contiguous
 * CMCL (index i_d <-> k_d = i_d - N_d/2), x fastest. */
void orc_direct_t2(int dim, int sign, const i64* N, const double* f, i64 S, const double* x0,
                   const double* x1, const double* x2, double* out) {
  const i64 N0 = N[0], N1 = dim >= 2 ? N[1] : 1, N2 = dim >= 3 ? N[2] : 1;
#pragma omp parallel for schedule(dynamic, 1)
  for (i64 s = 0; s < S; ++s) {
    double* t0r = (double*)malloc(sizeof(double) * (size_t)N0);
    double* t0i = (double*)malloc(sizeof(double) * (size_t)N0);
    double* t1r = (double*)malloc(sizeof(double) * (size_t)N1);
    double* t1i = (double*)malloc(sizeof(double) * (size_t)N1);
    double* t2r = (double*)malloc(sizeof(double) * (size_t)N2);
    double* t2i = (double*)malloc(sizeof(double) * (size_t)N2);
    orc_power_row(x0[s], sign, -(N0 / 2), N0, t0r, t0i);
    if (dim >= 2) orc_power_row(x1[s], sign, -(N1 / 2), N1, t1r, t1i);
    else { t1r[0] = 1.0; t1i[0] = 0.0; }
    if (dim >= 3) orc_power_row(x2[s], sign, -(N2 / 2), N2, t2r, t2i);
    else { t2r[0] = 1.0; t2i[0] = 0.0; }
    double sr = 0.0, si = 0.0;
    for (i64 i2 = 0; i2 < N2; ++i2) {
      double ar = 0.0, ai = 0.0;
      for (i64 i1 = 0; i1 < N1; ++i1) {
        const double* row = f + 2 * (i2 * N1 + i1) * N0;
        double br = 0.0, bi = 0.0;
        for (i64 i0 = 0; i0 < N0; ++i0) {
          const double fr = row[2 * i0], fi = row[2 * i0 + 1];
          br += fr * t0r[i0] - fi * t0i[i0];
          bi += fr * t0i[i0] + fi * t0r[i0];
        }
        ar += br * t1r[i1] - bi * t1i[i1];
        ai += br * t1i[i1] + bi * t1r[i1];
      }
      sr += ar * t2r[i2] - ai * t2i[i2];
      si += ar * t2i[i2] + ai * t2r[i2];
    }
    out[2 * s] = sr;
    out[2 * s + 1] = si;
    free(t0r); free(t0i); free(t1r); free(t1i); free(t2r); free(t2i);
  }
}

/* Type 3 at S targets: out[s] = sum_j c_j exp(i sign s_s . x_j). */
void orc_direct_t3(int dim, int sign, i64 m, const double* x0, const double* x1, const double* x2,
                   const double* c, i64 S, const double* s0, const double* s1, const double* s2,
                   double* out) {
#pragma omp parallel for schedule(dynamic, 1)
  for (i64 s = 0; s < S; ++s) {
    double sr = 0.0, si = 0.0;
    for (i64 j = 0; j < m; ++j) {
      double ph = x0[j] * s0[s];
      if (dim >= 2) ph += x1[j] * s1[s];
      if (dim >= 3) ph += x2[j] * s2[s];
      ph *= sign;
      const double cs = cos(ph), sn = sin(ph);
      sr += c[2 * j] * cs - c[2 * j + 1] * sn;
      si += c[2 * j] * sn + c[2 * j + 1] * cs;
    }
    out[2 * s] = sr;
    out[2 * s + 1] = si;
  }
}

int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
