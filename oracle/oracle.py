"""FP64 CPU restatement of the NeoNUFFT plan pipeline, used as the checker.

TEST INFRASTRUCTURE ONLY: tests/, ``__graft_entry__.smoke()`` and the
cpu_baseline leg of bench.py may import this module; the product package
``neonufft_amd`` never does.

The heavy loops (spreading, interpolation, direct sums) live in
``nufft_oracle.c`` (built to ``liboracle.so`` by ``oracle/Makefile``); the
uniform FFT step uses ``numpy.fft`` in FP64. Each function names the reference
file:line it restates (paths relative to the reference repository root).
"""
from __future__ import annotations

import ctypes
import math
import os
import subprocess
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_i64 = ctypes.c_int64
_dp = ctypes.POINTER(ctypes.c_double)
_lp = ctypes.POINTER(ctypes.c_int64)


def build() -> str:
    """Compile liboracle.so with gcc (oracle/Makefile)."""
    subprocess.run(["make", "-s", "-C", _HERE, "liboracle.so"], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_kernel_params.argtypes = [ctypes.c_double, ctypes.c_double,
                                        ctypes.POINTER(ctypes.c_int), _dp]
        L.orc_es_kernel.argtypes = [ctypes.c_double, ctypes.c_int, ctypes.c_double]
        L.orc_es_kernel.restype = ctypes.c_double
        L.orc_gauss_legendre.argtypes = [ctypes.c_int, _dp, _dp]
        L.orc_phihat.argtypes = [ctypes.c_double, ctypes.c_int, ctypes.c_double, ctypes.c_int]
        L.orc_phihat.restype = ctypes.c_double
        L.orc_fseries.argtypes = [_i64, ctypes.c_int, ctypes.c_double, ctypes.c_int, _dp]
        L.orc_fold.argtypes = [ctypes.c_double]
        L.orc_fold.restype = ctypes.c_double
        L.orc_fold_array.argtypes = [_i64, _dp, ctypes.c_double, ctypes.c_double, _dp]
        L.orc_spread.argtypes = [ctypes.c_int, _lp, ctypes.c_int, ctypes.c_double, _i64,
                                 _dp, _dp, _dp, _dp, _dp, _dp]
        L.orc_interp.argtypes = [ctypes.c_int, _lp, ctypes.c_int, ctypes.c_double, _i64,
                                 _dp, _dp, _dp, _dp, _dp, _dp]
        L.orc_downsample.argtypes = [ctypes.c_int, _lp, _lp, _dp, _dp, _dp, ctypes.c_int,
                                     _dp, _lp, _dp]
        L.orc_upsample.argtypes = [ctypes.c_int, _lp, _lp, _dp, _dp, _dp, ctypes.c_int,
                                   _dp, _lp, _dp]
        L.orc_next235even.argtypes = [_i64]
        L.orc_next235even.restype = _i64
        L.orc_t3_grid_size.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                       ctypes.c_int]
        L.orc_t3_grid_size.restype = _i64
        L.orc_set_nhg.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double, _i64,
                                  _dp, _dp]
        L.orc_prephase.argtypes = [ctypes.c_int, ctypes.c_int, _i64, _dp, _dp, _dp, _dp, _dp]
        L.orc_postphase.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                    ctypes.c_int, _i64, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
        L.orc_direct_t1.argtypes = [ctypes.c_int, ctypes.c_int, _i64, _dp, _dp, _dp, _dp,
                                    _i64, _lp, _dp]
        L.orc_direct_t2.argtypes = [ctypes.c_int, ctypes.c_int, _lp, _dp, _i64, _dp, _dp, _dp,
                                    _dp]
        L.orc_direct_t3.argtypes = [ctypes.c_int, ctypes.c_int, _i64, _dp, _dp, _dp, _dp,
                                    _i64, _dp, _dp, _dp, _dp]
        L.orc_num_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def _d(a: Optional[np.ndarray]):
    if a is None:
        return ctypes.cast(None, _dp)
    assert a.dtype == np.float64 and a.flags.c_contiguous, (a.dtype, a.flags)
    return a.ctypes.data_as(_dp)


def _l(a: np.ndarray):
    assert a.dtype == np.int64 and a.flags.c_contiguous
    return a.ctypes.data_as(_lp)


def _cplx_view(a: np.ndarray) -> np.ndarray:
    """complex128 array -> float64 view of interleaved (re, im)."""
    a = np.ascontiguousarray(a, dtype=np.complex128)
    return a.view(np.float64)


def _coords(x: Sequence[np.ndarray], dim: int) -> List[np.ndarray]:
    out = [np.ascontiguousarray(np.asarray(x[d], dtype=np.float64)) for d in range(dim)]
    while len(out) < 3:
        out.append(None)
    return out


# --------------------------------------------------------------------------
# Kernel parameters and the ES kernel
# --------------------------------------------------------------------------

def kernel_params(tol: float, upsampfac: float = 2.0):
    """(n_spread, es_beta): src/neonufft/es_kernel_param.hpp:24-56."""
    ns = ctypes.c_int()
    beta = ctypes.c_double()
    lib().orc_kernel_params(tol, upsampfac, ctypes.byref(ns), ctypes.byref(beta))
    return ns.value, beta.value


def es_kernel(t, ns: int, beta: float):
    """phi(t) = exp(beta (sqrt(1-(2t/ns)^2) - 1)); es_kernel_eval.cuh:36-46."""
    f = lib().orc_es_kernel
    return np.vectorize(lambda v: f(float(v), ns, beta))(np.asarray(t, dtype=np.float64))


def gauss_legendre(n: int):
    x = np.zeros(n)
    w = np.zeros(n)
    lib().orc_gauss_legendre(n, _d(x), _d(w))
    return x, w


def default_nquad(ns: int) -> int:
    """Quadrature nodes on (0,1) for phihat: ns + ns%2 + 10, the count of the
    reference CPU path (src/neonufft/kernels/fseries_kernel.cpp:35)."""
    return ns + ns % 2 + 10


def phihat(freq: float, ns: int, beta: float, nq: Optional[int] = None) -> float:
    return lib().orc_phihat(freq, ns, beta, nq or default_nquad(ns))


def fseries(n: int, ns: int, beta: float, nq: Optional[int] = None) -> np.ndarray:
    """fs[k] = (-1)^k / phihat(2 pi k / n), k=0..n/2; fseries_kernel.cu:24-67."""
    out = np.zeros(n // 2 + 1)
    lib().orc_fseries(n, ns, beta, nq or default_nquad(ns), _d(out))
    return out


def fold(x, offset: float = 0.0, scale: float = 1.0) -> np.ndarray:
    """u = frac(((x-offset)*scale + pi) / 2pi); rescale_loc_kernel.cu:104-118,172-178."""
    x = np.ascontiguousarray(np.asarray(x, dtype=np.float64))
    u = np.empty_like(x)
    lib().orc_fold_array(x.size, _d(x), offset, scale, _d(u))
    return u


# --------------------------------------------------------------------------
# Pipeline stages
# --------------------------------------------------------------------------

def fine_grid_size(modes: Sequence[int], ns: int, upsampfac: float) -> List[int]:
    """n_d = max(2 ns, (size_t)(N_d * sigma)); gpu/plan_impl.hpp:56-60."""
    return [max(2 * ns, int(N * upsampfac)) for N in modes]


def spread(n: Sequence[int], ns: int, beta: float, u: Sequence[np.ndarray], c: np.ndarray,
           pre: Optional[np.ndarray] = None, grid: Optional[np.ndarray] = None) -> np.ndarray:
    """Fine grid (complex128, numpy shape reversed: x fastest) after spreading."""
    dim = len(n)
    nn = np.array(list(n) + [1] * (3 - dim), dtype=np.int64)
    if grid is None:
        grid = np.zeros(tuple(reversed(list(n))), dtype=np.complex128)
    uu = _coords(u, dim)
    cv = _cplx_view(c)
    pv = _cplx_view(pre) if pre is not None else None
    gv = grid.view(np.float64)
    lib().orc_spread(dim, _l(nn), ns, beta, len(uu[0]), _d(uu[0]), _d(uu[1]), _d(uu[2]),
                     _d(cv), _d(pv), _d(gv))
    return grid


def interp(n: Sequence[int], ns: int, beta: float, u: Sequence[np.ndarray], grid: np.ndarray,
           post: Optional[np.ndarray] = None) -> np.ndarray:
    dim = len(n)
    nn = np.array(list(n) + [1] * (3 - dim), dtype=np.int64)
    uu = _coords(u, dim)
    m = len(uu[0])
    out = np.zeros(m, dtype=np.complex128)
    gv = np.ascontiguousarray(grid, dtype=np.complex128).view(np.float64)
    pv = _cplx_view(post) if post is not None else None
    lib().orc_interp(dim, _l(nn), ns, beta, m, _d(uu[0]), _d(uu[1]), _d(uu[2]), _d(gv),
                     _d(pv), _d(out.view(np.float64)))
    return out


def fft_grid(grid: np.ndarray, sign: int) -> np.ndarray:
    """Unnormalised C2C over all axes with exp(sign 2 pi i m g / n);
    gpu/util/fft_grid.cpp:59-70 (sign<0 forward, else backward)."""
    if sign < 0:
        return np.fft.fftn(grid)
    return np.fft.ifftn(grid) * grid.size


def contiguous_strides(modes: Sequence[int]) -> List[int]:
    s = [1]
    for N in modes[:-1]:
        s.append(s[-1] * N)
    return s


def downsample(n, modes, fs, order_fft: bool, grid: np.ndarray,
               out_strides: Optional[Sequence[int]] = None, out: Optional[np.ndarray] = None):
    """Type-1 deconvolve + truncate + reorder; downsample_kernel.cu:133-191."""
    dim = len(n)
    nn = np.array(list(n) + [1] * (3 - dim), dtype=np.int64)
    NN = np.array(list(modes) + [1] * (3 - dim), dtype=np.int64)
    st = np.array(list(out_strides or contiguous_strides(modes)) + [0] * (3 - dim), dtype=np.int64)
    if out is None:
        size = 1 + sum((modes[d] - 1) * st[d] for d in range(dim))
        out = np.zeros(size, dtype=np.complex128)
    fsl = [np.ascontiguousarray(f, dtype=np.float64) for f in fs] + [None] * (3 - dim)
    gv = np.ascontiguousarray(grid, dtype=np.complex128).view(np.float64)
    lib().orc_downsample(dim, _l(nn), _l(NN), _d(fsl[0]), _d(fsl[1]), _d(fsl[2]),
                         int(order_fft), _d(gv), _l(st), _d(out.view(np.float64)))
    return out


def upsample(n, modes, fs, order_fft: bool, f: np.ndarray,
             in_strides: Optional[Sequence[int]] = None) -> np.ndarray:
    """Type-2 pre-correction into a zeroed fine grid; upsample_kernel.cu:18-154."""
    dim = len(n)
    nn = np.array(list(n) + [1] * (3 - dim), dtype=np.int64)
    NN = np.array(list(modes) + [1] * (3 - dim), dtype=np.int64)
    st = np.array(list(in_strides or contiguous_strides(modes)) + [0] * (3 - dim), dtype=np.int64)
    grid = np.zeros(tuple(reversed(list(n))), dtype=np.complex128)
    fsl = [np.ascontiguousarray(v, dtype=np.float64) for v in fs] + [None] * (3 - dim)
    fv = _cplx_view(f)
    lib().orc_upsample(dim, _l(nn), _l(NN), _d(fsl[0]), _d(fsl[1]), _d(fsl[2]), int(order_fft),
                       _d(fv), _l(st), _d(grid.view(np.float64)))
    return grid


# --------------------------------------------------------------------------
# Full type 1 / type 2 pipelines (gpu/plan_impl.hpp:91-136)
# --------------------------------------------------------------------------

def nufft1(x: Sequence[np.ndarray], c: np.ndarray, modes: Sequence[int], sign: int, tol: float,
           upsampfac: float = 2.0, order_fft: bool = False, nq: Optional[int] = None):
    dim = len(modes)
    ns, beta = kernel_params(tol, upsampfac)
    n = fine_grid_size(modes, ns, upsampfac)
    u = [fold(x[d]) for d in range(dim)]
    grid = spread(n, ns, beta, u, c)
    F = fft_grid(grid, sign)
    fs = [fseries(n[d], ns, beta, nq) for d in range(dim)]
    return downsample(n, modes, fs, order_fft, F)


def nufft2(x: Sequence[np.ndarray], f: np.ndarray, modes: Sequence[int], sign: int, tol: float,
           upsampfac: float = 2.0, order_fft: bool = False, nq: Optional[int] = None):
    dim = len(modes)
    ns, beta = kernel_params(tol, upsampfac)
    n = fine_grid_size(modes, ns, upsampfac)
    fs = [fseries(n[d], ns, beta, nq) for d in range(dim)]
    grid = upsample(n, modes, fs, order_fft, f)
    F = fft_grid(grid, sign)
    u = [fold(x[d]) for d in range(dim)]
    return interp(n, ns, beta, u, F)


# --------------------------------------------------------------------------
# Type 3 (gpu/plan_t3_impl.hpp, util/grid_info_t3.hpp)
# --------------------------------------------------------------------------

def next235even(n: int) -> int:
    return int(lib().orc_next235even(int(n)))


def t3_grid_size(S: float, X: float, upsampfac: float, ns: int) -> int:
    return int(lib().orc_t3_grid_size(S, X, upsampfac, ns))


def set_nhg(S: float, X: float, upsampfac: float, n: int):
    h = ctypes.c_double()
    g = ctypes.c_double()
    lib().orc_set_nhg(S, X, upsampfac, n, ctypes.byref(h), ctypes.byref(g))
    return h.value, g.value


@dataclass
class GridInfoT3:
    """Restates GridInfoT3 (src/neonufft/util/grid_info_t3.hpp:18-52) in FP64."""
    dim: int
    ns: int
    upsampfac: float
    input_offsets: List[float] = field(default_factory=list)
    output_offsets: List[float] = field(default_factory=list)
    input_half_extent: List[float] = field(default_factory=list)
    output_half_extent: List[float] = field(default_factory=list)
    spread_grid_size: List[int] = field(default_factory=list)
    fft_grid_size: List[int] = field(default_factory=list)

    @staticmethod
    def make(ns, upsampfac, recenter_threshold, in_min, in_max, out_min, out_max):
        dim = len(in_min)
        gi = GridInfoT3(dim, ns, upsampfac)
        for d in range(dim):
            xo = (in_max[d] + in_min[d]) / 2
            so = (out_max[d] + out_min[d]) / 2
            X = max(abs(in_max[d] - xo), abs(in_min[d] - xo))
            S = max(abs(out_max[d] - so), abs(out_min[d] - so))
            if abs(xo) < recenter_threshold * X:
                xo = 0.0
                X = max(abs(in_max[d]), abs(in_min[d]))
            if abs(so) < recenter_threshold * S:
                so = 0.0
                S = max(abs(out_max[d]), abs(out_min[d]))
            n = t3_grid_size(S, X, upsampfac, ns)
            gi.input_offsets.append(xo)
            gi.output_offsets.append(so)
            gi.input_half_extent.append(X)
            gi.output_half_extent.append(S)
            gi.spread_grid_size.append(n)
            gi.fft_grid_size.append(next235even(max(2 * ns, int(n * upsampfac))))
        return gi


def prephase(x: Sequence[np.ndarray], sbar: Sequence[float], sign: int) -> np.ndarray:
    dim = len(sbar)
    xx = _coords(x, dim)
    sb = np.array(list(sbar) + [0.0] * (3 - dim))
    out = np.zeros(len(xx[0]), dtype=np.complex128)
    lib().orc_prephase(dim, sign, len(xx[0]), _d(xx[0]), _d(xx[1]), _d(xx[2]), _d(sb),
                       _d(out.view(np.float64)))
    return out


def postphase(s: Sequence[np.ndarray], xbar, sbar, scale, sign: int, ns: int, beta: float,
              nq: Optional[int] = None) -> np.ndarray:
    dim = len(sbar)
    ss = _coords(s, dim)
    xb = np.array(list(xbar) + [0.0] * (3 - dim))
    sb = np.array(list(sbar) + [0.0] * (3 - dim))
    sc = np.array(list(scale) + [0.0] * (3 - dim))
    out = np.zeros(len(ss[0]), dtype=np.complex128)
    lib().orc_postphase(dim, sign, ns, beta, nq or default_nquad(ns), len(ss[0]), _d(ss[0]),
                        _d(ss[1]), _d(ss[2]), _d(xb), _d(sb), _d(sc), _d(out.view(np.float64)))
    return out


def nufft3(x: Sequence[np.ndarray], c: np.ndarray, s: Sequence[np.ndarray], sign: int,
           tol: float, upsampfac: float = 2.0, recenter_threshold: float = 0.1,
           nq: Optional[int] = None, bounds=None):
    """Type 3 via prephase x spread x upsample/FFT/interp x postphase
    (gpu/plan_t3_impl.hpp:74-252). bounds = (input_min, input_max,
    output_min, output_max) sizes the grids as the bounds constructor does
    (gpu/plan_t3_impl.hpp:121-138); None takes the points' own extents (the
    points constructor, gpu/plan_t3_impl.hpp:74-119)."""
    dim = len(x)
    ns, beta = kernel_params(tol, upsampfac)
    xs = [np.asarray(x[d], dtype=np.float64) for d in range(dim)]
    ss = [np.asarray(s[d], dtype=np.float64) for d in range(dim)]
    if bounds is None:
        bounds = ([v.min() for v in xs], [v.max() for v in xs],
                  [v.min() for v in ss], [v.max() for v in ss])
    gi = GridInfoT3.make(ns, upsampfac, recenter_threshold, *[list(b) for b in bounds])
    h = []
    gam = []
    for d in range(dim):
        hd, gd = set_nhg(gi.output_half_extent[d], gi.input_half_extent[d], upsampfac,
                         gi.spread_grid_size[d])
        h.append(hd)
        gam.append(gd)
    pre = None
    if any(o != 0 for o in gi.output_offsets):
        pre = prephase(xs, gi.output_offsets, sign)
    u_in = [fold(xs[d], gi.input_offsets[d], 1.0 / gam[d]) for d in range(dim)]
    grid = spread(gi.spread_grid_size, ns, beta, u_in, c, pre=pre)
    nf = gi.fft_grid_size
    fs = [fseries(nf[d], ns, beta, nq) for d in range(dim)]
    big = upsample(nf, gi.spread_grid_size, fs, False, grid.reshape(-1))
    F = fft_grid(big, sign)
    scale = [gam[d] * h[d] for d in range(dim)]
    u_out = [fold(ss[d], gi.output_offsets[d], scale[d]) for d in range(dim)]
    post = postphase(ss, gi.input_offsets, gi.output_offsets, scale, sign, ns, beta, nq)
    return interp(nf, ns, beta, u_out, F, post=post)


# --------------------------------------------------------------------------
# Direct sums (ground truth) and error metrics
# --------------------------------------------------------------------------

def mode_vectors(modes: Sequence[int], linear_idx: np.ndarray, order_fft: bool = False):
    """Integer mode vectors (S x 3) for linear indices into a contiguous mode array."""
    dim = len(modes)
    kk = np.zeros((len(linear_idx), 3), dtype=np.int64)
    rem = np.asarray(linear_idx, dtype=np.int64).copy()
    for d in range(dim):
        i = rem % modes[d]
        rem //= modes[d]
        N = modes[d]
        if order_fft:
            kk[:, d] = np.where(i < N - N // 2, i, i - N)
        else:
            kk[:, d] = i - N // 2
    return kk


def direct_type1(x: Sequence[np.ndarray], c: np.ndarray, kk: np.ndarray, sign: int) -> np.ndarray:
    dim = len(x)
    xx = _coords(x, dim)
    kk = np.ascontiguousarray(kk, dtype=np.int64)
    assert kk.shape[1] == 3
    out = np.zeros(len(kk), dtype=np.complex128)
    lib().orc_direct_t1(dim, sign, len(xx[0]), _d(xx[0]), _d(xx[1]), _d(xx[2]),
                        _d(_cplx_view(c)), len(kk), _l(kk), _d(out.view(np.float64)))
    return out


def direct_type2(x: Sequence[np.ndarray], f: np.ndarray, modes: Sequence[int],
                 sign: int) -> np.ndarray:
    """f contiguous CMCL order."""
    dim = len(modes)
    xx = _coords(x, dim)
    NN = np.array(list(modes) + [1] * (3 - dim), dtype=np.int64)
    out = np.zeros(len(xx[0]), dtype=np.complex128)
    lib().orc_direct_t2(dim, sign, _l(NN), _d(_cplx_view(f)), len(xx[0]), _d(xx[0]),
                        _d(xx[1]), _d(xx[2]), _d(out.view(np.float64)))
    return out


def direct_type3(x: Sequence[np.ndarray], c: np.ndarray, s: Sequence[np.ndarray],
                 sign: int) -> np.ndarray:
    dim = len(x)
    xx = _coords(x, dim)
    ss = _coords(s, dim)
    out = np.zeros(len(ss[0]), dtype=np.complex128)
    lib().orc_direct_t3(dim, sign, len(xx[0]), _d(xx[0]), _d(xx[1]), _d(xx[2]),
                        _d(_cplx_view(c)), len(ss[0]), _d(ss[0]), _d(ss[1]), _d(ss[2]),
                        _d(out.view(np.float64)))
    return out


def rel_l2(approx: np.ndarray, exact: np.ndarray) -> float:
    approx = np.asarray(approx, dtype=np.complex128)
    exact = np.asarray(exact, dtype=np.complex128)
    den = np.linalg.norm(exact)
    return float(np.linalg.norm(approx - exact) / (den if den > 0 else 1.0))


def num_threads() -> int:
    return int(lib().orc_num_threads())
