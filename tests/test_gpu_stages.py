"""GPU tests of the spreading and interpolation stages alone, through the C ABI
(stage entry points nnf_plan_spread / _spread_tiles / _reduce_tiles /
_interp / _fine_grid), against the FP64 restatement of the reference's
spread and interpolation (oracle/, following
src/neonufft/gpu/kernels/spreading_kernel.cu:477-660 and
interpolation_kernel.cu:27-319).

Geometries cover both forms of the tile reduction: regular bins (every cell
covered by its own bin's tile and at most one neighbour per dimension) and
small grids where one tile wraps onto itself (general cover tables). 2D plans
are also run with row-band spreading (one workgroup per band of whole grid
rows) for several band heights, down to one row and up to the largest band a
grid allows (n1 - ns + 1 rows, whose point range wraps below row 0).
"""
import ctypes

import numpy as np
import pytest

torch = pytest.importorskip("torch")

import neonufft_amd as nnf
from oracle import oracle as O

from test_gpu_plan import CDT, RDT, cplx, dev, points

pytestmark = pytest.mark.gpu


def _hip():
    for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
        try:
            L = ctypes.CDLL(name)
            L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
            L.hipMemcpy.restype = ctypes.c_int
            return L
        except OSError:
            continue
    pytest.skip("HIP runtime library not found")


def d2h(ptr, count, dtype):
    torch.cuda.synchronize()
    a = np.empty(count, dtype=dtype)
    assert _hip().hipMemcpy(a.ctypes.data, ptr, a.nbytes, 2) == 0
    return a


def h2d(ptr, a):
    torch.cuda.synchronize()
    a = np.ascontiguousarray(a)
    assert _hip().hipMemcpy(ptr, a.ctypes.data, a.nbytes, 1) == 0


CASES = [
    # (dim, modes, points, bin-owned (regular) tile reduction expected for
    # f32 / f64; fp64 tiles are twice as large, so the LDS budget shrinks 3D
    # bins below ns and the general reduction is used)
    (3, (32, 32, 32), 3000, (True, False)),
    (2, (128, 100), 5000, (True, True)),
    (1, (700,), 2000, (True, True)),
    (3, (10, 12, 9), 500, (False, False)),
    (2, (23, 2), 100, (False, False)),
    (1, (100,), 300, (False, False)),
]

# fp32: Horner-fitted kernel (fit error ~1e-6 relative at ns = 6) and fp32
# arithmetic; fp64: exact kernel, 62-bit fixed-point tile accumulation
TOL = {"f32": 2e-5, "f64": 1e-12}
NPC = {"f32": np.complex64, "f64": np.complex128}


def _regular(info):
    """Whether the plan's bins take the bin-owned tile reduction: per
    dimension at least two bins, bin >= ns, last bin full or >= ns wide."""
    ns = info["n_spread"]
    for n, b in zip(info["grid"], info["bin_size"]):
        if -(-n // b) < 2 or b < ns or (n % b and n % b < ns):
            return False
    return True


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _plan(prec, dim, modes, m, seed, **tuning):
    rng = np.random.default_rng(seed)
    x = points(rng, dim, m, prec)
    opt = nnf.Options(tol=1e-5 if prec == "f32" else 1e-10)
    plan = nnf.Plan(opt, -1, m, [dev(v, RDT[prec]) for v in x], modes, dtype=RDT[prec],
                    **tuning)
    info = plan.info()
    u = [O.fold(v.astype(np.float64)) for v in x]
    return rng, plan, info, u


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("case", CASES, ids=[f"d{c[0]}_{'x'.join(map(str, c[1]))}" for c in CASES])
def test_spread_and_accumulate(case, prec):
    dim, modes, m, regular = case
    rng, plan, info, u = _plan(prec, dim, modes, m, 7)
    assert _regular(info) == regular[0 if prec == "f32" else 1]
    n, ns, beta = info["grid"], info["n_spread"], info["es_beta"]
    cells = int(np.prod(n))
    c1 = cplx(rng, m, prec)
    c2 = cplx(rng, m, prec)
    g1 = c1.astype(np.complex128)
    g2 = c2.astype(np.complex128)

    plan.spread(dev(c1, CDT[prec]))
    got1 = d2h(plan.fine_grid_ptr(), cells, NPC[prec])
    ref1 = O.spread(n, ns, beta, u, g1).ravel()
    assert _rel(got1, ref1) < TOL[prec]

    # accumulating spread adds: grid = spread(c1) + spread(c2)
    plan.spread(dev(c2, CDT[prec]), accumulate=True)
    got12 = d2h(plan.fine_grid_ptr(), cells, got1.dtype)
    ref12 = O.spread(n, ns, beta, u, g1 + g2).ravel()
    assert _rel(got12, ref12) < TOL[prec]

    # the two kernels separately give the same grid as spread()
    plan.spread_tiles(dev(c2, CDT[prec]))
    plan.reduce_tiles(accumulate=False)
    got2 = d2h(plan.fine_grid_ptr(), cells, got1.dtype)
    plan.spread(dev(c2, CDT[prec]))
    got2b = d2h(plan.fine_grid_ptr(), cells, got1.dtype)
    assert np.array_equal(got2, got2b)


# 2D grids for the row-band spreading: (modes, points, band rows; 0 = default,
# -1 = the largest band the grid allows)
BAND_CASES = [
    ((128, 100), 5000, 0),
    ((128, 100), 5000, 1),
    ((128, 100), 5000, 3),
    ((128, 100), 5000, -1),
    ((33, 47), 2000, 5),
    ((23, 2), 100, 0),
    ((23, 2), 100, -1),
    ((256, 64), 0, 0),
    ((4, 300), 3000, 7),
]


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("case", BAND_CASES,
                         ids=[f"{'x'.join(map(str, c[0]))}_m{c[1]}_r{c[2]}" for c in BAND_CASES])
def test_band_spread(case, prec):
    modes, m, rows = case
    ns = 6 if prec == "f32" else 11
    n0 = max(2 * ns, 2 * modes[0])
    n1 = max(2 * ns, 2 * modes[1])
    fused = n0 >= 32 and (n0 & (n0 - 1)) == 0
    if rows == -1:
        rows = n1 - ns + 1
    # requested rows, capped by the 68 KiB LDS band (two int32 / int64
    # accumulator planes per row of n0 + 24 words: halo cells and a bank
    # shift) and rounded down to a power of two with the fused x-FFT
    pitch = n0 + 24
    expect = min(rows, n1 - ns + 1, (68 * 1024) // (2 * pitch * (4 if prec == "f32" else 8)))
    if fused and expect > 0:
        expect = 1 << (expect.bit_length() - 1)
    rng, plan, info, u = _plan(prec, 2, modes, m, 5, spread_mode="bands", band_rows=rows)
    assert info["spread_mode"] == ("bands" if fused else "bands_grid")
    if rows:
        assert info["band_rows"] == expect
    n, ns, beta = info["grid"], info["n_spread"], info["es_beta"]
    cells = int(np.prod(n))
    c1 = cplx(rng, m, prec)
    c2 = cplx(rng, m, prec)
    plan.spread(dev(c1, CDT[prec]))
    got1 = d2h(plan.fine_grid_ptr(), cells, NPC[prec])
    if m == 0:
        assert not np.any(got1)
        return
    ref1 = O.spread(n, ns, beta, u, c1.astype(np.complex128)).ravel()
    assert _rel(got1, ref1) < TOL[prec]
    plan.spread(dev(c2, CDT[prec]), accumulate=True)
    got12 = d2h(plan.fine_grid_ptr(), cells, NPC[prec])
    ref12 = O.spread(n, ns, beta, u, (c1 + c2).astype(np.complex128)).ravel()
    assert _rel(got12, ref12) < TOL[prec]
    # the same plan with subgrid tiles spreads the same grid
    _, tplan, tinfo, _ = _plan(prec, 2, modes, m, 5, spread_mode="tiles")
    assert tinfo["spread_mode"] == "tiles"
    tplan.spread(dev(c1, CDT[prec]))
    got_t = d2h(tplan.fine_grid_ptr(), cells, NPC[prec])
    assert _rel(got1, got_t) < TOL[prec]


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("case", CASES, ids=[f"d{c[0]}_{'x'.join(map(str, c[1]))}" for c in CASES])
def test_interp(case, prec):
    dim, modes, m, _ = case
    rng, plan, info, u = _plan(prec, dim, modes, m, 11)
    n, ns, beta = info["grid"], info["n_spread"], info["es_beta"]
    cells = int(np.prod(n))
    grid = (rng.standard_normal(cells) + 1j * rng.standard_normal(cells)).astype(NPC[prec])
    h2d(plan.fine_grid_ptr(), grid)
    out = torch.empty(m, dtype=CDT[prec], device="cuda")
    plan.interp(out)
    torch.cuda.synchronize()
    ref = O.interp(n, ns, beta, u, grid.astype(np.complex128).reshape(tuple(reversed(n))))
    assert _rel(out.cpu().numpy(), ref) < TOL[prec]
