"""nnf_plan_set_points_noalloc: set_points with no device allocation.

The entry point has no counterpart in the reference; it is a variant of
nnf_plan_set_points (reference gpu::Plan::set_points,
include/neonufft/gpu/plan.hpp:42) for callers that cannot allocate. Its
temporaries live in a fixed per-device array inside the library, its results
in the memory earlier (allocating) set_points calls left in the plan. These
tests call it through the C ABI and compare every transform after it with
the same transform after nnf_plan_set_points on the same points (same
kernels; only the summation order inside a subproblem may differ), on two
calls in a row with the second one smaller, and with the FP64 direct NUDFT.
"""
import ctypes

import numpy as np
import pytest

torch = pytest.importorskip("torch")

import neonufft_amd as nnf
from oracle import oracle as O

from test_gpu_plan import CDT, RDT, cplx, dev, points

pytestmark = pytest.mark.gpu

CASES = [
    # (dim, modes, capacity M, first M, second M, prec, tol)
    (3, (24, 20, 16), 30000, 30000, 11000, "f32", 1e-5),
    (2, (64, 48), 40000, 25000, 9000, "f32", 1e-5),    # row bands with the fused x-FFT
    (2, (50, 37), 20000, 20000, 7000, "f64", 1e-9),    # row bands writing the grid
    (1, (300,), 8000, 6000, 500, "f64", 1e-10),
    (3, (16, 16, 16), 12000, 12000, 4000, "f64", 1e-8),
]


def _cid(c):
    return f"d{c[0]}_{'x'.join(map(str, c[1]))}_{c[5]}"


def _vpp(ts):
    return (ctypes.c_void_p * len(ts))(*[t.data_ptr() for t in ts])


def _set_noalloc(L, plan, m, loc):
    f = L.nnf_plan_set_points_noalloc
    f.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.POINTER(ctypes.c_void_p)]
    f.restype = ctypes.c_int
    return f(plan._h, m, _vpp(loc))


def _type1(plan, c, prec):
    out = torch.zeros(int(np.prod(plan.modes)), dtype=CDT[prec], device="cuda")
    plan.transform_type_1(c, out)
    return out


def _type2(plan, f, m, prec):
    out = torch.zeros(m, dtype=CDT[prec], device="cuda")
    plan.transform_type_2(f, None, out)
    return out


@pytest.mark.parametrize("case", CASES, ids=[_cid(c) for c in CASES])
def test_set_points_noalloc_matches_set_points(case):
    dim, modes, cap, m1, m2, prec, tol = case
    L = nnf._L.lib()
    rng = np.random.default_rng(90 + dim)
    opt = nnf.Options(tol=tol)
    nm = int(np.prod(modes))
    x0 = points(rng, dim, cap, prec)
    # the plan under test gets its memory from one allocating call
    plan = nnf.Plan(opt, -1, cap, [dev(v, RDT[prec]) for v in x0], modes)
    fin = dev(cplx(rng, nm, prec), CDT[prec])
    for m in (m1, m2):
        x = points(rng, dim, m, prec)
        loc = [dev(v, RDT[prec]) for v in x]
        c = dev(cplx(rng, m, prec), CDT[prec])
        torch.cuda.synchronize()
        print(f"noalloc set_points dim={dim} modes={modes} m={m} {prec}", flush=True)
        rc = _set_noalloc(L, plan, m, loc)
        assert rc == 0, L.nnf_last_error().decode()
        ref_plan = nnf.Plan(opt, -1, m, loc, modes)
        assert plan.info()["spread_mode"] == ref_plan.info()["spread_mode"]
        assert plan.info()["num_points"] == m
        f_got = _type1(plan, c, prec)
        f_ref = _type1(ref_plan, c, prec)
        b_got = _type2(plan, fin, m, prec)
        b_ref = _type2(ref_plan, fin, m, prec)
        torch.cuda.synchronize()
        lim = 1e-6 if prec == "f32" else 1e-13
        for got, ref in ((f_got, f_ref), (b_got, b_ref)):
            g = got.cpu().numpy().astype(np.complex128)
            r = ref.cpu().numpy().astype(np.complex128)
            assert O.rel_l2(g, r) < lim, O.rel_l2(g, r)
        # and the result itself against the FP64 direct sum
        xs = [v.astype(np.float64) for v in x]
        idx = np.arange(0, nm, max(1, nm // 400))
        ref1 = O.direct_type1(xs, c.cpu().numpy().astype(np.complex128),
                              O.mode_vectors(modes, idx), -1)
        alg1 = O.nufft1(xs, c.cpu().numpy().astype(np.complex128), modes, -1, tol)
        e = O.rel_l2(f_got.cpu().numpy().astype(np.complex128)[idx], ref1)
        assert e < 1.5 * O.rel_l2(alg1[idx], ref1) + (3e-5 if prec == "f32" else 1e-12), e


def test_set_points_noalloc_rejects_beyond_capacity():
    """More points than the plan holds memory for: a clear InputError before
    any work; the plan then refuses transforms until a set_points succeeds."""
    L = nnf._L.lib()
    rng = np.random.default_rng(95)
    modes, prec = (20, 18, 12), "f32"
    x = points(rng, 3, 5000, prec)
    plan = nnf.Plan(nnf.Options(tol=1e-5), 1, 3000, [dev(v[:3000], RDT[prec]) for v in x], modes)
    loc = [dev(v, RDT[prec]) for v in x]
    torch.cuda.synchronize()
    rc = _set_noalloc(L, plan, 5000, loc)
    assert rc == 5, rc  # NEONUFFT_INPUT_ERROR
    assert b"without allocation" in L.nnf_last_error()
    c = dev(cplx(rng, 5000, prec), CDT[prec])
    with pytest.raises(nnf.InputError):
        _type1(plan, c, prec)
    # an allocating set_points makes the plan usable (and larger) again
    plan.set_points(5000, loc)
    assert _set_noalloc(L, plan, 5000, loc) == 0, L.nnf_last_error().decode()
    got = _type1(plan, c, prec)
    ref = _type1(nnf.Plan(nnf.Options(tol=1e-5), 1, 5000, loc, modes), c, prec)
    torch.cuda.synchronize()
    assert O.rel_l2(got.cpu().numpy().astype(np.complex128),
                    ref.cpu().numpy().astype(np.complex128)) < 1e-6


def test_set_points_noalloc_rejects_above_scratch_maximum():
    """A point count whose temporaries exceed the fixed scratch array (64 MiB,
    about 16 million points) is refused with InputError, before any work."""
    L = nnf._L.lib()
    rng = np.random.default_rng(96)
    m = 17_000_000
    x = [torch.from_numpy(rng.uniform(-np.pi, np.pi, m).astype(np.float32)).cuda()
         for _ in range(1)]
    plan = nnf.Plan(nnf.Options(tol=1e-4), 1, m, x, (64,))
    torch.cuda.synchronize()
    rc = _set_noalloc(L, plan, m, x)
    assert rc == 5, rc
    assert b"scratch" in L.nnf_last_error()
    # the refused call left the plan's points as they were
    c = torch.zeros(m, dtype=torch.complex64, device="cuda")
    out = _type1(plan, c, "f32")
    torch.cuda.synchronize()
    assert float(out.abs().max()) == 0.0
