"""GPU parity of the type-3 plan's streaming and batched surface.

The reference's PlanT3 (include/neonufft/gpu/plan.hpp:48-82,
src/neonufft/gpu/plan_t3_impl.hpp:121-252) can be built from bounds alone,
takes its source and target points later through set_input_points /
set_output_points, accumulates every add_input into its spread grid until
reset() (set_input_points does not clear it), and runs `batch_size`
independent transforms whose inputs / outputs sit `bdist` elements apart.
Each step below is checked against the FP64 direct sum of what the plan has
accumulated, and against the FP64 restatement (oracle/) run on the grids the
bounds fix.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import neonufft_amd as nnf
from oracle import oracle as O

from test_gpu_plan import cplx

pytestmark = pytest.mark.gpu

PREC = {
    # plan dtype, coordinate dtype, complex dtype, tol, fp32 rounding floor
    "f32": ("float32", np.float32, np.complex64, 1e-5, 2e-5),
    "f64": ("float64", np.float64, np.complex128, 1e-9, 1e-12),
    "mixed": ("mixed", np.float64, np.complex128, 1e-6, 2e-5),
}


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _pts(rng, lo, hi, m, dt):
    return [rng.uniform(lo[d], hi[d], m).astype(dt) for d in range(len(lo))]


def _direct(x, c, s, sign):
    return O.direct_type3([v.astype(np.float64) for v in x], c.astype(np.complex128),
                          [v.astype(np.float64) for v in s], sign)


def _alg(x, c, s, sign, tol, bounds):
    return O.nufft3([v.astype(np.float64) for v in x], c.astype(np.complex128),
                    [v.astype(np.float64) for v in s], sign, tol, bounds=bounds)


def _check(got, ref, alg, tol, floor, prec):
    alg_err = O.rel_l2(alg, ref)
    assert alg_err < 20 * tol
    err = O.rel_l2(got, ref)
    assert err < 1.2 * alg_err + floor, (err, alg_err)
    assert O.rel_l2(got, alg) < (1e-9 if prec == "f64" else 5e-5 + 0.05 * tol)


DIMS = [1, 2, 3]


@pytest.mark.parametrize("prec", ["f32", "f64", "mixed"])
@pytest.mark.parametrize("dim", DIMS)
def test_t3_bounds_batch_accumulate_reset(dim, prec):
    dtype, rdt, cdt, tol, floor = PREC[prec]
    rng = np.random.default_rng(31 * dim + len(prec))
    sign = 1 if dim % 2 else -1
    # bounds enclose both source sets and both target sets; off-centre so
    # the plan's prephase and recentering both run
    in_lo, in_hi = [-3.0, 1.0, -20.0][:dim], [5.0, 9.0, -2.0][:dim]
    out_lo, out_hi = [-6.0, 2.0, -4.0][:dim], [8.0, 7.0, 3.0][:dim]
    bounds = (in_lo, in_hi, out_lo, out_hi)
    batch = 3
    mA, mB, n1, n2 = 700, 450, 300, 520
    A = _pts(rng, in_lo, in_hi, mA, rdt)
    B = _pts(rng, in_lo, in_hi, mB, rdt)
    S1 = _pts(rng, out_lo, out_hi, n1, rdt)
    S2 = _pts(rng, out_lo, out_hi, n2, rdt)

    opt = nnf.Options(tol=tol)
    plan = nnf.PlanT3.from_bounds(opt, sign, in_lo, in_hi, out_lo, out_hi, batch_size=batch,
                                  dtype=dtype)
    plan.set_input_points(mA, [_dev(v) for v in A])
    plan.set_output_points(n1, [_dev(v) for v in S1])

    # 1. two add_input calls (batched, bdist > num_in) then transform
    #    (bdist > num_out): each batch member is the sum of both inputs
    bd_in, bd_out = mA + 5, n1 + 3
    c1 = [cplx(rng, mA, prec if prec != "mixed" else "f64") for _ in range(batch)]
    c2 = [cplx(rng, mA, prec if prec != "mixed" else "f64") for _ in range(batch)]
    junk = np.complex128(1e6 - 1e6j)
    buf1 = np.full(batch * bd_in, junk, dtype=cdt)
    buf2 = np.full(batch * bd_in, junk, dtype=cdt)
    for b in range(batch):
        buf1[b * bd_in:b * bd_in + mA] = c1[b]
        buf2[b * bd_in:b * bd_in + mA] = c2[b]
    sentinel = 3.0 + 4.0j
    out = torch.full((batch * bd_out,), sentinel, dtype=torch.from_numpy(buf1).dtype,
                     device="cuda")
    plan.add_input(_dev(buf1), bd_in)
    plan.add_input(_dev(buf2), bd_in)
    plan.transform(out, bd_out)
    torch.cuda.synchronize()
    o = out.cpu().numpy().astype(np.complex128)
    for b in range(batch):
        seg = o[b * bd_out:b * bd_out + n1]
        csum = c1[b].astype(np.complex128) + c2[b].astype(np.complex128)
        _check(seg, _direct(A, csum, S1, sign), _alg(A, csum, S1, sign, tol, bounds), tol,
               floor, prec)
        assert np.all(o[b * bd_out + n1:(b + 1) * bd_out] == sentinel), "wrote into the gap"

    # 2. reset, new source points (new count) and new targets (new count),
    #    contiguous batch members (bdist 0 = the point counts)
    plan.reset()
    plan.set_input_points(mB, [_dev(v) for v in B])
    plan.set_output_points(n2, [_dev(v) for v in S2])
    c3 = [cplx(rng, mB, prec if prec != "mixed" else "f64") for _ in range(batch)]
    out2 = torch.zeros(batch * n2, dtype=out.dtype, device="cuda")
    plan.add_input(_dev(np.concatenate(c3).astype(cdt)))
    plan.transform(out2)
    torch.cuda.synchronize()
    o2 = out2.cpu().numpy().astype(np.complex128)
    for b in range(batch):
        _check(o2[b * n2:(b + 1) * n2], _direct(B, c3[b], S2, sign),
               _alg(B, c3[b], S2, sign, tol, bounds), tol, floor, prec)

    # 3. streaming: inputs of two source sets accumulate across
    #    set_input_points until reset (the grid is not cleared by it)
    plan.reset()
    plan.set_input_points(mA, [_dev(v) for v in A])
    plan.add_input(_dev(buf1), bd_in)
    plan.set_input_points(mB, [_dev(v) for v in B])
    plan.add_input(_dev(np.concatenate(c3).astype(cdt)))
    out3 = torch.zeros(batch * n2, dtype=out.dtype, device="cuda")
    plan.transform(out3)
    torch.cuda.synchronize()
    o3 = out3.cpu().numpy().astype(np.complex128)
    for b in range(batch):
        ref = _direct(A, c1[b], S2, sign) + _direct(B, c3[b], S2, sign)
        alg = _alg(A, c1[b], S2, sign, tol, bounds) + _alg(B, c3[b], S2, sign, tol, bounds)
        _check(o3[b * n2:(b + 1) * n2], ref, alg, tol, floor, prec)

    # 4. transform does not consume the accumulated grid: a second call
    #    gives the same result
    plan.transform(out2)
    torch.cuda.synchronize()
    assert O.rel_l2(out2.cpu().numpy().astype(np.complex128), o3) == 0.0


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_t3_bounds_grid_matches_points_constructor(prec):
    """A bounds plan whose bounds are the points' own extents sizes its grids
    exactly as the points constructor does, so both give the same result."""
    dtype, rdt, cdt, tol, floor = PREC[prec]
    rng = np.random.default_rng(5)
    x = _pts(rng, [-2.0, 0.5, 3.0], [4.0, 6.0, 9.0], 900, rdt)
    s = _pts(rng, [-5.0, -1.0, 0.0], [5.0, 2.0, 4.0], 400, rdt)
    c = cplx(rng, 900, prec)
    opt = nnf.Options(tol=tol)
    a = nnf.PlanT3(opt, -1, 900, [_dev(v) for v in x], 400, [_dev(v) for v in s])
    lo = [float(v.min()) for v in x], [float(v.max()) for v in x]
    so = [float(v.min()) for v in s], [float(v.max()) for v in s]
    b = nnf.PlanT3.from_bounds(opt, -1, lo[0], lo[1], so[0], so[1], dtype=dtype)
    assert a.info()["grid"] == b.info()["grid"]
    assert a.info()["spread_grid"] == b.info()["spread_grid"]
    b.set_input_points(900, [_dev(v) for v in x])
    b.set_output_points(400, [_dev(v) for v in s])
    oa = torch.zeros(400, dtype=torch.from_numpy(c).dtype, device="cuda")
    ob = torch.zeros_like(oa)
    a.add_input(_dev(c))
    a.transform(oa)
    b.add_input(_dev(c))
    b.transform(ob)
    torch.cuda.synchronize()
    ga, gb = oa.cpu().numpy(), ob.cpu().numpy()
    assert O.rel_l2(gb, ga) < (1e-12 if prec == "f64" else 1e-6)
    _check(gb.astype(np.complex128), _direct(x, c, s, -1), _alg(x, c, s, -1, tol, None), tol,
           floor, prec)


def test_t3_input_errors():
    opt = nnf.Options(tol=1e-5)
    with pytest.raises(nnf.InputError):
        nnf.PlanT3.from_bounds(opt, 1, [1.0], [0.0], [0.0], [1.0])
    with pytest.raises(nnf.InputError):
        nnf.PlanT3.from_bounds(opt, 1, [0.0], [1.0], [0.0], [1.0], batch_size=0)
    p = nnf.PlanT3.from_bounds(opt, 1, [0.0], [1.0], [0.0], [1.0])
    with pytest.raises(nnf.InputError):
        p.set_input_points(-1, [0])
