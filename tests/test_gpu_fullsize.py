"""GPU checks at the full size of the 3D workload: 1e7 uniform points in
[-pi, pi)^3, 128^3 modes, tol 1e-5, upsampling 2, fp32 (fine grid 256^3,
8192 subproblems, the row kernels and the LDS y / z passes).

- Type 1 and type 2 against the FP64 direct sum on sampled outputs (the zero
  mode of type 1 is the plain sum of the strengths) and against the FP64
  restatement of the same algorithm at full size.
- Linearity of type 1: T1(a c1 + b c2) = a T1(c1) + b T1(c2). Each call
  scales its fixed-point tiles by its own strengths, so this holds only to
  the fp32 share of the error.
- The adjoint identity <T1_s c, f> = <c, T2_-s f> between a type-1 plan of
  sign s and a type-2 plan of sign -s. Each side differs from the exact
  bilinear form by the transform's error (~tol relative) projected onto a
  random vector, i.e. by ~tol |T1 c| |f| / sqrt(modes); the wrong sign makes
  them differ by their own size, ~|T1 c| |f| / sqrt(modes).

The clustered type-2 workload (2e7 points, 256^3 modes), the type-3 workload
(2e6 sources, 1e6 targets) and the batched workload (256 uneven 2D type-1
transforms) are checked at their full sizes as well, on sampled outputs.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import neonufft_amd as nnf
from oracle import oracle as O

pytestmark = pytest.mark.gpu

M = 10_000_000
N = 128
MODES = (N, N, N)
TOL = 1e-5
SIGN = -1


def _dev(a, dtype):
    return torch.from_numpy(np.ascontiguousarray(a)).to(device="cuda", dtype=dtype)


def _host(t):
    return t.cpu().numpy().astype(np.complex128)


def _cplx(rng, n):
    return (rng.standard_normal(n, dtype=np.float32)
            + 1j * rng.standard_normal(n, dtype=np.float32)).astype(np.complex64)


@pytest.fixture(scope="module")
def data():
    rng = np.random.default_rng(20261203)
    x = [rng.uniform(-np.pi, np.pi, M).astype(np.float32) for _ in range(3)]
    loc = [_dev(v, torch.float32) for v in x]
    opt = nnf.Options(tol=TOL, upsampfac=2.0)
    plan = nnf.Plan(opt, SIGN, M, loc, MODES)
    info = plan.info()
    assert tuple(info["grid"]) == (256, 256, 256)
    return {"rng": rng, "x": x, "xs": [v.astype(np.float64) for v in x], "loc": loc,
            "opt": opt, "plan": plan}


def _bound(alg_err):
    # fp32 share at sigma = 2: rounding 3e-6 plus the Horner fit 0.04 tol
    return 1.2 * alg_err + 3e-6 + 0.04 * TOL


def test_fullsize_type1(data):
    rng, plan = data["rng"], data["plan"]
    c = _cplx(rng, M)
    out = torch.empty(N ** 3, dtype=torch.complex64, device="cuda")
    plan.transform_type_1(_dev(c, torch.complex64), out)
    torch.cuda.synchronize()
    got = _host(out)
    c64 = c.astype(np.complex128)
    # sampled modes, and the zero mode: the plain sum of the strengths
    zero = (N // 2) + N * (N // 2) + N * N * (N // 2)
    idx = np.unique(np.concatenate([rng.choice(N ** 3, size=199, replace=False), [zero]]))
    ref = O.direct_type1(data["xs"], c64, O.mode_vectors(MODES, idx), SIGN)
    assert abs(ref[np.searchsorted(idx, zero)] - c64.sum()) < 1e-9 * np.abs(c64).sum()
    alg = O.nufft1(data["xs"], c64, MODES, SIGN, TOL, 2.0)
    alg_err = O.rel_l2(alg[idx], ref)
    assert alg_err < 2 * TOL
    err = O.rel_l2(got[idx], ref)
    assert err < _bound(alg_err), (err, alg_err)
    assert O.rel_l2(got, alg) < 3e-6 + 0.04 * TOL


def test_fullsize_type2(data):
    rng, plan = data["rng"], data["plan"]
    f = _cplx(rng, N ** 3)
    out = torch.empty(M, dtype=torch.complex64, device="cuda")
    plan.transform_type_2(_dev(f, torch.complex64), None, out)
    torch.cuda.synchronize()
    got = _host(out)
    f64 = f.astype(np.complex128)
    idx = np.sort(rng.choice(M, size=200, replace=False))
    ref = O.direct_type2([v[idx] for v in data["xs"]], f64, MODES, SIGN)
    alg = O.nufft2(data["xs"], f64, MODES, SIGN, TOL, 2.0)
    alg_err = O.rel_l2(alg[idx], ref)
    assert alg_err < 2 * TOL
    err = O.rel_l2(got[idx], ref)
    assert err < _bound(alg_err), (err, alg_err)
    assert O.rel_l2(got, alg) < 3e-6 + 0.04 * TOL


def test_fullsize_type1_linear(data):
    rng, plan = data["rng"], data["plan"]
    c1, c2 = _cplx(rng, M), _cplx(rng, M)
    a, b = 0.75 - 0.5j, -1.25 + 2.0j
    comb = (a * c1.astype(np.complex128) + b * c2.astype(np.complex128)).astype(np.complex64)
    outs = []
    for c in (c1, c2, comb):
        o = torch.empty(N ** 3, dtype=torch.complex64, device="cuda")
        plan.transform_type_1(_dev(c, torch.complex64), o)
        outs.append(o)
    torch.cuda.synchronize()
    g1, g2, gc = (_host(o) for o in outs)
    lin = a * g1 + b * g2
    assert O.rel_l2(gc, lin) < 1e-5


def test_fullsize_adjoint(data):
    rng = data["rng"]
    plan1 = data["plan"]
    plan2 = nnf.Plan(data["opt"], -SIGN, M, data["loc"], MODES)
    c = _cplx(rng, M)
    f = _cplx(rng, N ** 3)
    t1 = torch.empty(N ** 3, dtype=torch.complex64, device="cuda")
    plan1.transform_type_1(_dev(c, torch.complex64), t1)
    t2 = torch.empty(M, dtype=torch.complex64, device="cuda")
    plan2.transform_type_2(_dev(f, torch.complex64), None, t2)
    torch.cuda.synchronize()
    g1, g2 = _host(t1), _host(t2)
    f64, c64 = f.astype(np.complex128), c.astype(np.complex128)
    lhs = np.vdot(f64, g1)   # <T1 c, f> = sum_k T1(c)_k conj(f_k)
    rhs = np.vdot(g2, c64)   # <c, T2 f> = sum_j c_j conj(T2(f)_j)
    scale = np.linalg.norm(g1) * np.linalg.norm(f64) / np.sqrt(N ** 3)
    assert abs(lhs) > 0.05 * scale  # the form itself is of the expected size
    assert abs(lhs - rhs) < 20 * TOL * scale, (abs(lhs - rhs) / scale)


@pytest.mark.parametrize("tol,tuned", [(1e-4, False), (1e-6, True)], ids=["tol1e-4_default",
                                                                           "tol1e-6_auto"])
def test_fullsize_clustered_type2(tol, tuned):
    """The clustered type-2 workload at full size: 2e7 points in 16 Gaussians
    (sigma 0.05), 256^3 modes (fine grid 512^3), the default plan at tol 1e-4
    and the autotuned one at 1e-6; 200 sampled points against the FP64
    direct sum and the FP64 restatement (type 2 at a point depends on that
    point alone, so both run on the sample)."""
    from test_gpu_clustered import clustered
    from test_gpu_plan import alg_bound, amplification, bound

    m, modes = 20_000_000, (256, 256, 256)
    rng = np.random.default_rng(0xC3 + int(tuned))
    x = clustered(rng, m, np.float32)
    f = _cplx(rng, int(np.prod(modes)))
    plan = nnf.Plan(nnf.Options(tol=tol), SIGN, m, [_dev(v, torch.float32) for v in x], modes,
                    bin_size="auto" if tuned else None)
    assert tuple(plan.info()["grid"]) == (512, 512, 512)
    out = torch.empty(m, dtype=torch.complex64, device="cuda")
    plan.transform_type_2(_dev(f, torch.complex64), None, out)
    torch.cuda.synchronize()
    idx = np.sort(rng.choice(m, size=200, replace=False))
    got = out[torch.from_numpy(idx).to("cuda")].cpu().numpy().astype(np.complex128)
    del plan, out
    xs = [v[idx].astype(np.float64) for v in x]
    f64 = f.astype(np.complex128)
    ref = O.direct_type2(xs, f64, modes, SIGN)
    alg = O.nufft2(xs, f64, modes, SIGN, tol)
    alg_err = O.rel_l2(alg, ref)
    amp = amplification(modes, tol, 2.0)
    assert alg_err < 20 * tol
    err = O.rel_l2(got, ref)
    assert err < bound("f32", alg_err, tol, amp), (err, alg_err)
    assert O.rel_l2(got, alg) < alg_bound("f32", tol, amp)


def test_fullsize_type3_mixed():
    """The type-3 workload at full size: 2e6 sources with x, y in [-1e4, 1e4]
    and z in [-50, 50], 1e6 targets in [-0.02, 0.02]^2 x [-1, 1], tol 1e-6, the
    mixed plan (double coordinates and complex128 I/O, FP64 fold and phases,
    fp32 spreading / FFT / interpolation). 200 sampled targets against the
    direct sum over the double coordinates and the FP64 restatement with the
    grids sized from the full point sets' extents, as the plan's are; the
    fp32 share is held to 5e-6 (measured at this size: ~1.5e-6)."""
    m, n, tol = 2_000_000, 1_000_000, 1e-6
    rng = np.random.default_rng(0x7733)
    x = [rng.uniform(-1e4, 1e4, m), rng.uniform(-1e4, 1e4, m), rng.uniform(-50, 50, m)]
    s = [rng.uniform(-0.02, 0.02, n), rng.uniform(-0.02, 0.02, n), rng.uniform(-1, 1, n)]
    c = rng.standard_normal(m) + 1j * rng.standard_normal(m)
    plan = nnf.PlanT3(nnf.Options(tol=tol), SIGN, m, [_dev(v, torch.float64) for v in x], n,
                      [_dev(v, torch.float64) for v in s], dtype="mixed")
    assert plan.info()["precision"] == "mixed"
    out = torch.empty(n, dtype=torch.complex128, device="cuda")
    plan.add_input(_dev(c, torch.complex128))
    plan.transform(out)
    torch.cuda.synchronize()
    got_all = out.cpu().numpy()
    del plan, out
    tidx = np.sort(rng.choice(n, size=200, replace=False))
    st = [v[tidx] for v in s]
    ref = O.direct_type3(x, c, st, SIGN)
    bnd = ([float(v.min()) for v in x], [float(v.max()) for v in x],
           [float(v.min()) for v in s], [float(v.max()) for v in s])
    alg = O.nufft3(x, c, st, SIGN, tol, bounds=bnd)
    alg_err = O.rel_l2(alg, ref)
    assert alg_err < 20 * tol
    got = got_all[tidx]
    err = O.rel_l2(got, ref)
    assert err < 1.2 * alg_err + 5e-6, (err, alg_err)
    assert O.rel_l2(got, alg) < 5e-6


def test_fullsize_batched_2d():
    """The batched workload at full size: 256 2D type-1 transforms, point
    counts floor(10^U[4, 6]), modes {256, 512, 1024}^2, tol 1e-5, one
    BatchPlan replayed through its HIP graph. Every transform against its plain
    plan (the same kernels), and the sparsest and densest transforms against
    the FP64 direct sum on sampled modes and the FP64 restatement."""
    from test_gpu_plan import alg_bound, amplification, bound

    count, tol = 256, 1e-5
    rng = np.random.default_rng(0xBA7C)
    ms = np.floor(10.0 ** rng.uniform(4.0, 6.0, count)).astype(np.int64)
    nmodes = rng.choice([256, 512, 1024], size=count)
    xs, cs = [], []
    for t in range(count):
        xs.append([rng.uniform(-np.pi, np.pi, int(ms[t])).astype(np.float32) for _ in range(2)])
        cs.append(_cplx(rng, int(ms[t])))
    locs = [[_dev(v, torch.float32) for v in x] for x in xs]
    cin = [_dev(c, torch.complex64) for c in cs]
    modes = [[int(n), int(n)] for n in nmodes]
    opt = nnf.Options(tol=tol, upsampfac=2.0)
    batch = nnf.BatchPlan(opt, SIGN, [int(v) for v in ms], locs, modes, use_graph=True)
    outs = [torch.empty(int(n) ** 2, dtype=torch.complex64, device="cuda") for n in nmodes]
    batch.transform_type_1(cin, outs)
    batch.transform_type_1(cin, outs)  # the graph replay
    torch.cuda.synchronize()
    assert batch.info()["graph"]
    worst = 0.0
    for t in range(count):
        plan = nnf.Plan(opt, SIGN, int(ms[t]), locs[t], modes[t])
        ref_t = torch.empty_like(outs[t])
        plan.transform_type_1(cin[t], ref_t)
        torch.cuda.synchronize()
        worst = max(worst, float(torch.linalg.vector_norm(outs[t] - ref_t)
                                 / torch.linalg.vector_norm(ref_t)))
        del plan
    assert worst < 5e-6, worst
    for t in (int(np.argmin(ms)), int(np.argmax(ms))):
        md = tuple(modes[t])
        x64 = [v.astype(np.float64) for v in xs[t]]
        c64 = cs[t].astype(np.complex128)
        idx = np.sort(rng.choice(md[0] * md[1], size=200, replace=False))
        ref = O.direct_type1(x64, c64, O.mode_vectors(md, idx), SIGN)
        alg = O.nufft1(x64, c64, md, SIGN, tol, 2.0)
        alg_err = O.rel_l2(alg[idx], ref)
        amp = amplification(md, tol, 2.0)
        got = outs[t].cpu().numpy().astype(np.complex128)
        assert alg_err < 20 * tol
        assert O.rel_l2(got[idx], ref) < bound("f32", alg_err, tol, amp), t
        assert O.rel_l2(got, alg) < alg_bound("f32", tol, amp), t
