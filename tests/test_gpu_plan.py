"""GPU parity tests of the type 1/2/3 plans through the C ABI.

Each case mirrors a case family of the reference's tests
(tests/test_type_1_2.cpp:339-392, tests/test_type_3.cpp:387-429): points far
outside [-pi, pi) (x in [20,30], y in [-20,20], z in [-20,-15]) to exercise
the periodic fold, both signs, several upsampling factors and tolerances.
Results are compared with the FP64 direct NUDFT (the closed-form truth) by
relative l2 error, and with the FP64 restatement of the same algorithm
(oracle/), which must agree far more tightly.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import neonufft_amd as nnf
from oracle import oracle as O

pytestmark = pytest.mark.gpu

RDT = {"f32": torch.float32, "f64": torch.float64}
CDT = {"f32": torch.complex64, "f64": torch.complex128}
NPR = {"f32": np.float32, "f64": np.float64}
NPC = {"f32": np.complex64, "f64": np.complex128}


def dev(a, dtype):
    return torch.from_numpy(np.ascontiguousarray(a)).to(device="cuda", dtype=dtype)


def points(rng, dim, m, prec):
    lo_hi = [(20.0, 30.0), (-20.0, 20.0), (-20.0, -15.0)]
    return [rng.uniform(*lo_hi[d], m).astype(NPR[prec]) for d in range(dim)]


def cplx(rng, m, prec):
    return (rng.standard_normal(m) + 1j * rng.standard_normal(m)).astype(NPC[prec])


def amplification(modes, tol, sigma):
    """RMS growth of fine-grid rounding through the deconvolution
    1/phihat(k) over the kept modes, multiplied over dimensions. It is ~2.5 in
    3D at sigma = 2 but reaches ~4e3 at sigma = 1.25 (k/n up to 0.4), which sets
    the fp32 floor of those plans."""
    ns, beta = O.kernel_params(tol, sigma)
    a = 1.0
    for N in modes:
        n = max(2 * ns, int(N * sigma))
        fs = np.abs(O.fseries(n, ns, beta))
        k = np.abs(np.arange(-(N // 2), N - N // 2))
        a *= float(np.sqrt(np.mean((fs[k] / fs[0]) ** 2)))
    return a


def bound(prec, alg_err, tol, amp=1.0):
    """Allowed relative l2 error against the FP64 direct sum.

    alg_err is the error of the same algorithm run in FP64 (oracle/): the
    kernel truncation error that the reference's (tol, upsampfac) -> (ns, beta)
    rule leaves (es_kernel_param.hpp:24-56). By the triangle inequality the
    plan's error is at most alg_err plus its distance to that FP64 result,
    which alg_bound caps (1.2 alg_err: the sampled subset)."""
    return 1.2 * alg_err + alg_bound(prec, tol, amp)


def alg_bound(prec, tol, amp):
    """Allowed relative l2 distance to the same algorithm run in FP64 (the
    fp32 share of the error): fp64 rounding; or fp32 rounding, 3e-6, plus the
    same rounding amplified by the deconvolution 1/phihat (amp: up to ~4e3 at
    sigma = 1.25), plus the fp32 Horner fit of the ES kernel (~0.02 tol at
    ns = 4, far less beyond). Measured on MI355X over the type 1 / 2 parity
    cases: at most 1e-6 at sigma = 2 and tol <= 1e-4, 3.2e-5 at sigma = 1.25
    (amp 4180), 2.2e-5 at tol 1e-3 (ns = 4)."""
    if prec == "f64":
        return 1e-10
    return 3e-6 + 1.5e-8 * amp + 0.04 * tol


def sample_modes(modes, rng, s=400):
    tot = int(np.prod(modes))
    if tot <= s:
        return np.arange(tot)
    return np.sort(rng.choice(tot, size=s, replace=False))


T12_CASES = [
    # (dim, modes, M, sigma, tol, sign)
    (1, (1,), 100, 2.0, 1e-4, -1),
    (1, (2,), 100, 2.0, 1e-4, -1),
    (1, (23,), 100, 2.0, 1e-4, -1),
    (1, (256,), 100, 2.0, 1e-4, -1),
    (2, (23, 2), 100, 2.0, 1e-4, -1),
    (2, (256, 23), 100, 2.0, 1e-4, -1),
    (2, (1, 256), 100, 2.0, 1e-4, -1),
    (3, (23, 1, 2), 100, 2.0, 1e-4, -1),
    (3, (23, 23, 23), 100, 2.0, 1e-4, -1),
    (3, (256, 2, 23), 100, 2.0, 1e-4, -1),
    (1, (16,), 10, 2.0, 1e-5, 1),
    (2, (16, 16), 10, 2.0, 1e-5, 1),
    (3, (16, 16, 16), 10, 2.0, 1e-5, 1),
    (3, (16, 16, 16), 10, 2.0, 1e-5, -1),
    (1, (16,), 10, 1.25, 1e-5, -1),
    (2, (16, 16), 10, 1.6, 1e-5, -1),
    (3, (16, 16, 16), 10, 1.25, 1e-5, -1),
    (3, (32, 32, 32), 2000, 2.0, 1e-5, -1),
    (3, (45, 45, 45), 500, 2.0, 1e-3, -1),
    (2, (45, 45), 2000, 2.0, 1e-8, -1),
    (3, (24, 20, 18), 3000, 2.0, 1e-6, 1),
]


def _ids(c):
    dim, modes, m, sig, tol, sign = c
    return f"d{dim}_m{'x'.join(map(str, modes))}_n{m}_up{sig}_tol{tol:g}_s{sign}"


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("case", T12_CASES, ids=[_ids(c) for c in T12_CASES])
def test_type1(case, prec):
    dim, modes, m, sigma, tol, sign = case
    if prec == "f32" and tol < 1e-7:
        pytest.skip("below fp32 resolution")
    rng = np.random.default_rng(hash((dim, modes, m, sigma, tol, sign)) & 0xFFFF)
    x = points(rng, dim, m, prec)
    c = cplx(rng, m, prec)
    opt = nnf.Options(tol=tol, upsampfac=sigma)
    plan = nnf.Plan(opt, sign, m, [dev(v, RDT[prec]) for v in x], modes)
    out = torch.zeros(int(np.prod(modes)), dtype=CDT[prec], device="cuda")
    plan.transform_type_1(dev(c, CDT[prec]), out)
    torch.cuda.synchronize()
    got = out.cpu().numpy().astype(np.complex128)
    xs = [v.astype(np.float64) for v in x]
    idx = sample_modes(modes, rng)
    ref = O.direct_type1(xs, c.astype(np.complex128), O.mode_vectors(modes, idx), sign)
    err = O.rel_l2(got[idx], ref)
    alg = O.nufft1(xs, c.astype(np.complex128), modes, sign, tol, sigma)
    alg_err = O.rel_l2(alg[idx], ref)
    amp = amplification(modes, tol, sigma)
    assert alg_err < 20 * tol
    assert err < bound(prec, alg_err, tol, amp), (err, alg_err, amp)
    # same algorithm in FP64: only Horner / fp32 rounding separate them
    assert O.rel_l2(got, alg) < alg_bound(prec, tol, amp)


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("case", T12_CASES, ids=[_ids(c) for c in T12_CASES])
def test_type2(case, prec):
    dim, modes, m, sigma, tol, sign = case
    if prec == "f32" and tol < 1e-7:
        pytest.skip("below fp32 resolution")
    rng = np.random.default_rng((hash((dim, modes, m, sigma, tol, sign)) + 7) & 0xFFFF)
    x = points(rng, dim, m, prec)
    f = cplx(rng, int(np.prod(modes)), prec)
    opt = nnf.Options(tol=tol, upsampfac=sigma)
    plan = nnf.Plan(opt, sign, m, [dev(v, RDT[prec]) for v in x], modes)
    out = torch.zeros(m, dtype=CDT[prec], device="cuda")
    plan.transform_type_2(dev(f, CDT[prec]), None, out)
    torch.cuda.synchronize()
    got = out.cpu().numpy().astype(np.complex128)
    xs = [v.astype(np.float64) for v in x]
    ref = O.direct_type2(xs, f.astype(np.complex128), modes, sign)
    err = O.rel_l2(got, ref)
    alg = O.nufft2(xs, f.astype(np.complex128), modes, sign, tol, sigma)
    alg_err = O.rel_l2(alg, ref)
    amp = amplification(modes, tol, sigma)
    assert alg_err < 20 * tol
    assert err < bound(prec, alg_err, tol, amp), (err, alg_err, amp)
    assert O.rel_l2(got, alg) < alg_bound(prec, tol, amp)


T3_CASES = [
    # (dim, M, N, tol, sign)
    (1, 1, 1, 1e-4, 1),
    (1, 10, 200, 1e-4, -1),
    (1, 503, 503, 1e-7, 1),
    (2, 200, 10, 1e-4, -1),
    (2, 503, 503, 1e-7, -1),
    (2, 1, 503, 1e-7, 1),
    (3, 10, 99, 1e-4, 1),
    (3, 99, 99, 1e-7, 1),
    (3, 99, 1, 1e-4, 1),
]


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("case", T3_CASES, ids=[f"d{c[0]}_m{c[1]}_n{c[2]}_tol{c[3]:g}_s{c[4]}" for c in T3_CASES])
def test_type3(case, prec):
    dim, m, n, tol, sign = case
    rng = np.random.default_rng((dim * 1000 + m * 7 + n) & 0xFFFF)
    x = points(rng, dim, m, prec)
    s = points(rng, dim, n, prec)
    c = cplx(rng, m, prec)
    opt = nnf.Options(tol=tol)
    plan = nnf.PlanT3(opt, sign, m, [dev(v, RDT[prec]) for v in x], n,
                      [dev(v, RDT[prec]) for v in s])
    out = torch.zeros(n, dtype=CDT[prec], device="cuda")
    plan.add_input(dev(c, CDT[prec]))
    plan.transform(out)
    torch.cuda.synchronize()
    got = out.cpu().numpy().astype(np.complex128)
    xs = [v.astype(np.float64) for v in x]
    ss = [v.astype(np.float64) for v in s]
    ref = O.direct_type3(xs, c.astype(np.complex128), ss, sign)
    err = O.rel_l2(got, ref)
    alg = O.nufft3(xs, c.astype(np.complex128), ss, sign, tol)
    alg_err = O.rel_l2(alg, ref)
    assert alg_err < 20 * tol
    # |x.s| reaches ~1e3 rad here; the plan keeps phase arguments in FP64
    lim = 1.2 * alg_err + (1e-12 if prec == "f64" else 2e-5)
    assert err < lim, (err, alg_err)
    assert O.rel_l2(got, alg) < (1e-9 if prec == "f64" else 5e-5 + 0.05 * tol)


T3_MIXED_CASES = T3_CASES + [
    # large-magnitude sources and small frequencies, the layout where fp32
    # coordinates alone would cost ~1e-5 of phase (|x| ~ 1e4, |x.s| ~ 200)
    ("wide", 3000, 700, 1e-6, -1),
]


@pytest.mark.parametrize("case", T3_MIXED_CASES,
                         ids=[f"d{c[0]}_m{c[1]}_n{c[2]}_tol{c[3]:g}_s{c[4]}" for c in T3_MIXED_CASES])
def test_type3_mixed(case):
    """NNF_MIXED type 3 (extension): double coordinates and complex128 I/O with
    fp32 spreading / FFT / interpolation. Error bound as for fp32 plans, but
    against the direct sum over the unrounded double coordinates."""
    dim, m, n, tol, sign = case
    rng = np.random.default_rng(4242 + m + n)
    if dim == "wide":
        dim = 3
        x = [rng.uniform(-1e4, 1e4, m), rng.uniform(-1e4, 1e4, m), rng.uniform(-50, 50, m)]
        s = [rng.uniform(-0.02, 0.02, n), rng.uniform(-0.02, 0.02, n), rng.uniform(-1, 1, n)]
    else:
        x = points(rng, dim, m, "f64")
        s = points(rng, dim, n, "f64")
    c = cplx(rng, m, "f64")
    opt = nnf.Options(tol=tol)
    plan = nnf.PlanT3(opt, sign, m, [dev(v, torch.float64) for v in x], n,
                      [dev(v, torch.float64) for v in s], dtype="mixed")
    assert plan.info()["precision"] == "mixed"
    out = torch.zeros(n, dtype=torch.complex128, device="cuda")
    plan.add_input(dev(c, torch.complex128))
    plan.transform(out)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    ref = O.direct_type3(x, c, s, sign)
    err = O.rel_l2(got, ref)
    alg = O.nufft3(x, c, s, sign, tol)
    alg_err = O.rel_l2(alg, ref)
    assert alg_err < 20 * tol
    assert err < 1.2 * alg_err + 2e-5, (err, alg_err)
    assert O.rel_l2(got, alg) < 5e-5 + 0.05 * tol


# 2D type 1 through each spreading decomposition: row bands with the x-FFT
# fused into the band kernel (power-of-two n0 >= 32), row bands writing the
# fine grid, and subgrid tiles. All three must give the direct sum's answer
# to the plan's tolerance and agree with each other to rounding.
BAND_T1 = [
    # (modes, M, tol, sign, band_rows)
    ((64, 48), 3000, 1e-5, -1, 0),
    ((16, 100), 2000, 1e-5, 1, 0),
    ((512, 40), 5000, 1e-5, -1, 0),
    ((1024, 12), 4000, 1e-4, 1, 0),
    ((64, 48), 3000, 1e-5, -1, 1),
    ((64, 48), 3000, 1e-6, 1, 2),
    ((128, 30), 1000, 1e-9, -1, 0),
    # power-of-two n1 >= 32 as well: y-FFT + deconvolution fused too
    ((64, 64), 3000, 1e-5, -1, 0),
    ((16, 128), 2000, 1e-6, 1, 0),
    ((256, 16), 4000, 1e-5, 1, 0),
    ((64, 64), 3000, 1e-10, 1, 0),
    # fp64 rows of 4096 cells: one row per band (LDS budget)
    ((2048, 8), 2000, 1e-9, 1, 0),
]


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("order", [nnf.ModeOrder.CMCL, nnf.ModeOrder.FFT])
@pytest.mark.parametrize("case", BAND_T1,
                         ids=[f"m{'x'.join(map(str, c[0]))}_n{c[1]}_tol{c[2]:g}_s{c[3]}_r{c[4]}"
                              for c in BAND_T1])
def test_type1_spread_modes(case, order, prec):
    modes, m, tol, sign, rows = case
    if prec == "f32" and tol < 1e-7:
        pytest.skip("below fp32 resolution")
    rng = np.random.default_rng(hash(case) & 0xFFFF)
    x = points(rng, 2, m, prec)
    c = cplx(rng, m, prec)
    opt = nnf.Options(tol=tol, upsampfac=2.0, order=order)
    loc = [dev(v, RDT[prec]) for v in x]
    cin = dev(c, CDT[prec])
    outs = {}
    for mode in ("bands", "bands_grid", "tiles"):
        plan = nnf.Plan(opt, sign, m, loc, modes, spread_mode=mode,
                        band_rows=rows if mode != "tiles" else 0)
        assert plan.info()["spread_mode"] == mode
        out = torch.zeros(int(np.prod(modes)), dtype=CDT[prec], device="cuda")
        plan.transform_type_1(cin, out)
        plan.transform_type_1(cin, out)  # a second call gives the same result
        torch.cuda.synchronize()
        outs[mode] = out.cpu().numpy().astype(np.complex128)
    xs = [v.astype(np.float64) for v in x]
    idx = sample_modes(modes, rng)
    offt = order == nnf.ModeOrder.FFT
    ref = O.direct_type1(xs, c.astype(np.complex128), O.mode_vectors(modes, idx, offt), sign)
    amp = amplification(modes, tol, 2.0)
    alg = O.nufft1(xs, c.astype(np.complex128), modes, sign, tol, 2.0, order_fft=offt)
    alg_err = O.rel_l2(alg[idx], ref)
    for mode, got in outs.items():
        assert O.rel_l2(got[idx], ref) < bound(prec, alg_err, tol, amp), mode
    same = 1e-12 if prec == "f64" else 5e-6
    assert O.rel_l2(outs["bands"], outs["tiles"]) < same
    assert O.rel_l2(outs["bands_grid"], outs["tiles"]) < same
    # caller strides: y modes fastest (transposed output)
    plan = nnf.Plan(opt, sign, m, loc, modes, spread_mode="bands", band_rows=rows)
    out_t = torch.zeros(int(np.prod(modes)), dtype=CDT[prec], device="cuda")
    plan.transform_type_1(cin, out_t, [modes[1], 1])
    torch.cuda.synchronize()
    got_t = out_t.cpu().numpy().astype(np.complex128).reshape(modes[0], modes[1]).T.ravel()
    assert O.rel_l2(got_t, outs["bands"]) < same


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_type1_clustered_rows_fall_back_to_tiles(prec):
    """Points packed into a few fine-grid rows (thousands per cell): 2D plans
    use subgrid tiles for them, also when row bands are asked for (the band's
    fixed-point bound would count ~1e4 points per cell and lose precision);
    uniform points set later switch the plan back to row bands, and a batch
    holding such a transform falls back as a whole. The result stays within
    the tolerance of the direct sum."""
    rng = np.random.default_rng(77)
    m, modes = 200_000, (64, 64)
    x = rng.uniform(-np.pi, np.pi, m).astype(NPR[prec])
    y = rng.uniform(0.0, 0.02, m).astype(NPR[prec])  # ~1 of 128 fine rows
    c = cplx(rng, m, prec)
    opt = nnf.Options(tol=1e-5 if prec == "f32" else 1e-9)
    loc = [dev(x, RDT[prec]), dev(y, RDT[prec])]
    cin = dev(c, CDT[prec])
    outs = {}
    for mode in ("auto", "bands"):
        plan = nnf.Plan(opt, -1, m, loc, modes, spread_mode=mode)
        assert plan.info()["spread_mode"] == "tiles"
        out = torch.zeros(int(np.prod(modes)), dtype=CDT[prec], device="cuda")
        plan.transform_type_1(cin, out)
        torch.cuda.synchronize()
        outs[mode] = out.cpu().numpy().astype(np.complex128)
        xu_h = [rng.uniform(-np.pi, np.pi, 1000).astype(NPR[prec]) for _ in range(2)]
        xu = [dev(v, RDT[prec]) for v in xu_h]
        plan.set_points(1000, xu)
        assert plan.info()["spread_mode"] == "bands"
        # the plan switched from tiles to bands: its result follows the new points
        cu = cplx(rng, 1000, prec)
        ou = torch.zeros(int(np.prod(modes)), dtype=CDT[prec], device="cuda")
        plan.transform_type_1(dev(cu, CDT[prec]), ou)
        torch.cuda.synchronize()
        ref_u = O.direct_type1([v.astype(np.float64) for v in xu_h], cu.astype(np.complex128),
                               O.mode_vectors(modes, np.arange(int(np.prod(modes)))), -1)
        assert O.rel_l2(ou.cpu().numpy().astype(np.complex128), ref_u) < (
            4e-5 if prec == "f32" else 1e-8)
        # and back to the clustered points: tiles again, same result as before
        plan.set_points(m, loc)
        assert plan.info()["spread_mode"] == "tiles"
        plan.transform_type_1(cin, ou)
        torch.cuda.synchronize()
        again = ou.cpu().numpy().astype(np.complex128)
        assert O.rel_l2(again, outs[mode]) < (1e-12 if prec == "f64" else 5e-6)
    # same decomposition; points within a bin are ranked by atomics, so the
    # subproblems (and their fixed-point scales) may differ between plans
    same = 1e-12 if prec == "f64" else 5e-6
    assert O.rel_l2(outs["auto"], outs["bands"]) < same
    idx = sample_modes(modes, rng, 200)
    ref = O.direct_type1([x.astype(np.float64), y.astype(np.float64)],
                         c.astype(np.complex128), O.mode_vectors(modes, idx), -1)
    assert O.rel_l2(outs["auto"][idx], ref) < (4e-5 if prec == "f32" else 1e-8)
    batch = nnf.BatchPlan(opt, -1, [m, 500], [loc, [loc[0][:500], loc[0][500:1000]]],
                          [list(modes), [32, 32]], use_graph=False)
    assert batch.info()["spread_mode"] == "tiles"
    bo = [torch.zeros(64 * 64, dtype=CDT[prec], device="cuda"),
          torch.zeros(32 * 32, dtype=CDT[prec], device="cuda")]
    batch.transform_type_1([cin, cin[:500]], bo)
    torch.cuda.synchronize()
    assert O.rel_l2(bo[0].cpu().numpy().astype(np.complex128), outs["auto"]) < same


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_axis_fft_odd_modes(prec):
    """3D plan on a power-of-two fine grid with odd mode counts (43^3 modes
    at upsampling 1.5: 64^3 cells), so type 1 and type 2 run the y and z
    passes in LDS with the retained-mode maps of odd N, in FFT mode order
    with permuted caller strides (z fastest): against the FP64 direct sum,
    the FP64 restatement and the plan without the row kernels (rocFFT passes
    and transposes)."""
    modes = (43, 43, 43)
    sigma = 1.5
    m = 5000
    tol = 1e-5 if prec == "f32" else 1e-8
    sign = 1
    rng = np.random.default_rng(4343)
    x = points(rng, 3, m, prec)
    c = cplx(rng, m, prec)
    nm = int(np.prod(modes))
    f = cplx(rng, nm, prec)  # logical array (x fastest) in FFT mode order
    loc = [dev(v, RDT[prec]) for v in x]
    opt = nnf.Options(tol=tol, upsampfac=sigma, order=nnf.ModeOrder.FFT)
    fused = nnf.Plan(opt, sign, m, loc, modes)
    sep = nnf.Plan(opt, sign, m, loc, modes, spread_mode="tiles")
    assert tuple(fused.info()["grid"]) == (64, 64, 64)
    st = [modes[1] * modes[2], modes[2], 1]  # z fastest
    lin = np.arange(nm)
    off = (lin % 43) * st[0] + ((lin // 43) % 43) * st[1] + (lin // (43 * 43)) * st[2]
    fbuf = np.empty(nm, dtype=NPC[prec])
    fbuf[off] = f
    outs = {}
    for name, plan in (("fused", fused), ("sep", sep)):
        g1 = torch.zeros(nm, dtype=CDT[prec], device="cuda")
        plan.transform_type_1(dev(c, CDT[prec]), g1, st)
        g2 = torch.zeros(m, dtype=CDT[prec], device="cuda")
        plan.transform_type_2(dev(fbuf, CDT[prec]), st, g2)
        torch.cuda.synchronize()
        outs[name] = (g1.cpu().numpy().astype(np.complex128)[off],
                      g2.cpu().numpy().astype(np.complex128))
    xs = [v.astype(np.float64) for v in x]
    c64, f64 = c.astype(np.complex128), f.astype(np.complex128)
    amp = amplification(modes, tol, sigma)
    idx = sample_modes(modes, rng)
    ref1 = O.direct_type1(xs, c64, O.mode_vectors(modes, idx, order_fft=True), sign)
    alg1 = O.nufft1(xs, c64, modes, sign, tol, upsampfac=sigma, order_fft=True)
    # the direct type-2 sum takes CMCL order: roll every axis by N // 2
    f_cmcl = f64.reshape(tuple(reversed(modes)))
    for ax, n_ax in enumerate(reversed(modes)):
        f_cmcl = np.roll(f_cmcl, n_ax // 2, axis=ax)
    ref2 = O.direct_type2(xs, f_cmcl.ravel(), modes, sign)
    alg2 = O.nufft2(xs, f64, modes, sign, tol, upsampfac=sigma, order_fft=True)
    for name, (g1, g2) in outs.items():
        assert O.rel_l2(g1[idx], ref1) < bound(prec, O.rel_l2(alg1[idx], ref1), tol, amp), name
        assert O.rel_l2(g1, alg1) < alg_bound(prec, tol, amp), name
        assert O.rel_l2(g2, ref2) < bound(prec, O.rel_l2(alg2, ref2), tol, amp), name
        assert O.rel_l2(g2, alg2) < alg_bound(prec, tol, amp), name
    # the two FFT paths agree to rounding
    for t in (0, 1):
        d = O.rel_l2(outs["fused"][t], outs["sep"][t])
        assert d < (1e-5 if prec == "f32" else 1e-12), (t, d)


Y_LDS_CASES = [
    # (modes, upsampfac, fine grid): x' groups of R = 8 / 4 / 16 lines, the
    # last group partial for 43 x modes; odd retained y counts
    ((43, 683), 1.5, (64, 1024)),
    ((32, 1024), 2.0, (64, 2048)),
    ((43, 171), 1.5, (64, 256)),
]


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("order", [nnf.ModeOrder.CMCL, nnf.ModeOrder.FFT])
@pytest.mark.parametrize("case", Y_LDS_CASES, ids=[f"m{c[0][0]}x{c[0][1]}" for c in Y_LDS_CASES])
def test_y_pass_in_lds(case, order, prec):
    """2D plans on fine grids with power-of-two n1: the y pass and the
    deconvolution run in LDS (type 1: y_out2_kernel after the band kernel's
    x-FFT; type 2: y_in2_kernel before the row kernel's x-FFT), over full and
    partial groups of x modes, with caller strides in both orders (x
    fastest, y fastest): against the FP64 direct sum, the FP64 restatement
    and plans on rocFFT passes with transposes (type 1: bands writing the
    fine grid; type 2: subgrid tiles without the row kernels)."""
    modes, sigma, grid = case
    m = 4000
    tol = 1e-5 if prec == "f32" else 1e-9
    sign = -1
    rng = np.random.default_rng(hash(case) & 0xFFFF)
    x = points(rng, 2, m, prec)
    c = cplx(rng, m, prec)
    loc = [dev(v, RDT[prec]) for v in x]
    cin = dev(c, CDT[prec])
    opt = nnf.Options(tol=tol, upsampfac=sigma, order=order)
    nm = int(np.prod(modes))
    outs = {}
    for mode in ("bands", "bands_grid"):
        plan = nnf.Plan(opt, sign, m, loc, modes, spread_mode=mode)
        assert tuple(plan.info()["grid"]) == grid
        assert plan.info()["spread_mode"] == mode
        out = torch.zeros(nm, dtype=CDT[prec], device="cuda")
        plan.transform_type_1(cin, out)
        out_t = torch.zeros(nm, dtype=CDT[prec], device="cuda")
        plan.transform_type_1(cin, out_t, [modes[1], 1])  # y fastest
        torch.cuda.synchronize()
        got = out.cpu().numpy().astype(np.complex128)
        got_t = out_t.cpu().numpy().astype(np.complex128).reshape(modes[0], modes[1]).T.ravel()
        outs[mode] = (got, got_t)
    xs = [v.astype(np.float64) for v in x]
    c64 = c.astype(np.complex128)
    offt = order == nnf.ModeOrder.FFT
    idx = sample_modes(modes, rng)
    ref = O.direct_type1(xs, c64, O.mode_vectors(modes, idx, offt), sign)
    alg = O.nufft1(xs, c64, modes, sign, tol, upsampfac=sigma, order_fft=offt)
    amp = amplification(modes, tol, sigma)
    same = 1e-12 if prec == "f64" else 5e-6
    for mode, (got, got_t) in outs.items():
        assert O.rel_l2(got[idx], ref) < bound(prec, O.rel_l2(alg[idx], ref), tol, amp), mode
        assert O.rel_l2(got, alg) < alg_bound(prec, tol, amp), mode
        assert O.rel_l2(got_t, got) < same, mode
    assert O.rel_l2(outs["bands"][0], outs["bands_grid"][0]) < same
    # type 2: modes in `order`, x fastest, then the same array y fastest
    f = cplx(rng, nm, prec)
    f_t = f.reshape(modes[1], modes[0]).T.ravel()
    res = {}
    for mode in ("auto", "tiles"):
        plan = nnf.Plan(opt, sign, m, loc, modes, spread_mode=mode)
        g = torch.zeros(m, dtype=CDT[prec], device="cuda")
        plan.transform_type_2(dev(f, CDT[prec]), None, g)
        g_t = torch.zeros(m, dtype=CDT[prec], device="cuda")
        plan.transform_type_2(dev(f_t, CDT[prec]), [modes[1], 1], g_t)
        torch.cuda.synchronize()
        res[mode] = (g.cpu().numpy().astype(np.complex128), g_t.cpu().numpy().astype(np.complex128))
    f64 = f.astype(np.complex128)
    f_cmcl = f64.reshape(modes[1], modes[0])
    if offt:  # the direct sum takes CMCL order
        f_cmcl = np.roll(np.roll(f_cmcl, modes[1] // 2, axis=0), modes[0] // 2, axis=1)
    ref2 = O.direct_type2(xs, f_cmcl.ravel(), modes, sign)
    alg2 = O.nufft2(xs, f64, modes, sign, tol, upsampfac=sigma, order_fft=offt)
    for mode, (g, g_t) in res.items():
        assert O.rel_l2(g, ref2) < bound(prec, O.rel_l2(alg2, ref2), tol, amp), mode
        assert O.rel_l2(g, alg2) < alg_bound(prec, tol, amp), mode
        assert O.rel_l2(g_t, g) < same, mode
    assert O.rel_l2(res["auto"][0], res["tiles"][0]) < same


ROW_CASES = [
    # (dim, modes, M, tol, sign): fine grids with a power-of-two x extent, so
    # type 1 joins the tile reduction to the x-FFT and type 2 the zero padding
    # (x extent >= 2 bins, so the tile reduction is the bin-regular one)
    (3, (32, 12, 20), 4000, 1e-5, -1),
    (3, (64, 40, 24), 20000, 1e-6, 1),
    # powers of two on every axis (64 x 64 x 32): the y and z passes run in
    # LDS too (against rocFFT passes and transposes in the "tiles" plan)
    (3, (32, 31, 16), 6000, 1e-5, -1),
    (2, (128, 50), 40000, 1e-5, 1),
    (2, (64, 200), 60000, 1e-4, -1),
]


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("case", ROW_CASES,
                         ids=[f"d{c[0]}_m{'x'.join(map(str, c[1]))}_tol{c[3]:g}" for c in ROW_CASES])
def test_row_fft_paths(case, prec):
    """Type 1 / type 2 through the row kernels (x-FFT in LDS joined to the
    tile reduction and to the zero padding) against the FP64 direct sum and
    the FP64 restatement, and against the separate reduction + rocFFT x pass
    (spread_mode="tiles") of the same points. 2D plans use subgrid tiles
    for these points (clustered in a few rows), so type 1 takes the row
    reduction there too."""
    dim, modes, m, tol, sign = case
    if prec == "f64":
        tol *= 1e-3
    rng = np.random.default_rng(int(np.prod(modes)) + m)
    x = points(rng, dim, m, prec)
    if dim == 2:
        # all points within 3 rows of the 2D grid: the row bands refuse them
        x[1] = (rng.uniform(-0.02, 0.02, m) + 1.0).astype(NPR[prec])
    c = cplx(rng, m, prec)
    f = cplx(rng, int(np.prod(modes)), prec)
    loc = [dev(v, RDT[prec]) for v in x]
    opt = nnf.Options(tol=tol)
    fused = nnf.Plan(opt, sign, m, loc, modes)
    sep = nnf.Plan(opt, sign, m, loc, modes, spread_mode="tiles")
    assert fused.info()["spread_mode"] == "tiles"
    nm = int(np.prod(modes))
    outs = {}
    for name, plan in (("fused", fused), ("sep", sep)):
        f1 = torch.zeros(nm, dtype=CDT[prec], device="cuda")
        plan.transform_type_1(dev(c, CDT[prec]), f1)
        c2 = torch.zeros(m, dtype=CDT[prec], device="cuda")
        plan.transform_type_2(dev(f, CDT[prec]), None, c2)
        torch.cuda.synchronize()
        outs[name] = (f1.cpu().numpy().astype(np.complex128), c2.cpu().numpy().astype(np.complex128))
    xs = [v.astype(np.float64) for v in x]
    amp = amplification(modes, tol, 2.0)
    idx = sample_modes(modes, rng)
    ref1 = O.direct_type1(xs, c.astype(np.complex128), O.mode_vectors(modes, idx), sign)
    alg1 = O.nufft1(xs, c.astype(np.complex128), modes, sign, tol)
    ref2 = O.direct_type2(xs, f.astype(np.complex128), modes, sign)
    alg2 = O.nufft2(xs, f.astype(np.complex128), modes, sign, tol)
    for name, (g1, g2) in outs.items():
        assert O.rel_l2(g1[idx], ref1) < bound(prec, O.rel_l2(alg1[idx], ref1), tol, amp), name
        assert O.rel_l2(g1, alg1) < alg_bound(prec, tol, amp), name
        assert O.rel_l2(g2, ref2) < bound(prec, O.rel_l2(alg2, ref2), tol, amp), name
        assert O.rel_l2(g2, alg2) < alg_bound(prec, tol, amp), name
    same = 1e-12 if prec == "f64" else 2e-6
    assert O.rel_l2(outs["fused"][0], outs["sep"][0]) < same
    assert O.rel_l2(outs["fused"][1], outs["sep"][1]) < same


@pytest.mark.parametrize("dim,prec", [(3, "f32"), (2, "f64"), (2, "f32")])
def test_row_fft_odd_bin_count(dim, prec):
    """Row reduction over an odd number of x bins (three passes: the last
    bin's wrapped tile row meets bin 0's) and a partial last bin, against the
    FP64 direct sum and the separate reduction. (2D: points in a few rows, so
    the plan uses subgrid tiles; fp64 3D tiles of this width exceed the LDS
    budget.)"""
    tol = 1e-5 if prec == "f32" else 1e-8
    modes = (64, 20, 12) if dim == 3 else (64, 90)
    m, sign = (6000 if dim == 3 else 60000), 1
    rng = np.random.default_rng(505 + dim)
    x = points(rng, dim, m, prec)
    if dim == 2:
        x[1] = (rng.uniform(-0.02, 0.02, m) - 2.0).astype(NPR[prec])
    c = cplx(rng, m, prec)
    loc = [dev(v, RDT[prec]) for v in x]
    opt = nnf.Options(tol=tol)
    bins = [48, 8, 8] if dim == 3 else [48, 32]
    # 128 fine cells in x cut into 48-wide bins: 3 bins, the last 32 wide
    fused = nnf.Plan(opt, sign, m, loc, modes, bin_size=bins)
    sep = nnf.Plan(opt, sign, m, loc, modes, bin_size=bins, spread_mode="tiles")
    info = fused.info()
    assert info["grid"][0] == 128 and info["bin_size"][0] == 48, info
    assert info["spread_mode"] == "tiles"
    nm = int(np.prod(modes))
    outs = []
    for plan in (fused, sep):
        f1 = torch.zeros(nm, dtype=CDT[prec], device="cuda")
        plan.transform_type_1(dev(c, CDT[prec]), f1)
        torch.cuda.synchronize()
        outs.append(f1.cpu().numpy().astype(np.complex128))
    xs = [v.astype(np.float64) for v in x]
    idx = sample_modes(modes, rng)
    ref = O.direct_type1(xs, c.astype(np.complex128), O.mode_vectors(modes, idx), sign)
    alg = O.nufft1(xs, c.astype(np.complex128), modes, sign, tol)
    amp = amplification(modes, tol, 2.0)
    assert O.rel_l2(outs[0][idx], ref) < bound(prec, O.rel_l2(alg[idx], ref), tol, amp)
    assert O.rel_l2(outs[0], alg) < alg_bound(prec, tol, amp)
    assert O.rel_l2(outs[0], outs[1]) < (1e-12 if prec == "f64" else 2e-6)
