"""GPU tests of the batched plan (many uneven type 1 / type 2 transforms in one
set of launches), through the C ABI.

The batched plan has no reference counterpart; its contract is that each
transform gives what a separate plan of the reference's interface
(include/neonufft/gpu/plan.hpp:25-46) gives for the same points, modes and
options. So every transform is compared with a plain ``Plan`` run alone
(same kernels, so agreement is to rounding of the float atomics' order), and
with the FP64 direct sum (closed-form truth) and the same algorithm in FP64
(oracle/), type 1 and type 2.
Graph replays are checked against eager runs with changing buffers and
unrelated eager work between the replays.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import neonufft_amd as nnf
from oracle import oracle as O

from test_gpu_layouts import to_cmcl
from test_gpu_plan import CDT, RDT, alg_bound, amplification, bound, cplx, dev, points

pytestmark = pytest.mark.gpu

# per transform: (modes, points); repeated shapes share one batched FFT
SHAPES = {
    1: [((64,), 300), ((17,), 0), ((64,), 1000), ((200,), 57), ((5,), 40)],
    2: [((32, 24), 300), ((8, 8), 0), ((32, 24), 1500), ((60, 2), 77), ((8, 8), 900),
        ((1, 40), 33)],
    3: [((12, 10, 8), 400), ((16, 16, 16), 2500), ((12, 10, 8), 0), ((4, 20, 6), 300)],
    # 2D with power-of-two fine grids >= 32 per dimension: type 1 runs the
    # fused path (band spreading + x-FFT in LDS, then y-FFT + deconvolution
    # in one launch over all transforms)
    "2f": [((32, 32), 500), ((64, 16), 800), ((32, 32), 0), ((16, 64), 300), ((128, 32), 1200)],
}

# relative l2 distance allowed between a batched transform and a plain plan
SAME = {"f32": 2e-6, "f64": 1e-13}


def _rel(a, b):
    a = a.ravel()
    b = b.ravel()
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))


def _setup(dim, prec, seed):
    rng = np.random.default_rng(seed)
    spec = SHAPES[dim]
    dim = 2 if dim == "2f" else dim
    locs, cs, fs = [], [], []
    for modes, m in spec:
        x = points(rng, dim, m, prec)
        locs.append([dev(v, RDT[prec]) for v in x])
        cs.append(dev(cplx(rng, m, prec), CDT[prec]))
        nm = int(np.prod(modes))
        fs.append(dev(cplx(rng, nm, prec), CDT[prec]))
    return spec, locs, cs, fs


def _to_np(t):
    return t.detach().cpu().numpy()


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("dim", [1, 2, "2f", 3])
@pytest.mark.parametrize("order", [nnf.ModeOrder.CMCL, nnf.ModeOrder.FFT])
@pytest.mark.parametrize("sign", [-1, 1])
def test_batch_matches_plans(dim, prec, order, sign):
    tol = 1e-5 if prec == "f32" else 1e-9
    opt = nnf.Options(tol=tol, order=order)
    key = dim
    dim = 2 if key == "2f" else key
    spec, locs, cs, fs = _setup(key, prec, 10 * dim + (sign > 0) + (key == "2f"))
    num = [m for _, m in spec]
    modes = [list(md) for md, _ in spec]
    batch = nnf.BatchPlan(opt, sign, num, locs, modes, dtype=RDT[prec], use_graph=False)
    info = batch.info()
    if key == "2f":
        assert info["spread_mode"] == "bands"
    assert info["count"] == len(spec)
    assert info["num_points"] == sum(num)
    assert info["num_shapes"] == len({tuple(md) for md, _ in spec})

    outs1 = [torch.empty(int(np.prod(md)), dtype=CDT[prec], device="cuda") for md, _ in spec]
    outs2 = [torch.empty(max(m, 1), dtype=CDT[prec], device="cuda") for _, m in spec]
    batch.transform_type_1(cs, outs1)
    batch.transform_type_2(fs, outs2)
    torch.cuda.synchronize()

    for t, (md, m) in enumerate(spec):
        plan = nnf.Plan(opt, sign, m, locs[t], md, dtype=RDT[prec])
        r1 = torch.empty_like(outs1[t])
        plan.transform_type_1(cs[t], r1)
        r2 = torch.empty_like(outs2[t])
        plan.transform_type_2(fs[t], None, r2)
        torch.cuda.synchronize()
        if m == 0:
            assert np.all(_to_np(outs1[t]) == 0), f"transform {t}: no points must give zeros"
        else:
            assert _rel(_to_np(outs1[t]), _to_np(r1)) < SAME[prec], f"type 1 transform {t}"
            assert _rel(_to_np(outs2[t])[:m], _to_np(r2)[:m]) < SAME[prec], f"type 2 transform {t}"

    # every transform, type 1 and type 2, against the FP64 direct sum and the
    # same algorithm in FP64 (oracle/)
    offt = nnf.ModeOrder(order) == nnf.ModeOrder.FFT
    for t, (md, m) in enumerate(spec):
        if m == 0:
            continue
        xs = [_to_np(v).astype(np.float64) for v in locs[t]]
        amp = amplification(md, tol, 2.0)
        c64 = _to_np(cs[t]).astype(np.complex128)
        kk = O.mode_vectors(md, np.arange(int(np.prod(md))), order_fft=offt)
        truth1 = O.direct_type1(xs, c64, kk, sign)
        alg1 = O.nufft1(xs, c64, md, sign, tol, order_fft=offt)
        got1 = _to_np(outs1[t]).astype(np.complex128)
        alg_err = _rel(alg1, truth1)
        assert _rel(got1, truth1) < bound(prec, alg_err, tol, amp), f"type 1 transform {t}"
        assert _rel(got1, alg1) < alg_bound(prec, tol, amp), f"type 1 transform {t}"
        f64 = _to_np(fs[t]).astype(np.complex128)
        truth2 = O.direct_type2(xs, to_cmcl(f64, md, offt), md, sign)
        alg2 = O.nufft2(xs, f64, md, sign, tol, order_fft=offt)
        got2 = _to_np(outs2[t])[:m].astype(np.complex128)
        alg_err = _rel(alg2, truth2)
        assert _rel(got2, truth2) < bound(prec, alg_err, tol, amp), f"type 2 transform {t}"
        assert _rel(got2, alg2) < alg_bound(prec, tol, amp), f"type 2 transform {t}"


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("shapes", [2, "2f"])
def test_batch_graph_replays(prec, shapes):
    """Captured graphs replay correctly when the caller's buffers change
    between calls, with unrelated eager work in between."""
    opt = nnf.Options(tol=1e-5 if prec == "f32" else 1e-9)
    spec, locs, cs, fs = _setup(shapes, prec, 99)
    num = [m for _, m in spec]
    modes = [list(md) for md, _ in spec]
    gb = nnf.BatchPlan(opt, -1, num, locs, modes, dtype=RDT[prec], use_graph=True)
    eb = nnf.BatchPlan(opt, -1, num, locs, modes, dtype=RDT[prec], use_graph=False)
    rng = np.random.default_rng(5)
    for rep in range(4):
        cin = [dev(cplx(rng, max(m, 0), prec), CDT[prec]) for _, m in spec]
        fin = [dev(cplx(rng, int(np.prod(md)), prec), CDT[prec]) for md, _ in spec]
        g1 = [torch.empty(int(np.prod(md)), dtype=CDT[prec], device="cuda") for md, _ in spec]
        e1 = [torch.empty_like(v) for v in g1]
        g2 = [torch.empty(max(m, 1), dtype=CDT[prec], device="cuda") for _, m in spec]
        e2 = [torch.empty_like(v) for v in g2]
        gb.transform_type_1(cin, g1)
        # unrelated eager work on the same stream between the replays
        junk = torch.randn(1 << 16, device="cuda").cumsum(0)
        gb.transform_type_2(fin, g2)
        eb.transform_type_1(cin, e1)
        eb.transform_type_2(fin, e2)
        torch.cuda.synchronize()
        assert float(junk[-1].abs()) >= 0.0
        for t, (md, m) in enumerate(spec):
            assert _rel(_to_np(g1[t]), _to_np(e1[t])) < SAME[prec], f"replay {rep} type 1 t={t}"
            if m:
                assert _rel(_to_np(g2[t])[:m], _to_np(e2[t])[:m]) < SAME[prec], \
                    f"replay {rep} type 2 t={t}"
        assert gb.info()["graph"], "graph capture did not happen"
        assert not eb.info()["graph"]


# new point counts per transform for test_batch_set_points (0 empties one,
# larger counts regrow the point views)
NUM2 = {3: [120, 0, 3000, 45], "2f": [0, 5000, 700, 40, 2500]}


@pytest.mark.parametrize("dim,prec,use_graph",
                         [(3, "f32", False), ("2f", "f32", False), ("2f", "f32", True),
                          ("2f", "f64", True)])
def test_batch_set_points(dim, prec, use_graph):
    """New points (and point counts) through set_points match a fresh plan,
    also after a graph was captured for the old points."""
    opt = nnf.Options(tol=1e-5 if prec == "f32" else 1e-9)
    spec, locs, cs, fs = _setup(dim, prec, 3)
    pdim = 2 if dim == "2f" else dim
    modes = [list(md) for md, _ in spec]
    batch = nnf.BatchPlan(opt, 1, [m for _, m in spec], locs, modes, dtype=RDT[prec],
                          use_graph=use_graph)
    outs = [torch.empty(int(np.prod(md)), dtype=CDT[prec], device="cuda") for md, _ in spec]
    batch.transform_type_1(cs, outs)
    torch.cuda.synchronize()
    rng = np.random.default_rng(8)
    num2 = NUM2[dim]
    locs2 = [[dev(v, RDT[prec]) for v in points(rng, pdim, m, prec)] for m in num2]
    cs2 = [dev(cplx(rng, m, prec), CDT[prec]) for m in num2]
    batch.set_points(num2, locs2)
    batch.transform_type_1(cs2, outs)
    for t, (md, _) in enumerate(spec):
        plan = nnf.Plan(opt, 1, num2[t], locs2[t], md, dtype=RDT[prec])
        r = torch.empty_like(outs[t])
        plan.transform_type_1(cs2[t], r)
        torch.cuda.synchronize()
        if num2[t]:
            assert _rel(_to_np(outs[t]), _to_np(r)) < SAME[prec]
        else:
            assert np.all(_to_np(outs[t]) == 0)
    assert batch.info()["num_points"] == sum(num2)


def test_batch_input_errors():
    opt = nnf.Options(tol=1e-5)
    x = torch.zeros(4, device="cuda")
    with pytest.raises(nnf.InputError):
        nnf.BatchPlan(opt, -1, [4], [[x, x]], [[8, 0]])
    with pytest.raises(nnf.InputError):
        nnf.BatchPlan(opt, -1, [4, 4], [[x]], [[8]])
    b = nnf.BatchPlan(opt, -1, [4], [[x]], [[8]])
    with pytest.raises(nnf.InputError):
        b.transform_type_1([x], [])
