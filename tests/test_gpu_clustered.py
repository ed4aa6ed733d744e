"""GPU parity on clustered 3D points: the binning / subproblem-split path.

A mixture of 16 isotropic Gaussians (sigma 0.05, centres uniform in
[-pi, pi)^3, folded periodically), the point distribution of the clustered
type-2 workload, scaled down so that the FP64 checks finish in seconds. Each
cluster is about one fine-grid cell wide, so its bins hold ~10x the
per-workgroup point cap and are split into several subproblems. Default and
autotuned (bin_size="auto") plans are compared with the FP64 direct NUDFT on
sampled outputs and with the FP64 restatement (oracle/) of the same
algorithm.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import neonufft_amd as nnf
from oracle import oracle as O

from test_gpu_plan import CDT, NPR, RDT, alg_bound, amplification, bound, cplx, dev


def clustered(rng, m, dt, k=16, sigma=0.05):
    centres = rng.uniform(-np.pi, np.pi, (k, 3))
    which = rng.integers(0, k, m)
    out = []
    for d in range(3):
        v = centres[which, d] + sigma * rng.standard_normal(m)
        out.append((np.mod(v + np.pi, 2 * np.pi) - np.pi).astype(dt))
    return out


def occupied_bins(x, n, bins):
    """Distinct bins (tile edge `bins` per dim) holding points, from the
    oracle's FP64 fold u = frac((x + pi) / 2 pi) and the containing fine
    cell floor(u n)."""
    key = np.zeros(len(x[0]), dtype=np.int64)
    mult = 1
    for d in range(3):
        cell = np.floor(O.fold(x[d].astype(np.float64)) * n[d]).astype(np.int64)
        cell = np.minimum(cell, n[d] - 1)
        b = cell // bins[d]
        key += b * mult
        mult *= (n[d] + bins[d] - 1) // bins[d]
    return len(np.unique(key))


M = 300_000
MODES = (48, 48, 48)


@pytest.fixture(scope="module")
def data():
    rng = np.random.default_rng(0xC1)
    x = clustered(rng, M, np.float32)
    f = cplx(rng, int(np.prod(MODES)), "f32")
    c = cplx(rng, M, "f32")
    pidx = np.sort(rng.choice(M, size=800, replace=False))
    midx = np.sort(rng.choice(int(np.prod(MODES)), size=600, replace=False))
    return x, f, c, pidx, midx


@pytest.mark.gpu
@pytest.mark.parametrize("tuned", [False, True], ids=["default", "auto"])
@pytest.mark.parametrize("tol", [1e-4, 1e-6])
def test_clustered_type2(data, tol, tuned):
    x, f, c, pidx, _ = data
    prec = "f32"
    opt = nnf.Options(tol=tol)
    loc = [dev(v, RDT[prec]) for v in x]
    plan = nnf.Plan(opt, -1, M, loc, MODES, bin_size="auto" if tuned else None)
    info = plan.info()
    assert info["num_subproblems"] > occupied_bins(x, info["grid"], info["bin_size"]), \
        "no bin was split into several subproblems"
    out = torch.zeros(M, dtype=CDT[prec], device="cuda")
    plan.transform_type_2(dev(f, CDT[prec]), None, out)
    torch.cuda.synchronize()
    got = out.cpu().numpy().astype(np.complex128)
    xs = [v.astype(np.float64)[pidx] for v in x]
    f64 = f.astype(np.complex128)
    ref = O.direct_type2(xs, f64, MODES, -1)
    alg = O.nufft2(xs, f64, MODES, -1, tol)
    alg_err = O.rel_l2(alg, ref)
    amp = amplification(MODES, tol, 2.0)
    assert alg_err < 20 * tol
    err = O.rel_l2(got[pidx], ref)
    assert err < bound(prec, alg_err, tol, amp), (err, alg_err)
    assert O.rel_l2(got[pidx], alg) < alg_bound(prec, tol, amp)


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("tuned", [False, True], ids=["default", "auto"])
def test_clustered_type1(data, tuned, prec):
    """Spreading of clustered points: split bins accumulate several
    subproblem tiles per bin in the tile reduction."""
    x, _, c, _, midx = data
    tol = 1e-5 if prec == "f32" else 1e-9
    xx = [v.astype(NPR[prec]) for v in x]
    cc = c.astype(np.complex64 if prec == "f32" else np.complex128)
    opt = nnf.Options(tol=tol)
    plan = nnf.Plan(opt, 1, M, [dev(v, RDT[prec]) for v in xx], MODES,
                    bin_size="auto" if tuned else None)
    out = torch.zeros(int(np.prod(MODES)), dtype=CDT[prec], device="cuda")
    plan.transform_type_1(dev(cc, CDT[prec]), out)
    torch.cuda.synchronize()
    got = out.cpu().numpy().astype(np.complex128)
    xs = [v.astype(np.float64) for v in xx]
    c64 = cc.astype(np.complex128)
    ref = O.direct_type1(xs, c64, O.mode_vectors(MODES, midx), 1)
    alg = O.nufft1(xs, c64, MODES, 1, tol)
    alg_err = O.rel_l2(alg[midx], ref)
    amp = amplification(MODES, tol, 2.0)
    assert alg_err < 20 * tol
    err = O.rel_l2(got[midx], ref)
    assert err < bound(prec, alg_err, tol, amp), (err, alg_err)
    assert O.rel_l2(got, alg) < alg_bound(prec, tol, amp)


def test_occupied_bins_counts_clusters():
    """CPU check of the helper: 16 tight clusters occupy at least 16 and far
    fewer than all bins."""
    rng = np.random.default_rng(3)
    x = clustered(rng, 50_000, np.float64)
    k = occupied_bins(x, [96, 96, 96], [32, 8, 8])
    assert 16 <= k < 3 * 12 * 12 // 2
