"""GPU parity of the mode-array layouts of type 1 and type 2 plans.

The reference reads type-2 modes and writes type-1 modes with caller strides
in either mode order (src/neonufft/gpu/plan_impl.hpp:91-136, upsample_kernel.cu:18-154,
downsample_kernel.cu:24-191): NEONUFFT_MODE_ORDER_CMCL puts k = -N/2 first,
NEONUFFT_MODE_ORDER_FFT puts k = 0 first and the negative modes last. Every case
here runs through the C ABI with a non-contiguous layout (padded rows and
planes, or a permuted axis order) and is checked against the FP64 direct NUDFT
and against the FP64 restatement of the same algorithm run on the contiguous
array.
"""
import zlib

import numpy as np
import pytest

torch = pytest.importorskip("torch")

import neonufft_amd as nnf
from oracle import oracle as O

from test_gpu_plan import CDT, NPC, RDT, alg_bound, amplification, bound, cplx, dev, points


def seed_of(*parts):
    """Seed that does not depend on the interpreter's string hashing."""
    return zlib.crc32(repr(parts).encode())


def layout_strides(modes, kind):
    """Element strides of a mode array: 'contig' (x fastest), 'padded'
    (x fastest, rows / planes padded by 3 / 2 elements) or 'permuted' (the
    last axis fastest, the reverse of the default; in 1D, where one axis has
    one order, every third element)."""
    dim = len(modes)
    if kind == "contig":
        s, acc = [], 1
        for m in modes:
            s.append(acc)
            acc *= m
        return s
    if kind == "padded":
        pad = [3, 2, 0]
        s, acc = [], 1
        for d, m in enumerate(modes):
            s.append(acc)
            acc *= m + (pad[d] if d < dim - 1 else 0)
        return s
    if dim == 1:
        return [3]
    s, acc = [0] * dim, 1
    for d in reversed(range(dim)):
        s[d] = acc
        acc *= modes[d]
    return s


def flat_offsets(modes, strides):
    """Memory offset of each element of the logical (x fastest) mode array."""
    idx = np.zeros(int(np.prod(modes)), dtype=np.int64)
    rem = np.arange(idx.size, dtype=np.int64)
    for d, m in enumerate(modes):
        idx += (rem % m) * strides[d]
        rem //= m
    return idx


def to_cmcl(f, modes, order_fft):
    """Logical array (x fastest) in the given mode order -> CMCL order."""
    if not order_fft:
        return f
    a = f.reshape(tuple(reversed(modes)))
    for ax, m in enumerate(reversed(modes)):
        a = np.roll(a, m // 2, axis=ax)
    return a.ravel()


CASES = [
    # (dim, modes, M, tol, sign)
    (1, (37,), 400, 1e-5, 1),
    (1, (64,), 500, 1e-6, -1),
    (2, (24, 17), 800, 1e-5, -1),
    (2, (64, 64), 1500, 1e-5, 1),
    (3, (12, 9, 10), 1200, 1e-5, 1),
    (3, (16, 16, 16), 2000, 1e-4, -1),
]


def _cid(c):
    return f"d{c[0]}_m{'x'.join(map(str, c[1]))}_n{c[2]}_tol{c[3]:g}_s{c[4]}"


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("order", [nnf.ModeOrder.FFT, nnf.ModeOrder.CMCL])
@pytest.mark.parametrize("kind", ["padded", "permuted"])
@pytest.mark.parametrize("case", CASES, ids=[_cid(c) for c in CASES])
def test_type2_layouts(case, kind, order, prec):
    dim, modes, m, tol, sign = case
    if prec == "f64":
        tol = tol * 1e-3
    rng = np.random.default_rng(seed_of(case, kind, int(order), prec))
    x = points(rng, dim, m, prec)
    nm = int(np.prod(modes))
    f = cplx(rng, nm, prec)  # logical array in `order`, x fastest
    strides = layout_strides(modes, kind)
    off = flat_offsets(modes, strides)
    # junk everywhere the layout leaves a gap: reading it would show
    buf = np.full(int(off.max()) + 1, 1e3 + 1e3j, dtype=NPC[prec])
    buf[off] = f
    opt = nnf.Options(tol=tol, order=order)
    plan = nnf.Plan(opt, sign, m, [dev(v, RDT[prec]) for v in x], modes)
    out = torch.zeros(m, dtype=CDT[prec], device="cuda")
    plan.transform_type_2(dev(buf, CDT[prec]), strides, out)
    torch.cuda.synchronize()
    got = out.cpu().numpy().astype(np.complex128)
    offt = order == nnf.ModeOrder.FFT
    xs = [v.astype(np.float64) for v in x]
    f64 = f.astype(np.complex128)
    ref = O.direct_type2(xs, to_cmcl(f64, modes, offt), modes, sign)
    alg = O.nufft2(xs, f64, modes, sign, tol, order_fft=offt)
    alg_err = O.rel_l2(alg, ref)
    amp = amplification(modes, tol, 2.0)
    assert alg_err < 20 * tol
    err = O.rel_l2(got, ref)
    assert err < bound(prec, alg_err, tol, amp), (err, alg_err)
    assert O.rel_l2(got, alg) < alg_bound(prec, tol, amp)


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("order", [nnf.ModeOrder.FFT, nnf.ModeOrder.CMCL])
@pytest.mark.parametrize("kind", ["padded", "permuted"])
@pytest.mark.parametrize("case", CASES, ids=[_cid(c) for c in CASES])
def test_type1_layouts(case, kind, order, prec):
    dim, modes, m, tol, sign = case
    if prec == "f64":
        tol = tol * 1e-3
    rng = np.random.default_rng(seed_of(case, kind, int(order), prec, 1))
    x = points(rng, dim, m, prec)
    c = cplx(rng, m, prec)
    strides = layout_strides(modes, kind)
    off = flat_offsets(modes, strides)
    sentinel = 7.0 - 5.0j
    out = torch.full((int(off.max()) + 1,), sentinel, dtype=CDT[prec], device="cuda")
    opt = nnf.Options(tol=tol, order=order)
    plan = nnf.Plan(opt, sign, m, [dev(v, RDT[prec]) for v in x], modes)
    plan.transform_type_1(dev(c, CDT[prec]), out, strides)
    torch.cuda.synchronize()
    buf = out.cpu().numpy().astype(np.complex128)
    gaps = np.ones(buf.size, dtype=bool)
    gaps[off] = False
    assert np.all(buf[gaps] == sentinel), "wrote outside the caller's layout"
    got = buf[off]
    offt = order == nnf.ModeOrder.FFT
    xs = [v.astype(np.float64) for v in x]
    kk = O.mode_vectors(modes, np.arange(int(np.prod(modes))), order_fft=offt)
    ref = O.direct_type1(xs, c.astype(np.complex128), kk, sign)
    alg = O.nufft1(xs, c.astype(np.complex128), modes, sign, tol, order_fft=offt)
    alg_err = O.rel_l2(alg, ref)
    amp = amplification(modes, tol, 2.0)
    assert alg_err < 20 * tol
    err = O.rel_l2(got, ref)
    assert err < bound(prec, alg_err, tol, amp), (err, alg_err)
    assert O.rel_l2(got, alg) < alg_bound(prec, tol, amp)


def test_to_cmcl_matches_mode_vectors():
    """The helper's FFT -> CMCL reordering agrees with the oracle's mode
    indexing (CPU check of the test itself)."""
    modes = (5, 4, 3)
    idx = np.arange(60)
    k_fft = O.mode_vectors(modes, idx, order_fft=True)
    k_cm = O.mode_vectors(modes, idx, order_fft=False)
    lin = {tuple(k): i for i, k in enumerate(k_cm)}
    vals = np.arange(60).astype(np.complex128)
    moved = to_cmcl(vals, modes, True)
    for i, k in enumerate(k_fft):
        assert moved[lin[tuple(k)]] == vals[i]
