"""The reference's own C++ interface, run on the GPU.

tests/cpp/dropin_example.cpp is written against the reference's public GPU
API (include/neonufft/gpu/plan.hpp:25-82 of the reference) and built as
tests/cpp/libdropin.so against this library. Its extern "C" hooks drive
neonufft::gpu::Plan<float,3> (constructor, transform_type_1/2, set_points),
neonufft::gpu::PlanT3<double,3> (bounds constructor, set_input_points,
set_output_points, add_input twice, transform, reset) and
neonufft::gpu::PlanT3<float,2> (points constructor, batch of two with bdist)
on device buffers owned by this test. Every result is checked against the
FP64 direct NUDFT and the FP64 restatement (oracle/).
"""
import ctypes
import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")

import neonufft_amd as nnf  # noqa: F401  (loads libneonufft_amd.so first)
from oracle import oracle as O

from test_gpu_plan import cplx

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "cpp", "libdropin.so")


@pytest.fixture(scope="module")
def hooks():
    assert os.path.exists(LIB), "tests/cpp/libdropin.so is not built (run build())"
    nnf._L.lib()
    return ctypes.CDLL(LIB)


def _ptrs(ts):
    arr = (ctypes.c_void_p * len(ts))(*[t.data_ptr() for t in ts])
    return arr


def _msg():
    return ctypes.create_string_buffer(512)


def test_dropin_plan_f32_3d(hooks):
    rng = np.random.default_rng(71)
    modes = (20, 16, 12)
    nm = int(np.prod(modes))
    ma, mb, tol, sign = 3000, 1800, 1e-5, -1
    xa = [rng.uniform(-np.pi, np.pi, ma).astype(np.float32) for _ in range(3)]
    xb = [rng.uniform(-7.0, 7.0, mb).astype(np.float32) for _ in range(3)]
    c1, c2 = cplx(rng, ma, "f32"), cplx(rng, mb, "f32")
    fin = cplx(rng, nm, "f32")
    dxa = [torch.from_numpy(v).cuda() for v in xa]
    dxb = [torch.from_numpy(v).cuda() for v in xb]
    dc1, dc2, dfin = (torch.from_numpy(v).cuda() for v in (c1, c2, fin))
    f1 = torch.zeros(nm, dtype=torch.complex64, device="cuda")
    f2 = torch.zeros_like(f1)
    cout = torch.zeros(ma, dtype=torch.complex64, device="cuda")
    torch.cuda.synchronize()
    md = (ctypes.c_int64 * 3)(*modes)
    msg = _msg()
    rc = hooks.dropin_plan_f32_3d(ctypes.c_int64(ma), _ptrs(dxa), ctypes.c_int64(mb), _ptrs(dxb),
                                  md, ctypes.c_double(tol), sign,
                                  ctypes.c_void_p(dc1.data_ptr()), ctypes.c_void_p(f1.data_ptr()),
                                  ctypes.c_void_p(dfin.data_ptr()),
                                  ctypes.c_void_p(cout.data_ptr()),
                                  ctypes.c_void_p(dc2.data_ptr()), ctypes.c_void_p(f2.data_ptr()),
                                  msg, 512)
    assert rc == 0, msg.value.decode()
    kk = O.mode_vectors(modes, np.arange(nm))
    for x, c, f in ((xa, c1, f1), (xb, c2, f2)):
        xs = [v.astype(np.float64) for v in x]
        ref = O.direct_type1(xs, c.astype(np.complex128), kk, sign)
        alg = O.nufft1(xs, c.astype(np.complex128), modes, sign, tol)
        got = f.cpu().numpy().astype(np.complex128)
        assert O.rel_l2(got, ref) < 1.2 * O.rel_l2(alg, ref) + 2e-5
        assert O.rel_l2(got, alg) < 2.5e-5
    xs = [v.astype(np.float64) for v in xa]
    ref2 = O.direct_type2(xs, fin.astype(np.complex128), modes, sign)
    alg2 = O.nufft2(xs, fin.astype(np.complex128), modes, sign, tol)
    got2 = cout.cpu().numpy().astype(np.complex128)
    assert O.rel_l2(got2, ref2) < 1.2 * O.rel_l2(alg2, ref2) + 1e-5
    assert O.rel_l2(got2, alg2) < 2.5e-5


def test_dropin_t3_f64_3d(hooks):
    rng = np.random.default_rng(72)
    tol, sign = 1e-9, 1
    in_lo, in_hi = [-4.0, 0.0, 2.0], [3.0, 5.0, 9.0]
    out_lo, out_hi = [-3.0, -6.0, 1.0], [5.0, 2.0, 4.0]
    ma, mb, n = 900, 600, 500
    xa = [rng.uniform(in_lo[d], in_hi[d], ma) for d in range(3)]
    xb = [rng.uniform(in_lo[d], in_hi[d], mb) for d in range(3)]
    s = [rng.uniform(out_lo[d], out_hi[d], n) for d in range(3)]
    c1, c2, c3 = cplx(rng, ma, "f64"), cplx(rng, ma, "f64"), cplx(rng, mb, "f64")
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    dxa, dxb, ds = [d(v) for v in xa], [d(v) for v in xb], [d(v) for v in s]
    dc1, dc2, dc3 = d(c1), d(c2), d(c3)
    o1 = torch.zeros(n, dtype=torch.complex128, device="cuda")
    o2 = torch.zeros_like(o1)
    bounds = (ctypes.c_double * 12)(*(in_lo + in_hi + out_lo + out_hi))
    gbytes = ctypes.c_uint64()
    torch.cuda.synchronize()
    msg = _msg()
    rc = hooks.dropin_t3_f64_3d(ctypes.c_int64(ma), _ptrs(dxa), ctypes.c_int64(mb), _ptrs(dxb),
                                ctypes.c_int64(n), _ptrs(ds), bounds, ctypes.c_double(tol), sign,
                                ctypes.c_void_p(dc1.data_ptr()), ctypes.c_void_p(dc2.data_ptr()),
                                ctypes.c_void_p(dc3.data_ptr()), ctypes.c_void_p(o1.data_ptr()),
                                ctypes.c_void_p(o2.data_ptr()), ctypes.byref(gbytes), msg, 512)
    assert rc == 0, msg.value.decode()
    bnd = (in_lo, in_hi, out_lo, out_hi)
    ns, _ = O.kernel_params(tol, 2.0)
    gi = O.GridInfoT3.make(ns, 2.0, 0.1, *bnd)
    assert gbytes.value == 16 * (int(np.prod(gi.spread_grid_size))
                                 + 2 * int(np.prod(gi.fft_grid_size)))
    for x, c, o in ((xa, c1 + c2, o1), (xb, c3, o2)):
        ref = O.direct_type3(x, c, s, sign)
        alg = O.nufft3(x, c, s, sign, tol, bounds=bnd)
        got = o.cpu().numpy()
        assert O.rel_l2(alg, ref) < 20 * tol
        assert O.rel_l2(got, ref) < 1.2 * O.rel_l2(alg, ref) + 1e-12
        assert O.rel_l2(got, alg) < 1e-9


def test_dropin_t3_f32_2d_batch(hooks):
    rng = np.random.default_rng(73)
    tol, sign, m, n = 1e-5, -1, 700, 400
    x = [rng.uniform(-3.0, 8.0, m).astype(np.float32) for _ in range(2)]
    s = [rng.uniform(-5.0, 5.0, n).astype(np.float32) for _ in range(2)]
    cb = [cplx(rng, m, "f32") for _ in range(2)]
    buf = np.full(2 * (m + 1), 1e6, dtype=np.complex64)
    buf[:m], buf[m + 1:2 * m + 1] = cb[0], cb[1]
    dx = [torch.from_numpy(v).cuda() for v in x]
    dsx = [torch.from_numpy(v).cuda() for v in s]
    dbuf = torch.from_numpy(buf).cuda()
    out = torch.full((2 * (n + 1),), 9.0 + 0j, dtype=torch.complex64, device="cuda")
    torch.cuda.synchronize()
    msg = _msg()
    rc = hooks.dropin_t3_f32_2d(ctypes.c_int64(m), _ptrs(dx), ctypes.c_int64(n), _ptrs(dsx),
                                ctypes.c_double(tol), sign, ctypes.c_void_p(dbuf.data_ptr()),
                                ctypes.c_void_p(out.data_ptr()), msg, 512)
    assert rc == 0, msg.value.decode()
    o = out.cpu().numpy().astype(np.complex128)
    assert o[n] == 9.0 and o[2 * n + 1] == 9.0, "wrote into the gap between batch members"
    xs = [v.astype(np.float64) for v in x]
    ss = [v.astype(np.float64) for v in s]
    for b in range(2):
        got = o[b * (n + 1):b * (n + 1) + n]
        ref = O.direct_type3(xs, cb[b].astype(np.complex128), ss, sign)
        alg = O.nufft3(xs, cb[b].astype(np.complex128), ss, sign, tol)
        assert O.rel_l2(got, ref) < 1.2 * O.rel_l2(alg, ref) + 2e-5
        assert O.rel_l2(got, alg) < 5e-5


def test_dropin_input_error(hooks):
    msg = _msg()
    rc = hooks.dropin_input_error(msg, 512)
    assert rc == 5, (rc, msg.value.decode())  # NEONUFFT_INPUT_ERROR


def _host_ptrs(arrs):
    return (ctypes.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])


def _hp(a):
    return ctypes.c_void_p(a.ctypes.data)


def test_dropin_host_plan_f64_2d(hooks):
    """neonufft::Plan<double, 2> on host arrays (the reference's CPU plan
    interface): strided type-1 output and type-2 input, then set_points."""
    rng = np.random.default_rng(74)
    modes = (24, 17)
    nm = modes[0] * modes[1]
    ma, mb, tol, sign = 900, 500, 1e-9, 1
    xa = [rng.uniform(-9.0, 4.0, ma) for _ in range(2)]
    xb = [rng.uniform(-np.pi, np.pi, mb) for _ in range(2)]
    c1, c2 = cplx(rng, ma, "f64"), cplx(rng, mb, "f64")
    fin = cplx(rng, nm, "f64")
    # type-1 output padded (rows of 27), type-2 input transposed (y fastest)
    st1, st2 = (1, 27), (17, 1)
    off1 = np.array([(k % 24) * st1[0] + (k // 24) * st1[1] for k in range(nm)])
    off2 = np.array([(k % 24) * st2[0] + (k // 24) * st2[1] for k in range(nm)])
    sentinel = 7.0 - 5.0j
    f1 = np.full(int(off1.max()) + 1, sentinel, dtype=np.complex128)
    fin_buf = np.zeros(nm, dtype=np.complex128)
    fin_buf[off2] = fin
    cout = np.zeros(ma, dtype=np.complex128)
    f2 = np.zeros(nm, dtype=np.complex128)
    md = (ctypes.c_int64 * 2)(*modes)
    s1 = (ctypes.c_int64 * 2)(*st1)
    s2 = (ctypes.c_int64 * 2)(*st2)
    msg = _msg()
    rc = hooks.dropin_host_plan_f64_2d(ctypes.c_int64(ma), _host_ptrs(xa), ctypes.c_int64(mb),
                                       _host_ptrs(xb), md, ctypes.c_double(tol), sign, _hp(c1),
                                       _hp(f1), s1, _hp(fin_buf), s2, _hp(cout), _hp(c2), _hp(f2),
                                       msg, 512)
    assert rc == 0, msg.value.decode()
    gaps = np.ones(f1.size, dtype=bool)
    gaps[off1] = False
    assert np.all(f1[gaps] == sentinel), "wrote outside the strided positions"
    kk = O.mode_vectors(modes, np.arange(nm))
    for x, c, got in ((xa, c1, f1[off1]), (xb, c2, f2)):
        ref = O.direct_type1(x, c, kk, sign)
        alg = O.nufft1(x, c, modes, sign, tol)
        assert O.rel_l2(alg, ref) < 20 * tol
        assert O.rel_l2(got, ref) < 1.2 * O.rel_l2(alg, ref) + 1e-12
        assert O.rel_l2(got, alg) < 1e-10
    ref2 = O.direct_type2(xa, fin, modes, sign)
    alg2 = O.nufft2(xa, fin, modes, sign, tol)
    assert O.rel_l2(cout, ref2) < 1.2 * O.rel_l2(alg2, ref2) + 1e-12
    assert O.rel_l2(cout, alg2) < 1e-10


def test_dropin_host_t3_f32_3d(hooks):
    """neonufft::PlanT3<float, 3> on host arrays from bounds: batch of two
    with gaps between the members, two add_input calls accumulating, reset;
    and grid_memory_size against the reference CPU plan's formula."""
    rng = np.random.default_rng(75)
    tol, sign, m, n = 1e-5, -1, 700, 400
    in_lo, in_hi = [-3.0, 1.0, -8.0], [4.0, 6.0, -2.0]
    out_lo, out_hi = [-5.0, -2.0, 0.5], [2.0, 3.0, 6.0]
    x = [rng.uniform(in_lo[d], in_hi[d], m).astype(np.float32) for d in range(3)]
    s = [rng.uniform(out_lo[d], out_hi[d], n).astype(np.float32) for d in range(3)]
    bi, bo = m + 3, n + 2
    cs = [[cplx(rng, m, "f32") for _ in range(2)] for _ in range(3)]
    bufs = []
    for call in cs:
        b = np.full(2 * bi, 1e6 + 1e6j, dtype=np.complex64)
        b[:m], b[bi:bi + m] = call[0], call[1]
        bufs.append(b)
    out1 = np.full(2 * bo, 9.0 + 0j, dtype=np.complex64)
    out2 = np.full(2 * bo, 9.0 + 0j, dtype=np.complex64)
    bounds = np.array(in_lo + in_hi + out_lo + out_hi, dtype=np.float32)
    gbytes = ctypes.c_uint64()
    msg = _msg()
    rc = hooks.dropin_host_t3_f32_3d(ctypes.c_int64(m), _host_ptrs(x), ctypes.c_int64(n),
                                     _host_ptrs(s), _hp(bounds), ctypes.c_double(tol), sign,
                                     _hp(bufs[0]), _hp(bufs[1]), _hp(bufs[2]),
                                     ctypes.c_int64(bi), ctypes.c_int64(bo), _hp(out1), _hp(out2),
                                     ctypes.byref(gbytes), msg, 512)
    assert rc == 0, msg.value.decode()
    ns, _ = O.kernel_params(tol, 2.0)
    b64 = [bounds[0:3].astype(np.float64), bounds[3:6].astype(np.float64),
           bounds[6:9].astype(np.float64), bounds[9:12].astype(np.float64)]
    gi = O.GridInfoT3.make(ns, 2.0, 0.1, *b64)
    pad = 2 * (ns // 2 + 2)
    assert gbytes.value == 8 * (int(np.prod([v + pad for v in gi.spread_grid_size]))
                                + int(np.prod(gi.fft_grid_size)))
    for o in (out1, out2):
        assert o[n] == 9.0 and o[bo + n] == 9.0, "wrote into the gap between batch members"
    xs = [v.astype(np.float64) for v in x]
    ss = [v.astype(np.float64) for v in s]
    for b in range(2):
        for got, c in ((out1[b * bo:b * bo + n], cs[0][b].astype(np.complex128)
                        + cs[1][b].astype(np.complex128)),
                       (out2[b * bo:b * bo + n], cs[2][b].astype(np.complex128))):
            ref = O.direct_type3(xs, c, ss, sign)
            alg = O.nufft3(xs, c, ss, sign, tol, bounds=b64)
            got = got.astype(np.complex128)
            assert O.rel_l2(got, ref) < 1.2 * O.rel_l2(alg, ref) + 2e-5
            assert O.rel_l2(got, alg) < 5e-5
