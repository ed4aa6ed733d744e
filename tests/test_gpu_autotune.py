"""GPU tests of the autotuner (extension; no reference counterpart): it
returns one of its candidate launch shapes, caches per key, and a plan built
with the tuned shape gives the same transform as the default plan to fp32
rounding (the tile shape changes the fixed-point scale of each subproblem and
the order of the float atomics)."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import neonufft_amd as nnf

from test_gpu_plan import CDT, RDT, cplx, dev, points

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("dim,modes,m", [(1, (300,), 20000), (2, (64, 48), 50000),
                                         (3, (32, 24, 20), 100000)])
def test_autotune_result_and_plan(dim, modes, m):
    rng = np.random.default_rng(dim)
    x = [dev(v, torch.float32) for v in points(rng, dim, m, "f32")]
    c = dev(cplx(rng, m, "f32"), torch.complex64)
    opt = nnf.Options(tol=1e-5)
    nnf._L.lib().nnf_autotune_clear_cache()
    t = nnf.autotune(opt, -1, m, x, modes, types=3, reps=2)
    assert t["ms"] > 0.0
    assert t["max_points"] in (1024, 2048)
    # 2D plans spread with row bands: their height is tuned too
    assert (t["band_rows"] > 0) == (dim == 2)
    assert len(t["bin_size"]) == dim and all(b > 0 for b in t["bin_size"])
    # cached: same answer, same time
    t2 = nnf.autotune(opt, -1, m, x, modes, types=3, reps=2)
    assert t2 == t

    ref = nnf.Plan(opt, -1, m, x, modes)
    tuned = nnf.Plan(opt, -1, m, x, modes, bin_size="auto")
    assert tuned.info()["bin_size"] == t["bin_size"]
    if dim == 2:
        assert tuned.info()["band_rows"] == t["band_rows"]
    nm = int(np.prod(modes))
    f0 = torch.empty(nm, dtype=torch.complex64, device="cuda")
    f1 = torch.empty_like(f0)
    ref.transform_type_1(c, f0)
    tuned.transform_type_1(c, f1)
    o0 = torch.empty(m, dtype=torch.complex64, device="cuda")
    o1 = torch.empty_like(o0)
    ref.transform_type_2(f0, None, o0)
    tuned.transform_type_2(f0, None, o1)
    torch.cuda.synchronize()
    # a different tiling changes each subproblem's fixed-point scale (a bound on
    # the sum of its strengths), so type 1 agrees to fp32 rounding, not bitwise
    assert _rel(f1.cpu().numpy(), f0.cpu().numpy()) < 1e-5
    assert _rel(o1.cpu().numpy(), o0.cpu().numpy()) < 1e-5


def test_autotune_input_errors():
    x = torch.zeros(8, device="cuda")
    opt = nnf.Options(tol=1e-5)
    with pytest.raises(nnf.InputError):
        nnf.autotune(opt, -1, 8, [x], (16,), types=0)
    with pytest.raises(nnf.InputError):
        nnf.autotune(opt, -1, 8, [x], (16,), reps=0)
    with pytest.raises(nnf.InputError):
        nnf.Plan(opt, -1, 8, [x], (16,), bin_size="fast")
