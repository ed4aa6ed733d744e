"""GPU tests of the autotuner (extension; no reference counterpart): it
returns one of its candidate launch shapes, caches per key (including the
point distribution) in the process and in a file later processes read, and a
plan built with the tuned shape gives the same transform as the default plan
to fp32 rounding (the tile shape changes the fixed-point scale of each
subproblem). Tuned plans against the FP64 direct sum: tests/test_gpu_clustered.py."""
import json
import os
import subprocess
import sys

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
    assert t["source"] == 0
    # cached: same answer, same time
    t2 = nnf.autotune(opt, -1, m, x, modes, types=3, reps=2)
    assert t2["source"] == 1
    t2["source"] = 0
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


HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)

# child process: tune the same seeded points against the same cache file and
# report where the answer came from
_CHILD = r"""
import json, sys
import numpy as np
import torch
sys.path.insert(0, sys.argv[1])
import neonufft_amd as nnf
kind, m = sys.argv[2], int(sys.argv[3])
sys.path.insert(0, sys.argv[1] + "/tests")
from test_gpu_autotune import _points3
x = [torch.from_numpy(v).cuda() for v in _points3(kind, m)]
t = nnf.autotune(nnf.Options(tol=1e-5), -1, m, x, (32, 32, 32), types=2, reps=2)
print(json.dumps(t))
"""


def _points3(kind, m):
    """Seeded 3D float32 points: uniform in [-pi, pi)^3, or 16 Gaussian
    clusters (sigma 0.05) folded into [-pi, pi)."""
    rng = np.random.default_rng(77)
    if kind == "uniform":
        return [rng.uniform(-np.pi, np.pi, m).astype(np.float32) for _ in range(3)]
    centres = rng.uniform(-np.pi, np.pi, (16, 3))
    which = rng.integers(0, 16, m)
    return [(np.mod(centres[which, d] + 0.05 * rng.standard_normal(m) + np.pi, 2 * np.pi)
             - np.pi).astype(np.float32) for d in range(3)]


def test_autotune_cache_keys_distribution_and_persists(tmp_path, monkeypatch):
    """Uniform and clustered points of one size and grid get separate cache
    entries (each is timed), both are written to the cache file, and a fresh
    process reads them back instead of timing again."""
    cache = tmp_path / "tune.txt"
    monkeypatch.setenv("NEONUFFT_AMD_TUNE_CACHE", str(cache))
    nnf._L.lib().nnf_autotune_clear_cache()
    m, modes, opt = 200_000, (32, 32, 32), nnf.Options(tol=1e-5)
    got = {}
    for kind in ("uniform", "clustered"):
        x = [torch.from_numpy(v).cuda() for v in _points3(kind, m)]
        t = nnf.autotune(opt, -1, m, x, modes, types=2, reps=2)
        assert t["source"] == 0, (kind, t)  # timed: the other kind's entry does not match
        got[kind] = t
    lines = [ln for ln in cache.read_text().splitlines() if ln.startswith("nnf-tune ")]
    assert len(lines) == 2, lines
    # the distribution fields (occupancy, empty-bin buckets) differ
    assert lines[0].split()[11:13] != lines[1].split()[11:13], lines
    torch.cuda.synchronize()
    for kind in ("uniform", "clustered"):
        r = subprocess.run([sys.executable, "-c", _CHILD, ROOT, kind, str(m)],
                           capture_output=True, text=True, timeout=300,
                           env=dict(os.environ, NEONUFFT_AMD_TUNE_CACHE=str(cache)))
        assert r.returncode == 0, r.stderr[-2000:]
        child = json.loads(r.stdout.strip().splitlines()[-1])
        assert child["source"] == 2, child  # read from the file
        assert child["bin_size"] == got[kind]["bin_size"]
        assert child["max_points"] == got[kind]["max_points"]
