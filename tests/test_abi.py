"""CPU-only checks of the C ABI boundary: the library loads, exports every
symbol include/neonufft_amd.h declares, and its host-only entry points
(options, type-3 sizing, argument validation) behave like the reference's."""
import ctypes
import os
import re

import pytest

import neonufft_amd as nnf
from neonufft_amd import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "neonufft_amd.h")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|void|const char\*)\s+(nnf_[a-z0-9_]+)\s*\(", text,
                                 flags=re.M)))


def test_library_builds_and_loads():
    assert os.path.exists(L.LIB_PATH), "run python -m neonufft_amd.build"
    L.lib()


def test_every_declared_symbol_is_exported():
    syms = declared_symbols()
    assert len(syms) >= 20
    lib = ctypes.CDLL(L.LIB_PATH)
    for s in syms:
        assert hasattr(lib, s), s
    assert set(syms) == set(L.EXPORTED)


def test_exports_are_only_the_api():
    out = os.popen(f"nm -D --defined-only {L.LIB_PATH}").read()
    names = [ln.split()[-1] for ln in out.splitlines() if " T " in ln or " W " in ln]
    c_api = [n for n in names if n.startswith("nnf_")]
    assert set(c_api) == set(L.EXPORTED)
    # the C++ drop-in classes are exported too (mangled neonufft::gpu::Plan / PlanT3)
    assert any("neonufft3gpu4PlanIfLm3EE" in n for n in names)
    assert any("neonufft3gpu6PlanT3IdLm2EE" in n for n in names)


def test_options_defaults_match_reference():
    o = L.NnfOptions()
    L.lib().nnf_options_default(ctypes.byref(o))
    # include/neonufft/plan.hpp:19-29 of the reference
    assert o.tol == 0.001 and o.upsampfac == 2.0 and o.recenter_threshold == 0.1
    assert o.num_threads == 0 and o.sort_input == 1 and o.sort_output == 1
    assert o.kernel_approximation == 0 and o.order == 1 and o.kernel_type == 0
    d = nnf.Options()
    assert (d.tol, d.upsampfac, d.recenter_threshold, int(d.order)) == (0.001, 2.0, 0.1, 1)


def test_version():
    assert nnf.version().startswith("0.1.0")


def test_invalid_precision_and_dim_are_input_errors():
    lib = L.lib()
    h = ctypes.c_void_p()
    modes = (ctypes.c_int64 * 3)(8, 8, 8)
    rc = lib.nnf_plan_create(ctypes.byref(h), 7, 3, None, -1, 0, None, modes, None)
    assert rc == 5  # NEONUFFT_INPUT_ERROR
    assert "precision" in lib.nnf_last_error().decode()
    rc = lib.nnf_plan_create(ctypes.byref(h), 0, 4, None, -1, 0, None, modes, None)
    assert rc == 5
    assert lib.nnf_plan_transform_type_1(None, None, None, None) == 5
    assert "null plan" in lib.nnf_last_error().decode()


def test_t3_grid_memory_size_matches_sizing(orc):
    # the wide type-3 layout: x,y in [-1e4,1e4], z in [-50,50]; s_xy in [-0.02,0.02], s_z in [-1,1]
    opt = nnf.Options(tol=1e-6)
    size = nnf.PlanT3.grid_memory_size(opt, [-1e4, -1e4, -50], [1e4, 1e4, 50],
                                       [-0.02, -0.02, -1], [0.02, 0.02, 1], dtype="float32")
    ns, _ = orc.kernel_params(1e-6, 2.0)
    gi = orc.GridInfoT3.make(ns, 2.0, 0.1, [-1e4, -1e4, -50], [1e4, 1e4, 50],
                             [-0.02, -0.02, -1], [0.02, 0.02, 1])
    assert gi.spread_grid_size == [262, 262, 72]
    assert gi.fft_grid_size == [540, 540, 144]
    assert size == 8 * (262 * 262 * 72 + 2 * 540 * 540 * 144)
    size64 = nnf.PlanT3.grid_memory_size(opt, [-1e4, -1e4, -50], [1e4, 1e4, 50],
                                         [-0.02, -0.02, -1], [0.02, 0.02, 1], dtype="float64")
    assert size64 == 2 * size


def test_t3_invalid_bounds_is_input_error():
    opt = nnf.Options(tol=1e-4)
    with pytest.raises(nnf.InputError):
        nnf.PlanT3.grid_memory_size(opt, [1.0], [0.0], [0.0], [1.0])
    with pytest.raises(nnf.InputError):
        nnf.PlanT3.grid_memory_size(opt, [0.0], [1.0], [2.0], [1.0])


def test_cpp_dropin_example_compiles_and_links(tmp_path):
    """A program written against the reference's C++ GPU API builds against
    include/neonufft and links against libneonufft_amd.so."""
    import shutil
    import subprocess
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("no host C++ compiler")
    exe = tmp_path / "dropin"
    cmd = [cxx, "-std=c++17", f"-I{os.path.join(ROOT, 'include')}", "-I/opt/rocm/include",
           "-D__HIP_PLATFORM_AMD__", os.path.join(ROOT, "tests", "cpp", "dropin_example.cpp"),
           "-o", str(exe), f"-L{os.path.dirname(L.LIB_PATH)}", "-lneonufft_amd",
           f"-Wl,-rpath,{os.path.dirname(L.LIB_PATH)}", "-L/opt/rocm/lib", "-lamdhip64",
           "-Wl,-rpath,/opt/rocm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True)  # no arguments: no GPU calls
    assert r.returncode == 0 and "usage" in r.stdout


def test_cpp_dropin_hooks_load():
    """The drop-in caller built as a shared object (tests/cpp/Makefile, run by
    build()) loads on the CPU and exports the hooks the GPU tests call."""
    path = os.path.join(ROOT, "tests", "cpp", "libdropin.so")
    assert os.path.exists(path), "run build()"
    L.lib()
    lib = ctypes.CDLL(path)
    for s in ("dropin_plan_f32_3d", "dropin_t3_f64_3d", "dropin_t3_f32_2d", "dropin_input_error"):
        assert hasattr(lib, s), s
