"""Builds the HIP library libneonufft_amd.so in-tree for gfx950 with hipcc.

Usage: python -m neonufft_amd.build  (or neonufft_amd.build.build()).
Objects go to neonufft_amd/_build/, the shared library to neonufft_amd/.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libneonufft_amd.so")

ARCH = "gfx950"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# host-only translation units (no kernels) go through the host compiler
HOSTCXX = os.environ.get("HOSTCXX", "/opt/rocm/llvm/bin/clang++")

COMMON = [
    "-O3", "-std=c++17", "-fPIC", "-fvisibility=hidden",
    f"-I{os.path.join(ROOT, 'include')}",
    "-Wall", "-Wno-unused-function",
]
DEVICE = [
    f"--offload-arch={ARCH}",
    "-munsafe-fp-atomics",
    # keep the compiler from forming packed fp32 instructions (v_pk_*_f32)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]

SOURCES = [
    ("points.hip", True),
    ("spread_interp.hip", True),
    ("band.hip", True),
    ("grid.hip", True),
    ("pruned.hip", True),
    ("noalloc.hip", True),
    ("rows.hip", True),
    ("params.cpp", False),
    ("fft.cpp", False),
    ("plan.cpp", False),
    ("host_plan.cpp", False),
    ("autotune.cpp", False),
    ("c_api.cpp", False),
]


def _compile(src: str, device: bool, csrc: str, build_dir: str, extra) -> str:
    obj = os.path.join(build_dir, os.path.splitext(src)[0] + ".o")
    if device:
        cmd = [HIPCC] + COMMON + list(extra) + DEVICE
    else:
        cmd = [HOSTCXX] + COMMON + list(extra) + ["-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include"]
    cmd += ["-c", os.path.join(csrc, src), "-o", obj]
    t0 = time.time()
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return f"{src}: {time.time() - t0:.1f}s"


def build(verbose: bool = True, jobs: int | None = None, csrc: str = CSRC,
          build_dir: str = BUILD, lib: str = LIB, extra_flags=()) -> str:
    """Compiles every source and links the shared library. Always rebuilds.

    csrc / build_dir / lib / extra_flags let a developer build a modified copy
    of the sources elsewhere; the defaults build the package in-tree."""
    if shutil.which(HIPCC) is None and not os.path.exists(HIPCC):
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    os.makedirs(build_dir, exist_ok=True)
    jobs = jobs or min(len(SOURCES), os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = [ex.submit(_compile, s, d, csrc, build_dir, extra_flags) for s, d in SOURCES]
        for f in futs:
            msg = f.result()
            if verbose:
                print("  built", msg, flush=True)
    objs = [os.path.join(build_dir, os.path.splitext(s)[0] + ".o") for s, _ in SOURCES]
    tmp = lib + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs + [
        "-L/opt/rocm/lib", "-lrocfft", "-Wl,-rpath,/opt/rocm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, lib)
    if verbose:
        print("  linked", lib, flush=True)
    return lib


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
