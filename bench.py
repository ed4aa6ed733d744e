"""Benchmark of the 3D type-1/2 NUFFT on one MI355X.

Workload (default): 3D type 1 and type 2, complex64, M = 10,000,000 points
uniform in [-pi, pi)^3, strengths / coefficients i.i.d. complex N(0,1),
128 x 128 x 128 modes, tol 1e-5, upsampling 2.0 (fine grid 256^3). FP32 bulk
arithmetic, FP64 fold. One step = one type-1 transform (M strengths -> modes)
plus one type-2 transform (modes -> M values) on a plan whose points are set;
inputs are resident in HBM. value = NU points processed per second
(2 M per step) over the whole job.

Prints ONE JSON line. Extra keys: the spreading / interpolation kernel
rooflines (HIP events on the plan's stream), a bounded CPU baseline (the FP64
C restatement in oracle/, OpenMP), and the relative l2 error of both
transforms on 1000 sampled outputs against the FP64 direct NUDFT.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s
SEED = 0x4E554654


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--points", type=int, default=10_000_000)
    p.add_argument("--modes", type=int, default=128)
    p.add_argument("--tol", type=float, default=1e-5)
    p.add_argument("--kernel-reps", type=int, default=10)
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--cpu-sample", type=int, default=4_000_000,
                   help="points in the bounded CPU-baseline sample")
    p.add_argument("--calib-copy-mib", type=int, default=0,
                   help="also run a plain device copy of this many MiB (profiler calibration)")
    return p.parse_args()


def cpu_baseline_leg(x, c, f_gpu, out2_gpu, modes, tol, sample):
    """Checker work on the host: FP64 direct-sum error of the GPU outputs and
    the timed CPU restatement (oracle/) on a bounded sample."""
    import numpy as np
    from oracle import oracle as O

    rng = np.random.default_rng(SEED + 99)
    m = x[0].shape[0]
    xs = [v.astype(np.float64) for v in x]
    # accuracy: 1000 sampled modes (type 1) and 1000 sampled points (type 2)
    ntot = int(np.prod(modes))
    idx = np.sort(rng.choice(ntot, size=min(1000, ntot), replace=False))
    ref1 = O.direct_type1(xs, c.astype(np.complex128), O.mode_vectors(modes, idx), -1)
    err1 = O.rel_l2(f_gpu[idx], ref1)
    pidx = np.sort(rng.choice(m, size=min(1000, m), replace=False))
    ref2 = O.direct_type2([v[pidx] for v in xs], f_gpu.astype(np.complex128), modes, -1)
    err2 = O.rel_l2(out2_gpu[pidx], ref2)

    # timed CPU restatement on the first `sample` points of the same workload
    ms = min(sample, m)
    xsub = [v[:ms] for v in xs]
    csub = c[:ms].astype(np.complex128)
    t0 = time.perf_counter()
    f_cpu = O.nufft1(xsub, csub, modes, -1, tol)
    t1 = time.perf_counter()
    O.nufft2(xsub, f_cpu, modes, -1, tol)
    t2 = time.perf_counter()
    cpu = {
        "value": 2.0 * ms / (t2 - t0),
        "unit": "NU pts/s",
        "cores": O.num_threads(),
        "kind": "port",
        "sample": (f"FP64 C restatement (oracle/nufft_oracle.c, OpenMP) + numpy FFT, "
                   f"type 1 + type 2 on the first {ms} points of the workload, same grid; "
                   f"type1 {t1 - t0:.2f}s type2 {t2 - t1:.2f}s"),
    }
    return cpu, err1, err2


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU only: --gpus must be 1 "
                          "(multi-GPU sharding is a design note, not a measured path)"}))
        sys.exit(2)

    import numpy as np
    import torch

    import neonufft_amd as nnf

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    m = args.points
    N = args.modes
    modes = (N, N, N)
    rng = np.random.default_rng(SEED + 1)
    x = [rng.uniform(-math.pi, math.pi, m).astype(np.float32) for _ in range(3)]
    c = (rng.standard_normal(m, dtype=np.float32)
         + 1j * rng.standard_normal(m, dtype=np.float32)).astype(np.complex64)
    dx = [torch.from_numpy(v).to(dev) for v in x]
    dc = torch.from_numpy(c).to(dev)
    f = torch.empty(N ** 3, dtype=torch.complex64, device=dev)
    out2 = torch.empty(m, dtype=torch.complex64, device=dev)
    stream = torch.cuda.current_stream()

    opt = nnf.Options(tol=args.tol, upsampfac=2.0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plan = nnf.Plan(opt, -1, m, dx, modes, stream=stream)
    torch.cuda.synchronize()
    create_ms = (time.perf_counter() - t0) * 1e3
    t0 = time.perf_counter()
    plan.set_points(m, dx)
    torch.cuda.synchronize()
    set_points_ms = (time.perf_counter() - t0) * 1e3
    info = plan.info()

    def step():
        plan.transform_type_1(dc, f)
        plan.transform_type_2(f, None, out2)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ms_step = dt / args.steps * 1e3
    value = 2.0 * m * args.steps / dt

    # per-transform times
    def timed(fn, reps):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        fn()
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        e1.synchronize()
        return e0.elapsed_time(e1) / reps

    reps = args.kernel_reps
    t1_ms = timed(lambda: plan.transform_type_1(dc, f), reps)
    t2_ms = timed(lambda: plan.transform_type_2(f, None, out2), reps)
    # dominant kernels alone: spread (accumulating, no zero fill) and interp
    spread_ms = timed(lambda: plan.spread(dc, accumulate=True), reps)
    interp_ms = timed(lambda: plan.interp(out2), reps)

    # restore valid outputs for the accuracy check
    plan.transform_type_1(dc, f)
    plan.transform_type_2(f, None, out2)
    torch.cuda.synchronize()

    if args.calib_copy_mib:
        a = torch.empty(args.calib_copy_mib * (1 << 20) // 4, dtype=torch.float32, device=dev)
        b = torch.empty_like(a)
        a.fill_(1.0)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()

    G = 1
    for n in info["grid"]:
        G *= n
    alg_bytes = m * (3 * 4 + 8) + G * 8  # coords + strength, fine grid once
    spread_gbs = alg_bytes / (spread_ms * 1e-3)
    interp_gbs = alg_bytes / (interp_ms * 1e-3)
    dom = ("spread_kernel", spread_ms, spread_gbs) if spread_ms >= interp_ms else \
        ("interp_kernel", interp_ms, interp_gbs)

    f_host = f.cpu().numpy()
    out2_host = out2.cpu().numpy()

    cpu = None
    err1 = err2 = None
    if not args.no_cpu_baseline:
        cpu, err1, err2 = cpu_baseline_leg(x, c, f_host, out2_host, modes, args.tol,
                                           args.cpu_sample)

    traffic = None
    tfile = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tfile):
        try:
            tj = json.load(open(tfile))
            traffic = tj.get(dom[0], {}).get("hbm_bytes_per_launch")
        except Exception:
            traffic = None

    line = {
        "metric": "NU pts/s, 3D type-1/2 fp32 @1e-5 vs fp64 oracle; batched vs plain, 1 GPU",
        "value": value,
        "unit": "NU pts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded numpy: uniform points, complex N(0,1) strengths)",
        "config": {
            "workload": f"3D type-1 + type-2, complex64, M={m}, modes {N}^3, tol {args.tol:g}, "
                        f"upsampfac 2.0, fine grid {'x'.join(map(str, info['grid']))}",
            "points": m, "modes": [N, N, N], "tol": args.tol, "upsampfac": 2.0,
            "n_spread": info["n_spread"], "bin_size": info["bin_size"],
            "subproblems": info["num_subproblems"],
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dom[0],
            "achieved": dom[2] / 1e9,
            "peak": HBM_PEAK / 1e9,
            "unit": "GB/s",
            "frac": dom[2] / HBM_PEAK,
            "traffic": traffic,
            "alg_bytes_per_launch": alg_bytes,
            "avg_launch_ms": dom[1],
        },
        "kernels": {
            "spread_ms": spread_ms, "spread_GBps": spread_gbs / 1e9,
            "interp_ms": interp_ms, "interp_GBps": interp_gbs / 1e9,
        },
        "stages": {
            "type1_ms": t1_ms, "type2_ms": t2_ms, "set_points_ms": set_points_ms,
            "plan_create_ms": create_ms,
            "type1_pts_per_s": m / (t1_ms * 1e-3), "type2_pts_per_s": m / (t2_ms * 1e-3),
        },
        "accuracy": {"type1_rel_l2": err1, "type2_rel_l2": err2,
                     "samples": 1000, "reference": "FP64 direct NUDFT"},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
