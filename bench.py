"""Benchmark of the 3D type-1/2 NUFFT on one MI355X.

Workload (default): 3D type 1 and type 2, complex64, M = 10,000,000 points
uniform in [-pi, pi)^3, strengths / coefficients i.i.d. complex N(0,1),
128 x 128 x 128 modes, tol 1e-5, upsampling 2.0 (fine grid 256^3). FP32 bulk
arithmetic, FP64 fold. One step = one type-1 transform (M strengths -> modes)
plus one type-2 transform (modes -> M values) on a plan whose points are set;
inputs are resident in HBM. value = NU points processed per second
(2 M per step) over the whole job.

Prints ONE JSON line. Extra keys: the roofline of the slower of the spread
stage (spreading + tile reduction) and interpolation (HIP events on the plan's
stream) against HBM and against LDS, with the rocprofv3 counters committed
under profiles/; a bounded CPU baseline (the FP64 C restatement in oracle/,
OpenMP) with the host's core count and CPU model; the relative l2 error of
both transforms on 1000 sampled outputs against the FP64 direct NUDFT, beside
the error of the same algorithm in FP64 on the same samples; and three more
workloads, each with its own accuracy and CPU baseline:
  batched: 256 independent 2D type-1 transforms (complex64, point counts
    log-uniform in [1e4, 1e6], modes drawn from {256^2, 512^2, 1024^2}, tol
    1e-5) through one BatchPlan (HIP graph) against 256 plain plans;
  clustered: 3D type 2, 2e7 points in 16 Gaussian clusters, 256^3 modes, tol
    1e-4 and 1e-6, autotuned tiles;
  type3: 3D type 3, 2e6 sources (|x| up to 1e4), 1e6 targets, tol 1e-6, in
    the mixed, fp32 and fp64 plans.
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
# LDS: 256 bytes per clock per CU (twice CDNA3's 128; 16-byte ds_read_b128 was
# measured at 221 B/clk/CU and 8-byte ds_read_b64 at 175 on this GPU, counting
# clocks at the nominal rate), 256 CUs, 2.4 GHz peak engine clock
LDS_PEAK = 256 * 256 * 2.4e9
# VALU: one wave64 instruction per clock per CU (4 SIMDs x 16 lanes), 256 CUs
VALU_PEAK = 256 * 2.4e9
SEED = 0x4E554654


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--workload", default="uniform",
                   choices=["uniform", "clustered", "type3", "batched"],
                   help="uniform: 3D type 1+2, 1e7 points, 128^3 modes (default, plus the "
                        "batched comparison); clustered: 3D type 2, 2e7 points in 16 "
                        "Gaussian clusters, 256^3 modes, tol 1e-4 and 1e-6; type3: 3D type 3, "
                        "2e6 sources, 1e6 targets, tol 1e-6; batched: the batched comparison "
                        "alone")
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--points", type=int, default=10_000_000)
    p.add_argument("--modes", type=int, default=128)
    p.add_argument("--tol", type=float, default=1e-5)
    p.add_argument("--kernel-reps", type=int, default=10)
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--cpu-sample", type=int, default=4_000_000,
                   help="points in the bounded CPU-baseline sample")
    p.add_argument("--batch-count", type=int, default=256,
                   help="transforms in the batched-uneven comparison (0 = skip it)")
    p.add_argument("--batch-reps", type=int, default=5)
    p.add_argument("--no-clustered", action="store_true",
                   help="leave the clustered type-2 leg out of the default line")
    p.add_argument("--no-type3", action="store_true",
                   help="leave the type-3 leg out of the default line")
    p.add_argument("--calib-copy-mib", type=int, default=0,
                   help="also run a plain device copy of this many MiB (profiler calibration)")
    return p.parse_args()


def host_cpu():
    """Threads the CPU legs use and what the host offers: the logical CPUs this
    process may run on, the physical cores (distinct (package, core) pairs in
    /proc/cpuinfo) and the CPU model."""
    from oracle import oracle as O
    model, cores = None, set()
    phys = core = None
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                k, _, v = line.partition(":")
                k, v = k.strip(), v.strip()
                if k == "model name" and model is None:
                    model = v
                elif k == "physical id":
                    phys = v
                elif k == "core id":
                    core = v
                elif not k and phys is not None:
                    cores.add((phys, core))
                    phys = core = None
        if phys is not None:
            cores.add((phys, core))
    except OSError:
        pass
    try:
        visible = len(os.sched_getaffinity(0))
    except AttributeError:
        visible = os.cpu_count()
    return {"cores": O.num_threads(), "physical_cores_host": len(cores) or None,
            "logical_cpus_visible": visible, "cpu_model": model}


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
    # the same algorithm in FP64 (oracle/) on the same samples: the ns-wide
    # kernel's own error, which the reference shares; the rest of the GPU
    # error is fp32 rounding (Horner fit, fixed-point sums, fp32 FFT)
    t_alg = time.perf_counter()
    alg1 = O.nufft1(xs, c.astype(np.complex128), modes, -1, tol)
    alg2 = O.nufft2([v[pidx] for v in xs], f_gpu.astype(np.complex128), modes, -1, tol)
    acc = {
        "type1_rel_l2": err1, "type2_rel_l2": err2,
        "fp64_restatement_type1_rel_l2": O.rel_l2(alg1[idx], ref1),
        "fp64_restatement_type2_rel_l2": O.rel_l2(alg2, ref2),
        "fp32_share_type1": O.rel_l2(f_gpu[idx], alg1[idx]),
        "fp32_share_type2": O.rel_l2(out2_gpu[pidx], alg2),
        "samples": 1000, "reference": "FP64 direct NUDFT",
        "restatement_s": time.perf_counter() - t_alg,
    }

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
        **host_cpu(),
        "kind": "port",
        "sample": (f"FP64 C restatement (oracle/nufft_oracle.c, OpenMP) + numpy FFT, "
                   f"type 1 + type 2 on the first {ms} points of the workload, same grid; "
                   f"type1 {t1 - t0:.2f}s type2 {t2 - t1:.2f}s"),
    }
    return cpu, acc


def batched_leg(args, dev, stream):
    """Batched uneven transforms vs the same transforms as sequential plain
    plans. Every timed call has its points set and inputs resident in HBM."""
    import numpy as np
    import torch

    import neonufft_amd as nnf

    count = args.batch_count
    rng = np.random.default_rng(SEED + 5)
    ms = np.floor(10.0 ** rng.uniform(4.0, 6.0, count)).astype(np.int64)
    nmodes = rng.choice([256, 512, 1024], size=count)
    gen = torch.Generator(device=dev)
    gen.manual_seed(SEED + 5)
    locs, cs, outs_b, outs_p = [], [], [], []
    for t in range(count):
        m = int(ms[t])
        xy = torch.rand(2, m, generator=gen, device=dev) * (2 * math.pi) - math.pi
        locs.append([xy[0], xy[1]])
        cs.append(torch.view_as_complex(torch.randn(m, 2, generator=gen, device=dev)))
        outs_b.append(torch.empty(int(nmodes[t]) ** 2, dtype=torch.complex64, device=dev))
        outs_p.append(torch.empty_like(outs_b[-1]))
    opt = nnf.Options(tol=1e-5, upsampfac=2.0)
    modes = [[int(n), int(n)] for n in nmodes]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    batch = nnf.BatchPlan(opt, -1, [int(v) for v in ms], locs, modes, stream=stream,
                          use_graph=True)
    torch.cuda.synchronize()
    batch_create_ms = (time.perf_counter() - t0) * 1e3
    plans = [nnf.Plan(opt, -1, int(ms[t]), locs[t], modes[t], stream=stream)
             for t in range(count)]
    torch.cuda.synchronize()

    def run_batched():
        batch.transform_type_1(cs, outs_b)

    def run_plain():
        for t in range(count):
            plans[t].transform_type_1(cs[t], outs_p[t])

    def timed(fn):
        for _ in range(2):
            fn()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record(stream)
        for _ in range(args.batch_reps):
            fn()
        e1.record(stream)
        e1.synchronize()
        return e0.elapsed_time(e1) / args.batch_reps

    batched_ms = timed(run_batched)
    plain_ms = timed(run_plain)
    torch.cuda.synchronize()
    info = batch.info()

    # every batched transform against its plain plan (same kernels)
    max_diff = 0.0
    for t in range(count):
        a, b = outs_b[t], outs_p[t]
        max_diff = max(max_diff, float(torch.linalg.vector_norm(a - b)
                                       / torch.linalg.vector_norm(b)))
    def direct(t, samples):
        """Relative l2 error of batched transform t on sampled modes against
        the FP64 direct sum."""
        n = int(nmodes[t])
        x = locs[t][0].double().cpu().numpy()
        y = locs[t][1].double().cpu().numpy()
        c = cs[t].cpu().numpy().astype(np.complex128)
        idx = np.sort(rng.choice(n * n, size=samples, replace=False))
        kx = (idx % n - n // 2).astype(np.float64)
        ky = (idx // n - n // 2).astype(np.float64)
        ref = np.zeros(len(idx), dtype=np.complex128)
        for j0 in range(0, len(x), 4096):
            sl = slice(j0, j0 + 4096)
            ref += np.exp(-1j * (np.outer(kx, x[sl]) + np.outer(ky, y[sl]))) @ c[sl]
        got = outs_b[t].cpu().numpy()[idx].astype(np.complex128)
        return float(np.linalg.norm(got - ref) / np.linalg.norm(ref))

    # the transform with fewest points (1000 modes) and the one with most
    # points per fine-grid cell (100 modes: the fixed-point sums are densest)
    t = int(np.argmin(ms))
    direct_err = direct(t, 1000)
    t_dense = int(np.argmax(ms / (nmodes.astype(np.float64) ** 2)))
    direct_err_dense = direct(t_dense, 100)

    total = int(ms.sum())
    out = {
        "workload": (f"{count} independent 2D type-1 transforms, complex64, point counts "
                     f"floor(10^U[4,6]) (sum {total}), modes from {{256^2, 512^2, 1024^2}}, "
                     f"tol 1e-5, upsampfac 2.0"),
        "transforms": count,
        "sum_points": total,
        "shapes": info["num_shapes"],
        "grid_cells": info["grid_cells"],
        "subproblems": info["num_subproblems"],
        "graph": info["graph"],
        "batched_ms": batched_ms,
        "plain_ms": plain_ms,
        "speedup": plain_ms / batched_ms,
        "batched_pts_per_s": total / (batched_ms * 1e-3),
        "plain_pts_per_s": total / (plain_ms * 1e-3),
        "batch_create_ms": batch_create_ms,
        "max_rel_diff_vs_plain": max_diff,
        "direct_rel_l2": direct_err,
        "direct_rel_l2_densest": direct_err_dense,
        "direct_check_densest": (f"100 sampled modes of transform {t_dense} "
                                 f"(M={int(ms[t_dense])}, {int(nmodes[t_dense])}^2 modes)"),
        "direct_check": (f"1000 sampled modes of transform {t} (M={int(ms[t])}, "
                         f"{int(nmodes[t])}^2 modes)"),
    }
    del plans, batch
    torch.cuda.synchronize()
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU only: --gpus must be 1 "
                          "(multi-GPU sharding is a design note, not a measured path)"}))
        sys.exit(2)
    if args.workload == "clustered":
        return run_clustered(args)
    if args.workload == "type3":
        return run_type3(args)
    if args.workload == "batched":
        return run_batched(args)
    return run_uniform(args)


def _timed(stream, fn, reps):
    import torch
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    fn()
    e0.record(stream)
    for _ in range(reps):
        fn()
    e1.record(stream)
    e1.synchronize()
    return e0.elapsed_time(e1) / reps


PROFILE_COUNTERS = "profiles/counters_3d_c64_M1e7_modes128.json"


def _profile_counters():
    """Per-dispatch rocprofv3 counters of the default workload's kernels
    (HBM bytes from FETCH_SIZE / WRITE_SIZE scaled by a copy of known size,
    LDS busy and bank-conflict fractions), committed under profiles/."""
    path = os.path.join(ROOT, PROFILE_COUNTERS)
    if not os.path.exists(path):
        return None
    try:
        with open(path) as fh:
            return json.load(fh)["kernels"]
    except (OSError, ValueError, KeyError):
        return None


def _common(args, metric, value, ms_step, dtype, config):
    return {
        "metric": metric, "value": value, "unit": "NU pts/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": dtype,
        "data": "synthetic (seeded)", "config": config,
    }


def _roofline(kernel, alg_bytes, ms, traffic=None):
    gbs = alg_bytes / (ms * 1e-3)
    return {"bound": "hbm", "kernel": kernel, "achieved": gbs / 1e9, "peak": HBM_PEAK / 1e9,
            "unit": "GB/s", "frac": gbs / HBM_PEAK, "traffic": traffic,
            "alg_bytes_per_launch": alg_bytes, "avg_launch_ms": ms}


def clustered_points(rng, m):
    """Mixture of 16 isotropic Gaussians (sigma 0.05, centres uniform in
    [-pi, pi)^3), folded periodically into [-pi, pi)."""
    import numpy as np
    centres = rng.uniform(-math.pi, math.pi, (16, 3))
    which = rng.integers(0, 16, m)
    out = []
    for d in range(3):
        v = centres[which, d] + 0.05 * rng.standard_normal(m)
        v = np.mod(v + math.pi, 2 * math.pi) - math.pi
        out.append(v.astype(np.float32))
    return out


def cpu_leg_type2(x, f, out_gpu, modes, tol, sample, nacc):
    """Checker work on the host for a type-2 workload: FP64 direct-sum error
    on `nacc` sampled points and the timed CPU restatement (oracle/) on the
    first `sample` points."""
    import numpy as np
    from oracle import oracle as O

    rng = np.random.default_rng(SEED + 98)
    m = x[0].shape[0]
    xs = [v.astype(np.float64) for v in x]
    f64 = f.astype(np.complex128)
    ms = min(sample, m)
    # accuracy samples among the points the timed restatement covers, so its
    # output gives the same algorithm's FP64 error on the same samples
    pidx = np.sort(rng.choice(ms, size=min(nacc, ms), replace=False))
    ref = O.direct_type2([v[pidx] for v in xs], f64, modes, -1)
    err = O.rel_l2(out_gpu[pidx], ref)
    t0 = time.perf_counter()
    alg = O.nufft2([v[:ms] for v in xs], f64, modes, -1, tol)
    t1 = time.perf_counter()
    cpu = {"value": ms / (t1 - t0), "unit": "NU pts/s", **host_cpu(),
           "kind": "port",
           "sample": (f"FP64 C restatement (oracle/, OpenMP) + numpy FFT, type 2 on the first "
                      f"{ms} points of the workload, same grid: {t1 - t0:.2f}s")}
    return cpu, err, O.rel_l2(alg[pidx], ref)


def clustered_leg(args):
    """Clustered workload: 3D type 2 on 2e7 clustered points, 256^3 modes,
    tol 1e-4 and 1e-6, tiles chosen by the autotuner."""
    import numpy as np
    import torch

    import neonufft_amd as nnf

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    stream = torch.cuda.current_stream()
    m = 20_000_000
    N = 256
    modes = (N, N, N)
    rng = np.random.default_rng(SEED + 3)
    x = clustered_points(rng, m)
    f = (rng.standard_normal(N ** 3, dtype=np.float32)
         + 1j * rng.standard_normal(N ** 3, dtype=np.float32)).astype(np.complex64)
    dx = [torch.from_numpy(v).to(dev) for v in x]
    df = torch.from_numpy(f).to(dev)
    out = torch.empty(m, dtype=torch.complex64, device=dev)
    results = []
    for tol in (1e-4, 1e-6):
        opt = nnf.Options(tol=tol, upsampfac=2.0)
        t0 = time.perf_counter()
        tune = nnf.autotune(opt, -1, m, dx, modes, types=2, reps=3, stream=stream)
        tune_s = time.perf_counter() - t0
        plan = nnf.Plan(opt, -1, m, dx, modes, stream=stream, bin_size=tune["bin_size"],
                        max_points=tune["max_points"])
        base = nnf.Plan(opt, -1, m, dx, modes, stream=stream)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        plan.set_points(m, dx)
        torch.cuda.synchronize()
        set_ms = (time.perf_counter() - t0) * 1e3
        step = lambda: plan.transform_type_2(df, None, out)  # noqa: E731
        for _ in range(args.warmup):
            step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step()
        torch.cuda.synchronize()
        ms_step = (time.perf_counter() - t0) / args.steps * 1e3
        interp_ms = _timed(stream, lambda: plan.interp(out), args.kernel_reps)
        base_interp_ms = _timed(stream, lambda: base.interp(out), args.kernel_reps)
        plan.transform_type_2(df, None, out)
        torch.cuda.synchronize()
        info = plan.info()
        G = int(np.prod(info["grid"]))
        results.append({
            "tol": tol, "n_spread": info["n_spread"], "ms_per_step": ms_step,
            "pts_per_s": m / (ms_step * 1e-3), "set_points_ms": set_ms,
            "interp_ms": interp_ms, "default_shape_interp_ms": base_interp_ms,
            "tuned": {"bin_size": tune["bin_size"], "max_points": tune["max_points"],
                      "tuning_s": tune_s},
            "subproblems": info["num_subproblems"],
            "roofline": _roofline("interp_kernel", m * (3 * 4 + 8) + G * 8, interp_ms),
            "out": out.cpu().numpy() if not args.no_cpu_baseline else None,
        })
        del plan, base
    cpu = None
    for r in results:
        o = r.pop("out")
        if o is not None:
            c, err, alg_err = cpu_leg_type2(x, f, o, modes, r["tol"], args.cpu_sample // 2, 200)
            r["rel_l2_vs_direct"] = err
            r["fp64_restatement_rel_l2"] = alg_err
            r["accuracy_samples"] = 200
            if cpu is None:
                cpu = c
    del dx, df, out
    torch.cuda.synchronize()
    return {"workload": "3D type-2, complex64, M=2e7 points in 16 Gaussian clusters "
                        "(sigma 0.05, folded), modes 256^3 (fine grid 512^3), "
                        "tol 1e-4 and 1e-6, autotuned tiles",
            "points": m, "modes": list(modes), "sweep": results, "cpu_baseline": cpu}


def run_clustered(args):
    leg = clustered_leg(args)
    head = leg["sweep"][0]
    line = _common(args, "NU pts/s, 3D type-2 fp32, 2e7 clustered points, 256^3 modes, 1 GPU",
                   head["pts_per_s"], head["ms_per_step"], "f32",
                   {"workload": leg["workload"], "points": leg["points"],
                    "modes": leg["modes"]})
    line["roofline"] = head["roofline"]
    line["sweep"] = leg["sweep"]
    line["cpu_baseline"] = leg["cpu_baseline"]
    print(json.dumps(line), flush=True)


def type3_leg(args):
    """Type-3 workload: 3D type 3, 2e6 sources, 1e6 targets, tol 1e-6, in
    the mixed, fp32 and fp64 plans."""
    import numpy as np
    import torch

    import neonufft_amd as nnf

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    stream = torch.cuda.current_stream()
    m, n = 2_000_000, 1_000_000
    rng = np.random.default_rng(SEED + 4)
    x = [rng.uniform(-1e4, 1e4, m), rng.uniform(-1e4, 1e4, m), rng.uniform(-50, 50, m)]
    s = [rng.uniform(-0.02, 0.02, n), rng.uniform(-0.02, 0.02, n), rng.uniform(-1, 1, n)]
    c = rng.standard_normal(m) + 1j * rng.standard_normal(m)
    res = {}
    outs = {}
    for prec, rdt, cdt in (("mixed", torch.float64, torch.complex128),
                           ("f32", torch.float32, torch.complex64),
                           ("f64", torch.float64, torch.complex128)):
        dx = [torch.from_numpy(v).to(dev, rdt) for v in x]
        ds = [torch.from_numpy(v).to(dev, rdt) for v in s]
        dc = torch.from_numpy(c).to(dev, cdt)
        out = torch.empty(n, dtype=cdt, device=dev)
        opt = nnf.Options(tol=1e-6, upsampfac=2.0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        plan = nnf.PlanT3(opt, -1, m, dx, n, ds, stream=stream,
                          dtype="mixed" if prec == "mixed" else None)
        torch.cuda.synchronize()
        create_ms = (time.perf_counter() - t0) * 1e3

        def step():
            plan.reset()
            plan.add_input(dc)
            plan.transform(out)

        for _ in range(args.warmup):
            step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step()
        torch.cuda.synchronize()
        ms_step = (time.perf_counter() - t0) / args.steps * 1e3
        add_ms = _timed(stream, lambda: plan.add_input(dc), args.kernel_reps)
        tr_ms = _timed(stream, lambda: plan.transform(out), args.kernel_reps)
        step()
        torch.cuda.synchronize()
        info = plan.info()
        Gs = int(np.prod(info["spread_grid"]))
        cb = 4 if prec == "f32" else 8  # bytes per coordinate and per input complex part
        gb = 16 if prec == "f64" else 8  # bytes per spread-grid cell
        res[prec] = {
            "ms_per_step": ms_step, "pts_per_s": (m + n) / (ms_step * 1e-3),
            "plan_create_ms": create_ms, "add_input_ms": add_ms, "transform_ms": tr_ms,
            "n_spread": info["n_spread"], "spread_grid": info["spread_grid"],
            "fft_grid": info["grid"],
            "roofline": _roofline("spread_kernel (add_input)",
                                  m * (3 * cb + 2 * cb) + Gs * gb, add_ms),
        }
        outs[prec] = out.cpu().numpy()
        del plan
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import oracle as O
        rng2 = np.random.default_rng(SEED + 97)
        tidx = np.sort(rng2.choice(n, size=1000, replace=False))
        for prec, dt in (("mixed", np.float64), ("f32", np.float32), ("f64", np.float64)):
            xs = [v.astype(dt).astype(np.float64) for v in x]
            ss = [v[tidx].astype(dt).astype(np.float64) for v in s]
            cc = c.astype(np.complex64 if prec == "f32" else np.complex128).astype(np.complex128)
            ref = O.direct_type3(xs, cc, ss, -1)
            res[prec]["rel_l2_vs_direct"] = O.rel_l2(outs[prec][tidx], ref)
            if prec == "mixed":
                # the same algorithm in FP64 on the same targets (grids sized
                # from the full point sets' extents, as the plan's are)
                bnd = ([float(v.min()) for v in x], [float(v.max()) for v in x],
                       [float(v.min()) for v in s], [float(v.max()) for v in s])
                alg = O.nufft3(xs, cc, ss, -1, 1e-6, bounds=bnd)
                fp64_alg_err = O.rel_l2(alg, ref)
                res["mixed"]["fp64_restatement_rel_l2"] = fp64_alg_err
                res["f64"]["fp64_restatement_rel_l2"] = fp64_alg_err
        ms_, ns_ = args.cpu_sample // 20, args.cpu_sample // 40
        t0 = time.perf_counter()
        O.nufft3([v[:ms_] for v in x], c[:ms_], [v[:ns_] for v in s], -1, 1e-6)
        t1 = time.perf_counter()
        cpu = {"value": (ms_ + ns_) / (t1 - t0), "unit": "NU pts/s", **host_cpu(),
               "kind": "port",
               "sample": (f"FP64 C restatement (oracle/, OpenMP) + numpy FFT, type 3 with the "
                          f"first {ms_} sources and {ns_} targets: {t1 - t0:.2f}s")}
    return {"workload": "3D type-3, M=2e6 sources x,y in U[-1e4,1e4], z in U[-50,50]; "
                        "N=1e6 targets s_x,s_y in U[-0.02,0.02], s_z in U[-1,1]; "
                        "tol 1e-6; mixed plan (double coordinates, complex128 I/O, "
                        "FP64 fold and phases, fp32 spreading / FFT / interpolation); "
                        "fp32 and fp64 plans alongside",
            "sources": m, "targets": n, "precisions": res, "cpu_baseline": cpu}


def run_type3(args):
    leg = type3_leg(args)
    mixed = leg["precisions"]["mixed"]
    line = _common(args, "NU pts/s (sources + targets), 3D type-3 @1e-6, 1 GPU",
                   mixed["pts_per_s"], mixed["ms_per_step"], "f32",
                   {"workload": leg["workload"], "sources": leg["sources"],
                    "targets": leg["targets"]})
    line["roofline"] = mixed["roofline"]
    line["precisions"] = leg["precisions"]
    line["cpu_baseline"] = leg["cpu_baseline"]
    print(json.dumps(line), flush=True)


def run_batched(args):
    import torch
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    b = batched_leg(args, dev, torch.cuda.current_stream())
    line = _common(args, "NU pts/s, 256 batched uneven 2D type-1 fp32 @1e-5 vs plain, 1 GPU",
                   b["batched_pts_per_s"], b["batched_ms"], "f32",
                   {"workload": b["workload"]})
    line["batched"] = b
    line["roofline"] = None
    line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)


def run_uniform(args):
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
    # the spread stage (strengths -> fine grid: spreading into subgrid tiles
    # plus the tile reduction), its two kernels alone, and interpolation
    stage_ms = timed(lambda: plan.spread(dc), reps)
    spread_ms = timed(lambda: plan.spread_tiles(dc), reps)
    reduce_ms = timed(lambda: plan.reduce_tiles(accumulate=True), reps)
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
    ns = info["n_spread"]
    alg_bytes = m * (3 * 4 + 8) + G * 8  # coords + strength, fine grid once
    # bytes the reduction moves: every stored tile cell read once, grid written
    tile_cells = 1
    for d, nb in enumerate(info["bin_size"]):
        tile_cells *= nb + ns
    reduce_bytes = info["num_subproblems"] * tile_cells * 8 + G * 8
    # LDS work of the algorithm: spreading adds re and im (one packed 8-byte
    # integer atomic) into each of ns^3 cells per point; interpolation reads
    # each of ns^3 cells per point (one 8-byte complex read)
    lds_spread = m * ns ** 3 * 8
    lds_interp = m * ns ** 3 * 8
    prof = _profile_counters()
    legs = {
        "spread_stage": {"ms": stage_ms, "lds_bytes": lds_spread,
                         "kernels": ["spread_kernel", "bin_reduce_kernel"]},
        "interp_kernel": {"ms": interp_ms, "lds_bytes": lds_interp, "kernels": ["interp_kernel"]},
    }
    dom_name = max(legs, key=lambda k: legs[k]["ms"])
    dom = legs[dom_name]
    traffic = None
    if prof:
        try:
            traffic = sum(prof[k]["hbm_bytes_per_launch"] for k in dom["kernels"])
        except KeyError:
            traffic = None
    hbm_gbs = alg_bytes / (dom["ms"] * 1e-3)
    lds_bps = dom["lds_bytes"] / (dom["ms"] * 1e-3)
    lds = {"achieved": lds_bps / 1e12, "peak": LDS_PEAK / 1e12, "unit": "TB/s",
           "frac": lds_bps / LDS_PEAK, "lds_bytes_per_launch": dom["lds_bytes"],
           "note": ("algorithmic LDS data bytes (spreading: one packed 8-byte atomic per point and "
                    "cell, each a read-modify-write of its bank; interpolation: one 8-byte "
                    "read per point and cell) / launch time, against 256 B/clk/CU x 256 CUs "
                    "x 2.4 GHz")}
    if prof:
        busy = {}
        for k in dom["kernels"]:
            ctr = prof.get(k, {})
            if "lds_busy_frac" in ctr:
                busy[k] = {"lds_busy_frac": ctr["lds_busy_frac"],
                           "bank_conflict_frac_of_busy": ctr.get("bank_conflict_frac_of_busy")}
        lds["rocprof"] = busy
    # VALU issue rate of each kernel of the leg, from the committed counters
    # (SQ_INSTS_VALU over the kernel-trace duration of the same profiled run)
    valu = None
    valu_frac = 0.0
    if prof:
        valu = {"peak_wave_instr_per_s": VALU_PEAK, "source": PROFILE_COUNTERS, "kernels": {}}
        for k in dom["kernels"]:
            ctr = prof.get(k, {})
            if "sq_insts_valu" in ctr and "avg_us_kernel_trace" in ctr:
                rate = ctr["sq_insts_valu"] / (ctr["avg_us_kernel_trace"] * 1e-6)
                valu["kernels"][k] = {"wave_instr": ctr["sq_insts_valu"],
                                      "wave_instr_per_s": rate, "frac": rate / VALU_PEAK}
                valu_frac = max(valu_frac, rate / VALU_PEAK)
    fracs = {"hbm": hbm_gbs / HBM_PEAK, "lds": lds_bps / LDS_PEAK, "valu": valu_frac}
    bound = max(fracs, key=fracs.get)

    f_host = f.cpu().numpy()
    out2_host = out2.cpu().numpy()

    cpu = None
    acc = {"samples": 1000, "reference": "FP64 direct NUDFT"}
    if not args.no_cpu_baseline:
        cpu, acc = cpu_baseline_leg(x, c, f_host, out2_host, modes, args.tol, args.cpu_sample)

    del plan
    torch.cuda.synchronize()
    batched = batched_leg(args, dev, stream) if args.batch_count > 0 else None
    clustered = None if args.no_clustered else clustered_leg(args)
    type3 = None if args.no_type3 else type3_leg(args)

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
            "n_spread": ns, "bin_size": info["bin_size"],
            "subproblems": info["num_subproblems"],
        },
        "roofline": {
            "bound": bound,
            "kernel": dom_name,
            "achieved": hbm_gbs / 1e9,
            "peak": HBM_PEAK / 1e9,
            "unit": "GB/s",
            "frac": hbm_gbs / HBM_PEAK,
            "traffic": traffic,
            "traffic_source": PROFILE_COUNTERS if traffic is not None else None,
            "alg_bytes_per_launch": alg_bytes,
            "avg_launch_ms": dom["ms"],
            "lds": lds,
            "valu": valu,
            "ceiling_fracs": fracs,
        },
        "kernels": {
            "spread_stage_ms": stage_ms,
            "spread_ms": spread_ms, "spread_GBps": alg_bytes / (spread_ms * 1e-3) / 1e9,
            "reduce_ms": reduce_ms, "reduce_GBps": reduce_bytes / (reduce_ms * 1e-3) / 1e9,
            "reduce_bytes": reduce_bytes,
            "interp_ms": interp_ms, "interp_GBps": alg_bytes / (interp_ms * 1e-3) / 1e9,
            "interp_lds_frac": lds_interp / (interp_ms * 1e-3) / LDS_PEAK,
            "spread_stage_lds_frac": lds_spread / (stage_ms * 1e-3) / LDS_PEAK,
        },
        "stages": {
            "type1_ms": t1_ms, "type2_ms": t2_ms, "set_points_ms": set_points_ms,
            "plan_create_ms": create_ms,
            "type1_pts_per_s": m / (t1_ms * 1e-3), "type2_pts_per_s": m / (t2_ms * 1e-3),
        },
        "accuracy": acc,
        "batched": batched,
        "clustered": clustered,
        "type3": type3,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
