"""Benchmark of the explicit finite-volume hot path on one MI355X.

Workload (the configuration the headline metric is quoted on): first-order
acoustic system (p, u, v, w) on a 256^3 uniform hex mesh of the unit cube,
FP32, Gaussian pressure pulse (amplitude 1, sigma 0.05) at a seeded centre,
u = v = w = 0, c = 1, reflective walls, dt = 0.4 dx / c.  One step = one fused
launch that advances all four fields (equations in order, as Equation::evaluate
runs them).  Inputs are resident in HBM before the timed region starts.

Prints one JSON line with MCUPS (million cell-updates per second), the
roofline position of the dominant kernel (HIP events over the timed region on
the launch stream) and a CPU baseline (the oracle, on a bounded sample of the
same workload, on this host's cores).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "MCUPS + % HBM roofline (1 GPU); adjoint/forward cost ratio; batched vs plain"
HBM_PEAK_GBS = 8000.0
KERNELS = {0: "wave_uniform_k", 1: "wave_march_k"}


def _traffic(n, dtype):
    """HBM bytes per launch from the committed rocprofv3 PMC summary, if any."""
    path = os.path.join(ROOT, "profiles", f"wave{n}cube_{dtype}_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        rec = json.load(f)
    return rec.get("hbm_bytes_per_launch")


def cpu_baseline(case, dtype, steps):
    """The oracle (reference-faithful CPU evaluation: every operator materialised
    over the whole mesh, equations in order) on `steps` steps of the same mesh."""
    import numpy as np

    from oracle import pyoracle

    npdt = np.float32 if dtype == "f32" else np.float64
    om = pyoracle.OracleMesh(*case.n, *case.spacing, dtype=npdt)
    z = np.zeros(om.N, dtype=npdt)
    p0 = case.p0.astype(npdt)
    t0 = time.perf_counter()
    om.wave_run(p0, z, z, z, -case.c ** 2, case.dt, steps)
    t = time.perf_counter() - t0
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": om.N * steps / t / 1e6, "unit": "MCUPS", "cores": cores, "kind": "port",
            "sample": f"oracle wave_run, {case.n[0]}^3 {dtype}, {steps} steps of the benchmark workload, "
                      f"{t:.2f} s"}


def calibrate_copy(args):
    """A plain device-to-device copy of 4 fields' worth of bytes (what one fused
    launch reads), repeated --steps times: the known-size reference for
    scaling rocprofv3's FETCH_SIZE / WRITE_SIZE counters."""
    import torch

    esize = 4 if args.dtype == "f32" else 8
    nbytes = 4 * esize * args.n ** 3
    src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(args.steps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"calibrate_copy_bytes_per_launch": nbytes, "launches": args.steps}), flush=True)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--n", type=int, default=256, help="cube side")
    ap.add_argument("--dtype", choices=["f32", "f64"], default="f32")
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--shape", default="auto",
                    help="fused-kernel shape 'variant,tile_rows,chunk,xcd_remap' or 'auto' (autotuner)")
    ap.add_argument("--tune-verbose", action="store_true")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run only a plain device copy of the benchmark's read volume (PMC calibration)")
    args = ap.parse_args(argv)

    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (--gpus {args.gpus} requested); "
              "the multi-GPU slab decomposition is design-only", file=sys.stderr)
        sys.exit(2)

    import torch

    from cfdarcov2_amd import capi, workloads

    if args.calibrate_copy:
        calibrate_copy(args)
        return

    case = workloads.wave_config(n=args.n, steps=args.steps)
    tdt = torch.float32 if args.dtype == "f32" else torch.float64
    ctx = capi.Context(0)
    mesh = capi.HexMesh(ctx, *case.n, *case.spacing, dtype=args.dtype)
    N = mesh.N
    from cfdarcov2_amd import autotune
    if args.shape == "auto":
        shape = autotune.tune_wave(mesh, use_cache=False, verbose=args.tune_verbose)
    else:
        v_, ty_, L_, r_ = (int(x) for x in args.shape.split(","))
        mesh.set_wave_shape(v_, ty_, L_, r_)
        shape = {"variant": v_, "tile_rows": ty_, "chunk": L_, "xcd_remap": r_}
    A = [torch.from_numpy(case.p0).to(tdt).cuda()] + [torch.zeros(N, dtype=tdt, device="cuda") for _ in range(3)]
    B = [torch.empty_like(A[0]) for _ in range(4)]
    K = -case.c ** 2

    cur, nxt = A, B
    if args.warmup:
        res = mesh.wave_run(cur, nxt, K, case.dt, args.warmup)
        if res[0] is not cur[0]:
            cur, nxt = nxt, cur
    torch.cuda.synchronize()

    stream = torch.cuda.current_stream()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev0.record(stream)
    res = mesh.wave_run(cur, nxt, K, case.dt, args.steps)
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kernel_ms = ev0.elapsed_time(ev1) / max(args.steps, 1)

    bad = sum(ctx.count_nonfinite(t) for t in res)
    if bad:
        raise RuntimeError(f"non-finite state after the timed run ({bad} entries)")

    esize = 4 if args.dtype == "f32" else 8
    bytes_per_launch = 8 * esize * N          # 4 fields read + 4 written
    achieved = bytes_per_launch / (kernel_ms * 1e-3) / 1e9
    traffic = _traffic(args.n, args.dtype)
    out = {
        "metric": METRIC,
        "value": N * args.steps / wall / 1e6,
        "unit": "MCUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.dtype,
        "data": "synthetic",
        "config": {"workload": f"3D first-order acoustic wave (p,u,v,w), {args.n}^3 uniform hex mesh, "
                               f"{args.dtype.upper()}, CFL 0.4, reflective walls, seeded Gaussian pulse",
                   "cells": N, "parallelism": "single GPU",
                   "kernel_shape": {k: shape[k] for k in ("variant", "tile_rows", "chunk", "xcd_remap")}},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": KERNELS[shape["variant"]], "kernel_ms": kernel_ms,
                     "algorithmic_bytes_per_launch": bytes_per_launch},
    }
    mesh.close()
    ctx.close()
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(case, args.dtype, args.cpu_steps)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
