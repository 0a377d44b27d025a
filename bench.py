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
KERNELS = {0: "wave_uniform_k", 1: "wave_march_k", 7: "wave_row_k", 8: "wave_row_k", 9: "wave_row_k",
           11: "wave_row2_k", 12: "wave_row2_k"}


def _traffic(n, dtype, shape):
    """HBM bytes per launch from the committed rocprofv3 PMC summary, if one
    exists for this workload and kernel shape."""
    path = os.path.join(ROOT, "profiles", f"wave{n}cube_{dtype}_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        rec = json.load(f)
    for entry in rec.get("shapes", [rec]):
        prof_shape = entry.get("kernel_shape")
        if prof_shape is None or all(prof_shape.get(k) == shape.get(k) for k in prof_shape):
            return entry.get("hbm_bytes_per_launch")
    return None


def cpu_baseline(case, dtype, steps, seconds=7.0):
    """The CPU restatement of the path on the same mesh, timed on this host
    in three forms: the fused loop (one pass per stage) and the operator-by-
    operator evaluation (every operator materialised over the whole mesh, as
    the reference's CPU path evaluates), both with every core, and the fused
    loop on one core.  The reported value is the faster all-core form (which
    one wins depends on the host: the fused loop computes fewer face values,
    the operator form streams simpler loops).  Each variant runs a sample of
    the benchmark's steps sized to about `seconds`."""
    import numpy as np

    from oracle import pyoracle

    npdt = np.float32 if dtype == "f32" else np.float64
    om = pyoracle.OracleMesh(*case.n, *case.spacing, dtype=npdt)
    z = np.zeros(om.N, dtype=npdt)
    p0 = case.p0.astype(npdt)
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))

    def timed(fn, base_steps):
        t0 = time.perf_counter()
        fn(base_steps)
        probe = time.perf_counter() - t0
        n = max(base_steps, min(500, int(seconds * base_steps / max(probe, 1e-9))))
        t0 = time.perf_counter()
        fn(n)
        t = time.perf_counter() - t0
        return om.N * n / t / 1e6, n, t

    run_op = lambda n: om.wave_run(p0, z, z, z, -case.c ** 2, case.dt, n)
    run_fused = lambda n: om.wave_run_fused(p0, z, z, z, -case.c ** 2, case.dt, n)
    fused, nf, tf = timed(run_fused, steps)
    op, no, to = timed(run_op, steps)
    prev = pyoracle.set_threads(1)
    try:
        one, n1, t1 = timed(run_fused, 1)
    finally:
        pyoracle.set_threads(prev)
    if fused >= op:
        best, sample = fused, f"fused CPU loop (oracle wave_run_fused), first {nf} steps, {tf:.1f} s"
    else:
        best, sample = op, f"operator-by-operator CPU loop (oracle wave_run), first {no} steps, {to:.1f} s"
    return {"value": best, "unit": "MCUPS", "cores": cores, "kind": "port",
            "sample": f"{sample}, of the benchmark workload ({case.n[0]}^3 {dtype}); the faster all-core form",
            "variants": {"fused_all_cores": fused, "operator_by_operator_all_cores": op,
                         "fused_one_core": one,
                         "samples": {"fused_all_cores": f"{nf} steps, {tf:.1f} s",
                                     "operator_by_operator_all_cores": f"{no} steps, {to:.1f} s",
                                     "fused_one_core": f"{n1} steps, {t1:.1f} s"}}}


def _events():
    import torch
    return torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)


def _timed(fn):
    """(wall seconds, event milliseconds) of fn() on the current stream."""
    import torch
    stream = torch.cuda.current_stream()
    e0, e1 = _events()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e0.record(stream)
    fn()
    e1.record(stream)
    torch.cuda.synchronize()
    return time.perf_counter() - t0, e0.elapsed_time(e1)


def copy_ceiling(nbytes=256 << 20, reps=20):
    """Device-to-device copy rate (GB/s, read + write bytes) of one buffer of
    `nbytes`: the practical streaming ceiling the stencil is compared with."""
    import torch
    a = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    stream = torch.cuda.current_stream()
    e0, e1 = _events()
    torch.cuda.synchronize()
    e0.record(stream)
    for _ in range(reps):
        b.copy_(a)
    e1.record(stream)
    torch.cuda.synchronize()
    t = e0.elapsed_time(e1) * 1e-3 / reps
    del a, b
    return 2 * nbytes / t / 1e9


def stream_ceiling(nfloat=256 ** 3, reps=30):
    """Best rate (GB/s, 8 fields x 4 B per cell) of the acoustic step's access
    pattern without its arithmetic: 4 fields read and 4 written per pass,
    passes ping-ponging between two sets, over the cache policies and grid
    sizes of the stream probe.  Returns (GB/s, description)."""
    import ctypes

    from cfdarcov2_amd import build as hipbuild
    lib = ctypes.CDLL(hipbuild.STREAM_PROBE)
    lib.stream_probe_pingpong.argtypes = [ctypes.c_long, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                          ctypes.POINTER(ctypes.c_double)]
    names = {0: "plain loads and stores", 1: "non-temporal stores", 2: "non-temporal loads",
             3: "non-temporal loads and stores"}
    best = (0.0, "")
    for policy in range(4):
        for grid in (2048, 8192, 16384):
            us = ctypes.c_double()
            if lib.stream_probe_pingpong(nfloat, reps, grid, policy, ctypes.byref(us)) != 0:
                raise RuntimeError("stream probe failed")
            gbs = 8 * 4 * nfloat / (us.value * 1e-6) / 1e9
            if gbs > best[0]:
                best = (gbs, f"ping-pong 4 reads + 4 writes of {nfloat * 4 >> 20} MiB fields, {names[policy]}, "
                              f"grid {grid}")
    return best


def suite_heat(ctx, cpu=True, verbose=False):
    """Heat 64^3 FP64, 200 steps at dt = 0.2 dx^2/alpha (the reference CPU
    configuration); 16 algorithmic bytes per cell-update; 4 MiB of state stays
    in the last-level cache, so the fraction is of HBM peak but the data is
    cache-resident."""
    import numpy as np
    import torch

    from cfdarcov2_amd import autotune, capi, workloads
    case = workloads.heat_config()
    m = capi.HexMesh(ctx, *case.n, *case.spacing, dtype="f64")
    shape = autotune.tune_heat(m, case.alpha, case.dt, verbose=verbose)
    a = torch.from_numpy(case.T0).cuda()
    b = torch.empty_like(a)
    # untimed warm-up of the same run (captures the run's block of steps)
    m.heat_run(a, b, case.alpha, case.dt, case.steps)
    a.copy_(torch.from_numpy(case.T0))
    wall, ev = _timed(lambda: m.heat_run(a, b, case.alpha, case.dt, case.steps))
    N = m.N
    rec = {"workload": "heat 64^3 FP64, 200 steps, dt = 0.2 dx^2/alpha", "mcups": N * case.steps / wall / 1e6,
           "kernel_us_per_step": ev * 1e3 / case.steps,
           "gbs_algorithmic": 16 * N / (ev * 1e-3 / case.steps) / 1e9,
           "kernel_shape": {k: shape[k] for k in ("variant", "waves", "chunk")}}
    rec["roofline_frac"] = rec["gbs_algorithmic"] / HBM_PEAK_GBS
    m.close()
    if cpu:
        from oracle import pyoracle
        om = pyoracle.OracleMesh(*case.n, *case.spacing, dtype=np.float64)
        om.heat_run(case.T0, case.alpha, case.dt, 2)  # thread pool and first-touch warm-up
        # the whole 200-step run, repeated until about 2 s of CPU work (one run
        # takes a few hundredths of a second, too short to time alone)
        runs, t = 0, 0.0
        while t < 2.0 and runs < 200:
            t0 = time.perf_counter()
            om.heat_run(case.T0, case.alpha, case.dt, case.steps)
            t += time.perf_counter() - t0
            runs += 1
        rec["cpu_oracle_mcups"] = N * case.steps * runs / t / 1e6
        rec["cpu_sample"] = (f"oracle heat_run 64^3 FP64, the {case.steps}-step run {runs} times after 2 warm-up "
                             f"steps, {t:.2f} s")
    return rec


def suite_wave_f64(ctx, steps=200):
    import torch

    from cfdarcov2_amd import autotune, capi, workloads
    case = workloads.wave_config(n=256, steps=steps)
    m = capi.HexMesh(ctx, *case.n, *case.spacing, dtype="f64")
    shape = autotune.tune_wave(m, steps=12, use_cache=False,
                               candidates=[(1, ty, L, r) for ty in (4, 8) for L in (8, 16, 32) for r in (0, 1)]
                               + [(v, ty, L, r) for v in (7, 9) for ty in (4, 8) for L in (8, 16, 32, 64) for r in (1, 3)]
                               + [(v, 2, L, r) for v in (11, 12) for L in (8, 16, 32, 64) for r in (1, 3)])
    N = m.N
    A = [torch.from_numpy(case.p0).cuda()] + [torch.zeros(N, dtype=torch.float64, device="cuda") for _ in range(3)]
    B = [torch.empty_like(A[0]) for _ in range(4)]
    wall, ev = _timed(lambda: m.wave_run(A, B, -1.0, case.dt, steps))
    gbs = 64 * N / (ev * 1e-3 / steps) / 1e9
    m.close()
    del A, B
    # the same access pattern without arithmetic: 8 fields of 2N floats (the
    # bytes of N doubles), ping-ponging; 1 GiB per pass, twice the last-level
    # cache, so the FP64 step streams from HBM where the FP32 one partly hits
    ceiling, how = stream_ceiling(nfloat=2 * N, reps=20)
    return {"workload": "acoustic wave 256^3 FP64", "mcups": N * steps / wall / 1e6,
            "kernel_us_per_step": ev * 1e3 / steps, "gbs_algorithmic": gbs, "roofline_frac": gbs / HBM_PEAK_GBS,
            "stream_ceiling_gbs": ceiling, "stream_ceiling": how, "frac_of_stream_ceiling": gbs / ceiling,
            "kernel_shape": shape}


def suite_sod(ctx, cpu=True, verbose=False):
    """Sod 512x64x64, interior vertices moved by up to 20% of dx, FP32, 500
    Rusanov steps at CFL 0.3.  92 algorithmic bytes per cell-update: 5 fields
    read and written, 3 owned faces x (normal + area) and the volume."""
    import numpy as np
    import torch

    from cfdarcov2_amd import capi, workloads
    cfg = workloads.sod_config()
    m = capi.HexMesh(ctx, *cfg.n, *cfg.spacing, dtype="f32", vertex_offsets=cfg.vertex_offsets)
    U0 = workloads.sod_state(m.table("ccx"), dtype=np.float32)
    A = [torch.from_numpy(U0[k].copy()).cuda() for k in range(5)]
    B = [torch.empty_like(a) for a in A]
    from cfdarcov2_amd import autotune
    shape = autotune.tune_euler(m, A, cfg.gamma, cfg.cfl, steps=10, verbose=verbose)  # also builds the geometry
    wall, ev = _timed(lambda: m.euler_run(A, B, cfg.gamma, cfg.cfl, cfg.steps))
    res = A if cfg.steps % 2 == 0 else B
    bad = sum(ctx.count_nonfinite(t) for t in res)
    N = m.N
    gbs = 92 * N / (ev * 1e-3 / cfg.steps) / 1e9
    rec = {"workload": "Sod 512x64x64 perturbed (20% dx), FP32, Rusanov, CFL 0.3, 500 steps",
           "mcups": N * cfg.steps / wall / 1e6, "kernel_us_per_step": ev * 1e3 / cfg.steps,
           "gbs_algorithmic": gbs, "roofline_frac": gbs / HBM_PEAK_GBS, "nonfinite": int(bad),
           "kernel": "euler_march_k" if shape["variant"] == 1 else "euler_step_k",
           "kernel_shape": {k: shape[k] for k in ("variant", "tile_rows", "chunk")}}
    m.close()
    if cpu:
        from oracle import pyoracle
        om = pyoracle.OracleMesh(*cfg.n, *cfg.spacing, dtype=np.float32, perturb=cfg.vertex_offsets)
        U = workloads.sod_state(om.table("ccx"), dtype=np.float32)
        om.euler_run(U, cfg.gamma, cfg.cfl, 1)  # thread pool and first-touch warm-up
        t0 = time.perf_counter()
        om.euler_run(U, cfg.gamma, cfg.cfl, 20)
        t = time.perf_counter() - t0
        rec["cpu_oracle_mcups"] = N * 20 / t / 1e6
        rec["cpu_sample"] = f"oracle euler_run 512x64x64 FP32, first 20 steps after a warm-up step, {t:.2f} s"
    return rec


PLAIN_WAVE_CANDIDATES = ([(0, 4, 8, 0)] + [(1, 4, L, r) for L in (4, 8, 16) for r in (0, 1)]
                         + [(2, 8, L, 0) for L in (8, 16)] + [(3, 2, L, 0) for L in (8, 16, 32)]
                         + [(4, 4, L, 0) for L in (8, 16)]
                         + [(v, ty, L, 3) for v in (7, 9) for ty in (4, 8) for L in (8, 16)]
                         + [(v, ty, L, 3) for v in (11, 12) for ty in (4, 8, 10) for L in (8, 16, 32)])


def ensemble_api_seconds(specs, steps, timeout=900):
    """Time of the same cases solved one after another through the C++
    operator API (Variable / Equation::evaluate with hipRTC kernels), by the
    helper program cfdarcov2_amd/_build/bench/ensemble_api."""
    import subprocess
    import tempfile

    import numpy as np
    prog = os.path.join(ROOT, "cfdarcov2_amd", "_build", "bench", "ensemble_api")
    if not os.path.exists(prog):
        return None
    with tempfile.TemporaryDirectory() as tmp:
        path = os.path.join(tmp, "cases.bin")
        with open(path, "wb") as f:
            np.array([len(specs), steps], dtype=np.int64).tofile(f)
            for sp in specs:
                kind = 0 if sp.kind == "heat" else 1
                coef = sp.coef if kind == 0 else -sp.coef      # c^2 for the acoustic system
                np.array([kind, sp.n], dtype=np.int64).tofile(f)
                np.array([sp.spacing, coef, sp.dt], dtype=np.float64).tofile(f)
                np.ascontiguousarray(sp.field0, dtype=np.float32).tofile(f)
        res = subprocess.run([prog, path], capture_output=True, text=True, timeout=timeout, cwd=tmp)
    if res.returncode != 0:
        raise RuntimeError("ensemble_api failed: " + res.stdout[-2000:] + res.stderr[-2000:])
    line = [l for l in res.stdout.splitlines() if l.startswith("{")][-1]
    return json.loads(line)["ensemble_api_seconds"]


def suite_ensemble(ctx, steps=500, verbose=False):
    """64 heat/acoustic cases of seeded uneven sides, FP32, 500 steps: the
    autotuned batched graph run against 64 sequential single-case runs (each
    case with the fused kernel shape tuned for its own size), and against the
    same cases solved one by one through the operator API."""
    from cfdarcov2_amd import autotune, workloads
    from cfdarcov2_amd.ensemble import Ensemble, autotune as tune_batch, time_gpu
    specs, steps = workloads.ensemble_config(steps=steps)
    ens = Ensemble(ctx, specs)
    plain_shapes, plain_heat = {}, {}
    for c in ens.cases:
        if c.spec.kind != "wave":
            # each heat case with its fastest form (per-thread, or two steps per launch)
            if c.spec.n not in plain_heat:
                plain_heat[c.spec.n] = autotune.tune_heat(c.mesh, c.spec.coef, c.spec.dt)
            h = plain_heat[c.spec.n]
            c.mesh.set_heat_shape(h["variant"], h["waves"], h["chunk"])
            continue
        if c.spec.n not in plain_shapes:
            plain_shapes[c.spec.n] = autotune.tune_wave(c.mesh, steps=20, use_cache=False,
                                                        candidates=PLAIN_WAVE_CANDIDATES)
        b = plain_shapes[c.spec.n]
        c.mesh.set_wave_shape(b["variant"], b["tile_rows"], b["chunk"], b["xcd_remap"])
    ens.reset()
    ens.run_sequential(64)  # untimed warm-up (captures each heat case's block of steps)
    ens.reset()
    t_seq = time_gpu(lambda: ens.run_sequential(steps), reps=1)
    best = tune_batch(ens, steps=20, verbose=verbose)
    ens.run(steps)  # capture the full-length graph
    ens.reset()
    t_b = time_gpu(lambda: ens.run(steps), reps=1)
    cu = ens.cell_updates(steps)
    rec = {"workload": f"{len(specs)} cases (heat even, wave odd), seeded sides in {workloads.ENSEMBLE_SIDES}, "
                       f"FP32, {steps} steps", "cells": ens.cell_updates(1),
           "batched_mcups": cu / t_b / 1e6, "sequential_mcups": cu / t_seq / 1e6, "speedup": t_seq / t_b,
           "batched_ms": t_b * 1e3, "sequential_ms": t_seq * 1e3, "batch_shape": best,
           "plain_wave_shapes": {n: {k: v[k] for k in ("variant", "tile_rows", "chunk", "xcd_remap")}
                                 for n, v in plain_shapes.items()},
           "plain_heat_shapes": {n: {k: v[k] for k in ("variant", "waves", "chunk")} for n, v in plain_heat.items()},
           "batched_gbs_algorithmic": ens.algorithmic_bytes(steps) / t_b / 1e9}
    ens.close()
    t_api = ensemble_api_seconds(specs, steps)
    if t_api is not None:
        rec["operator_api_sequential_ms"] = t_api * 1e3
        rec["speedup_vs_operator_api"] = t_api / t_b
    return rec


def suite_adjoint(ctx, verbose=False):
    """Adjoint 128^3 FP32, 2000 steps, checkpoints every 45, 32 sensors,
    constant K = -1 (c = 1).  Mean time per step of: the autotuned fused
    forward (cfd_wave_run, the plain solve), the traced forward (traces
    recorded inside the fused step, as the autograd forward runs it), the
    transposed step (dL/dp0) and the transposed step with dL/dK through
    checkpoint-and-recompute.  The ratios divide by the fused forward; the
    ratio against the traced forward is reported beside them."""
    import numpy as np
    import torch

    from cfdarcov2_amd import adjoint, autotune, capi, workloads
    cfg = workloads.adjoint_config()
    m = capi.HexMesh(ctx, *cfg.n, *cfg.spacing, dtype="f32")
    steps = cfg.steps
    # the plain solve with the tuner's shape
    fwd_shape = autotune.tune_wave(m, use_cache=False, verbose=verbose)
    A = [torch.from_numpy(cfg.p0_guess.astype(np.float32)).cuda()] + [
        torch.zeros(m.N, dtype=torch.float32, device="cuda") for _ in range(3)]
    B = [torch.empty_like(A[0]) for _ in range(4)]
    m.wave_run(A, B, -1.0, cfg.dt, 4)
    _, ev_plain = _timed(lambda: m.wave_run(A, B, -1.0, cfg.dt, steps))
    del A, B
    pr = adjoint.WaveProblem(m, cfg.sensors, cfg.dt, steps, K=-1.0, every=cfg.every)
    adj_shape = autotune.tune_adjoint(pr, verbose=verbose)
    p_true = torch.from_numpy(cfg.p0_true.astype(np.float32)).cuda()
    obs, _, _ = adjoint.forward_traces(pr, p_true, checkpoints=False)
    p0 = torch.from_numpy(cfg.p0_guess.astype(np.float32)).cuda()
    out = {}
    # the autograd forward of dL/dp0: traces only
    _, ev_f = _timed(lambda: out.__setitem__("f", adjoint.forward_traces(pr, p0, checkpoints=False)))
    tr = out["f"][0]
    g = 2.0 * (tr - obs)
    _, ev_a = _timed(lambda: adjoint.adjoint(pr, g))
    # dL/dK as well: forward with checkpoints, reverse pass with recomputation
    _, ev_fc = _timed(lambda: out.__setitem__("fc", adjoint.forward_traces(pr, p0, checkpoints=True)))
    ck = out["fc"][2]
    _, ev_k = _timed(lambda: adjoint.adjoint(pr, g, None, ck, want_K=True))
    us = lambda ms: ms * 1e3 / steps
    adj_bytes = 32.0 * m.N           # read and write p, u, v, w once per transposed step
    rec = {"workload": "adjoint acoustic 128^3 FP32, 2000 steps, checkpoints every 45, 32 sensors, K = -1",
           "fused_forward_us_per_step": us(ev_plain),
           "traced_forward_us_per_step": us(ev_f),
           "traced_forward_with_checkpoints_us_per_step": us(ev_fc),
           "adjoint_us_per_step": us(ev_a),
           "adjoint_with_K_recompute_us_per_step": us(ev_k),
           "ratio": ev_a / ev_plain,
           "ratio_vs_traced_forward": ev_a / ev_f,
           "ratio_with_K_recompute": ev_k / ev_plain,
           "adjoint_gbs_algorithmic": adj_bytes / (us(ev_a) * 1e-6) / 1e9,
           "forward_shape": {k: fwd_shape[k] for k in ("variant", "tile_rows", "chunk", "xcd_remap")},
           "adjoint_shape": {k: adj_shape[k] for k in ("form", "tile_rows", "chunk")}}
    m.close()
    return rec


# The reference's published A100 figures (Mpts/s, KernelBuilder on) for the
# same drivers and sizes: scripts/v2_results/
# report_history_a100_cuda_full_no_data_move_even_better.csv:7,9,17,19,28,32.
A100_MPTS = {("heat", 200): 6440.3, ("heat", 300): 6441.7, ("euler", 200): 264.7, ("euler", 300): 267.1,
             ("wave", 200): 6559.7, ("wave", 400): 7258.6}


def suite_unstructured(ctx, n=128, steps=200):
    """Heat on a general hex mesh given by its connectivity
    (cfd_mesh_create_unstructured): the n^3 lattice of the unit cube with its
    cells, faces and vertices numbered at random, FP64, alpha 1e-3,
    dt = 0.1 dx^2 / alpha.  The vertices stay on the lattice: the reference's
    cell volume, built on sum-of-components face areas (src/mesh3d.cpp:61,
    101-104), is not translation-invariant, and at 128^3 a 1% vertex move
    already gives cells of negative volume (down to -3 dx^3) on which the heat
    step diverges.  The kernel reads every face's geometry all the same.  The
    library renumbers the cells by reverse Cuthill-McKee on upload; the same
    mesh kept in its random numbering is timed beside it.  Algorithmic bytes
    per step: per cell the state read and written (16 B), six neighbour and
    six face ids (48 B) and the volume (8 B); per face its normal, area and
    alpha_d (40 B), each face counted once.  The fused acoustic step and the
    Rusanov step on the renumbered mesh are timed after it (`wave_rcm`,
    `euler_rcm`)."""
    import numpy as np
    import torch

    from cfdarcov2_amd import capi, meshio, workloads
    d = 1.0 / n
    conn, _ = meshio.shuffled(meshio.lattice_connectivity(n, n, n, d, d, d, dtype=np.float64), 7)
    rng = np.random.default_rng(9)
    T0 = rng.random(n ** 3)
    alpha, dt = 1e-3, 0.1 * d * d / 1e-3
    rec = {"workload": f"heat on an unstructured hex mesh: {n}^3 lattice connectivity in a random "
                       f"cell/face/vertex numbering, FP64, {steps} steps"}
    for reorder in (True, False):
        um = capi.UnstructuredMesh(ctx, conn, dtype="f64", reorder=reorder)
        a = torch.from_numpy(um.to_mesh_order(T0)).cuda()
        b = torch.empty_like(a)
        um.heat_run(a, b, alpha, dt, 10)
        a.copy_(torch.from_numpy(um.to_mesh_order(T0)))
        wall, ev = _timed(lambda: um.heat_run(a, b, alpha, dt, steps))
        res = a if steps % 2 == 0 else b
        if ctx.count_nonfinite(res):
            raise RuntimeError("non-finite heat state on the unstructured mesh")
        N, F = um.N, um.F
        nbytes = N * (16 + 48 + 8) + F * 40
        us = ev * 1e3 / steps
        key = "rcm" if reorder else "random_order"
        rec[key] = {"mcups": N * steps / wall / 1e6, "kernel_us_per_step": us,
                    "gbs_algorithmic": nbytes / (us * 1e-6) / 1e9,
                    "roofline_frac": nbytes / (us * 1e-6) / 1e9 / HBM_PEAK_GBS}
        if reorder:
            rec["algorithmic_bytes_per_step"] = nbytes
            rec["kernel"] = "uheat_k"
            # the acoustic step on the same mesh (K = -1, dt = 0.4 dx), two
            # gathers per step: the velocities, then the pressure from them.
            # Bytes: the state's 4 fields read and written (64 B per cell),
            # and per pass the cell's ids, signs and volume (62 B) and each
            # face's normal and area (32 B)
            st = [torch.from_numpy(um.to_mesh_order(T0)).cuda()] + [torch.zeros_like(a) for _ in range(3)]
            sb = [torch.empty_like(x) for x in st]
            um.wave_run(st, sb, -1.0, 0.4 * d, 4)
            wsteps = steps // 2 * 2
            wwall, wev = _timed(lambda: um.wave_run(st, sb, -1.0, 0.4 * d, wsteps))
            if ctx.count_nonfinite(st[0]):
                raise RuntimeError("non-finite acoustic state on the unstructured mesh")
            wbytes = N * (64 + 2 * 62) + F * 2 * 32
            wus = wev * 1e3 / wsteps
            rec["wave_rcm"] = {"mcups": N * wsteps / wwall / 1e6, "kernel_us_per_step": wus,
                               "algorithmic_bytes_per_step": wbytes,
                               "gbs_algorithmic": wbytes / (wus * 1e-6) / 1e9,
                               "roofline_frac": wbytes / (wus * 1e-6) / 1e9 / HBM_PEAK_GBS,
                               "kernels": "uwave_vel_k + uwave_pres_k"}
            del st, sb
            # the Rusanov step on the same mesh: a Sod tube along x, CFL 0.3.
            # Bytes: the five state fields read and written (80 B per cell),
            # the cell's ids and signs (54 B) and each face's normal and area
            # (32 B)
            ccx = um.to_mesh_order(np.asarray(conn["vertices"])[np.asarray(conn["cell_verts"])][:, :, 0].mean(axis=1))
            U0 = workloads.sod_state(ccx, dtype=np.float64)
            ea = [torch.from_numpy(np.ascontiguousarray(U0[k])).cuda() for k in range(5)]
            eb = [torch.empty_like(x) for x in ea]
            um.euler_run(ea, eb, 1.4, 0.3, 2)
            esteps = 100
            ewall, eev = _timed(lambda: um.euler_run(ea, eb, 1.4, 0.3, esteps))
            if any(ctx.count_nonfinite(x) for x in ea):
                raise RuntimeError("non-finite Rusanov state on the unstructured mesh")
            ebytes = N * (80 + 54) + F * 32
            eus = eev * 1e3 / esteps
            rec["euler_rcm"] = {"mcups": N * esteps / ewall / 1e6, "kernel_us_per_step": eus,
                                "algorithmic_bytes_per_step": ebytes,
                                "gbs_algorithmic": ebytes / (eus * 1e-6) / 1e9,
                                "roofline_frac": ebytes / (eus * 1e-6) / 1e9 / HBM_PEAK_GBS,
                                "kernel": "ueuler_step_k"}
            del ea, eb
        um.close()
        del a, b
    rec["rcm_speedup"] = rec["random_order"]["kernel_us_per_step"] / rec["rcm"]["kernel_us_per_step"]
    return rec


def suite_drivers(reps=3):
    """The reference's example drivers, compiled unchanged against the operator
    API, run the way the reference's harness runs them
    (scripts/benchmark_cfdarchov2.py:44: `-Lx n -Ly n -Lz n -dx 1 -dy 1 -dz 1
    --visualize -t 200 -dt 0.005 -c`, the "Time difference" line, best of 3;
    Mpts/s = n^3 * 200 / microseconds).  The wave example ships with a host-only
    boundary callback (a host round trip of the field every step);
    point_source_wave solves the same problem with its sources set on the
    device as well."""
    import re
    import subprocess

    from cfdarcov2_amd import build as hipbuild
    runs = [("heat", "drivers/cfdARCO_heat", 200), ("heat", "drivers/cfdARCO_heat", 300),
            ("euler", "drivers/cfdARCO_euler", 200), ("euler", "drivers/cfdARCO_euler", 300),
            ("wave", "drivers/cfdARCO_wave", 200), ("wave", "drivers/cfdARCO_wave", 400),
            ("wave", "bench/point_source_wave", 200), ("wave", "bench/point_source_wave", 400)]
    out = []
    for exp, rel, n in runs:
        exe = os.path.join(hipbuild.BUILD_DIR, rel)
        if not os.path.exists(exe):
            out.append({"program": os.path.basename(rel), "n": n, "skipped": "not built"})
            continue
        best = None
        for _ in range(reps):
            print(f"[drivers] {os.path.basename(rel)} {n}^3", file=sys.stderr, flush=True)
            res = subprocess.run([exe, "-Lx", str(n), "-Ly", str(n), "-Lz", str(n), "-dx", "1", "-dy", "1", "-dz", "1",
                                  "-t", "200", "-dt", "0.005", "-c"], capture_output=True, text=True, timeout=300)
            if res.returncode != 0:
                raise RuntimeError(f"{rel} {n}^3 failed:\n{res.stderr[-2000:]}")
            us = int(re.search(r"Time difference = (\d+)\[microseconds\]", res.stdout).group(1))
            best = us if best is None else min(best, us)
        sums = re.findall(r"At (\S+) Current sum = (\S+) Current max = (\S+)", res.stdout)
        mpts = n ** 3 * 200 / best
        out.append({"program": os.path.basename(rel), "n": n, "best_us": best, "mpts": mpts,
                    "a100_published_mpts": A100_MPTS[(exp, n)], "vs_a100": mpts / A100_MPTS[(exp, n)],
                    "final": {name: [float(s), float(m)] for name, s, m in sums}})
    return {"workload": "reference example drivers, harness parameters, 200 steps, best of 3", "runs": out}


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
    ap.add_argument("--only", choices=["heat64_f64", "wave256_f64", "sod512x64x64_f32", "ensemble64_f32",
                                       "adjoint128_f32", "unstructured128_f64"],
                    help="run just one secondary workload and print its record (for profiling)")
    ap.add_argument("--no-suite", action="store_true",
                    help="skip the secondary workloads (heat FP64, wave FP64, Sod, ensemble, adjoint, "
                         "unstructured heat)")
    ap.add_argument("--drivers", action="store_true",
                    help="run only the reference's example drivers with the reference harness's parameters")
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

    if args.drivers:
        print(json.dumps(suite_drivers()), flush=True)
        return

    if args.only:
        ctx = capi.Context(0)
        fns = {"heat64_f64": lambda: suite_heat(ctx, cpu=False, verbose=args.tune_verbose), "wave256_f64": lambda: suite_wave_f64(ctx),
               "sod512x64x64_f32": lambda: suite_sod(ctx, cpu=False, verbose=args.tune_verbose),
               "ensemble64_f32": lambda: suite_ensemble(ctx, verbose=args.tune_verbose),
               "adjoint128_f32": lambda: suite_adjoint(ctx, verbose=args.tune_verbose),
               "unstructured128_f64": lambda: suite_unstructured(ctx)}
        rec = fns[args.only]()
        ctx.close()
        print(json.dumps({"workload_only": args.only, **rec}), flush=True)
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
    # the two-step kernels advance two steps per launch (in pairs of launches;
    # a K field or a step count that is not a multiple of 4 leaves single steps)
    two_step = shape["variant"] >= 11 and args.steps % 4 == 0
    steps_per_launch = 2 if two_step else 1
    launch_ms = kernel_ms * steps_per_launch
    # algorithmic bytes: 32 B (FP32) per cell-update, so per launch 8 fields x N
    # x the steps one launch advances
    bytes_per_launch = 8 * esize * N * steps_per_launch
    achieved = bytes_per_launch / (launch_ms * 1e-3) / 1e9
    traffic = _traffic(args.n, args.dtype, shape)
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
                     "kernel": KERNELS[shape["variant"]], "kernel_ms": launch_ms,
                     "steps_per_launch": steps_per_launch,
                     **({"note": "two steps per launch: every field crosses HBM once per two steps, so the rate at "
                                 "32 B per cell-update can pass what a one-step kernel streams"}
                        if steps_per_launch == 2 else {}),
                     "algorithmic_bytes_per_launch": bytes_per_launch},
    }
    mesh.close()
    del A, B, cur, nxt, res
    torch.cuda.empty_cache()
    ceil = copy_ceiling()
    out["roofline"]["copy_ceiling_gbs"] = ceil
    out["roofline"]["frac_of_copy_ceiling"] = achieved / ceil
    sc, sdesc = stream_ceiling()
    out["roofline"]["stream_ceiling_gbs"] = sc
    out["roofline"]["stream_ceiling"] = sdesc
    out["roofline"]["frac_of_stream_ceiling"] = achieved / sc
    if not args.no_suite:
        suite = {}
        for name, fn in (("heat64_f64", lambda: suite_heat(ctx, cpu=not args.no_cpu)),
                         ("wave256_f64", lambda: suite_wave_f64(ctx)),
                         ("sod512x64x64_f32", lambda: suite_sod(ctx, cpu=not args.no_cpu,
                                                                verbose=args.tune_verbose)),
                         ("ensemble64_f32", lambda: suite_ensemble(ctx)),
                         ("adjoint128_f32", lambda: suite_adjoint(ctx)),
                         ("unstructured128_f64", lambda: suite_unstructured(ctx))):
            print(f"[bench] {name}", file=sys.stderr, flush=True)
            suite[name] = fn()
            torch.cuda.empty_cache()
        out["workloads"] = suite
        # dL/dp0 (the workload's gradient) per step against the fused forward
        # step; the dL/dK figure with recomputation beside it
        out["adjoint_forward_ratio"] = suite["adjoint128_f32"]["ratio"]
        out["adjoint_forward_ratio_with_K"] = suite["adjoint128_f32"]["ratio_with_K_recompute"]
        out["batched_vs_plain"] = suite["ensemble64_f32"]["speedup"]
    ctx.close()
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(case, args.dtype, args.cpu_steps)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
