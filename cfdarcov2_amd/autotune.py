"""Autotuner for the fused kernels' block and tile shapes.

For a given mesh it times every candidate shape of the fused acoustic step
(per-thread kernel; 2.5D marching kernel over tile rows x x-chunk x XCD
placement) with HIP events on the launch stream, keeps the fastest, and caches
the choice per (kernel, dtype, mesh extents) in a JSON file
(``$CFDARCO_TUNE_CACHE``, default ``~/.cache/cfdarco/tune.json``).  Results are
state-independent (the kernels do the same work for any data), so tuning runs
on a scratch state and never touches the caller's fields.
"""
from __future__ import annotations

import itertools
import json
import os

DEFAULT_CANDIDATES = [(0, 16, 32, 1)] + [
    (1, ty, L, r) for ty, L, r in itertools.product((4, 8, 16, 32), (4, 8, 16, 32, 64), (0, 1))
] + [
    (v, ty, L, r) for v, ty, L, r in itertools.product((7, 8, 9), (4, 8, 14), (4, 8, 16, 32, 64), (0, 1, 2, 3))
] + [
    (v, ty, L, r) for v, ty, L, r in itertools.product((11, 12), (2, 4, 8, 10), (8, 16, 32, 64), (0, 1, 2, 3))
]


def _cache_path():
    return os.environ.get("CFDARCO_TUNE_CACHE", os.path.expanduser("~/.cache/cfdarco/tune.json"))


def _load_cache():
    try:
        with open(_cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store_cache(cache):
    path = _cache_path()
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        with open(path, "w") as f:
            json.dump(cache, f, indent=1, sort_keys=True)
    except OSError:
        pass


def wave_key(mesh) -> str:
    s = mesh.shape
    return f"wave/{'f32' if mesh.code == 0 else 'f64'}/{s.nx}x{s.ny}x{s.nz}"


def tune_wave(mesh, steps: int = 20, candidates=None, use_cache: bool = True, verbose: bool = False) -> dict:
    """Pick the fastest fused acoustic-step shape for `mesh` and apply it."""
    import torch

    key = wave_key(mesh)
    cache = _load_cache() if use_cache else {}
    if use_cache and key in cache:
        best = cache[key]
        mesh.set_wave_shape(best["variant"], best["tile_rows"], best["chunk"], best["xcd_remap"])
        return best
    cands = candidates or DEFAULT_CANDIDATES
    N = mesh.N
    tdt = mesh.torch_dtype
    A = [torch.zeros(N, dtype=tdt, device="cuda") for _ in range(4)]
    B = [torch.zeros(N, dtype=tdt, device="cuda") for _ in range(4)]
    stream = torch.cuda.current_stream()
    table = []
    from .capi import CfdError
    for shape in cands:
        variant, ty, L, remap = shape
        try:
            mesh.set_wave_shape(variant, ty, L, remap)
        except CfdError:  # a form this mesh's extents do not allow
            continue
        mesh.wave_run(A, B, -1.0, 1e-3, 2)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        mesh.wave_run(A, B, -1.0, 1e-3, steps)
        e1.record(stream)
        e1.synchronize()
        ms = e0.elapsed_time(e1) / steps
        table.append({"variant": variant, "tile_rows": ty, "chunk": L, "xcd_remap": remap, "ms_per_step": ms})
        if verbose:
            print(f"  wave shape variant={variant} ty={ty} L={L} remap={remap}: {ms * 1e3:.1f} us/step", flush=True)
    best = min(table, key=lambda r: r["ms_per_step"])
    mesh.set_wave_shape(best["variant"], best["tile_rows"], best["chunk"], best["xcd_remap"])
    if use_cache:
        cache[key] = best
        _store_cache(cache)
    del A, B
    return dict(best, table=table)


EULER_CANDIDATES = [(0, 4, 16)] + [(1, ty, L) for ty in (2, 4, 8) for L in (8, 16, 32, 64)]


def tune_euler(mesh, U, gamma=1.4, cfl=0.3, steps=10, candidates=None, verbose=False) -> dict:
    """Pick the fastest Rusanov-step shape for `mesh` from runs on a copy of
    the state U (list of five device fields); the caller's fields are not
    touched.  Returns the best shape and the table."""
    import torch

    from .capi import CfdError
    cands = candidates or EULER_CANDIDATES
    A = [u.clone() for u in U]
    B = [torch.empty_like(u) for u in U]
    stream = torch.cuda.current_stream()
    table = []
    for variant, ty, L in cands:
        try:
            mesh.set_euler_shape(variant, ty, L)
        except CfdError:   # a shape this mesh's precision does not offer
            continue
        for a, u in zip(A, U):
            a.copy_(u)
        mesh.euler_run(A, B, gamma, cfl, 2)
        for a, u in zip(A, U):
            a.copy_(u)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        mesh.euler_run(A, B, gamma, cfl, steps)
        e1.record(stream)
        e1.synchronize()
        ms = e0.elapsed_time(e1) / steps
        table.append({"variant": variant, "tile_rows": ty, "chunk": L, "ms_per_step": ms})
        if verbose:
            print(f"  euler shape variant={variant} ty={ty} L={L}: {ms * 1e3:.1f} us/step", flush=True)
    best = min(table, key=lambda r: r["ms_per_step"])
    mesh.set_euler_shape(best["variant"], best["tile_rows"], best["chunk"])
    del A, B
    return dict(best, table=table)


HEAT_CANDIDATES = [(0, 8, 16)] + [(1, w, L) for w in (8, 12, 16) for L in (1, 2, 4, 8, 16)] + [
    (2, B, K) for B in (4, 8) for K in (2, 4)]


def tune_heat(mesh, alpha, dt, steps=100, candidates=None, verbose=False) -> dict:
    """Pick the fastest heat-step form for `mesh` (per-thread, two steps per
    launch in row form with its waves and chunk, or several steps per launch
    in LDS boxes) from runs on scratch fields, each timed twice (the faster
    run counts: a step can be as short as a few microseconds); forms the mesh
    does not allow are skipped."""
    import torch

    from .capi import CfdError
    cands = candidates or HEAT_CANDIDATES
    N = mesh.N
    tdt = mesh.torch_dtype
    a = torch.zeros(N, dtype=tdt, device="cuda")
    b = torch.zeros(N, dtype=tdt, device="cuda")
    stream = torch.cuda.current_stream()
    table = []
    for variant, nw, L in cands:
        try:
            mesh.set_heat_shape(variant, nw, L)
        except CfdError:
            continue
        mesh.heat_run(a, b, alpha, dt, 4)
        ms = float("inf")
        for _ in range(2):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            mesh.heat_run(a, b, alpha, dt, steps)
            e1.record(stream)
            e1.synchronize()
            ms = min(ms, e0.elapsed_time(e1) / steps)
        table.append({"variant": variant, "waves": nw, "chunk": L, "ms_per_step": ms})
        if verbose:
            print(f"  heat shape variant={variant} waves={nw} chunk={L}: {ms * 1e3:.2f} us/step", flush=True)
    best = min(table, key=lambda r: r["ms_per_step"])
    mesh.set_heat_shape(best["variant"], best["waves"], best["chunk"])
    del a, b
    return dict(best, table=table)


ADJOINT_CANDIDATES = [(0, 4, 8)] + [(1, ty, L) for ty in (4, 8) for L in (2, 4, 8, 16, 32)]


def tune_adjoint(problem, steps=40, candidates=None, verbose=False, repeats=3) -> dict:
    """Pick the fastest kernel shape of the transposed acoustic step for the
    problem's mesh (constant K, no K gradient: the dL/dp0 path), from runs on
    scratch fields (best of `repeats` timed runs per shape, the shapes being
    a few per cent apart); forms the mesh does not allow are skipped."""
    import torch

    from . import adjoint
    from .capi import CfdError
    cands = candidates or ADJOINT_CANDIDATES
    mesh = problem.mesh
    short = adjoint.WaveProblem(mesh, problem.sensors, problem.dt, steps, K=problem.K, every=problem.every)
    g = torch.zeros((steps, problem.sensors.size), dtype=mesh.torch_dtype, device="cuda")
    stream = torch.cuda.current_stream()
    table = []
    for form, ty, L in cands:
        try:
            mesh.set_adjoint_shape(form, ty, L)
        except CfdError:
            continue
        adjoint.adjoint(short, g)
        ms = float("inf")
        for _ in range(repeats):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            adjoint.adjoint(short, g)
            e1.record(stream)
            e1.synchronize()
            ms = min(ms, e0.elapsed_time(e1) / steps)
        table.append({"form": form, "tile_rows": ty, "chunk": L, "ms_per_step": ms})
        if verbose:
            print(f"  adjoint shape form={form} ty={ty} L={L}: {ms * 1e3:.1f} us/step", flush=True)
    best = min(table, key=lambda r: r["ms_per_step"])
    mesh.set_adjoint_shape(best["form"], best["tile_rows"], best["chunk"])
    return dict(best, table=table)
