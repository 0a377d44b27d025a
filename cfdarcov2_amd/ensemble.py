"""Batched ensemble of independent heat and acoustic cases (FP32).

The cases are packed into groups (each a cfd_batch_t): a group advances all
of its cases with one batched launch per step, and an n-step run of a group is
one HIP graph launch.  Groups run one after another; grouping by working set
keeps a group's fields resident in the GPU's last-level cache across steps.
Tile shape and grouping come from `autotune`.

The plain baseline (`run_sequential`) advances the same cases one at a time
with the single-case fused kernels (cfd_heat_run / cfd_wave_run).
"""
from __future__ import annotations

import ctypes
import time
from dataclasses import dataclass

import numpy as np

from . import capi


@dataclass
class _Case:
    spec: object             # workloads.EnsembleCase
    mesh: "capi.HexMesh"
    a: list                  # device tensors (heat: [T]; wave: [p, u, v, w])
    b: list

    @property
    def nbytes(self) -> int:
        return sum(t.numel() * t.element_size() for t in self.a + self.b)


class Ensemble:
    def __init__(self, ctx: "capi.Context", specs):
        import torch
        self.ctx = ctx
        self.lib = ctx.lib
        self.cases = []
        for sp in specs:
            m = capi.HexMesh(ctx, sp.n, sp.n, sp.n, sp.spacing, sp.spacing, sp.spacing, dtype="f32")
            if m.info().uniform != 1:
                raise ValueError("ensemble cases need exactly uniform meshes (dyadic spacing)")
            f0 = torch.from_numpy(np.ascontiguousarray(sp.field0, dtype=np.float32)).cuda()
            if sp.kind == "heat":
                a, b = [f0], [torch.empty_like(f0)]
            else:
                a = [f0] + [torch.zeros_like(f0) for _ in range(3)]
                b = [torch.empty_like(f0) for _ in range(4)]
            self.cases.append(_Case(sp, m, a, b))
        self.initial = [[t.clone() for t in c.a] for c in self.cases]
        self.groups = []
        self.shape = (4, 8)
        self.group_bytes = 0

    # ---- state --------------------------------------------------------------
    def reset(self):
        for c, init in zip(self.cases, self.initial):
            for t, i in zip(c.a, init):
                t.copy_(i)

    def state(self, i, nsteps):
        c = self.cases[i]
        return c.a if nsteps % 2 == 0 else c.b

    # ---- grouping -----------------------------------------------------------
    def _destroy_groups(self):
        for h in self.groups:
            capi.check(self.lib.cfd_batch_destroy(h))
        self.groups = []

    def configure(self, tile_rows=4, chunk=8, group_bytes=0, form=0):
        """Pack cases into groups of at most `group_bytes` of fields (0: one
        group), largest first, and set the tile form (0: 2.5D march tiles, 1:
        row form, 2: row form with two steps per launch) and shape."""
        self._destroy_groups()
        order = sorted(range(len(self.cases)), key=lambda i: -self.cases[i].nbytes)
        packs, cur, cur_b = [], [], 0
        for i in order:
            nb = self.cases[i].nbytes
            if cur and group_bytes and cur_b + nb > group_bytes:
                packs.append(cur)
                cur, cur_b = [], 0
            cur.append(i)
            cur_b += nb
        if cur:
            packs.append(cur)
        for pack in packs:
            arr = (capi.BatchCase * len(pack))()
            for k, i in enumerate(pack):
                c = self.cases[i]
                arr[k].mesh = c.mesh.h
                arr[k].kind = 0 if c.spec.kind == "heat" else 1
                arr[k].coef = float(c.spec.coef)
                arr[k].dt = float(c.spec.dt)
                for f in range(4):
                    arr[k].state_a[f] = c.a[f].data_ptr() if f < len(c.a) else None
                    arr[k].state_b[f] = c.b[f].data_ptr() if f < len(c.b) else None
            h = ctypes.c_void_p()
            capi.check(self.lib.cfd_batch_create(self.ctx.h, arr, len(pack), ctypes.byref(h)))
            self.groups.append(h)
            capi.check(self.lib.cfd_batch_set_form(h, int(form)))
            capi.check(self.lib.cfd_batch_set_shape(h, int(tile_rows), int(chunk)))
        self.shape = (tile_rows, chunk)
        self.form = form
        self.group_bytes = group_bytes
        self.packs = packs

    # ---- runs ---------------------------------------------------------------
    def run(self, nsteps, use_graph=True):
        if not self.groups:
            self.configure()
        for h in self.groups:
            capi.check(self.lib.cfd_batch_run(h, int(nsteps), int(bool(use_graph)), capi._stream_ptr()))

    def run_sequential(self, nsteps):
        """The plain baseline: every case alone through its single-case fused run."""
        for c in self.cases:
            if c.spec.kind == "heat":
                c.mesh.heat_run(c.a[0], c.b[0], c.spec.coef, c.spec.dt, nsteps)
            else:
                c.mesh.wave_run(c.a, c.b, c.spec.coef, c.spec.dt, nsteps)

    def cell_updates(self, nsteps) -> int:
        return sum(c.spec.n ** 3 for c in self.cases) * int(nsteps)

    def algorithmic_bytes(self, nsteps) -> int:
        per = sum(c.spec.n ** 3 * (8 if c.spec.kind == "heat" else 32) for c in self.cases)
        return per * int(nsteps)

    def close(self):
        self._destroy_groups()
        for c in self.cases:
            c.mesh.close()


def time_gpu(fn, reps=1):
    """Wall time of fn() bracketed by device synchronisations, best of reps."""
    import torch
    best = float("inf")
    for _ in range(reps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fn()
        torch.cuda.synchronize()
        best = min(best, time.perf_counter() - t0)
    return best


def autotune(ens: Ensemble, steps=20, candidates=None, verbose=False):
    """Pick (tile_rows, chunk, group_bytes) by timing graph runs of `steps`
    steps; leaves the ensemble configured with the winner and its state reset."""
    if candidates is None:
        candidates = [(ty, L, gb, 0) for ty in (4, 8, 16) for L in (8, 16, 32)
                      for gb in (0, 96 << 20, 192 << 20)]
        candidates += [(ty, L, gb, 1) for ty in (4, 8) for L in (4, 8, 16, 32)
                       for gb in (0, 96 << 20, 192 << 20)]
        candidates += [(nw, L, gb, 2) for nw in (8, 10, 12) for L in (8, 16, 32)
                       for gb in (0, 192 << 20)]
    results = []
    for cand in candidates:
        ty, L, gb = cand[:3]
        form = cand[3] if len(cand) > 3 else 0
        try:
            ens.configure(ty, L, gb, form)
        except capi.CfdError:  # a form these cases do not allow
            continue
        ens.run(steps)  # capture + warm
        t = time_gpu(lambda: ens.run(steps), reps=2)
        results.append((t, ty, L, gb, form))
        if verbose:
            print(f"  form={form} tile_rows={ty:2d} chunk={L:2d} group={gb >> 20:4d} MiB: "
                  f"{t / steps * 1e6:8.1f} us/step", flush=True)
    results.sort()
    _, ty, L, gb, form = results[0]
    ens.configure(ty, L, gb, form)
    ens.reset()
    return {"form": form, "tile_rows": ty, "chunk": L, "group_mib": gb >> 20,
            "us_per_step": results[0][0] / steps * 1e6}
