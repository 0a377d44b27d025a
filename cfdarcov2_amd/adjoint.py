"""Differentiable sensor traces of the acoustic system and of the heat
equation (uniform meshes).

``WaveTraces.apply(p0, K_field, problem)`` runs ``problem.steps`` fused
acoustic steps from (p0, 0, 0, 0) on the device and returns the pressure at
the sensor cells after every step, shape (steps, n_sensors).  Its backward
runs the hand-written transposed step (cfd_wave_adjoint) in reverse time and
returns dL/dp0 and, when the K field requires a gradient, dL/dK per cell;
the K gradient recomputes each checkpoint segment of the forward solve
(checkpoints every ``problem.every`` steps, sqrt(steps) by default).

``HeatTraces.apply(T0, alpha_field, problem)`` does the same for the heat
step (cfd_heat_traces / cfd_heat_adjoint): T at the sensors after every step,
dL/dT0 and, when the alpha field requires a gradient, dL/dalpha per cell.

Tensors are plumbing here: the arithmetic runs in libcfdarco_hip.so.
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass

import numpy as np

from . import capi


@dataclass
class WaveProblem:
    mesh: "capi.HexMesh"       # uniform mesh (FP32 or FP64)
    sensors: np.ndarray        # int64 cell ids (at most 64, distinct)
    dt: float
    steps: int
    K: float = -1.0            # -c^2 when no K field is given
    every: int = 0             # checkpoint interval; 0 = ceil(sqrt(steps))

    def __post_init__(self):
        self.sensors = np.ascontiguousarray(self.sensors, dtype=np.int64)
        if self.every <= 0:
            self.every = max(1, int(math.ceil(math.sqrt(self.steps))))

    @property
    def n_checkpoints(self) -> int:
        return (self.steps + self.every - 1) // self.every

    def _sens(self):
        return self.sensors.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), int(self.sensors.size)


def _arr4(ts):
    return (ctypes.c_void_p * 4)(*[t.data_ptr() for t in ts])


def forward_traces(problem: WaveProblem, p0, K_field=None, checkpoints: bool = True):
    """Forward solve from (p0, 0, 0, 0).  Returns (traces, final state,
    checkpoint tensor or None)."""
    import torch
    m = problem.mesh
    dtype = m.torch_dtype
    N = m.N
    p0 = p0.detach().to(dtype=dtype).contiguous()
    A = [p0.clone()] + [torch.zeros(N, dtype=dtype, device=p0.device) for _ in range(3)]
    B = [torch.empty_like(A[0]) for _ in range(4)]
    traces = torch.empty((problem.steps, problem.sensors.size), dtype=dtype, device=p0.device)
    ck = None
    if checkpoints:
        ck = torch.empty((problem.n_checkpoints, 4, N), dtype=dtype, device=p0.device)
    kf = None if K_field is None else K_field.detach().to(dtype=dtype).contiguous()
    sp, ns = problem._sens()
    capi.check(m.lib.cfd_wave_traces(m.h, _arr4(A), _arr4(B), float(problem.K), capi._dptr(kf), float(problem.dt),
                                     int(problem.steps), sp, ns, capi._dptr(traces), capi._dptr(ck),
                                     int(problem.every), capi._stream_ptr()))
    final = A if problem.steps % 2 == 0 else B
    return traces, final, ck


def adjoint(problem: WaveProblem, trace_grad, K_field=None, ck=None, want_K: bool = False):
    """Reverse solve.  Returns (dL/dp0, dL/dK or None)."""
    import torch
    m = problem.mesh
    dtype = m.torch_dtype
    N = m.N
    dev = trace_grad.device
    g = trace_grad.detach().to(dtype=dtype).contiguous()
    A = [torch.zeros(N, dtype=dtype, device=dev) for _ in range(4)]
    B = [torch.empty_like(A[0]) for _ in range(4)]
    kf = None if K_field is None else K_field.detach().to(dtype=dtype).contiguous()
    gK = torch.zeros(N, dtype=dtype, device=dev) if want_K else None
    if want_K and ck is None:
        raise ValueError("the K gradient needs the forward checkpoints")
    sp, ns = problem._sens()
    capi.check(m.lib.cfd_wave_adjoint(m.h, _arr4(A), _arr4(B), float(problem.K), capi._dptr(kf), float(problem.dt),
                                      int(problem.steps), sp, ns, capi._dptr(g), capi._dptr(ck),
                                      int(problem.every), capi._dptr(gK), capi._stream_ptr()))
    res = A if problem.steps % 2 == 0 else B
    return res[0], gK


def _wave_traces_function():
    import torch

    class WaveTraces(torch.autograd.Function):
        """traces = WaveTraces.apply(p0, K_field, problem); K_field may be None."""

        @staticmethod
        def forward(ctx, p0, K_field, problem):
            need_K = K_field is not None and K_field.requires_grad
            traces, _, ck = forward_traces(problem, p0, K_field, checkpoints=need_K)
            ctx.problem = problem
            ctx.need_K = need_K
            ctx.ck = ck
            ctx.save_for_backward(K_field if K_field is not None else torch.empty(0))
            ctx.has_K = K_field is not None
            return traces

        @staticmethod
        def backward(ctx, grad_traces):
            (kf,) = ctx.saved_tensors
            kf = kf if ctx.has_K else None
            gp, gK = adjoint(ctx.problem, grad_traces, kf, ctx.ck, want_K=ctx.need_K)
            ctx.ck = None
            return gp, gK, None

    return WaveTraces


_WT = None


def wave_traces(p0, K_field, problem: WaveProblem):
    """Differentiable sensor traces (see WaveTraces)."""
    global _WT
    if _WT is None:
        _WT = _wave_traces_function()
    return _WT.apply(p0, K_field, problem)


@dataclass
class HeatProblem:
    mesh: "capi.HexMesh"       # uniform mesh (FP32 or FP64)
    sensors: np.ndarray        # int64 cell ids (at most 64, distinct)
    dt: float
    steps: int
    alpha: float = 1.0         # diffusivity when no alpha field is given
    every: int = 0             # checkpoint interval; 0 = ceil(sqrt(steps))

    def __post_init__(self):
        self.sensors = np.ascontiguousarray(self.sensors, dtype=np.int64)
        if self.every <= 0:
            self.every = max(1, int(math.ceil(math.sqrt(self.steps))))

    @property
    def n_checkpoints(self) -> int:
        return (self.steps + self.every - 1) // self.every

    def _sens(self):
        return self.sensors.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), int(self.sensors.size)


def heat_forward_traces(problem: HeatProblem, T0, alpha_field=None, checkpoints: bool = True):
    """Forward heat solve from T0.  Returns (traces, final state, checkpoint
    tensor or None)."""
    import torch
    m = problem.mesh
    dtype = m.torch_dtype
    N = m.N
    a = T0.detach().to(dtype=dtype).contiguous().clone()
    b = torch.empty_like(a)
    traces = torch.empty((problem.steps, problem.sensors.size), dtype=dtype, device=a.device)
    ck = torch.empty((problem.n_checkpoints, N), dtype=dtype, device=a.device) if checkpoints else None
    af = None if alpha_field is None else alpha_field.detach().to(dtype=dtype).contiguous()
    sp, ns = problem._sens()
    capi.check(m.lib.cfd_heat_traces(m.h, capi._dptr(a), capi._dptr(b), float(problem.alpha), capi._dptr(af),
                                     float(problem.dt), int(problem.steps), sp, ns, capi._dptr(traces),
                                     capi._dptr(ck), int(problem.every), capi._stream_ptr()))
    return traces, (a if problem.steps % 2 == 0 else b), ck


def heat_adjoint(problem: HeatProblem, trace_grad, alpha_field=None, ck=None, want_alpha: bool = False,
                 final_grad=None):
    """Reverse heat solve.  Returns (dL/dT0, dL/dalpha or None)."""
    import torch
    m = problem.mesh
    dtype = m.torch_dtype
    N = m.N
    dev = trace_grad.device if trace_grad is not None else final_grad.device
    g = None if trace_grad is None else trace_grad.detach().to(dtype=dtype).contiguous()
    a = (torch.zeros(N, dtype=dtype, device=dev) if final_grad is None
         else final_grad.detach().to(dtype=dtype).contiguous().clone())
    b = torch.empty_like(a)
    af = None if alpha_field is None else alpha_field.detach().to(dtype=dtype).contiguous()
    gA = torch.zeros(N, dtype=dtype, device=dev) if want_alpha else None
    if want_alpha and ck is None:
        raise ValueError("the alpha gradient needs the forward checkpoints")
    sp, ns = problem._sens()
    capi.check(m.lib.cfd_heat_adjoint(m.h, capi._dptr(a), capi._dptr(b), float(problem.alpha), capi._dptr(af),
                                      float(problem.dt), int(problem.steps), sp, ns, capi._dptr(g), capi._dptr(ck),
                                      int(problem.every), capi._dptr(gA), capi._stream_ptr()))
    return (a if problem.steps % 2 == 0 else b), gA


def _heat_traces_function():
    import torch

    class HeatTraces(torch.autograd.Function):
        """traces = HeatTraces.apply(T0, alpha_field, problem); alpha_field may be None."""

        @staticmethod
        def forward(ctx, T0, alpha_field, problem):
            need_a = alpha_field is not None and alpha_field.requires_grad
            traces, _, ck = heat_forward_traces(problem, T0, alpha_field, checkpoints=need_a)
            ctx.problem = problem
            ctx.need_a = need_a
            ctx.ck = ck
            ctx.has_a = alpha_field is not None
            ctx.save_for_backward(alpha_field if alpha_field is not None else torch.empty(0))
            return traces

        @staticmethod
        def backward(ctx, grad_traces):
            (af,) = ctx.saved_tensors
            af = af if ctx.has_a else None
            gT, gA = heat_adjoint(ctx.problem, grad_traces, af, ctx.ck, want_alpha=ctx.need_a)
            ctx.ck = None
            return gT, gA, None

    return HeatTraces


_HT = None


def heat_traces(T0, alpha_field, problem: HeatProblem):
    """Differentiable heat sensor traces (see HeatTraces)."""
    global _HT
    if _HT is None:
        _HT = _heat_traces_function()
    return _HT.apply(T0, alpha_field, problem)
