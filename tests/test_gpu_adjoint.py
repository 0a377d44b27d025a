"""GPU checks of the hand-written adjoint (cfd_wave_traces / cfd_wave_adjoint
behind torch.autograd.Function):

* the traced forward equals the plain fused step bit for bit;
* the transposed step satisfies <S x, y> = <x, S^T y> to FP64 round-off;
* dL/dp0 and dL/dK equal autograd through the plain FP64 torch stepper;
* on the full workload (128^3, 2000 steps, checkpoints every 45, 32 sensors)
  dL/dp0 matches FP64 central finite differences on 16 random directions
  (h = 1e-2 and 1e-3 of |p0|) to 1e-4 relative, and the FP32 gradient is
  reported against the FP64 one;
* dL/dK through checkpoint-and-recompute against finite differences."""
import numpy as np
import pytest

from cfdarcov2_amd import workloads
from tests import torch_wave

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _problem(hip_ctx, n, d, dt, steps, sensors, every=0, dtype=np.float64, K=-1.0):
    from cfdarcov2_amd import adjoint, capi
    m = capi.HexMesh(hip_ctx, *n, *d, dtype=dtype)
    assert m.info().uniform == 1
    return adjoint.WaveProblem(m, np.asarray(sensors), dt, steps, K=K, every=every)


def test_traced_forward_equals_fused_step(hip_ctx):
    from cfdarcov2_amd import adjoint
    n, d = (20, 18, 70), (1 / 16, 1 / 16, 1 / 16)
    N = 20 * 18 * 70
    rng = np.random.default_rng(0)
    sensors = rng.choice(N, 40, replace=False)
    pr = _problem(hip_ctx, n, d, 0.02, 9, sensors, every=4)
    p0 = torch.from_numpy(rng.standard_normal(N)).cuda()
    print("\n[case] adjoint traced forward", flush=True)
    tr, final, ck = adjoint.forward_traces(pr, p0)
    A = [p0.clone()] + [torch.zeros_like(p0) for _ in range(3)]
    B = [torch.empty_like(p0) for _ in range(4)]
    pr.mesh.set_wave_shape(0, 4, 1, 0)
    ref = pr.mesh.wave_run(A, B, -1.0, 0.02, 9)
    torch.cuda.synchronize()
    for a, b in zip(final, ref):
        assert torch.equal(a, b)
    assert torch.equal(tr[-1], ref[0][torch.from_numpy(sensors).cuda()])
    assert torch.equal(ck[0, 0], p0) and torch.all(ck[0, 1:] == 0)
    assert ck.shape[0] == 3


def test_transpose_identity(hip_ctx):
    from cfdarcov2_amd import capi
    n, d = (9, 11, 13), (0.125, 0.25, 0.0625)
    N = 9 * 11 * 13
    m = capi.HexMesh(hip_ctx, *n, *d, dtype=np.float64)
    rng = np.random.default_rng(5)
    Kf = torch.from_numpy(-(0.5 + rng.random(N))).cuda()
    for steps in (1, 4):
        x = [torch.from_numpy(rng.standard_normal(N)).cuda() for _ in range(4)]
        y = [torch.from_numpy(rng.standard_normal(N)).cuda() for _ in range(4)]
        xa = [t.clone() for t in x]
        xb = [torch.empty_like(t) for t in x]
        Sx = m.wave_run(xa, xb, 0.0, 0.01, steps, K_field=Kf)
        ya = [t.clone() for t in y]
        yb = [torch.empty_like(t) for t in y]
        arr = lambda ts: (capi.ctypes.c_void_p * 4)(*[t.data_ptr() for t in ts])
        capi.check(m.lib.cfd_wave_adjoint(m.h, arr(ya), arr(yb), 0.0, capi._dptr(Kf), 0.01, steps, None, 0, None,
                                          None, 1, None, capi._stream_ptr()))
        STy = ya if steps % 2 == 0 else yb
        lhs = sum(float(torch.dot(a, b)) for a, b in zip(Sx, y))
        rhs = sum(float(torch.dot(a, b)) for a, b in zip(x, STy))
        scale = sum(float(a.norm() * b.norm()) for a, b in zip(Sx, y))
        assert abs(lhs - rhs) <= 1e-13 * scale, (steps, lhs, rhs)


def test_gradients_match_autograd_of_plain_stepper(hip_ctx):
    from cfdarcov2_amd import adjoint
    n, d = (10, 9, 12), (0.125, 0.125, 0.0625)
    N = 10 * 9 * 12
    rng = np.random.default_rng(8)
    sensors = rng.choice(N, 7, replace=False)
    steps, dt = 23, 0.03
    obs = rng.standard_normal((steps, 7))
    p0 = rng.standard_normal(N)
    Kf = -(0.5 + rng.random(N))
    pr = _problem(hip_ctx, n, d, dt, steps, sensors, every=5)
    # device adjoint
    tp = torch.from_numpy(p0).cuda().requires_grad_(True)
    tk = torch.from_numpy(Kf).cuda().requires_grad_(True)
    L = ((adjoint.wave_traces(tp, tk, pr) - torch.from_numpy(obs).cuda()) ** 2).sum()
    L.backward()
    # autograd through the plain FP64 stepper on the CPU
    cp = torch.from_numpy(p0).requires_grad_(True)
    ck = torch.from_numpy(Kf).requires_grad_(True)
    Lr = ((torch_wave.traces(cp, ck, dt, steps, torch.from_numpy(sensors), n, d) - torch.from_numpy(obs)) ** 2).sum()
    Lr.backward()
    assert float(L.detach()) == pytest.approx(float(Lr.detach()), rel=1e-14)
    for got, ref in ((tp.grad, cp.grad), (tk.grad, ck.grad)):
        g = got.cpu().numpy()
        r = ref.numpy()
        assert np.linalg.norm(g - r) <= 1e-12 * np.linalg.norm(r)


def _loss(pr, p0, obs, K_field=None):
    from cfdarcov2_amd import adjoint
    tr, _, _ = adjoint.forward_traces(pr, p0, K_field, checkpoints=False)
    return float(((tr - obs) ** 2).sum())


def test_full_workload_gradient_against_finite_differences(hip_ctx):
    """128^3, 2000 steps, checkpoints every 45, 32 sensors."""
    from cfdarcov2_amd import adjoint
    cfg = workloads.adjoint_config()
    print(f"\n[case] adjoint full workload n={cfg.n} steps={cfg.steps}", flush=True)
    pr64 = _problem(hip_ctx, cfg.n, cfg.spacing, cfg.dt, cfg.steps, cfg.sensors, cfg.every, np.float64)
    obs, _, _ = adjoint.forward_traces(pr64, torch.from_numpy(cfg.p0_true).cuda(), checkpoints=False)
    p0 = torch.from_numpy(cfg.p0_guess).cuda().requires_grad_(True)
    L = ((adjoint.wave_traces(p0, None, pr64) - obs) ** 2).sum()
    L.backward()
    g = p0.grad.detach()
    rng = np.random.default_rng(16)
    pnorm = float(p0.detach().norm())
    worst = 0.0
    for k in range(16):
        e = torch.from_numpy(rng.standard_normal(p0.numel())).cuda()
        e = e / e.norm()
        dd = float(torch.dot(g, e))
        for hf in (1e-2, 1e-3):
            h = hf * pnorm
            fd = (_loss(pr64, p0.detach() + h * e, obs) - _loss(pr64, p0.detach() - h * e, obs)) / (2 * h)
            rel = abs(fd - dd) / max(abs(dd), 1e-300)
            worst = max(worst, rel)
            assert rel <= 1e-4, (k, hf, fd, dd)
    print(f"[adjoint] worst FD relative error over 16 directions x 2 steps: {worst:.3e}")
    # FP32 adjoint on the same workload, against the FP64 gradient
    pr32 = _problem(hip_ctx, cfg.n, cfg.spacing, cfg.dt, cfg.steps, cfg.sensors, cfg.every, np.float32)
    p32 = torch.from_numpy(cfg.p0_guess.astype(np.float32)).cuda().requires_grad_(True)
    L32 = ((adjoint.wave_traces(p32, None, pr32) - obs.float()) ** 2).sum()
    L32.backward()
    rel32 = float((p32.grad.double() - g).norm() / g.norm())
    print(f"[adjoint] FP32 gradient vs FP64: relative L2 {rel32:.3e}")
    assert rel32 < 1e-3


def test_K_gradient_with_recompute_against_finite_differences(hip_ctx):
    from cfdarcov2_amd import adjoint
    n = (32, 32, 32)
    d = (1 / 32,) * 3
    N = 32 ** 3
    rng = np.random.default_rng(21)
    sensors = rng.choice(N, 12, replace=False)
    steps = 150
    pr = _problem(hip_ctx, n, d, 0.4 / 32, steps, sensors, every=13)
    p0 = torch.from_numpy(workloads.gaussian(*n, *d, (0.4, 0.5, 0.6), 0.08)).cuda()
    K0 = -(1.0 + 0.3 * torch.from_numpy(rng.random(N))).cuda()
    obs = torch.from_numpy(rng.standard_normal((steps, 12)) * 0.01).cuda()
    K = K0.clone().requires_grad_(True)
    L = ((adjoint.wave_traces(p0, K, pr) - obs) ** 2).sum()
    L.backward()
    gK = K.grad.detach()
    for k in range(4):
        e = torch.from_numpy(rng.standard_normal(N)).cuda()
        e = e / e.norm()
        # L is not quadratic in K: a small step keeps the O(h^2) truncation
        # below the tolerance (FP64 round-off in L stays far smaller)
        h = 2e-6 * float(K0.norm())
        fd = (_loss(pr, p0, obs, K0 + h * e) - _loss(pr, p0, obs, K0 - h * e)) / (2 * h)
        dd = float(torch.dot(gK, e))
        assert fd == pytest.approx(dd, rel=1e-6), k


ROW_CASES = [((20, 18, 64), (4, 8)), ((13, 11, 256), (8, 3)), ((9, 30, 32), (4, 1)), ((17, 5, 4), (8, 16)),
             ((6, 7, 128), (4, 2))]


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("n,shape", ROW_CASES)
def test_row_form_equals_per_thread_form(hip_ctx, dtype, n, shape):
    """The row-form forward (traces recorded in the fused step) and transposed
    step (trace gradients added in the step, K gradient from one divergence
    field per recomputed step) against the per-thread kernels, bit for bit:
    traces, checkpoints, dL/dp0 and dL/dK, constant K, with sensors on the
    walls, the corners and chunk borders, an uneven last checkpoint segment
    and an odd step count."""
    from cfdarcov2_amd import adjoint
    N = n[0] * n[1] * n[2]
    d = (1 / 32,) * 3
    rng = np.random.default_rng(N)
    X, Y, Z = n
    cid = lambda x, y, z: (x * Y + y) * Z + z
    fixed = {0, N - 1, cid(X - 1, 0, Z - 1), cid(0, Y - 1, 0), cid(min(shape[1], X - 1), Y // 2, Z // 2),
             cid(X // 2, Y - 1, Z - 1), cid(X - 1, Y - 1, 3 % Z)}
    sens = sorted(fixed | set(rng.choice(N, 30, replace=False).tolist()))[:64]
    steps, dt, K = 11, 0.4 / 32, -1.3
    p0 = torch.from_numpy(rng.standard_normal(N).astype(dtype)).cuda()
    res = {}
    if dtype == np.float64:
        shape = (4, shape[1])  # FP64 row tiles have 4 output waves (the 8-row builds spill)
    for form in (0, 1):
        print(f"\n[case] adjoint row vs per-thread n={n} shape={shape} dtype={dtype.__name__} form={form}", flush=True)
        pr = _problem(hip_ctx, n, d, dt, steps, sens, every=4, dtype=dtype, K=K)
        pr.mesh.set_adjoint_shape(form, *shape)
        tr, final, ck = adjoint.forward_traces(pr, p0, checkpoints=True)
        g = torch.from_numpy(rng.standard_normal(tuple(tr.shape)).astype(dtype)).cuda() if form == 0 else res[0][4]
        gp, gK = adjoint.adjoint(pr, g, None, ck, want_K=True)
        gp2, _ = adjoint.adjoint(pr, g)      # no K gradient: no recomputation
        torch.cuda.synchronize()
        res[form] = (tr, [t.clone() for t in final], ck, gp.clone(), g, gK, gp2)
        pr.mesh.close()
    a, b = res[0], res[1]
    assert torch.equal(a[0], b[0])
    for x, y in zip(a[1], b[1]):
        assert torch.equal(x, y)
    assert torch.equal(a[2], b[2])
    assert torch.equal(a[3], b[3])
    assert torch.equal(a[5], b[5])
    assert torch.equal(a[6], b[6]) and torch.equal(a[3], a[6])


def test_constant_K_gradient_equals_K_field_gradient(hip_ctx):
    """dL/dK at a constant K through the row form (constant-K transposed
    steps) equals the gradient at the equal K field (per-thread transposed
    steps with a K field), bit for bit."""
    from cfdarcov2_amd import adjoint
    n, d = (24, 20, 32), (1 / 32,) * 3
    N = n[0] * n[1] * n[2]
    rng = np.random.default_rng(3)
    sens = rng.choice(N, 20, replace=False)
    K = -0.9
    pr = _problem(hip_ctx, n, d, 0.4 / 32, 16, sens, every=5, dtype=np.float64, K=K)
    p0 = torch.from_numpy(rng.standard_normal(N)).cuda()
    Kf = torch.full((N,), K, dtype=torch.float64, device="cuda")
    print("\n[case] constant-K gradient vs K-field gradient", flush=True)
    tr, _, ck = adjoint.forward_traces(pr, p0, checkpoints=True)
    trf, _, ckf = adjoint.forward_traces(pr, p0, Kf, checkpoints=True)
    assert torch.equal(tr, trf) and torch.equal(ck, ckf)
    g = torch.from_numpy(rng.standard_normal(tuple(tr.shape))).cuda()
    gp, gK = adjoint.adjoint(pr, g, None, ck, want_K=True)
    gpf, gKf = adjoint.adjoint(pr, g, Kf, ckf, want_K=True)
    torch.cuda.synchronize()
    assert torch.equal(gp, gpf) and torch.equal(gK, gKf)



def test_adjoint_shape_errors(hip_ctx):
    from cfdarcov2_amd import capi
    m64 = capi.HexMesh(hip_ctx, 8, 8, 64, 1 / 8, 1 / 8, 1 / 8, dtype=np.float64)
    with pytest.raises(capi.CfdError):   # FP64 row tiles: 4 output waves only (the 8-row builds spill)
        m64.set_adjoint_shape(1, 8, 8)
    m64.set_adjoint_shape(1, 4, 8)
    m64.close()
    m = capi.HexMesh(hip_ctx, 8, 8, 256, 1 / 8, 1 / 8, 1 / 8, dtype=np.float32)
    for bad in ((2, 4, 8), (1, 10, 8), (1, 2, 8), (1, 4, 0)):
        with pytest.raises(capi.CfdError):
            m.set_adjoint_shape(*bad)
    m.set_adjoint_shape(1, 8, 8)
    m.close()
