"""GPU checks of the adjoint of the heat step (cfd_heat_traces /
cfd_heat_adjoint behind torch.autograd.Function):

* the traces equal the oracle's heat_run sampled at the sensors, bit for bit;
* the transposed steps satisfy <H^n x, y> = <x, (H^T)^n y> to FP64 round-off,
  with a constant diffusivity and with a diffusivity field;
* dL/dT0 and dL/dalpha equal autograd through dense FP64 matrices built from
  the oracle's Laplacian (H = I + dt diag(k) L, columns L e_j);
* dL/dT0 and dL/dalpha match central finite differences;
* the FP32 gradient is close to the FP64 one."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

N_SMALL = (6, 5, 7)
D_SMALL = (1 / 8, 1 / 8, 1 / 8)


def _problem(hip_ctx, n, d, dt, steps, sensors, every=0, dtype=np.float64, alpha=1.0):
    from cfdarcov2_amd import adjoint, capi
    m = capi.HexMesh(hip_ctx, *n, *d, dtype=dtype)
    assert m.info().uniform == 1
    return adjoint.HeatProblem(m, np.asarray(sensors), dt, steps, alpha=alpha, every=every)


def _dense_L(oracle_lib, n, d):
    om = oracle_lib.OracleMesh(*n, *d, dtype=np.float64)
    N = om.N
    L = np.zeros((N, N))
    for j in range(N):
        e = np.zeros(N)
        e[j] = 1.0
        L[:, j] = om.lapl(e, 7)
    return om, L


def _dense_traces(L, T0, k, dt, steps, sensors):
    """Forward of the dense FP64 model (differentiable in T0 and k)."""
    Lt = torch.from_numpy(L)
    T = T0
    out = []
    for _ in range(steps):
        T = T + dt * (k * (Lt @ T))
        out.append(T[torch.from_numpy(np.asarray(sensors))])
    return torch.stack(out)


def test_traces_equal_oracle(hip_ctx, oracle_lib):
    from cfdarcov2_amd import adjoint
    n, d = (12, 10, 9), (1 / 16,) * 3
    rng = np.random.default_rng(3)
    N = int(np.prod(n))
    sensors = rng.choice(N, 11, replace=False)
    om = oracle_lib.OracleMesh(*n, *d, dtype=np.float64)
    T0 = rng.standard_normal(N)
    alpha, dt, steps = 1e-2, 0.1 * d[0] ** 2 / 1e-2, 9
    print("\n[case] heat traces vs oracle", flush=True)
    for af in (None, 1e-2 * (1 + rng.random(N))):
        pr = _problem(hip_ctx, n, d, dt, steps, sensors, every=4, alpha=alpha)
        kf = None if af is None else torch.from_numpy(af).cuda()
        tr, final, ck = adjoint.heat_forward_traces(pr, torch.from_numpy(T0).cuda(), kf)
        tr = tr.cpu().numpy()
        for t in range(1, steps + 1):
            ref = om.heat_run(T0, alpha, dt, t, alpha_field=af)
            assert np.array_equal(tr[t - 1], ref[sensors]), t
        assert np.array_equal(final.cpu().numpy(), om.heat_run(T0, alpha, dt, steps, alpha_field=af))
        ckn = ck.cpu().numpy()
        for k in range(pr.n_checkpoints):
            assert np.array_equal(ckn[k], om.heat_run(T0, alpha, dt, k * pr.every, alpha_field=af)), k
        pr.mesh.close()


@pytest.mark.parametrize("field", [False, True])
def test_transpose_identity(hip_ctx, field):
    """<H^n x, y> = <x, (H^T)^n y> in FP64 (no sensors: adj_a = y)."""
    from cfdarcov2_amd import adjoint
    n, d = (12, 10, 9), (1 / 16,) * 3
    N = int(np.prod(n))
    rng = np.random.default_rng(5)
    dt, steps, alpha = 0.05 * d[0] ** 2, 7, 1.0
    pr = _problem(hip_ctx, n, d, dt, steps, [], alpha=alpha)
    kf = torch.from_numpy(1 + rng.random(N)).cuda() if field else None
    x = torch.from_numpy(rng.standard_normal(N)).cuda()
    y = torch.from_numpy(rng.standard_normal(N)).cuda()
    print(f"\n[case] heat transpose identity field={field}", flush=True)
    _, Hx, _ = adjoint.heat_forward_traces(pr, x, kf, checkpoints=False)
    HTy, _ = adjoint.heat_adjoint(pr, None, kf, final_grad=y)
    lhs, rhs = float(torch.dot(Hx, y)), float(torch.dot(x, HTy))
    assert abs(lhs - rhs) <= 1e-12 * (abs(lhs) + float(Hx.abs().sum() * y.abs().max())), (lhs, rhs)
    pr.mesh.close()


def test_gradients_equal_dense_autograd(hip_ctx, oracle_lib):
    from cfdarcov2_amd import adjoint
    n, d = N_SMALL, D_SMALL
    N = int(np.prod(n))
    rng = np.random.default_rng(9)
    sensors = rng.choice(N, 6, replace=False)
    dt, steps = 0.05 * d[0] ** 2, 9
    om, L = _dense_L(oracle_lib, n, d)
    T0 = rng.standard_normal(N)
    k = 1 + rng.random(N)
    W = rng.standard_normal((steps, sensors.size))
    print("\n[case] heat adjoint vs dense autograd", flush=True)
    # device: traces, loss = sum(W * tr) + 0.5 sum(tr^2)
    pr = _problem(hip_ctx, n, d, dt, steps, sensors, every=4)
    T0d = torch.from_numpy(T0).cuda().requires_grad_(True)
    kd = torch.from_numpy(k).cuda().requires_grad_(True)
    tr = adjoint.heat_traces(T0d, kd, pr)
    loss = (torch.from_numpy(W).cuda() * tr).sum() + 0.5 * (tr * tr).sum()
    loss.backward()
    # dense model on the host
    T0h = torch.from_numpy(T0).requires_grad_(True)
    kh = torch.from_numpy(k).requires_grad_(True)
    trh = _dense_traces(L, T0h, kh, dt, steps, sensors)
    lh = (torch.from_numpy(W) * trh).sum() + 0.5 * (trh * trh).sum()
    lh.backward()
    np.testing.assert_allclose(tr.detach().cpu().numpy(), trh.detach().numpy(), rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(T0d.grad.cpu().numpy(), T0h.grad.numpy(), rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(kd.grad.cpu().numpy(), kh.grad.numpy(), rtol=1e-10, atol=1e-12)
    # constant diffusivity: no field, gradient in T0 only
    pr2 = _problem(hip_ctx, n, d, dt, steps, sensors, alpha=1.5)
    T0d.grad = None
    tr2 = adjoint.heat_traces(T0d, None, pr2)
    ((torch.from_numpy(W).cuda() * tr2).sum()).backward()
    T0h.grad = None
    trh2 = _dense_traces(L, T0h, torch.full((N,), 1.5, dtype=torch.float64), dt, steps, sensors)
    ((torch.from_numpy(W) * trh2).sum()).backward()
    np.testing.assert_allclose(T0d.grad.cpu().numpy(), T0h.grad.numpy(), rtol=1e-10, atol=1e-12)
    pr.mesh.close()
    pr2.mesh.close()


def test_gradients_match_finite_differences(hip_ctx):
    from cfdarcov2_amd import adjoint
    n, d = (10, 8, 12), (1 / 16,) * 3
    N = int(np.prod(n))
    rng = np.random.default_rng(13)
    sensors = rng.choice(N, 8, replace=False)
    dt, steps = 0.05 * d[0] ** 2, 20
    pr = _problem(hip_ctx, n, d, dt, steps, sensors, every=6)
    T0 = torch.from_numpy(rng.standard_normal(N)).cuda()
    k = torch.from_numpy(1 + rng.random(N)).cuda()
    obs = torch.from_numpy(rng.standard_normal((steps, sensors.size))).cuda()

    def loss_of(T, kk):
        tr, _, _ = adjoint.heat_forward_traces(pr, T, kk, checkpoints=False)
        return float(((tr - obs) ** 2).sum())

    print("\n[case] heat adjoint vs finite differences", flush=True)
    T0r = T0.clone().requires_grad_(True)
    kr = k.clone().requires_grad_(True)
    tr = adjoint.heat_traces(T0r, kr, pr)
    ((tr - obs) ** 2).sum().backward()
    for _ in range(4):
        for which in ("T0", "k"):
            v = torch.from_numpy(rng.standard_normal(N)).cuda()
            v = v / v.norm()
            h = 1e-6 * (T0.norm() if which == "T0" else k.norm())
            if which == "T0":
                fd = (loss_of(T0 + h * v, k) - loss_of(T0 - h * v, k)) / (2 * float(h))
                ad = float(torch.dot(T0r.grad, v))
            else:
                fd = (loss_of(T0, k + h * v) - loss_of(T0, k - h * v)) / (2 * float(h))
                ad = float(torch.dot(kr.grad, v))
            assert abs(fd - ad) <= 1e-6 * max(1.0, abs(ad)), (which, fd, ad)
    pr.mesh.close()


def test_fp32_gradient_close_to_fp64(hip_ctx):
    from cfdarcov2_amd import adjoint
    n, d = (16, 16, 16), (1 / 16,) * 3
    N = int(np.prod(n))
    rng = np.random.default_rng(17)
    sensors = rng.choice(N, 16, replace=False)
    dt, steps = 0.1 * d[0] ** 2, 30
    T0 = rng.standard_normal(N)
    W = rng.standard_normal((steps, sensors.size))
    grads = {}
    print("\n[case] heat adjoint FP32 vs FP64", flush=True)
    for dtype in (np.float64, np.float32):
        pr = _problem(hip_ctx, n, d, dt, steps, sensors, dtype=dtype)
        tdt = torch.float64 if dtype == np.float64 else torch.float32
        T = torch.from_numpy(T0).to(tdt).cuda().requires_grad_(True)
        tr = adjoint.heat_traces(T, None, pr)
        (torch.from_numpy(W).to(tdt).cuda() * tr).sum().backward()
        grads[dtype] = T.grad.double().cpu().numpy()
        pr.mesh.close()
    rel = np.linalg.norm(grads[np.float32] - grads[np.float64]) / np.linalg.norm(grads[np.float64])
    assert rel < 1e-5, rel


def test_heat_adjoint_errors(hip_ctx):
    from cfdarcov2_amd import adjoint, capi, workloads
    n, d = (6, 5, 7), (0.25, 0.25, 0.25)
    off = workloads.perturbed_vertex_offsets(*n, *d, seed=1, frac=0.1)
    m = capi.HexMesh(hip_ctx, *n, *d, dtype=np.float64, vertex_offsets=off)
    assert m.info().uniform == 0
    pr = adjoint.HeatProblem(m, np.array([1, 2]), 1e-3, 4)
    with pytest.raises(capi.CfdError) as e:
        adjoint.heat_forward_traces(pr, torch.zeros(m.N, dtype=torch.float64, device="cuda"))
    assert e.value.code == -4
    m.close()
    pr = _problem(hip_ctx, n, d, 1e-3, 4, [1, 2])
    with pytest.raises(capi.CfdError):   # the alpha gradient without checkpoints
        capi.check(pr.mesh.lib.cfd_heat_adjoint(
            pr.mesh.h, capi._dptr(torch.zeros(pr.mesh.N, dtype=torch.float64, device="cuda")),
            capi._dptr(torch.zeros(pr.mesh.N, dtype=torch.float64, device="cuda")), 1.0, None, 1e-3, 4,
            *pr._sens(), None, None, 0, capi._dptr(torch.zeros(pr.mesh.N, dtype=torch.float64, device="cuda")),
            capi._stream_ptr()))
    pr.mesh.close()
