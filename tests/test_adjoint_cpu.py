"""CPU pins behind the adjoint of the fused acoustic step: the FP64 torch
stepper equals the oracle's wave_run, the summation-by-parts identities the
transposed step rests on (G^T = -D, D^T = -G on a uniform mesh, by brute-force
matrices), and autograd through the torch stepper against central finite
differences."""
import numpy as np
import pytest
import torch

from cfdarcov2_amd import workloads
from tests import torch_wave


def test_torch_stepper_equals_oracle(oracle_lib):
    n, d = (7, 6, 5), (0.125, 0.25, 0.5)
    om = oracle_lib.OracleMesh(*n, *d, dtype=np.float64)
    rng = np.random.default_rng(1)
    p0 = rng.standard_normal(om.N)
    Kf = -(0.5 + rng.random(om.N))
    dt, steps = 0.01, 7
    ref = om.wave_run(p0, np.zeros(om.N), np.zeros(om.N), np.zeros(om.N), 0.0, dt, steps, K_field=Kf)
    sensors = torch.arange(om.N)
    tr = torch_wave.traces(torch.from_numpy(p0), torch.from_numpy(Kf), dt, steps, sensors, n, d)
    assert np.array_equal(tr[-1].numpy(), ref[0])


def _matrix(fn, n):
    cols = []
    for j in range(n):
        e = torch.zeros(n, dtype=torch.float64)
        e[j] = 1.0
        cols.append(fn(e))
    return torch.stack(cols, dim=1)


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_summation_by_parts(axis):
    shape = (5, 4, 3)
    A, V = 0.3, 0.7
    N = 60
    G = _matrix(lambda q: torch_wave.grad_axis(q.reshape(shape), axis, A, V).reshape(-1), N)
    D = _matrix(lambda q: torch_wave.grad_axis(q.reshape(shape), axis, A, V, True).reshape(-1), N)
    assert torch.allclose(G.T, -D, rtol=0, atol=1e-15)
    assert torch.allclose(D.T, -G, rtol=0, atol=1e-15)


def test_autograd_of_torch_stepper_against_fd():
    n, d = (6, 5, 4), (0.2, 0.2, 0.25)
    N = 120
    rng = np.random.default_rng(3)
    p0 = torch.from_numpy(rng.standard_normal(N)).requires_grad_(True)
    Kf = torch.from_numpy(-(0.5 + rng.random(N))).requires_grad_(True)
    obs = torch.from_numpy(rng.standard_normal((9, 4)))
    sensors = torch.tensor([3, 50, 77, 119])

    def loss(p, k):
        return ((torch_wave.traces(p, k, 0.03, 9, sensors, n, d) - obs) ** 2).sum()

    L = loss(p0, Kf)
    gp, gk = torch.autograd.grad(L, (p0, Kf))
    for g, which in ((gp, 0), (gk, 1)):
        e = torch.from_numpy(rng.standard_normal(N))
        e = e / e.norm()
        h = 1e-5
        args_p = [p0.detach(), Kf.detach()]
        args_m = [p0.detach(), Kf.detach()]
        args_p[which] = args_p[which] + h * e
        args_m[which] = args_m[which] - h * e
        fd = (loss(*args_p) - loss(*args_m)) / (2 * h)
        assert float(fd) == pytest.approx(float(g @ e), rel=1e-7, abs=1e-10)


def test_adjoint_config_is_seeded():
    a = workloads.adjoint_config(n=16, steps=4, every=2)
    b = workloads.adjoint_config(n=16, steps=4, every=2)
    assert np.array_equal(a.sensors, b.sensors) and np.array_equal(a.p0_true, b.p0_true)
    assert len(set(a.sensors.tolist())) == 32 and a.sensors.max() < 16 ** 3
