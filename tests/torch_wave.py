"""Plain FP64 torch restatement of the fused acoustic step on a uniform mesh,
for reference gradients by automatic differentiation (test infrastructure; it
is pinned against the oracle's wave_run in test_adjoint_cpu.py).

Step, in the oracle's order (oracle/cfd_oracle.hpp wave_run):
  u' = dt*(-Gx p) + u, v' = dt*(-Gy p) + v, w' = dt*(-Gz p) + w
  p' = dt*(K*((Dx u' + Dy v') + Dz w')) + p
G: linear face values with neighbour = self on the walls; D: zero face value
on the walls normal to the axis; face pair sum ((-f_lo) A + f_hi A) / V."""
import torch


def _shift(q, axis, lo: bool, wall_zero: bool):
    """Face value on the lower (lo) or upper face along `axis` of every cell."""
    n = q.shape[axis]
    idx = torch.arange(n, device=q.device)
    nb = (idx - 1).clamp(min=0) if lo else (idx + 1).clamp(max=n - 1)
    qn = q.index_select(axis, nb)
    f = (q + qn) * 0.5
    if wall_zero:
        edge = 0 if lo else n - 1
        mask = torch.ones(n, dtype=q.dtype, device=q.device)
        mask[edge] = 0.0
        shape = [1, 1, 1]
        shape[axis] = n
        f = f * mask.reshape(shape)
    return f


def grad_axis(q, axis, A, V, wall_zero=False):
    f_lo = _shift(q, axis, True, wall_zero)
    f_hi = _shift(q, axis, False, wall_zero)
    return ((-f_lo) * A + f_hi * A) / V


def step(p, u, v, w, K, dt, area, vol):
    """One fused step on (X, Y, Z) tensors; K scalar or (X, Y, Z) tensor."""
    u = dt * (-grad_axis(p, 0, area[0], vol)) + u
    v = dt * (-grad_axis(p, 1, area[1], vol)) + v
    w = dt * (-grad_axis(p, 2, area[2], vol)) + w
    S = (grad_axis(u, 0, area[0], vol, True) + grad_axis(v, 1, area[1], vol, True)) + grad_axis(w, 2, area[2], vol,
                                                                                              True)
    p = dt * (K * S) + p
    return p, u, v, w


def traces(p0, K, dt, steps, sensors, shape, spacing):
    """Pressure at the sensor cells (flat ids) after every step: (steps, n)."""
    X, Y, Z = shape
    dx, dy, dz = spacing
    area = (dy * dz, dx * dz, dx * dy)
    vol = dx * dy * dz
    p = p0.reshape(X, Y, Z)
    u = torch.zeros_like(p)
    v = torch.zeros_like(p)
    w = torch.zeros_like(p)
    Kt = K.reshape(X, Y, Z) if torch.is_tensor(K) and K.dim() > 0 else K
    out = []
    for _ in range(steps):
        p, u, v, w = step(p, u, v, w, Kt, dt, area, vol)
        out.append(p.reshape(-1)[sensors])
    return torch.stack(out)
