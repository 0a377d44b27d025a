"""Exact solution of the 1D Riemann problem for the ideal-gas Euler equations
(the textbook solver: Toro, "Riemann Solvers and Numerical Methods for Fluid
Dynamics", 3rd ed., ch. 4 - pressure function f_K, Newton iteration for the
star pressure, sampling of the self-similar solution).  Test infrastructure:
the closed form the Rusanov stepper's Sod runs are checked against."""
import numpy as np


def _fk(p, rho, pk, ak, g):
    if p > pk:  # shock
        A = 2.0 / ((g + 1.0) * rho)
        B = (g - 1.0) / (g + 1.0) * pk
        s = np.sqrt(A / (p + B))
        return (p - pk) * s, s * (1.0 - 0.5 * (p - pk) / (p + B))
    # rarefaction
    e = (g - 1.0) / (2.0 * g)
    f = 2.0 * ak / (g - 1.0) * ((p / pk) ** e - 1.0)
    df = 1.0 / (rho * ak) * (p / pk) ** (-(g + 1.0) / (2.0 * g))
    return f, df


def star_state(rl, ul, pl, rr, ur, pr, g):
    al, ar = np.sqrt(g * pl / rl), np.sqrt(g * pr / rr)
    p = max(1e-10, 0.5 * (pl + pr))
    for _ in range(100):
        fl, dfl = _fk(p, rl, pl, al, g)
        fr, dfr = _fk(p, rr, pr, ar, g)
        dp = (fl + fr + (ur - ul)) / (dfl + dfr)
        pn = max(1e-12, p - dp)
        if abs(pn - p) / (0.5 * (pn + p)) < 1e-14:
            p = pn
            break
        p = pn
    fl, _ = _fk(p, rl, pl, al, g)
    fr, _ = _fk(p, rr, pr, ar, g)
    u = 0.5 * (ul + ur) + 0.5 * (fr - fl)
    return p, u


def sample(xi, rl, ul, pl, rr, ur, pr, g):
    """(rho, u, p) at similarity coordinates xi = (x - x0) / t."""
    ps, us = star_state(rl, ul, pl, rr, ur, pr, g)
    al, ar = np.sqrt(g * pl / rl), np.sqrt(g * pr / rr)
    xi = np.asarray(xi, dtype=np.float64)
    rho, u, p = np.empty_like(xi), np.empty_like(xi), np.empty_like(xi)
    gm, gp = g - 1.0, g + 1.0
    for i, s in enumerate(xi):
        if s <= us:  # left of the contact
            if ps > pl:  # left shock
                sl = ul - al * np.sqrt(gp / (2 * g) * ps / pl + gm / (2 * g))
                if s <= sl:
                    rho[i], u[i], p[i] = rl, ul, pl
                else:
                    rho[i] = rl * ((ps / pl + gm / gp) / (gm / gp * ps / pl + 1.0))
                    u[i], p[i] = us, ps
            else:  # left rarefaction
                shl = ul - al
                asl = al * (ps / pl) ** (gm / (2 * g))
                stl = us - asl
                if s <= shl:
                    rho[i], u[i], p[i] = rl, ul, pl
                elif s >= stl:
                    rho[i] = rl * (ps / pl) ** (1.0 / g)
                    u[i], p[i] = us, ps
                else:
                    c = 2.0 / gp + gm / (gp * al) * (ul - s)
                    rho[i] = rl * c ** (2.0 / gm)
                    u[i] = 2.0 / gp * (al + gm / 2.0 * ul + s)
                    p[i] = pl * c ** (2.0 * g / gm)
        else:  # right of the contact
            if ps > pr:  # right shock
                sr = ur + ar * np.sqrt(gp / (2 * g) * ps / pr + gm / (2 * g))
                if s >= sr:
                    rho[i], u[i], p[i] = rr, ur, pr
                else:
                    rho[i] = rr * ((ps / pr + gm / gp) / (gm / gp * ps / pr + 1.0))
                    u[i], p[i] = us, ps
            else:  # right rarefaction
                shr = ur + ar
                asr = ar * (ps / pr) ** (gm / (2 * g))
                str_ = us + asr
                if s >= shr:
                    rho[i], u[i], p[i] = rr, ur, pr
                elif s <= str_:
                    rho[i] = rr * (ps / pr) ** (1.0 / g)
                    u[i], p[i] = us, ps
                else:
                    c = 2.0 / gp - gm / (gp * ar) * (ur - s)
                    rho[i] = rr * c ** (2.0 / gm)
                    u[i] = 2.0 / gp * (-ar + gm / 2.0 * ur + s)
                    p[i] = pr * c ** (2.0 * g / gm)
    return rho, u, p


def sod(x, t, x0=0.5, g=1.4, left=(1.0, 0.0, 1.0), right=(0.125, 0.0, 0.1)):
    """Sod's problem: (rho, u, p) at positions x and time t > 0."""
    return sample((np.asarray(x) - x0) / t, *left, *right, g)
