"""Synthetic inputs for the benchmark workloads.

All randomness comes from one project-defined generator, SplitMix64
(Steele, Lea & Flood 2014), with doubles drawn as (x >> 11) * 2**-53, so the
Python and C++ sides produce identical streams.  Initial states are evaluated
in float64 at cell centres ((i + 1/2) * d on the unperturbed lattice) and cast
to the run dtype once.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

_MASK = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed: int):
        self.state = seed & _MASK

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & _MASK
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _MASK
        return z ^ (z >> 31)

    def uniform(self) -> float:
        return (self.next_u64() >> 11) * (1.0 / (1 << 53))

    def uniform_array(self, n: int) -> np.ndarray:
        """n draws, vectorised (same stream as n calls of uniform())."""
        with np.errstate(over="ignore"):
            k = np.arange(1, n + 1, dtype=np.uint64)
            z = np.uint64(self.state) + k * np.uint64(0x9E3779B97F4A7C15)
            self.state = int(z[-1]) if n else self.state
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            z = z ^ (z >> np.uint64(31))
        return (z >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))


def cell_centres(nx, ny, nz, dx, dy, dz):
    """Cell-centre coordinates in the reference cell order (z fastest), float64."""
    x = (np.arange(nx) + 0.5) * dx
    y = (np.arange(ny) + 0.5) * dy
    z = (np.arange(nz) + 0.5) * dz
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
    return X.ravel(), Y.ravel(), Z.ravel()


def gaussian(nx, ny, nz, dx, dy, dz, centre, sigma, amplitude=1.0):
    X, Y, Z = cell_centres(nx, ny, nz, dx, dy, dz)
    r2 = (X - centre[0]) ** 2 + (Y - centre[1]) ** 2 + (Z - centre[2]) ** 2
    return amplitude * np.exp(-r2 / (2.0 * sigma * sigma))


def perturbed_vertex_offsets(nx, ny, nz, dx, dy, dz, seed, frac):
    """Seeded offsets, uniform in [-frac*d, frac*d] per component, for interior
    vertices only (boundary vertices stay on the box), in the reference vertex
    order (z fastest).  Returns (V, 3) float64."""
    V = (nx + 1) * (ny + 1) * (nz + 1)
    g = SplitMix64(seed)
    r = g.uniform_array(3 * V).reshape(V, 3)
    off = (2.0 * r - 1.0) * frac * np.array([dx, dy, dz])
    ix, iy, iz = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1), indexing="ij")
    interior = ((ix > 0) & (ix < nx) & (iy > 0) & (iy < ny) & (iz > 0) & (iz < nz)).ravel()
    off[~interior] = 0.0
    return off


@dataclass
class HeatCase:
    n: tuple
    spacing: tuple
    alpha: float
    dt: float
    steps: int
    T0: np.ndarray


def heat_config(n=64, steps=200, dt_factor=0.2, alpha=1e-3, sigma=0.1):
    """Plain 3D heat on n^3 over the unit cube, Gaussian amplitude 1, sigma 0.1,
    centred at (0.5, 0.5, 0.5); dt = dt_factor * dx^2 / alpha."""
    d = 1.0 / n
    T0 = gaussian(n, n, n, d, d, d, (0.5, 0.5, 0.5), sigma)
    return HeatCase((n, n, n), (d, d, d), alpha, dt_factor * d * d / alpha, steps, T0)


@dataclass
class WaveCase:
    n: tuple
    spacing: tuple
    c: float
    dt: float
    steps: int
    p0: np.ndarray
    centre: tuple


def wave_config(n=256, steps=1000, cfl=0.4, c=1.0, sigma=0.05, seed=0):
    """First-order acoustic system on n^3 over the unit cube: Gaussian p pulse
    (amplitude 1, sigma 0.05) at a seeded centre in [0.25, 0.75]^3, u=v=w=0,
    dt = cfl * dx / c."""
    d = 1.0 / n
    g = SplitMix64(seed)
    centre = tuple(0.25 + 0.5 * g.uniform() for _ in range(3))
    p0 = gaussian(n, n, n, d, d, d, centre, sigma)
    return WaveCase((n, n, n), (d, d, d), c, cfl * d / c, steps, p0, centre)


def wave_energy(p, u, v, w, c):
    """Discrete acoustic energy 0.5 * sum(u^2 + v^2 + w^2 + p^2 / c^2)."""
    return 0.5 * float(np.sum(u * u + v * v + w * w + (p * p) / (c * c)))


def heat_free_space(nx, ny, nz, dx, dy, dz, centre, sigma, alpha, t):
    """Free-space heat kernel: a Gaussian of variance sigma^2 spreads to
    sigma^2 + 2 alpha t with amplitude (sigma^2 / (sigma^2 + 2 alpha t))^(3/2)."""
    s2 = sigma * sigma + 2.0 * alpha * t
    amp = (sigma * sigma / s2) ** 1.5
    return gaussian(nx, ny, nz, dx, dy, dz, centre, math.sqrt(s2), amp)


def sod_state(ccx, gamma=1.4, x0=0.5, left=(1.0, 1.0), right=(0.125, 0.1), dtype=np.float64):
    """Conserved Sod state (rho, rho u, rho v, rho w, E) at cell-centre x:
    (rho, p) = left for x < x0, right otherwise, at rest; returns (5, N)."""
    ccx = np.asarray(ccx, dtype=np.float64)
    lt = ccx < x0
    rho = np.where(lt, left[0], right[0])
    p = np.where(lt, left[1], right[1])
    U = np.zeros((5, ccx.size), dtype=np.float64)
    U[0] = rho
    U[4] = p / (gamma - 1.0)
    return U.astype(dtype)


@dataclass
class SodCase:
    n: tuple
    spacing: tuple
    gamma: float
    cfl: float
    steps: int
    vertex_offsets: np.ndarray  # (V, 3) float64, None for the uniform mesh
    seed: int
    frac: float


def sod_config(nx=512, ny=64, nz=64, steps=500, cfl=0.3, gamma=1.4, seed=0, frac=0.2):
    """Sod shock tube on [0, 1] x [0, ny/nx] x [0, nz/nx] with dx = 1/nx;
    interior vertices moved by seeded offsets up to frac * dx per component
    (frac = 0 gives the uniform mesh).  The initial state needs the mesh's cell
    centres: sod_state(mesh.table('ccx'))."""
    d = 1.0 / nx
    off = perturbed_vertex_offsets(nx, ny, nz, d, d, d, seed, frac) if frac > 0 else None
    return SodCase((nx, ny, nz), (d, d, d), gamma, cfl, steps, off, seed, frac)


@dataclass
class AdjointCase:
    n: tuple
    spacing: tuple
    dt: float
    steps: int
    every: int
    sensors: np.ndarray      # int64 cell ids
    p0_true: np.ndarray      # hidden initial pressure that makes the observations
    p0_guess: np.ndarray     # initial pressure the gradient is taken at


def adjoint_config(n=128, steps=2000, every=45, n_sensors=32, cfl=0.4, c=1.0, seed=0):
    """Adjoint acoustic workload on n^3 over the unit cube: a hidden
    4-component Gaussian mixture (sigma ~ U[0.03, 0.08], amplitude ~ U[0.5, 1],
    centre ~ U[0.2, 0.8]^3) makes the observed traces at `n_sensors` distinct
    seeded cells; the gradient is taken at a single centred Gaussian (sigma 0.1,
    amplitude 0.5).  dt = cfl * dx / c."""
    d = 1.0 / n
    g = SplitMix64(seed)
    p_true = np.zeros(n * n * n)
    for _ in range(4):
        sigma = 0.03 + 0.05 * g.uniform()
        amp = 0.5 + 0.5 * g.uniform()
        centre = tuple(0.2 + 0.6 * g.uniform() for _ in range(3))
        p_true += gaussian(n, n, n, d, d, d, centre, sigma, amp)
    cells = []
    while len(cells) < n_sensors:
        cid = int(g.next_u64() % (n * n * n))
        if cid not in cells:
            cells.append(cid)
    guess = gaussian(n, n, n, d, d, d, (0.5, 0.5, 0.5), 0.1, 0.5)
    return AdjointCase((n, n, n), (d, d, d), cfl * d / c, steps, every, np.array(cells, dtype=np.int64), p_true,
                       guess)


@dataclass
class EnsembleCase:
    kind: str                # "heat" or "wave"
    n: int                   # cube side in cells
    spacing: float           # cell size (dyadic, so the mesh is exactly uniform)
    coef: float              # heat: alpha; wave: K = -c^2
    dt: float
    field0: np.ndarray       # heat: T0; wave: p0 (u = v = w = 0)


ENSEMBLE_SIDES = (16, 24, 32, 48, 64, 96, 128)


def ensemble_config(n_cases=64, steps=500, seed=0, spacing=1.0 / 128, sides=ENSEMBLE_SIDES,
                    heat_dt_factor=0.15, wave_cfl=0.4):
    """Batched ensemble: case i is heat for even i and acoustic for odd i; the
    cube side is drawn uniformly from `sides`; alpha ~ log-uniform [1e-4, 1e-2],
    c ~ U[0.5, 2]; a Gaussian pulse (sigma 0.1 of the side for heat, 0.05 for
    waves, amplitude 1) at a seeded centre in [0.25, 0.75]^3 of the box.  Every
    case uses the same dyadic cell size so the meshes are exactly uniform;
    dt = 0.15 dx^2 / alpha (heat, stable) or 0.4 dx / c (wave)."""
    g = SplitMix64(seed)
    cases = []
    for i in range(n_cases):
        side = sides[int(g.uniform() * len(sides)) % len(sides)]
        box = side * spacing
        centre = tuple(box * (0.25 + 0.5 * g.uniform()) for _ in range(3))
        if i % 2 == 0:
            alpha = 10.0 ** (-4.0 + 2.0 * g.uniform())
            f0 = gaussian(side, side, side, spacing, spacing, spacing, centre, 0.1 * box)
            cases.append(EnsembleCase("heat", side, spacing, alpha, heat_dt_factor * spacing ** 2 / alpha, f0))
        else:
            c = 0.5 + 1.5 * g.uniform()
            f0 = gaussian(side, side, side, spacing, spacing, spacing, centre, 0.05 * box)
            cases.append(EnsembleCase("wave", side, spacing, -c * c, wave_cfl * spacing / c, f0))
    return cases, steps
