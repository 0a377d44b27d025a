"""Pins of the oracle's Rusanov stepper for the compressible Euler system
(CPU only): the textbook exact Riemann solution, conservation, exact
preservation of uniform and one-dimensional states, mirror symmetry and the
FP32-vs-FP64 error that calibrates the GPU tolerance."""
import numpy as np
import pytest

from cfdarcov2_amd import workloads
from tests import riemann_exact


def test_exact_solver_star_state():
    # Toro, 3rd ed., Table 4.3, test 1 (Sod): p* = 0.30313, u* = 0.92745
    p, u = riemann_exact.star_state(1.0, 0.0, 1.0, 0.125, 0.0, 0.1, 1.4)
    assert p == pytest.approx(0.30313, abs=5e-6)
    assert u == pytest.approx(0.92745, abs=5e-6)
    rho, _, _ = riemann_exact.sod(np.array([0.5 + 0.2 * 0.5, 0.5 + 0.2 * 1.5]), 0.2)
    assert rho[0] == pytest.approx(0.42632, abs=5e-5)   # left star density
    assert rho[1] == pytest.approx(0.26557, abs=5e-5)   # right star density


def _mesh(oracle_lib, n, d, off=None, dtype=np.float64):
    return oracle_lib.OracleMesh(*n, *d, dtype=dtype, perturb=off)


def _sod_1d(oracle_lib, nx, steps):
    om = _mesh(oracle_lib, (nx, 1, 1), (1.0 / nx,) * 3)
    U0 = workloads.sod_state(om.table("ccx"))
    U, hist = om.euler_run(U0, 1.4, 0.3, steps)
    return om, U, float(np.sum(hist))


def test_sod_converges_to_exact_solution(oracle_lib):
    errs = []
    for nx, steps in ((100, 60), (200, 120), (400, 240)):
        om, U, t = _sod_1d(oracle_lib, nx, steps)
        rho_ex, _, _ = riemann_exact.sod(om.table("ccx"), t)
        errs.append(np.mean(np.abs(U[0] - rho_ex)))
    # first-order scheme: the smeared contact dominates the L1 error, which
    # falls about like h^0.4 here (0.0070, 0.0051, 0.0040)
    assert errs[0] < 0.01 and errs[2] < 0.005, errs
    assert errs[1] < errs[0] / 1.2 and errs[2] < errs[1] / 1.2, errs


def test_rusanov_flux_signs_on_a_stationary_contact(oracle_lib):
    """A pure density jump at rest with equal pressure: the velocity stays zero
    to round-off and the jump only diffuses (Rusanov dissipation), with the mass
    between the walls exactly conserved to round-off."""
    nx = 64
    om = _mesh(oracle_lib, (nx, 1, 1), (1.0 / nx,) * 3)
    x = om.table("ccx")
    U0 = np.zeros((5, nx))
    U0[0] = np.where(x < 0.5, 1.0, 0.5)
    U0[4] = 1.0 / 0.4
    U, _ = om.euler_run(U0, 1.4, 0.3, 20)
    assert np.max(np.abs(U[1])) < 1e-12
    assert np.all(np.diff(U[0]) <= 1e-15)          # still monotone
    assert U[0][0] <= 1.0 and U[0][-1] >= 0.5
    assert abs(np.sum(U[0]) - np.sum(U0[0])) < 1e-12 * np.sum(U0[0])


def test_uniform_state_preserved_on_uniform_mesh(oracle_lib):
    n, d = (6, 5, 4), (0.25, 0.25, 0.25)
    om = _mesh(oracle_lib, n, d)
    U0 = np.zeros((5, om.N))
    U0[0], U0[1], U0[2], U0[3] = 1.3, 0.2 * 1.3, -0.1 * 1.3, 0.05 * 1.3
    U0[4] = 2.0
    U, _ = om.euler_run(U0, 1.4, 0.3, 5)
    # wall faces reflect the normal momentum, so only the interior stays put
    # exactly for a moving state; a state at rest is preserved everywhere
    U0r = U0.copy()
    U0r[1:4] = 0.0
    Ur, _ = om.euler_run(U0r, 1.4, 0.3, 5)
    assert np.array_equal(Ur, U0r)
    assert np.all(np.isfinite(U))


def test_one_dimensional_data_stays_one_dimensional(oracle_lib):
    """Sod on a 3D uniform mesh: every (y, z) column is identical bit for bit
    and the transverse momenta stay exactly zero."""
    n = (32, 3, 4)
    om = _mesh(oracle_lib, n, (1 / 32,) * 3)
    U0 = workloads.sod_state(om.table("ccx"))
    U, _ = om.euler_run(U0, 1.4, 0.3, 25)
    assert np.all(U[2] == 0) and np.all(U[3] == 0)
    cols = U.reshape(5, n[0], n[1] * n[2])
    assert np.all(cols == cols[:, :, :1])
    _, U1, _ = _sod_1d(oracle_lib, 32, 25)
    assert np.array_equal(cols[:, :, 0], U1)


def test_mirror_symmetry(oracle_lib):
    nx, steps = 80, 40
    om = _mesh(oracle_lib, (nx, 1, 1), (1.0 / nx,) * 3)
    x = om.table("ccx")
    U0 = workloads.sod_state(x)
    U0m = U0[:, ::-1].copy()
    U, _ = om.euler_run(U0, 1.4, 0.3, steps)
    Um, _ = om.euler_run(U0m, 1.4, 0.3, steps)
    assert np.allclose(Um[0][::-1], U[0], rtol=1e-12, atol=1e-14)
    assert np.allclose(Um[1][::-1], -U[1], rtol=1e-12, atol=1e-14)
    assert np.allclose(Um[4][::-1], U[4], rtol=1e-12, atol=1e-14)


def test_conservation_on_perturbed_mesh(oracle_lib):
    n = (24, 6, 5)
    d = (1 / 24,) * 3
    off = workloads.perturbed_vertex_offsets(*n, *d, seed=3, frac=0.2)
    om = _mesh(oracle_lib, n, d, off)
    vol = om.table("svol")
    assert np.all(vol > 0) and abs(np.sum(vol) - 0.25 * 0.2083333333333333) < 1e-12
    U0 = workloads.sod_state(om.table("ccx"))
    U, hist = om.euler_run(U0, 1.4, 0.3, 30)
    for k in (0, 4):  # mass and energy: reflective walls let none out
        before, after = np.sum(vol * U0[k]), np.sum(vol * U[k])
        assert abs(after - before) <= 1e-13 * abs(before), k
    assert np.all(hist > 0) and np.all(np.isfinite(U))
    # the first step's dt from the closed form on the initial state
    a = np.sqrt(1.4 * np.where(U0[0] > 0.5, 1.0, 0.1) / U0[0])
    area = om.table("sva")[om.table("cell_faces").reshape(-1, 6)]
    s = np.sum(area * a[:, None], axis=1)
    assert hist[0] == pytest.approx(0.3 * np.min(vol / s), rel=1e-14)


def test_vector_area_geometry(oracle_lib):
    """Closed forms: on a uniform mesh the stepper geometry is the axis normal,
    dy*dz-type areas and dx*dy*dz volumes; on a perturbed mesh the six outward
    vector areas of every cell sum to zero to round-off and the volumes tile
    the box exactly (boundary vertices stay on the box)."""
    n, d = (5, 4, 3), (0.5, 0.25, 0.125)
    om = _mesh(oracle_lib, n, d)
    nzf, nyf = 5 * 4 * 4, 5 * 5 * 3
    sva = om.table("sva")
    assert np.allclose(sva[:nzf], 0.5 * 0.25) and np.allclose(sva[nzf:nzf + nyf], 0.5 * 0.125)
    assert np.allclose(sva[nzf + nyf:], 0.25 * 0.125)
    assert np.all(om.table("svz")[:nzf] == 1) and np.all(om.table("svx")[nzf + nyf:] == 1)
    assert np.allclose(om.table("svol"), 0.5 * 0.25 * 0.125, rtol=1e-15)
    n = (6, 5, 4)
    d = (0.2, 0.2, 0.2)
    off = workloads.perturbed_vertex_offsets(*n, *d, seed=9, frac=0.2)
    om = _mesh(oracle_lib, n, d, off)
    cf = om.table("cell_faces").reshape(-1, 6)
    S = np.stack([om.table(k) * om.table("sva") for k in ("svx", "svy", "svz")], axis=1)
    sgn = np.array([-1, 1, -1, 1, -1, 1])[None, :, None]
    closed = np.sum(S[cf] * sgn, axis=1)
    assert np.max(np.abs(closed)) < 1e-15
    assert np.sum(om.table("svol")) == pytest.approx(1.2 * 1.0 * 0.8, rel=1e-14)
    assert np.all(om.table("svol") > 0)
    # by hand: 2x2x2 unit cells, the one interior vertex (1,1,1) moved by
    # delta; the x-face at x=1, y=0, z=0 has diagonals (dx, 1+dy, 1+dz) and
    # (0, -1, 1), so S = (1 + (dy+dz)/2, -dx/2, -dx/2)
    off = np.zeros((27, 3))
    off[13] = (0.1, 0.05, -0.02)
    om = _mesh(oracle_lib, (2, 2, 2), (1.0, 1.0, 1.0), off)
    f = 12 + 12 + 1
    S = np.array([1.015, -0.05, -0.05])
    A = np.sqrt(S @ S)
    assert om.table("sva")[f] == pytest.approx(A, rel=1e-15)
    got = np.array([om.table(k)[f] for k in ("svx", "svy", "svz")])
    assert np.allclose(got, S / A, rtol=1e-15, atol=1e-16)
    assert np.sum(om.table("svol")) == pytest.approx(8.0, rel=1e-15)


def test_uniform_state_at_rest_on_perturbed_mesh(oracle_lib):
    n, d = (8, 6, 5), (0.1, 0.1, 0.1)
    off = workloads.perturbed_vertex_offsets(*n, *d, seed=4, frac=0.2)
    om = _mesh(oracle_lib, n, d, off)
    U0 = np.zeros((5, om.N))
    U0[0], U0[4] = 0.7, 1.9
    U, _ = om.euler_run(U0, 1.4, 0.3, 10)
    assert np.max(np.abs(U[1:4])) < 1e-14
    assert np.max(np.abs(U[0] - 0.7)) < 1e-14


def test_fp32_error_against_fp64(oracle_lib):
    """The FP32 oracle's deviation from the FP64 oracle on a perturbed Sod
    mesh: the calibration behind the GPU FP32 tolerance."""
    n = (48, 6, 6)
    d = (1 / 48,) * 3
    off = workloads.perturbed_vertex_offsets(*n, *d, seed=1, frac=0.2)
    o64 = _mesh(oracle_lib, n, d, off, np.float64)
    o32 = _mesh(oracle_lib, n, d, off, np.float32)
    U64, _ = o64.euler_run(workloads.sod_state(o64.table("ccx")), 1.4, 0.3, 40)
    U32, _ = o32.euler_run(workloads.sod_state(o32.table("ccx"), dtype=np.float32), 1.4, 0.3, 40)
    for k in (0, 1, 4):
        rel_l1 = np.sum(np.abs(U32[k] - U64[k])) / np.sum(np.abs(U64[k]))
        assert rel_l1 < 1e-5, (k, rel_l1)
