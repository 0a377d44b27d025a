"""Parity of the C++ operator API (Variable / Equation::evaluate / KernelBuilder)
against the CPU oracle.

tests/cpp/fvm3d_parity.cpp runs equation systems through the API on the
device - fused hipRTC kernels (flag true) and operator-by-operator evaluation
(flag false) - and writes the final fields; here the same systems are composed
from the oracle's operators and float32 numpy arithmetic in the reference
evaluation order, and the fields must agree exactly.  The reference's own
example drivers, compiled unchanged against the API, are run as well.
"""
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

from cfdarcov2_amd import workloads

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BUILD = os.path.join(ROOT, "cfdarcov2_amd", "_build")
PROG = os.path.join(BUILD, "tests", "fvm3d_parity")
DRIVERS = os.path.join(BUILD, "drivers")
f32 = np.float32


def _run_case(case, n, d, dt, steps, fields, nout):
    """Runs one system of tests/cpp/fvm3d_parity.cpp.  A case ending in
    "_perop" is the same system with CFDARCO_PER_OP=1: equations flagged
    `false` are then evaluated operator by operator instead of fused."""
    if not os.path.exists(PROG):
        pytest.fail(f"{PROG} not built (run __graft_entry__.build())")
    env = dict(os.environ)
    env.pop("CFDARCO_PER_OP", None)
    if case.endswith("_perop"):
        case = case[: -len("_perop")]
        env["CFDARCO_PER_OP"] = "1"
    N = n[0] * n[1] * n[2]
    with tempfile.TemporaryDirectory() as tmp:
        inp, out = os.path.join(tmp, "in.bin"), os.path.join(tmp, "out.bin")
        with open(inp, "wb") as f:
            np.array([*n, steps], dtype=np.int64).tofile(f)
            np.array([*d, dt], dtype=np.float64).tofile(f)
            np.array([len(fields)], dtype=np.int64).tofile(f)
            for a in fields:
                np.ascontiguousarray(a, dtype=f32).tofile(f)
        print(f"[case] fvm3d {case} n={n} steps={steps}", flush=True)
        res = subprocess.run([PROG, case, inp, out], capture_output=True, text=True, timeout=300, env=env)
        assert res.returncode == 0, res.stdout + res.stderr
        got = np.fromfile(out, dtype=f32)
    assert got.size == nout * N
    return got.reshape(nout, N)


def _mesh(oracle_lib, n, d, perturbed=False):
    off = None
    if perturbed:
        off = workloads.perturbed_vertex_offsets(*n, *d, seed=5, frac=0.02)
    return oracle_lib.OracleMesh(*n, *d, dtype=f32, perturb=off)


@pytest.mark.parametrize("case", ["heat", "heat_nojit", "heat_nojit_perop"])
def test_heat_driver_structure(oracle_lib, case):
    n, d = (12, 10, 16), (1 / 16, 1 / 16, 1 / 16)
    om = _mesh(oracle_lib, n, d)
    X, Y, Z = workloads.cell_centres(*n, *d)
    T0 = np.where(Z < 2 * d[2], 100.0, 0.0).astype(f32)
    dt, steps = f32(0.02 * d[0] ** 2), 15
    got = _run_case(case, n, d, float(dt), steps, [T0], 1)
    bnd = om.table("is_bnd").astype(f32)
    rev = (f32(1) - bnd).astype(f32)
    T = T0.copy()
    for _ in range(steps):
        T = dt * (f32(3.0) * om.lapl(T)) + T
        T = T * rev + T0 * bnd
    assert np.array_equal(got[0], T)


@pytest.mark.parametrize("case", ["wave", "wave_nojit", "wave_perturbed", "wave_perturbed_nojit",
                                  "wave_perturbed_nojit_perop"])
def test_acoustic_system_through_api(oracle_lib, case):
    n, d = (10, 12, 14), (1 / 16, 1 / 16, 1 / 16)
    perturbed = "perturbed" in case
    om = _mesh(oracle_lib, n, d, perturbed)
    X, Y, Z = workloads.cell_centres(*n, *d)
    rng = np.random.default_rng(4)
    p0 = np.exp(-((X - 0.3) ** 2 + (Y - 0.4) ** 2 + (Z - 0.5) ** 2) / 0.02).astype(f32)
    u0, v0, w0 = (0.1 * rng.standard_normal(p0.size)).astype(f32), np.zeros_like(p0), np.zeros_like(p0)
    dt, steps = f32(0.2 / 16), 12
    got = _run_case(case, n, d, float(dt), steps, [p0, u0, v0, w0], 4)
    p, u, v, w = p0.copy(), u0.copy(), v0.copy(), w0.copy()
    c = f32(1.25)
    K = (-c) * c
    for _ in range(steps):
        u = dt * (-om.grad(p, 1)) + u
        v = dt * (-om.grad(p, 2)) + v
        w = dt * (-om.grad(p, 4)) + w
        p = dt * (K * ((om.grad(u, 1) + om.grad(v, 2)) + om.grad(w, 4))) + p
    for name, g, r in zip("puvw", got, (p, u, v, w)):
        assert np.all(np.isfinite(r)), name
        assert np.array_equal(g, r), name


@pytest.mark.parametrize("case", ["expabs", "expabs_nojit"])
def test_exp_abs_through_api(oracle_lib, case):
    """exp and abs on the device.  abs is exact, so x and z, which do not
    depend on exp, match bit for bit; y = exp(-|x| / 2) * y takes the
    device exp, within 1 ulp of the exact value, against exp in float64
    rounded to float32: each of the 4 steps adds at most 1 ulp (exp) and 1/2
    ulp (rounded product) of relative error on each side, so |got - ref| <=
    16 ulp of the result bounds it."""
    n, d = (10, 12, 8), (1 / 16, 1 / 16, 1 / 16)
    om = _mesh(oracle_lib, n, d)
    rng = np.random.default_rng(21)
    N = n[0] * n[1] * n[2]
    x0 = (3.0 * rng.standard_normal(N)).astype(f32)
    y0, z0 = rng.standard_normal(N).astype(f32), rng.standard_normal(N).astype(f32)
    dt, steps = f32(0.01), 4
    got = _run_case(case, n, d, float(dt), steps, [x0, y0, z0], 3)
    x, y, z = x0.copy(), y0.copy(), z0.copy()
    for _ in range(steps):
        e = np.exp(((-np.abs(x)) * f32(0.5)).astype(np.float64)).astype(f32)
        y = e * y
        z = np.abs(om.grad(x, 1)) - np.abs(x) * z
        x = dt * (-np.abs(om.grad(x, 2))) + x
    assert np.array_equal(got[0], x)
    assert np.array_equal(got[2], z)
    ulps = np.abs(got[1] - y) / np.spacing(np.abs(y))
    assert np.all(ulps <= 16), float(ulps.max())


@pytest.mark.parametrize("case,rule", [("d2t", 0), ("d2t_nojit", 2), ("d2t_nojit_store", 1)])
def test_second_order_time(oracle_lib, case, rule):
    """d2t: fused equations run two Euler-form steps then leapfrog; operator-by-
    operator equations leapfrog only with stored history (--store, here the
    "store" case), else stay in the Euler form (the reference's behaviours,
    restated by the oracle's D2tVar rules)."""
    n, d = (16, 14, 6), (1 / 16, 1 / 16, 1 / 16)
    om = _mesh(oracle_lib, n, d)
    X, Y, Z = workloads.cell_centres(*n, *d)
    u0 = (np.sin(np.pi * X) * np.sin(np.pi * Y)).astype(f32)
    dt, steps = f32(0.05), 20
    got = _run_case(case, n, d, float(dt), steps, [u0], 1)
    c = f32(0.3)
    u = om.d2t_run(u0, float(c * c), 3, float(dt), steps, rule)
    assert np.array_equal(got[0], u)


@pytest.mark.parametrize("case", ["euler1d", "euler1d_nojit", "euler1d_perturbed", "euler1d_nojit_perop"])
def test_euler_like_system(oracle_lib, case):
    """Assignment equations, DT inside expressions, products inside stencil
    operands, stabilisation, a stencil of a stencil and the device CFL policy."""
    n, d = (24, 4, 4), (1 / 32, 1 / 32, 1 / 32)
    om = _mesh(oracle_lib, n, d, "perturbed" in case)
    X, Y, Z = workloads.cell_centres(*n, *d)
    rho0 = np.where(X > 0.4, 1.0, 0.125).astype(f32)
    u0 = np.zeros_like(rho0)
    p0 = np.where(X > 0.4, 1.0, 0.1).astype(f32)
    vol = om.table("vol")
    gamma = f32(1.4)
    g1 = gamma - f32(1)
    mass0 = rho0 * vol
    rho_u0 = rho0 * u0 * vol
    E0 = (p0 / g1 + f32(0.5) * rho0 * (u0 * u0)) * vol
    cfl, steps = 0.2, 6
    got = _run_case(case, n, d, cfl, steps, [rho0, u0, p0, mass0, rho_u0, E0], 7)

    rho, u, p, mass, rho_u, rho_e = rho0.copy(), u0.copy(), p0.copy(), mass0.copy(), rho_u0.copy(), E0.copy()
    mass_tmp, rho_u_tmp, rho_e_tmp = mass0.copy(), rho_u0.copy(), E0.copy()
    G = lambda a, m=1: om.grad(a, m)
    half, zero, two = f32(0.5), f32(0.0), f32(2)
    for _ in range(steps):
        dt = f32(om.cfl1d(cfl, u, p, rho))
        E = lambda: (p / g1) + ((half * rho) * (u * u))
        rho = mass / vol
        u = (rho_u / rho) / vol
        p = ((rho_e / vol) - ((half * rho) * (u * u))) * g1
        rho = rho - (half * dt) * ((u * G(rho)) + (rho * G(u)))
        u = u - (half * dt) * ((u * G(u)) + ((f32(1) / rho) * G(p)))
        p = p - (half * dt) * (((gamma * p) * G(u)) + (u * G(p)))
        mass_tmp = (zero * mass_tmp) - G(rho * u)
        rho_u_tmp = (zero * rho_u_tmp) - G(((rho * u) * u) + p)
        rho_e_tmp = (zero * rho_e_tmp) - G((E() + p) * u)
        mass = dt * (mass_tmp - (om.stab(rho, 1) * two)) + mass
        rho_u = dt * (rho_u_tmp - (om.stab(rho * u, 1) * two)) + rho_u
        rho_e = dt * (rho_e_tmp - (om.stab(E(), 7) * two)) + rho_e
        mass_tmp = om.lapl(G(rho), 7) + om.lapl(mass, 2)
    for name, g, r in zip(["rho", "u", "p", "mass", "rho_u", "rho_e", "mass_tmp"], got,
                          (rho, u, p, mass, rho_u, rho_e, mass_tmp)):
        assert np.array_equal(g, r, equal_nan=True), name


def _sum(a):
    """The sum the driver prints, to within its rounding: the operator API's
    host vectors sum pairwise, so an FP64 sum of the FP32 values is within a
    few ulp of it."""
    return np.sum(a, dtype=np.float64)


def _run_driver(name, args, per_op=False, stab_in_kernel=False, cwd=None):
    path = os.path.join(DRIVERS, name)
    if not os.path.exists(path):
        pytest.skip(f"{name}: reference driver binaries are built only where the reference sources exist")
    print(f"[case] driver {name} {' '.join(args)}{' (per operator)' if per_op else ''}"
          f"{' (stabilisation in kernel)' if stab_in_kernel else ''}", flush=True)
    env = dict(os.environ)
    env.pop("CFDARCO_PER_OP", None)
    env.pop("CFDARCO_STAB_IN_KERNEL", None)
    if per_op:
        env["CFDARCO_PER_OP"] = "1"
    if stab_in_kernel:
        env["CFDARCO_STAB_IN_KERNEL"] = "1"
    if cwd is not None:
        res = subprocess.run([path, *args], capture_output=True, text=True, timeout=300, cwd=cwd, env=env)
    else:
        with tempfile.TemporaryDirectory() as tmp:
            res = subprocess.run([path, *args], capture_output=True, text=True, timeout=300, cwd=tmp, env=env)
    assert res.returncode == 0, res.stdout + res.stderr
    return res.stdout


def test_reference_heat_driver_unchanged(oracle_lib):
    """examples/main_3d_heat.cpp compiled unchanged: its printed sum and max
    equal the oracle's run of the same equations."""
    n, d, steps, dt = (16, 16, 16), (1.0, 1.0, 1.0), 25, 0.01
    out = _run_driver("cfdARCO_heat", ["-Lx", "16", "-Ly", "16", "-Lz", "16", "-dx", "1", "-dy", "1", "-dz", "1",
                                       "-t", str(steps), "-dt", str(dt), "-c"])
    m = re.search(r"At T Current sum = (\S+) Current max = (\S+)", out)
    assert m, out
    om = _mesh(oracle_lib, n, d)
    X, Y, Z = workloads.cell_centres(*n, *d)
    T0 = np.where(Z.astype(f32) < f32(2) * f32(1.0), f32(100), f32(0)).astype(f32)
    bnd = om.table("is_bnd").astype(f32)
    rev = f32(1) - bnd
    T = T0.copy()
    dtf = f32(dt)
    for _ in range(steps):
        T = dtf * (f32(3.0) * om.lapl(T)) + T
        T = T * rev + T0 * bnd
    # the driver prints with the stream's default six significant digits
    assert float(m.group(1)) == pytest.approx(float(_sum(T)), rel=1e-5)
    assert float(m.group(2)) == pytest.approx(float(T.max()), rel=1e-5)
    assert "Time difference" in out


@pytest.mark.parametrize("stepping", [False, True])
def test_reference_heat_driver_store(oracle_lib, stepping):
    """examples/main_3d_heat.cpp with -s (history as <var>/res_<i>.vti at the
    end, every entry but the last) and with -s -st (mesh.json first, then
    <var>/<i>.bin after every step, written from asynchronous snapshots):
    every file read back equals the oracle's state after that many steps, bit
    for bit, and dumps/run_latest points at the run."""
    import base64
    import json
    n, d, steps, dt = (12, 10, 16), (1.0, 1.0, 1.0), 9, 0.01
    X, Y, Z = n
    om = _mesh(oracle_lib, n, d)
    Xc, Yc, Zc = workloads.cell_centres(*n, *d)
    T0 = np.where(Zc.astype(f32) < f32(2) * f32(1.0), f32(100), f32(0)).astype(f32)
    bnd = om.table("is_bnd").astype(f32)
    rev = f32(1) - bnd
    states, T, dtf = [], T0.copy(), f32(dt)
    for _ in range(steps):
        T = dtf * (f32(3.0) * om.lapl(T)) + T
        T = T * rev + T0 * bnd
        states.append(T.copy())
    args = ["-Lx", str(X), "-Ly", str(Y), "-Lz", str(Z), "-t", str(steps), "-dt", str(dt), "-c", "-s"]
    with tempfile.TemporaryDirectory() as tmp:
        _run_driver("cfdARCO_heat", args + (["-st"] if stepping else []), cwd=tmp)
        latest = os.path.join(tmp, "dumps", "run_latest")
        assert os.path.islink(latest)
        run = os.path.realpath(latest)
        assert os.path.basename(run).startswith("run_")
        tdir = os.path.join(run, "T")
        if stepping:
            with open(os.path.join(run, "mesh.json")) as f:
                mj = json.load(f)
            assert (mj["x"], mj["y"], mj["z"]) == n and len(mj["nodes"]) == X * Y * Z
            assert sorted(os.listdir(tdir)) == sorted(f"{i}.bin" for i in range(steps))
            for i in range(steps):
                got = np.fromfile(os.path.join(tdir, f"{i}.bin"), dtype=f32)
                assert np.array_equal(got, states[i]), i
        else:
            assert sorted(os.listdir(tdir)) == sorted(f"res_{i}.vti" for i in range(steps - 1))
            for i in range(steps - 1):
                with open(os.path.join(tdir, f"res_{i}.vti")) as f:
                    blob = re.search(r'format="binary">\s*(\S+)\s*</DataArray>', f.read()).group(1)
                img = np.frombuffer(base64.b64decode(blob[12:]), dtype=f32).reshape(Z, Y, X)
                assert np.array_equal(img.transpose(2, 1, 0).ravel(), states[i]), i


def test_reference_heat_driver_mesh_input():
    """-m/--mesh: the heat driver on a perturbed mesh writes mesh.json with
    -s -st; run again from that file (-m, extents, spacing and vertices from
    it), it writes the same fields after every step, bit for bit."""
    steps = 6
    common = ["-t", str(steps), "-dt", "0.001", "-c", "-s", "-st"]
    with tempfile.TemporaryDirectory() as tmp:
        a, b = os.path.join(tmp, "a"), os.path.join(tmp, "b")
        os.makedirs(a)
        os.makedirs(b)
        out_a = _run_driver("cfdARCO_heat", ["-Lx", "10", "-Ly", "8", "-Lz", "12", "-dx", "0.125", "-dy", "0.125",
                                             "-dz", "0.0625", "--perturb", "0.2", "--perturb_seed", "3"] + common,
                            cwd=a)
        run_a = os.path.realpath(os.path.join(a, "dumps", "run_latest"))
        out_b = _run_driver("cfdARCO_heat", ["-m", os.path.join(run_a, "mesh.json")] + common, cwd=b)
        run_b = os.path.realpath(os.path.join(b, "dumps", "run_latest"))
        assert _printed(out_a) == _printed(out_b)
        for i in range(steps):
            fa = np.fromfile(os.path.join(run_a, "T", f"{i}.bin"), dtype=f32)
            fb = np.fromfile(os.path.join(run_b, "T", f"{i}.bin"), dtype=f32)
            assert fa.size == 10 * 8 * 12 and np.array_equal(fa, fb), i


def _printed(out):
    return {v: (float(s), float(m)) for v, s, m in re.findall(r"At (\S+) Current sum = (\S+) Current max = (\S+)",
                                                             out)}


def _check_printed(printed, name, ref):
    """The driver prints its FP32 sum and maximum with six significant digits:
    allow half a unit in the sixth digit plus the FP32 sum's own rounding
    (bounded by a few ulp of the sum of magnitudes)."""
    s, mx = printed[name]
    ref = np.asarray(ref, dtype=f32)
    assert np.all(np.isfinite(ref)), name
    want_s, want_m = float(_sum(ref)), float(ref.max())
    mag = float(np.sum(np.abs(ref), dtype=np.float64))
    assert abs(s - want_s) <= 5e-6 * abs(want_s) + 1e-6 * mag + 1e-30, (name, s, want_s)
    assert abs(mx - want_m) <= 5e-6 * abs(want_m) + 1e-30, (name, mx, want_m)


def test_reference_wave_driver_unchanged(oracle_lib):
    """examples/main_3d_wave.cu compiled unchanged: d2t(u) = c*c*(d2dx(u) +
    d2dy(u)) fused (two Euler-form steps, then leapfrog), dt = 0.1 fixed by
    the driver, point sources at cells (0.1 n) and (0.9 n) set by its host
    callback to sin(float(step) * 0.2) after every step.  Printed sum and
    maximum equal the oracle's run of the same equation and callback."""
    n, steps = (20, 18, 16), 30
    out = _run_driver("cfdARCO_wave", ["-Lx", str(n[0]), "-Ly", str(n[1]), "-Lz", str(n[2]), "-t", str(steps),
                                       "-dt", "0.7", "-c"])
    om = _mesh(oracle_lib, n, (1.0, 1.0, 1.0))
    X, Y, Z = n
    cid = lambda x, y, z: Z * Y * x + Z * y + z
    src = [cid(int(X * 0.1), int(Y * 0.1), int(Z * 0.1)), cid(int(X * 0.9), int(Y * 0.9), int(Z * 0.9))]

    def sources(u, t):  # _current_time_step_int is t + 1 when the callback runs
        u = u.copy()
        u[src] = f32(np.sin(np.float64(f32(t + 1)) * 0.2))
        return u

    c = f32(0.3)
    u = om.d2t_run(np.zeros(om.N, f32), float(c * c), 3, float(f32(0.1)), steps, oracle_lib.D2T_FUSED, bc=sources)
    # the operator has no z term, so the sources spread only within their two
    # z layers (2 * 20 * 18 cells); after 30 steps they cover most of them
    zlayer = np.arange(om.N) % Z
    assert np.all(u[(zlayer != src[0] % Z) & (zlayer != src[1] % Z)] == 0)
    assert np.count_nonzero(u) > 300
    _check_printed(_printed(out), "u", u)
    assert "Time difference" in out


@pytest.mark.parametrize("store", [False, True])
def test_reference_gauss_pulse_driver_unchanged(oracle_lib, store):
    """examples/main_3d_wave_gauss_pulse.cu compiled unchanged, 2D via one z
    layer: the standing-wave initial state sampled at i*dx, the x/y boundary
    layer zeroed (boundary_zero, by cell centre against dx + 0.001 and
    lx - dx - 0.001), d2t flagged `false`.  Without --store the per-operator
    D2tVar keeps the Euler form; with it, it leapfrogs from the third step.
    Printed sum and maximum equal the oracle's run of the same rule."""
    n, d, steps, dt = (40, 36, 1), 1.0 / 32, 60, 0.002
    args = ["-Lx", str(n[0]), "-Ly", str(n[1]), "-Lz", "1", "-dx", str(d), "-dy", str(d), "-dz", str(d),
            "-t", str(steps), "-dt", str(dt), "-c"] + (["-s"] if store else [])
    out = _run_driver("cfdARCO_gauss_pulse", args)
    om = _mesh(oracle_lib, n, (d, d, d))
    X, Y = n[0], n[1]
    dxf = f32(d)
    lx, ly = f32(dxf * f32(X)), f32(dxf * f32(Y))
    ix, iy = (a.ravel() for a in np.meshgrid(np.arange(X), np.arange(Y), indexing="ij"))
    xv = f32(lx / f32(X)) * ix.astype(f32)
    yv = f32(ly / f32(Y)) * iy.astype(f32)
    u0 = (np.sin(np.pi * xv.astype(np.float64)) * np.sin(np.pi * yv.astype(np.float64))).astype(f32)
    xl, xu = f32(float(dxf) + 0.001), f32(float(f32(lx - dxf)) - 0.001)
    yl, yu = f32(float(dxf) + 0.001), f32(float(f32(ly - dxf)) - 0.001)
    cx, cy = om.table("ccx"), om.table("ccy")
    keep = (xl < cx) & (cx < xu) & (yl < cy) & (cy < yu)
    zero_bc = lambda u, t: np.where(keep, u, f32(0)).astype(f32)
    u0 = zero_bc(u0, -1)
    c = f32(0.3)
    rule = oracle_lib.D2T_HISTORY if store else oracle_lib.D2T_NO_HISTORY
    u = om.d2t_run(u0, float(c * c), 3, float(f32(dt)), steps, rule, bc=zero_bc)
    _check_printed(_printed(out), "u", u)


def test_reference_euler_driver_unchanged(oracle_lib):
    """examples/main_3d.cpp compiled unchanged (10 variables, 15 fused
    equations and 5 stabilised d1t equations flagged `false`, 3D CFL policy
    with gamma 5/3, boundary layers held at their initial values): every
    printed sum and maximum equals the oracle's run of the same twenty
    equations in the same order."""
    n, steps, cfl = (12, 10, 14), 8, 0.1
    out = _run_driver("cfdARCO_euler", ["-Lx", str(n[0]), "-Ly", str(n[1]), "-Lz", str(n[2]), "-t", str(steps),
                                        "-dt", str(cfl), "-c"])
    printed = _printed(out)
    om = _mesh(oracle_lib, n, (1.0, 1.0, 1.0))
    cx, cy, cz, vol = om.table("ccx"), om.table("ccy"), om.table("ccz"), om.table("vol")
    lx, ly, lz = (f32(f32(1.0) * f32(k)) for k in n)
    lim = lambda a, L: f32(a * float(L))
    rho0 = np.where((lim(0.4, ly) < cy) & (cy < lim(0.6, ly)) & (lim(0.4, lx) < cx) & (cx < lim(0.6, lx))
                    & (cz < lim(1.0, lz)), f32(5), f32(1)).astype(f32)
    # u and v alternate -0.1, +0.1 in ten z bands: band b (1..10) is the first
    # with z < 0.b * lz (compared in double), value (-1)^b * 0.1
    zc = cz.astype(np.float64)
    band = np.full(om.N, 10)
    for b in range(9, 0, -1):
        band = np.where(zc < b / 10 * float(lz), b, band)
    u0 = np.where(band % 2 == 1, f32(-0.1), f32(0.1)).astype(f32)
    v0 = u0.copy()
    w0 = np.full(om.N, f32(0.6), f32)
    p0 = np.full(om.N, f32(2.5), f32)
    gamma = f32(5.0 / 3.0)
    g1 = gamma - f32(1)
    half, zero, two = f32(0.5), f32(0), f32(2)
    mass0 = rho0 * vol
    ru0, rv0, rw0 = (rho0 * u0) * vol, (rho0 * v0) * vol, (rho0 * w0) * vol
    E0 = ((p0 / g1) + ((half * rho0) * (((u0 * u0) + (v0 * v0)) + (w0 * w0)))) * vol
    init = dict(rho=rho0, u=u0, v=v0, w=w0, p=p0, mass=mass0, rho_u=ru0, rho_v=rv0, rho_w=rw0, rho_e=E0)
    S = {k: a.copy() for k, a in init.items()}
    tmp = dict(mt=mass0.copy(), rut=ru0.copy(), rvt=rv0.copy(), rwt=rw0.copy(), ret=E0.copy())
    bnd = om.table("is_bnd").astype(f32)
    rev = f32(1) - bnd
    Gx, Gy, Gz = (lambda a: om.grad(a, 1)), (lambda a: om.grad(a, 2)), (lambda a: om.grad(a, 4))
    for _ in range(steps):
        rho, u, v, w, p = S["rho"], S["u"], S["v"], S["w"], S["p"]
        dt = f32(om.cfl3d(cfl, u, v, w, p, rho))
        hd = half * dt
        rho = S["mass"] / vol
        u = (S["rho_u"] / rho) / vol
        v = (S["rho_v"] / rho) / vol
        w = (S["rho_w"] / rho) / vol
        p = ((S["rho_e"] / vol) - ((half * rho) * (((u * u) + (v * v)) + (w * w)))) * g1
        rho = rho - hd * ((((((u * Gx(rho)) + (rho * Gx(u))) + (v * Gy(rho))) + (rho * Gy(v))) + (w * Gz(rho)))
                          + (rho * Gz(w)))
        u = u - hd * ((((u * Gx(u)) + (v * Gy(u))) + (w * Gz(u))) + ((f32(1) / rho) * Gx(p)))
        v = v - hd * ((((u * Gx(v)) + (v * Gy(v))) + (w * Gz(v))) + ((f32(1) / rho) * Gy(p)))
        w = w - hd * ((((u * Gx(w)) + (v * Gy(w))) + (w * Gz(w))) + ((f32(1) / rho) * Gz(p)))
        p = p - hd * (((((gamma * p) * ((Gx(u) + Gy(v)) + Gz(w))) + (u * Gx(p))) + (v * Gy(p))) + (w * Gz(p)))
        E = (p / g1) + ((half * rho) * (((u * u) + (v * v)) + (w * w)))
        tmp["mt"] = (zero * tmp["mt"]) - ((Gx(rho * u) + Gy(rho * v)) + Gz(rho * w))
        tmp["rut"] = (zero * tmp["rut"]) - ((Gx(((rho * u) * u) + p) + Gy((rho * v) * u)) + Gz((rho * w) * u))
        tmp["rvt"] = (zero * tmp["rvt"]) - ((Gx((rho * v) * u) + Gy(((rho * v) * v) + p)) + Gz((rho * v) * w))
        tmp["rwt"] = (zero * tmp["rwt"]) - ((Gx((rho * w) * u) + Gy((rho * w) * v)) + Gz(((rho * w) * w) + p))
        tmp["ret"] = (zero * tmp["ret"]) - ((Gx((E + p) * u) + Gy((E + p) * v)) + Gz((E + p) * w))
        S["mass"] = dt * (tmp["mt"] - (om.stab(rho, 7) * two)) + S["mass"]
        S["rho_u"] = dt * (tmp["rut"] - (om.stab(rho * u, 7) * two)) + S["rho_u"]
        S["rho_v"] = dt * (tmp["rvt"] - (om.stab(rho * v, 7) * two)) + S["rho_v"]
        S["rho_w"] = dt * (tmp["rwt"] - (om.stab(rho * w, 7) * two)) + S["rho_w"]
        S["rho_e"] = dt * (tmp["ret"] - (om.stab(E, 7) * two)) + S["rho_e"]
        S.update(rho=rho, u=u, v=v, w=w, p=p)
        for k in S:   # boundary_copy_cu: arr * rev + initial * bnd
            S[k] = (S[k] * rev) + (init[k] * bnd)
    assert len(printed) == 10
    for k, ref in S.items():
        _check_printed(printed, k, ref)
    assert "Time difference" in out


@pytest.mark.parametrize("name,args", [
    ("cfdARCO_euler", ["-Lx", "12", "-Ly", "10", "-Lz", "14", "-t", "12", "-dt", "0.1", "-c"]),
    ("cfdARCO_sod_shock", ["-Lx", "64", "-Ly", "2", "-Lz", "2", "-t", "30", "-dt", "0.2", "-c"]),
])
def test_false_flagged_equations_fused_equal_per_operator(name, args):
    """The drivers' equations flagged `false` (d1t with stabilisation) run as
    fused kernels, with multi-operator stabilised operands from kernels of
    their own or (CFDARCO_STAB_IN_KERNEL=1) evaluated inside the fused kernel;
    every printed sum and maximum of both equals the run that evaluates them
    operator by operator."""
    def lines(out):
        return re.findall(r"At \S+ Current sum = \S+ Current max = \S+", out)
    fused = lines(_run_driver(name, args))
    in_kernel = lines(_run_driver(name, args, stab_in_kernel=True))
    per_op = lines(_run_driver(name, args, per_op=True))
    assert fused and fused == per_op and in_kernel == per_op


def test_device_point_sources_equal_host_callback():
    """cpp/bench/point_source_wave.cpp (the wave example's problem with its
    point sources set by a device callback) prints the same final sum and
    maximum as the example as shipped (host callback, a host round trip of the
    field every step)."""
    psw = os.path.join(BUILD, "bench", "point_source_wave")
    if not os.path.exists(psw):
        pytest.fail(f"{psw} not built (run __graft_entry__.build())")
    # fields of 1.3 MiB: the host vectors take the large-block path
    # (page-locked once the device is enabled)
    args = ["-Lx", "72", "-Ly", "72", "-Lz", "64", "-t", "30", "-c"]
    with tempfile.TemporaryDirectory() as tmp:
        res = subprocess.run([psw, *args], capture_output=True, text=True, timeout=300, cwd=tmp)
    assert res.returncode == 0, res.stdout + res.stderr
    dev = re.findall(r"At u Current sum = \S+ Current max = \S+", res.stdout)
    assert dev, res.stdout
    host = re.findall(r"At u Current sum = \S+ Current max = \S+", _run_driver("cfdARCO_wave", args))
    assert dev == host


@pytest.mark.parametrize("in_kernel", [False, True])
def test_stabilised_operand_placement(in_kernel):
    """The Euler driver's stab_tot(E) (E of five fields and eight operators)
    has its operand computed by one generated kernel, while rho * u (one
    operator) runs operator by operator; with CFDARCO_STAB_IN_KERNEL=1 on a
    uniform mesh, E is evaluated inside the fused kernel instead.  The printed
    results equal the per-operator run (test above)."""
    path = os.path.join(DRIVERS, "cfdARCO_euler")
    if not os.path.exists(path):
        pytest.skip("reference driver binaries are built only where the reference sources exist")
    env = dict(os.environ, CFDARCO_JIT_PRINT="1")
    env.pop("CFDARCO_PER_OP", None)
    env.pop("CFDARCO_STAB_IN_KERNEL", None)
    if in_kernel:
        env["CFDARCO_STAB_IN_KERNEL"] = "1"
    with tempfile.TemporaryDirectory() as tmp:
        res = subprocess.run([path, "-Lx", "12", "-Ly", "10", "-Lz", "14", "-t", "3", "-dt", "0.1", "-c"],
                             capture_output=True, text=True, timeout=300, cwd=tmp, env=env)
    assert res.returncode == 0, res.stdout + res.stderr
    sources = res.stdout.split("Building kernel:")[1:]
    # kernel bodies that call the in-kernel stabilisation (the prelude defines it in every source)
    assert sum(" = jstab_axis(g, T" in src for src in sources) == (1 if in_kernel else 0)
    assert sum("void __launch_bounds__(256) cfd_operand_" in src for src in sources) == (0 if in_kernel else 1)


def test_reference_sod_driver_unchanged(oracle_lib):
    """examples/main_3d_sod_shock.cpp compiled unchanged (unit cells: the
    driver adds per-volume fluxes to cell totals, so it is written for dx = 1):
    its printed sums and maxima of rho, u, p equal the oracle's run of the same
    twelve equations, boundary copy and 1D CFL policy."""
    n, steps, cfl = (64, 2, 2), 40, 0.2
    out = _run_driver("cfdARCO_sod_shock", ["-Lx", "64", "-Ly", "2", "-Lz", "2", "-t", str(steps), "-dt", str(cfl),
                                            "-c"])
    printed = {v: (float(s), float(m)) for v, s, m in re.findall(r"At (\S+) Current sum = (\S+) Current max = (\S+)",
                                                                 out)}
    om = _mesh(oracle_lib, n, (1.0, 1.0, 1.0))
    ccx, vol = om.table("ccx"), om.table("vol")
    lx = f32(64.0)
    rho0 = np.where(f32(0.5) * lx < ccx, f32(1), f32(0.125)).astype(f32)
    p0 = np.where(f32(0.5) * lx < ccx, f32(1), f32(0.1)).astype(f32)
    u0 = np.zeros_like(rho0)
    gamma = f32(1.4)
    g1 = gamma - f32(1)
    half, zero, two = f32(0.5), f32(0), f32(2)
    mass, rho_u = rho0 * vol, rho0 * u0 * vol
    rho_e = (p0 / g1 + half * rho0 * (u0 * u0)) * vol
    rho, u, p = rho0.copy(), u0.copy(), p0.copy()
    mt, rt, et = mass.copy(), rho_u.copy(), rho_e.copy()
    G = lambda a: om.grad(a, 1)

    def bc(a):  # examples/main_3d_sod_shock.cpp:70-104: cells 0 and N-1 copy their neighbour
        r = a.copy()
        r[0], r[-1] = a[1], a[-2]
        return r

    for _ in range(steps):
        dt = f32(om.cfl1d(cfl, u, p, rho))
        rho = mass / vol
        u = (rho_u / rho) / vol
        p = ((rho_e / vol) - ((half * rho) * (u * u))) * g1
        rho = rho - (half * dt) * ((u * G(rho)) + (rho * G(u)))
        u = u - (half * dt) * ((u * G(u)) + ((f32(1) / rho) * G(p)))
        p = p - (half * dt) * (((gamma * p) * G(u)) + (u * G(p)))
        E = lambda: (p / g1) + ((half * rho) * (u * u))
        mt = (zero * mt) - G(rho * u)
        rt = (zero * rt) - G(((rho * u) * u) + p)
        et = (zero * et) - G((E() + p) * u)
        mass = dt * (mt - (om.stab(rho, 1) * two)) + mass
        rho_u = dt * (rt - (om.stab(rho * u, 1) * two)) + rho_u
        rho_e = dt * (et - (om.stab(E(), 1) * two)) + rho_e
        rho, u, p, mass, rho_u, rho_e = map(bc, (rho, u, p, mass, rho_u, rho_e))
    assert np.all(np.isfinite(rho)) and rho.min() > 0.1
    for name, ref in (("rho", rho), ("u", u), ("p", p)):
        s, mx = printed[name]
        assert s == pytest.approx(float(_sum(ref)), rel=1e-5, abs=1e-5), name
        assert mx == pytest.approx(float(ref.max()), rel=1e-5, abs=1e-6), name
