"""GPU parity: the HIP library (through the C-ABI) against the CPU oracle on
the same seeded inputs.

Tolerance policy: the library and the oracle evaluate the same operation
sequence with no FMA contraction, so results are compared for exact equality
(==) in FP64 and FP32 alike, which is stronger than the 1e-12 relative L2
required for FP64.  FP32-vs-FP64 error is reported separately.
"""
import numpy as np
import pytest

from cfdarcov2_amd import workloads

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _t(a, dtype):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=dtype)).cuda()


def _np(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


MESHES = [
    # (name, shape, spacing, perturb seed, force_general)
    ("uniform", (9, 7, 10), (0.5, 0.25, 0.125), None, False),
    # one- and two-cell extents: a cell's axis neighbours are itself or each other
    ("uniform-thin", (1, 2, 5), (0.5, 0.25, 0.125), None, False),
    # several 64-cell z tiles, 4-row y tiles and 8-plane x chunks of the
    # marching stabilisation, with partial ones at the ends
    ("uniform-wide", (19, 11, 133), (0.5, 0.25, 0.125), None, False),
    ("uniform-as-general", (9, 7, 10), (0.5, 0.25, 0.125), None, True),
    ("non-dyadic", (6, 8, 5), (0.1, 0.3, 0.7), None, False),
    ("perturbed", (8, 6, 7), (0.125, 0.25, 0.5), 42, False),
]


def _pair(hip_ctx, oracle_lib, case, dtype):
    from cfdarcov2_amd import capi
    name, n, d, seed, force = case
    off = None if seed is None else workloads.perturbed_vertex_offsets(*n, *d, seed=seed, frac=0.2)
    gm = capi.HexMesh(hip_ctx, *n, *d, dtype=dtype, vertex_offsets=off, force_general=force)
    om = oracle_lib.OracleMesh(*n, *d, dtype=dtype, perturb=off)
    return gm, om


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("case", MESHES, ids=[c[0] for c in MESHES])
def test_operators_match_oracle(hip_ctx, oracle_lib, case, dtype):
    print(f"\n[case] operators {case[0]} {np.dtype(dtype).name}", flush=True)
    gm, om = _pair(hip_ctx, oracle_lib, case, dtype)
    info = gm.info()
    if case[0] == "uniform":
        assert info.uniform == 1
    if case[0] in ("uniform-as-general", "perturbed"):
        assert info.uniform == 0
    rng = np.random.default_rng(7)
    phi = rng.standard_normal(om.N).astype(dtype)
    tphi = _t(phi, dtype)
    for mask in (1, 2, 4, 7, 5):
        for wall in (-1, 0, 1, 2):
            got = _np(gm.grad(tphi, mask, wall))
            ref = om.grad(phi, mask, wall)
            assert np.array_equal(got, ref), f"grad mask={mask} wall={wall}"
        assert np.array_equal(_np(gm.lapl(tphi, mask)), om.lapl(phi, mask)), f"lapl mask={mask}"
        assert np.array_equal(_np(gm.stab(tphi, mask)), om.stab(phi, mask)), f"stab mask={mask}"
    gm.close()


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_heat_config_bit_exact(hip_ctx, oracle_lib, dtype):
    """Heat 64^3, Gaussian, alpha 1e-3, dt = 0.2 dx^2/alpha, 200 steps.  r = 0.2
    exceeds the 3D explicit limit 1/6, so round-off grows ~1.4x per step; only
    an identical operation order keeps GPU and oracle together."""
    from cfdarcov2_amd import capi
    print(f"\n[case] heat 64^3 {np.dtype(dtype).name}", flush=True)
    case = workloads.heat_config()
    gm = capi.HexMesh(hip_ctx, *case.n, *case.spacing, dtype=dtype)
    om = oracle_lib.OracleMesh(*case.n, *case.spacing, dtype=dtype)
    T0 = case.T0.astype(dtype)
    a, b = _t(T0, dtype), torch.empty_like(_t(T0, dtype))
    res = gm.heat_run(a, b, case.alpha, case.dt, case.steps)
    got = _np(res)
    ref = om.heat_run(T0, case.alpha, case.dt, case.steps)
    assert np.array_equal(got, ref)
    if dtype == np.float32:
        o64 = oracle_lib.OracleMesh(*case.n, *case.spacing, dtype=np.float64)
        r64 = o64.heat_run(case.T0, case.alpha, case.dt, case.steps)
        err = np.linalg.norm(got - r64) / np.linalg.norm(r64)
        print(f"heat FP32 vs FP64 oracle rel L2 after {case.steps} steps (unstable dt): {err:.3e}")


def test_heat_alpha_field_and_general_mesh(hip_ctx, oracle_lib):
    from cfdarcov2_amd import capi
    n, d = (10, 9, 8), (0.1, 0.1, 0.1)
    off = workloads.perturbed_vertex_offsets(*n, *d, seed=9, frac=0.15)
    gm = capi.HexMesh(hip_ctx, *n, *d, dtype=np.float64, vertex_offsets=off)
    om = oracle_lib.OracleMesh(*n, *d, dtype=np.float64, perturb=off)
    rng = np.random.default_rng(3)
    T0 = rng.random(om.N)
    af = 1e-3 * (1 + rng.random(om.N))
    a, b = _t(T0, np.float64), torch.empty(om.N, dtype=torch.float64, device="cuda")
    got = _np(gm.heat_run(a, b, 0.0, 0.5, 25, alpha_field=_t(af, np.float64)))
    ref = om.heat_run(T0, 0.0, 0.5, 25, alpha_field=af)
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("n", [(32, 32, 32), (17, 9, 70)])
def test_wave_fused_matches_sequential_oracle(hip_ctx, oracle_lib, dtype, n):
    from cfdarcov2_amd import capi
    print(f"\n[case] wave {n} {np.dtype(dtype).name}", flush=True)
    # dyadic spacings keep x*dx exact, so the mesh is exactly uniform
    d = (1.0 / 32, 1.0 / 16, 1.0 / 64)
    X, Y, Z = workloads.cell_centres(*n, *d)
    p0 = np.exp(-((X - 0.4) ** 2 + (Y - 0.55) ** 2 + (Z - 0.5) ** 2) / (2 * 0.1 ** 2))
    dt = 0.4 * min(d)
    gm = capi.HexMesh(hip_ctx, *n, *d, dtype=dtype)
    om = oracle_lib.OracleMesh(*n, *d, dtype=dtype)
    N = om.N
    z = np.zeros(N, dtype=dtype)
    A = [_t(p0.astype(dtype), dtype)] + [_t(z, dtype) for _ in range(3)]
    B = [torch.empty_like(A[0]) for _ in range(4)]
    steps = 60
    got = [_np(t) for t in gm.wave_run(A, B, -1.0, dt, steps)]
    ref = om.wave_run(p0.astype(dtype), z, z, z, -1.0, dt, steps)
    for name, g, r in zip("puvw", got, ref):
        assert np.array_equal(g, r), f"field {name}: max diff {np.max(np.abs(g - r))}"


SHAPES = [(0, 16, 32, 1), (1, 4, 1, 0), (1, 8, 3, 1), (1, 16, 32, 1), (1, 4, 7, 1), (1, 16, 5, 0),
          (1, 32, 4, 1), (1, 32, 64, 0), (1, 8, 16, 0)]


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("shape", SHAPES, ids=[f"v{a}-ty{b}-L{c}-r{d}" for a, b, c, d in SHAPES])
def test_wave_kernel_shapes(hip_ctx, oracle_lib, dtype, shape):
    """Every tile shape of the fused acoustic kernel (per-thread form and the
    2.5D marching form with partial tiles in y and z and chunks that do not
    divide x) gives the oracle's bits."""
    from cfdarcov2_amd import capi
    print(f"\n[case] wave shape {shape} {np.dtype(dtype).name}", flush=True)
    n = (19, 21, 70)
    d = (1.0 / 32, 1.0 / 32, 1.0 / 64)
    X, Y, Z = workloads.cell_centres(*n, *d)
    rng = np.random.default_rng(17)
    p0 = np.exp(-((X - 0.3) ** 2 + (Y - 0.3) ** 2 + (Z - 0.6) ** 2) / (2 * 0.08 ** 2))
    u0, v0, w0 = (0.1 * rng.standard_normal(p0.size) for _ in range(3))
    gm = capi.HexMesh(hip_ctx, *n, *d, dtype=dtype)
    gm.set_wave_shape(*shape)
    om = oracle_lib.OracleMesh(*n, *d, dtype=dtype)
    A = [_t(a, dtype) for a in (p0, u0, v0, w0)]
    B = [torch.empty_like(A[0]) for _ in range(4)]
    dt = 0.4 / 64
    got = [_np(t) for t in gm.wave_run(A, B, -1.0, dt, 7)]
    ref = om.wave_run(*(a.astype(dtype) for a in (p0, u0, v0, w0)), -1.0, dt, 7)
    for name, g, r in zip("puvw", got, ref):
        assert np.array_equal(g, r), f"field {name}"


@pytest.mark.parametrize("kind", ["uniform", "non-dyadic", "perturbed"])
def test_wave_K_field(hip_ctx, oracle_lib, kind):
    from cfdarcov2_amd import capi
    n = (12, 10, 14)
    d = (0.125, 0.125, 0.0625) if kind == "uniform" else (0.1, 0.1, 0.1)
    off = workloads.perturbed_vertex_offsets(*n, *d, seed=2, frac=0.2) if kind == "perturbed" else None
    rng = np.random.default_rng(5)
    gm = capi.HexMesh(hip_ctx, *n, *d, dtype=np.float64, vertex_offsets=off)
    assert gm.info().uniform == (kind == "uniform")
    om = oracle_lib.OracleMesh(*n, *d, dtype=np.float64, perturb=off)
    p0, u0, v0, w0 = (rng.standard_normal(om.N) for _ in range(4))
    K = -(0.5 + rng.random(om.N)) ** 2
    A = [_t(x, np.float64) for x in (p0, u0, v0, w0)]
    B = [torch.empty_like(A[0]) for _ in range(4)]
    got = [_np(t) for t in gm.wave_run(A, B, 0.0, 0.02, 9, K_field=_t(K, np.float64))]
    ref = om.wave_run(p0, u0, v0, w0, 0.0, 0.02, 9, K_field=K)
    for g, r in zip(got, ref):
        assert np.array_equal(g, r)


def test_wave_full_size_properties(hip_ctx, oracle_lib):
    """256^3 benchmark case: (a) FP32, two steps equal the oracle bit for bit
    at full size; (b) FP64 linearity, run(2 p0) == 2 run(p0) exactly over 100
    steps (scaling by 2 is exact while every value stays in the normal range,
    which holds in FP64; in FP32 ulp differences born in subnormals near the
    walls ride the steep wave front into larger values, so FP32 is checked to
    1e-6); (c) FP32 tracks FP64 and the energy stays bounded."""
    from cfdarcov2_amd import capi
    print("\n[case] wave 256^3 properties", flush=True)
    case = workloads.wave_config(n=256, steps=100)
    N = 256 ** 3

    def run(mesh, p0, tdt, steps):
        A = [torch.from_numpy(p0).to(tdt).cuda()] + [torch.zeros(N, dtype=tdt, device="cuda") for _ in range(3)]
        B = [torch.empty_like(A[0]) for _ in range(4)]
        return mesh.wave_run(A, B, -case.c ** 2, case.dt, steps)

    g32 = capi.HexMesh(hip_ctx, *case.n, *case.spacing, dtype=np.float32)
    p0f = case.p0.astype(np.float32)
    # (a) full-size exactness over two steps
    om = oracle_lib.OracleMesh(*case.n, *case.spacing, dtype=np.float32)
    zf = np.zeros(N, dtype=np.float32)
    ref2 = om.wave_run(p0f, zf, zf, zf, -1.0, case.dt, 2)
    del om
    got2 = [_np(t) for t in run(g32, p0f, torch.float32, 2)]
    for name, g, r in zip("puvw", got2, ref2):
        assert np.array_equal(g, r), f"256^3 field {name}"
    for shape in ((7, 4, 4, 0), (8, 8, 16, 3), (9, 4, 4, 2)):  # row forms, forward and reversed tile order
        g32.set_wave_shape(*shape)
        got2 = [_np(t) for t in run(g32, p0f, torch.float32, 2)]
        for name, g, r in zip("puvw", got2, ref2):
            assert np.array_equal(g, r), f"256^3 field {name}, shape {shape}"
    om = oracle_lib.OracleMesh(*case.n, *case.spacing, dtype=np.float32)
    ref4 = om.wave_run(p0f, zf, zf, zf, -1.0, case.dt, 4)
    del om
    g32.set_wave_shape(11, 8, 16, 3)  # two passes of two steps each
    got4 = [_np(t) for t in run(g32, p0f, torch.float32, 4)]
    for name, g, r in zip("puvw", got4, ref4):
        assert np.array_equal(g, r), f"256^3 field {name}, two steps per pass"
    g32.set_wave_shape(1, 16, 32, 1)
    del got2, ref2, got4, ref4
    # (b) linearity
    g64 = capi.HexMesh(hip_ctx, *case.n, *case.spacing, dtype=np.float64)
    d1 = run(g64, case.p0, torch.float64, case.steps)
    d2 = run(g64, 2 * case.p0, torch.float64, case.steps)
    for a, b in zip(d1, d2):
        assert torch.equal(2 * a, b)
    del d2
    r1 = run(g32, p0f, torch.float32, case.steps)
    r2 = run(g32, 2 * p0f, torch.float32, case.steps)
    for a, b in zip(r1, r2):
        assert float(torch.max(torch.abs(2 * a - b))) <= 1e-6 * float(torch.max(torch.abs(b)))
    del r2
    # (c) FP32 vs FP64, energy
    E0 = workloads.wave_energy(case.p0, 0, 0, 0, 1.0)
    e32 = workloads.wave_energy(*[_np(t).astype(np.float64) for t in r1], 1.0)
    f64 = [_np(t) for t in d1]
    e64 = workloads.wave_energy(*f64, 1.0)
    assert abs(e32 - e64) < 1e-5 * E0
    assert 0.8 * E0 < e64 < 1.1 * E0
    rel = [np.linalg.norm(_np(a).astype(np.float64) - b) / np.linalg.norm(f64[0]) for a, b in zip(r1, f64)]
    print("wave 256^3 FP32 vs FP64 rel L2 (p,u,v,w) after 100 steps:", ", ".join(f"{x:.2e}" for x in rel))
    assert max(rel) < 1e-4


def test_wave_two_step_f64_full_size(hip_ctx, oracle_lib):
    """256^3 benchmark case in FP64: two passes of the two-step row form
    (2-row tiles), forward and reversed tile order, equal four oracle steps
    bit for bit."""
    from cfdarcov2_amd import capi
    print("\n[case] wave 256^3 FP64 two-step", flush=True)
    case = workloads.wave_config(n=256, steps=4)
    N = 256 ** 3
    om = oracle_lib.OracleMesh(*case.n, *case.spacing, dtype=np.float64)
    z = np.zeros(N)
    ref = om.wave_run(case.p0, z, z, z, -1.0, case.dt, 4)
    del om
    g = capi.HexMesh(hip_ctx, *case.n, *case.spacing, dtype=np.float64)
    for shape in ((11, 2, 16, 3), (12, 2, 32, 1)):
        g.set_wave_shape(*shape)
        A = [torch.from_numpy(case.p0).cuda()] + [torch.zeros(N, dtype=torch.float64, device="cuda") for _ in range(3)]
        B = [torch.empty_like(A[0]) for _ in range(4)]
        got = [_np(t) for t in g.wave_run(A, B, -case.c ** 2, case.dt, 4)]
        for name, a, r in zip("puvw", got, ref):
            assert np.array_equal(a, r), f"256^3 FP64 field {name}, shape {shape}"


def test_ewise_and_reductions(hip_ctx, oracle_lib):
    from cfdarcov2_amd import capi
    rng = np.random.default_rng(11)
    # lengths that fill whole 16-byte vectors and ones that leave a tail;
    # offset views (not 16-byte aligned) take the one-element-per-thread kernel
    for dtype, n, off in [(np.float32, 1000, 0), (np.float64, 1000, 0), (np.float32, 1003, 0),
                          (np.float64, 1001, 0), (np.float32, 7, 0), (np.float32, 999, 1), (np.float64, 999, 1)]:
        a, b, c = (rng.standard_normal(n + off).astype(dtype) for _ in range(3))
        b = np.where(np.abs(b) < 0.1, 0.5, b).astype(dtype)
        ta, tb, tc = _t(a, dtype)[off:], _t(b, dtype)[off:], _t(c, dtype)[off:]
        a, b, c = a[off:], b[off:], c[off:]
        out = torch.empty(n + off, dtype=ta.dtype, device="cuda")[off:]
        s = dtype(0.37)
        expect = {"add": a + b, "sub": a - b, "mul": a * b, "div": a / b, "neg": -a, "muls": a * s,
                  "divs": a / s, "smul": s * a, "axpy": s * a + b, "leap": ((s * s) * a + dtype(2) * b) - c,
                  "fill": np.full_like(a, s), "copy": a, "adds": a + s, "sadd": s + a, "subs": a - s,
                  "ssub": s - a, "sdiv": s / b}
        for op, ref in expect.items():
            if op == "sdiv":
                hip_ctx.ewise(op, tb, None, None, float(s), out)
            else:
                hip_ctx.ewise(op, ta, tb, tc, float(s), out)
            assert np.array_equal(_np(out), ref.astype(dtype)), (op, np.dtype(dtype).name, n, off)
    # CFL minimum against the oracle (3D and 1D forms)
    n, d = (8, 8, 8), (0.5, 0.25, 0.125)
    gm = capi.HexMesh(hip_ctx, *n, *d, dtype=np.float64)
    om = oracle_lib.OracleMesh(*n, *d, dtype=np.float64)
    rho, p = 1 + rng.random(om.N), 1 + rng.random(om.N)
    u, v, w = (rng.standard_normal(om.N) for _ in range(3))
    T = lambda x: _t(x, np.float64)
    assert gm.cfl_min(T(u), T(p), T(rho), T(v), T(w), 0.4) == om.cfl3d(0.4, u, v, w, p, rho)
    assert gm.cfl_min(T(u), T(p), T(rho), None, None, 0.4) == om.cfl1d(0.4, u, p, rho)
    x = rng.standard_normal(5000)
    x[[3, 77, 4999]] = [np.nan, np.inf, -np.inf]
    assert hip_ctx.count_nonfinite(T(x)) == 3


def _ulps(got, ref):
    """Distance in units in the last place, counted on the ordered integer
    image of the floats (0 for equal values, NaN == NaN, +0 == -0)."""
    it = np.int32 if got.dtype == np.float32 else np.int64
    def key(a):
        i = a.view(it).astype(np.int64)
        return np.where(i < 0, np.iinfo(it).min - i, i)
    d = np.abs(key(got) - key(ref))
    d[np.isnan(got) & np.isnan(ref)] = 0
    return d


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_ewise_exp_abs(hip_ctx, dtype):
    """exp and abs (no device form in the reference, whose exp/abs run on the
    host only).  abs is exact: bit for bit against numpy.  exp is the device
    libm's, within 1 ulp of the exact value; it is held to 2 ulp of numpy's
    exp (in float64, itself within 1 ulp; for FP32 rounded from float64), and
    exp(0) = 1, exp(-inf) = 0, exp(inf) = inf, overflow to inf and NaN
    propagation exactly."""
    rng = np.random.default_rng(17)
    lo, hi = (-87.0, 88.0) if dtype == np.float32 else (-708.0, 709.0)
    for n, off in ((4099, 0), (999, 1)):  # vector kernel with a tail; unaligned element kernel
        a = rng.uniform(lo, hi, n + off)
        a[off: off + 200] = rng.standard_normal(200)  # the common range as well
        special = [0.0, -0.0, np.inf, -np.inf, np.nan, 1e4, -1e4, 1.0, -1.0]
        a[off + 200: off + 200 + len(special)] = special
        a = a.astype(dtype)
        ta = _t(a, dtype)[off:]
        a = a[off:]
        out = torch.empty(n + off, dtype=ta.dtype, device="cuda")[off:]
        hip_ctx.ewise("abs", ta, None, None, 0.0, out)
        got = _np(out)
        assert np.array_equal(got.view(np.uint8), np.abs(a).view(np.uint8)), (np.dtype(dtype).name, n, off)
        hip_ctx.ewise("exp", ta, None, None, 0.0, out)
        got = _np(out)
        with np.errstate(over="ignore"):
            ref = np.exp(a.astype(np.float64)).astype(dtype)
        u = _ulps(got, ref)
        assert u.max() <= 2, (np.dtype(dtype).name, n, off, int(u.max()), a[np.argmax(u)])
        for x, want in ((0.0, 1.0), (-0.0, 1.0), (np.inf, np.inf), (-np.inf, 0.0), (1e4, np.inf), (-1e4, 0.0)):
            i = np.flatnonzero((a == x) & (np.signbit(a) == np.signbit(x)))
            assert i.size and np.all(got[i] == want), (x, got[i])
        assert np.all(np.isnan(got[np.isnan(a)]))


def test_error_paths(hip_ctx):
    """C-ABI misuse and pool exhaustion come back as status codes."""
    from cfdarcov2_amd import capi
    with pytest.raises(capi.CfdError) as e:
        capi.HexMesh(hip_ctx, 0, 4, 4, 1.0, 1.0, 1.0)
    assert e.value.code == -1
    gm = capi.HexMesh(hip_ctx, 4, 4, 4, 1.0, 1.0, 1.0, dtype=np.float32)
    t = torch.zeros(64, device="cuda")
    assert gm.lib.cfd_grad(gm.h, None, None, 1, -1, None) == -1
    rc = gm.lib.cfd_grad(gm.h, capi._dptr(t), capi._dptr(t), 9, -1, None)
    assert rc == -1 and b"axis_mask" in gm.lib.cfd_last_error()
    with pytest.raises(capi.CfdError) as e:
        gm.wave_run([t, t, t, t], [t, t, t, t], -1.0, 0.1, 1)
    assert e.value.code == -1
    used, _ = hip_ctx.pool_stats()
    hip_ctx.pool_set_limit(used + 1024)
    try:
        with pytest.raises(capi.CfdError) as e:
            hip_ctx.alloc(1 << 20)
        assert e.value.code == -3
    finally:
        hip_ctx.pool_set_limit(0)
    p = hip_ctx.alloc(1 << 20)
    hip_ctx.free(p)


@pytest.mark.parametrize("shape", [(1, 4, 8, 0), (1, 8, 3, 1), (0, 16, 32, 1)])
def test_wave_uniform_volume_not_power_of_two(hip_ctx, oracle_lib, shape):
    """A uniform mesh whose cell volume is not a power of two (the kernels
    then divide by the volume; powers of two multiply by the exact reciprocal):
    the oracle's bits in both cases, heat included."""
    from cfdarcov2_amd import capi
    n, d = (13, 11, 70), (0.375, 0.25, 0.125)      # V = 3/256
    gm = capi.HexMesh(hip_ctx, *n, *d, dtype=np.float32)
    assert gm.info().uniform == 1
    gm.set_wave_shape(*shape)
    om = oracle_lib.OracleMesh(*n, *d, dtype=np.float32)
    rng = np.random.default_rng(23)
    f = [rng.standard_normal(om.N).astype(np.float32) for _ in range(4)]
    A = [_t(a, np.float32) for a in f]
    B = [torch.empty_like(A[0]) for _ in range(4)]
    got = [_np(t) for t in gm.wave_run(A, B, -1.0, 0.01, 5)]
    ref = om.wave_run(*f, -1.0, 0.01, 5)
    for name, g, r in zip("puvw", got, ref):
        assert np.array_equal(g, r), name
    a, b = _t(f[0], np.float32), torch.empty_like(_t(f[0], np.float32))
    assert np.array_equal(_np(gm.heat_run(a, b, 0.01, 0.5, 4)), om.heat_run(f[0], 0.01, 0.5, 4))


ROW_CASES = [
    # (extents, spacing, shape): partial y tiles, chunks that do not divide x,
    # idle lanes (z < 256), one-lane rows, full 256-cell rows, a volume that is
    # not a power of two
    ((19, 21, 68), (1.0 / 32, 1.0 / 32, 1.0 / 64), (7, 4, 5, 0)),
    ((19, 21, 68), (1.0 / 32, 1.0 / 32, 1.0 / 64), (7, 8, 3, 1)),
    ((19, 21, 68), (1.0 / 32, 1.0 / 32, 1.0 / 64), (7, 14, 1, 0)),
    ((19, 21, 68), (1.0 / 32, 1.0 / 32, 1.0 / 64), (7, 14, 3, 1)),
    ((19, 21, 68), (1.0 / 32, 1.0 / 32, 1.0 / 64), (7, 8, 1, 0)),
    ((19, 21, 68), (1.0 / 32, 1.0 / 32, 1.0 / 64), (7, 4, 1, 1)),
    ((9, 30, 256), (1.0 / 16, 1.0 / 32, 1.0 / 256), (7, 8, 4, 1)),
    ((6, 5, 4), (0.25, 0.25, 0.25), (7, 4, 2, 0)),
    ((1, 3, 8), (0.25, 0.25, 0.25), (7, 4, 16, 0)),
    ((13, 11, 60), (0.375, 0.25, 0.125), (7, 8, 16, 0)),
    ((19, 21, 68), (1.0 / 32, 1.0 / 32, 1.0 / 64), (8, 4, 5, 2)),
    ((19, 21, 68), (1.0 / 32, 1.0 / 32, 1.0 / 64), (8, 8, 3, 3)),
    ((19, 21, 68), (1.0 / 32, 1.0 / 32, 1.0 / 64), (7, 14, 4, 2)),
    ((19, 21, 68), (1.0 / 32, 1.0 / 32, 1.0 / 64), (9, 4, 6, 2)),
    ((9, 30, 256), (1.0 / 16, 1.0 / 32, 1.0 / 256), (9, 8, 4, 3)),
    # two steps per pass (FP32): 7 steps = one pair of two-step passes + 3 single steps
    ((19, 21, 68), (1.0 / 32, 1.0 / 32, 1.0 / 64), (11, 4, 5, 0)),
    ((19, 21, 68), (1.0 / 32, 1.0 / 32, 1.0 / 64), (11, 8, 3, 3)),
    ((19, 21, 68), (1.0 / 32, 1.0 / 32, 1.0 / 64), (12, 10, 16, 2)),
    ((30, 27, 256), (1.0 / 32, 1.0 / 32, 1.0 / 256), (11, 8, 4, 1)),
    ((6, 5, 4), (0.25, 0.25, 0.25), (11, 4, 1, 0)),
    ((13, 11, 60), (0.375, 0.25, 0.125), (12, 8, 2, 3)),
    # two steps per pass with several rows per wave
    ((10, 70, 128), (1.0 / 16, 1.0 / 128, 1.0 / 128), (11, 8, 6, 1)),
    ((12, 50, 96), (1.0 / 16, 1.0 / 64, 1.0 / 128), (12, 10, 3, 2)),
    ((7, 45, 32), (0.125, 1.0 / 64, 1.0 / 32), (11, 4, 16, 3)),
    # two steps per pass, FP64 tiles (2 rows of full rows; 8 waves of several rows)
    ((19, 21, 68), (1.0 / 32, 1.0 / 32, 1.0 / 64), (11, 2, 5, 0)),
    ((30, 27, 256), (1.0 / 32, 1.0 / 32, 1.0 / 256), (12, 2, 4, 1)),
    ((9, 30, 256), (1.0 / 16, 1.0 / 32, 1.0 / 256), (11, 2, 16, 2)),
    ((6, 5, 4), (0.25, 0.25, 0.25), (12, 2, 1, 0)),
    ((13, 11, 60), (0.375, 0.25, 0.125), (11, 2, 2, 3)),
    ((10, 70, 128), (1.0 / 16, 1.0 / 128, 1.0 / 128), (12, 2, 6, 1)),
]


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("case", ROW_CASES, ids=[f"{c[0][0]}x{c[0][1]}x{c[0][2]}-v{c[2][0]}-ty{c[2][1]}-L{c[2][2]}-r{c[2][3]}" for c in ROW_CASES])
def test_wave_row_form(hip_ctx, oracle_lib, dtype, case):
    """Row form of the acoustic step (one wave per z row, four cells per lane,
    z-neighbours by lane shifts): the oracle's bits."""
    from cfdarcov2_amd import capi
    n, d, shape = case
    if dtype == np.float64 and (shape[1] == 14 or (shape[0] >= 11 and shape[1] != 2)):
        pytest.skip("14-row tiles and two-step tiles of 4 to 10 rows are FP32 only")
    if dtype == np.float32 and shape[0] >= 11 and shape[1] == 2:
        pytest.skip("two-step tiles of 2 rows are FP64 only")
    print(f"\n[case] wave row form n={n} shape={shape} {np.dtype(dtype).name}", flush=True)
    rng = np.random.default_rng(41)
    gm = capi.HexMesh(hip_ctx, *n, *d, dtype=dtype)
    gm.set_wave_shape(*shape)
    om = oracle_lib.OracleMesh(*n, *d, dtype=dtype)
    f = [rng.standard_normal(om.N).astype(dtype) for _ in range(4)]
    A = [_t(a, dtype) for a in f]
    B = [torch.empty_like(A[0]) for _ in range(4)]
    dt = 0.3 * min(d)
    got = [_np(t) for t in gm.wave_run(A, B, -1.0, dt, 7)]
    ref = om.wave_run(*f, -1.0, dt, 7)
    for name, g, r in zip("puvw", got, ref):
        assert np.array_equal(g, r), f"field {name}: max diff {np.max(np.abs(g - r))}"


@pytest.mark.parametrize("shape", [(7, 4, 3, 0), (11, 4, 3, 0)])
def test_wave_row_form_K_field_and_misaligned(hip_ctx, oracle_lib, shape):
    """Row form with a per-cell K field, then with field pointers that are not
    16-byte aligned (the launch takes the 2.5D tile instead): the oracle's bits
    in both."""
    from cfdarcov2_amd import capi
    n, d = (10, 9, 36), (0.125, 0.125, 0.0625)
    rng = np.random.default_rng(7)
    gm = capi.HexMesh(hip_ctx, *n, *d, dtype=np.float32)
    gm.set_wave_shape(*shape)
    om = oracle_lib.OracleMesh(*n, *d, dtype=np.float32)
    f = [rng.standard_normal(om.N).astype(np.float32) for _ in range(4)]
    K = (-(0.5 + rng.random(om.N)) ** 2).astype(np.float32)
    A = [_t(a, np.float32) for a in f]
    B = [torch.empty_like(A[0]) for _ in range(4)]
    got = [_np(t) for t in gm.wave_run(A, B, 0.0, 0.02, 9, K_field=_t(K, np.float32))]
    ref = om.wave_run(*f, 0.0, 0.02, 9, K_field=K)
    for name, g, r in zip("puvw", got, ref):
        assert np.array_equal(g, r), f"K field, field {name}"
    # views starting one element into a larger buffer
    big = [torch.zeros(om.N + 1, dtype=torch.float32, device="cuda") for _ in range(8)]
    A = [b[1:] for b in big[:4]]
    B = [b[1:] for b in big[4:]]
    for a, x in zip(A, f):
        a.copy_(_t(x, np.float32))
    got = [_np(t) for t in gm.wave_run(A, B, -1.0, 0.02, 9)]
    ref = om.wave_run(*f, -1.0, 0.02, 9)
    for name, g, r in zip("puvw", got, ref):
        assert np.array_equal(g, r), f"misaligned, field {name}"


def test_wave_row_form_rejects_unsupported_rows(hip_ctx):
    from cfdarcov2_amd import capi
    gm = capi.HexMesh(hip_ctx, 8, 8, 70, 0.125, 0.125, 0.125, dtype=np.float32)
    with pytest.raises(capi.CfdError) as e:
        gm.set_wave_shape(7, 4, 8, 0)
    assert e.value.code == -4
    with pytest.raises(capi.CfdError):
        capi.HexMesh(hip_ctx, 8, 8, 64, 0.125, 0.125, 0.125, dtype=np.float32).set_wave_shape(7, 6, 8, 0)
    with pytest.raises(capi.CfdError):
        capi.HexMesh(hip_ctx, 8, 8, 64, 0.125, 0.125, 0.125, dtype=np.float64).set_wave_shape(7, 14, 8, 0)
    # two-step tiles: 4, 8 or 10 rows in FP32, 2 rows in FP64
    with pytest.raises(capi.CfdError):
        capi.HexMesh(hip_ctx, 8, 8, 64, 0.125, 0.125, 0.125, dtype=np.float64).set_wave_shape(11, 8, 8, 0)
    with pytest.raises(capi.CfdError):
        capi.HexMesh(hip_ctx, 8, 8, 64, 0.125, 0.125, 0.125, dtype=np.float32).set_wave_shape(12, 2, 8, 0)
    with pytest.raises(capi.CfdError) as e:
        capi.HexMesh(hip_ctx, 8, 8, 70, 0.125, 0.125, 0.125, dtype=np.float64).set_wave_shape(11, 2, 8, 0)
    assert e.value.code == -4


def test_wave_retired_variants_rejected(hip_ctx):
    """Variant numbers outside 0, 1, 7-9, 11, 12 are refused as invalid."""
    from cfdarcov2_amd import capi
    gm = capi.HexMesh(hip_ctx, 8, 8, 64, 0.125, 0.125, 0.125, dtype=np.float32)
    for v in (2, 3, 4, 5, 6, 10, 13, -1):
        with pytest.raises(capi.CfdError) as e:
            gm.set_wave_shape(v, 4, 8, 0)
        assert e.value.code == -1, v
    gm.set_wave_shape(9, 4, 8, 0)


HEAT_ROW_CASES = [
    # (extents, spacing, (waves, chunk)): rows per wave 3, 1, 64, 4, 4 (volume
    # not a power of two) and 32 with two planes; partial tiles, chunks that do
    # not divide x
    ((13, 11, 68), (1.0 / 16, 1.0 / 16, 1.0 / 64), (8, 3)),
    ((9, 30, 256), (1.0 / 16, 1.0 / 32, 1.0 / 256), (16, 4)),
    ((6, 5, 4), (0.25, 0.25, 0.25), (8, 1)),
    ((40, 33, 64), (1.0 / 64, 1.0 / 64, 1.0 / 64), (12, 16)),
    ((13, 11, 60), (0.375, 0.25, 0.125), (8, 5)),
    ((2, 7, 8), (0.125, 0.125, 0.125), (16, 2)),
]


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("case", HEAT_ROW_CASES, ids=[f"{c[0][0]}x{c[0][1]}x{c[0][2]}-w{c[2][0]}-L{c[2][1]}" for c in HEAT_ROW_CASES])
def test_heat_two_step_row_form(hip_ctx, oracle_lib, dtype, case):
    """Two heat steps per launch in row form: 9 steps (two pairs of two-step
    launches and one single step) equal the oracle bit for bit."""
    from cfdarcov2_amd import capi
    n, d, (nw, L) = case
    print(f"\n[case] heat two-step row form n={n} waves={nw} chunk={L} {np.dtype(dtype).name}", flush=True)
    rng = np.random.default_rng(13)
    gm = capi.HexMesh(hip_ctx, *n, *d, dtype=dtype)
    gm.set_heat_shape(1, nw, L)
    om = oracle_lib.OracleMesh(*n, *d, dtype=dtype)
    T0 = rng.standard_normal(om.N).astype(dtype)
    alpha = 1e-2
    dt = 0.15 * min(d) ** 2 / alpha
    a, b = _t(T0, dtype), torch.empty(om.N, dtype=_t(T0, dtype).dtype, device="cuda")
    got = _np(gm.heat_run(a, b, alpha, dt, 9))
    ref = om.heat_run(T0, alpha, dt, 9)
    assert np.array_equal(got, ref), float(np.max(np.abs(got - ref)))


def test_heat_shape_errors(hip_ctx):
    from cfdarcov2_amd import capi
    gm = capi.HexMesh(hip_ctx, 8, 8, 70, 0.125, 0.125, 0.125, dtype=np.float32)
    with pytest.raises(capi.CfdError) as e:
        gm.set_heat_shape(1, 8, 8)
    assert e.value.code == -4
    with pytest.raises(capi.CfdError):
        capi.HexMesh(hip_ctx, 8, 8, 64, 0.125, 0.125, 0.125, dtype=np.float32).set_heat_shape(1, 10, 8)


HEAT_MULTI_CASES = [
    ((64, 64, 64), (1.0 / 64,) * 3, (4, 4), 200),      # the 64^3 workload's extents and length
    ((13, 11, 100), (1.0 / 16, 1.0 / 16, 1.0 / 64), (4, 2), 7),
    ((9, 10, 130), (0.375, 0.25, 0.125), (8, 2), 9),
    ((5, 3, 4), (0.25, 0.25, 0.25), (4, 4), 6),
    ((20, 17, 61), (1.0 / 32,) * 3, (8, 4), 11),
    ((8, 8, 64), (0.125,) * 3, (4, 2), 1),
    ((8, 8, 64), (0.125,) * 3, (4, 4), 2),
]


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("case", HEAT_MULTI_CASES,
                         ids=[f"{c[0][0]}x{c[0][1]}x{c[0][2]}-B{c[2][0]}-K{c[2][1]}-s{c[3]}" for c in HEAT_MULTI_CASES])
def test_heat_multi_step_form(hip_ctx, oracle_lib, dtype, case):
    """Up to K heat steps per launch in LDS boxes: boxes that reach one, two or
    no walls per axis, z extents below, at and above 64, odd and even step
    counts (the launches split so the state ends in the right buffer): equal
    to the oracle bit for bit."""
    from cfdarcov2_amd import capi
    n, d, (B, K), steps = case
    print(f"\n[case] heat multi-step n={n} B={B} K={K} steps={steps} {np.dtype(dtype).name}", flush=True)
    rng = np.random.default_rng(17)
    gm = capi.HexMesh(hip_ctx, *n, *d, dtype=dtype)
    gm.set_heat_shape(2, B, K)
    om = oracle_lib.OracleMesh(*n, *d, dtype=dtype)
    T0 = rng.standard_normal(om.N).astype(dtype)
    alpha = 1e-2
    dt = 0.15 * min(d) ** 2 / alpha
    a, b = _t(T0, dtype), torch.empty(om.N, dtype=_t(T0, dtype).dtype, device="cuda")
    got = _np(gm.heat_run(a, b, alpha, dt, steps))
    ref = om.heat_run(T0, alpha, dt, steps)
    assert np.array_equal(got, ref), float(np.max(np.abs(got - ref)))


def test_heat_run_step_blocks_replayed(hip_ctx, oracle_lib):
    """Runs of 32 steps or more replay a captured block of 32 steps (a HIP
    graph): repeated runs on the same fields, with eager work and
    synchronisations between them, each equal to the oracle; a new shape, dt
    or pair of fields captures again; step counts that leave a remainder
    (33, 70) and exact multiples (64)."""
    from cfdarcov2_amd import capi
    n, d = (20, 18, 24), (1.0 / 16,) * 3
    rng = np.random.default_rng(29)
    for dtype in (np.float64, np.float32):
        om = oracle_lib.OracleMesh(*n, *d, dtype=dtype)
        gm = capi.HexMesh(hip_ctx, *n, *d, dtype=dtype)
        T0 = rng.standard_normal(om.N).astype(dtype)
        alpha = 1e-2
        for shape in ((0, 8, 16), (2, 4, 2), (2, 8, 4)):
            gm.set_heat_shape(*shape)
            for dt in (0.1 * d[0] ** 2 / alpha, 0.15 * d[0] ** 2 / alpha):
                a = _t(T0, dtype)
                b = torch.empty_like(a)
                for steps in (70, 64, 33, 70):
                    print(f"\n[case] heat blocks {np.dtype(dtype).name} shape={shape} steps={steps}", flush=True)
                    a.copy_(_t(T0, dtype))
                    got = _np(gm.heat_run(a, b, alpha, dt, steps))
                    torch.randn(1 << 16, device="cuda").sum().item()
                    ref = om.heat_run(T0, alpha, dt, steps)
                    assert np.array_equal(got, ref), (np.dtype(dtype).name, shape, dt, steps)


def test_c_abi_autotuner(hip_ctx, oracle_lib):
    """cfd_mesh_autotune picks and applies a shape the mesh allows, and runs
    with it still equal the oracle."""
    from cfdarcov2_amd import capi
    rng = np.random.default_rng(31)
    n, d = (40, 36, 64), (1.0 / 64,) * 3
    gm = capi.HexMesh(hip_ctx, *n, *d, dtype=np.float32)
    om = oracle_lib.OracleMesh(*n, *d, dtype=np.float32)
    print("\n[case] C-ABI autotuner, wave", flush=True)
    best = gm.autotune("wave", steps=4)
    assert best["variant"] in (0, 1, 7, 8, 9, 11, 12) and best["us_per_step"] > 0, best
    f = [rng.standard_normal(om.N).astype(np.float32) for _ in range(4)]
    A = [_t(a, np.float32) for a in f]
    B = [torch.empty_like(A[0]) for _ in range(4)]
    got = [_np(t) for t in gm.wave_run(A, B, -1.0, 0.4 / 64, 8)]
    for name, g, r in zip("puvw", got, om.wave_run(*f, -1.0, 0.4 / 64, 8)):
        assert np.array_equal(g, r), (name, best)
    gh = capi.HexMesh(hip_ctx, 32, 32, 32, 1 / 32, 1 / 32, 1 / 32, dtype=np.float64)
    oh = oracle_lib.OracleMesh(32, 32, 32, 1 / 32, 1 / 32, 1 / 32, dtype=np.float64)
    print("\n[case] C-ABI autotuner, heat", flush=True)
    best = gh.autotune("heat", steps=6)
    assert best["variant"] in (0, 1, 2), best
    T0 = rng.standard_normal(oh.N)
    a = _t(T0, np.float64)
    b = torch.empty_like(a)
    dt = 0.1 / 32 ** 2
    assert np.array_equal(_np(gh.heat_run(a, b, 1.0, dt, 9)), oh.heat_run(T0, 1.0, dt, 9)), best
    gm.close()
    gh.close()


def test_heat_multi_step_shape_errors(hip_ctx):
    from cfdarcov2_amd import capi
    gm = capi.HexMesh(hip_ctx, 8, 8, 8, 0.125, 0.125, 0.125, dtype=np.float64)
    for shape in ((2, 6, 2), (2, 4, 3), (2, 8, 8)):
        with pytest.raises(capi.CfdError):
            gm.set_heat_shape(*shape)
