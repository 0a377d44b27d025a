"""Unstructured hex meshes on the device (cfd_mesh_create_unstructured): the
gradient, Laplacian, stabilisation, heat step, acoustic step and Rusanov step
through the connectivity form, with the
cells renumbered by reverse Cuthill-McKee, equal the oracle built from the
same connectivity bit for bit after mapping back to the caller's numbering;
on a perturbed lattice, in a shuffled numbering, and at 64^3."""
import numpy as np
import pytest

from cfdarcov2_amd import meshio, workloads
from tests.test_unstructured_cpu import shuffle

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _mesh_case(kind, dtype):
    n, d = (6, 5, 7), (1 / 8, 1 / 8, 1 / 8)
    if kind == "big":
        n, d = (64, 64, 64), (1 / 64,) * 3
    off = None if kind == "lattice" else workloads.perturbed_vertex_offsets(*n, *d, seed=3, frac=0.2)
    conn = meshio.lattice_connectivity(*n, *d, vertex_offsets=off, dtype=dtype)
    if kind == "shuffled":
        conn, _ = shuffle(conn, 11)
    return conn


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("kind,reorder", [("lattice", True), ("perturbed", False), ("perturbed", True),
                                          ("shuffled", True), ("big", True)])
def test_unstructured_operators_equal_oracle(hip_ctx, oracle_lib, dtype, kind, reorder):
    from cfdarcov2_amd import capi
    conn = _mesh_case(kind, dtype)
    print(f"\n[case] unstructured {kind} reorder={reorder} {np.dtype(dtype).name}", flush=True)
    um = capi.UnstructuredMesh(hip_ctx, conn, dtype=dtype, reorder=reorder)
    om = oracle_lib.OracleMesh.from_connectivity(conn, dtype=dtype)
    assert um.N == om.N
    if not reorder:
        assert np.array_equal(um.perm, np.arange(um.N))
    else:
        assert sorted(um.perm.tolist()) == list(range(um.N))
    rng = np.random.default_rng(5)
    phi = rng.standard_normal(um.N).astype(dtype)
    dphi = torch.from_numpy(um.to_mesh_order(phi)).cuda()
    back = lambda t: um.to_caller_order(t.cpu().numpy())
    for mask in (1, 2, 4, 7):
        for wall in (-1, 0, 2):
            assert np.array_equal(back(um.grad(dphi, mask, wall)), om.grad(phi, mask, wall)), (mask, wall)
        assert np.array_equal(back(um.lapl(dphi, mask)), om.lapl(phi, mask)), mask
        assert np.array_equal(back(um.stab(dphi, mask)), om.stab(phi, mask)), ("stab", mask)
    # the mesh's tables, in its numbering
    for t in ("vol", "alpha", "nax"):
        got = um.table(t)
        ref = om.table(t)
        assert np.array_equal(got, ref[um.perm] if ref.ndim == 1 else ref[:, um.perm]), t
    ids = um.table("ids")
    assert np.array_equal(um.perm[ids], om.table("ids")[:, um.perm])
    # heat steps: odd and even counts
    dt = 0.1 * min(1 / 8, 1 / 64 if kind == "big" else 1 / 8) ** 2 / 1e-2
    for steps in (5, 6):
        a = torch.from_numpy(um.to_mesh_order(phi)).cuda()
        b = torch.empty_like(a)
        got = back(um.heat_run(a, b, 1e-2, dt, steps))
        assert np.array_equal(got, om.heat_run(phi, 1e-2, dt, steps)), steps
    um.close()


def test_unstructured_entry_points_refuse_other_operators(hip_ctx):
    from cfdarcov2_amd import capi
    conn = meshio.lattice_connectivity(4, 4, 4, 0.25, 0.25, 0.25)
    um = capi.UnstructuredMesh(hip_ctx, conn, dtype="f32")
    phi = torch.zeros(um.N, dtype=torch.float32, device="cuda")
    out = torch.empty_like(phi)
    with pytest.raises(capi.CfdError) as e:
        um.stab_noalloc(phi, out)
    assert e.value.code == -4
    with pytest.raises(capi.CfdError) as e:
        um.set_euler_shape(1, 4, 16)
    assert e.value.code == -4
    with pytest.raises(capi.CfdError):
        um.set_wave_shape(7, 4, 8, 0)
    um.close()


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("kind,reorder", [("lattice", True), ("perturbed", True), ("shuffled", True),
                                          ("shuffled", False)])
def test_unstructured_wave_equals_oracle(hip_ctx, oracle_lib, dtype, kind, reorder):
    """The fused acoustic step on an unstructured mesh (velocities, then the
    pressure from the new velocities, reflective walls): equal to the
    oracle's operator stepper on the same connectivity, odd and even step
    counts, constant K and a K field."""
    from cfdarcov2_amd import capi
    conn = _mesh_case(kind, dtype)
    print(f"\n[case] unstructured wave {kind} reorder={reorder} {np.dtype(dtype).name}", flush=True)
    um = capi.UnstructuredMesh(hip_ctx, conn, dtype=dtype, reorder=reorder)
    om = oracle_lib.OracleMesh.from_connectivity(conn, dtype=dtype)
    rng = np.random.default_rng(8)
    f = [rng.standard_normal(um.N).astype(dtype) for _ in range(4)]
    kf = (-1.0 - rng.random(um.N)).astype(dtype)
    dt = 0.2 / 8
    for steps, K_field in ((3, None), (4, kf)):
        A = [torch.from_numpy(um.to_mesh_order(x)).cuda() for x in f]
        B = [torch.empty_like(t) for t in A]
        kd = None if K_field is None else torch.from_numpy(um.to_mesh_order(K_field)).cuda()
        got = um.wave_run(A, B, -1.3, dt, steps, K_field=kd)
        ref = om.wave_run(*[x.copy() for x in f], -1.3, dt, steps, K_field=K_field)
        for name, g_, r in zip("puvw", got, ref):
            assert np.array_equal(um.to_caller_order(g_.cpu().numpy()), r), (steps, name)
    um.close()


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("kind,reorder", [("lattice", True), ("perturbed", True), ("perturbed", False),
                                          ("shuffled", True)])
def test_unstructured_euler_equals_oracle(hip_ctx, oracle_lib, dtype, kind, reorder):
    """The Rusanov stepper on an unstructured mesh (face axis from the first
    cell's local face pair, walls mirrored, the time step from the device
    minimum): fields and every step's dt equal the oracle's on the same
    connectivity."""
    from cfdarcov2_amd import capi
    conn = _mesh_case(kind, dtype)
    print(f"\n[case] unstructured euler {kind} reorder={reorder} {np.dtype(dtype).name}", flush=True)
    um = capi.UnstructuredMesh(hip_ctx, conn, dtype=dtype, reorder=reorder)
    om = oracle_lib.OracleMesh.from_connectivity(conn, dtype=dtype)
    U0 = workloads.sod_state(om.table("ccx"), x0=0.4, dtype=dtype)
    rng = np.random.default_rng(12)
    U0[1:4] = (0.05 * rng.standard_normal((3, om.N)) * U0[0]).astype(dtype)
    steps = 7
    ref, ref_dt = om.euler_run(U0, 1.4, 0.3, steps)
    A = [torch.from_numpy(um.to_mesh_order(U0[k])).cuda() for k in range(5)]
    B = [torch.empty_like(t) for t in A]
    hist = torch.empty(steps, dtype=A[0].dtype, device="cuda")
    got = um.euler_run(A, B, 1.4, 0.3, steps, dt_hist=hist)
    assert np.array_equal(hist.cpu().numpy(), ref_dt)
    for k in range(5):
        assert np.array_equal(um.to_caller_order(got[k].cpu().numpy()), ref[k]), k
    um.close()
