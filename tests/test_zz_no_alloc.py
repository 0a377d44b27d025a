"""cfd_stab_noalloc: the stabilisation with the library's fixed per-device
workspace instead of a pool block.

Checked through the C-ABI against cfd_stab (which takes its workspace from the
pool) bit for bit on general (perturbed) and uniform meshes in FP32 and FP64,
for back-to-back calls whose second input is smaller than the first (the
workspace then still holds the larger call's values), at the largest FP64
general mesh the workspace takes, and against the oracle.  A general mesh one
plane beyond that size is refused before anything runs."""
import numpy as np
import pytest

from cfdarcov2_amd import capi, workloads

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _phi(n, d, dtype, seed):
    X, Y, Z = workloads.cell_centres(*n, *d)
    rng = np.random.default_rng(seed)
    a = np.exp(-((X - 0.4) ** 2 + (Y - 0.5) ** 2 + (Z - 0.3) ** 2) / 0.05) + 0.3 * X * Y
    a = a + 0.01 * rng.standard_normal(a.size)
    return a.astype(np.float32 if dtype == "f32" else np.float64)


def _mesh(ctx, n, dtype, perturbed, force_general=False):
    d = (1.0 / 64,) * 3                     # dyadic: the lattice mesh is exactly uniform
    off = workloads.perturbed_vertex_offsets(*n, *d, seed=7, frac=0.15) if perturbed else None
    return capi.HexMesh(ctx, *n, *d, dtype=dtype, vertex_offsets=off, force_general=force_general), d, off


@pytest.mark.parametrize("dtype", ["f64", "f32"])
@pytest.mark.parametrize("perturbed", [True, False])
def test_noalloc_equals_stab_on_back_to_back_calls(hip_ctx, oracle_lib, dtype, perturbed):
    sizes = [(40, 36, 30), (12, 10, 7)]     # the second call's input is smaller
    meshes, phis, refs = [], [], []
    for k, n in enumerate(sizes):
        print(f"\n[case] stab_noalloc dtype={dtype} perturbed={perturbed} n={n}", flush=True)
        mesh, d, off = _mesh(hip_ctx, n, dtype, perturbed)
        assert bool(mesh.info().uniform) == (not perturbed)
        phi = torch.from_numpy(_phi(n, d, dtype, k)).cuda()
        meshes.append((mesh, d, off))
        phis.append(phi)
        refs.append({mask: mesh.stab(phi, mask) for mask in (7, 1, 6)})
    torch.cuda.synchronize()
    for rnd in range(2):                    # large, small, then large and small again
        outs = []
        for (mesh, _, _), phi in zip(meshes, phis):
            o = {}
            for mask in (7, 1, 6):
                o[mask] = torch.full_like(phi, float("nan"))
                mesh.stab_noalloc(phi, o[mask], mask)
            outs.append(o)
        torch.cuda.synchronize()
        for k, (o, r) in enumerate(zip(outs, refs)):
            for mask in (7, 1, 6):
                assert torch.equal(o[mask], r[mask]), (rnd, k, mask)
    for (mesh, d, off), phi, n in zip(meshes, phis, sizes):
        om = oracle_lib.OracleMesh(*n, *d, dtype=np.float32 if dtype == "f32" else np.float64, perturb=off)
        got = torch.full_like(phi, float("nan"))
        mesh.stab_noalloc(phi, got, 7)
        assert np.array_equal(got.cpu().numpy(), om.stab(phi.cpu().numpy(), 7)), n
        mesh.close()


def test_noalloc_at_the_workspace_limit_and_beyond(hip_ctx):
    # 128^3 = 2,097,152 FP64 cells fill the 48 MiB workspace exactly
    n = (128, 128, 128)
    print(f"\n[case] stab_noalloc at the limit n={n} f64 general", flush=True)
    mesh, d, _ = _mesh(hip_ctx, n, "f64", False, force_general=True)
    assert not mesh.info().uniform
    phi = torch.from_numpy(_phi(n, d, "f64", 3)).cuda()
    ref = mesh.stab(phi, 7)
    got = torch.full_like(phi, float("nan"))
    mesh.stab_noalloc(phi, got, 7)
    torch.cuda.synchronize()
    assert torch.equal(got, ref)
    mesh.close()
    del phi, ref, got
    # one x-plane more: refused with CFD_E_INVALID, output left untouched
    n = (129, 128, 128)
    print(f"[case] stab_noalloc beyond the limit n={n} f64 general", flush=True)
    mesh, d, _ = _mesh(hip_ctx, n, "f64", False, force_general=True)
    phi = torch.zeros(mesh.N, dtype=torch.float64, device="cuda")
    out = torch.full_like(phi, 5.0)
    with pytest.raises(capi.CfdError) as ei:
        mesh.stab_noalloc(phi, out, 7)
    assert ei.value.code == -1 and "workspace" in str(ei.value)
    torch.cuda.synchronize()
    assert bool(torch.all(out == 5.0))
    mesh.close()
