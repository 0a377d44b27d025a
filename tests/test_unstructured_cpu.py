"""Unstructured hex meshes on the CPU (no device): the library's reference
tables computed from connectivity alone (cfd_unstructured_table) against the
oracle built from the same connectivity, for the lattice mesh (where they
also equal the structured tables), a perturbed one, the same mesh in a
shuffled numbering (tables must follow the renumbering exactly), and a mesh
read back from a mesh.json file."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from cfdarcov2_amd import capi, meshio

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TABLES = ["ids", "ids_bf", "nx", "ny", "nz", "nax", "nay", "naz", "area", "alpha", "len", "vol", "is_bnd",
          "ccx", "ccy", "ccz"]


def shuffle(conn, seed):
    """The same mesh in a random numbering (meshio.shuffled)."""
    return meshio.shuffled(conn, seed)


def _check_all(conn, oracle_lib, dtype):
    om = oracle_lib.OracleMesh.from_connectivity(conn, dtype=dtype)
    for t in TABLES:
        got = capi.unstructured_table(conn, t, "f32" if dtype == np.float32 else "f64")
        assert np.array_equal(got, om.table(t)), t


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_lattice_tables_equal_structured(oracle_lib, dtype):
    n, d = (5, 4, 6), (0.25, 0.5, 0.125)
    conn = meshio.lattice_connectivity(*n, *d, dtype=dtype)
    om = oracle_lib.OracleMesh(*n, *d, dtype=dtype)
    for t in TABLES:
        got = capi.unstructured_table(conn, t, "f32" if dtype == np.float32 else "f64")
        assert np.array_equal(got, om.table(t)), t


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_perturbed_and_shuffled_tables(oracle_lib, dtype):
    from cfdarcov2_amd import workloads
    n, d = (6, 5, 7), (1 / 8, 1 / 8, 1 / 8)
    off = workloads.perturbed_vertex_offsets(*n, *d, seed=9, frac=0.2)
    conn = meshio.lattice_connectivity(*n, *d, vertex_offsets=off, dtype=dtype)
    _check_all(conn, oracle_lib, dtype)
    sh, pc = shuffle(conn, 4)
    _check_all(sh, oracle_lib, dtype)
    # the shuffled mesh's tables are the original's, renumbered
    tag = "f32" if dtype == np.float32 else "f64"
    for t in ("vol", "ccx", "area", "alpha", "nax", "len"):
        a, b = capi.unstructured_table(conn, t, tag), capi.unstructured_table(sh, t, tag)
        if a.ndim == 1:
            assert np.array_equal(b[pc], a), t
        else:
            assert np.array_equal(b[:, pc], a), t
    ids_a, ids_b = capi.unstructured_table(conn, "ids", tag), capi.unstructured_table(sh, "ids", tag)
    assert np.array_equal(ids_b[:, pc], pc[ids_a])


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_oracle_stab_from_connectivity(oracle_lib, dtype):
    """The oracle's stabilisation on a mesh built from connectivity: equal to
    the structured oracle's on the lattice (whose wall rule and upwind
    extrapolation are pinned by hand), and renumbering-covariant on a
    perturbed mesh given in a shuffled numbering."""
    from cfdarcov2_amd import workloads
    n, d = (5, 4, 6), (0.25, 0.5, 0.125)
    rng = np.random.default_rng(2)
    conn = meshio.lattice_connectivity(*n, *d, dtype=dtype)
    om_c = oracle_lib.OracleMesh.from_connectivity(conn, dtype=dtype)
    om_s = oracle_lib.OracleMesh(*n, *d, dtype=dtype)
    phi = rng.standard_normal(om_s.N).astype(dtype)
    for mask in (1, 2, 4, 7):
        assert np.array_equal(om_c.stab(phi, mask), om_s.stab(phi, mask)), mask
    off = workloads.perturbed_vertex_offsets(*n, *d, seed=5, frac=0.2)
    pconn = meshio.lattice_connectivity(*n, *d, vertex_offsets=off, dtype=dtype)
    sh, pc = shuffle(pconn, 7)
    a = oracle_lib.OracleMesh.from_connectivity(pconn, dtype=dtype)
    b = oracle_lib.OracleMesh.from_connectivity(sh, dtype=dtype)
    phi_b = np.empty_like(phi)
    phi_b[pc] = phi
    for mask in (1, 7):
        assert np.array_equal(b.stab(phi_b, mask)[pc], a.stab(phi, mask)), mask


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_oracle_euler_from_connectivity(oracle_lib, dtype):
    """The oracle's Rusanov stepper on a mesh built from connectivity (face
    axis from the first cell's local face pair) equals the structured
    oracle's (face axis from the face family) on the same perturbed lattice:
    fields and every step's dt; and follows a shuffled numbering exactly."""
    from cfdarcov2_amd import workloads
    n, d = (7, 5, 6), (1 / 8, 1 / 8, 1 / 8)
    off = workloads.perturbed_vertex_offsets(*n, *d, seed=6, frac=0.2)
    om_s = oracle_lib.OracleMesh(*n, *d, dtype=dtype, perturb=off)
    conn = meshio.lattice_connectivity(*n, *d, vertex_offsets=off, dtype=dtype)
    om_c = oracle_lib.OracleMesh.from_connectivity(conn, dtype=dtype)
    U0 = workloads.sod_state(om_s.table("ccx"), dtype=dtype)
    rng = np.random.default_rng(3)
    U0[1:4] = (0.05 * rng.standard_normal((3, om_s.N)) * U0[0]).astype(dtype)
    Us, hs = om_s.euler_run(U0, 1.4, 0.3, 6)
    Uc, hc = om_c.euler_run(U0, 1.4, 0.3, 6)
    assert np.array_equal(hs, hc)
    assert np.array_equal(Us, Uc)
    sh, pc = shuffle(conn, 5)
    om_h = oracle_lib.OracleMesh.from_connectivity(sh, dtype=dtype)
    U0h = np.empty_like(U0)
    U0h[:, pc] = U0
    Uh, hh = om_h.euler_run(U0h, 1.4, 0.3, 6)
    assert np.array_equal(hh, hs)
    assert np.array_equal(Uh[:, pc], Us)


def test_mesh_json_round_trip(oracle_lib):
    """mesh.json written by the drivers' writer (tests/cpp/io_layout) reads back
    to the connectivity of the same perturbed mesh."""
    prog = os.path.join(ROOT, "cfdarcov2_amd", "_build", "tests", "io_layout")
    if not os.path.exists(prog):
        pytest.fail(f"{prog} not built (run __graft_entry__.build())")
    with tempfile.TemporaryDirectory() as tmp:
        res = subprocess.run([prog, tmp], capture_output=True, text=True, timeout=120)
        assert res.returncode == 0, res.stdout + res.stderr
        conn = meshio.read_mesh_json(os.path.join(tmp, "run_test", "mesh.json"))
        off = np.fromfile(os.path.join(tmp, "offsets.f64"), dtype=np.float64)
    assert conn["extents"] == (3, 2, 4)
    want = meshio.lattice_connectivity(3, 2, 4, 0.5, 0.25, 0.125, vertex_offsets=off, dtype=np.float32)
    for k in meshio.KEYS:
        assert np.array_equal(conn[k], want[k]), k
    _check_all(conn, oracle_lib, np.float32)


def test_bad_connectivity_is_refused():
    conn = meshio.lattice_connectivity(2, 2, 2, 0.5, 0.5, 0.5)
    bad = {k: v.copy() for k, v in conn.items()}
    bad["cell_faces"][0, 0] = bad["cell_faces"][1, 5]     # a face that does not list the cell
    with pytest.raises(capi.CfdError):
        capi.unstructured_table(bad, "vol")
    bad = {k: v.copy() for k, v in conn.items()}
    bad["face_verts"][3, 1] = 10 ** 6
    with pytest.raises(capi.CfdError):
        capi.unstructured_table(bad, "vol")
