"""On-disk output in the reference's layout (include/io_operators_tmpl.hpp:44-212),
read back on the CPU (no device): tests/cpp/io_layout.cpp writes mesh.json of a
perturbed 3 x 2 x 4 mesh, a .vti and a .bin of a field that encodes each
cell's lattice position, and the run_latest link.

mesh.json is checked against the reference's construction loop restated here
(src/mesh3d.cpp:323-379: faces take their vertex quads and cells as the cells
are visited in id order), with the face and vertex ids of its closed forms
(src/mesh3d.cpp:391-416)."""
import base64
import json
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PROG = os.path.join(ROOT, "cfdarcov2_amd", "_build", "tests", "io_layout")
X, Y, Z = 3, 2, 4
D = (0.5, 0.25, 0.125)


@pytest.fixture(scope="module")
def out():
    if not os.path.exists(PROG):
        pytest.fail(f"{PROG} not built (run __graft_entry__.build())")
    with tempfile.TemporaryDirectory() as tmp:
        res = subprocess.run([PROG, tmp], capture_output=True, text=True, timeout=120)
        assert res.returncode == 0, res.stdout + res.stderr
        yield tmp


def _face_ids(x, y, z):
    nz = X * Y * (Z + 1)
    ny = X * (Y + 1) * Z
    return [(Z + 1) * Y * x + (Z + 1) * y + z, (Z + 1) * Y * x + (Z + 1) * y + z + 1,
            nz + X * (Y + 1) * z + (Y + 1) * x + y, nz + X * (Y + 1) * z + (Y + 1) * x + y + 1,
            nz + ny + (X + 1) * Z * y + (X + 1) * z + x, nz + ny + (X + 1) * Z * y + (X + 1) * z + x + 1]


def _vertex_ids(x, y, z):
    return [(x + a) * (Z + 1) * (Y + 1) + (y + b) * (Z + 1) + (z + c) for a in (0, 1) for b in (0, 1) for c in (0, 1)]


def _reference_construction():
    F = X * Y * (Z + 1) + X * (Y + 1) * Z + (X + 1) * Y * Z
    quads = [None] * F
    cells = [[] for _ in range(F)]
    nodes = []
    for x in range(X):
        for y in range(Y):
            for z in range(Z):
                node = Z * Y * x + Z * y + z
                v = _vertex_ids(x, y, z)
                f = _face_ids(x, y, z)
                quads[f[0]] = [v[0], v[2], v[6], v[4]]
                quads[f[1]] = [v[1], v[3], v[7], v[5]]
                quads[f[2]] = [v[0], v[1], v[5], v[4]]
                quads[f[3]] = [v[2], v[3], v[7], v[6]]
                quads[f[4]] = [v[1], v[3], v[2], v[0]]
                quads[f[5]] = [v[5], v[7], v[6], v[4]]
                for fid in f:
                    cells[fid].append(node)
                assert len(nodes) == node
                nodes.append({"face": f, "vertexes": v})
    return quads, cells, nodes


def test_mesh_json_matches_the_reference_construction(out):
    with open(os.path.join(out, "run_test", "mesh.json")) as f:
        m = json.load(f)
    assert (m["x"], m["y"], m["z"]) == (X, Y, Z)
    assert (m["lx"], m["ly"], m["lz"]) == (1.5, 0.5, 0.5)
    quads, cells, nodes = _reference_construction()
    assert len(m["face"]) == len(quads)
    for i, fc in enumerate(m["face"]):
        assert fc["vertexes_id"] == quads[i], i
        assert fc["nodes_id"] == cells[i], i
    assert m["nodes"] == nodes
    off = np.fromfile(os.path.join(out, "offsets.f64"), dtype=np.float64).reshape(-1, 3)
    assert off.shape[0] == (X + 1) * (Y + 1) * (Z + 1)
    assert np.any(off != 0)
    v = np.array(m["vertexes"])
    k = 0
    for x in range(X + 1):
        for y in range(Y + 1):
            for z in range(Z + 1):
                want = [np.float32(np.float32(i) * np.float32(d)) + np.float32(o)
                        for i, d, o in zip((x, y, z), D, off[k])]
                assert np.array_equal(v[k], np.array(want, dtype=np.float32).astype(np.float64)), k
                k += 1


def test_vti_holds_the_field_as_point_data(out):
    with open(os.path.join(out, "run_test", "T", "res_0.vti")) as f:
        text = f.read()
    assert 'type="ImageData"' in text and 'header_type="UInt64"' in text
    assert re.search(r'WholeExtent="0 2 0 1 0 3"', text)
    blob = re.search(r'format="binary">\s*(\S+)\s*</DataArray>', text).group(1)
    header = np.frombuffer(base64.b64decode(blob[:12]), dtype=np.uint64)[0]
    data = np.frombuffer(base64.b64decode(blob[12:]), dtype=np.float32)
    assert header == X * Y * Z * 4 and data.size == X * Y * Z
    img = data.reshape(Z, Y, X)           # x fastest
    for x in range(X):
        for y in range(Y):
            for z in range(Z):
                assert img[z, y, x] == 100 * x + 10 * y + z


def test_bin_and_run_latest(out):
    data = np.fromfile(os.path.join(out, "run_test", "T", "0.bin"), dtype=np.float32)
    want = np.array([100 * x + 10 * y + z for x in range(X) for y in range(Y) for z in range(Z)], dtype=np.float32)
    assert np.array_equal(data, want)    # the mesh's cell order, z fastest
    link = os.path.join(out, "run_latest")
    assert os.path.islink(link)
    assert os.readlink(link) == os.path.abspath(os.path.join(out, "run_test"))
