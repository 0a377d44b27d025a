"""Hex-mesh connectivity for unstructured meshes (cfd_mesh_create_unstructured).

The layout is the reference's mesh.json (include/io_operators_tmpl.hpp:161-212,
written by the drivers with -s -st and read back by the -m readers,
include/io_operators_tmpl.hpp:215-247, legacy/io_operators.cpp:170-205):
vertex coordinates, each face's four vertices and its cells, each cell's six
faces in the local order z-, z+, y-, y+, x-, x+ and its eight vertices.

`lattice_connectivity` builds it for Mesh3D(x, y, z, dx, dy, dz) in the
reference numbering (src/mesh3d.cpp:323-416), `shuffled` renumbers a mesh at
random, `read_mesh_json` reads a mesh.json file.
"""
from __future__ import annotations

import json

import numpy as np

KEYS = ("vertices", "face_verts", "face_cells", "cell_faces", "cell_verts")


def lattice_connectivity(nx, ny, nz, dx, dy, dz, vertex_offsets=None, dtype=np.float32) -> dict:
    """Connectivity of the lattice mesh.  Vertex coordinates are x * dx (plus
    the optional V x 3 offsets) evaluated in `dtype`, the way the structured
    mesh computes them, and returned as float64 (exact)."""
    X, Y, Z = int(nx), int(ny), int(nz)
    T = np.dtype(dtype).type
    ix, iy, iz = np.meshgrid(np.arange(X + 1), np.arange(Y + 1), np.arange(Z + 1), indexing="ij")
    v = np.stack([ix.ravel().astype(T) * T(dx), iy.ravel().astype(T) * T(dy), iz.ravel().astype(T) * T(dz)], axis=1)
    if vertex_offsets is not None:
        v = v + np.asarray(vertex_offsets, dtype=np.float64).reshape(-1, 3).astype(T)
    V = lambda x, y, z: (x * (Y + 1) + y) * (Z + 1) + z
    cx, cy, cz = (a.ravel() for a in np.meshgrid(np.arange(X), np.arange(Y), np.arange(Z), indexing="ij"))
    cell_verts = np.stack([V(cx + a, cy + b, cz + c) for a in (0, 1) for b in (0, 1) for c in (0, 1)], axis=1)
    nzf, nyf = X * Y * (Z + 1), X * (Y + 1) * Z
    cell_faces = np.stack([(Z + 1) * Y * cx + (Z + 1) * cy + cz, (Z + 1) * Y * cx + (Z + 1) * cy + cz + 1,
                           nzf + X * (Y + 1) * cz + (Y + 1) * cx + cy, nzf + X * (Y + 1) * cz + (Y + 1) * cx + cy + 1,
                           nzf + nyf + (X + 1) * Z * cy + (X + 1) * cz + cx,
                           nzf + nyf + (X + 1) * Z * cy + (X + 1) * cz + cx + 1], axis=1)
    F = nzf + nyf + (X + 1) * Y * Z
    face_verts = np.zeros((F, 4), dtype=np.int64)
    face_cells = np.full((F, 2), -1, dtype=np.int64)
    quad = ((0, 2, 6, 4), (1, 3, 7, 5), (0, 1, 5, 4), (2, 3, 7, 6), (1, 3, 2, 0), (5, 7, 6, 4))
    cells = np.arange(X * Y * Z)
    for fc in range(6):
        face_verts[cell_faces[:, fc]] = cell_verts[:, list(quad[fc])]
    # cells in id order: the lower cell of a face (fc odd side) comes first
    for fc in (1, 3, 5):
        face_cells[cell_faces[:, fc], 0] = cells
    for fc in (0, 2, 4):
        f = cell_faces[:, fc]
        first = face_cells[f, 0] >= 0
        face_cells[f[first], 1] = cells[first]
        face_cells[f[~first], 0] = cells[~first]
    return {"vertices": v.astype(np.float64), "face_verts": face_verts, "face_cells": face_cells,
            "cell_faces": cell_faces.astype(np.int64), "cell_verts": cell_verts.astype(np.int64)}


def shuffled(conn, seed) -> tuple[dict, np.ndarray]:
    """The same mesh with cells, faces and vertices renumbered at random (each
    face keeps its vertex order and each cell its local face order).
    Returns the new connectivity and the cell map old -> new."""
    rng = np.random.default_rng(seed)
    N, F, V = len(conn["cell_faces"]), len(conn["face_verts"]), len(conn["vertices"])
    pc, pf, pv = rng.permutation(N), rng.permutation(F), rng.permutation(V)   # old -> new
    out = {k: np.empty_like(conn[k]) for k in KEYS}
    out["vertices"][pv] = conn["vertices"]
    out["face_verts"][pf] = pv[conn["face_verts"]]
    fcells = conn["face_cells"]
    out["face_cells"][pf] = np.where(fcells >= 0, pc[np.maximum(fcells, 0)], -1)
    out["cell_faces"][pc] = pf[conn["cell_faces"]]
    out["cell_verts"][pc] = pv[conn["cell_verts"]]
    return out, pc


def read_mesh_json(path) -> dict:
    """Connectivity from a mesh.json file (keys "vertexes", "face" with
    "vertexes_id" and "nodes_id", "nodes" with "vertexes" and "face")."""
    with open(path) as f:
        m = json.load(f)
    faces = m["face"]
    face_cells = np.full((len(faces), 2), -1, dtype=np.int64)
    for i, fc in enumerate(faces):
        ids = fc["nodes_id"]
        if not 1 <= len(ids) <= 2:
            raise ValueError(f"face {i} has {len(ids)} cells")
        face_cells[i, :len(ids)] = ids
    return {"vertices": np.asarray(m["vertexes"], dtype=np.float64).reshape(-1, 3),
            "face_verts": np.asarray([fc["vertexes_id"] for fc in faces], dtype=np.int64).reshape(-1, 4),
            "face_cells": face_cells,
            "cell_faces": np.asarray([n["face"] for n in m["nodes"]], dtype=np.int64).reshape(-1, 6),
            "cell_verts": np.asarray([n["vertexes"] for n in m["nodes"]], dtype=np.int64).reshape(-1, 8),
            "extents": tuple(m.get(k) for k in ("x", "y", "z"))}
