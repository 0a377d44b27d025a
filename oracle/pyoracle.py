"""ORACLE — test infrastructure only.

ctypes binding of the CPU oracle (``liboracle.so``, built from
``oracle/cfd_oracle.hpp`` by ``oracle/Makefile``).  Only tests, the smoke check
and the CPU-baseline leg of the benchmark import this module; the product path
never does.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_F32, _F64 = 0, 1


def build(force: bool = False) -> str:
    """Compile the oracle with make (the recipe sits beside the sources)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE] + (["-B"] if force else []), check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        vp, i64, d, i = ctypes.c_void_p, ctypes.c_int64, ctypes.c_double, ctypes.c_int
        L.oracle_mesh_create.restype = vp
        L.oracle_mesh_create.argtypes = [i64, i64, i64, d, d, d, i, vp]
        L.oracle_mesh_create_conn.restype = vp
        L.oracle_mesh_create_conn.argtypes = [i64, i64, i64, vp, vp, vp, vp, vp, i]
        L.oracle_mesh_destroy.argtypes = [vp]
        L.oracle_mesh_sizes.argtypes = [vp, vp]
        L.oracle_mesh_table.argtypes = [vp, ctypes.c_char_p, vp]
        L.oracle_mesh_table.restype = i
        L.oracle_grad.argtypes = [vp, vp, i, i, vp]
        L.oracle_lapl.argtypes = [vp, vp, i, vp]
        L.oracle_stab.argtypes = [vp, vp, i, vp]
        L.oracle_heat_run.argtypes = [vp, vp, d, vp, d, i64]
        L.oracle_wave_run.argtypes = [vp, vp, vp, vp, vp, d, vp, d, i64]
        L.oracle_wave_run_fused.argtypes = [vp, vp, vp, vp, vp, d, vp, d, i64]
        L.oracle_set_threads.argtypes = [i]
        L.oracle_set_threads.restype = i
        L.oracle_cfl3d.restype = d
        L.oracle_cfl3d.argtypes = [vp, d, vp, vp, vp, vp, vp]
        L.oracle_cfl1d.restype = d
        L.oracle_cfl1d.argtypes = [vp, d, vp, vp, vp]
        L.oracle_euler_run.argtypes = [vp, vp, d, d, i64, vp]
        L.oracle_euler_dt.argtypes = [vp, vp, d, d]
        L.oracle_euler_dt.restype = d
        L.oracle_d2t_leapfrogs.argtypes = [i64, i]
        L.oracle_d2t_leapfrogs.restype = i
        L.oracle_d2t_step.argtypes = [vp, vp, vp, d, i, d, i, vp]
        _lib = L
    return _lib


def set_threads(n: int) -> int:
    """OpenMP threads of the oracle's loops; returns the previous setting."""
    return int(lib().oracle_set_threads(int(n)))


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


# D2tVar rules (cfd_oracle.hpp d2t_leapfrogs): fused kernels, per-operator
# evaluation with stored history (--store), per-operator without history
D2T_FUSED, D2T_HISTORY, D2T_NO_HISTORY = 0, 1, 2


def d2t_leapfrogs(t: int, rule: int) -> bool:
    """Whether step t (0-based) of a d2t equation leapfrogs under `rule`."""
    return bool(lib().oracle_d2t_leapfrogs(int(t), int(rule)))


_INT_TABLES = {"ids", "ids_bf", "face_verts", "face_cells", "cell_faces", "cell_verts"}


class OracleMesh:
    """Reference-semantics hex mesh (Mesh3D::init_basic_internals + compute)."""

    def __init__(self, X, Y, Z, dx, dy, dz, dtype=np.float64, perturb=None):
        self.dtype = np.dtype(dtype)
        self.shape = (int(X), int(Y), int(Z))
        self.spacing = (float(dx), float(dy), float(dz))
        pert = None
        if perturb is not None:
            pert = np.ascontiguousarray(perturb, dtype=np.float64)
        code = _F32 if self.dtype == np.float32 else _F64
        self._h = lib().oracle_mesh_create(X, Y, Z, dx, dy, dz, code, _ptr(pert))
        sizes = np.zeros(3, dtype=np.int64)
        lib().oracle_mesh_sizes(self._h, _ptr(sizes))
        self.N, self.F, self.V = (int(s) for s in sizes)

    @classmethod
    def from_connectivity(cls, conn, dtype=np.float64):
        """A mesh from its connectivity: a dict (or object with the same
        attributes) holding vertices (V x 3), face_verts (F x 4), face_cells
        (F x 2, -1 when absent), cell_faces (N x 6), cell_verts (N x 8)."""
        get = (lambda k: conn[k]) if isinstance(conn, dict) else (lambda k: getattr(conn, k))
        arrs = {k: np.ascontiguousarray(get(k), dtype=np.float64 if k == "vertices" else np.int64)
                for k in ("vertices", "face_verts", "face_cells", "cell_faces", "cell_verts")}
        self = cls.__new__(cls)
        self.dtype = np.dtype(dtype)
        self.shape = None
        self.spacing = None
        code = _F32 if self.dtype == np.float32 else _F64
        N, F, V = arrs["cell_faces"].size // 6, arrs["face_verts"].size // 4, arrs["vertices"].size // 3
        self._h = lib().oracle_mesh_create_conn(N, F, V, _ptr(arrs["vertices"]), _ptr(arrs["face_verts"]),
                                                _ptr(arrs["face_cells"]), _ptr(arrs["cell_faces"]),
                                                _ptr(arrs["cell_verts"]), code)
        self.N, self.F, self.V = N, F, V
        return self

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _lib is not None:
            _lib.oracle_mesh_destroy(h)
            self._h = None

    def table(self, name: str) -> np.ndarray:
        n6 = 6 * self.N
        sizes = {"face_verts": 4 * self.F, "face_cells": 2 * self.F, "cell_faces": 6 * self.N,
                 "cell_verts": 8 * self.N, "vx": self.V, "vy": self.V, "vz": self.V,
                 "fnx": self.F, "fny": self.F, "fnz": self.F, "farea": self.F,
                 "svx": self.F, "svy": self.F, "svz": self.F, "sva": self.F, "svol": self.N,
                 "vol": self.N, "is_bnd": self.N, "ccx": self.N, "ccy": self.N, "ccz": self.N}
        n = sizes.get(name, n6)
        dt = np.int64 if name in _INT_TABLES else self.dtype
        out = np.empty(n, dtype=dt)
        rc = lib().oracle_mesh_table(self._h, name.encode(), _ptr(out))
        if rc != 0:
            raise KeyError(name)
        if n == n6 and name not in ("cell_faces",):
            return out.reshape(6, self.N)  # column-major N x 6: row fc holds face fc of every cell
        return out

    def _arr(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.shape == (self.N,), a.shape
        return a

    def grad(self, phi, mask=7, wall_axis=-1):
        phi = self._arr(phi)
        out = np.empty_like(phi)
        lib().oracle_grad(self._h, _ptr(phi), int(mask), int(wall_axis), _ptr(out))
        return out

    def lapl(self, phi, mask=7):
        phi = self._arr(phi)
        out = np.empty_like(phi)
        lib().oracle_lapl(self._h, _ptr(phi), int(mask), _ptr(out))
        return out

    def stab(self, phi, mask=7):
        phi = self._arr(phi)
        out = np.empty_like(phi)
        lib().oracle_stab(self._h, _ptr(phi), int(mask), _ptr(out))
        return out

    def heat_run(self, T, alpha, dt, steps, alpha_field=None):
        T = self._arr(T).copy()
        af = None if alpha_field is None else self._arr(alpha_field)
        lib().oracle_heat_run(self._h, _ptr(T), float(alpha), _ptr(af), float(dt), int(steps))
        return T

    def wave_run(self, p, u, v, w, K, dt, steps, K_field=None):
        p, u, v, w = (self._arr(a).copy() for a in (p, u, v, w))
        kf = None if K_field is None else self._arr(K_field)
        lib().oracle_wave_run(self._h, _ptr(p), _ptr(u), _ptr(v), _ptr(w), float(K), _ptr(kf),
                              float(dt), int(steps))
        return p, u, v, w

    def wave_run_fused(self, p, u, v, w, K, dt, steps, K_field=None):
        """wave_run's steps, one fused pass per stage (bit-identical)."""
        p, u, v, w = (self._arr(a).copy() for a in (p, u, v, w))
        kf = None if K_field is None else self._arr(K_field)
        lib().oracle_wave_run_fused(self._h, _ptr(p), _ptr(u), _ptr(v), _ptr(w), float(K), _ptr(kf),
                                    float(dt), int(steps))
        return p, u, v, w

    def cfl3d(self, CFL, u, v, w, p, rho):
        u, v, w, p, rho = (self._arr(a) for a in (u, v, w, p, rho))
        return lib().oracle_cfl3d(self._h, float(CFL), _ptr(u), _ptr(v), _ptr(w), _ptr(p), _ptr(rho))

    def cfl1d(self, CFL, u, p, rho):
        u, p, rho = (self._arr(a) for a in (u, p, rho))
        return lib().oracle_cfl1d(self._h, float(CFL), _ptr(u), _ptr(p), _ptr(rho))

    def euler_run(self, U, gamma, CFL, steps):
        """Rusanov steps of the Euler system; U is (5, N) = (rho, rho u, rho v,
        rho w, E).  Returns (U after `steps` steps, the dt of every step)."""
        U = np.ascontiguousarray(U, dtype=self.dtype).copy()
        assert U.shape == (5, self.N), U.shape
        hist = np.empty(int(steps), dtype=self.dtype)
        lib().oracle_euler_run(self._h, _ptr(U), float(gamma), float(CFL), int(steps), _ptr(hist))
        return U, hist

    def euler_dt(self, U, gamma, CFL):
        U = np.ascontiguousarray(U, dtype=self.dtype)
        assert U.shape == (5, self.N), U.shape
        return lib().oracle_euler_dt(self._h, _ptr(U), float(gamma), float(CFL))

    def d2t_step(self, u, prev, k, mask, dt, leapfrog):
        """One step of {d2t(u), '=', k * (sum of d2d<axis>(u) over mask)}:
        leapfrog ((dt*dt)*rhs + 2u) - prev, or the Euler form dt*rhs + u."""
        u = self._arr(u)
        prev = self._arr(prev) if prev is not None else np.zeros_like(u)
        out = np.empty_like(u)
        lib().oracle_d2t_step(self._h, _ptr(u), _ptr(prev), float(k), int(mask), float(dt), int(bool(leapfrog)),
                              _ptr(out))
        return out

    def d2t_run(self, u0, k, mask, dt, steps, rule, bc=None, prev0=None):
        """`steps` steps of a d2t equation under a D2tVar rule.  bc(u, t)
        (optional) is the boundary callback applied after step t; the state at
        the start of a step is the next step's "previous".  prev0 replaces the
        previous state of step 0 (for a leapfrog started from two levels)."""
        u = self._arr(u0).copy()
        prev = np.zeros_like(u) if prev0 is None else self._arr(prev0).copy()
        for t in range(int(steps)):
            leap = rule is True or (rule is not False and d2t_leapfrogs(t, rule))
            new = self.d2t_step(u, prev, k, mask, dt, leap)
            prev, u = u, new
            if bc is not None:
                u = self._arr(bc(u, t))
        return u
