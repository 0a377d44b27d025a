"""ctypes binding of the C-ABI in include/cfdarco.h (libcfdarco_hip.so).

PyTorch supplies device memory and streams only: field arrays are CUDA
(HIP) tensors whose data pointers go straight to the C-ABI, and calls are
ordered on torch's current stream.  There is no CPU fallback: if the HIP
library is missing or no GPU is visible, loading fails with an error.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcfdarco_hip.so")

F32, F64 = 0, 1
H2D, D2H, D2D = 1, 2, 3

EW = {"add": 0, "sub": 1, "mul": 2, "div": 3, "neg": 4, "muls": 5, "divs": 6, "smul": 7,
      "axpy": 8, "leap": 9, "fill": 10, "copy": 11, "adds": 12, "sadd": 13, "subs": 14, "ssub": 15, "sdiv": 16,
      "exp": 17, "abs": 18}

TABLES = {"ids": 0, "ids_bf": 1, "nx": 2, "ny": 3, "nz": 4, "nax": 5, "nay": 6, "naz": 7,
          "area": 8, "alpha": 9, "len": 10, "vol": 11, "is_bnd": 12, "ccx": 13, "ccy": 14, "ccz": 15,
          "fv_nx": 16, "fv_ny": 17, "fv_nz": 18, "fv_area": 19, "fv_vol": 20}
_N_TABLES = ("vol", "is_bnd", "ccx", "ccy", "ccz", "fv_vol")

# Every symbol declared in include/cfdarco.h (checked by the CPU test suite).
EXPORTS = [
    "cfd_init", "cfd_destroy", "cfd_last_error", "cfd_version",
    "cfd_alloc", "cfd_free", "cfd_pool_set_limit", "cfd_pool_stats",
    "cfd_memcpy", "cfd_stream_sync",
    "cfd_mesh_create_hex", "cfd_mesh_destroy", "cfd_mesh_get_info", "cfd_mesh_table", "cfd_hex_table",
    "cfd_grad", "cfd_lapl", "cfd_stab", "cfd_ewise", "cfd_cfl_min", "cfd_count_nonfinite",
    "cfd_heat_run", "cfd_wave_run", "cfd_euler_run", "cfd_mesh_set_wave_shape", "cfd_mesh_set_euler_shape",
    "cfd_mesh_set_heat_shape", "cfd_mesh_set_adjoint_shape",
    "cfd_wave_traces", "cfd_wave_adjoint", "cfd_heat_traces", "cfd_heat_adjoint", "cfd_mesh_autotune",
    "cfd_batch_create", "cfd_batch_set_form", "cfd_batch_set_shape", "cfd_batch_run", "cfd_batch_destroy",
    "cfd_program_compile", "cfd_program_compile_only", "cfd_program_launch", "cfd_program_destroy",
    "cfd_mesh_geometry", "cfd_stab_noalloc",
    "cfd_mesh_create_unstructured", "cfd_mesh_permutation", "cfd_unstructured_table",
]


class CfdError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class BatchCase(ctypes.Structure):
    _fields_ = [("mesh", ctypes.c_void_p), ("kind", ctypes.c_int), ("coef", ctypes.c_double),
                ("dt", ctypes.c_double), ("state_a", ctypes.c_void_p * 4), ("state_b", ctypes.c_void_p * 4)]


class HexDesc(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int64), ("ny", ctypes.c_int64), ("nz", ctypes.c_int64),
                ("dx", ctypes.c_double), ("dy", ctypes.c_double), ("dz", ctypes.c_double),
                ("dtype", ctypes.c_int), ("force_general", ctypes.c_int),
                ("vertex_offsets", ctypes.c_void_p)]


class UMeshDesc(ctypes.Structure):
    _fields_ = [("n_cells", ctypes.c_int64), ("n_faces", ctypes.c_int64), ("n_vertices", ctypes.c_int64),
                ("vertices", ctypes.c_void_p), ("face_verts", ctypes.c_void_p), ("face_cells", ctypes.c_void_p),
                ("cell_faces", ctypes.c_void_p), ("cell_verts", ctypes.c_void_p), ("dtype", ctypes.c_int),
                ("reorder", ctypes.c_int)]


def umesh_desc(conn, dtype, reorder):
    """cfd_umesh_desc of a connectivity dict (cfdarcov2_amd.meshio layout); the
    returned arrays must outlive the descriptor's use."""
    arrs = {k: np.ascontiguousarray(conn[k], dtype=np.float64 if k == "vertices" else np.int64)
            for k in ("vertices", "face_verts", "face_cells", "cell_faces", "cell_verts")}
    N, F, V = arrs["cell_faces"].size // 6, arrs["face_verts"].size // 4, arrs["vertices"].size // 3
    d = UMeshDesc(N, F, V, arrs["vertices"].ctypes.data, arrs["face_verts"].ctypes.data,
                  arrs["face_cells"].ctypes.data, arrs["cell_faces"].ctypes.data, arrs["cell_verts"].ctypes.data,
                  _dtype_code(dtype), int(bool(reorder)))
    return d, arrs


def unstructured_table(conn, name, dtype="f64"):
    """Reference-layout table of an unstructured mesh in the caller's numbering,
    computed on the host (no device)."""
    L = load_library()
    d, keep = umesh_desc(conn, dtype, False)
    N = d.n_cells
    n = N if name in _N_TABLES else 6 * N
    code = _dtype_code(dtype)
    dt = np.int64 if name in ("ids", "ids_bf") else (np.float32 if code == F32 else np.float64)
    out = np.empty(n, dtype=dt)
    check(L.cfd_unstructured_table(ctypes.byref(d), TABLES[name], out.ctypes.data))
    return out if n == N else out.reshape(6, N)


class MeshInfo(ctypes.Structure):
    _fields_ = [("n_cells", ctypes.c_int64), ("n_faces", ctypes.c_int64), ("n_vertices", ctypes.c_int64),
                ("nx", ctypes.c_int64), ("ny", ctypes.c_int64), ("nz", ctypes.c_int64),
                ("dtype", ctypes.c_int), ("uniform", ctypes.c_int), ("device_bytes", ctypes.c_int64)]


_lib = None


def load_library(path: str = LIB_PATH):
    """Load libcfdarco_hip.so and declare argument types (no device calls)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"HIP library not built: {path} is missing (run __graft_entry__.build())")
    L = ctypes.CDLL(path)
    vp, i64, i, d = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_double
    pp = ctypes.POINTER(ctypes.c_void_p)
    sig = {
        "cfd_init": [i, pp], "cfd_destroy": [vp], "cfd_version": [],
        "cfd_alloc": [vp, i64, pp, vp], "cfd_free": [vp, vp, vp],
        "cfd_pool_set_limit": [vp, i64], "cfd_pool_stats": [vp, vp, vp],
        "cfd_memcpy": [vp, vp, vp, i64, i, vp], "cfd_stream_sync": [vp],
        "cfd_mesh_create_hex": [vp, ctypes.POINTER(HexDesc), pp], "cfd_mesh_destroy": [vp],
        "cfd_mesh_get_info": [vp, ctypes.POINTER(MeshInfo)], "cfd_mesh_table": [vp, i, vp],
        "cfd_hex_table": [ctypes.POINTER(HexDesc), i, vp],
        "cfd_grad": [vp, vp, vp, i, i, vp], "cfd_lapl": [vp, vp, vp, i, vp], "cfd_stab": [vp, vp, vp, i, vp],
        "cfd_stab_noalloc": [vp, vp, vp, i, vp],
        "cfd_mesh_create_unstructured": [vp, ctypes.POINTER(UMeshDesc), pp],
        "cfd_mesh_permutation": [vp, vp, vp],
        "cfd_unstructured_table": [ctypes.POINTER(UMeshDesc), i, vp],
        "cfd_ewise": [vp, i, i, vp, vp, vp, d, vp, i64, vp],
        "cfd_cfl_min": [vp, i, vp, vp, vp, vp, vp, i64, d, d, d, ctypes.POINTER(ctypes.c_double), vp],
        "cfd_mesh_geometry": [vp, vp],
        "cfd_count_nonfinite": [vp, i, vp, i64, ctypes.POINTER(ctypes.c_int64), vp],
        "cfd_heat_run": [vp, vp, vp, d, vp, d, i64, vp],
        "cfd_wave_run": [vp, ctypes.c_void_p * 4, ctypes.c_void_p * 4, d, vp, d, i64, vp],
        "cfd_euler_run": [vp, ctypes.c_void_p * 5, ctypes.c_void_p * 5, d, d, i64, vp, vp],
        "cfd_wave_traces": [vp, ctypes.c_void_p * 4, ctypes.c_void_p * 4, d, vp, d, i64,
                            ctypes.POINTER(ctypes.c_int64), i, vp, vp, i64, vp],
        "cfd_wave_adjoint": [vp, ctypes.c_void_p * 4, ctypes.c_void_p * 4, d, vp, d, i64,
                             ctypes.POINTER(ctypes.c_int64), i, vp, vp, i64, vp, vp],
        "cfd_heat_traces": [vp, vp, vp, d, vp, d, i64, ctypes.POINTER(ctypes.c_int64), i, vp, vp, i64, vp],
        "cfd_heat_adjoint": [vp, vp, vp, d, vp, d, i64, ctypes.POINTER(ctypes.c_int64), i, vp, vp, i64, vp, vp],
        "cfd_mesh_autotune": [vp, i, i, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double), vp],
        "cfd_mesh_set_wave_shape": [vp, i, i, i, i],
        "cfd_mesh_set_euler_shape": [vp, i, i, i],
        "cfd_mesh_set_heat_shape": [vp, i, i, i],
        "cfd_mesh_set_adjoint_shape": [vp, i, i, i],
        "cfd_program_compile": [vp, ctypes.c_char_p, ctypes.c_char_p, pp],
        "cfd_program_compile_only": [ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(ctypes.c_int64)],
        "cfd_program_launch": [vp, ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint32), pp, vp],
        "cfd_program_destroy": [vp],
        "cfd_batch_create": [vp, ctypes.POINTER(BatchCase), i, pp],
        "cfd_batch_set_form": [vp, i],
        "cfd_batch_set_shape": [vp, i, i],
        "cfd_batch_run": [vp, i64, i, vp],
        "cfd_batch_destroy": [vp],
    }
    for name, args in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = ctypes.c_int
    L.cfd_last_error.restype = ctypes.c_char_p
    L.cfd_last_error.argtypes = []
    _lib = L
    return L


def check(rc: int):
    if rc != 0:
        raise CfdError(rc, load_library().cfd_last_error().decode(errors="replace"))


def _dtype_code(dtype) -> int:
    import torch
    if dtype in (np.float32, torch.float32, "f32", "float32"):
        return F32
    if dtype in (np.float64, torch.float64, "f64", "float64"):
        return F64
    raise ValueError(f"unsupported dtype {dtype!r}")


def hex_table(nx, ny, nz, dx, dy, dz, dtype, name, vertex_offsets=None):
    """Reference-layout mesh table computed on the host (no device needed)."""
    L = load_library()
    code = _dtype_code(dtype)
    off = None
    if vertex_offsets is not None:
        off = np.ascontiguousarray(vertex_offsets, dtype=np.float64)
    desc = HexDesc(nx, ny, nz, dx, dy, dz, code, 0, None if off is None else off.ctypes.data)
    n = nx * ny * nz
    tid = TABLES[name]
    np_dt = np.int64 if name in ("ids", "ids_bf") else (np.float32 if code == F32 else np.float64)
    count = n if name in _N_TABLES else 6 * n
    out = np.empty(count, dtype=np_dt)
    check(L.cfd_hex_table(ctypes.byref(desc), tid, out.ctypes.data))
    return out.reshape(6, n) if count == 6 * n else out


def compile_only(src: str, kernel_name: str) -> int:
    """hipRTC-compile a source on the host (no device); returns the code-object size."""
    L = load_library()
    n = ctypes.c_int64()
    check(L.cfd_program_compile_only(src.encode(), kernel_name.encode(), ctypes.byref(n)))
    return n.value


def _stream_ptr():
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _dptr(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


class Context:
    """cfd_init / cfd_destroy on one device."""

    def __init__(self, device: int = 0):
        import torch
        if not torch.cuda.is_available():
            raise RuntimeError("no GPU visible: the HIP path has no CPU fallback")
        self.lib = load_library()
        self.device = device
        h = ctypes.c_void_p()
        check(self.lib.cfd_init(device, ctypes.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            check(self.lib.cfd_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def pool_set_limit(self, nbytes: int):
        check(self.lib.cfd_pool_set_limit(self.h, int(nbytes)))

    def pool_stats(self):
        a, b = ctypes.c_int64(), ctypes.c_int64()
        check(self.lib.cfd_pool_stats(self.h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def alloc(self, nbytes: int):
        p = ctypes.c_void_p()
        check(self.lib.cfd_alloc(self.h, int(nbytes), ctypes.byref(p), _stream_ptr()))
        return p

    def free(self, p):
        check(self.lib.cfd_free(self.h, p, _stream_ptr()))

    def ewise(self, op: str, a=None, b=None, aux=None, s=0.0, out=None, n=None):
        import torch
        ref = out if out is not None else a
        code = _dtype_code(ref.dtype)
        n = ref.numel() if n is None else n
        check(self.lib.cfd_ewise(self.h, EW[op], code, _dptr(a), _dptr(b), _dptr(aux), float(s), _dptr(out),
                                 int(n), _stream_ptr()))
        return out

    def count_nonfinite(self, x) -> int:
        bad = ctypes.c_int64()
        check(self.lib.cfd_count_nonfinite(self.h, _dtype_code(x.dtype), _dptr(x), x.numel(), ctypes.byref(bad),
                                           _stream_ptr()))
        return bad.value


@dataclass
class MeshShape:
    nx: int
    ny: int
    nz: int


class HexMesh:
    """cfd_mesh_create_hex: reference geometry of Mesh3D(nx,ny,nz,dx,dy,dz)."""

    def __init__(self, ctx: Context, nx, ny, nz, dx, dy, dz, dtype="f64", vertex_offsets=None,
                 force_general=False):
        self.ctx = ctx
        self.lib = ctx.lib
        self.code = _dtype_code(dtype)
        self.shape = MeshShape(int(nx), int(ny), int(nz))
        self.spacing = (float(dx), float(dy), float(dz))
        self.N = int(nx) * int(ny) * int(nz)
        off = None
        if vertex_offsets is not None:
            off = np.ascontiguousarray(vertex_offsets, dtype=np.float64)
            assert off.size == 3 * (nx + 1) * (ny + 1) * (nz + 1)
        self._off = off
        desc = HexDesc(int(nx), int(ny), int(nz), float(dx), float(dy), float(dz), self.code, int(force_general),
                       None if off is None else off.ctypes.data)
        h = ctypes.c_void_p()
        check(self.lib.cfd_mesh_create_hex(ctx.h, ctypes.byref(desc), ctypes.byref(h)))
        self.h = h

    @property
    def torch_dtype(self):
        import torch
        return torch.float32 if self.code == F32 else torch.float64

    def info(self) -> MeshInfo:
        inf = MeshInfo()
        check(self.lib.cfd_mesh_get_info(self.h, ctypes.byref(inf)))
        return inf

    def close(self):
        if getattr(self, "h", None):
            check(self.lib.cfd_mesh_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def table(self, name):
        n = self.N
        np_dt = np.int64 if name in ("ids", "ids_bf") else (np.float32 if self.code == F32 else np.float64)
        count = n if name in _N_TABLES else 6 * n
        out = np.empty(count, dtype=np_dt)
        check(self.lib.cfd_mesh_table(self.h, TABLES[name], out.ctypes.data))
        return out.reshape(6, n) if count == 6 * n else out

    def _check_field(self, t):
        assert t.is_cuda and t.dtype == self.torch_dtype and t.is_contiguous() and t.numel() == self.N, \
            f"field must be a contiguous {self.torch_dtype} CUDA tensor of {self.N} cells"

    def grad(self, phi, axis_mask=7, wall_axis=-1, out=None):
        import torch
        self._check_field(phi)
        out = torch.empty_like(phi) if out is None else out
        check(self.lib.cfd_grad(self.h, _dptr(phi), _dptr(out), int(axis_mask), int(wall_axis), _stream_ptr()))
        return out

    def lapl(self, phi, axis_mask=7, out=None):
        import torch
        self._check_field(phi)
        out = torch.empty_like(phi) if out is None else out
        check(self.lib.cfd_lapl(self.h, _dptr(phi), _dptr(out), int(axis_mask), _stream_ptr()))
        return out

    def stab(self, phi, axis_mask=7, out=None):
        import torch
        self._check_field(phi)
        out = torch.empty_like(phi) if out is None else out
        check(self.lib.cfd_stab(self.h, _dptr(phi), _dptr(out), int(axis_mask), _stream_ptr()))
        return out

    def stab_noalloc(self, phi, out, axis_mask=7):
        """cfd_stab_noalloc: the stabilisation into a caller-given `out`, with
        the library's fixed per-device workspace instead of a pool block."""
        self._check_field(phi)
        self._check_field(out)
        check(self.lib.cfd_stab_noalloc(self.h, _dptr(phi), _dptr(out), int(axis_mask), _stream_ptr()))
        return out

    def cfl_min(self, u, p, rho, v=None, w=None, cfl=1.0, gamma=5.0 / 3.0) -> float:
        """cfl * min dl / (sqrt(gamma p / rho) + |u|), dl = min spacing (in the mesh dtype)."""
        dt = ctypes.c_double()
        dl = min(self.spacing)
        if self.code == F32:
            dl = float(np.float32(min(np.float32(s) for s in self.spacing)))
        check(self.lib.cfd_cfl_min(self.ctx.h, self.code, _dptr(u), _dptr(v), _dptr(w), _dptr(p), _dptr(rho),
                                   self.N, float(dl), float(cfl), float(gamma), ctypes.byref(dt), _stream_ptr()))
        return dt.value

    def heat_run(self, T_a, T_b, alpha, dt, nsteps, alpha_field=None):
        """Fused heat steps; returns the tensor holding the final state."""
        for t in (T_a, T_b):
            self._check_field(t)
        check(self.lib.cfd_heat_run(self.h, _dptr(T_a), _dptr(T_b), float(alpha), _dptr(alpha_field), float(dt),
                                    int(nsteps), _stream_ptr()))
        return T_a if nsteps % 2 == 0 else T_b

    def set_euler_shape(self, variant=1, tile_rows=4, chunk=16):
        check(self.lib.cfd_mesh_set_euler_shape(self.h, int(variant), int(tile_rows), int(chunk)))

    def autotune(self, kind="wave", steps=10):
        """cfd_mesh_autotune: times every kernel shape of the acoustic or heat
        step on this mesh and applies the fastest; returns the shape and its
        time per step."""
        shape = (ctypes.c_int * 4)()
        us = ctypes.c_double()
        check(self.lib.cfd_mesh_autotune(self.h, {"wave": 0, "heat": 1}[kind], int(steps), shape, ctypes.byref(us),
                                         _stream_ptr()))
        keys = ("variant", "tile_rows", "chunk", "xcd_remap") if kind == "wave" else ("variant", "waves", "chunk")
        return dict(zip(keys, list(shape)), us_per_step=us.value)

    def set_wave_shape(self, variant=1, tile_rows=16, chunk=32, xcd_remap=1):
        """xcd_remap: bit 0 XCD-contiguous tiles; bit 1 (row forms) every other
        step walks its tiles in reverse."""
        check(self.lib.cfd_mesh_set_wave_shape(self.h, int(variant), int(tile_rows), int(chunk), int(xcd_remap)))

    def set_heat_shape(self, variant=1, waves=8, chunk=16):
        check(self.lib.cfd_mesh_set_heat_shape(self.h, int(variant), int(waves), int(chunk)))

    def set_adjoint_shape(self, form=1, tile_rows=4, chunk=8):
        """Kernel form of wave_traces / wave_adjoint: 0 = one thread per cell,
        1 = row form."""
        check(self.lib.cfd_mesh_set_adjoint_shape(self.h, int(form), int(tile_rows), int(chunk)))

    def wave_run(self, state_a, state_b, K, dt, nsteps, K_field=None):
        """Fused acoustic steps on (p, u, v, w); returns the tuple holding the final state."""
        for t in list(state_a) + list(state_b):
            self._check_field(t)
        A = (ctypes.c_void_p * 4)(*[t.data_ptr() for t in state_a])
        B = (ctypes.c_void_p * 4)(*[t.data_ptr() for t in state_b])
        check(self.lib.cfd_wave_run(self.h, A, B, float(K), _dptr(K_field), float(dt), int(nsteps),
                                    _stream_ptr()))
        return tuple(state_a) if nsteps % 2 == 0 else tuple(state_b)

    def euler_run(self, state_a, state_b, gamma, cfl, nsteps, dt_hist=None):
        """Rusanov steps on (rho, rho u, rho v, rho w, E); returns the tuple
        holding the final state.  dt_hist (device tensor, nsteps long) gets
        every step's dt."""
        for t in list(state_a) + list(state_b):
            self._check_field(t)
        if dt_hist is not None:
            assert dt_hist.is_cuda and dt_hist.dtype == self.torch_dtype and dt_hist.numel() >= nsteps
        A = (ctypes.c_void_p * 5)(*[t.data_ptr() for t in state_a])
        B = (ctypes.c_void_p * 5)(*[t.data_ptr() for t in state_b])
        check(self.lib.cfd_euler_run(self.h, A, B, float(gamma), float(cfl), int(nsteps), _dptr(dt_hist),
                                     _stream_ptr()))
        return tuple(state_a) if nsteps % 2 == 0 else tuple(state_b)


class UnstructuredMesh(HexMesh):
    """cfd_mesh_create_unstructured: a hex mesh given by its connectivity (a
    cfdarcov2_amd.meshio dict).  With reorder, the cells are renumbered by
    reverse Cuthill-McKee: fields are in that numbering, `perm[new] = old`
    maps them to the caller's.  grad, lapl, stab, heat_run, wave_run,
    euler_run and table apply."""

    def __init__(self, ctx: Context, conn, dtype="f64", reorder=True):
        self.ctx = ctx
        self.lib = ctx.lib
        self.code = _dtype_code(dtype)
        d, self._arrs = umesh_desc(conn, dtype, reorder)
        self.N, self.F = int(d.n_cells), int(d.n_faces)
        self.shape = None
        h = ctypes.c_void_p()
        check(self.lib.cfd_mesh_create_unstructured(ctx.h, ctypes.byref(d), ctypes.byref(h)))
        self.h = h
        self.perm = np.empty(self.N, dtype=np.int64)
        self.face_perm = np.empty(self.F, dtype=np.int64)
        check(self.lib.cfd_mesh_permutation(self.h, self.perm.ctypes.data, self.face_perm.ctypes.data))
        self.iperm = np.empty_like(self.perm)
        self.iperm[self.perm] = np.arange(self.N)

    def to_mesh_order(self, a):
        """Caller's numbering -> the mesh's."""
        return a[self.perm]

    def to_caller_order(self, a):
        """The mesh's numbering -> the caller's."""
        return a[self.iperm]
