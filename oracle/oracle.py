"""ctypes wrapper of the CPU FP64 oracle (bem_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.  Every function
mirrors a reference function; the citations are in bem_oracle.h.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libbem_oracle.so")

OP_V, OP_K, OP_KADJ, OP_W = 0, 1, 2, 3
LAPLACE, HELMHOLTZ, HELMHOLTZ_TRANSMISSION = 0, 1, 2
SEPARATE, COINCIDE, COMMON_VERTEX, COMMON_EDGE = 0, 1, 2, 3
FLAG_ROTATE_NEAR, FLAG_SYM_COINCIDE = 1, 2


class BoOp(ctypes.Structure):
    _fields_ = [
        ("op", ctypes.c_int32), ("kernel", ctypes.c_int32),
        ("k", ctypes.c_double), ("sqrt_n", ctypes.c_double),
        ("q", ctypes.c_int32), ("deg_test", ctypes.c_int32),
        ("deg_trial", ctypes.c_int32), ("flags", ctypes.c_int32),
    ]


def build(force: bool = False) -> str:
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        build()
        _lib = ctypes.CDLL(_LIB_PATH)
    return _lib


def _dp(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))


def _mesh(verts, tris):
    v = np.ascontiguousarray(verts, dtype=np.float64)
    t = np.ascontiguousarray(tris, dtype=np.int32)
    return v, t


def max_threads() -> int:
    return int(lib().bo_max_threads())


def gauss01(q: int) -> tuple[np.ndarray, np.ndarray]:
    p = np.zeros(12)
    w = np.zeros(12)
    if lib().bo_gauss01(int(q), _dp(p), _dp(w)) != 0:
        raise NotImplementedError("Please refer N as 2~8 for gaussion interation")
    return p[:q].copy(), w[:q].copy()


def geometry(verts, tris):
    v, t = _mesh(verts, tris)
    nv, nf = len(v), len(t)
    area = np.zeros(nf)
    normal = np.zeros((nf, 3))
    va = np.zeros(nv)
    vn = np.zeros((nv, 3))
    lib().bo_geometry(nv, nf, _dp(v), _ip(t), _dp(area), _dp(normal), _dp(va), _dp(vn))
    return area, normal, va, vn


def relation(tris, i: int, j: int) -> int:
    t = np.ascontiguousarray(tris, dtype=np.int32)
    return int(lib().bo_relation(_ip(t), int(i), int(j)))


def relation_counts(tris) -> np.ndarray:
    t = np.ascontiguousarray(tris, dtype=np.int32)
    c = np.zeros(4, dtype=np.int64)
    lib().bo_relation_counts(len(t), _ip(t), c.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)))
    return c


def make_op(op: int, kernel: int, k: float = 0.0, sqrt_n: float = 1.0, q: int = 5,
            deg_test: int = 0, deg_trial: int = 0, flags: int = 0) -> BoOp:
    return BoOp(op, kernel, float(k), float(sqrt_n), int(q), int(deg_test), int(deg_trial), int(flags))


def pair(op: BoOp, verts, tris, i: int, j: int, rel: int | None = None, geom=None) -> np.ndarray:
    v, t = _mesh(verts, tris)
    area, normal = (geom or geometry(v, t))[:2]
    if rel is None:
        rel = relation(t, i, j)
    ns = 3 if op.deg_test else 1
    nt = 3 if op.deg_trial else 1
    out = np.zeros(ns * nt * 2)
    if lib().bo_pair(ctypes.byref(op), _dp(v), _ip(t), _dp(np.ascontiguousarray(area)),
                     _dp(np.ascontiguousarray(normal)), int(i), int(j), int(rel), _dp(out)) != 0:
        raise NotImplementedError("unsupported GaussQR")
    return (out[0::2] + 1j * out[1::2]).reshape(ns, nt)


def assemble(op: BoOp, verts, tris, nthreads: int = 0, geom=None) -> np.ndarray:
    """Dense operator matrix (complex128; Laplace has zero imaginary part)."""
    v, t = _mesh(verts, tris)
    area, normal = (geom or geometry(v, t))[:2]
    rows = len(v) if op.deg_test else len(t)
    cols = len(v) if op.deg_trial else len(t)
    out = np.zeros((rows, cols, 2))
    rc = lib().bo_assemble(ctypes.byref(op), len(v), len(t), _dp(v), _ip(t),
                           _dp(np.ascontiguousarray(area)), _dp(np.ascontiguousarray(normal)),
                           _dp(out), int(nthreads))
    if rc != 0:
        raise NotImplementedError("unsupported GaussQR")
    return out[..., 0] + 1j * out[..., 1]


def assemble_rows(op: BoOp, verts, tris, row_begin: int, row_end: int, nthreads: int = 0,
                  geom=None) -> np.ndarray:
    """Rows [row_begin, row_end) of a P0-test operator (CPU baseline samples)."""
    assert not op.deg_test
    v, t = _mesh(verts, tris)
    area, normal = (geom or geometry(v, t))[:2]
    cols = len(v) if op.deg_trial else len(t)
    out = np.zeros((row_end - row_begin, cols, 2))
    lib().bo_assemble_rows(ctypes.byref(op), len(v), len(t), _dp(v), _ip(t),
                           _dp(np.ascontiguousarray(area)), _dp(np.ascontiguousarray(normal)),
                           int(row_begin), int(row_end), _dp(out), int(nthreads))
    return out[..., 0] + 1j * out[..., 1]


def apply(op: BoOp, verts, tris, x: np.ndarray, conj_x: bool = True, alpha: float = 1.0,
          nthreads: int = 0, geom=None) -> np.ndarray:
    v, t = _mesh(verts, tris)
    area, normal = (geom or geometry(v, t))[:2]
    rows = len(v) if op.deg_test else len(t)
    xx = np.ascontiguousarray(np.stack([np.real(x), np.imag(x)], -1), dtype=np.float64)
    y = np.zeros((rows, 2))
    lib().bo_apply(ctypes.byref(op), len(v), len(t), _dp(v), _ip(t),
                   _dp(np.ascontiguousarray(area)), _dp(np.ascontiguousarray(normal)),
                   _dp(xx), int(bool(conj_x)), ctypes.c_double(alpha), _dp(y), int(nthreads))
    return y[:, 0] + 1j * y[:, 1]


def apply_rows(op: BoOp, verts, tris, x: np.ndarray, row_begin: int, row_end: int, conj_x: bool = True,
               alpha: float = 1.0, nthreads: int = 0, geom=None) -> np.ndarray:
    """Matrix-free product for test panels [row_begin, row_end) only (P0 test)."""
    v, t = _mesh(verts, tris)
    area, normal = (geom or geometry(v, t))[:2]
    rows = len(v) if op.deg_test else len(t)
    xx = np.ascontiguousarray(np.stack([np.real(x), np.imag(x)], -1), dtype=np.float64)
    y = np.zeros((rows, 2))
    lib().bo_apply_rows(ctypes.byref(op), len(v), len(t), _dp(v), _ip(t),
                        _dp(np.ascontiguousarray(area)), _dp(np.ascontiguousarray(normal)),
                        _dp(xx), int(bool(conj_x)), ctypes.c_double(alpha), int(row_begin), int(row_end),
                        _dp(y), int(nthreads))
    return y[:, 0] + 1j * y[:, 1]


def mass(q: int, deg_test: int, deg_trial: int, verts, tris, alpha: float = 1.0, geom=None) -> np.ndarray:
    v, t = _mesh(verts, tris)
    area = np.ascontiguousarray((geom or geometry(v, t))[0])
    rows = len(v) if deg_test else len(t)
    cols = len(v) if deg_trial else len(t)
    out = np.zeros((rows, cols, 2))
    lib().bo_mass(int(q), int(deg_test), int(deg_trial), len(v), len(t), _ip(t), _dp(area),
                  ctypes.c_double(alpha), _dp(out))
    return out[..., 0]


def rhs_dirichlet(kernel: int, k: float, q: int, verts, tris, g: np.ndarray, conj_g: bool = True,
                  s: float = 1.0, sqrt_n: float = 1.0, flags: int = 0, nthreads: int = 0,
                  geom=None) -> np.ndarray:
    v, t = _mesh(verts, tris)
    area, normal = (geom or geometry(v, t))[:2]
    op = make_op(OP_K, kernel, k, sqrt_n, q, 0, 1, flags)
    gg = np.ascontiguousarray(np.stack([np.real(g), np.imag(g)], -1), dtype=np.float64)
    rhs = np.zeros((len(t), 2))
    lib().bo_rhs_dirichlet(ctypes.byref(op), len(v), len(t), _dp(v), _ip(t),
                           _dp(np.ascontiguousarray(area)), _dp(np.ascontiguousarray(normal)),
                           _dp(gg), int(bool(conj_g)), ctypes.c_double(s), _dp(rhs), int(nthreads))
    return rhs[:, 0] + 1j * rhs[:, 1]
