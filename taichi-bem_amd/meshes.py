"""Synthetic surface meshes and the OBJ reader for the BEM hot path.

The reference loads Wavefront OBJ files through pywavefront and normalises
them (src/managers/mesh_manager.py:479-494); its benchmark configurations use
generated icospheres instead.  Everything here is host-side preprocessing that
produces the two arrays the device path consumes: vertices [nv, 3] float64 and
panels [nf, 3] int32 with outward counter-clockwise orientation, so that the
reference normal (x2 - x1) x (x3 - x2) points outward
(src/managers/mesh_manager.py:121-125).
"""
from __future__ import annotations

import math
import os

import numpy as np


def icosahedron() -> tuple[np.ndarray, np.ndarray]:
    """The 12-vertex icosahedron on the unit sphere, 20 outward CCW faces."""
    p = (1.0 + math.sqrt(5.0)) / 2.0
    v = np.array(
        [[-1, p, 0], [1, p, 0], [-1, -p, 0], [1, -p, 0],
         [0, -1, p], [0, 1, p], [0, -1, -p], [0, 1, -p],
         [p, 0, -1], [p, 0, 1], [-p, 0, -1], [-p, 0, 1]],
        dtype=np.float64,
    )
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.array(
        [[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
         [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
         [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
         [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]],
        dtype=np.int64,
    )
    return v, _orient_outward(v, f)


def _orient_outward(v: np.ndarray, f: np.ndarray) -> np.ndarray:
    f = f.copy()
    x1, x2, x3 = v[f[:, 0]], v[f[:, 1]], v[f[:, 2]]
    n = np.cross(x2 - x1, x3 - x2)
    flip = np.einsum("ij,ij->i", n, (x1 + x2 + x3)) < 0
    f[flip] = f[flip][:, [0, 2, 1]]
    return f


def icosphere(level: int) -> tuple[np.ndarray, np.ndarray]:
    """Unit icosphere: each subdivision splits (a,b,c) into (a,ab,ca),
    (b,bc,ab), (c,ca,bc), (ab,bc,ca); midpoints are cached by sorted edge key
    and projected to the sphere.  Level L has 10*4^L+2 vertices and 20*4^L
    panels (L=3: 642/1280, L=5: 10242/20480)."""
    v, f = icosahedron()
    verts = [row for row in v]
    faces = [tuple(int(i) for i in row) for row in f]
    for _ in range(level):
        cache: dict[tuple[int, int], int] = {}

        def mid(a: int, b: int) -> int:
            key = (a, b) if a < b else (b, a)
            idx = cache.get(key)
            if idx is None:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                idx = len(verts) - 1
                cache[key] = idx
            return idx

        new_faces = []
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            new_faces += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = new_faces
    return np.asarray(verts, dtype=np.float64), np.asarray(faces, dtype=np.int32)


def real_sph_harm(l: int, m: int, theta: np.ndarray, phi: np.ndarray) -> np.ndarray:
    """Orthonormal real spherical harmonic Y_lm (theta polar, phi azimuth)."""
    from scipy.special import lpmv

    am = abs(m)
    norm = math.sqrt((2 * l + 1) / (4 * math.pi) * math.factorial(l - am) / math.factorial(l + am))
    p = lpmv(am, l, np.cos(theta))
    if m == 0:
        return norm * p
    if m > 0:
        return math.sqrt(2.0) * norm * p * np.cos(am * phi)
    return math.sqrt(2.0) * norm * p * np.sin(am * phi)


def perturbed_icosphere(level: int, seed: int = 0, amplitude: float = 0.05, lmax: int = 4,
                        degrees: tuple[int, ...] | None = None) -> tuple[np.ndarray, np.ndarray]:
    """Radially perturbed icosphere r = 1 + amplitude * sum a_lm Y_lm with
    a_lm ~ N(0,1)/(l+1)^2 drawn from numpy.random.default_rng(seed) in the
    order l = 0..lmax, m = -l..l (all degrees are drawn; only those in
    `degrees` are used when it is given)."""
    v, f = icosphere(level)
    rng = np.random.default_rng(seed)
    theta = np.arccos(np.clip(v[:, 2], -1.0, 1.0))
    phi = np.arctan2(v[:, 1], v[:, 0])
    r = np.ones(len(v))
    for l in range(lmax + 1):
        for m in range(-l, l + 1):
            a = rng.standard_normal() / (l + 1) ** 2
            if degrees is None or l in degrees:
                r += amplitude * a * real_sph_harm(l, m, theta, phi)
    return v * r[:, None], f


def seeded_direction(seed: int = 2) -> np.ndarray:
    """Unit plane-wave direction d = normalize(N(0, I3)) from default_rng(seed)."""
    d = np.random.default_rng(seed).standard_normal(3)
    return d / np.linalg.norm(d)


def load_obj(path: str | os.PathLike, normalise: bool = True) -> tuple[np.ndarray, np.ndarray]:
    """Minimal Wavefront OBJ reader (v / f records, polygons fanned into
    triangles, 1-based and negative indices).  With normalise=True it applies
    the reference's load_asset transform: translate by minus the vertex mean,
    then scale by 1/max|coord| (src/managers/mesh_manager.py:479-494)."""
    verts: list[list[float]] = []
    faces: list[tuple[int, int, int]] = []
    with open(path, "r", encoding="utf-8", errors="replace") as fh:
        for line in fh:
            parts = line.split()
            if not parts:
                continue
            if parts[0] == "v":
                verts.append([float(t) for t in parts[1:4]])
            elif parts[0] == "f":
                idx = []
                for tok in parts[1:]:
                    i = int(tok.split("/")[0])
                    idx.append(i - 1 if i > 0 else len(verts) + i)
                for t in range(1, len(idx) - 1):
                    faces.append((idx[0], idx[t], idx[t + 1]))
    v = np.asarray(verts, dtype=np.float64)
    f = np.asarray(faces, dtype=np.int32)
    if normalise:
        v = v - v.mean(axis=0, keepdims=True)
        v = v * (1.0 / np.abs(v).max())
    return v, f
