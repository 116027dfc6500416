"""Writes tests/golden/oracle_goldens.npz: small FP64 fixtures from the CPU
oracle (oracle/, the C restatement of the reference's quadrature), frozen so
that a later change of either side shows up against a fixed record.

The reference itself cannot run here (Taichi is not installed), so these
vectors come from the restatement, whose functions are pinned by the closed
forms in tests/test_oracle.py.  Contents (GaussQR 5):

* two-triangle meshes whose second panel shares an edge, one vertex or
  nothing with the first, each with its vertex order rotated three ways: the
  near-field rules of every relation under every slot permutation;
* the octahedron (8 panels) and the icosahedron (20 panels);
* for each of those meshes V (P0 x P0), K (P0 x P1), K' (P1 x P0) and
  W (P1 x P1) for Laplace, Helmholtz k = 2 and Helmholtz-transmission
  k = 3, sqrt(n) = 1.5, with both quadrature presets (flags 0 and 3);
* the icosphere L1 (80 panels): V, K, K', W of Helmholtz k = 3 (flags 0) and
  the P0 x P1 mass;
* the config-1 problem on the icosahedron and on L1 (Laplace interior
  Dirichlet, g = 1/|x - (2, 0.5, 0.3)|): right-hand side 1/2 M g + K g and
  the solution of V t = rhs.

Run from the repository root:  python tests/golden/make_golden.py
"""
import importlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)
sys.path.insert(0, os.path.join(REPO, "oracle"))

import oracle  # noqa: E402

OPS = (("V", 0, 0, 0), ("K", 1, 0, 1), ("KADJ", 2, 1, 0), ("W", 3, 1, 1))
KERNELS = (("lap", 0, 0.0, 1.0), ("helm2", 1, 2.0, 1.0), ("trans3", 2, 3.0, 1.5))
FLAGS = (0, 3)
X0 = np.array([2.0, 0.5, 0.3])


def two_triangles(kind: str, rot: int):
    """Panel 0 = (0, 1, 2); panel 1 shares an edge, a vertex or nothing with
    it, its vertex order rotated rot times (orientation kept)."""
    v = [[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0]]
    if kind == "edge":
        v.append([1.0, 1.0, 0.3])
        t1 = [1, 3, 2]
    elif kind == "vertex":
        v += [[1.5, 0.2, 0.1], [1.4, 0.9, -0.2]]
        t1 = [1, 3, 4]
    else:
        v += [[2.0, 0.1, 0.2], [2.6, 0.3, 0.0], [2.1, 0.8, 0.4]]
        t1 = [3, 4, 5]
    t1 = t1[rot:] + t1[:rot]
    return np.array(v, dtype=np.float64), np.array([[0, 1, 2], t1], dtype=np.int32)


def octahedron():
    v = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]], dtype=np.float64)
    f = np.array([[0, 2, 4], [2, 1, 4], [1, 3, 4], [3, 0, 4], [2, 0, 5], [1, 2, 5], [3, 1, 5], [0, 3, 5]],
                 dtype=np.int32)
    return v, f


def operators(out: dict, name: str, v, f, kernels=KERNELS, flags=FLAGS):
    out[f"{name}_verts"] = v
    out[f"{name}_tris"] = f
    for kname, kernel, k, sn in kernels:
        for fl in flags:
            for oname, op, dt, dr in OPS:
                out[f"{name}_{kname}_f{fl}_{oname}"] = oracle.assemble(oracle.make_op(op, kernel, k, sn, 5, dt, dr, fl),
                                                                       v, f)


def config1(out: dict, name: str, v, f):
    g = 1.0 / np.linalg.norm(v - X0, axis=1)
    V = oracle.assemble(oracle.make_op(0, 0, 0.0, 1.0, 5, 0, 0, 0), v, f).real
    rhs = oracle.rhs_dirichlet(0, 0.0, 5, v, f, g.astype(complex), True, 1.0).real
    out[f"{name}_config1_rhs"] = rhs
    out[f"{name}_config1_solution"] = np.linalg.solve(V, rhs)


def main():
    oracle.build()
    meshes = importlib.import_module("taichi-bem_amd.meshes")
    out = {}
    for kind in ("edge", "vertex", "separate"):
        for rot in range(3):
            v, f = two_triangles(kind, rot)
            operators(out, f"tri_{kind}_r{rot}", v, f)
    operators(out, "octa", *octahedron())
    v0, f0 = meshes.icosphere(0)
    operators(out, "ico0", v0, f0)
    config1(out, "ico0", v0, f0)
    v1, f1 = meshes.icosphere(1)
    operators(out, "ico1", v1, f1, kernels=(("helm3", 1, 3.0, 1.0),), flags=(0,))
    out["ico1_mass_p0p1"] = oracle.mass(5, 0, 1, v1, f1)
    config1(out, "ico1", v1, f1)
    path = os.path.join(HERE, "oracle_goldens.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {len(out)} arrays to {path} ({os.path.getsize(path)} bytes)")


if __name__ == "__main__":
    main()
