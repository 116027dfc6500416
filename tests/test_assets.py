"""The reference's mesh assets (tests/golden/meshes, copied from its assets/
directory) read by the OBJ loader, and oracle identities on them: closure and
orientation of the closed meshes, the Gauss solid-angle identity of K and the
constant null vector of W on non-smooth and non-convex geometry."""
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
MESHES = os.path.join(HERE, "golden", "meshes")
COUNTS = {"cube": (8, 12), "suzan": (507, 968), "hemisphere": (650, 1296), "sphere": (642, 1280)}
CLOSED = ("cube", "hemisphere", "sphere")


def load(bem, name):
    return bem.meshes.load_obj(os.path.join(MESHES, name + ".obj"))


@pytest.mark.parametrize("name", sorted(COUNTS))
def test_asset_loads_like_the_reference(bem, name):
    """Vertex and (fanned) triangle counts; the reference's normalisation
    (mesh_manager.py:479-494): zero vertex mean, largest |coordinate| 1."""
    v, f = load(bem, name)
    assert (len(v), len(f)) == COUNTS[name]
    assert np.abs(v.mean(axis=0)).max() < 1e-13
    assert abs(np.abs(v).max() - 1.0) < 1e-15
    assert f.min() == 0 and f.max() == len(v) - 1


@pytest.mark.parametrize("name", CLOSED)
def test_closed_assets_are_closed_and_outward(bem, oracle, name):
    """sum_i A_i n_i = 0 on a closed surface, and the enclosed volume
    (1/3) sum_i A_i c_i . n_i is positive for outward normals (the oracle's
    geometry, which the GPU geometry matches)."""
    v, f = load(bem, name)
    area, normal, _, _ = oracle.geometry(v, f)
    s = (area[:, None] * normal).sum(axis=0)
    assert np.abs(s).max() < 1e-13 * area.sum()
    c = v[f].mean(axis=1)
    assert (area * (c * normal).sum(axis=1)).sum() / 3.0 > 0.1


@pytest.mark.parametrize("name,tol", [("cube", 1e-4), ("sphere", 1e-6)])
def test_solid_angle_identity_on_assets(bem, oracle, name, tol):
    """sum_v K[i, v] = -A_i / 2 with the singularity-adapted rules on a cube
    (coplanar neighbours, right-angle edges) and on the reference's sphere."""
    v, f = load(bem, name)
    geom = oracle.geometry(v, f)
    K = oracle.assemble(oracle.make_op(oracle.OP_K, oracle.LAPLACE, q=5, deg_trial=1, flags=3), v, f,
                        geom=geom).real
    assert np.abs(K.sum(1) / (-0.5 * geom[0]) - 1).max() < tol


@pytest.mark.parametrize("name", ["cube", "suzan"])
def test_hypersingular_constant_null_vector_on_assets(bem, oracle, name):
    """W 1 = 0 (the surface curl of a constant vanishes), also on the
    non-convex, open Suzanne."""
    v, f = load(bem, name)
    W = oracle.assemble(oracle.make_op(oracle.OP_W, oracle.LAPLACE, q=4, deg_test=1, deg_trial=1, flags=0), v, f)
    assert np.abs(W.real.sum(1)).max() < 1e-13 * np.abs(W).max()
