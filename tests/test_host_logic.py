"""Host-side logic: meshes, enums, Gauss table, dtype/error mapping."""
import numpy as np
import pytest
import torch


@pytest.mark.parametrize("level", [0, 1, 2, 3])
def test_icosphere_topology(bem, level):
    v, f = bem.meshes.icosphere(level)
    assert len(v) == 10 * 4 ** level + 2 and len(f) == 20 * 4 ** level
    assert np.allclose(np.linalg.norm(v, axis=1), 1.0)
    edges = {tuple(sorted((a, b))) for t in f for a, b in ((t[0], t[1]), (t[1], t[2]), (t[2], t[0]))}
    assert len(v) - len(edges) + len(f) == 2  # Euler characteristic of the sphere
    x1, x2, x3 = v[f[:, 0]], v[f[:, 1]], v[f[:, 2]]
    n = np.cross(x2 - x1, x3 - x2)
    assert np.all(np.einsum("ij,ij->i", n, x1 + x2 + x3) > 0)  # outward (mesh_manager.py:121-125)


def test_perturbed_icosphere_seeded(bem):
    a, fa = bem.meshes.perturbed_icosphere(2, seed=0)
    b, fb = bem.meshes.perturbed_icosphere(2, seed=0)
    c, _ = bem.meshes.perturbed_icosphere(2, seed=1)
    assert np.array_equal(a, b) and np.array_equal(fa, fb)
    assert not np.allclose(a, c)
    r = np.linalg.norm(a, axis=1)
    assert 0.8 < r.min() < 1.0 < r.max() < 1.2
    d = bem.meshes.seeded_direction(2)
    assert abs(np.linalg.norm(d) - 1) < 1e-15


def test_real_spherical_harmonics_orthonormal(bem):
    g, w = np.polynomial.legendre.leggauss(40)
    th = np.arccos(g)
    ph = np.linspace(0, 2 * np.pi, 80, endpoint=False)
    T, P = np.meshgrid(th, ph, indexing="ij")
    W = np.outer(w, np.full(80, 2 * np.pi / 80))
    ys = [bem.meshes.real_sph_harm(l, m, T, P) for l in range(4) for m in range(-l, l + 1)]
    G = np.array([[np.sum(W * a * b) for b in ys] for a in ys])
    assert np.abs(G - np.eye(len(ys))).max() < 1e-12


def test_obj_loader(tmp_path, bem):
    p = tmp_path / "oct.obj"
    p.write_text("# octahedron\nv 1 0 0\nv -1 0 0\nv 0 1 0\nv 0 -1 0\nv 0 0 1\nv 0 0 -1\n"
                 "f 1 3 5\nf 3 2 5\nf 2 4 5\nf 4 1 5\nf 3 1 6\nf 2 3 6\nf 4 2 6\nf 1/1/1 4/2/2 6/3/3\n")
    v, f = bem.meshes.load_obj(p)
    assert v.shape == (6, 3) and f.shape == (8, 3)
    assert np.abs(v).max() == 1.0 and np.allclose(v.mean(0), 0)


def test_enum_values(bem):
    assert [int(e) for e in bem.KernelType] == [0, 1, 2]
    assert [int(e) for e in bem.PanelsRelation] == [0, 1, 2, 3]
    assert [int(e) for e in bem.AssembleType] == [0, 1, 2, 3, 4]
    assert [int(e) for e in bem.ScopeType] == [0, 1, 2]
    assert [int(e) for e in bem.CellFluxType] == [0, 1, 2, 3]
    assert [int(e) for e in bem.VertAttachType] == [0, 1, 2]


def test_gauss_table_and_errors(bem):
    p, w = bem.get_gaussion_integration_points_and_weights(5, np.float64)
    assert abs(w.sum() - 1.0) < 1e-15 and np.all((p > 0) & (p < 1))
    with pytest.raises(NotImplementedError):
        bem.get_gaussion_integration_points_and_weights(9)


def test_warp_unwarp(bem):
    x = torch.randn(4, 5, 2, dtype=torch.float64)
    c = bem.warp_tensor(x)
    assert c.dtype == torch.complex128 and torch.equal(bem.unwarp_tensor(c), x)
    r = torch.randn(3, 1)
    assert bem.warp_tensor(r).shape == (3,)


def test_cpu_tensors_rejected(bem):
    with pytest.raises(bem.BEMError):
        bem._lib._require_cuda(torch.zeros(3))


def test_status_mapping(bem, monkeypatch):
    class FakeLib:
        def bem_last_error(self):
            return b"msg"

    monkeypatch.setattr(bem._lib, "_lib", FakeLib())
    with pytest.raises(NotImplementedError):
        bem._lib.check(bem._lib.BEM_ERR_UNSUPPORTED)
    with pytest.raises(bem.SingularMatrixError):
        bem._lib.check(bem._lib.BEM_ERR_SINGULAR)
    with pytest.raises(RuntimeError):
        bem._lib.check(bem._lib.BEM_ERR_INVALID)


def test_bench_refuses_more_than_one_gpu():
    """bench.py --gpus N > 1 exits with status 2 and a JSON line saying the
    build runs on one GPU (before anything touches a device)."""
    import json
    import os
    import subprocess
    import sys

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    msg = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in msg["error"]
