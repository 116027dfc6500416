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


def test_lazy_conj_views_are_rejected_at_the_boundary(bem):
    """torch's conj() and neg() of a complex tensor are views over the
    unchanged storage; the library reads storage through data_ptr(), so such
    views are refused before any pointer is taken."""
    L = bem._lib
    x = torch.ones(4, dtype=torch.complex128)
    for bad in (x.conj(), x.conj().contiguous(), x.conj().imag):
        with pytest.raises(L.BEMError, match="lazy conj"):
            L._require_cuda(bad)
    with pytest.raises(L.BEMError, match="device"):
        L._require_cuda(x.conj().resolve_conj())


def _splat_manager(bem, nf_kind, transmission, panel_type, vert_type, v, f):
    """A BEMManager holding only what splat_u_from_raw_solved_to_vertices
    reads, on CPU tensors (no device call is made)."""
    from types import SimpleNamespace

    M = bem.managers.bem_manager.BEMManager
    bm = M.__new__(M)
    nv, nf = len(v), len(f)
    area = 0.5 * np.linalg.norm(np.cross(v[f[:, 1]] - v[f[:, 0]], v[f[:, 2]] - v[f[:, 1]]), axis=1)
    va = np.zeros(nv)
    np.add.at(va, f.ravel(), np.repeat(area / 3, 3))
    mm = SimpleNamespace(
        panels=torch.tensor(f, dtype=torch.int32),
        get_panel_areas=lambda: torch.tensor(area),
        get_vert_areas=lambda: torch.tensor(va),
        get_panel_types=lambda: torch.full((nf,), panel_type, dtype=torch.int32),
        get_vertice_types=lambda: torch.full((nv,), vert_type, dtype=torch.int32))
    bm._core_manager = SimpleNamespace(_mesh_manager=mm)
    bm._is_transmission = transmission
    bm.N_Neumann, bm.M_Dirichlet = nf_kind
    bm.shape_func_degree_Neumann, bm.shape_func_degree_Dirichlet = 0, 1
    bm._Dirichlet_offset_j, bm._Neumann_offset_j = 0, bm.M_Dirichlet
    n = bm.N_Neumann + bm.M_Dirichlet
    rng = np.random.default_rng(7)
    bm.raw_solved = torch.tensor(rng.standard_normal(n) + 1j * rng.standard_normal(n))
    bm.raw_analytical_solved = torch.tensor(rng.standard_normal(n) + 0j)
    bm.solved = torch.zeros(nv, dtype=torch.complex128)
    bm.analytical_solved = torch.zeros(nv, dtype=torch.complex128)
    return bm, area, va


@pytest.mark.parametrize("case", ["dirichlet_bvp", "neumann_bvp", "transmission", "both_fields"])
def test_splat_follows_reference_show_branches(bem, case):
    """BEM_manager.py:753-801: the P0 Neumann values are spread with
    area/3/vertex-area weights (per panel, per vertex, as the reference loops
    do), the P1 Dirichlet values are added as they are; transmission shows the
    Neumann field only, other problems add both."""
    CF, VA = bem.CellFluxType, bem.VertAttachType
    v, f = bem.meshes.icosphere(1)
    nv, nf = len(v), len(f)
    kind, tr, pt, vt = {
        "dirichlet_bvp": ((nf, 0), 0, CF.NEUMANN_TOBESOLVED, VA.NEUMANN_TOBESOLVED),
        "neumann_bvp": ((0, nv), 0, CF.DIRICHLET_TOBESOLVED, VA.DIRICHLET_TOBESOLVED),
        "transmission": ((nf, nv), 1, CF.BOTH_TOBESOLVED, VA.BOTH_TOBESOLVED),
        "both_fields": ((nf, nv), 0, CF.BOTH_TOBESOLVED, VA.BOTH_TOBESOLVED),
    }[case]
    bm, area, va = _splat_manager(bem, kind, tr, int(pt), int(vt), v, f)
    bm.splat_u_from_raw_solved_to_vertices()
    raw = bm.raw_solved.numpy()
    ref = np.zeros(nv, dtype=complex)
    if bm.N_Neumann:
        for i in range(nf):
            for ii in range(3):
                ref[f[i, ii]] += raw[bm.M_Dirichlet + i] * area[i] / 3.0 / va[f[i, ii]]
    if bm.M_Dirichlet and not tr:
        ref += raw[:nv]
    assert np.abs(bm.solved.numpy() - ref).max() < 1e-13
    # a constant P0 field stays constant after the area-weighted spread
    if case == "dirichlet_bvp":
        bm.raw_analytical_solved[:] = 2.5
        bm.splat_u_from_raw_solved_to_vertices()
        assert np.abs(bm.analytical_solved.numpy() - 2.5).max() < 1e-13
