"""Pins the CPU FP64 oracle against what the mathematics fixes, independent
of the oracle's own formulas: Gauss-Legendre nodes, icosphere relation
counts, exact mass integrals, the Gauss solid-angle identity, the curl of a
constant, sphere eigenvalues of V/K/W, and an independent closed-form
triangle potential for the singular (Sauter-Schwab) rules."""
import math
import os

import numpy as np
import pytest
from scipy import special


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


# ---------------------------------------------------------------- Gauss rule
@pytest.mark.parametrize("q", [1, 2, 3, 4, 5, 6, 7, 8, 12])
def test_gauss_matches_legendre_roots(oracle, q):
    p, w = oracle.gauss01(q)
    xp, xw = np.polynomial.legendre.leggauss(q)
    assert np.abs(p - (xp + 1) / 2).max() < 1e-15
    assert np.abs(w - xw / 2).max() < 1e-15
    # exact for monomials of degree 2q-1 on [0, 1]
    for d in range(2 * q):
        assert abs(np.dot(w, p ** d) - 1.0 / (d + 1)) < 1e-14


@pytest.mark.parametrize("q", [0, 9, 10, 11, 13])
def test_gauss_unsupported(oracle, q):
    with pytest.raises(NotImplementedError):
        oracle.gauss01(q)


# ---------------------------------------------------------------- relations
@pytest.mark.parametrize("level", [0, 1, 2])
def test_relation_counts_icosphere(bem, oracle, level):
    v, f = bem.meshes.icosphere(level)
    c = oracle.relation_counts(f)
    nf = len(f)
    deg = np.bincount(f.ravel(), minlength=len(v))
    assert c[1] == nf                      # coincident
    assert c[3] == 3 * nf                  # each panel has 3 edge neighbours
    assert c[2] == int(np.sum(deg * (deg - 3)))  # vertex-only neighbours
    assert c.sum() == nf * nf


def test_relation_two_triangle_fixtures(oracle):
    t = np.array([[0, 1, 2], [0, 1, 2], [1, 0, 3], [0, 1, 3], [2, 4, 5], [6, 7, 8], [4, 2, 0]], dtype=np.int32)
    assert oracle.relation(t, 0, 0) == oracle.COINCIDE
    assert oracle.relation(t, 0, 2) == oracle.COMMON_EDGE   # reversed shared edge
    assert oracle.relation(t, 0, 3) == oracle.COMMON_EDGE   # same-direction shared edge
    assert oracle.relation(t, 0, 4) == oracle.COMMON_VERTEX
    assert oracle.relation(t, 0, 5) == oracle.SEPARATE
    assert oracle.relation(t, 0, 6) == oracle.COMMON_EDGE   # edge (2,0) vs (0,4)->(4,2)->(2,0)? shares 0 and 2
    assert oracle.relation(t, 0, 1) == oracle.COMMON_EDGE   # identical vertex lists, different panels


# ---------------------------------------------------------------- geometry / mass
def test_geometry_closed_forms(bem, oracle):
    v = np.array([[0, 0, 0], [2, 0, 0], [0, 3, 0], [0, 0, 1]], dtype=float)
    f = np.array([[0, 1, 2], [0, 3, 1]], dtype=np.int32)
    area, normal, va, vn = oracle.geometry(v, f)
    assert np.allclose(area, [3.0, 1.0])
    assert np.allclose(normal[0], [0, 0, 1]) and np.allclose(normal[1], [0, 1, 0])
    assert np.allclose(va, [4 / 3, 4 / 3, 1.0, 1 / 3])


def test_mass_exact(bem, oracle):
    v, f = bem.meshes.perturbed_icosphere(1, seed=3)
    area = oracle.geometry(v, f)[0]
    M01 = oracle.mass(5, 0, 1, v, f)
    for i in range(len(f)):
        assert np.allclose(M01[i, f[i]], area[i] / 3, rtol=0, atol=1e-15)
    M11 = oracle.mass(5, 1, 1, v, f)
    ref = np.zeros_like(M11)
    for i, t in enumerate(f):
        for a in range(3):
            for b in range(3):
                ref[t[a], t[b]] += area[i] * (2.0 if a == b else 1.0) / 12.0
    assert rel(M11, ref) < 1e-14
    M00 = oracle.mass(5, 0, 0, v, f)
    assert rel(np.diag(M00), area) < 1e-15


# ---------------------------------------------------------------- identities
@pytest.fixture(scope="module")
def l2(bem, oracle):
    v, f = bem.meshes.icosphere(2)
    return v, f, oracle.geometry(v, f)


def test_double_layer_gauss_identity(oracle, l2):
    """sum_v K[i, v] = int_{T_i} int_S dG/dn_y = -A_i/2 (solid angle)."""
    v, f, geom = l2
    K = oracle.assemble(oracle.make_op(oracle.OP_K, oracle.LAPLACE, q=5, deg_trial=1), v, f, geom=geom).real
    area = geom[0]
    assert np.abs(K.sum(1) / (-0.5 * area) - 1).max() < 0.02
    # with the singularity-adapted rules the identity holds much better
    Kc = oracle.assemble(oracle.make_op(oracle.OP_K, oracle.LAPLACE, q=5, deg_trial=1, flags=3), v, f,
                         geom=geom).real
    assert np.abs(Kc.sum(1) / (-0.5 * area) - 1).max() < 2e-3


def test_adjoint_double_layer_is_transpose(oracle, l2):
    """K' (P1 test x P0 trial) equals K^T for regular pairs (same integrand
    with x and y exchanged); coincident pairs are zero in both."""
    v, f, geom = l2
    for kern, k in [(oracle.LAPLACE, 0.0), (oracle.HELMHOLTZ, 2.0)]:
        K = oracle.assemble(oracle.make_op(oracle.OP_K, kern, k, q=4, deg_trial=1, flags=3), v, f, geom=geom)
        Ka = oracle.assemble(oracle.make_op(oracle.OP_KADJ, kern, k, q=4, deg_test=1, flags=3), v, f, geom=geom)
        assert rel(Ka, K.T) < 5e-3


def test_hypersingular_constant_nullspace(oracle, l2):
    v, f, geom = l2
    W = oracle.assemble(oracle.make_op(oracle.OP_W, oracle.LAPLACE, q=4, deg_test=1, deg_trial=1), v, f,
                        geom=geom).real
    assert np.abs(W.sum(1)).max() < 1e-13 * np.abs(W).max()
    assert np.abs(W - W.T).max() < 0.05 * np.abs(W).max()


def _ylm_vertex(v, l, m):
    theta = np.arccos(np.clip(v[:, 2], -1, 1))
    phi = np.arctan2(v[:, 1], v[:, 0])
    from importlib import import_module
    return import_module("taichi-bem_amd").meshes.real_sph_harm(l, m, theta, phi)


def test_sphere_eigenvalues_laplace(bem, oracle):
    """Unit sphere: V Y_l = Y_l/(2l+1), K Y_l = -Y_l/(2(2l+1)),
    W Y_l = l(l+1)/(2l+1) Y_l (Rayleigh quotients of the Galerkin matrices
    with P1 interpolants, corrected near-field rules; O(h^2) at L3)."""
    v, f = bem.meshes.icosphere(3)
    geom = oracle.geometry(v, f)
    M = oracle.mass(4, 1, 1, v, f, geom=geom)
    V = oracle.assemble(oracle.make_op(oracle.OP_V, oracle.LAPLACE, q=4, deg_test=1, deg_trial=1, flags=3),
                        v, f, geom=geom).real
    K = oracle.assemble(oracle.make_op(oracle.OP_KADJ, oracle.LAPLACE, q=4, deg_test=1, deg_trial=1, flags=3),
                        v, f, geom=geom).real
    W = oracle.assemble(oracle.make_op(oracle.OP_W, oracle.LAPLACE, q=4, deg_test=1, deg_trial=1, flags=3),
                        v, f, geom=geom).real
    for l in (1, 2, 3):
        y = _ylm_vertex(v, l, 1)
        den = y @ M @ y
        assert abs((y @ V @ y) / den - 1 / (2 * l + 1)) < 0.03 / (2 * l + 1)
        assert abs((y @ K.T @ y) / den + 1 / (2 * (2 * l + 1))) < 0.05 / (2 * l + 1)
        assert abs((y @ W @ y) / den - l * (l + 1) / (2 * l + 1)) < 0.05 * l * (l + 1) / (2 * l + 1)


def test_sphere_eigenvalues_helmholtz_single_layer(bem, oracle):
    """V Y_l = i k j_l(k) h_l(k) Y_l on the unit sphere (k = 2)."""
    v, f = bem.meshes.icosphere(3)
    geom = oracle.geometry(v, f)
    k = 2.0
    M = oracle.mass(4, 1, 1, v, f, geom=geom)
    V = oracle.assemble(oracle.make_op(oracle.OP_V, oracle.HELMHOLTZ, k, q=4, deg_test=1, deg_trial=1, flags=3),
                        v, f, geom=geom)
    for l in (0, 1, 2):
        y = _ylm_vertex(v, l, 0)
        lam = (y @ V @ y) / (y @ M @ y)
        h = special.spherical_jn(l, k) + 1j * special.spherical_yn(l, k)
        exact = 1j * k * special.spherical_jn(l, k) * h
        assert abs(lam - exact) < 0.03 * abs(exact)


def _helmholtz_sphere_eigs(l, kap):
    """Eigenvalues of V, K (= K' on the sphere, where (x - y).n_y =
    (y - x).n_x = x.y - 1), and W on the unit sphere for G = e^{i kap r}/(4 pi r)
    and outward normals, from the interior and exterior single-layer
    potentials i kap j_l(kap |x|) h_l(kap) Y_l and their jump relations:
        V: i kap j_l h_l,   K: 1/2 + i kap^2 j_l h_l',   W: -i kap^3 j_l' h_l'
    (h_l = h_l^(1), W = -d/dn of the double-layer potential)."""
    jl, jlp = special.spherical_jn(l, kap), special.spherical_jn(l, kap, derivative=True)
    hl = jl + 1j * special.spherical_yn(l, kap)
    hlp = jlp + 1j * special.spherical_yn(l, kap, derivative=True)
    return 1j * kap * jl * hl, 0.5 + 1j * kap ** 2 * jl * hlp, -1j * kap ** 3 * jlp * hlp


@pytest.mark.parametrize("l", [0, 1, 2, 3])
def test_helmholtz_sphere_eigenvalues_reduce_to_laplace(l):
    """The closed forms used below tend to the Laplace eigenvalues 1/(2l+1),
    -1/(2(2l+1)) and l(l+1)/(2l+1) as kap -> 0."""
    v, kk, w = _helmholtz_sphere_eigs(l, 1e-6)
    assert abs(v - 1 / (2 * l + 1)) < 1e-5
    assert abs(kk + 1 / (2 * (2 * l + 1))) < 1e-5
    assert abs(w - l * (l + 1) / (2 * l + 1)) < 1e-5


@pytest.fixture(scope="module")
def l3_helmholtz(bem, oracle):
    """Unit icosphere L3, Helmholtz k = 2 with the adapted near-field rules:
    V (P0 x P0), K (P0 x P1), K' (P1 x P0), W (P1 x P1) and the mass blocks."""
    v, f = bem.meshes.icosphere(3)
    geom = oracle.geometry(v, f)
    k = 2.0
    mk = lambda op, dt, dr: oracle.make_op(op, oracle.HELMHOLTZ, k, q=4, deg_test=dt, deg_trial=dr, flags=3)
    ops = {name: oracle.assemble(mk(op, dt, dr), v, f, geom=geom)
           for name, (op, dt, dr) in {"V": (oracle.OP_V, 0, 0), "K": (oracle.OP_K, 0, 1),
                                      "KADJ": (oracle.OP_KADJ, 1, 0), "W": (oracle.OP_W, 1, 1)}.items()}
    mass = {(dt, dr): oracle.mass(4, dt, dr, v, f, geom=geom) for dt in (0, 1) for dr in (0, 1)}
    return v, f, k, ops, mass


@pytest.mark.parametrize("l", [0, 1, 2, 3])
def test_sphere_eigenvalues_helmholtz_all_operators(l, l3_helmholtz):
    """Generalised Rayleigh quotients <psi, Op phi> / <psi, phi> with psi, phi
    the P0 (centroid) or P1 (vertex) interpolants of Y_l against the closed
    forms above: they pin the Helmholtz dG/dn_y factor (1 - i kap r) e^{i kap r}
    of K and K' and the k^2 (n_x . n_y) term of W.  Discretisation error is
    O(h^2) (below 1 % at L3); the W bar is taken on the size of its two terms,
    which nearly cancel at l = 1."""
    v, f, k, ops, mass = l3_helmholtz
    c = v[f].mean(1)
    c = c / np.linalg.norm(c, axis=1, keepdims=True)
    y1 = _ylm_vertex(v, l, 0 if l == 0 else 1)
    y0 = _ylm_vertex(c, l, 0 if l == 0 else 1)
    ev, ek, ew = _helmholtz_sphere_eigs(l, k)
    q = lambda a, M, b, B: (a @ M @ b) / (a @ B @ b)
    assert abs(q(y0, ops["V"], y0, mass[0, 0]) - ev) < 0.02 * abs(ev)
    assert abs(q(y0, ops["K"], y1, mass[0, 1]) - ek) < 0.02 * abs(ek)
    assert abs(q(y1, ops["KADJ"], y0, mass[1, 0]) - ek) < 0.02 * abs(ek)
    scale_w = l * (l + 1) / (2 * l + 1) + k * k * abs(ev)
    assert abs(q(y1, ops["W"], y1, mass[1, 1]) - ew) < 0.01 * scale_w


@pytest.mark.parametrize("flags", [0, 3])
def test_transmission_kernel_is_helmholtz_at_scaled_wavenumber(bem, oracle, flags):
    """The transmission kernel at (k, sqrt_n) is the Helmholtz kernel at
    kap = k sqrt_n in G, dG/dn_y and W's k^2 n term (BEM_manager.py:452-525,
    hypersingular_layer_3d.py:171): all four operators agree to rounding."""
    v, f = bem.meshes.perturbed_icosphere(1, seed=5)
    geom = oracle.geometry(v, f)
    k, sn = 2.0, 1.5
    for op, dt, dr in ((oracle.OP_V, 0, 0), (oracle.OP_K, 0, 1), (oracle.OP_KADJ, 1, 0), (oracle.OP_W, 1, 1)):
        a = oracle.assemble(oracle.make_op(op, oracle.HELMHOLTZ_TRANSMISSION, k, sn, 5, dt, dr, flags), v, f,
                            geom=geom)
        b = oracle.assemble(oracle.make_op(op, oracle.HELMHOLTZ, k * sn, 1.0, 5, dt, dr, flags), v, f, geom=geom)
        c = oracle.assemble(oracle.make_op(op, oracle.HELMHOLTZ, k, 1.0, 5, dt, dr, flags), v, f, geom=geom)
        assert rel(a, b) < 1e-13, op
        assert rel(a, c) > 1e-2, op


def test_helmholtz_interior_dirichlet_converges_to_analytic_trace(bem, oracle):
    """Interior Helmholtz Dirichlet problem (k = 3) with a point source
    outside the unit sphere, x0 = (2, 0.5, 0.3): u = e^{ik|x-x0|}/(4 pi |x-x0|)
    solves it, so V t = (M/2 + K) g must converge to t = du/dn at the panel
    centroids.  The data are passed conjugated, since the reference's right-
    hand side conjugates them (rhs_constructor_3d.py:361-372).  Adapted
    near-field rules, O(h^2) from L1 to L3."""
    k = 3.0
    x0 = np.array([2.0, 0.5, 0.3])
    errs = []
    for L in (1, 2, 3):
        v, f = bem.meshes.icosphere(L)
        geom = oracle.geometry(v, f)
        V = oracle.assemble(oracle.make_op(oracle.OP_V, oracle.HELMHOLTZ, k, q=5, flags=3), v, f, geom=geom)
        r = np.linalg.norm(v - x0, axis=1)
        u = np.exp(1j * k * r) / (4 * math.pi * r)
        rhs = oracle.rhs_dirichlet(oracle.HELMHOLTZ, k, 5, v, f, np.conj(u), True, 1.0, flags=3, geom=geom)
        t = np.linalg.solve(V, rhs)
        c = v[f].mean(1)
        d = c - x0
        rc = np.linalg.norm(d, axis=1)
        texact = (1j * k * rc - 1) * np.exp(1j * k * rc) / (4 * math.pi * rc ** 3) * np.einsum("ij,ij->i", d,
                                                                                                  geom[1])
        errs.append(rel(t, texact))
    assert errs[1] < errs[0] / 2.5 and errs[2] < errs[1] / 2.5, errs
    assert errs[2] < 0.06, errs


# --------------------------------------------- singular rules vs closed form
def _triangle_potential(x, p):
    """int_T 1/|x - y| dy for a flat triangle p[3,3] and any point x
    (closed form of the potential of a uniform triangle: edge log terms plus
    the solid-angle term of the out-of-plane distance)."""
    n = np.cross(p[1] - p[0], p[2] - p[0])
    n /= np.linalg.norm(n)
    w = np.dot(x - p[0], n)
    xp = x - w * n
    total = 0.0
    for i in range(3):
        a, b = p[i], p[(i + 1) % 3]
        L = np.linalg.norm(b - a)
        s = (b - a) / L
        m = np.cross(s, n)       # outward in-plane edge normal for CCW p
        t0 = np.dot(a - xp, m)   # signed distance of the edge line
        sm, sp = np.dot(a - xp, s), np.dot(b - xp, s)
        Rm, Rp = np.linalg.norm(x - a), np.linalg.norm(x - b)
        R0 = math.sqrt(t0 * t0 + w * w)
        if R0 > 1e-14:
            total += t0 * math.log((Rp + sp) / (Rm + sm)) if (Rm + sm) > 1e-300 else 0.0
        if abs(w) > 1e-14:
            total -= abs(w) * (math.atan2(t0 * sp, R0 * R0 + abs(w) * Rp) - math.atan2(t0 * sm, R0 * R0 + abs(w) * Rm))
    return total


def _outer(p, fn, order=40):
    """int_T fn(x) dx with a collapsed Gauss rule of `order`^2 points."""
    g, w = np.polynomial.legendre.leggauss(order)
    g = (g + 1) / 2
    w = w / 2
    area2 = np.linalg.norm(np.cross(p[1] - p[0], p[2] - p[0]))
    s = 0.0
    for a in range(order):
        for b in range(order):
            r1, r2 = g[a], g[b] * g[a]
            x = (1 - r1) * p[0] + (r1 - r2) * p[1] + r2 * p[2]
            s += w[a] * w[b] * r1 * area2 * fn(x)
    return s


def _pair_mesh():
    v = np.array([[0.0, 0, 0], [1.0, 0, 0], [0.2, 0.9, 0], [0.6, -0.8, 0.3], [-0.9, -0.2, 0.4]])
    f = np.array([[0, 1, 2], [1, 0, 3], [0, 4, 3]], dtype=np.int32)
    return v, f


@pytest.mark.parametrize("j,rel_expected", [(0, 1), (1, 3), (2, 2)])
def test_singular_rules_match_closed_form(oracle, j, rel_expected):
    """Laplace V on coincident / edge / vertex pairs with the adapted rules
    (rotation + symmetric coincident) at GaussQR 12 against the closed-form
    inner potential integrated over the test panel."""
    v, f = _pair_mesh()
    assert oracle.relation(f, 0, j) == rel_expected
    op = oracle.make_op(oracle.OP_V, oracle.LAPLACE, q=12, flags=3)
    val = oracle.pair(op, v, f, 0, j)[0, 0].real
    py = v[f[j]]
    ref = _outer(v[f[0]], lambda x: _triangle_potential(x, py)) / (4 * math.pi)
    assert abs(val - ref) < 2e-6 * abs(ref)


def test_reference_rules_equal_adapted_when_order_matches(oracle):
    """With the shared vertex already at slot 0 of both panels the
    reference's stored-order vertex rule is the adapted rule."""
    v, f = _pair_mesh()
    op0 = oracle.make_op(oracle.OP_V, oracle.HELMHOLTZ, 3.0, q=5, flags=0)
    op1 = oracle.make_op(oracle.OP_V, oracle.HELMHOLTZ, 3.0, q=5, flags=1)
    a = oracle.pair(op0, v, f, 0, 2)
    b = oracle.pair(op1, v, f, 0, 2)
    assert abs(a - b).max() == 0.0
    # but not for the edge pair, whose shared edge runs 0->1 vs 1->0
    c = oracle.pair(op0, v, f, 0, 1)
    d = oracle.pair(op1, v, f, 0, 1)
    assert abs(c - d).max() > 1e-6 * abs(d).max()


def test_helmholtz_smooth_part(oracle):
    """V_k - V_0 has the smooth kernel (e^{ikr}-1)/(4 pi r): a plain high-order
    product rule reproduces the oracle's adapted coincident value."""
    v, f = _pair_mesh()
    k = 2.5
    vk = oracle.pair(oracle.make_op(oracle.OP_V, oracle.HELMHOLTZ, k, q=12, flags=3), v, f, 0, 0)[0, 0]
    v0 = oracle.pair(oracle.make_op(oracle.OP_V, oracle.LAPLACE, 0.0, q=12, flags=3), v, f, 0, 0)[0, 0]
    p = v[f[0]]

    def inner(x):
        def fn(y):
            r = np.linalg.norm(x - y)
            return (np.expm1(1j * k * r) / r) if r > 1e-12 else 1j * k
        return _outer(p, fn, order=14)

    ref = _outer(p, inner, order=14) / (4 * math.pi)
    # sin(kr)/r is analytic in r^2, so the product rule is spectrally accurate
    # for the imaginary part; (cos kr - 1)/r has an |r| kink on the diagonal
    # and converges only algebraically.
    assert abs((vk - v0).imag - ref.imag) < 1e-10 * abs(ref.imag)
    assert abs((vk - v0).real - ref.real) < 1e-3 * abs(ref.real)


def test_coincident_quirk_is_exact_for_p0(oracle):
    """The reference coincident rule (f(x,y) twice) equals the symmetric one
    for P0 x P0 V, and differs for P1 x P1 pairs of different bases."""
    v, f = _pair_mesh()
    a = oracle.pair(oracle.make_op(oracle.OP_V, oracle.LAPLACE, q=6, flags=0), v, f, 0, 0)
    b = oracle.pair(oracle.make_op(oracle.OP_V, oracle.LAPLACE, q=6, flags=2), v, f, 0, 0)
    assert abs(a - b).max() < 1e-14 * abs(b).max()
    a1 = oracle.pair(oracle.make_op(oracle.OP_V, oracle.LAPLACE, q=6, deg_test=1, deg_trial=1), v, f, 0, 0)
    b1 = oracle.pair(oracle.make_op(oracle.OP_V, oracle.LAPLACE, q=6, deg_test=1, deg_trial=1, flags=2), v, f, 0, 0)
    assert abs(np.diag(a1) - np.diag(b1)).max() < 1e-13 * abs(b1).max()
    assert abs(a1 - b1).max() > 1e-3 * abs(b1).max()


def test_config1_quadrature_presets(bem, oracle):
    """Config 1 shape at L1/L2: the reference's stored-order rules do not
    converge to the analytic trace; the adapted rules do (O(h^2))."""
    x0 = np.array([2.0, 0.5, 0.3])
    errs = {}
    for flags in (0, 3):
        for L in (1, 2):
            v, f = bem.meshes.icosphere(L)
            geom = oracle.geometry(v, f)
            V = oracle.assemble(oracle.make_op(oracle.OP_V, oracle.LAPLACE, q=5, flags=flags), v, f, geom=geom).real
            g = 1 / np.linalg.norm(v - x0, axis=1)
            rhs = oracle.rhs_dirichlet(oracle.LAPLACE, 0.0, 5, v, f, g.astype(complex), True, 1.0, flags=flags,
                                       geom=geom).real
            t = np.linalg.solve(V, rhs)
            c = v[f].mean(1)
            d = c - x0
            texact = -np.einsum("ij,ij->i", d, geom[1]) / np.linalg.norm(d, axis=1) ** 3
            errs[(flags, L)] = rel(t, texact)
    assert errs[(3, 2)] < errs[(3, 1)] / 2.5
    assert errs[(0, 2)] > 0.15


GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "oracle_goldens.npz")
_OPS = {"V": (0, 0, 0), "K": (1, 0, 1), "KADJ": (2, 1, 0), "W": (3, 1, 1)}
_KERNELS = {"lap": (0, 0.0, 1.0), "helm2": (1, 2.0, 1.0), "trans3": (2, 3.0, 1.5), "helm3": (1, 3.0, 1.0)}


def test_oracle_reproduces_goldens(oracle):
    """The committed fixtures (tests/golden/make_golden.py) are what the
    oracle computes today: a change of the restatement shows up here."""
    d = np.load(GOLDEN)
    meshes = sorted({k[: -len("_verts")] for k in d.files if k.endswith("_verts")})
    checked = 0
    for m in meshes:
        v, f = d[m + "_verts"], d[m + "_tris"]
        for key in d.files:
            if not key.startswith(m + "_"):
                continue
            parts = key[len(m) + 1:].split("_")
            if len(parts) != 3 or parts[0] not in _KERNELS:
                continue
            kernel, k, sn = _KERNELS[parts[0]]
            op, dt, dr = _OPS[parts[2]]
            got = oracle.assemble(oracle.make_op(op, kernel, k, sn, 5, dt, dr, int(parts[1][1:])), v, f)
            ref = d[key]
            assert np.abs(got - ref).max() <= 1e-14 * max(np.abs(ref).max(), 1e-300), key
            checked += 1
    assert checked >= 250
    v, f = d["ico1_verts"], d["ico1_tris"]
    assert np.abs(oracle.mass(5, 0, 1, v, f) - d["ico1_mass_p0p1"]).max() <= 1e-16
    for name, rel_expected in (("edge", oracle.COMMON_EDGE), ("vertex", oracle.COMMON_VERTEX),
                               ("separate", oracle.SEPARATE)):
        for rot in range(3):
            assert oracle.relation(d[f"tri_{name}_r{rot}_tris"], 0, 1) == rel_expected
