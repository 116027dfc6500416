"""Derivatives on the GPU: the k-derivative products (BEM_FLAG_DK), the
adjoint contraction bem_shape_gradient, the area gradient, the transposed LU
solve, and the NeumannTrace autograd function (shape gradient of the
boundary-data misfit against central finite differences, on the GPU and
against the FP64 oracle)."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


def make_ctx(bem, v, f, q=5):
    print(f"[ctx] nv={len(v)} nf={len(f)} q={q}", flush=True)
    ctx = bem.Context(0)
    ctx.set_mesh(torch.as_tensor(v, device=DEV), torch.as_tensor(f, device=DEV), q)
    return ctx


def crandn(rng, n):
    return torch.tensor(rng.standard_normal(n) + 1j * rng.standard_normal(n), device=DEV)


OPS = [(0, 0, 0), (1, 0, 1), (2, 1, 0), (3, 1, 1), (0, 1, 1), (1, 1, 1)]  # (op, deg_test, deg_trial)


@pytest.mark.parametrize("kernel,k,sqrt_n", [(1, 2.0, 1.0), (2, 5.0, 1.5)])
@pytest.mark.parametrize("op,dt,dr", OPS)
def test_dk_apply_matches_fd(bem, op, dt, dr, kernel, k, sqrt_n):
    """dOp/dk . x against the central difference of Op(k +- h) . x."""
    v, f = bem.meshes.perturbed_icosphere(1, seed=0)
    ctx = make_ctx(bem, v, f)
    rng = np.random.default_rng(7)
    nt = len(v) if dt else len(f)
    nr = len(v) if dr else len(f)
    x = crandn(rng, nr)
    print(f"[dk] op={op} dt={dt} dr={dr} kernel={kernel}", flush=True)

    def app(kk, flags):
        y = torch.zeros(nt, dtype=torch.complex128, device=DEV)
        d = bem.Context.op_desc(op, kernel, kk, sqrt_n, dt, dr, bem._lib.PREC_FP64, 1, flags)
        ctx.apply(d, x, y, conj_x=False)
        return y

    h = 1e-5
    fd = (app(k + h, 0) - app(k - h, 0)) / (2 * h)
    got = app(k, bem._lib.FLAG_DK)
    err = float(torch.linalg.vector_norm(got - fd) / torch.linalg.vector_norm(fd))
    assert err < 1e-7, err


def test_dk_laplace_is_zero_and_assemble_rejects(bem):
    v, f = bem.meshes.icosphere(1)
    ctx = make_ctx(bem, v, f)
    x = torch.ones(len(f), dtype=torch.complex128, device=DEV)
    y = torch.zeros(len(f), dtype=torch.complex128, device=DEV)
    ctx.apply(bem.Context.op_desc(0, 0, 0.0, 1.0, 0, 0, 0, 1, bem._lib.FLAG_DK), x, y)
    assert float(y.abs().max()) == 0.0
    A = torch.zeros(len(f), len(f), dtype=torch.complex128, device=DEV)
    with pytest.raises(NotImplementedError):
        ctx.assemble(bem.Context.op_desc(0, 1, 2.0, 1.0, 0, 0, 0, 1, bem._lib.FLAG_DK), A)


@pytest.mark.parametrize("flags", [0, 3])
@pytest.mark.parametrize("kernel,k", [(0, 0.0), (1, 2.0)])
@pytest.mark.parametrize("op,dr", [(0, 0), (1, 1)])
def test_shape_gradient_matches_fd(bem, op, dr, kernel, k, flags):
    """Re(lam^T dA/dX u) . e against Re(lam^T (A(x+he) - A(x-he)) u) / 2h with
    assembled matrices, over regular and near-field pairs."""
    v, f = bem.meshes.perturbed_icosphere(2, seed=0)
    rng = np.random.default_rng(11)
    nf, nv = len(f), len(v)
    lam = crandn(rng, nf)
    u = crandn(rng, nv if dr else nf)
    e = rng.standard_normal(v.shape)
    e /= np.linalg.norm(e)
    print(f"[shape] op={op} kernel={kernel} flags={flags}", flush=True)

    def form(vv, kk):
        ctx = make_ctx(bem, vv, f)
        A = torch.zeros(nf, nv if dr else nf, dtype=torch.complex128, device=DEV)
        ctx.assemble(bem.Context.op_desc(op, kernel, kk, 1.0, 0, dr, 0, 1, flags), A)
        return float((lam @ (A @ u)).real)

    h = 1e-6
    fd = (form(v + h * e, k) - form(v - h * e, k)) / (2 * h)
    ctx = make_ctx(bem, v, f)
    grad = torch.zeros(nv, 3, dtype=torch.float64, device=DEV)
    gk = torch.zeros(1, dtype=torch.float64, device=DEV)
    ctx.shape_gradient(bem.Context.op_desc(op, kernel, k, 1.0, 0, dr, 0, 1, flags), lam, u, grad, gk)
    got = float((grad.cpu().numpy() * e).sum())
    assert abs(got - fd) <= 1e-6 * abs(fd), (got, fd)
    if kernel:
        hk = 1e-5
        fdk = (form(v, k + hk) - form(v, k - hk)) / (2 * hk)
        assert abs(float(gk) - fdk) <= 1e-7 * abs(fdk), (float(gk), fdk)


def test_shape_gradient_translation_invariance(bem):
    """A rigid translation leaves every entry unchanged: the vertex gradient
    sums to zero."""
    v, f = bem.meshes.perturbed_icosphere(2, seed=0)
    ctx = make_ctx(bem, v, f)
    rng = np.random.default_rng(3)
    grad = torch.zeros(len(v), 3, dtype=torch.float64, device=DEV)
    for op, dr in ((0, 0), (1, 1)):
        u = crandn(rng, len(v) if dr else len(f))
        ctx.shape_gradient(bem.Context.op_desc(op, 1, 2.0, 1.0, 0, dr, 0, 1, 3), crandn(rng, len(f)), u, grad)
    s = grad.sum(0).abs().max().item()
    assert s <= 1e-11 * grad.abs().max().item()


@pytest.mark.parametrize("kernel,k,flags", [(1, 2.0, 0), (1, 3.0, 3), (0, 0.0, 0)])
def test_shape_gradient_kt(bem, kernel, k, flags):
    """bem_shape_gradient_kt: kt = K^T lam equals the assembled K, transposed,
    times lam; its vertex and k gradients equal bem_shape_gradient's.  L3
    (1280 panels): five trial blocks of the far kernel, near pairs of every
    relation."""
    v, f = bem.meshes.perturbed_icosphere(3, seed=4)
    ctx = make_ctx(bem, v, f)
    rng = np.random.default_rng(9)
    lam, u = crandn(rng, len(f)), crandn(rng, len(v))
    desc = bem.Context.op_desc(1, kernel, k, 1.0, 0, 1, 0, 1, flags)
    g1 = torch.zeros(len(v), 3, dtype=torch.float64, device=DEV)
    k1 = torch.zeros(1, dtype=torch.float64, device=DEV)
    ctx.shape_gradient(desc, lam, u, g1, k1, alpha=0.5 - 1.5j)
    g2 = torch.zeros_like(g1)
    k2 = torch.zeros_like(k1)
    kt = torch.zeros(len(v), dtype=torch.complex128, device=DEV)
    ctx.shape_gradient(desc, lam, u, g2, k2, alpha=0.5 - 1.5j, kt=kt)
    assert (g2 - g1).abs().max().item() <= 1e-12 * g1.abs().max().item()
    if kernel:
        assert abs(float(k2) - float(k1)) <= 1e-12 * abs(float(k1))
    K = torch.zeros(len(f), len(v), dtype=torch.complex128, device=DEV)
    ctx.assemble(desc, K)
    ref = K.transpose(0, 1) @ lam
    err = (kt - ref).norm().item() / ref.norm().item()
    assert err <= 1e-13, err
    with pytest.raises(bem.BEMError):  # kt is for K only
        ctx.shape_gradient(bem.Context.op_desc(0, kernel, k, 1.0, 0, 0, 0, 1, flags), lam, crandn(rng, len(f)),
                           g1, k1, kt=kt)


def test_area_gradient(bem):
    v, f = bem.meshes.perturbed_icosphere(2, seed=0)
    rng = np.random.default_rng(5)
    coef = rng.standard_normal(len(f))
    e = rng.standard_normal(v.shape)
    ctx = make_ctx(bem, v, f)
    grad = torch.zeros(len(v), 3, dtype=torch.float64, device=DEV)
    ctx.area_gradient(torch.tensor(coef, device=DEV), grad)

    def areas(vv):
        a, b, c = vv[f[:, 0]], vv[f[:, 1]], vv[f[:, 2]]
        return 0.5 * np.linalg.norm(np.cross(b - a, c - a), axis=1)

    h = 1e-6
    fd = (coef @ areas(v + h * e) - coef @ areas(v - h * e)) / (2 * h)
    got = float((grad.cpu().numpy() * e).sum())
    assert abs(got - fd) <= 1e-7 * abs(fd)


@pytest.mark.parametrize("dtype", [torch.complex128, torch.float64, torch.complex64, torch.float32])
@pytest.mark.parametrize("n", [100, 257])
def test_lu_solve_transpose(bem, dtype, n):
    rng = np.random.default_rng(n)
    cplx = dtype.is_complex
    a = rng.standard_normal((n, n)) + (1j * rng.standard_normal((n, n)) if cplx else 0)
    b = rng.standard_normal(n) + (1j * rng.standard_normal(n) if cplx else 0)
    A = torch.tensor(a, dtype=dtype, device=DEV)
    ctx = bem.Context(0)
    LU = A.clone()
    ipiv = torch.empty(n, dtype=torch.int32, device=DEV)
    print(f"[lu_t] n={n} dtype={dtype}", flush=True)
    ctx.lu_factor(LU, ipiv)
    x = torch.tensor(b, dtype=dtype, device=DEV)
    ctx.lu_solve_transpose(LU, ipiv, x)
    ref = np.linalg.solve(a.T, b)
    err = np.linalg.norm(x.cpu().numpy() - ref) / np.linalg.norm(ref)
    tol = 1e-11 if dtype in (torch.complex128, torch.float64) else 2e-3
    assert err <= tol, err


# ---------------------------------------------------------------------------
# NeumannTrace: J(x) = ||t(x) - t_target||^2_M
# ---------------------------------------------------------------------------
X0 = (2.0, 0.5, 0.3)


def misfit_setup(bem, level, kernel, k, flags=0):
    v, f = bem.meshes.icosphere(level)
    vt, _ = bem.meshes.perturbed_icosphere(level, seed=1, degrees=(3,))
    A = bem.adjoint
    prob = A.DirichletProblem(f, kernel=kernel, gauss_qr=5, flags=flags)
    kt = torch.tensor(float(k), dtype=torch.float64, device=DEV)
    vt_t = torch.tensor(vt, device=DEV)
    with torch.no_grad():
        t_target = A.neumann_trace(vt_t, kt, A.point_source(vt_t, kt, X0, kernel), prob)

    def J(vv, kk, grad=False):
        vv = torch.as_tensor(vv, device=DEV).clone().requires_grad_(grad)
        kk = torch.tensor(float(kk), dtype=torch.float64, device=DEV, requires_grad=grad)
        t = A.neumann_trace(vv, kk, A.point_source(vv, kk, X0, kernel), prob)
        j = A.misfit(t, t_target, vv, prob.panels)
        if not grad:
            return float(j)
        j.backward()
        return float(j.detach()), vv.grad.cpu().numpy(), float(kk.grad) if kernel else 0.0

    return v, f, vt, J


@pytest.mark.parametrize("kernel,k", [(0, 0.0), (1, 2.0)])
@pytest.mark.parametrize("flags", [0, 3])
def test_neumann_trace_gradient_matches_oracle_fd(bem, oracle, kernel, k, flags):
    """FP64 reference gradient: central differences of the misfit computed
    entirely by the CPU oracle (its own operators, numpy solve) at L1."""
    v, f, vt, J = misfit_setup(bem, 1, kernel, k, flags)
    j0, gx, gk = J(v, k, grad=True)
    okern = oracle.LAPLACE if kernel == 0 else oracle.HELMHOLTZ

    def g_of(vv, kk):
        r = np.linalg.norm(vv - np.array(X0), axis=1)
        return (1 / r) if kernel == 0 else np.exp(1j * kk * r) / r

    def trace(vv, kk):
        geom = oracle.geometry(vv, f)
        Vm = oracle.assemble(oracle.make_op(oracle.OP_V, okern, kk, 1.0, 5, 0, 0, flags), vv, f, geom=geom)
        rhs = oracle.rhs_dirichlet(okern, kk, 5, vv, f, g_of(vv, kk).astype(complex), False, 1.0, flags=flags,
                                   geom=geom)
        return np.linalg.solve(Vm, rhs), geom[0]

    t_star, _ = trace(vt, k)

    def Jo(vv, kk):
        t, area = trace(vv, kk)
        return float((area * np.abs(t - t_star) ** 2).sum())

    assert abs(Jo(v, k) - j0) <= 1e-10 * abs(j0)
    rng = np.random.default_rng(3)
    h = 1e-6
    for _ in range(4):
        d = rng.standard_normal(v.shape)
        d /= np.linalg.norm(d)
        fd = (Jo(v + h * d, k) - Jo(v - h * d, k)) / (2 * h)
        got = float((gx * d).sum())
        print(f"[oracle-fd] kernel={kernel} flags={flags} adj={got:.12e} fd={fd:.12e}", flush=True)
        assert abs(got - fd) <= 1e-5 * abs(fd)
    if kernel:
        fdk = (Jo(v, k + 1e-6) - Jo(v, k - 1e-6)) / 2e-6
        assert abs(gk - fdk) <= 1e-5 * abs(fdk)


@pytest.mark.parametrize("kernel,k", [(0, 0.0), (1, 2.0)])
def test_config4_shape_gradient_l4(bem, kernel, k):
    """Icosphere L4 (2562 vertices): adjoint gradient against central
    differences (h = 1e-6) along 32 seeded unit directions in R^(3*2562),
    and dJ/dk, to 1e-5."""
    v, f, vt, J = misfit_setup(bem, 4, kernel, k, flags=0)
    j0, gx, gk = J(v, k, grad=True)
    assert gx.shape == (2562, 3)
    rng = np.random.default_rng(3)
    h = 1e-6
    gnorm = float(np.linalg.norm(gx))
    worst = worst_raw = 0.0
    for n in range(32):
        d = rng.standard_normal(v.shape)
        d /= np.linalg.norm(d)
        fd = (J(v + h * d, k) - J(v - h * d, k)) / (2 * h)
        got = float((gx * d).sum())
        # error relative to the directional derivative itself; beside it the
        # error relative to 1e-4 |g| |d| (a random unit direction in R^7686
        # sees about 1e-2 |g|).  The forward pass is bitwise reproducible
        # (every assembly and LU sum runs in a fixed order), so J(x + h d) and
        # J(x - h d) carry no run-to-run noise.
        worst = max(worst, abs(got - fd) / max(abs(fd), 1e-4 * gnorm))
        worst_raw = max(worst_raw, abs(got - fd) / abs(fd))
    print(f"[config4] kernel={kernel} J={j0:.6e} worst rel err over 32 directions {worst_raw:.3e} "
          f"(relative to max(|fd|, 1e-4 |g|): {worst:.3e})", flush=True)
    assert worst_raw <= 1e-5
    if kernel:
        fdk = (J(v, k + 1e-6) - J(v, k - 1e-6)) / 2e-6
        print(f"[config4] dJ/dk adjoint {gk:.12e} fd {fdk:.12e}", flush=True)
        assert abs(gk - fdk) <= 1e-5 * abs(fdk)


def test_adjoint_gradients_bitwise_reproducible(bem):
    """Every sum of the backward pass (regular and near pairs, panel corners,
    vertices, dJ/dk, K^T mu, the P1 x P0 mass product) runs in a fixed order:
    two backward passes give identical gradients for the vertices, k and g."""
    v, f = bem.meshes.perturbed_icosphere(3, seed=2)
    A = bem.adjoint
    prob = A.DirichletProblem(f, kernel=1, gauss_qr=5, flags=0)
    outs = []
    for rep in range(2):
        print(f"[repro] backward {rep}", flush=True)
        vv = torch.tensor(v, device=DEV, requires_grad=True)
        kk = torch.tensor(2.0, dtype=torch.float64, device=DEV, requires_grad=True)
        g = A.point_source(vv.detach(), kk.detach(), X0, 1).detach().clone().requires_grad_(True)
        t = A.neumann_trace(vv, kk, g, prob)
        (t.abs() ** 2).sum().backward()
        outs.append((vv.grad.clone(), kk.grad.clone(), g.grad.clone()))
    for name, a, b in zip(("vertices", "k", "g"), *outs):
        assert bool(torch.isfinite(torch.view_as_real(a) if a.is_complex() else a).all()), name
        assert float(a.abs().max()) > 0.0, name
        assert torch.equal(a, b), name


# ---------------------------------------------------------------------------
# Transmission: dJ/dk_i through the interior Calderon blocks
# ---------------------------------------------------------------------------
def transmission_manager(bem, v, f, k, ni, quadrature="reference"):
    def fD(x, sqrt_n):
        return torch.exp(1j * k * sqrt_n * x[:, 0])

    def fN(x, n, sqrt_n):
        return 1j * k * sqrt_n * n[:, 0] * torch.exp(1j * k * sqrt_n * x[:, 0])

    params = {"kernel": "Helmholtz_Transmission", "k": k, "n_i": ni, "n_o": 1.0, "boundary": "Full",
              "GaussQR": 5, "Q_Neumann": 0, "Q_Dirichlet": 1, "vertices": v, "panels": f,
              "quadrature": quadrature}
    cm = bem.CoreManager(params)
    cm.initialization(fD, fN)
    return cm


@pytest.mark.parametrize("quadrature", ["reference", "corrected"])
def test_transmission_k_gradient_matches_fd(bem, quadrature):
    """J = ||x - x*||^2 of the transmission solution with the incident data
    b held fixed: adjoint dJ/dk_i against central differences in k_i."""
    v, f = bem.meshes.perturbed_icosphere(1, seed=0)
    k, ni = 5.0, 2.25
    cm = transmission_manager(bem, v, f, k, ni, quadrature)
    cm.run()
    bm = cm._BEM_manager
    n = bm.N_Neumann + bm.M_Dirichlet
    b = bm.rhs[:n].clone()
    rng = np.random.default_rng(9)
    xs = torch.tensor(rng.standard_normal(n) + 1j * rng.standard_normal(n), device=DEV) * bm.raw_solved.abs().mean()
    x0 = bm.raw_solved.clone()
    got = bem.adjoint.transmission_k_gradient(bm, 2 * (x0 - xs))

    def J(ki):
        c2 = transmission_manager(bem, v, f, k, (ki / k) ** 2, quadrature)
        c2.run()
        b2 = c2._BEM_manager
        b2.rhs[:n] = b
        x = b2.solve()
        return float(((x - xs).abs() ** 2).sum())

    ki = k * np.sqrt(ni)
    h = 1e-6
    fd = (J(ki + h) - J(ki - h)) / (2 * h)
    print(f"[transmission dk] adjoint {got:.12e} fd {fd:.12e}", flush=True)
    assert abs(got - fd) <= 1e-6 * abs(fd)


@pytest.mark.parametrize("flags", [0, 3])
@pytest.mark.parametrize("kernel,k,sqrt_n", [(1, 3.0, 1.0), (2, 5.0, 1.5)])
def test_vkw_dk_forms_match_products(bem, kernel, k, sqrt_n, flags):
    """The fused k-derivative forms (one evaluation per unordered point pair,
    contracted in both orders) equal dots with the matrix-free BEM_FLAG_DK
    products of V, K and W, which are themselves checked against finite
    differences of assembled matrices."""
    v, f = bem.meshes.perturbed_icosphere(2, seed=4)
    ctx = bem.Context(0)
    ctx.set_mesh(torch.tensor(v, device=DEV), torch.tensor(f, device=DEV), 5)
    nv, nf = len(v), len(f)
    rng = np.random.default_rng(11)

    def cvec(n):
        return torch.tensor(rng.standard_normal(n) + 1j * rng.standard_normal(n), device=DEV)

    aN, bN, aD, bD = cvec(nf), cvec(nf), cvec(nv), cvec(nv)
    L = bem._lib
    got = ctx.vkw_dk_forms(bem.Context.op_desc(0, kernel, k, sqrt_n, 0, 0, L.PREC_FP64, L.C128, flags),
                           aN, bN, aD, bD).cpu().numpy()

    def dprod(op, dt, dr, x, n):
        y = torch.zeros(n, dtype=torch.complex128, device=DEV)
        ctx.apply(bem.Context.op_desc(op, kernel, k, sqrt_n, dt, dr, L.PREC_FP64, L.C128, flags | L.FLAG_DK), x, y,
                  alpha=1.0, conj_x=False)
        return y

    ref = np.array([
        (aN * dprod(L.OP_V, 0, 0, bN, nf)).sum().item(),
        (aN * dprod(L.OP_K, 0, 1, bD, nf)).sum().item(),
        (bN * dprod(L.OP_K, 0, 1, aD, nf)).sum().item(),
        (aD * dprod(L.OP_W, 1, 1, bD, nv)).sum().item(),
    ])
    err = np.abs(got - ref) / np.abs(ref)
    print(f"[dk forms] kernel={kernel} flags={flags} rel errors {err}", flush=True)
    assert err.max() <= 1e-12


def test_transmission_dA_contract_fused_equals_products(bem):
    v, f = bem.meshes.perturbed_icosphere(2, seed=0)
    cm = transmission_manager(bem, v, f, 5.0, 2.25)
    cm.run()
    bm = cm._BEM_manager
    n = bm.N_Neumann + bm.M_Dirichlet
    rng = np.random.default_rng(5)
    mu = torch.tensor(rng.standard_normal(n) + 1j * rng.standard_normal(n), device=DEV)
    fused = bem.adjoint.transmission_dA_contract(bm, mu, bm.raw_solved)
    bm.fuse_layers = False
    sep = bem.adjoint.transmission_dA_contract(bm, mu, bm.raw_solved)
    print(f"[dA contract] fused {fused} separate {sep}", flush=True)
    assert abs(fused - sep) <= 1e-12 * abs(sep)


def test_vkw_dk_forms_octahedron(bem):
    """Eight panels: four regular pairs, most tiles of the pass empty."""
    v = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]], dtype=np.float64)
    f = np.array([[0, 2, 4], [2, 1, 4], [1, 3, 4], [3, 0, 4], [2, 0, 5], [1, 2, 5], [3, 1, 5], [0, 3, 5]],
                 dtype=np.int32)
    ctx = bem.Context(0)
    ctx.set_mesh(torch.tensor(v, device=DEV), torch.tensor(f, device=DEV), 5)
    rng = np.random.default_rng(12)
    vec = [torch.tensor(rng.standard_normal(n) + 1j * rng.standard_normal(n), device=DEV) for n in (8, 8, 6, 6)]
    L = bem._lib
    got = ctx.vkw_dk_forms(bem.Context.op_desc(0, 1, 2.5, 1.0, 0, 0, L.PREC_FP64, L.C128, 0), *vec).cpu().numpy()

    def dprod(op, dt, dr, x, n):
        y = torch.zeros(n, dtype=torch.complex128, device=DEV)
        ctx.apply(bem.Context.op_desc(op, 1, 2.5, 1.0, dt, dr, L.PREC_FP64, L.C128, L.FLAG_DK), x, y, alpha=1.0,
                  conj_x=False)
        return y

    aN, bN, aD, bD = vec
    ref = np.array([(aN * dprod(L.OP_V, 0, 0, bN, 8)).sum().item(), (aN * dprod(L.OP_K, 0, 1, bD, 8)).sum().item(),
                    (bN * dprod(L.OP_K, 0, 1, aD, 8)).sum().item(), (aD * dprod(L.OP_W, 1, 1, bD, 6)).sum().item()])
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()
