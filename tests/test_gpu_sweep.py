"""Batched wave-number sweeps on the GPU: the persistent combined-field
assembly against the FP64 oracle, the batched LU against the single-problem
LU (identical factors), batched solves with FP64 refinement, batched 1-norm
condition estimates against numpy, graph capture and replay, and the
autotuner's cache."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def make_ctx(bem, v, f, q=5):
    print(f"[ctx] nv={len(v)} nf={len(f)}", flush=True)
    c = bem.Context(0)
    c.set_mesh(torch.as_tensor(v, device=DEV), torch.as_tensor(f, device=DEV), q)
    return c


def oracle_cf(oracle, v, f, k, eta, flags):
    K = oracle.assemble(oracle.make_op(oracle.OP_K, oracle.HELMHOLTZ, k, 1.0, 5, 1, 1, flags), v, f)
    V = oracle.assemble(oracle.make_op(oracle.OP_V, oracle.HELMHOLTZ, k, 1.0, 5, 1, 1, flags), v, f)
    M = oracle.mass(5, 1, 1, v, f)
    return 0.5 * M + K - 1j * eta * V


@pytest.mark.parametrize("flags", [0, 3])
def test_combined_batch_matches_oracle(bem, oracle, flags):
    """Two meshes of different size and three wave numbers in one batch."""
    meshes = [bem.meshes.perturbed_icosphere(1, seed=0), bem.meshes.icosphere(1)]
    ctxs = [make_ctx(bem, v, f) for v, f in meshes]
    probs = [(0, 3.0), (1, 2.0), (0, 4.5)]
    mats = [torch.zeros(len(meshes[m][0]), len(meshes[m][0]), dtype=torch.complex128, device=DEV)
            for m, _ in probs]
    batch = bem._lib.Batch(bem.Context(0))
    print(f"[cf] flags={flags}", flush=True)
    batch.assemble_combined([(ctxs[m], k, k, A) for (m, k), A in zip(probs, mats)], bem._lib.C128,
                            bem._lib.PREC_FP64, flags)
    for (m, k), A in zip(probs, mats):
        v, f = meshes[m]
        assert rel(A.cpu().numpy(), oracle_cf(oracle, v, f, k, k, flags)) <= 1e-11


@pytest.mark.parametrize("tuning", [(1, 1), (2, 4), (8, 2), (16, 4)])
def test_combined_fp32_and_tunings(bem, oracle, tuning):
    v, f = bem.meshes.perturbed_icosphere(2, seed=0)
    c = make_ctx(bem, v, f)
    A = torch.zeros(len(v), len(v), dtype=torch.complex64, device=DEV)
    batch = bem._lib.Batch(bem.Context(0))
    print(f"[cf32] tuning={tuning}", flush=True)
    batch.assemble_combined([(c, 3.0, 3.0, A)], bem._lib.C64, bem._lib.PREC_FP32, 0, tuning)
    assert rel(A.cpu().numpy(), oracle_cf(oracle, v, f, 3.0, 3.0, 0)) <= 5e-6


def rand_mats(sizes, dtype, seed):
    rng = np.random.default_rng(seed)
    out = []
    for n in sizes:
        a = rng.standard_normal((n, n)) + n ** 0.5 * np.eye(n) * 0.1
        if dtype.is_complex:
            a = a + 1j * rng.standard_normal((n, n))
        out.append(torch.tensor(a, dtype=dtype, device=DEV))
    return out


@pytest.mark.parametrize("dtype", [torch.complex128, torch.complex64, torch.float64, torch.float32])
def test_lu_batched_identical_to_single(bem, dtype):
    sizes = [130, 64, 257, 1, 200, 65]
    mats = rand_mats(sizes, dtype, 1)
    ctx = bem.Context(0)
    # batches factor their panels one launch per column; the single-problem
    # reference does the same (its persistent panel rounds differently)
    ctx.set_option(bem._lib.OPT_LU_PANEL, 0)
    singles, spiv = [], []
    for A in mats:
        LU = A.clone()
        ip = torch.empty(A.shape[0], dtype=torch.int32, device=DEV)
        ctx.lu_factor(LU, ip)
        singles.append(LU)
        spiv.append(ip)
    lus = [A.clone() for A in mats]
    ipivs = [torch.empty(A.shape[0], dtype=torch.int32, device=DEV) for A in mats]
    info = torch.full((len(mats),), -1, dtype=torch.int32, device=DEV)
    print(f"[lu batched] dtype={dtype}", flush=True)
    bem._lib.Batch(ctx).lu_factor(lus, ipivs, info)
    torch.cuda.synchronize()
    assert info.cpu().tolist() == [0] * len(mats)
    for a, b, p, q in zip(lus, singles, ipivs, spiv):
        assert torch.equal(a, b)
        assert torch.equal(p, q)


@pytest.mark.parametrize("pivoting", [0, 1])
@pytest.mark.parametrize("dtype,tol", [(torch.complex128, 1e-14), (torch.complex64, 2e-6), (torch.float64, 1e-14)])
def test_lu_batched_lone_tail(bem, dtype, tol, pivoting):
    """The largest problem outlives the others by several outer blocks, so
    its last columns take the single-problem path (persistent partial-pivoting
    panels) after the batch's steps: its factors must still solve the system."""
    sizes = [1500, 300, 64, 1]
    mats = rand_mats(sizes, dtype, 7)
    ctx = bem.Context(0)
    ctx.set_option(bem._lib.OPT_LU_PIVOTING, pivoting)
    lus = [A.clone() for A in mats]
    ipivs = [torch.empty(A.shape[0], dtype=torch.int32, device=DEV) for A in mats]
    info = torch.full((len(mats),), -1, dtype=torch.int32, device=DEV)
    print(f"[lu lone tail] dtype={dtype} pivoting={pivoting}", flush=True)
    bem._lib.Batch(ctx).lu_factor(lus, ipivs, info)
    torch.cuda.synchronize()
    assert info.cpu().tolist() == [0] * len(mats)
    rng = np.random.default_rng(8)
    for A, LU, ip in zip(mats, lus, ipivs):
        n = A.shape[0]
        b = torch.tensor(rng.standard_normal(n), dtype=dtype, device=DEV)
        x = b.clone()
        ctx.lu_solve(LU, ip, x)
        A64, x64, b64 = (t.to(torch.complex128) for t in (A, x, b))
        r = (torch.linalg.vector_norm(A64 @ x64 - b64) /
             (torch.linalg.matrix_norm(A64) * torch.linalg.vector_norm(x64))).item()
        assert r <= tol, (n, r)


def test_lu_batched_reports_singular(bem):
    mats = rand_mats([70, 40], torch.complex128, 2)
    mats[1][:, 5] = 0
    ipivs = [torch.empty(A.shape[0], dtype=torch.int32, device=DEV) for A in mats]
    info = torch.full((2,), -1, dtype=torch.int32, device=DEV)
    bem._lib.Batch(bem.Context(0)).lu_factor(mats, ipivs, info)
    assert info.cpu().tolist() == [0, 6]


@pytest.mark.parametrize("dtype", [torch.complex64, torch.complex128])
def test_batch_solve_with_refinement(bem, dtype):
    sizes = [100, 257, 64]
    mats = rand_mats(sizes, dtype, 3)
    rng = np.random.default_rng(4)
    ctx = bem.Context(0)
    batch = bem._lib.Batch(ctx)
    lus = [A.clone() for A in mats]
    ipivs = [torch.empty(n, dtype=torch.int32, device=DEV) for n in sizes]
    info = torch.zeros(len(sizes), dtype=torch.int32, device=DEV)
    batch.lu_factor(lus, ipivs, info)
    bs = [torch.tensor(rng.standard_normal(n) + 1j * rng.standard_normal(n), device=DEV) for n in sizes]
    xs = [torch.zeros(n, dtype=torch.complex128, device=DEV) for n in sizes]
    resid = torch.zeros(len(sizes), dtype=torch.float64, device=DEV)
    print(f"[batch solve] dtype={dtype}", flush=True)
    batch.solve(lus, ipivs, mats, bs, xs, refine_iters=3, resid=resid)
    torch.cuda.synchronize()
    for A, b, x, r in zip(mats, bs, xs, resid.cpu().tolist()):
        a = A.cpu().numpy().astype(np.complex128)
        ref = np.linalg.solve(a, b.cpu().numpy())
        assert rel(x.cpu().numpy(), ref) <= 1e-11
        true_r = np.linalg.norm(b.cpu().numpy() - a @ x.cpu().numpy()) / np.linalg.norm(b.cpu().numpy())
        assert r <= 1e-12 and abs(r - true_r) <= 1e-13
    # without refinement and matrices
    xs2 = [torch.zeros_like(x) for x in xs]
    batch.solve(lus, ipivs, None, bs, xs2)
    tol = 1e-11 if dtype == torch.complex128 else 1e-3
    for A, b, x in zip(mats, bs, xs2):
        ref = np.linalg.solve(A.cpu().numpy().astype(np.complex128), b.cpu().numpy())
        assert rel(x.cpu().numpy(), ref) <= tol


@pytest.mark.parametrize("dtype", [torch.complex128, torch.float64])
def test_batch_condest(bem, dtype):
    sizes = [50, 120, 257]
    mats = rand_mats(sizes, dtype, 5)
    ctx = bem.Context(0)
    batch = bem._lib.Batch(ctx)
    lus = [A.clone() for A in mats]
    ipivs = [torch.empty(n, dtype=torch.int32, device=DEV) for n in sizes]
    batch.lu_factor(lus, ipivs, torch.zeros(3, dtype=torch.int32, device=DEV))
    out = torch.zeros(3, 2, dtype=torch.float64, device=DEV)
    print(f"[condest] dtype={dtype}", flush=True)
    batch.condest(lus, ipivs, mats, out, iters=4)
    o = out.cpu().numpy()
    for q, A in enumerate(mats):
        a = A.cpu().numpy()
        n1 = np.linalg.norm(a, 1)
        ninv = np.linalg.norm(np.linalg.inv(a), 1)
        assert abs(o[q, 0] - n1) <= 1e-12 * n1
        assert o[q, 1] <= ninv * (1 + 1e-10)
        assert o[q, 1] >= 0.3 * ninv


def small_sweep(bem, dtype=torch.complex64, precision=1, refine=2):
    meshes = [bem.meshes.icosphere(1), bem.meshes.icosphere(2)]
    probs = [(0, 2.0), (1, 2.0), (1, 2.5), (0, 3.0)]
    return bem.sweep.WavenumberSweep(meshes, probs, dtype=dtype, precision=precision, refine_iters=refine,
                                     direction=bem.meshes.seeded_direction(2))


def test_sweep_graph_replay_matches_eager(bem):
    sw = small_sweep(bem)
    print("[sweep] eager", flush=True)
    sw.run()
    torch.cuda.synchronize()
    x0 = [x.cpu().numpy() for x in sw.x]
    r0 = sw.results()
    assert all(r["info"] == 0 for r in r0)
    assert max(r["rel_residual"] for r in r0) <= 1e-10
    print("[sweep] capture", flush=True)
    sw.capture()
    for rep in range(3):
        for x in sw.x:
            x.fill_(7.0)  # eager work between replays; the replay must overwrite it
        sw.cond.zero_()
        torch.cuda.synchronize()
        sw.replay()
        torch.cuda.synchronize()
        r = sw.results()
        for a, b in zip(x0, sw.x):
            assert rel(b.cpu().numpy(), a) <= 1e-4
        for p, q in zip(r0, r):
            assert q["info"] == 0 and q["rel_residual"] <= 1e-10
            assert abs(q["cond1_est"] - p["cond1_est"]) <= 1e-3 * p["cond1_est"]


def test_sweep_mixed_matches_fp64(bem):
    """c64 assembly (FP32 regular pairs) + c64 LU + FP64 refinement against
    the all-FP64 complex128 sweep."""
    s32 = small_sweep(bem)
    s64 = small_sweep(bem, dtype=torch.complex128, precision=0, refine=0)
    s32.run()
    s64.run()
    torch.cuda.synchronize()
    for a, b in zip(s32.x, s64.x):
        e = rel(a.cpu().numpy(), b.cpu().numpy())
        print(f"[sweep mixed] rel err {e:.3e}", flush=True)
        assert e <= 2e-5


def test_autotune_cache(bem, tmp_path):
    sw = small_sweep(bem)
    path = str(tmp_path / "tune.json")
    r1 = bem.sweep.autotune(sw, candidates=((1, 2), (2, 4)), cache_path=path)
    assert not r1["cached"] and tuple(r1["tuning"]) in ((1, 2), (2, 4))
    sw.tuning = None
    r2 = bem.sweep.autotune(sw, candidates=((1, 2), (2, 4)), cache_path=path)
    assert r2["cached"] and tuple(sw.tuning) == tuple(r1["tuning"])


def test_condest_on_l3_combined_field_matrices(bem):
    """The batched 1-norm condition estimate on the sweep's own operator:
    combined-field matrices of icosphere L3 (642 P1 unknowns) at k = 16.5 and
    17.9, complex128, against the exact kappa_1 from numpy (an estimate is a
    lower bound, and Hager-Higham is usually within a small factor)."""
    mesh = bem.meshes.icosphere(3)
    sw = bem.sweep.WavenumberSweep([mesh], [(0, 16.5), (0, 17.9)], dtype=torch.complex128, precision=0,
                                   refine_iters=0, condest_iters=4, direction=bem.meshes.seeded_direction(2))
    print("[condest L3] combined field", flush=True)
    sw.run()
    torch.cuda.synchronize()
    for q, r in enumerate(sw.results()):
        a = sw.A[q].cpu().numpy()
        exact = np.linalg.cond(a, 1)
        print(f"[condest L3] k={r['k']}: estimate {r['cond1_est']:.4e} exact {exact:.4e}", flush=True)
        assert abs(r["norm1_A"] - np.linalg.norm(a, 1)) <= 1e-12 * np.linalg.norm(a, 1)
        assert r["cond1_est"] <= exact * (1 + 1e-10)
        assert r["cond1_est"] >= 0.3 * exact
    sw.close()
