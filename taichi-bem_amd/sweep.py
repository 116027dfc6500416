"""Batched wave-number sweeps of the exterior Dirichlet (sound-soft)
scattering problem with the combined-field operator

    (1/2 M + K - i eta V) phi = M g,   g = -u_inc = -exp(i k d.x),   eta = k,

on P1 unknowns, for many (mesh, k) problems at once: one persistent
assembly launch set, one batched c64 LU, batched solves with FP64
refinement and batched 1-norm condition estimates, all of which can be
captured once in a HIP graph and replayed.

The reference runs the same kind of sweep as a Python loop, one wave number
at a time (demo/demo_3d_plot_Sio.py:67-76, 60 wave numbers in [16, 18] at
step 1/30 per readme.md:34), assembling with the Taichi kernels and taking a
dense SVD per k (BEMManager.get_mat_A1_inv_norm, BEM_manager.py:874-936).
It has no combined-field operator; this one is built from its K, V and mass
blocks.
"""
from __future__ import annotations

import json
import os
import time

import numpy as np
import torch

from . import _lib


class WavenumberSweep:
    def __init__(self, meshes, problems, dtype=torch.complex64, precision: int = _lib.PREC_FP32, flags: int = 0,
                 direction=None, gauss_qr: int = 5, device: int = 0, eta_factor: float = 1.0,
                 refine_iters: int = 2, condest_iters: int = 3, tuning: tuple[int, int] | None = None,
                 pivoting: int = 2):
        """meshes: list of (vertices [nv,3], panels [nf,3]) numpy arrays;
        problems: list of (mesh index, k)."""
        self.dev = torch.device("cuda", device)
        self.dtype = dtype
        self.precision = int(precision)
        self.flags = int(flags)
        self.refine_iters = int(refine_iters)
        self.condest_iters = int(condest_iters)
        self.tuning = tuning
        self.batch_ctx = _lib.Context(device)
        # the combined-field operator is well conditioned (cond_1 ~ 1e2 on
        # the sweep's meshes) and its c64 factors only precondition the FP64
        # refinement, so they use block-local pivoting (2: no waits between
        # workgroups); 1 selects tournament, 0 partial pivoting
        self.batch_ctx.set_option(_lib.OPT_LU_PIVOTING, int(pivoting))
        self.batch = _lib.Batch(self.batch_ctx)
        self.ctxs = []
        self.nv = []
        for v, f in meshes:
            c = _lib.Context(device)
            c.set_mesh(torch.as_tensor(np.asarray(v, dtype=np.float64), device=self.dev),
                       torch.as_tensor(np.asarray(f, dtype=np.int32), device=self.dev), gauss_qr)
            self.ctxs.append(c)
            self.nv.append(int(len(v)))
        self.problems = [(int(m), float(k)) for m, k in problems]
        self.eta_factor = float(eta_factor)
        d = np.asarray(direction if direction is not None else (0.0, 0.0, 1.0), dtype=np.float64)
        d = d / np.linalg.norm(d)
        sizes = [self.nv[m] for m, _ in self.problems]
        tot = int(sum(n * n for n in sizes))
        self.A_flat = torch.zeros(tot, dtype=dtype, device=self.dev)
        self.LU_flat = torch.zeros(tot, dtype=dtype, device=self.dev)
        self.ipiv_flat = torch.zeros(int(sum(sizes)), dtype=torch.int32, device=self.dev)
        self.A, self.LU, self.ipiv, self.b, self.x = [], [], [], [], []
        off = poff = 0
        for (m, k), n in zip(self.problems, sizes):
            self.A.append(self.A_flat[off:off + n * n].view(n, n))
            self.LU.append(self.LU_flat[off:off + n * n].view(n, n))
            self.ipiv.append(self.ipiv_flat[poff:poff + n])
            off += n * n
            poff += n
            v = torch.as_tensor(np.asarray(meshes[m][0], dtype=np.float64), device=self.dev)
            g = -torch.exp(1j * k * (v @ torch.as_tensor(d, device=self.dev))).to(torch.complex128)
            b = torch.zeros(n, dtype=torch.complex128, device=self.dev)
            self.ctxs[m].mass_apply(1, 1, g, b, 1.0, conj_x=False)
            self.b.append(b)
            self.x.append(torch.zeros(n, dtype=torch.complex128, device=self.dev))
        self.info = torch.zeros(len(self.problems), dtype=torch.int32, device=self.dev)
        self.resid = torch.zeros(len(self.problems), dtype=torch.float64, device=self.dev)
        self.cond = torch.zeros(len(self.problems), 2, dtype=torch.float64, device=self.dev)
        self.graph = None

    # -- the launch sequence ------------------------------------------------
    def cf_problems(self):
        return [(self.ctxs[m], k, self.eta_factor * k, A) for (m, k), A in zip(self.problems, self.A)]

    def assemble(self, tuning=None) -> None:
        self.A_flat.zero_()
        self.batch.assemble_combined(self.cf_problems(), _lib.dtype_code(self.A[0]), self.precision, self.flags,
                                     tuning or self.tuning)

    def run(self) -> None:
        """assembly, factorisation, refined solves and condition estimates"""
        self.assemble()
        _lib.convert(self.A_flat, self.LU_flat)
        self.batch.lu_factor(self.LU, self.ipiv, self.info)
        self.batch.solve(self.LU, self.ipiv, self.A, self.b, self.x, self.refine_iters, self.resid)
        self.batch.condest(self.LU, self.ipiv, self.A, self.cond, self.condest_iters)

    def capture(self) -> torch.cuda.CUDAGraph:
        """One eager run (uploads the batch tables), then the same sequence
        captured in a graph."""
        self.run()
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self.run()
        torch.cuda.synchronize()
        self.graph = g
        return g

    def replay(self) -> None:
        if self.graph is None:
            raise RuntimeError("capture() first")
        self.graph.replay()

    def close(self) -> None:
        """Releases the contexts of the batch and of every mesh."""
        self.graph = None
        for c in self.ctxs:
            c.close()
        self.batch_ctx.close()

    def results(self) -> list[dict]:
        info = self.info.cpu().tolist()
        res = self.resid.cpu().tolist()
        cond = self.cond.cpu().numpy()
        out = []
        for q, (m, k) in enumerate(self.problems):
            out.append({"mesh": m, "k": k, "n": self.nv[m], "info": info[q], "rel_residual": res[q],
                        "norm1_A": float(cond[q, 0]), "norm1_Ainv_est": float(cond[q, 1]),
                        "cond1_est": float(cond[q, 0] * cond[q, 1])})
        return out


# ---------------------------------------------------------------------------
# Autotuner of the persistent far-field launch shape
# ---------------------------------------------------------------------------
DEFAULT_CANDIDATES = ((1, 2), (1, 4), (2, 4), (2, 8), (4, 4), (4, 8), (8, 2), (8, 4), (8, 8), (16, 2), (16, 4),
                      (16, 8))


def default_cache_path() -> str:
    return os.environ.get("TAICHI_BEM_AMD_TUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "taichi_bem_amd", "autotune.json"))


def tuning_key(sweep: WavenumberSweep) -> str:
    name = torch.cuda.get_device_properties(sweep.dev).name
    sig = {}
    for m, _ in sweep.problems:
        key = f"nf{int(sweep.ctxs[m].nf)}"
        sig[key] = sig.get(key, 0) + 1
    return f"{name}|cf|{sweep.dtype}|prec{sweep.precision}|" + ",".join(f"{k}x{v}" for k, v in sorted(sig.items()))


def autotune(sweep: WavenumberSweep, candidates=DEFAULT_CANDIDATES, cache_path: str | None = None,
             reps: int = 1, use_cache: bool = True) -> dict:
    """Times the combined-field assembly of this batch for each candidate
    (test panels per block, blocks per CU), keeps the fastest in
    sweep.tuning and records it in a JSON cache keyed by device and batch
    signature (a cached entry is reused without timing)."""
    path = cache_path or default_cache_path()
    key = tuning_key(sweep)
    cache = {}
    if os.path.exists(path):
        try:
            with open(path) as fh:
                cache = json.load(fh)
        except (OSError, ValueError):
            cache = {}
    if use_cache and key in cache:
        sweep.tuning = tuple(cache[key]["tuning"])
        return dict(cache[key], cached=True)
    times = {}
    for cand in candidates:
        ts = []
        for _ in range(reps + 1):  # first run warms up
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            sweep.assemble(tuning=cand)
            e1.record()
            torch.cuda.synchronize()
            ts.append(e0.elapsed_time(e1))
        times[f"{cand[0]}x{cand[1]}"] = min(ts[1:])
    best = min(times, key=times.get)
    tuning = tuple(int(x) for x in best.split("x"))
    sweep.tuning = tuning
    entry = {"tuning": list(tuning), "ms": times[best], "candidates_ms": times, "time": time.time()}
    cache[key] = entry
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        with open(path, "w") as fh:
            json.dump(cache, fh, indent=1, sort_keys=True)
    except OSError:
        pass
    return dict(entry, cached=False)
