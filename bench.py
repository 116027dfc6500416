"""Benchmark of the MI355X BEM hot path on the mixed-precision Helmholtz
workload: interior Dirichlet problem, k = 3, seeded perturbed icosphere level
5 (20,480 panels, 10,242 vertices), plane-wave data, GaussQR 5, P0 Neumann
unknowns and P1 Dirichlet data.

One step = the FP64 time-to-solution pipeline on resident inputs: assemble V
(c128, far + near field) and the right-hand side 1/2 M conj(g) + K conj(g)
(matrix-free, fused with V), LU factorisation and solve.  `value` is BIO
entries/s of that assembly: (F^2 V entries + F*Nv K entries) / device time of
its launches.  The mixed-precision
pipelines (FP32 regular pairs + FP64 near field in c64, c64 LU, FP64
refinement) are timed after the steps and reported beside it with their error
against the FP64 solution.

Prints one JSON line.  --gpus above 1 is not supported (one GPU per box).
"""
from __future__ import annotations

import argparse
import importlib
import json
import os
import subprocess
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
# BEM_OPT_C64_LU_GEMM values the mixed pipelines take
LU_GEMM_NAMES = {0: "FP32 MFMA", 2: "fp16 MFMA on split operand planes (three part products, FP32 accumulation)"}
sys.path.insert(0, REPO)

PEAK_FP64_TFLOPS = 78.6   # MI355X FP64 vector (= FP64 matrix) peak
PEAK_FP32_TFLOPS = 157.3
# Algorithmic flops per quadrature-point pair of the fused far-field pass,
# which evaluates each unordered panel pair once (V is symmetric) and serves
# both orderings (add/mul = 1, FMA = 2; rsqrt and sin/cos are transcendentals,
# not counted), for the direct evaluation: d = x - y 3, r^2 5, w/r 1, w/r^3 2,
# (x-y).n_j 6, (y-x).n_i 6, kr 2, V accumulate 4, (cos + kr sin,
# sin - kr cos) 4, row-i K accumulate 4, row-j K times the test-point data
# 2 + 8.  The kernel's moment form does fewer operations for the same
# integrals; the count is kept so that kernel versions compare on one unit.
FLOPS_PER_EVAL_FUSED = 47
# HBM bytes per launch of the FP64 Dirichlet far-field kernel
# (far_dirichlet_f64_kernel<Helmholtz, 25>) on the L5 workload from
# rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, each scaled by the same counter on
# a 512 MiB copy (FETCH x 2.00, WRITE x 1.00), and the hardware FP64 VALU
# flops of that launch (SQ_INSTS_VALU_FLOPS_FP64 x 64 lanes).  Measured:
# 105.2 GB fetched + 8.4 GB written; 8.335e12 flops.
# Residual tolerance (relative to |b|) of the compliant mixed pipeline's
# FP64 refinement: two orders under the 1e-8 error bar of the metric.
REFINE_TOL = 1e-10
FAR_L5_TRAFFIC_BYTES = 113_612_336_928
FAR_L5_HW_FLOPS = 8_203_729_469_440


def cpu_threads() -> int:
    """The host cores this process may run on (its affinity mask), capped by
    OMP_NUM_THREADS where the job sets it: all the cores the job is allotted."""
    try:
        n = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        n = os.cpu_count() or 1
    omp = os.environ.get("OMP_NUM_THREADS", "")
    if omp.isdigit() and int(omp) > 0:
        n = min(n, int(omp))
    return max(1, n)


def cpu_host() -> dict:
    """nproc, the affinity mask and the CPU model of the host, for the CPU
    baseline's record."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        aff = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        aff = None
    return {"nproc": os.cpu_count(), "affinity_cores": aff, "omp_num_threads": os.environ.get("OMP_NUM_THREADS"),
            "cpu_model": model}


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--level", type=int, default=5)
    ap.add_argument("--k", type=float, default=3.0)
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="target CPU-baseline sample time")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--workload", choices=("dirichlet", "sweep", "transmission", "gradient"), default="dirichlet",
                    help="dirichlet: the L5 Helmholtz Dirichlet solve (default); sweep: the batched "
                         "combined-field wave-number sweep (60 k at L4 plus L3 and L5); transmission: the "
                         "L5 2x2 Calderon transmission solve with the interior-k gradient; gradient: the adjoint "
                         "shape and wave-number gradient of the Dirichlet misfit (L4, and L5 beside it)")
    ap.add_argument("--no-fd", action="store_true", help="transmission: skip the finite-difference check")
    ap.add_argument("--lu-lookahead", type=int, default=None,
                    help="compute units of the LU's panel stream (0: no look-ahead; default: the library's)")
    ap.add_argument("--c64-lu-gemm", type=int, default=2, choices=(0, 2),
                    help="trailing updates of the mixed pipelines' complex64 LU: 2 = fp16 MFMA on split operand "
                         "planes (default), 0 = FP32 MFMA")
    ap.add_argument("--no-other-configs", action="store_true",
                    help="dirichlet: skip the one-step runs of the sweep, transmission and gradient workloads "
                         "(child processes) whose summaries the line carries under other_configs")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="first run one plain device copy of this many MiB (the reference kernel that scales "
                         "FETCH_SIZE/WRITE_SIZE counters to bytes)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 (multi-GPU sharding is a "
                                   "design note only)"}))
        return 2

    import numpy as np
    import torch

    bem = importlib.import_module("taichi-bem_amd")
    L = bem._lib
    dev = torch.device("cuda", 0)
    if args.workload == "sweep":
        return sweep_main(args, bem)
    if args.workload == "transmission":
        return transmission_main(args, bem)
    if args.workload == "gradient":
        return gradient_main(args, bem)
    if args.calibrate_copy_mib > 0:
        n = args.calibrate_copy_mib * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copy calibration: {args.calibrate_copy_mib} MiB read + {args.calibrate_copy_mib} MiB written")
        del src, dst
    k = args.k
    v, f = bem.meshes.perturbed_icosphere(args.level, seed=0)
    d = bem.meshes.seeded_direction(2)
    nv, nf = len(v), len(f)
    log(f"mesh level {args.level}: nv={nv} nf={nf}")
    ctx = bem.Context(0)
    if args.lu_lookahead is not None:
        ctx.set_option(L.OPT_LU_LOOKAHEAD, args.lu_lookahead)
    ctx.set_mesh(torch.tensor(v, device=dev), torch.tensor(f, device=dev), 5)
    counts = ctx.nearfield_counts()
    g = torch.exp(1j * k * (torch.tensor(v, device=dev) @ torch.tensor(d, device=dev))).to(torch.complex128)

    def desc(op, dt, dr, prec, parts=0):
        return bem.Context.op_desc(op, 1, k, 1.0, dt, dr, prec, 0, 0, parts)  # kernel 1 = Helmholtz

    A64 = torch.zeros(nf, nf, dtype=torch.complex128, device=dev)
    LU64 = torch.empty_like(A64)
    b64 = torch.zeros(nf, dtype=torch.complex128, device=dev)
    ipiv = torch.empty(nf, dtype=torch.int32, device=dev)
    ev = lambda: torch.cuda.Event(enable_timing=True)  # noqa: E731

    def fp64_step():
        # A = V and b = 1/2 M conj(g) + K conj(g): far field in one fused pass
        # (V and K conj(g) share each point-pair evaluation), then the near field
        e = [ev() for _ in range(7)]
        e[0].record()
        A64.zero_()
        b64.zero_()
        e[6].record()
        ctx.assemble_dirichlet(desc(L.OP_V, 0, 0, L.PREC_FP64, L.PART_FAR), A64, g, b64, 1.0, 1.0, 0.5)
        e[1].record()
        ctx.assemble_dirichlet(desc(L.OP_V, 0, 0, L.PREC_FP64, L.PART_NEAR), A64, g, b64, 1.0, 1.0, 0.5)
        e[2].record()
        e[3].record()
        LU64.copy_(A64)
        ctx.lu_factor(LU64, ipiv)
        e[4].record()
        x = b64.clone()
        ctx.lu_solve(LU64, ipiv, x)
        e[5].record()
        torch.cuda.synchronize()
        t = [e[i].elapsed_time(e[i + 1]) for i in range(5)]
        # far_ms: the far-field launches alone (zeroing the matrix is zero_ms)
        return x, {"zero_ms": e[0].elapsed_time(e[6]), "far_ms": e[6].elapsed_time(e[1]), "near_ms": t[1],
                   "rhs_extra_ms": t[2], "lu_ms": t[3], "solve_ms": t[4], "total_ms": e[0].elapsed_time(e[5])}

    for w in range(args.warmup):
        log(f"warmup {w}")
        fp64_step()
    torch.cuda.synchronize()
    steps = []
    t0 = time.perf_counter()
    for s in range(args.steps):
        x64, tm = fp64_step()
        steps.append(tm)
        log(f"step {s}: " + ", ".join(f"{kk}={vv:.1f}" for kk, vv in tm.items()))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    mean = {kk: float(np.mean([s[kk] for s in steps])) for kk in steps[0]}
    asm_ms = mean["far_ms"] + mean["near_ms"]
    # entries produced per assembly: V (F x F, into the matrix) and K (F x Nv,
    # contracted with conj(g) on the fly for the right-hand side)
    entries = nf * nf + nf * nv
    value = entries / (asm_ms * 1e-3)

    # roofline of the dominant assembly kernel (far field, FP64 Helmholtz V)
    far_pairs = counts[0]
    far_flops = far_pairs // 2 * 625 * FLOPS_PER_EVAL_FUSED  # unordered pairs
    achieved = far_flops / (mean["far_ms"] * 1e-3) / 1e12
    roofline = {"bound": "valu", "kernel": "far_dirichlet_f64_kernel<Helmholtz> (V + K conj(g), moment form)",
                "achieved": round(achieved, 3), "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                "frac": round(achieved / PEAK_FP64_TFLOPS, 4),
                "traffic": FAR_L5_TRAFFIC_BYTES if args.level == 5 and FAR_L5_TRAFFIC_BYTES else None,
                "flops_per_launch": far_flops, "launch_ms": round(mean["far_ms"], 3)}
    if args.level == 5 and FAR_L5_HW_FLOPS:
        # hardware VALU rate (counters), beside the algorithmic one above
        roofline["hw_valu_tflops"] = round(FAR_L5_HW_FLOPS / (mean["far_ms"] * 1e-3) / 1e12, 3)
        roofline["hw_valu_frac"] = round(roofline["hw_valu_tflops"] / PEAK_FP64_TFLOPS, 4)

    result = {
        "metric": "BIO entries/s assembled; time-to-solution at 1e-8 rel err, mixed vs FP64 GPU",
        "value": value, "unit": "entries/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": mean["total_ms"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "c128", "data": "synthetic (seeded perturbed icosphere, plane wave)",
        "config": {"workload": f"Helmholtz interior Dirichlet k={k}, perturbed icosphere L{args.level} "
                               f"({nf} panels, {nv} vertices), GaussQR 5, P0/P1",
                   "panels": nf, "vertices": nv, "near_pairs": counts[1:], "far_pairs": far_pairs},
        "stages_ms_fp64": {kk: round(vv, 3) for kk, vv in mean.items()},
        # V alone: the entries stored in the matrix (K conj(g) is contracted
        # on the fly, as the reference's matrix-free apply_K_dot_D_boundary)
        "value_V_only": nf * nf / (asm_ms * 1e-3),
        "roofline": roofline,
        "wall_s_steps": round(wall, 3),
    }

    if not args.no_mixed:
        tts = {"fp64_ms": round(mean["total_ms"], 3)}
        lu_gemm_name = LU_GEMM_NAMES[args.c64_lu_gemm]
        x64n = torch.linalg.vector_norm(x64).item()
        # mixed (fast): FP32 regular pairs + FP64 near field into c64, c64 LU,
        # refinement against that c64 matrix with FP64 residuals
        A32 = torch.zeros(nf, nf, dtype=torch.complex64, device=dev)
        LU32 = torch.empty_like(A32)
        b = torch.zeros(nf, dtype=torch.complex128, device=dev)
        x = torch.zeros(nf, dtype=torch.complex128, device=dev)
        for rep in range(2):
            e = [ev() for _ in range(6)]
            e[0].record()
            A32.zero_()
            b.zero_()
            ctx.assemble_dirichlet(desc(L.OP_V, 0, 0, L.PREC_FP32, L.PART_FAR), A32, g, b, 1.0, 1.0, 0.5)
            e[1].record()
            ctx.assemble_dirichlet(desc(L.OP_V, 0, 0, L.PREC_FP64, L.PART_NEAR), A32, g, b, 1.0, 1.0, 0.5)
            e[2].record()
            LU32.copy_(A32)
            ctx.set_option(L.OPT_LU_PIVOTING, 2)  # block-local pivoting (the factor preconditions refinement)
            ctx.set_option(L.OPT_C64_LU_GEMM, args.c64_lu_gemm)
            ctx.lu_factor(LU32, ipiv)
            ctx.set_option(L.OPT_LU_PIVOTING, 0)
            ctx.set_option(L.OPT_C64_LU_GEMM, 0)
            e[3].record()
            st = ctx.refine_solve(A32, LU32, ipiv, b, x, max_iter=30, tol=1e-13)
            e[4].record()
            torch.cuda.synchronize()
        err = (torch.linalg.vector_norm(x - x64).item()) / x64n
        tts["mixed_fp32asm"] = {"ms": round(e[0].elapsed_time(e[4]), 3), "far_fp32_ms": round(e[0].elapsed_time(e[1]), 3),
                                "assembly_ms": round(e[0].elapsed_time(e[2]), 3), "lu_c64_ms": round(e[2].elapsed_time(e[3]), 3),
                                "refine_ms": round(e[3].elapsed_time(e[4]), 3), "refine_iters": st.iterations,
                                "lu_pivoting": "block-local", "lu_gemm": lu_gemm_name, "rel_err_vs_fp64": err,
                                "assembly_entries_per_s": entries / (e[0].elapsed_time(e[2]) * 1e-3)}
        # mixed (accurate): FP64 assembly, c64 LU, FP64 refinement against A64
        for rep in range(2):
            e = [ev() for _ in range(5)]
            e[0].record()
            A64.zero_()
            b64.zero_()
            ctx.assemble_dirichlet(desc(L.OP_V, 0, 0, L.PREC_FP64), A64, g, b64, 1.0, 1.0, 0.5)
            e[1].record()
            bem.convert(A64, LU32)
            ctx.set_option(L.OPT_LU_PIVOTING, 2)  # block-local pivoting (the factor preconditions refinement)
            ctx.set_option(L.OPT_C64_LU_GEMM, args.c64_lu_gemm)
            ctx.lu_factor(LU32, ipiv)
            ctx.set_option(L.OPT_LU_PIVOTING, 0)
            ctx.set_option(L.OPT_C64_LU_GEMM, 0)
            e[2].record()
            # the metric's bar is a 1e-8 relative error: refinement stops at a
            # residual of 1e-10 |b| (the error/residual ratio of this system
            # measures about 34; one step reaches a 2.5e-13 residual and a
            # 8.6e-12 error, a second 4e-14); rel_err_vs_fp64 reports the error
            st = ctx.refine_solve(A64, LU32, ipiv, b64, x, max_iter=30, tol=REFINE_TOL)
            e[3].record()
            torch.cuda.synchronize()
        err2 = (torch.linalg.vector_norm(x - x64).item()) / x64n
        tts["mixed_fp64asm"] = {"ms": round(e[0].elapsed_time(e[3]), 3), "assembly_ms": round(e[0].elapsed_time(e[1]), 3),
                                "lu_c64_ms": round(e[1].elapsed_time(e[2]), 3), "refine_ms": round(e[2].elapsed_time(e[3]), 3),
                                "refine_iters": st.iterations, "converged": st.converged, "refine_tol": REFINE_TOL,
                                "lu_pivoting": "block-local", "lu_gemm": lu_gemm_name, "rel_err_vs_fp64": err2}
        tts["speedup_mixed_fp32asm"] = round(mean["total_ms"] / tts["mixed_fp32asm"]["ms"], 3)
        tts["speedup_mixed_fp64asm"] = round(mean["total_ms"] / tts["mixed_fp64asm"]["ms"], 3)
        result["time_to_solution"] = tts
        del A32, LU32
        log("mixed: " + json.dumps(tts))

    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(REPO, "oracle"))
        import oracle as O

        threads = cpu_threads()
        opv = O.make_op(O.OP_V, O.HELMHOLTZ, k, 1.0, 5, 0, 0, 0)
        opk = O.make_op(O.OP_K, O.HELMHOLTZ, k, 1.0, 5, 0, 1, 0)
        geom = O.geometry(v, f)
        gh = g.cpu().numpy()

        def cpu_rows(r):
            O.assemble_rows(opv, v, f, 0, r, nthreads=threads, geom=geom)
            O.apply_rows(opk, v, f, gh, 0, r, True, 1.0, nthreads=threads, geom=geom)

        rows = 2 * threads  # calibration sample: two test panels per thread
        t = time.perf_counter()
        cpu_rows(rows)
        dt = time.perf_counter() - t
        rows2 = max(rows, min(nf, int(rows * args.cpu_seconds / max(dt, 1e-3))))
        t = time.perf_counter()
        cpu_rows(rows2)
        dt2 = time.perf_counter() - t
        result["cpu_baseline"] = {"value": rows2 * (nf + nv) / dt2, "unit": "entries/s", "cores": threads,
                                  "kind": "port",
                                  "sample": f"test panels 0..{rows2 - 1} of {nf}: V Helmholtz row (all {nf} "
                                            f"columns) + K conj(g) row ({nv} vertex columns), FP64, GaussQR 5, "
                                            f"{dt2:.1f} s, C/OpenMP oracle on {threads} threads",
                                  "host": cpu_host()}
        result["cpu_baseline"]["gpu_over_cpu"] = value / result["cpu_baseline"]["value"]
        log(f"cpu baseline: {rows2} rows in {dt2:.1f}s")

    # release the library's streams and device memory before the other
    # configurations run in their own processes
    ctx.close()
    if not args.no_other_configs:
        result["other_configs"] = other_configs()
    print(json.dumps(result), flush=True)
    return 0


# Keys of the other workloads' lines that the Dirichlet line repeats.
OTHER_KEYS = ("metric", "value", "unit", "ms_per_step", "dtype", "config", "stages_ms_eager", "graph_replay_ms",
              "roofline", "max_rel_residual", "rel_err_vs_fp64", "cond1_est_L3_L5", "fp64",
              "mixed_c64lu_fp64asm", "mixed_fp32asm", "speedup_fp32asm", "gradient", "levels")


def other_configs() -> dict:
    """One warm step of each other workload (the batched sweep, the
    transmission solve with its interior-k gradient, the adjoint shape
    gradient), each in a child process of its own, so that one bench run
    records every configuration; a child's failure is recorded, not raised."""
    runs = {"sweep": ["--steps", "1", "--warmup", "1"],
            "transmission": ["--steps", "1", "--warmup", "1", "--no-fd"],
            "gradient": ["--steps", "1", "--warmup", "1"]}
    out = {}
    for w, extra in runs.items():
        cmd = [sys.executable, "-u", os.path.abspath(__file__), "--workload", w, "--no-cpu-baseline"] + extra
        log(f"other configuration: {w}")
        t = time.perf_counter()
        try:
            p = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
        except subprocess.TimeoutExpired:
            out[w] = {"error": "timed out after 300 s"}
            continue
        lines = [ln for ln in p.stdout.splitlines() if ln.startswith("{")]
        if p.returncode != 0 or not lines:
            out[w] = {"error": f"exit {p.returncode}", "stderr_tail": p.stderr[-400:]}
            continue
        d = json.loads(lines[-1])
        out[w] = {k: d[k] for k in OTHER_KEYS if k in d}
        out[w]["wall_s"] = round(time.perf_counter() - t, 1)
    return out


# Algorithmic flops per point pair of the combined-field far kernel
# (cf_far_kernel, FP32): d 3, r^2 5, w/r 1, (x-y).n 5, w (x-y).n / r^3 3,
# kr 2, phase reduction 2, eta w/r 1, K - i eta V value 8, accumulation of
# the combination into the three test bases 12.
FLOPS_PER_EVAL_CF = 42


def sweep_main(args, bem) -> int:
    """Batched exterior combined-field sweep: 60 wave numbers k = 16 + i/30
    on icosphere L4 plus L3 and L5 at k = 16.5, c64 with FP32 regular pairs,
    FP64 near field and FP64 refinement, condition estimates per k.  One step
    = one replay of the captured graph (assembly, LU, refined solves,
    condition estimates of all 62 problems)."""
    import numpy as np
    import torch

    L = bem._lib
    S = bem.sweep
    meshes = [bem.meshes.icosphere(4), bem.meshes.icosphere(3), bem.meshes.icosphere(5)]
    problems = [(0, 16.0 + i / 30.0) for i in range(60)] + [(1, 16.5), (2, 16.5)]
    log(f"sweep: {len(problems)} problems, vertices {[len(m[0]) for m in meshes]}")
    sw = S.WavenumberSweep(meshes, problems, dtype=torch.complex64, precision=L.PREC_FP32, refine_iters=2,
                           condest_iters=3, direction=bem.meshes.seeded_direction(2))
    tune = S.autotune(sw)
    log(f"tuning {sw.tuning} ({'cached' if tune.get('cached') else 'timed'}): {tune.get('candidates_ms')}")
    ev = lambda: torch.cuda.Event(enable_timing=True)  # noqa: E731
    # eager stage timing
    stages = {}
    for w in range(2):
        e = [ev() for _ in range(6)]
        e[0].record()
        sw.assemble()
        e[1].record()
        L.convert(sw.A_flat, sw.LU_flat)
        sw.batch.lu_factor(sw.LU, sw.ipiv, sw.info)
        e[2].record()
        sw.batch.solve(sw.LU, sw.ipiv, sw.A, sw.b, sw.x, sw.refine_iters, sw.resid)
        e[3].record()
        sw.batch.condest(sw.LU, sw.ipiv, sw.A, sw.cond, sw.condest_iters)
        e[4].record()
        torch.cuda.synchronize()
        stages = {"assembly_ms": e[0].elapsed_time(e[1]), "lu_ms": e[1].elapsed_time(e[2]),
                  "solve_refine_ms": e[2].elapsed_time(e[3]), "condest_ms": e[3].elapsed_time(e[4])}
        log(f"eager {w}: " + ", ".join(f"{kk}={vv:.1f}" for kk, vv in stages.items()))
    sw.capture()
    for _ in range(args.warmup):
        sw.replay()
    torch.cuda.synchronize()
    e0, e1 = ev(), ev()
    t0 = time.perf_counter()
    e0.record()
    for _ in range(args.steps):
        sw.replay()
    e1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = e0.elapsed_time(e1) / args.steps
    res = sw.results()
    entries = sum(2 * r["n"] ** 2 for r in res)  # V and K entries of every problem
    far_pairs = 0
    for m, _ in problems:
        c = sw.ctxs[m]
        far_pairs += c.nf * c.nf - sum(c.nearfield_counts()[1:])
    # cf_far_kernel evaluates each unordered panel pair once for both orders
    far_flops = far_pairs // 2 * 625 * FLOPS_PER_EVAL_CF
    # accuracy: the k = 16.5 problems against all-FP64 complex128 solves
    errs = {}
    for q in (len(problems) - 2, len(problems) - 1, 15):
        m, k = problems[q]
        ref = S.WavenumberSweep([meshes[m]], [(0, k)], dtype=torch.complex128, precision=L.PREC_FP64,
                                refine_iters=0, condest_iters=1, direction=bem.meshes.seeded_direction(2))
        ref.run()
        torch.cuda.synchronize()
        errs[f"L{[4, 3, 5][m]} k={k:.4f}"] = float(torch.linalg.vector_norm(sw.x[q] - ref.x[0]) /
                                                   torch.linalg.vector_norm(ref.x[0]))
        del ref
    # CPU baseline: the oracle assembling V and K (P1 x P1) of the L3 problem
    cpu = None
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(REPO, "oracle"))
        import oracle as O

        threads = cpu_threads()
        v3, f3 = meshes[1]
        t = time.perf_counter()
        for op in (O.OP_V, O.OP_K):
            O.assemble(O.make_op(op, O.HELMHOLTZ, 16.5, 1.0, 5, 1, 1, 0), v3, f3, nthreads=threads)
        dt = time.perf_counter() - t
        cpu = {"value": 2 * len(v3) ** 2 / dt, "unit": "entries/s", "cores": threads, "kind": "port",
               "sample": f"V and K (P1 x P1, Helmholtz k=16.5, GaussQR 5, FP64) of icosphere L3 "
                         f"({len(f3)} panels), {dt:.1f} s, C/OpenMP oracle on {threads} threads",
               "host": cpu_host()}
    asm_s = stages["assembly_ms"] * 1e-3
    result = {
        "metric": "BIO entries/s assembled; time-to-solution of the batched combined-field sweep",
        "value": entries / asm_s, "unit": "entries/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "c64",
        "data": "synthetic (icospheres, seeded plane wave)",
        "config": {"workload": "exterior combined-field sweep: 60 k in [16, 18) step 1/30 on icosphere L4 "
                               "(5120 panels, 2562 vertices) + L3 and L5 at k=16.5, P1, GaussQR 5, c64 with "
                               "FP32 regular pairs, FP64 near field, 2 FP64 refinement steps",
                   "problems": len(problems), "tuning": list(sw.tuning)},
        "stages_ms_eager": {kk: round(vv, 3) for kk, vv in stages.items()},
        "graph_replay_ms": round(ms, 3),
        "roofline": {"bound": "valu", "kernel": "cf_far_kernel<float> (V + K, P1 x P1)",
                     "achieved": round(far_flops / asm_s / 1e12, 3), "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(far_flops / asm_s / 1e12 / PEAK_FP32_TFLOPS, 4), "traffic": None,
                     "note": "unordered far point pairs x 42 flops over the whole assembly time (far + near "
                             "+ mass launches), so a lower bound of the far kernel's own rate"},
        "max_rel_residual": max(r["rel_residual"] for r in res),
        "rel_err_vs_fp64": errs,
        "cond1_est_per_k": [round(r["cond1_est"], 3) for r in res[:60]],
        "cond1_est_L3_L5": [round(r["cond1_est"], 3) for r in res[60:]],
        "wall_s_steps": round(wall, 3),
    }
    if cpu:
        result["cpu_baseline"] = cpu
    sw.close()
    print(json.dumps(result), flush=True)
    return 0


def gradient_main(args, bem) -> int:
    """Adjoint gradient of the Dirichlet misfit J = ||t - t*||^2_M with
    respect to the vertex positions and k (Helmholtz k = 2, point-source data
    g = exp(i k r)/r from (2, 0.5, 0.3), target trace t* of a seeded
    perturbed sphere), through the torch.autograd.Function NeumannTrace, on
    icosphere L4 (2562 vertices, the config) and L5.  One step = the forward
    solve (fused V + K g assembly, LU, solve) plus the backward (transposed
    solve, the adjoint contractions of dV/dX, dK/dX and dOp/dk, the area
    term).  The two contraction launches are also timed alone.  A central
    difference along one seeded direction checks the L4 gradient."""
    import numpy as np
    import torch

    A = bem.adjoint
    L = bem._lib
    dev = torch.device("cuda", 0)
    k, x0 = 2.0, (2.0, 0.5, 0.3)
    ev = lambda: torch.cuda.Event(enable_timing=True)  # noqa: E731
    out = {}
    for level in (4, 5):
        v, f = bem.meshes.icosphere(level)
        vt, _ = bem.meshes.perturbed_icosphere(level, seed=1, degrees=(3,))
        prob = A.DirichletProblem(f, kernel=1, gauss_qr=5)
        kt = torch.tensor(k, dtype=torch.float64, device=dev)
        vt_t = torch.tensor(vt, device=dev)
        with torch.no_grad():
            t_target = A.neumann_trace(vt_t, kt, A.point_source(vt_t, kt, x0, 1), prob)

        def J(vv, kk, grad):
            vv = torch.as_tensor(vv, device=dev).clone().requires_grad_(grad)
            kk = torch.tensor(float(kk), dtype=torch.float64, device=dev, requires_grad=grad)
            e = [ev() for _ in range(3)]
            e[0].record()
            t = A.neumann_trace(vv, kk, A.point_source(vv, kk, x0, 1), prob)
            j = A.misfit(t, t_target, vv, prob.panels)
            e[1].record()
            if grad:
                j.backward()
            e[2].record()
            torch.cuda.synchronize()
            tm = (e[0].elapsed_time(e[1]), e[1].elapsed_time(e[2]))
            if not grad:
                return float(j.detach()), tm
            return float(j.detach()), tm, vv.grad, kk.grad

        for _ in range(max(1, args.warmup)):
            J(v, k, True)
        fw, bw = [], []
        for _ in range(args.steps):
            j0, tm, gx, gk = J(v, k, True)
            fw.append(tm[0])
            bw.append(tm[1])
        # the two adjoint contraction launches alone (V with t, K with g)
        nf = len(f)
        C = prob.ctx
        prob.set_vertices(torch.tensor(v, device=dev))
        lam = torch.randn(nf, dtype=torch.complex128, device=dev)
        u_t = torch.randn(nf, dtype=torch.complex128, device=dev)
        u_g = torch.randn(len(v), dtype=torch.complex128, device=dev)
        grad = torch.zeros(len(v), 3, dtype=torch.float64, device=dev)
        grad_k = torch.zeros(1, dtype=torch.float64, device=dev)
        ts = []
        for rep in range(3):
            e = [ev() for _ in range(3)]
            e[0].record()
            C.shape_gradient(prob.desc(L.OP_V, k, 0), lam, u_t, grad, grad_k, alpha=-1.0)
            e[1].record()
            C.shape_gradient(prob.desc(L.OP_K, k, 1), lam, u_g, grad, grad_k, alpha=1.0)
            e[2].record()
            torch.cuda.synchronize()
            ts.append((e[0].elapsed_time(e[1]), e[1].elapsed_time(e[2])))
        rec = {"vertices": len(v), "panels": nf, "forward_ms": round(float(np.median(fw)), 3),
               "backward_ms": round(float(np.median(bw)), 3),
               "adjoint_V_ms": round(float(np.median([t[0] for t in ts[1:]])), 3),
               "adjoint_K_ms": round(float(np.median([t[1] for t in ts[1:]])), 3),
               "separate_ordered_panel_pairs": int(C.nearfield_counts()[0])}
        if level == 4:
            rng = np.random.default_rng(3)
            d = rng.standard_normal(v.shape)
            d /= np.linalg.norm(d)
            h = 1e-6
            fd = (J(v + h * d, k, False)[0] - J(v - h * d, k, False)[0]) / (2 * h)
            got = float((gx.cpu().numpy() * d).sum())
            rec["fd_check"] = {"adjoint": got, "central_difference": fd, "rel_err": abs(got - fd) / abs(fd)}
            rec["dJ_dk"] = float(gk)
        out[f"L{level}"] = rec
        log(f"gradient L{level}: {rec}")
        prob.ctx.close()
    l4 = out["L4"]
    line = {
        "metric": "adjoint shape + wave-number gradient of the Dirichlet misfit (forward + backward)",
        "value": round(l4["forward_ms"] + l4["backward_ms"], 3), "unit": "ms", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(l4["forward_ms"] + l4["backward_ms"], 3),
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "c128",
        "data": "synthetic (icosphere, seeded perturbed target, point source)",
        "config": {"workload": "Helmholtz k=2 interior Dirichlet misfit gradient, icosphere L4 (2562 vertices, "
                               "5120 panels), GaussQR 5; L5 beside it"},
        "levels": out,
    }
    print(json.dumps(line))
    return 0


def transmission_main(args, bem) -> int:
    """Helmholtz transmission problem: 2 x 2 Calderon block system of V, K, K^T,
    W at exterior k = 5 and interior k = 5 sqrt(2.25) on the seeded perturbed
    icosphere L5 (P1 Dirichlet + P0 Neumann unknowns, n = 30722), incident
    plane wave, GaussQR 5, through the reference-shaped CoreManager /
    BEMManager API.  FP64 pipeline (c128) against the mixed ones (c64 LU with
    FP64 refinement against the c128 matrix; FP32 regular-pair assembly in c64
    with refinement against that matrix), and the adjoint gradient of
    J = ||x - x*||^2 with respect to the interior wave number (incident data
    held fixed), checked by central differences."""
    import numpy as np
    import torch

    L = bem._lib
    dev = torch.device("cuda", 0)
    k, ni = 5.0, 2.25
    v, f = bem.meshes.perturbed_icosphere(args.level, seed=0)
    d = torch.tensor(bem.meshes.seeded_direction(2), device=dev)
    nv, nf = len(v), len(f)

    def fD(x, sqrt_n):
        return torch.exp(1j * k * sqrt_n * (x @ d))

    def fN(x, n, sqrt_n):
        return 1j * k * sqrt_n * (n @ d) * torch.exp(1j * k * sqrt_n * (x @ d))

    def manager(prec, n_i=ni):
        params = {"kernel": "Helmholtz_Transmission", "k": k, "n_i": n_i, "n_o": 1.0, "boundary": "Full",
                  "GaussQR": 5, "Q_Neumann": 0, "Q_Dirichlet": 1, "vertices": v, "panels": f, "precision": prec}
        cm = bem.CoreManager(params)
        cm.initialization(fD, fN)
        return cm

    ev = lambda: torch.cuda.Event(enable_timing=True)  # noqa: E731
    ST = bem.ScopeType
    AT = bem.AssembleType

    def assemble(bm):
        bm._mat_A.zero_()
        bm.matrix_layer_forward(k=bm._k, sqrt_n=bm._sqrt_no, scope_type=int(ST.EXTERIOR))
        bm.assemble_matA(assemble_type=int(AT.ADD_P_MINUS), multiplier=1.0)
        bm.rhs_constructor.set_boundaries(scope_type=int(ST.INTERIOR), sqrt_ni=bm._sqrt_ni, sqrt_no=bm._sqrt_no)
        bm.assemble_rhs()
        bm.matrix_layer_forward(k=bm._k, sqrt_n=bm._sqrt_ni, scope_type=int(ST.INTERIOR))
        bm.assemble_matA(assemble_type=int(AT.ADD_P_PLUS), multiplier=1.0)

    n = nf + nv
    log(f"transmission L{args.level}: n = {n}")
    cm = manager("fp64")
    bm = cm._BEM_manager
    times = []
    for rep in range(1 + max(1, args.steps)):
        e = [ev() for _ in range(3)]
        e[0].record()
        assemble(bm)
        e[1].record()
        bm.solve()
        e[2].record()
        torch.cuda.synchronize()
        times.append((e[0].elapsed_time(e[1]), e[1].elapsed_time(e[2])))
        log(f"fp64 run {rep}: assembly {times[-1][0]:.1f} ms, LU + solve {times[-1][1]:.1f} ms")
    asm64, sol64 = (float(np.mean([t[q] for t in times[1:]])) for q in (0, 1))
    x64 = bm.raw_solved.clone()
    b = bm.rhs[:n].clone()
    xn = torch.linalg.vector_norm(x64).item()
    # mixed A: c64 LU of the FP64 matrix, FP64 refinement against it
    A64 = bm._mat_A
    LU32 = torch.empty(n, n, dtype=torch.complex64, device=dev)
    ipiv = torch.empty(n, dtype=torch.int32, device=dev)
    x = torch.zeros(n, dtype=torch.complex128, device=dev)
    for rep in range(2):
        e = [ev() for _ in range(3)]
        e[0].record()
        # partial pivoting: block-local pivoting (BEM_OPT_LU_PIVOTING = 2) does
        # not give a converging refinement on this system (kappa about 1e6)
        L.convert(A64, LU32)
        bm.ctx.set_option(L.OPT_C64_LU_GEMM, args.c64_lu_gemm)
        bm.ctx.lu_factor(LU32, ipiv)
        bm.ctx.set_option(L.OPT_C64_LU_GEMM, 0)
        st = bm.ctx.refine_solve(A64, LU32, ipiv, b, x, max_iter=30, tol=1e-14)
        e[2].record()
        torch.cuda.synchronize()
    mixA = {"lu_and_refine_ms": e[0].elapsed_time(e[2]), "lu_pivoting": "partial",
            "lu_gemm": LU_GEMM_NAMES[args.c64_lu_gemm],
            "refine_iters": int(st.iterations), "converged": int(st.converged),
            "rel_err_vs_fp64": torch.linalg.vector_norm(x - x64).item() / xn}
    mixA["ms"] = asm64 + mixA["lu_and_refine_ms"]
    del LU32
    # gradient with respect to the interior wave number
    rng = np.random.default_rng(9)
    xs = torch.tensor(rng.standard_normal(n) + 1j * rng.standard_normal(n), device=dev) * x64.abs().mean()
    e = [ev() for _ in range(2)]
    e[0].record()
    gk = bem.adjoint.transmission_k_gradient(bm, 2 * (x64 - xs))
    e[1].record()
    torch.cuda.synchronize()
    grad = {"dJ_dki": gk, "ms": e[0].elapsed_time(e[1])}
    cm.kill()
    del bm, cm, A64
    torch.cuda.empty_cache()
    # mixed B: FP32 regular pairs assembled into c64, c64 LU, refinement against it
    cm32 = manager("fp32")
    b32 = cm32._BEM_manager
    for rep in range(2):
        e = [ev() for _ in range(4)]
        e[0].record()
        assemble(b32)
        e[1].record()
        LU = b32._mat_A.clone()
        b32.ctx.set_option(L.OPT_C64_LU_GEMM, args.c64_lu_gemm)
        b32.ctx.lu_factor(LU, ipiv)
        b32.ctx.set_option(L.OPT_C64_LU_GEMM, 0)
        st = b32.ctx.refine_solve(b32._mat_A, LU, ipiv, b, x, max_iter=30, tol=1e-12)
        e[3].record()
        torch.cuda.synchronize()
        del LU
    mixB = {"assembly_ms": e[0].elapsed_time(e[1]), "lu_and_refine_ms": e[1].elapsed_time(e[3]),
            "lu_pivoting": "partial", "lu_gemm": LU_GEMM_NAMES[args.c64_lu_gemm], "refine_iters": int(st.iterations),
            "rel_err_vs_fp64": torch.linalg.vector_norm(x - x64).item() / xn}
    mixB["ms"] = mixB["assembly_ms"] + mixB["lu_and_refine_ms"]
    cm32.kill()
    del b32, cm32
    torch.cuda.empty_cache()
    if not args.no_fd:
        h = 1e-6
        js = []
        for sgn in (1, -1):
            ki = k * np.sqrt(ni) + sgn * h
            c2 = manager("fp64", (ki / k) ** 2)
            b2 = c2._BEM_manager
            assemble(b2)
            b2.rhs[:n] = b
            xx = b2.solve()
            js.append(float(((xx - xs).abs() ** 2).sum()))
            c2.kill()
            del b2, c2, xx
            torch.cuda.empty_cache()
        fd = (js[0] - js[1]) / (2 * h)
        grad["fd"] = fd
        grad["rel_err_vs_fd"] = abs(gk - fd) / abs(fd)
    tts64 = asm64 + sol64
    result = {
        "metric": "time-to-solution of the transmission system, mixed vs FP64 GPU; dJ/dk_i by adjoint",
        "value": tts64 / mixA["ms"], "unit": "x (FP64 time / mixed time, c64 LU + FP64 refinement)",
        "n_gpus": 1, "steps": args.steps, "warmup": 1, "ms_per_step": tts64, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "c128", "data": "synthetic (seeded perturbed icosphere, "
        "plane wave)",
        "config": {"workload": f"Helmholtz transmission k_o=5, k_i=7.5, perturbed icosphere L{args.level} "
                               f"({nf} panels, {nv} vertices), n={n}, GaussQR 5, P1 Dirichlet + P0 Neumann"},
        "fp64": {"assembly_ms": round(asm64, 3), "lu_solve_ms": round(sol64, 3), "ms": round(tts64, 3)},
        "mixed_c64lu_fp64asm": mixA, "mixed_fp32asm": mixB, "speedup_fp32asm": tts64 / mixB["ms"],
        "gradient": grad,
    }
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
