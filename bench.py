"""Benchmark: Newton time steps per second of the FEM soft-body step on one MI355X (FP64).

Workload: cantilever bar of 32×32×128 Freudenthal cubes (140,481 vertices, 786,432 tets), E = 1e5 Pa, ν = 0.45,
ρ = 1000 kg/m³, x = 0 face clamped, gravity −9.8 m/s² in y, seeded jitter ±0.05·h, dt = 1/60 s, Newton ≤ 10
iterations with Jacobi-PCG rtol 1e-8 and backtracking Armijo (c = 0.3).  The Newton stopping rule is the
reference's ‖g‖₂ < threshold × total mass (time_step.hpp:482-492) with threshold 1e-6: at 786k tets the
reference drivers' 1e-3 (tests/fem_cu/main.cu:60) stops most steps at the predictor, 1e-6 makes every step
converge the nonlinear solve (about 2 Newton iterations and 1,700 PCG iterations per step).  All inputs are resident in HBM when the timed
region starts; the timed region is K whole steps (prepare + Newton solve + step_next).

Prints one JSON line.  ``roofline`` is for the dominant kernel (the PCG SpMV), whose average duration is
measured with hipEvents inside the timed region; ``cpu_baseline`` is the CPU FP64 oracle timed on a bounded
sample of the same workload on this host (see ``cpu_sample`` below).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Newton steps/s @786k tets FP64 (plain vs batched/adjoint) + HBM GB/s % peak"
HBM_PEAK_GBS = 8000.0


def build_case(nx, ny, nz, seed):
    from ssim_amd import mesh

    return mesh.cantilever_bar(nx, ny, nz, h=0.01, E=1e5, nu=0.45, rho=1000.0, dt=1.0 / 60, seed=seed)


def spmv_algorithmic_bytes(info):
    """Bytes one launch of the PCG SpMV kernel must move, for the format it reads (symmetric BSR3: only blocks
    with row <= column are stored).  76 B per stored block (72 B of values + a 4 B column index), 8 B per
    lower-block reference (column + value offset), the gathered input z once (n·8 B), and the lane's own rows
    of p and q read and written (4·n·8 B).  Padding slots are not counted."""
    n = 3 * info["nv"]
    upper = info["upper_blocks"]
    lower = info["nonzero_blocks"] - upper
    return 76 * upper + 8 * lower + 5 * n * 8


def load_traffic():
    path = os.path.join(ROOT, "profiles", "k_pcg_spmv_786k_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("hbm_bytes_per_launch")


def cpu_sample(case, counts, threads, rtol):
    """Times the oracle's phases on the same mesh (one Hessian assembly, two energy+gradient evaluations, 20 PCG
    iterations) and composes the CPU time of the GPU-timed steps from their Newton / PCG / evaluation counts."""
    os.environ.setdefault("OMP_NUM_THREADS", str(threads))
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    ts = oracle.TimeStep.from_case(case)
    ts.set_threshold(rtol)
    ts.prepare_step()
    u = case.u0.copy()
    t0 = time.perf_counter()
    E, g = ts.energy_grad(u)
    E, g = ts.energy_grad(u)
    t_eval = (time.perf_counter() - t0) / 2
    t0 = time.perf_counter()
    vals = ts.hessian(u, True, True)
    t_hess = time.perf_counter() - t0
    rowptr, colidx = ts.csr_pattern()
    t0 = time.perf_counter()
    _, it, _ = oracle.pcg(rowptr, colidx, vals, -g, rtol=0.0, maxit=20)
    t_pcg = (time.perf_counter() - t0) / max(it, 1)
    steps = counts["steps"]
    total = (counts["energy_evals"] * t_eval + counts["hessian_evals"] * t_hess + counts["pcg_iterations"] * t_pcg)
    per_step = total / steps
    return {
        "value": 1.0 / per_step,
        "unit": "steps/s",
        "cores": oracle.lib().orc_num_threads(),
        "kind": "port",
        "sample": (f"oracle phases on the same 786k-tet mesh: energy+grad {t_eval * 1e3:.1f} ms, SPD Hessian assembly "
                   f"{t_hess * 1e3:.1f} ms, PCG iteration {t_pcg * 1e3:.2f} ms; composed with the timed GPU steps' "
                   f"counts ({counts['energy_evals']} evals, {counts['hessian_evals']} Hessians, "
                   f"{counts['pcg_iterations']} PCG iterations over {steps} steps)"),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--rtol", type=float, default=1e-6, help="Newton threshold (× total mass)")
    ap.add_argument("--cells", type=str, default="128,32,32", help="cubes along x,y,z")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=16)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only: this build's time step runs on a single device "
              f"(--gpus {args.gpus} requested)", file=sys.stderr)
        sys.exit(2)

    import torch

    from ssim_amd import fem

    nx, ny, nz = (int(v) for v in args.cells.split(","))
    case = build_case(nx, ny, nz, args.seed)
    print(f"[bench] mesh {nx}x{ny}x{nz}: nv={case.nv} nt={case.nt}", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    solver = fem.TetFiniteElementSolver_Hip.from_case(case, rtol=args.rtol)
    print(f"[bench] setup {time.perf_counter() - t0:.2f} s", file=sys.stderr, flush=True)
    info = solver.layout_info()
    for k in range(args.warmup):
        solver.step()
        print(f"[bench] warmup step {k}: {solver.last_stats()}", file=sys.stderr, flush=True)
    solver.set_kernel_timing(True)
    counts = dict(steps=args.steps, newton_iterations=0, pcg_iterations=0, energy_evals=0, hessian_evals=0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        solver.step()
        st = solver.last_stats()
        for key in ("newton_iterations", "pcg_iterations", "energy_evals", "hessian_evals"):
            counts[key] += st[key]
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ktimes = solver.kernel_timing()
    solver.set_kernel_timing(False)
    print(f"[bench] timed {args.steps} steps in {elapsed:.3f} s; counts {counts}", file=sys.stderr, flush=True)
    print(f"[bench] kernel timing (avg ms, samples): {ktimes}", file=sys.stderr, flush=True)

    spmv_ms, spmv_n = ktimes["pcg_spmv"]
    bytes_per_launch = spmv_algorithmic_bytes(info)
    achieved = bytes_per_launch / (spmv_ms * 1e-3) / 1e9 if spmv_ms > 0 else None
    traffic = load_traffic()
    roofline = {
        "bound": "hbm",
        "achieved": achieved,
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": (achieved / HBM_PEAK_GBS) if achieved else None,
        "traffic": traffic,
        "kernel": "k_pcg_spmv (symmetric blocked-ELL 3x3 SpMV, p/q update, p·q partials)",
        "bytes_per_launch": bytes_per_launch,
        "avg_launch_ms": spmv_ms,
        "sampled_launches": spmv_n,
    }
    cpu = None
    if not args.no_cpu:
        print("[bench] timing the CPU oracle sample", file=sys.stderr, flush=True)
        cpu = cpu_sample(case, counts, args.cpu_threads, args.rtol)
    value = args.steps / elapsed
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": (f"cantilever bar {nx}x{ny}x{nz} Freudenthal cubes ({case.nv} vertices, {case.nt} tets), "
                         "stable Neo-Hookean E=1e5 nu=0.45 rho=1000, dt=1/60, Newton<=10 + Jacobi-PCG rtol 1e-8 + "
                         f"Armijo backtracking, Newton tol {args.rtol:g}*mass"),
            "seed": args.seed,
            "newton_iterations": counts["newton_iterations"],
            "pcg_iterations": counts["pcg_iterations"],
            "energy_evals": counts["energy_evals"],
            "hessian_evals": counts["hessian_evals"],
            "kernel_avg_ms": {k: v[0] for k, v in ktimes.items()},
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
