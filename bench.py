"""Benchmark: Newton time steps per second of the FEM soft-body step on one MI355X (FP64).

Workload: cantilever bar of 32×32×128 Freudenthal cubes (140,481 vertices, 786,432 tets), E = 1e5 Pa, ν = 0.45,
ρ = 1000 kg/m³, x = 0 face clamped, gravity −9.8 m/s² in y, seeded jitter ±0.05·h, dt = 1/60 s, Newton ≤ 10
iterations with Jacobi-PCG rtol 1e-8 and backtracking Armijo (c = 0.3).  The Newton stopping rule is the
reference's ‖g‖₂ < threshold × total mass (time_step.hpp:482-492) with threshold 1e-6: at 786k tets the
reference drivers' 1e-3 (tests/fem_cu/main.cu:60) stops most steps at the predictor, 1e-6 makes every step
converge the nonlinear solve (about 3 Newton iterations and 2,800 PCG iterations per step).  All inputs are
resident in HBM when the timed region starts; the timed region is K whole steps (prepare + Newton solve + step_next).

Prints one JSON line.  ``roofline`` is for the dominant kernel (the PCG SpMV), whose average duration is
measured with hipEvents inside the timed region (stamped at the sampled launches' own start and end by
hipExtLaunchKernelGGL); ``cpu_baseline`` is the CPU FP64 oracle timed on a bounded
sample of the same workload on this host (see ``cpu_sample`` below).

Two more objects in the same line cover the batched and differentiable variants of the step:
``batched``: 256 bars of uneven size (dims uniform in 4..24, per-bar E log-uniform in [1e4, 1e6] Pa, ν uniform in
[0.3, 0.48], x = 0 end clamped, gravity) stepped by one set of launches whose knobs the autotuner picks, against
stepping the same 256 meshes one after another through the single-mesh path on the GPU, with the per-mesh
agreement of the two after one step, and the same batch split into concurrent size classes and into sequential
groups of meshes with similar step-1 PCG iterations (``best`` names the fastest of these over the same steps);
``adjoint``: a 20-step rollout of 32 bars of 16×16×64 cubes (per-tet log E uniform in [ln 1e4, ln 1e6], ν = 0.45)
with the loss Σ‖u − t‖² over the free-end vertices, timed forward and backward (adjoint PCG per step);
``lbfgs``: L-BFGS steps of the twisted 64³-cube (1,572,864 tets, per-tet material);
``fp32_elements``: 60 steps of the same bar with FP32 element kernels against FP64, both timed, with the largest
position difference; ``step_accounting``: the bytes a step must move by the per-step accounting (K1, K2, K3+K4
times the timed counts) over the step time.
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
    with row <= column are stored).  The fixed layout (``info["fixed_layout"]``, the one regular meshes get): 72 B
    of values + a 2 B column offset per stored off-diagonal block, 48 B per diagonal block (its upper triangle),
    2 B per lower-block reference and a 4 B word of the references' partner slots per row.  The general layout:
    76 B per stored block (72 B of values + a 4 B column index) and 8 B per lower-block reference (column +
    value offset).  Both: the gathered input z once (n·8 B), and the lane's own rows of p, q and x read and
    written (6·n·8 B; the kernel applies the previous iteration's x += α·p).  Padding slots are not counted."""
    nv = info["nv"]
    n = 3 * nv
    upper = info["upper_blocks"]
    lower = info["nonzero_blocks"] - upper
    if info.get("fixed_layout"):
        return 74 * (upper - nv) + 48 * nv + 2 * lower + 4 * nv + 7 * n * 8
    return 76 * upper + 8 * lower + 7 * n * 8


def element_bytes(info, uniform):
    """Algorithmic bytes of one energy + gradient evaluation (K1) and one SPD Hessian assembly (K2): per tet the
    rest geometry B (72 B), V (8 B) and cells (16 B), plus (λ, μ) (16 B) when the material varies per tet;
    K1 adds X, u, û read and g written (4·n·8), K2 X and u read (2·n·8), the nnz values written and the dof flags.
    These are the accounting's fixed figures: the energy + gradient kernel forms B and V from the gathered rest
    positions instead of loading them (SSIM_TUNE_REST_FROM_X), and so does the Hessian tile kernel: they stream
    80 B/tet less than K1 and K2 count."""
    n, nt = 3 * info["nv"], info["nt"]
    nnz = 9 * info["nonzero_blocks"]
    per_tet = 96 if uniform else 112
    return per_tet * nt + 4 * n * 8, per_tet * nt + 2 * n * 8 + nnz * 8 + n


def step_accounting(info, counts, elapsed, uniform):
    """Bytes one time step must move by the per-step accounting sum over Newton iterations of
    [K2 + N_pcg·(K3 + K4) + N_ls·K1] + K1, with K1 and K2 from element_bytes, K3 + K4 (one PCG iteration) = the
    SpMV kernel's bytes + 5·n·8 for the update kernel, and the counts of the timed steps."""
    n = 3 * info["nv"]
    k1, k2 = element_bytes(info, uniform)
    k34 = spmv_algorithmic_bytes(info) + 5 * n * 8
    total = counts["pcg_iterations"] * k34 + counts["hessian_evals"] * k2 + counts["energy_evals"] * k1
    per_step = total / counts["steps"]
    achieved = total / elapsed / 1e9
    return {"bytes_per_step": per_step, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "k1": k1, "k2": k2, "k3_k4": k34}


def element_rooflines(info, ktimes, uniform):
    """K1 (k_elem_grad_tile + k_grad_gather) and K2 (k_elem_hess_tile + k_hess_gather, the tile-reduced assembly)
    against their algorithmic bytes, from the same hipEvent kernel timers as the SpMV (timer slots elem_grad /
    grad_reduce and elem_hess / hess_reduce)."""
    k1, k2 = element_bytes(info, uniform)
    out = {}
    for name, b, pair, kernels in (
            ("k1_energy_gradient", k1, ("elem_grad", "grad_reduce"), ["k_elem_grad_tile", "k_grad_gather"]),
            ("k2_hessian_assembly", k2, ("elem_hess", "hess_reduce"), ["k_elem_hess_tile", "k_hess_gather"])):
        ms = sum(ktimes[k][0] for k in pair)
        gbs = b / (ms * 1e-3) / 1e9 if ms > 0 else None
        out[name] = {"bytes": b, "avg_ms": ms, "achieved": gbs, "unit": "GB/s",
                     "frac": gbs / HBM_PEAK_GBS if gbs else None, "kernels": kernels,
                     "per_kernel_ms": [ktimes[k][0] for k in pair]}
    return out


def load_traffic():
    path = os.path.join(ROOT, "profiles", "k_pcg_spmv_786k_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("hbm_bytes_per_launch")


def host_cpu():
    """Model, sockets and physical cores of the host (from /proc/cpuinfo) and the CPUs this process may use."""
    model, phys = None, set()
    pid = core = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                k, _, v = line.partition(":")
                k, v = k.strip(), v.strip()
                if k == "model name" and model is None:
                    model = v
                elif k == "physical id":
                    pid = v
                elif k == "core id":
                    core = v
                elif not k and pid is not None:
                    phys.add((pid, core))
                    pid = core = None
        if pid is not None:
            phys.add((pid, core))
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = os.cpu_count()
    return {"model": model, "sockets": len({p for p, _ in phys}) or None, "physical_cores": len(phys) or None,
            "usable_cpus": usable}


def cpu_steps(case, state, threads, rtol, nsteps=3):
    """Times whole oracle time steps (prepare + Newton + step_next, the reference's step(), fem_host.cpp:102-111)
    of the same bar on `threads` OpenMP threads, starting from the GPU solver's state at the start of its timed
    window, so that the CPU takes the first `nsteps` of the steps the GPU timed; the median of steps 2..nsteps
    after the first (warm-up) one, as BASELINE.md's CPU plan asks.  Returns (median seconds, per-step records)."""
    os.environ.setdefault("OMP_NUM_THREADS", str(threads))
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    ts = oracle.TimeStep.from_case(case)
    ts.set_threshold(rtol)
    ts.set_state(*state)
    rec = []
    for k in range(nsteps):
        t0 = time.perf_counter()
        _, st = ts.step()
        rec.append({"s": time.perf_counter() - t0, "newton": st["newton"], "pcg": st["pcg"]})
        print(f"[bench] CPU oracle step {k + 1}/{nsteps}: {rec[-1]['s']:.2f} s, {st['newton']} Newton, "
              f"{st['pcg']} PCG", file=sys.stderr, flush=True)
    return float(np.median([r["s"] for r in rec[1:]])), rec


def cpu_sample(case, counts, threads, rtol):
    """Times the oracle's phases on the same mesh (one Hessian assembly, two energy+gradient evaluations, 20 PCG
    iterations) and composes the CPU time of the GPU-timed steps from their Newton / PCG / evaluation counts; the
    same phases once more on one thread (one evaluation, one assembly, 5 PCG iterations).  A cross-check of the
    timed whole steps of cpu_steps."""
    os.environ.setdefault("OMP_NUM_THREADS", str(threads))
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    ts = oracle.TimeStep.from_case(case)
    ts.set_threshold(rtol)
    ts.prepare_step()
    u = case.u0.copy()
    rowptr, colidx = ts.csr_pattern()

    def phases(n_eval, pcg_its):
        t0 = time.perf_counter()
        for _ in range(n_eval):
            E, g = ts.energy_grad(u)
        t_eval = (time.perf_counter() - t0) / n_eval
        t0 = time.perf_counter()
        vals = ts.hessian(u, True, True)
        t_hess = time.perf_counter() - t0
        t0 = time.perf_counter()
        _, it, _ = oracle.pcg(rowptr, colidx, vals, -g, rtol=0.0, maxit=pcg_its)
        t_pcg = (time.perf_counter() - t0) / max(it, 1)
        return t_eval, t_hess, t_pcg

    steps = counts["steps"]

    def per_step(t_eval, t_hess, t_pcg):
        return (counts["energy_evals"] * t_eval + counts["hessian_evals"] * t_hess
                + counts["pcg_iterations"] * t_pcg) / steps

    cores = oracle.lib().orc_num_threads()
    t_eval, t_hess, t_pcg = phases(2, 20)
    oracle.lib().orc_set_num_threads(1)
    t1 = phases(1, 5)
    oracle.lib().orc_set_num_threads(cores)
    return {
        "value": 1.0 / per_step(t_eval, t_hess, t_pcg),
        "unit": "steps/s",
        "cores": cores,
        "kind": "port",
        "sample": (f"oracle phases on the same 786k-tet mesh: energy+grad {t_eval * 1e3:.1f} ms, SPD Hessian assembly "
                   f"{t_hess * 1e3:.1f} ms, PCG iteration {t_pcg * 1e3:.2f} ms; composed with the timed GPU steps' "
                   f"counts ({counts['energy_evals']} evals, {counts['hessian_evals']} Hessians, "
                   f"{counts['pcg_iterations']} PCG iterations over {steps} steps)"),
        "one_thread": {"value": 1.0 / per_step(*t1), "unit": "steps/s",
                       "sample": (f"the same phases on 1 thread: energy+grad {t1[0] * 1e3:.1f} ms, SPD Hessian "
                                  f"assembly {t1[1] * 1e3:.1f} ms, PCG iteration {t1[2] * 1e3:.2f} ms")},
        "host": host_cpu(),
    }


def _oracle(threads):
    os.environ.setdefault("OMP_NUM_THREADS", str(threads))
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    return oracle


def cpu_batched(cases, rtol, threads, sample=16):
    """The oracle's Newton step on the first `sample` meshes of the batch, scaled to the batch by tet count."""
    oracle = _oracle(threads)
    t, tets = 0.0, 0
    for c in cases[:sample]:
        ts = oracle.TimeStep.from_case(c)
        ts.set_threshold(rtol)
        t0 = time.perf_counter()
        ts.step()
        t += time.perf_counter() - t0
        tets += c.nt
    total = sum(c.nt for c in cases)
    per_step = t * total / tets
    return {"value": 1.0 / per_step, "unit": "batch steps/s", "cores": oracle.lib().orc_num_threads(), "kind": "port",
            "sample": f"oracle Newton step of the first {sample} meshes ({tets} tets, {t:.2f} s), scaled by tet count "
                      f"to the {total} tets of the batch"}


def cpu_lbfgs(case, rtol, threads, its_per_step, sample_its=20):
    """The oracle's L-BFGS on the same cube for `sample_its` iterations; composed with the GPU's iterations per
    step."""
    oracle = _oracle(threads)
    ts = oracle.TimeStep.from_case(case)
    ts.set_threshold(rtol)
    ts.set_lbfgs(max_iter=1)  # builds the rest-Hessian H0 once, outside the timed sample
    ts.prepare_step()
    ts.compute_step("lbfgs")
    ts.set_lbfgs(max_iter=sample_its)
    ts.prepare_step()
    t0 = time.perf_counter()
    _, st = ts.compute_step("lbfgs")
    t = time.perf_counter() - t0
    per_it = t / max(st["newton"], 1)
    return {"value": 1.0 / (per_it * its_per_step), "unit": "steps/s", "cores": oracle.lib().orc_num_threads(),
            "kind": "port",
            "sample": f"oracle L-BFGS, {st['newton']} iterations on the same cube ({per_it * 1e3:.1f} ms each), "
                      f"composed with the GPU's {its_per_step:.0f} iterations per step"}


def cpu_adjoint(rc, rtol, threads, adj_pcg_per_step, steps=2):
    """CPU time of the rollout and its adjoint, composed from the oracle on one bar: `steps` measured Newton steps
    (forward), and per step one unprojected Hessian assembly plus the GPU adjoint's PCG iterations per step at the
    oracle's time per PCG iteration (backward); scaled to the rollout's bars and steps."""
    oracle = _oracle(threads)
    from ssim_amd import adjoint

    case = rc.cases[0]
    lam, mu = adjoint.lame_from_log_e(rc.log_e[:case.nt], rc.nu[:case.nt])
    ts = oracle.TimeStep(case.X, case.cells, case.dt, lam, mu, case.rho)
    ts.set_dirichlet(case.dof_type, case.dbc)
    ts.set_ext_accel(case.accel)
    ts.set_state(case.u0, None, case.v0)
    ts.set_threshold(rtol)
    ts.set_solver(pcg_rtol=FD_SETTINGS["pcg_rtol"])
    ts.set_polish(FD_SETTINGS["newton_polish"])
    t0 = time.perf_counter()
    for _ in range(steps):
        ts.step()
    t_fwd = (time.perf_counter() - t0) / steps
    u = ts.get_state()[0]
    t0 = time.perf_counter()
    vals = ts.hessian(u, False, True)
    t_hess = time.perf_counter() - t0
    rowptr, colidx = ts.csr_pattern()
    _, g = ts.energy_grad(u)
    t0 = time.perf_counter()
    _, it, _ = oracle.pcg(rowptr, colidx, vals, g, rtol=0.0, maxit=20)
    t_pcg = (time.perf_counter() - t0) / max(it, 1)
    t_bwd = t_hess + adj_pcg_per_step * t_pcg
    total = (t_fwd + t_bwd) * len(rc.cases) * rc.steps
    return {"value": total, "unit": "s per rollout (forward + adjoint)", "cores": oracle.lib().orc_num_threads(),
            "kind": "port",
            "sample": f"oracle on one bar: Newton step {t_fwd:.2f} s (measured over {steps}), adjoint step = Hessian "
                      f"{t_hess:.2f} s + {adj_pcg_per_step:.0f} PCG iterations (the GPU adjoint's per step) x "
                      f"{t_pcg * 1e3:.1f} ms; scaled to {len(rc.cases)} bars x {rc.steps} steps"}


def bench_batched(rtol, steps=3, seed=0, cpu_threads=None):
    """Batched vs sequential per-mesh GPU stepping of 256 uneven bars; per-mesh agreement after one step."""
    import torch

    from ssim_amd import batch, fem, mesh

    cases = mesh.bar_batch(256, seed=seed)
    t0 = time.perf_counter()
    B = batch.TetFiniteElementBatch_Hip.from_cases(cases, rtol=rtol)
    setup_b = time.perf_counter() - t0
    singles = [fem.TetFiniteElementSolver_Hip.from_case(c, rtol=rtol) for c in cases]
    B.step()
    for s in singles:
        s.step()
    B1 = [B.get_array("deformation", m) for m in range(len(cases))]  # step-1 state, for the grouped runs' check
    pcg1 = [B.last_stats(m)["pcg_iterations"] for m in range(len(cases))]  # the grouping key of sequential groups
    worst, mism = 0.0, 0
    for m, (c, s) in enumerate(zip(cases, singles)):
        xb = c.X.reshape(-1) + B.get_array("deformation", m)
        xs = c.X.reshape(-1) + s.get_array("deformation")
        worst = max(worst, float(np.abs(xb - xs).max() / np.abs(xs).max()))
        mism += int(B.last_stats(m)["newton_iterations"] != s.last_stats()["newton_iterations"])
    print(f"[bench] batched: per-mesh max position rel err {worst:.3e}, Newton count mismatches {mism}",
          file=sys.stderr, flush=True)
    # launch knobs of the batch (SpMV variant, XCD mapping, PCG chunking, graph replay, grid compaction) chosen by
    # the autotuner: one step per candidate from the same state, which it restores afterwards
    from ssim_amd import autotune

    t0 = time.perf_counter()
    tuned = autotune.tune(B, repeats=1, cache=None)
    tune_s = time.perf_counter() - t0
    print(f"[bench] batched: autotuned {tuned} in {tune_s:.1f} s", file=sys.stderr, flush=True)
    # both paths take the same steps 2..steps+1 from the same state (the autotuner restored the batch's)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    tot = dict(newton_iterations=0, pcg_iterations=0)
    for _ in range(steps):
        B.step()
        st = B.last_stats()
        tot["newton_iterations"] += st["newton_iterations"]
        tot["pcg_iterations"] += st["pcg_iterations"]
    torch.cuda.synchronize()
    tb = (time.perf_counter() - t0) / steps
    seq_newton = seq_pcg = 0
    t0 = time.perf_counter()
    for _ in range(steps):
        for s in singles:
            s.step()
            st = s.last_stats()
            seq_newton += st["newton_iterations"]
            seq_pcg += st["pcg_iterations"]
    torch.cuda.synchronize()
    ts = (time.perf_counter() - t0) / steps
    print(f"[bench] batched {tb:.3f} s/step, sequential {ts:.3f} s/step (same {steps} steps)", file=sys.stderr,
          flush=True)
    # the same batch split into size classes stepped concurrently on their own streams, and into groups of meshes
    # with similar step-1 PCG iterations stepped one after another: step 1 untimed, then the same steps
    # 2..steps+1 as above; each grouping is one candidate of the tuning
    grouped, sequential = {}, {}
    for ngroups, concurrent in ((2, True), (4, True), (3, False), (4, False)):
        G = (batch.BatchGroups(cases, groups=ngroups, rtol=rtol) if concurrent else
             batch.BatchGroups(cases, groups=ngroups, rtol=rtol, key=pcg1, concurrent=False))
        G.set_tuning(**tuned)
        G.step()
        gworst = max(float(np.abs(G.get_array("deformation", m) - B1[m]).max() / max(np.abs(B1[m]).max(), 1e-300))
                     for m in range(len(cases)))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            G.step()
        torch.cuda.synchronize()
        tg = (time.perf_counter() - t0) / steps
        rec = {"s_per_step": tg, "speedup_vs_sequential": ts / tg, "step1_max_rel_diff_vs_batch": gworst,
               "group_sizes": [len(ms) for ms in G.members],
               "group_tets": [int(sum(cases[m].nt for m in ms)) for ms in G.members]}
        (grouped if concurrent else sequential)[ngroups] = rec
        kind = "concurrent size classes" if concurrent else "sequential groups by step-1 PCG iterations"
        print(f"[bench] batched in {ngroups} {kind}: {tg:.3f} s/step ({ts / tg:.1f}x sequential), step-1 "
              f"difference to the single batch {gworst:.1e}", file=sys.stderr, flush=True)
        del G
    cands = [("lock-step", 1, tb)] + [("concurrent size classes", g, r["s_per_step"]) for g, r in grouped.items()] + \
        [("sequential groups by step-1 PCG iterations", g, r["s_per_step"]) for g, r in sequential.items()]
    best = min(cands, key=lambda c: c[2])
    return {
        "workload": "256 bars, dims uniform in 4..24, per-bar E log-uniform [1e4,1e6], nu uniform [0.3,0.48]",
        "tuning": tuned,
        "tune_s": tune_s,
        "meshes": len(cases),
        "tets": int(sum(c.nt for c in cases)),
        "vertices": int(sum(c.nv for c in cases)),
        "newton_tol": f"{rtol:g}*mass per mesh",
        "batched_steps_per_s": 1.0 / tb,
        "sequential_steps_per_s": 1.0 / ts,
        "speedup_vs_sequential": ts / tb,
        "concurrent_groups": grouped,
        "sequential_groups_by_pcg": sequential,
        "best": {"grouping": best[0], "groups": best[1], "batched_steps_per_s": 1.0 / best[2],
                 "speedup_vs_sequential": ts / best[2]},
        "timed_steps": steps,
        "timed_window": f"steps 2-{steps + 1} of every mesh, batched and sequential",
        "lockstep_newton_iterations": tot["newton_iterations"],
        "lockstep_pcg_iterations": tot["pcg_iterations"],
        "sequential_newton_iterations": seq_newton,
        "sequential_pcg_iterations": seq_pcg,
        "setup_s": setup_b,
        "per_mesh_max_position_rel_err": worst,
        "newton_count_mismatches": mism,
        "cpu_baseline": cpu_batched(cases, rtol, cpu_threads) if cpu_threads else None,
    }


def bench_lbfgs(rtol, n=64, steps=3, seed=0, cpu_threads=None):
    """L-BFGS steps/s on the twisted cube of n³ cubes (per-tet E log-uniform, ν uniform, z = 0 face clamped)."""
    import torch

    from ssim_amd import fem, mesh

    case = mesh.twisted_cube(n=n, seed=seed)
    s = fem.TetFiniteElementSolver_Hip.from_case(case, method="lbfgs", rtol=rtol)
    s.step()
    its = evals = 0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        s.step()
        st = s.last_stats()
        its += st["newton_iterations"]
        evals += st["energy_evals"]
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    print(f"[bench] lbfgs: {dt:.3f} s/step, {its} iterations", file=sys.stderr, flush=True)
    return {
        "workload": (f"twisted cube {n}^3 Freudenthal cubes ({case.nv} vertices, {case.nt} tets), per-tet E log-uniform "
                     "[1e4,1e6], nu uniform [0.3,0.48], z=0 clamped, L-BFGS m=8 with H0 = Jacobi of the rest Hessian, "
                     "Armijo c=1e-4"),
        "newton_tol": f"{rtol:g}*mass",
        "steps_per_s": 1.0 / dt,
        "timed_steps": steps,
        "lbfgs_iterations": its,
        "energy_evals": evals,
        "cpu_baseline": cpu_lbfgs(case, rtol, cpu_threads, its / steps) if cpu_threads else None,
    }


def bench_fp32(case, rtol, total=60):
    """The FP32 element-kernel variant on the same bar: an FP64 and an FP32-element solver each take ``total`` steps
    from the same initial state (both timed); the positions after the last step are compared."""
    import numpy as np
    import torch

    from ssim_amd import fem

    out, rate, pcg = {}, {}, {}
    for fp32 in (0, 1):
        s = fem.TetFiniteElementSolver_Hip.from_case(case, rtol=rtol)
        s.set_solver_params(fp32_elements=fp32)
        its = 0
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(total):
            s.step()
            its += s.last_stats()["pcg_iterations"]
        torch.cuda.synchronize()
        rate[fp32] = total / (time.perf_counter() - t0)
        pcg[fp32] = its
        out[fp32] = case.X.reshape(-1) + s.get_array("deformation")
        del s
    err = float(np.abs(out[1] - out[0]).max())
    print(f"[bench] fp32 elements: {rate[1]:.2f} steps/s (FP64 {rate[0]:.2f}) over {total} steps, max position error "
          f"{err:.3e} m", file=sys.stderr, flush=True)
    return {
        "workload": (f"the same bar, steps 1-{total} from the same initial state with the element kernels (F, Psi, P, H, "
                     "H+, K) in FP32; reductions, assembly sums, PCG and positions in FP64"),
        "steps": total,
        "steps_per_s": rate[1],
        "fp64_steps_per_s": rate[0],
        "pcg_iterations": pcg[1],
        "fp64_pcg_iterations": pcg[0],
        "max_position_error_m": err,
        "max_position_error_rel": err / float(np.abs(out[0]).max()),
    }


# the settings at which the adjoint gradient of configuration 5 is checked against central differences
# (tests/test_gpu_configs.py): Newton threshold 1e-8·mass, PCG to 1e-12, two polishing Newton steps
FD_SETTINGS = dict(rtol=1e-8, pcg_rtol=1e-12, newton_polish=2)


def bench_adjoint(nbar=32, seed=0, cpu_threads=None, checkpoint_every=1, rc=None):
    """Forward and adjoint (backward) time of a 20-step differentiable rollout of 32 bars of 16x16x64 cubes at the
    finite-difference-verified solver settings; with checkpoint_every > 1 the backward pass recomputes the states
    between checkpoints."""
    import torch

    from ssim_amd import adjoint, batch, mesh

    rtol = FD_SETTINGS["rtol"]
    rc = rc or mesh.rollout_batch(nbar, seed=seed)
    B = batch.TetFiniteElementBatch_Hip.from_cases(rc.cases, rtol=rtol)
    B.set_solver_params(pcg_rtol=FD_SETTINGS["pcg_rtol"], newton_polish=FD_SETTINGS["newton_polish"])
    runner = adjoint.Rollout(B, rc.nu, rc.steps, rc.cases[0].dt, checkpoint_every=checkpoint_every)
    idx, tgt, off = [], [], 0
    for c, fe, t in zip(rc.cases, rc.free_end, rc.targets):
        idx.append(off + fe)
        tgt.append(np.broadcast_to(t, (fe.size, 3)))
        off += c.nv
    tidx = torch.as_tensor(np.concatenate(idx), device="cuda")
    ttgt = torch.as_tensor(np.concatenate(tgt), device="cuda")
    le = torch.as_tensor(rc.log_e, device="cuda").requires_grad_(True)
    v0 = torch.zeros(off * 3, dtype=torch.float64, device="cuda").requires_grad_(True)
    u0 = torch.as_tensor(np.concatenate([c.u0 for c in rc.cases]), device="cuda")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    uN = adjoint.rollout(runner, le, v0, u0)
    d = uN.view(-1, 3)[tidx] - ttgt
    L = (d * d).sum()
    loss = L.item()
    torch.cuda.synchronize()
    fwd = time.perf_counter() - t0
    t0 = time.perf_counter()
    L.backward()
    torch.cuda.synchronize()
    bwd = time.perf_counter() - t0
    print(f"[bench] adjoint (checkpoint every {checkpoint_every}): forward {fwd:.2f} s, backward {bwd:.2f} s",
          file=sys.stderr, flush=True)
    return {
        "workload": (f"{len(rc.cases)} bars of 16x16x64 cubes ({rc.cases[0].nt} tets each), per-tet log E uniform in "
                     "[ln 1e4, ln 1e6], nu=0.45, 20 steps, loss sum over free-end vertices of |u - t|^2, |t|=0.02 m"),
        "tets": int(sum(c.nt for c in rc.cases)),
        "steps": rc.steps,
        "settings": (f"Newton tol {rtol:g}*mass per mesh, PCG rtol {FD_SETTINGS['pcg_rtol']:g}, "
                     f"{FD_SETTINGS['newton_polish']} polishing Newton steps (the finite-difference-checked settings)"),
        "checkpoint_every": checkpoint_every,
        "recomputed_forward_steps": rc.steps - len(range(0, rc.steps, checkpoint_every)),
        "loss": loss,
        "forward_s": fwd,
        "backward_s": bwd,
        "backward_over_forward": bwd / fwd,
        "adjoint_pcg_iterations_lockstep": runner.adjoint_pcg_iterations,
        "gradient_norm_log_e": float(le.grad.norm().item()),
        "gradient_norm_v0": float(v0.grad.norm().item()),
        "cpu_baseline": (cpu_adjoint(rc, rtol, cpu_threads, runner.adjoint_pcg_iterations / rc.steps)
                         if cpu_threads else None),
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
    ap.add_argument("--cpu-steps", type=int, default=3, help="whole oracle steps timed for cpu_baseline")
    ap.add_argument("--no-batched", action="store_true", help="skip the 256-mesh batched leg")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the differentiable-rollout leg")
    ap.add_argument("--adjoint-checkpoints", type=str, default="1,5",
                    help="checkpoint intervals of the adjoint leg: the first is the reported run, a second one is "
                         "reported as with_recompute")
    ap.add_argument("--no-lbfgs", action="store_true", help="skip the L-BFGS twisted-cube leg")
    ap.add_argument("--no-fp32", action="store_true", help="skip the FP32 element-variant leg")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only: the time step runs on a single device "
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
    # state at the start of the timed window: the CPU baseline steps from here (copied before timing starts)
    state0 = tuple(solver.get_array(a) for a in ("deformation", "prev_deform", "velocity"))
    solver.set_kernel_timing(True)
    counts = dict(steps=args.steps, newton_iterations=0, pcg_iterations=0, energy_evals=0, hessian_evals=0)
    step_ends, step_stats = [], []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        solver.step()  # returns after the step's last host read-back, so the step's wall time is its own
        st = solver.last_stats()
        step_ends.append(time.perf_counter())
        step_stats.append(st)
        for key in ("newton_iterations", "pcg_iterations", "energy_evals", "hessian_evals"):
            counts[key] += st[key]
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    step_s = np.diff([t0] + step_ends)
    ktimes = solver.kernel_timing()
    solver.set_kernel_timing(False)
    print(f"[bench] timed {args.steps} steps in {elapsed:.3f} s; counts {counts}", file=sys.stderr, flush=True)
    print(f"[bench] kernel timing (avg ms, samples): {ktimes}", file=sys.stderr, flush=True)

    uniform = bool(np.all(case.lam == case.lam[0]) and np.all(case.mu == case.mu[0]))
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
        "kernel": "k_pcg_spmv (symmetric blocked-ELL 3x3 SpMV, x/p/q update, p·q partials)",
        "bytes_per_launch": bytes_per_launch,
        "avg_launch_ms": spmv_ms,
        "sampled_launches": spmv_n,
    }
    cpu = None
    if not args.no_cpu:
        print("[bench] timing the CPU oracle: whole steps from the timed window's start", file=sys.stderr, flush=True)
        nsteps = min(args.cpu_steps, args.steps)
        med, rec = cpu_steps(case, state0, args.cpu_threads, args.rtol, nsteps)
        composed = cpu_sample(case, counts, args.cpu_threads, args.rtol)
        gpu_med = float(np.median(step_s[1:nsteps]))
        cpu = {
            "value": 1.0 / med,
            "unit": "steps/s",
            "cores": composed["cores"],
            "kind": "port",
            "sample": (f"oracle whole time steps 1-{nsteps} of the GPU's timed window (started from the GPU state "
                       f"after warm-up), {composed['cores']} OpenMP threads: median of steps 2-{nsteps} = {med:.2f} s "
                       f"per step; per step {[round(r['s'], 2) for r in rec]} s, Newton {[r['newton'] for r in rec]}, "
                       f"PCG {[r['pcg'] for r in rec]}"),
            "gpu_same_steps": {"median_s": gpu_med, "newton": [st["newton_iterations"] for st in step_stats[:nsteps]],
                               "pcg": [st["pcg_iterations"] for st in step_stats[:nsteps]],
                               "gpu_over_cpu": med / gpu_med},
            "composed": {k: composed[k] for k in ("value", "unit", "sample", "one_thread")},
            "host": composed["host"],
        }
    del solver
    fp32 = None if args.no_fp32 else bench_fp32(case, args.rtol)
    cpu_threads = None if args.no_cpu else args.cpu_threads
    batched = None if args.no_batched else bench_batched(args.rtol, cpu_threads=cpu_threads)
    adj = None
    if not args.no_adjoint:
        from ssim_amd import mesh as _mesh

        rc = _mesh.rollout_batch(32, seed=args.seed)
        every = [int(v) for v in args.adjoint_checkpoints.split(",") if v]
        adj = bench_adjoint(cpu_threads=cpu_threads, rc=rc, checkpoint_every=every[0])
        if len(every) > 1:
            adj["with_recompute"] = bench_adjoint(checkpoint_every=every[1], rc=rc)
    lbfgs = None if args.no_lbfgs else bench_lbfgs(args.rtol, cpu_threads=cpu_threads)
    # GPU / CPU ratios of the extension legs (their cpu_baseline objects are bounded samples, as for the main line)
    if batched and batched.get("cpu_baseline"):
        batched["gpu_over_cpu"] = batched["batched_steps_per_s"] / batched["cpu_baseline"]["value"]
    if lbfgs and lbfgs.get("cpu_baseline"):
        lbfgs["gpu_over_cpu"] = lbfgs["steps_per_s"] / lbfgs["cpu_baseline"]["value"]
    if adj and adj.get("cpu_baseline"):
        adj["gpu_over_cpu"] = adj["cpu_baseline"]["value"] / (adj["forward_s"] + adj["backward_s"])
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
        "step_accounting": step_accounting(info, counts, elapsed, uniform),
        "element_kernels": element_rooflines(info, ktimes, uniform),
        "cpu_baseline": cpu,
        "batched": batched,
        "adjoint": adj,
        "lbfgs": lbfgs,
        "fp32_elements": fp32,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
