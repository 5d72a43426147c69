"""Summaries of rocprofv3 output for this library's kernels, and the copy kernel that calibrates its counters.

  python -m ssim_amd.profiling trace <rocprofv3 -d dir>   per-kernel durations (--kernel-trace database)
  python -m ssim_amd.profiling pmc <rocprofv3 -d dir>...  per-kernel, per-grid counter means (--pmc databases)
  python -m ssim_amd.profiling copy [MiB]                 device-to-device copy of a buffer larger than the
                                                          caches (default 512 MiB), three times, for calibrating
                                                          FETCH_SIZE / WRITE_SIZE against known bytes
  python -m ssim_amd.profiling evaluators                 the element evaluators (ssim_tri_eval on ~1.0M
                                                          triangles with every output, ssim_material_eval on 2^20
                                                          items with Hessians), 50 timed launches per case
  python -m ssim_amd.profiling batch-steps                the 256-bar batch of bench.py: one step, the autotuner,
                                                          two more steps (per-mesh PCG work of the last one)
  python -m ssim_amd.profiling timeline <rocprofv3 -d dir> the last time step of a kernel trace (from its last
                                                          k_prepare): busy time per kernel, launch gaps, and the
                                                          batched PCG launches by grid size
"""
from __future__ import annotations

import glob
import sqlite3
import statistics
import sys
from collections import defaultdict

if __package__ in (None, ""):  # run as a script: make the package importable
    import os

    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))


def _dbs(path):
    return sorted(glob.glob(path.rstrip("/") + "/**/*.db", recursive=True)) or [path]


def _short(name):
    return name.replace("(anonymous namespace)::", "").split("(")[0].replace("void ", "")


def trace(path):
    for db in _dbs(path):
        cur = sqlite3.connect(db).cursor()
        rows = list(cur.execute("select name, start, end from kernels order by start"))
        d = defaultdict(list)
        for n, s, e in rows:
            d[_short(n)].append((e - s) / 1e3)
        print(f"# {db}: {len(rows)} dispatches")
        print(f"{'kernel':40s} {'calls':>6s} {'total_ms':>9s} {'avg_us':>8s} {'median_us':>9s} {'avg_busy_us':>11s}")
        for n, v in sorted(d.items(), key=lambda x: -sum(x[1])):
            busy = [t for t in v if t > 2.0]  # early-exit launches of a converged PCG chunk take ~1.5 us
            print(f"{n:40s} {len(v):6d} {sum(v) / 1e3:9.2f} {statistics.fmean(v):8.2f} {statistics.median(v):9.2f} "
                  f"{(statistics.fmean(busy) if busy else 0):11.2f}")


def pmc(paths):
    for path in paths:
        for db in _dbs(path):
            cur = sqlite3.connect(db).cursor()
            rows = list(cur.execute("select kernel_name, grid_size, counter_name, value, duration "
                                    "from counters_collection"))
            d = defaultdict(list)
            for kn, gs, cn, v, dur in rows:
                d[(_short(kn), gs, cn)].append((v, dur))
            print(f"# {db}")
            for (kn, gs, cn), vals in sorted(d.items()):
                v = [x[0] for x in vals]
                durs = sorted(x[1] for x in vals)
                cut = durs[len(durs) // 2] / 4 if durs else 0
                busy = [x[0] for x in vals if x[1] >= cut]
                print(f"{kn:40s} grid={gs:9d} {cn:12s} n={len(v):6d} mean={statistics.fmean(v):13.1f} "
                      f"median={statistics.median(v):13.1f} busy_mean={(statistics.fmean(busy) if busy else 0):13.1f}")


def copy(mib=512):
    import torch

    n = mib * (1 << 20) // 8
    a = torch.rand(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(3):
        b.copy_(a)
        a.copy_(b)  # alternate so that no source is still cached from a write
    torch.cuda.synchronize()
    print(f"copied {mib} MiB ({n * 8} bytes) x6")


def evaluators(reps=50):
    """Average launch time (hipEvents on the launching stream) and algorithmic bytes of the element evaluators:
    per triangle 12 B of cells, 16 B of λ, μ and 70 doubles of outputs, plus X and u once; per material item
    F, λ, μ, Ψ, P and H once."""
    import ctypes

    import numpy as np
    import torch

    from ssim_amd import _lib

    _lib.load()
    dev = torch.device("cuda")
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    ptr = lambda x: ctypes.c_void_p(x.data_ptr())
    dev_t = lambda a: torch.as_tensor(np.ascontiguousarray(a), device=dev)

    def timed(call):
        for _ in range(5):
            call()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record()
        for _ in range(reps):
            call()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps * 1e3

    print(f"{'case':44s} {'items':>9s} {'us/launch':>10s} {'MB':>8s} {'GB/s':>8s} {'% of 8 TB/s':>11s}", flush=True)
    N = 708  # (N-1)^2 squares, two triangles each
    xs, ys = np.meshgrid(np.linspace(0, 1, N), np.linspace(0, 1, N), indexing="ij")
    X = np.stack([xs.ravel(), ys.ravel()], 1)
    i, j = np.meshgrid(np.arange(N - 1), np.arange(N - 1), indexing="ij")
    a = (i * N + j).ravel()
    C = np.concatenate([np.stack([a, a + N, a + N + 1], 1), np.stack([a, a + N + 1, a + 1], 1)]).astype(np.int32)
    rng = np.random.default_rng(0)
    nv, nt = len(X), len(C)
    Xd, Cd, ud = dev_t(X), dev_t(C), dev_t(rng.standard_normal(X.shape) * 1e-4)
    lam, mu = dev_t(np.full(nt, 1e4)), dev_t(np.full(nt, 1e3))
    widths = (1, 4, 4, 1, 6, 36, 9, 9)  # area, dminv, F, energy, grad, stiff, mass, laplacian
    outs = [torch.empty(nt * w, dtype=torch.float64, device=dev) for w in widths]
    tri_bytes = nt * (12 + 16 + 8 * sum(widths)) + nv * 32
    for model, name in ((0, "stable_neohookean"), (1, "linear")):
        for spsd in (0, 1):
            us = timed(lambda: _lib.call("ssim_tri_eval", model, nv, nt, ptr(Xd), ptr(Cd), ptr(ud), ptr(lam),
                                         ptr(mu), 1000.0, spsd, *[ptr(o) for o in outs], stream))
            print(f"{'tri ' + name + (' projected' if spsd else ''):44s} {nt:9d} {us:10.1f} {tri_bytes / 1e6:8.1f} "
                  f"{tri_bytes / us / 1e3:8.0f} {tri_bytes / us / 8e4:10.1f}%", flush=True)
    n = 1 << 20
    for dim in (3, 2):
        F = dev_t(np.tile(np.eye(dim).ravel(), n) + 1e-2 * rng.standard_normal(n * dim * dim))
        lm, mm = dev_t(np.full(n, 1e4)), dev_t(np.full(n, 1e3))
        psi = torch.empty(n, dtype=torch.float64, device=dev)
        P = torch.empty(n * dim * dim, dtype=torch.float64, device=dev)
        H = torch.empty(n * dim ** 4, dtype=torch.float64, device=dev)
        mat_bytes = n * 8 * (2 * dim * dim + 3 + dim ** 4)
        for model, name in ((0, "stable_neohookean"), (1, "linear")):
            for spsd in ((0, 1) if dim == 3 else (0,)):
                us = timed(lambda: _lib.call("ssim_material_eval", dim, model, ctypes.c_longlong(n), ptr(F), ptr(lm),
                                             ptr(mm), spsd, ptr(psi), ptr(P), ptr(H), stream))
                label = f"material {dim}D {name}" + (" projected" if spsd else "")
                print(f"{label:44s} {n:9d} {us:10.1f} {mat_bytes / 1e6:8.1f} {mat_bytes / us / 1e3:8.0f} "
                      f"{mat_bytes / us / 8e4:10.1f}%", flush=True)


def batch_steps():
    """The batch of bench.py's batched leg (256 bars, dims 4..24): step 1, the autotuner's trials, steps 2 and 3;
    prints the lock-step counts, the per-mesh PCG iterations of step 3 and the batch's PCG work in whole-batch
    iterations, Σ_m iterations_m · tets_m / Σ_m tets_m."""
    import time

    import numpy as np
    import torch

    from ssim_amd import autotune, batch, mesh

    cases = mesh.bar_batch(256, seed=0)
    B = batch.TetFiniteElementBatch_Hip.from_cases(cases, rtol=1e-6)
    B.step()
    print("tuned", autotune.tune(B, repeats=1, cache=None), flush=True)
    for k in (2, 3):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        B.step()
        torch.cuda.synchronize()
        st = B.last_stats()
        print(f"step {k}: {(time.perf_counter() - t0) * 1e3:.1f} ms, lock-step Newton {st['newton_iterations']}, "
              f"PCG {st['pcg_iterations']}", flush=True)
    nts = np.array([c.nt for c in cases])
    pcg = np.array([B.last_stats(m)["pcg_iterations"] for m in range(len(cases))])
    print(f"step 3 per-mesh PCG iterations min/median/max {pcg.min()} {np.median(pcg):.0f} {pcg.max()}; "
          f"PCG work in whole-batch iterations {(pcg * nts).sum() / nts.sum():.1f}", flush=True)
    slow = np.argsort(-pcg)[:5]
    print("slowest meshes (tets, PCG iterations):", [(int(nts[i]), int(pcg[i])) for i in slow], flush=True)


def timeline(path):
    for db in _dbs(path):
        cur = sqlite3.connect(db).cursor()
        rows = [(_short(n), s, e, g) for n, s, e, g in
                cur.execute("select name, start, end, grid_x from kernels order by start")]
        starts = [i for i, r in enumerate(rows) if r[0].endswith("k_prepare")]
        if not starts:
            continue
        seg = rows[starts[-1]:]
        gaps = sum(max(0, b[1] - a[2]) for a, b in zip(seg, seg[1:])) / 1e3
        print(f"# {db}: last step {len(seg)} launches over {(seg[-1][2] - seg[0][1]) / 1e3:.1f} us, "
              f"gaps between launches {gaps:.1f} us")
        busy = defaultdict(list)
        for n, s0, e0, _ in seg:
            busy[n].append((e0 - s0) / 1e3)
        for n, v in sorted(busy.items(), key=lambda x: -sum(x[1])):
            print(f"{n:40s} {len(v):6d} {sum(v) / 1e3:9.2f} ms")
        for kern in ("k_bpcg_spmv", "k_bpcg_update"):
            by = defaultdict(list)
            for n, s0, e0, g in seg:
                if kern in n:
                    by[g].append((e0 - s0) / 1e3)
            if not by:
                continue
            print(f"{kern} by grid (work-items): launches, mean us, total ms")
            for g in sorted(by, reverse=True):
                print(f"  {g:9d} {len(by[g]):5d} {statistics.fmean(by[g]):8.2f} {sum(by[g]) / 1e3:8.2f}")


if __name__ == "__main__":
    cmd = sys.argv[1] if len(sys.argv) > 1 else ""
    if cmd == "trace":
        trace(sys.argv[2])
    elif cmd == "pmc":
        pmc(sys.argv[2:])
    elif cmd == "copy":
        copy(int(sys.argv[2]) if len(sys.argv) > 2 else 512)
    elif cmd == "evaluators":
        evaluators()
    elif cmd == "batch-steps":
        batch_steps()
    elif cmd == "timeline":
        timeline(sys.argv[2])
    else:
        print(__doc__)
        sys.exit(2)
