"""Summaries of rocprofv3 output for this library's kernels, and the copy kernel that calibrates its counters.

  python -m ssim_amd.profiling trace <rocprofv3 -d dir>   per-kernel durations (--kernel-trace database)
  python -m ssim_amd.profiling pmc <rocprofv3 -d dir>...  per-kernel, per-grid counter means (--pmc databases)
  python -m ssim_amd.profiling copy [MiB]                 device-to-device copy of a buffer larger than the
                                                          caches (default 512 MiB), three times, for calibrating
                                                          FETCH_SIZE / WRITE_SIZE against known bytes
"""
from __future__ import annotations

import glob
import sqlite3
import statistics
import sys
from collections import defaultdict


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


if __name__ == "__main__":
    cmd = sys.argv[1] if len(sys.argv) > 1 else ""
    if cmd == "trace":
        trace(sys.argv[2])
    elif cmd == "pmc":
        pmc(sys.argv[2:])
    elif cmd == "copy":
        copy(int(sys.argv[2]) if len(sys.argv) > 2 else 512)
    else:
        print(__doc__)
        sys.exit(2)
