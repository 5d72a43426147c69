"""Autotuner for the launch knobs of a solver or a batch (``SSIM_TUNE_*`` in ssim_hip.h).

Knobs: XCD remapping of the SpMV rows, the SpMV variant (fixed-width unrolled / looped / four lanes per row),
PCG iterations per done-flag poll, graph replay of PCG chunks, the tile size of the tile-reduced energy + gradient
(single meshes) and, for batches, the threshold at which the batched PCG grids are compacted to the meshes still
iterating.  None of them changes the algorithm; the
variants sum in different orders, so results can differ in the last bits.

The tuner snapshots the state, times one full time step per candidate from that same state (so every
candidate does the same work), restores the state, and applies the fastest setting.  It searches one knob at a
time, in the order above, keeping the best value of each before moving to the next (coordinate descent), and
repeats each timing ``repeats`` times (median); a candidate must beat the incumbent by a margin (1 %) to replace it.  Results are cached in a JSON file keyed by the mesh layout, so a
later process with the same mesh applies them without timing.
"""
from __future__ import annotations

import json
import os
import statistics
import time

DEFAULT_CACHE = os.environ.get("SSIM_TUNE_CACHE", os.path.join(os.path.expanduser("~"), ".cache", "ssim_amd",
                                                                  "tuning.json"))

# knob -> candidate values, in search order
SINGLE_SPACE = {
    "xcd_remap": [1, 0],
    "spmv_lanes": [0, 1, 4],
    "pcg_chunk_max": [64, 32, 128, 256],
    "pcg_graph": [0, 1],
    "grad_tile": [256, 128, 64],
}
BATCH_SPACE = {
    "xcd_remap": [1, 0],
    "spmv_lanes": [0, 1],
    "pcg_chunk_max": [64, 32, 128, 256],
    "compact_pct": [70, 0, 50, 90],
    "pcg_graph": [0, 1],
}
_STATE = ("deformation", "prev_deform", "velocity")


def signature(obj) -> str:
    """Cache key: kind and layout sizes of the solver or batch."""
    info = obj.layout_info()
    kind = "batch%d" % obj.num_meshes if hasattr(obj, "num_meshes") else "single"
    return f"{kind}:nv={info['nv']}:nt={info['nt']}:blocks={info['nonzero_blocks']}:slices={info['slices']}"


def coordinate_search(space: dict, measure, start: dict | None = None, margin: float = 0.0):
    """Minimises measure(setting) one knob at a time.  Returns (best setting, {setting repr: time}).

    A candidate replaces the incumbent only if it is faster by more than the fraction `margin`, so that timing
    noise does not move a knob off its first (expected best) value."""
    best = dict(start or {k: v[0] for k, v in space.items()})
    seen = {}

    def timed(setting):
        key = json.dumps(setting, sort_keys=True)
        if key not in seen:
            seen[key] = measure(setting)
        return seen[key]

    best_t = timed(best)
    for knob, values in space.items():
        for v in values:
            if v == best[knob]:
                continue
            trial = dict(best, **{knob: v})
            t = timed(trial)
            if t < best_t * (1.0 - margin):
                best, best_t = trial, t
    return best, seen


def _valid(obj, setting):
    try:
        obj.set_tuning(**setting)
        return True
    except ValueError:  # e.g. the fixed-width SpMV on a layout whose rows do not fit
        return False


def tune(obj, repeats: int = 2, cache: str | None = DEFAULT_CACHE, use_cache: bool = True, verbose: bool = False,
         margin: float = 0.01):
    """Tunes a TetFiniteElementSolver_Hip or TetFiniteElementBatch_Hip in place and returns the chosen setting.
    A knob leaves its current value only for a candidate faster by more than `margin` (1 %: one time step varies
    by about that much from run to run)."""
    import torch

    key = signature(obj)
    if cache and use_cache and os.path.exists(cache):
        with open(cache) as f:
            entry = json.load(f).get(key)
        if entry and _valid(obj, entry["setting"]):
            return entry["setting"]
    space = BATCH_SPACE if hasattr(obj, "num_meshes") else SINGLE_SPACE
    original = obj.tuning()
    snap = {name: obj.get_array(name) for name in _STATE}

    def restore():
        for name, values in snap.items():
            obj.set_array(name, values)

    def measure(setting):
        if not _valid(obj, setting):
            return float("inf")
        times = []
        for _ in range(repeats):
            restore()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            obj.step()
            torch.cuda.synchronize()
            times.append(time.perf_counter() - t0)
        t = statistics.median(times)
        if verbose:
            print(f"[tune] {setting}: {t * 1e3:.2f} ms", flush=True)
        return t

    start = {k: original.get(k, v[0]) for k, v in space.items()}
    best, seen = coordinate_search(space, measure, start, margin=margin)
    restore()
    obj.set_tuning(**best)
    if cache:
        os.makedirs(os.path.dirname(cache) or ".", exist_ok=True)
        data = {}
        if os.path.exists(cache):
            with open(cache) as f:
                data = json.load(f)
        data[key] = {"setting": best, "seconds": min(seen.values()), "candidates": len(seen)}
        with open(cache, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)
    return best
