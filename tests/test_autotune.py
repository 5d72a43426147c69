"""The autotuner's search logic (CPU) and, on the GPU, tuning a solver and a batch in place."""
import json

import numpy as np
import pytest

from ssim_amd import autotune


def test_coordinate_search_finds_the_best_value_of_each_knob():
    space = {"a": [0, 1, 2], "b": [5, 3, 4], "c": [0, 1]}
    cost = {0: 3.0, 1: 1.0, 2: 2.0}
    calls = []

    def measure(s):
        calls.append(dict(s))
        return cost[s["a"]] + abs(s["b"] - 4) + (0.5 if s["c"] else 0.0)

    best, seen = autotune.coordinate_search(space, measure)
    assert best == {"a": 1, "b": 4, "c": 0}
    assert len(seen) == len(calls)  # every setting timed once
    assert len(calls) <= 1 + sum(len(v) - 1 for v in space.values())


def test_a_knob_moves_only_for_a_gain_beyond_the_margin():
    space = {"x": [1, 0], "y": [0, 1]}
    cost = {(1, 0): 1.000, (0, 0): 0.995, (1, 1): 0.97}  # x = 0 is 0.5 % faster: within the margin

    best, _ = autotune.coordinate_search(space, lambda s: cost.get((s["x"], s["y"]), 2.0), margin=0.01)
    assert best == {"x": 1, "y": 1}
    best, _ = autotune.coordinate_search(space, lambda s: cost.get((s["x"], s["y"]), 2.0))
    assert best == {"x": 0, "y": 0}  # without a margin any gain moves the knob (then y = 1 is not tried with x = 0)


def test_infeasible_settings_never_win():
    space = {"x": [0, 1, 2]}
    best, _ = autotune.coordinate_search(space, lambda s: float("inf") if s["x"] == 2 else 1.0 + s["x"])
    assert best == {"x": 0}


@pytest.mark.gpu
def test_tuning_a_solver_applies_the_setting_restores_the_state_and_caches(tmp_path):
    from ssim_amd import batch, fem, mesh

    cache = str(tmp_path / "tune.json")
    case = mesh.cantilever_bar(16, 4, 4, seed=2)
    s = fem.TetFiniteElementSolver_Hip.from_case(case, rtol=1e-4)
    s.step()
    u_before = s.get_array("deformation")
    best = autotune.tune(s, repeats=1, cache=cache)
    assert {k: s.tuning()[k] for k in best} == best
    assert np.array_equal(s.get_array("deformation"), u_before)
    with open(cache) as f:
        data = json.load(f)
    assert data[autotune.signature(s)]["setting"] == best
    s2 = fem.TetFiniteElementSolver_Hip.from_case(case, rtol=1e-4)
    assert autotune.tune(s2, cache=cache) == best  # from the cache
    s.step()
    s2.step()
    print(f"[gpu] tuned single-mesh setting {best}", flush=True)

    cases = mesh.bar_batch(6, seed=3, lo=3, hi=7)
    B = batch.TetFiniteElementBatch_Hip.from_cases(cases, rtol=1e-4)
    B.step()
    ub = B.get_array("deformation")
    bb = autotune.tune(B, repeats=1, cache=cache)
    assert {k: B.tuning()[k] for k in bb} == bb
    assert np.array_equal(B.get_array("deformation"), ub)
    print(f"[gpu] tuned batch setting {bb}", flush=True)


def test_size_groups_partition_by_tet_count():
    from ssim_amd.batch import size_groups

    nts = [5, 100, 7, 50, 50, 1, 300, 8]
    for g in (1, 2, 3, 4, 8, 20):
        parts = size_groups(nts, g)
        flat = sorted(m for p in parts for m in p)
        assert flat == list(range(len(nts)))  # every mesh once
        assert len(parts) <= min(g, len(nts))
        # size classes: every mesh of an earlier group is no larger than every mesh of a later one
        for a, b in zip(parts, parts[1:]):
            assert max(nts[m] for m in a) <= min(nts[m] for m in b)
    assert size_groups(nts, 1) == [list(range(len(nts)))]
    # a sort key (e.g. each mesh's PCG iterations of an earlier step): classes follow the key, cut at about equal
    # total tet count
    key = [40, 900, 35, 300, 310, 20, 1200, 500]
    for g in (2, 3, 4):
        parts = size_groups(nts, g, key=key)
        assert sorted(m for p in parts for m in p) == list(range(len(nts)))
        for a, b in zip(parts, parts[1:]):
            assert max(key[m] for m in a) <= min(key[m] for m in b)
    with pytest.raises(ValueError):
        size_groups(nts, 2, key=[1, 2])
