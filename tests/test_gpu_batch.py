"""Batched stepping on MI355X: every mesh of a batch against the single-mesh path and the oracle.

Bars: positions after every step within 1e-8 of the single-mesh HIP path (‖x_batch − x_single‖∞ /
‖x_single‖∞, x = X + u) with identical Newton iteration counts per mesh; the same bar against the CPU oracle
within 1e-8.  The batched and single-mesh PCGs sum their dot products over different workgroup shapes, so the
two agree to rounding, not bitwise.
"""
import numpy as np
import pytest

import oracle
from helpers import rel_err
from ssim_amd import batch, fem, mesh

pytestmark = pytest.mark.gpu


def _x(case, u):
    return case.X.reshape(-1) + np.asarray(u).reshape(-1)


def _single(case, rtol):
    return fem.TetFiniteElementSolver_Hip.from_case(case, rtol=rtol)


def test_batch_matches_single_mesh_path_per_mesh():
    cases = mesh.bar_batch(7, seed=3, lo=3, hi=9)
    print("[gpu] batch vs single:", [c.name for c in cases], flush=True)
    B = batch.TetFiniteElementBatch_Hip.from_cases(cases, rtol=1e-3)
    singles = [_single(c, 1e-3) for c in cases]
    worst = 0.0
    for step in range(6):
        B.step()
        for m, (c, s) in enumerate(zip(cases, singles)):
            s.step()
            assert B.last_stats(m)["newton_iterations"] == s.last_stats()["newton_iterations"], (step, m)
            err = rel_err(_x(c, B.get_array("deformation", m)), _x(c, s.get_array("deformation")))
            worst = max(worst, err)
            assert err <= 1e-8, (step, m, err)
    tot = B.last_stats()
    assert tot["newton_iterations"] == max(B.last_stats(m)["newton_iterations"] for m in range(len(cases)))
    print(f"[gpu] batch of {len(cases)}: max position rel err vs single-mesh path {worst:.3e}", flush=True)


def test_batch_matches_oracle():
    cases = mesh.bar_batch(3, seed=11, lo=3, hi=6)
    B = batch.TetFiniteElementBatch_Hip.from_cases(cases, rtol=1e-3)
    orcs = []
    for c in cases:
        o = oracle.TimeStep.from_case(c)
        o.set_threshold(1e-3)
        orcs.append(o)
    for step in range(4):
        B.step()
        for m, (c, o) in enumerate(zip(cases, orcs)):
            _, st = o.step()
            assert B.last_stats(m)["newton_iterations"] == st["newton"], (step, m)
            xo = _x(c, o.get_state()[0])
            assert rel_err(_x(c, B.get_array("deformation", m)), xo) <= 1e-8, (step, m)


def test_meshes_converging_at_different_iterations():
    """One bar at rest without load converges at the predictor (0 Newton iterations) while the others
    iterate: the finished mesh must stay untouched by the later lock-step iterations."""
    cases = mesh.bar_batch(4, seed=21, lo=3, hi=7)
    quiet = cases[1]
    quiet.accel[:] = 0.0
    quiet.u0[:] = 0.0
    B = batch.TetFiniteElementBatch_Hip.from_cases(cases, rtol=1e-3)
    B.step()
    assert B.last_stats(1)["newton_iterations"] == 0
    assert max(B.last_stats(m)["newton_iterations"] for m in range(4)) > 0
    assert not B.get_array("deformation", 1).any()
    s = _single(cases[2], 1e-3)
    s.step()
    assert rel_err(_x(cases[2], B.get_array("deformation", 2)), _x(cases[2], s.get_array("deformation"))) <= 1e-8


def test_per_mesh_dirichlet_targets():
    cases = mesh.bar_batch(3, seed=5, lo=3, hi=6)
    B = batch.TetFiniteElementBatch_Hip.from_cases(cases, rtol=1e-3)
    c = cases[2]
    free_end = np.nonzero(c.X[:, 0] == c.X[:, 0].max())[0].astype(np.int32)
    targ = np.zeros((3, free_end.size))
    targ[2] = 0.003
    B.mark_dirichlet_batched(2, free_end, targ)
    B.reset()
    B.step()
    d = B.deformation(2)
    assert np.array_equal(d[:, free_end], targ)
    other = B.deformation(1)
    assert not np.allclose(other[2], 0.003)
    with pytest.raises(ValueError):
        B.mark_dirichlet_batched(3, free_end, targ)
    with pytest.raises(ValueError):
        B.mark_general_batched(0, np.array([c.nv + 10**6], dtype=np.int32))


def test_graph_replay_and_compaction_give_identical_results():
    """PCG chunks replayed from captured graphs (many replays per solve, with the host's own polls between them) and
    compacted grids produce bitwise the same trajectory as plain launches over the full grid."""
    cases = mesh.bar_batch(5, seed=13, lo=3, hi=9)
    runs = []
    for graph, compact in ((0, 0), (1, 70), (1, 100), (0, 70)):
        B = batch.TetFiniteElementBatch_Hip.from_cases(cases, rtol=1e-4)
        B.set_tuning(pcg_graph=graph, compact_pct=compact, pcg_chunk_max=16)
        for _ in range(3):
            B.step()
        runs.append(B.get_array("deformation"))
    for r in runs[1:]:
        assert np.array_equal(r, runs[0])


@pytest.mark.parametrize("concurrent", [True, False])
def test_concurrent_groups_match_the_single_batch(concurrent):
    """BatchGroups steps size classes of the batch concurrently on their own streams from their own host threads,
    or groups keyed by each mesh's PCG iterations of a first step one after another.  Every mesh takes the same
    Newton iterations as in the single batch and lands within 1e-8 of its positions (the bar of the batch tests):
    the blocked-ELL slot order of a row follows the column offsets that most rows of the whole batch share, so a
    row's products are summed in an order that depends on the batch's other meshes, and the PCG solutions differ by
    rounding amplified up to its tolerance."""
    cases = mesh.bar_batch(24, seed=5, lo=3, hi=10)
    B = batch.TetFiniteElementBatch_Hip.from_cases(cases, rtol=1e-4)
    if concurrent:
        G = batch.BatchGroups(cases, groups=3, rtol=1e-4)
    else:
        P = batch.TetFiniteElementBatch_Hip.from_cases(cases, rtol=1e-4)
        P.step()
        key = [P.last_stats(m)["pcg_iterations"] for m in range(len(cases))]
        G = batch.BatchGroups(cases, groups=3, rtol=1e-4, key=key, concurrent=False)
        assert not hasattr(G, "_streams")
    assert sorted(m for ms in G.members for m in ms) == list(range(len(cases)))
    for step in range(3):
        B.step()
        G.step()
        for m in range(len(cases)):
            assert G.last_stats(m)["newton_iterations"] == B.last_stats(m)["newton_iterations"], (step, m)
            xg = _x(cases[m], G.get_array("deformation", m))
            assert rel_err(xg, _x(cases[m], B.get_array("deformation", m))) <= 1e-8, (step, m)
