"""Parity of the HIP path (through the C ABI) with the CPU FP64 oracle on identical seeded inputs.

Bars: element values (Ψ, P, H, H⁺, K) ≤ 1e-10 relative per element; energy, gradient, Hessian and mass
≤ 1e-10 relative (norm-wise); PCG solutions ≤ 1e-8; positions after N Newton steps ≤ 1e-8 relative to the
oracle's positions (‖x_gpu − x_orc‖∞ / ‖x_orc‖∞ with x = X + u).
"""
import json
import os

import numpy as np
import pytest
import scipy.sparse as sp

import oracle
from helpers import per_item_rel_err, rel_err
from ssim_amd import fem, mesh

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "snh_autograd.json")


def _pair(case, method="newton", rtol=1e-3, **solver):
    g = fem.TetFiniteElementSolver_Hip.from_case(case, method=method, rtol=rtol)
    o = oracle.TimeStep.from_case(case)
    o.set_threshold(rtol)
    if solver:
        g.set_solver_params(**solver)
        o.set_solver(**{k: v for k, v in solver.items() if k in ("max_newton", "pcg_rtol", "pcg_maxit", "ls_c",
                                                                "ls_beta", "ls_max")})
    return g, o


@pytest.fixture(scope="module")
def bar():
    return mesh.cantilever_bar(6, 3, 2, seed=11)


def _strained(case, seed, amp):
    rng = np.random.default_rng(seed)
    u = case.u0 + amp * 0.01 * rng.standard_normal(3 * case.nv)
    return u


@pytest.mark.parametrize("amp", [0.05, 0.4, 1.5])
def test_element_values_match_oracle(bar, amp):
    print(f"[gpu] elements amp={amp} nt={bar.nt}", flush=True)
    g, o = _pair(bar)
    u = _strained(bar, 1, amp)
    ref = o.elements(u, make_spsd=True, want_K=True)
    for spsd in (True, False):
        got = g.eval_elements(u, make_spsd=spsd, want_H=True, want_K=True)
        assert per_item_rel_err(got["psi"][:, None], ref["psi"][:, None]) <= 1e-10 or \
            np.abs(got["psi"] - ref["psi"]).max() <= 1e-10 * np.abs(ref["psi"]).max()
        assert per_item_rel_err(got["P"], ref["P"]) <= 1e-10
        assert per_item_rel_err(got["H"], ref["Hp"] if spsd else ref["H"]) <= 1e-10
        if spsd:
            assert per_item_rel_err(got["K"], ref["K"]) <= 1e-10


def test_projection_on_golden_deformation_gradients():
    """Single tets deformed to the fixture F's (rest, large strain, compressed, inverted, rotated): H⁺ vs
    numpy eigh of the torch-autograd Hessian."""
    with open(GOLDEN) as f:
        gold = json.load(f)
    X, C = mesh.tetra()
    for k, c in enumerate(gold["cases"]):
        F = np.array(c["F"]).reshape(3, 3)
        u = (X @ F.T - X).reshape(-1)
        nu = c["lam"] / (2 * (c["lam"] + c["mu"]))
        E = c["mu"] * 2 * (1 + nu)
        s = fem.TetFiniteElementSolver_Hip(X, C, 1e-2, E, nu, 1.0)
        s.set_material_per_tet([c["lam"]], [c["mu"]])
        got = s.eval_elements(u, make_spsd=True)
        assert rel_err(got["H"][0].reshape(-1), c["Hp"]) <= 1e-10, k
        raw = s.eval_elements(u, make_spsd=False)
        assert rel_err(raw["H"][0].reshape(-1), c["H"]) <= 1e-10, k
        assert rel_err(raw["P"][0].reshape(-1), c["P"]) <= 1e-10, k


def test_energy_gradient_hessian_mass_match_oracle(bar):
    print("[gpu] assembly", flush=True)
    g, o = _pair(bar)
    rng = np.random.default_rng(2)
    u = _strained(bar, 3, 0.3)
    uhat = u + 1e-3 * rng.standard_normal(3 * bar.nv)
    uhat[bar.dof_type == 1] = 0
    g.set_array("next_deform", u)
    g.set_array("inertia", uhat)
    o.set_vectors(u_hat=uhat)
    E = g.update_energy_and_gradients()
    Eo, go = o.energy_grad(u)
    assert abs(E - Eo) <= 1e-10 * abs(Eo)
    assert rel_err(g.forces().T.reshape(-1), go) <= 1e-10
    for spsd in (True, False):
        for filt in (True, False):
            if filt:
                g.update_hessian(spsd)
            else:
                g.update_hessian_unfiltered(spsd)
            Hg = g.hessian()
            Ho = o.hessian_csr(u, spsd, filt)
            assert np.array_equal(Hg.indptr, Ho.indptr) and np.array_equal(Hg.indices, Ho.indices)
            assert rel_err(Hg.data, Ho.data) <= 1e-10
            assert np.array_equal(g.hessian_nonzeros(), Hg.data)
    M = g.mass_matrix()
    assert np.array_equal(M.indptr, Ho.indptr) and np.array_equal(M.indices, Ho.indices)
    assert rel_err(M.data, o.mass_values()) <= 1e-12
    assert abs(g.total_mass() - o.total_mass()) <= 1e-12 * o.total_mass()


def test_gradient_tile_sizes_match_oracle():
    """The tile-reduced energy + gradient with 64-, 128- and 256-tet tiles (SSIM_TUNE_GRAD_TILE) against the oracle
    (≤1e-10) and against each other (≤1e-13: the same contributions summed in other groupings), on a bar whose tets
    fill several tiles of every size; then three Newton steps with 64-tet tiles against the oracle's positions."""
    case = mesh.cantilever_bar(12, 4, 4, seed=2)
    g, o = _pair(case)
    assert case.nt > 4 * 256 and g.tuning()["grad_tile"] == 256
    rng = np.random.default_rng(4)
    u = _strained(case, 6, 0.4)
    uhat = u + 1e-3 * rng.standard_normal(3 * case.nv)
    uhat[case.dof_type == 1] = 0
    g.set_array("next_deform", u)
    g.set_array("inertia", uhat)
    o.set_vectors(u_hat=uhat)
    Eo, go = o.energy_grad(u)
    got = {}
    for tile in (256, 128, 64):
        g.set_tuning(grad_tile=tile)
        print(f"[gpu] gradient tiles of {tile} tets", flush=True)
        got[tile] = (g.update_energy_and_gradients(), g.forces().T.reshape(-1).copy())
        assert abs(got[tile][0] - Eo) <= 1e-10 * abs(Eo)
        assert rel_err(got[tile][1], go) <= 1e-10
    for tile in (128, 64):
        assert abs(got[tile][0] - got[256][0]) <= 1e-13 * abs(got[256][0])
        assert rel_err(got[tile][1], got[256][1]) <= 1e-13
    with pytest.raises(ValueError):
        g.set_tuning(grad_tile=100)
    g2, o2 = _pair(case)
    g2.set_tuning(grad_tile=64)
    for step in range(3):
        g2.step()
        _, st = o2.step()
        assert g2.last_stats()["newton_iterations"] == st["newton"], step
        assert rel_err(_positions(case, g2.deformation().T.reshape(-1)), _positions(case, o2.get_state()[0])) <= 1e-8


def test_rest_shape_and_filtered_mass_match(bar):
    """dminv / rest_vol against numpy (Dm⁻¹ and |det Dm|/6, ΣV = the bar's volume) and mass_matrix_filtered against
    the oracle's mass matrix with identity Dirichlet rows and columns (time_step.hpp:460-466)."""
    g, o = _pair(bar)
    dminv, vol = g.rest_shape()
    X, C = bar.X, bar.cells
    Dm = np.stack([X[C[:, 1]] - X[C[:, 0]], X[C[:, 2]] - X[C[:, 0]], X[C[:, 3]] - X[C[:, 0]]], axis=2)
    assert rel_err(dminv.reshape(-1), np.linalg.inv(Dm).reshape(-1)) <= 1e-13
    assert rel_err(vol, np.abs(np.linalg.det(Dm)) / 6) <= 1e-13
    ext = X.max(axis=0) - X.min(axis=0)
    assert abs(vol.sum() - ext.prod()) <= 1e-12 * ext.prod()
    M = o.mass_csr().tocsr()
    dof = bar.dof_type.astype(bool)
    keep = sp.diags((~dof).astype(float))
    Mf = (keep @ M @ keep + sp.diags(dof.astype(float))).tocsr()
    Gf = g.mass_matrix_filtered()
    assert np.array_equal(Gf.indptr, M.indptr) and np.array_equal(Gf.indices, M.indices)
    assert rel_err(Gf.toarray().reshape(-1), Mf.toarray().reshape(-1)) <= 1e-12
    assert np.array_equal(Gf.diagonal()[dof], np.ones(dof.sum()))


def test_spmv_and_pcg_match_oracle(bar):
    print("[gpu] pcg", flush=True)
    g, o = _pair(bar)
    u = _strained(bar, 5, 0.2)
    g.set_array("next_deform", u)
    g.update_hessian(True)
    H = o.hessian_csr(u, True, True)
    rng = np.random.default_rng(7)
    x = rng.standard_normal(3 * bar.nv)
    assert rel_err(g.spmv(x), H @ x) <= 1e-13
    b = rng.standard_normal(3 * bar.nv)
    xg, itg = g.pcg_solve(b, rtol=1e-8)
    xo, ito, rel = oracle.pcg(H.indptr, H.indices, H.data, b, rtol=1e-8)
    assert abs(itg - ito) <= 1
    assert rel_err(xg, xo) <= 1e-8
    assert np.linalg.norm(H @ xg - b) <= 1.01e-8 * np.linalg.norm(b)
    xz, itz = g.pcg_solve(np.zeros_like(b))
    assert itz == 0 and not xz.any()


def _positions(case, u):
    return (case.X.reshape(-1) + u)


def test_newton_trajectory_config1_60_steps():
    """Configuration 1 of the benchmark (8x8x32 bar, 60 steps, E=1e5, ν=0.45, ρ=1000, dt=1/60): positions after
    every step within 1e-8 of the oracle, and identical Newton counts."""
    case = mesh.cantilever_bar(32, 8, 8, seed=0)
    g, o = _pair(case)
    worst = 0.0
    for step in range(60):
        g.step()
        _, st = o.step()
        sg = g.last_stats()
        assert sg["newton_iterations"] == st["newton"], (step, sg, st)
        xo = _positions(case, o.get_state()[0])
        xg = _positions(case, g.deformation().T.reshape(-1))
        worst = max(worst, rel_err(xg, xo))
        assert worst <= 1e-8, (step, worst)
    print(f"[gpu] config1 60 steps: max position rel err {worst:.3e}", flush=True)


def test_lbfgs_trajectory_matches_oracle():
    # tightly converged steps: L-BFGS iterates are rounding-sensitive inside the tolerance ball
    case = mesh.twisted_cube(n=4, seed=3)
    g, o = _pair(case, method="lbfgs", rtol=1e-10)
    for step in range(3):
        g.step()
        o.step("lbfgs")
        xo = _positions(case, o.get_state()[0])
        xg = _positions(case, g.deformation().T.reshape(-1))
        assert rel_err(xg, xo) <= 1e-8, step


def test_ncg_stepper_reaches_the_oracle_solution():
    """Nonlinear CG (Dai-Kou): GPU and oracle run the same iteration; over thousands of iterations rounding
    differences change the path, so both are compared where they converge (threshold 1e-8 × mass, the GPU
    positions within 1e-6 of the oracle's; the oracle's NCG is pinned to the Newton minimiser)."""
    case = mesh.cantilever_bar(3, 1, 1, h=0.05, E=1e4, nu=0.3, rho=100.0, dt=0.02, seed=2)
    g = fem.TetFiniteElementSolver_Hip.from_case(case, method="ncg", rtol=1e-8)
    g.set_solver_params(lbfgs_maxit=20000)
    o = oracle.TimeStep.from_case(case)
    o.set_threshold(1e-8)
    o.set_lbfgs(max_iter=20000)
    for step in range(2):
        g.step()
        _, st = o.step("ncg")
        sg = g.last_stats()
        assert sg["converged"] and st["converged"], (sg, st)
        assert rel_err(_positions(case, g.deformation().T.reshape(-1)), _positions(case, o.get_state()[0])) <= 1e-6
    print(f"[gpu] NCG: {sg['newton_iterations']} iterations (oracle {st['newton']})", flush=True)


def test_repeat_runs_are_bitwise_identical(bar):
    outs = []
    for _ in range(2):
        g, _o = _pair(bar)
        for _s in range(3):
            g.step()
        outs.append((g.deformation().copy(), g.forces().copy()))
    assert np.array_equal(outs[0][0], outs[1][0]) and np.array_equal(outs[0][1], outs[1][1])


def test_dirichlet_marking_and_release(bar):
    g, o = _pair(bar)
    free_end = np.nonzero(bar.X[:, 0] == bar.X[:, 0].max())[0].astype(np.int32)
    targ = np.zeros((3, free_end.size))
    targ[1] = 0.002
    g.mark_dirichlet_batched(free_end, targ)
    g.reset()
    g.step()
    d = g.deformation()
    assert np.allclose(d[:, free_end], targ, rtol=0, atol=0)
    g.mark_general_batched(free_end)
    g.step()
    assert not np.allclose(g.deformation()[:, free_end], targ)


def test_setters_validate_like_the_reference(bar):
    g, _ = _pair(bar)
    with pytest.raises(ValueError):
        g.set_rtol(0.0)
    with pytest.raises(ValueError):
        g.set_solver_params(ls_c=2.0)
    with pytest.raises(ValueError):
        g.set_material_per_tet(np.zeros(bar.nt), np.ones(bar.nt))


def test_fp32_element_variant_error_is_bounded(bar):
    """FP32 element kernels (FP64 reductions and PCG): per-element stress error vs the FP64 oracle."""
    g, o = _pair(bar)
    g.set_solver_params(fp32_elements=1)
    u = _strained(bar, 9, 0.3)
    got = g.eval_elements(u, make_spsd=True, want_H=True)
    ref = o.elements(u, make_spsd=True)
    errP = per_item_rel_err(got["P"], ref["P"])
    errH = per_item_rel_err(got["H"], ref["Hp"])
    print(f"[gpu] fp32 element variant: P rel err {errP:.2e}, H+ rel err {errH:.2e}", flush=True)
    assert errP <= 1e-4 and errH <= 1e-3


def test_graph_replayed_pcg_is_bitwise_identical(bar):
    out = []
    for graph in (0, 1, 1):
        g, _ = _pair(bar)
        g.set_tuning(pcg_graph=graph, pcg_chunk_max=16)
        for _ in range(3):
            g.step()
        out.append(g.deformation().copy())
    assert np.array_equal(out[0], out[1]) and np.array_equal(out[0], out[2])


def _unit_box_case(nx=6, ny=5, nz=9, seed=4):
    """The reference's own 5-tet unit_box (mesh/common_shapes.hpp:67-132): irregular vertex degrees, so its
    rows do not fit the fixed-width layout and the looped / four-lane SpMV variants run."""
    X, cells = mesh.unit_box_rows(nx, ny, nz)
    nv, nt = X.shape[0], cells.shape[0]
    lam, mu = mesh.lame(np.full(nt, 1e5), np.full(nt, 0.4))
    clamped = X[:, 0] == X[:, 0].min()
    dof = np.repeat(clamped.astype(np.uint8), 3)
    rng = np.random.default_rng(seed)
    u0 = 0.002 * rng.standard_normal(3 * nv)
    u0[dof == 1] = 0.0
    accel = np.zeros(3 * nv)
    accel[1::3] = -9.8
    return mesh.BarCase(X, cells, lam, mu, 1000.0, 1.0 / 60, dof, np.zeros(3 * nv), accel, u0, np.zeros(3 * nv),
                        name="unit_box")


@pytest.mark.parametrize("lanes", [1, 4])
def test_irregular_mesh_spmv_variants_match_oracle(lanes):
    case = _unit_box_case()
    g, o = _pair(case)
    assert g.tuning()["spmv_lanes"] == 1  # rows too wide for the fixed-width layout
    with pytest.raises(ValueError):
        g.set_tuning(spmv_lanes=0)
    g.set_tuning(spmv_lanes=lanes)
    print(f"[gpu] unit_box {case.nv} vertices, SpMV variant {lanes}", flush=True)
    u = _strained(case, 3, 0.3)
    g.set_array("next_deform", u)
    g.update_hessian(True)
    H = o.hessian_csr(u, True, True)
    x = np.random.default_rng(1).standard_normal(3 * case.nv)
    assert rel_err(g.spmv(x), H @ x) <= 1e-13
    g2, o2 = _pair(case)
    g2.set_tuning(spmv_lanes=lanes)
    for step in range(5):
        g2.step()
        _, st = o2.step()
        assert g2.last_stats()["newton_iterations"] == st["newton"], step
        assert rel_err(_positions(case, g2.deformation().T.reshape(-1)), _positions(case, o2.get_state()[0])) <= 1e-8


def _spread_case(case, gap=70_000):
    """The same bar with `gap` Dirichlet-fixed vertices in no tet inserted in the middle of the numbering: the
    bar's middle edges then join vertices more than 65,535 apart, which the fixed layout's 16-bit offsets cannot
    hold."""
    nv = case.nv
    half = nv // 2
    new = np.arange(nv) + np.where(np.arange(nv) < half, 0, gap)
    N = nv + gap

    def spread(a, fill=0.0):
        out = np.full((N, 3), fill, dtype=a.dtype)
        out[new] = a.reshape(nv, 3)
        return out.reshape(-1)

    X = spread(case.X.reshape(-1)).reshape(N, 3)
    dof = spread(case.dof_type, 1).astype(np.uint8)
    return mesh.BarCase(X, new[case.cells].astype(case.cells.dtype), case.lam, case.mu, case.rho, case.dt, dof,
                        spread(case.dbc), spread(case.accel), spread(case.u0), spread(case.v0),
                        name=case.name + "-spread"), new


def test_fixed_and_general_layouts_agree(bar):
    """A Freudenthal bar gets the fixed layout (7 + 7 off-diagonal slots, 6-value diagonals, 16-bit offsets);
    the same bar renumbered with column offsets beyond 65,535 gets the general layout.  The two SpMVs agree to
    rounding on the bar's rows (identity rows elsewhere), so do the exported Hessians (the fixed layout keeps the
    upper triangle of each diagonal block), and three Newton steps take the same iterations and
    agree within 1e-10 (positions)."""
    big, new = _spread_case(bar)
    gs = fem.TetFiniteElementSolver_Hip.from_case(bar)
    gb = fem.TetFiniteElementSolver_Hip.from_case(big)
    print(f"[gpu] layouts: bar {bar.nv} vertices, spread {big.nv} vertices", flush=True)
    assert gs.layout_info()["fixed_layout"] == 1 and gs.tuning()["spmv_lanes"] == 0
    assert gb.layout_info()["fixed_layout"] == 0 and gb.tuning()["spmv_lanes"] == 1
    with pytest.raises(ValueError):
        gs.set_tuning(spmv_lanes=1)
    u = _strained(bar, 3, 0.3)
    ub = np.zeros(3 * big.nv)
    ub.reshape(-1, 3)[new] = u.reshape(-1, 3)
    gs.set_array("next_deform", u)
    gb.set_array("next_deform", ub)
    gs.update_hessian(True)
    gb.update_hessian(True)
    x = np.random.default_rng(2).standard_normal(3 * big.nv)
    ys, yb = gs.spmv(x.reshape(-1, 3)[new].reshape(-1)), gb.spmv(x)
    assert rel_err(yb.reshape(-1, 3)[new].reshape(-1), ys) <= 1e-13
    rest = np.setdiff1d(np.arange(big.nv), new)
    assert np.array_equal(yb.reshape(-1, 3)[rest], x.reshape(-1, 3)[rest])
    Hs, Hb = gs.hessian(), gb.hessian()
    perm = (3 * new[:, None] + np.arange(3)).reshape(-1)
    assert rel_err(Hb[perm][:, perm].toarray().reshape(-1), Hs.toarray().reshape(-1)) <= 1e-13
    for step in range(3):
        gs.step()
        gb.step()
        assert gs.last_stats()["newton_iterations"] == gb.last_stats()["newton_iterations"], step
    ds = gs.deformation().T.reshape(-1)
    db = gb.deformation().T.reshape(-1, 3)[new].reshape(-1)
    assert rel_err(_positions(bar, db), _positions(bar, ds)) <= 1e-10


def test_lbfgs_gram_and_vector_recursions_agree():
    """The two-loop recursion on Gram matrices (the default) and vector by vector are the same arithmetic up to
    the order of the sums: at threshold 1e-8 they take the same iterations and agree within 1e-9 (positions).
    (Near the limit of double precision, threshold 1e-10 on this cube, both iterate erratically.)"""
    case = mesh.twisted_cube(n=4, seed=3)
    runs = []
    for gram in (1, 0):
        g = fem.TetFiniteElementSolver_Hip.from_case(case, method="lbfgs", rtol=1e-8)
        g.set_tuning(lbfgs_gram=gram)
        its = []
        for _ in range(3):
            g.step()
            its.append(g.last_stats()["newton_iterations"])
        runs.append((_positions(case, g.deformation().T.reshape(-1)), its))
    assert rel_err(runs[0][0], runs[1][0]) <= 1e-9
    assert all(abs(a - b) <= 2 for a, b in zip(runs[0][1], runs[1][1])), (runs[0][1], runs[1][1])
    print(f"[gpu] L-BFGS iterations per step, Gram {runs[0][1]} vs vector {runs[1][1]}", flush=True)


def test_lbfgs_after_failed_line_searches_matches_oracle():
    """A strict Armijo constant (lbfgs_c = 0.5) and at most 4 trials per line search (α down to 1/8): the search
    fails now and then on this cube at threshold 1e-9 (the oracle counts 5, 14 and 12 failures over the 3 steps and
    converges), and a failed search still takes its last, shortest step along a descent direction.  After its
    first failure the Gram-form solve continues vector by vector on the same history, from an iterate that may sit
    in the line search's trial buffers at that point.  Positions agree with the oracle run with the same parameters
    within 2e-8 (the oracle's own 1e-9 solutions lie within 5e-9 of its 1e-10 ones), and the gradient left in
    `forces` is the converged one.  (With one trial per iteration every step is taken at α = 1 even when it raises
    the energy, and the path depends on the last bits: the oracle converges there and the GPU's path overflows,
    which the solve reports as a numeric error.)"""
    case = mesh.twisted_cube(n=4, seed=3)
    g, o = _pair(case, method="lbfgs", rtol=1e-9)
    g.set_solver_params(lbfgs_c=0.5, lbfgs_ls_max=4)
    o.set_lbfgs(ls_c=0.5, ls_max=4)
    fails = []
    for step in range(3):
        g.step()
        o.step("lbfgs")
        st = g.last_stats()
        fails.append(st["ls_failures"])
        assert st["converged"], step
        assert np.linalg.norm(g.forces()) < g.grad_threshold_abs(), step
        xo = _positions(case, o.get_state()[0])
        xg = _positions(case, g.deformation().T.reshape(-1))
        assert rel_err(xg, xo) <= 2e-8, step
    print(f"[gpu] L-BFGS with lbfgs_c = 0.5, 4 trials: line-search failures per step {fails}", flush=True)
    assert sum(fails) > 0
