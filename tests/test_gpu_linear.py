"""Linear elasticity (elast/linear.hpp) and the forward-Euler stepper (finite_elements/time_step.hpp:717-736)
on MI355X against the CPU oracle.

Bars as for the stable Neo-Hookean path: element values (Ψ, P, H, H⁺, K) ≤ 1e-10 relative per element;
energy, gradient and Hessian ≤ 1e-10 (norm-wise); positions after every step ≤ 1e-8 relative with identical
Newton counts; the batched path within 1e-8 of the single-mesh path; the adjoint step within 1e-8 of the
oracle's direct solve.  Forward Euler is one gradient evaluation and an axpy, so its trajectory is held to 1e-12.
"""
import ctypes

import numpy as np
import pytest

import oracle
from helpers import per_item_rel_err, rel_err
from ssim_amd import _lib, batch, fem, mesh

pytestmark = pytest.mark.gpu


def _pair(case, method="newton", rtol=1e-3, model="linear"):
    g = fem.TetFiniteElementSolver_Hip.from_case(case, method=method, rtol=rtol, material=model)
    o = oracle.TimeStep.from_case(case)
    o.set_model(model)
    o.set_threshold(rtol)
    return g, o


def _x(case, u):
    return case.X.reshape(-1) + np.asarray(u).reshape(-1)


@pytest.mark.parametrize("amp", [0.05, 1.5])
def test_linear_element_values_match_oracle(amp):
    case = mesh.cantilever_bar(6, 3, 2, seed=11)
    print(f"[gpu] linear elements amp={amp} nt={case.nt}", flush=True)
    g, o = _pair(case)
    rng = np.random.default_rng(1)
    u = case.u0 + amp * 0.01 * rng.standard_normal(3 * case.nv)
    ref = o.elements(u, make_spsd=True, want_K=True)
    for spsd in (True, False):
        got = g.eval_elements(u, make_spsd=spsd, want_H=True, want_K=True)
        assert np.abs(got["psi"] - ref["psi"]).max() <= 1e-10 * np.abs(ref["psi"]).max()
        assert per_item_rel_err(got["P"], ref["P"]) <= 1e-10
        assert per_item_rel_err(got["H"], ref["Hp"] if spsd else ref["H"]) <= 1e-10
        if spsd:
            assert per_item_rel_err(got["K"], ref["K"]) <= 1e-10


def test_linear_assembly_matches_oracle():
    case = mesh.cantilever_bar(6, 3, 2, seed=12)
    g, o = _pair(case)
    rng = np.random.default_rng(2)
    u = case.u0 + 0.003 * rng.standard_normal(3 * case.nv)
    uhat = u + 1e-3 * rng.standard_normal(3 * case.nv)
    uhat[case.dof_type == 1] = 0
    g.set_array("next_deform", u)
    g.set_array("inertia", uhat)
    o.set_vectors(u_hat=uhat)
    E = g.update_energy_and_gradients()
    Eo, go = o.energy_grad(u)
    assert abs(E - Eo) <= 1e-10 * abs(Eo)
    assert rel_err(g.forces().T.reshape(-1), go) <= 1e-10
    for spsd in (True, False):
        g.update_hessian(spsd)
        Hg = g.hessian()
        Ho = o.hessian_csr(u, spsd, True)
        assert rel_err(Hg.data, Ho.data) <= 1e-10


def test_linear_newton_trajectory_matches_oracle():
    case = mesh.cantilever_bar(16, 4, 4, seed=0)
    g, o = _pair(case)
    worst = 0.0
    for step in range(20):
        g.step()
        _, st = o.step()
        assert g.last_stats()["newton_iterations"] == st["newton"], step
        worst = max(worst, rel_err(_x(case, g.get_array("deformation")), _x(case, o.get_state()[0])))
        assert worst <= 1e-8, (step, worst)
    print(f"[gpu] linear model 20 steps: max position rel err {worst:.3e}", flush=True)


@pytest.mark.parametrize("model", ["stable_neohookean", "linear"])
def test_forward_euler_trajectory_matches_oracle(model):
    case = mesh.cantilever_bar(8, 2, 2, seed=4)
    g, o = _pair(case, method="forward_euler", model=model)
    for step in range(10):
        r = g.step()
        gn, st = o.step("forward_euler")
        assert g.last_stats()["newton_iterations"] == 1
        assert rel_err(_x(case, g.get_array("deformation")), _x(case, o.get_state()[0])) <= 1e-12, step
        assert rel_err(g.get_array("velocity"), o.get_state()[2]) <= 1e-9, step
    assert g.last_stats()["grad_norm"] == pytest.approx(gn, rel=1e-9)


def test_material_switch_takes_effect_and_rejects_unknown_models():
    case = mesh.cantilever_bar(4, 2, 2, seed=6)
    g = fem.TetFiniteElementSolver_Hip.from_case(case)
    rng = np.random.default_rng(3)
    u = case.u0 + 0.01 * rng.standard_normal(3 * case.nv)
    a = g.eval_elements(u)["psi"]
    _lib.call("ssim_set_material_model", g.handle, 1)
    b = g.eval_elements(u)["psi"]
    o = oracle.TimeStep.from_case(case)
    o.set_model("linear")
    assert np.abs(b - o.elements(u)["psi"]).max() <= 1e-10 * np.abs(b).max()
    assert not np.allclose(a, b)
    m = ctypes.c_int()
    _lib.call("ssim_get_material_model", g.handle, ctypes.byref(m))
    assert m.value == 1
    with pytest.raises(ValueError):
        _lib.call("ssim_set_material_model", g.handle, 7)
    with pytest.raises(ValueError):
        fem.TetFiniteElementSolver_Hip(case.X, case.cells, material="corotated")


def test_linear_batch_matches_single_mesh_path():
    cases = mesh.bar_batch(4, seed=17, lo=3, hi=8)
    B = batch.TetFiniteElementBatch_Hip.from_cases(cases, rtol=1e-3, material="linear")
    singles = [fem.TetFiniteElementSolver_Hip.from_case(c, rtol=1e-3, material="linear") for c in cases]
    for step in range(4):
        B.step()
        for m, (c, s) in enumerate(zip(cases, singles)):
            s.step()
            assert B.last_stats(m)["newton_iterations"] == s.last_stats()["newton_iterations"], (step, m)
            assert rel_err(_x(c, B.get_array("deformation", m)), _x(c, s.get_array("deformation"))) <= 1e-8


def test_linear_adjoint_step_matches_oracle():
    torch = pytest.importorskip("torch")
    case = mesh.cantilever_bar(5, 2, 3, seed=2, E=1e5, nu=0.45)
    g = fem.TetFiniteElementSolver_Hip.from_case(case, rtol=1e-8, material="linear")
    g.set_solver_params(pcg_rtol=1e-12, newton_polish=2)
    g.step()
    u_star = g.get_array("next_deform")
    rng = np.random.default_rng(4)
    gu = rng.standard_normal(3 * case.nv)
    dev = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device="cuda")
    ptr = lambda t: ctypes.c_void_p(t.data_ptr())
    gu_d = dev(gu)
    gh_d = torch.empty(3 * case.nv, dtype=torch.float64, device="cuda")
    lv_d = torch.empty(3 * case.nv, dtype=torch.float64, device="cuda")
    gl_d = torch.empty(case.nt, dtype=torch.float64, device="cuda")
    gm_d = torch.empty(case.nt, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    _lib.call("ssim_adjoint_step", g.handle, ptr(gu_d), ptr(gh_d), ptr(gl_d), ptr(gm_d), ptr(lv_d), None)
    o = oracle.TimeStep.from_case(case)
    o.set_model("linear")
    lv, guhat, gl, gm = o.adjoint_step(u_star, gu)
    errs = [rel_err(lv_d.cpu().numpy(), lv), rel_err(gh_d.cpu().numpy(), guhat),
            rel_err(gl_d.cpu().numpy(), gl), rel_err(gm_d.cpu().numpy(), gm)]
    print(f"[gpu] linear adjoint step vs oracle: {['%.2e' % e for e in errs]}", flush=True)
    assert max(errs) <= 1e-8
