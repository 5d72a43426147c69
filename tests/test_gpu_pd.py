"""The projective-dynamics stepper ("lbfgs_pd", the reference's tet_lbfgs_pd, bindings/src/fem_host.cpp:223,
stepper/projective_dynamics.hpp:9-88) on MI355X against the oracle.

Bars: the matrix M + dt²κK within 1e-12 of the oracle's (same CSR pattern); the first step (−A⁻¹g, inner
Jacobi-PCG to 1e-12 on both sides) within 1e-9; converged positions within 1e-6 of the oracle's Newton
minimiser, with iteration counts of GPU and oracle within 3.
"""
import numpy as np
import pytest

import oracle
from helpers import rel_err
from ssim_amd import fem, mesh
from test_oracle_pd import pd_kappa

pytestmark = pytest.mark.gpu


def _case():
    return mesh.cantilever_bar(4, 2, 2, h=0.05, E=1e4, nu=0.3, rho=100.0, dt=0.02, seed=2)


def _e_nu(case):
    # the (E, ν) from_case hands to the constructor
    lam0, mu0 = float(case.lam[0]), float(case.mu[0])
    nu = lam0 / (2 * (lam0 + mu0))
    return mu0 * 2 * (1 + nu), nu


def _pair(case, maxit=5000, rtol=1e-8):
    g = fem.TetFiniteElementSolver_Hip.from_case(case, method="lbfgs_pd", rtol=rtol)
    g.set_solver_params(lbfgs_maxit=maxit)
    g.reset()
    o = oracle.TimeStep.from_case(case)
    o.set_threshold(rtol)
    o.set_lbfgs(max_iter=maxit)
    o.build_pd(pd_kappa(*_e_nu(case)))
    return g, o


def test_pd_matrix_matches_oracle():
    case = _case()
    g, o = _pair(case)
    Ag, Ao = g.pd_matrix(), o.pd_csr()
    assert np.array_equal(Ag.indptr, Ao.indptr) and np.array_equal(Ag.indices, Ao.indices)
    assert rel_err(Ag.data, Ao.data) <= 1e-12


def test_first_step_matches_oracle():
    case = _case()
    g, o = _pair(case, maxit=1)
    g.prepare_step()
    g.compute_step()
    o.prepare_step()
    o.compute_step("lbfgs_pd")
    uh = o.get_vectors()[1]
    assert rel_err(g.get_array("next_deform") - uh, o.get_vectors()[0] - uh) <= 1e-9


def test_lbfgs_pd_converges_like_the_oracle():
    case = _case()
    g, o = _pair(case)
    g.prepare_step()
    g.compute_step()
    sg = g.last_stats()
    o.prepare_step()
    _, so = o.compute_step("lbfgs_pd")
    n = oracle.TimeStep.from_case(case)
    n.set_threshold(1e-8)
    n.set_solver(max_newton=50, pcg_rtol=1e-12)
    n.prepare_step()
    n.compute_step("newton")
    assert sg["converged"] and so["converged"]
    X = case.X.reshape(-1)
    assert rel_err(X + g.get_array("next_deform"), X + n.get_vectors()[0]) <= 1e-6
    assert abs(sg["newton_iterations"] - so["newton"]) <= 3, (sg["newton_iterations"], so["newton"])
    print(f"[gpu] lbfgs_pd: {sg['newton_iterations']} iterations, {sg['pcg_iterations']} inner PCG iterations "
          f"(oracle {so['newton']}, {so['pcg']})", flush=True)


def test_trajectory_with_the_reset_built_matrix():
    """Several steps: the matrix built at reset is reused; positions stay within 1e-6 of the oracle's."""
    case = _case()
    g, o = _pair(case, rtol=1e-6)
    X = case.X.reshape(-1)
    for step in range(3):
        g.step()
        o.step("lbfgs_pd")
        assert rel_err(X + g.get_array("deformation"), X + o.get_state()[0]) <= 1e-6, step


def test_lbfgs_h0_matches_oracle():
    case = _case()
    for maxit in (1, 5000):
        g = fem.TetFiniteElementSolver_Hip.from_case(case, method="lbfgs_h0", rtol=1e-8)
        g.set_solver_params(lbfgs_maxit=maxit)
        o = oracle.TimeStep.from_case(case)
        o.set_threshold(1e-8)
        o.set_lbfgs(max_iter=maxit)
        g.prepare_step()
        g.compute_step()
        o.prepare_step()
        _, so = o.compute_step("lbfgs_h0")
        uh = o.get_vectors()[1]
        sg = g.last_stats()
        if maxit == 1:
            assert rel_err(g.get_array("next_deform") - uh, o.get_vectors()[0] - uh) <= 1e-9
        else:
            assert sg["converged"] and so["converged"]
            X = case.X.reshape(-1)
            assert rel_err(X + g.get_array("next_deform"), X + o.get_vectors()[0]) <= 1e-6
            assert abs(sg["newton_iterations"] - so["newton"]) <= 3
            print(f"[gpu] lbfgs_h0: {sg['newton_iterations']} iterations (oracle {so['newton']})", flush=True)


def test_pd_matrix_follows_the_material_model():
    """κ is the model's approximated diffusivity: E + 3.5ν for stable Neo-Hookean, E + 2ν for linear elasticity."""
    case = _case()
    E, nu = _e_nu(case)
    for model in ("stable_neohookean", "linear"):
        g = fem.TetFiniteElementSolver_Hip.from_case(case, method="lbfgs_pd", material=model)
        g.reset()
        o = oracle.TimeStep.from_case(case)
        o.build_pd(pd_kappa(E, nu, model))
        assert rel_err(g.pd_matrix().data, o.pd_csr().data) <= 1e-12, model
