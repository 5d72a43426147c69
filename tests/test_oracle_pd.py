"""Oracle checks of the projective-dynamics stepper (time_step_solver_pd, stepper/projective_dynamics.hpp:9-88;
bound as tet_lbfgs_pd, bindings/src/fem_host.cpp:223).

Pins: the matrix M + dt²κK against the closed form on the reference tet (K = V·G·Gᵀ with the P1 gradients G),
zero row sums and the patch test of K (K·x = 0 on interior rows for linear x), symmetry and the Dirichlet
identity rows; the first L-BFGS step against −A⁻¹g from a sparse direct solve; convergence to the Newton
minimiser.  The diffusivity κ follows the reference's uniform_elasticity(), which builds the material from
(E, ν) in the (λ, μ) slots (finite_elements/time_step.hpp:356), so κ = λ + 5μ/6 + 2·(4μ/3) of the stable
Neo-Hookean model reads E + 3.5ν.
"""
import numpy as np
import scipy.sparse.linalg as spla

import oracle
from helpers import rel_err
from ssim_amd import mesh


def pd_kappa(E, nu, model="stable_neohookean"):
    """approximated_diffusivity of ElastModel{E, ν}, i.e. λ = E, μ = ν: stable Neo-Hookean overrides it with
    λ + 5μ/6 + 2·(4μ/3) (elast/stable_neohookean.hpp:109-114), linear elasticity keeps λ + 2μ
    (elast/basic_elast.hpp:225)."""
    if model == "linear":
        return E + 2.0 * nu
    return (E + 5.0 / 6.0 * nu) + 2.0 * (4.0 / 3.0 * nu)


def test_pd_matrix_closed_form_on_the_reference_tet():
    X, C = mesh.tetra()
    lam, mu = mesh.lame(1e5, 0.3)
    rho, dt, kappa = 50.0, 0.1, 7.0
    ts = oracle.TimeStep(X, C, dt, np.array([lam]), np.array([mu]), rho)
    ts.build_pd(kappa)
    A = ts.pd_csr().toarray()
    V = 1.0 / 6.0
    G = np.array([[-1.0, -1.0, -1.0], [1.0, 0, 0], [0, 1.0, 0], [0, 0, 1.0]])
    K4 = V * G @ G.T
    M4 = rho * V / 20.0 * (np.ones((4, 4)) + np.eye(4))
    expect = np.kron(M4 + dt * dt * kappa * K4, np.eye(3))
    assert np.abs(A - expect).max() <= 1e-14 * np.abs(expect).max()


def test_pd_matrix_invariants():
    X, C = mesh.unit_box_rows(4, 3, 5)
    lam, mu = mesh.lame(1e5, 0.3)
    rho, dt, kappa = 100.0, 0.05, 3.0
    ts = oracle.TimeStep(X, C, dt, np.full(len(C), lam), np.full(len(C), mu), rho)
    ts.build_pd(kappa)
    A = ts.pd_csr()
    K = (A - ts.mass_csr()) / (dt * dt * kappa)
    assert abs(A - A.T).max() == 0.0
    ones = np.ones(A.shape[0])
    assert np.abs(K @ ones).max() <= 1e-12 * abs(K).max()
    lo, hi = X.min(axis=0), X.max(axis=0)
    interior = np.all((X > lo + 1e-12) & (X < hi - 1e-12), axis=1)
    assert interior.any()
    for a in range(3):
        lin = np.zeros(A.shape[0])
        lin[a::3] = X[:, a] * 2.0 - X[:, (a + 1) % 3]
        r = (K @ lin).reshape(-1, 3)
        assert np.abs(r[interior]).max() <= 1e-12 * abs(K).max()
    # Dirichlet rows and columns become the identity
    dof = np.zeros(A.shape[0], dtype=np.uint8)
    dof[:9] = 1
    ts.set_dirichlet(dof, np.zeros(A.shape[0]))
    ts.build_pd(kappa)
    Af = ts.pd_csr().toarray()
    assert np.array_equal(Af[:9, :9], np.eye(9)) and not Af[:9, 9:].any() and not Af[9:, :9].any()


def _case():
    return mesh.cantilever_bar(3, 1, 1, h=0.05, E=1e4, nu=0.3, rho=100.0, dt=0.02, seed=2)


def _ts(case, rtol=1e-8, max_iter=5000):
    ts = oracle.TimeStep.from_case(case)
    ts.set_solver(max_newton=50, pcg_rtol=1e-12)
    ts.set_lbfgs(m=8, max_iter=max_iter)
    ts.set_threshold(rtol)
    ts.build_pd(pd_kappa(1e4, 0.3))
    ts.prepare_step()
    return ts


def test_first_lbfgs_pd_step_is_the_preconditioned_gradient():
    case = _case()
    ts = _ts(case, max_iter=1)
    _, uh, _ = ts.get_vectors()
    _, g0 = ts.energy_grad(uh)
    ts.compute_step("lbfgs_pd")
    step = ts.get_vectors()[0] - uh
    d0 = -spla.spsolve(ts.pd_csr().tocsc(), g0)
    alpha = (step @ d0) / (d0 @ d0)
    assert rel_err(step, alpha * d0) <= 1e-10
    k = np.log2(1.0 / alpha)
    assert abs(k - round(k)) < 1e-9 and round(k) >= 0


def test_lbfgs_pd_reaches_the_newton_minimiser():
    case = _case()
    newton, pd = _ts(case), _ts(case)
    _, sn = newton.compute_step("newton")
    _, sp_ = pd.compute_step("lbfgs_pd")
    assert sn["converged"] and sp_["converged"], sp_
    X = case.X.reshape(-1)
    assert rel_err(X + pd.get_vectors()[0], X + newton.get_vectors()[0]) <= 1e-6


def test_lbfgs_h0_first_step_is_the_newton_direction():
    """time_step_solver_h0 (stepper/projective_dynamics.hpp:91-148): the two-loop middle step solves with the
    SPD Hessian at the start of the step, so the first step is −α·H⁻¹g (sparse direct solve), and the iteration
    reaches the Newton minimiser."""
    case = _case()
    ts = _ts(case, max_iter=1)
    _, uh, _ = ts.get_vectors()
    _, g0 = ts.energy_grad(uh)
    H = ts.hessian_csr(uh, True, True)
    ts.compute_step("lbfgs_h0")
    step = ts.get_vectors()[0] - uh
    d0 = -spla.spsolve(H.tocsc(), g0)
    alpha = (step @ d0) / (d0 @ d0)
    assert rel_err(step, alpha * d0) <= 1e-10
    newton, h0 = _ts(case), _ts(case)
    newton.compute_step("newton")
    _, s = h0.compute_step("lbfgs_h0")
    assert s["converged"]
    X = case.X.reshape(-1)
    assert rel_err(X + h0.get_vectors()[0], X + newton.get_vectors()[0]) <= 1e-6
