"""Oracle checks of the linear-elasticity model (elast/linear.hpp) and the forward-Euler stepper
(finite_elements/time_step.hpp:717-736).

Pins: the reference's own linear3 test (tests/fem_intro/main.cpp:13-84: rest-state known answers and finite
differences at its test point), an identity against the independently pinned stable Neo-Hookean code (its
Hessian at F = I is the linear Hessian with Lamé (λ − 5μ/8, 3μ/4)), exact quadratic-form identities, the SPD
projection's spectrum from numpy's eigh, the implicit step against a sparse direct solve (the incremental
potential is quadratic, so one linear solve is the exact step) and the adjoint against central differences.
"""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

import oracle
from helpers import rel_err
from ssim_amd import mesh


def test_linear_rest_state_known_answers():
    # tests/fem_intro/main.cpp:14-33: Ψ(I) = 0 and P(I) = 0 (E = 1e6, ν = 0.4)
    lam, mu = mesh.lame(1e6, 0.4)
    e, P = oracle.lin_stress(np.eye(3), lam, mu)
    assert e == 0.0 and np.abs(P).max() == 0.0


def test_linear_finite_differences_at_reference_test_point():
    # tests/fem_intro/main.cpp:35-83 (same point, E, ν; central differences, exact for a quadratic energy up to
    # roundoff, so the bars are far below the reference's 1e-3)
    lam, mu = mesh.lame(1e6, 0.4)
    F0 = np.array([[1, 0.3, 0.2], [0.2, 2, 0.1], [0.1, 0.3, 1]])
    _, P = oracle.lin_stress(F0, lam, mu)
    H = oracle.lin_hessian(lam, mu)
    d = 1e-3
    for i in range(3):
        for j in range(3):
            Fp, Fm = F0.copy(), F0.copy()
            Fp[i, j] += d
            Fm[i, j] -= d
            fd = (oracle.lin_energy(Fp, lam, mu) - oracle.lin_energy(Fm, lam, mu)) / (2 * d)
            assert abs(P[i, j] - fd) <= 1e-8 * np.abs(P).max()
            dP = (oracle.lin_stress(Fp, lam, mu)[1] - oracle.lin_stress(Fm, lam, mu)[1]) / (2 * d)
            assert rel_err(H[3 * i + j], dP.reshape(-1)) <= 1e-9


def test_linear_hessian_is_the_rest_hessian_of_stable_neohookean():
    """At F = I the stable Neo-Hookean Hessian (pinned by its own golden vectors) is
    0.75μ·(I + S) + (λ − 5μ/8)·t tᵀ, the linear Hessian with Lamé (λ − 5μ/8, 3μ/4); a dropped transposition term,
    a misplaced λ block or a wrong factor on either fails this."""
    for E, nu in ((1e6, 0.3), (1e5, 0.45), (3e4, 0.2)):
        lam, mu = mesh.lame(E, nu)
        H_lin = oracle.lin_hessian(lam - 5 * mu / 8, 0.75 * mu)
        H_snh = oracle.snh_hessian(np.eye(3), lam, mu)
        assert rel_err(H_lin, H_snh) <= 1e-14


def test_linear_energy_and_stress_are_the_quadratic_form_of_the_hessian():
    rng = np.random.default_rng(7)
    lam, mu = mesh.lame(1e6, 0.4)
    H = oracle.lin_hessian(lam, mu)
    assert np.array_equal(H, H.T)
    for _ in range(20):
        D = 0.3 * rng.standard_normal((3, 3))
        e, P = oracle.lin_stress(np.eye(3) + D, lam, mu)
        d = D.reshape(-1)
        assert abs(e - 0.5 * d @ H @ d) <= 1e-12 * abs(e)
        assert rel_err(P.reshape(-1), H @ d) <= 1e-12
        assert oracle.lin_energy(np.eye(3) + D, lam, mu) == e


@pytest.mark.parametrize("lam,mu", [(5.7e5, 3.6e5), (1e-4, 1e-4), (2e-4, 3e-4)])
def test_linear_projection_spectrum(lam, mu):
    """Eigenvalues of the linear Hessian: 0 on the skew matrices (3), 2μ on the trace-free symmetric ones (5),
    2μ + 3λ on multiples of I; the projection raises those below the 1e-3 floor to it and keeps the eigenvectors."""
    H = oracle.lin_hessian(lam, mu)
    expect = np.sort(np.maximum(np.array([0.0] * 3 + [2 * mu] * 5 + [2 * mu + 3 * lam]), 1e-3))
    Hp = oracle.spd_project9(H)
    w = np.linalg.eigvalsh(Hp)
    assert np.abs(w - expect).max() <= 1e-12 * max(1.0, expect.max())
    # eigenvectors kept: the skew matrix basis is an eigenspace of the projected matrix
    K = np.zeros((3, 3))
    K[0, 1], K[1, 0] = 1.0, -1.0
    k = K.reshape(-1)
    assert np.abs(Hp @ k - 1e-3 * k).max() <= 1e-14 * max(1.0, np.abs(H).max())


def _linear_case(seed=2):
    case = mesh.cantilever_bar(3, 1, 1, h=0.05, E=1e4, nu=0.3, rho=100.0, dt=0.02, seed=seed)
    return case


def test_linear_implicit_step_is_one_linear_solve():
    """With a quadratic incremental potential the exact step is û − H⁻¹g(û) (H unprojected, Dirichlet-filtered):
    the Newton stepper (projected Hessian, PCG, Armijo) must reach it."""
    case = _linear_case()
    ts = oracle.TimeStep.from_case(case)
    ts.set_model("linear")
    ts.set_solver(max_newton=50, pcg_rtol=1e-12)
    ts.set_threshold(1e-11)
    ts.prepare_step()
    _, u_hat, _ = ts.get_vectors()
    _, g = ts.energy_grad(u_hat)
    H = ts.hessian_csr(u_hat, make_spsd=False, filter_dbc=True)
    u_star = u_hat - spla.spsolve(H.tocsc(), g)
    gn, st = ts.compute_step()
    assert st["converged"]
    u_next, _, _ = ts.get_vectors()
    X = case.X.reshape(-1)
    assert rel_err(X + u_next, X + u_star) <= 1e-10
    free = case.dof_type == 0
    assert np.array_equal(u_next[~free], u_hat[~free])


def test_forward_euler_step():
    """next = û + ½dt²·forces(û) (time_step.hpp:717-736), forces the Dirichlet-filtered gradient; a rigid
    translation has no elastic force and zero inertia residual at the predictor, so it stays on the predictor."""
    for model in ("stable_neohookean", "linear"):
        case = _linear_case(seed=5)
        ts = oracle.TimeStep.from_case(case)
        ts.set_model(model)
        ts.prepare_step()
        _, u_hat, _ = ts.get_vectors()
        _, g = ts.energy_grad(u_hat)
        gn, st = ts.compute_step("forward_euler")
        u_next, _, _ = ts.get_vectors()
        assert np.array_equal(u_next, u_hat + 0.5 * case.dt ** 2 * g)
        assert gn == pytest.approx(np.linalg.norm(ts.energy_grad(u_next)[1]), rel=1e-14)
        free = case.dof_type == 0
        assert np.array_equal(u_next[~free], u_hat[~free])
    X, C = mesh.unit_box_rows(3, 3, 3)
    lam, mu = mesh.lame(1e5, 0.3)
    ts = oracle.TimeStep(X, C, 1e-2, np.full(len(C), lam), np.full(len(C), mu), 1e3)
    ts.set_ext_accel(np.tile([0.0, 0.0, -9.8], len(X)))
    ts.set_state(np.tile([0.1, -0.2, 0.3], len(X)), None, np.tile([1.0, 0.5, 0.0], len(X)))
    ts.prepare_step()
    _, u_hat, _ = ts.get_vectors()
    ts.compute_step("forward_euler")
    u_next, _, _ = ts.get_vectors()
    assert np.abs(u_next - u_hat).max() <= 1e-15


def test_linear_rollout_adjoint_matches_central_differences():
    """∂L/∂(log E) and ∂L/∂v0 of a rollout under the linear model (∂P/∂λ = tr(ε)·I, ∂P/∂μ = 2ε) against central
    differences of the rollout loss."""
    case = mesh.cantilever_bar(4, 2, 2, seed=3, E=1e5, nu=0.45)
    rng = np.random.default_rng(0)
    case.v0 = 0.01 * rng.standard_normal(case.v0.shape)
    case.v0[case.dof_type == 1] = 0
    fe = np.nonzero(case.X[:, 0] == case.X[:, 0].max())[0]
    t = np.array([0.0, 0.012, 0.016])

    def loss_grad(u):
        d = u.reshape(-1, 3)[fe] - t
        g = np.zeros((case.nv, 3))
        g[fe] = 2 * d
        return float(np.sum(d * d)), g.reshape(-1)

    steps, h = 3, 1e-6
    _, g = oracle.rollout_gradient(case, steps, loss_grad, model="linear")
    for k in range(2):
        d = rng.standard_normal(case.nt)
        # λ and μ separately: a mistake in either derivative shows
        for which in ("lam", "mu"):
            lp = dict(lam=case.lam, mu=case.mu)
            lm = dict(lam=case.lam, mu=case.mu)
            lp[which] = getattr(case, which) * np.exp(h * d)
            lm[which] = getattr(case, which) * np.exp(-h * d)
            fd = (oracle.rollout_loss(case, steps, loss_grad, model="linear", **lp) -
                  oracle.rollout_loss(case, steps, loss_grad, model="linear", **lm)) / (2 * h)
            an = (getattr(case, which) * g[which]) @ d
            assert abs(fd - an) <= 1e-5 * abs(fd), (which, fd, an)
    dv = rng.standard_normal(case.nv * 3)
    dv[case.dof_type == 1] = 0
    fd = (oracle.rollout_loss(case, steps, loss_grad, v0=case.v0 + h * dv, model="linear") -
          oracle.rollout_loss(case, steps, loss_grad, v0=case.v0 - h * dv, model="linear")) / (2 * h)
    assert abs(fd - g["v0"] @ dv) <= 1e-5 * abs(fd)
