"""Oracle checks of the 2D material models (Ndim = 2 of elast/stable_neohookean.hpp and elast/linear.hpp).

Pins: the reference's own 2D element tests (tests/fem_intro/main.cpp:155-300: rest-state known answers and
central differences at its test point), frame indifference Ψ(RF) = Ψ(F), P(RF) = R·P(F) (a wrong determinant or
cofactor breaks it) and symmetric Hessians.  The δ of the 2D stable Neo-Hookean model (the positive root of the
reference's quadratic, elast/stable_neohookean.hpp:96-107) is pinned by the property that quadratic encodes: with
the stress-free α, the Hessian at the collapsed state F = 0 has zero curvature along pure shear
(test_stable_neohookean_2d_delta_is_the_zero_shear_curvature_root).
"""
import numpy as np
import pytest

import oracle
from helpers import rel_err
from ssim_amd import mesh

F0 = np.array([[1.0, 0.3], [0.2, 3.0]])  # tests/fem_intro/main.cpp:178, 252


@pytest.mark.parametrize("model,E,nu", [("stable_neohookean", 1e6, 0.3), ("linear", 1e6, 0.4)])
def test_rest_state_known_answers(model, E, nu):
    lam, mu = mesh.lame(E, nu)
    e, P, _ = oracle.material2(model, np.eye(2), lam, mu)
    assert abs(e) < 1e-7 and np.abs(P).max() < 1e-7


@pytest.mark.parametrize("model,E,nu,d", [("stable_neohookean", 1e6, 0.3, 1e-6), ("linear", 1e6, 0.4, 1e-3)])
def test_finite_differences_at_reference_test_point(model, E, nu, d):
    lam, mu = mesh.lame(E, nu)
    _, P, H = oracle.material2(model, F0, lam, mu)
    assert np.abs(H - H.T).max() <= 1e-15 * np.abs(H).max()
    for i in range(2):
        for j in range(2):
            Fp, Fm = F0.copy(), F0.copy()
            Fp[i, j] += d
            Fm[i, j] -= d
            fd = (oracle.energy2(model, Fp, lam, mu) - oracle.energy2(model, Fm, lam, mu)) / (2 * d)
            assert abs(P[i, j] - fd) <= 1e-6 * np.abs(P).max()
            dP = (oracle.material2(model, Fp, lam, mu)[1] - oracle.material2(model, Fm, lam, mu)[1]) / (2 * d)
            assert rel_err(H[2 * i + j], dP.reshape(-1)) <= 1e-6


def test_stable_neohookean_2d_is_frame_indifferent():
    lam, mu = mesh.lame(1e5, 0.35)
    rng = np.random.default_rng(3)
    for _ in range(20):
        th = rng.uniform(0, 2 * np.pi)
        R = np.array([[np.cos(th), -np.sin(th)], [np.sin(th), np.cos(th)]])
        F = np.eye(2) + 0.4 * rng.standard_normal((2, 2))
        e, P, _ = oracle.material2("stable_neohookean", F, lam, mu)
        er, Pr, _ = oracle.material2("stable_neohookean", R @ F, lam, mu)
        assert abs(er - e) <= 1e-9 * max(1.0, abs(e))
        assert rel_err(Pr, R @ P) <= 1e-12
        # a pure rotation carries no energy
        assert abs(oracle.energy2("stable_neohookean", R, lam, mu)) <= 1e-9 * mu


@pytest.mark.parametrize("E,nu", [(1e5, 0.3), (1e6, 0.45), (2e4, 0.1), (1e6, 0.49), (3e5, 0.01)])
def test_stable_neohookean_2d_delta_is_the_zero_shear_curvature_root(E, nu):
    """The reference's 2D δ, the positive root of (2t+1)δ² + (2t+2)δ − 2t = 0 with t = μ/λ
    (elast/stable_neohookean.hpp:96-107), is the δ for which the Hessian at the collapsed state F = 0 has zero
    curvature along pure shear.  At F = 0 the Hessian is μ(1 − 1/δ)·I − λα·∂²J/∂F², and ∂²J/∂F² maps a symmetric
    traceless S to −S (J(S) = −|S|²/2), so H(0)·vec S = (μ(1 − 1/δ) + λα)·vec S; with the stress-free
    α = 1 + t(1 − 1/(2+δ)), μ(1 − 1/δ) + λα = 0 multiplied by δ(2+δ)/λ is the reference's quadratic.  The check
    uses the oracle's Hessian and stress functions, not its δ: H(0) annihilates both pure-shear directions,
    P(εS)/ε falls like ε² (P(εS)/ε = ε·[λ(α + ε²) + μ(1 − 1/(2ε² + δ))]/ε·S, whose ε → 0 limit is the curvature
    above; with a δ off by 1 % it would stay near 1e-2·μ/δ), the other two curvatures at F = 0 are nonzero, and the
    energy is finite at F = 0 (δ > 0, the root the reference takes)."""
    lam, mu = mesh.lame(E, nu)
    e0, P0, H0 = oracle.material2("stable_neohookean", np.zeros((2, 2)), lam, mu)
    assert np.isfinite(e0) and np.abs(P0).max() == 0.0
    scale = np.abs(H0).max()
    shears = (np.array([[1.0, 0.0], [0.0, -1.0]]), np.array([[0.0, 1.0], [1.0, 0.0]]))
    for S in shears:
        assert np.abs(H0 @ S.reshape(-1)).max() <= 1e-12 * scale
        slope = []
        for eps in (1e-3, 1e-4):
            _, P, _ = oracle.material2("stable_neohookean", eps * S, lam, mu)
            slope.append(np.abs(P).max() / eps)
        assert slope[1] <= 0.02 * slope[0]  # ∝ ε²: a factor 100 per decade (a wrong δ: a factor ~1)
    # the complement (rotation and dilation directions) keeps a nonzero curvature: the check is not vacuous
    for D in (np.array([[0.0, 1.0], [-1.0, 0.0]]), np.eye(2)):
        assert np.abs(H0 @ D.reshape(-1)).max() >= 1e-3 * scale
