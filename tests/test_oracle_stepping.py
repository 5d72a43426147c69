"""Pins the oracle's time stepping (prepare_step, step_next, gravity, damping and the dt² scaling of the elastic
energy) by closed forms that do not restate its code.

* Free fall: an unclamped bar released from rest under gravity g moves rigidly (F = I, no elastic force), so
  backward Euler gives u_n = ½·dt²·g·n(n+1) and v_n = n·dt·g (finite_elements/time_step.hpp:499-543); with
  velocity damping d, v_{n+1} = (1 − d)(v_n + dt·g), whose closed form is v_n = v*(1 − (1 − d)^n) with
  v* = (1 − d)·dt·g/d.
* Static equilibrium: with the linear model the incremental potential is exactly ½(u − û)ᵀM(u − û) + dt²·½uᵀKu,
  so a clamped, damped bar settles where K_ff u_f = M_ff g_f, independent of dt.  K and M are assembled here
  densely from textbook P1 formulas (isotropic linear elasticity through the shape-function gradients, consistent
  mass ρV(1 + δ_ab)/20), not from the oracle's routines.  A dt or a dt³ in place of dt² on either term moves the
  fixed point by a factor 60 at dt = 1/60.
"""
import numpy as np
import pytest

import oracle
from ssim_amd import mesh


def dense_linear_system(case):
    """Dense stiffness K and consistent mass M of a P1 tet mesh with the linear model (per-tet λ, μ)."""
    X, cells = case.X, case.cells
    n = 3 * case.nv
    K = np.zeros((n, n))
    M = np.zeros((n, n))
    for e, c in enumerate(cells):
        Dm = np.stack([X[c[i]] - X[c[0]] for i in (1, 2, 3)], axis=1)
        B = np.linalg.inv(Dm)
        V = abs(np.linalg.det(Dm)) / 6.0
        G = np.zeros((4, 3))  # rows: ∇N_a
        G[1:] = B
        G[0] = -B.sum(axis=0)
        lam, mu = case.lam[e], case.mu[e]
        for a in range(4):
            for b in range(4):
                Kab = V * (mu * (G[a] @ G[b]) * np.eye(3) + mu * np.outer(G[b], G[a]) + lam * np.outer(G[a], G[b]))
                K[3 * c[a]:3 * c[a] + 3, 3 * c[b]:3 * c[b] + 3] += Kab
                M[3 * c[a]:3 * c[a] + 3, 3 * c[b]:3 * c[b] + 3] += case.rho * V * (2.0 if a == b else 1.0) / 20.0 * np.eye(3)
    return K, M


def equilibrium(case):
    """u* with K_ff u_f = M_ff g_f and u_d = 0."""
    K, M = dense_linear_system(case)
    free = case.dof_type == 0
    u = np.zeros(3 * case.nv)
    u[free] = np.linalg.solve(K[np.ix_(free, free)], M[np.ix_(free, free)] @ case.accel[free])
    return u


def stiff_bar():
    """4×2×2 cubes, E = 1e6, ν = 0.3: the slowest mode has dt²ω² ≈ 45 at dt = 1/60, so the damped bar settles
    within a few tens of steps."""
    return mesh.cantilever_bar(4, 2, 2, h=0.01, E=1e6, nu=0.3, rho=1000.0, jitter=0.0)


def free_bar(gravity=-9.8):
    case = mesh.cantilever_bar(4, 2, 2, h=0.01, E=1e5, nu=0.45, rho=1000.0, jitter=0.0, gravity=gravity)
    case.dof_type[:] = 0
    return case


def free_fall_closed_form(n, dt, g, damping):
    """(u_n, v_n) of a rigid body released from rest under acceleration g."""
    if damping == 0.0:
        return 0.5 * dt * dt * g * n * (n + 1), n * dt * g
    q = 1.0 - damping
    vstar = q * dt * g / damping
    v = vstar * (1.0 - q ** n)
    # u_n = Σ_{k<n} dt·(v_k + dt·g) = n·dt²·g + dt·Σ_{k<n} v_k,  Σ_{k<n} v_k = v*·(n − (1 − q^n)/(1 − q))
    u = n * dt * dt * g + dt * vstar * (n - (1.0 - q ** n) / (1.0 - q))
    return u, v


@pytest.mark.parametrize("damping", [0.0, 0.3])
def test_free_fall_matches_closed_form(damping):
    case = free_bar()
    ts = oracle.TimeStep.from_case(case)
    ts.set_damping(damping)
    g = case.accel.reshape(-1, 3)[0]  # (0, −9.8, 0) on every vertex
    for n in range(1, 21):
        _, st = ts.step()
        assert st["newton"] == 0 and st["converged"], (n, st)  # rigid motion: the predictor already solves
        u, _, v = ts.get_state()
        uc, vc = free_fall_closed_form(n, case.dt, g, damping)
        np.testing.assert_allclose(u.reshape(-1, 3), np.broadcast_to(uc, (case.nv, 3)), rtol=1e-12, atol=1e-18)
        np.testing.assert_allclose(v.reshape(-1, 3), np.broadcast_to(vc, (case.nv, 3)), rtol=1e-11, atol=1e-16)


def test_damping_outside_its_range_is_rejected():
    ts = oracle.TimeStep.from_case(free_bar())
    for bad in (-0.1, 1.0, 2.0):
        with pytest.raises(ValueError):
            ts.set_damping(bad)


def test_damped_linear_bar_settles_at_static_equilibrium():
    case = stiff_bar()
    ustar = equilibrium(case)
    ts = oracle.TimeStep.from_case(case)
    ts.set_model("linear")
    ts.set_damping(0.9)
    ts.set_threshold(1e-11)
    ts.set_solver(pcg_rtol=1e-12)
    for _ in range(30):
        ts.step()
    u, _, v = ts.get_state()
    err = np.abs(u - ustar).max() / np.abs(ustar).max()
    assert err <= 1e-9, err
    # the check separates the dt² scaling from its plausible slips: a dt (or dt³) on the elastic energy, or on
    # gravity in the predictor, puts the fixed point at K u = (dt or 1/dt)·M g
    for wrong in (ustar * case.dt, ustar / case.dt):
        assert np.abs(u - wrong).max() / np.abs(ustar).max() > 0.5
    assert np.abs(v).max() <= 1e-9 * np.abs(ustar).max() / case.dt
