"""Oracle checks of the P1 triangle elements in 2D (orc_tri_eval): basic_time_step with PhysicalDim 2,
TopologyDim 3, element by element (finite_elements/def_grad.hpp:53-160, rest_vol.hpp:33-50, time_step.hpp:41-83,
bilinears.hpp:27-91, 191-264).  The reference instantiates no 2D time step, so these pins come from mathematics:

* area and Dm⁻¹ against numpy (|det|/2, numpy.linalg.inv); F of an affine displacement is its matrix;
* the mass against ρA/12·(1 + δ_ab), the Laplacian against the cotangent formula −½·cot θ (off-diagonal) with
  zero row sums — both independent of the reference-triangle tables the oracle integrates with;
* the gradient against central differences of the energy, zero net force and torque; Ψ(rest) = 0;
* the unprojected stiffness against central differences of the gradient; the projected one against
  A·Bᵀ·H⁺·B with B = ∂F/∂x from finite differences of F (exact: F is linear in x) and H⁺ from numpy.linalg.eigh
  clamped at 1e-3;
* rotation invariance of the stable Neo-Hookean energy; NaN for a triangle with an out-of-range index.
"""
import numpy as np
import pytest

import oracle
from helpers import rel_err

MODELS = ["stable_neohookean", "linear"]


def _mesh(nx=5, ny=4, seed=0, amp=0.2):
    """A grid over [0, 1] × [0, 0.8] with jittered interior vertices, split into triangles (alternating
    diagonals), x-slowest vertex ids."""
    rng = np.random.default_rng(seed)
    xs, ys = np.meshgrid(np.linspace(0, 1, nx + 1), np.linspace(0, 0.8, ny + 1), indexing="ij")
    X = np.stack([xs.ravel(), ys.ravel()], axis=1)
    inner = (xs.ravel() > 0) & (xs.ravel() < 1) & (ys.ravel() > 0) & (ys.ravel() < 0.8)
    X[inner] += rng.uniform(-0.25, 0.25, (inner.sum(), 2)) * (1.0 / nx)  # the boundary stays the rectangle's
    cells = []
    vid = lambda i, j: i * (ny + 1) + j
    for i in range(nx):
        for j in range(ny):
            a, b, c, d = vid(i, j), vid(i + 1, j), vid(i + 1, j + 1), vid(i, j + 1)
            cells += [[a, b, c], [a, c, d]] if (i + j) % 2 == 0 else [[a, b, d], [b, c, d]]
    u = rng.standard_normal(X.shape) * amp * (1.0 / nx)
    return X, np.array(cells, dtype=np.int32), u


def _one(X, c, u, model, lam=2e4, mu=3e3, **kw):
    return oracle.tri_eval(X, np.asarray(c).reshape(1, 3), u, lam, mu, model=model, **kw)


def test_rest_shape_and_affine_deformation():
    X, C, _ = _mesh()
    G = np.array([[1.2, 0.3], [-0.1, 0.9]])
    t = np.array([0.5, -2.0])
    u = X @ G.T + t - X
    o = oracle.tri_eval(X, C, u, 1e4, 1e3)
    for e, c in enumerate(C):
        Dm = np.stack([X[c[1]] - X[c[0]], X[c[2]] - X[c[0]]], axis=1)
        assert abs(o["area"][e] - abs(np.linalg.det(Dm)) / 2) <= 1e-15
        assert rel_err(o["dminv"][e], np.linalg.inv(Dm)) <= 1e-13
    assert np.abs(o["F"] - G).max() <= 1e-13
    assert abs(o["area"].sum() - 0.8) <= 1e-13  # the triangles tile the rectangle

def test_mass_and_laplacian_closed_forms():
    X, C, _ = _mesh(seed=3)
    rho = 7.5
    o = oracle.tri_eval(X, C, np.zeros_like(X), 1e4, 1e3, density=rho)
    for e, c in enumerate(C):
        A = o["area"][e]
        assert np.abs(o["mass"][e] - rho * A / 12 * (np.ones((3, 3)) + np.eye(3))).max() <= 1e-14 * rho
        L = o["laplacian"][e]
        p = X[c]
        for a in range(3):
            i, j = (a + 1) % 3, (a + 2) % 3
            e1, e2 = p[i] - p[a], p[j] - p[a]
            cot = (e1 @ e2) / abs(e1[0] * e2[1] - e1[1] * e2[0])
            assert abs(L[i, j] + 0.5 * cot) <= 1e-12 * np.abs(L).max()
        assert np.abs(L.sum(axis=1)).max() <= 1e-12 * np.abs(L).max()
        assert np.array_equal(L, L.T)


@pytest.mark.parametrize("model", MODELS)
def test_energy_gradient_and_stiffness_by_finite_differences(model):
    X, C, u = _mesh(nx=3, ny=2, seed=1, amp=0.3)
    h = 1e-6
    for e, c in enumerate(C):
        Xe = X[c]
        ue = u[c].copy()
        o = _one(Xe, [0, 1, 2], ue, model)
        g, K = o["grad"][0].reshape(6), o["stiff"][0]
        gfd, Kfd = np.zeros(6), np.zeros((6, 6))
        for k in range(6):
            up, um = ue.copy(), ue.copy()
            up.reshape(6)[k] += h
            um.reshape(6)[k] -= h
            op, om = _one(Xe, [0, 1, 2], up, model), _one(Xe, [0, 1, 2], um, model)
            gfd[k] = (op["energy"][0] - om["energy"][0]) / (2 * h)
            Kfd[:, k] = (op["grad"][0].reshape(6) - om["grad"][0].reshape(6)) / (2 * h)
        assert rel_err(g, gfd) <= 1e-6, (e, model)
        assert rel_err(K, Kfd) <= 1e-6, (e, model)
        assert rel_err(K, K.T) <= 1e-14
        # no net force, no net torque (the energy is invariant to translations and, at first order, rotations)
        gv = g.reshape(3, 2)
        x = Xe + ue
        scale = np.abs(gv).max()
        assert np.abs(gv.sum(axis=0)).max() <= 1e-10 * scale
        if model == "stable_neohookean":
            torque = np.sum(x[:, 0] * gv[:, 1] - x[:, 1] * gv[:, 0])
            assert abs(torque) <= 1e-10 * scale


@pytest.mark.parametrize("model", MODELS)
def test_projected_stiffness_against_numpy_eigh(model):
    X = np.array([[0.0, 0.0], [1.0, 0.1], [0.2, 0.9]])
    # an inverted and strongly compressed state, where the clamp changes the Hessian
    for u in (np.array([[0.0, 0.0], [-0.7, 0.0], [0.0, -1.5]]), np.array([[0.1, 0.0], [0.3, -0.2], [0.0, 0.4]])):
        o = _one(X, [0, 1, 2], u, model, make_spsd=True)
        lam, mu = 2e4, 3e3
        B = np.zeros((4, 6))
        for k in range(6):
            up = u.copy()
            up.reshape(6)[k] += 1.0
            B[:, k] = (_one(X, [0, 1, 2], up, model)["F"][0] - _one(X, [0, 1, 2], u, model)["F"][0]).reshape(4)
        _, _, H = oracle.material2(model, o["F"][0], lam, mu)
        w, Q = np.linalg.eigh(H)
        Hp = Q @ np.diag(np.maximum(w, 1e-3)) @ Q.T
        expect = o["area"][0] * B.T @ Hp @ B
        assert rel_err(o["stiff"][0], expect) <= 1e-11, model
        assert np.linalg.eigvalsh(o["stiff"][0]).min() >= -1e-9 * np.abs(expect).max()
    if model == "stable_neohookean":
        unproj = _one(X, [0, 1, 2], u, model)["stiff"][0]
        assert rel_err(unproj, o["stiff"][0]) > 1e-3  # the clamp did act on the last state


def test_rest_energy_and_rotation_invariance():
    X, C, u = _mesh(seed=5)
    for model in MODELS:
        o = oracle.tri_eval(X, C, np.zeros_like(X), 1e4, 1e3, model=model)
        assert np.abs(o["energy"]).max() <= 1e-9 * 1e4 * o["area"].max()
    th = 0.7
    R = np.array([[np.cos(th), -np.sin(th)], [np.sin(th), np.cos(th)]])
    x = X + u
    o1 = oracle.tri_eval(X, C, u, 1e4, 1e3)
    o2 = oracle.tri_eval(X, C, x @ R.T + np.array([3.0, 1.0]) - X, 1e4, 1e3)
    assert rel_err(o2["energy"], o1["energy"]) <= 1e-12


def test_out_of_range_index_gives_nan():
    X, C, u = _mesh()
    C = C.copy()
    C[3, 1] = len(X)
    C[5, 0] = -1
    o = oracle.tri_eval(X, C, u, 1e4, 1e3, make_spsd=True)
    for k, v in o.items():
        assert np.isnan(v[3]).all() and np.isnan(v[5]).all(), k
        assert np.isfinite(v[np.r_[0:3, 4, 6:len(C)]]).all(), k
