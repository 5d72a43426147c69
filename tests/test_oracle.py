"""Pins the CPU FP64 oracle (oracle/ssim_oracle.c) before the HIP path is tested against it.

Each check is independent of the oracle's own formulas: reference known-answer tests, torch autograd fixtures
(tests/golden/snh_autograd.json), numpy eigh, central finite differences, closed forms, dense brute force and
scipy solvers.
"""
import json
import os

import numpy as np
import pytest
import scipy.optimize
import scipy.sparse.linalg

import oracle
from helpers import rel_err
from ssim_amd import mesh

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "snh_autograd.json")


@pytest.fixture(scope="module")
def golden():
    with open(GOLDEN) as f:
        return json.load(f)


# ---- element level ------------------------------------------------------------------------------------

def test_snh_rest_state_known_answers():
    # tests/fem_intro/main.cpp:100-106: Ψ(I) = 0 and P(I) = 0 (E = 1e6, ν = 0.3)
    lam, mu = oracle.lib().orc_lambda(1e6, 0.3), oracle.lib().orc_mu(1e6, 0.3)
    e, P = oracle.snh_stress(np.eye(3), lam, mu)
    assert abs(e) < 1e-7 and np.abs(P).max() < 1e-7


def test_lame_parameters_closed_form():
    # elast/basic_elast.hpp:11-20
    assert oracle.lib().orc_mu(1e5, 0.45) == pytest.approx(1e5 / 2.9, rel=1e-15)
    assert oracle.lib().orc_lambda(1e5, 0.45) == pytest.approx(1e5 * 0.45 / (1.45 * 0.1), rel=1e-15)


def test_snh_energy_stress_hessian_match_autograd(golden):
    for c in golden["cases"]:
        F = np.array(c["F"])
        e, P = oracle.snh_stress(F, c["lam"], c["mu"])
        assert oracle.snh_energy(F, c["lam"], c["mu"]) == e
        scale = max(abs(c["psi"]), c["mu"])
        assert abs(e - c["psi"]) <= 1e-12 * scale
        assert rel_err(P.reshape(-1), c["P"]) <= 1e-12
        H = oracle.snh_hessian(F, c["lam"], c["mu"])
        assert rel_err(H.reshape(-1), c["H"]) <= 1e-12


def test_snh_finite_differences_at_reference_test_point():
    # tests/fem_intro/main.cpp:86-154 (same point, E, ν, step; tighter tolerances in FP64)
    lam, mu = oracle.lib().orc_lambda(1e6, 0.3), oracle.lib().orc_mu(1e6, 0.3)
    F0 = np.array([[1, 0.3, 0.2], [0.2, 2, 0.1], [0.1, 0.3, 1]])
    _, P = oracle.snh_stress(F0, lam, mu)
    H = oracle.snh_hessian(F0, lam, mu)
    d = 1e-6
    for i in range(3):
        for j in range(3):
            Fp, Fm = F0.copy(), F0.copy()
            Fp[i, j] += d
            Fm[i, j] -= d
            fd = (oracle.snh_energy(Fp, lam, mu) - oracle.snh_energy(Fm, lam, mu)) / (2 * d)
            assert abs(P[i, j] - fd) <= 1e-3
            dP = (oracle.snh_stress(Fp, lam, mu)[1] - oracle.snh_stress(Fm, lam, mu)[1]) / (2 * d)
            # H(3i+j, 3k+l) = ∂P_kl/∂F_ij
            assert rel_err(H[3 * i + j], dP.reshape(-1)) <= 1e-6


def test_sym_eig_is_an_eigendecomposition():
    rng = np.random.default_rng(0)
    for n in (3, 9):
        A = rng.standard_normal((n, n))
        A = A + A.T
        w, Q = oracle.sym_eig(A)
        assert np.abs(Q.T @ Q - np.eye(n)).max() < 1e-13
        assert np.abs(A @ Q - Q * w).max() < 1e-12 * np.abs(A).max()
        np.testing.assert_allclose(np.sort(w), np.linalg.eigvalsh(A), atol=1e-12 * np.abs(A).max())


def test_spd_projection_matches_eigh_fixtures(golden):
    for c in golden["cases"]:
        Hp = oracle.spd_project9(np.array(c["H"]), golden["floor"])
        assert rel_err(Hp.reshape(-1), c["Hp"]) <= 1e-11


def test_spd_projection_invariants():
    rng = np.random.default_rng(1)
    # already SPD with λ_min ≥ floor: unchanged
    A = rng.standard_normal((9, 9))
    H = A @ A.T + np.eye(9)
    assert rel_err(oracle.spd_project9(H), H) <= 1e-13
    # indefinite: eigenvalues clamped at the floor, eigenspaces kept
    H = A + A.T
    w, Q = np.linalg.eigh(H)
    Hp = oracle.spd_project9(H, 1e-3)
    wp = np.linalg.eigvalsh(Hp)
    np.testing.assert_allclose(np.sort(wp), np.sort(np.maximum(w, 1e-3)), atol=1e-12 * np.abs(w).max())
    assert np.abs(Q.T @ Hp @ Q - np.diag(np.maximum(w, 1e-3))).max() < 1e-12 * np.abs(w).max()


def test_local_force_and_stiffness_are_chain_rule_of_F():
    # f = pfpxᵀ vec P and K = pfpxᵀ H pfpx with pfpx = ∂F/∂x checked by finite differences of F(x)
    rng = np.random.default_rng(2)
    X = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float) + 0.1 * rng.standard_normal((4, 3))
    Dm = np.stack([X[1] - X[0], X[2] - X[0], X[3] - X[0]], axis=1)
    B = np.linalg.inv(Dm)

    def F_of(x):
        Ds = np.stack([x[1] - x[0], x[2] - x[0], x[3] - x[0]], axis=1)
        return Ds @ B

    x = X + 0.05 * rng.standard_normal((4, 3))
    pf = np.zeros((9, 12))
    d = 1e-7
    for v in range(4):
        for b in range(3):
            xp, xm = x.copy(), x.copy()
            xp[v, b] += d
            xm[v, b] -= d
            pf[:, 3 * v + b] = ((F_of(xp) - F_of(xm)) / (2 * d)).reshape(-1)
    ref_pf = np.zeros(108)
    oracle.lib().orc_pfpx(B.reshape(-1).ctypes.data_as(oracle._D), ref_pf.ctypes.data_as(oracle._D))
    assert np.abs(ref_pf.reshape(9, 12) - pf).max() < 1e-6
    lam, mu = mesh.lame(1e5, 0.4)
    _, P = oracle.snh_stress(F_of(x), lam, mu)
    f = oracle.local_force(P, B)
    assert rel_err(f.reshape(-1), pf.T @ P.reshape(-1)) <= 1e-6
    H = oracle.snh_hessian(F_of(x), lam, mu)
    K = oracle.local_stiffness(H, B)
    assert rel_err(K, pf.T @ H @ pf) <= 1e-6


# ---- problem level ----------------------------------------------------------------------------------

def _ts(X, cells, dt=1e-2, E=1e6, nu=0.33, rho=1e3):
    lam, mu = mesh.lame(E, nu)
    return oracle.TimeStep(X, cells, dt, np.full(len(cells), lam), np.full(len(cells), mu), rho)


def test_rest_state_deformation_gradient_and_volume():
    # tests/fem_intro/main.cpp:311-333: F = I at u = 0, Σ V = 1 on unit_box(2,2,2)
    X, C = mesh.unit_box_rows(2, 2, 2)
    ts = _ts(X, C)
    el = ts.elements(np.zeros(3 * len(X)))
    assert np.abs(el["psi"]).max() < 1e-9 and np.abs(el["P"]).max() < 1e-9
    B, V = ts.rest_data()
    assert abs(V.sum() - 1.0) < 1e-14
    # B·Dm = I for every tet
    for e, t in enumerate(C):
        Dm = np.stack([X[t[1]] - X[t[0]], X[t[2]] - X[t[0]], X[t[3]] - X[t[0]]], axis=1)
        assert np.abs(B[e] @ Dm - np.eye(3)).max() < 1e-13


def test_affine_deformation_gives_constant_F():
    X, C = mesh.bar_freudenthal(2, 2, 2, 0.5)
    A = np.array([[1.1, 0.2, 0.0], [0.0, 0.9, 0.1], [0.05, 0.0, 1.2]])
    u = (X @ A.T - X).reshape(-1)
    lam, mu = mesh.lame(1e5, 0.3)
    ts = _ts(X, C, E=1e5, nu=0.3)
    el = ts.elements(u)
    _, Pa = oracle.snh_stress(A, lam, mu)
    assert np.abs(el["P"] - Pa).max() <= 1e-9 * np.abs(Pa).max()


def test_mass_matrix_closed_forms():
    # tests/fem_local_global/main.cpp:166-184: the consistent mass of the unit box sums to ρ (per dof block)
    X, C = mesh.unit_box_rows(3, 3, 3)
    ts = _ts(X, C, rho=7.0)
    M = ts.mass_values()
    assert abs(M.sum() - 3 * 7.0) < 1e-12
    assert abs(ts.total_mass() - 7.0) < 1e-12
    # per tet: M_e = ρ V (1 + δ_ij) / 20 on the diagonal of each 3x3 block, zeros off the block diagonal
    Xt, Ct = mesh.tetra()
    ts1 = _ts(Xt, Ct, rho=3.0)
    rowptr, colidx = ts1.csr_pattern()
    import scipy.sparse as sp

    Md = sp.csr_matrix((ts1.mass_values(), colidx, rowptr), shape=(12, 12)).toarray()
    V = 1.0 / 6
    ref = np.kron((np.ones((4, 4)) + np.eye(4)) * 3.0 * V / 20, np.eye(3))
    assert np.abs(Md - ref).max() < 1e-15


def _dense_H(ts, u, make_spsd, filter_dbc=True):
    rowptr, colidx = ts.csr_pattern()
    import scipy.sparse as sp

    return sp.csr_matrix((ts.hessian(u, make_spsd, filter_dbc), colidx, rowptr), shape=(ts.n, ts.n)).toarray()


def test_brute_force_dense_assembly():
    # global_composer / gather semantics (tests/fem_local_global/main.cpp:85-164): H = M + dt² Σ_e V_e scatter(K_e)
    X, C = mesh.bar_freudenthal(2, 1, 1, 0.1)
    rng = np.random.default_rng(3)
    ts = _ts(X, C, dt=0.05, E=1e4, nu=0.4, rho=100.0)
    u = 0.02 * rng.standard_normal(3 * len(X))
    el = ts.elements(u, make_spsd=True, want_K=True)
    B, V = ts.rest_data()
    n = ts.n
    Hd = np.zeros((n, n))
    for e, t in enumerate(C):
        idx = np.concatenate([3 * t[i] + np.arange(3) for i in range(4)])
        Me = np.kron((np.ones((4, 4)) + np.eye(4)) * 100.0 * V[e] / 20, np.eye(3))
        Hd[np.ix_(idx, idx)] += Me + 0.05 ** 2 * V[e] * el["K"][e]
    assert rel_err(_dense_H(ts, u, True, False), Hd) <= 1e-13


def _fd_check(ts, u, step, tol):
    E, g = ts.energy_grad(u)
    fd = np.zeros_like(g)
    for k in range(len(u)):
        up, um = u.copy(), u.copy()
        up[k] += step
        um[k] -= step
        fd[k] = (ts.energy_grad(up)[0] - ts.energy_grad(um)[0]) / (2 * step)
    assert rel_err(g, fd) <= tol
    H = _dense_H(ts, u, make_spsd=False)
    Hfd = np.zeros_like(H)
    for k in range(len(u)):
        up, um = u.copy(), u.copy()
        up[k] += step
        um[k] -= step
        Hfd[:, k] = (ts.energy_grad(up)[1] - ts.energy_grad(um)[1]) / (2 * step)
    assert rel_err(H, Hfd) <= tol


def test_incremental_potential_fd_tetra():
    # tests/fem_time_step_diff/main.cpp:19-92: one tet, dt = 1e-2, E = 1e6, ν = .33, ρ = 0, fixed u
    X, C = mesh.tetra()
    ts = _ts(X, C, dt=1e-2, E=1e6, nu=0.33, rho=0.0)
    u = np.array([[-0.1, -0.1, -0.1], [0.1, -0.1, -0.1], [-0.1, 0.1, -0.1], [0.1, 0.1, -0.1]]).reshape(-1)
    ts.set_vectors(u_hat=np.random.default_rng(4).uniform(-1, 1, 12))
    _fd_check(ts, u, 1e-6, 1e-6)
    # with inertia (ρ > 0)
    ts = _ts(X, C, dt=1e-2, E=1e6, nu=0.33, rho=10.0)
    ts.set_vectors(u_hat=np.random.default_rng(5).uniform(-1, 1, 12))
    _fd_check(ts, u, 1e-6, 1e-6)


def test_incremental_potential_fd_unit_box():
    # tests/fem_time_step_diff/main.cpp:95-172 (7x7x7 box, u = Random·0.3): here a 3x3x3 box to stay fast
    X, C = mesh.unit_box_rows(3, 3, 3)
    ts = _ts(X, C, dt=1e-2, E=1e6, nu=0.33, rho=1e3)
    rng = np.random.default_rng(6)
    ts.set_vectors(u_hat=0.01 * rng.standard_normal(3 * len(X)))
    _fd_check(ts, 0.03 * rng.standard_normal(3 * len(X)), 1e-6, 1e-6)


def test_dirichlet_filter_semantics():
    X, C = mesh.unit_box_rows(2, 2, 3)
    ts = _ts(X, C)
    dof = np.zeros(3 * len(X), np.uint8)
    dof[[0, 1, 2, 7]] = 1
    ts.set_dirichlet(dof, np.zeros(3 * len(X)))
    u = 0.01 * np.random.default_rng(7).standard_normal(3 * len(X))
    Hf = _dense_H(ts, u, True, True)
    Hu = _dense_H(ts, u, True, False)
    rowptr, colidx = ts.csr_pattern()
    for r in range(ts.n):
        for k in range(rowptr[r], rowptr[r + 1]):
            c = colidx[k]
            if dof[r] or dof[c]:
                assert Hf[r, c] == (1.0 if r == c else 0.0)
            else:
                assert Hf[r, c] == Hu[r, c]
    _, g = ts.energy_grad(u)
    assert np.all(g[dof == 1] == 0)


def test_pcg_matches_dense_solve():
    X, C = mesh.bar_freudenthal(3, 2, 2, 0.1)
    ts = _ts(X, C, dt=1 / 60, E=1e5, nu=0.45, rho=1000.0)
    u = 0.005 * np.random.default_rng(8).standard_normal(3 * len(X))
    rowptr, colidx = ts.csr_pattern()
    vals = ts.hessian(u, True, True)
    import scipy.sparse as sp

    H = sp.csr_matrix((vals, colidx, rowptr), shape=(ts.n, ts.n))
    b = np.random.default_rng(9).standard_normal(ts.n)
    x, it, rel = oracle.pcg(rowptr, colidx, vals, b, rtol=1e-12, maxit=10 * ts.n)
    xd = np.linalg.solve(H.toarray(), b)
    assert rel <= 1e-12 and 0 < it <= 10 * ts.n
    assert rel_err(x, xd) <= 1e-9
    # zero right-hand side: zero iterations, zero solution
    x0, it0, _ = oracle.pcg(rowptr, colidx, vals, np.zeros(ts.n))
    assert it0 == 0 and not x0.any()


def _tiny_case():
    case = mesh.cantilever_bar(3, 1, 1, h=0.05, E=1e4, nu=0.3, rho=100.0, dt=0.02, seed=2)
    return case


def test_newton_step_reaches_the_minimiser_of_the_incremental_potential():
    case = _tiny_case()
    ts = oracle.TimeStep.from_case(case)
    ts.set_solver(max_newton=50, pcg_rtol=1e-12)
    ts.set_threshold(1e-10)
    ts.prepare_step()
    u_next0, u_hat, _ = ts.get_vectors()
    free = case.dof_type == 0

    E0, g0 = ts.energy_grad(u_hat)
    scale = 1.0 / np.abs(g0).max()  # scipy's stopping rules assume O(1) gradients

    def fun(z):
        u = u_hat.copy()
        u[free] = z
        E, g = ts.energy_grad(u)
        return E * scale, g[free] * scale

    res = scipy.optimize.minimize(fun, u_hat[free], jac=True, method="L-BFGS-B",
                                  options=dict(maxiter=20000, ftol=0.0, gtol=1e-13, maxcor=30))
    gn, st = ts.compute_step()
    assert st["converged"] and gn < 1e-10 * ts.total_mass() * 10
    u_next, _, _ = ts.get_vectors()
    assert rel_err(u_next[free], res.x) <= 1e-6
    assert np.all(u_next[~free] == u_hat[~free])


def test_lbfgs_step_agrees_with_newton():
    case = _tiny_case()
    a = oracle.TimeStep.from_case(case)
    b = oracle.TimeStep.from_case(case)
    for ts in (a, b):
        ts.set_solver(max_newton=50, pcg_rtol=1e-12)
        ts.set_lbfgs(m=8, max_iter=5000)
        ts.set_threshold(1e-9)
    a.prepare_step()
    b.prepare_step()
    _, sa = a.compute_step("newton")
    _, sb = b.compute_step("lbfgs")
    assert sa["converged"] and sb["converged"]
    ua, _, _ = a.get_vectors()
    ub, _, _ = b.get_vectors()
    X = case.X.reshape(-1)
    assert rel_err(X + ub, X + ua) <= 1e-8


def test_ncg_step_agrees_with_newton():
    """The nonlinear-CG stepper converges to the minimiser the Newton stepper finds, to the accuracy its gradient
    threshold allows (unpreconditioned NCG stalls on Armijo roundoff below ~1e-10, so the threshold is 1e-8 × mass).
    The Dai-Kou iteration path itself is parity unpinned: mathprim's ncg_optimizer is not available."""
    case = _tiny_case()
    a = oracle.TimeStep.from_case(case)
    b = oracle.TimeStep.from_case(case)
    for ts in (a, b):
        ts.set_solver(max_newton=50, pcg_rtol=1e-12)
        ts.set_lbfgs(m=8, max_iter=20000)
        ts.set_threshold(1e-8)
    a.prepare_step()
    b.prepare_step()
    _, sa = a.compute_step("newton")
    _, sb = b.compute_step("ncg")
    assert sa["converged"] and sb["converged"], sb
    ua, _, _ = a.get_vectors()
    ub, _, _ = b.get_vectors()
    X = case.X.reshape(-1)
    assert rel_err(X + ub, X + ua) <= 1e-6


def test_step_next_and_prepare_step_formulas():
    case = _tiny_case()
    ts = oracle.TimeStep.from_case(case)
    u0, _, v0 = ts.get_state()
    v0 = np.random.default_rng(10).standard_normal(len(u0)) * 0.01
    ts.set_state(vel=v0)
    ts.prepare_step()
    _, u_hat, _ = ts.get_vectors()
    ref = u0 + case.dt * v0 + case.dt ** 2 * case.accel
    ref[case.dof_type == 1] = case.dbc[case.dof_type == 1]
    assert np.array_equal(u_hat, ref)
    ts.compute_step()
    un, _, _ = ts.get_vectors()
    ts.step_next()
    u, up, v = ts.get_state()
    assert np.array_equal(u, un) and np.array_equal(up, u0)
    assert rel_err(v, (un - u0) / case.dt) <= 1e-15


def test_planted_mistake_is_caught():
    """A dropped term in Ψ (the −μ·ln(Ic+δ) barrier) must fail the autograd fixtures: the checks have teeth."""
    with open(GOLDEN) as f:
        cases = json.load(f)["cases"]
    bad = 0
    for c in cases:
        F = np.array(c["F"])
        Ic = (F * F).sum()
        wrong = oracle.snh_energy(F, c["lam"], c["mu"]) + 0.5 * c["mu"] * np.log(Ic + 1)
        bad += abs(wrong - c["psi"]) > 1e-12 * max(abs(c["psi"]), c["mu"])
    assert bad >= len(cases) - 2


# ---- adjoint of the implicit step (extension; pinned by finite differences and closed forms) ----

def test_snh_stress_lame_derivatives_match_finite_differences(golden):
    for c in golden["cases"][:12]:
        F = np.array(c["F"]).reshape(3, 3)
        lam, mu = c["lam"], c["mu"]
        dl, dm = oracle.snh_dP_dlame(F)
        h = 1e-6
        fd_l = (oracle.snh_stress(F, lam * (1 + h), mu)[1] - oracle.snh_stress(F, lam * (1 - h), mu)[1]) / (2 * h * lam)
        fd_m = (oracle.snh_stress(F, lam, mu * (1 + h))[1] - oracle.snh_stress(F, lam, mu * (1 - h))[1]) / (2 * h * mu)
        scale = max(np.abs(F).max(), np.abs(oracle.cofactor(F[None])[0]).max())  # the derivatives' natural size
        assert np.abs(dl - fd_l).max() <= 1e-6 * scale
        assert np.abs(dm - fd_m).max() <= 1e-6 * scale
        # P is homogeneous of degree 1 in (λ, μ): λ ∂P/∂λ + μ ∂P/∂μ = P
        P = oracle.snh_stress(F, lam, mu)[1]
        assert np.abs(lam * dl + mu * dm - P).max() <= 1e-12 * (lam + mu) * scale
    Fr = np.eye(3)
    C = oracle.cofactor(Fr[None])[0]
    assert np.allclose(C, np.eye(3))
    R = np.array([[0.0, -1, 0], [1, 0, 0], [0, 0, 1]]) @ np.diag([1.2, 0.9, 1.1])
    assert np.allclose(oracle.cofactor(R[None])[0], np.linalg.det(R) * np.linalg.inv(R).T)


def _adjoint_case():
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

    return case, loss_grad


def test_adjoint_step_solves_the_linearised_step():
    case, _ = _adjoint_case()
    ts = oracle.TimeStep.from_case(case)
    ts.set_threshold(1e-8)
    ts.set_solver(pcg_rtol=1e-12)
    ts.set_polish(2)
    ts.step()
    u = ts.get_state()[0]
    rng = np.random.default_rng(1)
    gu = rng.standard_normal(ts.n)
    lv, guhat, gl, gm = ts.adjoint_step(u, gu)
    free = case.dof_type == 0
    H = ts.hessian_csr(u, make_spsd=False, filter_dbc=True)
    assert np.linalg.norm((H @ lv - np.where(free, gu, 0))) <= 1e-10 * np.linalg.norm(gu)
    assert not lv[~free].any() and not guhat[~free].any()
    # ∂L/∂û = M l for L = gu·u*: central differences of the step's solution in the predictor û
    d_uhat = np.where(free, rng.standard_normal(ts.n), 0.0)
    ts2 = oracle.TimeStep.from_case(case)
    ts2.set_threshold(1e-8)
    ts2.set_solver(pcg_rtol=1e-12)
    ts2.set_polish(2)
    ts2.prepare_step()
    _, uhat, _ = ts2.get_vectors()
    h = 1e-6
    sol = []
    for sgn in (1, -1):
        ts2.set_vectors(u_next=uhat + sgn * h * d_uhat, u_hat=uhat + sgn * h * d_uhat)
        ts2.compute_step()
        sol.append(ts2.get_vectors()[0])
    lv2, guhat2, _, _ = ts2.adjoint_step(u, gu)
    # identity gu·H⁻¹(M δû) = (M l)·δû (H symmetric), against the tangent solve (two sparse solves: the bar is
    # their roundoff times the condition number of H)
    du_lin = scipy.sparse.linalg.spsolve(H.tocsc(), np.where(free, ts2.mass_csr() @ d_uhat, 0.0))
    assert abs(gu @ du_lin - guhat2 @ d_uhat) <= 1e-10 * abs(gu @ du_lin)
    # and the tangent is the step's: central differences of the converged solution (a random functional
    # has cancellation, so the differences resolve it to a few 1e-6)
    fd_du = (sol[0] - sol[1]) / (2 * h)
    assert np.abs(fd_du - du_lin).max() <= 1e-6 * np.abs(du_lin).max()
    assert abs(gu @ fd_du - guhat2 @ d_uhat) <= 1e-5 * abs(gu @ du_lin)


def test_rollout_adjoint_matches_central_differences():
    case, loss_grad = _adjoint_case()
    steps = 4
    L, g = oracle.rollout_gradient(case, steps, loss_grad)
    g_log_e = case.lam * g["lam"] + case.mu * g["mu"]
    rng = np.random.default_rng(2)
    h = 1e-6
    for _ in range(3):
        d = rng.standard_normal(case.nt)
        Lp = oracle.rollout_loss(case, steps, loss_grad, lam=case.lam * np.exp(h * d), mu=case.mu * np.exp(h * d))
        Lm = oracle.rollout_loss(case, steps, loss_grad, lam=case.lam * np.exp(-h * d), mu=case.mu * np.exp(-h * d))
        fd = (Lp - Lm) / (2 * h)
        assert abs(fd - g_log_e @ d) <= 1e-5 * abs(fd)
        dv = rng.standard_normal(case.nv * 3)
        dv[case.dof_type == 1] = 0
        Lp = oracle.rollout_loss(case, steps, loss_grad, v0=case.v0 + h * dv)
        Lm = oracle.rollout_loss(case, steps, loss_grad, v0=case.v0 - h * dv)
        fd = (Lp - Lm) / (2 * h)
        assert abs(fd - g["v0"] @ dv) <= 1e-5 * abs(fd)


def test_planted_adjoint_mistake_fails_the_difference_check():
    """Dropping the velocity path of the recursion (ū_n = M l instead of M l − v̄/dt) must break agreement
    with central differences: the check has the power to see a missing term."""
    case, loss_grad = _adjoint_case()
    steps = 4
    ts = oracle.TimeStep.from_case(case)
    ts.set_threshold(1e-8)
    ts.set_solver(pcg_rtol=1e-12)
    ts.set_polish(2)
    us = [case.u0.copy()]
    for _ in range(steps):
        ts.step()
        us.append(ts.get_state()[0])
    _, gu = loss_grad(us[-1])
    ubar, vbar = gu.copy(), np.zeros(ts.n)
    for k in range(steps - 1, -1, -1):
        _, guhat, _, _ = ts.adjoint_step(us[k + 1], ubar + vbar / ts.dt)
        ubar, vbar = guhat, ts.dt * guhat  # planted: − v̄/dt dropped
    rng = np.random.default_rng(3)
    dv = rng.standard_normal(case.nv * 3)
    dv[case.dof_type == 1] = 0
    h = 1e-6
    fd = (oracle.rollout_loss(case, steps, loss_grad, v0=case.v0 + h * dv)
          - oracle.rollout_loss(case, steps, loss_grad, v0=case.v0 - h * dv)) / (2 * h)
    assert abs(fd - vbar @ dv) > 1e-3 * abs(fd)
