"""Oracle checks of the externally preconditioned NCG stepper (time_step_solver_ncg_with_ext_prec,
stepper/nonlinear_cg.hpp:41-97; bound as tet_ncg_cuda_ext_ai, bindings/src/fem_cuda.cu:172-244).

Pins: with L = I the preconditioned iteration is the plain one bitwise; the first step is along −L·Lᵀ·g computed
by numpy for a non-symmetric L (the block-Jacobi factor C⁻ᵀ of the diagonal blocks, so applying Lᵀ·L instead of
L·Lᵀ fails); the iteration converges to the Newton minimiser, in fewer iterations than the plain NCG.
mathprim's preconditioned ncg_optimizer is not available, so the exact iteration path is parity unpinned.
"""
import numpy as np
import scipy.sparse as sp

import oracle
from helpers import rel_err
from ssim_amd import mesh


def _case():
    return mesh.cantilever_bar(3, 1, 1, h=0.05, E=1e4, nu=0.3, rho=100.0, dt=0.02, seed=2)


def block_jacobi_factor(H):
    """CSR values (H's pattern and order) of L with diagonal blocks C⁻ᵀ, C the Cholesky factor of H's diagonal
    block, and zero elsewhere: L·Lᵀ = blockdiag(H)⁻¹."""
    H = sp.csr_matrix(H)
    n = H.shape[0]
    vals = np.zeros(H.nnz)
    for v in range(n // 3):
        Hb = H[3 * v:3 * v + 3, 3 * v:3 * v + 3].toarray()
        C = np.linalg.cholesky(Hb)
        Lb = np.linalg.inv(C).T
        for b in range(3):
            r = 3 * v + b
            cols = H.indices[H.indptr[r]:H.indptr[r + 1]]
            for d in range(3):
                k = np.nonzero(cols == 3 * v + d)[0][0]
                vals[H.indptr[r] + k] = Lb[b, d]
    return vals


def identity_values(H):
    H = sp.csr_matrix(H)
    vals = np.zeros(H.nnz)
    for r in range(H.shape[0]):
        cols = H.indices[H.indptr[r]:H.indptr[r + 1]]
        vals[H.indptr[r] + np.nonzero(cols == r)[0][0]] = 1.0
    return vals


def _ts(case, rtol=1e-8, max_iter=20000):
    ts = oracle.TimeStep.from_case(case)
    ts.set_solver(max_newton=50, pcg_rtol=1e-12)
    ts.set_lbfgs(m=8, max_iter=max_iter)
    ts.set_threshold(rtol)
    ts.prepare_step()
    return ts


def test_identity_preconditioner_is_the_plain_iteration():
    case = _case()
    a, b = _ts(case), _ts(case)
    _, uh, _ = a.get_vectors()
    b.set_precond(identity_values(b.hessian_csr(uh, True, True)))
    _, sa = a.compute_step("ncg")
    _, sb = b.compute_step("ncg")
    assert sa == sb
    assert np.array_equal(a.get_vectors()[0], b.get_vectors()[0])


def test_first_step_is_along_the_preconditioned_gradient():
    case = _case()
    ts = _ts(case, max_iter=1)
    _, uh, _ = ts.get_vectors()
    H = ts.hessian_csr(uh, True, True)
    vals = block_jacobi_factor(H)
    L = sp.csr_matrix((vals, H.indices, H.indptr), shape=H.shape)
    assert abs(L - L.T).max() > 1e-3  # non-symmetric: Lᵀ·L ≠ L·Lᵀ
    ts.set_precond(vals)
    _, g0 = ts.energy_grad(uh)
    ts.compute_step("ncg")
    step = ts.get_vectors()[0] - uh
    d0 = -(L @ (L.T @ g0))
    wrong = -(L.T @ (L @ g0))
    alpha = (step @ d0) / (d0 @ d0)
    assert rel_err(step, alpha * d0) <= 1e-12
    assert rel_err(wrong, d0) > 1e-3
    # backtracking from the first trial 1e-2 by halving
    k = np.log2(1e-2 / alpha)
    assert abs(k - round(k)) < 1e-9 and round(k) >= 0


def test_preconditioned_ncg_reaches_the_newton_minimiser_faster():
    case = _case()
    newton, plain, prec = _ts(case), _ts(case), _ts(case)
    _, uh, _ = prec.get_vectors()
    prec.set_precond(block_jacobi_factor(prec.hessian_csr(uh, True, True)))
    _, sn = newton.compute_step("newton")
    _, s0 = plain.compute_step("ncg")
    _, s1 = prec.compute_step("ncg")
    assert sn["converged"] and s0["converged"] and s1["converged"]
    X = case.X.reshape(-1)
    xn = X + newton.get_vectors()[0]
    assert rel_err(X + prec.get_vectors()[0], xn) <= 1e-6
    assert s1["newton"] < s0["newton"], (s1["newton"], s0["newton"])
