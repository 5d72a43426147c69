"""Batched material evaluation on MI355X (``ssim_material_eval``, the vmap over a model's stress and Hessian
operators) against the oracle, in 3D and 2D, for both models.

Bars: Ψ, P and H within 1e-12 relative per item (norm-wise over the item); the SPD-projected Hessians within
1e-10 (3D: the device projects in closed form in the SVD frame, the oracle by a 9x9 Jacobi eigen-decomposition;
2D: against numpy.linalg.eigh).
"""
import numpy as np
import pytest

import oracle
from helpers import per_item_rel_err
from ssim_amd import material, mesh

pytestmark = pytest.mark.gpu


def _inputs(d, n, seed):
    rng = np.random.default_rng(seed)
    F = np.eye(d) + 0.3 * rng.standard_normal((n, d, d))
    E = np.exp(rng.uniform(np.log(1e4), np.log(1e6), n))
    nu = rng.uniform(0.2, 0.48, n)
    lam, mu = mesh.lame(E, nu)
    return F, lam, mu


@pytest.mark.parametrize("model", ["stable_neohookean", "linear"])
def test_2d_models_match_oracle(model):
    F, lam, mu = _inputs(2, 500, 1)
    got = material.evaluate(F, lam, mu, model)
    ref = [oracle.material2(model, F[i], lam[i], mu[i]) for i in range(len(F))]
    psi = np.array([r[0] for r in ref])
    assert np.abs(got["psi"] - psi).max() <= 1e-12 * np.abs(psi).max()
    assert per_item_rel_err(got["P"], np.array([r[1] for r in ref])) <= 1e-12
    assert per_item_rel_err(got["H"], np.array([r[2] for r in ref])) <= 1e-12


@pytest.mark.parametrize("model", ["stable_neohookean", "linear"])
def test_3d_models_match_oracle(model):
    F, lam, mu = _inputs(3, 500, 2)
    for spsd in (False, True):
        got = material.evaluate(F, lam, mu, model, make_spsd=spsd)
        if model == "linear":
            ref = [(oracle.lin_stress(F[i], lam[i], mu[i]), oracle.lin_hessian(lam[i], mu[i])) for i in range(len(F))]
        else:
            ref = [(oracle.snh_stress(F[i], lam[i], mu[i]), oracle.snh_hessian(F[i], lam[i], mu[i]))
                   for i in range(len(F))]
        psi = np.array([r[0][0] for r in ref])
        assert np.abs(got["psi"] - psi).max() <= 1e-12 * np.abs(psi).max()
        assert per_item_rel_err(got["P"], np.array([r[0][1] for r in ref])) <= 1e-12
        H = np.array([oracle.spd_project9(r[1]) if spsd else r[1] for r in ref])
        assert per_item_rel_err(got["H"], H) <= (1e-10 if spsd else 1e-12), spsd
        # without the Hessian output the stress is the same, bit for bit
        lean = material.evaluate(F, lam, mu, model, make_spsd=spsd, want_H=False)
        assert "H" not in lean and np.array_equal(lean["P"], got["P"]) and np.array_equal(lean["psi"], got["psi"])


@pytest.mark.parametrize("model", ["stable_neohookean", "linear"])
def test_2d_projection_matches_numpy_eigh(model):
    """The 2D clamp (closed form for stable Neo-Hookean, Jacobi for linear) against numpy.linalg.eigh of the
    oracle's Hessian with eigenvalues raised to 1e-3, including F = I, a rotation, F = 0 and a reflection."""
    F, lam, mu = _inputs(2, 500, 4)
    F[:4] = [np.eye(2), [[np.cos(0.4), -np.sin(0.4)], [np.sin(0.4), np.cos(0.4)]], np.zeros((2, 2)), np.diag([1, -1])]
    got = material.evaluate(F, lam, mu, model, make_spsd=True)
    ref = []
    for i in range(len(F)):
        w, Q = np.linalg.eigh(oracle.material2(model, F[i], lam[i], mu[i])[2])
        ref.append(Q @ np.diag(np.maximum(w, 1e-3)) @ Q.T)
    assert per_item_rel_err(got["H"], np.array(ref)) <= 1e-10
    plain = material.evaluate(F, lam, mu, model)
    assert np.array_equal(got["P"], plain["P"]) and np.array_equal(got["psi"], plain["psi"])


def test_invalid_requests_are_rejected():
    F, lam, mu = _inputs(2, 4, 3)
    with pytest.raises(ValueError):
        material.evaluate(F, lam, mu, model="corotated")
    with pytest.raises(ValueError):
        material.evaluate(np.zeros((4, 4, 4)), 1.0, 1.0)
