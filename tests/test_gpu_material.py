"""Batched material evaluation on MI355X (``ssim_material_eval``, the vmap over a model's stress and Hessian
operators) against the oracle, in 3D and 2D, for both models.

Bars: Ψ, P and H within 1e-12 relative per item (norm-wise over the item); the SPD-projected 3D Hessian within
1e-10 (the device projects in closed form in the SVD frame, the oracle by a 9x9 Jacobi eigen-decomposition).
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


def test_invalid_requests_are_rejected():
    F, lam, mu = _inputs(2, 4, 3)
    with pytest.raises(ValueError):
        material.evaluate(F, lam, mu, make_spsd=True)
    with pytest.raises(ValueError):
        material.evaluate(F, lam, mu, model="corotated")
    with pytest.raises(ValueError):
        material.evaluate(np.zeros((4, 4, 4)), 1.0, 1.0)
