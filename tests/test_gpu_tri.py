"""P1 triangles in 2D on MI355X (ssim_tri_eval) against the oracle's orc_tri_eval: every per-element output for
both material models, with and without the SPD projection, on a jittered 40×30 grid (2,400 triangles) with
random displacements large enough to invert some triangles.  Bars: area, Dm⁻¹, F, mass and Laplacian within
1e-13; energy, gradient and stiffness within 1e-10 (relative, over the whole array).  Also: an out-of-range vertex
index gives NaN outputs for that triangle only, a degenerate triangle the oracle's pattern of non-finite outputs,
NULL outputs are skipped, and zero triangles are accepted."""
import numpy as np
import pytest

import oracle
from helpers import rel_err
from ssim_amd import tri
from test_oracle_tri import _mesh

pytestmark = pytest.mark.gpu

TIGHT = ("area", "dminv", "F", "mass", "laplacian")


@pytest.mark.parametrize("model", ["stable_neohookean", "linear"])
@pytest.mark.parametrize("make_spsd", [False, True])
def test_triangles_match_oracle(model, make_spsd):
    X, C, u = _mesh(nx=40, ny=30, seed=11, amp=0.6)
    rng = np.random.default_rng(2)
    lam = rng.uniform(1e3, 1e5, len(C))
    mu = rng.uniform(1e2, 1e4, len(C))
    print(f"[gpu] triangles {model} spsd={make_spsd}: {len(C)} tris", flush=True)
    g = tri.evaluate(X, C, u, lam, mu, model=model, make_spsd=make_spsd, density=1200.0)
    o = oracle.tri_eval(X, C, u, lam, mu, model=model, make_spsd=make_spsd, density=1200.0)
    assert (o["area"] > 0).all()
    F = o["F"]
    assert (F[:, 0, 0] * F[:, 1, 1] - F[:, 0, 1] * F[:, 1, 0] < 0).any()  # some triangles are inverted
    for k in tri.FIELDS:
        assert g[k].shape == o[k].shape, k
        bar = 1e-13 if k in TIGHT else 1e-10
        assert rel_err(g[k], o[k]) <= bar, (k, rel_err(g[k], o[k]))


def test_bad_index_null_outputs_and_empty():
    X, C, u = _mesh(nx=6, ny=5, seed=4)
    C = C.copy()
    C[7, 2] = len(X)
    C[9, 0] = -3
    C[11] = [0, 6, 12]  # three collinear boundary vertices on y = 0: det Dm = 0
    g = tri.evaluate(X, C, u, 1e4, 1e3, make_spsd=True)
    o = oracle.tri_eval(X, C, u, 1e4, 1e3, make_spsd=True)
    ok = np.ones(len(C), dtype=bool)
    ok[[7, 9, 11]] = False
    for k in tri.FIELDS:
        assert np.isnan(g[k][[7, 9]]).all(), k
        assert np.array_equal(np.isfinite(g[k][11]), np.isfinite(o[k][11])), k
        assert rel_err(g[k][ok], o[k][ok]) <= 1e-10, k
    assert g["area"][11] == 0.0 and not g["mass"][11].any() and not np.isfinite(g["stiff"][11]).any()
    part = tri.evaluate(X, C, u, 1e4, 1e3, want=("energy", "stiff"))
    assert set(part) == {"energy", "stiff"}
    assert rel_err(part["stiff"][ok], oracle.tri_eval(X, C, u, 1e4, 1e3)["stiff"][ok]) <= 1e-10
    empty = tri.evaluate(X, np.zeros((0, 3), dtype=np.int32), u, 1e4, 1e3)
    assert all(v.shape[0] == 0 for v in empty.values())


@pytest.mark.parametrize("model", ["stable_neohookean", "linear"])
def test_projection_at_special_deformations(model):
    """The clamp at F = I, a rotation, F = 0, a reflection and a uniform scaling (where the stable Neo-Hookean
    closed form meets vanishing or repeated components) against the oracle's Jacobi route."""
    X = np.array([[0.0, 0.0], [1.0, 0.0], [0.0, 1.0]])
    th = 0.3
    Fs = [np.eye(2), np.array([[np.cos(th), -np.sin(th)], [np.sin(th), np.cos(th)]]), np.zeros((2, 2)),
          np.diag([1.0, -1.0]), 2.0 * np.eye(2), np.array([[0.0, 1.0], [1.0, 0.0]])]
    for F in Fs:
        u = X @ F.T - X  # Dm = I, so F = Ds
        g = tri.evaluate(X, [[0, 1, 2]], u, 2e4, 3e3, model=model, make_spsd=True, want=("F", "stiff"))
        o = oracle.tri_eval(X, [[0, 1, 2]], u, 2e4, 3e3, model=model, make_spsd=True)
        assert np.abs(g["F"][0] - F).max() <= 1e-15
        assert rel_err(g["stiff"], o["stiff"]) <= 1e-12, (model, F)
        assert np.linalg.eigvalsh(g["stiff"][0]).min() >= -1e-9 * np.abs(g["stiff"]).max()
