"""Mesh generators and seeded inputs (CPU)."""
import numpy as np

from ssim_amd import mesh


def _volumes(X, cells):
    x0, x1, x2, x3 = (X[cells[:, i]] for i in range(4))
    Dm = np.stack([x1 - x0, x2 - x0, x3 - x0], axis=2)
    return np.linalg.det(Dm) / 6


def test_freudenthal_counts_match_benchmark_meshes():
    X, C = mesh.bar_freudenthal(32, 8, 8)
    assert X.shape == (2673, 3) and C.shape == (12288, 4)
    X, C = mesh.bar_freudenthal(32, 32, 128, 0.01)
    assert X.shape == (33 * 33 * 129, 3) == (140481, 3)
    assert C.shape == (786432, 4)


def test_freudenthal_orientation_volume_and_conformity():
    h = 0.01
    X, C = mesh.bar_freudenthal(3, 2, 4, h)
    vol = _volumes(X, C)
    assert np.all(vol > 0)
    np.testing.assert_allclose(vol, h ** 3 / 6, rtol=1e-12)
    # x-slowest ids
    assert np.allclose(X[1] - X[0], [0, 0, h]) and np.allclose(X[5] - X[0], [0, h, 0])
    # conforming: every interior face is shared by exactly two tets, boundary faces by one
    faces = {}
    for t in C:
        for f in ((0, 1, 2), (0, 1, 3), (0, 2, 3), (1, 2, 3)):
            key = tuple(sorted(t[list(f)]))
            faces[key] = faces.get(key, 0) + 1
    counts = np.array(list(faces.values()))
    assert counts.max() == 2
    nb = (counts == 1).sum()
    # boundary area / face area: 2*(3*2 + 3*4 + 2*4) squares * 2 triangles each
    assert nb == 2 * 2 * (3 * 2 + 3 * 4 + 2 * 4)


def test_unit_box_matches_reference_counts_and_volume():
    verts, cells = mesh.unit_box(3, 4, 5)
    assert verts.shape == (3, 60) and cells.shape == (4, 5 * 2 * 3 * 4)
    X, C = mesh.unit_box_rows(3, 4, 5)
    np.testing.assert_allclose(np.abs(_volumes(X, C)).sum(), 1.0, rtol=1e-12)


def test_rng_is_reproducible_and_uniform():
    a = mesh.uniform(7, 1, 1000)
    b = mesh.uniform(7, 1, 1000)
    c = mesh.uniform(7, 2, 1000)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    assert a.min() >= 0 and a.max() < 1 and abs(a.mean() - 0.5) < 0.05
    # splitmix64 reference value: first output of key 0 (Vigna's splitmix64 with state advanced once)
    assert mesh.uniform(0, 0, 1)[0] == (0xE220A8397B1DCDAF >> 11) * 2.0 ** -53


def test_cantilever_case_boundary_and_jitter():
    case = mesh.cantilever_bar(4, 2, 2, seed=3)
    clamped = case.X[:, 0] == 0
    dof = case.dof_type.reshape(-1, 3)
    assert np.array_equal(dof.any(axis=1), clamped) and dof[clamped].all()
    u = case.u0.reshape(-1, 3)
    assert np.all(u[clamped] == 0)
    assert np.abs(u).max() <= 0.05 * 0.01
    assert np.all(case.accel.reshape(-1, 3)[:, 1] == -9.8)


def test_twisted_cube_material_ranges():
    case = mesh.twisted_cube(n=4, seed=1)
    E = case.mu * (3 * case.lam + 2 * case.mu) / (case.lam + case.mu)
    nu = case.lam / (2 * (case.lam + case.mu))
    assert E.min() >= 1e4 * (1 - 1e-12) and E.max() <= 1e6 * (1 + 1e-12)
    assert nu.min() >= 0.3 - 1e-12 and nu.max() <= 0.48 + 1e-12
    assert np.all(case.u0.reshape(-1, 3)[case.X[:, 2] == 0] == 0)


def test_bar_batch_is_seeded_and_within_ranges():
    a = mesh.bar_batch(16, seed=4)
    b = mesh.bar_batch(16, seed=4)
    c = mesh.bar_batch(16, seed=5)
    assert [x.name for x in a] == [x.name for x in b]
    assert [x.name for x in a] != [x.name for x in c]
    for x, y in zip(a, b):
        assert np.array_equal(x.u0, y.u0) and np.array_equal(x.lam, y.lam)
    for x in a:
        nx, ny, nz = x.meta["nx"], x.meta["ny"], x.meta["nz"]
        assert 4 <= min(nx, ny, nz) and max(nx, ny, nz) <= 24
        assert x.nv == (nx + 1) * (ny + 1) * (nz + 1) and x.nt == 6 * nx * ny * nz
        assert 1e4 <= x.meta["E"] <= 1e6 and 0.3 <= x.meta["nu"] <= 0.48
        assert np.unique(x.lam).size == 1 and np.unique(x.mu).size == 1  # one material per bar
        clamped = x.dof_type.reshape(-1, 3).any(axis=1)
        assert np.array_equal(clamped, x.X[:, 0] == 0.0)
        assert not x.u0[x.dof_type == 1].any()
        assert np.all(np.abs(x.u0) <= 0.05 * 0.01 + 1e-18)
    # the full 256-bar batch: dims cover the range, several million tets
    full = mesh.bar_batch(256, seed=0)
    dims = np.array([[x.meta["nx"], x.meta["ny"], x.meta["nz"]] for x in full])
    assert dims.min() == 4 and dims.max() == 24
    assert 3e6 <= sum(x.nt for x in full) <= 6e6
