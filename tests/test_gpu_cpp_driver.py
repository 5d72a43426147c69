"""The header-only C++ front end (include/ssim_hip.hpp) driven by a C++ program, against the Python front end
on the same bar: identical positions (bitwise) and iteration counts after 5 steps of the 8×8×32 bar, for the
Newton, L-BFGS, nonlinear-CG and forward-Euler steppers; the remaining views of the header (mesh, dof_type, dminv,
rest_vol, local_stiff, sysmatrix, mass_matrix, mass_matrix_filtered) agree with the Python front end's; the
per-step deform_<step>.npy frames the driver writes hold X + u and equal the Python front end's save_frame."""
import os
import subprocess

import numpy as np
import pytest

from ssim_amd import fem, mesh

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(os.path.dirname(__file__), "cpp", "build", "driver")


@pytest.mark.parametrize("stepper", [0, 1, 2, 3])
def test_cpp_driver_matches_python_front_end(tmp_path, stepper):
    if not os.path.exists(DRIVER):
        pytest.fail("tests/cpp/build/driver is not built (python tests/cpp/build.py)")
    case = mesh.cantilever_bar(32, 8, 8, seed=0)
    steps, dt, E, nu, rho, rtol, g = 5, case.dt, 1e5, 0.45, case.rho, 1e-4, -9.8
    clamped = case.dof_type.reshape(-1, 3).any(axis=1)
    inp, out = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(inp, "wb") as f:
        np.array([case.nv, case.nt, steps, stepper], dtype=np.int32).tofile(f)
        np.array([dt, E, nu, rho, rtol, g], dtype=np.float64).tofile(f)
        case.X.astype(np.float64).tofile(f)
        case.cells.astype(np.int32).tofile(f)
        clamped.astype(np.uint8).tofile(f)
        case.u0.astype(np.float64).tofile(f)
    print(f"[gpu] C++ driver, stepper {stepper}", flush=True)
    frames = tmp_path / "frames"
    frames.mkdir()
    args = [DRIVER, str(inp), str(out)] + ([str(frames)] if stepper == 0 else [])
    res = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    raw = np.fromfile(out, dtype=np.uint8)
    u_cpp = raw[:8 * 3 * case.nv].view(np.float64)
    off = 8 * 3 * case.nv
    counts = raw[off:off + 8 * steps].view(np.int32).reshape(steps, 2)
    views = raw[off + 8 * steps:].view(np.float64)

    method = ["newton", "lbfgs", "ncg", "forward_euler"][stepper]
    s = fem.TetFiniteElementSolver_Hip(case.X, case.cells, dt, E, nu, rho, method)
    idx = np.nonzero(clamped)[0].astype(np.int32)
    s.mark_dirichlet_batched(idx, np.zeros((3, idx.size)))
    s.add_gravity(np.array([0.0, g, 0.0]))
    s.set_array("deformation", case.u0)
    s.set_rtol(rtol)
    s.reset()
    py_counts = []
    for _ in range(steps):
        s.step()
        st = s.last_stats()
        py_counts.append((st["newton_iterations"], st["pcg_iterations"]))
    assert np.array_equal(counts, np.array(py_counts, dtype=np.int32))
    assert np.array_equal(u_cpp, s.get_array("deformation"))
    if stepper == 0:
        # the deform_<step>.npy frames: (nv, 3) positions X + u, the last one equal to the Python front end's
        names = sorted(os.listdir(frames))
        assert names == sorted(f"deform_{i}.npy" for i in range(steps))
        last = np.load(frames / f"deform_{steps - 1}.npy")
        assert last.shape == (case.nv, 3) and last.dtype == np.float64
        assert np.array_equal(last, s.save_frame(tmp_path / "py_frame.npy"))
        assert np.array_equal(last, np.load(tmp_path / "py_frame.npy"))
        assert np.array_equal(last.reshape(-1), case.X.reshape(-1) + u_cpp)
    dminv, vol = s.rest_shape()
    H = s.hessian()
    K = s.eval_elements(s.get_array("next_deform"), make_spsd=True, want_H=False, want_K=True)["K"]
    dof = np.zeros(3 * case.nv, dtype=np.uint8)
    dof[np.repeat(clamped, 3)] = 1
    a = lambda v: np.abs(v).sum()
    expect = [a(vol), a(dminv), H.nnz, a(H.data), a(s.mass_matrix().data), a(s.mass_matrix_filtered().data), a(dof),
              a(K)]
    assert views.shape == (8,)
    np.testing.assert_allclose(views, expect, rtol=1e-10, atol=0)  # sequential vs pairwise sums of ~3e5 terms
