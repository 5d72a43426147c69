"""ssim::fem::basic_time_step<double, device::hip, 3, 4, ...> (include/ssim, include/mathprim) driven by a HIP
program with the reference device driver's call sequence (tests/cpp/time_step_driver.hip; the sequence of
tests/fem_cu/main.cu:22-102), against the Python front end on the same stretched unit box: per-step energies,
residuals and iteration counts, the mass sum and threshold, and the final positions, for each stepper type the
header maps (Newton with Jacobi-PCG and with plain CG, L-BFGS, nonlinear CG, forward Euler)."""
import os
import subprocess

import numpy as np
import pytest

from ssim_amd import fem

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(os.path.dirname(__file__), "cpp", "build", "time_step_driver")
METHODS = {"newton": "newton", "newton_plain": "newton", "lbfgs": "lbfgs", "ncg": "ncg",
           "forward_euler": "forward_euler"}


def python_run(nx, steps, method):
    verts, cells = fem.unit_box(4 * nx, nx, nx)
    X = np.ascontiguousarray(verts.T)
    X[:, 0] *= 4.0
    s = fem.TetFiniteElementSolver_Hip(X, cells.T, 1e-2, 1e6, 0.4, 10, method=METHODS[method])
    s.mark_dirichlet_batched(np.arange(nx * nx), np.zeros((3, nx * nx)))
    s.add_gravity(np.array([0.0, -9.8, 0.0]))
    s.set_rtol(1e-3)
    if method == "newton_plain":
        s.set_solver_params(pcg_precond=1)
    s.reset()
    mass = s.mass_matrix().sum() / 3
    thr = s.grad_threshold_abs()
    s.update_hessian(make_spsd=False)
    rows = []
    for i in range(steps):
        s.prepare_step()
        e = s.update_energy_and_gradients()
        r = s.compute_step()
        s.step_next()
        st = s.last_stats()
        rows.append((i, e, r, st["newton_iterations"], st["converged"]))
    return mass, thr, rows, X + s.get_array("deformation").reshape(-1, 3)


@pytest.mark.parametrize("method", list(METHODS))
def test_basic_time_step_driver_matches_python_front_end(tmp_path, method):
    if not os.path.exists(DRIVER):
        pytest.fail("tests/cpp/build/time_step_driver is not built (python tests/cpp/build.py)")
    nx, steps = 4, 20
    out, npy = tmp_path / "out.txt", tmp_path / "final.npy"
    print(f"[gpu] basic_time_step driver, {method}", flush=True)
    res = subprocess.run([DRIVER, str(nx), str(steps), method, str(out), str(npy)], capture_output=True, text=True,
                         timeout=120)
    assert res.returncode == 0, res.stderr
    lines = out.read_text().split("\n")
    mass_c = float(lines[0].split()[1])
    thr_c = float(lines[1].split()[1])
    rows_c = [l.split() for l in lines[2:] if l.strip()]
    mass, thr, rows, x = python_run(nx, steps, method)
    assert mass_c == pytest.approx(mass, rel=1e-13) and thr_c == pytest.approx(thr, rel=1e-13)
    assert len(rows_c) == steps
    for rc, rp in zip(rows_c, rows):
        step = int(rc[0])
        assert float(rc[1]) == pytest.approx(rp[1], rel=1e-12, abs=1e-300), (step, "energy")
        assert float(rc[2]) == pytest.approx(rp[2], rel=1e-9, abs=1e-300), (step, "residual")
        assert int(rc[3]) == rp[3] and int(rc[4]) == rp[4], step
    xc = np.load(npy)
    assert xc.shape == x.shape
    assert np.abs(xc - x).max() <= 1e-12 * np.abs(x).max()
