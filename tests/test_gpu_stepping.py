"""The GPU time stepping against the closed forms that pin the oracle's (tests/test_oracle_stepping.py): free
fall with and without velocity damping, and the damped linear bar settling at K_ff u_f = M_ff g_f; plus the same
trajectories against the oracle step by step."""
import numpy as np
import pytest

import oracle
from helpers import rel_err
from ssim_amd import fem
from test_oracle_stepping import equilibrium, free_bar, free_fall_closed_form, stiff_bar

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("damping", [0.0, 0.3])
def test_free_fall_matches_closed_form(damping):
    case = free_bar()
    s = fem.TetFiniteElementSolver_Hip.from_case(case)
    s.set_damping(damping)
    o = oracle.TimeStep.from_case(case)
    o.set_damping(damping)
    g = case.accel.reshape(-1, 3)[0]
    for n in range(1, 21):
        s.step()
        o.step()
        st = s.last_stats()
        assert st["newton_iterations"] == 0 and st["converged"], (n, st)
        u = s.get_array("deformation").reshape(-1, 3)
        v = s.get_array("velocity").reshape(-1, 3)
        uc, vc = free_fall_closed_form(n, case.dt, g, damping)
        np.testing.assert_allclose(u, np.broadcast_to(uc, u.shape), rtol=1e-12, atol=1e-18)
        np.testing.assert_allclose(v, np.broadcast_to(vc, v.shape), rtol=1e-11, atol=1e-16)
        uo, _, vo = o.get_state()
        assert rel_err(u.reshape(-1), uo) <= 1e-14 and rel_err(v.reshape(-1), vo) <= 1e-13, n


def test_damped_linear_bar_settles_at_static_equilibrium():
    case = stiff_bar()
    ustar = equilibrium(case)
    s = fem.TetFiniteElementSolver_Hip.from_case(case, rtol=1e-11, material="linear")
    s.set_damping(0.9)
    s.set_solver_params(pcg_rtol=1e-12)
    o = oracle.TimeStep.from_case(case)
    o.set_model("linear")
    o.set_damping(0.9)
    o.set_threshold(1e-11)
    o.set_solver(pcg_rtol=1e-12)
    # the steps are solved to 1e-11·mass, where the line search compares energy differences at roundoff, so the
    # two solvers may stop a step at different iterates; both settle at the same equilibrium
    for step in range(30):
        s.step()
        o.step()
    u = s.get_array("deformation")
    err = np.abs(u - ustar).max() / np.abs(ustar).max()
    print(f"[gpu] damped linear bar after 30 steps: |u - u*| / |u*| = {err:.2e}", flush=True)
    assert err <= 1e-9
    assert rel_err(u, o.get_state()[0]) <= 1e-9
