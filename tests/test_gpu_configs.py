"""Benchmark configurations as parity cases on MI355X.

* FP32 element variant at the 786k-tet bar: max position error after 60 steps against the FP64 run of the same
  trajectory (reported; bounded at 1e-5 relative to the positions).
* Batched stepping of the 256-bar batch: every mesh against the single-mesh path after one step (≤1e-8).
* The differentiable rollout of configuration 5 (4 reduced bars, and the full 32 bars of 16×16×64 cubes with 64
  directions): gradients against central differences (≤1e-5) and the oracle (≤1e-6).
* L-BFGS on a twisted cube with per-tet E (log-uniform in [1e4, 1e6]) and ν (uniform in [0.3, 0.48]), z = 0 face
  clamped: positions after 3 steps within 1e-8 of the oracle.  L-BFGS stops anywhere inside ‖g‖ < tol and its
  iterates depend on rounding (history, line-search decisions), so GPU and oracle are compared on tightly
  converged steps (threshold 1e-10 × mass), where both sit at the same minimiser.
"""
import numpy as np
import pytest

import oracle
from helpers import rel_err
from ssim_amd import fem, mesh

pytestmark = pytest.mark.gpu


def _positions(case, s):
    return case.X.reshape(-1) + s.deformation().T.reshape(-1)


def test_fp32_element_variant_position_error_after_60_steps_786k():
    case = mesh.cantilever_bar(128, 32, 32, seed=0)
    runs = {}
    for fp32 in (0, 1):
        s = fem.TetFiniteElementSolver_Hip.from_case(case, rtol=1e-6)
        s.set_solver_params(fp32_elements=fp32)
        for _ in range(60):
            s.step()
        runs[fp32] = _positions(case, s)
    err = rel_err(runs[1], runs[0])
    abs_err = np.abs(runs[1] - runs[0]).max()
    print(f"[gpu] FP32 element variant, 786k tets, 60 steps: max position error {abs_err:.3e} m "
          f"({err:.3e} relative to max |x|)", flush=True)
    assert err <= 1e-5


def test_lbfgs_twisted_cube_per_tet_material_matches_oracle():
    case = mesh.twisted_cube(n=6, seed=7)
    g = fem.TetFiniteElementSolver_Hip.from_case(case, method="lbfgs", rtol=1e-10)
    g.set_solver_params(lbfgs_maxit=5000)
    o = oracle.TimeStep.from_case(case)
    o.set_threshold(1e-10)
    o.set_lbfgs(max_iter=5000)
    for step in range(3):
        g.step()
        o.step("lbfgs")
        xo = case.X.reshape(-1) + o.get_state()[0]
        assert rel_err(_positions(case, g), xo) <= 1e-8, step
    print(f"[gpu] L-BFGS twisted cube n=6: {g.last_stats()}", flush=True)


def test_config4_full_batch_agrees_with_single_mesh_path():
    """The 256-bar batch (dims 4..24): one batched step against each mesh stepped alone, positions within 1e-8 and
    identical Newton counts."""
    from ssim_amd import batch

    cases = mesh.bar_batch(256, seed=0)
    B = batch.TetFiniteElementBatch_Hip.from_cases(cases, rtol=1e-6)
    B.step()
    worst = 0.0
    for m, c in enumerate(cases):
        s = fem.TetFiniteElementSolver_Hip.from_case(c, rtol=1e-6)
        s.step()
        assert B.last_stats(m)["newton_iterations"] == s.last_stats()["newton_iterations"], m
        xs = c.X.reshape(-1) + s.get_array("deformation")
        worst = max(worst, rel_err(c.X.reshape(-1) + B.get_array("deformation", m), xs))
        del s
    print(f"[gpu] 256-bar batch ({sum(c.nt for c in cases)} tets): max position rel err {worst:.3e}", flush=True)
    assert worst <= 1e-8


def _rollout_fd_and_oracle(rc, ndir, oracle_bars, h=1e-6, verbose=False):
    """Rollout gradient of the batch (per-tet log E, v₀) against central differences along ndir random
    directions for bar 0 (2·ndir perturbed copies stepped as one batch) and against the oracle's gradient for
    the first oracle_bars bars.  Returns (worst FD relative difference, worst oracle relative error)."""
    import torch

    from ssim_amd import adjoint, batch

    tight = dict(pcg_rtol=1e-12, newton_polish=2)
    nb, nv, nt = len(rc.cases), rc.cases[0].nv, rc.cases[0].nt
    dt = rc.cases[0].dt

    def run(cases, log_e, v0, nu, grad):
        B = batch.TetFiniteElementBatch_Hip.from_cases(cases, rtol=1e-8)
        B.set_solver_params(**tight)
        prog = (lambda phase, k: print(f"[gpu]   {phase} step {k} ({len(cases)} meshes)", flush=True)) if verbose else None
        r = adjoint.Rollout(B, nu, rc.steps, dt, progress=prog)
        le = torch.as_tensor(log_e, device="cuda").requires_grad_(grad)
        vv = torch.as_tensor(v0, device="cuda").requires_grad_(grad)
        u0 = torch.as_tensor(np.concatenate([c.u0 for c in cases]), device="cuda")
        uN = adjoint.rollout(r, le, vv, u0).view(len(cases), nv, 3)
        run.last_stats = [B.last_stats(m) for m in range(min(2, len(cases)))]
        fe = torch.as_tensor(rc.free_end[0], device="cuda")
        return uN[:, fe, :], le, vv

    U, le, vv = run(rc.cases, rc.log_e, np.zeros(3 * nv * nb), rc.nu, True)
    print(f"[gpu] rollout last-step stats (batch): {run.last_stats}", flush=True)
    tg = torch.as_tensor(rc.targets, device="cuda")[:, None, :]
    L = ((U - tg) ** 2).sum()
    L.backward()
    g_le = le.grad.view(nb, nt).cpu().numpy()
    g_v0 = vv.grad.view(nb, 3 * nv).cpu().numpy()

    rng = np.random.default_rng(0)
    D_le = rng.standard_normal((ndir, nt))
    D_v = rng.standard_normal((ndir, 3 * nv))
    D_v[:, rc.cases[0].dof_type == 1] = 0.0
    le2 = np.concatenate([rc.log_e[:nt] + s * h * D_le[k] for k in range(ndir) for s in (1, -1)])
    v02 = np.concatenate([s * h * D_v[k] for k in range(ndir) for s in (1, -1)])
    with torch.no_grad():
        U2, _, _ = run([rc.cases[0]] * (2 * ndir), le2, v02, np.tile(rc.nu[:nt], 2 * ndir), False)
        Ls = ((U2 - tg[0]) ** 2).sum(dim=(1, 2)).cpu().numpy()
    fd = (Ls[0::2] - Ls[1::2]) / (2 * h)
    ad = D_le @ g_le[0] + D_v @ g_v0[0]
    rel = np.abs(fd - ad) / np.abs(ad)
    fd_worst = float(np.max(rel))
    print(f"[gpu] FD directions: |ad| {np.array2string(np.abs(ad), precision=3, max_line_width=10**4)}, "
          f"rel {np.array2string(rel, precision=2, max_line_width=10**4)}, loss {float(Ls.mean()):.6e}", flush=True)

    orc_worst = 0.0
    for b in range(oracle_bars):
        c, fe, t = rc.cases[b], rc.free_end[b], rc.targets[b]
        lam, mu = c.lam, c.mu

        def lg(u, fe=fe, t=t, c=c):
            d = u.reshape(-1, 3)[fe] - t
            gg = np.zeros((c.nv, 3))
            gg[fe] = 2 * d
            return float(np.sum(d * d)), gg.reshape(-1)

        _, og = oracle.rollout_gradient(c, rc.steps, lg, lam=lam, mu=mu)
        orc_worst = max(orc_worst, rel_err(g_le[b], lam * og["lam"] + mu * og["mu"]), rel_err(g_v0[b], og["v0"]))
    return fd_worst, orc_worst


def test_config5_structure_rollout_gradient_reduced():
    """Configuration 5's rollout (20 steps, per-tet log E, free-end target loss) on 4 bars of 16×4×4 cubes:
    gradients within 1e-5 of central differences (8 directions) and 1e-6 of the oracle's gradient."""
    rc = mesh.rollout_batch(4, nx=16, ny=4, nz=4)
    fd, orc = _rollout_fd_and_oracle(rc, ndir=8, oracle_bars=1)
    print(f"[gpu] config-5 structure, 4 bars 16x4x4, 20 steps: FD {fd:.2e}, oracle {orc:.2e}", flush=True)
    assert fd <= 1e-5 and orc <= 1e-6


def test_config5_full_rollout_gradient_fd():
    """Configuration 5 at full size: 32 bars of 16×16×64 cubes, 20 steps; the rollout gradient of bar 0 against
    central differences along 64 random directions of (log E, v₀)."""
    rc = mesh.rollout_batch(32)
    fd, _ = _rollout_fd_and_oracle(rc, ndir=64, oracle_bars=0, verbose=True)
    print(f"[gpu] configuration 5 full size: FD (64 directions, 20 steps) {fd:.2e}", flush=True)
    assert fd <= 1e-5


def test_config5_full_bar_gradient_matches_oracle():
    """One full-size configuration-5 bar (16×16×64 cubes) over its first 2 steps: the rollout gradient against
    the oracle's (CPU, about a minute per step at this size)."""
    short = mesh.rollout_batch(1, steps=2)
    print("[gpu] oracle gradient of one full-size bar over 2 steps (CPU)", flush=True)
    _, orc = _rollout_fd_and_oracle(short, ndir=2, oracle_bars=1, verbose=True)
    print(f"[gpu] configuration 5 full size: oracle (bar 0, 2 steps) {orc:.2e}", flush=True)
    assert orc <= 1e-6
