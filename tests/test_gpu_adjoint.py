"""Adjoint of the implicit step on MI355X, against the CPU oracle's adjoint (sparse direct solve) and central
finite differences of the GPU forward rollout.

Bars: one adjoint step (adjoint vector, ∂L/∂û, per-tet ∂L/∂λ and ∂L/∂μ) within 1e-8 of the oracle at the same
converged state (PCG at rtol 1e-12 against a direct solve, so the bar carries the condition number of H); the
batched adjoint within 1e-9 of the single-mesh one per mesh; rollout gradients (per-tet log E, v₀) within 1e-6 of
the oracle's rollout gradient and within 1e-5 of central differences (step 1e-6) along random directions.  Steps are converged with two polish Newton
steps so that the finite differences resolve the loss.
"""
import ctypes

import numpy as np
import pytest

import oracle
from helpers import rel_err
from ssim_amd import _lib, batch, fem, mesh

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

TIGHT = dict(pcg_rtol=1e-12, newton_polish=2)


def _dev(a):
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device="cuda")


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _bar(nx, ny, nz, seed, E=1e5):
    c = mesh.cantilever_bar(nx, ny, nz, seed=seed, E=E, nu=0.45)
    rng = np.random.default_rng(seed)
    c.v0 = 0.01 * rng.standard_normal(c.v0.shape)
    c.v0[c.dof_type == 1] = 0.0
    return c


def test_adjoint_step_matches_oracle():
    case = _bar(5, 2, 3, seed=2)
    g = fem.TetFiniteElementSolver_Hip.from_case(case, rtol=1e-8)
    g.set_solver_params(**TIGHT)
    g.step()
    u_star = g.get_array("next_deform")
    rng = np.random.default_rng(4)
    gu = rng.standard_normal(3 * case.nv)
    gu_d, guhat_d, lv_d = _dev(gu), torch.empty(3 * case.nv, dtype=torch.float64, device="cuda"), \
        torch.empty(3 * case.nv, dtype=torch.float64, device="cuda")
    gl_d = torch.empty(case.nt, dtype=torch.float64, device="cuda")
    gm_d = torch.empty(case.nt, dtype=torch.float64, device="cuda")
    it = ctypes.c_int()
    torch.cuda.synchronize()
    _lib.call("ssim_adjoint_step", g.handle, _ptr(gu_d), _ptr(guhat_d), _ptr(gl_d), _ptr(gm_d), _ptr(lv_d),
              ctypes.byref(it))
    o = oracle.TimeStep.from_case(case)
    lv, guhat, gl, gm = o.adjoint_step(u_star, gu)
    print(f"[gpu] adjoint step: {it.value} PCG iterations", flush=True)
    errs = [rel_err(lv_d.cpu().numpy(), lv), rel_err(guhat_d.cpu().numpy(), guhat),
            rel_err(gl_d.cpu().numpy(), gl), rel_err(gm_d.cpu().numpy(), gm)]
    print(f"[gpu] adjoint step vs oracle: l {errs[0]:.2e}, M l {errs[1]:.2e}, dL/dlam {errs[2]:.2e}, "
          f"dL/dmu {errs[3]:.2e}", flush=True)
    assert max(errs) <= 1e-8
    assert not guhat_d.cpu().numpy()[case.dof_type == 1].any()


def test_batched_adjoint_matches_single_mesh_adjoint():
    cases = mesh.bar_batch(3, seed=8, lo=3, hi=6)
    B = batch.TetFiniteElementBatch_Hip.from_cases(cases, rtol=1e-8)
    B.set_solver_params(**TIGHT)
    B.step()
    n = sum(3 * c.nv for c in cases)
    nt = sum(c.nt for c in cases)
    rng = np.random.default_rng(1)
    gu = _dev(rng.standard_normal(n))
    guhat = torch.empty(n, dtype=torch.float64, device="cuda")
    gl = torch.empty(nt, dtype=torch.float64, device="cuda")
    gm = torch.empty(nt, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    _lib.call("ssim_batch_adjoint_step", B.handle, _ptr(gu), _ptr(guhat), _ptr(gl), _ptr(gm), None, None)
    vo = to = 0
    for m, c in enumerate(cases):
        s = fem.TetFiniteElementSolver_Hip.from_case(c, rtol=1e-8)
        s.set_solver_params(**TIGHT)
        s.set_array("next_deform", B.get_array("next_deform", m))
        gu_m = gu[vo:vo + 3 * c.nv].contiguous()
        gh_m = torch.empty(3 * c.nv, dtype=torch.float64, device="cuda")
        gl_m = torch.empty(c.nt, dtype=torch.float64, device="cuda")
        gm_m = torch.empty(c.nt, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        _lib.call("ssim_adjoint_step", s.handle, _ptr(gu_m), _ptr(gh_m), _ptr(gl_m), _ptr(gm_m), None, None)
        assert rel_err(guhat[vo:vo + 3 * c.nv].cpu().numpy(), gh_m.cpu().numpy()) <= 1e-9, m
        assert rel_err(gl[to:to + c.nt].cpu().numpy(), gl_m.cpu().numpy()) <= 1e-9, m
        assert rel_err(gm[to:to + c.nt].cpu().numpy(), gm_m.cpu().numpy()) <= 1e-9, m
        vo += 3 * c.nv
        to += c.nt


def _free_end_loss(cases, targets):
    """L = Σ_meshes Σ_{free-end vertices} ‖u_v − t_m‖² on the compact concatenated u."""
    idx, tgt = [], []
    off = 0
    for c, t in zip(cases, targets):
        fe = np.nonzero(c.X[:, 0] == c.X[:, 0].max())[0]
        idx.append(off + fe)
        tgt.append(np.broadcast_to(t, (fe.size, 3)))
        off += c.nv
    idx = np.concatenate(idx)
    tgt = np.concatenate(tgt)
    return idx, tgt


def test_rollout_gradient_matches_oracle_and_finite_differences():
    from ssim_amd import adjoint

    cases = [_bar(4, 2, 2, seed=3), _bar(3, 2, 3, seed=5, E=3e5)]
    steps = 4
    rng = np.random.default_rng(9)
    targets = [np.array([0.0, 0.012, 0.016]), np.array([0.01, -0.01, 0.0])]
    idx, tgt = _free_end_loss(cases, targets)
    B = batch.TetFiniteElementBatch_Hip.from_cases(cases, rtol=1e-8)
    B.set_solver_params(**TIGHT)
    nu = np.concatenate([np.full(c.nt, 0.45) for c in cases])
    log_e0 = np.concatenate([np.log(1e5) + 0.3 * rng.standard_normal(c.nt) for c in cases])
    v0 = np.concatenate([c.v0 for c in cases])
    u0 = np.concatenate([c.u0 for c in cases])
    runner = adjoint.Rollout(B, nu, steps, cases[0].dt)
    tidx = torch.as_tensor(idx, device="cuda")
    ttgt = _dev(tgt)

    def loss_of(log_e, v):
        uN = adjoint.rollout(runner, log_e, v, _dev(u0))
        d = uN.view(-1, 3)[tidx] - ttgt
        return (d * d).sum()

    le = _dev(log_e0).requires_grad_(True)
    vv = _dev(v0).requires_grad_(True)
    L = loss_of(le, vv)
    L.backward()
    g_le = le.grad.cpu().numpy()
    g_v0 = vv.grad.cpu().numpy()
    print(f"[gpu] rollout loss {L.item():.6e}, adjoint PCG iterations {runner.adjoint_pcg_iterations}", flush=True)

    # oracle gradient, mesh by mesh (meshes are independent)
    vo = to = 0
    for m, (c, t) in enumerate(zip(cases, targets)):
        lam, mu = adjoint.lame_from_log_e(log_e0[to:to + c.nt], nu[to:to + c.nt])
        fe = np.nonzero(c.X[:, 0] == c.X[:, 0].max())[0]

        def lg(u, fe=fe, t=t):
            d = u.reshape(-1, 3)[fe] - t
            gg = np.zeros((c.nv, 3))
            gg[fe] = 2 * d
            return float(np.sum(d * d)), gg.reshape(-1)

        _, og = oracle.rollout_gradient(c, steps, lg, lam=lam, mu=mu)
        o_le = lam * og["lam"] + mu * og["mu"]
        assert rel_err(g_le[to:to + c.nt], o_le) <= 1e-6, m
        assert rel_err(g_v0[vo:vo + 3 * c.nv], og["v0"]) <= 1e-6, m
        vo += 3 * c.nv
        to += c.nt

    # central differences of the GPU forward along random directions
    h = 1e-6
    worst = 0.0
    with torch.no_grad():
        for k in range(4):
            d_le = rng.standard_normal(log_e0.size)
            d_v = rng.standard_normal(v0.size)
            d_v[np.concatenate([c.dof_type for c in cases]) == 1] = 0.0
            Lp = loss_of(_dev(log_e0 + h * d_le), _dev(v0 + h * d_v)).item()
            Lm = loss_of(_dev(log_e0 - h * d_le), _dev(v0 - h * d_v)).item()
            fd = (Lp - Lm) / (2 * h)
            ad = g_le @ d_le + g_v0 @ d_v
            worst = max(worst, abs(fd - ad) / abs(ad))
    print(f"[gpu] rollout FD agreement: worst relative difference {worst:.2e}", flush=True)
    assert worst <= 1e-5


def test_checkpoint_segments_give_the_same_gradient():
    from ssim_amd import adjoint

    cases = [_bar(3, 2, 2, seed=6)]
    B = batch.TetFiniteElementBatch_Hip.from_cases(cases, rtol=1e-8)
    B.set_solver_params(**TIGHT)
    nu = np.full(cases[0].nt, 0.45)
    log_e = _dev(np.full(cases[0].nt, np.log(1e5)))
    v0, u0 = _dev(cases[0].v0), _dev(cases[0].u0)
    grads = []
    for k in (1, 3):
        r = adjoint.Rollout(B, nu, 5, cases[0].dt, checkpoint_every=k)
        le = log_e.clone().requires_grad_(True)
        uN = adjoint.rollout(r, le, v0, u0)
        uN.pow(2).sum().backward()
        grads.append(le.grad.cpu().numpy())
    assert np.array_equal(grads[0], grads[1])
