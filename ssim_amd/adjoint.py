"""Differentiable rollouts of the implicit FEM time step (extension: the reference is forward-only).

``Rollout`` runs ``steps`` backward-Euler steps of a ``TetFiniteElementBatch_Hip`` (a single mesh is a batch of
one) from an initial state (u₀, v₀) with per-tet material log E (fixed per-tet ν), and returns the final
deformation u_N.  ``rollout(...)`` wraps it in a ``torch.autograd.Function``: the backward pass is the adjoint
recursion over the steps, on the GPU through ``ssim_batch_adjoint_step``:

  step n:   û_n = u_n + dt·v_n + dt²·a (free dofs),  u_{n+1} solves g(u; û_n, λ, μ) = 0,  v_{n+1} = (u_{n+1} − u_n)/dt
  adjoint:  w = ū_{n+1} + v̄_{n+1}/dt;  H l = w;  ∂L/∂û_n = M l;  ū_n = M l − v̄_{n+1}/dt;  v̄_n = dt·M l;
            ∂L/∂λ_e += −dt²V_e (∂P/∂λ : Ds(l)B_e), ∂L/∂μ_e likewise;  ∂L/∂log E_e = λ_e ∂L/∂λ_e + μ_e ∂L/∂μ_e

(time_step.hpp:499-543 for the forward step).  Checkpoints: the state (u_n, v_n) is kept every
``checkpoint_every`` steps; the backward pass recomputes each segment's other states from its checkpoint and
re-assembles every step's Hessian at its converged u_{n+1} instead of storing it.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from .batch import TetFiniteElementBatch_Hip

_I = ctypes.POINTER(ctypes.c_int)
_D = ctypes.POINTER(ctypes.c_double)


def lame_from_log_e(log_e: np.ndarray, nu: np.ndarray):
    """(λ, μ) from per-tet log E and ν (elast/basic_elast.hpp:11-20)."""
    E = np.exp(log_e)
    return E * nu / ((1 + nu) * (1 - 2 * nu)), E / (2 * (1 + nu))


class Rollout:
    """Forward rollout and adjoint of a batch of meshes (see the module docstring)."""

    def __init__(self, batch: TetFiniteElementBatch_Hip, nu: np.ndarray, steps: int, dt: float,
                 checkpoint_every: int = 1, progress=None):
        if steps < 1 or checkpoint_every < 1:
            raise ValueError("steps and checkpoint_every must be positive")
        self.b = batch
        self.nu = np.ascontiguousarray(nu, dtype=np.float64).reshape(-1)
        self.steps = int(steps)
        self.dt = float(dt)
        self.k = int(checkpoint_every)
        self.n = int(3 * batch._nv.sum())
        self.nt = int(batch._nt.sum())
        if self.nu.size != self.nt:
            raise ValueError("nu needs one entry per tet")
        self.dev = torch.device("cuda")
        _lib.call("ssim_batch_set_stream", batch.handle, ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
        self.ckpt = {}
        self.lam = self.mu = None
        self.adjoint_pcg_iterations = 0
        self.progress = progress  # optional callable(phase, step) after each forward / adjoint step

    # ---- device state ----
    def _set(self, name, t):
        t = t.detach().to(self.dev, torch.float64).contiguous()
        _lib.call("ssim_batch_set_array_device", self.b.handle, _lib.ARRAYS[name], ctypes.c_void_p(t.data_ptr()))

    def _get(self, name):
        out = torch.empty(self.n, dtype=torch.float64, device=self.dev)
        _lib.call("ssim_batch_get_array_device", self.b.handle, _lib.ARRAYS[name], ctypes.c_void_p(out.data_ptr()))
        return out

    def _material(self, log_e):
        lam, mu = lame_from_log_e(np.asarray(log_e.detach().cpu().numpy(), dtype=np.float64), self.nu)
        self.lam, self.mu = np.ascontiguousarray(lam), np.ascontiguousarray(mu)
        _lib.call("ssim_batch_set_material", self.b.handle, self.lam.ctypes.data_as(_D),
                  self.mu.ctypes.data_as(_D))

    def _restore(self, n):
        u, v = self.ckpt[n]
        self._set("deformation", u)
        self._set("velocity", v)

    # ---- forward / backward ----
    def forward(self, log_e: torch.Tensor, v0: torch.Tensor, u0: torch.Tensor) -> torch.Tensor:
        self._material(log_e)
        self.ckpt = {0: (u0.detach().to(self.dev, torch.float64).clone(), v0.detach().to(self.dev, torch.float64).clone())}
        self._restore(0)
        for n in range(self.steps):
            self.b.step()
            if (n + 1) % self.k == 0 or n + 1 == self.steps:
                self.ckpt[n + 1] = (self._get("deformation"), self._get("velocity"))
            if self.progress:
                self.progress("forward", n)
        return self.ckpt[self.steps][0].clone()

    def backward(self, grad_uN: torch.Tensor):
        dev, n, nt = self.dev, self.n, self.nt
        ubar = grad_uN.detach().to(dev, torch.float64).contiguous().clone()
        vbar = torch.zeros(n, dtype=torch.float64, device=dev)
        glam = torch.zeros(nt, dtype=torch.float64, device=dev)
        gmu = torch.zeros(nt, dtype=torch.float64, device=dev)
        guhat = torch.empty(n, dtype=torch.float64, device=dev)
        gl = torch.empty(nt, dtype=torch.float64, device=dev)
        gm = torch.empty(nt, dtype=torch.float64, device=dev)
        iters = np.zeros(self.b.num_meshes, dtype=np.int32)
        self.adjoint_pcg_iterations = 0
        starts = list(range(0, self.steps, self.k))
        for s0 in reversed(starts):
            s1 = min(s0 + self.k, self.steps)
            # the segment's converged states u_{s0+1..s1}: checkpoint s1, the others recomputed from s0
            seg = {s1: self.ckpt[s1][0]}
            if s1 - s0 > 1:
                self._restore(s0)
                for m in range(s0, s1 - 1):
                    self.b.step()
                    seg[m + 1] = self._get("deformation")
            for m in range(s1 - 1, s0 - 1, -1):
                self._set("next_deform", seg[m + 1])
                w = ubar + vbar / self.dt
                _lib.call("ssim_batch_adjoint_step", self.b.handle, ctypes.c_void_p(w.data_ptr()),
                          ctypes.c_void_p(guhat.data_ptr()), ctypes.c_void_p(gl.data_ptr()),
                          ctypes.c_void_p(gm.data_ptr()), None, iters.ctypes.data_as(_I))
                self.adjoint_pcg_iterations += int(iters.max())
                glam += gl
                gmu += gm
                ubar, vbar = guhat - vbar / self.dt, self.dt * guhat
                if self.progress:
                    self.progress("adjoint", m)
        lam = torch.from_numpy(self.lam).to(dev)
        mu = torch.from_numpy(self.mu).to(dev)
        return lam * glam + mu * gmu, vbar, ubar


class _RolloutFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, log_e, v0, u0, runner):
        ctx.runner = runner
        return runner.forward(log_e, v0, u0)

    @staticmethod
    def backward(ctx, grad_uN):
        g_log_e, g_v0, g_u0 = ctx.runner.backward(grad_uN)
        return g_log_e, g_v0, g_u0, None


def rollout(runner: Rollout, log_e: torch.Tensor, v0: torch.Tensor, u0: torch.Tensor) -> torch.Tensor:
    """u_N as a differentiable function of (per-tet log E, v₀, u₀); gradients through the adjoint."""
    return _RolloutFn.apply(log_e, v0, u0, runner)
