"""Batched material evaluation on the GPU: the reference's vmap over an elastic model's stress / Hessian operators
(elast/basic_elast.hpp:144-219; tests/fem_intro/main.cpp:302-309), for 3D (9x9 Hessians) and 2D (4x4) deformation
gradients, with the stable Neo-Hookean (elast/stable_neohookean.hpp) or linear (elast/linear.hpp) model.
Every evaluation runs in libssim_hip.so (``ssim_material_eval``); torch only provides the device buffers.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib

MODELS = {"stable_neohookean": 0, "linear": 1}


def evaluate(F, lam, mu, model: str = "stable_neohookean", make_spsd: bool = False, want_H: bool = True):
    """F: (n, d, d) with d = 2 or 3, row-major matrices; lam, mu: scalars or (n,) arrays.  Returns a dict with
    "psi" (n,), "P" (n, d, d) and, if requested, "H" (n, d², d²) with H[:, a·d+c, b·d+e] = ∂²Ψ/∂F_ac∂F_be; with
    make_spsd the Hessians have their eigenvalues raised to at least 1e-3 (the time step's SPD projection)."""
    import torch

    if model not in MODELS:
        raise ValueError(f"Unknown material: {model}")
    F = np.ascontiguousarray(F, dtype=np.float64)
    if F.ndim != 3 or F.shape[1] != F.shape[2] or F.shape[1] not in (2, 3):
        raise ValueError(f"F must have shape (n, 2, 2) or (n, 3, 3), got {F.shape}")
    n, d = F.shape[0], F.shape[1]
    lam = np.array(np.broadcast_to(np.asarray(lam, dtype=np.float64), (n,)))  # writable copies
    mu = np.array(np.broadcast_to(np.asarray(mu, dtype=np.float64), (n,)))
    _lib.load()
    dev = torch.device("cuda")
    t = lambda a: torch.as_tensor(a, device=dev)
    Fd, ld, md = t(F), t(lam), t(mu)
    psi = torch.empty(n, dtype=torch.float64, device=dev)
    P = torch.empty((n, d, d), dtype=torch.float64, device=dev)
    H = torch.empty((n, d * d, d * d), dtype=torch.float64, device=dev) if want_H else None
    p = lambda x: ctypes.c_void_p(x.data_ptr()) if x is not None else None
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    _lib.call("ssim_material_eval", d, MODELS[model], ctypes.c_longlong(n), p(Fd), p(ld), p(md),
              1 if make_spsd else 0, p(psi), p(P), p(H), stream)
    torch.cuda.current_stream().synchronize()
    out = {"psi": psi.cpu().numpy(), "P": P.cpu().numpy()}
    if want_H:
        out["H"] = H.cpu().numpy()
    return out
