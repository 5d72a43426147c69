"""P1 triangles in 2D on the GPU: the per-element quantities of the reference's basic_time_step with
PhysicalDim 2, TopologyDim 3 (finite_elements/time_step.hpp, def_grad.hpp, rest_vol.hpp, bilinears.hpp), which the
reference templates but never instantiates.  Every evaluation runs in libssim_hip.so (``ssim_tri_eval``); torch
only provides the device buffers.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .material import MODELS

FIELDS = ("area", "dminv", "F", "energy", "grad", "stiff", "mass", "laplacian")
SHAPES = {"area": (), "dminv": (2, 2), "F": (2, 2), "energy": (), "grad": (3, 2), "stiff": (6, 6), "mass": (3, 3),
          "laplacian": (3, 3)}


def evaluate(X, cells, u, lam, mu, model: str = "stable_neohookean", make_spsd: bool = False, density: float = 1.0,
             want=FIELDS):
    """X: (nv, 2) rest positions, cells: (nt, 3), u: (nv, 2) displacements, lam / mu: scalars or (nt,) arrays.
    Returns a dict with the requested fields, each with leading axis nt: "area", "dminv" (Dm⁻¹) and "F" (2, 2)
    row-major, "energy" (A·Ψ), "grad" (3, 2) per vertex, "stiff" (6, 6) over dofs 2·vertex + component (H
    SPD-projected if make_spsd), "mass" (3, 3) times density, "laplacian" (3, 3).  A triangle with a vertex index
    outside [0, nv) gets NaN."""
    import torch

    if model not in MODELS:
        raise ValueError(f"Unknown material: {model}")
    unknown = set(want) - set(FIELDS)
    if unknown:
        raise ValueError(f"Unknown fields: {sorted(unknown)}")
    X = np.ascontiguousarray(X, dtype=np.float64)
    u = np.ascontiguousarray(u, dtype=np.float64)
    C = np.ascontiguousarray(cells, dtype=np.int32)
    if X.ndim != 2 or X.shape[1] != 2:
        raise ValueError(f"X must have shape (nv, 2), got {X.shape}")
    if u.shape != X.shape:
        raise ValueError(f"u must have shape {X.shape}, got {u.shape}")
    if C.ndim != 2 or C.shape[1] != 3:
        raise ValueError(f"cells must have shape (nt, 3), got {C.shape}")
    nv, nt = X.shape[0], C.shape[0]
    lam = np.array(np.broadcast_to(np.asarray(lam, dtype=np.float64), (nt,)))
    mu = np.array(np.broadcast_to(np.asarray(mu, dtype=np.float64), (nt,)))
    _lib.load()
    dev = torch.device("cuda")
    t = lambda a: torch.as_tensor(a, device=dev)
    Xd, Cd, ud, ld, md = t(X), t(C), t(u), t(lam), t(mu)
    out = {k: torch.empty((nt,) + SHAPES[k], dtype=torch.float64, device=dev) for k in want}
    p = lambda x: ctypes.c_void_p(x.data_ptr()) if x is not None else None
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    _lib.call("ssim_tri_eval", MODELS[model], nv, nt, p(Xd), p(Cd), p(ud), p(ld), p(md), float(density),
              1 if make_spsd else 0, *[p(out.get(k)) for k in FIELDS], stream)
    torch.cuda.current_stream().synchronize()
    return {k: v.cpu().numpy() for k, v in out.items()}
