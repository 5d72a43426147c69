"""Writes tests/golden/snh_autograd.json: stable Neo-Hookean Ψ, P = ∂Ψ/∂F and H = ∂²Ψ/∂F² by torch FP64
autograd of the closed-form energy (ssim/include/ssim/elast/stable_neohookean.hpp:23-36, δ = 1), and the SPD
projection H⁺ = Q·max(Λ, 1e-3)·Qᵀ (finite_elements/time_step.hpp:50-56) by numpy.linalg.eigh.

The autograd derivatives are independent of the hand-derived P and H of the oracle and of the HIP kernels, so
they pin both.  Run:  python tests/golden/make_golden.py
"""
import json
import os

import numpy as np
import torch

HERE = os.path.dirname(os.path.abspath(__file__))


def psi(F, lam, mu):
    delta = 1.0
    Ic = (F * F).sum()
    J = torch.det(F)
    alpha = (1 - 1 / (3 + delta)) * mu / lam + 1
    zero = 0.5 * (lam * (1 - alpha) ** 2 - mu * np.log(3 + delta))
    return 0.5 * (mu * (Ic - 3) + lam * (J - alpha) ** 2 - mu * torch.log(Ic + delta)) - zero


def lame(E, nu):
    return E * nu / ((1 + nu) * (1 - 2 * nu)), E / (2 * (1 + nu))


def main():
    rng = np.random.default_rng(20251203)
    cases = []
    # the reference's own test point (tests/fem_intro/main.cpp:90, :109)
    Fs = [np.eye(3), np.array([[1, 0.3, 0.2], [0.2, 2, 0.1], [0.1, 0.3, 1]])]
    mats = [lame(1e6, 0.3), lame(1e6, 0.3)]
    for k in range(30):
        kind = k % 5
        if kind == 0:
            F = np.eye(3) + 0.1 * rng.standard_normal((3, 3))  # near rest
        elif kind == 1:
            F = np.eye(3) + 0.5 * rng.standard_normal((3, 3))  # large strain
        elif kind == 2:
            F = np.diag([0.5, 0.7, 0.6]) + 0.05 * rng.standard_normal((3, 3))  # compressed
        elif kind == 3:
            F = np.diag([1.0, 1.0, -0.3]) + 0.1 * rng.standard_normal((3, 3))  # inverted
        else:
            Q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
            F = Q @ np.diag([1.2, 0.9, 0.8])  # rotated stretch
        Fs.append(F)
        E = float(np.exp(rng.uniform(np.log(1e4), np.log(1e6))))
        nu = float(rng.uniform(0.3, 0.48))
        mats.append(lame(E, nu))
    for F, (lam, mu) in zip(Fs, mats):
        Ft = torch.tensor(F, dtype=torch.float64, requires_grad=True)
        e = psi(Ft, lam, mu)
        (P,) = torch.autograd.grad(e, Ft, create_graph=True)
        H = torch.zeros(9, 9, dtype=torch.float64)
        Pf = P.reshape(-1)
        for i in range(9):
            (gi,) = torch.autograd.grad(Pf[i], Ft, retain_graph=True)
            H[i] = gi.reshape(-1)
        Hn = H.detach().numpy()
        Hn = 0.5 * (Hn + Hn.T)
        w, Q = np.linalg.eigh(Hn)
        Hp = (Q * np.maximum(w, 1e-3)) @ Q.T
        cases.append(dict(F=F.reshape(-1).tolist(), lam=lam, mu=mu, psi=float(e.detach()),
                          P=P.detach().numpy().reshape(-1).tolist(), H=Hn.reshape(-1).tolist(),
                          Hp=Hp.reshape(-1).tolist(), eig=w.tolist()))
    out = dict(generator="tests/golden/make_golden.py (torch FP64 autograd + numpy eigh)", floor=1e-3, cases=cases)
    with open(os.path.join(HERE, "snh_autograd.json"), "w") as f:
        json.dump(out, f, indent=0)
    print(f"wrote {len(cases)} cases")


if __name__ == "__main__":
    main()
