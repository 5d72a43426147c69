"""ctypes front end of the CPU FP64 oracle (oracle/ssim_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg, which
compare the HIP path against it or time it.  The product (ssim_amd) never imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libssim_oracle.so")
_lib = None

_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)
_U8 = ctypes.POINTER(ctypes.c_ubyte)
_V = ctypes.c_void_p


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        sig = {
            "orc_mu": (ctypes.c_double, [ctypes.c_double, ctypes.c_double]),
            "orc_lambda": (ctypes.c_double, [ctypes.c_double, ctypes.c_double]),
            "orc_snh_energy": (ctypes.c_double, [_D, ctypes.c_double, ctypes.c_double]),
            "orc_snh_stress": (ctypes.c_double, [_D, ctypes.c_double, ctypes.c_double, _D]),
            "orc_snh_hessian": (None, [_D, ctypes.c_double, ctypes.c_double, _D]),
            "orc_sym_eig": (ctypes.c_int, [ctypes.c_int, _D, _D, _D]),
            "orc_spd_project9": (None, [_D, ctypes.c_double, _D]),
            "orc_local_force": (None, [_D, _D, _D]),
            "orc_local_stiffness": (None, [_D, _D, _D]),
            "orc_pfpx": (None, [_D, _D]),
            "orc_ts_create": (_V, [_D, ctypes.c_int, _I, ctypes.c_int, ctypes.c_double, _D, _D, ctypes.c_double]),
            "orc_ts_destroy": (None, [_V]),
            "orc_ts_nnz": (ctypes.c_int, [_V]),
            "orc_ts_csr_pattern": (None, [_V, _I, _I]),
            "orc_ts_mass_values": (None, [_V, _D]),
            "orc_ts_total_mass": (ctypes.c_double, [_V]),
            "orc_ts_rest_data": (None, [_V, _D, _D]),
            "orc_ts_set_dirichlet": (None, [_V, _U8, _D]),
            "orc_ts_set_ext_accel": (None, [_V, _D]),
            "orc_ts_set_threshold": (None, [_V, ctypes.c_double]),
            "orc_ts_set_state": (None, [_V, _D, _D, _D]),
            "orc_ts_get_state": (None, [_V, _D, _D, _D]),
            "orc_ts_set_vectors": (None, [_V, _D, _D]),
            "orc_ts_get_vectors": (None, [_V, _D, _D, _D]),
            "orc_ts_set_solver": (None, [_V, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_double,
                                         ctypes.c_double, ctypes.c_int]),
            "orc_ts_set_polish": (None, [_V, ctypes.c_int]),
            "orc_ts_set_lbfgs": (None, [_V, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                        ctypes.c_int]),
            "orc_ts_elements": (None, [_V, _D, _D, _D, _D, _D, _D, ctypes.c_int]),
            "orc_ts_energy_grad": (ctypes.c_double, [_V, _D, _D]),
            "orc_ts_hessian": (None, [_V, _D, ctypes.c_int, ctypes.c_int, _D]),
            "orc_pcg": (ctypes.c_int, [ctypes.c_int, _I, _I, _D, _D, _D, ctypes.c_double, ctypes.c_int, _D]),
            "orc_ts_prepare_step": (None, [_V]),
            "orc_ts_compute_step": (ctypes.c_double, [_V, _D]),
            "orc_ts_compute_step_lbfgs": (ctypes.c_double, [_V, _D]),
            "orc_ts_compute_step_ncg": (ctypes.c_double, [_V, _D]),
            "orc_ts_compute_step_fe": (ctypes.c_double, [_V, _D]),
            "orc_ts_set_model": (None, [_V, ctypes.c_int]),
            "orc_ts_set_precond": (None, [_V, _D]),
            "orc_ts_build_pd": (None, [_V, ctypes.c_double]),
            "orc_ts_pd_values": (None, [_V, _D]),
            "orc_ts_compute_step_lbfgs_pd": (ctypes.c_double, [_V, _D]),
            "orc_ts_compute_step_lbfgs_h0": (ctypes.c_double, [_V, _D]),
            "orc_lin_energy": (ctypes.c_double, [_D, ctypes.c_double, ctypes.c_double]),
            "orc_lin_stress": (ctypes.c_double, [_D, ctypes.c_double, ctypes.c_double, _D]),
            "orc_lin_hessian": (None, [ctypes.c_double, ctypes.c_double, _D]),
            "orc_snh2_energy": (ctypes.c_double, [_D, ctypes.c_double, ctypes.c_double]),
            "orc_snh2_stress": (ctypes.c_double, [_D, ctypes.c_double, ctypes.c_double, _D]),
            "orc_snh2_hessian": (None, [_D, ctypes.c_double, ctypes.c_double, _D]),
            "orc_lin2_energy": (ctypes.c_double, [_D, ctypes.c_double, ctypes.c_double]),
            "orc_lin2_stress": (ctypes.c_double, [_D, ctypes.c_double, ctypes.c_double, _D]),
            "orc_lin2_hessian": (None, [ctypes.c_double, ctypes.c_double, _D]),
            "orc_ts_step_next": (None, [_V]),
            "orc_ts_set_damping": (ctypes.c_int, [_V, ctypes.c_double]),
            "orc_ts_rest_hessian_diag": (None, [_V, _D]),
            "orc_ts_spmv": (None, [_V, _D, _D, _D]),
            "orc_tri_eval": (None, [ctypes.c_int, ctypes.c_int, ctypes.c_int, _D, _I, _D, _D, _D, ctypes.c_double,
                                    ctypes.c_int] + [_D] * 8),
            "orc_num_threads": (ctypes.c_int, []),
            "orc_set_num_threads": (None, [ctypes.c_int]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(_D) if a is not None else None


def _ip(a):
    return a.ctypes.data_as(_I) if a is not None else None


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


# ---- element level ----
def snh_energy(F, lam, mu):
    F = _f64(F).reshape(9)
    return lib().orc_snh_energy(_dp(F), lam, mu)


def snh_stress(F, lam, mu):
    F = _f64(F).reshape(9)
    P = np.zeros(9)
    e = lib().orc_snh_stress(_dp(F), lam, mu, _dp(P))
    return e, P.reshape(3, 3)


def lin_stress(F, lam, mu):
    F = _f64(F).reshape(9)
    P = np.zeros(9)
    e = lib().orc_lin_stress(_dp(F), lam, mu, _dp(P))
    return e, P.reshape(3, 3)


def lin_energy(F, lam, mu):
    return lib().orc_lin_energy(_dp(_f64(F).reshape(9)), lam, mu)


def lin_hessian(lam, mu):
    H = np.zeros(81)
    lib().orc_lin_hessian(lam, mu, _dp(H))
    return H.reshape(9, 9)


def material2(model, F, lam, mu):
    """2D model ("stable_neohookean" or "linear"): (Ψ, P (2,2), H (4,4)) at the row-major 2x2 F."""
    F = _f64(F).reshape(4)
    P, H = np.zeros(4), np.zeros(16)
    if model == "linear":
        e = lib().orc_lin2_stress(_dp(F), lam, mu, _dp(P))
        lib().orc_lin2_hessian(lam, mu, _dp(H))
    else:
        e = lib().orc_snh2_stress(_dp(F), lam, mu, _dp(P))
        lib().orc_snh2_hessian(_dp(F), lam, mu, _dp(H))
    return e, P.reshape(2, 2), H.reshape(4, 4)


TRI_FIELDS = ("area", "dminv", "F", "energy", "grad", "stiff", "mass", "laplacian")
TRI_SHAPES = {"area": (), "dminv": (2, 2), "F": (2, 2), "energy": (), "grad": (3, 2), "stiff": (6, 6),
              "mass": (3, 3), "laplacian": (3, 3)}


def tri_eval(X, cells, u, lam, mu, model="stable_neohookean", make_spsd=False, density=1.0):
    """P1 triangles in 2D, per element: X (nv, 2), cells (nt, 3), u (nv, 2); lam, mu scalars or (nt,).  Returns
    a dict of TRI_FIELDS with the shapes TRI_SHAPES (leading axis nt)."""
    X, u = _f64(X).reshape(-1, 2), _f64(u).reshape(-1, 2)
    C = np.ascontiguousarray(cells, dtype=np.int32).reshape(-1, 3)
    nv, nt = X.shape[0], C.shape[0]
    lam = np.array(np.broadcast_to(_f64(lam), (nt,)))
    mu = np.array(np.broadcast_to(_f64(mu), (nt,)))
    out = {k: np.zeros((nt,) + TRI_SHAPES[k]) for k in TRI_FIELDS}
    lib().orc_tri_eval(1 if model == "linear" else 0, nv, nt, _dp(X), _ip(C), _dp(u), _dp(lam), _dp(mu),
                       float(density), 1 if make_spsd else 0, *[_dp(out[k]) for k in TRI_FIELDS])
    return out


def energy2(model, F, lam, mu):
    F = _f64(F).reshape(4)
    return (lib().orc_lin2_energy if model == "linear" else lib().orc_snh2_energy)(_dp(F), lam, mu)


def snh_hessian(F, lam, mu):
    F = _f64(F).reshape(9)
    H = np.zeros(81)
    lib().orc_snh_hessian(_dp(F), lam, mu, _dp(H))
    return H.reshape(9, 9)


def sym_eig(A):
    A = _f64(A)
    n = A.shape[0]
    w = np.zeros(n)
    Q = np.zeros((n, n))
    lib().orc_sym_eig(n, _dp(A.reshape(-1)), _dp(w), _dp(Q.reshape(-1)))
    return w, Q


def spd_project9(H, floor=1e-3):
    H = _f64(H).reshape(81)
    Hp = np.zeros(81)
    lib().orc_spd_project9(_dp(H), floor, _dp(Hp))
    return Hp.reshape(9, 9)


def local_force(P, B):
    f = np.zeros(12)
    lib().orc_local_force(_dp(_f64(P).reshape(9)), _dp(_f64(B).reshape(9)), _dp(f))
    return f.reshape(4, 3)


def local_stiffness(H, B):
    K = np.zeros(144)
    lib().orc_local_stiffness(_dp(_f64(H).reshape(81)), _dp(_f64(B).reshape(9)), _dp(K))
    return K.reshape(12, 12)


def pcg(rowptr, colidx, vals, b, rtol=1e-8, maxit=10000):
    rowptr = np.ascontiguousarray(rowptr, dtype=np.int32)
    colidx = np.ascontiguousarray(colidx, dtype=np.int32)
    vals = _f64(vals)
    b = _f64(b)
    x = np.zeros_like(b)
    rel = ctypes.c_double(0)
    it = lib().orc_pcg(b.shape[0], _ip(rowptr), _ip(colidx), _dp(vals), _dp(b), _dp(x), rtol, maxit,
                       ctypes.byref(rel))
    return x, it, rel.value


def cofactor(F):
    """cof F = ∂det F/∂F for F (..., 3, 3): columns f1×f2, f2×f0, f0×f1 of F's columns f0, f1, f2."""
    F = np.asarray(F, dtype=np.float64)
    f0, f1, f2 = F[..., :, 0], F[..., :, 1], F[..., :, 2]
    return np.stack([np.cross(f1, f2), np.cross(f2, f0), np.cross(f0, f1)], axis=-1)


def snh_dP_dlame(F):
    """(∂P/∂λ, ∂P/∂μ) of the stable Neo-Hookean first Piola stress (elast/stable_neohookean.hpp:23-62):
    P = λ(J − α)·cof F + μ(1 − 1/(Ic+δ))·F with α = 1 + (1 − 1/(3+δ))·μ/λ, δ = 1, is linear in (λ, μ), so
    ∂P/∂λ = (J − 1)·cof F and ∂P/∂μ = −(1 − 1/(3+δ))·cof F + (1 − 1/(Ic+δ))·F, independent of (λ, μ)."""
    F = np.asarray(F, dtype=np.float64)
    C = cofactor(F)
    J = np.linalg.det(F)
    Ic = np.sum(F * F, axis=(-2, -1))
    dl = (J - 1.0)[..., None, None] * C
    dm = -(1.0 - 1.0 / 4.0) * C + (1.0 - 1.0 / (Ic + 1.0))[..., None, None] * F
    return dl, dm


def lin_dP_dlame(F):
    """(∂P/∂λ, ∂P/∂μ) of linear elasticity, P = 2μ·ε + λ·tr(ε)·I with ε = ½(F + Fᵀ) − I (elast/linear.hpp:37-50):
    ∂P/∂λ = tr(ε)·I, ∂P/∂μ = 2ε."""
    F = np.asarray(F, dtype=np.float64)
    eps = 0.5 * (F + np.swapaxes(F, -1, -2)) - np.eye(3)
    tr = np.trace(eps, axis1=-2, axis2=-1)
    return tr[..., None, None] * np.eye(3), 2.0 * eps


def _edge_matrix(x, cells):
    """Ds per tet: columns x1 − x0, x2 − x0, x3 − x0 (def_grad.hpp:270-288), shape (nt, 3, 3)."""
    x0 = x[cells[:, 0]]
    return np.stack([x[cells[:, 1]] - x0, x[cells[:, 2]] - x0, x[cells[:, 3]] - x0], axis=2)


def rollout_gradient(case, steps, loss_grad, lam=None, mu=None, rtol=1e-8, pcg_rtol=1e-12, polish=2,
                     model="stable_neohookean"):
    """Forward rollout of `steps` Newton steps from case's initial state, then the adjoint recursion backwards.

    loss_grad(u_N) -> (L, ∂L/∂u_N).  Step n: û = u_n + dt·v_n + dt²·a (free dofs), u_{n+1} = solution,
    v_{n+1} = (u_{n+1} − u_n)/dt (time_step.hpp:499-543).  Backwards, with w = ū_{n+1} + v̄_{n+1}/dt:
    (l, ∂L/∂û) = adjoint_step(u_{n+1}, w), ū_n = ∂L/∂û − v̄_{n+1}/dt, v̄_n = dt·∂L/∂û.
    Returns (L, dict(u0, v0, lam, mu)) with per-tet Lamé gradients summed over steps."""
    lam = case.lam if lam is None else lam
    mu = case.mu if mu is None else mu
    ts = TimeStep(case.X, case.cells, case.dt, lam, mu, case.rho)
    ts.set_dirichlet(case.dof_type, case.dbc)
    ts.set_ext_accel(case.accel)
    ts.set_state(case.u0, np.zeros_like(case.u0), case.v0)
    ts.set_threshold(rtol)
    ts.set_solver(pcg_rtol=pcg_rtol)
    ts.set_polish(polish)
    ts.set_model(model)
    us = [np.array(case.u0, dtype=np.float64)]
    for _ in range(steps):
        ts.step()
        us.append(ts.get_state()[0])
    L, gu = loss_grad(us[-1])
    ubar, vbar = np.array(gu, dtype=np.float64), np.zeros(ts.n)
    glam, gmu = np.zeros(ts.nt), np.zeros(ts.nt)
    dt = ts.dt
    for k in range(steps - 1, -1, -1):
        _, guhat, gl, gm = ts.adjoint_step(us[k + 1], ubar + vbar / dt)
        glam += gl
        gmu += gm
        ubar, vbar = guhat - vbar / dt, dt * guhat
    return L, dict(u0=ubar, v0=vbar, lam=glam, mu=gmu)


def rollout_loss(case, steps, loss_grad, lam=None, mu=None, v0=None, rtol=1e-8, pcg_rtol=1e-12, polish=2,
                 model="stable_neohookean"):
    """Loss of the forward rollout only (finite-difference checks)."""
    lam = case.lam if lam is None else lam
    mu = case.mu if mu is None else mu
    ts = TimeStep(case.X, case.cells, case.dt, lam, mu, case.rho)
    ts.set_dirichlet(case.dof_type, case.dbc)
    ts.set_ext_accel(case.accel)
    ts.set_state(case.u0, np.zeros_like(case.u0), case.v0 if v0 is None else v0)
    ts.set_threshold(rtol)
    ts.set_solver(pcg_rtol=pcg_rtol)
    ts.set_polish(polish)
    ts.set_model(model)
    for _ in range(steps):
        ts.step()
    return loss_grad(ts.get_state()[0])[0]


class TimeStep:
    """Oracle mirror of ssim::fem::basic_time_step (finite_elements/time_step.hpp:171-715)."""

    def __init__(self, X, cells, dt, lam, mu, rho):
        self.dt = float(dt)
        self.dof = None
        self.model = "stable_neohookean"
        self.X = _f64(X).reshape(-1, 3)
        self.cells = np.ascontiguousarray(cells, dtype=np.int32).reshape(-1, 4)
        self.nv, self.nt = self.X.shape[0], self.cells.shape[0]
        self.n = 3 * self.nv
        lam = np.broadcast_to(_f64(lam), (self.nt,)).copy()
        mu = np.broadcast_to(_f64(mu), (self.nt,)).copy()
        self._h = lib().orc_ts_create(_dp(self.X), self.nv, _ip(self.cells), self.nt, dt, _dp(lam), _dp(mu), rho)
        self.nnz = lib().orc_ts_nnz(self._h)

    @classmethod
    def from_case(cls, case):
        ts = cls(case.X, case.cells, case.dt, case.lam, case.mu, case.rho)
        ts.set_dirichlet(case.dof_type, case.dbc)
        ts.set_ext_accel(case.accel)
        ts.set_state(case.u0, np.zeros_like(case.u0), case.v0)
        return ts

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_ts_destroy(self._h)
            self._h = None

    def csr_pattern(self):
        rowptr = np.zeros(self.n + 1, dtype=np.int32)
        colidx = np.zeros(self.nnz, dtype=np.int32)
        lib().orc_ts_csr_pattern(self._h, _ip(rowptr), _ip(colidx))
        return rowptr, colidx

    def mass_values(self):
        v = np.zeros(self.nnz)
        lib().orc_ts_mass_values(self._h, _dp(v))
        return v

    def total_mass(self):
        return lib().orc_ts_total_mass(self._h)

    def rest_data(self):
        B = np.zeros((self.nt, 3, 3))
        V = np.zeros(self.nt)
        lib().orc_ts_rest_data(self._h, _dp(B), _dp(V))
        return B, V

    def set_dirichlet(self, dof_type, dbc):
        d = np.ascontiguousarray(dof_type, dtype=np.uint8).reshape(-1)
        self.dof = d.copy()
        lib().orc_ts_set_dirichlet(self._h, d.ctypes.data_as(_U8), _dp(_f64(dbc).reshape(-1)))

    def set_ext_accel(self, a):
        lib().orc_ts_set_ext_accel(self._h, _dp(_f64(a).reshape(-1)))

    def set_threshold(self, rtol):
        lib().orc_ts_set_threshold(self._h, rtol)

    def set_state(self, u=None, u_prev=None, vel=None):
        conv = lambda a: None if a is None else _f64(a).reshape(-1)
        u, u_prev, vel = conv(u), conv(u_prev), conv(vel)
        lib().orc_ts_set_state(self._h, _dp(u), _dp(u_prev), _dp(vel))

    def get_state(self):
        u, up, v = np.zeros(self.n), np.zeros(self.n), np.zeros(self.n)
        lib().orc_ts_get_state(self._h, _dp(u), _dp(up), _dp(v))
        return u, up, v

    def set_vectors(self, u_next=None, u_hat=None):
        conv = lambda a: None if a is None else _f64(a).reshape(-1)
        u_next, u_hat = conv(u_next), conv(u_hat)
        lib().orc_ts_set_vectors(self._h, _dp(u_next), _dp(u_hat))

    def get_vectors(self):
        a, b, c = np.zeros(self.n), np.zeros(self.n), np.zeros(self.n)
        lib().orc_ts_get_vectors(self._h, _dp(a), _dp(b), _dp(c))
        return a, b, c

    def set_solver(self, max_newton=10, pcg_rtol=1e-8, pcg_maxit=10000, ls_c=0.3, ls_beta=0.5, ls_max=20):
        lib().orc_ts_set_solver(self._h, max_newton, pcg_rtol, pcg_maxit, ls_c, ls_beta, ls_max)

    def set_model(self, model):
        """"stable_neohookean" or "linear"."""
        lib().orc_ts_set_model(self._h, {"stable_neohookean": 0, "linear": 1}[model])
        self.model = model

    def set_precond(self, values):
        """NCG preconditioner factor L (z = L·Lᵀ·g) as values in the Hessian's CSR order; None removes it."""
        if values is None:
            lib().orc_ts_set_precond(self._h, None)
            return
        v = _f64(values).reshape(-1)
        if v.size != self.nnz:
            raise ValueError(f"expected {self.nnz} values, got {v.size}")
        self._prec = v
        lib().orc_ts_set_precond(self._h, _dp(v))

    def build_pd(self, kappa):
        """Projective-dynamics matrix M + dt²κK with the diffusivity κ (Dirichlet-filtered, current marks)."""
        lib().orc_ts_build_pd(self._h, float(kappa))

    def pd_csr(self):
        import scipy.sparse as sp

        rowptr, colidx = self.csr_pattern()
        vals = np.zeros(self.nnz)
        lib().orc_ts_pd_values(self._h, _dp(vals))
        return sp.csr_matrix((vals, colidx, rowptr), shape=(self.n, self.n))

    def set_polish(self, steps):
        lib().orc_ts_set_polish(self._h, int(steps))

    def set_lbfgs(self, m=8, max_iter=1000, ls_c=1e-4, ls_beta=0.5, ls_max=30):
        lib().orc_ts_set_lbfgs(self._h, m, max_iter, ls_c, ls_beta, ls_max)

    def elements(self, u, make_spsd=True, want_K=False):
        u = _f64(u).reshape(-1)
        nt = self.nt
        psi, P = np.zeros(nt), np.zeros((nt, 3, 3))
        H, Hp = np.zeros((nt, 9, 9)), np.zeros((nt, 9, 9))
        K = np.zeros((nt, 12, 12)) if want_K else None
        lib().orc_ts_elements(self._h, _dp(u), _dp(psi), _dp(P), _dp(H), _dp(Hp), _dp(K), int(make_spsd))
        return dict(psi=psi, P=P, H=H, Hp=Hp, K=K)

    def energy_grad(self, u):
        g = np.zeros(self.n)
        E = lib().orc_ts_energy_grad(self._h, _dp(_f64(u).reshape(-1)), _dp(g))
        return E, g

    def hessian(self, u, make_spsd=True, filter_dbc=True):
        vals = np.zeros(self.nnz)
        lib().orc_ts_hessian(self._h, _dp(_f64(u).reshape(-1)), int(make_spsd), int(filter_dbc), _dp(vals))
        return vals

    def hessian_csr(self, u, make_spsd=True, filter_dbc=True):
        import scipy.sparse as sp

        rowptr, colidx = self.csr_pattern()
        return sp.csr_matrix((self.hessian(u, make_spsd, filter_dbc), colidx, rowptr), shape=(self.n, self.n))

    def spmv(self, vals, x):
        y = np.zeros(self.n)
        lib().orc_ts_spmv(self._h, _dp(_f64(vals)), _dp(_f64(x)), _dp(y))
        return y

    def prepare_step(self):
        lib().orc_ts_prepare_step(self._h)

    def compute_step(self, method="newton"):
        st = np.zeros(8)
        fn = {"newton": lib().orc_ts_compute_step, "lbfgs": lib().orc_ts_compute_step_lbfgs,
              "ncg": lib().orc_ts_compute_step_ncg, "forward_euler": lib().orc_ts_compute_step_fe,
              "lbfgs_pd": lib().orc_ts_compute_step_lbfgs_pd, "lbfgs_h0": lib().orc_ts_compute_step_lbfgs_h0}[method]
        gn = fn(self._h, _dp(st))
        return gn, dict(newton=int(st[0]), pcg=int(st[1]), evals=int(st[2]), converged=bool(st[3]),
                        grad_norm=float(st[4]), ls_fail=int(st[5]), last_change=float(st[6]),
                        ls_rejected=int(st[7]))

    def set_damping(self, damping):
        """Velocity damping of step_next (time_step.hpp:337-341, 541-542); ValueError outside [0, 1)."""
        if lib().orc_ts_set_damping(self._h, float(damping)) != 0:
            raise ValueError("velocity damping must be in [0, 1).")

    def step_next(self):
        lib().orc_ts_step_next(self._h)

    def step(self, method="newton"):
        self.prepare_step()
        out = self.compute_step(method)
        self.step_next()
        return out

    def mass_csr(self):
        import scipy.sparse as sp

        rowptr, colidx = self.csr_pattern()
        return sp.csr_matrix((self.mass_values(), colidx, rowptr), shape=(self.n, self.n))

    def deformation_gradients(self, u):
        """F = Ds(X + u)·B per tet (def_grad.hpp:93-132), shape (nt, 3, 3)."""
        B, _ = self.rest_data()
        x = self.X + _f64(u).reshape(-1, 3)
        return _edge_matrix(x, self.cells) @ B

    def adjoint_step(self, u_star, grad_u):
        """Adjoint of one implicit step at its converged solution u_star, by a sparse direct solve.

        The step solves g(u; û, λ, μ) = M(u − û) + dt²·Σ_e V_e f_e(u) = 0 on the free dofs
        (time_step.hpp:577-624).  With H the unprojected, Dirichlet-filtered Hessian (time_step.hpp:626-658):
        H l = grad_u (free dofs), ∂L/∂û = M l (free dofs), ∂L/∂λ_e = −dt²·V_e·(∂P/∂λ : Ds(l)·B_e), likewise μ.
        Returns (l, grad_uhat, grad_lambda, grad_mu)."""
        import scipy.sparse.linalg as spla

        free = np.ones(self.n, dtype=bool) if self.dof is None else self.dof == 0
        rhs = np.where(free, _f64(grad_u).reshape(-1), 0.0)
        H = self.hessian_csr(u_star, make_spsd=False, filter_dbc=True)
        lv = spla.spsolve(H.tocsc(), rhs)
        g_uhat = np.where(free, self.mass_csr() @ lv, 0.0)
        B, V = self.rest_data()
        F = self.deformation_gradients(u_star)
        G = _edge_matrix(lv.reshape(-1, 3), self.cells) @ B
        dl, dm = (lin_dP_dlame if self.model == "linear" else snh_dP_dlame)(F)
        sc = -self.dt * self.dt * V
        return lv, g_uhat, sc * np.sum(dl * G, axis=(1, 2)), sc * np.sum(dm * G, axis=(1, 2))

    def rest_hessian_diag(self):
        d = np.zeros(self.n)
        lib().orc_ts_rest_hessian_diag(self._h, _dp(d))
        return d
