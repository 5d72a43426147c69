"""Python front end mirroring ``pyssim.fem`` (reference pyssim/fem/__init__.py, bindings/src/fem_host.cpp).

``TetFiniteElementSolver_Hip`` has the constructor and methods of the reference's
``TetFiniteElementSolver_Host`` / ``_Cuda`` classes with the same argument meaning, layouts ((3, N) vertex
outputs, (3, count) Dirichlet targets, scipy CSR matrices) and error behaviour (``ValueError`` where the
reference throws ``std::invalid_argument``).  Every numeric operation runs in libssim_hip.so on the GPU.
"""
from __future__ import annotations

import ctypes
import time
from typing import Literal, Tuple

import numpy as np

from . import _lib
from . import mesh as _mesh

_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)


def _dp(a):
    return a.ctypes.data_as(_D)


def _ip(a):
    return a.ctypes.data_as(_I)


def unit_box(nx: int, ny: int, nz: int) -> Tuple[np.ndarray, np.ndarray]:
    """Reference ``unit_box`` (5-tet split, vertex counts per axis): (verts (3,N), cells (4,M))."""
    return _mesh.unit_box(nx, ny, nz)


class TetFiniteElementSolver_Hip:
    """Implicit (backward Euler) tetrahedral FEM time step on one MI355X (FP64).

    method: "newton" (Newton + Jacobi-PCG + backtracking Armijo), "lbfgs" (L-BFGS, m = 8, H0 = diagonal
    of the rest Hessian), "ncg" (nonlinear CG, Dai-Kou β, the L-BFGS line-search constants) or
    "forward_euler" (explicit: next = û + ½dt²·forces(û), finite_elements/time_step.hpp:717-736) or "ncg_ext_ai"
    (NCG preconditioned by z = L·Lᵀ·g with L set by ``set_matrix``, the reference's tet_ncg_cuda_ext_ai,
    bindings/src/fem_cuda.cu:172-244; L is the identity after ``reset``) or "lbfgs_pd" (L-BFGS preconditioned by
    solves with the projective-dynamics matrix M + dt²κK built at ``reset``, the reference's tet_lbfgs_pd,
    bindings/src/fem_host.cpp:223, stepper/projective_dynamics.hpp:9-88) or "lbfgs_h0" (L-BFGS preconditioned by
    solves with the SPD Hessian at the start of each step, time_step_solver_h0, projective_dynamics.hpp:91-148).
    material: "stable_neohookean" (elast/stable_neohookean.hpp, the reference bindings' model) or "linear"
    (elast/linear.hpp).
    """

    METHODS = {"newton": 0, "lbfgs": 1, "ncg": 2, "forward_euler": 3, "ncg_ext_ai": 4, "lbfgs_pd": 5, "lbfgs_h0": 6}
    MATERIALS = {"stable_neohookean": 0, "linear": 1}

    def __init__(self, vertices: np.ndarray, elements: np.ndarray, time_step: float = 1e-2,
                 young_modulus: float = 1e6, poisson_ratio: float = 0.33, density: float = 1e2,
                 method: Literal["newton", "lbfgs", "ncg", "forward_euler", "ncg_ext_ai", "lbfgs_pd",
                                 "lbfgs_h0"] = "newton",
                 material: Literal["stable_neohookean", "linear"] = "stable_neohookean"):
        if method not in self.METHODS:
            raise ValueError(f"Unknown method: {method}")
        if material not in self.MATERIALS:
            raise ValueError(f"Unknown material: {material}")
        X = np.ascontiguousarray(vertices, dtype=np.float64)
        C = np.ascontiguousarray(elements, dtype=np.int32)
        if X.ndim != 2 or X.shape[1] != 3:
            raise ValueError(f"vertices must have shape (N, 3), got {X.shape}")
        if C.ndim != 2 or C.shape[1] != 4:
            raise ValueError(f"elements must have shape (M, 4), got {C.shape}")
        self._nv, self._nt = X.shape[0], C.shape[0]
        self._X = X
        self._cells = C
        h = ctypes.c_void_p()
        _lib.call("ssim_create", ctypes.byref(h), _dp(X), self._nv, _ip(C), self._nt, float(time_step),
                  float(young_modulus), float(poisson_ratio), float(density))
        self._h = h
        self._has_reset = False
        _lib.call("ssim_set_stepper", self._h, self.METHODS[method])
        self.method = method
        _lib.call("ssim_set_material_model", self._h, self.MATERIALS[material])
        self.material = material

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and _lib is not None and _lib._lib is not None:
            _lib._lib.ssim_destroy(h)
            self._h = None

    # ---- lifecycle (fem_host.cpp:83-111) ----
    def reset(self):
        _lib.call("ssim_reset", self._h)
        self._has_reset = True

    def prepare_step(self):
        if not self._has_reset:
            self.reset()
            import sys

            print("Warning: prepare_step() called before reset(), bad logic.", file=sys.stderr)
        _lib.call("ssim_prepare_step", self._h)

    def compute_step(self):
        r = ctypes.c_double()
        _lib.call("ssim_compute_step", self._h, ctypes.byref(r))
        return r.value

    def step_next(self):
        _lib.call("ssim_step_next", self._h)

    def step(self) -> float:
        """prepare + compute + step_next; returns the wall time in milliseconds (fem_host.cpp:102-111)."""
        t0 = time.perf_counter()
        self.prepare_step()
        self.compute_step()
        self.step_next()
        return (time.perf_counter() - t0) * 1e3

    # ---- boundary conditions and loads (fem_host.cpp:113-152) ----
    def mark_dirichlet(self, node_idx: int, targ_deform: np.ndarray):
        t = np.ascontiguousarray(targ_deform, dtype=np.float64).reshape(1, 3)
        v = np.array([node_idx], dtype=np.int32)
        _lib.call("ssim_mark_dirichlet", self._h, _ip(v), 1, _dp(t))

    def mark_dirichlet_batched(self, verts: np.ndarray, targ_deform: np.ndarray):
        v = np.ascontiguousarray(verts, dtype=np.int32).reshape(-1)
        t = np.asarray(targ_deform, dtype=np.float64)
        if t.shape != (3, v.size):
            raise ValueError(f"targ_deform must have shape (3, {v.size}), got {t.shape}")
        t = np.ascontiguousarray(t.T)
        _lib.call("ssim_mark_dirichlet", self._h, _ip(v), v.size, _dp(t))

    def mark_general_batched(self, verts: np.ndarray):
        v = np.ascontiguousarray(verts, dtype=np.int32).reshape(-1)
        _lib.call("ssim_mark_general", self._h, _ip(v), v.size)

    def set_matrix(self, values: np.ndarray):
        """Preconditioner factor L of the "ncg_ext_ai" method: the values of a CSR matrix with the system matrix's
        pattern (``hessian()`` order, pyssim/fem/__init__.py:151-157; fem_cuda.cu:204)."""
        v = np.ascontiguousarray(values, dtype=np.float64).reshape(-1)
        _lib.call("ssim_set_precond_matrix", self._h, _dp(v), ctypes.c_longlong(v.size))

    def set_rtol(self, rtol: float):
        _lib.call("ssim_set_threshold", self._h, float(rtol))

    def set_damping(self, damping: float):
        """Velocity damping of step_next, v ← (1 − damping)·v (basic_time_step::set_damping,
        finite_elements/time_step.hpp:337-341, 541-542; not bound by the reference's Python module).  ValueError
        outside [0, 1), as the reference's std::invalid_argument."""
        _lib.call("ssim_set_damping", self._h, float(damping))

    def add_gravity(self, gravity: np.ndarray):
        g = np.asarray(gravity, dtype=np.float64).reshape(3)
        for d in range(3):
            _lib.call("ssim_add_ext_force_dof", self._h, d, float(g[d]))

    # ---- state (fem_host.cpp:154-171) ----
    def vertices(self) -> np.ndarray:
        return np.ascontiguousarray(self._X.T)

    def cells(self) -> np.ndarray:
        return np.ascontiguousarray(self._cells.T)

    def _get(self, name) -> np.ndarray:
        out = np.empty(3 * self._nv)
        _lib.call("ssim_get_array", self._h, _lib.ARRAYS[name], _dp(out))
        return out

    def _set(self, name, values):
        a = np.ascontiguousarray(values, dtype=np.float64).reshape(-1)
        if a.size != 3 * self._nv:
            raise ValueError(f"{name} must hold {3 * self._nv} values")
        _lib.call("ssim_set_array", self._h, _lib.ARRAYS[name], _dp(a))

    def deformation(self) -> np.ndarray:
        return np.ascontiguousarray(self._get("deformation").reshape(-1, 3).T)

    def forces(self) -> np.ndarray:
        return np.ascontiguousarray(self._get("forces").reshape(-1, 3).T)

    def save_frame(self, path) -> np.ndarray:
        """Writes the positions X + u as an (N, 3) float64 .npy file and returns them: the deform_<step>.npy
        frames of tests/fem_cu/main.cu:86-92, which tests/fem_time_step/visualize.py:9-12 loads."""
        x = self._X + self._get("deformation").reshape(-1, 3)
        np.save(path, x)
        return x

    def update_energy_and_gradients(self) -> float:
        e = ctypes.c_double()
        _lib.call("ssim_update_energy_and_gradients", self._h, ctypes.byref(e))
        return e.value

    def update_hessian(self, make_spsd: bool = True):
        _lib.call("ssim_update_hessian", self._h, int(bool(make_spsd)), 1)

    def update_hessian_unfiltered(self, make_spsd: bool = True):
        _lib.call("ssim_update_hessian", self._h, int(bool(make_spsd)), 0)

    def _csr(self, which):
        import scipy.sparse as sp

        nnz = ctypes.c_longlong()
        _lib.call("ssim_csr_nnz", self._h, ctypes.byref(nnz))
        n = 3 * self._nv
        rowptr = np.empty(n + 1, dtype=np.int32)
        colidx = np.empty(nnz.value, dtype=np.int32)
        vals = np.empty(nnz.value)
        _lib.call(which, self._h, _ip(rowptr), _ip(colidx), _dp(vals))
        return sp.csr_matrix((vals, colidx, rowptr), shape=(n, n))

    def mass_matrix(self):
        return self._csr("ssim_export_mass_csr")

    def hessian(self):
        return self._csr("ssim_export_hessian_csr")

    def mass_matrix_filtered(self):
        """mass_matrix_filtered() (time_step.hpp:310-311): M with identity Dirichlet rows and columns."""
        return self._csr("ssim_export_mass_filtered_csr")

    def rest_shape(self):
        """(dminv [nt,3,3], rest_vol [nt]): Dm⁻¹ and |det Dm|/6 per tet (time_step.hpp:296-301)."""
        dminv = np.empty((self._nt, 3, 3))
        vol = np.empty(self._nt)
        _lib.call("ssim_get_rest_shape", self._h, _dp(dminv), _dp(vol))
        return dminv, vol

    def pd_matrix(self):
        """The projective-dynamics matrix of the "lbfgs_pd" method (scipy CSR, the Hessian's pattern)."""
        return self._csr("ssim_export_pd_csr")

    def hessian_nonzeros(self) -> np.ndarray:
        return self.hessian().data.copy()

    # ---- extensions beyond the reference Python API ----
    @property
    def handle(self):
        return self._h

    def set_array(self, name: str, values: np.ndarray):
        self._set(name, values)

    def get_array(self, name: str) -> np.ndarray:
        return self._get(name)

    def set_material_per_tet(self, lam, mu):
        lam = np.ascontiguousarray(lam, dtype=np.float64).reshape(-1)
        mu = np.ascontiguousarray(mu, dtype=np.float64).reshape(-1)
        if lam.size != self._nt or mu.size != self._nt:
            raise ValueError("per-tet material arrays must have one entry per tet")
        _lib.call("ssim_set_material_per_tet", self._h, _dp(lam), _dp(mu))

    def solver_params(self) -> dict:
        p = _lib.SolverParams()
        _lib.call("ssim_get_solver_params", self._h, ctypes.byref(p))
        return {k: getattr(p, k) for k, _ in p._fields_}

    def set_solver_params(self, **kw):
        p = _lib.SolverParams()
        _lib.call("ssim_get_solver_params", self._h, ctypes.byref(p))
        for k, v in kw.items():
            if not hasattr(p, k):
                raise ValueError(f"unknown solver parameter {k}")
            setattr(p, k, v)
        _lib.call("ssim_set_solver_params", self._h, ctypes.byref(p))

    def last_stats(self) -> dict:
        st = _lib.StepStats()
        _lib.call("ssim_get_step_stats", self._h, ctypes.byref(st))
        return st.as_dict()

    def total_mass(self) -> float:
        m = ctypes.c_double()
        _lib.call("ssim_total_mass", self._h, ctypes.byref(m))
        return m.value

    def grad_threshold_abs(self) -> float:
        m = ctypes.c_double()
        _lib.call("ssim_grad_threshold_abs", self._h, ctypes.byref(m))
        return m.value

    def eval_elements(self, u: np.ndarray, make_spsd: bool = True, want_H: bool = True, want_K: bool = False):
        u = np.ascontiguousarray(u, dtype=np.float64).reshape(-1)
        nt = self._nt
        psi, P = np.empty(nt), np.empty((nt, 3, 3))
        H = np.empty((nt, 9, 9)) if want_H else None
        K = np.empty((nt, 12, 12)) if want_K else None
        null = ctypes.cast(None, _D)
        _lib.call("ssim_eval_elements", self._h, _dp(u), int(bool(make_spsd)), _dp(psi), _dp(P),
                  _dp(H) if H is not None else null, _dp(K) if K is not None else null)
        return dict(psi=psi, P=P, H=H, K=K)

    def pcg_solve(self, b: np.ndarray, rtol: float = 1e-8, maxit: int = 10000):
        b = np.ascontiguousarray(b, dtype=np.float64).reshape(-1)
        x = np.empty_like(b)
        it = ctypes.c_int()
        _lib.call("ssim_pcg_solve", self._h, _dp(b), _dp(x), float(rtol), int(maxit), ctypes.byref(it))
        return x, it.value

    def spmv(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
        y = np.empty_like(x)
        _lib.call("ssim_spmv", self._h, _dp(x), _dp(y))
        return y

    def set_kernel_timing(self, enabled: bool = True):
        _lib.call("ssim_set_kernel_timing", self._h, int(bool(enabled)))

    def kernel_timing(self) -> dict:
        out = {}
        for name, k in _lib.KERNELS.items():
            ms = ctypes.c_double()
            cnt = ctypes.c_longlong()
            _lib.call("ssim_get_kernel_timing", self._h, k, ctypes.byref(ms), ctypes.byref(cnt))
            out[name] = (ms.value, cnt.value)
        return out

    def set_tuning(self, **kw):
        for k, v in kw.items():
            if k not in _lib.TUNING:
                raise ValueError(f"unknown tuning key {k}")
            _lib.call("ssim_set_tuning", self._h, _lib.TUNING[k], int(v))

    def tuning(self) -> dict:
        out = {}
        for k, key in _lib.TUNING.items():
            v = ctypes.c_int()
            _lib.call("ssim_get_tuning", self._h, key, ctypes.byref(v))
            out[k] = v.value
        return out

    def layout_info(self) -> dict:
        info = _lib.LayoutInfo()
        _lib.call("ssim_get_layout_info", self._h, ctypes.byref(info))
        return {k: getattr(info, k) for k, _ in info._fields_}

    @classmethod
    def from_case(cls, case, method="newton", rtol=None, material="stable_neohookean"):
        """Solver for a seeded ``ssim_amd.mesh.BarCase`` (per-tet material, Dirichlet set, loads, initial state)."""
        lam0, mu0 = float(case.lam[0]), float(case.mu[0])
        # any valid (E, ν) for the constructor; the per-tet arrays below define the material
        nu = lam0 / (2 * (lam0 + mu0))
        E = mu0 * 2 * (1 + nu)
        s = cls(case.X, case.cells, case.dt, E, nu, case.rho, method, material)
        s.set_material_per_tet(case.lam, case.mu)  # exact per-tet (λ, μ), uniform or not
        clamped = np.nonzero(case.dof_type.reshape(-1, 3).any(axis=1))[0].astype(np.int32)
        if clamped.size:
            s.mark_dirichlet_batched(clamped, case.dbc.reshape(-1, 3)[clamped].T)
        s.set_array("ext_accel", case.accel)
        s.set_array("deformation", case.u0)
        s.set_array("velocity", case.v0)
        if rtol is not None:
            s.set_rtol(rtol)
        s.reset()
        return s


# The reference's class names (pyssim/fem/__init__.py:11 and :97) map onto the one GPU class, so code written
# against either runs unchanged; the "Host" name keeps its interface, and its arithmetic runs on the GPU.
TetFiniteElementSolver_Cuda = TetFiniteElementSolver_Hip
TetFiniteElementSolver_Host = TetFiniteElementSolver_Hip
