"""Batched stepping: many independent tetrahedral meshes of uneven size advanced by one set of GPU launches.

``TetFiniteElementBatch_Hip`` holds a list of meshes, each with the state and Newton step of one
``TetFiniteElementSolver_Hip`` (reference: one ``TetFiniteElementSolver_*`` object per mesh,
bindings/src/fem_host.cpp).  Per-mesh methods take a mesh index; vertex outputs are (3, N) like the reference's.
Meshes carry per-tet Lamé parameters; time step and density are shared by the batch.
"""
from __future__ import annotations

import ctypes
from typing import Sequence

import numpy as np

from . import _lib

_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)


def _dp(a):
    return a.ctypes.data_as(_D)


def _ip(a):
    return a.ctypes.data_as(_I)


class TetFiniteElementBatch_Hip:
    """Backward-Euler FEM time step of a batch of meshes on one MI355X (FP64, Newton + Jacobi-PCG)."""

    def __init__(self, vertices: Sequence[np.ndarray], elements: Sequence[np.ndarray], lam: Sequence[np.ndarray],
                 mu: Sequence[np.ndarray], time_step: float = 1e-2, density: float = 1e2,
                 material: str = "stable_neohookean"):
        if material not in ("stable_neohookean", "linear"):
            raise ValueError(f"Unknown material: {material}")
        if not (len(vertices) == len(elements) == len(lam) == len(mu)) or len(vertices) == 0:
            raise ValueError("vertices, elements, lam and mu must be non-empty lists of equal length")
        Xs = [np.ascontiguousarray(x, dtype=np.float64) for x in vertices]
        Cs = [np.ascontiguousarray(c, dtype=np.int32) for c in elements]
        for X, C in zip(Xs, Cs):
            if X.ndim != 2 or X.shape[1] != 3 or C.ndim != 2 or C.shape[1] != 4:
                raise ValueError("each mesh needs vertices (N, 3) and elements (M, 4)")
        self._nv = np.array([x.shape[0] for x in Xs], dtype=np.int32)
        self._nt = np.array([c.shape[0] for c in Cs], dtype=np.int32)
        lam = np.ascontiguousarray(np.concatenate([np.asarray(v, np.float64).reshape(-1) for v in lam]))
        mu = np.ascontiguousarray(np.concatenate([np.asarray(v, np.float64).reshape(-1) for v in mu]))
        if lam.size != self._nt.sum() or mu.size != self._nt.sum():
            raise ValueError("lam and mu need one entry per tet of every mesh")
        X = np.ascontiguousarray(np.concatenate(Xs))
        C = np.ascontiguousarray(np.concatenate(Cs))
        self._off = np.concatenate([[0], np.cumsum(3 * self._nv.astype(np.int64))])
        h = ctypes.c_void_p()
        _lib.call("ssim_batch_create", ctypes.byref(h), len(Xs), _ip(self._nv), _ip(self._nt), _dp(X), _ip(C),
                  _dp(lam), _dp(mu), float(density), float(time_step))
        self._h = h
        _lib.call("ssim_batch_set_material_model", self._h, {"stable_neohookean": 0, "linear": 1}[material])
        self.material = material

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and _lib is not None and _lib._lib is not None:
            _lib._lib.ssim_batch_destroy(h)
            self._h = None

    @property
    def handle(self):
        return self._h

    @property
    def num_meshes(self) -> int:
        return int(self._nv.size)

    def num_vertices(self, mesh: int) -> int:
        return int(self._nv[mesh])

    # ---- lifecycle ----
    def reset(self):
        _lib.call("ssim_batch_reset", self._h)

    def prepare_step(self):
        _lib.call("ssim_batch_prepare_step", self._h)

    def compute_step(self):
        _lib.call("ssim_batch_compute_step", self._h)

    def step_next(self):
        _lib.call("ssim_batch_step_next", self._h)

    def step(self):
        _lib.call("ssim_batch_step", self._h)

    # ---- boundary conditions and parameters ----
    def mark_dirichlet_batched(self, mesh: int, verts: np.ndarray, targ_deform: np.ndarray):
        v = np.ascontiguousarray(verts, dtype=np.int32).reshape(-1)
        t = np.asarray(targ_deform, dtype=np.float64)
        if t.shape != (3, v.size):
            raise ValueError(f"targ_deform must have shape (3, {v.size}), got {t.shape}")
        t = np.ascontiguousarray(t.T)
        _lib.call("ssim_batch_mark_dirichlet", self._h, int(mesh), _ip(v), v.size, _dp(t))

    def mark_general_batched(self, mesh: int, verts: np.ndarray):
        v = np.ascontiguousarray(verts, dtype=np.int32).reshape(-1)
        _lib.call("ssim_batch_mark_general", self._h, int(mesh), _ip(v), v.size)

    def set_rtol(self, rtol: float):
        _lib.call("ssim_batch_set_threshold", self._h, float(rtol))

    def solver_params(self) -> dict:
        p = _lib.SolverParams()
        _lib.call("ssim_batch_get_solver_params", self._h, ctypes.byref(p))
        return {k: getattr(p, k) for k, _ in p._fields_}

    def set_solver_params(self, **kw):
        p = _lib.SolverParams()
        _lib.call("ssim_batch_get_solver_params", self._h, ctypes.byref(p))
        for k, v in kw.items():
            if not hasattr(p, k):
                raise ValueError(f"unknown solver parameter {k}")
            setattr(p, k, v)
        _lib.call("ssim_batch_set_solver_params", self._h, ctypes.byref(p))

    # ---- state ----
    def get_array(self, name: str, mesh: int = -1) -> np.ndarray:
        n = int(self._off[-1]) if mesh < 0 else 3 * int(self._nv[mesh])
        out = np.empty(n)
        _lib.call("ssim_batch_get_array", self._h, _lib.ARRAYS[name], int(mesh), _dp(out))
        return out

    def set_array(self, name: str, values: np.ndarray, mesh: int = -1):
        a = np.ascontiguousarray(values, dtype=np.float64).reshape(-1)
        n = int(self._off[-1]) if mesh < 0 else 3 * int(self._nv[mesh])
        if a.size != n:
            raise ValueError(f"{name} must hold {n} values")
        _lib.call("ssim_batch_set_array", self._h, _lib.ARRAYS[name], int(mesh), _dp(a))

    def deformation(self, mesh: int) -> np.ndarray:
        return np.ascontiguousarray(self.get_array("deformation", mesh).reshape(-1, 3).T)

    def forces(self, mesh: int) -> np.ndarray:
        return np.ascontiguousarray(self.get_array("forces", mesh).reshape(-1, 3).T)

    def last_stats(self, mesh: int = -1) -> dict:
        st = _lib.StepStats()
        _lib.call("ssim_batch_get_step_stats", self._h, int(mesh), ctypes.byref(st))
        return st.as_dict()

    def total_mass(self, mesh: int) -> float:
        m = ctypes.c_double()
        _lib.call("ssim_batch_total_mass", self._h, int(mesh), ctypes.byref(m))
        return m.value

    # ---- launch tuning and instrumentation ----
    def set_tuning(self, **kw):
        for k, v in kw.items():
            if k not in _lib.BATCH_TUNING:
                raise ValueError(f"unknown tuning key {k}")
            _lib.call("ssim_batch_set_tuning", self._h, _lib.BATCH_TUNING[k], int(v))

    def tuning(self) -> dict:
        out = {}
        for k, key in _lib.BATCH_TUNING.items():
            v = ctypes.c_int()
            _lib.call("ssim_batch_get_tuning", self._h, key, ctypes.byref(v))
            out[k] = v.value
        return out

    def set_kernel_timing(self, enabled: bool = True):
        _lib.call("ssim_batch_set_kernel_timing", self._h, int(bool(enabled)))

    def kernel_timing(self) -> dict:
        out = {}
        for name, k in _lib.KERNELS.items():
            ms = ctypes.c_double()
            cnt = ctypes.c_longlong()
            _lib.call("ssim_batch_get_kernel_timing", self._h, k, ctypes.byref(ms), ctypes.byref(cnt))
            out[name] = (ms.value, cnt.value)
        return out

    def layout_info(self) -> dict:
        info = _lib.LayoutInfo()
        _lib.call("ssim_batch_get_layout_info", self._h, ctypes.byref(info))
        return {k: getattr(info, k) for k, _ in info._fields_}

    @classmethod
    def from_cases(cls, cases, rtol=None, material="stable_neohookean"):
        """Batch of seeded ``ssim_amd.mesh.BarCase`` problems (all with the same dt and density)."""
        dt, rho = cases[0].dt, cases[0].rho
        if any(c.dt != dt or c.rho != rho for c in cases):
            raise ValueError("a batch shares one time step and one density")
        b = cls([c.X for c in cases], [c.cells for c in cases], [c.lam for c in cases], [c.mu for c in cases],
                dt, rho, material)
        for m, c in enumerate(cases):
            clamped = np.nonzero(c.dof_type.reshape(-1, 3).any(axis=1))[0].astype(np.int32)
            if clamped.size:
                b.mark_dirichlet_batched(m, clamped, c.dbc.reshape(-1, 3)[clamped].T)
        b.set_array("ext_accel", np.concatenate([c.accel for c in cases]))
        b.set_array("deformation", np.concatenate([c.u0 for c in cases]))
        b.set_array("velocity", np.concatenate([c.v0 for c in cases]))
        if rtol is not None:
            b.set_rtol(rtol)
        b.reset()
        return b


def size_groups(nts, groups: int, key=None):
    """Splits meshes with tet counts ``nts`` into at most ``groups`` classes: meshes sorted by ``key`` (default: the
    tet count; ties by index) are cut into contiguous runs of about equal total tet count.  Returns the index lists,
    each ascending."""
    groups = max(1, min(int(groups), len(nts)))
    key = nts if key is None else key
    if len(key) != len(nts):
        raise ValueError("one key per mesh")
    order = sorted(range(len(nts)), key=lambda m: (key[m], m))
    total = float(sum(nts))
    members, acc, g = [[] for _ in range(groups)], 0.0, 0
    for m in order:
        if g < groups - 1 and members[g] and acc >= total * (g + 1) / groups:
            g += 1
        members[g].append(m)
        acc += nts[m]
    return [sorted(ms) for ms in members if ms]


class BatchGroups:
    """A batch split into ``groups`` sub-batches (no reference counterpart: an extension of batching), stepped
    either concurrently, each on its own HIP stream from its own host thread, or one after another.

    Lock-step batching runs every PCG iteration as launches over all meshes still iterating, so the last
    iterations of the slowest meshes are small launches whose dispatch latency the GPU cannot hide, and the
    full-batch launches stream a working set (matrix and vectors) far larger than the 256 MB Infinity Cache.  Two
    ways to group, by the sort key ``key`` (default: the tet count; cuts of about equal total tet count):

    * concurrent size classes (``concurrent=True``): one group's long tail overlaps the others' work;
    * sequential groups keyed by each mesh's PCG iterations of an earlier step (``concurrent=False``): meshes that
      iterate alike share a group, so the group's grid stays full until it is done, and a group whose working set
      fits the Infinity Cache runs its SpMV from it.  With three groups of ~1.6M tets (~240 MB each) out of the
      256-bar benchmark batch, a step takes ~218 ms against ~252 ms lock-step.

    Each mesh still follows the single-mesh Newton step exactly, and a mesh's arithmetic does not depend on the
    group it is in.  Mesh indices are those of the case list.
    """

    def __init__(self, cases, groups: int = 2, rtol=None, material: str = "stable_neohookean", key=None,
                 concurrent: bool = True):
        import torch

        if groups < 1:
            raise ValueError("groups must be positive")
        self.concurrent = bool(concurrent)
        self.members = size_groups([c.nt for c in cases], groups, key)
        self.batches = [TetFiniteElementBatch_Hip.from_cases([cases[m] for m in ms], rtol=rtol, material=material)
                        for ms in self.members]
        self.where = {}
        for gi, ms in enumerate(self.members):
            for li, m in enumerate(ms):
                self.where[m] = (gi, li)
        if self.concurrent:
            self._streams = [torch.cuda.Stream() for _ in self.batches]  # default priority
            for b, st in zip(self.batches, self._streams):
                _lib.call("ssim_batch_set_stream", b.handle, ctypes.c_void_p(st.cuda_stream))
        self._pool = None
        self._ncases = len(cases)

    @property
    def num_meshes(self) -> int:
        return self._ncases

    def _each(self, fn):
        if len(self.batches) == 1 or not self.concurrent:
            return [fn(b) for b in self.batches]
        if self._pool is None:
            from concurrent.futures import ThreadPoolExecutor

            self._pool = ThreadPoolExecutor(max_workers=len(self.batches))
        return [f.result() for f in [self._pool.submit(fn, b) for b in self.batches]]

    def step(self):
        self._each(lambda b: b.step())

    def set_solver_params(self, **kw):
        for b in self.batches:
            b.set_solver_params(**kw)

    def set_tuning(self, **kw):
        for b in self.batches:
            b.set_tuning(**kw)

    def get_array(self, name: str, mesh: int) -> np.ndarray:
        g, m = self.where[mesh]
        return self.batches[g].get_array(name, m)

    def last_stats(self, mesh: int) -> dict:
        g, m = self.where[mesh]
        return self.batches[g].last_stats(m)

    def totals(self) -> dict:
        """Lock-step Newton and PCG iterations of the last step, per group."""
        return [{k: b.last_stats()[k] for k in ("newton_iterations", "pcg_iterations")} for b in self.batches]

    def __del__(self):
        if getattr(self, "_pool", None) is not None:
            self._pool.shutdown(wait=True)
