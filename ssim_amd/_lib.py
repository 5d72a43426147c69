"""ctypes binding of libssim_hip.so (the C ABI of include/ssim_hip.h).

The library is built in-tree by ``ssim_amd.build`` (hipcc, gfx950).  Loading fails loudly when it is
missing: there is no CPU fallback for any entry point.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libssim_hip.so")

_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)
_U8 = ctypes.POINTER(ctypes.c_ubyte)
_V = ctypes.c_void_p
_S = ctypes.c_int  # ssim_status_t

ARRAYS = {
    "deformation": 0, "prev_deform": 1, "inertia": 2, "next_deform": 3, "velocity": 4, "ext_accel": 5,
    "forces": 6, "dbc_values": 7,
}


class SolverParams(ctypes.Structure):
    _fields_ = [
        ("max_newton", ctypes.c_int), ("pcg_rtol", ctypes.c_double), ("pcg_maxit", ctypes.c_int),
        ("ls_c", ctypes.c_double), ("ls_beta", ctypes.c_double), ("ls_max", ctypes.c_int),
        ("lbfgs_m", ctypes.c_int), ("lbfgs_maxit", ctypes.c_int), ("lbfgs_c", ctypes.c_double),
        ("lbfgs_ls_max", ctypes.c_int), ("fp32_elements", ctypes.c_int), ("newton_polish", ctypes.c_int),
        ("pcg_precond", ctypes.c_int),
    ]


class StepStats(ctypes.Structure):
    _fields_ = [
        ("converged", ctypes.c_int), ("newton_iterations", ctypes.c_int), ("pcg_iterations", ctypes.c_int),
        ("energy_evals", ctypes.c_int), ("hessian_evals", ctypes.c_int), ("ls_failures", ctypes.c_int),
        ("grad_norm", ctypes.c_double), ("energy", ctypes.c_double), ("last_change", ctypes.c_double),
        ("ls_rejected", ctypes.c_int), ("reserved", ctypes.c_int),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


class LayoutInfo(ctypes.Structure):
    _fields_ = [("nv", ctypes.c_int), ("nt", ctypes.c_int), ("nonzero_blocks", ctypes.c_longlong),
                ("padded_slots", ctypes.c_longlong), ("lower_slots", ctypes.c_longlong),
                ("upper_blocks", ctypes.c_int), ("slices", ctypes.c_int), ("fixed_layout", ctypes.c_int),
                ("reserved", ctypes.c_int)]


KERNELS = {"elem_grad": 0, "grad_reduce": 1, "elem_hess": 2, "hess_reduce": 3, "pcg_spmv": 4, "pcg_update": 5,
           "pcg_p": 6}

# name -> argtypes (all return ssim_status_t except ssim_last_error)
SIGNATURES = {
    "ssim_create": [ctypes.POINTER(_V), _D, ctypes.c_int, _I, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                    ctypes.c_double, ctypes.c_double],
    "ssim_destroy": [_V],
    "ssim_set_stream": [_V, _V],
    "ssim_set_time_step": [_V, ctypes.c_double],
    "ssim_set_youngs": [_V, ctypes.c_double],
    "ssim_set_poisson": [_V, ctypes.c_double],
    "ssim_set_density": [_V, ctypes.c_double],
    "ssim_set_damping": [_V, ctypes.c_double],
    "ssim_set_threshold": [_V, ctypes.c_double],
    "ssim_set_material_per_tet": [_V, _D, _D],
    "ssim_set_solver_params": [_V, ctypes.POINTER(SolverParams)],
    "ssim_get_solver_params": [_V, ctypes.POINTER(SolverParams)],
    "ssim_set_stepper": [_V, ctypes.c_int],
    "ssim_set_material_model": [_V, ctypes.c_int],
    "ssim_material_eval": [ctypes.c_int, ctypes.c_int, ctypes.c_longlong, _V, _V, _V, ctypes.c_int, _V, _V, _V, _V],
    "ssim_tri_eval": [ctypes.c_int, ctypes.c_int, ctypes.c_int, _V, _V, _V, _V, _V, ctypes.c_double, ctypes.c_int]
    + [_V] * 9,
    "ssim_set_precond_matrix": [_V, _D, ctypes.c_longlong],
    "ssim_get_material_model": [_V, ctypes.POINTER(ctypes.c_int)],
    "ssim_mark_dirichlet": [_V, _I, ctypes.c_int, _D],
    "ssim_mark_general": [_V, _I, ctypes.c_int],
    "ssim_add_ext_force_dof": [_V, ctypes.c_int, ctypes.c_double],
    "ssim_reset": [_V],
    "ssim_prepare_step": [_V],
    "ssim_compute_step": [_V, _D],
    "ssim_step_next": [_V],
    "ssim_step": [_V, _D],
    "ssim_restore": [_V],
    "ssim_get_step_stats": [_V, ctypes.POINTER(StepStats)],
    "ssim_update_energy_and_gradients": [_V, _D],
    "ssim_update_hessian": [_V, ctypes.c_int, ctypes.c_int],
    "ssim_check_convergence": [_V, _I],
    "ssim_get_array": [_V, ctypes.c_int, _D],
    "ssim_set_array": [_V, ctypes.c_int, _D],
    "ssim_device_array": [_V, ctypes.c_int, ctypes.POINTER(_D)],
    "ssim_get_array_device": [_V, ctypes.c_int, _V],
    "ssim_set_array_device": [_V, ctypes.c_int, _V],
    "ssim_adjoint_step": [_V, _V, _V, _V, _V, _V, _I],
    "ssim_get_dof_type": [_V, _U8],
    "ssim_device_dof_type": [_V, ctypes.POINTER(ctypes.POINTER(ctypes.c_ubyte))],
    "ssim_num_vertices": [_V, _I],
    "ssim_num_cells": [_V, _I],
    "ssim_total_mass": [_V, _D],
    "ssim_grad_threshold_abs": [_V, _D],
    "ssim_csr_nnz": [_V, ctypes.POINTER(ctypes.c_longlong)],
    "ssim_export_hessian_csr": [_V, _I, _I, _D],
    "ssim_export_pd_csr": [_V, _I, _I, _D],
    "ssim_export_mass_csr": [_V, _I, _I, _D],
    "ssim_export_mass_filtered_csr": [_V, _I, _I, _D],
    "ssim_get_rest_shape": [_V, _D, _D],
    "ssim_get_physical_params": [_V, _D, _D, _D, _D, _D, _D],
    "ssim_device_mesh": [_V, ctypes.POINTER(_D), ctypes.POINTER(_I)],
    "ssim_eval_elements": [_V, _D, ctypes.c_int, _D, _D, _D, _D],
    "ssim_pcg_solve": [_V, _D, _D, ctypes.c_double, ctypes.c_int, _I],
    "ssim_pcg_solve_noalloc": [_V, _V, _V, ctypes.c_double, ctypes.c_int, _I],
    "ssim_spmv": [_V, _D, _D],
    "ssim_set_kernel_timing": [_V, ctypes.c_int],
    "ssim_get_kernel_timing": [_V, ctypes.c_int, _D, ctypes.POINTER(ctypes.c_longlong)],
    "ssim_get_layout_info": [_V, ctypes.POINTER(LayoutInfo)],
    "ssim_set_tuning": [_V, ctypes.c_int, ctypes.c_int],
    "ssim_get_tuning": [_V, ctypes.c_int, _I],
    # batches
    "ssim_batch_create": [ctypes.POINTER(_V), ctypes.c_int, _I, _I, _D, _I, _D, _D, ctypes.c_double,
                          ctypes.c_double],
    "ssim_batch_destroy": [_V],
    "ssim_batch_set_stream": [_V, _V],
    "ssim_batch_num_meshes": [_V, _I],
    "ssim_batch_set_threshold": [_V, ctypes.c_double],
    "ssim_batch_set_solver_params": [_V, ctypes.POINTER(SolverParams)],
    "ssim_batch_get_solver_params": [_V, ctypes.POINTER(SolverParams)],
    "ssim_batch_mark_dirichlet": [_V, ctypes.c_int, _I, ctypes.c_int, _D],
    "ssim_batch_mark_general": [_V, ctypes.c_int, _I, ctypes.c_int],
    "ssim_batch_reset": [_V],
    "ssim_batch_prepare_step": [_V],
    "ssim_batch_compute_step": [_V],
    "ssim_batch_step_next": [_V],
    "ssim_batch_step": [_V],
    "ssim_batch_get_step_stats": [_V, ctypes.c_int, ctypes.POINTER(StepStats)],
    "ssim_batch_get_array": [_V, ctypes.c_int, ctypes.c_int, _D],
    "ssim_batch_set_array": [_V, ctypes.c_int, ctypes.c_int, _D],
    "ssim_batch_total_mass": [_V, ctypes.c_int, _D],
    "ssim_batch_get_array_device": [_V, ctypes.c_int, _V],
    "ssim_batch_set_array_device": [_V, ctypes.c_int, _V],
    "ssim_batch_set_material": [_V, _D, _D],
    "ssim_batch_set_material_model": [_V, ctypes.c_int],
    "ssim_batch_adjoint_step": [_V, _V, _V, _V, _V, _V, _I],
    "ssim_batch_get_layout_info": [_V, ctypes.POINTER(LayoutInfo)],
    "ssim_batch_set_tuning": [_V, ctypes.c_int, ctypes.c_int],
    "ssim_batch_get_tuning": [_V, ctypes.c_int, _I],
    "ssim_batch_set_kernel_timing": [_V, ctypes.c_int],
    "ssim_batch_get_kernel_timing": [_V, ctypes.c_int, _D, ctypes.POINTER(ctypes.c_longlong)],
}

TUNING = {"xcd_remap": 0, "pcg_chunk_max": 1, "spmv_lanes": 2, "pcg_graph": 3, "lbfgs_gram": 5, "grad_tile": 6,
          "hess_tile": 7, "rest_from_x": 8}
BATCH_TUNING = dict(TUNING, compact_pct=4)

_lib = None


class SsimError(RuntimeError):
    """Raised for a non-zero ssim_status_t; ``code`` is the status."""

    def __init__(self, code, msg):
        super().__init__(f"[ssim status {code}] {msg}")
        self.code = code


class InvalidArgument(SsimError, ValueError):
    pass


def _share_torch_runtime():
    """PyTorch-ROCm wheels carry their own HIP and HSA runtime libraries.  If libssim_hip.so were loaded first,
    it would bring in the system's HIP runtime and a later ``import torch`` would start a second runtime in the
    same process, which then finds no GPU.  Importing torch first (when it is installed) makes the library bind
    to the runtime torch already loaded, so the process has exactly one."""
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def load(path: str = LIB_PATH):
    """Load libssim_hip.so.  Raises if it has not been built (no fallback exists)."""
    global _lib
    if _lib is None:
        if not os.path.exists(path):
            raise ImportError(f"libssim_hip.so not found at {path}: build it with `python -m ssim_amd.build`")
        _share_torch_runtime()
        L = ctypes.CDLL(path)
        L.ssim_last_error.restype = ctypes.c_char_p
        L.ssim_last_error.argtypes = []
        for name, args in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = _S
            f.argtypes = args
        _lib = L
    return _lib


def check(status: int):
    if status != 0:
        msg = load().ssim_last_error().decode("utf-8", "replace")
        if status == 1:
            raise InvalidArgument(status, msg)
        raise SsimError(status, msg)


def call(name, *args):
    check(getattr(load(), name)(*args))
