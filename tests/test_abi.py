"""The C ABI library: it loads, exports every entry point include/ssim_hip.h declares, and validates its
arguments before touching a device (CPU; no compute call needs a GPU here)."""
import ctypes
import os
import re

import numpy as np
import pytest

from ssim_amd import _lib, fem, mesh

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    with open(os.path.join(ROOT, "include", "ssim_hip.h")) as f:
        text = f.read()
    return sorted(set(re.findall(r"\b(ssim_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = _lib.load()
    names = _declared()
    assert len(names) >= 60
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # and the ctypes table covers all of them
    assert set(names) - {"ssim_last_error"} == set(_lib.SIGNATURES)


def test_invalid_arguments_are_rejected_before_device_use():
    X, C = mesh.tetra()
    bad = C.copy()
    bad[0, 3] = 4  # vertex out of range
    with pytest.raises(ValueError, match="outside"):
        fem.TetFiniteElementSolver_Hip(X, bad)
    with pytest.raises(ValueError, match="Poisson"):
        fem.TetFiniteElementSolver_Hip(X, C, 1e-2, 1e6, 0.5, 1.0)
    with pytest.raises(ValueError, match="Time step"):
        fem.TetFiniteElementSolver_Hip(X, C, 0.0, 1e6, 0.3, 1.0)
    with pytest.raises(ValueError, match="Unknown method"):
        fem.TetFiniteElementSolver_Hip(X, C, method="gauss_seidel")
    with pytest.raises(ValueError, match="shape"):
        fem.TetFiniteElementSolver_Hip(X.T, C)
    h = ctypes.c_void_p()
    st = _lib.load().ssim_create(None, None, 0, None, 0, 1.0, 1.0, 0.3, 1.0)
    assert st == 1 and b"null" in _lib.load().ssim_last_error()
    assert _lib.load().ssim_destroy(None) == 0
    assert h.value is None


def test_status_codes_map_to_python_exceptions():
    with pytest.raises(_lib.InvalidArgument):
        _lib.check(1)
    with pytest.raises(_lib.SsimError):
        _lib.check(2)
    _lib.check(0)


def test_batch_arguments_are_rejected_before_device_use():
    from ssim_amd import batch

    X, C = mesh.tetra()
    lam, mu = np.ones(1), np.ones(1)
    bad = C.copy()
    bad[0, 0] = 7
    with pytest.raises(ValueError, match="outside"):
        batch.TetFiniteElementBatch_Hip([X, X], [C, bad], [lam, lam], [mu, mu])
    with pytest.raises(ValueError, match="Lamé"):
        batch.TetFiniteElementBatch_Hip([X], [C], [-lam], [mu])
    with pytest.raises(ValueError, match="one entry per tet"):
        batch.TetFiniteElementBatch_Hip([X], [C], [np.ones(2)], [mu])
    with pytest.raises(ValueError, match="Time step"):
        batch.TetFiniteElementBatch_Hip([X], [C], [lam], [mu], time_step=0.0)
    with pytest.raises(ValueError, match="non-empty"):
        batch.TetFiniteElementBatch_Hip([], [], [], [])
    L = _lib.load()
    assert L.ssim_batch_create(None, 1, None, None, None, None, None, None, 1.0, 1.0) == 1
    assert L.ssim_batch_destroy(None) == 0
    assert L.ssim_batch_step(None) == 1 and b"null batch" in L.ssim_last_error()


def test_one_hip_runtime_per_process():
    """Loading the library and torch in either order leaves a single HIP runtime mapped (two runtimes in one
    process leave the second without a GPU)."""
    import subprocess
    import sys

    code = ("import sys; sys.path.insert(0, %r)\n"
            "from ssim_amd import _lib\n_lib.load()\nimport torch\n"
            "maps = open('/proc/self/maps').read().splitlines()\n"
            "print(len({l.split()[-1] for l in maps if 'libamdhip64' in l}))\n") % ROOT
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, check=True).stdout
    assert out.strip() == "1", out


def test_reference_class_names_map_to_the_gpu_class():
    """pyssim.fem's TetFiniteElementSolver_Host / _Cuda (pyssim/fem/__init__.py:11, 97) are the GPU class."""
    from ssim_amd import fem

    assert fem.TetFiniteElementSolver_Cuda is fem.TetFiniteElementSolver_Hip
    assert fem.TetFiniteElementSolver_Host is fem.TetFiniteElementSolver_Hip
    assert callable(fem.unit_box)
