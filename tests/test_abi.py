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
    assert len(names) >= 40
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
        fem.TetFiniteElementSolver_Hip(X, C, method="ncg")
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
