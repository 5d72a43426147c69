"""The host side of the header-only API (include/mathprim, include/ssim/mesh), compiled with g++ and run on the
CPU: the C++ unit_box equals the Python generator (the reference's 5-tet split, common_shapes.hpp:67-132) vertex
for vertex and cell for cell; views, buffers, par::seq, the CSR visit task and the .npy writer behave as
documented (the written file equals numpy.save's bytes)."""
import io
import os
import subprocess

import numpy as np
import pytest

from ssim_amd import mesh

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


@pytest.fixture(scope="module")
def host_check(tmp_path_factory):
    exe = tmp_path_factory.mktemp("host_api") / "host_api_check"
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                    os.path.join(HERE, "cpp", "host_api_check.cpp"), "-o", str(exe)], check=True)
    return exe


@pytest.mark.parametrize("dims", [(2, 2, 2), (5, 3, 4), (8, 2, 2)])
def test_unit_box_and_views_match(tmp_path, host_check, dims):
    out, npy = tmp_path / "o.bin", tmp_path / "o.npy"
    subprocess.run([str(host_check), *map(str, dims), str(out), str(npy)], check=True)
    raw = out.read_bytes()
    nv, nt = np.frombuffer(raw[:8], dtype=np.int32)
    off = 8
    X = np.frombuffer(raw[off:off + 24 * nv], dtype=np.float64).reshape(nv, 3)
    off += 24 * nv
    C = np.frombuffer(raw[off:off + 16 * nt], dtype=np.int32).reshape(nt, 4)
    off += 16 * nt
    checks = np.frombuffer(raw[off:off + 32], dtype=np.float64)
    Xp, Cp = mesh.unit_box_rows(*dims)
    Xp = Xp.copy()
    Xp[:, 0] *= 4.0
    assert np.array_equal(X, Xp) and np.array_equal(C, Cp)
    assert checks[0] == 1 * 1 + 2 * 11 + 3 * 12  # Σ v·(1 + r + 10c) over (0,0), (0,1), (1,1)
    assert checks[1] == pytest.approx(Xp.sum(), rel=1e-14)
    assert checks[2] == Xp.reshape(3, -1)[1, 0]
    assert checks[3] == Xp[1].sum()
    buf = io.BytesIO()
    np.save(buf, Xp)
    assert npy.read_bytes() == buf.getvalue()
