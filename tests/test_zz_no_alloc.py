"""ssim_pcg_solve_noalloc (the entry point that allocates nothing; its work buffers live in a static device array)
against ssim_pcg_solve on the same matrix and right-hand side: bitwise-identical solutions and iteration counts,
on two calls in a row, the second on a smaller mesh (so the shared array holds the first call's contents), and a
mesh too large for the array is rejected."""
import ctypes

import numpy as np
import pytest

from ssim_amd import _lib, fem, mesh

pytestmark = pytest.mark.gpu


def _solve_both(case, seed):
    import torch

    s = fem.TetFiniteElementSolver_Hip.from_case(case, rtol=1e-6)
    s.set_array("next_deform", case.u0 * 2.0)
    s.update_hessian(make_spsd=True)
    b = np.random.default_rng(seed).standard_normal(3 * case.nv)
    b[case.dof_type == 1] = 0.0
    x_ref, it_ref = s.pcg_solve(b, rtol=1e-10)
    bd = torch.as_tensor(b, device="cuda")
    xd = torch.full_like(bd, float("nan"))
    torch.cuda.synchronize()
    it = ctypes.c_int(-1)
    _lib.call("ssim_pcg_solve_noalloc", s.handle, ctypes.c_void_p(bd.data_ptr()), ctypes.c_void_p(xd.data_ptr()),
              1e-10, 10000, ctypes.byref(it))
    return x_ref, it_ref, xd.cpu().numpy(), it.value


def test_noalloc_pcg_matches_the_allocating_entry_point_on_two_calls():
    big = mesh.cantilever_bar(32, 8, 8, seed=1)
    small = mesh.cantilever_bar(8, 2, 2, seed=2)
    for case, seed in ((big, 5), (small, 6)):
        print(f"[gpu] no-alloc PCG on {case.name} ({case.nv} vertices)", flush=True)
        x_ref, it_ref, x, it = _solve_both(case, seed)
        assert it == it_ref > 0
        assert np.array_equal(x, x_ref)


def test_noalloc_pcg_rejects_meshes_beyond_its_scratch():
    import torch

    # one tet plus 750,000 vertices in no tet: 2.25M dofs, whose four work vectors alone exceed 64 MiB
    nv = 750_004
    X = np.zeros((nv, 3))
    X[:4] = mesh.tetra()[0]
    X[4:, 0] = 2.0 + np.arange(nv - 4)
    s = fem.TetFiniteElementSolver_Hip(X, np.array([[0, 1, 2, 3]], dtype=np.int32))
    s.reset()
    bd = torch.zeros(3 * nv, dtype=torch.float64, device="cuda")
    xd = torch.zeros_like(bd)
    st = _lib.load().ssim_pcg_solve_noalloc(s.handle, ctypes.c_void_p(bd.data_ptr()), ctypes.c_void_p(xd.data_ptr()),
                                            1e-8, 100, None)
    assert st == 1 and b"scratch" in _lib.load().ssim_last_error()
