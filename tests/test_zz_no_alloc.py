"""cholespy_solver_solve_noalloc through the C ABI (ctypes), against the
existing entry points cholespy_solver_solve_f32/f64 on the same handle:
bit-identical results (the same kernels, deterministic reduction orders),
on two calls in a row where the second is smaller (a smaller system and
fewer right-hand sides), then back to the first layout, and a request that
exceeds the fixed array is refused with a clear error."""
import ctypes
from pathlib import Path

import numpy as np
import pytest

torch = pytest.importorskip("torch")

import cholespy_amd  # noqa: F401  (torch's HIP runtime first, then the library)
from cholespy_amd import CholeskySolverD, MatrixType, workloads

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parent.parent
P, I = ctypes.c_void_p, ctypes.c_int


@pytest.fixture(scope="module")
def lib():
    L = ctypes.CDLL(str(ROOT / "cholespy_amd" / "libcholespy_amd.so"))
    L.cholespy_solver_solve_noalloc.argtypes = [P, I, P, P, P]
    L.cholespy_solver_solve_f32.argtypes = [P, I, P, P, P]
    L.cholespy_solver_solve_f64.argtypes = [P, I, P, P, P]
    L.cholespy_solver_create.argtypes = [I, I, ctypes.POINTER(P)]
    L.cholespy_solver_destroy.argtypes = [P]
    L.cholespy_solver_upload_factor.argtypes = [P, P, P, P, P, ctypes.c_int64]
    L.cholespy_solver_check.argtypes = [P]
    L.cholespy_last_error.restype = ctypes.c_char_p
    return L


class Handle:
    """A C-ABI solver handle for icosphere(level), factored on the host."""

    def __init__(self, lib, level, fbytes):
        n, faces = workloads.icosphere(level)
        values, idx = workloads.mesh_laplacian(n, faces, 2.0)
        f = CholeskySolverD(n, idx[0], idx[1], values, MatrixType.COO)._factor()   # CPU mode
        self.lib, self.n, self.h = lib, n, P()
        self.arrays = [np.ascontiguousarray(f[key], dtype=dt) for key, dt in
                       (("perm", np.int32), ("colptr", np.int32), ("rowidx", np.int32), ("val", np.float64))]
        assert lib.cholespy_solver_create(n, fbytes, ctypes.byref(self.h)) == 0
        perm, colptr, rowidx, val = self.arrays
        assert lib.cholespy_solver_upload_factor(self.h, perm.ctypes.data, colptr.ctypes.data,
                                                 rowidx.ctypes.data, val.ctypes.data, len(val)) == 0, \
            lib.cholespy_last_error()

    def close(self):
        self.lib.cholespy_solver_destroy(self.h)


@pytest.mark.parametrize("fbytes,dtype", [(4, torch.float32), (8, torch.float64)])
def test_noalloc_matches_solve(lib, fbytes, dtype):
    big, small = Handle(lib, 5, fbytes), Handle(lib, 4, fbytes)
    try:
        solve = lib.cholespy_solver_solve_f64 if fbytes == 8 else lib.cholespy_solver_solve_f32
        gen = torch.Generator(device="cuda").manual_seed(0)
        # second call smaller (system and n_rhs), then the first layout again,
        # then a wide n_rhs (MFMA panel kernels) and a narrow one
        for hd, k in ((big, 8), (small, 3), (big, 8), (big, 3), (small, 40)):
            print(f"no-alloc solve: n={hd.n} k={k} {dtype}", flush=True)
            b = torch.rand(hd.n, k, device="cuda", dtype=dtype, generator=gen)
            x_ref = torch.zeros_like(b)
            x = torch.full_like(b, float("nan"))
            assert solve(hd.h, k, b.data_ptr(), x_ref.data_ptr(), None) == 0, lib.cholespy_last_error()
            assert lib.cholespy_solver_solve_noalloc(hd.h, k, b.data_ptr(), x.data_ptr(), None) == 0, \
                lib.cholespy_last_error()
            torch.cuda.synchronize()
            assert lib.cholespy_solver_check(hd.h) == 0, lib.cholespy_last_error()
            assert torch.equal(x, x_ref), (hd.n, k)
    finally:
        big.close()
        small.close()


def test_noalloc_refuses_oversized_work(lib):
    hd = Handle(lib, 6, 8)          # n = 40,962: 128 FP64 columns need ~130 MB
    try:
        b = torch.rand(hd.n, 128, device="cuda", dtype=torch.float64)
        x = torch.zeros_like(b)
        assert lib.cholespy_solver_solve_noalloc(hd.h, 128, b.data_ptr(), x.data_ptr(), None) == 1
        assert b"no-allocation solve needs" in lib.cholespy_last_error()
    finally:
        hd.close()
