"""The C ABI driven the way a reference maintainer would bind it (the ctypes
stub of INTEGRATION.md, section 3), on the committed fixtures
tests/golden/*.npz: create -> upload_factor (the fixture's own L and perm,
not this project's host factor) -> analyze -> get_levels, which must equal
the fixture's level_lower / level_upper and order_* (written from the
elimination-tree closed forms), then solve_f64 against the fixture's x64
(SciPy spsolve on A), solve_f32 against the oracle's FP32 restatement, and
destroy. Replaces src/cholesky_solver.cpp:299-432 (analyze_cuda,
launch_kernel, solve_cuda)."""
import ctypes
from pathlib import Path

import numpy as np
import pytest

torch = pytest.importorskip("torch")

import cholespy_amd  # noqa: F401  (loads torch's HIP runtime first, then the library)
from oracle import oracle

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parent.parent
GOLDEN = ROOT / "tests" / "golden"
P, I, I64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64


@pytest.fixture(scope="module")
def lib():
    L = ctypes.CDLL(str(ROOT / "cholespy_amd" / "libcholespy_amd.so"))
    L.cholespy_solver_create.argtypes = [I, I, ctypes.POINTER(P)]
    L.cholespy_solver_destroy.argtypes = [P]
    L.cholespy_solver_upload_factor.argtypes = [P, P, P, P, P, I64]
    L.cholespy_solver_analyze.argtypes = [P, I, ctypes.POINTER(I)]
    L.cholespy_solver_get_levels.argtypes = [P, I, P, P]
    L.cholespy_solver_set_path.argtypes = [P, I]
    L.cholespy_solver_solve_f32.argtypes = [P, I, P, P, P]
    L.cholespy_solver_solve_f64.argtypes = [P, I, P, P, P]
    L.cholespy_solver_check.argtypes = [P]
    L.cholespy_last_error.restype = ctypes.c_char_p
    return L


def _ok(rc, L):
    assert rc == 0, L.cholespy_last_error().decode()


@pytest.mark.parametrize("name", ["cube", "ico3", "grid64"])
@pytest.mark.parametrize("path", [0, 1])   # CHOLESPY_PATH_ROWS, CHOLESPY_PATH_BLOCKS
def test_golden_fixture_through_the_c_abi(lib, name, path):
    g = np.load(GOLDEN / f"{name}.npz")
    n = int(g["n"])
    perm = np.ascontiguousarray(g["perm"], dtype=np.int32)
    colptr = np.ascontiguousarray(g["colptr"], dtype=np.int32)
    rowidx = np.ascontiguousarray(g["rowidx"], dtype=np.int32)
    val = np.ascontiguousarray(g["val"], dtype=np.float64)
    b = np.ascontiguousarray(g["b"])
    k = b.shape[1]
    fac = {"perm": perm, "colptr": colptr, "rowidx": rowidx, "val": val}
    for fbytes, dtype, tdtype in ((8, np.float64, torch.float64), (4, np.float32, torch.float32)):
        h = P()
        _ok(lib.cholespy_solver_create(n, fbytes, ctypes.byref(h)), lib)
        try:
            _ok(lib.cholespy_solver_upload_factor(h, perm.ctypes.data, colptr.ctypes.data,
                                                   rowidx.ctypes.data, val.ctypes.data, len(val)), lib)
            for lower, key in ((1, "lower"), (0, "upper")):
                nl = I(0)
                _ok(lib.cholespy_solver_analyze(h, lower, ctypes.byref(nl)), lib)
                order = np.zeros(n, dtype=np.int32)
                level = np.zeros(n, dtype=np.int32)
                _ok(lib.cholespy_solver_get_levels(h, lower, order.ctypes.data, level.ctypes.data), lib)
                assert np.array_equal(level, g[f"level_{key}"]), (name, key)
                assert np.array_equal(order, g[f"order_{key}"]), (name, key)
                assert nl.value == int(g[f"level_{key}"].max()) + 1
            _ok(lib.cholespy_solver_set_path(h, path), lib)
            bt = torch.tensor(b, dtype=tdtype, device="cuda")
            xt = torch.zeros_like(bt)
            solve = lib.cholespy_solver_solve_f64 if fbytes == 8 else lib.cholespy_solver_solve_f32
            _ok(solve(h, k, bt.data_ptr(), xt.data_ptr(), None), lib)
            torch.cuda.synchronize()
            _ok(lib.cholespy_solver_check(h), lib)
            x = xt.cpu().numpy()
            if fbytes == 8:
                x64 = g["x64"]
                assert np.abs(x - x64).max() / np.abs(x64).max() <= 1e-12, name
            else:
                ref = oracle.solve(fac, b.astype(np.float32), np.float32)
                assert np.abs(x - ref).max() / np.abs(ref).max() < 2e-5, name
        finally:
            _ok(lib.cholespy_solver_destroy(h), lib)
