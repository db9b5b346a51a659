"""GPU parity: the HIP analysis and solves against the oracle on the same factor.

Bar: level arrays bit-exact (integer work); solves within a stated relative
tolerance of the oracle's restatement of the reference kernels in the same
precision (the HIP kernels reduce a row's products in a different order),
and within 1e-12 of the FP64 host solve for double precision.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from cholespy_amd import CholeskySolverD, CholeskySolverF, MatrixType, workloads
from oracle import oracle

pytestmark = pytest.mark.gpu

DEV = "cuda"

CASES = {
    "cube": lambda: (lambda n, f: (n,) + workloads.mesh_laplacian(n, f, 2.0))(*workloads.cube()),
    "ico3": lambda: (lambda n, f: (n,) + workloads.mesh_laplacian(n, f, 2.0))(*workloads.icosphere(3)),
    "ico5": lambda: (lambda n, f: (n,) + workloads.mesh_laplacian(n, f, 2.0))(*workloads.icosphere(5)),
    "grid64": lambda: (4096,) + workloads.grid2d_laplacian(64, 64, 1e-2),
}

# Relative (max-norm) tolerance against the same-precision oracle.
TOL = {np.float32: 2e-5, np.float64: 1e-12}


def _gpu_solver(cls, n, values, idx):
    return cls(n, torch.tensor(idx[0], device=DEV), torch.tensor(idx[1], device=DEV),
               torch.tensor(values, device=DEV), MatrixType.COO)


@pytest.mark.parametrize("case", list(CASES))
def test_levels_bit_exact(case):
    n, values, idx = CASES[case]()
    solver = _gpu_solver(CholeskySolverD, n, values, idx)
    fac = solver._factor()
    rowptr, cols, _ = oracle.factor_rows(fac["colptr"], fac["rowidx"], fac["val"])
    for lower, rp, cl in ((True, rowptr, cols), (False, fac["colptr"], fac["rowidx"])):
        ref_level, mx = oracle.levels(rp, cl, lower)
        order, row_level = solver._levels(lower)
        assert np.array_equal(row_level, ref_level), (case, lower)
        assert np.array_equal(order, oracle.level_sort(ref_level, mx)), (case, lower)
        assert solver._n_levels(lower) == mx + 1


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("cls,dtype", [(CholeskySolverF, np.float32), (CholeskySolverD, np.float64)])
def test_solve_matches_oracle(case, cls, dtype):
    n, values, idx = CASES[case]()
    solver = _gpu_solver(cls, n, values, idx)
    fac = solver._factor()
    rng = np.random.default_rng(1)
    for k in (1, 3, 4, 16, 64, 65, 128):
        b = rng.random((n, k)).astype(dtype)
        bt = torch.tensor(b, device=DEV)
        xt = torch.zeros_like(bt)
        solver.solve(bt, xt)
        torch.cuda.synchronize()
        solver._check()
        x = xt.cpu().numpy()
        ref = oracle.solve(fac, b, dtype)
        err = np.abs(x - ref).max() / np.abs(ref).max()
        assert err < TOL[dtype], (case, k, err)
        if dtype == np.float64:
            host = oracle.host_solve(fac, b)
            assert np.abs(x - host).max() / np.abs(host).max() < 1e-12


def test_repeated_solves_and_rhs_change():
    """Work buffers are reused across solves and re-laid-out when k changes."""
    n, values, idx = CASES["ico5"]()
    solver = _gpu_solver(CholeskySolverF, n, values, idx)
    fac = solver._factor()
    rng = np.random.default_rng(2)
    for k in (3, 3, 1, 128, 3, 7, 7):
        b = rng.standard_normal((n, k)).astype(np.float32)
        bt = torch.tensor(b, device=DEV)
        xt = torch.zeros_like(bt)
        solver.solve(bt, xt)
        torch.cuda.synchronize()
        solver._check()
        ref = oracle.solve(fac, b, np.float32)
        assert np.abs(xt.cpu().numpy() - ref).max() / np.abs(ref).max() < 2e-5, k


def test_vector_rhs_and_identity():
    """README example: A = I gives x = b (README.md:104-121), 1-D b."""
    n = 20
    rows = torch.arange(n, device=DEV)
    solver = CholeskySolverF(n, rows, rows, torch.ones(n, device=DEV), MatrixType.COO)
    b = torch.ones(n, device=DEV)
    x = torch.zeros_like(b)
    solver.solve(b, x)
    torch.cuda.synchronize()
    assert torch.equal(x, b)


def test_too_many_rhs_raises():
    n, values, idx = CASES["cube"]()
    solver = _gpu_solver(CholeskySolverF, n, values, idx)
    b = torch.ones((n, 129), device=DEV)
    with pytest.raises(ValueError, match="The number of RHS should be less than 128."):
        solver.solve(b, torch.zeros_like(b))


def test_full_size_mesh_smoothing_residual():
    """icosphere(7), n = 163,842, k = 3: size-independent checks at full size.
    FP64 GPU solve vs the FP64 host solve on the same factor, and the FP32
    solve's relative residual."""
    n, values, idx, rhs = workloads.smoothing_system(7, 10.0, seed=0)
    A = oracle.symmetric_from_lower(n, *oracle.normalized_lower_csc(n, idx[0], idx[1], values, "COO"))
    for cls, dtype, tol in ((CholeskySolverD, np.float64, 1e-12), (CholeskySolverF, np.float32, 1e-5)):
        solver = _gpu_solver(cls, n, values, idx)
        b = rhs.astype(dtype)
        bt = torch.tensor(b, device=DEV)
        xt = torch.zeros_like(bt)
        solver.solve(bt, xt)
        torch.cuda.synchronize()
        solver._check()
        x = xt.cpu().numpy().astype(np.float64)
        host = oracle.host_solve(solver._factor(), b.astype(np.float64))
        rel = np.linalg.norm(x - host) / np.linalg.norm(host)
        res = np.linalg.norm(b - A @ x) / np.linalg.norm(b)
        assert rel < tol and res < tol, (dtype, rel, res)
