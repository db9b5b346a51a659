"""Failure handling on the public path, forced from the host so that every
kernel still exits normally: the watchdog's give-up (a tiny poll budget per
wait, or the abort word raised from the host) is reported by the next call
as RuntimeError (the reference prints device errors, src/cuda_driver.cpp:
48-56), the results of the call that gave up are NaN rather than finite and
wrong, and the call after the report solves correctly again. Also: any
order of single passes is valid (the reference clears its flags at every
launch, src/cholesky_solver.cpp:380-382), and gradient indices outside the
matrix are rejected."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import cholespy_amd
from cholespy_amd import CholeskySolverD, CholeskySolverF, MatrixType, workloads
from cholespy_amd.autograd import coo_indices, solve
from oracle import oracle

pytestmark = pytest.mark.gpu
DEV = "cuda"
PATHS = {"rows": cholespy_amd._core._PATH_ROWS, "blocks": cholespy_amd._core._PATH_BLOCKS}


def _ico(level, cls):
    n, faces = workloads.icosphere(level)
    values, idx = workloads.mesh_laplacian(n, faces, 2.0)
    s = cls(n, torch.tensor(idx[0], device=DEV), torch.tensor(idx[1], device=DEV),
            torch.tensor(values, device=DEV), MatrixType.COO)
    return n, s


def _solve_checked(solver, b, dtype, tol):
    bt = torch.tensor(b, device=DEV)
    xt = torch.zeros_like(bt)
    solver.solve(bt, xt)
    torch.cuda.synchronize()
    ref = oracle.solve(solver._factor(), b, dtype)
    err = np.abs(xt.cpu().numpy() - ref).max() / np.abs(ref).max()
    assert err < tol, err


def test_abort_word_is_reported_then_cleared():
    n, s = _ico(4, CholeskySolverF)
    b = np.random.default_rng(0).random((n, 3)).astype(np.float32)
    _solve_checked(s, b, np.float32, 2e-5)
    s._raise_abort()
    bt = torch.tensor(b, device=DEV)
    with pytest.raises(RuntimeError, match="abort"):
        s.solve(bt, torch.zeros_like(bt))
    _solve_checked(s, b, np.float32, 2e-5)
    s._check()


@pytest.mark.parametrize("path", ["rows", "blocks"])
def test_give_up_gives_nan_then_reports_and_recovers(path):
    """With a poll budget of one poll, the first wait that finds its value
    missing gives up; the solve's output then holds NaN, the next call raises
    RuntimeError, and the one after it matches the oracle."""
    n, s = _ico(5, CholeskySolverD)
    s._set_path(PATHS[path])
    rng = np.random.default_rng(1)
    b = rng.random((n, 3))
    bt = torch.tensor(b, device=DEV)
    s._set_spin_limit(1)
    gave_up = False
    for _ in range(8):   # a wait has to find its input missing once; this is near certain
        xt = torch.zeros_like(bt)
        s.solve(bt, xt)
        torch.cuda.synchronize()
        if torch.isnan(xt).any():
            gave_up = True
            break
        # a solve in which every value was ready at its first poll is exact
        ref = oracle.host_solve(s._factor(), b)
        assert np.abs(xt.cpu().numpy() - ref).max() / np.abs(ref).max() < 1e-12
    s._set_spin_limit(0)
    assert gave_up
    with pytest.raises(RuntimeError, match="bounded wait"):
        s.solve(bt, torch.zeros_like(bt))
    _solve_checked(s, b, np.float64, 1e-12)
    _solve_checked(s, b, np.float64, 1e-12)
    s._check()


@pytest.mark.parametrize("path", ["rows", "blocks"])
@pytest.mark.parametrize("cls,dtype", [(CholeskySolverF, torch.float32), (CholeskySolverD, torch.float64)])
def test_any_pass_order(cls, dtype, path):
    n, s = _ico(5, cls)
    s._set_path(PATHS[path])
    b1 = torch.rand(n, 3, device=DEV, dtype=dtype)
    b2 = torch.rand(n, 3, device=DEV, dtype=dtype)
    x_ref = torch.zeros_like(b2)
    s.solve(b2, x_ref)
    # two forward passes, then the backward pass: the second forward pass wins
    x = torch.zeros_like(b2)
    s._solve_pass(0, b1, x)
    s._solve_pass(0, b2, x)
    s._solve_pass(1, b2, x)
    torch.cuda.synchronize()
    assert torch.equal(x, x_ref)
    # a forward pass followed by a whole solve
    s._solve_pass(0, b1, x)
    x2 = torch.zeros_like(b2)
    s.solve(b2, x2)
    torch.cuda.synchronize()
    assert torch.equal(x2, x_ref)
    # a lone backward pass: the block schedule reads the y of the last
    # forward pass (that solve's); the row schedule consumed it. Either way
    # it ends, and the next solve is exact.
    x3 = torch.zeros_like(b2)
    s._solve_pass(1, b1, x3)
    torch.cuda.synchronize()
    if path == "blocks":
        assert torch.equal(x3, x_ref)
    x4 = torch.zeros_like(b2)
    s.solve(b2, x4)
    torch.cuda.synchronize()
    assert torch.equal(x4, x_ref)
    s._check()


def test_gradient_indices_out_of_range():
    n, faces = workloads.icosphere(3)
    values, idx = workloads.mesh_laplacian(n, faces, 2.0)
    ii = torch.tensor(idx[0], device=DEV)
    jj = torch.tensor(idx[1], device=DEV)
    vals = torch.tensor(values, device=DEV, requires_grad=True)
    s = CholeskySolverD(n, ii, jj, vals.detach(), MatrixType.COO)
    rows, cols = coo_indices(n, ii, jj, MatrixType.COO)
    bad = rows.clone()
    bad[5] = n + 7
    b = torch.rand(n, 3, device=DEV, dtype=torch.float64)
    x = solve(s, b, vals, bad, cols)
    with pytest.raises(IndexError):
        x.sum().backward()
    # the kernel itself never reads through an out-of-range index: NaN there
    lam = torch.rand(n, 3, device=DEV, dtype=torch.float64)
    xs = torch.rand(n, 3, device=DEV, dtype=torch.float64)
    g = torch.empty(rows.shape[0], device=DEV, dtype=torch.float64)
    cholespy_amd._core._entry_gradient(bad, cols, lam, xs, g)
    torch.cuda.synchronize()
    gh = g.cpu().numpy()
    assert np.isnan(gh[5]) and np.isfinite(np.delete(gh, 5)).all()
