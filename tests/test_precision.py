"""The precision workload and the long-chain block plans it triggers.

Full size: the 3D 7-point Neumann Laplacian on a 64^3 grid + 1e-4 I
(n = 262,144), factored once in FP64 on the host and cast to FP32 (as
src/cholesky_solver.cpp:254-271 casts the CHOLMOD factor), 32 right-hand
sides ~ N(0, 1): FP32 block solves + FP64 residual refinement until the
relative residual max_j ||b_j - A x_j|| / ||b_j|| <= 1e-12, checked on the
host with SciPy, with the error against the oracle's FP64 host solve on the
same factor printed and bounded.

Reduced size: a 40^3 grid (+ I), whose elimination tree has a 2,327-row chain, so
the planner cuts chain blocks of more than 1024 rows (2048 by default here;
4096 forced), at k = 3 (vector-FMA kernels) and k = 32 (MFMA panel kernels),
FP32 and FP64, against the oracle's restatement of the reference's solves."""
import numpy as np
import pytest
import scipy.sparse as sp

torch = pytest.importorskip("torch")

from cholespy_amd import CholeskySolverD, CholeskySolverF, MatrixType, workloads
from oracle import oracle

pytestmark = pytest.mark.gpu
DEV = "cuda"
TOL = {np.float32: 2e-5, np.float64: 1e-12}


def _solver(cls, n, values, idx):
    return cls(n, torch.tensor(idx[0], device=DEV), torch.tensor(idx[1], device=DEV),
               torch.tensor(values, device=DEV), MatrixType.COO)


def test_precision_workload_refines_to_1e_12():
    m, k = 64, 32
    n = m ** 3
    values, idx = workloads.grid3d_laplacian(m, 1e-4)
    rhs = np.random.default_rng(0).standard_normal((n, k))
    print(f"64^3 grid: n={n} k={k}", flush=True)
    sf = _solver(CholeskySolverF, n, values, idx)
    plan = sf._plan_info()
    print("plan", plan, flush=True)
    A = sp.coo_matrix((values, (idx[0], idx[1])), shape=(n, n)).tocsr()
    b = torch.tensor(rhs, device=DEV)
    x = torch.zeros_like(b)
    steps, res_dev = sf._solve_refined_tol(b, x, 1e-12, 10)
    torch.cuda.synchronize()
    sf._check()
    xh = x.cpu().numpy()
    res = float((np.linalg.norm(rhs - A @ xh, axis=0) / np.linalg.norm(rhs, axis=0)).max())
    x_ref = oracle.host_solve(sf._factor(), rhs)
    err = float(np.linalg.norm(xh - x_ref) / np.linalg.norm(x_ref))
    print(f"refinement steps {steps}, residual {res:.2e} (device {res_dev:.2e}), "
          f"error vs FP64 host solve {err:.2e}", flush=True)
    assert res <= 1e-12 and res_dev <= 1e-12
    assert abs(res - res_dev) <= 0.1 * res + 1e-16
    assert 1 <= steps <= 4
    # kappa ~ 1.2e5: the error is bounded by kappa times the residual
    assert err <= 1.2e5 * 1e-12
    # the plain FP32 solve alone is far from that: refinement did the work
    x32 = torch.zeros(n, k, device=DEV, dtype=torch.float32)
    sf.solve(b.float(), x32)
    torch.cuda.synchronize()
    r32 = float((np.linalg.norm(rhs - A @ x32.cpu().numpy().astype(np.float64), axis=0)
                 / np.linalg.norm(rhs, axis=0)).max())
    assert r32 > 1e-8


@pytest.fixture(scope="module")
def grid40():
    # The chain structure depends on the pattern only; a unit shift keeps
    # kappa ~ 13, so FP32 solves that round in different orders agree to
    # the same tolerance as on the mesh cases.
    m = 40
    values, idx = workloads.grid3d_laplacian(m, 1.0)
    return m ** 3, values, idx


@pytest.mark.parametrize("max_block", [0, 2048, 4096])
@pytest.mark.parametrize("cls,dtype", [(CholeskySolverF, np.float32), (CholeskySolverD, np.float64)])
def test_long_chain_blocks_match_oracle(grid40, cls, dtype, max_block, monkeypatch):
    n, values, idx = grid40
    if max_block:
        monkeypatch.setenv("CHOLESPY_MAX_BLOCK", str(max_block))
    print(f"40^3 grid n={n} max_block={max_block or 'auto'} {dtype.__name__}", flush=True)
    s = _solver(cls, n, values, idx)
    pa = s._plan_arrays()
    biggest = int(np.diff(pa["start"]).max())
    # the longest chain (2,327 rows) gives one block of more than 1024 rows
    assert biggest > 1024, biggest
    if max_block == 4096:
        assert biggest > 2048
    fac = s._factor()
    rng = np.random.default_rng(3)
    for k in (3, 32):
        b = rng.standard_normal((n, k)).astype(dtype)
        bt = torch.tensor(b, device=DEV)
        xt = torch.zeros_like(bt)
        s.solve(bt, xt)
        torch.cuda.synchronize()
        s._check()
        ref = oracle.solve(fac, b, dtype)
        err = np.abs(xt.cpu().numpy() - ref).max() / np.abs(ref).max()
        assert err < TOL[dtype], (k, err)
