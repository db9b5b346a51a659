"""Differentiation (mesh-smoothing system x = (I + lam L)^{-1} b, loss
||x - t||^2): the hand-written backward against a dense FP64 gradient and
FP64 central finite differences computed with SciPy (independent of this
project's solver)."""
import numpy as np
import pytest
import scipy.sparse as sp
from scipy.sparse.linalg import spsolve

torch = pytest.importorskip("torch")

from cholespy_amd import CholeskySolverD, CholeskySolverF, MatrixType, workloads
from cholespy_amd.autograd import coo_indices, solve

DEVICES = ["cpu", pytest.param("cuda", marks=pytest.mark.gpu)]


def system(level, lam=10.0, seed=0):
    n, faces = workloads.icosphere(level)
    values, idx = workloads.mesh_laplacian(n, faces, lam)
    rng = np.random.default_rng(seed)
    b = rng.standard_normal((n, 3))
    t = rng.standard_normal((n, 3))
    return n, values, idx, b, t


def dense_reference(n, values, idx, b, t):
    """FP64 dense gradients: x = A^{-1} b, g = 2 (x - t), lam = A^{-1} g,
    dL/db = lam, dL/dv_e by the lower-triangle rule."""
    A = np.zeros((n, n))
    i, j = idx
    low = i >= j
    np.add.at(A, (i[low], j[low]), values[low])
    strict = i > j
    np.add.at(A, (j[strict], i[strict]), values[strict])
    x = np.linalg.solve(A, b)
    lam = np.linalg.solve(A, 2.0 * (x - t))
    gv = np.where(i == j, -(lam[i] * x[i]).sum(1),
                  np.where(i > j, -((lam[i] * x[j]).sum(1) + (lam[j] * x[i]).sum(1)), 0.0))
    return x, lam, gv


@pytest.mark.parametrize("device", DEVICES)
def test_gradient_vs_dense_fp64(device):
    n, values, idx, b, t = system(4)
    _, lam_ref, gv_ref = dense_reference(n, values, idx, b, t)
    ii = torch.tensor(idx[0], device=device)
    jj = torch.tensor(idx[1], device=device)
    vals = torch.tensor(values, device=device, requires_grad=True)
    solver = CholeskySolverD(n, ii, jj, vals.detach(), MatrixType.COO)
    rows, cols = coo_indices(n, ii, jj, MatrixType.COO)
    bt = torch.tensor(b, device=device, requires_grad=True)
    x = solve(solver, bt, vals, rows, cols)
    loss = ((x - torch.tensor(t, device=device)) ** 2).sum()
    loss.backward()
    gv = vals.grad.cpu().numpy()
    gb = bt.grad.cpu().numpy()
    assert np.abs(gv - gv_ref).max() / np.abs(gv_ref).max() < 1e-10
    assert np.abs(gb - lam_ref).max() / np.abs(lam_ref).max() < 1e-10
    assert np.all(gv[idx[0] < idx[1]] == 0)


@pytest.mark.parametrize("device", DEVICES)
def test_gradient_fp32_solver(device):
    """Single-precision solver: same gradient to FP32 accuracy."""
    n, values, idx, b, t = system(4)
    _, _, gv_ref = dense_reference(n, values, idx, b, t)
    ii = torch.tensor(idx[0], device=device)
    jj = torch.tensor(idx[1], device=device)
    vals = torch.tensor(values, device=device, requires_grad=True)
    solver = CholeskySolverF(n, ii, jj, vals.detach(), MatrixType.COO)
    rows, cols = coo_indices(n, ii, jj, MatrixType.COO)
    bt = torch.tensor(b, device=device, dtype=torch.float32)
    x = solve(solver, bt, vals, rows, cols)
    ((x - torch.tensor(t, device=device, dtype=torch.float32)) ** 2).sum().backward()
    gv = vals.grad.cpu().numpy()
    assert np.abs(gv - gv_ref).max() / np.abs(gv_ref).max() < 1e-4


@pytest.mark.parametrize("device", DEVICES)
def test_gradient_vs_central_differences(device):
    """>= 200 seeded stored entries (diagonal, below and above the diagonal,
    duplicates included) against FP64 central differences of the loss
    evaluated with SciPy's spsolve; step h = 1e-6 * max(1, |v_e|)."""
    n, values, idx, b, t = system(4)
    ii = torch.tensor(idx[0], device=device)
    jj = torch.tensor(idx[1], device=device)
    vals = torch.tensor(values, device=device, requires_grad=True)
    solver = CholeskySolverD(n, ii, jj, vals.detach(), MatrixType.COO)
    rows, cols = coo_indices(n, ii, jj, MatrixType.COO)
    x = solve(solver, torch.tensor(b, device=device), vals, rows, cols)
    ((x - torch.tensor(t, device=device)) ** 2).sum().backward()
    g = vals.grad.cpu().numpy()

    def loss(v):
        i, j = idx
        low = i >= j
        A = sp.coo_matrix((v[low], (i[low], j[low])), shape=(n, n)).tocsc()
        A = A + sp.tril(A, -1).T
        return float(((spsolve(A.tocsc(), b) - t) ** 2).sum())

    rng = np.random.default_rng(11)
    picks = rng.choice(len(values), size=200, replace=False)
    for e in picks:
        h = 1e-6 * max(1.0, abs(values[e]))
        vp, vm = values.copy(), values.copy()
        vp[e] += h
        vm[e] -= h
        fd = (loss(vp) - loss(vm)) / (2 * h)
        assert abs(fd - g[e]) <= 1e-6 * max(1.0, abs(g[e])) + 1e-7, (e, fd, g[e])


@pytest.mark.gpu
def test_full_size_gradient_fp32_vs_fp64():
    """The differentiation workload at full size: icosphere(6) (n = 40,962),
    lambda = 10, 3 right-hand sides, loss ||x - t||^2. The FP32 solver's
    gradients (b and every stored entry, HIP scatter) agree with the FP64
    solver's to FP32 accuracy, and the FP64 gradient of b is the adjoint
    solution (residual of A lam = 2 (x - t) at FP64 level)."""
    n, values, idx, b, t = system(6)
    ii = torch.tensor(idx[0], device="cuda")
    jj = torch.tensor(idx[1], device="cuda")
    rows, cols = coo_indices(n, ii, jj, MatrixType.COO)
    grads = {}
    for cls, dt in ((CholeskySolverF, torch.float32), (CholeskySolverD, torch.float64)):
        vals = torch.tensor(values, device="cuda", requires_grad=True)
        solver = cls(n, ii, jj, vals.detach(), MatrixType.COO)
        bt = torch.tensor(b, device="cuda", dtype=dt, requires_grad=True)
        x = solve(solver, bt, vals, rows, cols)
        ((x - torch.tensor(t, device="cuda", dtype=dt)) ** 2).sum().backward()
        torch.cuda.synchronize()
        solver._check()
        grads[dt] = (vals.grad.double().cpu().numpy(), bt.grad.double().cpu().numpy(), x.detach().double().cpu().numpy())
    gv32, gb32, _ = grads[torch.float32]
    gv64, gb64, x64 = grads[torch.float64]
    assert np.abs(gv32 - gv64).max() / np.abs(gv64).max() < 1e-4
    assert np.abs(gb32 - gb64).max() / np.abs(gb64).max() < 1e-4
    A = sp.coo_matrix((values, (idx[0], idx[1])), shape=(n, n)).tocsr()
    r = A @ gb64 - 2.0 * (x64 - t)
    assert np.linalg.norm(r) / np.linalg.norm(2.0 * (x64 - t)) < 1e-12


def test_csr_and_csc_indices():
    n, values, idx, b, t = system(2)
    A = sp.coo_matrix((values, (idx[0], idx[1])), shape=(n, n))
    R = A.tocsr()
    C = A.tocsc()
    r1, c1 = coo_indices(n, R.indptr, R.indices, MatrixType.CSR)
    Rc = R.tocoo()
    assert np.array_equal(r1.numpy(), Rc.row) and np.array_equal(c1.numpy(), Rc.col)
    r2, c2 = coo_indices(n, C.indptr, C.indices, MatrixType.CSC)
    Cc = C.tocoo()
    assert np.array_equal(r2.numpy(), Cc.row) and np.array_equal(c2.numpy(), Cc.col)
