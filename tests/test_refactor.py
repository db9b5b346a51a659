"""Numeric refactorisation (refactorize): new values for the constructor's
pattern reuse the ordering, the symbolic factor, the level analysis and the
block plan's structure. The result must equal, bit for bit, a new solver
built from the same indices and the new values (the ordering depends on the
pattern only), and solve the new system (checked against SciPy's FP64
sparse LU, independent of this build)."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import cholespy_amd
from cholespy_amd import CholeskySolverD, CholeskySolverF, MatrixType, workloads


def coo_with_noise(level, lam=2.0, seed=0):
    """Mesh Laplacian COO plus duplicate diagonal entries and entries above
    the diagonal (ignored: only the lower triangle is read)."""
    rng = np.random.default_rng(seed)
    n, f = workloads.icosphere(level)
    v, idx = workloads.mesh_laplacian(n, f, lam)
    dup = rng.integers(0, n, 50)
    up_r = rng.integers(0, n - 1, 20)
    up_c = up_r + 1 + rng.integers(0, 3, 20)
    up_c = np.minimum(up_c, n - 1)
    keep = up_c > up_r
    ii = np.concatenate([idx[0], dup, up_r[keep]])
    jj = np.concatenate([idx[1], dup, up_c[keep]])
    vv = np.concatenate([v, np.full(50, 0.25), rng.standard_normal(int(keep.sum())) * 10])
    return n, ii.astype(np.int32), jj.astype(np.int32), vv


def new_values(ii, jj, x, seed):
    """Values of another SPD matrix with the same pattern: the off-diagonal
    part scaled by a random factor per entry pair, the diagonal made
    dominant again."""
    rng = np.random.default_rng(seed)
    x1 = x.copy()
    off = ii != jj
    x1[off] *= rng.uniform(0.5, 1.5, int(off.sum()))
    n = int(max(ii.max(), jj.max())) + 1
    rowsum = np.zeros(n)
    lower_off = off & (ii > jj)
    np.add.at(rowsum, ii[lower_off], np.abs(x1[lower_off]))
    np.add.at(rowsum, jj[lower_off], np.abs(x1[lower_off]))
    diag = np.where(~off)[0]
    first = {}
    for e in diag:
        first.setdefault(int(ii[e]), e)
    for r, e in first.items():
        x1[e] = rowsum[r] + rng.uniform(0.1, 1.0)
    return x1


def lower_matrix(n, ii, jj, x):
    keep = ii >= jj
    L = sp.coo_matrix((x[keep], (ii[keep], jj[keep])), shape=(n, n)).tocsc()
    return (L + sp.tril(L, -1).T).tocsc()


@pytest.mark.parametrize("method", ["simplicial", "supernodal"])
@pytest.mark.parametrize("cls", [CholeskySolverF, CholeskySolverD])
def test_refactorize_equals_new_solver_cpu(cls, method, monkeypatch):
    monkeypatch.setenv("CHOLESPY_FACTOR", method)
    n, ii, jj, x0 = coo_with_noise(4)
    x1 = new_values(ii, jj, x0, 1)
    s = cls(n, ii, jj, x0, MatrixType.COO)
    s.refactorize(x1)
    fresh = cls(n, ii, jj, x1, MatrixType.COO)
    fa, fb = s._factor(), fresh._factor()
    for key in ("perm", "colptr", "rowidx", "parent", "val"):
        assert np.array_equal(fa[key], fb[key]), key
    assert np.array_equal(s._matrix()["val"], fresh._matrix()["val"])
    dt = np.float32 if cls is CholeskySolverF else np.float64
    b = np.random.default_rng(2).standard_normal((n, 3)).astype(dt)
    xa, xb = np.zeros_like(b), np.zeros_like(b)
    s.solve(b, xa)
    fresh.solve(b, xb)
    assert np.array_equal(xa, xb)
    # ... and it solves the new system (SciPy FP64 LU as the independent check).
    want = spla.spsolve(lower_matrix(n, ii, jj, x1), b.astype(np.float64))
    tol = 1e-5 if dt is np.float32 else 1e-11
    assert np.abs(xa - want).max() <= tol * np.abs(want).max()


@pytest.mark.parametrize("mt", ["CSR", "CSC"])
def test_refactorize_compressed_inputs(mt):
    n, ii, jj, x0 = coo_with_noise(3)
    x1 = new_values(ii, jj, x0, 4)
    A0 = sp.coo_matrix((x0, (ii, jj)), shape=(n, n))
    A1 = sp.coo_matrix((x1, (ii, jj)), shape=(n, n))
    conv = (lambda A: A.tocsr()) if mt == "CSR" else (lambda A: A.tocsc())
    M0, M1 = conv(A0), conv(A1)   # duplicates summed by SciPy, same pattern
    assert np.array_equal(M0.indptr, M1.indptr) and np.array_equal(M0.indices, M1.indices)
    kind = MatrixType.CSR if mt == "CSR" else MatrixType.CSC
    s = CholeskySolverD(n, M0.indptr, M0.indices, M0.data, kind)
    s.refactorize(M1.data)
    fresh = CholeskySolverD(n, M1.indptr, M1.indices, M1.data, kind)
    assert np.array_equal(s._factor()["val"], fresh._factor()["val"])


def test_refactorize_same_values_is_identity():
    n, ii, jj, x0 = coo_with_noise(3)
    s = CholeskySolverD(n, ii, jj, x0, MatrixType.COO)
    before = s._factor()["val"].copy()
    s.refactorize(x0)
    assert np.array_equal(s._factor()["val"], before)


@pytest.mark.parametrize("method", ["simplicial", "supernodal"])
def test_refactorize_not_positive_definite_keeps_solver(method, monkeypatch):
    monkeypatch.setenv("CHOLESPY_FACTOR", method)
    n, ii, jj, x0 = coo_with_noise(3)
    s = CholeskySolverD(n, ii, jj, x0, MatrixType.COO)
    f0 = s._factor()["val"].copy()
    bad = x0.copy()
    bad[(ii == jj) & (ii == n // 2)] = -5.0
    with pytest.raises(ValueError, match="Matrix is not positive definite!"):
        s.refactorize(bad)
    assert np.array_equal(s._factor()["val"], f0)


def test_refactorize_argument_errors():
    n, ii, jj, x0 = coo_with_noise(2)
    s = CholeskySolverD(n, ii, jj, x0, MatrixType.COO)
    with pytest.raises(ValueError, match="as many entries"):
        s.refactorize(x0[:-1])
    with pytest.raises(TypeError):
        s.refactorize(x0.reshape(1, -1))


@pytest.mark.gpu
@pytest.mark.parametrize("cls", [CholeskySolverF, CholeskySolverD])
def test_refactorize_gpu_equals_new_solver(cls):
    """On the GPU: levels unchanged, and both schedules' solves bit-identical
    to a new solver's (same plan structure, same values, deterministic
    reduction order)."""
    torch = pytest.importorskip("torch")
    n, ii, jj, x0 = coo_with_noise(5)
    x1 = new_values(ii, jj, x0, 7)
    dev = "cuda"
    t = lambda a: torch.tensor(a, device=dev)
    s = cls(n, t(ii), t(jj), t(x0), MatrixType.COO)
    dt = torch.float32 if cls is CholeskySolverF else torch.float64
    b = torch.randn(n, 16, device=dev, dtype=dt, generator=torch.Generator(dev).manual_seed(0))
    x = torch.zeros_like(b)
    s.solve(b, x)                       # the old values are in use before the refactorisation
    lv0 = s._levels(True)
    s.refactorize(t(x1))
    fresh = cls(n, t(ii), t(jj), t(x1), MatrixType.COO)
    assert np.array_equal(s._factor()["val"], fresh._factor()["val"])
    for lower in (True, False):
        assert all(np.array_equal(a, c) for a, c in zip(s._levels(lower), fresh._levels(lower)))
    assert all(np.array_equal(a, c) for a, c in zip(s._levels(True), lv0))
    for path in (cholespy_amd._core._PATH_BLOCKS, cholespy_amd._core._PATH_ROWS):
        s._set_path(path)
        fresh._set_path(path)
        for k in (1, 3, 16):
            bk = b[:, :k].contiguous()
            xa, xb = torch.zeros_like(bk), torch.zeros_like(bk)
            s.solve(bk, xa)
            fresh.solve(bk, xb)
            torch.cuda.synchronize()
            s._check()
            assert torch.equal(xa, xb), (path, k)
    want = spla.spsolve(lower_matrix(n, ii, jj, x1), b[:, :3].double().cpu().numpy())
    got = xa[:, :3].double().cpu().numpy()
    tol = 1e-5 if cls is CholeskySolverF else 1e-11
    assert np.abs(got - want).max() <= tol * np.abs(want).max()


@pytest.mark.gpu
def test_refactorize_gpu_refined_uses_new_matrix():
    """The FP64 residual of the refinement path uses the new A."""
    torch = pytest.importorskip("torch")
    n, ii, jj, x0 = coo_with_noise(5)
    x1 = new_values(ii, jj, x0, 11)
    t = lambda a: torch.tensor(a, device="cuda")
    s = CholeskySolverF(n, t(ii), t(jj), t(x0), MatrixType.COO)
    s.refactorize(x1)
    b = torch.randn(n, 3, device="cuda", dtype=torch.float64, generator=torch.Generator("cuda").manual_seed(1))
    x = torch.zeros_like(b)
    s._solve_refined(b, x, 2)
    torch.cuda.synchronize()
    s._check()
    want = spla.spsolve(lower_matrix(n, ii, jj, x1), b.cpu().numpy())
    got = x.cpu().numpy()
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()
