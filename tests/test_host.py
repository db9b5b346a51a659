"""Host side of the drop-in: input normalisation, the host factor, the CPU
solve mode and the Python boundary's validation (errors and messages follow
src/main.cpp:28-106 and src/cholesky_solver.cpp:241-242, 419-420)."""
import numpy as np
import pytest
import scipy.sparse as sp

from cholespy_amd import CholeskySolverD, CholeskySolverF, MatrixType, workloads
from oracle import oracle


def mesh(level):
    n, f = workloads.icosphere(level) if level is not None else workloads.cube()
    v, idx = workloads.mesh_laplacian(n, f, 2.0)
    return n, v, idx


def test_normalisation_matches_oracle():
    """COO with duplicates and entries above the diagonal (ignored, CHOLMOD
    stype = -1), CSR and CSC all give the oracle's lower CSC."""
    rng = np.random.default_rng(3)
    n, v, idx = mesh(2)
    # add strictly-upper noise that must be ignored
    up = np.array([[0, 1, 2], [5, 9, 7]])
    vv = np.concatenate([v, rng.standard_normal(3) * 100])
    ii = np.concatenate([idx[0], up[0]])
    jj = np.concatenate([idx[1], up[1]])
    want = oracle.normalized_lower_csc(n, ii, jj, vv, "COO")
    s = CholeskySolverD(n, ii, jj, vv, MatrixType.COO)
    got = s._matrix()
    assert np.array_equal(got["colptr"], want[0])
    assert np.array_equal(got["rowidx"], want[1])
    assert np.allclose(got["val"], want[2], rtol=0, atol=1e-15)
    A = sp.coo_matrix((v, (idx[0], idx[1])), shape=(n, n))
    for kind, M in (("CSR", A.tocsr()), ("CSC", A.tocsc())):
        mt = MatrixType.CSR if kind == "CSR" else MatrixType.CSC
        s2 = CholeskySolverD(n, M.indptr, M.indices, M.data, mt)
        g2 = s2._matrix()
        w2 = oracle.normalized_lower_csc(n, M.indptr, M.indices, M.data, kind)
        assert np.array_equal(g2["colptr"], w2[0]) and np.array_equal(g2["rowidx"], w2[1])
        assert np.allclose(g2["val"], w2[2])


@pytest.mark.parametrize("level", [None, 2, 4])
def test_factor_reproduces_permuted_matrix(level):
    """L L' = P A P' to rounding, L lower with a positive diagonal first in
    each column, perm a permutation, elimination tree postordered."""
    n, v, idx = mesh(level)
    s = CholeskySolverD(n, idx[0], idx[1], v, MatrixType.COO)
    f = s._factor()
    perm = f["perm"]
    assert np.array_equal(np.sort(perm), np.arange(n))
    L = sp.csc_matrix((f["val"], f["rowidx"], f["colptr"]), shape=(n, n))
    assert np.all(f["rowidx"][f["colptr"][:-1]] == np.arange(n))
    assert np.all(f["val"][f["colptr"][:-1]] > 0)
    assert sp.triu(L, 1).nnz == 0
    A = sp.coo_matrix((v, (idx[0], idx[1])), shape=(n, n)).tocsc()
    PAP = A[perm][:, perm]
    err = abs(PAP - L @ L.T).max() / abs(A).max()
    assert err < 1e-14
    parent = f["parent"]
    assert np.all((parent == -1) | (parent > np.arange(n)))
    # postorder: every subtree occupies a contiguous index range ending at its root
    size = np.ones(n, dtype=np.int64)
    for j in range(n):
        if parent[j] >= 0:
            size[parent[j]] += size[j]
    for j in range(n):
        if parent[j] >= 0:
            p = parent[j]
            assert p - size[p] + 1 <= j - size[j] + 1


def test_factor_matches_dense_cholesky():
    """For the fixed P the positive-diagonal Cholesky factor is unique:
    compare with numpy.linalg.cholesky of P A P'."""
    n, v, idx = mesh(1)
    s = CholeskySolverD(n, idx[0], idx[1], v, MatrixType.COO)
    f = s._factor()
    A = sp.coo_matrix((v, (idx[0], idx[1])), shape=(n, n)).toarray()
    P = f["perm"]
    Ld = np.linalg.cholesky(A[np.ix_(P, P)])
    L = sp.csc_matrix((f["val"], f["rowidx"], f["colptr"]), shape=(n, n)).toarray()
    assert np.allclose(L, Ld, rtol=1e-13, atol=1e-14)


@pytest.mark.parametrize("method", ["simplicial", "supernodal"])
@pytest.mark.parametrize("case", ["ico4", "grid", "cube"])
def test_numeric_methods_agree(case, method, monkeypatch):
    """The supernodal multifrontal numeric phase and the simplicial up-looking
    one give the same pattern of L (fundamental supernodes add no entries)
    and values equal to rounding, and both reproduce P A P' = L L'."""
    if case == "ico4":
        n, v, idx = mesh(4)
    elif case == "grid":
        n = 48 * 40
        v, idx = workloads.grid2d_laplacian(48, 40, 1e-2)
    else:
        n = 14 ** 3
        v, idx = workloads.grid3d_laplacian(14, 1e-3)
    monkeypatch.setenv("CHOLESPY_FACTOR", "simplicial")
    ref = CholeskySolverD(n, idx[0], idx[1], v, MatrixType.COO)._factor()
    monkeypatch.setenv("CHOLESPY_FACTOR", method)
    f = CholeskySolverD(n, idx[0], idx[1], v, MatrixType.COO)._factor()
    for key in ("perm", "colptr", "rowidx", "parent"):
        assert np.array_equal(f[key], ref[key]), key
    assert np.abs(f["val"] - ref["val"]).max() <= 1e-12 * np.abs(ref["val"]).max()
    L = sp.csc_matrix((f["val"], f["rowidx"], f["colptr"]), shape=(n, n))
    A = sp.coo_matrix((v, (idx[0], idx[1])), shape=(n, n)).tocsc()
    P = f["perm"]
    assert abs(A[P][:, P] - L @ L.T).max() / abs(A).max() < 1e-13


@pytest.mark.parametrize("method", ["simplicial", "supernodal"])
def test_not_positive_definite_large(method, monkeypatch):
    """An indefinite matrix large enough for dense fronts raises the same
    error from either numeric phase."""
    monkeypatch.setenv("CHOLESPY_FACTOR", method)
    n = 40 * 40
    v, idx = workloads.grid2d_laplacian(40, 40, 1e-2)
    v = v.copy()
    v[idx[0] == idx[1]] -= 2.5   # shifts the spectrum below zero
    with pytest.raises(ValueError, match="Matrix is not positive definite!"):
        CholeskySolverD(n, idx[0], idx[1], v, MatrixType.COO)


def test_not_positive_definite():
    n = 3
    ii = np.array([0, 1, 2, 1])
    jj = np.array([0, 1, 2, 0])
    x = np.array([1.0, 1.0, 1.0, 2.0])   # [[1,2],[2,1]] block is indefinite
    with pytest.raises(ValueError, match="Matrix is not positive definite!"):
        CholeskySolverD(n, ii, jj, x, MatrixType.COO)


@pytest.mark.parametrize("cls,dtype", [(CholeskySolverF, np.float32), (CholeskySolverD, np.float64)])
def test_cpu_solve_matches_host_oracle(cls, dtype):
    n, v, idx = mesh(4)
    s = cls(n, idx[0], idx[1], v, MatrixType.COO)
    b = np.random.default_rng(0).random((n, 5)).astype(dtype)
    x = np.zeros_like(b)
    s.solve(b, x)
    ref = oracle.host_solve(s._factor(), b.astype(np.float64)).astype(dtype)
    assert np.array_equal(x, ref)   # same FP64 algorithm, cast once


def test_int64_indices_and_float32_values_accepted():
    """nanobind's implicit conversion of the constructor arrays."""
    n, v, idx = mesh(2)
    s1 = CholeskySolverD(n, idx[0].astype(np.int64), idx[1].astype(np.int64), v, MatrixType.COO)
    s2 = CholeskySolverD(n, idx[0].astype(np.int32), idx[1].astype(np.int32), v.astype(np.float32),
                         MatrixType.COO)
    s3 = CholeskySolverD(n, list(idx[0]), list(idx[1]), list(v), MatrixType.COO)
    b = np.ones(n)
    xs = []
    for s in (s1, s2, s3):
        x = np.zeros(n)
        s.solve(b, x)
        xs.append(x)
    assert np.array_equal(xs[0], xs[2])
    assert np.allclose(xs[0], xs[1], rtol=1e-6)


def test_strided_constructor_arrays():
    n, v, idx = mesh(2)
    big = np.zeros((2, len(v) * 2), dtype=np.int64)
    big[:, ::2] = idx
    vb = np.zeros(len(v) * 3)
    vb[::3] = v
    s = CholeskySolverD(n, big[0, ::2], big[1, ::2], vb[::3], MatrixType.COO)
    s_ref = CholeskySolverD(n, idx[0], idx[1], v, MatrixType.COO)
    b = np.arange(n, dtype=np.float64)
    x1, x2 = np.zeros(n), np.zeros(n)
    s.solve(b, x1)
    s_ref.solve(b, x2)
    assert np.array_equal(x1, x2)


# --- validation messages (src/main.cpp:28-45) ------------------------------------------
def test_constructor_shape_errors():
    n, v, idx = mesh(None)
    with pytest.raises(ValueError, match="Sparse COO matrix: the two index arrays should have the same size."):
        CholeskySolverF(n, idx[0], idx[1][:-1], v, MatrixType.COO)
    with pytest.raises(ValueError, match="Sparse COO matrix: the index and data arrays should have the same size."):
        CholeskySolverF(n, idx[0], idx[1], v[:-1], MatrixType.COO)
    A = sp.coo_matrix((v, (idx[0], idx[1])), shape=(n, n))
    R, C = A.tocsr(), A.tocsc()
    with pytest.raises(ValueError, match="Sparse CSR matrix: the column index and data arrays should have the same size."):
        CholeskySolverF(n, R.indptr, R.indices, R.data[:-1], MatrixType.CSR)
    with pytest.raises(ValueError, match="Sparse CSR matrix: Invalid size for row pointer array."):
        CholeskySolverF(n, R.indptr[:-1], R.indices, R.data, MatrixType.CSR)
    with pytest.raises(ValueError, match="Sparse CSC matrix: the row index and data arrays should have the same size."):
        CholeskySolverF(n, C.indptr, C.indices, C.data[:-1], MatrixType.CSC)
    with pytest.raises(ValueError, match="Sparse CSC matrix: Invalid size for column pointer array."):
        CholeskySolverF(n, C.indptr[:-1], C.indices, C.data, MatrixType.CSC)


def test_constructor_index_out_of_range():
    with pytest.raises(ValueError, match="out of range"):
        CholeskySolverD(3, np.array([0, 1, 3]), np.array([0, 1, 2]), np.ones(3), MatrixType.COO)


def test_solve_errors():
    n, v, idx = mesh(None)
    s = CholeskySolverF(n, idx[0], idx[1], v, MatrixType.COO)
    b = np.ones((n, 2), dtype=np.float32)
    with pytest.raises(ValueError, match="Expected 1D or 2D tensors as input."):
        s.solve(np.ones((n, 2, 2), dtype=np.float32), np.ones((n, 2, 2), dtype=np.float32))
    with pytest.raises(ValueError, match="x and b should have the same dimensions."):
        s.solve(b, np.zeros((n, 3), dtype=np.float32))
    with pytest.raises(ValueError, match="x and b should have the same dimensions."):
        s.solve(b, np.zeros((n + 1, 2), dtype=np.float32))
    with pytest.raises(TypeError):   # noconvert: wrong dtype
        s.solve(b.astype(np.float64), np.zeros((n, 2)))
    with pytest.raises(TypeError):   # noconvert: non-contiguous
        s.solve(np.ones((n, 4), dtype=np.float32)[:, ::2], np.zeros((n, 2), dtype=np.float32))
    with pytest.raises(ValueError, match="b and x should have n_rows rows."):
        s.solve(np.ones(n + 1, dtype=np.float32), np.zeros(n + 1, dtype=np.float32))


def test_torch_cpu_tensors():
    torch = pytest.importorskip("torch")
    n, v, idx = mesh(3)
    s = CholeskySolverF(n, torch.tensor(idx[0]), torch.tensor(idx[1]), torch.tensor(v), MatrixType.COO)
    b = torch.rand(n, 3)
    x = torch.zeros_like(b)
    s.solve(b, x)
    A = sp.coo_matrix((v, (idx[0], idx[1])), shape=(n, n)).tocsc()
    assert np.allclose(x.numpy(), sp.linalg.spsolve(A, b.numpy().astype(np.float64)), rtol=1e-5, atol=1e-7)


def test_readonly_b_accepted():
    n, v, idx = mesh(None)
    s = CholeskySolverD(n, idx[0], idx[1], v, MatrixType.COO)
    b = np.ones(n)
    b.setflags(write=False)
    x = np.zeros(n)
    s.solve(b, x)
    assert np.all(np.isfinite(x))
    with pytest.raises(TypeError):
        s.solve(np.ones(n), b)   # x must be writable


def test_module_surface():
    import cholespy
    import cholespy_amd
    assert cholespy.CholeskySolverF is cholespy_amd.CholeskySolverF
    assert cholespy.__version__ == "2.2.2"
    assert int(cholespy.MatrixType.CSC) == 0 and int(cholespy.MatrixType.CSR) == 1
    assert int(cholespy.MatrixType.COO) == 2
