"""Device-side input normalisation (cholespy_normalize_input): GPU inputs to
the constructor are reduced to the lower CSC on the GPU. The result must
equal the oracle's restatement of the reference's host path
(oracle.normalized_lower_csc: src/cholesky_solver.cpp:8-149, sort and sum
duplicates; CHOLMOD stype = -1 keeps row >= col, 211-220): pattern bit-exact,
values to the rounding of a duplicate sum in another order. It must also be
bit-identical to this project's host normalisation of the same arrays
(CPU-mode solver built from them), with the same errors for invalid indices
and pointers (src/main.cpp:28-45 checks sizes before; the index checks are
the host normalisation's)."""
import numpy as np
import pytest
import scipy.sparse as sp

torch = pytest.importorskip("torch")

from cholespy_amd import CholeskySolverD, CholeskySolverF, MatrixType, workloads
from oracle import oracle

pytestmark = pytest.mark.gpu


def noisy_coo(level=3, seed=0):
    """Mesh Laplacian COO with extra duplicate entries (diagonal and
    off-diagonal, in both orders) and entries above the diagonal."""
    rng = np.random.default_rng(seed)
    n, f = workloads.icosphere(level)
    v, idx = workloads.mesh_laplacian(n, f, 2.0)
    ii, jj = idx[0].astype(np.int64), idx[1].astype(np.int64)
    pick = rng.integers(0, len(v), 300)
    ii = np.concatenate([ii, ii[pick], rng.integers(0, n - 1, 40)])
    jj = np.concatenate([jj, jj[pick], rng.integers(0, n - 1, 40)])
    vv = np.concatenate([v, rng.uniform(-0.01, 0.01, 300), rng.standard_normal(40)])
    # keep the matrix positive definite: upper-triangle noise is ignored by
    # the normalisation, lower-triangle noise stays small
    lower_noise = ii[-40:] > jj[-40:]
    vv[-40:][lower_noise] *= 1e-3
    return n, ii, jj, vv


def host_matrix(cls, n, ii, jj, vv, mt):
    s = cls(n, np.asarray(ii), np.asarray(jj), np.asarray(vv, dtype=np.float64), mt)
    return s._matrix()


def gpu(a, dtype=None):
    t = torch.as_tensor(np.asarray(a))
    if dtype is not None:
        t = t.to(dtype)
    return t.to("cuda")


def assert_same(got, want):
    for key in ("colptr", "rowidx", "val"):
        assert np.array_equal(got[key], want[key]), key


def assert_matches_oracle(got, n, ii, jj, vv, kind):
    colptr, rowidx, val = oracle.normalized_lower_csc(n, ii, jj, vv, kind)
    assert np.array_equal(got["colptr"], colptr)
    assert np.array_equal(got["rowidx"], rowidx)
    # duplicates are summed in input order here; SciPy may add them in
    # another order: agreement to a few units in the last place of the sum
    np.testing.assert_allclose(got["val"], val, rtol=1e-14, atol=1e-15 * np.abs(val).max())


@pytest.mark.parametrize("idx_dtype", [torch.int32, torch.int64])
@pytest.mark.parametrize("val_dtype", [torch.float32, torch.float64])
def test_coo_matches_host(idx_dtype, val_dtype):
    n, ii, jj, vv = noisy_coo()
    vcast = torch.as_tensor(vv).to(val_dtype).double().numpy()   # the values the GPU path sees
    want = host_matrix(CholeskySolverD, n, ii, jj, vcast, MatrixType.COO)
    s = CholeskySolverD(n, gpu(ii, idx_dtype), gpu(jj, idx_dtype), gpu(vv, val_dtype), MatrixType.COO)
    assert_same(s._matrix(), want)
    assert_matches_oracle(s._matrix(), n, ii, jj, vcast, "COO")


@pytest.mark.parametrize("kind", ["CSR", "CSC"])
@pytest.mark.parametrize("idx_dtype", [torch.int32, torch.int64])
def test_compressed_matches_host(kind, idx_dtype):
    n, ii, jj, vv = noisy_coo(seed=1)
    A = sp.coo_matrix((vv, (ii, jj)), shape=(n, n))
    M = A.tocsr() if kind == "CSR" else A.tocsc()
    M.sum_duplicates()
    # put duplicates back into one compressed row / column, unsorted
    mt = MatrixType.CSR if kind == "CSR" else MatrixType.CSC
    ptr, ind, val = M.indptr.copy(), M.indices.copy(), M.data.copy()
    seg = int(np.argmax(np.diff(ptr)))
    a, b = ptr[seg], ptr[seg + 1]
    ind[a:b] = ind[a:b][::-1].copy()
    val[a:b] = val[a:b][::-1].copy()
    want = host_matrix(CholeskySolverD, n, ptr, ind, val, mt)
    s = CholeskySolverF(n, gpu(ptr, idx_dtype), gpu(ind, idx_dtype), gpu(val), mt)
    assert_same(s._matrix(), want)
    assert_matches_oracle(s._matrix(), n, ptr, ind, val, kind)


@pytest.mark.parametrize("kind", ["COO", "CSR", "CSC"])
@pytest.mark.parametrize("val_dtype", [torch.float32, torch.float64])
def test_device_normalisation_matches_oracle(kind, val_dtype):
    """Straight against the oracle (not the host path): duplicates spread
    over the input, entries above the diagonal, int64 indices, both value
    widths, every matrix type."""
    n, ii, jj, vv = noisy_coo(level=2, seed=4)
    vv = torch.as_tensor(vv).to(val_dtype).double().numpy()
    if kind == "COO":
        a, b_, c = ii, jj, vv
    else:
        A = sp.coo_matrix((vv, (ii, jj)), shape=(n, n))
        # compressed form with duplicates kept (no sum_duplicates): rows /
        # columns sorted by their outer index, inner order as given
        outer, inner = (ii, jj) if kind == "CSR" else (jj, ii)
        order = np.argsort(outer, kind="stable")
        ptr = np.zeros(n + 1, dtype=np.int64)
        np.add.at(ptr, outer + 1, 1)
        a, b_, c = np.cumsum(ptr), inner[order], vv[order]
        del A
    mt = getattr(MatrixType, kind)
    s = CholeskySolverD(n, gpu(a, torch.int64), gpu(b_, torch.int64), gpu(c, val_dtype), mt)
    assert_matches_oracle(s._matrix(), n, a, b_, c, kind)


def test_solution_matches_host_path():
    """Same normalised matrix -> same factor -> same solve as a GPU solver
    built through the host copy path (strided inputs take that path)."""
    n, ii, jj, vv = noisy_coo(level=4, seed=2)
    dev = CholeskySolverD(n, gpu(ii), gpu(jj), gpu(vv), MatrixType.COO)
    # strided views: not eligible for the device path
    ii2 = torch.as_tensor(np.repeat(ii, 2)).cuda()[::2]
    jj2 = torch.as_tensor(np.repeat(jj, 2)).cuda()[::2]
    vv2 = torch.as_tensor(np.repeat(vv, 2)).cuda()[::2]
    assert not ii2.is_contiguous()
    hostpath = CholeskySolverD(n, ii2, jj2, vv2, MatrixType.COO)
    assert_same(dev._matrix(), hostpath._matrix())
    fa, fb = dev._factor(), hostpath._factor()
    assert np.array_equal(fa["val"], fb["val"])
    b = torch.rand(n, 3, device="cuda", dtype=torch.float64)
    x1, x2 = torch.zeros_like(b), torch.zeros_like(b)
    dev.solve(b, x1)
    hostpath.solve(b, x2)
    torch.cuda.synchronize()
    assert torch.equal(x1, x2)


def test_refactorize_uses_device_entry_map():
    n, ii, jj, vv = noisy_coo(seed=3)
    s = CholeskySolverD(n, gpu(ii), gpu(jj), gpu(vv), MatrixType.COO)
    v2 = vv.copy()
    v2[ii == jj] *= 1.5
    s.refactorize(gpu(v2))
    fresh = CholeskySolverD(n, ii, jj, v2, MatrixType.COO)   # CPU mode: host normalisation
    assert np.array_equal(s._factor()["val"], fresh._factor()["val"])


@pytest.mark.parametrize("case,msg", [
    ("coo_row", "Sparse matrix: row index out of range [0, n_rows)."),
    ("coo_col", "Sparse matrix: column index out of range [0, n_rows)."),
    ("coo_row_first", "Sparse matrix: row index out of range [0, n_rows)."),
    ("csr_end", "Sparse matrix: row pointer array must start at 0 and end at nnz."),
    ("csr_order", "Sparse matrix: row pointer array must be non-decreasing."),
    ("csr_col", "Sparse matrix: column index out of range [0, n_rows)."),
    ("csc_row", "Sparse matrix: row index out of range [0, n_rows)."),
])
def test_invalid_inputs_raise_like_host(case, msg):
    n = 50
    v, idx = workloads.grid2d_laplacian(10, 5, 1e-2)
    ii, jj, vv = idx[0].copy(), idx[1].copy(), v.copy()
    mt = MatrixType.COO
    if case == "coo_row":
        ii[7] = n
    elif case == "coo_col":
        jj[7] = -1
    elif case == "coo_row_first":
        jj[3] = n + 5      # earlier entry: column out of range
        ii[9] = n          # later entry: row out of range
        msg = "Sparse matrix: column index out of range [0, n_rows)."
    else:
        A = sp.coo_matrix((vv, (ii, jj)), shape=(n, n))
        M = A.tocsr() if case.startswith("csr") else A.tocsc()
        ii, jj, vv = M.indptr.copy(), M.indices.copy(), M.data.copy()
        mt = MatrixType.CSR if case.startswith("csr") else MatrixType.CSC
        if case == "csr_end":
            ii[-1] -= 1
        elif case == "csr_order":
            ii[10], ii[11] = ii[11], ii[10] - 1
        else:
            jj[4] = n
    with pytest.raises(ValueError) as e:
        CholeskySolverD(n, gpu(ii), gpu(jj), gpu(vv), mt)
    assert str(e.value) == msg
    with pytest.raises(ValueError) as e2:     # the host path says the same
        CholeskySolverD(n, ii, jj, vv, mt)
    assert str(e2.value) == msg
