"""Pins the oracle (oracle/oracle.c, oracle/oracle.py) before it is trusted.

Each oracle function is checked against something independent of its own
formula: the golden fixtures (levels from elimination-tree closed forms,
solutions from SciPy's spsolve on A), SciPy's triangular solver on the same
L, brute force on tiny inputs, and invariants. One test plants a mistake in
a copy of the oracle's level function and shows the checks catch it.
"""
from pathlib import Path

import numpy as np
import pytest
import scipy.sparse as sp
from scipy.sparse.linalg import spsolve, spsolve_triangular

from oracle import oracle

GOLDEN = Path(__file__).resolve().parent / "golden"
FIXTURES = ["cube", "ico3", "grid64"]


def load(name):
    with np.load(GOLDEN / f"{name}.npz", allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def factor_of(g):
    return {"perm": g["perm"], "colptr": g["colptr"], "rowidx": g["rowidx"], "val": g["val"]}


def lower_matrix(g):
    n = int(g["n"])
    return sp.csc_matrix((g["val"], g["rowidx"], g["colptr"]), shape=(n, n))


@pytest.mark.parametrize("name", FIXTURES)
def test_levels_match_closed_forms(name):
    """analysis_lower == subtree height, analysis_upper == depth in the
    elimination tree (closed forms; kernels/kernels.cu:8-60)."""
    g = load(name)
    rowptr, cols, _ = oracle.factor_rows(g["colptr"], g["rowidx"], g["val"])
    lo, mx_lo = oracle.levels(rowptr, cols, True)
    up, mx_up = oracle.levels(g["colptr"], g["rowidx"], False)
    assert np.array_equal(lo, g["level_lower"])
    assert np.array_equal(up, g["level_upper"])
    assert mx_lo == mx_up == g["level_lower"].max()   # both = etree height
    assert np.array_equal(oracle.level_sort(lo, mx_lo), g["order_lower"])
    assert np.array_equal(oracle.level_sort(up, mx_up), g["order_upper"])


@pytest.mark.parametrize("name", FIXTURES)
def test_level_invariants(name):
    """Every dependency of a row is in a strictly lower level; the sorted
    order is a permutation with non-decreasing levels, stable in row index."""
    g = load(name)
    n = int(g["n"])
    rowptr, cols, _ = oracle.factor_rows(g["colptr"], g["rowidx"], g["val"])
    for lower, rp, cl in ((True, rowptr, cols), (False, g["colptr"], g["rowidx"])):
        lv, mx = oracle.levels(rp, cl, lower)
        rows = np.repeat(np.arange(n), np.diff(rp))
        off = rows != cl
        assert np.all(lv[cl[off]] < lv[rows[off]])
        order = oracle.level_sort(lv, mx)
        assert np.array_equal(np.sort(order), np.arange(n))
        assert np.all(np.diff(lv[order]) >= 0)
        same = np.diff(lv[order]) == 0
        assert np.all(np.diff(order)[same] > 0)


def test_planted_level_mistake_is_caught():
    """A plausible mistake (the reference's `<=` written as `<`, i.e. taking
    max(level) instead of max(level + 1)) fails the closed-form check."""
    g = load("ico3")
    rowptr, cols, _ = oracle.factor_rows(g["colptr"], g["rowidx"], g["val"])
    n = int(g["n"])
    wrong = np.zeros(n, dtype=np.int64)
    for r in range(n):
        lvl = 0
        for i in range(rowptr[r], rowptr[r + 1] - 1):
            lvl = max(lvl, wrong[cols[i]])   # missing + 1
        wrong[r] = lvl
    assert not np.array_equal(wrong, g["level_lower"])


@pytest.mark.parametrize("name", FIXTURES)
def test_f64_solves_match_scipy(name):
    """Forward / backward passes in FP64 against SciPy's triangular solver on
    the same L, and the full solve against spsolve(A, b) from the fixture."""
    g = load(name)
    L = lower_matrix(g)
    perm = g["perm"]
    b = g["b"]
    rowptr, cols, rvals = oracle.factor_rows(g["colptr"], g["rowidx"], g["val"])
    y = oracle.forward(rowptr, cols, rvals, perm, b, np.float64)
    y_ref = spsolve_triangular(L.tocsr(), b[perm], lower=True)
    assert np.allclose(y, y_ref.reshape(y.shape), rtol=1e-12, atol=1e-14)
    x = oracle.backward(g["colptr"], g["rowidx"], g["val"], perm, y, np.float64)
    z_ref = spsolve_triangular(L.T.tocsr(), y_ref, lower=False)
    x_ref = np.empty_like(z_ref)
    x_ref[perm] = z_ref
    assert np.allclose(x, x_ref.reshape(x.shape), rtol=1e-12, atol=1e-14)
    assert np.abs(x - g["x64"]).max() / np.abs(g["x64"]).max() < 1e-12


@pytest.mark.parametrize("name", FIXTURES)
def test_host_solve_matches_spsolve(name):
    g = load(name)
    x = oracle.host_solve(factor_of(g), g["b"])
    assert np.abs(x - g["x64"]).max() / np.abs(g["x64"]).max() < 1e-12


@pytest.mark.parametrize("name", FIXTURES)
def test_f32_solve_accuracy(name):
    """FP32 restatement: within the FP32 error bound of the FP64 solution
    (relative error ~ condition number x 2^-24)."""
    g = load(name)
    x32 = oracle.solve(factor_of(g), g["b"].astype(np.float32), np.float32)
    assert x32.dtype == np.float32
    err = np.abs(x32 - g["x64"]).max() / np.abs(g["x64"]).max()
    assert err < 1e-4
    assert np.allclose(x32, g["x64"], rtol=1e-4, atol=1e-6)


def test_f32_identity_is_exact():
    """README example (README.md:104-121): A = I gives x = b exactly."""
    n = 20
    fac = {"perm": np.arange(n), "colptr": np.arange(n + 1), "rowidx": np.arange(n),
           "val": np.ones(n)}
    b = np.random.default_rng(0).random((n, 3)).astype(np.float32)
    assert np.array_equal(oracle.solve(fac, b, np.float32), b)


def test_f32_diagonal_single_division():
    """With a diagonal factor each result is one IEEE division b / d."""
    n = 50
    rng = np.random.default_rng(1)
    d = rng.uniform(0.5, 2.0, n)
    fac = {"perm": rng.permutation(n), "colptr": np.arange(n + 1), "rowidx": np.arange(n), "val": d}
    b = rng.random((n, 2)).astype(np.float32)
    x = oracle.solve(fac, b, np.float32)
    d32 = d.astype(np.float32)
    p = fac["perm"]
    expect = np.empty_like(b)
    expect[p] = (b[p] / d32[:, None]) / d32[:, None]
    assert np.array_equal(x, expect)


def test_normalized_lower_csc_brute_force():
    """Input semantics (duplicates summed, lower triangle only) against a
    dense brute-force accumulation on random tiny COO / CSR / CSC inputs."""
    rng = np.random.default_rng(7)
    for trial in range(20):
        n = int(rng.integers(1, 9))
        nnz = int(rng.integers(0, 30))
        i = rng.integers(0, n, nnz)
        j = rng.integers(0, n, nnz)
        v = rng.standard_normal(nnz)
        dense = np.zeros((n, n))
        np.add.at(dense, (i, j), v)
        want = np.tril(dense)
        for kind in ("COO", "CSR", "CSC"):
            if kind == "COO":
                args = (i, j, v)
            elif kind == "CSR":
                M = sp.coo_matrix((v, (i, j)), shape=(n, n)).tocsr()
                M.has_canonical_format = False
                args = (M.indptr, M.indices, M.data)
            else:
                M = sp.coo_matrix((v, (i, j)), shape=(n, n)).tocsc()
                args = (M.indptr, M.indices, M.data)
            cp, ri, val = oracle.normalized_lower_csc(n, *args, kind)
            got = sp.csc_matrix((val, ri, cp), shape=(n, n)).toarray()
            assert np.allclose(got, want)
            assert np.all(ri >= np.repeat(np.arange(n), np.diff(cp)))   # lower only


def test_spsolve_end_to_end_small():
    """Whole solve on the tiny cube system against a dense solve."""
    g = load("cube")
    n = int(g["n"])
    A = sp.coo_matrix((g["coo_values"], (g["coo_idx"][0], g["coo_idx"][1])), shape=(n, n)).toarray()
    x = oracle.host_solve(factor_of(g), g["b"])
    assert np.allclose(x, np.linalg.solve(A, g["b"]), rtol=1e-13, atol=1e-14)


@pytest.mark.parametrize("name", ["ico3", "grid64"])
def test_parallel_host_solve_is_the_same_solve(name):
    """The all-cores CPU baseline shares the columns out over threads; every
    column is the same sequential solve, so the result is bit-identical."""
    g = load(name)
    rng = np.random.default_rng(9)
    b = rng.standard_normal((int(g["n"]), 5))
    f = factor_of(g)
    assert np.array_equal(oracle.host_solve_parallel(f, b, 3), oracle.host_solve(f, b))
