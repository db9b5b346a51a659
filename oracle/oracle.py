"""ORACLE -- test infrastructure only.

Python side of the CPU restatement of the reference's solve phase
(oracle/oracle.c, loaded with ctypes) plus NumPy/SciPy restatements of the
host input semantics. Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg may import this module, and only as the checker / the
timed CPU baseline; the product never routes through it.

Reference routines restated here:
  normalized_lower_csc   coo_to_csc / csr_to_csc / csc_sort_indices /
                         csc_sum_duplicates + CHOLMOD stype = -1
                         (src/cholesky_solver.cpp:8-149, 211-220)
  levels / level_sort    kernels/kernels.cu:8-60, src/cholesky_solver.cpp:348-367
  forward / backward     kernels/kernels.cu:64-202
  host_solve             cholmod_solve(CHOLMOD_A), src/cholesky_solver.cpp:434-463
  etree_height_depth     closed forms of the two level functions (used to
                         pin `levels`; not a reference routine)
"""
from __future__ import annotations

import ctypes
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"

_lib = None


def build() -> Path:
    """Compiles oracle.c into oracle/liboracle.so (gcc, no fast-math)."""
    cmd = ["gcc", "-O2", "-fPIC", "-shared", "-std=c99", "-ffp-contract=off",
           str(HERE / "oracle.c"), "-o", str(LIB), "-lm"]
    subprocess.run(cmd, check=True)
    return LIB


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            raise RuntimeError(f"oracle library not built: {LIB} (run oracle.build())")
        L = ctypes.CDLL(str(LIB))
        P = ctypes.c_void_p
        i = ctypes.c_int
        L.oracle_levels.argtypes = [i, P, P, i, P]
        L.oracle_levels.restype = i
        L.oracle_level_sort.argtypes = [i, P, i, P]
        for name in ("oracle_lower_f", "oracle_lower_d", "oracle_upper_f", "oracle_upper_d"):
            getattr(L, name).argtypes = [i, i, P, P, P, P, P, P]
        L.oracle_host_solve.argtypes = [i, i, P, P, P, P, P]
        _lib = L
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def _i32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.int32)


# ---------------------------------------------------------------- factor forms
def factor_rows(colptr, rowidx, val):
    """Row-compressed L (rows ascending, diagonal last) from column form, the
    role cholmod_transpose plays at src/cholesky_solver.cpp:252."""
    n = len(colptr) - 1
    nnz = int(colptr[-1])
    cols = np.repeat(np.arange(n, dtype=np.int64), np.diff(colptr))
    order = np.lexsort((cols, np.asarray(rowidx)))
    rowptr = np.zeros(n + 1, dtype=np.int64)
    np.add.at(rowptr, np.asarray(rowidx, dtype=np.int64) + 1, 1)
    rowptr = np.cumsum(rowptr)
    assert rowptr[-1] == nnz
    return _i32(rowptr), _i32(cols[order]), np.asarray(val, dtype=np.float64)[order]


# ---------------------------------------------------------------- analysis
def levels(rowptr, cols, lower: bool):
    """Per-row levels and max level (kernels/kernels.cu:8-60)."""
    rowptr, cols = _i32(rowptr), _i32(cols)
    n = len(rowptr) - 1
    out = np.zeros(n, dtype=np.int32)
    mx = lib().oracle_levels(n, _p(rowptr), _p(cols), 1 if lower else 0, _p(out))
    return out, int(mx)


def level_sort(row_level, max_lvl: int):
    row_level = _i32(row_level)
    out = np.zeros(len(row_level), dtype=np.int32)
    lib().oracle_level_sort(len(row_level), _p(row_level), int(max_lvl), _p(out))
    return out


def etree_height_depth(parent):
    """(subtree height, depth) of every node of an elimination tree given by
    parent[] with parent[j] > j (roots: -1). Leaf height 0, root depth 0."""
    parent = np.asarray(parent)
    n = len(parent)
    height = np.zeros(n, dtype=np.int64)
    for j in range(n):   # children precede parents
        p = parent[j]
        if p >= 0 and height[p] < height[j] + 1:
            height[p] = height[j] + 1
    depth = np.zeros(n, dtype=np.int64)
    for j in range(n - 1, -1, -1):
        p = parent[j]
        if p >= 0:
            depth[j] = depth[p] + 1
    return height, depth


# ---------------------------------------------------------------- solves
def forward(rowptr, cols, vals, perm, b, dtype):
    """tmp = L^{-1} P b in `dtype` (row-major (n, k))."""
    dtype = np.dtype(dtype)
    b2 = np.ascontiguousarray(b, dtype=dtype).reshape(len(perm), -1)
    n, k = b2.shape
    v = np.ascontiguousarray(vals, dtype=dtype)
    tmp = np.zeros((n, k), dtype=dtype)
    fn = lib().oracle_lower_f if dtype == np.float32 else lib().oracle_lower_d
    rowptr, cols, perm = _i32(rowptr), _i32(cols), _i32(perm)
    fn(n, k, _p(rowptr), _p(cols), _p(v), _p(perm), _p(b2), _p(tmp))
    return tmp


def backward(colptr, rowidx, vals, perm, tmp, dtype):
    """x with x[perm[r]] = (L'^{-1} tmp)[r] in `dtype`."""
    dtype = np.dtype(dtype)
    t2 = np.array(tmp, dtype=dtype, copy=True).reshape(len(perm), -1)
    n, k = t2.shape
    v = np.ascontiguousarray(vals, dtype=dtype)
    x = np.zeros((n, k), dtype=dtype)
    fn = lib().oracle_upper_f if dtype == np.float32 else lib().oracle_upper_d
    colptr, rowidx, perm = _i32(colptr), _i32(rowidx), _i32(perm)
    fn(n, k, _p(colptr), _p(rowidx), _p(v), _p(perm), _p(t2), _p(x))
    return x


def solve(factor: dict, b, dtype):
    """Both passes of the reference GPU solve, restated, on the given factor
    (dict with perm, colptr, rowidx, val). Returns x with b's shape."""
    b = np.asarray(b)
    rowptr, cols, rvals = factor_rows(factor["colptr"], factor["rowidx"], factor["val"])
    y = forward(rowptr, cols, rvals, factor["perm"], b, dtype)
    x = backward(factor["colptr"], factor["rowidx"], factor["val"], factor["perm"], y, dtype)
    return x.reshape(b.shape)


def host_solve(factor: dict, b):
    """FP64 solve as cholmod_solve(CHOLMOD_A) does it for a simplicial LL'
    factor, with the reference's row-major <-> column-major conversions
    (src/cholesky_solver.cpp:449-460). Returns float64 with b's shape."""
    b = np.asarray(b, dtype=np.float64)
    n = len(factor["perm"])
    b2 = b.reshape(n, -1)
    k = b2.shape[1]
    bx = np.asfortranarray(b2).copy(order="F")
    colptr, rowidx, perm = _i32(factor["colptr"]), _i32(factor["rowidx"]), _i32(factor["perm"])
    val = np.ascontiguousarray(factor["val"], dtype=np.float64)
    lib().oracle_host_solve(n, k, _p(colptr), _p(rowidx), _p(val), _p(perm), _p(bx))
    return np.ascontiguousarray(bx).reshape(b.shape)


def host_solve_parallel(factor: dict, b, threads: int):
    """host_solve with the right-hand-side columns shared out over `threads`
    host threads (each column is an independent solve; ctypes releases the
    GIL for the C call). Same result as host_solve, bit for bit."""
    from concurrent.futures import ThreadPoolExecutor
    b = np.asarray(b, dtype=np.float64)
    n = len(factor["perm"])
    b2 = b.reshape(n, -1)
    k = b2.shape[1]
    bx = np.asfortranarray(b2).copy(order="F")
    colptr, rowidx, perm = _i32(factor["colptr"]), _i32(factor["rowidx"]), _i32(factor["perm"])
    val = np.ascontiguousarray(factor["val"], dtype=np.float64)
    L = lib()
    base = bx.ctypes.data
    cols = [list(range(t, k, threads)) for t in range(min(threads, k))]

    def run(ts):
        for t in ts:
            L.oracle_host_solve(n, 1, _p(colptr), _p(rowidx), _p(val), _p(perm),
                                ctypes.c_void_p(base + 8 * n * t))

    with ThreadPoolExecutor(max_workers=len(cols)) as ex:
        list(ex.map(run, cols))
    return np.ascontiguousarray(bx).reshape(b.shape)


# ---------------------------------------------------------------- input semantics
def normalized_lower_csc(n, ii, jj, x, kind: str):
    """(colptr, rowidx, val) of the lower triangle of A as CHOLMOD reads it
    (stype = -1), duplicates summed, rows sorted; kind in {COO, CSR, CSC}."""
    import scipy.sparse as sp
    x = np.asarray(x, dtype=np.float64)
    if kind == "COO":
        A = sp.coo_matrix((x, (np.asarray(ii), np.asarray(jj))), shape=(n, n))
    elif kind == "CSR":
        A = sp.csr_matrix((x, np.asarray(jj), np.asarray(ii)), shape=(n, n))
    else:
        A = sp.csc_matrix((x, np.asarray(jj), np.asarray(ii)), shape=(n, n))
    A = A.tocoo()
    keep = A.row >= A.col
    L = sp.coo_matrix((A.data[keep], (A.row[keep], A.col[keep])), shape=(n, n)).tocsc()
    L.sum_duplicates()
    L.sort_indices()
    return L.indptr.astype(np.int64), L.indices.astype(np.int64), L.data


def symmetric_from_lower(n, colptr, rowidx, val):
    """Full symmetric scipy matrix from its lower CSC."""
    import scipy.sparse as sp
    L = sp.csc_matrix((val, rowidx, colptr), shape=(n, n))
    return (L + sp.tril(L, -1).T).tocsc()
