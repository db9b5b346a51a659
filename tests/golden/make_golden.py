"""Writes tests/golden/*.npz: fixed factors with the expected outputs of the
solve phase on them. Run from the repository root:

    python tests/golden/make_golden.py

Each fixture stores a Cholesky factor L (column form, rows ascending,
diagonal first), its permutation, a seeded right-hand side and:
  level_lower / level_upper   level of every row (analysis_lower / _upper,
                              kernels/kernels.cu:8-60), computed here from the
                              elimination-tree closed forms (subtree height /
                              depth), NOT by the oracle it later checks;
  order_lower / order_upper   rows sorted by level, stable (numpy stable argsort);
  x64                         FP64 solution from SciPy's spsolve on A itself
                              (the reference tests' own oracle,
                              tests/test_cholesky.py:29-31).
The factors come from this project's host factorisation; CHOLMOD's factor
is not available, so the fixtures pin the solve phase on a fixed L, not the
factorisation.
"""
import sys
from pathlib import Path

import numpy as np
import scipy.sparse as sp
from scipy.sparse.linalg import spsolve

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from cholespy_amd import CholeskySolverD, MatrixType, workloads  # noqa: E402


def etree_levels(parent):
    n = len(parent)
    height = np.zeros(n, dtype=np.int32)
    for j in range(n):
        p = parent[j]
        if p >= 0:
            height[p] = max(height[p], height[j] + 1)
    depth = np.zeros(n, dtype=np.int32)
    for j in range(n - 1, -1, -1):
        if parent[j] >= 0:
            depth[j] = depth[parent[j]] + 1
    return height, depth


def make(name, n, values, idx, k, seed):
    solver = CholeskySolverD(n, idx[0], idx[1], values, MatrixType.COO)
    f = solver._factor()
    colptr, rowidx, val, perm = f["colptr"], f["rowidx"], f["val"], f["perm"]
    # Elimination tree from L itself: parent(j) = first off-diagonal row of column j.
    parent = np.full(n, -1, dtype=np.int64)
    has = np.diff(colptr) > 1
    parent[has] = rowidx[colptr[:-1][has] + 1]
    height, depth = etree_levels(parent)
    A = sp.coo_matrix((values, (idx[0], idx[1])), shape=(n, n)).tocsc()
    rng = np.random.default_rng(seed)
    b = rng.uniform(-1.0, 1.0, size=(n, k))
    x64 = spsolve(A, b).reshape(n, k)
    np.savez_compressed(
        ROOT / "tests" / "golden" / f"{name}.npz",
        n=n, perm=perm.astype(np.int32), colptr=colptr.astype(np.int32),
        rowidx=rowidx.astype(np.int32), val=val,
        coo_values=values, coo_idx=idx.astype(np.int32),
        level_lower=height, level_upper=depth,
        order_lower=np.argsort(height, kind="stable").astype(np.int32),
        order_upper=np.argsort(depth, kind="stable").astype(np.int32),
        b=b, x64=x64)
    print(name, n, len(val), "levels", height.max() + 1)


if __name__ == "__main__":
    make("cube", *((lambda n, f: (n,) + workloads.mesh_laplacian(n, f, 2.0))(*workloads.cube())), k=4, seed=0)
    make("ico3", *((lambda n, f: (n,) + workloads.mesh_laplacian(n, f, 2.0))(*workloads.icosphere(3))), k=3, seed=1)
    make("grid64", 4096, *workloads.grid2d_laplacian(64, 64, 1e-2), k=1, seed=2)
