"""Synthetic sparse SPD systems for tests and benchmarks.

Meshes and their Laplacian systems reproduce the reference test matrices
entry for entry (tests/utils.py:3-97): the same icosphere subdivision with
the same vertex numbering, and the same COO layout A = I + lam * (D - W)
with duplicated diagonal entries (one +lam per neighbour, plus 1). The
icosphere additionally gets vertex positions (edge midpoints projected to
the unit sphere), which the mesh-smoothing workload uses as right-hand side.

Grid workloads:
  grid2d_laplacian  5-point Dirichlet Laplacian on nx x ny (diagonal 4) + shift*I
  grid3d_laplacian  7-point graph (Neumann) Laplacian on m^3 + shift*I
All generators return (values float64, idx int64 of shape (2, nnz)) in COO.
"""
from __future__ import annotations

import numpy as np

# Base icosahedron: the 20 faces of the reference generator (tests/utils.py:26-49)
# and the matching golden-ratio vertex coordinates.
_ICO_FACES = np.array([
    [0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
    [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
    [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
    [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], dtype=np.int64)
_PHI = (1.0 + 5.0 ** 0.5) / 2.0
_ICO_POS = np.array([
    [-1, _PHI, 0], [1, _PHI, 0], [-1, -_PHI, 0], [1, -_PHI, 0],
    [0, -1, _PHI], [0, 1, _PHI], [0, -1, -_PHI], [0, 1, -_PHI],
    [_PHI, 0, -1], [_PHI, 0, 1], [-_PHI, 0, -1], [-_PHI, 0, 1]], dtype=np.float64)

# Cube surface triangulation of the reference tests (tests/utils.py:84-97).
_CUBE_FACES = np.array([
    [0, 1, 3], [0, 3, 2], [2, 3, 7], [2, 7, 6], [4, 5, 7], [4, 7, 6],
    [0, 1, 5], [0, 5, 4], [1, 5, 7], [1, 7, 3], [0, 4, 6], [0, 6, 2]], dtype=np.int64)


def cube():
    """(n_verts, faces) of the 8-vertex cube mesh."""
    return 8, _CUBE_FACES.copy()


def icosphere(level: int = 0, with_positions: bool = False):
    """Loop-style 1-to-4 subdivision of the icosahedron, `level` times.

    A new vertex is created for every edge the first time the edge is met
    while visiting faces in order and, inside a face (a, b, c), the edges
    (a, b), (b, c), (c, a); it gets the next free index. Face (a, b, c) is
    replaced by (a, m_ab, m_ca), (b, m_bc, m_ab), (c, m_ca, m_bc),
    (m_ab, m_bc, m_ca). n = 10 * 4**level + 2.
    """
    faces = _ICO_FACES.copy()
    pos = _ICO_POS / np.linalg.norm(_ICO_POS, axis=1, keepdims=True)
    n = 12
    for _ in range(level):
        a, b, c = faces[:, 0], faces[:, 1], faces[:, 2]
        edges = np.stack([np.stack([a, b], 1), np.stack([b, c], 1), np.stack([c, a], 1)],
                         axis=1).reshape(-1, 2)
        lo, hi = edges.min(1), edges.max(1)
        key = lo * n + hi
        uniq, first, inv = np.unique(key, return_index=True, return_inverse=True)
        by_first = np.argsort(first, kind="stable")
        rank = np.empty(len(uniq), dtype=np.int64)
        rank[by_first] = np.arange(len(uniq))
        mid = (n + rank[inv]).reshape(-1, 3)
        m_ab, m_bc, m_ca = mid[:, 0], mid[:, 1], mid[:, 2]
        faces = np.stack([np.stack([a, m_ab, m_ca], 1), np.stack([b, m_bc, m_ab], 1),
                          np.stack([c, m_ca, m_bc], 1), np.stack([m_ab, m_bc, m_ca], 1)],
                         axis=1).reshape(-1, 3)
        if with_positions:
            e_first = first[by_first]
            p = 0.5 * (pos[lo[e_first]] + pos[hi[e_first]])
            pos = np.concatenate([pos, p / np.linalg.norm(p, axis=1, keepdims=True)])
        n += len(uniq)
    if with_positions:
        return n, faces, pos
    return n, faces


def mesh_laplacian(n_verts: int, faces, lam: float):
    """COO of I + lam * (D - W) for the mesh's vertex graph, laid out as the
    reference builds it: the sorted directed edges with -lam, then one +lam
    diagonal entry per directed edge, then the identity."""
    f = np.asarray(faces, dtype=np.int64)
    src = f[:, [1, 2, 0]].ravel()
    dst = f[:, [2, 0, 1]].ravel()
    r = np.concatenate([src, dst])
    c = np.concatenate([dst, src])
    keys = np.unique(r * n_verts + c)
    rows, cols = keys // n_verts, keys % n_verts
    diag = np.arange(n_verts, dtype=np.int64)
    idx = np.concatenate([np.stack([rows, cols]), np.stack([rows, rows]), np.stack([diag, diag])],
                         axis=1)
    values = np.concatenate([np.full(len(keys), -float(lam)), np.full(len(keys), float(lam)),
                             np.ones(n_verts)])
    return values, idx


def grid2d_laplacian(nx: int, ny: int, shift: float):
    """5-point Laplacian with Dirichlet boundary (diagonal 4, -1 per grid
    neighbour) on an nx x ny grid, plus shift * I; vertex (i, j) -> i*ny + j."""
    n = nx * ny
    ids = np.arange(n, dtype=np.int64).reshape(nx, ny)
    rows = [ids.ravel()]
    cols = [ids.ravel()]
    vals = [np.full(n, 4.0 + shift)]
    for a, b in ((ids[:-1, :], ids[1:, :]), (ids[:, :-1], ids[:, 1:])):
        a, b = a.ravel(), b.ravel()
        rows += [a, b]
        cols += [b, a]
        vals += [np.full(len(a), -1.0), np.full(len(a), -1.0)]
    return np.concatenate(vals), np.stack([np.concatenate(rows), np.concatenate(cols)])


def grid3d_laplacian(m: int, shift: float):
    """7-point graph Laplacian (diagonal = number of grid neighbours, Neumann)
    on an m x m x m grid, plus shift * I."""
    n = m ** 3
    ids = np.arange(n, dtype=np.int64).reshape(m, m, m)
    deg = np.zeros((m, m, m))
    rows, cols, vals = [], [], []
    for axis in range(3):
        sl_a = [slice(None)] * 3
        sl_b = [slice(None)] * 3
        sl_a[axis] = slice(None, -1)
        sl_b[axis] = slice(1, None)
        a, b = ids[tuple(sl_a)].ravel(), ids[tuple(sl_b)].ravel()
        deg[tuple(sl_a)] += 1
        deg[tuple(sl_b)] += 1
        rows += [a, b]
        cols += [b, a]
        vals += [np.full(len(a), -1.0), np.full(len(a), -1.0)]
    rows.insert(0, ids.ravel())
    cols.insert(0, ids.ravel())
    vals.insert(0, deg.ravel() + shift)
    return np.concatenate(vals), np.stack([np.concatenate(rows), np.concatenate(cols)])


def smoothing_system(level: int = 7, lam: float = 10.0, seed: int = 0):
    """Mesh-smoothing system on icosphere(level): returns (n, values, idx, rhs)
    with rhs = vertex positions + N(0, 0.01) noise, shape (n, 3)."""
    n, faces, pos = icosphere(level, with_positions=True)
    values, idx = mesh_laplacian(n, faces, lam)
    rng = np.random.default_rng(seed)
    rhs = pos + rng.normal(0.0, 0.01, size=pos.shape)
    return n, values, idx, rhs


def batch_systems(count: int = 256, seed: int = 0, n_min: int = 1000, n_max: int = 50000,
                  k_max: int = 128):
    """Independent grid systems for batched solves: n log-uniform in
    [n_min, n_max], a 2D 5-point grid nx x ny with aspect ratio log-uniform
    in [1/4, 4], plus shift * I with shift ~ U[1e-2, 1], and a right-hand-side
    count uniform in {1..k_max}. Returns a list of (n, values, idx, k)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        n = float(np.exp(rng.uniform(np.log(n_min), np.log(n_max))))
        r = float(np.exp(rng.uniform(np.log(0.25), np.log(4.0))))
        nx = max(2, int(round(np.sqrt(n * r))))
        ny = max(2, int(round(n / nx)))
        shift = float(rng.uniform(1e-2, 1.0))
        values, idx = grid2d_laplacian(nx, ny, shift)
        out.append((nx * ny, values, idx, int(rng.integers(1, k_max + 1))))
    return out
