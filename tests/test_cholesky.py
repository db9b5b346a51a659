"""The reference's own test module (tests/test_cholesky.py of cholespy),
restated against the drop-in: same meshes, seed, lambda, RHS sweep, tolerance
(np.allclose defaults) and error message. "cuda" cases run on the ROCm
device and carry the gpu marker."""
import numpy as np
import pytest
import scipy.sparse as sp
from scipy.sparse.linalg import spsolve

from cholespy import CholeskySolverD, CholeskySolverF, MatrixType
from cholespy_amd.workloads import cube, icosphere, mesh_laplacian

RHS = [1, 4, 16, 64, 128, 256]
DEVICES = ["cpu", pytest.param("cuda", marks=pytest.mark.gpu)]


@pytest.mark.parametrize("n_verts, faces", [cube(), icosphere(3), icosphere(5)], ids=["cube", "ico3", "ico5"])
@pytest.mark.parametrize("variant", ["float", "double"])
@pytest.mark.parametrize("device", DEVICES)
def test_solver(device, variant, n_verts, faces):
    torch = pytest.importorskip("torch")
    np.random.seed(45)
    CholeskySolver = CholeskySolverF if variant == "float" else CholeskySolverD
    dtype = np.float32 if variant == "float" else np.float64
    lambda_ = 2.0
    values, idx = mesh_laplacian(n_verts, faces, lambda_)
    L_csc = sp.csc_matrix((values, idx))
    b = np.random.random(size=(n_verts, max(RHS))).astype(dtype)
    x = spsolve(L_csc, b)
    solver = CholeskySolver(n_verts, torch.tensor(idx[0], device=device), torch.tensor(idx[1], device=device),
                            torch.tensor(values, device=device), MatrixType.COO)
    for n_rhs in RHS:
        b_torch = torch.tensor(b[:, :n_rhs], device=device)
        x_torch = torch.zeros_like(b_torch)
        try:
            solver.solve(b_torch, x_torch)
            assert np.allclose(x_torch.cpu().numpy(), x[:, :n_rhs])
        except ValueError as e:
            assert device == "cuda" and n_rhs > 128
            assert str(e) == "The number of RHS should be less than 128."


@pytest.mark.parametrize("device", DEVICES)
def test_matrices(device):
    torch = pytest.importorskip("torch")
    np.random.seed(45)
    n_verts, faces = cube()
    values, idx = mesh_laplacian(n_verts, faces, 2.0)
    L_csc = sp.csc_matrix((values, idx))
    L_csr = sp.csr_matrix((values, idx))
    b = np.random.random(size=(n_verts, 32)).astype(np.float32)
    b_torch = torch.tensor(b, device=device)
    x_torch = torch.zeros_like(b_torch)
    x_ref = spsolve(L_csc, b)
    for ii, jj, xx, kind in ((idx[0], idx[1], values, MatrixType.COO),
                             (L_csr.indptr, L_csr.indices, L_csr.data, MatrixType.CSR),
                             (L_csc.indptr, L_csc.indices, L_csc.data, MatrixType.CSC)):
        solver = CholeskySolverF(n_verts, torch.tensor(ii, device=device), torch.tensor(jj, device=device),
                                 torch.tensor(xx, device=device), kind)
        solver.solve(b_torch, x_torch)
        assert np.allclose(x_torch.cpu().numpy(), x_ref)


@pytest.mark.parametrize("framework", ["numpy", "torch", "tensorflow", "jax", "cupy", "drjit"])
def test_frameworks(framework):
    import importlib
    try:
        importlib.import_module(framework)
    except ModuleNotFoundError:
        pytest.skip(f"Module {framework} is not installed.")
    np.random.seed(45)
    n_verts, faces = cube()
    values, idx = mesh_laplacian(n_verts, faces, 2.0)
    b = np.random.random(size=(n_verts, 32)).astype(np.float32)
    x_ref = spsolve(sp.csc_matrix((values, idx)), b)
    if framework == "numpy":
        solver = CholeskySolverF(n_verts, idx[0], idx[1], values, MatrixType.COO)
        x = np.zeros_like(b)
        solver.solve(b, x)
        assert np.allclose(x, x_ref)
    elif framework == "torch":
        import torch
        solver = CholeskySolverF(n_verts, torch.tensor(idx[0]), torch.tensor(idx[1]), torch.tensor(values),
                                 MatrixType.COO)
        b_torch = torch.tensor(b)
        x_torch = torch.zeros_like(b_torch)
        solver.solve(b_torch, x_torch)
        assert np.allclose(x_torch.numpy(), x_ref)
    else:
        pytest.skip(f"{framework} interop is not exercised in this environment.")


@pytest.mark.gpu
def test_frameworks_torch_gpu():
    """The CUDA half of the reference's torch framework test."""
    import torch
    np.random.seed(45)
    n_verts, faces = cube()
    values, idx = mesh_laplacian(n_verts, faces, 2.0)
    b = np.random.random(size=(n_verts, 32)).astype(np.float32)
    x_ref = spsolve(sp.csc_matrix((values, idx)), b)
    solver = CholeskySolverF(n_verts, torch.tensor(idx[0], device="cuda"), torch.tensor(idx[1], device="cuda"),
                             torch.tensor(values, device="cuda"), MatrixType.COO)
    b_torch = torch.tensor(b, device="cuda")
    x_torch = torch.zeros_like(b_torch)
    solver.solve(b_torch, x_torch)
    assert np.allclose(x_torch.cpu().numpy(), x_ref)


def test_mesh_generators_match_reference_vectors():
    """cholespy_amd.workloads reproduces the reference generators'
    outputs (tests/golden/mesh_vectors.npz, written from tests/utils.py)."""
    from pathlib import Path
    with np.load(Path(__file__).parent / "golden" / "mesh_vectors.npz", allow_pickle=False) as z:
        n, f = cube()
        assert n == z["cube_n"] and np.array_equal(f, z["cube_faces"])
        v, idx = mesh_laplacian(n, f, 2.0)
        assert np.array_equal(v, z["cube_coo_values"]) and np.array_equal(idx, z["cube_coo_idx"])
        for level in range(4):
            n, f = icosphere(level)
            assert n == z[f"ico{level}_n"] and np.array_equal(f, z[f"ico{level}_faces"])
        n, f = icosphere(2)
        v, idx = mesh_laplacian(n, f, 2.0)
        assert np.array_equal(v, z["ico2_coo_values"]) and np.array_equal(idx, z["ico2_coo_idx"])
