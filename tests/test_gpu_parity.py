"""GPU parity: the HIP analysis and solves against the oracle on the same factor.

Bar: level arrays bit-exact (integer work); solves within a stated relative
tolerance of the oracle's restatement of the reference kernels in the same
precision (the HIP kernels reduce a row's products in a different order),
and within 1e-12 of the FP64 host solve for double precision.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import cholespy_amd
from cholespy_amd import CholeskySolverD, CholeskySolverF, MatrixType, workloads
from oracle import oracle

PATHS = {"rows": cholespy_amd._core._PATH_ROWS, "blocks": cholespy_amd._core._PATH_BLOCKS}

pytestmark = pytest.mark.gpu

DEV = "cuda"

CASES = {
    "cube": lambda: (lambda n, f: (n,) + workloads.mesh_laplacian(n, f, 2.0))(*workloads.cube()),
    "ico3": lambda: (lambda n, f: (n,) + workloads.mesh_laplacian(n, f, 2.0))(*workloads.icosphere(3)),
    "ico5": lambda: (lambda n, f: (n,) + workloads.mesh_laplacian(n, f, 2.0))(*workloads.icosphere(5)),
    "grid64": lambda: (4096,) + workloads.grid2d_laplacian(64, 64, 1e-2),
}

# Relative (max-norm) tolerance against the same-precision oracle.
TOL = {np.float32: 2e-5, np.float64: 1e-12}


def _gpu_solver(cls, n, values, idx):
    return cls(n, torch.tensor(idx[0], device=DEV), torch.tensor(idx[1], device=DEV),
               torch.tensor(values, device=DEV), MatrixType.COO)


@pytest.mark.parametrize("case", list(CASES))
def test_levels_bit_exact(case):
    n, values, idx = CASES[case]()
    solver = _gpu_solver(CholeskySolverD, n, values, idx)
    fac = solver._factor()
    rowptr, cols, _ = oracle.factor_rows(fac["colptr"], fac["rowidx"], fac["val"])
    for lower, rp, cl in ((True, rowptr, cols), (False, fac["colptr"], fac["rowidx"])):
        ref_level, mx = oracle.levels(rp, cl, lower)
        order, row_level = solver._levels(lower)
        assert np.array_equal(row_level, ref_level), (case, lower)
        assert np.array_equal(order, oracle.level_sort(ref_level, mx)), (case, lower)
        assert solver._n_levels(lower) == mx + 1


@pytest.mark.parametrize("path", list(PATHS))
@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("cls,dtype", [(CholeskySolverF, np.float32), (CholeskySolverD, np.float64)])
def test_solve_matches_oracle(case, cls, dtype, path):
    n, values, idx = CASES[case]()
    solver = _gpu_solver(cls, n, values, idx)
    solver._set_path(PATHS[path])
    fac = solver._factor()
    rng = np.random.default_rng(1)
    for k in (1, 3, 4, 5, 16, 17, 33, 64, 65, 128):
        b = rng.random((n, k)).astype(dtype)
        bt = torch.tensor(b, device=DEV)
        xt = torch.zeros_like(bt)
        solver.solve(bt, xt)
        torch.cuda.synchronize()
        solver._check()
        x = xt.cpu().numpy()
        ref = oracle.solve(fac, b, dtype)
        err = np.abs(x - ref).max() / np.abs(ref).max()
        assert err < TOL[dtype], (case, k, err)
        if dtype == np.float64:
            host = oracle.host_solve(fac, b)
            assert np.abs(x - host).max() / np.abs(host).max() < 1e-12


def test_repeated_solves_and_rhs_change():
    """Work buffers are reused across solves, re-laid-out when k changes and
    re-initialised when the schedule changes."""
    n, values, idx = CASES["ico5"]()
    solver = _gpu_solver(CholeskySolverF, n, values, idx)
    fac = solver._factor()
    rng = np.random.default_rng(2)
    for i, k in enumerate((3, 3, 1, 128, 3, 7, 7, 3, 3)):
        solver._set_path(PATHS["rows"] if i in (5, 7) else PATHS["blocks"])
        b = rng.standard_normal((n, k)).astype(np.float32)
        bt = torch.tensor(b, device=DEV)
        xt = torch.zeros_like(bt)
        solver.solve(bt, xt)
        torch.cuda.synchronize()
        solver._check()
        ref = oracle.solve(fac, b, np.float32)
        assert np.abs(xt.cpu().numpy() - ref).max() / np.abs(ref).max() < 2e-5, k


def test_vector_rhs_and_identity():
    """README example: A = I gives x = b (README.md:104-121), 1-D b."""
    n = 20
    rows = torch.arange(n, device=DEV)
    solver = CholeskySolverF(n, rows, rows, torch.ones(n, device=DEV), MatrixType.COO)
    b = torch.ones(n, device=DEV)
    x = torch.zeros_like(b)
    solver.solve(b, x)
    torch.cuda.synchronize()
    assert torch.equal(x, b)


def test_too_many_rhs_raises():
    n, values, idx = CASES["cube"]()
    solver = _gpu_solver(CholeskySolverF, n, values, idx)
    b = torch.ones((n, 129), device=DEV)
    with pytest.raises(ValueError, match="The number of RHS should be less than 128."):
        solver.solve(b, torch.zeros_like(b))


def test_full_size_mesh_smoothing_residual():
    """icosphere(7), n = 163,842, k = 3: size-independent checks at full size.
    FP64 GPU solve vs the FP64 host solve on the same factor, and the FP32
    solve's relative residual."""
    n, values, idx, rhs = workloads.smoothing_system(7, 10.0, seed=0)
    A = oracle.symmetric_from_lower(n, *oracle.normalized_lower_csc(n, idx[0], idx[1], values, "COO"))
    for cls, dtype, tol, path in ((CholeskySolverD, np.float64, 1e-12, "blocks"),
                                  (CholeskySolverF, np.float32, 1e-5, "blocks"),
                                  (CholeskySolverD, np.float64, 1e-12, "rows")):
        solver = _gpu_solver(cls, n, values, idx)
        solver._set_path(PATHS[path])
        b = rhs.astype(dtype)
        bt = torch.tensor(b, device=DEV)
        xt = torch.zeros_like(bt)
        solver.solve(bt, xt)
        torch.cuda.synchronize()
        solver._check()
        x = xt.cpu().numpy().astype(np.float64)
        host = oracle.host_solve(solver._factor(), b.astype(np.float64))
        rel = np.linalg.norm(x - host) / np.linalg.norm(host)
        res = np.linalg.norm(b - A @ x) / np.linalg.norm(b)
        assert rel < tol and res < tol, (dtype, rel, res)


@pytest.mark.parametrize("case", ["ico5", "grid64"])
def test_refined_reaches_fp64_accuracy(case):
    """FP32 block solve + FP64 refinement: <= 1e-10 relative error against
    the FP64 host solve on the same factor after 2 steps, and the error
    shrinks with each step."""
    n, values, idx = CASES[case]()
    solver = _gpu_solver(CholeskySolverF, n, values, idx)
    rng = np.random.default_rng(5)
    b = rng.standard_normal((n, 3))
    host = oracle.host_solve(solver._factor(), b)
    bt = torch.tensor(b, device=DEV)
    errs = []
    for it in (0, 1, 2):
        xt = torch.zeros_like(bt)
        solver._solve_refined(bt, xt, it)
        torch.cuda.synchronize()
        solver._check()
        errs.append(np.linalg.norm(xt.cpu().numpy() - host) / np.linalg.norm(host))
    assert errs[2] <= 1e-10, errs
    assert errs[1] < errs[0] * 1e-2, errs


def test_single_passes_compose_to_a_solve():
    """solve_pass(0) then solve_pass(1) equals solve() (launch_kernel(lower)
    then launch_kernel(upper), src/cholesky_solver.cpp:428-431)."""
    n, values, idx = CASES["ico5"]()
    for cls, dtype in ((CholeskySolverF, torch.float32), (CholeskySolverD, torch.float64)):
        solver = _gpu_solver(cls, n, values, idx)
        b = torch.rand(n, 3, device=DEV, dtype=dtype)
        x1 = torch.zeros_like(b)
        x2 = torch.zeros_like(b)
        solver.solve(b, x1)
        solver._solve_pass(0, b, x2)
        solver._solve_pass(1, b, x2)
        torch.cuda.synchronize()
        solver._check()
        assert torch.equal(x1, x2)


@pytest.mark.parametrize("split_entries", [512, 2048])
@pytest.mark.parametrize("case", ["ico5", "grid64"])
@pytest.mark.parametrize("cls,dtype", [(CholeskySolverF, np.float32), (CholeskySolverD, np.float64)])
def test_split_reductions_match_oracle(case, cls, dtype, split_entries, monkeypatch):
    """Plans whose large product tasks are split along their reduction (split
    tasks + combine tasks in both passes) give the same solutions; both the
    vector-FMA (k <= 4) and the MFMA (k > 4) kernels run them."""
    monkeypatch.setenv("CHOLESPY_SPLIT_ENTRIES", str(split_entries))
    n, values, idx = CASES[case]()
    solver = _gpu_solver(cls, n, values, idx)
    fac = solver._factor()
    tasks = solver._plan_arrays()
    assert (tasks["fwd_task"][:, 3] == 3).any() and (tasks["bwd_task"][:, 3] == 1).any()
    rng = np.random.default_rng(7)
    for k in (3, 17, 64, 3):
        b = rng.random((n, k)).astype(dtype)
        bt = torch.tensor(b, device=DEV)
        xt = torch.zeros_like(bt)
        solver.solve(bt, xt)
        torch.cuda.synchronize()
        solver._check()
        ref = oracle.solve(fac, b, dtype)
        err = np.abs(xt.cpu().numpy() - ref).max() / np.abs(ref).max()
        assert err < TOL[dtype], (case, k, err)


@pytest.mark.parametrize("width", [32, 64])
@pytest.mark.parametrize("case", ["ico5", "grid64"])
@pytest.mark.parametrize("cls,dtype", [(CholeskySolverF, np.float32), (CholeskySolverD, np.float64)])
def test_wide_panel_tasks_match_oracle(case, cls, dtype, width, monkeypatch):
    """The MFMA panel kernels on 32- and 64-wide tasks of the large blocks
    (forward rows, backward columns: two and four 16-wide groups, one per
    wavefront at 64), split along their reductions, for 1 to 8 tiles of
    right-hand sides."""
    monkeypatch.setenv("CHOLESPY_WIDE_FINE_BWD_COLS_BIG", str(width))
    monkeypatch.setenv("CHOLESPY_WIDE_FWD_ROWS_BIG", str(width))
    monkeypatch.setenv("CHOLESPY_SPLIT_ENTRIES", "1024")
    n, values, idx = CASES[case]()
    solver = _gpu_solver(cls, n, values, idx)
    fac = solver._factor()
    tasks = solver._plan_arrays()
    assert ((tasks["bwd_task_wide"][:, 2] - tasks["bwd_task_wide"][:, 1]) == width).any()
    assert ((tasks["fwd_task_wide"][:, 2] - tasks["fwd_task_wide"][:, 1]) == width).any()
    rng = np.random.default_rng(11)
    for k in (5, 32, 33, 128):
        b = rng.random((n, k)).astype(dtype)
        bt = torch.tensor(b, device=DEV)
        xt = torch.zeros_like(bt)
        solver.solve(bt, xt)
        torch.cuda.synchronize()
        solver._check()
        ref = oracle.solve(fac, b, dtype)
        err = np.abs(xt.cpu().numpy() - ref).max() / np.abs(ref).max()
        assert err < TOL[dtype], (case, k, err)


@pytest.mark.parametrize("case", ["ico5", "grid64"])
def test_refined_to_tolerance(case):
    """Refinement until the device-computed relative residual is <= tol: the
    reported residual matches SciPy's on the host, the error against the
    FP64 host solve is <= 1e-10, and no more steps are used than the fixed
    schedule needs for the same residual."""
    import scipy.sparse as sp
    n, values, idx = CASES[case]()
    solver = _gpu_solver(CholeskySolverF, n, values, idx)
    A = sp.coo_matrix((values, (idx[0], idx[1])), shape=(n, n)).tocsr()
    rng = np.random.default_rng(6)
    b = rng.standard_normal((n, 3))
    host = oracle.host_solve(solver._factor(), b)
    bt = torch.tensor(b, device=DEV)
    xt = torch.zeros_like(bt)
    steps, res = solver._solve_refined_tol(bt, xt, 1e-12, 8)
    torch.cuda.synchronize()
    solver._check()
    x = xt.cpu().numpy()
    res_host = float((np.linalg.norm(b - A @ x, axis=0) / np.linalg.norm(b, axis=0)).max())
    assert res <= 1e-12 and abs(res - res_host) <= 0.05 * res_host + 1e-17, (res, res_host)
    assert np.linalg.norm(x - host) / np.linalg.norm(host) <= 1e-10
    assert 1 <= steps <= 3
    # one step fewer does not reach the bar
    xt2 = torch.zeros_like(bt)
    solver._solve_refined(bt, xt2, steps - 1)
    torch.cuda.synchronize()
    x2 = xt2.cpu().numpy()
    assert float((np.linalg.norm(b - A @ x2, axis=0) / np.linalg.norm(b, axis=0)).max()) > 1e-12


@pytest.mark.parametrize("cls,dtype", [(CholeskySolverF, torch.float32), (CholeskySolverD, torch.float64)])
def test_solve_on_a_stream_and_in_a_graph(cls, dtype):
    """solve(b, x, stream=...) runs on the caller's stream, and a solve
    captured in a graph replays bit-identically to eager solves of the same
    right-hand sides, with eager solves (other widths too) between replays.
    Both kernel families: k = 3 (narrow) and k = 32 (panel)."""
    n, values, idx = CASES["ico5"]()
    solver = _gpu_solver(cls, n, values, idx)
    gen = torch.Generator(device=DEV).manual_seed(11)
    for k in (3, 32):
        b = torch.rand(n, k, device=DEV, dtype=dtype, generator=gen)
        ref = torch.zeros_like(b)
        solver.solve(b, ref)
        torch.cuda.synchronize()
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        x = torch.zeros_like(b)
        with torch.cuda.stream(side):
            solver.solve(b, x, stream=side.cuda_stream)
        torch.cuda.current_stream().wait_stream(side)
        torch.cuda.synchronize()
        solver._check()
        assert torch.equal(x, ref), k
        bg = b.clone()
        xg = torch.zeros_like(b)
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph):
            solver.solve(bg, xg, stream=torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        for rep in range(4):
            bn = torch.rand(n, k, device=DEV, dtype=dtype, generator=gen)
            want = torch.zeros_like(bn)
            solver.solve(bn, want)
            other = torch.rand(n, 5 + rep, device=DEV, dtype=dtype, generator=gen)
            solver.solve(other, torch.zeros_like(other))
            torch.cuda.synchronize()
            bg.copy_(bn)
            xg.zero_()
            graph.replay()
            torch.cuda.synchronize()
            solver._check()
            assert torch.equal(xg, want), (k, rep)
        ref_np = oracle.solve(solver._factor(), bn.cpu().numpy(), np.float32 if dtype == torch.float32 else np.float64)
        err = np.abs(xg.cpu().numpy() - ref_np).max() / np.abs(ref_np).max()
        assert err < TOL[np.float32 if dtype == torch.float32 else np.float64], (k, err)
