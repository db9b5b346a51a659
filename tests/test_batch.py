"""Batched solves: many independent systems, one launch per pass. Each
system's slice of the packed solution is checked against the oracle on that
system's own factor, in the solver's precision."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from cholespy_amd import BatchedSolverD, BatchedSolverF, CholeskySolverD, MatrixType, workloads
from oracle import oracle


def test_batch_generator_shapes():
    systems = workloads.batch_systems(count=16, seed=1, n_min=100, n_max=2000)
    for n, values, idx, k in systems:
        assert 1 <= k <= 128 and idx.shape[1] == len(values) and idx.max() < n


def test_batch_constructor_factors_each_system():
    """CPU side: the batch factors each system like a single solver."""
    systems = workloads.batch_systems(count=3, seed=2, n_min=50, n_max=300)
    if torch.cuda.device_count() > 0:
        pytest.skip("covered by the GPU test on a GPU box")
    with pytest.raises(RuntimeError):   # no device here: fails loudly, no CPU fallback
        BatchedSolverF([(n, idx[0], idx[1], v, MatrixType.COO) for n, v, idx, k in systems])


@pytest.mark.gpu
@pytest.mark.parametrize("cls,dtype,tol", [(BatchedSolverF, np.float32, 2e-5), (BatchedSolverD, np.float64, 1e-12)])
def test_batch_matches_oracle(cls, dtype, tol):
    systems = workloads.batch_systems(count=24, seed=3, n_min=50, n_max=3000)
    batch = cls([(n, idx[0], idx[1], v, MatrixType.COO) for n, v, idx, k in systems])
    rng = np.random.default_rng(4)
    for ks in ([k for *_, k in systems], [1 + (i % 5) for i in range(len(systems))]):
        bs = [rng.uniform(-1, 1, (n, k)).astype(dtype) for (n, v, idx, _), k in zip(systems, ks)]
        b = torch.tensor(np.concatenate([x.ravel() for x in bs]), device="cuda")
        x = torch.zeros_like(b)
        batch.solve(ks, b, x)
        torch.cuda.synchronize()
        batch._check()
        xs = x.cpu().numpy()
        off = 0
        for i, ((n, v, idx, _), k) in enumerate(zip(systems, ks)):
            got = xs[off:off + n * k].reshape(n, k)
            off += n * k
            ref = oracle.solve(batch._factor(i), bs[i], dtype)
            err = np.abs(got - ref).max() / np.abs(ref).max()
            assert err < tol, (i, n, k, err)


@pytest.mark.gpu
def test_batch_of_one_equals_single_solver():
    n = 4096
    v, idx = workloads.grid2d_laplacian(64, 64, 1e-2)
    batch = BatchedSolverD([(n, idx[0], idx[1], v, MatrixType.COO)])
    single = CholeskySolverD(n, torch.tensor(idx[0], device="cuda"), torch.tensor(idx[1], device="cuda"),
                             torch.tensor(v, device="cuda"), MatrixType.COO)
    b = torch.rand(n, 7, device="cuda", dtype=torch.float64)
    x1 = torch.zeros_like(b)
    x2 = torch.zeros_like(b)
    single.solve(b, x1)
    batch.solve([7], b.reshape(-1), x2.reshape(-1))
    torch.cuda.synchronize()
    assert torch.equal(x1, x2)


@pytest.mark.gpu
def test_batch_graph_replays_match_eager():
    """One batched solve captured in a hipGraph and replayed three times,
    with new right-hand sides written eagerly between the replays; every
    replay equals the eager solve of the same right-hand sides bit for bit
    (the kernels reset their own queue counters and sentinels, so a replay
    needs no memset node)."""
    systems = workloads.batch_systems(count=12, seed=5, n_min=200, n_max=4000)
    ks = [k for *_, k in systems]
    batch = BatchedSolverF([(n, idx[0], idx[1], v, MatrixType.COO) for n, v, idx, k in systems])
    total = sum(n * k for (n, *_), k in zip(systems, ks))
    gen = torch.Generator(device="cuda").manual_seed(6)
    b = torch.rand(total, device="cuda", generator=gen)
    x = torch.zeros_like(b)
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        batch.solve(ks, b, x, torch.cuda.current_stream().cuda_stream)
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        batch.solve(ks, b, x, torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    for rep in range(3):
        b.copy_(torch.rand(total, device="cuda", generator=gen))
        x.fill_(0.0)
        g.replay()
        torch.cuda.synchronize()
        batch._check()
        got = x.clone()
        x_eager = torch.zeros_like(b)
        batch.solve(ks, b, x_eager)
        torch.cuda.synchronize()
        batch._check()
        assert torch.equal(got, x_eager), rep


@pytest.mark.gpu
def test_batch_with_split_reductions(monkeypatch):
    """A batch whose plans contain split and combine tasks (small split size)
    matches the oracle per system, with per-system right-hand-side counts."""
    monkeypatch.setenv("CHOLESPY_SPLIT_ENTRIES", "512")
    monkeypatch.setenv("CHOLESPY_SPLIT_GATHER", "65")
    systems = workloads.batch_systems(count=10, seed=8, n_min=500, n_max=4000)
    batch = BatchedSolverD([(n, idx[0], idx[1], v, MatrixType.COO) for n, v, idx, k in systems])
    assert (batch._plan_arrays()["fwd_task"][:, 3] == 3).any()
    rng = np.random.default_rng(9)
    ks = [1 + (i * 13) % 40 for i in range(len(systems))]
    bs = [rng.uniform(-1, 1, (n, k)) for (n, *_), k in zip(systems, ks)]
    b = torch.tensor(np.concatenate([x.ravel() for x in bs]), device="cuda")
    x = torch.zeros_like(b)
    for _ in range(2):
        batch.solve(ks, b, x)
        torch.cuda.synchronize()
        batch._check()
        xs = x.cpu().numpy()
        off = 0
        for i, ((n, *_), k) in enumerate(zip(systems, ks)):
            got = xs[off:off + n * k].reshape(n, k)
            off += n * k
            ref = oracle.solve(batch._factor(i), bs[i], np.float64)
            assert np.abs(got - ref).max() / np.abs(ref).max() < 1e-12, (i, n, k)


@pytest.mark.gpu
def test_full_size_batch_of_256_grid_systems():
    """The batch workload at full size: 256 grid systems, n in [1k, 50k],
    k_i in 1..128, FP32, one launch per pass. Every system's relative
    residual in FP64 is small for an FP32 solve, and every 16th system is
    within 2e-5 of the oracle on its own factor."""
    systems = workloads.batch_systems(count=256, seed=0)
    ks = [k for *_, k in systems]
    print(f"batch of {len(systems)} systems, n total {sum(n for n, *_ in systems)}, sum k {sum(ks)}",
          flush=True)
    batch = BatchedSolverF([(n, idx[0], idx[1], v, MatrixType.COO) for n, v, idx, _ in systems])
    rng = np.random.default_rng(9)
    bs = [rng.uniform(-1, 1, (n, k)).astype(np.float32) for (n, *_), k in zip(systems, ks)]
    b = torch.tensor(np.concatenate([x.ravel() for x in bs]), device="cuda")
    x = torch.zeros_like(b)
    batch.solve(ks, b, x)
    torch.cuda.synchronize()
    batch._check()
    xs = x.cpu().numpy()
    off = 0
    for i, ((n, v, idx, _), k) in enumerate(zip(systems, ks)):
        got = xs[off:off + n * k].reshape(n, k).astype(np.float64)
        off += n * k
        A = oracle.symmetric_from_lower(n, *oracle.normalized_lower_csc(n, idx[0], idx[1], v, "COO"))
        res = np.linalg.norm(bs[i] - A @ got) / np.linalg.norm(bs[i])
        assert res < 1e-5, (i, n, k, res)
        if i % 16 == 0:
            ref = oracle.solve(batch._factor(i), bs[i], np.float32)
            assert np.abs(got - ref).max() / np.abs(ref).max() < 2e-5, (i, n, k)
