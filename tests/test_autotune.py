"""The autotuner (cholespy_solver_autotune): it times candidate
configurations on the device and keeps the fastest, never slower than the
hand-set defaults as measured; the tuned solver still matches the oracle;
the choice is cached on disk and re-applied without timing; environment
knobs stay overrides (not varied)."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from cholespy_amd import CholeskySolverD, CholeskySolverF, MatrixType, workloads
from oracle import oracle

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _ico(level, cls):
    n, faces = workloads.icosphere(level)
    values, idx = workloads.mesh_laplacian(n, faces, 2.0)
    return n, cls(n, torch.tensor(idx[0], device=DEV), torch.tensor(idx[1], device=DEV),
                  torch.tensor(values, device=DEV), MatrixType.COO)


def _check(s, n, k, dtype, tol):
    b = np.random.default_rng(k).random((n, k)).astype(dtype)
    bt = torch.tensor(b, device=DEV)
    xt = torch.zeros_like(bt)
    s.solve(bt, xt)
    torch.cuda.synchronize()
    s._check()
    ref = oracle.solve(s._factor(), b, dtype)
    assert np.abs(xt.cpu().numpy() - ref).max() / np.abs(ref).max() < tol


@pytest.mark.parametrize("k", [3, 40])
@pytest.mark.parametrize("cls,dtype,tol", [(CholeskySolverF, np.float32, 2e-5), (CholeskySolverD, np.float64, 1e-12)])
def test_autotune_then_solve(tmp_path, cls, dtype, tol, k):
    n, s = _ico(5, cls)
    cache = tmp_path / "tune.txt"
    r = s.autotune(k, cache=str(cache))
    print(r, flush=True)
    assert r["source"] == "tuned" and r["candidates"] >= 5
    assert r["ms"] <= r["default_ms"] * 1.0001
    _check(s, n, k, dtype, tol)
    line = cache.read_text().strip().splitlines()[-1]
    assert line.startswith(r["key"])
    # a second solver of the same system reads the cache: no timing
    n2, s2 = _ico(5, cls)
    r2 = s2.autotune(k, cache=str(cache))
    assert r2["source"] == "cached" and r2["candidates"] == 0
    for key in ("path", "tile", "fused", "queues", "max_block", "split_entries", "max_split"):
        assert r2[key] == r[key], key
    _check(s2, n2, k, dtype, tol)


def test_environment_knobs_stay_overrides(tmp_path, monkeypatch):
    monkeypatch.setenv("CHOLESPY_MAX_SPLIT", "4")
    monkeypatch.setenv("CHOLESPY_FUSED_LEAVES", "2")
    n, s = _ico(4, CholeskySolverF)
    r = s.autotune(3, cache=str(tmp_path / "t.txt"), write=False)
    assert r["fused"] == 2
    assert r["max_split"] == 8 or r["max_split"] == 4   # the base value is kept; the override applies on top
    assert not (tmp_path / "t.txt").exists()
    _check(s, n, 3, np.float32, 2e-5)
