"""The hand-off microbenchmark (cholespy_probe_handoff) runs, every wait
ends, and the figures are in the ranges the block schedule is designed
around: a value crosses workgroups in about a microsecond or less, a
dependent agent-scope load and a barrier cost far less."""
import pytest

pytest.importorskip("torch")

import cholespy_amd

pytestmark = pytest.mark.gpu


def test_probe_handoff():
    r = cholespy_amd._core._probe_handoff()
    print(r, flush=True)
    assert 0 < r["handoff_relaxed_us"] < 20 and 0 < r["handoff_fenced_us"] < 20
    assert 0 < r["agent_load_ns"] < r["handoff_relaxed_us"] * 1e3
    assert 0 < r["barrier_ns"] < 1000
    assert 0 < r["ticket_atomic_ns"] < 1000
    assert 0 < r["launch_us"] < 100
