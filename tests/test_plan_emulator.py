"""CPU check of the block plan's task lists (cholespy_amd/csrc/device/block_plan.cpp)
by emulation (tests/plan_emulator.cpp, built here with g++).

The emulator runs each list's forward and backward tasks in ticket order
with every published value starting as the sentinel. It fails if a task
reads a value that no earlier ticket wrote. On the GPU, where workgroups
take tickets in order, that would be a wait that never ends. It also fails
on split-group mistakes, on a leaf launch's assumption that does not hold,
or on a solution that differs from the sequential solve with the same L.
The factors come from the host factorisation (CPU mode), and the option
sets include the aggressive splits and thresholds the GPU parity tests use.
"""
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

from cholespy_amd import CholeskySolverD, MatrixType, workloads

ROOT = Path(__file__).resolve().parents[1]
DEVICE = ROOT / "cholespy_amd" / "csrc" / "device"


@pytest.fixture(scope="module")
def emulator(tmp_path_factory):
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("g++ is not available")
    exe = tmp_path_factory.mktemp("plan") / "plan_emulator"
    subprocess.run([cxx, "-O2", "-std=c++17", "-pthread", "-I", str(DEVICE),
                    str(ROOT / "tests" / "plan_emulator.cpp"), str(DEVICE / "block_plan.cpp"),
                    "-o", str(exe)], check=True)
    return exe


def _mesh(level):
    n, f = workloads.icosphere(level)
    return (n,) + workloads.mesh_laplacian(n, f, 2.0)


CASES = {
    "ico4": lambda: _mesh(4),
    "grid64": lambda: (4096,) + workloads.grid2d_laplacian(64, 64, 1e-2),
    "cube": lambda: (lambda n, f: (n,) + workloads.mesh_laplacian(n, f, 2.0))(*workloads.cube()),
}


@pytest.fixture(scope="module")
def factors(tmp_path_factory):
    out = {}
    d = tmp_path_factory.mktemp("factors")
    for name, make in CASES.items():
        n, values, idx = make()
        fac = CholeskySolverD(n, idx[0], idx[1], values, MatrixType.COO)._factor()
        path = d / f"{name}.bin"
        with open(path, "wb") as fh:
            fh.write(np.array([n, len(fac["rowidx"])], dtype=np.int64).tobytes())
            fh.write(np.ascontiguousarray(fac["colptr"], dtype=np.int32).tobytes())
            fh.write(np.ascontiguousarray(fac["rowidx"], dtype=np.int32).tobytes())
            fh.write(np.ascontiguousarray(fac["val"], dtype=np.float64).tobytes())
        out[name] = path
    return out


OPTIONS = [
    [],
    ["split_entries=512"],
    ["split_entries=512", "max_split=16"],
    ["split_entries=2048", "level_tasks=256"],
    ["split_gather=33"],
    ["split_entries=0"],
    ["wide_list=0"],
    ["leaf_max=128", "max_block=256"],
    ["critical_order=0"],
    ["wide_fine_bwd_cols_big=64", "wide_fwd_rows_big=64"],
]


@pytest.fixture(scope="module")
def grid3d_factor(tmp_path_factory):
    """32^3 7-point grid: its elimination tree has a 1,411-row chain, so the
    forced caps below give one block of more than 1024 rows."""
    m = 32
    values, idx = workloads.grid3d_laplacian(m, 1.0)
    fac = CholeskySolverD(m ** 3, idx[0], idx[1], values, MatrixType.COO)._factor()
    path = tmp_path_factory.mktemp("factors3d") / "grid32.bin"
    with open(path, "wb") as fh:
        fh.write(np.array([m ** 3, len(fac["rowidx"])], dtype=np.int64).tobytes())
        fh.write(np.ascontiguousarray(fac["colptr"], dtype=np.int32).tobytes())
        fh.write(np.ascontiguousarray(fac["rowidx"], dtype=np.int32).tobytes())
        fh.write(np.ascontiguousarray(fac["val"], dtype=np.float64).tobytes())
    return path


@pytest.mark.parametrize("opts", [["max_block=2048"], ["max_block=4096"],
                                  ["max_block=4096", "split_entries=512"]],
                         ids=lambda o: ",".join(o))
@pytest.mark.parametrize("task_list", ["narrow", "wide"])
def test_long_chain_blocks_run_in_ticket_order(emulator, grid3d_factor, task_list, opts):
    res = subprocess.run([str(emulator), str(grid3d_factor), task_list, *opts],
                         capture_output=True, text=True, timeout=300)
    assert res.returncode == 0 and res.stdout.startswith("ok"), res.stdout + res.stderr


@pytest.mark.parametrize("opts", OPTIONS, ids=lambda o: ",".join(o) or "default")
@pytest.mark.parametrize("task_list", ["narrow", "wide"])
@pytest.mark.parametrize("case", list(CASES))
def test_task_list_runs_in_ticket_order(emulator, factors, case, task_list, opts):
    res = subprocess.run([str(emulator), str(factors[case]), task_list, *opts],
                         capture_output=True, text=True, timeout=120)
    assert res.returncode == 0 and res.stdout.startswith("ok"), res.stdout + res.stderr


def test_emulator_reports_a_planted_schedule_mistake(emulator, factors):
    """The check has teeth: the same plan run in reverse ticket order (a
    mistake planted in the emulator, on the CPU) reads values no earlier
    task wrote, and the emulator says so; unknown options are refused."""
    res = subprocess.run([str(emulator), str(factors["grid64"]), "narrow", "reverse=1"],
                         capture_output=True, text=True, timeout=120)
    assert res.returncode == 1 and "reads an unwritten" in res.stdout, res.stdout
    res = subprocess.run([str(emulator), str(factors["grid64"]), "narrow", "no_such_option=1"],
                         capture_output=True, text=True, timeout=120)
    assert res.returncode == 2
