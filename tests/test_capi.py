"""The C ABI library loads and exports every entry point include/cholespy_amd.h
declares; device-free calls behave without a GPU."""
import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "cholespy_amd.h"
LIB = ROOT / "cholespy_amd" / "libcholespy_amd.so"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(?:int|const char \*)\s*(cholespy_\w+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ("cholespy_init", "cholespy_shutdown", "cholespy_solver_create", "cholespy_solver_destroy",
                 "cholespy_solver_upload_factor", "cholespy_solver_analyze", "cholespy_solver_solve_f32",
                 "cholespy_solver_solve_f64"):
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(LIB))
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_device_free_calls():
    lib = ctypes.CDLL(str(LIB))
    lib.cholespy_last_error.restype = ctypes.c_char_p
    count = ctypes.c_int(-1)
    assert lib.cholespy_device_count(ctypes.byref(count)) == 0
    assert count.value >= 0
    assert lib.cholespy_shutdown() == 0
    assert lib.cholespy_solver_destroy(None) == 0
    # invalid arguments are rejected before any device work
    assert lib.cholespy_solver_create(4, 3, ctypes.byref(ctypes.c_void_p())) == 1
    assert lib.cholespy_device_count(None) == 1
    assert lib.cholespy_last_error()


def test_no_device_fails_loudly():
    """Without a visible GPU the library reports it; nothing falls back to the CPU."""
    import torch
    if torch.cuda.device_count() > 0:
        pytest.skip("a GPU is visible")
    lib = ctypes.CDLL(str(LIB))
    h = ctypes.c_void_p()
    assert lib.cholespy_solver_create(4, 4, ctypes.byref(h)) == 3   # CHOLESPY_ERR_NO_DEVICE
