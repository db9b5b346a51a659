"""In-tree build of the cholespy_amd native libraries.

  libcholespy_amd.so   HIP kernels for gfx950 + the C ABI (include/cholespy_amd.h),
                       compiled by hipcc; no run-time code loading.
  _cholespy_core*.so   pybind11 module: CholeskySolverF/D, MatrixType (host C++,
                       compiled by g++ and linked against libcholespy_amd.so).

Run `python cholespy_amd/build.py` or call build(). Every call recompiles.
"""
from __future__ import annotations

import os
import subprocess
import sys
import sysconfig
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
OBJ = PKG / "_build"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")
ARCH = "gfx950"

HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # No packed FP32 instructions in device code.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]
HOST_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-sign-compare", "-pthread"]

DEVICE_SOURCES = ["analysis.hip", "trsv_rows.hip", "trsv_blocks.hip", "trsv_panel.hip", "refine.hip", "grad.hip", "normalize.hip",
                  "probe.hip", "capi.hip"]
PLAN_SOURCES = ["block_plan.cpp"]   # host code linked into libcholespy_amd.so
HOST_SOURCES = ["sparse_input.cpp", "ordering.cpp", "factor.cpp", "supernodal.cpp", "cholesky_solver.cpp",
                "bindings.cpp"]

LIB_NAME = "libcholespy_amd.so"
EXT_NAME = "_cholespy_core" + sysconfig.get_config_var("EXT_SUFFIX")


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    # hipcc's host pass does not know the packed-fp32 feature; that note is noise.
    out = "\n".join(l for l in proc.stdout.splitlines()
                    if "packed-fp32-ops' is not a recognized feature" not in l)
    if proc.returncode != 0:
        raise RuntimeError(f"build step failed ({proc.returncode}): {' '.join(cmd)}\n{out}")
    if out.strip():
        print(out, file=sys.stderr)


def _pybind_includes() -> list[str]:
    import pybind11
    return [f"-I{pybind11.get_include()}", f"-I{sysconfig.get_paths()['include']}"]


def build(verbose: bool = False) -> dict:
    OBJ.mkdir(exist_ok=True)
    jobs = []
    for src in DEVICE_SOURCES:
        obj = OBJ / (Path(src).stem + ".hip.o")
        jobs.append([HIPCC, *HIP_FLAGS, "-c", str(CSRC / "device" / src), "-o", str(obj)])
    for src in PLAN_SOURCES:
        obj = OBJ / (Path(src).stem + ".plan.o")
        jobs.append([CXX, *HOST_FLAGS, "-c", str(CSRC / "device" / src), "-o", str(obj)])
    for src in HOST_SOURCES:
        obj = OBJ / (Path(src).stem + ".o")
        extra = _pybind_includes() + ["-fvisibility=hidden"] if src == "bindings.cpp" else []
        jobs.append([CXX, *HOST_FLAGS, *extra, "-c", str(CSRC / "host" / src), "-o", str(obj)])
    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 4)) as ex:
        for cmd, fut in [(c, ex.submit(_run, c)) for c in jobs]:
            if verbose:
                print(" ".join(cmd))
            fut.result()

    lib = PKG / LIB_NAME
    dev_objs = [str(OBJ / (Path(s).stem + ".hip.o")) for s in DEVICE_SOURCES]
    dev_objs += [str(OBJ / (Path(s).stem + ".plan.o")) for s in PLAN_SOURCES]
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-fPIC", *dev_objs, "-o", str(lib)])

    ext = PKG / EXT_NAME
    host_objs = [str(OBJ / (Path(s).stem + ".o")) for s in HOST_SOURCES]
    _run([CXX, "-shared", *host_objs, "-o", str(ext), f"-L{PKG}", "-lcholespy_amd",
          "-Wl,-rpath,$ORIGIN", "-pthread"])
    return {"lib": str(lib), "ext": str(ext)}


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
