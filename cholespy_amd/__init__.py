"""cholespy_amd: cholespy's API with an MI355X-native (gfx950) GPU solve phase.

Exports the same names as cholespy (MatrixType, CholeskySolverF,
CholeskySolverD, __version__). The native module is built in-tree
(`python cholespy_amd/build.py`); importing without it raises ImportError,
there is no Python fallback.
"""
from importlib import import_module as _import_module

# PyTorch-ROCm bundles its own HIP runtime under the same SONAME
# (libamdhip64.so.7) as the system ROCm one this extension links. Only the
# first one loaded is used by the whole process, and a process that mixes
# them sees no device; loading torch first makes its runtime the one both
# use. Without torch the system runtime is loaded.
try:  # pragma: no cover - depends on the environment
    import torch as _torch  # noqa: F401
except ImportError:  # pragma: no cover
    _torch = None

try:
    _core = _import_module("cholespy_amd._cholespy_core")
except ImportError as exc:  # pragma: no cover - exercised only on a broken build
    raise ImportError(
        "cholespy_amd native extension is not built; run `python cholespy_amd/build.py` "
        f"({exc})") from exc

MatrixType = _core.MatrixType
CholeskySolverF = _core.CholeskySolverF
CholeskySolverD = _core.CholeskySolverD
BatchedSolverF = _core.BatchedSolverF
BatchedSolverD = _core.BatchedSolverD
__version__ = _core.__version__

__all__ = ["MatrixType", "CholeskySolverF", "CholeskySolverD", "BatchedSolverF", "BatchedSolverD",
           "__version__"]
