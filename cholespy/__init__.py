"""Drop-in import name: `from cholespy import CholeskySolverF, MatrixType`
resolves to the MI355X implementation in cholespy_amd."""
from cholespy_amd import CholeskySolverD, CholeskySolverF, MatrixType, __version__

__all__ = ["MatrixType", "CholeskySolverF", "CholeskySolverD", "__version__"]
