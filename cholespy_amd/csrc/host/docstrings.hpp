// Python docstrings of the binding. Same API as the reference
// (src/docstr.h:1-48), described in this project's own words.
#pragma once

static const char *doc_constructor = R"(
Factor a sparse symmetric positive-definite matrix given in COO, CSR or CSC
form and prepare its triangular solves.

Parameters
----------

- n_rows - Number of rows (and columns) of the square matrix.
- ii - COO: row index of every entry. CSR / CSC: the row / column pointer
       array of length n_rows + 1; the entries of row (column) k live at
       positions ii[k]:ii[k+1] of jj and x.
- jj - COO: column index of every entry. CSR: column indices. CSC: row indices.
- x - Value of every entry. Repeated (i, j) pairs are summed; only the lower
      triangle (i >= j) is read, the matrix is taken to be symmetric.
- type - A MatrixType value: MatrixType.COO, MatrixType.CSR or MatrixType.CSC.

The device of the input arrays decides where solve() runs for the lifetime of
the object: CPU arrays give a host solver, GPU (ROCm) arrays a GPU solver.
)";

static const char *doc_matrix_type =
    "Sparse storage format of the arrays given to CholeskySolverF / CholeskySolverD.";

static const char *doc_cholesky_f = "Sparse Cholesky solver whose solves run in single precision.";

static const char *doc_cholesky_d = "Sparse Cholesky solver whose solves run in double precision.";

static const char *doc_solve = R"(

Solve A x = b for one or more right-hand sides.

Parameters
----------

- `b` - Right-hand side(s): a vector of length n_rows or a matrix of shape
        (n_rows, n_rhs), on the device the solver was built for. GPU solves
        accept at most 128 columns.
- `x` - Output array with the shape and device of `b`; it is overwritten.
- `stream` - GPU solves only: the HIP stream to launch on, as an integer
        handle (`torch.cuda.current_stream().cuda_stream`). 0, the
        default, is the null stream, as in cholespy. A solve on a stream
        can be captured in a CUDA/HIP graph once the same number of
        right-hand sides has been solved outside the capture (the first
        solve at a new width sizes the work buffers).

Both arrays must already have the solver's dtype (float32 for
CholeskySolverF, float64 for CholeskySolverD) and be C-contiguous: `x` is
written in place, so no conversion is made.
)";

static const char *doc_refactorize = R"(

Refactorize for new values of the matrix with the same pattern.

Parameters
----------

- `x` - New nonzero values: as many as the constructor's `x`, in the same
        order (the constructor's `ii`, `jj` and matrix type are kept). CPU or
        GPU array, converted to float64.

The ordering, the symbolic factor, the GPU level analysis and the block
plan's structure are reused; only the numeric factor and the dense block
inverses are recomputed. The result equals a new solver built from the same
`ii`, `jj` and this `x`. Waits for pending solves of this device. Raises
ValueError("Matrix is not positive definite!") and leaves the solver
unchanged when the new matrix is not positive definite.
)";
