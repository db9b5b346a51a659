// Input normalisation for CholeskySolver: COO / CSR / CSC -> lower-triangular CSC.
//
// Semantics follow the reference's host path:
//   coo_to_csc / csr_to_csc        src/cholesky_solver.cpp:8-92
//   csc_sort_indices               src/cholesky_solver.cpp:95-116
//   csc_sum_duplicates             src/cholesky_solver.cpp:119-149
//   CHOLMOD stype = -1             src/cholesky_solver.cpp:211-220
// With stype = -1 the factorisation only reads the lower triangle (row >= col),
// so entries strictly above the diagonal are dropped here. Unlike the
// reference, the caller's arrays are never sorted or reallocated in place.
#pragma once
#include <cstdint>
#include <vector>

namespace cholespy_amd {

enum class MatrixType : int { CSC = 0, CSR = 1, COO = 2 };

// Lower triangle of a symmetric matrix in compressed-column form. Row indices
// are strictly increasing inside every column and duplicates are summed.
struct LowerCSC {
    int n = 0;
    std::vector<int> colptr;   // n + 1
    std::vector<int> rowidx;   // nnz
    std::vector<double> val;   // nnz
    int64_t nnz() const { return (int64_t)rowidx.size(); }
};

// Builds the lower CSC of A from the caller's arrays.
//   COO: ii = row indices, jj = column indices, both of length nnz
//   CSR: ii = row pointer (n + 1), jj = column indices (nnz)
//   CSC: ii = column pointer (n + 1), jj = row indices (nnz)
// Throws std::invalid_argument on indices outside [0, n) or malformed pointers.
LowerCSC normalize_input(int n, int64_t nnz, const int *ii, const int *jj,
                         const double *x, MatrixType type);

}  // namespace cholespy_amd
