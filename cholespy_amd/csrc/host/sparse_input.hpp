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
// When entry_dst is given it receives, for every input entry (position in x),
// the position in the result's val that the entry is summed into, or -1 for
// entries above the diagonal: the map a numeric refactorisation uses to
// rebuild val from new values in the same order (refill_values).
LowerCSC normalize_input(int n, int64_t nnz, const int *ii, const int *jj,
                         const double *x, MatrixType type,
                         std::vector<int64_t> *entry_dst = nullptr);

// val[q] := sum of x[e] over the entries e with entry_dst[e] == q, added in
// input order starting from the first such entry: the same sums, bit for
// bit, that normalize_input forms from these values.
void refill_values(const std::vector<int64_t> &entry_dst, const double *x, std::vector<double> &val);

}  // namespace cholespy_amd
