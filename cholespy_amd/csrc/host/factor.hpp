// Host sparse Cholesky factorisation in the role CHOLMOD plays for the
// reference (src/cholesky_solver.cpp:200-243): simplicial L L' = P A P' with a
// fill-reducing nested-dissection ordering followed by an elimination-tree
// postorder. Written from the textbook algorithms (Liu 1986 elimination tree;
// Davis, "Direct Methods for Sparse Linear Systems", SIAM 2006, ch. 4
// up-looking Cholesky with row subtrees); CHOLMOD is not available here.
#pragma once
#include <cstdint>
#include <vector>

#include "sparse_input.hpp"

namespace cholespy_amd {

struct Factor {
    int n = 0;
    // perm[k] = row of A that becomes row k of P A P' (CHOLMOD's L->Perm).
    std::vector<int> perm;
    // L in compressed-column form; rows ascending, diagonal first per column.
    std::vector<int> colptr;
    std::vector<int> rowidx;
    std::vector<double> val;
    // Elimination tree of P A P' (parent[j] = -1 for roots).
    std::vector<int> parent;
    int64_t nnz() const { return (int64_t)rowidx.size(); }
};

struct FactorOptions {
    // Use the identity ordering instead of nested dissection (tests only).
    bool natural = false;
    int nd_leaf_size = 48;
    double nd_min_balance = 0.25;
    // Numeric phase: simplicial up-looking (exact row-by-row), supernodal
    // multifrontal (supernodal.hpp), or auto = supernodal once L has at
    // least supernodal_min_nnz entries. CHOLESPY_FACTOR=simplicial|supernodal
    // overrides auto. Both give the same pattern of L and values equal to
    // rounding.
    enum class Method { Auto, Simplicial, Supernodal } method = Method::Auto;
    int64_t supernodal_min_nnz = 200000;
    int threads = 0;   // supernodal worker threads (0 = hardware count, at most 16)
};

// Everything of the factorisation that depends only on the pattern of A
// (CHOLMOD's symbolic factor, cholmod_analyze at src/cholesky_solver.cpp:226):
// the ordering, the elimination tree, the column counts of L and the maps
// that place A's values in the permuted matrix. A numeric factorisation
// (numeric_factor) of any matrix with A's pattern reuses it.
struct SymbolicFactor {
    int n = 0;
    std::vector<int> perm, parent;      // as in Factor
    std::vector<int> colptr;            // column pointers of L
    std::vector<int64_t> count;         // column counts of L (diagonal included)
    // Upper CSC C of P A P' (column k holds the rows i <= k) and, for each
    // entry of A's lower CSC, its position in C.
    std::vector<int> c_colptr, c_rowidx, a_to_c;
    // Supernodal numeric phase: the lower CSC of P A P' (rows ascending) and,
    // for each entry of C, its position there.
    std::vector<int> l_colptr, l_rowidx, c_to_l;
    bool supernodal = false;
    int threads = 0;
};

SymbolicFactor symbolic_factor(const LowerCSC &A, const FactorOptions &opt = FactorOptions());

// L of P A P' for the values a_val of a matrix with the symbolic factor's
// pattern (a_val in the order of A's lower CSC). Throws
// std::invalid_argument("Matrix is not positive definite!") when a pivot is
// not strictly positive (src/cholesky_solver.cpp:241-242).
Factor numeric_factor(const SymbolicFactor &S, const std::vector<double> &a_val);

// symbolic_factor followed by numeric_factor.
Factor cholesky_factor(const LowerCSC &A, const FactorOptions &opt = FactorOptions());

// Row-compressed copy of L: rows ascending, columns ascending, diagonal last.
void factor_rows(const Factor &f, std::vector<int> &rowptr, std::vector<int> &colidx,
                 std::vector<double> &val);

// x := A^{-1} b on the host in FP64 with column-major right-hand sides
// (n x k, leading dimension n): permute, forward L, backward L', un-permute.
// This is the work cholmod_solve(CHOLMOD_A) does for a simplicial LL' factor
// (src/cholesky_solver.cpp:455).
void factor_solve_colmajor(const Factor &f, int k, double *bx);

}  // namespace cholespy_amd
