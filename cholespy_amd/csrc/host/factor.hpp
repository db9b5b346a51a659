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

// Throws std::invalid_argument("Matrix is not positive definite!") when a
// pivot is not strictly positive (src/cholesky_solver.cpp:241-242).
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
