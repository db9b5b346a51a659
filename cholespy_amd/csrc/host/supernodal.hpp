// Supernodal multifrontal numeric Cholesky for the host factorisation
// (the numeric phase CHOLMOD runs for the reference, src/cholesky_solver.cpp:
// 235-243; CHOLMOD itself is not available here). Written from the textbook
// method (Duff & Reid 1983; Liu, "The multifrontal method for sparse matrix
// solution", SIAM Review 34, 1992; Davis, "Direct Methods for Sparse Linear
// Systems", 2006, ch. 4):
//   * fundamental supernodes: column j+1 joins column j's supernode when
//     j+1 is j's etree parent, j is its only child and the column counts
//     differ by one, so the supernodal L has exactly the simplicial pattern;
//   * one dense frontal matrix per supernode, assembled from the permuted
//     matrix and the children's update matrices (extend-add), partially
//     factored by blocked POTRF / TRSM / SYRK on a register-tiled FMA kernel;
//   * independent subtrees factored by worker threads; the trailing update of
//     large fronts is split across threads.
#pragma once
#include <cstdint>
#include <vector>

namespace cholespy_amd {

struct Factor;

// Fills f.rowidx / f.val (f.colptr already sized from the column counts) with
// L of P A P', given the lower CSC of P A P' (rows ascending), the
// elimination tree `parent` and the column counts (diagonal included).
// Throws std::invalid_argument("Matrix is not positive definite!") when a
// pivot is not strictly positive. threads <= 0 picks the hardware count.
void supernodal_numeric(int n, const std::vector<int> &colptr, const std::vector<int> &rowidx,
                        const std::vector<double> &val, const std::vector<int> &parent,
                        const std::vector<int64_t> &count, Factor &f, int threads = 0);

}  // namespace cholespy_amd
