// CholeskySolver<Float>: the host object behind CholeskySolverF / D.
// Mirrors src/cholesky_solver.h:22-99: the constructor normalises the input
// and factors on the host; in GPU mode the factor goes through the C ABI of
// libcholespy_amd (include/cholespy_amd.h) for the level analysis and the
// triangular solves; in CPU mode solves stay on the host in FP64.
#pragma once
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/cholespy_amd.h"
#include "factor.hpp"
#include "sparse_input.hpp"

namespace cholespy_amd {

template <typename Float>
class CholeskySolver {
public:
    // src/cholesky_solver.h:36 / src/cholesky_solver.cpp:151-198
    CholeskySolver(int n_rows, int64_t nnz, const int *ii, const int *jj, const double *x,
                   MatrixType type, bool cpu);
    // GPU mode from an input already normalised on the device
    // (cholespy_normalize_input): A's lower CSC and the input entry map.
    CholeskySolver(LowerCSC A, std::vector<int64_t> entry_dst);
    ~CholeskySolver();
    CholeskySolver(const CholeskySolver &) = delete;
    CholeskySolver &operator=(const CholeskySolver &) = delete;

    // src/cholesky_solver.cpp:415-432. b and x are device addresses of
    // row-major (n x n_rhs) arrays; asynchronous on `stream` (a
    // hipStream_t, nullptr = the null stream, as the reference).
    void solve_cuda(int n_rhs, const void *b, void *x, void *stream = nullptr);

    // src/cholesky_solver.cpp:434-463: row-major Float in/out, FP64 inside.
    void solve_cpu(int n_rhs, const Float *b, Float *x);

    // Numeric refactorisation for new values x of the constructor's pattern:
    // nnz values in the order of the constructor's x (same ii / jj, same
    // matrix type). Reuses the ordering, the symbolic factor, the level
    // analysis and the block plan's structure; the result equals a new
    // solver built from the same ii / jj and x. No counterpart in the
    // reference, which builds a new solver for every new matrix
    // (tests/Benchmark.ipynb:238-254). A matrix that is not positive definite
    // raises and leaves the solver unchanged.
    void refactorize(int64_t nnz, const double *x);
    int64_t input_nnz() const { return (int64_t)entry_dst_.size(); }

    bool is_cpu() const { return cpu_; }
    int n() const { return n_; }
    const Factor &factor() const { return factor_; }
    const LowerCSC &matrix() const { return A_; }
    int n_levels(bool lower) const {
        analyse();
        return lower ? n_levels_lower_ : n_levels_upper_;
    }
    // Level-sorted rows and per-row levels as produced by the GPU analysis.
    void levels(bool lower, std::vector<int> &order, std::vector<int> &row_level) const;
    // Waits for the device and raises if a kernel reported a timeout.
    void check() const;
    // One pass of solve_cuda (0 = forward, 1 = backward).
    void solve_pass(int pass, int n_rhs, const void *b, void *x);
    // FP32 solve + FP64 refinement on the GPU (float32 solvers): b, x are
    // device addresses of row-major double arrays.
    void solve_refined(int n_rhs, const double *b, double *x, int iterations);
    // Device autotuning for n_rhs right-hand sides (cholespy_solver_autotune);
    // returns the summary line.
    std::string autotune(int n_rhs, const std::string &cache_path, int flags);
    // Selects the GPU schedule (CHOLESPY_PATH_ROWS / CHOLESPY_PATH_BLOCKS).
    void set_path(int path);
    std::vector<int64_t> plan_info() const;
    // The C ABI handle (GPU mode only; throws in CPU mode).
    cholespy_gpu_solver *gpu() const {
        if (!gpu_) throw std::invalid_argument("Only GPU solvers have a device handle.");
        return gpu_;
    }

private:
    int n_ = 0;
    bool cpu_ = true;
    LowerCSC A_;
    std::vector<int64_t> entry_dst_;   // input entry -> position in A_.val (-1: dropped)
    SymbolicFactor symbolic_;
    Factor factor_;
    cholespy_gpu_solver *gpu_ = nullptr;
    // The level analysis of L and L' (src/cholesky_solver.cpp:299-376) is
    // only needed by the row schedule and by levels(): it runs on first use.
    void analyse() const;
    void setup();   // host factor and, in GPU mode, the device upload
    mutable bool analysed_ = false;
    mutable int n_levels_lower_ = 0, n_levels_upper_ = 0;
};

// One system of a batch, as given to the constructor (host arrays).
struct SystemInput {
    int n = 0;
    MatrixType type = MatrixType::COO;
    std::vector<int> ii, jj;
    std::vector<double> x;
};

// Many independent systems factored on host threads and solved together on
// the GPU, one launch per pass (cholespy_batch_* in the C ABI).
template <typename Float>
class BatchedCholeskySolver {
public:
    explicit BatchedCholeskySolver(const std::vector<SystemInput> &systems);
    ~BatchedCholeskySolver();
    BatchedCholeskySolver(const BatchedCholeskySolver &) = delete;
    BatchedCholeskySolver &operator=(const BatchedCholeskySolver &) = delete;
    // b, x: packed device arrays (system i: row-major n_i x n_rhs[i]);
    // enqueued on `stream` (nullptr = null stream).
    void solve(const std::vector<int> &n_rhs, const void *b, void *x, void *stream = nullptr);
    int count() const { return (int)n_.size(); }
    const std::vector<int> &sizes() const { return n_; }
    const Factor &factor(int i) const { return factors_.at(i); }
    cholespy_gpu_solver *gpu() const { return gpu_; }

private:
    std::vector<int> n_;
    std::vector<Factor> factors_;
    cholespy_gpu_solver *gpu_ = nullptr;
};

// Raises the C++ exception matching a C ABI status (invalid argument ->
// std::invalid_argument, anything else -> std::runtime_error).
void throw_on_error(int status);

}  // namespace cholespy_amd
