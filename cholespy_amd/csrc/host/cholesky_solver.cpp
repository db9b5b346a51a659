#include "cholesky_solver.hpp"

#include <algorithm>
#include <exception>
#include <stdexcept>
#include <string>
#include <thread>
#include <type_traits>

namespace cholespy_amd {

void throw_on_error(int status) {
    if (status == CHOLESPY_OK) return;
    const std::string msg = cholespy_last_error();
    if (status == CHOLESPY_ERR_INVALID_ARGUMENT) throw std::invalid_argument(msg);
    throw std::runtime_error(msg);
}

template <typename Float>
CholeskySolver<Float>::CholeskySolver(int n_rows, int64_t nnz, const int *ii, const int *jj,
                                      const double *x, MatrixType type, bool cpu)
    : n_(n_rows), cpu_(cpu) {
    A_ = normalize_input(n_rows, nnz, ii, jj, x, type, &entry_dst_);
    setup();
}

template <typename Float>
CholeskySolver<Float>::CholeskySolver(LowerCSC A, std::vector<int64_t> entry_dst)
    : n_(A.n), cpu_(false), A_(std::move(A)), entry_dst_(std::move(entry_dst)) {
    setup();
}

template <typename Float>
void CholeskySolver<Float>::setup() {
    symbolic_ = symbolic_factor(A_);
    factor_ = numeric_factor(symbolic_, A_.val);
    if (cpu_) return;

    // GPU tail of the reference's factorize() (src/cholesky_solver.cpp:244-297):
    // upload perm and L (and the block plan built from it). The level
    // analysis of L and L' runs on first use (analyse()): the default block
    // schedule does not need it.
    throw_on_error(cholespy_solver_create(n_, (int)sizeof(Float), &gpu_));
    try {
        throw_on_error(cholespy_solver_upload_factor(gpu_, factor_.perm.data(), factor_.colptr.data(),
                                                     factor_.rowidx.data(), factor_.val.data(),
                                                     factor_.nnz()));
        throw_on_error(cholespy_solver_upload_matrix(gpu_, A_.colptr.data(), A_.rowidx.data(),
                                                     A_.val.data(), A_.nnz()));
    } catch (...) {
        cholespy_solver_destroy(gpu_);
        gpu_ = nullptr;
        throw;
    }
}

template <typename Float>
void CholeskySolver<Float>::analyse() const {
    if (analysed_ || !gpu_) return;
    throw_on_error(cholespy_solver_analyze(gpu_, 1, &n_levels_lower_));
    throw_on_error(cholespy_solver_analyze(gpu_, 0, &n_levels_upper_));
    analysed_ = true;
}

template <typename Float>
CholeskySolver<Float>::~CholeskySolver() {
    if (gpu_) cholespy_solver_destroy(gpu_);
}

template <typename Float>
void CholeskySolver<Float>::solve_cuda(int n_rhs, const void *b, void *x, void *stream) {
    if (n_rhs > 128) throw std::invalid_argument("The number of RHS should be less than 128.");
    if (std::is_same<Float, float>::value)
        throw_on_error(cholespy_solver_solve_f32(gpu_, n_rhs, static_cast<const float *>(b),
                                                 static_cast<float *>(x), stream));
    else
        throw_on_error(cholespy_solver_solve_f64(gpu_, n_rhs, static_cast<const double *>(b),
                                                 static_cast<double *>(x), stream));
}

template <typename Float>
void CholeskySolver<Float>::solve_cpu(int n_rhs, const Float *b, Float *x) {
    if (n_rhs <= 0 || n_ == 0) return;
    // Row-major Float -> column-major FP64 (src/cholesky_solver.cpp:449-453).
    std::vector<double> buf((size_t)n_ * (size_t)n_rhs);
    for (int i = 0; i < n_; ++i)
        for (int j = 0; j < n_rhs; ++j) buf[(size_t)j * n_ + i] = (double)b[(size_t)i * n_rhs + j];
    factor_solve_colmajor(factor_, n_rhs, buf.data());
    for (int i = 0; i < n_; ++i)
        for (int j = 0; j < n_rhs; ++j) x[(size_t)i * n_rhs + j] = (Float)buf[(size_t)j * n_ + i];
}

template <typename Float>
void CholeskySolver<Float>::refactorize(int64_t nnz, const double *x) {
    if (nnz != input_nnz())
        throw std::invalid_argument("refactorize(): x must have as many entries as at construction.");
    std::vector<double> aval(A_.val.size(), 0.0);
    refill_values(entry_dst_, x, aval);
    Factor f = numeric_factor(symbolic_, aval);
    if (gpu_) {
        throw_on_error(cholespy_solver_refactor(gpu_, f.val.data(), f.nnz()));
        throw_on_error(cholespy_solver_upload_matrix(gpu_, A_.colptr.data(), A_.rowidx.data(), aval.data(),
                                                     A_.nnz()));
    }
    A_.val.swap(aval);
    factor_.val.swap(f.val);
}

template <typename Float>
void CholeskySolver<Float>::levels(bool lower, std::vector<int> &order,
                                   std::vector<int> &row_level) const {
    if (!gpu_) throw std::invalid_argument("Level analysis only exists for GPU solvers.");
    analyse();
    order.resize((size_t)n_);
    row_level.resize((size_t)n_);
    throw_on_error(cholespy_solver_get_levels(gpu_, lower ? 1 : 0, order.data(), row_level.data()));
}

template <typename Float>
void CholeskySolver<Float>::check() const {
    if (gpu_) throw_on_error(cholespy_solver_check(gpu_));
}

template <typename Float>
void CholeskySolver<Float>::solve_pass(int pass, int n_rhs, const void *b, void *x) {
    if (!gpu_) throw std::invalid_argument("Passes only exist for GPU solvers.");
    if (n_rhs > 128) throw std::invalid_argument("The number of RHS should be less than 128.");
    throw_on_error(cholespy_solver_solve_pass(gpu_, pass, n_rhs, b, x, nullptr));
}

template <typename Float>
void CholeskySolver<Float>::solve_refined(int n_rhs, const double *b, double *x, int iterations) {
    if (!gpu_) throw std::invalid_argument("Refined solves only exist for GPU solvers.");
    if (n_rhs > 128) throw std::invalid_argument("The number of RHS should be less than 128.");
    throw_on_error(cholespy_solver_solve_refined(gpu_, n_rhs, b, x, iterations, nullptr));
}

template <typename Float>
std::string CholeskySolver<Float>::autotune(int n_rhs, const std::string &cache_path, int flags) {
    if (!gpu_) throw std::invalid_argument("Autotuning only exists for GPU solvers.");
    if (n_rhs > 128) throw std::invalid_argument("The number of RHS should be less than 128.");
    char report[512] = {0};
    throw_on_error(cholespy_solver_autotune(gpu_, factor_.val.data(), factor_.nnz(), n_rhs,
                                            cache_path.empty() ? nullptr : cache_path.c_str(), flags, report,
                                            sizeof report));
    return report;
}

template <typename Float>
void CholeskySolver<Float>::set_path(int path) {
    if (!gpu_) throw std::invalid_argument("Solve paths only exist for GPU solvers.");
    if (path == CHOLESPY_PATH_ROWS) analyse();
    throw_on_error(cholespy_solver_set_path(gpu_, path));
}

template <typename Float>
std::vector<int64_t> CholeskySolver<Float>::plan_info() const {
    if (!gpu_) throw std::invalid_argument("Block plans only exist for GPU solvers.");
    std::vector<int64_t> info(10, 0);
    throw_on_error(cholespy_solver_plan_info(gpu_, info.data(), (int)info.size()));
    return info;
}

template class CholeskySolver<float>;
template class CholeskySolver<double>;

template <typename Float>
BatchedCholeskySolver<Float>::BatchedCholeskySolver(const std::vector<SystemInput> &systems) {
    const int ns = (int)systems.size();
    if (ns == 0) throw std::invalid_argument("A batch needs at least one system.");
    n_.resize(ns);
    factors_.resize(ns);
    for (int i = 0; i < ns; ++i) n_[i] = systems[i].n;
    // Normalise and factor on host threads; the first failure is re-raised.
    std::vector<std::exception_ptr> errors(ns);
    const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> pool;
    for (unsigned w = 0; w < hw; ++w)
        pool.emplace_back([&, w]() {
            for (int i = (int)w; i < ns; i += (int)hw) {
                try {
                    const SystemInput &in = systems[i];
                    LowerCSC A = normalize_input(in.n, (int64_t)in.x.size(), in.ii.data(), in.jj.data(),
                                                 in.x.data(), in.type);
                    factors_[i] = cholesky_factor(A);
                } catch (...) {
                    errors[i] = std::current_exception();
                }
            }
        });
    for (auto &t : pool) t.join();
    for (auto &e : errors)
        if (e) std::rethrow_exception(e);
    throw_on_error(cholespy_batch_create(ns, n_.data(), (int)sizeof(Float), &gpu_));
    try {
        for (int i = 0; i < ns; ++i) {
            const Factor &f = factors_[i];
            throw_on_error(cholespy_batch_set_factor(gpu_, i, f.perm.data(), f.colptr.data(), f.rowidx.data(),
                                                     f.val.data(), f.nnz()));
        }
        throw_on_error(cholespy_batch_finalize(gpu_));
    } catch (...) {
        cholespy_solver_destroy(gpu_);
        gpu_ = nullptr;
        throw;
    }
}

template <typename Float>
BatchedCholeskySolver<Float>::~BatchedCholeskySolver() {
    if (gpu_) cholespy_solver_destroy(gpu_);
}

template <typename Float>
void BatchedCholeskySolver<Float>::solve(const std::vector<int> &n_rhs, const void *b, void *x, void *stream) {
    if ((int)n_rhs.size() != count()) throw std::invalid_argument("n_rhs needs one entry per system.");
    for (int k : n_rhs)
        if (k > 128) throw std::invalid_argument("The number of RHS should be less than 128.");
    if (std::is_same<Float, float>::value)
        throw_on_error(cholespy_batch_solve_f32(gpu_, n_rhs.data(), static_cast<const float *>(b),
                                                static_cast<float *>(x), stream));
    else
        throw_on_error(cholespy_batch_solve_f64(gpu_, n_rhs.data(), static_cast<const double *>(b),
                                                static_cast<double *>(x), stream));
}

template class BatchedCholeskySolver<float>;
template class BatchedCholeskySolver<double>;

}  // namespace cholespy_amd
