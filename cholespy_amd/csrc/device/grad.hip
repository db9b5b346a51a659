// Gradient of a loss with respect to the stored entries of A, for
// x = A^{-1} b with adjoint lam = A^{-1} dL/dx (A symmetric, so the adjoint
// solve reuses the same factor):  dL/dA = -lam x'.
//
// The factorisation reads only the lower triangle of what the caller passes
// (CHOLMOD stype = -1, src/cholesky_solver.cpp:211-220), i.e. the effective
// matrix is A_eff = tril(A) + tril(A, -1)'. For a stored entry e at (i, j):
//   i >  j : it is both A_eff(i, j) and A_eff(j, i):
//            dL/dv_e = -sum_t (lam[i,t] x[j,t] + lam[j,t] x[i,t])
//   i == j : dL/dv_e = -sum_t lam[i,t] x[i,t]
//   i <  j : never read, dL/dv_e = 0
// Duplicate entries each get the full gradient of their position (A_eff sums
// them). One thread per stored entry; lam and x are row-major (n x k). An
// entry whose row or column lies outside [0, n) is not read through: its
// gradient is NaN (the callers check the index range first).
#include "common.hpp"
#include "launchers.hpp"

namespace cholespy_amd {

namespace {

template <typename V, typename G>
__global__ void entry_gradient(int n, int64_t nnz, const int *__restrict__ rows, const int *__restrict__ cols,
                               int k, const V *__restrict__ lam, const V *__restrict__ x,
                               G *__restrict__ grad) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < nnz;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int i = rows[e], j = cols[e];
        double g = 0.0;
        if ((unsigned)i >= (unsigned)n || (unsigned)j >= (unsigned)n) {
            g = __builtin_nan("");
        } else if (i >= j) {
            const V *li = lam + (size_t)i * k, *xi = x + (size_t)i * k;
            const V *lj = lam + (size_t)j * k, *xj = x + (size_t)j * k;
            if (i == j) {
                for (int t = 0; t < k; ++t) g -= (double)li[t] * (double)xi[t];
            } else {
                for (int t = 0; t < k; ++t) g -= (double)li[t] * (double)xj[t] + (double)lj[t] * (double)xi[t];
            }
        }
        grad[e] = (G)g;
    }
}

template <typename V, typename G>
hipError_t launch(int n, int64_t nnz, const int *rows, const int *cols, int k, const void *lam, const void *x,
                  void *grad, hipStream_t stream) {
    const int64_t want = (nnz + 255) / 256;
    const int grid = (int)(want < 8192 ? (want > 0 ? want : 1) : 8192);
    entry_gradient<V, G><<<grid, 256, 0, stream>>>(n, nnz, rows, cols, k, static_cast<const V *>(lam),
                                                    static_cast<const V *>(x), static_cast<G *>(grad));
    return hipGetLastError();
}

}  // namespace

hipError_t launch_entry_gradient(int n, int64_t nnz, const int *rows, const int *cols, int k, const void *lam,
                                 const void *x, int vec_bytes, void *grad, int grad_bytes,
                                 hipStream_t stream) {
    if (nnz == 0) return hipSuccess;
    if (vec_bytes == 4)
        return grad_bytes == 4 ? launch<float, float>(n, nnz, rows, cols, k, lam, x, grad, stream)
                               : launch<float, double>(n, nnz, rows, cols, k, lam, x, grad, stream);
    return grad_bytes == 4 ? launch<double, float>(n, nnz, rows, cols, k, lam, x, grad, stream)
                           : launch<double, double>(n, nnz, rows, cols, k, lam, x, grad, stream);
}

}  // namespace cholespy_amd
