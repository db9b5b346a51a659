// Dense column-major double kernels for the host factorisation
// (supernodal.cpp) and the block plan (block_plan.cpp). The product kernel
// is an 8 x 4 register tile of C updated with one rank-1 term per step (two
// 4-wide loads of A, four broadcasts of B, eight FMAs), with row panels of A
// blocked to stay in cache. It is compiled twice, for AVX2+FMA (chosen at
// run time) and for the baseline instruction set.
#pragma once
#include <algorithm>
#include <cstring>
#include <functional>
#include <thread>
#include <vector>

namespace cholespy_amd {
namespace dense {

// Runs fn(part) for part in [0, parts) on up to `threads` threads.
inline void parallel_parts(int parts, int threads, const std::function<void(int)> &fn) {
    if (threads <= 1 || parts <= 1) {
        for (int p = 0; p < parts; ++p) fn(p);
        return;
    }
    std::vector<std::thread> pool;
    const int t = std::min(threads, parts);
    for (int w = 1; w < t; ++w)
        pool.emplace_back([&, w] {
            for (int p = w; p < parts; p += t) fn(p);
        });
    for (int p = 0; p < parts; p += t) fn(p);
    for (auto &th : pool) th.join();
}

// Bounds of part `part` of `parts` equal ranges of [0, total), rounded to
// multiples of `align`.
inline int part_bound(int total, int parts, int part, int align) {
    if (part >= parts) return total;
    const int b = (int)((int64_t)total * part / parts);
    return std::min(total, (b + align - 1) / align * align);
}

typedef double v4d __attribute__((vector_size(32)));

constexpr int kMR = 8, kNR = 4, kMC = 128;

// C(m x n) -= A(m x k) B(n x k)^T. Lower: C is a diagonal block (m == n) and
// only tiles that reach the lower triangle are computed (tiles straddling the
// diagonal also write above it; callers ignore the strict upper triangle).
template <bool Lower>
static inline __attribute__((always_inline)) void gemm_nt_body(int m, int n, int k, const double *A, int lda,
                                                              const double *B, int ldb, double *C, int ldc) {
    for (int ib = 0; ib < m; ib += kMC) {
        const int ie = std::min(m, ib + kMC);
        const int jend = Lower ? std::min(n, ie) : n;
        int j0 = 0;
        for (; j0 + kNR <= jend; j0 += kNR) {
            int i0 = ib;
            if (Lower) i0 = std::max(ib, (j0 / kMR) * kMR);
            for (; i0 + kMR <= ie; i0 += kMR) {
                v4d c0[kNR], c1[kNR];
                for (int q = 0; q < kNR; ++q) c0[q] = c1[q] = v4d{0.0, 0.0, 0.0, 0.0};
                const double *a = A + i0;
                const double *b = B + j0;
                for (int p = 0; p < k; ++p) {
                    v4d a0, a1;
                    std::memcpy(&a0, a + (size_t)p * lda, sizeof(v4d));
                    std::memcpy(&a1, a + (size_t)p * lda + 4, sizeof(v4d));
                    const double *bp = b + (size_t)p * ldb;
                    for (int q = 0; q < kNR; ++q) {
                        const double bv = bp[q];
                        c0[q] += a0 * bv;
                        c1[q] += a1 * bv;
                    }
                }
                for (int q = 0; q < kNR; ++q) {
                    double *cp = C + (size_t)(j0 + q) * ldc + i0;
                    v4d x0, x1;
                    std::memcpy(&x0, cp, sizeof(v4d));
                    std::memcpy(&x1, cp + 4, sizeof(v4d));
                    x0 -= c0[q];
                    x1 -= c1[q];
                    std::memcpy(cp, &x0, sizeof(v4d));
                    std::memcpy(cp + 4, &x1, sizeof(v4d));
                }
            }
            for (; i0 < ie; ++i0)
                for (int q = 0; q < kNR; ++q) {
                    double s = 0.0;
                    for (int p = 0; p < k; ++p) s += A[i0 + (size_t)p * lda] * B[j0 + q + (size_t)p * ldb];
                    C[i0 + (size_t)(j0 + q) * ldc] -= s;
                }
        }
        for (; j0 < jend; ++j0)
            for (int i = Lower ? std::max(ib, j0) : ib; i < ie; ++i) {
                double s = 0.0;
                for (int p = 0; p < k; ++p) s += A[i + (size_t)p * lda] * B[j0 + (size_t)p * ldb];
                C[i + (size_t)j0 * ldc] -= s;
            }
    }
}

__attribute__((target("avx2,fma"))) inline void gemm_nt_avx2(bool lower, int m, int n, int k, const double *A,
                                                             int lda, const double *B, int ldb, double *C, int ldc) {
    if (lower)
        gemm_nt_body<true>(m, n, k, A, lda, B, ldb, C, ldc);
    else
        gemm_nt_body<false>(m, n, k, A, lda, B, ldb, C, ldc);
}
inline void gemm_nt_base(bool lower, int m, int n, int k, const double *A, int lda, const double *B, int ldb,
                         double *C, int ldc) {
    if (lower)
        gemm_nt_body<true>(m, n, k, A, lda, B, ldb, C, ldc);
    else
        gemm_nt_body<false>(m, n, k, A, lda, B, ldb, C, ldc);
}

using GemmFn = void (*)(bool, int, int, int, const double *, int, const double *, int, double *, int);
inline GemmFn gemm_nt() {
    static const GemmFn fn = [] {
        __builtin_cpu_init();
        return (__builtin_cpu_supports("avx2") && __builtin_cpu_supports("fma")) ? &gemm_nt_avx2 : &gemm_nt_base;
    }();
    return fn;
}

// X(m x n) := X L^{-T}, L (n x n) lower with a nonzero diagonal. Blocked:
// kBlock columns of X solved by column operations, then the remaining
// columns updated by one product.
inline void trsm_right_lower_trans(int m, int n, const double *L, int ldl, double *X, int ldx) {
    constexpr int kBlock = 64;
    const GemmFn gemm = gemm_nt();
    for (int j0 = 0; j0 < n; j0 += kBlock) {
        const int j1 = std::min(n, j0 + kBlock);
        for (int j = j0; j < j1; ++j) {
            double *xj = X + (size_t)j * ldx;
            for (int p = j0; p < j; ++p) {
                const double l = L[j + (size_t)p * ldl];
                const double *xp = X + (size_t)p * ldx;
                for (int i = 0; i < m; ++i) xj[i] -= l * xp[i];
            }
            const double d = L[j + (size_t)j * ldl];
            for (int i = 0; i < m; ++i) xj[i] /= d;
        }
        if (j1 < n) gemm(false, m, n - j1, j1 - j0, X + (size_t)j0 * ldx, ldx, L + j1 + (size_t)j0 * ldl, ldl,
                         X + (size_t)j1 * ldx, ldx);
    }
}

// X := L^{-T} for L (m x m) lower with a nonzero diagonal, X holding the
// identity on entry. Column j of the result is zero below row j, so the
// column operations and the block updates touch rows 0..j only: a third of
// the work of trsm_right_lower_trans on the identity, with the same
// operations on every entry that is not structurally zero.
// threads > 1 splits each block update by columns (the same operations per
// entry, so the result does not depend on the thread count).
inline void lower_inverse_trans(int m, const double *L, int ldl, double *X, int ldx, int threads = 1) {
    constexpr int kBlock = 64;
    const GemmFn gemm = gemm_nt();
    for (int j0 = 0; j0 < m; j0 += kBlock) {
        const int j1 = std::min(m, j0 + kBlock);
        for (int j = j0; j < j1; ++j) {
            double *xj = X + (size_t)j * ldx;
            for (int p = j0; p < j; ++p) {
                const double l = L[j + (size_t)p * ldl];
                const double *xp = X + (size_t)p * ldx;
                for (int i = 0; i <= p; ++i) xj[i] -= l * xp[i];
            }
            const double d = L[j + (size_t)j * ldl];
            for (int i = 0; i <= j; ++i) xj[i] /= d;
        }
        if (j1 >= m) continue;
        const int cols = m - j1;
        const int parts = threads > 1 ? std::min(threads, (cols + 63) / 64) : 1;
        parallel_parts(parts, parts, [&](int q) {
            const int c0 = part_bound(cols, parts, q, kNR), c1 = part_bound(cols, parts, q + 1, kNR);
            if (c1 > c0)
                gemm(false, j1, c1 - c0, j1 - j0, X + (size_t)j0 * ldx, ldx, L + j1 + c0 + (size_t)j0 * ldl, ldl,
                     X + (size_t)(j1 + c0) * ldx, ldx);
        });
    }
}

// C(r x m) -= A(r x m) L^{-1}, with L^{-1} given as X = L^{-T} (column-major,
// as lower_inverse_trans leaves it): entry (p, j) of L^{-1} is X[j + p ldx]
// and zero for p < j, so the column panel [j0, j0 + w) of C reduces over
// p >= j0 only (half the work of the full product).
// threads > 1 splits the rows of C.
inline void gemm_lower_inverse(int r, int m, const double *A, int lda, const double *X, int ldx, double *C,
                               int ldc, int threads = 1) {
    constexpr int kPanel = 64;
    const GemmFn gemm = gemm_nt();
    const int parts = threads > 1 ? std::min(threads, (r + 63) / 64) : 1;
    parallel_parts(parts, parts, [&](int q) {
        const int r0 = part_bound(r, parts, q, kMR), r1 = part_bound(r, parts, q + 1, kMR);
        if (r1 <= r0) return;
        for (int j0 = 0; j0 < m; j0 += kPanel) {
            const int w = std::min(m, j0 + kPanel) - j0;
            gemm(false, r1 - r0, w, m - j0, A + r0 + (size_t)j0 * lda, lda, X + j0 + (size_t)j0 * ldx, ldx,
                 C + r0 + (size_t)j0 * ldc, ldc);
        }
    });
}

}  // namespace dense
}  // namespace cholespy_amd
