/*
 * ORACLE -- test infrastructure only. CPU restatement of the reference's
 * solve-phase algorithms, used as the checker by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg. No product
 * code path may load, call or link it.
 *
 * Every function restates one reference routine, sequentially:
 *   oracle_levels        analysis_lower / analysis_upper   kernels/kernels.cu:8-60
 *   oracle_level_sort    level counting sort                src/cholesky_solver.cpp:348-367
 *   oracle_lower_{f,d}   solve_lower<Float>                 kernels/kernels.cu:64-131
 *   oracle_upper_{f,d}   solve_upper<Float>                 kernels/kernels.cu:133-202
 *   oracle_host_solve    cholmod_solve(CHOLMOD_A) on a simplicial LL' factor,
 *                        as called by solve_cpu           src/cholesky_solver.cpp:434-463
 *
 * Processing rows sequentially in ticket order is one valid schedule of the
 * reference's concurrent kernels; because every row's arithmetic only reads
 * rows with smaller tickets, the results equal those of any schedule.
 * The per-row accumulation mirrors the reference: sequential in ascending
 * CSR order for L (kernels/kernels.cu:90-116), descending for L'
 * (kernels/kernels.cu:158-184), each step `r -= val * x` contracted to a
 * fused multiply-add (nvcc's default), then one IEEE division by the
 * diagonal (kernels/kernels.cu:120, 188).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* Level of every row of a row-compressed triangle. lower != 0: rows of L
 * with the diagonal last, rows taken 0..n-1; lower == 0: rows of L' with the
 * diagonal first, rows taken n-1..0 (kernels/kernels.cu:9, 36). Returns the
 * maximum level (the reference's max_lvl, kernels/kernels.cu:27, 54). */
int oracle_levels(int n, const int *rowptr, const int *cols, int lower, int *row_level) {
    int max_lvl = 0;
    for (int t = 0; t < n; ++t) {
        const int row = lower ? t : n - 1 - t;
        const int start = rowptr[row];
        const int end = rowptr[row + 1] - 1;
        int lvl = 0;
        if (lower) {
            for (int i = start; i < end; ++i) {
                const int cl = row_level[cols[i]];
                if (lvl <= cl) lvl = cl + 1;
            }
        } else {
            for (int i = end; i > start; --i) {
                const int cl = row_level[cols[i]];
                if (lvl <= cl) lvl = cl + 1;
            }
        }
        if (lvl > max_lvl) max_lvl = lvl;
        row_level[row] = lvl;
    }
    return max_lvl;
}

/* Rows grouped by level, stable in row index: levels[k] = k-th row. */
void oracle_level_sort(int n, const int *row_level, int max_lvl, int *levels) {
    const int nl = max_lvl + 1;
    int *ptr = (int *)calloc((size_t)nl + 1, sizeof(int));
    for (int i = 0; i < n; ++i) ptr[1 + row_level[i]]++;
    for (int l = 0; l < nl; ++l) ptr[l + 1] += ptr[l];
    for (int i = 0; i < n; ++i) levels[ptr[row_level[i]]++] = i;
    free(ptr);
}

#define DEFINE_LOWER(NAME, T, FMA)                                                        \
    void NAME(int n, int k, const int *rowptr, const int *cols, const T *vals,           \
              const int *perm, const T *b, T *tmp) {                                       \
        for (int row = 0; row < n; ++row) {                                                \
            const int start = rowptr[row], end = rowptr[row + 1] - 1;                      \
            const T diag = vals[end];                                                      \
            for (int t = 0; t < k; ++t) {                                                  \
                T r = b[(size_t)perm[row] * k + t];                                        \
                for (int i = start; i < end; ++i)                                          \
                    r = FMA(-vals[i], tmp[(size_t)cols[i] * k + t], r);                    \
                tmp[(size_t)row * k + t] = r / diag;                                       \
            }                                                                              \
        }                                                                                  \
    }

#define DEFINE_UPPER(NAME, T, FMA)                                                        \
    void NAME(int n, int k, const int *rowptr, const int *cols, const T *vals,           \
              const int *perm, T *tmp, T *x) {                                             \
        for (int row = n - 1; row >= 0; --row) {                                           \
            const int start = rowptr[row], end = rowptr[row + 1] - 1;                      \
            const T diag = vals[start];                                                    \
            for (int t = 0; t < k; ++t) {                                                  \
                T r = tmp[(size_t)row * k + t];                                            \
                for (int i = end; i > start; --i)                                          \
                    r = FMA(-vals[i], tmp[(size_t)cols[i] * k + t], r);                    \
                const T res = r / diag;                                                    \
                tmp[(size_t)row * k + t] = res;                                            \
                x[(size_t)perm[row] * k + t] = res;                                        \
            }                                                                              \
        }                                                                                  \
    }

DEFINE_LOWER(oracle_lower_f, float, fmaf)
DEFINE_LOWER(oracle_lower_d, double, fma)
DEFINE_UPPER(oracle_upper_f, float, fmaf)
DEFINE_UPPER(oracle_upper_d, double, fma)

/* FP64 host solve with a simplicial LL' factor given column-compressed
 * (rows ascending, diagonal first): for each column-major right-hand side,
 * y = P b, forward with L by columns, backward with L' by columns,
 * x = P' y. bx (n x k, leading dimension n) is overwritten with x. */
void oracle_host_solve(int n, int k, const int *colptr, const int *rowidx, const double *val,
                       const int *perm, double *bx) {
    double *y = (double *)malloc((size_t)(n > 0 ? n : 1) * sizeof(double));
    for (int t = 0; t < k; ++t) {
        double *col = bx + (size_t)t * n;
        for (int i = 0; i < n; ++i) y[i] = col[perm[i]];
        for (int j = 0; j < n; ++j) {
            const int p0 = colptr[j];
            const double yj = y[j] / val[p0];
            y[j] = yj;
            for (int p = p0 + 1; p < colptr[j + 1]; ++p) y[rowidx[p]] -= val[p] * yj;
        }
        for (int j = n - 1; j >= 0; --j) {
            const int p0 = colptr[j];
            double s = y[j];
            for (int p = p0 + 1; p < colptr[j + 1]; ++p) s -= val[p] * y[rowidx[p]];
            y[j] = s / val[p0];
        }
        for (int i = 0; i < n; ++i) col[perm[i]] = y[i];
    }
    free(y);
}
