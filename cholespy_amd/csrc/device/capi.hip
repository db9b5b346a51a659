// C ABI of the gfx950 solve phase (declared in include/cholespy_amd.h).
//
// The handle owns the device state the reference keeps in CholeskySolver
// (src/cholesky_solver.h:60-98): perm, the row-compressed L and L', the two
// level orders, and the per-solve work buffers. Kernels are linked into this
// library, so there is no module loading (src/cuda_driver.cpp:127-133).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "../../../include/cholespy_amd.h"
#include "common.hpp"
#include "launchers.hpp"

using namespace cholespy_amd;

namespace {

thread_local std::string g_error;
bool g_initialised = false;
int g_device = -1;

int fail(int code, const std::string &msg) {
    g_error = msg;
    return code;
}

int hip_fail(hipError_t e, const char *what) {
    return fail(CHOLESPY_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define HIP_TRY(expr)                                   \
    do {                                                \
        hipError_t e_ = (expr);                         \
        if (e_ != hipSuccess) return hip_fail(e_, #expr); \
    } while (0)

template <typename T>
hipError_t dev_alloc(T **p, size_t count) {
    *p = nullptr;
    if (count == 0) count = 1;
    return hipMalloc(reinterpret_cast<void **>(p), count * sizeof(T));
}

template <typename T>
void dev_free(T *&p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

}  // namespace

struct cholespy_gpu_solver {
    int n = 0;
    int fbytes = 4;
    int64_t nnz = 0;
    int device = 0;
    bool uploaded = false;
    bool analysed[2] = {false, false};   // [0] = upper (L'), [1] = lower (L)
    int n_levels[2] = {0, 0};

    int *perm = nullptr;
    // Row-compressed L (diagonal last) and L' (diagonal first).
    int *lrowptr = nullptr, *lcols = nullptr;
    void *lvals = nullptr;
    int *urowptr = nullptr, *ucols = nullptr;
    void *uvals = nullptr;
    int *order[2] = {nullptr, nullptr};   // ticket -> row per pass
    int *lvl1 = nullptr;                  // analysis scratch: level + 1
    int *status = nullptr;                // [0] error code, [1] abort flag
    void *y = nullptr, *z = nullptr;      // per-solve work buffers (n x k)
    int k_cur = 0;                        // layout the work buffers hold
    size_t work_elems = 0;                // capacity of y and z in elements
    std::vector<int> host_levels[2], host_row_level[2];
};

extern "C" {

const char *cholespy_last_error(void) { return g_error.c_str(); }

int cholespy_device_count(int *count) {
    if (!count) return fail(CHOLESPY_ERR_INVALID_ARGUMENT, "count is NULL");
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
    *count = c;
    return CHOLESPY_OK;
}

int cholespy_init(void) {
    if (g_initialised) return CHOLESPY_OK;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
        return fail(CHOLESPY_ERR_NO_DEVICE, "no HIP device is visible");
    HIP_TRY(hipGetDevice(&g_device));
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, g_device));
    int coop = 0;
    HIP_TRY(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, g_device));
    if (!coop) return fail(CHOLESPY_ERR_NO_DEVICE, "device does not support cooperative launch");
    g_initialised = true;
    return CHOLESPY_OK;
}

int cholespy_shutdown(void) {
    g_initialised = false;
    g_device = -1;
    return CHOLESPY_OK;
}

int cholespy_copy_to_host(void *dst, const void *device_src, size_t bytes) {
    if (bytes == 0) return CHOLESPY_OK;
    if (!dst || !device_src) return fail(CHOLESPY_ERR_INVALID_ARGUMENT, "null pointer");
    HIP_TRY(hipMemcpy(dst, device_src, bytes, hipMemcpyDeviceToHost));
    return CHOLESPY_OK;
}

int cholespy_solver_create(int n, int float_bytes, cholespy_gpu_solver **out) {
    if (!out) return fail(CHOLESPY_ERR_INVALID_ARGUMENT, "out is NULL");
    *out = nullptr;
    if (n < 0) return fail(CHOLESPY_ERR_INVALID_ARGUMENT, "n must be non-negative");
    if (float_bytes != 4 && float_bytes != 8)
        return fail(CHOLESPY_ERR_INVALID_ARGUMENT, "float_bytes must be 4 or 8");
    int rc = cholespy_init();
    if (rc != CHOLESPY_OK) return rc;
    auto *s = new cholespy_gpu_solver();
    s->n = n;
    s->fbytes = float_bytes;
    s->device = g_device;
    hipError_t e = dev_alloc(&s->status, 2);
    if (e == hipSuccess) e = hipMemset(s->status, 0, 2 * sizeof(int));
    if (e == hipSuccess) e = dev_alloc(&s->perm, (size_t)n);
    if (e == hipSuccess) e = dev_alloc(&s->lvl1, (size_t)n);
    if (e != hipSuccess) {
        cholespy_solver_destroy(s);
        return hip_fail(e, "cholespy_solver_create");
    }
    *out = s;
    return CHOLESPY_OK;
}

int cholespy_solver_destroy(cholespy_gpu_solver *s) {
    if (!s) return CHOLESPY_OK;
    (void)hipDeviceSynchronize();
    dev_free(s->status);
    dev_free(s->perm);
    dev_free(s->lvl1);
    dev_free(s->lrowptr);
    dev_free(s->lcols);
    dev_free(s->lvals);
    dev_free(s->urowptr);
    dev_free(s->ucols);
    dev_free(s->uvals);
    dev_free(s->order[0]);
    dev_free(s->order[1]);
    dev_free(s->y);
    dev_free(s->z);
    delete s;
    return CHOLESPY_OK;
}

int cholespy_solver_upload_factor(cholespy_gpu_solver *s, const int *perm, const int *colptr,
                                  const int *rowidx, const double *val, int64_t nnz) {
    if (!s || !perm || !colptr || !rowidx || !val)
        return fail(CHOLESPY_ERR_INVALID_ARGUMENT, "null argument");
    const int n = s->n;
    if (nnz != (int64_t)colptr[n] || colptr[0] != 0 || nnz > 0x7fffffff)
        return fail(CHOLESPY_ERR_INVALID_ARGUMENT, "colptr does not match nnz");
    // Validate everything the kernels index with: perm is a permutation,
    // every column starts with its diagonal and rows ascend inside [0, n).
    {
        std::vector<char> seen((size_t)n, 0);
        for (int i = 0; i < n; ++i) {
            if (perm[i] < 0 || perm[i] >= n || seen[perm[i]])
                return fail(CHOLESPY_ERR_INVALID_ARGUMENT, "perm is not a permutation");
            seen[perm[i]] = 1;
        }
        for (int j = 0; j < n; ++j) {
            if (colptr[j + 1] <= colptr[j] || rowidx[colptr[j]] != j)
                return fail(CHOLESPY_ERR_INVALID_ARGUMENT, "column without leading diagonal");
            for (int p = colptr[j] + 1; p < colptr[j + 1]; ++p)
                if (rowidx[p] <= rowidx[p - 1] || rowidx[p] >= n)
                    return fail(CHOLESPY_ERR_INVALID_ARGUMENT, "row indices not ascending in range");
        }
    }
    // Row-compressed L (transpose of the column form): diagonal ends each row.
    std::vector<int> rowptr((size_t)n + 1, 0), cols((size_t)nnz);
    std::vector<double> rvals((size_t)nnz);
    for (int64_t p = 0; p < nnz; ++p) rowptr[(size_t)rowidx[p] + 1]++;
    for (int i = 0; i < n; ++i) rowptr[i + 1] += rowptr[i];
    {
        std::vector<int> next(rowptr.begin(), rowptr.end() - 1);
        for (int j = 0; j < n; ++j)
            for (int p = colptr[j]; p < colptr[j + 1]; ++p) {
                const int dst = next[rowidx[p]]++;
                cols[dst] = j;
                rvals[dst] = val[p];
            }
    }
    const size_t fb = (size_t)s->fbytes;
    std::vector<char> lv((size_t)nnz * fb), uv((size_t)nnz * fb);
    if (s->fbytes == 4) {
        float *l = reinterpret_cast<float *>(lv.data()), *u = reinterpret_cast<float *>(uv.data());
        for (int64_t p = 0; p < nnz; ++p) {
            l[p] = (float)rvals[p];
            u[p] = (float)val[p];
        }
    } else {
        std::memcpy(lv.data(), rvals.data(), (size_t)nnz * 8);
        std::memcpy(uv.data(), val, (size_t)nnz * 8);
    }
    dev_free(s->lrowptr);
    dev_free(s->lcols);
    dev_free(s->lvals);
    dev_free(s->urowptr);
    dev_free(s->ucols);
    dev_free(s->uvals);
    HIP_TRY(dev_alloc(&s->lrowptr, (size_t)n + 1));
    HIP_TRY(dev_alloc(&s->lcols, (size_t)nnz));
    HIP_TRY(dev_alloc(reinterpret_cast<char **>(&s->lvals), (size_t)nnz * fb));
    HIP_TRY(dev_alloc(&s->urowptr, (size_t)n + 1));
    HIP_TRY(dev_alloc(&s->ucols, (size_t)nnz));
    HIP_TRY(dev_alloc(reinterpret_cast<char **>(&s->uvals), (size_t)nnz * fb));
    HIP_TRY(hipMemcpy(s->perm, perm, (size_t)n * sizeof(int), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(s->lrowptr, rowptr.data(), ((size_t)n + 1) * sizeof(int), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(s->lcols, cols.data(), (size_t)nnz * sizeof(int), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(s->lvals, lv.data(), (size_t)nnz * fb, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(s->urowptr, colptr, ((size_t)n + 1) * sizeof(int), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(s->ucols, rowidx, (size_t)nnz * sizeof(int), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(s->uvals, uv.data(), (size_t)nnz * fb, hipMemcpyHostToDevice));
    s->nnz = nnz;
    s->uploaded = true;
    s->analysed[0] = s->analysed[1] = false;
    return CHOLESPY_OK;
}

int cholespy_solver_analyze(cholespy_gpu_solver *s, int lower, int *n_levels) {
    if (!s) return fail(CHOLESPY_ERR_INVALID_ARGUMENT, "null solver");
    if (!s->uploaded) return fail(CHOLESPY_ERR_STATE, "analyze called before upload_factor");
    const int n = s->n;
    const int li = lower ? 1 : 0;
    if (n == 0) {
        s->analysed[li] = true;
        s->n_levels[li] = 0;
        if (n_levels) *n_levels = 0;
        return CHOLESPY_OK;
    }
    const int *rowptr = lower ? s->lrowptr : s->urowptr;
    const int *cols = lower ? s->lcols : s->ucols;
    HIP_TRY(hipMemset(s->lvl1, 0, (size_t)n * sizeof(int)));
    HIP_TRY(hipMemset(s->status, 0, 2 * sizeof(int)));
    const int cap = max_coresident_blocks(level_analysis_kernel(lower != 0), kRowBlock, 0);
    if (cap <= 0) return fail(CHOLESPY_ERR_HIP, "occupancy query failed for level analysis");
    const int grid = std::max(1, std::min(cap, (n + 3) / 4));
    HIP_TRY(launch_level_analysis(lower != 0, n, rowptr, cols, s->lvl1, s->status, grid, nullptr));
    std::vector<int> lvl1((size_t)n);
    int status[2] = {0, 0};
    HIP_TRY(hipMemcpy(lvl1.data(), s->lvl1, (size_t)n * sizeof(int), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(status, s->status, sizeof(status), hipMemcpyDeviceToHost));
    if (status[0] != 0) {
        HIP_TRY(hipMemset(s->status, 0, 2 * sizeof(int)));
        return fail(CHOLESPY_ERR_TIMEOUT, "level analysis exceeded its bounded wait");
    }
    // Stable counting sort of rows by level (src/cholesky_solver.cpp:348-367).
    int max_lvl = 0;
    std::vector<int> &row_level = s->host_row_level[li];
    row_level.resize((size_t)n);
    for (int i = 0; i < n; ++i) {
        row_level[i] = lvl1[i] - 1;
        max_lvl = std::max(max_lvl, row_level[i]);
    }
    const int nl = max_lvl + 1;
    std::vector<int> start((size_t)nl + 1, 0);
    for (int i = 0; i < n; ++i) start[(size_t)row_level[i] + 1]++;
    for (int l = 0; l < nl; ++l) start[l + 1] += start[l];
    std::vector<int> &levels = s->host_levels[li];
    levels.resize((size_t)n);
    for (int i = 0; i < n; ++i) levels[start[row_level[i]]++] = i;
    dev_free(s->order[li]);
    HIP_TRY(dev_alloc(&s->order[li], (size_t)n));
    HIP_TRY(hipMemcpy(s->order[li], levels.data(), (size_t)n * sizeof(int), hipMemcpyHostToDevice));
    s->n_levels[li] = nl;
    s->analysed[li] = true;
    if (n_levels) *n_levels = nl;
    return CHOLESPY_OK;
}

int cholespy_solver_get_levels(const cholespy_gpu_solver *s, int lower, int *levels, int *row_level) {
    if (!s || !levels) return fail(CHOLESPY_ERR_INVALID_ARGUMENT, "null argument");
    const int li = lower ? 1 : 0;
    if (!s->analysed[li]) return fail(CHOLESPY_ERR_STATE, "pass not analysed");
    std::copy(s->host_levels[li].begin(), s->host_levels[li].end(), levels);
    if (row_level) std::copy(s->host_row_level[li].begin(), s->host_row_level[li].end(), row_level);
    return CHOLESPY_OK;
}

int cholespy_solver_check(cholespy_gpu_solver *s) {
    if (!s) return fail(CHOLESPY_ERR_INVALID_ARGUMENT, "null solver");
    int status[2] = {0, 0};
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemcpy(status, s->status, sizeof(status), hipMemcpyDeviceToHost));
    if (status[0] != 0 || status[1] != 0) {
        HIP_TRY(hipMemset(s->status, 0, 2 * sizeof(int)));
        // Work buffers may hold partial results: restore the sentinel state.
        if (s->work_elems) {
            HIP_TRY(launch_fill_sentinel(s->y, s->work_elems, s->fbytes, nullptr));
            HIP_TRY(launch_fill_sentinel(s->z, s->work_elems, s->fbytes, nullptr));
            HIP_TRY(hipDeviceSynchronize());
        }
        return fail(CHOLESPY_ERR_TIMEOUT, "a solve kernel exceeded its bounded wait");
    }
    return CHOLESPY_OK;
}

}  // extern "C"

namespace {

int solve_impl(cholespy_gpu_solver *s, int fbytes, int k, const void *b, void *x, void *stream) {
    if (!s) return fail(CHOLESPY_ERR_INVALID_ARGUMENT, "null solver");
    if (s->fbytes != fbytes)
        return fail(CHOLESPY_ERR_INVALID_ARGUMENT, "solver precision does not match the call");
    if (k > kMaxRhs) return fail(CHOLESPY_ERR_INVALID_ARGUMENT, "The number of RHS should be less than 128.");
    if (k < 0) return fail(CHOLESPY_ERR_INVALID_ARGUMENT, "negative number of RHS");
    if (!s->analysed[0] || !s->analysed[1])
        return fail(CHOLESPY_ERR_STATE, "solve called before both passes were analysed");
    const int n = s->n;
    if (k == 0 || n == 0) return CHOLESPY_OK;
    if (!b || !x) return fail(CHOLESPY_ERR_INVALID_ARGUMENT, "null b or x");
    hipStream_t st = static_cast<hipStream_t>(stream);
    const size_t elems = (size_t)n * (size_t)k;
    if (elems > s->work_elems) {
        dev_free(reinterpret_cast<char *&>(s->y));
        dev_free(reinterpret_cast<char *&>(s->z));
        HIP_TRY(dev_alloc(reinterpret_cast<char **>(&s->y), elems * (size_t)fbytes));
        HIP_TRY(dev_alloc(reinterpret_cast<char **>(&s->z), elems * (size_t)fbytes));
        s->work_elems = elems;
        s->k_cur = 0;
    }
    if (k != s->k_cur) {
        // The buffers are indexed row * k + col: refill them for the new layout.
        HIP_TRY(launch_fill_sentinel(s->y, elems, fbytes, st));
        HIP_TRY(launch_fill_sentinel(s->z, elems, fbytes, st));
        s->k_cur = k;
    }
    RowPassBuffers bf;
    bf.n = n;
    bf.k = k;
    bf.perm = s->perm;
    bf.b = b;
    bf.x = x;
    bf.y = s->y;
    bf.z = s->z;
    bf.status = s->status;
    for (int pass = 0; pass < 2; ++pass) {
        const bool lower = pass == 0;
        bf.order = s->order[lower ? 1 : 0];
        bf.rowptr = lower ? s->lrowptr : s->urowptr;
        bf.cols = lower ? s->lcols : s->ucols;
        bf.vals = lower ? s->lvals : s->uvals;
        const int cap = max_coresident_blocks(row_pass_kernel(fbytes, lower, k), kRowBlock, 0);
        if (cap <= 0) return fail(CHOLESPY_ERR_HIP, "occupancy query failed for the solve");
        const int grid = std::max(1, std::min(cap, (n + 3) / 4));
        HIP_TRY(launch_row_pass(fbytes, lower, bf, grid, st));
    }
    return CHOLESPY_OK;
}

}  // namespace

extern "C" {

int cholespy_solver_solve_f32(cholespy_gpu_solver *s, int n_rhs, const float *b, float *x, void *stream) {
    return solve_impl(s, 4, n_rhs, b, x, stream);
}

int cholespy_solver_solve_f64(cholespy_gpu_solver *s, int n_rhs, const double *b, double *x,
                              void *stream) {
    return solve_impl(s, 8, n_rhs, b, x, stream);
}

}  // extern "C"
