// Device-side input normalisation: the caller's COO / CSR / CSC arrays on
// the GPU -> the lower-triangular CSC that the host factorisation reads,
// without copying the raw input to the host first. The reference copies
// the three arrays to the host and normalises there (src/main.cpp:55-67,
// src/cholesky_solver.cpp:8-149, CHOLMOD stype = -1 at 211-220).
//
// Steps (one stream, synchronous at the end):
//   1. one thread per entry: its (row, col) (a binary search in the pointer
//      array for CSR / CSC), the index check, and a 64-bit key
//      col * n + row for entries on or below the diagonal (n * n for the
//      others, which sort last and are dropped);
//   2. a stable radix sort of (key, entry) pairs (hipcub), so equal keys
//      keep their input order;
//   3. one thread per distinct key adds its entries' values in input order,
//      starting from the first value: the same sums, bit for bit, as the
//      host normalisation (sparse_input.cpp), and records for each input
//      entry the output position it went to;
//   4. column pointers from the sorted columns.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <cstdint>
#include <string>
#include <vector>

#include "launchers.hpp"

namespace cholespy_amd {

namespace {

constexpr int kThreads = 256;

inline unsigned int grid_for(int64_t count) { return (unsigned int)((count + kThreads - 1) / kThreads); }

// bad[0]: bit 0 = pointer array does not start at 0 / end at nnz, bit 1 =
// pointer array decreases somewhere.
template <typename IT>
__global__ void check_pointer(int n, int64_t nnz, const IT *ptr, int *bad) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i == 0 && ((long long)ptr[0] != 0 || (long long)ptr[n] != nnz)) atomicOr(bad, 1);
    if (i < n && ptr[i + 1] < ptr[i]) atomicOr(bad, 2);
}

// Entry p -> key and value. first_bad = min over bad entries of
// 2 p + (0: row index, 1: column index), the first failure in input order
// with the host's check order (row before column).
template <typename IT, typename VT>
__global__ void expand_entries(int n, int64_t nnz, const IT *ii, const IT *jj, const VT *x, int type,
                               unsigned long long drop, unsigned long long *keys, long long *idx, double *vals,
                               unsigned long long *first_bad) {
    const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (p >= nnz) return;
    long long row, col;
    if (type == 2) {   // COO
        row = (long long)ii[p];
        col = (long long)jj[p];
    } else {           // CSR (ii = row pointer) or CSC (ii = column pointer), validated
        int lo = 0, hi = n;   // largest s with ii[s] <= p
        while (hi - lo > 1) {
            const int mid = lo + (hi - lo) / 2;
            if ((long long)ii[mid] <= p) lo = mid;
            else hi = mid;
        }
        if (type == 1) {
            row = lo;
            col = (long long)jj[p];
        } else {
            col = lo;
            row = (long long)jj[p];
        }
    }
    const bool rbad = row < 0 || row >= n, cbad = col < 0 || col >= n;
    idx[p] = p;
    if (rbad || cbad) {
        atomicMin(first_bad, 2ull * (unsigned long long)p + (rbad ? 0ull : 1ull));
        keys[p] = drop;
        vals[p] = 0.0;
        return;
    }
    keys[p] = row >= col ? (unsigned long long)col * (unsigned long long)n + (unsigned long long)row : drop;
    vals[p] = (double)x[p];
}

__global__ void mark_heads(int64_t nnz, const unsigned long long *keys, unsigned long long drop, long long *head) {
    const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (p >= nnz) return;
    head[p] = keys[p] != drop && (p == 0 || keys[p] != keys[p - 1]) ? 1 : 0;
}

__global__ void sum_segments(int64_t nnz, int n, const unsigned long long *keys, const long long *idx,
                             const double *vals, const long long *head, const long long *seg, int *rowidx,
                             int *colv, double *val, long long *entry_dst) {
    const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (p >= nnz || !head[p]) return;
    const long long s = seg[p];
    const unsigned long long key = keys[p];
    double acc = vals[idx[p]];
    entry_dst[idx[p]] = s;
    for (int64_t q = p + 1; q < nnz && keys[q] == key; ++q) {
        acc += vals[idx[q]];
        entry_dst[idx[q]] = s;
    }
    rowidx[s] = (int)(key % (unsigned long long)n);
    colv[s] = (int)(key / (unsigned long long)n);
    val[s] = acc;
}

// colptr[c] = first output position with column >= c (colv ascending).
__global__ void column_pointers(int64_t m, int n, const int *colv, int *colptr) {
    const int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (s > m) return;
    if (s == m) {
        for (int c = m > 0 ? colv[m - 1] + 1 : 0; c <= n; ++c) colptr[c] = (int)m;
        return;
    }
    for (int c = s > 0 ? colv[s - 1] + 1 : 0; c <= colv[s]; ++c) colptr[c] = (int)s;
}

__global__ void fill_i64(int64_t count, long long v, long long *p) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < count) p[i] = v;
}

// Device buffers freed on every return path.
struct Buffers {
    std::vector<void *> held;
    template <typename T>
    hipError_t alloc(T **p, size_t count) {
        *p = nullptr;
        const hipError_t e = hipMalloc(reinterpret_cast<void **>(p), (count ? count : 1) * sizeof(T));
        if (e == hipSuccess) held.push_back(*p);
        return e;
    }
    ~Buffers() {
        for (void *p : held) (void)hipFree(p);
    }
};

#define NORM_TRY(expr)                  \
    do {                                \
        const hipError_t e_ = (expr);   \
        if (e_ != hipSuccess) return e_; \
    } while (0)

template <typename IT, typename VT>
hipError_t normalize_impl(int n, int64_t nnz, const IT *ii, const IT *jj, const VT *x, int type, NormalizedInput &out) {
    static const char *kWhat[3][2] = {{"row", "row"}, {"column", "column"}, {"row", "column"}};
    out.invalid.clear();
    out.colptr.assign((size_t)n + 1, 0);
    out.rowidx.clear();
    out.val.clear();
    out.entry_dst.assign((size_t)nnz, -1);
    Buffers buf;
    int *bad = nullptr;
    unsigned long long *first_bad = nullptr;
    NORM_TRY(buf.alloc(&bad, 1));
    NORM_TRY(buf.alloc(&first_bad, 1));
    NORM_TRY(hipMemset(bad, 0, sizeof(int)));
    NORM_TRY(hipMemset(first_bad, 0xff, sizeof(unsigned long long)));
    if (type != 2) {
        check_pointer<IT><<<grid_for((int64_t)n + 1), kThreads>>>(n, nnz, ii, bad);
        NORM_TRY(hipGetLastError());
        int h = 0;
        NORM_TRY(hipMemcpy(&h, bad, sizeof(int), hipMemcpyDeviceToHost));
        const std::string which = type == 1 ? "row" : "column";
        if (h & 1) {
            out.invalid = "Sparse matrix: " + which + " pointer array must start at 0 and end at nnz.";
            return hipSuccess;
        }
        if (h & 2) {
            out.invalid = "Sparse matrix: " + which + " pointer array must be non-decreasing.";
            return hipSuccess;
        }
    }
    if (nnz == 0) return hipSuccess;
    const unsigned long long drop = (unsigned long long)n * (unsigned long long)n;
    int end_bit = 1;
    while (end_bit < 64 && (drop >> end_bit) != 0) ++end_bit;
    unsigned long long *keys = nullptr, *keys_s = nullptr;
    long long *idx = nullptr, *idx_s = nullptr, *head = nullptr, *seg = nullptr, *edst = nullptr;
    double *vals = nullptr;
    NORM_TRY(buf.alloc(&keys, (size_t)nnz));
    NORM_TRY(buf.alloc(&keys_s, (size_t)nnz));
    NORM_TRY(buf.alloc(&idx, (size_t)nnz));
    NORM_TRY(buf.alloc(&idx_s, (size_t)nnz));
    NORM_TRY(buf.alloc(&vals, (size_t)nnz));
    expand_entries<IT, VT><<<grid_for(nnz), kThreads>>>(n, nnz, ii, jj, x, type, drop, keys, idx, vals, first_bad);
    NORM_TRY(hipGetLastError());
    unsigned long long fb = 0;
    NORM_TRY(hipMemcpy(&fb, first_bad, sizeof(fb), hipMemcpyDeviceToHost));
    if (fb != ~0ull) {
        out.invalid = std::string("Sparse matrix: ") + kWhat[type][fb & 1ull] + " index out of range [0, n_rows).";
        return hipSuccess;
    }
    size_t temp_bytes = 0;
    NORM_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, temp_bytes, keys, keys_s, idx, idx_s, (int)nnz, 0, end_bit));
    unsigned char *temp = nullptr;
    NORM_TRY(buf.alloc(&temp, temp_bytes));
    NORM_TRY(hipcub::DeviceRadixSort::SortPairs(temp, temp_bytes, keys, keys_s, idx, idx_s, (int)nnz, 0, end_bit));
    NORM_TRY(buf.alloc(&head, (size_t)nnz));
    NORM_TRY(buf.alloc(&seg, (size_t)nnz));
    mark_heads<<<grid_for(nnz), kThreads>>>(nnz, keys_s, drop, head);
    NORM_TRY(hipGetLastError());
    size_t scan_bytes = 0;
    NORM_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, head, seg, (int)nnz));
    unsigned char *scan_temp = nullptr;
    NORM_TRY(buf.alloc(&scan_temp, scan_bytes));
    NORM_TRY(hipcub::DeviceScan::ExclusiveSum(scan_temp, scan_bytes, head, seg, (int)nnz));
    long long last_seg = 0, last_head = 0;
    NORM_TRY(hipMemcpy(&last_seg, seg + nnz - 1, sizeof(long long), hipMemcpyDeviceToHost));
    NORM_TRY(hipMemcpy(&last_head, head + nnz - 1, sizeof(long long), hipMemcpyDeviceToHost));
    const int64_t m = last_seg + last_head;
    int *rowidx = nullptr, *colv = nullptr, *colptr = nullptr;
    double *val = nullptr;
    NORM_TRY(buf.alloc(&rowidx, (size_t)m));
    NORM_TRY(buf.alloc(&colv, (size_t)m));
    NORM_TRY(buf.alloc(&val, (size_t)m));
    NORM_TRY(buf.alloc(&colptr, (size_t)n + 1));
    NORM_TRY(buf.alloc(&edst, (size_t)nnz));
    fill_i64<<<grid_for(nnz), kThreads>>>(nnz, -1, edst);
    NORM_TRY(hipGetLastError());
    sum_segments<<<grid_for(nnz), kThreads>>>(nnz, n, keys_s, idx_s, vals, head, seg, rowidx, colv, val, edst);
    NORM_TRY(hipGetLastError());
    column_pointers<<<grid_for(m + 1), kThreads>>>(m, n, colv, colptr);
    NORM_TRY(hipGetLastError());
    out.rowidx.resize((size_t)m);
    out.val.resize((size_t)m);
    NORM_TRY(hipMemcpy(out.colptr.data(), colptr, ((size_t)n + 1) * sizeof(int), hipMemcpyDeviceToHost));
    if (m > 0) {
        NORM_TRY(hipMemcpy(out.rowidx.data(), rowidx, (size_t)m * sizeof(int), hipMemcpyDeviceToHost));
        NORM_TRY(hipMemcpy(out.val.data(), val, (size_t)m * sizeof(double), hipMemcpyDeviceToHost));
    }
    static_assert(sizeof(long long) == sizeof(int64_t), "entry map element size");
    NORM_TRY(hipMemcpy(out.entry_dst.data(), edst, (size_t)nnz * sizeof(long long), hipMemcpyDeviceToHost));
    return hipSuccess;
}

}  // namespace

hipError_t normalize_on_device(int n, int64_t nnz, const void *ii, const void *jj, int index_bytes, const void *x,
                               int value_bytes, int type, NormalizedInput &out) {
    if (index_bytes == 4 && value_bytes == 8)
        return normalize_impl<int, double>(n, nnz, static_cast<const int *>(ii), static_cast<const int *>(jj),
                                           static_cast<const double *>(x), type, out);
    if (index_bytes == 8 && value_bytes == 8)
        return normalize_impl<long long, double>(n, nnz, static_cast<const long long *>(ii),
                                                 static_cast<const long long *>(jj), static_cast<const double *>(x),
                                                 type, out);
    if (index_bytes == 4 && value_bytes == 4)
        return normalize_impl<int, float>(n, nnz, static_cast<const int *>(ii), static_cast<const int *>(jj),
                                          static_cast<const float *>(x), type, out);
    return normalize_impl<long long, float>(n, nnz, static_cast<const long long *>(ii),
                                            static_cast<const long long *>(jj), static_cast<const float *>(x), type,
                                            out);
}

}  // namespace cholespy_amd
