// FP64 pieces of mixed-precision iterative refinement on gfx950:
//   r32 = float(b)                         (first right-hand side)
//   r32 = float(b - A x)                   (residual, FP64 SpMV on the full CSR of A)
//   norm2[t] += sum_i r[i,t]^2             (optional, the stopping test's residual norms)
// The FP32 block solve of r32 then adds its correction into the FP64 x in
// the backward pass's epilogue (trsv_blocks.hip), so the update costs no
// extra pass. A is symmetric; its full row-compressed form is stored with
// FP64 values, columns ascending in each row.
#include "common.hpp"
#include "launchers.hpp"

namespace cholespy_amd {

namespace {

__global__ void cast_to_f32(const double *__restrict__ in, float *__restrict__ out, size_t count) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < count;
         i += (size_t)gridDim.x * blockDim.x)
        out[i] = (float)in[i];
}

// Per-column sums of squares. The grid-stride loops below step by a
// multiple of k, so every thread stays on one column t and keeps its sum in
// a register; at the end each thread adds it to the workgroup's LDS sum of
// its column, and one global FP64 atomic per column and workgroup follows.
__device__ __forceinline__ size_t column_stride(int k) {
    const size_t threads = (size_t)gridDim.x * blockDim.x;
    return threads >= (size_t)k ? threads - threads % (size_t)k : (size_t)k;
}
__device__ __forceinline__ void add_column_squares(double *lsum, int k, int t, double mine, double *norm2) {
    if (t >= 0 && mine != 0.0) atomicAdd(&lsum[t], mine);
    __syncthreads();
    for (int c = threadIdx.x; c < k; c += blockDim.x)
        if (lsum[c] != 0.0) atomicAdd(&norm2[c], lsum[c]);
}

// One thread per (row, right-hand side): consecutive threads share a row,
// so cols / vals loads broadcast and x loads are contiguous across columns.
// A row's entries go in batches of kResBatch: their column indices and
// values are loaded together, then their x entries together, so a batch
// costs two dependent memory round trips instead of two per entry; the
// products are still added in ascending column order. With norm2 set, also
// accumulates the residual's squared column norms.
constexpr int kResBatch = 8;
__global__ void residual_f64(int n, int k, const int *__restrict__ rowptr, const int *__restrict__ cols,
                             const double *__restrict__ vals, const double *__restrict__ b,
                             const double *__restrict__ x, float *__restrict__ r32, double *norm2) {
    __shared__ double lsum[kMaxRhs];
    if (norm2) {
        for (int t = threadIdx.x; t < k; t += blockDim.x) lsum[t] = 0.0;
        __syncthreads();
    }
    const size_t total = (size_t)n * k, stride = column_stride(k);
    const size_t idx0 = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    double mine = 0.0;
    for (size_t idx = idx0; idx < total && idx0 < stride; idx += stride) {
        const int row = (int)(idx / k), t = (int)(idx % k);
        double s = b[idx];
        const int p1 = rowptr[row + 1];
        for (int p0 = rowptr[row]; p0 < p1; p0 += kResBatch) {
            int c[kResBatch];
            double v[kResBatch], xv[kResBatch];
#pragma unroll
            for (int j = 0; j < kResBatch; ++j) {
                c[j] = p0 + j < p1 ? cols[p0 + j] : 0;
                v[j] = p0 + j < p1 ? vals[p0 + j] : 0.0;
            }
#pragma unroll
            for (int j = 0; j < kResBatch; ++j) xv[j] = p0 + j < p1 ? x[(size_t)c[j] * k + t] : 0.0;
#pragma unroll
            for (int j = 0; j < kResBatch; ++j)
                if (p0 + j < p1) s = fma(-v[j], xv[j], s);
        }
        r32[idx] = (float)s;
        mine = fma(s, s, mine);
    }
    if (norm2) add_column_squares(lsum, k, idx0 < stride ? (int)(idx0 % k) : -1, mine, norm2);
}

// norm2[t] += sum_i v[i,t]^2 for a row-major (n x k) array.
__global__ void column_squares(size_t total, int k, const double *__restrict__ v, double *norm2) {
    __shared__ double lsum[kMaxRhs];
    for (int t = threadIdx.x; t < k; t += blockDim.x) lsum[t] = 0.0;
    __syncthreads();
    const size_t stride = column_stride(k), idx0 = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    double mine = 0.0;
    for (size_t idx = idx0; idx < total && idx0 < stride; idx += stride) mine = fma(v[idx], v[idx], mine);
    add_column_squares(lsum, k, idx0 < stride ? (int)(idx0 % k) : -1, mine, norm2);
}

// dst[i] = src[i]: device values into host memory mapped into the device.
__global__ void copy_out_f64(double *dst, const double *src, int count) {
    for (int i = threadIdx.x; i < count; i += blockDim.x)
        __hip_atomic_store(dst + i, src[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

__global__ void zero_f64(double *p, int count) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) p[i] = 0.0;
}

int grid_for(size_t count, int block) {
    const size_t want = (count + block - 1) / block;
    return (int)(want < 8192 ? (want > 0 ? want : 1) : 8192);
}

}  // namespace

hipError_t launch_cast_to_f32(const double *in, float *out, size_t count, hipStream_t stream) {
    if (count == 0) return hipSuccess;
    cast_to_f32<<<grid_for(count, 256), 256, 0, stream>>>(in, out, count);
    return hipGetLastError();
}

hipError_t launch_residual_f64(int n, int k, const int *rowptr, const int *cols, const double *vals,
                               const double *b, const double *x, float *r32, double *norm2, hipStream_t stream) {
    const size_t count = (size_t)n * k;
    if (count == 0) return hipSuccess;
    residual_f64<<<grid_for(count, 256), 256, 0, stream>>>(n, k, rowptr, cols, vals, b, x, r32, norm2);
    return hipGetLastError();
}

hipError_t launch_column_norms2(size_t rows, int k, const double *v, double *norm2, hipStream_t stream) {
    zero_f64<<<1, 128, 0, stream>>>(norm2, k);
    const size_t count = rows * (size_t)k;
    if (count == 0) return hipGetLastError();
    column_squares<<<grid_for(count, 256), 256, 0, stream>>>(count, k, v, norm2);
    return hipGetLastError();
}

hipError_t launch_copy_out_f64(double *host_mapped, const double *src, int count, hipStream_t stream) {
    copy_out_f64<<<1, 256, 0, stream>>>(host_mapped, src, count);
    return hipGetLastError();
}

hipError_t launch_zero_f64(double *p, int count, hipStream_t stream) {
    zero_f64<<<1, 128, 0, stream>>>(p, count);
    return hipGetLastError();
}

}  // namespace cholespy_amd
