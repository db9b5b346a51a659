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

// Per-column sums of squares: each thread adds its squares into the
// workgroup's LDS sums (one per column), then one global FP64 atomic per
// column and workgroup.
__device__ __forceinline__ void add_column_squares(double *lsum, int k, double *norm2) {
    __syncthreads();
    for (int t = threadIdx.x; t < k; t += blockDim.x)
        if (lsum[t] != 0.0) atomicAdd(&norm2[t], lsum[t]);
}

// One thread per (row, right-hand side): consecutive threads share a row,
// so cols / vals loads broadcast and x loads are contiguous across columns.
// With norm2 set, also accumulates the residual's squared column norms.
__global__ void residual_f64(int n, int k, const int *__restrict__ rowptr, const int *__restrict__ cols,
                             const double *__restrict__ vals, const double *__restrict__ b,
                             const double *__restrict__ x, float *__restrict__ r32, double *norm2) {
    __shared__ double lsum[kMaxRhs];
    if (norm2) {
        for (int t = threadIdx.x; t < k; t += blockDim.x) lsum[t] = 0.0;
        __syncthreads();
    }
    const size_t total = (size_t)n * k;
    for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < total;
         idx += (size_t)gridDim.x * blockDim.x) {
        const int row = (int)(idx / k), t = (int)(idx % k);
        double s = b[idx];
        for (int p = rowptr[row]; p < rowptr[row + 1]; ++p) s = fma(-vals[p], x[(size_t)cols[p] * k + t], s);
        r32[idx] = (float)s;
        if (norm2) atomicAdd(&lsum[t], s * s);
    }
    if (norm2) add_column_squares(lsum, k, norm2);
}

// norm2[t] += sum_i v[i,t]^2 for a row-major (n x k) array.
__global__ void column_squares(size_t total, int k, const double *__restrict__ v, double *norm2) {
    __shared__ double lsum[kMaxRhs];
    for (int t = threadIdx.x; t < k; t += blockDim.x) lsum[t] = 0.0;
    __syncthreads();
    for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < total;
         idx += (size_t)gridDim.x * blockDim.x) {
        const double e = v[idx];
        atomicAdd(&lsum[idx % k], e * e);
    }
    add_column_squares(lsum, k, norm2);
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

hipError_t launch_zero_f64(double *p, int count, hipStream_t stream) {
    zero_f64<<<1, 128, 0, stream>>>(p, count);
    return hipGetLastError();
}

}  // namespace cholespy_amd
