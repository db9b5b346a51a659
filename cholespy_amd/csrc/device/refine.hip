// FP64 pieces of mixed-precision iterative refinement on gfx950:
//   r32 = float(b)                         (first right-hand side)
//   r32 = float(b - A x)                   (residual, FP64 SpMV on the full CSR of A)
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

// One thread per (row, right-hand side): consecutive threads share a row,
// so cols / vals loads broadcast and x loads are contiguous across columns.
__global__ void residual_f64(int n, int k, const int *__restrict__ rowptr, const int *__restrict__ cols,
                             const double *__restrict__ vals, const double *__restrict__ b,
                             const double *__restrict__ x, float *__restrict__ r32) {
    const size_t total = (size_t)n * k;
    for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < total;
         idx += (size_t)gridDim.x * blockDim.x) {
        const int row = (int)(idx / k), t = (int)(idx % k);
        double s = b[idx];
        for (int p = rowptr[row]; p < rowptr[row + 1]; ++p) s = fma(-vals[p], x[(size_t)cols[p] * k + t], s);
        r32[idx] = (float)s;
    }
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
                               const double *b, const double *x, float *r32, hipStream_t stream) {
    const size_t count = (size_t)n * k;
    if (count == 0) return hipSuccess;
    residual_f64<<<grid_for(count, 256), 256, 0, stream>>>(n, k, rowptr, cols, vals, b, x, r32);
    return hipGetLastError();
}

}  // namespace cholespy_amd
