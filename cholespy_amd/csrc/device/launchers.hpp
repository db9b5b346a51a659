// Host-side launchers for the gfx950 kernels (one definition per .hip file).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace cholespy_amd {

// Threads per workgroup of the persistent row kernels (4 wavefronts).
constexpr int kRowBlock = 256;

// Largest co-resident grid (workgroups) for a persistent kernel.
int max_coresident_blocks(const void *kernel, int block, size_t lds_bytes);

// Level analysis (kernels/kernels.cu:8-60). lvl1[r] must be zero on entry;
// on exit lvl1[r] = level(r) + 1.
const void *level_analysis_kernel(bool lower);
hipError_t launch_level_analysis(bool lower, int n, const int *rowptr, const int *cols,
                                 int *lvl1, int *status, int grid, hipStream_t stream);

// Row-per-wavefront level-scheduled triangular pass (kernels/kernels.cu:64-202).
struct RowPassBuffers {
    int n = 0, k = 0;
    const int *order = nullptr;    // ticket -> row (level order)
    const int *rowptr = nullptr;
    const int *cols = nullptr;
    const void *vals = nullptr;
    const int *perm = nullptr;
    const void *b = nullptr;       // user right-hand side (forward pass)
    void *y = nullptr;             // forward result, waited on in the forward pass
    void *z = nullptr;             // backward result, waited on in the backward pass
    void *x = nullptr;             // user solution (backward pass)
    int *status = nullptr;
};
const void *row_pass_kernel(int float_bytes, bool lower, int k);
hipError_t launch_row_pass(int float_bytes, bool lower, const RowPassBuffers &buf, int grid,
                           hipStream_t stream);

// Fills n words of `bytes` (4 or 8) with the sentinel of that width.
hipError_t launch_fill_sentinel(void *p, size_t n, int bytes, hipStream_t stream);

}  // namespace cholespy_amd
