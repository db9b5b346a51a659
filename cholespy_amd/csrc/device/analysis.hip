// Level-set analysis of L and L' on gfx950.
//
// Computes, for every row r of a triangular factor stored row-compressed,
//     level(r) = 0                        if row r has no off-diagonal entry,
//     level(r) = 1 + max_c level(c)       over its off-diagonal columns c,
// the function analysis_lower / analysis_upper compute
// (kernels/kernels.cu:8-33 and 35-60): lower = rows of L, diagonal last,
// dependencies c < r, rows taken in increasing order; upper = rows of L'
// (columns of L), diagonal first, dependencies c > r, rows taken from n-1
// down (kernels/kernels.cu:36).
//
// MI355X design: one wavefront per row instead of one thread, so a row's
// dependencies are read 64 at a time (coalesced) and reduced with a
// wave-wide max; wavefronts of a persistent grid take chunks of rows in
// dependency order from the ticket queues of common.hpp (the reference takes
// one atomicAdd ticket per row, kernels/kernels.cu:9). The per-row "analyzed" flag and
// level of the reference (two arrays, a fence in between) collapse into one
// word lvl1[r] = level(r) + 1 (0 = not yet known), published with one
// agent-scope atomic store.
#include "common.hpp"
#include "launchers.hpp"

namespace cholespy_amd {

template <bool LOWER>
__global__ __launch_bounds__(kRowBlock) void level_analysis(int n, const int *__restrict__ rowptr,
                                                            const int *__restrict__ cols,
                                                            int *lvl1, int *status, QueueArgs qa) {
    const int lane = threadIdx.x & (kWave - 1);
    const int w = blockIdx.x * (kRowBlock / kWave) + (threadIdx.x / kWave);
    const int q = w % qa.nq;
    const int nchunks = (n + kRowChunk - 1) / kRowChunk;
    int next = lane == 0 ? take_ticket(qa, q, nchunks) : 0;
    next = __shfl(next, 0, kWave);
    while (next >= 0) {
        const int chunk = next;
        int pre = lane == 0 ? take_ticket(qa, q, nchunks) : 0;   // in flight during the chunk
        const int t_end = min(n, (chunk + 1) * kRowChunk);
        for (int t = chunk * kRowChunk; t < t_end; ++t) {
            const int row = LOWER ? t : n - 1 - t;
            const int rs = rowptr[row], re = rowptr[row + 1];
            const int lo = LOWER ? rs : rs + 1;
            const int hi = LOWER ? re - 1 : re;
            int lv = 0;
            for (int i = lo + lane; i < hi; i += kWave) lv = max(lv, wait_nonzero(&lvl1[cols[i]], status, qa));
            lv = wave_max(lv);
            if (lane == 0) store_int(&lvl1[row], lv + 1);
        }
        next = __shfl(pre, 0, kWave);
    }
    // One finish per workgroup, after all its wavefronts are done.
    __syncthreads();
    if (threadIdx.x == 0) finish_queues(qa);
}

const void *level_analysis_kernel(bool lower) {
    return lower ? reinterpret_cast<const void *>(&level_analysis<true>)
                 : reinterpret_cast<const void *>(&level_analysis<false>);
}

hipError_t launch_level_analysis(bool lower, int n, const int *rowptr, const int *cols, int *lvl1,
                                 int *status, int grid, QueueState &queues, hipStream_t stream) {
    QueueArgs qa;
    queues.fill(qa);
    if (lower)
        level_analysis<true><<<grid, kRowBlock, 0, stream>>>(n, rowptr, cols, lvl1, status, qa);
    else
        level_analysis<false><<<grid, kRowBlock, 0, stream>>>(n, rowptr, cols, lvl1, status, qa);
    return hipGetLastError();
}

void QueueState::fill(QueueArgs &qa) const {
    qa.ctr = ctr;
    qa.nq = nq;
    qa.spin_limit = spin_limit;
    qa.abort = abort;
}

int max_coresident_blocks(const void *kernel, int block, size_t lds_bytes) {
    int dev = 0, per_cu = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, lds_bytes) != hipSuccess)
        return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return 0;
    return per_cu * cus;
}

template <typename U>
__global__ void fill_words(U *p, size_t n, U v) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}

hipError_t launch_fill_sentinel(void *p, size_t n, int bytes, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    const int block = 256;
    const size_t want = (n + block - 1) / block;
    const int grid = (int)(want < 4096 ? want : 4096);
    if (bytes == 4)
        fill_words<uint32_t><<<grid, block, 0, stream>>>(static_cast<uint32_t *>(p), n,
                                                         Bits<float>::sentinel);
    else
        fill_words<uint64_t><<<grid, block, 0, stream>>>(static_cast<uint64_t *>(p), n,
                                                         Bits<double>::sentinel);
    return hipGetLastError();
}

}  // namespace cholespy_amd
