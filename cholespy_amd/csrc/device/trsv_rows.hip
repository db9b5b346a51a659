// Row-per-wavefront, level-scheduled sparse triangular passes on gfx950:
// the forward pass y = L^{-1} P b and the backward pass z = L'^{-1} y,
// x[perm[r]] = z[r] (kernels/kernels.cu:64-131 and 133-202).
//
// Work assignment. Rows are visited in level order (`order`, the sorted
// `levels` array of the analysis). Wavefronts of a persistent grid take
// chunks of kRowChunk consecutive rows from the ticket queues of common.hpp
// (the reference takes one atomicAdd per row, kernels/kernels.cu:73-75); a
// row's dependencies are in strictly lower levels, i.e. earlier tickets held
// by running wavefronts or earlier rows of the same chunk, so the schedule
// cannot deadlock.
//
// Lane layout for k right-hand sides. kp = min(64, next_pow2(k)) lanes form
// a group that covers the k columns of one nonzero (a lane covers columns t
// and t + 64 when k > 64); the wavefront's 64 / kp groups stream consecutive
// nonzeros of the row, so the reads of cols/vals are coalesced and no lane
// idles when k is small (the reference leaves 125 of 128 threads idle at
// k = 3). Group partial sums are combined with a butterfly of lane shuffles.
//
// Hand-off. y (forward) and z (backward) are pre-filled with a signalling-NaN
// sentinel; a value is published with one agent-scope atomic store and
// consumers wait on the value itself (common.hpp). Each pass also restores
// the sentinel in the buffer the next pass will wait on: the forward pass
// resets z[row], the backward pass resets y[row] after consuming it, so
// repeated solves need no memset (the reference issues two per pass,
// src/cholesky_solver.cpp:381-382).
//
// Accumulation order differs from the reference's sequential one (partial
// sums per group, then a tree), so results agree to rounding, not bitwise.
#include "common.hpp"
#include "launchers.hpp"

namespace cholespy_amd {

namespace {

template <typename F>
struct RowPassArgs {
    int n, k, kp_log2;
    const int *order;
    const int *rowptr;
    const int *cols;
    const F *vals;
    const int *perm;
    const F *b;
    F *y;
    F *z;
    F *x;
    int *status;
    QueueArgs qa;
};

constexpr int kUnroll = 4;   // nonzeros per group with loads in flight together

template <typename F, bool LOWER, int CPL>
__global__ __launch_bounds__(kRowBlock) void row_pass(RowPassArgs<F> a) {
    using U = typename Bits<F>::U;
    const int lane = threadIdx.x & (kWave - 1);
    const int w = blockIdx.x * (kRowBlock / kWave) + (threadIdx.x / kWave);
    const int qid = w % a.qa.nq;
    const int nchunks = (a.n + kRowChunk - 1) / kRowChunk;
    const int kp = 1 << a.kp_log2;
    const int groups = kWave >> a.kp_log2;
    const int g = lane >> a.kp_log2;
    const int t = lane & (kp - 1);
    const int k = a.k;
    F *dep = LOWER ? a.y : a.z;   // buffer this pass publishes to and waits on

    int next = lane == 0 ? take_ticket(a.qa, qid, nchunks) : 0;
    next = __shfl(next, 0, kWave);
    while (next >= 0) {
      const int chunk = next;
      int pre = lane == 0 ? take_ticket(a.qa, qid, nchunks) : 0;   // in flight during the chunk
      const int tk_end = min(a.n, (chunk + 1) * kRowChunk);
      for (int tk = chunk * kRowChunk; tk < tk_end; ++tk) {
        const int row = a.order[tk];
        const int rs = a.rowptr[row], re = a.rowptr[row + 1];
        const int lo = LOWER ? rs : rs + 1;
        const int hi = LOWER ? re - 1 : re;
        const int dpos = LOWER ? re - 1 : rs;

        F acc[CPL];
#pragma unroll
        for (int q = 0; q < CPL; ++q) acc[q] = F(0);

        for (int base = lo + g; base < hi; base += groups * kUnroll) {
            int c[kUnroll];
            F v[kUnroll];
            U xb[kUnroll][CPL];
#pragma unroll
            for (int j = 0; j < kUnroll; ++j) {
                const int i = base + j * groups;
                const bool in = i < hi;
                c[j] = in ? a.cols[i] : 0;
                v[j] = in ? a.vals[i] : F(0);
            }
#pragma unroll
            for (int j = 0; j < kUnroll; ++j) {
                const bool in = base + j * groups < hi;
#pragma unroll
                for (int q = 0; q < CPL; ++q) {
                    const int col = t + q * kWave;
                    xb[j][q] = (in && col < k) ? load_bits(&dep[(size_t)c[j] * k + col]) : U(0);
                }
            }
#pragma unroll
            for (int j = 0; j < kUnroll; ++j) {
#pragma unroll
                for (int q = 0; q < CPL; ++q) {
                    F xv = from_bits<F>(xb[j][q]);
                    if (xb[j][q] == Bits<F>::sentinel)
                        xv = wait_value(&dep[(size_t)c[j] * k + t + q * kWave], a.status, a.qa);
                    acc[q] = fma(v[j], xv, acc[q]);
                }
            }
        }
        for (int off = kp; off < kWave; off <<= 1) {
#pragma unroll
            for (int q = 0; q < CPL; ++q) acc[q] += __shfl_xor(acc[q], off, kWave);
        }
        if (g == 0) {
            const F d = a.vals[dpos];
#pragma unroll
            for (int q = 0; q < CPL; ++q) {
                const int col = t + q * kWave;
                if (col >= k) continue;
                const size_t own = (size_t)row * k + col;
                if (LOWER) {
                    const F r = a.b[(size_t)a.perm[row] * k + col];
                    publish(&a.y[own], (r - acc[q]) / d);
                    a.z[own] = from_bits<F>(Bits<F>::sentinel);
                } else {
                    const F r = a.y[own];
                    const F res = (r - acc[q]) / d;
                    publish(&a.z[own], res);
                    a.x[(size_t)a.perm[row] * k + col] = res;
                    a.y[own] = from_bits<F>(Bits<F>::sentinel);
                }
            }
        }
      }
      next = __shfl(pre, 0, kWave);
    }
    __syncthreads();
    if (threadIdx.x == 0) finish_queues(a.qa);
}

template <typename F>
const void *pick(bool lower, int k) {
    if (lower)
        return k > kWave ? reinterpret_cast<const void *>(&row_pass<F, true, 2>)
                         : reinterpret_cast<const void *>(&row_pass<F, true, 1>);
    return k > kWave ? reinterpret_cast<const void *>(&row_pass<F, false, 2>)
                     : reinterpret_cast<const void *>(&row_pass<F, false, 1>);
}

template <typename F>
hipError_t launch(bool lower, const RowPassBuffers &bf, int grid, QueueState &queues, hipStream_t stream) {
    RowPassArgs<F> a;
    a.n = bf.n;
    a.k = bf.k;
    int kp_log2 = 0;
    while ((1 << kp_log2) < bf.k && kp_log2 < 6) ++kp_log2;
    a.kp_log2 = kp_log2;
    a.order = bf.order;
    a.rowptr = bf.rowptr;
    a.cols = bf.cols;
    a.vals = static_cast<const F *>(bf.vals);
    a.perm = bf.perm;
    a.b = static_cast<const F *>(bf.b);
    a.y = static_cast<F *>(bf.y);
    a.z = static_cast<F *>(bf.z);
    a.x = static_cast<F *>(bf.x);
    a.status = bf.status;
    queues.fill(a.qa);
    void *args[] = {&a};
    return hipLaunchKernel(pick<F>(lower, bf.k), dim3(grid), dim3(kRowBlock), args, 0, stream);
}

}  // namespace

const void *row_pass_kernel(int float_bytes, bool lower, int k) {
    return float_bytes == 4 ? pick<float>(lower, k) : pick<double>(lower, k);
}

hipError_t launch_row_pass(int float_bytes, bool lower, const RowPassBuffers &buf, int grid,
                           QueueState &queues, hipStream_t stream) {
    return float_bytes == 4 ? launch<float>(lower, buf, grid, queues, stream)
                            : launch<double>(lower, buf, grid, queues, stream);
}

}  // namespace cholespy_amd
