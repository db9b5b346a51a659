// Partitioned-inverse block passes on gfx950 (the plan is in block_plan.hpp).
//
// Forward task (block B, rows [first, last) of M_B):
//   r_B[i] = b[perm[a+i]] + sum of the published partials for row a+i
//   out    = M_B[first:last, :] r_B
//   rows < m  -> y[a+rho]                   (read by the backward pass only)
//   rows >= m -> partial slot for row R_B[rho-m]   (published, waited on)
// Backward task (block B, columns [first, last) of M_B):
//   z_B[c] = sum_rho M_B[rho, c] * v[rho],  v = [y_B ; z_{R_B}]
//   z published (waited on by descendants), x[perm[a+c]] = z_B[c].
// No task waits on a task of its own block, and every wait targets a block of
// a strictly lower block level, i.e. an earlier ticket; workgroups take
// tickets in order from the queues of common.hpp (the next one prefetched
// while the current task runs), so every awaited task is running or done and
// the schedule cannot deadlock.
//
// Each task is one workgroup of 256 threads: 64 rows x 4 column segments
// (forward; each M load is a 256-byte wavefront access down one column) or
// 32 columns x 8 row segments (backward; 128-byte pieces of M_B's rows), so
// loads are coalesced; the segment sums are combined in LDS. Every thread
// keeps 16 M loads in flight. Right-hand sides are processed in tiles of KT
// columns held in LDS.
//
// Sentinel hand-off as in common.hpp. The forward pass resets z for its rows,
// the backward pass resets the partial slots of its block, so the next solve
// starts from the all-sentinel state without a memset.
#include "common.hpp"
#include "launchers.hpp"

namespace cholespy_amd {

namespace {

constexpr int kTaskThreads = 256;
constexpr int kLanes = 64;      // forward: rows of M_B per task
constexpr int kSegs = 4;        // forward: segments of the column range
constexpr int kBLanes = 32;     // backward: columns of M_B per task (128-byte row pieces)
constexpr int kBSegs = 8;       // backward: segments of the row range
constexpr int kGatherUnroll = 8;
constexpr int kMU = 16;         // M loads issued together per thread
constexpr int kPairs = 4;       // forward gather: (row, column) pairs walked together

template <typename F>
struct BlockPassArgs {
    int ntasks, k, max_rows;
    const int *task;        // (block, first, last) per task
    const int *start;       // block row ranges
    const int *rptr;        // R list / partial slot ranges
    const int *rlist;
    const int64_t *moff;
    const F *M;             // forward: column-major packed; backward: row-major packed
    const int *cptr;        // contribution lists (forward)
    const int *cslot;
    const int *perm;
    const F *b;
    F *y, *z, *part, *rbuf;
    void *x;
    int *status;
    int xmode;
    // Several independent systems in one launch: block -> system, and per
    // system its right-hand-side count and the (possibly negative) bases
    // that place row r / slot s at base + r * k + t in the packed vectors.
    // nullptr = one system with a.k columns and zero bases.
    const int *blk_sys;
    const int *sys_k;
    const long long *sys_vbase;
    const long long *sys_pbase;
    QueueArgs qa;
    // Optional task timeline (nullptr = off): per task, 4 words =
    // {start, after the gather phase, end, workgroup id} in wall-clock ticks (100 MHz).
    unsigned long long *trace;
};

// Right-hand-side layout of one task's system.
struct VecMap {
    int k;
    long long vbase, pbase;
};

template <typename F>
__device__ __forceinline__ VecMap vec_map(const BlockPassArgs<F> &a, int blk) {
    if (!a.blk_sys) return VecMap{a.k, 0, 0};
    const int sys = a.blk_sys[blk];
    return VecMap{a.sys_k[sys], a.sys_vbase[sys], a.sys_pbase[sys]};
}

// Ticket handling per workgroup: thread 0 takes tickets, the LDS word s_next
// broadcasts them at task boundaries (after a barrier).
__device__ __forceinline__ int wg_first_ticket(const QueueArgs &qa, int q, int n, int *s_next) {
    if (threadIdx.x == 0) *s_next = take_ticket(qa, q, n);
    __syncthreads();
    return *s_next;
}

__device__ __forceinline__ int64_t fwd_off(int64_t rows, int64_t c, int64_t rho) {
    return c * rows - c * (c - 1) / 2 + (rho - c);
}
__device__ __forceinline__ int64_t bwd_off(int64_t m, int64_t rho, int64_t c) {
    return rho < m ? rho * (rho + 1) / 2 + c : m * (m + 1) / 2 + (rho - m) * m + c;
}

template <typename F, int KT>
__global__ __launch_bounds__(kTaskThreads) void block_forward(BlockPassArgs<F> a) {
    using U = typename Bits<F>::U;
    extern __shared__ __align__(16) unsigned char smem[];
    F *rv = reinterpret_cast<F *>(smem);                   // max_rows x KT (only m used)
    F *red = rv + (size_t)a.max_rows * KT;                 // 256 x KT
    __shared__ int s_next;
    const int tid = threadIdx.x;
    const int lr = tid & (kLanes - 1), seg = tid / kLanes;
    const int qid = blockIdx.x % a.qa.nq;
    int tk = wg_first_ticket(a.qa, qid, a.ntasks, &s_next);
    while (tk >= 0) {
        int pre = 0;
        if (tid == 0) pre = take_ticket(a.qa, qid, a.ntasks);   // in flight during the task
        if (a.trace && tid == 0) a.trace[4 * (size_t)tk] = wall_clock64();
        const int blk = a.task[4 * tk], first = a.task[4 * tk + 1], last = a.task[4 * tk + 2];
        const int kind = a.task[4 * tk + 3];
        const int a0 = a.start[blk];
        const int m = a.start[blk + 1] - a0;
        const int R = a.rptr[blk + 1] - a.rptr[blk];
        const int rows = m + R;
        const F *M = a.M + a.moff[blk];
        const VecMap vm = vec_map(a, blk);
        const int k = vm.k;
        F *yv = a.y + vm.vbase, *zv = a.z + vm.vbase, *pv = a.part + vm.pbase, *rb = a.rbuf + vm.vbase;
        const F *bv = a.b + vm.vbase;
        // Rows of r_B this task gathers: all (kind 0), its slice (kind 1), none (kind 2).
        const int g0 = kind == 1 ? first : 0;
        const int g1 = kind == 1 ? last : (kind == 0 ? m : 0);
        for (int t0 = 0; t0 < k; t0 += KT) {
            const int kt = min(KT, k - t0);
            // r_B: each thread owns up to kPairs (row, column) pairs and
            // walks their partial-slot lists together, 4 slots per pair per
            // round, so up to 4 * kPairs slot loads are in flight at once.
            for (int base = g0 * KT + tid; base < g1 * KT; base += kTaskThreads * kPairs) {
                int p0[kPairs], p1[kPairs], col[kPairs];
                F v[kPairs];
#pragma unroll
                for (int q = 0; q < kPairs; ++q) {
                    const int idx = base + q * kTaskThreads;
                    const int i = idx / KT, t = idx % KT;
                    p0[q] = p1[q] = 0;
                    col[q] = t0 + t;
                    v[q] = F(0);
                    if (idx < g1 * KT && t < kt) {
                        const int row = a0 + i;
                        v[q] = bv[(long long)a.perm[row] * k + col[q]];
                        p0[q] = a.cptr[row];
                        p1[q] = a.cptr[row + 1];
                    }
                }
                for (int r = 0;; r += 4) {
                    bool any = false;
                    U u[kPairs][4];
                    int sl[kPairs][4];
#pragma unroll
                    for (int q = 0; q < kPairs; ++q)
#pragma unroll
                        for (int j = 0; j < 4; ++j) {
                            const int p = p0[q] + r + j;
                            sl[q][j] = p < p1[q] ? a.cslot[p] : -1;
                            u[q][j] = sl[q][j] >= 0 ? load_bits(&pv[(long long)sl[q][j] * k + col[q]]) : U(0);
                            any |= sl[q][j] >= 0;
                        }
                    if (!any) break;
#pragma unroll
                    for (int q = 0; q < kPairs; ++q)
#pragma unroll
                        for (int j = 0; j < 4; ++j) {
                            F x = from_bits<F>(u[q][j]);
                            if (sl[q][j] >= 0 && u[q][j] == Bits<F>::sentinel)
                                x = wait_value(&pv[(long long)sl[q][j] * k + col[q]], a.status);
                            v[q] += x;
                        }
                }
#pragma unroll
                for (int q = 0; q < kPairs; ++q) {
                    const int idx = base + q * kTaskThreads;
                    if (idx >= g1 * KT) continue;
                    if (kind == 1) {
                        if (idx % KT < kt) publish(&rb[(long long)(a0 + idx / KT) * k + col[q]], v[q]);
                    } else {
                        rv[idx] = v[q];
                    }
                }
            }
            if (kind == 2) {
                // r_B from the gather tasks of this block (published, waited on).
                for (int base = tid; base < m * KT; base += kTaskThreads * kGatherUnroll) {
                    U u[kGatherUnroll];
#pragma unroll
                    for (int j = 0; j < kGatherUnroll; ++j) {
                        const int idx = base + j * kTaskThreads;
                        const int i = idx / KT, t = idx % KT;
                        u[j] = (idx < m * KT && t < kt) ? load_bits(&rb[(long long)(a0 + i) * k + t0 + t]) : U(0);
                    }
#pragma unroll
                    for (int j = 0; j < kGatherUnroll; ++j) {
                        const int idx = base + j * kTaskThreads;
                        if (idx >= m * KT) break;
                        const int i = idx / KT, t = idx % KT;
                        F v = from_bits<F>(u[j]);
                        if (t < kt && u[j] == Bits<F>::sentinel)
                            v = wait_value(&rb[(long long)(a0 + i) * k + t0 + t], a.status);
                        rv[idx] = v;
                    }
                }
            }
            __syncthreads();
            if (a.trace && tid == 0 && t0 == 0) a.trace[4 * (size_t)tk + 1] = wall_clock64();
            if (kind == 1) {   // gather-only task: nothing to multiply
                __syncthreads();
                continue;
            }
            const int rho = first + lr;
            F acc[KT];
#pragma unroll
            for (int t = 0; t < KT; ++t) acc[t] = F(0);
            if (rho < last) {
                const int cend = rho < m ? rho + 1 : m;
                const int cq = (m + kSegs - 1) / kSegs;
                const int cb = seg * cq, ce = min(cend, cb + cq);
                // Column c of the packed column-major M_B starts (rows - c)
                // entries after column c - 1: walk the offsets incrementally
                // and keep kMU loads in flight before consuming them.
                int64_t o = fwd_off(rows, cb, rho);
                int c = cb;
                for (; c + kMU <= ce; c += kMU) {
                    F mv[kMU];
#pragma unroll
                    for (int j = 0; j < kMU; ++j) {
                        mv[j] = M[o];
                        o += rows - (c + j) - 1;
                    }
#pragma unroll
                    for (int j = 0; j < kMU; ++j)
#pragma unroll
                        for (int t = 0; t < KT; ++t) acc[t] = fma(mv[j], rv[(c + j) * KT + t], acc[t]);
                }
                for (; c < ce; ++c) {
                    const F mv = M[o];
                    o += rows - c - 1;
#pragma unroll
                    for (int t = 0; t < KT; ++t) acc[t] = fma(mv, rv[c * KT + t], acc[t]);
                }
            }
#pragma unroll
            for (int t = 0; t < KT; ++t) red[tid * KT + t] = acc[t];
            __syncthreads();
            if (seg == 0 && rho < last) {
#pragma unroll
                for (int t = 0; t < KT; ++t) {
                    if (t >= kt) break;
                    F s = red[lr * KT + t];
#pragma unroll
                    for (int g = 1; g < kSegs; ++g) s += red[(g * kLanes + lr) * KT + t];
                    const int col = t0 + t;
                    if (rho < m) {
                        const long long o = (long long)(a0 + rho) * k + col;
                        yv[o] = s;
                        zv[o] = from_bits<F>(Bits<F>::sentinel);
                    } else {
                        publish(&pv[(long long)(a.rptr[blk] + rho - m) * k + col], s);
                    }
                }
            }
            __syncthreads();
        }
        if (a.trace && tid == 0) {
            a.trace[4 * (size_t)tk + 2] = wall_clock64();
            a.trace[4 * (size_t)tk + 3] = blockIdx.x;
        }
        if (tid == 0) s_next = pre;
        __syncthreads();
        tk = s_next;
        __syncthreads();
    }
    if (tid == 0) finish_queues(a.qa);
}

template <typename F, int KT>
__global__ __launch_bounds__(kTaskThreads) void block_backward(BlockPassArgs<F> a) {
    using U = typename Bits<F>::U;
    extern __shared__ __align__(16) unsigned char smem[];
    F *vs = reinterpret_cast<F *>(smem);                   // max_rows x KT
    F *red = vs + (size_t)a.max_rows * KT;                 // 256 x KT
    __shared__ int s_next;
    const int tid = threadIdx.x;
    const int lr = tid & (kBLanes - 1), seg = tid / kBLanes;
    const int qid = blockIdx.x % a.qa.nq;
    int tk = wg_first_ticket(a.qa, qid, a.ntasks, &s_next);
    while (tk >= 0) {
        int pre = 0;
        if (tid == 0) pre = take_ticket(a.qa, qid, a.ntasks);   // in flight during the task
        if (a.trace && tid == 0) a.trace[4 * (size_t)tk] = wall_clock64();
        const int blk = a.task[4 * tk], first = a.task[4 * tk + 1], last = a.task[4 * tk + 2];
        const int a0 = a.start[blk];
        const int m = a.start[blk + 1] - a0;
        const int r0 = a.rptr[blk];
        const int R = a.rptr[blk + 1] - r0;
        const int rows = m + R;
        const F *M = a.M + a.moff[blk];
        const VecMap vm = vec_map(a, blk);
        const int k = vm.k;
        F *yv = a.y + vm.vbase, *zv = a.z + vm.vbase, *pv = a.part + vm.pbase;
        for (int t0 = 0; t0 < k; t0 += KT) {
            const int kt = min(KT, k - t0);
            // v = [y_B ; z_R]: y from the forward pass (complete), z waited on.
            for (int base = tid; base < rows * KT; base += kTaskThreads * kGatherUnroll) {
                U u[kGatherUnroll];
                long long addr[kGatherUnroll];
#pragma unroll
                for (int j = 0; j < kGatherUnroll; ++j) {
                    const int idx = base + j * kTaskThreads;
                    const int rho = idx / KT, t = idx % KT;
                    addr[j] = -1;
                    u[j] = U(0);
                    if (idx < rows * KT && t < kt) {
                        if (rho < m) {
                            u[j] = as_bits(yv[(long long)(a0 + rho) * k + t0 + t]);
                        } else {
                            addr[j] = (long long)a.rlist[r0 + rho - m] * k + t0 + t;
                            u[j] = load_bits(&zv[addr[j]]);
                        }
                    }
                }
#pragma unroll
                for (int j = 0; j < kGatherUnroll; ++j) {
                    const int idx = base + j * kTaskThreads;
                    if (idx >= rows * KT) break;
                    F v = from_bits<F>(u[j]);
                    if (addr[j] >= 0 && u[j] == Bits<F>::sentinel) v = wait_value(&zv[addr[j]], a.status);
                    vs[idx] = v;
                }
            }
            __syncthreads();
            if (a.trace && tid == 0 && t0 == 0) a.trace[4 * (size_t)tk + 1] = wall_clock64();
            const int c = first + lr;
            F acc[KT];
#pragma unroll
            for (int t = 0; t < KT; ++t) acc[t] = F(0);
            if (c < last) {
                const int rq = (rows + kBSegs - 1) / kBSegs;
                const int rb = max(seg * rq, c), re = min(rows, seg * rq + rq);
                // Row rho of the packed row-major M_B is min(rho + 1, m)
                // entries long: walk the offsets incrementally, kMU loads in flight.
                int64_t o = bwd_off(m, rb, c);
                int rho = rb;
                for (; rho + kMU <= re; rho += kMU) {
                    F mv[kMU];
#pragma unroll
                    for (int j = 0; j < kMU; ++j) {
                        mv[j] = M[o];
                        o += min(rho + j + 1, m);
                    }
#pragma unroll
                    for (int j = 0; j < kMU; ++j)
#pragma unroll
                        for (int t = 0; t < KT; ++t) acc[t] = fma(mv[j], vs[(rho + j) * KT + t], acc[t]);
                }
                for (; rho < re; ++rho) {
                    const F mv = M[o];
                    o += min(rho + 1, m);
#pragma unroll
                    for (int t = 0; t < KT; ++t) acc[t] = fma(mv, vs[rho * KT + t], acc[t]);
                }
            }
#pragma unroll
            for (int t = 0; t < KT; ++t) red[tid * KT + t] = acc[t];
            __syncthreads();
            if (seg == 0 && c < last) {
#pragma unroll
                for (int t = 0; t < KT; ++t) {
                    if (t >= kt) break;
                    F s = red[lr * KT + t];
#pragma unroll
                    for (int g = 1; g < kBSegs; ++g) s += red[(g * kBLanes + lr) * KT + t];
                    const int col = t0 + t;
                    publish(&zv[(long long)(a0 + c) * k + col], s);
                    const long long o = vm.vbase + (long long)a.perm[a0 + c] * k + col;
                    if (a.xmode == 0) {
                        static_cast<F *>(a.x)[o] = s;
                    } else {
                        double *x64 = static_cast<double *>(a.x);
                        x64[o] = (a.xmode == 2 ? x64[o] : 0.0) + (double)s;
                    }
                }
            }
            __syncthreads();
        }
        // Restore the sentinel in this block's partial slots for the next solve.
        if (first == 0) {
            const F sent = from_bits<F>(Bits<F>::sentinel);
            for (int idx = tid; idx < R * k; idx += kTaskThreads) pv[(long long)r0 * k + idx] = sent;
            F *rb = a.rbuf + vm.vbase;
            for (int idx = tid; idx < m * k; idx += kTaskThreads) rb[(long long)a0 * k + idx] = sent;
        }
        if (a.trace && tid == 0) {
            a.trace[4 * (size_t)tk + 2] = wall_clock64();
            a.trace[4 * (size_t)tk + 3] = blockIdx.x;
        }
        if (tid == 0) s_next = pre;
        __syncthreads();
        tk = s_next;
        __syncthreads();
    }
    if (tid == 0) finish_queues(a.qa);
}

template <typename F>
hipError_t launch_impl(bool forward, const BlockPassBuffers &bf, int grid, QueueState &queues,
                       hipStream_t stream) {
    BlockPassArgs<F> a;
    a.ntasks = bf.ntasks;
    a.k = bf.k;
    a.max_rows = bf.max_rows;
    a.task = bf.task;
    a.start = bf.start;
    a.rptr = bf.rptr;
    a.rlist = bf.rlist;
    a.moff = bf.moff;
    a.M = static_cast<const F *>(bf.M);
    a.cptr = bf.cptr;
    a.cslot = bf.cslot;
    a.perm = bf.perm;
    a.b = static_cast<const F *>(bf.b);
    a.y = static_cast<F *>(bf.y);
    a.z = static_cast<F *>(bf.z);
    a.part = static_cast<F *>(bf.part);
    a.rbuf = static_cast<F *>(bf.rbuf);
    a.x = bf.x;
    a.status = bf.status;
    a.xmode = bf.xmode;
    a.trace = bf.trace;
    a.blk_sys = bf.blk_sys;
    a.sys_k = bf.sys_k;
    a.sys_vbase = bf.sys_vbase;
    a.sys_pbase = bf.sys_pbase;
    queues.fill(a.qa);
    void *args[] = {&a};
    const int tile = block_tile(sizeof(F), bf.max_k);
    const size_t lds = block_pass_lds(sizeof(F), bf.max_rows, tile);
    return hipLaunchKernel(block_pass_kernel(sizeof(F), forward, tile), dim3(grid), dim3(kTaskThreads), args,
                           lds, stream);
}

}  // namespace

size_t block_pass_lds(int float_bytes, int max_rows, int tile) {
    return (size_t)(max_rows + kTaskThreads) * tile * (size_t)float_bytes;
}

const void *block_pass_kernel(int float_bytes, bool forward, int tile) {
    if (float_bytes == 4) {
        if (tile == 4)
            return forward ? reinterpret_cast<const void *>(&block_forward<float, 4>)
                           : reinterpret_cast<const void *>(&block_backward<float, 4>);
        return forward ? reinterpret_cast<const void *>(&block_forward<float, 16>)
                       : reinterpret_cast<const void *>(&block_backward<float, 16>);
    }
    if (tile == 4)
        return forward ? reinterpret_cast<const void *>(&block_forward<double, 4>)
                       : reinterpret_cast<const void *>(&block_backward<double, 4>);
    return forward ? reinterpret_cast<const void *>(&block_forward<double, 8>)
                   : reinterpret_cast<const void *>(&block_backward<double, 8>);
}

hipError_t launch_block_pass(int float_bytes, bool forward, const BlockPassBuffers &buf, int grid,
                             QueueState &queues, hipStream_t stream) {
    return float_bytes == 4 ? launch_impl<float>(forward, buf, grid, queues, stream)
                            : launch_impl<double>(forward, buf, grid, queues, stream);
}

}  // namespace cholespy_amd
