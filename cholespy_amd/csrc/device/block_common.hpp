// Task machinery shared by the block passes of trsv_blocks.hip (narrow
// right-hand-side tiles, vector FMAs) and trsv_panel.hip (all right-hand
// sides at once, MFMA): kernel arguments, the system layout of a task, ticket
// broadcast, the packed layouts of M_B and the gathers that stage r_B or
// [y_B ; z_R] in LDS.
#pragma once
#include "block_plan.hpp"
#include "common.hpp"
#include "launchers.hpp"

namespace cholespy_amd {

template <typename F>
struct BlockPassArgs {
    int ntasks, k;
    int chunk;              // narrow kernels: rows of a vector staged in LDS at a time
    int lds_bytes;          // panel kernels: dynamic LDS of the launch
    const int *task;        // kDevTaskInts ints per task (DeviceTask in launchers.hpp)
    const int *start;       // block row ranges
    const int *rptr;        // R list / partial slot ranges
    const int *rlist;
    const int64_t *moff;
    const F *M;             // forward: column-major packed; backward: row-major packed
    const int *cptr;        // contribution lists (forward)
    const int *cslot;
    const int *gslot;       // each row's first kGatherSlots slots, fixed stride, -1 padded
    const int *perm;
    const F *b;
    F *y, *z, *part, *rbuf;
    F *spart;               // split results (split rows x k)
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
    const long long *sys_sbase;
    QueueArgs qa;
    // Optional task timeline (nullptr = off): per task, 4 words =
    // {start, after the gather phase, end, workgroup id} in wall-clock ticks (100 MHz).
    unsigned long long *trace;
    // Leaf phase of the launch (narrow kernels; n_leaf = 0: none): n_leaf
    // leaf task records at leaf_task, taken per wavefront from the queues
    // lqa, before the block tasks (forward) or after them (backward).
    const int *leaf_task;
    int n_leaf;
    QueueArgs lqa;
    unsigned long long *leaf_trace;
};

// Right-hand-side layout of one task's system.
struct VecMap {
    int k;
    long long vbase, pbase, sbase;
};

// One task record as uploaded: the plan's record (block_plan.hpp) and its
// block's row range, R slots, M_B offset and system (DeviceTask).
struct Task {
    int blk, first, last, kind, lo, hi, sbase, nsplit;
    int a0, m, r0, R;
    int64_t moff;
    int sys;
};

// Panel (MFMA) variants of the block passes, trsv_panel.hip.
// nt = right-hand-side tiles of 16 the launch holds (2, 4 or 8).
int panel_lds_bytes();
const void *block_panel_kernel(int float_bytes, bool forward, int nt);
hipError_t launch_panel_f32(bool forward, const BlockPassArgs<float> &a, int nt, int grid, hipStream_t stream);
hipError_t launch_panel_f64(bool forward, const BlockPassArgs<double> &a, int nt, int grid, hipStream_t stream);

namespace {

constexpr int kTaskThreads = 256;
constexpr int kGatherUnroll = 8;

template <typename F>
__device__ __forceinline__ VecMap vec_map(const BlockPassArgs<F> &a, const Task &task) {
    if (!a.blk_sys) return VecMap{a.k, 0, 0, 0};
    const int sys = task.sys;
    return VecMap{a.sys_k[sys], a.sys_vbase[sys], a.sys_pbase[sys], a.sys_sbase[sys]};
}

// The whole record is one 64-byte line: four 16-byte loads in flight together.
__device__ __forceinline__ Task load_task_at(const int *tasks, int tk) {
    const int4 *t = reinterpret_cast<const int4 *>(tasks + (size_t)kDevTaskInts * tk);
    const int4 w0 = t[0], w1 = t[1], w2 = t[2], w3 = t[3];
    return Task{w0.x, w0.y, w0.z, w0.w, w1.x, w1.y, w1.z, w1.w,
                w2.x, w2.y, w2.z, w2.w,
                (int64_t)(((uint64_t)(uint32_t)w3.y << 32) | (uint32_t)w3.x), w3.z};
}
template <typename F>
__device__ __forceinline__ Task load_task(const BlockPassArgs<F> &a, int tk) {
    const int4 *t = reinterpret_cast<const int4 *>(a.task + (size_t)kDevTaskInts * tk);
    const int4 w0 = t[0], w1 = t[1], w2 = t[2], w3 = t[3];
    return Task{w0.x, w0.y, w0.z, w0.w, w1.x, w1.y, w1.z, w1.w,
                w2.x, w2.y, w2.z, w2.w,
                (int64_t)(((uint64_t)(uint32_t)w3.y << 32) | (uint32_t)w3.x), w3.z};
}

// Value of a published entry: one agent-scope load, and the bounded wait of
// common.hpp only when it still holds the sentinel.
template <typename F>
__device__ __forceinline__ F ready_value(const F *p, typename Bits<F>::U u, int *status, const QueueArgs &qa) {
    return u == Bits<F>::sentinel ? wait_value(p, status, qa) : from_bits<F>(u);
}

// Split epilogue: the split that adds its group's results is split
// nsplit - 1, the group's highest ticket. It waits on the other splits'
// published results (combine_splits), all of earlier tickets, so the
// schedule stays deadlock-free, and no counter round trip precedes the sum
// (a last-finisher counter measured 3 % slower on the icosphere(7) FP64
// solve). The barrier orders the task's own split stores and LDS use before
// the combine.
__device__ __forceinline__ bool combines_splits(int sidx, int nsplit) {
    __syncthreads();
    return sidx == nsplit - 1;
}

// Combine: the nsplit split results of w rows (split s, row i at split row
// sbase + s * w + i) added in split order, so the sum does not depend on
// which split finished last; each split value is reset to the sentinel after
// it is read (the combine is its only reader). Values of the other splits
// that are not yet visible are waited on. out(i, t, value) finishes row i,
// column t.
constexpr int kMaxSplit = 16;
template <typename F, typename Out>
__device__ __forceinline__ void combine_splits(const BlockPassArgs<F> &a, F *sp, int w, int k, int sbase, int nsplit,
                                               Out out) {
    using U = typename Bits<F>::U;
    const int total = w * k;
    for (int idx = threadIdx.x; idx < total; idx += kTaskThreads) {
        const int i = idx / k, t = idx - i * k;
        F *p0 = sp + (long long)(sbase + i) * k + t;
        const long long stride = (long long)w * k;
        U u[kMaxSplit];
#pragma unroll
        for (int s = 0; s < kMaxSplit; ++s) u[s] = s < nsplit ? peek_bits(p0 + s * stride) : U(0);
        F v = F(0);
#pragma unroll
        for (int s = 0; s < kMaxSplit; ++s)
            if (s < nsplit) v += ready_value(p0 + s * stride, u[s], a.status, a.qa);
        const F sent = from_bits<F>(Bits<F>::sentinel);
#pragma unroll
        for (int s = 0; s < kMaxSplit; ++s)
            if (s < nsplit) p0[s * stride] = sent;
        out(i, t, v);
    }
}

// Combine by the last split itself, whose own results (never published) are
// in LDS, own[i * KT + t] for row i, column t: the other nsplit - 1
// splits' results are waited on and added in split order, then the own
// value, so the sum is the same as combine_splits gives; the values read
// are reset to the sentinel. Saves the last split the round trip of
// publishing its own results and reading them back. out(i, t, value)
// finishes row i, column t.
template <typename F, int KT, typename Out>
__device__ __forceinline__ void combine_splits_own(const BlockPassArgs<F> &a, F *sp, int w, int k, int sbase,
                                                   int nsplit, const F *own, Out out) {
    using U = typename Bits<F>::U;
    const int total = w * k;
    for (int idx = threadIdx.x; idx < total; idx += kTaskThreads) {
        const int i = idx / k, t = idx - i * k;
        F *p0 = sp + (long long)(sbase + i) * k + t;
        const long long stride = (long long)w * k;
        U u[kMaxSplit - 1];
#pragma unroll
        for (int s = 0; s < kMaxSplit - 1; ++s) u[s] = s < nsplit - 1 ? peek_bits(p0 + s * stride) : U(0);
        F v = F(0);
#pragma unroll
        for (int s = 0; s < kMaxSplit - 1; ++s)
            if (s < nsplit - 1) v += ready_value(p0 + s * stride, u[s], a.status, a.qa);
        v += own[i * KT + t];
        const F sent = from_bits<F>(Bits<F>::sentinel);
#pragma unroll
        for (int s = 0; s < kMaxSplit - 1; ++s)
            if (s < nsplit - 1) p0[s * stride] = sent;
        out(i, t, v);
    }
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

// Right-hand side r_B for rows [r0, r1) of a block and columns [t0, t0 + kt):
// (P b) plus the partials other blocks published for those rows. Each thread
// owns one or two (row, column) pairs. Per pair it loads the slot indices of up
// to kSlotBatch contributions at once, then their values kValBatch at a time
// (all loads of a batch in flight together), so a row with up to 16
// contributions costs three dependent memory latencies (slot range, slot
// indices, values) plus one per further value batch; the sum runs in slot
// order, so it is deterministic.
// to_rbuf: publish the sums to rb (read by the block's product tasks);
// otherwise store them in LDS, row r0 + i at dst[i * stride + t] for
// t < width (zeros for kt <= t < width).
constexpr int kSlotBatch = 16;
constexpr int kValBatch = 8;
template <typename F>
__device__ __forceinline__ void gather_partials(const BlockPassArgs<F> &a, const F *bv, const F *pv, F *rb,
                                                F *dst, int a0, int r0, int r1, int t0, int kt, int k,
                                                int width, int stride, bool to_rbuf) {
    using U = typename Bits<F>::U;
    // One pair per thread in FP64 keeps the gather's registers within the
    // product loop's budget (occupancy), two in FP32.
    constexpr int kPairs = sizeof(F) == 4 ? 2 : 1;
    static_assert(kSlotBatch == kGatherSlots, "the first slot batch is a row of gslot");
    const int total = (r1 - r0) * width;
    for (int base = threadIdx.x; base < total; base += kTaskThreads * kPairs) {
        int p0[kPairs], p1[kPairs], col[kPairs], grow[kPairs];
        F v[kPairs];
#pragma unroll
        for (int q = 0; q < kPairs; ++q) {
            const int idx = base + q * kTaskThreads;
            const int i = idx / width, t = idx - i * width;
            p0[q] = p1[q] = 0;
            col[q] = t0 + t;
            grow[q] = -1;
            v[q] = F(0);
            if (idx < total && t < kt) {
                const int row = a0 + r0 + i;
                grow[q] = row;
                v[q] = bv[(long long)a.perm[row] * k + col[q]];
                p0[q] = a.cptr[row];
                p1[q] = a.cptr[row + 1];
            }
        }
        for (int r = 0;; r += kSlotBatch) {
            bool any = false;
            int sl[kPairs][kSlotBatch];
#pragma unroll
            for (int q = 0; q < kPairs; ++q) {
                if (r == 0) {
                    // The row's first slots from the fixed-stride table: no
                    // wait for cptr (which is only needed past kSlotBatch).
                    const int4 *g = reinterpret_cast<const int4 *>(a.gslot + (size_t)max(grow[q], 0) * kGatherSlots);
#pragma unroll
                    for (int h = 0; h < kSlotBatch / 4; ++h) {
                        const int4 w = grow[q] >= 0 ? g[h] : make_int4(-1, -1, -1, -1);
                        sl[q][4 * h] = w.x;
                        sl[q][4 * h + 1] = w.y;
                        sl[q][4 * h + 2] = w.z;
                        sl[q][4 * h + 3] = w.w;
                    }
                } else {
#pragma unroll
                    for (int j = 0; j < kSlotBatch; ++j) {
                        const int p = p0[q] + r + j;
                        sl[q][j] = p < p1[q] ? a.cslot[p] : -1;
                    }
                }
#pragma unroll
                for (int j = 0; j < kSlotBatch; ++j) any |= sl[q][j] >= 0;
            }
            if (!any) break;
#pragma unroll
            for (int h = 0; h < kSlotBatch; h += kValBatch) {
                bool more = false;
#pragma unroll
                for (int q = 0; q < kPairs; ++q) more |= sl[q][h] >= 0;
                if (!more) break;
                U u[kPairs][kValBatch];
#pragma unroll
                for (int q = 0; q < kPairs; ++q)
#pragma unroll
                    for (int j = 0; j < kValBatch; ++j)
                        u[q][j] = sl[q][h + j] >= 0 ? peek_bits(&pv[(long long)sl[q][h + j] * k + col[q]]) : U(0);
#pragma unroll
                for (int q = 0; q < kPairs; ++q)
#pragma unroll
                    for (int j = 0; j < kValBatch; ++j) {
                        F x = from_bits<F>(u[q][j]);
                        if (sl[q][h + j] >= 0 && u[q][j] == Bits<F>::sentinel)
                            x = wait_value(&pv[(long long)sl[q][h + j] * k + col[q]], a.status, a.qa);
                        v[q] += x;
                    }
            }
        }
#pragma unroll
        for (int q = 0; q < kPairs; ++q) {
            const int idx = base + q * kTaskThreads;
            if (idx >= total) continue;
            const int i = idx / width, t = idx - i * width;
            if (to_rbuf) {
                if (t < kt) publish(&rb[(long long)(a0 + r0 + i) * k + col[q]], v[q]);
            } else {
                dst[i * stride + t] = v[q];
            }
        }
    }
}

// Rows [r0, r1) of a vector into LDS (row r0 + i at dst[i * stride + t] for
// t < width, zeros for kt <= t < width). `at(row, t, published)` gives the
// address of an entry and whether it is published by another workgroup
// (then it is read with an agent-scope load and waited on while it holds
// the sentinel).
template <typename F, typename At>
__device__ __forceinline__ void gather_rows(const BlockPassArgs<F> &a, F *dst, int r0, int r1, int kt,
                                            int width, int stride, At at) {
    using U = typename Bits<F>::U;
    const int total = (r1 - r0) * width;
    for (int base = threadIdx.x; base < total; base += kTaskThreads * kGatherUnroll) {
        U u[kGatherUnroll];
        F *ad[kGatherUnroll];
        bool pub[kGatherUnroll];
#pragma unroll
        for (int j = 0; j < kGatherUnroll; ++j) {
            const int idx = base + j * kTaskThreads;
            const int i = idx / width, t = idx - i * width;
            u[j] = U(0);
            ad[j] = nullptr;
            pub[j] = false;
            if (idx < total && t < kt) {
                ad[j] = at(r0 + i, t, pub[j]);
                // One load form for published and complete rows (complete
                // rows never hold the sentinel): two forms picked per lane
                // would be two loads into one register, and the compiler
                // waits for the first before issuing the second.
                u[j] = peek_bits(ad[j]);
            }
        }
#pragma unroll
        for (int j = 0; j < kGatherUnroll; ++j) {
            const int idx = base + j * kTaskThreads;
            if (idx >= total) break;
            F v = from_bits<F>(u[j]);
            if (pub[j] && u[j] == Bits<F>::sentinel) v = wait_value(ad[j], a.status, a.qa);
            const int i = idx / width;
            dst[i * stride + (idx - i * width)] = v;
        }
    }
}

}  // namespace
}  // namespace cholespy_amd
