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
// tickets in order from the queues of common.hpp when they start a task (not
// earlier: a ticket taken ahead would hold a ready task back behind the
// current one), so every awaited task is running or done and the schedule
// cannot deadlock.
//
// Each task is one workgroup of 256 threads: up to 64 rows x 4 column
// segments (forward; a wavefront loads 64 consecutive entries of one column,
// or 16 or 32 entries of each of 4 or 2 columns for narrower tasks) or up to
// 32 columns x 8 row segments (backward; 16- or 32-entry pieces of M_B's
// rows), so loads are coalesced; the segment sums are combined in LDS. Every thread
// keeps 16 M loads in flight. Right-hand sides are processed in tiles of KT
// columns held in LDS.
//
// Sentinel hand-off as in common.hpp. The forward pass resets z for its rows,
// the backward pass resets the partial slots of its block, so the next solve
// starts from the all-sentinel state without a memset.
#include <algorithm>

#include "block_common.hpp"

namespace cholespy_amd {

namespace {

constexpr int kLanes = 64;      // forward: most rows of M_B per task
constexpr int kBLanes = 32;     // backward: most columns of M_B per task
#ifndef CHOLESPY_KMU
#define CHOLESPY_KMU 16
#endif
constexpr int kMU = CHOLESPY_KMU;   // M loads issued together per thread
#ifndef CHOLESPY_KPRE
#define CHOLESPY_KPRE 16
#endif
// M entries per thread loaded before the task's right-hand-side gather, which
// may wait on other tasks: their latency overlaps the wait and the gather.
constexpr int kPre = CHOLESPY_KPRE;

// Leaf blocks (block_plan.hpp: no child blocks, at most kLeafMaxM rows and
// kLeafMaxRows rows of M_B): one wavefront per task. Their forward tasks
// take nothing from other blocks (r_B = (P b)_B) and their backward tasks
// give nothing to other blocks, so they are the first forward tickets and
// the last backward tickets. They run either inside the block pass's launch
// (its leaf phase: every wavefront takes leaf tickets from queues of its
// own before (forward) or after (backward) the workgroup's block tasks, and
// a backward leaf task waits on the z_R the block tasks publish) or in a
// launch of their own before / after the block launch (no tickets and no
// waits). The tasks and the results are those of the block kernels; only
// the lane layout and the reduction order inside a task differ.
constexpr int kLeafWaves = kTaskThreads / kWave;
#ifndef CHOLESPY_KLEAFMU
#define CHOLESPY_KLEAFMU 16
#endif
constexpr int kLeafMU = CHOLESPY_KLEAFMU;   // M loads in flight per lane
#ifndef CHOLESPY_KLEAFMU_B
#define CHOLESPY_KLEAFMU_B CHOLESPY_KLEAFMU
#endif
constexpr int kLeafMUB = CHOLESPY_KLEAFMU_B;   // the same for backward leaf tasks

// Orders this wavefront's LDS writes before its later LDS reads by other lanes.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Forward leaf task tk of the records `tasks` (r: kLeafMaxM x KT of this
// wavefront's LDS).
template <typename F, int KT>
__device__ __forceinline__ void leaf_forward_task(const BlockPassArgs<F> &a, const int *tasks,
                                                  unsigned long long *trace, int tk, F *r) {
    const int lane = threadIdx.x % kWave;
    if (trace && lane == 0) trace[4 * (size_t)tk] = wall_clock64();
    const Task task = load_task_at(tasks, tk);
    const int first = task.first, last = task.last, a0 = task.a0, m = task.m;
    const int rows = m + task.R;
    const F *M = a.M + task.moff;
    const VecMap vm = vec_map(a, task);
    const int k = vm.k;
    // Lane = row rho of M_B; row rho uses columns [0, cend).
    const int rho = first + lane;
    const int cend = rho < last ? min(rho + 1, m) : 0;
    F mv[kLeafMU];
    int64_t o = fwd_off(rows, 0, rho);
#pragma unroll
    for (int j = 0; j < kLeafMU; ++j) {
        mv[j] = j < cend ? M[o] : F(0);
        o += rows - j - 1;
    }
    // r_B = (P b)_B, rows [0, m).
    if (lane < m) {
        const long long pr = vm.vbase + (long long)a.perm[a0 + lane] * k;
#pragma unroll
        for (int t = 0; t < KT; ++t) r[lane * KT + t] = t < k ? a.b[pr + t] : F(0);
    }
    wave_sync();
    if (trace && lane == 0) trace[4 * (size_t)tk + 1] = wall_clock64();
    F acc[KT];
#pragma unroll
    for (int t = 0; t < KT; ++t) acc[t] = F(0);
    for (int c = 0;;) {
#pragma unroll
        for (int j = 0; j < kLeafMU; ++j) {
            if (c + j < cend) {
                const F *rr = r + (c + j) * KT;
#pragma unroll
                for (int t = 0; t < KT; ++t) acc[t] = fma(mv[j], rr[t], acc[t]);
            }
        }
        c += kLeafMU;
        if (c >= cend) break;
#pragma unroll
        for (int j = 0; j < kLeafMU; ++j) {
            mv[j] = c + j < cend ? M[o] : F(0);
            o += rows - (c + j) - 1;
        }
    }
    if (rho < last) {
        F *yv = a.y + vm.vbase, *zv = a.z + vm.vbase, *pv = a.part + vm.pbase;
#pragma unroll
        for (int t = 0; t < KT; ++t) {
            if (t >= k) break;
            if (rho < m) {
                const long long q = (long long)(a0 + rho) * k + t;
                yv[q] = acc[t];
                zv[q] = from_bits<F>(Bits<F>::sentinel);
            } else {
                publish(&pv[(long long)(task.r0 + rho - m) * k + t], acc[t]);
            }
        }
    }
    if (trace && lane == 0) {
        trace[4 * (size_t)tk + 2] = wall_clock64();
        trace[4 * (size_t)tk + 3] = blockIdx.x;
    }
    // The next task reuses r.
    wave_sync();
}

// Backward leaf task tk (v: kLeafMaxRows x KT of this wavefront's LDS).
// WAIT: z_R is published by block tasks of the same launch (read with the
// bounded wait); otherwise it is complete before the launch.
template <typename F, int KT, bool WAIT>
__device__ __forceinline__ void leaf_backward_task(const BlockPassArgs<F> &a, const int *tasks,
                                                   unsigned long long *trace, int tk, F *v) {
    using U = typename Bits<F>::U;
    const int lane = threadIdx.x % kWave;
    if (trace && lane == 0) trace[4 * (size_t)tk] = wall_clock64();
    const Task task = load_task_at(tasks, tk);
    const int first = task.first, last = task.last, a0 = task.a0, m = task.m, r0 = task.r0, R = task.R;
    const int rows = m + R;
    const F *M = a.M + task.moff;
    const VecMap vm = vec_map(a, task);
    const int k = vm.k;
    F *yv = a.y + vm.vbase, *zv = a.z + vm.vbase, *pv = a.part + vm.pbase;
    // Lanes l and l + 32 share column c = first + l % 32 of M_B and take the
    // rows rho = first + l / 32 + 2 j, rho >= c (rows of one step are
    // consecutive entries of the row-major packed M_B across the lanes).
    const int c = first + (lane & 31), h = lane >> 5;
    const bool active = c < last;
    F mv[kLeafMUB];
#pragma unroll
    for (int j = 0; j < kLeafMUB; ++j) {
        const int rr = first + h + 2 * j;
        mv[j] = active && rr >= c && rr < rows ? M[bwd_off(m, rr, c)] : F(0);
    }
    // v = [y_B ; z_R], rows [first, rows): y from the forward pass, z_R from
    // the backward block tasks.
    for (int i = first + lane; i < rows; i += kWave) {
        if (i < m || !WAIT) {
            const F *src = i < m ? yv + (long long)(a0 + i) * k : zv + (long long)a.rlist[r0 + i - m] * k;
#pragma unroll
            for (int t = 0; t < KT; ++t) v[(i - first) * KT + t] = t < k ? src[t] : F(0);
        } else {
            const F *src = zv + (long long)a.rlist[r0 + i - m] * k;
            U u[KT];
#pragma unroll
            for (int t = 0; t < KT; ++t) u[t] = t < k ? peek_bits(src + t) : U(0);
#pragma unroll
            for (int t = 0; t < KT; ++t)
                v[(i - first) * KT + t] = t < k ? ready_value(src + t, u[t], a.status, a.qa) : F(0);
        }
    }
    wave_sync();
    if (trace && lane == 0) trace[4 * (size_t)tk + 1] = wall_clock64();
    F acc[KT];
#pragma unroll
    for (int t = 0; t < KT; ++t) acc[t] = F(0);
    for (int j0 = 0;; j0 += kLeafMUB) {
#pragma unroll
        for (int j = 0; j < kLeafMUB; ++j) {
            const int rr = first + h + 2 * (j0 + j);
            if (active && rr >= c && rr < rows) {
                const F *vv = v + (rr - first) * KT;
#pragma unroll
                for (int t = 0; t < KT; ++t) acc[t] = fma(mv[j], vv[t], acc[t]);
            }
        }
        if (first + h + 2 * (j0 + kLeafMUB) >= rows) break;
#pragma unroll
        for (int j = 0; j < kLeafMUB; ++j) {
            const int rr = first + h + 2 * (j0 + kLeafMUB + j);
            mv[j] = active && rr >= c && rr < rows ? M[bwd_off(m, rr, c)] : F(0);
        }
    }
#pragma unroll
    for (int t = 0; t < KT; ++t) acc[t] += __shfl_xor(acc[t], 32, kWave);
    if (h == 0 && active) {
#pragma unroll
        for (int t = 0; t < KT; ++t) {
            if (t >= k) break;
            publish(&zv[(long long)(a0 + c) * k + t], acc[t]);
            const long long q = vm.vbase + (long long)a.perm[a0 + c] * k + t;
            if (a.xmode == 0) {
                static_cast<F *>(a.x)[q] = acc[t];
            } else {
                double *x64 = static_cast<double *>(a.x);
                x64[q] = (a.xmode == 2 ? x64[q] : 0.0) + (double)acc[t];
            }
        }
    }
    // Restore the sentinel in the block's partial slots for the next solve
    // (once per block: its first column task). Its r_B rows in rbuf need no
    // reset: a leaf block has m < split_gather, so no forward task of the
    // narrow list (the only one the leaf tasks serve) gathers it into rbuf,
    // and a wide solve's backward pass resets what its forward pass wrote.
    if (first == 0) {
        const F sent = from_bits<F>(Bits<F>::sentinel);
        for (int idx = lane; idx < R * k; idx += kWave) pv[(long long)r0 * k + idx] = sent;
    }
    if (trace && lane == 0) {
        trace[4 * (size_t)tk + 2] = wall_clock64();
        trace[4 * (size_t)tk + 3] = blockIdx.x;
    }
    // The next task reuses v.
    wave_sync();
}

// Leaf phase of a fused block launch: every wavefront takes leaf tickets
// (a.lqa, nq queues; wavefront w serves queue w mod nq) until its queue is
// exhausted. The next ticket is requested while a task runs: a held leaf
// ticket is run by this running wavefront right after, so no wait can end
// on a ticket nobody runs. lds: this workgroup's dynamic LDS (kLeafWaves
// slices of kLeafMaxRows x KT).
template <typename F, int KT, bool FORWARD>
__device__ __forceinline__ void leaf_phase(const BlockPassArgs<F> &a, F *lds) {
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    F *mine = lds + (size_t)wave * (kLeafMaxRows * KT);
    const int q = (blockIdx.x * kLeafWaves + wave) % a.lqa.nq;
    int t = lane == 0 ? take_ticket(a.lqa, q, a.n_leaf) : 0;
    t = __shfl(t, 0, kWave);
    while (t >= 0) {
        int next = lane == 0 ? take_ticket(a.lqa, q, a.n_leaf) : 0;
        if (FORWARD)
            leaf_forward_task<F, KT>(a, a.leaf_task, a.leaf_trace, t, mine);
        else
            leaf_backward_task<F, KT, true>(a, a.leaf_task, a.leaf_trace, t, mine);
        t = __shfl(next, 0, kWave);
    }
}

template <typename F, int KT>
__global__ __launch_bounds__(kTaskThreads) void block_forward(BlockPassArgs<F> a) {
    extern __shared__ __align__(16) unsigned char smem[];
    // One chunk of r_B (a.chunk rows x KT); after the products the same
    // space holds the 256 x KT per-thread sums.
    F *rv = reinterpret_cast<F *>(smem);
    __shared__ int s_next;
    const int tid = threadIdx.x;
    const int qid = blockIdx.x % a.qa.nq;
    if (a.n_leaf > 0) {   // leaf tasks first: they wait on nothing
        leaf_phase<F, KT, true>(a, rv);
        __syncthreads();
    }
    int tk = wg_first_ticket(a.qa, qid, a.ntasks, &s_next);
    while (tk >= 0) {
        if (a.trace && tid == 0) a.trace[4 * (size_t)tk] = wall_clock64();
        const Task task = load_task(a, tk);
        const int first = task.first, last = task.last, kind = task.kind;
        // Lane layout of the product: `lanes` rows x `segs` column segments.
        const int lanes = last - first <= 16 ? 16 : (last - first <= 32 ? 32 : kLanes);
        const int segs = kTaskThreads / lanes;
        const int lr = tid & (lanes - 1), seg = tid / lanes;
        const int a0 = task.a0, m = task.m, R = task.R;
        const int rows = m + R;
        const F *M = a.M + task.moff;
        const VecMap vm = vec_map(a, task);
        const int k = vm.k;
        F *yv = a.y + vm.vbase, *zv = a.z + vm.vbase, *pv = a.part + vm.pbase, *rb = a.rbuf + vm.vbase;
        const F *bv = a.b + vm.vbase;
        F *sp = a.spart + vm.sbase;
        const int rho = first + lr;
        // Finishes row r, column t of the task: y (rows < m) or a partial slot.
        auto finish = [&](int r, int t, F v) {
            if (r < m) {
                const long long o = (long long)(a0 + r) * k + t;
                yv[o] = v;
                zv[o] = from_bits<F>(Bits<F>::sentinel);
            } else {
                publish(&pv[(long long)(task.r0 + r - m) * k + t], v);
            }
        };
        // Reduction range: all columns, or [lo, hi) for a split task, whose
        // rows go to split rows sbase + sidx * (last - first).
        const int clo = kind == 3 ? task.lo : 0, chi = kind == 3 ? task.hi : m;
        const int nsp = split_count(task.nsplit), sidx = split_index(task.nsplit);
        const int srow = task.sbase + sidx * (last - first) - first;
        // The last split of a group keeps its own results for the combine, in
        // LDS at [(rho - first) * KT + t]: the one entry of the products'
        // sums that only this row's reduction reads, after it has read it.
        const bool keep_own = kind == 3 && sidx == nsp - 1 && k <= KT;
        // This thread's columns of the first chunk (see the product below);
        // the first kPre of them are loaded before the gather.
        auto cols_of = [&](int c0, int c1, int &cb, int &ce) {
            const int cuse = min(c1, min(last, m));
            const int cend = min(rho < m ? rho + 1 : m, c1);
            const int cq = max(0, (cuse - c0 + segs - 1) / segs);
            cb = c0 + seg * cq;
            ce = min(cend, cb + cq);
        };
        F mpre[kPre];
        int npre = 0;
        if (kind != 1 && rho < last) {
            int cb, ce;
            cols_of(clo, min(chi, clo + a.chunk), cb, ce);
            npre = max(0, min(kPre, ce - cb));
            int64_t o = npre > 0 ? fwd_off(rows, cb, rho) : 0;
#pragma unroll
            for (int j = 0; j < kPre; ++j) {
                mpre[j] = j < npre ? M[o] : F(0);
                o += rows - (cb + j) - 1;
            }
        }
        for (int t0 = 0; t0 < k; t0 += KT) {
            const int kt = min(KT, k - t0);
            if (kind == 1) {   // gather-only task: rows [first, last) of r_B to rbuf
                gather_partials(a, bv, pv, rb, (F *)nullptr, a0, first, last, t0, kt, k, KT, KT, true);
                if (a.trace && tid == 0 && t0 == 0) a.trace[4 * (size_t)tk + 1] = wall_clock64();
                continue;
            }
            F acc[KT];
#pragma unroll
            for (int t = 0; t < KT; ++t) acc[t] = F(0);
            for (int c0 = clo; c0 < chi; c0 += a.chunk) {
                const int c1 = min(chi, c0 + a.chunk);
                if (kind == 0) {
                    gather_partials(a, bv, pv, rb, rv, a0, c0, c1, t0, kt, k, KT, KT, false);
                } else {   // r_B from the gather tasks of this block (published, waited on)
                    gather_rows(a, rv, c0, c1, kt, KT, KT, [&](int i, int t, bool &pub) {
                        pub = true;
                        return rb + (long long)(a0 + i) * k + t0 + t;
                    });
                }
                __syncthreads();
                if (a.trace && tid == 0 && t0 == 0 && c0 == clo) a.trace[4 * (size_t)tk + 1] = wall_clock64();
                if (rho < last) {
                    // The columns the task's rows use, [c0, min(c1, cuse)), split
                    // evenly over the segments (all lanes of a wavefront load the
                    // same column: one coalesced access); row rho uses c < cend.
                    int cb, ce;
                    cols_of(c0, c1, cb, ce);
                    int c = cb;
                    if (t0 == 0 && c0 == clo && npre > 0) {   // the prefetched columns
#pragma unroll
                        for (int j = 0; j < kPre; ++j) {
                            if (j < npre) {
                                const F *r = rv + (size_t)(cb + j - c0) * KT;
#pragma unroll
                                for (int t = 0; t < KT; ++t) acc[t] = fma(mpre[j], r[t], acc[t]);
                            }
                        }
                        c += npre;
                    }
                    // Column c of the packed column-major M_B starts (rows - c)
                    // entries after column c - 1: walk the offsets incrementally,
                    // kMU loads in flight per batch (the last batch predicated).
                    int64_t o = c < ce ? fwd_off(rows, c, rho) : 0;
                    for (; c < ce; c += kMU) {
                        F mv[kMU];
#pragma unroll
                        for (int j = 0; j < kMU; ++j) {
                            mv[j] = c + j < ce ? M[o] : F(0);
                            o += rows - (c + j) - 1;
                        }
#pragma unroll
                        for (int j = 0; j < kMU; ++j) {
                            if (c + j < ce) {
                                const F *r = rv + (size_t)(c + j - c0) * KT;
#pragma unroll
                                for (int t = 0; t < KT; ++t) acc[t] = fma(mv[j], r[t], acc[t]);
                            }
                        }
                    }
                }
                __syncthreads();
            }
#pragma unroll
            for (int t = 0; t < KT; ++t) rv[tid * KT + t] = acc[t];
            __syncthreads();
            if (seg == 0 && rho < last) {
#pragma unroll
                for (int t = 0; t < KT; ++t) {
                    if (t >= kt) break;
                    F s = rv[lr * KT + t];
                    for (int g = 1; g < segs; ++g) s += rv[(g * lanes + lr) * KT + t];
                    const int col = t0 + t;
                    if (keep_own)
                        rv[lr * KT + t] = s;
                    else if (kind == 3)
                        publish(&sp[(long long)(srow + rho) * k + col], s);
                    else
                        finish(rho, col, s);
                }
            }
            __syncthreads();
        }
        if (keep_own) {   // (k <= KT: one tile, its closing barrier orders the LDS)
            combine_splits_own<F, KT>(a, sp, last - first, k, task.sbase, nsp, rv,
                                      [&](int i, int t, F v) { finish(first + i, t, v); });
        } else if (kind == 3 && combines_splits(sidx, nsp)) {
            combine_splits(a, sp, last - first, k, task.sbase, nsp,
                           [&](int i, int t, F v) { finish(first + i, t, v); });
        }
        if (a.trace && tid == 0) {
            a.trace[4 * (size_t)tk + 2] = wall_clock64();
            a.trace[4 * (size_t)tk + 3] = blockIdx.x;
        }
        if (tid == 0) s_next = take_ticket(a.qa, qid, a.ntasks);
        __syncthreads();
        tk = s_next;
        __syncthreads();
    }
    __syncthreads();
    if (tid == 0) finish_queues(a.qa, a.lqa.ctr, a.n_leaf > 0 ? a.lqa.nq : 0);
}

template <typename F, int KT>
__global__ __launch_bounds__(kTaskThreads) void block_backward(BlockPassArgs<F> a) {
    extern __shared__ __align__(16) unsigned char smem[];
    // One chunk of v = [y_B ; z_R] (a.chunk rows x KT); after the products
    // the same space holds the 256 x KT per-thread sums.
    F *vs = reinterpret_cast<F *>(smem);
    __shared__ int s_next;
    const int tid = threadIdx.x;
    const int qid = blockIdx.x % a.qa.nq;
    int tk = wg_first_ticket(a.qa, qid, a.ntasks, &s_next);
    while (tk >= 0) {
        if (a.trace && tid == 0) a.trace[4 * (size_t)tk] = wall_clock64();
        const Task task = load_task(a, tk);
        const int first = task.first, last = task.last, kind = task.kind;
        // Lane layout of the product: `lanes` columns x `segs` row segments.
        const int lanes = last - first <= 16 ? 16 : kBLanes;
        const int segs = kTaskThreads / lanes;
        const int lr = tid & (lanes - 1), seg = tid / lanes;
        const int a0 = task.a0, m = task.m, r0 = task.r0, R = task.R;
        const int rows = m + R;
        const F *M = a.M + task.moff;
        const VecMap vm = vec_map(a, task);
        const int k = vm.k;
        F *yv = a.y + vm.vbase, *zv = a.z + vm.vbase, *pv = a.part + vm.pbase;
        F *sp = a.spart + vm.sbase;
        const int c = first + lr;
        // Finishes column cc, right-hand side col: z published, x written.
        auto finish = [&](int cc, int col, F s) {
            publish(&zv[(long long)(a0 + cc) * k + col], s);
            const long long o = vm.vbase + (long long)a.perm[a0 + cc] * k + col;
            if (a.xmode == 0) {
                static_cast<F *>(a.x)[o] = s;
            } else {
                double *x64 = static_cast<double *>(a.x);
                x64[o] = (a.xmode == 2 ? x64[o] : 0.0) + (double)s;
            }
        };
        // Reduction range: all rows, or [lo, hi) for a split task, whose
        // columns go to split rows sbase + sidx * (last - first).
        const int hlo = kind == 1 ? task.lo : 0, hhi = kind == 1 ? task.hi : rows;
        const int nsp = split_count(task.nsplit), sidx = split_index(task.nsplit);
        const int srow = task.sbase + sidx * (last - first) - first;
        // The last split of a group keeps its own results for the combine, in
        // LDS at [(c - first) * KT + t] (see block_forward).
        const bool keep_own = kind == 1 && sidx == nsp - 1 && k <= KT;
        // This thread's rows of a chunk (see the product below); the first
        // kPre of the first chunk are loaded before the gather.
        auto rows_of = [&](int h0, int h1, int &rb, int &re) {
            const int hb = max(h0, first);
            const int rq = max(0, (h1 - hb + segs - 1) / segs);
            rb = max(hb + seg * rq, c);
            re = min(h1, hb + seg * rq + rq);
        };
        F mpre[kPre];
        int npre = 0;
        if (c < last) {
            int rb, re;
            rows_of(hlo, min(hhi, hlo + a.chunk), rb, re);
            npre = max(0, min(kPre, re - rb));
            int64_t o = npre > 0 ? bwd_off(m, rb, c) : 0;
#pragma unroll
            for (int j = 0; j < kPre; ++j) {
                mpre[j] = j < npre ? M[o] : F(0);
                o += min(rb + j + 1, m);
            }
        }
        for (int t0 = 0; t0 < k; t0 += KT) {
            const int kt = min(KT, k - t0);
            F acc[KT];
#pragma unroll
            for (int t = 0; t < KT; ++t) acc[t] = F(0);
            for (int h0 = hlo; h0 < hhi; h0 += a.chunk) {
                const int h1 = min(hhi, h0 + a.chunk);
                // v = [y_B ; z_R]: y from the forward pass (complete), z waited on.
                gather_rows(a, vs, h0, h1, kt, KT, KT, [&](int rho, int t, bool &pub) {
                    pub = rho >= m;
                    return pub ? zv + (long long)a.rlist[r0 + rho - m] * k + t0 + t
                               : yv + (long long)(a0 + rho) * k + t0 + t;
                });
                __syncthreads();
                if (a.trace && tid == 0 && t0 == 0 && h0 == hlo) a.trace[4 * (size_t)tk + 1] = wall_clock64();
                if (c < last) {
                    // Rows [h0, h1) split evenly over the segments; column c
                    // uses rows rho >= c. Row rho of the packed row-major M_B
                    // is min(rho + 1, m) entries long: walk the offsets
                    // incrementally, kMU loads in flight per batch.
                    // Rows [max(h0, first), h1) split evenly over the segments
                    // (rows below the task's first column are zero in its columns).
                    int rb, re;
                    rows_of(h0, h1, rb, re);
                    int rho = rb;
                    if (t0 == 0 && h0 == hlo && npre > 0) {   // the prefetched rows
#pragma unroll
                        for (int j = 0; j < kPre; ++j) {
                            if (j < npre) {
                                const F *v = vs + (size_t)(rb + j - h0) * KT;
#pragma unroll
                                for (int t = 0; t < KT; ++t) acc[t] = fma(mpre[j], v[t], acc[t]);
                            }
                        }
                        rho += npre;
                    }
                    int64_t o = rho < re ? bwd_off(m, rho, c) : 0;
                    for (; rho < re; rho += kMU) {
                        F mv[kMU];
#pragma unroll
                        for (int j = 0; j < kMU; ++j) {
                            mv[j] = rho + j < re ? M[o] : F(0);
                            o += min(rho + j + 1, m);
                        }
#pragma unroll
                        for (int j = 0; j < kMU; ++j) {
                            if (rho + j < re) {
                                const F *v = vs + (size_t)(rho + j - h0) * KT;
#pragma unroll
                                for (int t = 0; t < KT; ++t) acc[t] = fma(mv[j], v[t], acc[t]);
                            }
                        }
                    }
                }
                __syncthreads();
            }
#pragma unroll
            for (int t = 0; t < KT; ++t) vs[tid * KT + t] = acc[t];
            __syncthreads();
            if (seg == 0 && c < last) {
#pragma unroll
                for (int t = 0; t < KT; ++t) {
                    if (t >= kt) break;
                    F s = vs[lr * KT + t];
                    for (int g = 1; g < segs; ++g) s += vs[(g * lanes + lr) * KT + t];
                    const int col = t0 + t;
                    if (keep_own)
                        vs[lr * KT + t] = s;
                    else if (kind == 1)
                        publish(&sp[(long long)(srow + c) * k + col], s);
                    else
                        finish(c, col, s);
                }
            }
            __syncthreads();
        }
        if (keep_own) {   // (k <= KT: one tile, its closing barrier orders the LDS)
            combine_splits_own<F, KT>(a, sp, last - first, k, task.sbase, nsp, vs,
                                      [&](int i, int t, F v) { finish(first + i, t, v); });
        } else if (kind == 1 && combines_splits(sidx, nsp)) {
            combine_splits(a, sp, last - first, k, task.sbase, nsp,
                           [&](int i, int t, F v) { finish(first + i, t, v); });
        }
        // Restore the sentinel in this block's partial slots for the next solve
        // (once per block: its first column task, or that task's first split).
        if (first == 0 && sidx == 0) {
            const F sent = from_bits<F>(Bits<F>::sentinel);
            for (int idx = tid; idx < R * k; idx += kTaskThreads) pv[(long long)r0 * k + idx] = sent;
            F *rb = a.rbuf + vm.vbase;
            for (int idx = tid; idx < m * k; idx += kTaskThreads) rb[(long long)a0 * k + idx] = sent;
        }
        if (a.trace && tid == 0) {
            a.trace[4 * (size_t)tk + 2] = wall_clock64();
            a.trace[4 * (size_t)tk + 3] = blockIdx.x;
        }
        if (tid == 0) s_next = take_ticket(a.qa, qid, a.ntasks);
        __syncthreads();
        tk = s_next;
        __syncthreads();
    }
    if (a.n_leaf > 0) {   // leaf tasks last: they wait on the block tasks' z
        leaf_phase<F, KT, false>(a, vs);
        __syncthreads();
    }
    if (tid == 0) finish_queues(a.qa, a.lqa.ctr, a.n_leaf > 0 ? a.lqa.nq : 0);
}

// Leaf tasks in a launch of their own: one wavefront per task, a.ntasks
// tasks at a.task.
template <typename F, int KT>
__global__ __launch_bounds__(kTaskThreads) void leaf_forward(BlockPassArgs<F> a) {
    __shared__ F rs[kLeafWaves][kLeafMaxM * KT];
    const int wave = threadIdx.x / kWave;
    const int tk = blockIdx.x * kLeafWaves + wave;
    if (tk >= a.ntasks) return;
    leaf_forward_task<F, KT>(a, a.task, a.trace, tk, rs[wave]);
}

template <typename F, int KT>
__global__ __launch_bounds__(kTaskThreads) void leaf_backward(BlockPassArgs<F> a) {
    __shared__ F vsh[kLeafWaves][kLeafMaxRows * KT];
    const int wave = threadIdx.x / kWave;
    const int tk = blockIdx.x * kLeafWaves + wave;
    if (tk >= a.ntasks) return;
    leaf_backward_task<F, KT, false>(a, a.task, a.trace, tk, vsh[wave]);
}

template <typename F>
hipError_t launch_narrow(bool forward, const BlockPassArgs<F> &a, int tile, size_t lds, int grid,
                         hipStream_t stream) {
    BlockPassArgs<F> args = a;
    void *p[] = {&args};
    return hipLaunchKernel(block_pass_kernel(sizeof(F), forward, tile), dim3(grid), dim3(kTaskThreads), p, lds,
                           stream);
}

template <typename F>
hipError_t launch_impl(bool forward, const BlockPassBuffers &bf, int grid, QueueState &queues,
                       hipStream_t stream) {
    BlockPassArgs<F> a;
    a.ntasks = bf.ntasks;
    a.k = bf.k;
    const int tile = bf.tile ? bf.tile : block_tile(sizeof(F), bf.max_k);
    a.chunk = block_chunk(sizeof(F), kNarrowTile);
    a.lds_bytes = 0;
    a.task = bf.task;
    a.start = bf.start;
    a.rptr = bf.rptr;
    a.rlist = bf.rlist;
    a.moff = bf.moff;
    a.M = static_cast<const F *>(bf.M);
    a.cptr = bf.cptr;
    a.cslot = bf.cslot;
    a.gslot = bf.gslot;
    a.perm = bf.perm;
    a.b = static_cast<const F *>(bf.b);
    a.y = static_cast<F *>(bf.y);
    a.z = static_cast<F *>(bf.z);
    a.part = static_cast<F *>(bf.part);
    a.rbuf = static_cast<F *>(bf.rbuf);
    a.spart = static_cast<F *>(bf.spart);
    a.x = bf.x;
    a.status = bf.status;
    a.xmode = bf.xmode;
    a.trace = bf.trace;
    a.blk_sys = bf.blk_sys;
    a.sys_k = bf.sys_k;
    a.sys_vbase = bf.sys_vbase;
    a.sys_pbase = bf.sys_pbase;
    a.sys_sbase = bf.sys_sbase;
    queues.fill(a.qa);
    a.leaf_task = nullptr;
    a.n_leaf = 0;
    a.leaf_trace = nullptr;
    if (bf.leaf_queues) bf.leaf_queues->fill(a.lqa);
    else a.lqa = QueueArgs{nullptr, 1, a.qa.spin_limit, a.qa.abort};
    if (is_panel_tile(tile)) {
        if constexpr (sizeof(F) == 4) return launch_panel_f32(forward, a, tile / 16, grid, stream);
        else return launch_panel_f64(forward, a, tile / 16, grid, stream);
    }
    size_t lds = block_pass_lds(sizeof(F), forward ? bf.max_m : bf.max_rows, tile);
    // Leaf tasks (forward: the first n_leaf, backward: the last n_leaf): in
    // the block launch's leaf phase (fused), or in their own launch before /
    // after the block pass over the others.
    const int n_leaf = std::min(bf.n_leaf, bf.ntasks);
    const int n_main = bf.ntasks - n_leaf;
    if (bf.fused_leaves && n_leaf > 0) {
        BlockPassArgs<F> f = a;
        const size_t leaf_off = forward ? 0 : (size_t)n_main;
        const size_t main_off = forward ? (size_t)n_leaf : 0;
        f.task = a.task + main_off * kDevTaskInts;
        f.ntasks = n_main;
        f.leaf_task = a.task + leaf_off * kDevTaskInts;
        f.n_leaf = n_leaf;
        if (a.trace) {
            f.trace = a.trace + 4 * main_off;
            f.leaf_trace = a.trace + 4 * leaf_off;
        }
        lds = std::max(lds, (size_t)kLeafWaves * kLeafMaxRows * kNarrowTile * sizeof(F));
        const int g = std::max(1, std::min(grid, std::max(n_main, (n_leaf + kLeafWaves - 1) / kLeafWaves)));
        // Every leaf queue needs a wavefront serving it (wavefront w serves
        // queue w mod nq, and the first workgroups are dispatched first).
        f.lqa.nq = std::max(1, std::min(f.lqa.nq, g * kLeafWaves));
        return launch_narrow<F>(forward, f, tile, lds, g, stream);
    }
    auto leaf = [&](int offset) -> hipError_t {
        BlockPassArgs<F> l = a;
        l.task = a.task + (size_t)offset * kDevTaskInts;
        l.ntasks = n_leaf;
        if (a.trace) l.trace = a.trace + 4 * (size_t)offset;
        void *p[] = {&l};
        const void *kern = forward ? reinterpret_cast<const void *>(&leaf_forward<F, kNarrowTile>)
                                   : reinterpret_cast<const void *>(&leaf_backward<F, kNarrowTile>);
        return hipLaunchKernel(kern, dim3((n_leaf + kLeafWaves - 1) / kLeafWaves), dim3(kTaskThreads), p, 0, stream);
    };
    auto main = [&](int offset) -> hipError_t {
        BlockPassArgs<F> b = a;
        b.task = a.task + (size_t)offset * kDevTaskInts;
        b.ntasks = n_main;
        if (a.trace) b.trace = a.trace + 4 * (size_t)offset;
        return launch_narrow<F>(forward, b, tile, lds, std::min(grid, n_main), stream);
    };
    hipError_t e = hipSuccess;
    if (forward) {
        if (n_leaf > 0) e = leaf(0);
        if (e == hipSuccess && n_main > 0) e = main(n_leaf);
    } else {
        if (n_main > 0) e = main(0);
        if (e == hipSuccess && n_leaf > 0) e = leaf(n_main);
    }
    return e;
}

}  // namespace

size_t block_pass_lds(int float_bytes, int need_rows, int tile) {
    if (is_panel_tile(tile)) return (size_t)panel_lds_bytes();
    const int rows = std::max(std::min(need_rows, block_chunk(float_bytes, tile)), kTaskThreads);
    return (size_t)rows * tile * (size_t)float_bytes;
}

const void *block_pass_kernel(int float_bytes, bool forward, int tile) {
    if (is_panel_tile(tile)) return block_panel_kernel(float_bytes, forward, tile / 16);
    if (float_bytes == 4)
        return forward ? reinterpret_cast<const void *>(&block_forward<float, kNarrowTile>)
                       : reinterpret_cast<const void *>(&block_backward<float, kNarrowTile>);
    return forward ? reinterpret_cast<const void *>(&block_forward<double, kNarrowTile>)
                   : reinterpret_cast<const void *>(&block_backward<double, kNarrowTile>);
}

hipError_t launch_block_pass(int float_bytes, bool forward, const BlockPassBuffers &buf, int grid,
                             QueueState &queues, hipStream_t stream) {
    return float_bytes == 4 ? launch_impl<float>(forward, buf, grid, queues, stream)
                            : launch_impl<double>(forward, buf, grid, queues, stream);
}

}  // namespace cholespy_amd
