// Partitioned-inverse block passes for wide right-hand sides (k > 4) on the
// gfx950 matrix cores. The tasks, hand-offs and buffers are those of
// trsv_blocks.hip (block_common.hpp); what changes is the product: a task
// multiplies its rows (forward) or columns (backward) of M_B by all k
// right-hand sides in one sweep over M_B, as 16 x 16 output tiles
// accumulated by v_mfma_{f32,f64}_16x16x4 (A = 16 x 4 piece of M_B from
// global memory, B = 4 x 16 piece of the staged vector from LDS), instead of
// one sweep per 4-column tile with vector FMAs.
//
// Lane layout of the 16x16x4 MFMAs (checked on gfx950): lane l supplies
// A[l % 16][l / 16] and B[l / 16][l % 16]; accumulator element v holds
// D[4 (l / 16) + v][l % 16] for f32 and D[l / 16 + 4 v][l % 16] for f64.
//
// Work split: a task's output is G groups of 16 rows (forward) or 16 columns
// (backward), G <= 4, times ceil(k / 16) tiles of 16 right-hand sides. The 4
// wavefronts of the workgroup take groups first, then groups of up to two
// consecutive tiles; when there are fewer (group, tile group) pairs than
// wavefronts, the wavefronts sharing a pair split the 4-wide reduction steps
// between them and add their partial tiles through LDS at the end, so a
// wavefront's A loads are its own.
//
// The staged vector (r_B forward, [y_B ; z_R] backward) is held in LDS in
// chunks of rows with a row stride of ceil(k / 16) * 16 (+16 when that is a
// multiple of 32) elements, so the four rows one B read touches start in
// different LDS bank quarters; 4 zero rows follow each chunk because the
// MFMA steps read B four rows at a time.
#include "block_common.hpp"

#include <algorithm>
#include <cstdlib>
#include <type_traits>

namespace cholespy_amd {

namespace {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));

template <typename F> struct Mfma;
template <> struct Mfma<float> {
    using Acc = f32x4;
    static __device__ __forceinline__ Acc step(float a, float b, Acc c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    // Output row of accumulator element v in lane group q = lane / 16.
    static __device__ __forceinline__ int row(int q, int v) { return 4 * q + v; }
};
template <> struct Mfma<double> {
    using Acc = f64x4;
    static __device__ __forceinline__ Acc step(double a, double b, Acc c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int q, int v) { return q + 4 * v; }
};

// Kernels are instantiated for NT = 2, 4 or 8 tiles of 16 right-hand sides
// (the launch picks the smallest that holds the largest k), so small k keeps
// few accumulators and loads more steps ahead: kSteps(NT) MFMA steps have
// their A (and B) loads issued together.
template <int NT> constexpr int kStepsFor() { return NT <= 2 ? 8 : 4; }
// Tasks whose B comes straight from global memory (large blocks forward,
// every backward task) run a wavefront's tiles in batches of kBatchTiles,
// each batch with kBatchSteps MFMA steps' A and B loads issued together:
// a step holds B registers for the batch's tiles only, not for all NT, so
// the loads in flight do not shrink as the tile count grows, and a task
// pays ~kBatchSteps times fewer dependent load rounds than one step at a
// time.
constexpr int kBatchTiles = 2;
// FP32: 8 steps (12 for the 8-tile kernels, whose occupancy floors leave
// room); FP64: 4 (more cost the forward kernels a wave per SIMD).
template <typename F, int NT> constexpr int batch_steps() { return sizeof(F) == 4 ? (NT >= 8 ? 12 : 8) : (NT <= 2 ? 8 : 4); }
// Waves per SIMD the compiler must allow (amdgpu_waves_per_eu), chosen per
// kernel as the most that compile with no register spills on gfx950. By
// default the compiler keeps more accumulators in AGPRs and gets 2-3 waves;
// more waves put more tasks in flight where many small tasks run at once
// (the leaf levels, batches). FP32, 8 tiles: 224 -> 166 registers forward
// (3 waves), 116 backward (4). FP32, up to 4 tiles: 4 waves forward, 5
// backward. FP64 forward with 4 tiles: 3 waves (left alone the compiler
// takes AGPRs and 2); other FP64 forward kernels would spill at more waves
// and keep the default; FP64 backward with up to 4 tiles fits 4 waves.
template <typename F, int NT, bool Forward> constexpr int kMinWavesFor() {
    if (sizeof(F) == 4) return NT >= 8 ? 3 : 4;
    if (!Forward) return NT <= 4 ? 4 : 1;
    return NT == 4 ? 3 : 1;
}
// Backward tasks stage the R-row indices they read in LDS (up to this many).
constexpr int kStageRlist = 4096;
// Entry i of an int array in LDS, read through an LDS-typed pointer: a load
// through a generic pointer (what a choice between an LDS and a global array
// compiles to) is a flat load, which counts against the vector-memory
// counter too, so waiting for it would also wait for every M load in flight.
__device__ __forceinline__ int lds_int(const int *p, int i) {
    return ((const __attribute__((address_space(3))) int *)p)[i];
}

// Packed offsets of M_B entries (32-bit: one block's M_B has < 2^31 entries).
__device__ __forceinline__ int fwd_off32(int rows, int c, int rho) {
    return c * rows - (c * (c - 1)) / 2 + (rho - c);
}
__device__ __forceinline__ int bwd_off32(int m, int rho, int c) {
    return rho < m ? (rho * (rho + 1)) / 2 + c : (m * (m + 1)) / 2 + (rho - m) * m + c;
}

struct WaveWork {
    int group;                   // 16-row / 16-column group of this wavefront, -1 = idle
    int tile0, tstride, ntiles;  // right-hand-side tiles tile0 + j * tstride, j < ntiles
    int split, nsplit;           // reduction steps s with s % nsplit == split
    int partner_stride;          // wavefront distance between the splits of one pair
};

__device__ __forceinline__ WaveWork wave_work(int wave, int width, int nT) {
    const int G = (width + 15) / 16;
    const int Gp = G <= 1 ? 1 : (G <= 2 ? 2 : 4);
    const int W = 4 / Gp;   // wavefronts per group
    // With fewer tile groups (up to kBatchTiles consecutive tiles) than
    // wavefronts, each wavefront takes one group and the wavefronts left over
    // split the group's reduction, so two wavefronts of a row group load the
    // same A (M_B) entries only when there are several tile groups (one tile
    // per wavefront with split pairs instead: the 64^3 k = 32 FP32 solve
    // 1.49 ms against 1.18). Otherwise the tiles are dealt round.
    const int tpw = nT < kBatchTiles ? max(nT, 1) : kBatchTiles;
    const int TG = (nT + tpw - 1) / tpw;
    WaveWork w;
    w.group = wave / W;
    const int u = wave % W;
    if (TG >= W) {   // each wavefront takes tiles u, u + W, ... (in batches)
        w.tile0 = u;
        w.tstride = W;
        w.ntiles = u < nT ? (nT - u + W - 1) / W : 0;
        w.split = 0;
        w.nsplit = 1;
        w.partner_stride = 1;
    } else {
        const int TGp = TG <= 1 ? 1 : (TG <= 2 ? 2 : 4);
        const int tg = u % TGp;
        w.tile0 = tg * tpw;
        w.tstride = 1;
        w.ntiles = tg < TG ? min(tpw, nT - w.tile0) : 0;
        w.split = u / TGp;
        w.nsplit = W / TGp;
        w.partner_stride = TGp;
    }
    if (w.group >= G) {
        w.group = -1;
        w.ntiles = 0;
    }
    return w;
}

// LDS row stride of a staged vector with k columns and the rows per chunk.
__device__ __forceinline__ int panel_stride(int k) {
    const int kp = (k + 15) / 16 * 16;
    return kp % 32 == 16 ? kp : kp + 16;
}
template <typename F>
__device__ __forceinline__ int panel_chunk(int lds_bytes, int ks) {
    return (lds_bytes / (ks * (int)sizeof(F)) - 4) & ~3;
}


// Staged gather of r_B rows [r0, r1) over all k columns (wide right-hand
// sides): per round of up to kStageRows rows, one thread per row loads its
// perm entry and up to kStageSlots contribution slots into LDS; then every
// thread issues the loads of kStageElems (row, column) elements at once
// (the b entry and the published partials, coalesced along the columns),
// so a round costs three dependent memory latencies however wide k is.
// to_rbuf: publish the sums to rb; otherwise store them to dst[i * stride + t].
constexpr int kStageRows = 128;
constexpr int kStageSlots = 16;
constexpr int kSlotGroup = 4;   // partial loads issued together per element
// Elements per thread per load round: 4 (f32) or 2 (f64) entries, each with
// its kStageSlots partials, bounds the registers this phase adds.
template <typename F> constexpr int stage_elems() { return sizeof(F) == 4 ? 4 : 2; }

struct GatherStage {
    int perm[kStageRows];
    int cnt[kStageRows];
    int first[kStageRows];
    int slot[kStageRows][kStageSlots];
};

// The element phase of gather_wide for one staged round of nr rows at q0,
// on 16-byte vectors: a thread takes kVecPer vectors (kVE consecutive
// columns of one row) at a time and loads the b vector and kVecSlots
// contribution vectors of each together: 4 (FP32) or 2 (FP64) elements per
// load instruction and per address instead of one. A lane of a vector read
// as the sentinel is waited on alone (ready_value), as in the scalar phase.
// (2 vectors a thread spill registers in the forward panel kernels.)
constexpr int kVecPer = 1;
constexpr int kVecSlots = 4;
template <typename F>
__device__ __forceinline__ void gather_wide_vec(const BlockPassArgs<F> &a, const GatherStage &st, const F *bv,
                                                const F *pv, F *rb, F *dst, int a0, int q0, int nr, int k, int stride,
                                                bool to_rbuf, int r0) {
    using U = typename Bits<F>::U;
    constexpr int kVE = 16 / (int)sizeof(F);
    typedef U V __attribute__((ext_vector_type(kVE)));
    const int nv = k / kVE;
    const int total = nr * nv;
    for (int base = threadIdx.x; base < total; base += kTaskThreads * kVecPer) {
        int ii[kVecPer], col[kVecPer], ncnt = 0;
        F sum[kVecPer][kVE];
#pragma unroll
        for (int e = 0; e < kVecPer; ++e) {
            const int idx = base + e * kTaskThreads;
            const bool ok = idx < total;
            ii[e] = ok ? idx / nv : -1;
            col[e] = ok ? (idx - ii[e] * nv) * kVE : 0;
            const V vb = ok ? *reinterpret_cast<const V *>(bv + (long long)st.perm[ii[e]] * k + col[e]) : V{};
#pragma unroll
            for (int c = 0; c < kVE; ++c) sum[e][c] = from_bits<F>(vb[c]);
            if (ok) ncnt = max(ncnt, min(st.cnt[ii[e]], kStageSlots));
        }
        for (int g = 0; g < ncnt; g += kVecSlots) {
            V us[kVecPer][kVecSlots];
#pragma unroll
            for (int e = 0; e < kVecPer; ++e)
#pragma unroll
                for (int j = 0; j < kVecSlots; ++j) {
                    const int sl = ii[e] >= 0 && g + j < kStageSlots ? st.slot[ii[e]][g + j] : -1;
                    us[e][j] = sl >= 0 ? *reinterpret_cast<const V *>(pv + (long long)sl * k + col[e]) : V{};
                }
#pragma unroll
            for (int e = 0; e < kVecPer; ++e)
#pragma unroll
                for (int j = 0; j < kVecSlots; ++j) {
                    const int sl = ii[e] >= 0 && g + j < kStageSlots ? st.slot[ii[e]][g + j] : -1;
                    if (sl >= 0) {
                        const F *p = pv + (long long)sl * k + col[e];
#pragma unroll
                        for (int c = 0; c < kVE; ++c) sum[e][c] += ready_value(p + c, us[e][j][c], a.status, a.qa);
                    }
                }
        }
#pragma unroll
        for (int e = 0; e < kVecPer; ++e) {
            const int i = ii[e];
            if (i < 0) break;
            for (int j = kStageSlots; j < st.cnt[i]; ++j) {   // rows with very many contributions
                const F *p = &pv[(long long)a.cslot[st.first[i] + j] * k + col[e]];
#pragma unroll
                for (int c = 0; c < kVE; ++c) sum[e][c] += ready_value(p + c, peek_bits(p + c), a.status, a.qa);
            }
#pragma unroll
            for (int c = 0; c < kVE; ++c) {
                if (to_rbuf)
                    publish(&rb[(long long)(a0 + q0 + i) * k + col[e] + c], sum[e][c]);
                else
                    dst[(q0 - r0 + i) * stride + col[e] + c] = sum[e][c];
            }
        }
    }
}

template <typename F>
__device__ __forceinline__ void gather_wide(const BlockPassArgs<F> &a, GatherStage &st, const F *bv, const F *pv,
                                            F *rb, F *dst, int a0, int r0, int r1, int k, int stride,
                                            bool to_rbuf) {
    using U = typename Bits<F>::U;
    constexpr int kStageElems = stage_elems<F>();
    const int tid = threadIdx.x;
    // 16-byte vectors along the columns when k and the vector bases allow.
    constexpr int kVE = 16 / (int)sizeof(F);
    const bool vec_ok = k % kVE == 0 && ((reinterpret_cast<uintptr_t>(bv) | reinterpret_cast<uintptr_t>(pv)) & 15) == 0;
    for (int q0 = r0; q0 < r1; q0 += kStageRows) {
        const int nr = min(kStageRows, r1 - q0);
        if (tid < nr) {
            const int row = a0 + q0 + tid;
            const int p0 = a.cptr[row], p1 = a.cptr[row + 1];
            st.perm[tid] = a.perm[row];
            st.cnt[tid] = p1 - p0;
            st.first[tid] = p0;
            // The row's first slots from the fixed-stride table (-1 padded),
            // loaded together with cptr instead of after it.
            static_assert(kStageSlots == kGatherSlots, "a staged row is a row of gslot");
            const int4 *g = reinterpret_cast<const int4 *>(a.gslot + (size_t)row * kGatherSlots);
#pragma unroll
            for (int h = 0; h < kStageSlots / 4; ++h) {
                const int4 w = g[h];
                st.slot[tid][4 * h] = w.x;
                st.slot[tid][4 * h + 1] = w.y;
                st.slot[tid][4 * h + 2] = w.z;
                st.slot[tid][4 * h + 3] = w.w;
            }
        }
        __syncthreads();
        if (vec_ok) {
            gather_wide_vec(a, st, bv, pv, rb, dst, a0, q0, nr, k, stride, to_rbuf, r0);
            __syncthreads();
            continue;
        }
        const int total = nr * k;
        for (int base = tid; base < total; base += kTaskThreads * kStageElems) {
            U ub[kStageElems];
            int ii[kStageElems], tt[kStageElems];
            F v[kStageElems];
            int ncnt = 0;
#pragma unroll
            for (int e = 0; e < kStageElems; ++e) {
                const int idx = base + e * kTaskThreads;
                const bool ok = idx < total;
                ii[e] = ok ? idx / k : 0;
                tt[e] = ok ? idx - ii[e] * k : 0;
                ub[e] = ok ? as_bits(bv[(long long)st.perm[ii[e]] * k + tt[e]]) : U(0);
                if (ok) ncnt = max(ncnt, min(st.cnt[ii[e]], kStageSlots));
            }
#pragma unroll
            for (int e = 0; e < kStageElems; ++e) v[e] = from_bits<F>(ub[e]);
            // Staged partials, kSlotGroup per element in flight at a time.
            for (int g = 0; g < ncnt; g += kSlotGroup) {
                U us[kStageElems][kSlotGroup];
#pragma unroll
                for (int e = 0; e < kStageElems; ++e)
#pragma unroll
                    for (int j = 0; j < kSlotGroup; ++j) {
                        const int idx = base + e * kTaskThreads;
                        const int sl = idx < total ? st.slot[ii[e]][g + j] : -1;
                        us[e][j] = sl >= 0 ? peek_bits(&pv[(long long)sl * k + tt[e]]) : U(0);
                    }
#pragma unroll
                for (int e = 0; e < kStageElems; ++e)
#pragma unroll
                    for (int j = 0; j < kSlotGroup; ++j) {
                        const int idx = base + e * kTaskThreads;
                        const int sl = idx < total ? st.slot[ii[e]][g + j] : -1;
                        if (sl >= 0) v[e] += ready_value(&pv[(long long)sl * k + tt[e]], us[e][j], a.status, a.qa);
                    }
            }
#pragma unroll
            for (int e = 0; e < kStageElems; ++e) {
                const int idx = base + e * kTaskThreads;
                if (idx >= total) break;
                const int i = ii[e], t = tt[e];
                for (int j = kStageSlots; j < st.cnt[i]; ++j) {   // rows with very many contributions
                    const F *p = &pv[(long long)a.cslot[st.first[i] + j] * k + t];
                    v[e] += ready_value(p, peek_bits(p), a.status, a.qa);
                }
                if (to_rbuf)
                    publish(&rb[(long long)(a0 + q0 + i) * k + t], v[e]);
                else
                    dst[(q0 - r0 + i) * stride + t] = v[e];
            }
        }
        __syncthreads();
    }
}


// Zero rows [len, len + 4) of a staged chunk.
template <typename F>
__device__ __forceinline__ void zero_tail(F *v, int len, int ks) {
    for (int idx = threadIdx.x; idx < 4 * ks; idx += kTaskThreads) v[len * ks + idx] = F(0);
}

// Adds the partial tiles of the other splits of this wavefront's tile group
// (LDS `red`, 64 x 4 values per wavefront and tile; a wavefront that splits
// a reduction holds at most kBatchTiles tiles). Every thread of the
// workgroup calls it.
template <typename F, typename Acc, int N>
__device__ __forceinline__ void reduce_splits(F *red, const WaveWork &w, int wave, int lane, Acc (&acc)[N]) {
    constexpr int kTiles = N < kBatchTiles ? N : kBatchTiles;
    if (w.ntiles > 0 && w.split > 0)
#pragma unroll
        for (int j = 0; j < kTiles; ++j)
            if (j < w.ntiles)
                for (int v = 0; v < 4; ++v) red[((j * 4 + wave) * 64 + lane) * 4 + v] = acc[j][v];
    __syncthreads();
    if (w.ntiles > 0 && w.split == 0)
        for (int s = 1; s < w.nsplit; ++s) {
            const int pw = wave + s * w.partner_stride;
#pragma unroll
            for (int j = 0; j < kTiles; ++j)
                if (j < w.ntiles)
                    for (int v = 0; v < 4; ++v) acc[j][v] += red[((j * 4 + pw) * 64 + lane) * 4 + v];
        }
}

template <typename F, int NT>
__global__ __launch_bounds__(kTaskThreads) __attribute__((amdgpu_waves_per_eu(kMinWavesFor<F, NT, true>())))
void block_forward_panel(BlockPassArgs<F> a) {
    using Acc = typename Mfma<F>::Acc;
    constexpr int kMaxTiles = NT;
    constexpr int kSteps = kStepsFor<NT>();
    constexpr int kBatchSteps = batch_steps<F, NT>();
    extern __shared__ __align__(16) unsigned char smem[];
    F *rv = reinterpret_cast<F *>(smem);
    __shared__ int s_next;
    __shared__ GatherStage stage;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int qid = blockIdx.x % a.qa.nq;
    int tk = wg_first_ticket(a.qa, qid, a.ntasks, &s_next);
    while (tk >= 0) {
        if (a.trace && tid == 0) a.trace[4 * (size_t)tk] = wall_clock64();
        const Task task = load_task(a, tk);
        const int first = task.first, last = task.last, kind = task.kind;
        const int a0 = task.a0, m = task.m;
        const int rows = m + task.R;
        const F *M = a.M + task.moff;
        const VecMap vm = vec_map(a, task);
        const int k = vm.k;
        F *yv = a.y + vm.vbase, *zv = a.z + vm.vbase, *pv = a.part + vm.pbase, *rb = a.rbuf + vm.vbase;
        F *sp = a.spart + vm.sbase;
        const F *bv = a.b + vm.vbase;
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
        const int nsp = split_count(task.nsplit), sidx = split_index(task.nsplit);
        const int srow = task.sbase + sidx * (last - first) - first;
        if (kind == 1) {   // gather-only task: rows [first, last) of r_B to rbuf
            gather_wide(a, stage, bv, pv, rb, (F *)nullptr, a0, first, last, k, k, true);
            if (a.trace && tid == 0) a.trace[4 * (size_t)tk + 1] = wall_clock64();
        } else if (k > 0) {
            const int nT = (k + 15) / 16;
            const int ks = panel_stride(k);
            const int chunk = panel_chunk<F>(a.lds_bytes, ks);
            const WaveWork w = wave_work(wave, last - first, nT);
            const int g0 = first + 16 * max(w.group, 0);
            const int row = g0 + (lane & 15);   // A row of this lane
            const int kk = lane >> 4;           // A column / B row within a step
            // Columns the group uses: rows < m use c <= row, rows >= m all c < m.
            const int glast = min(g0 + 16, last) - 1;
            const int cmax = glast < m ? glast + 1 : m;
            const int rowlim = row < m ? row + 1 : m;
            // Output of tile j (accumulator v of lane group kk: row g0 + row(kk, v)).
            auto out_tile = [&](const Acc &acc_j, int j) {
                const int t = 16 * (w.tile0 + j * w.tstride) + (lane & 15);
                if (t >= k) return;
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int rho = g0 + Mfma<F>::row(kk, v);
                    if (rho >= last) break;
                    const F s = acc_j[v];
                    if (kind == 3)
                        publish(&sp[(long long)(srow + rho) * k + t], s);
                    else
                        finish(rho, t, s);
                }
            };
            if (kind == 0) {
                Acc acc[kMaxTiles];
#pragma unroll
                for (int j = 0; j < kMaxTiles; ++j) acc[j] = Acc{};
                // Small block: r_B summed from the partials into LDS, in chunks.
                for (int c0 = 0; c0 < m; c0 += chunk) {
                    const int c1 = min(m, c0 + chunk);
                    gather_wide(a, stage, bv, pv, rb, rv, a0, c0, c1, k, ks, false);
                    zero_tail(rv, c1 - c0, ks);
                    __syncthreads();
                    if (a.trace && tid == 0 && c0 == 0) a.trace[4 * (size_t)tk + 1] = wall_clock64();
                    if (w.ntiles > 0) {
                        const int lim = min(rowlim, c1);
                        const int nsteps = (min(c1, cmax) - c0 + 3) / 4;
                        const F *bl = rv + kk * ks + (lane & 15) + 16 * w.tile0;
                        for (int s0 = w.split; s0 < nsteps; s0 += w.nsplit * kSteps) {
                            F av[kSteps];
#pragma unroll
                            for (int q = 0; q < kSteps; ++q) {
                                const int s = s0 + q * w.nsplit;
                                const int col = c0 + 4 * s + kk;
                                av[q] = (s < nsteps && row < last && col < lim) ? M[fwd_off32(rows, col, row)] : F(0);
                            }
#pragma unroll
                            for (int q = 0; q < kSteps; ++q) {
                                const int s = s0 + q * w.nsplit;
                                if (s < nsteps) {
                                    const F *bp = bl + 4 * s * ks;
#pragma unroll
                                    for (int j = 0; j < kMaxTiles; ++j)
                                        if (j < w.ntiles) acc[j] = Mfma<F>::step(av[q], bp[16 * j * w.tstride], acc[j]);
                                }
                            }
                        }
                    }
                    __syncthreads();
                }
                if (w.nsplit > 1) reduce_splits(rv, w, wave, lane, acc);
                if (w.ntiles > 0 && w.split == 0) {
#pragma unroll
                    for (int j = 0; j < kMaxTiles; ++j)
                        if (j < w.ntiles) out_tile(acc[j], j);
                }
            } else {
                // Large block: B = r_B read straight from rbuf (published by the
                // block's gather tasks), no staging; columns [lo, hi) only for a
                // split task. Tiles in batches (kBatchTiles, kBatchSteps).
                using U = typename Bits<F>::U;
                const int clo = kind == 3 ? task.lo : 0;
                const int chi = kind == 3 ? task.hi : m;
                const int nsteps = (min(chi, cmax) - clo + 3) / 4;
                const F *rb0 = rb + (long long)(a0 + clo) * k + (lane & 15) + 16 * w.tile0;
                // Waves that split a reduction (at most kBatchTiles tiles
                // each: one batch) meet in reduce_splits once, idle ones
                // included.
                const int nbt = NT <= kBatchTiles ? 1 : max(w.ntiles, w.nsplit > 1 ? 1 : 0);
                for (int tb = 0; tb < nbt; tb += kBatchTiles) {
                    Acc acc[kBatchTiles];
#pragma unroll
                    for (int jj = 0; jj < kBatchTiles; ++jj) acc[jj] = Acc{};
                    for (int s0 = w.split; w.ntiles > 0 && s0 < nsteps; s0 += w.nsplit * kBatchSteps) {
                        F av[kBatchSteps];
                        U bu[kBatchSteps][kBatchTiles];
#pragma unroll
                        for (int q = 0; q < kBatchSteps; ++q) {
                            const int s = s0 + q * w.nsplit;
                            const int col = clo + 4 * s + kk;
                            av[q] = (s < nsteps && row < last && col < rowlim && col < chi) ? M[fwd_off32(rows, col, row)]
                                                                                           : F(0);
                            const bool bok = s < nsteps && col < chi;
#pragma unroll
                            for (int jj = 0; jj < kBatchTiles; ++jj) {
                                const int j = tb + jj;
                                const int t = 16 * (w.tile0 + j * w.tstride) + (lane & 15);
                                bu[q][jj] = (j < w.ntiles && bok && t < k)
                                                ? peek_bits(rb0 + (long long)(col - clo) * k + 16 * j * w.tstride) : U(0);
                            }
                        }
#pragma unroll
                        for (int q = 0; q < kBatchSteps; ++q) {
                            const int s = s0 + q * w.nsplit;
                            if (s < nsteps) {
                                const int col = 4 * s + kk;   // relative to clo
#pragma unroll
                                for (int jj = 0; jj < kBatchTiles; ++jj) {
                                    const int j = tb + jj;
                                    if (j < w.ntiles)
                                        acc[jj] = Mfma<F>::step(
                                            av[q],
                                            ready_value(rb0 + (long long)col * k + 16 * j * w.tstride, bu[q][jj], a.status, a.qa),
                                            acc[jj]);
                                }
                            }
                        }
                    }
                    if (w.nsplit > 1) reduce_splits(rv, w, wave, lane, acc);
                    if (w.ntiles > 0 && w.split == 0) {
#pragma unroll
                        for (int jj = 0; jj < kBatchTiles; ++jj)
                            if (tb + jj < w.ntiles) out_tile(acc[jj], tb + jj);
                    }
                }
            }
            if (kind == 3 && combines_splits(sidx, nsp))
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
    if (tid == 0) finish_queues(a.qa);
}

template <typename F, int NT>
__global__ __launch_bounds__(kTaskThreads) __attribute__((amdgpu_waves_per_eu(kMinWavesFor<F, NT, false>())))
void block_backward_panel(BlockPassArgs<F> a) {
    using Acc = typename Mfma<F>::Acc;
    constexpr int kMaxTiles = NT;
    constexpr int kSteps = kStepsFor<NT>();
    constexpr int kBatchSteps = batch_steps<F, NT>();
    extern __shared__ __align__(16) unsigned char smem[];
    F *vs = reinterpret_cast<F *>(smem);
    __shared__ int s_next;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int qid = blockIdx.x % a.qa.nq;
    int tk = wg_first_ticket(a.qa, qid, a.ntasks, &s_next);
    while (tk >= 0) {
        if (a.trace && tid == 0) a.trace[4 * (size_t)tk] = wall_clock64();
        const Task task = load_task(a, tk);
        const int first = task.first, last = task.last, kind = task.kind;
        const int a0 = task.a0, m = task.m, r0 = task.r0, R = task.R;
        const int rows = m + R;
        const F *M = a.M + task.moff;
        const VecMap vm = vec_map(a, task);
        const int k = vm.k;
        F *yv = a.y + vm.vbase, *zv = a.z + vm.vbase, *pv = a.part + vm.pbase;
        F *sp = a.spart + vm.sbase;
        // Finishes column cc, right-hand side t: z published, x written.
        auto finish = [&](int cc, int t, F s) {
            publish(&zv[(long long)(a0 + cc) * k + t], s);
            const long long o = vm.vbase + (long long)a.perm[a0 + cc] * k + t;
            if (a.xmode == 0) {
                static_cast<F *>(a.x)[o] = s;
            } else {
                double *x64 = static_cast<double *>(a.x);
                x64[o] = (a.xmode == 2 ? x64[o] : 0.0) + (double)s;
            }
        };
        const int nsp = split_count(task.nsplit), sidx = split_index(task.nsplit);
        const int srow = task.sbase + sidx * (last - first) - first;
        if (k > 0) {
            const int nT = (k + 15) / 16;
            const WaveWork w = wave_work(wave, last - first, nT);
            const int cg0 = first + 16 * max(w.group, 0);
            const int c = cg0 + (lane & 15);   // A row (= column of M_B) of this lane
            const int kk = lane >> 4;          // A column / B row within a step
            // The R rows the task reads, [rb0, rb1), have their indices staged
            // in LDS (one load each, all in parallel) when they fit.
            const int he = kind == 1 ? task.hi : rows;
            const int rb0 = max(m, kind == 1 ? max(first, task.lo) : first);
            const bool staged = he - rb0 <= kStageRlist;
            int *srl = reinterpret_cast<int *>(vs);
            __syncthreads();   // the previous task's reads of vs are done
            if (staged)
                for (int r = rb0 + tid; r < he; r += kTaskThreads) srl[r - rb0] = a.rlist[r0 + r - m];
            __syncthreads();
            if (a.trace && tid == 0) a.trace[4 * (size_t)tk + 1] = wall_clock64();
            // B = v[rho] = y_B[rho] (rho < m, complete since the forward pass)
            // or z of R row rlist[rho - m] (published, waited on), read
            // straight from global memory. Rows below the group's first
            // column are zero in its columns. Tiles in batches (kBatchTiles,
            // kBatchSteps); waves that split a reduction (at most
            // kBatchTiles tiles each: one batch) meet in reduce_splits once,
            // idle ones included.
            using U = typename Bits<F>::U;
            const int hs = kind == 1 ? max(cg0, task.lo) : cg0;   // rows [hs, he)
            const int nsteps = (he - hs + 3) / 4;
            const int toff = (lane & 15) + 16 * w.tile0;
            const int nbt = NT <= kBatchTiles ? 1 : max(w.ntiles, w.nsplit > 1 ? 1 : 0);
            for (int tb = 0; tb < nbt; tb += kBatchTiles) {
                Acc acc[kBatchTiles];
#pragma unroll
                for (int jj = 0; jj < kBatchTiles; ++jj) acc[jj] = Acc{};
                // The batches, in two copies: R-row indices from LDS (staged) or
                // from global memory. One copy choosing per row would join an
                // LDS and a global load into one register, and the wait for it
                // would drain every M load in flight.
                auto batches = [&](auto staged_c) {
                constexpr bool kStaged = decltype(staged_c)::value;
                for (int s0 = w.split; w.ntiles > 0 && s0 < nsteps; s0 += w.nsplit * kBatchSteps) {
                    F av[kBatchSteps];
                    const F *bp[kBatchSteps];
                    bool pubr[kBatchSteps];
#pragma unroll
                    for (int q = 0; q < kBatchSteps; ++q) {
                        const int s = s0 + q * w.nsplit;
                        const int rho = hs + 4 * s + kk;
                        av[q] = (s < nsteps && c < last && rho < he && rho >= c) ? M[bwd_off32(m, rho, c)] : F(0);
                        pubr[q] = rho >= m;
                        bp[q] = nullptr;
                        if (s < nsteps && rho < he)
                            bp[q] = (rho < m ? yv + (long long)(a0 + rho) * k
                                             : zv + (long long)(kStaged ? lds_int(srl, rho - rb0) : a.rlist[r0 + rho - m]) * k) +
                                    toff;
                    }
                    U bu[kBatchSteps][kBatchTiles];
#pragma unroll
                    for (int q = 0; q < kBatchSteps; ++q)
#pragma unroll
                        for (int jj = 0; jj < kBatchTiles; ++jj) {
                            const int j = tb + jj;
                            const int t = toff + 16 * j * w.tstride;
                            const bool ok = j < w.ntiles && bp[q] && t < k;
                            // y rows never hold the sentinel: one load form for both.
                            bu[q][jj] = !ok ? U(0) : peek_bits(bp[q] + 16 * j * w.tstride);
                        }
#pragma unroll
                    for (int q = 0; q < kBatchSteps; ++q) {
                        const int s = s0 + q * w.nsplit;
                        if (s < nsteps) {
#pragma unroll
                            for (int jj = 0; jj < kBatchTiles; ++jj) {
                                const int j = tb + jj;
                                if (j < w.ntiles) {
                                    const F b = pubr[q] && bp[q]
                                                    ? ready_value(bp[q] + 16 * j * w.tstride, bu[q][jj], a.status, a.qa)
                                                    : from_bits<F>(bu[q][jj]);
                                    acc[jj] = Mfma<F>::step(av[q], b, acc[jj]);
                                }
                            }
                        }
                    }
                }
                };
                if (staged)
                    batches(std::true_type{});
                else
                    batches(std::false_type{});
                if (w.nsplit > 1) {
                    __syncthreads();   // the staged R-row indices are no longer read
                    reduce_splits(vs, w, wave, lane, acc);
                }
                if (w.ntiles > 0 && w.split == 0) {
#pragma unroll
                    for (int jj = 0; jj < kBatchTiles; ++jj) {
                        const int j = tb + jj;
                        if (j >= w.ntiles) break;
                        const int t = 16 * (w.tile0 + j * w.tstride) + (lane & 15);
                        if (t >= k) continue;
#pragma unroll
                        for (int v = 0; v < 4; ++v) {
                            const int cc = cg0 + Mfma<F>::row(kk, v);
                            if (cc >= last) break;
                            const F s = acc[jj][v];
                            if (kind == 1)
                                publish(&sp[(long long)(srow + cc) * k + t], s);
                            else
                                finish(cc, t, s);
                        }
                    }
                }
            }
            if (kind == 1 && combines_splits(sidx, nsp))
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
    if (tid == 0) finish_queues(a.qa);
}

}  // namespace

int panel_lds_bytes() {
    static const int bytes = [] {
        int v = kPanelLdsBytes;
        if (const char *e = std::getenv("CHOLESPY_PANEL_LDS")) v = std::atoi(e);
        return std::min(std::max(v, 16 * 1024), 64 * 1024);
    }();
    return bytes;
}

const void *block_panel_kernel(int float_bytes, bool forward, int nt) {
#define CHOLESPY_PANEL(F, NT)                                                                 \
    (forward ? reinterpret_cast<const void *>(&block_forward_panel<F, NT>)                   \
             : reinterpret_cast<const void *>(&block_backward_panel<F, NT>))
    if (float_bytes == 4) return nt <= 2 ? CHOLESPY_PANEL(float, 2) : nt <= 4 ? CHOLESPY_PANEL(float, 4) : CHOLESPY_PANEL(float, 8);
    return nt <= 2 ? CHOLESPY_PANEL(double, 2) : nt <= 4 ? CHOLESPY_PANEL(double, 4) : CHOLESPY_PANEL(double, 8);
#undef CHOLESPY_PANEL
}

template <typename F>
static hipError_t launch_panel_impl(bool forward, const BlockPassArgs<F> &args, int nt, int grid, hipStream_t stream) {
    BlockPassArgs<F> a = args;
    a.lds_bytes = panel_lds_bytes();
    void *p[] = {&a};
    return hipLaunchKernel(block_panel_kernel(sizeof(F), forward, nt), dim3(grid), dim3(kTaskThreads), p,
                           (size_t)a.lds_bytes, stream);
}

hipError_t launch_panel_f32(bool forward, const BlockPassArgs<float> &a, int nt, int grid, hipStream_t stream) {
    return launch_panel_impl<float>(forward, a, nt, grid, stream);
}
hipError_t launch_panel_f64(bool forward, const BlockPassArgs<double> &a, int nt, int grid, hipStream_t stream) {
    return launch_panel_impl<double>(forward, a, nt, grid, stream);
}

}  // namespace cholespy_amd
