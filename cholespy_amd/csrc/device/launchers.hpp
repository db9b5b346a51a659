// Host-side launchers for the gfx950 kernels (one definition per .hip file).
#pragma once
#include <cstdint>
#include <string>
#include <vector>
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace cholespy_amd {

struct QueueArgs;   // common.hpp

// Device counters of one kernel's ticket queues (see common.hpp): kMaxQueues
// counters 16 words apart, then a done counter; zero between launches.
struct QueueState {
    unsigned int *ctr = nullptr;
    int nq = 16;
    uint32_t spin_limit = 1u << 23;   // poll budget of each bounded wait (common.hpp kSpinLimit)
    int *abort = nullptr;             // device copy of the handle's abort word
    void fill(QueueArgs &qa) const;
};
constexpr size_t kQueueWords = 16 * 16 + 16;   // counters + done word (padded)

// Threads per workgroup of the persistent row kernels (4 wavefronts).
constexpr int kRowBlock = 256;
// Rows per ticket of the row kernels (consecutive in dependency order).
constexpr int kRowChunk = 4;

// Device-side input normalisation (normalize.hip): COO / CSR / CSC arrays in
// device memory (type as MatrixType: 0 CSC, 1 CSR, 2 COO; indices of
// index_bytes = 4 or 8, values of value_bytes = 4 or 8) -> the lower CSC
// on the host, bit-identical to the host normalisation, and for each input
// entry the output position it is summed into (-1: above the diagonal).
// An invalid input leaves `invalid` non-empty (the host path's message).
struct NormalizedInput {
    std::vector<int> colptr, rowidx;
    std::vector<double> val;
    std::vector<int64_t> entry_dst;
    std::string invalid;
};
hipError_t normalize_on_device(int n, int64_t nnz, const void *ii, const void *jj, int index_bytes, const void *x,
                               int value_bytes, int type, NormalizedInput &out);

// Largest co-resident grid (workgroups) for a persistent kernel.
int max_coresident_blocks(const void *kernel, int block, size_t lds_bytes);

// Level analysis (kernels/kernels.cu:8-60). lvl1[r] must be zero on entry;
// on exit lvl1[r] = level(r) + 1.
const void *level_analysis_kernel(bool lower);
hipError_t launch_level_analysis(bool lower, int n, const int *rowptr, const int *cols,
                                 int *lvl1, int *status, int grid, QueueState &queues,
                                 hipStream_t stream);

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
                           QueueState &queues, hipStream_t stream);

// Partitioned-inverse block passes (block_plan.hpp / trsv_blocks.hip,
// trsv_panel.hip). Up to 4 right-hand sides: vector-FMA kernels with one
// 4-column tile; more: the MFMA panel kernels (all columns in one sweep).
// block_tile returns kNarrowTile, or the right-hand-side columns (32, 64 or
// 128, i.e. 2, 4 or 8 tiles of 16) the panel kernel is instantiated for.
constexpr int kNarrowTile = 4;
inline int block_tile(int float_bytes, int k) {
    (void)float_bytes;
    if (k <= kNarrowTile) return kNarrowTile;
    return k <= 32 ? 32 : (k <= 64 ? 64 : 128);
}
inline bool is_panel_tile(int tile) { return tile > kNarrowTile; }
// Index of a tile choice in small per-handle caches (0 narrow, 1..3 panel).
inline int tile_slot(int tile) { return tile <= kNarrowTile ? 0 : (tile <= 32 ? 1 : (tile <= 64 ? 2 : 3)); }
// Rows of a vector a narrow block task stages in LDS at a time: a 32 KiB
// chunk, so four or five workgroups fit a CU's 160 KiB of LDS.
constexpr int kBlockLdsBytes = 32 * 1024;
inline int block_chunk(int float_bytes, int tile) { return kBlockLdsBytes / (tile * float_bytes); }
// Dynamic LDS of a panel launch (override: CHOLESPY_PANEL_LDS, 16-64 KiB);
// at least 16 KiB, the backward kernel's staged R-row indices. 28 KiB plus
// the forward kernel's ~10 KiB of static gather staging lets four
// workgroups share a CU's 160 KiB (32 KiB allowed three: the 64^3 k = 32
// FP32 solve 1.72 -> 1.53 ms).
constexpr int kPanelLdsBytes = 28 * 1024;
// Task records as uploaded (kDevTaskInts ints, one 64-byte line): the
// plan's kTaskInts ints (block, first, last, kind, lo, hi, sbase, split),
// then the block's first row a0, size m, first partial slot r0, |R|, the
// M_B offset (low and high 32 bits), its system (0 for a single system)
// and one unused int. A task needs one line of metadata, not a second round
// of loads indexed by its block.
constexpr int kDevTaskInts = 16;
struct BlockPassBuffers {
    int ntasks = 0, k = 0;
    // Leaf tasks of this pass (forward: the first n_leaf tasks, backward: the
    // last), run by the leaf kernels on the narrow path.
    int n_leaf = 0;
    int max_m = 0, max_rows = 0;   // largest block (m) and M_B height (m + |R|)
    int max_k = 0;                 // largest right-hand-side count of any system (tile choice)
    const int *task = nullptr;     // kDevTaskInts ints per task, in ticket order
    const int *start = nullptr, *rptr = nullptr, *rlist = nullptr;
    const int64_t *moff = nullptr;
    const void *M = nullptr;       // forward or backward layout of M_B
    const int *cptr = nullptr, *cslot = nullptr;
    const int *gslot = nullptr;    // first kGatherSlots contribution slots per row (-1 padded)
    const int *perm = nullptr;
    const void *b = nullptr;
    void *y = nullptr, *z = nullptr, *part = nullptr, *x = nullptr;
    void *spart = nullptr;         // split results of split tasks (split rows x k, sentinel-protected)
    void *rbuf = nullptr;          // gathered r_B of large blocks (n x k, sentinel-protected)
    int *status = nullptr;
    // Backward epilogue: 0 = x (solver precision) = z; 1 = x (double) = z;
    // 2 = x (double) += z (refinement update).
    int xmode = 0;
    unsigned long long *trace = nullptr;   // optional task timeline (4 words per task)
    // Batches of independent systems (nullptr for a single system).
    const int *blk_sys = nullptr, *sys_k = nullptr;
    const long long *sys_vbase = nullptr, *sys_pbase = nullptr, *sys_sbase = nullptr;
    // Leaf tasks inside the block launch (its leaf phase, ticket queues
    // leaf_queues with up to 64 queues) instead of a launch of their own.
    bool fused_leaves = false;
    QueueState *leaf_queues = nullptr;
    int tile = 0;                  // tile width of the launch (0: block_tile's rule for max_k)
};
constexpr size_t kLeafQueueWords = 64 * 16 + 16;   // leaf-phase queue counters (padded)
// Dynamic LDS of a block pass: one chunk of the vector it stages (r_B of
// up to need_rows = max m forward, [y_B; z_R] of up to max m + |R| backward),
// at least 256 rows for the per-thread sums that reuse the space.
size_t block_pass_lds(int float_bytes, int need_rows, int tile);
const void *block_pass_kernel(int float_bytes, bool forward, int tile);
hipError_t launch_block_pass(int float_bytes, bool forward, const BlockPassBuffers &buf, int grid,
                             QueueState &queues, hipStream_t stream);

// Mixed-precision refinement helpers (refine.hip).
hipError_t launch_cast_to_f32(const double *in, float *out, size_t count, hipStream_t stream);
// With norm2 (k doubles, zeroed by the caller) not null, also adds the
// residual's squared column norms into it.
hipError_t launch_residual_f64(int n, int k, const int *rowptr, const int *cols, const double *vals,
                               const double *b, const double *x, float *r32, double *norm2, hipStream_t stream);
// norm2[t] = sum_i v[i,t]^2 for a row-major (rows x k) FP64 array.
hipError_t launch_column_norms2(size_t rows, int k, const double *v, double *norm2, hipStream_t stream);
hipError_t launch_zero_f64(double *p, int count, hipStream_t stream);
// host_mapped[i] = src[i], i < count, into host memory mapped into the device.
hipError_t launch_copy_out_f64(double *host_mapped, const double *src, int count, hipStream_t stream);

// dL/d(stored entries of A) from the adjoint lam and the solution x (grad.hip).
// Entries with an index outside [0, n) get NaN and are not read through.
hipError_t launch_entry_gradient(int n, int64_t nnz, const int *rows, const int *cols, int k, const void *lam,
                                 const void *x, int vec_bytes, void *grad, int grad_bytes,
                                 hipStream_t stream);

// Hand-off microbenchmark (probe.hip): res[0..6) = one-way hand-off between
// two workgroups, relaxed and with release/acquire fences (us; -1 if a wait
// gave up), dependent agent-scope load (ns), workgroup barrier (ns), ticket
// atomic on one address under contention (ns per ticket), empty launch (us).
hipError_t probe_handoff(double *res);

// Fills n words of `bytes` (4 or 8) with the sentinel of that width.
hipError_t launch_fill_sentinel(void *p, size_t n, int bytes, hipStream_t stream);

}  // namespace cholespy_amd
