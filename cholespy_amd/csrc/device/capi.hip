// C ABI of the gfx950 solve phase (declared in include/cholespy_amd.h).
//
// The handle owns the device state the reference keeps in CholeskySolver
// (src/cholesky_solver.h:60-98): perm, the row-compressed L and L', the two
// level orders, and the per-solve work buffers. Kernels are linked into this
// library, so there is no module loading (src/cuda_driver.cpp:127-133).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <string>
#include <atomic>
#include <mutex>
#include <thread>
#include <sys/stat.h>
#include <vector>

#include "../../../include/cholespy_amd.h"
#include "block_plan.hpp"
#include "common.hpp"
#include "launchers.hpp"

using namespace cholespy_amd;

// Work memory of cholespy_solver_solve_noalloc: one array per GPU, in the
// library's code object, so that entry point needs no allocation. Its
// device address per GPU comes from hipGetSymbolAddress (the name itself
// is not a device address on the host).
constexpr size_t kScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

// Ticket counters of a handle: six queue sets of kQueueWords (analysis,
// row passes, block passes) and two leaf-phase sets of kLeafQueueWords.
// The device copy of the abort word follows them (its own 64-byte line).
constexpr size_t kAllQueueWords = 6 * kQueueWords + 2 * kLeafQueueWords + 16;
// Each set holds its queues' counters kQueueStride words apart and then the
// done counter (finish_queues), so the set sizes follow the queue limits.
static_assert(kQueueWords >= (size_t)(kMaxQueues + 1) * kQueueStride, "block queue counters overflow their set");
static_assert(kLeafQueueWords >= (size_t)(kMaxLeafQueues + 1) * kQueueStride, "leaf queue counters overflow their set");

thread_local std::string g_error;
bool g_initialised = false;
int g_device = -1;

// Per device: the scratch array's device address (null until the first
// no-allocation solve on that device) and the layout its work buffers hold
// in the all-sentinel state (0 = unknown: refill before use).
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char *g_scratch_ptr[kMaxDevices] = {};
uint64_t g_scratch_layout[kMaxDevices] = {};
std::atomic<uint64_t> g_next_uid{1};

int fail(int code, const std::string &msg) {
    g_error = msg;
    return code;
}

int hip_fail(hipError_t e, const char *what) {
    return fail(CHOLESPY_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define HIP_TRY(expr)                                   \
    do {                                                \
        hipError_t e_ = (expr);                         \
        if (e_ != hipSuccess) return hip_fail(e_, #expr); \
    } while (0)

template <typename T>
hipError_t dev_alloc(T **p, size_t count) {
    *p = nullptr;
    if (count == 0) count = 1;
    return hipMalloc(reinterpret_cast<void **>(p), count * sizeof(T));
}

template <typename T>
void dev_free(T *&p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

template <typename T>
hipError_t upload(T **dst, const T *src, size_t count) {
    hipError_t e = dev_alloc(dst, count);
    if (e != hipSuccess || count == 0) return e;
    return hipMemcpy(*dst, src, count * sizeof(T), hipMemcpyHostToDevice);
}

// Device copy of a BlockPlan (block_plan.hpp) in the solver's precision.
struct DeviceBlockPlan {
    int nblocks = 0, nslots = 0, nsplit_rows = 0, max_m = 0, max_rows = 0;
    int n_fwd_tasks = 0, n_bwd_tasks = 0;
    int n_leaf_fwd = 0, n_leaf_bwd = 0;   // leaf tasks: forward prefix, backward suffix
    int n_fwd_tasks_wide = 0, n_bwd_tasks_wide = 0;   // task lists of the panel (wide k) kernels
    int levels_fwd = 0, levels_bwd = 0;
    int64_t m_entries = 0;
    int *start = nullptr, *rptr = nullptr, *rlist = nullptr;
    int64_t *moff = nullptr;
    void *mfwd = nullptr, *mbwd = nullptr;
    int *cptr = nullptr, *cslot = nullptr, *gslot = nullptr;
    int *ftask = nullptr, *btask = nullptr;
    int *ftask_wide = nullptr, *btask_wide = nullptr;

    void release() {
        dev_free(start);
        dev_free(rptr);
        dev_free(rlist);
        dev_free(moff);
        dev_free(mfwd);
        dev_free(mbwd);
        dev_free(cptr);
        dev_free(cslot);
        dev_free(gslot);
        dev_free(ftask);
        dev_free(btask);
        dev_free(ftask_wide);
        dev_free(btask_wide);
    }
};

}  // namespace

struct cholespy_gpu_solver;
namespace {
BlockPlanOptions plan_options(const BlockPlanOptions &base = BlockPlanOptions());
int upload_plan(cholespy_gpu_solver *s, const BlockPlan &P, const std::vector<int> *blk_sys = nullptr);
}  // namespace

struct cholespy_gpu_solver {
    uint64_t uid = 0;   // unique per handle (keys the scratch layout)
    int n = 0;
    int fbytes = 4;
    int64_t nnz = 0;
    int device = 0;
    bool uploaded = false;
    bool analysed[2] = {false, false};   // [0] = upper (L'), [1] = lower (L)
    int n_levels[2] = {0, 0};

    int *perm = nullptr;
    // Row-compressed L (diagonal last) and L' (diagonal first).
    int *lrowptr = nullptr, *lcols = nullptr;
    void *lvals = nullptr;
    int *urowptr = nullptr, *ucols = nullptr;
    void *uvals = nullptr;
    int *order[2] = {nullptr, nullptr};   // ticket -> row per pass
    int *lvl1 = nullptr;                  // analysis scratch: level + 1
    // Watchdog words [0] error code, [1] abort flag: host memory mapped into
    // the device, so the host reads them with no device call. status is the
    // device address the kernels get, hstatus the host one.
    int *status = nullptr;
    volatile int *hstatus = nullptr;
    int *dabort = nullptr;                // device copy of the abort word (in qctr)
    // A forward pass (cholespy_solver_solve_pass 0) whose backward pass has
    // not run yet: its partial sums, gathered right-hand sides and y are
    // still published, so the next forward pass must start from a refill.
    bool fwd_pending = false;
    void *y = nullptr, *z = nullptr;      // per-solve work buffers (n x k)
    void *part = nullptr;                 // block path partial slots (nslots x k)
    void *spart = nullptr;                // block path split results (nsplit_rows x k)
    size_t spart_elems = 0;
    void *rbuf = nullptr;                 // block path gathered right-hand sides (n x k)
    int k_cur = 0;                        // layout the work buffers hold
    int path = CHOLESPY_PATH_BLOCKS;
    bool leaf_kernels = true;             // leaf tasks in launches of their own (CHOLESPY_LEAF_KERNELS=0: off)
    int buffers_path = -1;                // path whose invariants the buffers hold
    size_t work_elems = 0;                // capacity of y and z in elements
    size_t part_elems = 0;
    // Work buffers replaced by larger ones. They stay allocated until the
    // handle is destroyed, so a graph captured before the growth still
    // replays into memory that exists (every buffer holds the sentinel
    // between solves, whatever layout it last had).
    std::vector<void *> retired;
    DeviceBlockPlan plan;
    // Full row-compressed A in FP64 (residual of the refinement path).
    int *arowptr = nullptr, *acols = nullptr;
    double *avals = nullptr;
    bool has_matrix = false;
    float *r32 = nullptr;                 // FP32 right-hand side of correction solves
    size_t r32_elems = 0;
    double *norms = nullptr;              // refinement stopping test: [0, 128) residual, [128, 256) b
    double *hnorms = nullptr;             // their host-mapped copy (pinned)
    double *hnorms_dev = nullptr;         // its device address

    int grid_rows[2][2] = {{0, 0}, {0, 0}};   // [lower][k > 64]
    int grid_blocks[2][4] = {{0, 0, 0, 0}, {0, 0, 0, 0}};   // [forward][tile_slot]
    // Ticket queues per kernel: [0..1] analysis upper/lower, [2..3] row pass
    // backward/forward, [4..5] block pass backward/forward.
    unsigned int *qctr = nullptr;
    // [6..7] leaf phases of the fused block passes backward/forward (up to
    // kMaxLeafQueues queues each).
    QueueState queues[8];
    // Leaf tasks inside the block launches instead of launches of their own:
    // bit 0 forward, bit 1 backward (CHOLESPY_FUSED_LEAVES overrides).
    int fused_leaves = 3;
    // Plan options the handle plans with (before environment overrides),
    // and a tile width per right-hand-side class (0 = the rule of
    // block_tile): both set by cholespy_solver_autotune.
    BlockPlanOptions plan_base;
    int tile_for_class[4] = {0, 0, 0, 0};
    std::string tune_report;
    unsigned long long *trace[2] = {nullptr, nullptr};   // block pass timelines [backward, forward]
    std::vector<int> plan_copy[8];        // host copies of the plan for cholespy_solver_plan_arrays
    // Batches of independent systems (cholespy_batch_*): rows are
    // concatenated; per system its size, partial-slot count and host factor
    // until finalize; per call the RHS counts and vector bases.
    int nsys = 1;
    bool is_batch = false;
    std::vector<int> sys_n, sys_row0, sys_slots, sys_slot0, sys_splits, sys_split0;
    struct HostFactor {
        std::vector<int> perm, colptr, rowidx;
        std::vector<double> val;
    };
    std::vector<HostFactor> pending;
    int *blk_sys = nullptr, *sys_k = nullptr;
    long long *sys_vbase = nullptr, *sys_pbase = nullptr, *sys_sbase = nullptr;
    std::vector<int> last_k;   // RHS counts the per-system arrays hold
    std::vector<int> host_levels[2], host_row_level[2];
    // Pattern of the uploaded factor, kept for cholespy_solver_refactor:
    // column form of L, for each position of the row form the column-form
    // position it copies, and the block plan's structure (values cleared).
    std::vector<int> host_colptr, host_rowidx, csr_src;
    BlockPlan host_plan;
};

namespace {
// Plan options: `base` (the defaults, or what cholespy_solver_autotune
// chose) with the CHOLESPY_* environment knobs as overrides on top.
BlockPlanOptions plan_options(const BlockPlanOptions &base) {
    BlockPlanOptions opt = base;
    if (const char *e = std::getenv("CHOLESPY_MAX_BLOCK")) opt.max_block = std::max(1, std::atoi(e));
    if (const char *e = std::getenv("CHOLESPY_LEAF_MAX")) opt.leaf_max = std::max(1, std::atoi(e));
    if (const char *e = std::getenv("CHOLESPY_CRITICAL_ORDER")) opt.critical_order = std::atoi(e) != 0;
    // Task widths are limited to what the block kernels' lane layouts cover
    // (the panel kernels, which run the wide list: 64 rows or columns).
    auto width = [](const char *name, int &v, int hi) {
        if (const char *e = std::getenv(name)) v = std::min(hi, std::max(1, std::atoi(e)));
    };
    width("CHOLESPY_FWD_ROWS", opt.fwd_rows, 64);
    width("CHOLESPY_FWD_ROWS_BIG", opt.fwd_rows_big, 64);
    width("CHOLESPY_BWD_COLS", opt.bwd_cols, 32);
    width("CHOLESPY_BWD_COLS_BIG", opt.bwd_cols_big, 32);
    if (const char *e = std::getenv("CHOLESPY_SPLIT_GATHER")) opt.split_gather = std::max(1, std::atoi(e));
    if (const char *e = std::getenv("CHOLESPY_SPLIT_ENTRIES")) opt.split_entries = std::max(0, std::atoi(e));
    width("CHOLESPY_MAX_SPLIT", opt.max_split, 16);
    if (const char *e = std::getenv("CHOLESPY_LEVEL_TASKS")) opt.level_tasks = std::max(0, std::atoi(e));
    if (const char *e = std::getenv("CHOLESPY_MIN_SPLIT_ENTRIES")) opt.min_split_entries = std::max(1, std::atoi(e));
    width("CHOLESPY_WIDE_FWD_ROWS", opt.wide_fwd_rows, 64);
    width("CHOLESPY_WIDE_FWD_ROWS_BIG", opt.wide_fwd_rows_big, 64);
    if (const char *e = std::getenv("CHOLESPY_WIDE_SPLIT_GATHER")) opt.wide_split_gather = std::max(1, std::atoi(e));
    if (const char *e = std::getenv("CHOLESPY_WIDE_LEVEL_TASKS")) opt.wide_level_tasks = std::max(0, std::atoi(e));
    width("CHOLESPY_WIDE_FINE_BWD_COLS_BIG", opt.wide_fine_bwd_cols_big, 64);
    if (const char *e = std::getenv("CHOLESPY_WIDE_MIN_SPLIT_ENTRIES"))
        opt.wide_min_split_entries = std::max(1, std::atoi(e));
    return opt;
}

// Device task records (kDevTaskInts ints each) for the plan's tasks.
std::vector<int> device_tasks(const BlockPlan &P, const std::vector<int> &tasks, const std::vector<int> *blk_sys) {
    const size_t nt = tasks.size() / kTaskInts;
    std::vector<int> out(nt * kDevTaskInts, 0);
    for (size_t t = 0; t < nt; ++t) {
        const int *src = &tasks[t * kTaskInts];
        int *dst = &out[t * kDevTaskInts];
        for (int i = 0; i < kTaskInts; ++i) dst[i] = src[i];
        const int b = src[0];
        dst[8] = P.start[b];
        dst[9] = P.start[b + 1] - P.start[b];
        dst[10] = P.rptr[b];
        dst[11] = P.rptr[b + 1] - P.rptr[b];
        const uint64_t mo = (uint64_t)P.moff[b];
        dst[12] = (int)(uint32_t)(mo & 0xffffffffu);
        dst[13] = (int)(uint32_t)(mo >> 32);
        dst[14] = blk_sys ? (*blk_sys)[b] : 0;
    }
    return out;
}

// Copies a (possibly merged) block plan to the device in the handle's
// precision; blk_sys maps blocks to systems for a batch.
int upload_plan(cholespy_gpu_solver *s, const BlockPlan &P, const std::vector<int> *blk_sys) {
    DeviceBlockPlan &d = s->plan;
    d.release();
    d.nblocks = P.nblocks;
    d.nslots = P.rptr[P.nblocks];
    d.nsplit_rows = P.nsplit_rows;
    d.max_m = P.max_m;
    d.max_rows = P.max_rows;
    d.n_fwd_tasks = (int)(P.fwd_task.size() / kTaskInts);
    d.n_bwd_tasks = (int)(P.bwd_task.size() / kTaskInts);
    d.n_leaf_fwd = P.n_leaf_fwd;
    d.n_leaf_bwd = P.n_leaf_bwd;
    d.n_fwd_tasks_wide = (int)(P.fwd_task_wide.size() / kTaskInts);
    d.n_bwd_tasks_wide = (int)(P.bwd_task_wide.size() / kTaskInts);
    d.levels_fwd = d.levels_bwd = 0;
    for (int v : P.level_fwd) d.levels_fwd = std::max(d.levels_fwd, v + 1);
    for (int v : P.level_bwd) d.levels_bwd = std::max(d.levels_bwd, v + 1);
    d.m_entries = P.moff[P.nblocks];
    s->plan_copy[0] = P.start;
    s->plan_copy[1] = P.rptr;
    s->plan_copy[2] = P.level_fwd;
    s->plan_copy[3] = P.level_bwd;
    s->plan_copy[4] = P.fwd_task;
    s->plan_copy[5] = P.bwd_task;
    s->plan_copy[6] = P.fwd_task_wide;
    s->plan_copy[7] = P.bwd_task_wide;
    HIP_TRY(upload(&d.start, P.start.data(), P.start.size()));
    HIP_TRY(upload(&d.rptr, P.rptr.data(), P.rptr.size()));
    HIP_TRY(upload(&d.rlist, P.rlist.data(), P.rlist.size()));
    HIP_TRY(upload(&d.moff, P.moff.data(), P.moff.size()));
    HIP_TRY(upload(&d.cptr, P.contrib_ptr.data(), P.contrib_ptr.size()));
    HIP_TRY(upload(&d.cslot, P.contrib_slot.data(), P.contrib_slot.size()));
    HIP_TRY(upload(&d.gslot, P.gslot.data(), P.gslot.size()));
    {
        const std::vector<int> ft = device_tasks(P, P.fwd_task, blk_sys), bt = device_tasks(P, P.bwd_task, blk_sys);
        HIP_TRY(upload(&d.ftask, ft.data(), ft.size()));
        HIP_TRY(upload(&d.btask, bt.data(), bt.size()));
        const std::vector<int> fw = device_tasks(P, P.fwd_task_wide, blk_sys),
                               bw = device_tasks(P, P.bwd_task_wide, blk_sys);
        HIP_TRY(upload(&d.ftask_wide, fw.data(), fw.size()));
        HIP_TRY(upload(&d.btask_wide, bw.data(), bw.size()));
    }
    const size_t me = (size_t)d.m_entries;
    if (s->fbytes == 4) {
        std::vector<float> f(me), b(me);
        for (size_t i = 0; i < me; ++i) {
            f[i] = (float)P.mfwd[i];
            b[i] = (float)P.mbwd[i];
        }
        HIP_TRY(upload(reinterpret_cast<float **>(&d.mfwd), f.data(), me));
        HIP_TRY(upload(reinterpret_cast<float **>(&d.mbwd), b.data(), me));
    } else {
        HIP_TRY(upload(reinterpret_cast<double **>(&d.mfwd), P.mfwd.data(), me));
        HIP_TRY(upload(reinterpret_cast<double **>(&d.mbwd), P.mbwd.data(), me));
    }
    for (auto &row : s->grid_blocks)
        for (int &g : row) g = 0;
    s->buffers_path = -1;
    return CHOLESPY_OK;
}

// Copies the values of L (column form, in the order of host_rowidx) into the
// row and column forms on the device, in the handle's precision.
int upload_factor_values(cholespy_gpu_solver *s, const double *val) {
    const int64_t nnz = s->nnz;
    const size_t fb = (size_t)s->fbytes;
    std::vector<char> lv((size_t)nnz * fb), uv((size_t)nnz * fb);
    if (s->fbytes == 4) {
        float *l = reinterpret_cast<float *>(lv.data()), *u = reinterpret_cast<float *>(uv.data());
        for (int64_t p = 0; p < nnz; ++p) {
            l[p] = (float)val[s->csr_src[(size_t)p]];
            u[p] = (float)val[p];
        }
    } else {
        double *l = reinterpret_cast<double *>(lv.data());
        for (int64_t p = 0; p < nnz; ++p) l[p] = val[s->csr_src[(size_t)p]];
        std::memcpy(uv.data(), val, (size_t)nnz * 8);
    }
    HIP_TRY(hipMemcpy(s->lvals, lv.data(), (size_t)nnz * fb, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(s->uvals, uv.data(), (size_t)nnz * fb, hipMemcpyHostToDevice));
    return CHOLESPY_OK;
}

// Copies the values of M_B (both layouts) into the device plan, whose
// structure they must match.
int upload_plan_values(cholespy_gpu_solver *s, const BlockPlan &P) {
    DeviceBlockPlan &d = s->plan;
    const size_t me = (size_t)d.m_entries;
    if (P.mfwd.size() != me || P.mbwd.size() != me)
        return fail(CHOLESPY_ERR_STATE, "block plan values do not match the uploaded plan");
    if (s->fbytes == 4) {
        std::vector<float> f(me), b(me);
        for (size_t i = 0; i < me; ++i) {
            f[i] = (float)P.mfwd[i];
            b[i] = (float)P.mbwd[i];
        }
        HIP_TRY(hipMemcpy(d.mfwd, f.data(), me * sizeof(float), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(d.mbwd, b.data(), me * sizeof(float), hipMemcpyHostToDevice));
    } else {
        HIP_TRY(hipMemcpy(d.mfwd, P.mfwd.data(), me * sizeof(double), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(d.mbwd, P.mbwd.data(), me * sizeof(double), hipMemcpyHostToDevice));
    }
    return CHOLESPY_OK;
}

// Restores the all-sentinel state of every work buffer (the state a
// completed solve leaves behind).
int refill_work_buffers(cholespy_gpu_solver *s, hipStream_t st) {
    if (s->work_elems) {
        HIP_TRY(launch_fill_sentinel(s->y, s->work_elems, s->fbytes, st));
        HIP_TRY(launch_fill_sentinel(s->z, s->work_elems, s->fbytes, st));
        if (s->rbuf) HIP_TRY(launch_fill_sentinel(s->rbuf, s->work_elems, s->fbytes, st));
    }
    if (s->part && s->part_elems) HIP_TRY(launch_fill_sentinel(s->part, s->part_elems, s->fbytes, st));
    if (s->spart && s->spart_elems) HIP_TRY(launch_fill_sentinel(s->spart, s->spart_elems, s->fbytes, st));
    return CHOLESPY_OK;
}

// After a kernel of the handle gave up a wait (or the abort word was raised):
// once the device is idle, clears the watchdog words and the ticket queues,
// refills the work buffers and returns CHOLESPY_ERR_TIMEOUT, so the next call
// starts from a clean state. CHOLESPY_OK when the words are clear.
int report_give_up(cholespy_gpu_solver *s) {
    if (s->hstatus[0] == 0 && s->hstatus[1] == 0) return CHOLESPY_OK;
    HIP_TRY(hipDeviceSynchronize());
    const bool timed_out = s->hstatus[0] != 0;
    s->hstatus[0] = s->hstatus[1] = 0;
    {
        // A no-allocation solve may have given up with the scratch array's
        // buffers half consumed: refill them before their next use.
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        for (uint64_t &l : g_scratch_layout) l = 0;
    }
    HIP_TRY(hipMemset(s->qctr, 0, kAllQueueWords * sizeof(unsigned int)));
    int rc = refill_work_buffers(s, nullptr);
    if (rc != CHOLESPY_OK) return rc;
    HIP_TRY(hipDeviceSynchronize());
    s->fwd_pending = false;
    return fail(CHOLESPY_ERR_TIMEOUT, timed_out ? "a solve kernel exceeded its bounded wait; its results were NaN"
                                                : "the solver's abort word was raised; the last results may be NaN");
}

}  // namespace

extern "C" {

const char *cholespy_last_error(void) { return g_error.c_str(); }

int cholespy_device_count(int *count) {
    if (!count) return fail(CHOLESPY_ERR_INVALID_ARGUMENT, "count is NULL");
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
    *count = c;
    return CHOLESPY_OK;
}

int cholespy_init(void) {
    if (g_initialised) return CHOLESPY_OK;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
        return fail(CHOLESPY_ERR_NO_DEVICE, "no HIP device is visible");
    HIP_TRY(hipGetDevice(&g_device));
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, g_device));
    g_initialised = true;
    return CHOLESPY_OK;
}

int cholespy_shutdown(void) {
    g_initialised = false;
    g_device = -1;
    return CHOLESPY_OK;
}

int cholespy_copy_to_host(void *dst, const void *device_src, size_t bytes) {
    if (bytes == 0) return CHOLESPY_OK;
    if (!dst || !device_src) return fail(CHOLESPY_ERR_INVALID_ARGUMENT, "null pointer");
    HIP_TRY(hipMemcpy(dst, device_src, bytes, hipMemcpyDeviceToHost));
    return CHOLESPY_OK;
}

int cholespy_solver_create(int n, int float_bytes, cholespy_gpu_solver **out) {
    if (!out) return fail(CHOLESPY_ERR_INVALID_ARGUMENT, "out is NULL");
    *out = nullptr;
    if (n < 0) return fail(CHOLESPY_ERR_INVALID_ARGUMENT, "n must be non-negative");
    if (float_bytes != 4 && float_bytes != 8)
        return fail(CHOLESPY_ERR_INVALID_ARGUMENT, "float_bytes must be 4 or 8");
    int rc = cholespy_init();
    if (rc != CHOLESPY_OK) return rc;
    auto *s = new cholespy_gpu_solver();
    s->uid = g_next_uid.fetch_add(1);
    s->n = n;
    s->fbytes = float_bytes;
    s->device = g_device;
    void *hs = nullptr;
    hipError_t e = hipHostMalloc(&hs, 4 * sizeof(int), hipHostMallocMapped | hipHostMallocCoherent);
    if (e == hipSuccess) {
        s->hstatus = static_cast<volatile int *>(hs);
        for (int i = 0; i < 4; ++i) s->hstatus[i] = 0;
        void *dp = nullptr;
        e = hipHostGetDevicePointer(&dp, hs, 0);
        s->status = static_cast<int *>(dp);
    }
    if (e == hipSuccess) e = dev_alloc(&s->perm, (size_t)n);
    if (e == hipSuccess) e = dev_alloc(&s->lvl1, (size_t)n);
    if (e == hipSuccess) e = dev_alloc(&s->qctr, kAllQueueWords);
    if (e == hipSuccess) e = hipMemset(s->qctr, 0, kAllQueueWords * sizeof(unsigned int));
    if (e == hipSuccess) {
        for (int i = 0; i < 6; ++i) s->queues[i].ctr = s->qctr + (size_t)i * kQueueWords;
        for (int i = 0; i < 2; ++i) {
            s->queues[6 + i].ctr = s->qctr + 6 * kQueueWords + (size_t)i * kLeafQueueWords;
            s->queues[6 + i].nq = kMaxLeafQueues;
        }
        s->dabort = reinterpret_cast<int *>(s->qctr + kAllQueueWords - 16);
        for (QueueState &q : s->queues) q.abort = s->dabort;
    }
    // Block-pass ticket queues (tuning sweeps: CHOLESPY_BLOCK_QUEUES, 1..16).
    if (const char *v = std::getenv("CHOLESPY_BLOCK_QUEUES"))
        s->queues[4].nq = s->queues[5].nq = std::min(16, std::max(1, std::atoi(v)));
    if (const char *v = std::getenv("CHOLESPY_LEAF_KERNELS")) s->leaf_kernels = std::atoi(v) != 0;
    if (const char *v = std::getenv("CHOLESPY_FUSED_LEAVES")) s->fused_leaves = std::atoi(v) & 3;
    if (e != hipSuccess) {
        cholespy_solver_destroy(s);
        return hip_fail(e, "cholespy_solver_create");
    }
    *out = s;
    return CHOLESPY_OK;
}

int cholespy_solver_destroy(cholespy_gpu_solver *s) {
    if (!s) return CHOLESPY_OK;
    (void)hipDeviceSynchronize();
    if (s->hstatus) {
        // Nothing can raise from a destructor: report a give-up no solve call
        // has reported yet the way the reference reports device errors
        // (src/cuda_driver.cpp:48-56), on stderr.
        if (s->hstatus[0] != 0)
            std::fprintf(stderr, "cholespy_amd: a solve kernel exceeded its bounded wait (reported at destroy)\n");
        (void)hipHostFree(const_cast<int *>(s->hstatus));
    }
    s->hstatus = nullptr;
    s->status = nullptr;
    dev_free(s->perm);
    dev_free(s->lvl1);
    dev_free(s->qctr);
    dev_free(s->trace[0]);
    dev_free(s->trace[1]);
    dev_free(s->lrowptr);
    dev_free(s->lcols);
    dev_free(s->lvals);
    dev_free(s->urowptr);
    dev_free(s->ucols);
    dev_free(s->uvals);
    dev_free(s->order[0]);
    dev_free(s->order[1]);
    dev_free(s->y);
    dev_free(s->z);
    dev_free(s->part);
    dev_free(s->spart);
    dev_free(s->rbuf);
    for (void *p : s->retired) (void)hipFree(p);
    s->retired.clear();
    dev_free(s->arowptr);
    dev_free(s->acols);
    dev_free(s->avals);
    dev_free(s->r32);
    dev_free(s->norms);
    if (s->hnorms) (void)hipHostFree(s->hnorms);
    s->hnorms = nullptr;
    dev_free(s->blk_sys);
    dev_free(s->sys_k);
    dev_free(s->sys_vbase);
    dev_free(s->sys_pbase);
    dev_free(s->sys_sbase);
    s->plan.release();
    delete s;
    return CHOLESPY_OK;
}

int cholespy_solver_upload_factor(cholespy_gpu_solver *s, const int *perm, const int *colptr,
                                  const int *rowidx, const double *val, int64_t nnz) {
    if (!s || !perm || !colptr || !rowidx || !val)
        return fail(CHOLESPY_ERR_INVALID_ARGUMENT, "null argument");
    const int n = s->n;
    if (nnz != (int64_t)colptr[n] || colptr[0] != 0 || nnz > 0x7fffffff)
        return fail(CHOLESPY_ERR_INVALID_ARGUMENT, "colptr does not match nnz");
    // Validate everything the kernels index with: perm is a permutation,
    // every column starts with its diagonal and rows ascend inside [0, n).
    {
        std::vector<char> seen((size_t)n, 0);
        for (int i = 0; i < n; ++i) {
            if (perm[i] < 0 || perm[i] >= n || seen[perm[i]])
                return fail(CHOLESPY_ERR_INVALID_ARGUMENT, "perm is not a permutation");
            seen[perm[i]] = 1;
        }
        for (int j = 0; j < n; ++j) {
            if (colptr[j + 1] <= colptr[j] || rowidx[colptr[j]] != j)
                return fail(CHOLESPY_ERR_INVALID_ARGUMENT, "column without leading diagonal");
            for (int p = colptr[j] + 1; p < colptr[j + 1]; ++p)
                if (rowidx[p] <= rowidx[p - 1] || rowidx[p] >= n)
                    return fail(CHOLESPY_ERR_INVALID_ARGUMENT, "row indices not ascending in range");
        }
    }
    // Row-compressed L (transpose of the column form): diagonal ends each row.
    std::vector<int> rowptr((size_t)n + 1, 0), cols((size_t)nnz);
    std::vector<int> csr_src((size_t)nnz);
    for (int64_t p = 0; p < nnz; ++p) rowptr[(size_t)rowidx[p] + 1]++;
    for (int i = 0; i < n; ++i) rowptr[i + 1] += rowptr[i];
    {
        std::vector<int> next(rowptr.begin(), rowptr.end() - 1);
        for (int j = 0; j < n; ++j)
            for (int p = colptr[j]; p < colptr[j + 1]; ++p) {
                const int dst = next[rowidx[p]]++;
                cols[dst] = j;
                csr_src[dst] = p;
            }
    }
    const size_t fb = (size_t)s->fbytes;
    s->uploaded = false;
    dev_free(s->lrowptr);
    dev_free(s->lcols);
    dev_free(s->lvals);
    dev_free(s->urowptr);
    dev_free(s->ucols);
    dev_free(s->uvals);
    HIP_TRY(dev_alloc(&s->lrowptr, (size_t)n + 1));
    HIP_TRY(dev_alloc(&s->lcols, (size_t)nnz));
    HIP_TRY(dev_alloc(reinterpret_cast<char **>(&s->lvals), (size_t)nnz * fb));
    HIP_TRY(dev_alloc(&s->urowptr, (size_t)n + 1));
    HIP_TRY(dev_alloc(&s->ucols, (size_t)nnz));
    HIP_TRY(dev_alloc(reinterpret_cast<char **>(&s->uvals), (size_t)nnz * fb));
    HIP_TRY(hipMemcpy(s->perm, perm, (size_t)n * sizeof(int), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(s->lrowptr, rowptr.data(), ((size_t)n + 1) * sizeof(int), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(s->lcols, cols.data(), (size_t)nnz * sizeof(int), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(s->urowptr, colptr, ((size_t)n + 1) * sizeof(int), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(s->ucols, rowidx, (size_t)nnz * sizeof(int), hipMemcpyHostToDevice));
    s->nnz = nnz;
    s->csr_src.swap(csr_src);
    s->host_colptr.assign(colptr, colptr + n + 1);
    s->host_rowidx.assign(rowidx, rowidx + nnz);
    int rc = upload_factor_values(s, val);
    if (rc != CHOLESPY_OK) return rc;
    s->analysed[0] = s->analysed[1] = false;

    BlockPlanOptions opt = plan_options(s->plan_base);
    BlockPlan P;
    try {
        P = build_block_plan(n, colptr, rowidx, val, opt);
    } catch (const std::exception &ex) {
        return fail(CHOLESPY_ERR_INVALID_ARGUMENT, ex.what());
    }
    rc = upload_plan(s, P);
    if (rc != CHOLESPY_OK) return rc;
    // The structure stays on the host for refactorisations; the values are on the device.
    std::vector<double>().swap(P.mfwd);
    std::vector<double>().swap(P.mbwd);
    s->host_plan = std::move(P);
    s->uploaded = true;
    s->buffers_path = -1;
    return CHOLESPY_OK;
}

int cholespy_normalize_input(int n, int64_t nnz, const void *ii, const void *jj, int index_bytes, const void *x,
                             int value_bytes, int type, int *colptr, int *rowidx, double *val, int64_t *entry_dst,
                             int64_t *nnz_out) {
    if (n < 0 || nnz < 0) return fail(CHOLESPY_ERR_INVALID_ARGUMENT, "Sparse matrix: n_rows must be non-negative.");
    if (nnz > 0x7fffffff) return fail(CHOLESPY_ERR_INVALID_ARGUMENT, "more than 2^31 - 1 input entries");
    if (type < 0 || type > 2 || (index_bytes != 4 && index_bytes != 8) || (value_bytes != 4 && value_bytes != 8))
        return fail(CHOLESPY_ERR_INVALID_ARGUMENT, "bad matrix type or element size");
    if (!colptr || !nnz_out || (nnz > 0 && (!jj || !x || !rowidx || !val)) || (type != 2 && !ii) ||
        (type == 2 && nnz > 0 && !ii))
        return fail(CHOLESPY_ERR_INVALID_ARGUMENT, "null argument");
    int rc = cholespy_init();
    if (rc != CHOLESPY_OK) return rc;
    NormalizedInput out;
    HIP_TRY(normalize_on_device(n, nnz, ii, jj, index_bytes, x, value_bytes, type, out));
    if (!out.invalid.empty()) return fail(CHOLESPY_ERR_INVALID_ARGUMENT, out.invalid);
    std::memcpy(colptr, out.colptr.data(), ((size_t)n + 1) * sizeof(int));
    if (!out.rowidx.empty()) {
        std::memcpy(rowidx, out.rowidx.data(), out.rowidx.size() * sizeof(int));
        std::memcpy(val, out.val.data(), out.val.size() * sizeof(double));
    }
    if (entry_dst && nnz > 0) std::memcpy(entry_dst, out.entry_dst.data(), (size_t)nnz * sizeof(int64_t));
    *nnz_out = (int64_t)out.rowidx.size();
    return CHOLESPY_OK;
}

int cholespy_solver_refactor(cholespy_gpu_solver *s, const double *val, int64_t nnz) {
    if (!s || !val) return fail(CHOLESPY_ERR_INVALID_ARGUMENT, "null argument");
    if (s->is_batch) return fail(CHOLESPY_ERR_INVALID_ARGUMENT, "refactor is not available for batches");
    if (!s->uploaded) return fail(CHOLESPY_ERR_STATE, "refactor called before upload_factor");
    if (nnz != s->nnz) return fail(CHOLESPY_ERR_INVALID_ARGUMENT, "nnz does not match the uploaded factor");
    for (int j = 0; j < s->n; ++j) {
        const double d = val[s->host_colptr[j]];
        if (!(d > 0.0) || !std::isfinite(d))
            return fail(CHOLESPY_ERR_INVALID_ARGUMENT, "factor diagonal must be positive and finite");
    }
    BlockPlan &P = s->host_plan;
    fill_block_plan_values(P, s->host_colptr.data(), s->host_rowidx.data(), val, plan_options(s->plan_base).threads);
    // Solves still in flight read the old values: wait for them before
    // overwriting anything.
    HIP_TRY(hipDeviceSynchronize());
    int rc = upload_factor_values(s, val);
    if (rc == CHOLESPY_OK) rc = upload_plan_values(s, P);
    std::vector<double>().swap(P.mfwd);
    std::vector<double>().swap(P.mbwd);
    return rc;
}

int cholespy_solver_analyze(cholespy_gpu_solver *s, int lower, int *n_levels) {
    if (!s) return fail(CHOLESPY_ERR_INVALID_ARGUMENT, "null solver");
    if (!s->uploaded) return fail(CHOLESPY_ERR_STATE, "analyze called before upload_factor");
    const int n = s->n;
    const int li = lower ? 1 : 0;
    if (n == 0) {
        s->analysed[li] = true;
        s->n_levels[li] = 0;
        if (n_levels) *n_levels = 0;
        return CHOLESPY_OK;
    }
    const int *rowptr = lower ? s->lrowptr : s->urowptr;
    const int *cols = lower ? s->lcols : s->ucols;
    HIP_TRY(hipMemset(s->lvl1, 0, (size_t)n * sizeof(int)));
    // No kernel of this handle may be running while the watchdog words are
    // cleared (one could be giving up right now).
    HIP_TRY(hipDeviceSynchronize());
    s->hstatus[0] = s->hstatus[1] = 0;
    HIP_TRY(hipMemset(s->dabort, 0, sizeof(int)));
    const int cap = max_coresident_blocks(level_analysis_kernel(lower != 0), kRowBlock, 0);
    if (cap <= 0) return fail(CHOLESPY_ERR_HIP, "occupancy query failed for level analysis");
    const int grid = std::max(1, std::min(cap, (n + 4 * kRowChunk - 1) / (4 * kRowChunk)));
    HIP_TRY(launch_level_analysis(lower != 0, n, rowptr, cols, s->lvl1, s->status, grid,
                                  s->queues[li], nullptr));
    std::vector<int> lvl1((size_t)n);
    HIP_TRY(hipMemcpy(lvl1.data(), s->lvl1, (size_t)n * sizeof(int), hipMemcpyDeviceToHost));
    if (s->hstatus[0] != 0 || s->hstatus[1] != 0) {
        s->hstatus[0] = s->hstatus[1] = 0;
        HIP_TRY(hipMemset(s->qctr, 0, kAllQueueWords * sizeof(unsigned int)));
        return fail(CHOLESPY_ERR_TIMEOUT, "level analysis exceeded its bounded wait");
    }
    // Stable counting sort of rows by level (src/cholesky_solver.cpp:348-367).
    int max_lvl = 0;
    std::vector<int> &row_level = s->host_row_level[li];
    row_level.resize((size_t)n);
    for (int i = 0; i < n; ++i) {
        row_level[i] = lvl1[i] - 1;
        max_lvl = std::max(max_lvl, row_level[i]);
    }
    const int nl = max_lvl + 1;
    std::vector<int> start((size_t)nl + 1, 0);
    for (int i = 0; i < n; ++i) start[(size_t)row_level[i] + 1]++;
    for (int l = 0; l < nl; ++l) start[l + 1] += start[l];
    std::vector<int> &levels = s->host_levels[li];
    levels.resize((size_t)n);
    for (int i = 0; i < n; ++i) levels[start[row_level[i]]++] = i;
    dev_free(s->order[li]);
    HIP_TRY(dev_alloc(&s->order[li], (size_t)n));
    HIP_TRY(hipMemcpy(s->order[li], levels.data(), (size_t)n * sizeof(int), hipMemcpyHostToDevice));
    s->n_levels[li] = nl;
    s->analysed[li] = true;
    if (n_levels) *n_levels = nl;
    return CHOLESPY_OK;
}

int cholespy_solver_get_levels(const cholespy_gpu_solver *s, int lower, int *levels, int *row_level) {
    if (!s || !levels) return fail(CHOLESPY_ERR_INVALID_ARGUMENT, "null argument");
    const int li = lower ? 1 : 0;
    if (!s->analysed[li]) return fail(CHOLESPY_ERR_STATE, "pass not analysed");
    std::copy(s->host_levels[li].begin(), s->host_levels[li].end(), levels);
    if (row_level) std::copy(s->host_row_level[li].begin(), s->host_row_level[li].end(), row_level);
    return CHOLESPY_OK;
}

int cholespy_solver_check(cholespy_gpu_solver *s) {
    if (!s) return fail(CHOLESPY_ERR_INVALID_ARGUMENT, "null solver");
    HIP_TRY(hipDeviceSynchronize());
    return report_give_up(s);
}

int cholespy_solver_set_spin_limit(cholespy_gpu_solver *s, uint32_t limit) {
    if (!s) return fail(CHOLESPY_ERR_INVALID_ARGUMENT, "null solver");
    if (limit == 0) limit = kSpinLimit;
    for (QueueState &q : s->queues) q.spin_limit = limit;
    return CHOLESPY_OK;
}

int cholespy_solver_raise_abort(cholespy_gpu_solver *s) {
    if (!s) return fail(CHOLESPY_ERR_INVALID_ARGUMENT, "null solver");
    s->hstatus[1] = 1;
    const int one = 1;
    HIP_TRY(hipMemcpy(s->dabort, &one, sizeof one, hipMemcpyHostToDevice));
    return CHOLESPY_OK;
}

}  // extern "C"

namespace {

// Makes the work buffers hold n x k elements in the layout and sentinel
// state the current path expects.
// Capacity for a buffer of `rows` rows that must now hold k columns, having
// held `have` elements: at least twice the old capacity (at most 128
// columns), so the retired buffers of a handle add up to less than its
// final ones.
size_t grown_capacity(size_t rows, int k, size_t have) {
    const size_t need = rows * (size_t)k;
    return std::max(need, std::min(2 * have, rows * (size_t)kMaxRhs));
}

template <typename T>
void retire(cholespy_gpu_solver *s, T *&p) {
    if (p) s->retired.push_back(static_cast<void *>(p));
    p = nullptr;
}

int prepare_buffers(cholespy_gpu_solver *s, int k, hipStream_t st) {
    const int fbytes = s->fbytes;
    const size_t elems = (size_t)s->n * (size_t)k;
    if (elems > s->work_elems) {
        const size_t cap = grown_capacity((size_t)s->n, k, s->work_elems);
        retire(s, s->y);
        retire(s, s->z);
        retire(s, s->rbuf);
        HIP_TRY(dev_alloc(reinterpret_cast<char **>(&s->y), cap * (size_t)fbytes));
        HIP_TRY(dev_alloc(reinterpret_cast<char **>(&s->z), cap * (size_t)fbytes));
        HIP_TRY(dev_alloc(reinterpret_cast<char **>(&s->rbuf), cap * (size_t)fbytes));
        s->work_elems = cap;
        s->k_cur = 0;
    }
    const size_t pe = (size_t)s->plan.nslots * (size_t)k;
    if (s->path == CHOLESPY_PATH_BLOCKS && pe > s->part_elems) {
        const size_t cap = grown_capacity((size_t)s->plan.nslots, k, s->part_elems);
        retire(s, s->part);
        HIP_TRY(dev_alloc(reinterpret_cast<char **>(&s->part), std::max<size_t>(cap, 1) * (size_t)fbytes));
        s->part_elems = cap;
        s->k_cur = 0;
    }
    const size_t se = (size_t)s->plan.nsplit_rows * (size_t)k;
    if (s->path == CHOLESPY_PATH_BLOCKS && se > s->spart_elems) {
        const size_t cap = grown_capacity((size_t)s->plan.nsplit_rows, k, s->spart_elems);
        retire(s, s->spart);
        HIP_TRY(dev_alloc(reinterpret_cast<char **>(&s->spart), std::max<size_t>(cap, 1) * (size_t)fbytes));
        s->spart_elems = cap;
        s->k_cur = 0;
    }
    if (k != s->k_cur || s->buffers_path != s->path) {
        // Buffers are indexed row * k + col: refill them for the new layout.
        HIP_TRY(launch_fill_sentinel(s->y, elems, fbytes, st));
        HIP_TRY(launch_fill_sentinel(s->z, elems, fbytes, st));
        HIP_TRY(launch_fill_sentinel(s->rbuf, elems, fbytes, st));
        if (s->part && pe) HIP_TRY(launch_fill_sentinel(s->part, pe, fbytes, st));
        if (s->spart && se) HIP_TRY(launch_fill_sentinel(s->spart, se, fbytes, st));
        s->k_cur = k;
        s->buffers_path = s->path;
    }
    return CHOLESPY_OK;
}

int solve_rows(cholespy_gpu_solver *s, int k, const void *b, void *x, hipStream_t st, int only_pass = -1) {
    const int fbytes = s->fbytes, n = s->n;
    RowPassBuffers bf;
    bf.n = n;
    bf.k = k;
    bf.perm = s->perm;
    bf.b = b;
    bf.x = x;
    bf.y = s->y;
    bf.z = s->z;
    bf.status = s->status;
    for (int pass = 0; pass < 2; ++pass) {
        if (only_pass >= 0 && pass != only_pass) continue;
        const bool lower = pass == 0;
        bf.order = s->order[lower ? 1 : 0];
        bf.rowptr = lower ? s->lrowptr : s->urowptr;
        bf.cols = lower ? s->lcols : s->ucols;
        bf.vals = lower ? s->lvals : s->uvals;
        int &grid = s->grid_rows[lower ? 1 : 0][k > kWave ? 1 : 0];
        if (grid == 0) {
            const int cap = max_coresident_blocks(row_pass_kernel(fbytes, lower, k), kRowBlock, 0);
            if (cap <= 0) return fail(CHOLESPY_ERR_HIP, "occupancy query failed for the solve");
            grid = std::max(1, std::min(cap, (n + 4 * kRowChunk - 1) / (4 * kRowChunk)));
        }
        HIP_TRY(launch_row_pass(fbytes, lower, bf, grid, s->queues[2 + (lower ? 1 : 0)], st));
    }
    return CHOLESPY_OK;
}

// Tile width of a solve with k right-hand sides: the handle's tuned choice
// for k's class when it holds k (a panel tile at least k wide; the narrow
// tile only for k <= 4), else block_tile's rule.
int effective_tile(const cholespy_gpu_solver *s, int fbytes, int k) {
    const int rule = block_tile(fbytes, k);
    const int t = s->is_batch ? 0 : s->tile_for_class[tile_slot(rule)];
    if (t == kNarrowTile && k <= kNarrowTile) return t;
    if (t >= 32 && t >= k) return t;
    return rule;
}

// Work buffers of one block solve (the handle's own, or the library's
// scratch array for cholespy_solver_solve_noalloc).
struct WorkSet {
    void *y, *z, *part, *spart, *rbuf;
};

int solve_blocks(cholespy_gpu_solver *s, int k, const void *b, void *x, hipStream_t st, int xmode = 0,
                 int only_pass = -1, const WorkSet *work = nullptr) {
    const int fbytes = s->fbytes;
    const DeviceBlockPlan &d = s->plan;
    BlockPassBuffers bf;
    bf.k = k;
    bf.max_k = k;
    if (s->is_batch)
        for (int v : s->last_k) bf.max_k = std::max(bf.max_k, v);
    bf.max_m = d.max_m;
    bf.max_rows = d.max_rows;
    bf.start = d.start;
    bf.rptr = d.rptr;
    bf.rlist = d.rlist;
    bf.moff = d.moff;
    bf.cptr = d.cptr;
    bf.cslot = d.cslot;
    bf.gslot = d.gslot;
    bf.perm = s->perm;
    bf.b = b;
    bf.x = x;
    const WorkSet own{s->y, s->z, s->part, s->spart, s->rbuf};
    const WorkSet &w = work ? *work : own;
    bf.y = w.y;
    bf.z = w.z;
    bf.part = w.part;
    bf.spart = w.spart;
    bf.rbuf = w.rbuf;
    bf.status = s->status;
    bf.xmode = xmode;
    if (s->is_batch) {
        bf.blk_sys = s->blk_sys;
        bf.sys_k = s->sys_k;
        bf.sys_vbase = s->sys_vbase;
        bf.sys_pbase = s->sys_pbase;
        bf.sys_sbase = s->sys_sbase;
    }
    for (int pass = 0; pass < 2; ++pass) {
        if (only_pass >= 0 && pass != only_pass) continue;
        const bool forward = pass == 0;
        const int tile = effective_tile(s, fbytes, bf.max_k);
        if (is_panel_tile(tile)) {
            bf.ntasks = forward ? d.n_fwd_tasks_wide : d.n_bwd_tasks_wide;
            bf.n_leaf = 0;
            bf.task = forward ? d.ftask_wide : d.btask_wide;
        } else {
            bf.ntasks = forward ? d.n_fwd_tasks : d.n_bwd_tasks;
            bf.n_leaf = s->leaf_kernels ? (forward ? d.n_leaf_fwd : d.n_leaf_bwd) : 0;
            bf.task = forward ? d.ftask : d.btask;
        }
        bf.M = forward ? d.mfwd : d.mbwd;
        bf.trace = s->trace[forward ? 1 : 0];
        int &grid = s->grid_blocks[forward ? 1 : 0][tile_slot(tile)];
        if (grid == 0) {
            const size_t lds = block_pass_lds(fbytes, forward ? d.max_m : d.max_rows, tile);
            const int cap = max_coresident_blocks(block_pass_kernel(fbytes, forward, tile), 256, lds);
            if (cap <= 0) return fail(CHOLESPY_ERR_HIP, "occupancy query failed for the block solve");
            grid = std::max(1, std::min(cap, bf.ntasks));
        }
        bf.fused_leaves = (s->fused_leaves >> (forward ? 0 : 1)) & 1;
        bf.tile = tile;
        bf.leaf_queues = &s->queues[6 + (forward ? 1 : 0)];
        HIP_TRY(launch_block_pass(fbytes, forward, bf, grid, s->queues[4 + (forward ? 1 : 0)], st));
    }
    return CHOLESPY_OK;
}

int solve_impl(cholespy_gpu_solver *s, int fbytes, int k, const void *b, void *x, void *stream,
               int only_pass = -1) {
    if (!s) return fail(CHOLESPY_ERR_INVALID_ARGUMENT, "null solver");
    if (s->is_batch) return fail(CHOLESPY_ERR_INVALID_ARGUMENT, "batch handles solve with cholespy_batch_solve_*");
    if (s->fbytes != fbytes)
        return fail(CHOLESPY_ERR_INVALID_ARGUMENT, "solver precision does not match the call");
    if (k > kMaxRhs) return fail(CHOLESPY_ERR_INVALID_ARGUMENT, "The number of RHS should be less than 128.");
    if (k < 0) return fail(CHOLESPY_ERR_INVALID_ARGUMENT, "negative number of RHS");
    if (s->path == CHOLESPY_PATH_ROWS && (!s->analysed[0] || !s->analysed[1]))
        return fail(CHOLESPY_ERR_STATE, "the row schedule needs both passes analysed");
    if (!s->uploaded) return fail(CHOLESPY_ERR_STATE, "solve called before the factor was uploaded");
    if (k == 0 || s->n == 0) return CHOLESPY_OK;
    if (!b || !x) return fail(CHOLESPY_ERR_INVALID_ARGUMENT, "null b or x");
    // A wait that gave up in an earlier call of this handle is reported here
    // (RuntimeError in the binding); the handle is then clean again.
    int rc = report_give_up(s);
    if (rc != CHOLESPY_OK) return rc;
    hipStream_t st = static_cast<hipStream_t>(stream);
    rc = prepare_buffers(s, k, st);
    if (rc != CHOLESPY_OK) return rc;
    // Pass order. A forward pass publishes partial sums, gathered right-hand
    // sides and y, and only the backward pass resets them: a forward pass
    // (or a whole solve) that follows a forward pass with no backward pass
    // in between starts from refilled buffers. A lone backward pass waits on
    // z, which the forward pass resets: refill z before it.
    if (only_pass != 1 && s->fwd_pending) {
        rc = refill_work_buffers(s, st);
        if (rc != CHOLESPY_OK) return rc;
    } else if (only_pass == 1 && !s->fwd_pending) {
        HIP_TRY(launch_fill_sentinel(s->z, s->work_elems, s->fbytes, st));
    }
    rc = s->path == CHOLESPY_PATH_ROWS ? solve_rows(s, k, b, x, st, only_pass)
                                       : solve_blocks(s, k, b, x, st, 0, only_pass);
    if (rc == CHOLESPY_OK) s->fwd_pending = only_pass == 0;
    return rc;
}

}  // namespace

extern "C" {

int cholespy_solver_upload_matrix(cholespy_gpu_solver *s, const int *colptr, const int *rowidx,
                                  const double *val, int64_t nnz) {
    if (!s || !colptr || (!rowidx && nnz) || (!val && nnz))
        return fail(CHOLESPY_ERR_INVALID_ARGUMENT, "null argument");
    const int n = s->n;
    if (colptr[0] != 0 || (int64_t)colptr[n] != nnz)
        return fail(CHOLESPY_ERR_INVALID_ARGUMENT, "colptr does not match nnz");
    for (int j = 0; j < n; ++j) {
        if (colptr[j + 1] < colptr[j]) return fail(CHOLESPY_ERR_INVALID_ARGUMENT, "colptr not monotone");
        for (int p = colptr[j]; p < colptr[j + 1]; ++p)
            if (rowidx[p] < j || rowidx[p] >= n)
                return fail(CHOLESPY_ERR_INVALID_ARGUMENT, "matrix entry outside the lower triangle");
    }
    // Full symmetric rows: (i, j) and (j, i) for every stored i > j.
    std::vector<int> rowptr((size_t)n + 1, 0);
    for (int j = 0; j < n; ++j)
        for (int p = colptr[j]; p < colptr[j + 1]; ++p) {
            rowptr[(size_t)rowidx[p] + 1]++;
            if (rowidx[p] != j) rowptr[(size_t)j + 1]++;
        }
    for (int i = 0; i < n; ++i) rowptr[i + 1] += rowptr[i];
    std::vector<int> cols((size_t)rowptr[n]);
    std::vector<double> vals((size_t)rowptr[n]);
    {
        std::vector<int> next(rowptr.begin(), rowptr.end() - 1);
        // Column j's entries go to rows i (column index j, ascending j) and,
        // mirrored, to row j (column indices i > j, ascending): each row
        // ends up sorted by column.
        for (int j = 0; j < n; ++j)
            for (int p = colptr[j]; p < colptr[j + 1]; ++p) {
                const int i = rowidx[p];
                cols[next[i]] = j;
                vals[next[i]++] = val[p];
            }
        for (int j = 0; j < n; ++j)
            for (int p = colptr[j]; p < colptr[j + 1]; ++p) {
                const int i = rowidx[p];
                if (i == j) continue;
                cols[next[j]] = i;
                vals[next[j]++] = val[p];
            }
    }
    dev_free(s->arowptr);
    dev_free(s->acols);
    dev_free(s->avals);
    HIP_TRY(upload(&s->arowptr, rowptr.data(), rowptr.size()));
    HIP_TRY(upload(&s->acols, cols.data(), cols.size()));
    HIP_TRY(upload(&s->avals, vals.data(), vals.size()));
    s->has_matrix = true;
    return CHOLESPY_OK;
}

}  // extern "C"

namespace {

// Mixed-precision solve (both entry points): an FP32 block solve, then
// corrections x += A^{-1} (b - A x). tol < 0: exactly `iterations`
// corrections, all enqueued (no host synchronisation, graph-capturable).
// tol >= 0: before each correction the residual's column norms come back to
// the host, and refinement stops once max_j ||b_j - A x_j|| / ||b_j|| <= tol
// or after `iterations` corrections; *steps and *residual report the
// corrections applied and the last relative residual.
int refined_impl(cholespy_gpu_solver *s, int k, const double *b, double *x, int iterations, double tol,
                 void *stream, int *steps, double *residual) {
    if (!s) return fail(CHOLESPY_ERR_INVALID_ARGUMENT, "null solver");
    if (s->fbytes != 4) return fail(CHOLESPY_ERR_INVALID_ARGUMENT, "refined solves need a float32 solver");
    if (!s->has_matrix) return fail(CHOLESPY_ERR_STATE, "refined solve called before upload_matrix");
    if (k > kMaxRhs) return fail(CHOLESPY_ERR_INVALID_ARGUMENT, "The number of RHS should be less than 128.");
    if (k < 0 || iterations < 0) return fail(CHOLESPY_ERR_INVALID_ARGUMENT, "negative argument");
    if (k == 0 || s->n == 0) return CHOLESPY_OK;
    if (!b || !x) return fail(CHOLESPY_ERR_INVALID_ARGUMENT, "null b or x");
    if (s->path != CHOLESPY_PATH_BLOCKS) return fail(CHOLESPY_ERR_STATE, "refined solves use the block path");
    int rc = report_give_up(s);
    if (rc != CHOLESPY_OK) return rc;
    hipStream_t st = static_cast<hipStream_t>(stream);
    rc = prepare_buffers(s, k, st);
    if (rc != CHOLESPY_OK) return rc;
    if (s->fwd_pending) {
        rc = refill_work_buffers(s, st);
        if (rc != CHOLESPY_OK) return rc;
        s->fwd_pending = false;
    }
    const size_t elems = (size_t)s->n * (size_t)k;
    if (elems > s->r32_elems) {
        dev_free(s->r32);
        HIP_TRY(dev_alloc(&s->r32, elems));
        s->r32_elems = elems;
    }
    // x = A^{-1} b in FP32, then the corrections x += A^{-1}(b - A x).
    HIP_TRY(launch_cast_to_f32(b, s->r32, elems, st));
    rc = solve_blocks(s, k, s->r32, x, st, 1);
    if (rc != CHOLESPY_OK) return rc;
    if (tol < 0) {
        for (int it = 0; it < iterations; ++it) {
            HIP_TRY(launch_residual_f64(s->n, k, s->arowptr, s->acols, s->avals, b, x, s->r32, nullptr, st));
            rc = solve_blocks(s, k, s->r32, x, st, 2);
            if (rc != CHOLESPY_OK) return rc;
        }
        if (steps) *steps = iterations;
        return CHOLESPY_OK;
    }
    if (!s->norms) HIP_TRY(dev_alloc(&s->norms, 2 * (size_t)kMaxRhs));
    if (!s->hnorms) {
        void *hp = nullptr, *dp = nullptr;
        HIP_TRY(hipHostMalloc(&hp, 2 * (size_t)kMaxRhs * sizeof(double), hipHostMallocMapped | hipHostMallocCoherent));
        s->hnorms = static_cast<double *>(hp);
        HIP_TRY(hipHostGetDevicePointer(&dp, hp, 0));
        s->hnorms_dev = static_cast<double *>(dp);
    }
    double *rn = s->norms, *bn = s->norms + kMaxRhs;
    HIP_TRY(launch_column_norms2((size_t)s->n, k, b, bn, st));
    const volatile double *h = s->hnorms;
    int it = 0;
    double rel = 0.0;
    for (;; ++it) {
        if (it == 0 && tol < 1e-6 && iterations > 0) {
            // An FP32 solve is never that accurate: correct once before the
            // first check (one synchronisation fewer).
            HIP_TRY(launch_residual_f64(s->n, k, s->arowptr, s->acols, s->avals, b, x, s->r32, nullptr, st));
            rc = solve_blocks(s, k, s->r32, x, st, 2);
            if (rc != CHOLESPY_OK) return rc;
            continue;
        }
        HIP_TRY(launch_zero_f64(rn, k, st));
        HIP_TRY(launch_residual_f64(s->n, k, s->arowptr, s->acols, s->avals, b, x, s->r32, rn, st));
        // The norms straight into pinned host memory (no copy engine), then wait.
        HIP_TRY(launch_copy_out_f64(s->hnorms_dev, s->norms, 2 * kMaxRhs, st));
        HIP_TRY(hipStreamSynchronize(st));
        rc = report_give_up(s);
        if (rc != CHOLESPY_OK) return rc;
        rel = 0.0;
        for (int t = 0; t < k; ++t) {
            const double r = std::sqrt(h[t]), bb = std::sqrt(h[kMaxRhs + t]);
            rel = std::max(rel, bb > 0 ? r / bb : (r > 0 ? INFINITY : 0.0));
        }
        if (!(rel > tol) || it >= iterations) break;   // NaN residuals stop too
        rc = solve_blocks(s, k, s->r32, x, st, 2);
        if (rc != CHOLESPY_OK) return rc;
    }
    if (steps) *steps = it;
    if (residual) *residual = rel;
    return CHOLESPY_OK;
}

}  // namespace

extern "C" {

int cholespy_solver_solve_refined(cholespy_gpu_solver *s, int k, const double *b, double *x,
                                  int iterations, void *stream) {
    return refined_impl(s, k, b, x, iterations, -1.0, stream, nullptr, nullptr);
}

int cholespy_solver_solve_refined_tol(cholespy_gpu_solver *s, int k, const double *b, double *x, double tol,
                                      int max_steps, void *stream, int *steps, double *residual) {
    if (!(tol >= 0)) return fail(CHOLESPY_ERR_INVALID_ARGUMENT, "tol must be non-negative");
    return refined_impl(s, k, b, x, max_steps, tol, stream, steps, residual);
}

int cholespy_solver_solve_pass(cholespy_gpu_solver *s, int pass, int n_rhs, const void *b, void *x,
                               void *stream) {
    if (pass != 0 && pass != 1) return fail(CHOLESPY_ERR_INVALID_ARGUMENT, "pass must be 0 or 1");
    if (!s) return fail(CHOLESPY_ERR_INVALID_ARGUMENT, "null solver");
    return solve_impl(s, s->fbytes, n_rhs, b ? b : x, x ? x : const_cast<void *>(b), stream, pass);
}

int cholespy_entry_gradient(int n, int64_t nnz, const int *rows, const int *cols, int n_rhs, const void *lam,
                            const void *x, int vec_bytes, void *grad, int grad_bytes, void *stream) {
    if (n < 0 || nnz < 0 || n_rhs < 0) return fail(CHOLESPY_ERR_INVALID_ARGUMENT, "negative size");
    if ((vec_bytes != 4 && vec_bytes != 8) || (grad_bytes != 4 && grad_bytes != 8))
        return fail(CHOLESPY_ERR_INVALID_ARGUMENT, "element sizes must be 4 or 8");
    if (nnz == 0) return CHOLESPY_OK;
    if (!rows || !cols || !grad || (n_rhs > 0 && (!lam || !x)))
        return fail(CHOLESPY_ERR_INVALID_ARGUMENT, "null pointer");
    HIP_TRY(launch_entry_gradient(n, nnz, rows, cols, n_rhs, lam, x, vec_bytes, grad, grad_bytes,
                                  static_cast<hipStream_t>(stream)));
    return CHOLESPY_OK;
}

int cholespy_solver_set_trace(cholespy_gpu_solver *s, int enable) {
    if (!s) return fail(CHOLESPY_ERR_INVALID_ARGUMENT, "null solver");
    dev_free(s->trace[0]);
    dev_free(s->trace[1]);
    if (!enable) return CHOLESPY_OK;
    // Either task list (narrow or wide right-hand sides) may run while tracing.
    const size_t words[2] = {4 * (size_t)std::max({1, s->plan.n_bwd_tasks, s->plan.n_bwd_tasks_wide}),
                             4 * (size_t)std::max({1, s->plan.n_fwd_tasks, s->plan.n_fwd_tasks_wide})};
    for (int i = 0; i < 2; ++i) {
        HIP_TRY(dev_alloc(&s->trace[i], words[i]));
        HIP_TRY(hipMemset(s->trace[i], 0, words[i] * sizeof(unsigned long long)));
    }
    return CHOLESPY_OK;
}

int cholespy_solver_get_trace(const cholespy_gpu_solver *s, int forward, uint64_t *out, int64_t words) {
    if (!s || !out) return fail(CHOLESPY_ERR_INVALID_ARGUMENT, "null argument");
    const unsigned long long *t = s->trace[forward ? 1 : 0];
    if (!t) return fail(CHOLESPY_ERR_STATE, "tracing is off");
    const int64_t have = 4 * (int64_t)(forward ? std::max(s->plan.n_fwd_tasks, s->plan.n_fwd_tasks_wide)
                                               : std::max(s->plan.n_bwd_tasks, s->plan.n_bwd_tasks_wide));
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemcpy(out, t, (size_t)std::min(have, words) * sizeof(uint64_t), hipMemcpyDeviceToHost));
    return CHOLESPY_OK;
}

int cholespy_batch_create(int count, const int *n, int float_bytes, cholespy_gpu_solver **out) {
    if (!out || count < 1 || !n) return fail(CHOLESPY_ERR_INVALID_ARGUMENT, "bad batch arguments");
    int64_t total = 0;
    for (int i = 0; i < count; ++i) {
        if (n[i] < 0) return fail(CHOLESPY_ERR_INVALID_ARGUMENT, "negative system size");
        total += n[i];
    }
    if (total > 0x7fffffff) return fail(CHOLESPY_ERR_INVALID_ARGUMENT, "batch too large for 32-bit rows");
    int rc = cholespy_solver_create((int)total, float_bytes, out);
    if (rc != CHOLESPY_OK) return rc;
    cholespy_gpu_solver *s = *out;
    s->is_batch = true;
    s->nsys = count;
    s->sys_n.assign(n, n + count);
    s->sys_row0.assign((size_t)count, 0);
    for (int i = 1; i < count; ++i) s->sys_row0[i] = s->sys_row0[i - 1] + n[i - 1];
    s->pending.assign((size_t)count, {});
    return CHOLESPY_OK;
}

int cholespy_batch_set_factor(cholespy_gpu_solver *s, int index, const int *perm, const int *colptr,
                              const int *rowidx, const double *val, int64_t nnz) {
    if (!s || !s->is_batch) return fail(CHOLESPY_ERR_INVALID_ARGUMENT, "not a batch handle");
    if (index < 0 || index >= s->nsys) return fail(CHOLESPY_ERR_INVALID_ARGUMENT, "system index out of range");
    const int n = s->sys_n[index];
    if (!perm || !colptr || !rowidx || !val) return fail(CHOLESPY_ERR_INVALID_ARGUMENT, "null argument");
    if (nnz != (int64_t)colptr[n] || colptr[0] != 0) return fail(CHOLESPY_ERR_INVALID_ARGUMENT, "colptr does not match nnz");
    std::vector<char> seen((size_t)n, 0);
    for (int i = 0; i < n; ++i) {
        if (perm[i] < 0 || perm[i] >= n || seen[perm[i]])
            return fail(CHOLESPY_ERR_INVALID_ARGUMENT, "perm is not a permutation");
        seen[perm[i]] = 1;
    }
    for (int j = 0; j < n; ++j) {
        if (colptr[j + 1] <= colptr[j] || rowidx[colptr[j]] != j)
            return fail(CHOLESPY_ERR_INVALID_ARGUMENT, "column without leading diagonal");
        for (int p = colptr[j] + 1; p < colptr[j + 1]; ++p)
            if (rowidx[p] <= rowidx[p - 1] || rowidx[p] >= n)
                return fail(CHOLESPY_ERR_INVALID_ARGUMENT, "row indices not ascending in range");
    }
    auto &h = s->pending[index];
    h.perm.assign(perm, perm + n);
    h.colptr.assign(colptr, colptr + n + 1);
    h.rowidx.assign(rowidx, rowidx + nnz);
    h.val.assign(val, val + nnz);
    return CHOLESPY_OK;
}

int cholespy_batch_finalize(cholespy_gpu_solver *s) {
    if (!s || !s->is_batch) return fail(CHOLESPY_ERR_INVALID_ARGUMENT, "not a batch handle");
    const int ns = s->nsys;
    for (int i = 0; i < ns; ++i)
        if ((int)s->pending[i].colptr.size() != s->sys_n[i] + 1)
            return fail(CHOLESPY_ERR_STATE, "a system of the batch has no factor");
    // Per-system plans, built on host threads. Many systems already give
    // every level plenty of independent tasks, so their reductions are not
    // split (extra tasks would only add hand-offs) unless
    // CHOLESPY_SPLIT_ENTRIES asks for it, and gathers are split only past
    // 257 contributions unless CHOLESPY_SPLIT_GATHER is set (256 grid
    // systems, n 1k-50k, k 1-128: 10.53 ms against 10.77 ms at 65).
    BlockPlanOptions opt = plan_options();
    opt.threads = 1;
    if (ns >= 8 && !std::getenv("CHOLESPY_SPLIT_ENTRIES")) opt.split_entries = 0;
    if (ns >= 8 && !std::getenv("CHOLESPY_SPLIT_GATHER")) opt.split_gather = 257;
    if (ns >= 8) opt.wide_list = false;
    std::vector<BlockPlan> plans((size_t)ns);
    {
        const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        std::vector<std::thread> pool;
        std::vector<std::string> errors((size_t)ns);
        for (unsigned w = 0; w < hw; ++w)
            pool.emplace_back([&, w]() {
                for (int i = (int)w; i < ns; i += (int)hw) {
                    const auto &h = s->pending[i];
                    try {
                        plans[i] = build_block_plan(s->sys_n[i], h.colptr.data(), h.rowidx.data(), h.val.data(), opt);
                    } catch (const std::exception &ex) {
                        errors[i] = ex.what();
                    }
                }
            });
        for (auto &t : pool) t.join();
        for (const std::string &e : errors)
            if (!e.empty()) return fail(CHOLESPY_ERR_INVALID_ARGUMENT, e);
    }
    // Merge: blocks, slots, contributions and dense values concatenated with
    // offsets; tasks of all systems interleaved by block level.
    BlockPlan G;
    G.n = s->n;
    std::vector<int> blk_sys, gperm((size_t)s->n);
    s->sys_slots.assign((size_t)ns, 0);
    s->sys_slot0.assign((size_t)ns, 0);
    s->sys_splits.assign((size_t)ns, 0);
    s->sys_split0.assign((size_t)ns, 0);
    int B0 = 0, S0 = 0, C0 = 0, SP0 = 0;
    int64_t M0 = 0;
    struct TaskRef { int level, sys, idx; };
    std::vector<TaskRef> ftasks, btasks, ftasks_w, btasks_w;
    for (int i = 0; i < ns; ++i) {
        const BlockPlan &P = plans[i];
        const int R0 = s->sys_row0[i];
        s->sys_slot0[i] = S0;
        s->sys_slots[i] = P.rptr[P.nblocks];
        s->sys_split0[i] = SP0;
        s->sys_splits[i] = P.nsplit_rows;
        for (int r = 0; r < s->sys_n[i]; ++r) gperm[(size_t)R0 + r] = R0 + s->pending[i].perm[r];
        for (int b = 0; b < P.nblocks; ++b) {
            G.start.push_back(P.start[b] + R0);
            G.rptr.push_back(P.rptr[b] + S0);
            G.moff.push_back(P.moff[b] + M0);
            G.level_fwd.push_back(P.level_fwd[b]);
            G.level_bwd.push_back(P.level_bwd[b]);
            blk_sys.push_back(i);
        }
        for (int v : P.rlist) G.rlist.push_back(v + R0);
        G.mfwd.insert(G.mfwd.end(), P.mfwd.begin(), P.mfwd.end());
        G.mbwd.insert(G.mbwd.end(), P.mbwd.begin(), P.mbwd.end());
        for (int r = 0; r < s->sys_n[i]; ++r) G.contrib_ptr.push_back(P.contrib_ptr[r] + C0);
        for (int v : P.contrib_slot) G.contrib_slot.push_back(v + S0);
        for (int v : P.gslot) G.gslot.push_back(v >= 0 ? v + S0 : -1);
        // Sort keys: the level, with leaf tasks first (forward) or last
        // (backward) as in each system's plan.
        const size_t nf = P.fwd_task.size() / kTaskInts, nb = P.bwd_task.size() / kTaskInts;
        for (size_t t = 0; t < nf; ++t)
            ftasks.push_back({(int)t < P.n_leaf_fwd ? -1 : P.level_fwd[P.fwd_task[kTaskInts * t]], i, (int)t});
        for (size_t t = 0; t < nb; ++t)
            btasks.push_back({t >= nb - (size_t)P.n_leaf_bwd ? std::numeric_limits<int>::max()
                                                             : P.level_bwd[P.bwd_task[kTaskInts * t]],
                              i, (int)t});
        G.n_leaf_fwd += P.n_leaf_fwd;
        G.n_leaf_bwd += P.n_leaf_bwd;
        for (size_t t = 0; t < P.fwd_task_wide.size() / kTaskInts; ++t)
            ftasks_w.push_back({P.level_fwd[P.fwd_task_wide[kTaskInts * t]], i, (int)t});
        for (size_t t = 0; t < P.bwd_task_wide.size() / kTaskInts; ++t)
            btasks_w.push_back({P.level_bwd[P.bwd_task_wide[kTaskInts * t]], i, (int)t});

        G.max_m = std::max(G.max_m, P.max_m);
        G.max_rows = std::max(G.max_rows, P.max_rows);
        G.nnz_l += P.nnz_l;
        s->sys_row0[i] = R0;
        B0 += P.nblocks;
        S0 += P.rptr[P.nblocks];
        SP0 += P.nsplit_rows;
        C0 += P.contrib_ptr[s->sys_n[i]];
        M0 += P.moff[P.nblocks];
    }
    G.nblocks = B0;
    G.start.push_back(s->n);
    G.rptr.push_back(S0);
    G.moff.push_back(M0);
    G.contrib_ptr.push_back(C0);
    G.nsplit_rows = SP0;
    std::vector<int> blk0((size_t)ns + 1, 0);
    for (int i = 0; i < ns; ++i) blk0[i + 1] = blk0[i] + plans[i].nblocks;
    auto emit = [&](std::vector<TaskRef> &refs, bool fwd, bool wide, std::vector<int> &out) {
        std::stable_sort(refs.begin(), refs.end(), [](const TaskRef &x, const TaskRef &y) { return x.level < y.level; });
        for (const TaskRef &r : refs) {
            const BlockPlan &Q = plans[r.sys];
            const std::vector<int> &src = wide ? (fwd ? Q.fwd_task_wide : Q.bwd_task_wide) : (fwd ? Q.fwd_task : Q.bwd_task);
            const int *t = &src[(size_t)kTaskInts * r.idx];
            // block and split-row base move by the system's offsets
            const int rec[kTaskInts] = {t[0] + blk0[r.sys], t[1], t[2], t[3], t[4], t[5],
                                        t[7] > 0 ? t[6] + s->sys_split0[r.sys] : 0, t[7]};
            out.insert(out.end(), rec, rec + kTaskInts);
        }
    };
    // Batches of 8 or more systems plan without wide lists (each system's
    // copy is its narrow list): a wide list splits a system's top levels for
    // that system's own parallelism, which such a batch already has
    // (measured on the 256-system batch). The merged wide list then follows
    // the narrow order.
    emit(ftasks, true, false, G.fwd_task);
    emit(btasks, false, false, G.bwd_task);
    if (ns >= 8) {
        G.fwd_task_wide = G.fwd_task;
        G.bwd_task_wide = G.bwd_task;
    } else {
        emit(ftasks_w, true, true, G.fwd_task_wide);
        emit(btasks_w, false, true, G.bwd_task_wide);
    }
    plans.clear();
    HIP_TRY(hipMemcpy(s->perm, gperm.data(), (size_t)s->n * sizeof(int), hipMemcpyHostToDevice));
    int rc = upload_plan(s, G, &blk_sys);
    if (rc != CHOLESPY_OK) return rc;
    dev_free(s->blk_sys);
    HIP_TRY(upload(&s->blk_sys, blk_sys.data(), blk_sys.size()));
    dev_free(s->sys_k);
    dev_free(s->sys_vbase);
    dev_free(s->sys_pbase);
    dev_free(s->sys_sbase);
    HIP_TRY(dev_alloc(&s->sys_k, (size_t)ns));
    HIP_TRY(dev_alloc(&s->sys_vbase, (size_t)ns));
    HIP_TRY(dev_alloc(&s->sys_pbase, (size_t)ns));
    HIP_TRY(dev_alloc(&s->sys_sbase, (size_t)ns));
    s->last_k.clear();
    s->pending.clear();
    s->uploaded = true;
    s->analysed[0] = s->analysed[1] = true;   // the block path needs no level analysis
    s->nnz = G.nnz_l;
    return CHOLESPY_OK;
}

namespace {

int batch_solve(cholespy_gpu_solver *s, int fbytes, const int *n_rhs, const void *b, void *x, void *stream) {
    if (!s || !s->is_batch) return fail(CHOLESPY_ERR_INVALID_ARGUMENT, "not a batch handle");
    if (s->fbytes != fbytes) return fail(CHOLESPY_ERR_INVALID_ARGUMENT, "batch precision does not match the call");
    if (!n_rhs) return fail(CHOLESPY_ERR_INVALID_ARGUMENT, "null n_rhs");
    if (!s->uploaded) return fail(CHOLESPY_ERR_STATE, "batch not finalised");
    const int ns = s->nsys;
    std::vector<int> ks(n_rhs, n_rhs + ns);
    size_t elems = 0, pe = 0, se = 0;
    for (int i = 0; i < ns; ++i) {
        if (ks[i] < 0 || ks[i] > kMaxRhs)
            return fail(CHOLESPY_ERR_INVALID_ARGUMENT, "The number of RHS should be less than 128.");
        elems += (size_t)s->sys_n[i] * ks[i];
        pe += (size_t)s->sys_slots[i] * ks[i];
        se += (size_t)s->sys_splits[i] * ks[i];
    }
    if (elems == 0) return CHOLESPY_OK;
    if (!b || !x) return fail(CHOLESPY_ERR_INVALID_ARGUMENT, "null b or x");
    int rc = report_give_up(s);
    if (rc != CHOLESPY_OK) return rc;
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (ks != s->last_k) {
        // Solves already queued on the caller's stream read the per-system
        // arrays and buffers about to be rewritten: let them finish first.
        HIP_TRY(hipStreamSynchronize(st));
        std::vector<long long> vb((size_t)ns), pb((size_t)ns), sb((size_t)ns);
        long long voff = 0, poff = 0, soff = 0;
        for (int i = 0; i < ns; ++i) {
            vb[i] = voff - (long long)s->sys_row0[i] * ks[i];
            pb[i] = poff - (long long)s->sys_slot0[i] * ks[i];
            sb[i] = soff - (long long)s->sys_split0[i] * ks[i];
            voff += (long long)s->sys_n[i] * ks[i];
            poff += (long long)s->sys_slots[i] * ks[i];
            soff += (long long)s->sys_splits[i] * ks[i];
        }
        HIP_TRY(hipMemcpy(s->sys_k, ks.data(), (size_t)ns * sizeof(int), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(s->sys_vbase, vb.data(), (size_t)ns * sizeof(long long), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(s->sys_pbase, pb.data(), (size_t)ns * sizeof(long long), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(s->sys_sbase, sb.data(), (size_t)ns * sizeof(long long), hipMemcpyHostToDevice));
        if (elems > s->work_elems) {
            retire(s, s->y);
            retire(s, s->z);
            retire(s, s->rbuf);
            HIP_TRY(dev_alloc(reinterpret_cast<char **>(&s->y), elems * (size_t)fbytes));
            HIP_TRY(dev_alloc(reinterpret_cast<char **>(&s->z), elems * (size_t)fbytes));
            HIP_TRY(dev_alloc(reinterpret_cast<char **>(&s->rbuf), elems * (size_t)fbytes));
            s->work_elems = elems;
        }
        if (pe > s->part_elems) {
            retire(s, s->part);
            HIP_TRY(dev_alloc(reinterpret_cast<char **>(&s->part), pe * (size_t)fbytes));
            s->part_elems = pe;
        }
        if (se > s->spart_elems) {
            retire(s, s->spart);
            HIP_TRY(dev_alloc(reinterpret_cast<char **>(&s->spart), se * (size_t)fbytes));
            s->spart_elems = se;
        }
        HIP_TRY(launch_fill_sentinel(s->y, elems, fbytes, st));
        HIP_TRY(launch_fill_sentinel(s->z, elems, fbytes, st));
        HIP_TRY(launch_fill_sentinel(s->rbuf, elems, fbytes, st));
        if (pe) HIP_TRY(launch_fill_sentinel(s->part, pe, fbytes, st));
        if (se) HIP_TRY(launch_fill_sentinel(s->spart, se, fbytes, st));
        s->last_k = ks;
        s->buffers_path = CHOLESPY_PATH_BLOCKS;
    }
    return solve_blocks(s, 0, b, x, st);
}

}  // namespace

int cholespy_batch_solve_f32(cholespy_gpu_solver *s, const int *n_rhs, const float *b, float *x, void *stream) {
    return batch_solve(s, 4, n_rhs, b, x, stream);
}

int cholespy_batch_solve_f64(cholespy_gpu_solver *s, const int *n_rhs, const double *b, double *x,
                             void *stream) {
    return batch_solve(s, 8, n_rhs, b, x, stream);
}

int cholespy_solver_set_path(cholespy_gpu_solver *s, int path) {
    if (!s) return fail(CHOLESPY_ERR_INVALID_ARGUMENT, "null solver");
    if (path != CHOLESPY_PATH_ROWS && path != CHOLESPY_PATH_BLOCKS)
        return fail(CHOLESPY_ERR_INVALID_ARGUMENT, "unknown solve path");
    s->path = path;
    return CHOLESPY_OK;
}

int cholespy_solver_plan_info(const cholespy_gpu_solver *s, int64_t *info, int count) {
    if (!s || !info) return fail(CHOLESPY_ERR_INVALID_ARGUMENT, "null argument");
    const DeviceBlockPlan &d = s->plan;
    const int64_t vals[] = {d.nblocks,     d.m_entries,   d.levels_fwd, d.levels_bwd,       d.n_fwd_tasks,
                            d.n_bwd_tasks, d.nslots,      d.max_rows,   d.n_fwd_tasks_wide, d.n_bwd_tasks_wide};
    for (int i = 0; i < count && i < (int)(sizeof(vals) / sizeof(vals[0])); ++i) info[i] = vals[i];
    return CHOLESPY_OK;
}

int cholespy_solver_plan_arrays(const cholespy_gpu_solver *s, int which, int32_t *out, int64_t count,
                                int64_t *size) {
    if (!s || which < 0 || which > 7 || (count > 0 && !out)) return fail(CHOLESPY_ERR_INVALID_ARGUMENT, "bad argument");
    const std::vector<int> &v = s->plan_copy[which];
    if (size) *size = (int64_t)v.size();
    const int64_t nc = std::min<int64_t>(count, (int64_t)v.size());
    for (int64_t i = 0; i < nc; ++i) out[i] = v[(size_t)i];
    return CHOLESPY_OK;
}

int cholespy_solver_solve_f32(cholespy_gpu_solver *s, int n_rhs, const float *b, float *x, void *stream) {
    return solve_impl(s, 4, n_rhs, b, x, stream);
}

int cholespy_solver_solve_f64(cholespy_gpu_solver *s, int n_rhs, const double *b, double *x,
                              void *stream) {
    return solve_impl(s, 8, n_rhs, b, x, stream);
}

int cholespy_solver_solve_noalloc(cholespy_gpu_solver *s, int n_rhs, const void *b, void *x, void *stream) {
    if (!s) return fail(CHOLESPY_ERR_INVALID_ARGUMENT, "null solver");
    if (s->is_batch) return fail(CHOLESPY_ERR_INVALID_ARGUMENT, "batch handles solve with cholespy_batch_solve_*");
    if (n_rhs > kMaxRhs) return fail(CHOLESPY_ERR_INVALID_ARGUMENT, "The number of RHS should be less than 128.");
    if (n_rhs < 0) return fail(CHOLESPY_ERR_INVALID_ARGUMENT, "negative number of RHS");
    if (!s->uploaded) return fail(CHOLESPY_ERR_STATE, "solve called before the factor was uploaded");
    if (s->path != CHOLESPY_PATH_BLOCKS)
        return fail(CHOLESPY_ERR_STATE, "the no-allocation solve runs the block schedule only");
    if (n_rhs == 0 || s->n == 0) return CHOLESPY_OK;
    if (!b || !x) return fail(CHOLESPY_ERR_INVALID_ARGUMENT, "null b or x");
    int rc = report_give_up(s);
    if (rc != CHOLESPY_OK) return rc;
    // Layout in the scratch array: y, z, rbuf (n x k each), partial slots,
    // split rows, each region 256-byte aligned.
    const size_t fb = (size_t)s->fbytes, k = (size_t)n_rhs;
    const auto region = [](size_t bytes) { return (bytes + 255) & ~size_t(255); };
    const size_t vec = region((size_t)s->n * k * fb), part = region((size_t)s->plan.nslots * k * fb),
                 spart = region((size_t)s->plan.nsplit_rows * k * fb);
    const size_t need = 3 * vec + part + spart;
    if (need > kScratchBytes)
        return fail(CHOLESPY_ERR_INVALID_ARGUMENT,
                    "the no-allocation solve needs " + std::to_string(need) + " bytes of work memory for this system "
                    "and n_rhs; its fixed array holds " + std::to_string(kScratchBytes) +
                    " (use cholespy_solver_solve_f32/f64)");
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(CHOLESPY_ERR_INVALID_ARGUMENT, "device index out of range");
    hipStream_t st = static_cast<hipStream_t>(stream);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        // Once per device, with that device current: the array's address there.
        void *p = nullptr;
        HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(CHOLESPY_ERR_HIP, "hipGetSymbolAddress returned a null address for the scratch array");
        g_scratch_ptr[dev] = static_cast<unsigned char *>(p);
    }
    unsigned char *base = g_scratch_ptr[dev];
    const WorkSet w{base, base + vec, base + 3 * vec, base + 3 * vec + part, base + 2 * vec};
    // The array keeps the last call's contents. The kernels leave every
    // buffer they wait on in the sentinel state, so a call with the same
    // layout (handle, n_rhs) as the last one on this device starts as is;
    // any other refills its regions with fill kernels first.
    const uint64_t layout = (s->uid << 16) | (k << 4) | fb;
    if (g_scratch_layout[dev] != layout) {
        const size_t ne = (size_t)s->n * k;
        HIP_TRY(launch_fill_sentinel(w.y, ne, s->fbytes, st));
        HIP_TRY(launch_fill_sentinel(w.z, ne, s->fbytes, st));
        HIP_TRY(launch_fill_sentinel(w.rbuf, ne, s->fbytes, st));
        if (s->plan.nslots) HIP_TRY(launch_fill_sentinel(w.part, (size_t)s->plan.nslots * k, s->fbytes, st));
        if (s->plan.nsplit_rows)
            HIP_TRY(launch_fill_sentinel(w.spart, (size_t)s->plan.nsplit_rows * k, s->fbytes, st));
        g_scratch_layout[dev] = layout;
    }
    rc = solve_blocks(s, n_rhs, b, x, st, 0, -1, &w);
    if (rc != CHOLESPY_OK) g_scratch_layout[dev] = 0;
    return rc;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Autotuner. Per system (keyed by the factor's pattern, the precision, the
// right-hand-side class and the device) it times candidate configurations
// on the device and keeps the fastest: the schedule (row-per-wavefront
// passes or the block passes), the tile (vector-FMA or MFMA panel kernels,
// and the panel width), the block plan's shape (chain cap, split sizes,
// gather threshold, the wide list's level targets), the leaf phase and the
// ticket queues. Coordinate descent from the hand-set defaults: one knob at
// a time over its candidates, a change kept only when it is at least 1 %
// faster, so the result is never slower than the defaults as measured.
// Knobs set through CHOLESPY_* environment variables are not varied (they
// stay overrides). The winner is cached in a text file, one line per key.
namespace {

struct TuneConfig {
    int path = CHOLESPY_PATH_BLOCKS;
    int tile = 0;            // 0 = block_tile's rule
    int fused = 3;
    int queues = kMaxQueues;
    BlockPlanOptions opt;
};

uint64_t fnv1a(uint64_t h, const void *p, size_t bytes) {
    const unsigned char *c = static_cast<const unsigned char *>(p);
    for (size_t i = 0; i < bytes; ++i) {
        h ^= c[i];
        h *= 1099511628211ull;
    }
    return h;
}

std::string tune_key(const cholespy_gpu_solver *s, int k) {
    uint64_t h = 1469598103934665603ull;
    h = fnv1a(h, &s->n, sizeof(s->n));
    h = fnv1a(h, s->host_colptr.data(), s->host_colptr.size() * sizeof(int));
    h = fnv1a(h, s->host_rowidx.data(), s->host_rowidx.size() * sizeof(int));
    hipDeviceProp_t prop;
    std::string arch = "gpu";
    int cus = 0;
    if (hipGetDeviceProperties(&prop, s->device) == hipSuccess) {
        arch = prop.gcnArchName;
        cus = prop.multiProcessorCount;
        const size_t colon = arch.find(':');
        if (colon != std::string::npos) arch.resize(colon);
    }
    char buf[160];
    std::snprintf(buf, sizeof buf, "%016llx-n%d-f%d-t%d-%s-cu%d", (unsigned long long)h, s->n, s->fbytes * 8,
                  block_tile(s->fbytes, k), arch.c_str(), cus);
    return buf;
}

// Field order of a cache line after the key.
std::string format_config(const TuneConfig &c) {
    char buf[320];
    std::snprintf(buf, sizeof buf,
                  "path=%d tile=%d fused=%d queues=%d max_block=%d split_entries=%d max_split=%d split_gather=%d "
                  "wide_level_tasks=%d wide_min_split_entries=%d wide_fwd_rows_big=%d leaf_max=%d "
                  "wide_fine_bwd_cols_big=%d",
                  c.path, c.tile, c.fused, c.queues, c.opt.max_block, c.opt.split_entries, c.opt.max_split,
                  c.opt.split_gather, c.opt.wide_level_tasks, c.opt.wide_min_split_entries, c.opt.wide_fwd_rows_big,
                  c.opt.leaf_max, c.opt.wide_fine_bwd_cols_big);
    return buf;
}

bool parse_config(const std::string &text, TuneConfig &c) {
    int v[13];
    if (std::sscanf(text.c_str(),
                    "path=%d tile=%d fused=%d queues=%d max_block=%d split_entries=%d max_split=%d split_gather=%d "
                    "wide_level_tasks=%d wide_min_split_entries=%d wide_fwd_rows_big=%d leaf_max=%d "
                    "wide_fine_bwd_cols_big=%d",
                    &v[0], &v[1], &v[2], &v[3], &v[4], &v[5], &v[6], &v[7], &v[8], &v[9], &v[10], &v[11],
                    &v[12]) != 13)
        return false;
    if (v[11] < 1 || v[11] > 4096 || v[12] < 1 || v[12] > 64) return false;
    if ((v[0] != CHOLESPY_PATH_ROWS && v[0] != CHOLESPY_PATH_BLOCKS) || v[2] < 0 || v[2] > 3 || v[3] < 1 ||
        v[3] > kMaxQueues || v[4] < 0 || v[5] < 0 || v[6] < 1 || v[6] > 16 || v[7] < 1 || v[8] < 0 ||
        v[9] < 1 || v[10] < 1 || v[10] > 64)
        return false;
    c.path = v[0];
    c.tile = v[1];
    c.fused = v[2];
    c.queues = v[3];
    c.opt.max_block = v[4];
    c.opt.split_entries = v[5];
    c.opt.max_split = v[6];
    c.opt.split_gather = v[7];
    c.opt.wide_level_tasks = v[8];
    c.opt.wide_min_split_entries = v[9];
    c.opt.wide_fwd_rows_big = v[10];
    c.opt.leaf_max = v[11];
    c.opt.wide_fine_bwd_cols_big = v[12];
    return true;
}

bool same_plan(const BlockPlanOptions &a, const BlockPlanOptions &b) {
    return a.max_block == b.max_block && a.split_entries == b.split_entries && a.max_split == b.max_split &&
           a.split_gather == b.split_gather && a.wide_level_tasks == b.wide_level_tasks &&
           a.wide_min_split_entries == b.wide_min_split_entries && a.wide_fwd_rows_big == b.wide_fwd_rows_big &&
           a.leaf_max == b.leaf_max && a.wide_fine_bwd_cols_big == b.wide_fine_bwd_cols_big;
}

// Rebuilds and uploads the block plan for base options `base` (environment
// overrides on top) when they differ from the current ones.
int replan(cholespy_gpu_solver *s, const double *val, const BlockPlanOptions &base) {
    if (same_plan(base, s->plan_base) && s->uploaded) return CHOLESPY_OK;
    HIP_TRY(hipDeviceSynchronize());
    BlockPlan P;
    try {
        P = build_block_plan(s->n, s->host_colptr.data(), s->host_rowidx.data(), val, plan_options(base));
    } catch (const std::exception &ex) {
        return fail(CHOLESPY_ERR_INVALID_ARGUMENT, ex.what());
    }
    int rc = upload_plan(s, P);
    if (rc != CHOLESPY_OK) return rc;
    std::vector<double>().swap(P.mfwd);
    std::vector<double>().swap(P.mbwd);
    s->host_plan = std::move(P);
    s->plan_base = base;
    s->buffers_path = -1;
    return CHOLESPY_OK;
}

int apply_config(cholespy_gpu_solver *s, const double *val, int k, const TuneConfig &c) {
    int rc = replan(s, val, c.opt);
    if (rc != CHOLESPY_OK) return rc;
    if (c.path == CHOLESPY_PATH_ROWS && (!s->analysed[0] || !s->analysed[1])) {
        rc = cholespy_solver_analyze(s, 1, nullptr);
        if (rc == CHOLESPY_OK) rc = cholespy_solver_analyze(s, 0, nullptr);
        if (rc != CHOLESPY_OK) return rc;
    }
    s->path = c.path;
    s->tile_for_class[tile_slot(block_tile(s->fbytes, k))] = c.tile;
    if (!std::getenv("CHOLESPY_FUSED_LEAVES")) s->fused_leaves = c.fused;
    if (!std::getenv("CHOLESPY_BLOCK_QUEUES")) s->queues[4].nq = s->queues[5].nq = c.queues;
    for (auto &row : s->grid_blocks)
        for (int &g : row) g = 0;
    return CHOLESPY_OK;
}

// Milliseconds per solve of k zero right-hand sides: the fastest of five
// batches of solves (each batch about a millisecond long or more).
int time_config(cholespy_gpu_solver *s, int k, void *b, void *x, double &ms) {
    auto solve = [&]() { return solve_impl(s, s->fbytes, k, b, x, nullptr); };
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    int rc = CHOLESPY_OK;
    float t = 0.f;
    for (int i = 0; i < 2 && rc == CHOLESPY_OK; ++i) rc = solve();
    if (rc == CHOLESPY_OK && hipEventRecord(e0, nullptr) == hipSuccess) {
        rc = solve();
        if (rc == CHOLESPY_OK && hipEventRecord(e1, nullptr) == hipSuccess && hipEventSynchronize(e1) == hipSuccess)
            (void)hipEventElapsedTime(&t, e0, e1);
    }
    const int batch = std::max(1, std::min(50, (int)std::ceil(1.0 / std::max(1e-3, (double)t))));
    ms = 1e30;
    for (int r = 0; r < 5 && rc == CHOLESPY_OK; ++r) {
        if (hipEventRecord(e0, nullptr) != hipSuccess) break;
        for (int i = 0; i < batch && rc == CHOLESPY_OK; ++i) rc = solve();
        if (rc != CHOLESPY_OK || hipEventRecord(e1, nullptr) != hipSuccess || hipEventSynchronize(e1) != hipSuccess)
            break;
        float el = 0.f;
        if (hipEventElapsedTime(&el, e0, e1) == hipSuccess) ms = std::min(ms, (double)el / batch);
        if (t > 20.f) break;   // a slow candidate (the row schedule of a large system): one batch
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc == CHOLESPY_OK) rc = report_give_up(s);
    return rc;
}

std::string default_cache_path() {
    if (const char *p = std::getenv("CHOLESPY_TUNE_CACHE")) return p;
    std::string dir;
    if (const char *x = std::getenv("XDG_CACHE_HOME")) dir = x;
    else if (const char *h = std::getenv("HOME")) dir = std::string(h) + "/.cache";
    else dir = "/tmp";
    return dir + "/cholespy_amd_autotune.txt";
}

bool cache_lookup(const std::string &path, const std::string &key, TuneConfig &c, double &ms) {
    FILE *f = std::fopen(path.c_str(), "r");
    if (!f) return false;
    char line[1024];
    bool found = false;
    while (std::fgets(line, sizeof line, f)) {
        std::string l(line);
        const size_t sp = l.find(' ');
        if (sp == std::string::npos || l.compare(0, sp, key) != 0) continue;
        const size_t msp = l.rfind(" ms=");
        TuneConfig t;
        if (msp == std::string::npos || !parse_config(l.substr(sp + 1, msp - sp - 1), t)) continue;
        c = t;
        ms = std::atof(l.c_str() + msp + 4);
        found = true;   // the last matching line wins
    }
    std::fclose(f);
    return found;
}

void cache_store(const std::string &path, const std::string &key, const TuneConfig &c, double ms) {
    const size_t slash = path.rfind('/');
    if (slash != std::string::npos && slash > 0) (void)::mkdir(path.substr(0, slash).c_str(), 0755);
    FILE *f = std::fopen(path.c_str(), "a");
    if (!f) return;
    std::fprintf(f, "%s %s ms=%.6f\n", key.c_str(), format_config(c).c_str(), ms);
    std::fclose(f);
}

int longest_chain(const cholespy_gpu_solver *s) {
    int longest = s->n > 0 ? 1 : 0;
    for (int j = 0, run = 1; j + 1 < s->n; ++j) {
        const bool child = s->host_colptr[j + 1] - s->host_colptr[j] > 1 && s->host_rowidx[s->host_colptr[j] + 1] == j + 1;
        run = child ? run + 1 : 1;
        longest = std::max(longest, run);
    }
    return longest;
}

}  // namespace

extern "C" int cholespy_solver_autotune(cholespy_gpu_solver *s, const double *val, int64_t nnz, int n_rhs,
                                        const char *cache_path, int flags, char *report, size_t report_len) {
    if (!s || !val) return fail(CHOLESPY_ERR_INVALID_ARGUMENT, "null argument");
    if (s->is_batch) return fail(CHOLESPY_ERR_INVALID_ARGUMENT, "batches are not tuned");
    if (!s->uploaded) return fail(CHOLESPY_ERR_STATE, "autotune called before upload_factor");
    if (nnz != s->nnz) return fail(CHOLESPY_ERR_INVALID_ARGUMENT, "nnz does not match the uploaded factor");
    if (n_rhs < 1 || n_rhs > kMaxRhs) return fail(CHOLESPY_ERR_INVALID_ARGUMENT, "n_rhs must be in [1, 128]");
    int rc = report_give_up(s);
    if (rc != CHOLESPY_OK) return rc;
    const int k = n_rhs;
    const std::string key = tune_key(s, k);
    const std::string path = cache_path && *cache_path ? cache_path : default_cache_path();
    TuneConfig best;
    best.opt = s->plan_base;
    best.path = s->path;
    best.fused = s->fused_leaves;
    best.queues = s->queues[4].nq;
    double best_ms = 0;
    auto finish = [&](const char *how, int evaluated, double default_ms) {
        std::string r = std::string(how) + " " + format_config(best);
        char buf[160];
        std::snprintf(buf, sizeof buf, " ms=%.5f default_ms=%.5f candidates=%d key=%s", best_ms, default_ms,
                      evaluated, key.c_str());
        r += buf;
        s->tune_report = r;
        if (report && report_len) {
            std::strncpy(report, r.c_str(), report_len - 1);
            report[report_len - 1] = 0;
        }
        return CHOLESPY_OK;
    };
    if (!(flags & 1)) {
        TuneConfig cached;
        double ms = 0;
        if (cache_lookup(path, key, cached, ms)) {
            rc = apply_config(s, val, k, cached);
            if (rc != CHOLESPY_OK) return rc;
            best = cached;
            best_ms = ms;
            return finish("cached", 0, 0.0);
        }
    }
    // Work vectors for the timing solves (zeros).
    const size_t bytes = (size_t)s->n * (size_t)k * (size_t)s->fbytes;
    void *b = nullptr, *x = nullptr;
    HIP_TRY(hipMalloc(&b, std::max<size_t>(bytes, 1)));
    if (hipMalloc(&x, std::max<size_t>(bytes, 1)) != hipSuccess) {
        (void)hipFree(b);
        return fail(CHOLESPY_ERR_HIP, "autotune: allocation of the timing vectors failed");
    }
    (void)hipMemset(b, 0, bytes);
    (void)hipMemset(x, 0, bytes);
    int evaluated = 0;
    auto eval = [&](const TuneConfig &c, double &ms) {
        int r = apply_config(s, val, k, c);
        if (r == CHOLESPY_OK) r = time_config(s, k, b, x, ms);
        if (r == CHOLESPY_ERR_TIMEOUT) {   // a wait gave up: never choose this one
            ms = 1e30;
            r = CHOLESPY_OK;
        }
        ++evaluated;
        return r;
    };
    rc = eval(best, best_ms);
    const double default_ms = best_ms;
    // One knob at a time; a candidate replaces the best when >= 0.5 % faster
    // (repeated timings of one schedule agree to ~0.2 %).
    auto try_values = [&](const char *env, const std::vector<int> &values, auto set) {
        if (rc != CHOLESPY_OK || (env && std::getenv(env))) return;
        for (int v : values) {
            TuneConfig c = best;
            set(c, v);
            if (format_config(c) == format_config(best)) continue;
            double ms = 0;
            rc = eval(c, ms);
            if (rc != CHOLESPY_OK) return;
            if (ms < 0.995 * best_ms) {
                best = c;
                best_ms = ms;
            }
        }
    };
    const bool narrow = block_tile(s->fbytes, k) == kNarrowTile;
    // The schedule: level-scheduled rows against the block passes. The row
    // schedule walks the elimination tree one row per level and is only a
    // candidate for small factors: on large ones a row of a dense separator
    // takes milliseconds and the whole chain seconds.
    if (s->nnz <= (int64_t)1 << 20)
        try_values(nullptr, {CHOLESPY_PATH_ROWS}, [](TuneConfig &c, int v) { c.path = v; });
    if (best.path == CHOLESPY_PATH_BLOCKS) {
        // Tile: vector-FMA kernels or MFMA panels, and the panel width.
        std::vector<int> tiles;
        if (narrow) tiles = {32};
        else if (k <= 32) tiles = {64};
        else if (k <= 64) tiles = {128};
        try_values(nullptr, tiles, [](TuneConfig &c, int v) { c.tile = v; });
        const bool wide = !(narrow && best.tile == 0);
        // Chain cap: powers of two from 1024 up to the longest chain (4096 at most).
        std::vector<int> caps;
        const int chain = longest_chain(s);
        for (int cap = 1024; cap <= 4096 && cap <= std::max(1024, chain); cap *= 2) caps.push_back(cap);
        try_values("CHOLESPY_MAX_BLOCK", caps, [](TuneConfig &c, int v) { c.opt.max_block = v; });
        // Leaf subtrees: larger ones can save a block level.
        try_values("CHOLESPY_LEAF_MAX", {48, 56, 96}, [](TuneConfig &c, int v) { c.opt.leaf_max = v; });
        try_values("CHOLESPY_SPLIT_ENTRIES", {2048, 8192}, [](TuneConfig &c, int v) { c.opt.split_entries = v; });
        try_values("CHOLESPY_MAX_SPLIT", {4, 16}, [](TuneConfig &c, int v) { c.opt.max_split = v; });
        if (!wide) {
            try_values("CHOLESPY_SPLIT_GATHER", {33, 129}, [](TuneConfig &c, int v) { c.opt.split_gather = v; });
            try_values("CHOLESPY_FUSED_LEAVES", {0, 1, 2}, [](TuneConfig &c, int v) { c.fused = v; });
        } else {
            try_values("CHOLESPY_WIDE_LEVEL_TASKS", {256, 1024},
                       [](TuneConfig &c, int v) { c.opt.wide_level_tasks = v; });
            try_values("CHOLESPY_WIDE_MIN_SPLIT_ENTRIES", {256, 1024},
                       [](TuneConfig &c, int v) { c.opt.wide_min_split_entries = v; });
            try_values("CHOLESPY_WIDE_FWD_ROWS_BIG", {32, 64}, [](TuneConfig &c, int v) { c.opt.wide_fwd_rows_big = v; });
            try_values("CHOLESPY_WIDE_FINE_BWD_COLS_BIG", {32, 64},
                       [](TuneConfig &c, int v) { c.opt.wide_fine_bwd_cols_big = v; });
        }
        try_values("CHOLESPY_BLOCK_QUEUES", {8}, [](TuneConfig &c, int v) { c.queues = v; });
    }
    (void)hipDeviceSynchronize();
    (void)hipFree(b);
    (void)hipFree(x);
    if (rc != CHOLESPY_OK) return rc;
    rc = apply_config(s, val, k, best);
    if (rc != CHOLESPY_OK) return rc;
    if (!(flags & 2)) cache_store(path, key, best, best_ms);
    return finish("tuned", evaluated, default_ms);
}

extern "C" int cholespy_probe_handoff(double *results, int count) {
    if (!results || count < 6) return fail(CHOLESPY_ERR_INVALID_ARGUMENT, "results needs 6 doubles");
    int rc = cholespy_init();
    if (rc != CHOLESPY_OK) return rc;
    HIP_TRY(probe_handoff(results));
    return CHOLESPY_OK;
}

extern "C" int cholespy_solver_tune_report(const cholespy_gpu_solver *s, char *report, size_t report_len) {
    if (!s || !report || !report_len) return fail(CHOLESPY_ERR_INVALID_ARGUMENT, "null argument");
    std::strncpy(report, s->tune_report.c_str(), report_len - 1);
    report[report_len - 1] = 0;
    return CHOLESPY_OK;
}
