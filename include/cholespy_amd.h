/*
 * cholespy_amd C ABI: the device side of cholespy's GPU solve phase on
 * MI355X (gfx950), as one shared library (libcholespy_amd.so) holding the
 * hand-written HIP kernels and the entry points below.
 *
 * It replaces, in the reference:
 *   src/cuda_driver.h:16-61 / src/cuda_driver.cpp:58-149  (dlopen of libcuda,
 *       PTX module loading, 6 CUfunction handles, context push/pop), and
 *   the device half of CholeskySolver<Float> (src/cholesky_solver.h:60-98,
 *       src/cholesky_solver.cpp:180-188, 244-432, 465-486),
 * with kernels compiled into the library by hipcc --offload-arch=gfx950 (no
 * run-time module loading). Only plain C types cross this boundary: device
 * buffers are passed as pointers, streams as an opaque hipStream_t value
 * (NULL = the null stream, which is what the reference launches on).
 *
 * Conventions
 *   - Every function returns a cholespy_status; on failure a message is
 *     available from cholespy_last_error() (per calling thread).
 *   - A solver handle owns every device buffer it allocates. b is read only;
 *     x is overwritten. Calls on one handle must not overlap (the handle's
 *     scratch buffers are shared by its solves, as in the reference).
 *   - Dense right-hand sides are row-major (n x k), element (i, j) at
 *     [i * k + j], exactly as kernels/kernels.cu:87,114,120 index them.
 */
#ifndef CHOLESPY_AMD_H
#define CHOLESPY_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum cholespy_status {
    CHOLESPY_OK = 0,
    CHOLESPY_ERR_INVALID_ARGUMENT = 1, /* bad sizes, null pointers, k > 128 */
    CHOLESPY_ERR_HIP = 2,              /* a HIP runtime call failed */
    CHOLESPY_ERR_NO_DEVICE = 3,        /* no GPU visible */
    CHOLESPY_ERR_TIMEOUT = 4,          /* a kernel's bounded wait expired */
    CHOLESPY_ERR_STATE = 5             /* call out of order (e.g. solve before analyze) */
} cholespy_status;

typedef struct cholespy_gpu_solver cholespy_gpu_solver; /* opaque */

/* Solve schedules (cholespy_solver_set_path). */
enum {
    /* Level-scheduled row-per-wavefront passes over the row-compressed L and
     * L' in level order: the reference's algorithm (kernels/kernels.cu:64-202)
     * laid out for gfx950. The "plain" baseline. */
    CHOLESPY_PATH_ROWS = 0,
    /* Supernodal partitioned-inverse passes: one dense product per block of
     * the elimination tree (default). */
    CHOLESPY_PATH_BLOCKS = 1
};

/* Lazy, idempotent device initialisation. Replaces init_cuda()
 * (src/cuda_driver.cpp:58-136). Selects the calling thread's current HIP
 * device and records its compute-unit count for persistent launches. */
int cholespy_init(void);

/* Releases library-wide state. Replaces shutdown_cuda()
 * (src/cuda_driver.cpp:138-149); called when the Python module unloads
 * (src/main.cpp:127-129). Safe to call more than once. */
int cholespy_shutdown(void);

/* Message describing the last failure on this thread ("" if none).
 * Replaces cuda_check_impl()'s stderr print (src/cuda_driver.cpp:48-56). */
const char *cholespy_last_error(void);

/* Number of visible GPUs (0 when none); does not initialise a context. */
int cholespy_device_count(int *count);

/* Synchronous device-to-host copy, used by the binding for GPU inputs of the
 * constructor (src/main.cpp:55-61). */
int cholespy_copy_to_host(void *dst, const void *device_src, size_t bytes);

/* Creates a solver for an n x n factor whose device arithmetic uses
 * float_bytes = 4 (float, CholeskySolverF) or 8 (double, CholeskySolverD).
 * Replaces the GPU allocations of the constructor
 * (src/cholesky_solver.cpp:180-188). */
int cholespy_solver_create(int n, int float_bytes, cholespy_gpu_solver **out);

/* Frees every device buffer of the handle. Replaces ~CholeskySolver
 * (src/cholesky_solver.cpp:465-486). NULL is accepted. */
int cholespy_solver_destroy(cholespy_gpu_solver *solver);

/* Uploads the factor L (L L' = P A P', perm[k] = row of A placed at k) given
 * in compressed-column form with rows ascending and the diagonal first in
 * each column, values in double. The library derives the row-compressed
 * copy of L (diagonal last) for the forward pass and uses the column form as
 * the row-compressed L' for the backward pass, casting values to the
 * solver's precision. Replaces the GPU tail of factorize() and the uploads
 * of analyze_cuda() (src/cholesky_solver.cpp:244-288, 300-313). */
int cholespy_solver_upload_factor(cholespy_gpu_solver *solver, const int *perm,
                                  const int *colptr, const int *rowidx,
                                  const double *val, int64_t nnz);

/* Numeric refactorisation: replaces the values of the uploaded factor by
 * val (nnz doubles, same pattern and order as the colptr / rowidx given to
 * cholespy_solver_upload_factor; perm is kept). Keeps the row form's
 * pattern, the level analysis and the block plan's structure, recomputing
 * only the dense block inverses. Waits for all work on the device before
 * overwriting values. Single-system handles only. The reference has no
 * counterpart: its benchmark builds a new CholeskySolver for every new
 * matrix with the same pattern (tests/Benchmark.ipynb:238-254), redoing
 * the ordering, the symbolic factor and analyze_cuda
 * (src/cholesky_solver.cpp:151-198, 299-376). */
int cholespy_solver_refactor(cholespy_gpu_solver *solver, const double *val, int64_t nnz);

/* Device-side input normalisation. ii, jj, x are device pointers in the
 * layout of `type` (0 = CSC: ii column pointer; 1 = CSR: ii row pointer;
 * 2 = COO: ii rows, jj columns), indices of index_bytes = 4 or 8, values of
 * value_bytes = 4 (float) or 8 (double). Writes to host memory the lower
 * triangle (row >= col) in compressed columns, rows ascending and
 * duplicates summed in input order: colptr[n + 1], rowidx and val (room for
 * nnz entries; *nnz_out are written), and, when entry_dst is not NULL,
 * entry_dst[nnz] = the output position each input entry is summed into, or
 * -1 above the diagonal. The result is bit-identical to the reference's host
 * path, which copies the arrays to the host first (src/main.cpp:55-67) and
 * sorts and sums there (src/cholesky_solver.cpp:8-149; CHOLMOD stype = -1 at
 * 211-220). Invalid indices or pointers return CHOLESPY_ERR_INVALID_ARGUMENT
 * with the host path's message. Synchronous. */
int cholespy_normalize_input(int n, int64_t nnz, const void *ii, const void *jj, int index_bytes,
                             const void *x, int value_bytes, int type, int *colptr, int *rowidx,
                             double *val, int64_t *entry_dst, int64_t *nnz_out);

/* Level-set analysis of L (lower = 1) or L' (lower = 0) on the GPU followed
 * by the stable counting sort of rows by level. Replaces analyze_cuda()
 * (src/cholesky_solver.cpp:299-376) and the kernels analysis_lower /
 * analysis_upper (kernels/kernels.cu:8-60). Writes the number of levels
 * (max level + 1) to *n_levels when it is not NULL. */
int cholespy_solver_analyze(cholespy_gpu_solver *solver, int lower, int *n_levels);

/* Copies the analysis results to the host: levels[n] = rows sorted by level
 * (stable in row index; the `levels` array of analyze_cuda) and, when not
 * NULL, row_level[n] = level of each row. */
int cholespy_solver_get_levels(const cholespy_gpu_solver *solver, int lower, int *levels,
                               int *row_level);

/* x := A^{-1} b for n_rhs right-hand sides (1 <= n_rhs <= 128), b and x
 * device pointers to row-major (n x n_rhs) arrays of the solver's
 * precision. Asynchronous on `stream`. Replaces solve_cuda() and
 * launch_kernel() (src/cholesky_solver.cpp:378-432) and the kernels
 * solve_{lower,upper}_{float,double} (kernels/kernels.cu:64-218). Returns
 * CHOLESPY_ERR_INVALID_ARGUMENT for n_rhs > 128, as the reference throws
 * "The number of RHS should be less than 128." (src/cholesky_solver.cpp:419).
 * The default block schedule needs only upload_factor; the row schedule
 * (cholespy_solver_set_path) also needs both passes analysed
 * (CHOLESPY_ERR_STATE otherwise). */
int cholespy_solver_solve_f32(cholespy_gpu_solver *solver, int n_rhs, const float *b,
                              float *x, void *stream);
int cholespy_solver_solve_f64(cholespy_gpu_solver *solver, int n_rhs, const double *b,
                              double *x, void *stream);

/* x := A^{-1} b like cholespy_solver_solve_f32/f64 (the handle's precision:
 * b and x hold floats for a float_bytes = 4 handle, doubles for 8), with the
 * same kernels and the same results, bit for bit, but no device allocation
 * of any kind: the work vectors (y, z, gathered right-hand sides, partial
 * sums, split results) live in one fixed 64 MiB array in the library's
 * device code object, one copy per GPU, held until the process ends. No
 * counterpart in the reference; it is for callers that cannot allocate
 * (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc). Block schedule only, single-system handles only. Returns
 * CHOLESPY_ERR_INVALID_ARGUMENT when the system needs more than the array
 * holds (3 n n_rhs + (partial slots + split rows) n_rhs elements, each
 * region rounded up to 256 bytes; icosphere(7) fits up to n_rhs = 12 in
 * FP64). Every call on one GPU shares that GPU's array and finds the last
 * call's contents there: a call whose handle or n_rhs differs from the last
 * one on the device first refills the array with fill kernels on `stream`.
 * Calls that overlap on one GPU (concurrent streams, threads or processes
 * sharing the library) race on the array and must be serialised by the
 * caller. */
int cholespy_solver_solve_noalloc(cholespy_gpu_solver *solver, int n_rhs, const void *b, void *x,
                                  void *stream);

/* One triangular pass of a solve: pass 0 = forward (reads b, fills the
 * handle's intermediate y), pass 1 = backward (reads y, writes x). Running
 * pass 0 then pass 1 with the same n_rhs equals one solve. Replaces
 * launch_kernel(lower, x) (src/cholesky_solver.cpp:378-413); used to time
 * the two kernels separately. b is ignored by pass 1 and x by pass 0 (either
 * may then be NULL). Any order of calls is valid, as with the reference,
 * which clears its flags at every launch (src/cholesky_solver.cpp:380-382):
 * a forward pass (or a whole solve) that follows a forward pass with no
 * backward pass in between, and a backward pass with no forward pass before
 * it, first restore the handle's work buffers (extra fill kernels on
 * `stream`). A backward pass reads the y of the last forward pass. */
int cholespy_solver_solve_pass(cholespy_gpu_solver *solver, int pass, int n_rhs, const void *b,
                               void *x, void *stream);

/* Uploads A itself (its lower triangle, compressed columns, rows
 * ascending, FP64), as the refinement residual needs it. No counterpart in
 * the reference, whose GPU path keeps only L. */
int cholespy_solver_upload_matrix(cholespy_gpu_solver *solver, const int *colptr,
                                  const int *rowidx, const double *val, int64_t nnz);

/* Mixed-precision solve for a float32 handle: x (double, row-major n x
 * n_rhs, device) := A^{-1} b (double) by an FP32 block solve followed by
 * `iterations` FP64 refinement steps (r = b - A x in FP64, FP32 solve of r,
 * x += correction in FP64). Asynchronous on `stream`. No counterpart in the
 * reference (its CholeskySolverF solves in FP32 only). */
int cholespy_solver_solve_refined(cholespy_gpu_solver *solver, int n_rhs, const double *b,
                                  double *x, int iterations, void *stream);

/* The same, refined until the relative residual
 * max_j ||b_j - A x_j||_2 / ||b_j||_2 (FP64, on the device) is <= tol, or
 * after max_steps corrections. The residual norms come back to the host
 * before each correction (one synchronisation of `stream` per step, so this
 * variant cannot be captured in a graph). *steps (corrections applied) and
 * *residual (the last relative residual) are written when not NULL. No
 * counterpart in the reference. */
int cholespy_solver_solve_refined_tol(cholespy_gpu_solver *solver, int n_rhs, const double *b,
                                      double *x, double tol, int max_steps, void *stream, int *steps,
                                      double *residual);

/* Batches of independent systems solved by one launch per pass (no
 * counterpart in the reference, which solves one system per object).
 * create: `count` systems of sizes n[i]; set_factor: the factor of system
 * `index` (same form as cholespy_solver_upload_factor); finalize: plans all
 * systems (host threads), merges the plans and uploads them. solve: system i
 * has n_rhs[i] right-hand sides (0..128, host array); b and x are packed
 * device arrays, system i's row-major (n[i] x n_rhs[i]) block starting
 * after the blocks of systems 0..i-1. Destroy with cholespy_solver_destroy. */
int cholespy_batch_create(int count, const int *n, int float_bytes, cholespy_gpu_solver **out);
int cholespy_batch_set_factor(cholespy_gpu_solver *batch, int index, const int *perm,
                              const int *colptr, const int *rowidx, const double *val,
                              int64_t nnz);
int cholespy_batch_finalize(cholespy_gpu_solver *batch);
int cholespy_batch_solve_f32(cholespy_gpu_solver *batch, const int *n_rhs, const float *b,
                             float *x, void *stream);
int cholespy_batch_solve_f64(cholespy_gpu_solver *batch, const int *n_rhs, const double *b,
                             double *x, void *stream);

/* Selects the schedule used by later solves (default CHOLESPY_PATH_BLOCKS).
 * No counterpart in the reference, which has the row schedule only. */
int cholespy_solver_set_path(cholespy_gpu_solver *solver, int path);

/* Autotuner (the reference has no counterpart: it launches one fixed
 * shape, src/cholesky_solver.cpp:409-412). Times candidate configurations
 * of this handle for n_rhs right-hand sides on the device and keeps the
 * fastest: the schedule (CHOLESPY_PATH_ROWS or _BLOCKS), the tile
 * (vector-FMA or MFMA panel kernels and the panel width), the block plan's
 * shape (chain cap, split sizes, gather threshold, the wide list's level
 * targets; the plan is rebuilt from val, the values of L as given to
 * cholespy_solver_upload_factor), the leaf phase and the ticket queues.
 * Coordinate descent from the defaults, keeping a change only when it is
 * at least 1 % faster. Knobs set through CHOLESPY_* environment variables
 * are not varied. The winner is appended to a text cache (cache_path, or
 * $CHOLESPY_TUNE_CACHE, or ~/.cache/cholespy_amd_autotune.txt) under a key
 * of the factor's pattern, the precision, the right-hand-side class and the
 * device, and a later call with the same key applies the cached choice
 * without timing. flags: bit 0 = ignore the cache, bit 1 = do not write it.
 * report (if not NULL) receives a one-line summary. Synchronous; allocates
 * two n x n_rhs timing vectors. Single-system handles only. */
int cholespy_solver_autotune(cholespy_gpu_solver *solver, const double *val, int64_t nnz, int n_rhs,
                             const char *cache_path, int flags, char *report, size_t report_len);
/* Microbenchmark of the costs the block schedule is built around (no
 * counterpart in the reference; synchronous, allocates its own buffers):
 * results[0] one-way hand-off of a value between two workgroups through an
 * agent-scope relaxed store and loads of a sentinel word (us), results[1]
 * the same with release / acquire fences (us), results[2] one dependent
 * agent-scope load (ns), results[3] one 256-thread workgroup barrier (ns),
 * results[4] one ticket fetch-add on a shared address with 1024 workgroups
 * contending (ns, device-wide), results[5] one back-to-back empty launch
 * (us). A hand-off whose wait gave up reports -1. count >= 6. */
int cholespy_probe_handoff(double *results, int count);

/* The summary line of the handle's last autotune ("" if none). */
int cholespy_solver_tune_report(const cholespy_gpu_solver *solver, char *report, size_t report_len);

/* Block-plan statistics: info[0..count) = {blocks, dense entries of all
 * M_B, forward block levels, backward block levels, forward tasks, backward
 * tasks, partial slots, largest m + |R|, forward tasks and backward tasks of
 * the wide right-hand-side list}. No counterpart in the reference. */
int cholespy_solver_plan_info(const cholespy_gpu_solver *solver, int64_t *info, int count);

/* Copies of the block plan (diagnostics; no counterpart in the reference):
 * which = 0 block row starts (blocks + 1), 1 partial-slot ranges
 * (blocks + 1), 2 forward block levels (blocks), 3 backward block levels
 * (blocks), 4 forward tasks and 5 backward tasks (CHOLESPY_TASK_INTS ints
 * each: block, first, last, kind, lo, hi, split-row base, splits), 6 and 7
 * the forward and backward tasks of the wide right-hand-side list. Copies
 * min(count, size) ints to out; *size (if not null) receives the full size. */
#define CHOLESPY_TASK_INTS 8
int cholespy_solver_plan_arrays(const cholespy_gpu_solver *solver, int which, int32_t *out, int64_t count,
                                int64_t *size);

/* Gradient of a loss with respect to the caller's stored matrix entries
 * (rows[e], cols[e], e < nnz, device int32 arrays, indices in [0, n)) of an
 * n x n matrix A, for x = A^{-1} b, given
 * x and the adjoint lam = A^{-1} dL/dx (device, row-major n x n_rhs,
 * vec_bytes = 4 or 8): grad[e] = -sum_t (lam[i,t] x[j,t] + lam[j,t] x[i,t])
 * for i > j, -sum_t lam[i,t] x[i,t] for i == j, 0 for i < j (the strict
 * upper triangle is never read: CHOLMOD stype = -1,
 * src/cholesky_solver.cpp:211-220). grad has grad_bytes = 4 or 8 per
 * element. An entry with an index outside [0, n) is never read through; its
 * gradient is NaN (check the range first; the Python layer raises). No
 * counterpart in the reference (it has no backward). */
int cholespy_entry_gradient(int n, int64_t nnz, const int *rows, const int *cols, int n_rhs,
                            const void *lam, const void *x, int vec_bytes, void *grad,
                            int grad_bytes, void *stream);

/* Task timeline of the block passes (diagnostics; no counterpart in the
 * reference). With tracing on, every block-pass launch records per task
 * {start, end of the gather phase, end, workgroup} as 100 MHz wall-clock
 * ticks; get_trace copies min(words, 4 * tasks) words of the last forward
 * (forward = 1) or backward pass. */
int cholespy_solver_set_trace(cholespy_gpu_solver *solver, int enable);
int cholespy_solver_get_trace(const cholespy_gpu_solver *solver, int forward, uint64_t *out,
                              int64_t words);

/* Watchdog. Every wait of the solve and analysis kernels is bounded: a wait
 * that runs out of polls records CHOLESPY_ERR_TIMEOUT in the handle's status
 * words (host memory mapped into the device) and raises the handle's abort
 * word, and every wait that gives up returns NaN, so the results of that
 * call are NaN rather than finite and wrong. The next solve, refined solve,
 * pass or batch solve on the handle (or cholespy_solver_check) reports it:
 * it waits for the device, clears the words, restores the work buffers and
 * returns CHOLESPY_ERR_TIMEOUT without solving, and the call after it runs
 * normally. The reference prints driver errors and continues
 * (src/cuda_driver.cpp:48-56); its spin-waits have no bound.
 *
 * cholespy_solver_check synchronises the device and reports as above
 * (CHOLESPY_OK when nothing gave up). */
int cholespy_solver_check(cholespy_gpu_solver *solver);

/* Test hooks of the watchdog (no counterpart in the reference): set the
 * poll budget of every wait of the handle's later launches (0 = the default,
 * about 10 s), and raise the handle's abort word from the host as a give-up
 * would. Tests force the give-up path with these, so no kernel has to hang
 * for real. */
int cholespy_solver_set_spin_limit(cholespy_gpu_solver *solver, uint32_t limit);
int cholespy_solver_raise_abort(cholespy_gpu_solver *solver);

#ifdef __cplusplus
}
#endif

#endif /* CHOLESPY_AMD_H */
