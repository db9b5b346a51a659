// Block plan for the supernodal partitioned-inverse solve (host side).
//
// The postordered elimination tree of L is cut into contiguous index ranges
// ("blocks") B = [a, a + m). For each block the plan stores the dense matrix
//     M_B = [ L_BB^{-1} ; -L_RB L_BB^{-1} ]          ((m + |R_B|) x m)
// where R_B is the sorted set of rows below the block touched by its columns.
// With it both triangular passes become one dense product per block:
//   forward   r_B = (P b)_B + sum of the partials other blocks wrote for B's rows
//             [y_B ; part_B] = M_B r_B      (part_B: partials for the rows R_B)
//   backward  z_B = M_B' [y_B ; z_{R_B}]
// This is the partitioned-inverse idea (Alvarado & Schreiber 1993; Pothen &
// Alvarado 1992) applied to one dense diagonal block at a time: inside a
// block there is no sequential dependency left, so the only hand-offs between
// workgroups are between blocks, whose tree is far shallower than the
// elimination tree.
#pragma once
#include <cstdint>
#include <vector>

namespace cholespy_amd {

constexpr int kTaskInts = 8;   // ints per task record (CHOLESPY_TASK_INTS of the C ABI)
constexpr int kGatherSlots = 16;   // contribution slots per row in BlockPlan::gslot
// Leaf blocks: blocks without child blocks (forward level 0) of at most
// kLeafMaxM rows and kLeafMaxRows rows of M_B. They take nothing from other
// blocks in the forward pass and give nothing to other blocks in the
// backward pass, so their tasks can run in a launch of their own before the
// forward pass and after the backward pass (one wavefront per task).
constexpr int kLeafMaxM = 64;
constexpr int kLeafMaxRows = 160;
// Split index and split count of a split task, in one int.
constexpr int split_code(int s, int nsplit) { return nsplit | (s << 8); }
constexpr int split_count(int code) { return code & 255; }
constexpr int split_index(int code) { return code >> 8; }

struct BlockPlanOptions {
    // Rows per block of a chain (the dense inverse is m x m). 0 = automatic:
    // the largest power of two <= the elimination tree's longest chain,
    // within [1024, 4096]. Large 3D separators then take fewer block levels
    // (the 64^3 Laplacian's 4096-row chains: k = 32 refined 7.22 -> 6.29 ms
    // with 4096 against 1024), while icosphere(7), whose longest chain has
    // ~2,000 rows, keeps 1024 (2048 costs it 5 %).
    int max_block = 0;
    int leaf_max = 64;      // whole subtrees up to this size become one block
    int fwd_rows = 64;      // rows of M_B per forward task
    int bwd_cols = 32;      // columns of M_B per backward task
    int split_gather = 65;  // blocks with m >= this gather r_B in separate tasks
    // Blocks with m >= split_gather are few and sit at the top of the block
    // tree, where few tasks run at once: their product tasks are narrower so
    // more workgroups stream each block's M_B concurrently.
    int fwd_rows_big = 32;  // rows per forward product task of such blocks
    int bwd_cols_big = 32;  // columns per backward task of such blocks
    // A product task of such a block that covers more than split_entries
    // entries of M_B is split along its reduction (columns forward, rows
    // backward) into up to max_split tasks of about split_entries entries
    // whose partial results the group's last split adds (0 = no splitting).
    // 8 rather than 16: the icosphere(6) k = 128 FP32 solve 536 -> 517 us,
    // icosphere(7) k = 3 FP64 190 -> 188 us.
    int split_entries = 4096;
    int max_split = 8;
    // Levels whose large blocks hold fewer than level_tasks * split_entries
    // entries split them finer, down to min_split_entries per task, so the
    // level still has about level_tasks tasks (0 = fixed split_entries).
    int level_tasks = 0;
    int min_split_entries = 512;
    // The task list for wide right-hand sides (the MFMA panel kernels, whose
    // tasks carry k times more work per entry): on levels that would give
    // fewer than wide_level_tasks tasks with the narrow shape, narrower
    // gathers and small blocks, more blocks with gather tasks, narrower
    // backward tasks and finer splits; other levels keep the narrow shape.
    int wide_fwd_rows = 16;
    // Rows per forward product task of large blocks on the wide list's
    // sparse levels: 16 (one 16-row MFMA group, the tile columns spread over
    // the 4 wavefronts); busy levels keep fwd_rows_big.
    int wide_fwd_rows_big = 16;
    int wide_split_gather = 33;
    int wide_level_tasks = 512;
    int wide_min_split_entries = 512;
    int wide_fine_bwd_cols_big = 16;   // columns per backward task of large blocks on sparse levels
    // false: the wide list is a copy of the narrow one (batches, whose many
    // systems already give each level enough tasks).
    bool wide_list = true;
    // Order of the blocks inside a level: true = the blocks with the longest
    // chain of levels still ahead of them first (forward: most block
    // ancestors; backward: most block levels below), so the tasks on the
    // pass's critical path take the early tickets; false = block index.
    // Blocks of one level are independent, so either order is a valid
    // schedule.
    bool critical_order = true;
    // Host threads for the dense block inverses (0 = hardware count, at most
    // 16). Batches already build one plan per thread and pass 1.
    int threads = 0;
};

struct BlockPlan {
    int n = 0;
    int nblocks = 0;
    // Per block b: rows [start[b], start[b+1]); R rows at rlist[rptr[b]..rptr[b+1]).
    std::vector<int> start, rptr, rlist;
    // Partial-sum slots: block b owns slots [rptr[b], rptr[b+1]) (one per R row).
    // contrib_ptr / contrib_slot: for each row i, the slots whose partial sums
    // belong to row i (CSR over rows).
    std::vector<int> contrib_ptr, contrib_slot;
    // The first kGatherSlots entries of each row's contribution list at a
    // fixed stride (row i at gslot[i * kGatherSlots], -1 padded): a gather
    // loads them without first loading contrib_ptr[i].
    std::vector<int> gslot;
    // Values of M_B in two layouts (one per pass), offsets per block.
    //   forward : column-major packed  (column c holds rows c .. m+|R|-1)
    //   backward: row-major packed     (row rho holds columns 0 .. min(rho, m-1))
    std::vector<int64_t> moff;            // same size for both layouts
    std::vector<double> mfwd, mbwd;
    // Block level in the block tree: forward level (height) and backward level (depth).
    std::vector<int> level_fwd, level_bwd;
    // Tasks in schedule order, kTaskInts ints each:
    //   (block, first, last, kind, lo, hi, sbase, split).
    // Forward kinds: 0 = gather r_B and multiply rows [first, last) of M_B;
    // 1 = gather rows [first, last) of r_B only (published for kinds 2, 3);
    // 2 = multiply rows [first, last) of M_B by the published r_B;
    // 3 = split s of nsplit (split = split_code(s, nsplit)): as 2 over
    //     columns [lo, hi) only, result published to the split rows
    //     sbase + s * (last - first) + (rho - first) (an array of its own,
    //     k values a row); split nsplit - 1 (the group's last ticket)
    //     waits for the others, adds all nsplit results in split order
    //     and finishes the rows as kind 2 would.
    // Backward kinds: 0 = columns [first, last) of M_B; 1 = split s of
    // nsplit over rows [lo, hi), combined the same way and finished as 0.
    std::vector<int> fwd_task, bwd_task;
    // The leaf blocks' tasks are the first n_leaf_fwd forward tasks and the
    // last n_leaf_bwd backward tasks (the order stays valid for one launch
    // over all tasks).
    int n_leaf_fwd = 0, n_leaf_bwd = 0;
    // The same two passes as tasks for wide right-hand sides (same records,
    // BlockPlanOptions' wide_* shape, no leaf partition).
    std::vector<int> fwd_task_wide, bwd_task_wide;
    int nsplit_rows = 0;                   // rows of the split-result array (either list)
    int max_m = 0, max_rows = 0;           // largest m and m + |R|
    int64_t nnz_l = 0;
};

// Builds the plan from L in compressed-column form (rows ascending,
// diagonal first), postordered so that parent(j) > j.
BlockPlan build_block_plan(int n, const int *colptr, const int *rowidx, const double *val,
                           const BlockPlanOptions &opt);

// The two halves of build_block_plan. The structure (blocks, R rows, slots,
// levels, tasks) depends only on the pattern of L; fill_block_plan_values
// (re)computes mfwd / mbwd for the values val of L with that pattern, on up
// to `threads` host threads (0 = hardware count, at most 16). A numeric
// refactorisation keeps the structure and refills the values.
BlockPlan build_block_plan_structure(int n, const int *colptr, const int *rowidx, const BlockPlanOptions &opt);
void fill_block_plan_values(BlockPlan &P, const int *colptr, const int *rowidx, const double *val, int threads);

}  // namespace cholespy_amd
