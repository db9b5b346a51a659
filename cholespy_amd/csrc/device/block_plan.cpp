#include "block_plan.hpp"

#include "../host/dense.hpp"

#include <atomic>
#include <functional>
#include <thread>

#include <algorithm>
#include <numeric>
#include <stdexcept>

namespace cholespy_amd {

namespace {

// Packed offsets of element (rho, c) of M_B (rows = m + R) in both layouts.
inline int64_t fwd_off(int64_t rows, int64_t c, int64_t rho) {   // column-major, rows c..rows-1
    return c * rows - c * (c - 1) / 2 + (rho - c);
}
inline int64_t bwd_off(int64_t m, int64_t rho, int64_t c) {      // row-major, cols 0..min(rho,m-1)
    return rho < m ? rho * (rho + 1) / 2 + c : m * (m + 1) / 2 + (rho - m) * m + c;
}

}  // namespace

namespace {

struct TaskShape {
    int fwd_rows, bwd_cols, split_gather, fwd_rows_big, bwd_cols_big, split_entries, max_split;
    // level_tasks > 0: a level that gives at least level_tasks tasks with the
    // "busy" shape (busy_fwd_rows, busy_split_gather, busy_fwd_rows_big,
    // bwd_cols_big) uses it; a level with fewer uses the finer fwd_rows /
    // split_gather / fwd_rows_big / fine_bwd_cols_big and splits its large
    // blocks down to min_split_entries entries a task.
    int level_tasks, min_split_entries;
    int busy_fwd_rows, busy_split_gather, fine_bwd_cols_big;
    bool leaf_partition;
    int busy_fwd_rows_big;
    bool critical_order;
};

struct TaskList {
    std::vector<int> fwd, bwd;
    int nsplit_rows = 0, n_leaf_fwd = 0, n_leaf_bwd = 0;
};

TaskShape narrow_shape(const BlockPlanOptions &o) {
    return TaskShape{o.fwd_rows,          o.bwd_cols,          o.split_gather, o.fwd_rows_big,
                     o.bwd_cols_big,      o.split_entries,     o.max_split,    o.level_tasks,
                     o.min_split_entries, o.fwd_rows,          o.split_gather, o.bwd_cols_big,
                     true,                o.fwd_rows_big,      o.critical_order};
}
TaskShape wide_shape(const BlockPlanOptions &o) {
    return TaskShape{o.wide_fwd_rows,          o.bwd_cols,      o.wide_split_gather, o.wide_fwd_rows_big,
                     o.bwd_cols_big,           o.split_entries, o.max_split,         o.wide_level_tasks,
                     o.wide_min_split_entries, o.fwd_rows,      o.split_gather,      o.wide_fine_bwd_cols_big,
                     false,                    o.fwd_rows_big,  o.critical_order};
}

// Tasks of both passes in level order (stable in block index) for one task shape.
TaskList make_tasks(const BlockPlan &P, const TaskShape &opt) {
    TaskList L;
    const int nb = P.nblocks;
    const std::vector<int> &start = P.start;
    int next_split = 0;   // split rows: their own array, numbered per list
    auto push = [](std::vector<int> &v, int b, int f, int l, int kind, int lo = 0, int hi = 0, int sb = 0,
                   int ns = 0) {
        const int rec[kTaskInts] = {b, f, l, kind, lo, hi, sb, ns};
        v.insert(v.end(), rec, rec + kTaskInts);
    };
    // Cuts [lo, hi) into at most `parts` ranges of about equal weight.
    auto cut = [](int lo, int hi, int parts, const std::function<int64_t(int)> &w) {
        int64_t total = 0;
        for (int x = lo; x < hi; ++x) total += w(x);
        std::vector<int> b{lo};
        int64_t acc = 0;
        for (int x = lo; x < hi; ++x) {
            acc += w(x);
            if ((int)b.size() < parts && acc * parts >= total * (int64_t)b.size() && x + 1 < hi) b.push_back(x + 1);
        }
        b.push_back(hi);
        return b;
    };
    // Entries per split task at a level: split_entries, or finer (down to
    // min_split_entries) when the level's large blocks would otherwise give
    // fewer than level_tasks tasks.
    int level_split = opt.split_entries;
    auto nsplits = [&](int64_t entries) {
        if (level_split <= 0 || entries <= level_split) return 1;
        return (int)std::min<int64_t>(opt.max_split, (entries + level_split - 1) / level_split);
    };
    auto set_level_split = [&](const std::vector<int> &order, size_t lo, size_t hi, int sg) {
        level_split = opt.split_entries;
        if (opt.level_tasks <= 0 || opt.split_entries <= 0) return;
        int64_t entries = 0;
        for (size_t q = lo; q < hi; ++q) {
            const int b = order[q];
            const int64_t m = start[b + 1] - start[b], R = P.rptr[b + 1] - P.rptr[b];
            if (m >= sg) entries += m * (m + 1) / 2 + R * m;
        }
        const int64_t want = entries / opt.level_tasks;
        level_split = (int)std::max<int64_t>(opt.min_split_entries, std::min<int64_t>(opt.split_entries, want));
    };
    std::vector<int> order(nb);
    std::iota(order.begin(), order.end(), 0);
    // Level order; inside a level the deepest blocks (longest chain to the
    // root still ahead) first when critical_order is set.
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
        if (P.level_fwd[x] != P.level_fwd[y]) return P.level_fwd[x] < P.level_fwd[y];
        return opt.critical_order && P.level_bwd[x] > P.level_bwd[y];
    });
    for (size_t lo = 0; lo < order.size();) {
        // One level at a time: the gather tasks of its large blocks first,
        // then the product tasks, so every task's inputs come from tasks with
        // smaller tickets.
        size_t hi = lo;
        while (hi < order.size() && P.level_fwd[order[hi]] == P.level_fwd[order[lo]]) ++hi;
        // The level's shape: the busy one when it already gives enough tasks.
        int rows_l = opt.fwd_rows, sg_l = opt.split_gather, rows_big_l = opt.fwd_rows_big;
        if (opt.level_tasks > 0) {
            int64_t natural = 0;
            for (size_t q = lo; q < hi; ++q) {
                const int b = order[q], m = start[b + 1] - start[b];
                const int rows = m + (P.rptr[b + 1] - P.rptr[b]);
                natural += m < opt.busy_split_gather
                               ? (rows + opt.busy_fwd_rows - 1) / opt.busy_fwd_rows
                               : (m + opt.busy_fwd_rows - 1) / opt.busy_fwd_rows +
                                     (rows + opt.busy_fwd_rows_big - 1) / opt.busy_fwd_rows_big;
            }
            if (natural >= opt.level_tasks) {
                rows_l = opt.busy_fwd_rows;
                sg_l = opt.busy_split_gather;
                rows_big_l = opt.busy_fwd_rows_big;
            }
        }
        set_level_split(order, lo, hi, sg_l);
        for (size_t q = lo; q < hi; ++q) {
            const int b = order[q], m = start[b + 1] - start[b];
            if (m < sg_l) continue;
            for (int r0 = 0; r0 < m; r0 += rows_l) push(L.fwd, b, r0, std::min(m, r0 + rows_l), 1);
        }
        for (size_t q = lo; q < hi; ++q) {
            const int b = order[q], m = start[b + 1] - start[b];
            const int rows = m + (P.rptr[b + 1] - P.rptr[b]);
            if (m < sg_l) {
                for (int r0 = 0; r0 < rows; r0 += rows_l)
                    push(L.fwd, b, r0, std::min(rows, r0 + rows_l), 0);
                continue;
            }
            for (int r0 = 0; r0 < rows; r0 += rows_big_l) {
                const int r1 = std::min(rows, r0 + rows_big_l), w = r1 - r0;
                // Column c feeds the task rows rho >= c (rho < m) and all rows >= m.
                auto weight = [&](int c) -> int64_t { return r1 - std::max(r0, std::min(c, m)); };
                const int cuse = std::min(r1, m);
                int64_t entries = 0;
                for (int c = 0; c < cuse; ++c) entries += weight(c);
                const int ns = nsplits(entries);
                if (ns == 1) {
                    push(L.fwd, b, r0, r1, 2);
                    continue;
                }
                const std::vector<int> cb = cut(0, cuse, ns, weight);
                const int parts = (int)cb.size() - 1;
                const int base = next_split;
                for (int s2 = 0; s2 < parts; ++s2)
                    push(L.fwd, b, r0, r1, 3, cb[s2], cb[s2 + 1], base, split_code(s2, parts));
                next_split += parts * w;
            }
        }
        lo = hi;
    }
    // The backward pass reuses the split rows: each combine resets its rows
    // before its kernel ends.
    const int fwd_split_rows = next_split;
    next_split = 0;
    std::iota(order.begin(), order.end(), 0);
    // Depth order; inside a depth the blocks with the most levels below them
    // first when critical_order is set.
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
        if (P.level_bwd[x] != P.level_bwd[y]) return P.level_bwd[x] < P.level_bwd[y];
        return opt.critical_order && P.level_fwd[x] > P.level_fwd[y];
    });
    for (size_t lo = 0; lo < order.size();) {
        size_t hi = lo;
        while (hi < order.size() && P.level_bwd[order[hi]] == P.level_bwd[order[lo]]) ++hi;
        int sg_l = opt.split_gather, bcols_big_l = opt.bwd_cols_big;
        if (opt.level_tasks > 0) {
            int64_t natural = 0;
            for (size_t q = lo; q < hi; ++q) {
                const int m = start[order[q] + 1] - start[order[q]];
                const int w = m >= opt.busy_split_gather ? opt.bwd_cols_big : opt.bwd_cols;
                natural += (m + w - 1) / w;
            }
            if (natural >= opt.level_tasks) sg_l = opt.busy_split_gather;
            else bcols_big_l = opt.fine_bwd_cols_big;
        }
        set_level_split(order, lo, hi, sg_l);
        for (size_t q = lo; q < hi; ++q) {
            const int b = order[q], m = start[b + 1] - start[b];
            const int rows = m + (P.rptr[b + 1] - P.rptr[b]);
            const bool big = m >= sg_l;
            const int wmax = big ? bcols_big_l : opt.bwd_cols;
            for (int c0 = 0; c0 < m; c0 += wmax) {
                const int c1 = std::min(m, c0 + wmax), w = c1 - c0;
                // Row rho feeds the task columns c <= rho.
                auto weight = [&](int rho) -> int64_t { return std::min(rho + 1, c1) - c0; };
                int64_t entries = 0;
                for (int rho = c0; rho < rows; ++rho) entries += weight(rho);
                const int ns = big ? nsplits(entries) : 1;
                if (ns == 1) {
                    push(L.bwd, b, c0, c1, 0);
                    continue;
                }
                const std::vector<int> hb = cut(c0, rows, ns, weight);
                const int parts = (int)hb.size() - 1;
                const int base = next_split;
                for (int s2 = 0; s2 < parts; ++s2)
                    push(L.bwd, b, c0, c1, 1, hb[s2], hb[s2 + 1], base, split_code(s2, parts));
                next_split += parts * w;
            }
        }
        lo = hi;
    }
    L.nsplit_rows = std::max(fwd_split_rows, next_split);

    // 6. Leaf blocks' tasks first (forward) and last (backward), otherwise in
    //    the same order.
    auto is_leaf = [&](int b) {
        const int m = start[b + 1] - start[b], R = P.rptr[b + 1] - P.rptr[b];
        return opt.leaf_partition && P.level_fwd[b] == 0 && m <= kLeafMaxM && m + R <= kLeafMaxRows &&
               m < opt.split_gather && opt.fwd_rows <= 64 && opt.bwd_cols <= 32;
    };
    auto partition = [&](std::vector<int> &tasks, bool leaf_first, int &count) {
        std::vector<int> leaf, rest;
        const size_t nt = tasks.size() / kTaskInts;
        for (size_t t = 0; t < nt; ++t) {
            std::vector<int> &dst = is_leaf(tasks[t * kTaskInts]) ? leaf : rest;
            dst.insert(dst.end(), tasks.begin() + t * kTaskInts, tasks.begin() + (t + 1) * kTaskInts);
        }
        count = (int)(leaf.size() / kTaskInts);
        tasks = leaf_first ? leaf : rest;
        tasks.insert(tasks.end(), leaf_first ? rest.begin() : leaf.begin(), leaf_first ? rest.end() : leaf.end());
    };
    partition(L.fwd, true, L.n_leaf_fwd);
    partition(L.bwd, false, L.n_leaf_bwd);
    return L;
}


}  // namespace

BlockPlan build_block_plan_structure(int n, const int *colptr, const int *rowidx, const BlockPlanOptions &opt) {
    BlockPlan P;
    P.n = n;
    P.nnz_l = colptr[n];
    std::vector<int> parent(n, -1), size(n, 1);
    for (int j = 0; j < n; ++j)
        if (colptr[j + 1] - colptr[j] > 1) parent[j] = rowidx[colptr[j] + 1];
    for (int j = 0; j < n; ++j)
        if (parent[j] >= 0) size[parent[j]] += size[j];

    // 1. Partition into contiguous blocks: small whole subtrees, then chains
    //    (j and j + 1 share a block when parent(j) = j + 1), capped at max_block.
    int max_block = opt.max_block;
    if (max_block <= 0) {
        int longest = n > 0 ? 1 : 0;
        for (int j = 0, run = 1; j + 1 < n; ++j) {
            run = parent[j] == j + 1 ? run + 1 : 1;
            longest = std::max(longest, run);
        }
        max_block = 1024;
        while (max_block < 4096 && 2 * max_block <= longest) max_block *= 2;
    }
    std::vector<int> &start = P.start;
    int j = 0;
    while (j < n) {
        // Is j the first node of a small subtree whose root r has size <= leaf_max
        // and whose parent is not small? Subtree of r spans [r - size[r] + 1, r].
        int r = j;
        while (parent[r] >= 0 && size[parent[r]] <= opt.leaf_max &&
               parent[r] - size[parent[r]] + 1 == j)
            r = parent[r];
        if (size[r] <= opt.leaf_max && r - size[r] + 1 == j && size[r] > 1) {
            start.push_back(j);
            j = r + 1;
            continue;
        }
        start.push_back(j);
        int e = j;
        while (e + 1 < n && parent[e] == e + 1 && e + 1 - j + 1 <= max_block) ++e;
        j = e + 1;
    }
    const int nb = (int)start.size();
    start.push_back(n);
    P.nblocks = nb;
    std::vector<int> block_of(n);
    for (int b = 0; b < nb; ++b)
        for (int i = start[b]; i < start[b + 1]; ++i) block_of[i] = b;

    // 2. Per block: R rows and the offset of M_B.
    P.rptr.assign((size_t)nb + 1, 0);
    P.moff.assign((size_t)nb + 1, 0);
    std::vector<int> mark(n, -1);
    std::vector<int> rows_b;
    for (int b = 0; b < nb; ++b) {
        const int a = start[b], e = start[b + 1];
        rows_b.clear();
        for (int c = a; c < e; ++c)
            for (int p = colptr[c]; p < colptr[c + 1]; ++p) {
                const int i = rowidx[p];
                if (i >= e && mark[i] != b) {
                    mark[i] = b;
                    rows_b.push_back(i);
                }
            }
        std::sort(rows_b.begin(), rows_b.end());
        P.rlist.insert(P.rlist.end(), rows_b.begin(), rows_b.end());
        P.rptr[b + 1] = (int)P.rlist.size();
        const int64_t m = e - a, R = (int64_t)rows_b.size();
        // The panel kernels index one block's M_B with 32-bit offsets, and
        // their column-major offset c * (m + R) reaches m * (m + R).
        if (m * (m + R) >= (int64_t(1) << 31))
            throw std::length_error("block plan: a block's dense M_B would exceed 2^31 entries");
        P.moff[b + 1] = P.moff[b] + m * (m + 1) / 2 + R * m;
        P.max_m = std::max<int>(P.max_m, (int)m);
        P.max_rows = std::max<int>(P.max_rows, (int)(m + R));
    }
    // 3. Contribution lists: slot s (row rlist[s]) feeds row rlist[s].
    const int nslots = P.rptr[nb];
    P.contrib_ptr.assign((size_t)n + 1, 0);
    for (int s = 0; s < nslots; ++s) P.contrib_ptr[(size_t)P.rlist[s] + 1]++;
    for (int i = 0; i < n; ++i) P.contrib_ptr[i + 1] += P.contrib_ptr[i];
    P.contrib_slot.resize((size_t)nslots);
    {
        std::vector<int> next(P.contrib_ptr.begin(), P.contrib_ptr.end() - 1);
        for (int s = 0; s < nslots; ++s) P.contrib_slot[next[P.rlist[s]]++] = s;
    }
    P.gslot.assign((size_t)n * kGatherSlots, -1);
    for (int i = 0; i < n; ++i)
        for (int p = P.contrib_ptr[i], j = 0; p < P.contrib_ptr[i + 1] && j < kGatherSlots; ++p, ++j)
            P.gslot[(size_t)i * kGatherSlots + j] = P.contrib_slot[p];
    std::vector<int> slot_block(nslots);
    for (int b = 0; b < nb; ++b)
        for (int s = P.rptr[b]; s < P.rptr[b + 1]; ++s) slot_block[s] = b;

    // 4. Block levels: forward depends on blocks with slots for our rows,
    //    backward on the blocks holding our R rows.
    P.level_fwd.assign(nb, 0);
    for (int b = 0; b < nb; ++b) {
        int lv = 0;
        for (int i = start[b]; i < start[b + 1]; ++i)
            for (int p = P.contrib_ptr[i]; p < P.contrib_ptr[i + 1]; ++p)
                lv = std::max(lv, P.level_fwd[slot_block[P.contrib_slot[p]]] + 1);
        P.level_fwd[b] = lv;
    }
    P.level_bwd.assign(nb, 0);
    for (int b = nb - 1; b >= 0; --b) {
        int lv = 0;
        for (int s = P.rptr[b]; s < P.rptr[b + 1]; ++s)
            lv = std::max(lv, P.level_bwd[block_of[P.rlist[s]]] + 1);
        P.level_bwd[b] = lv;
    }

    // 5. Task lists: one for narrow right-hand sides, one for wide ones.
    TaskList narrow = make_tasks(P, narrow_shape(opt));
    TaskList wide = opt.wide_list ? make_tasks(P, wide_shape(opt)) : narrow;
    P.fwd_task.swap(narrow.fwd);
    P.bwd_task.swap(narrow.bwd);
    P.n_leaf_fwd = narrow.n_leaf_fwd;
    P.n_leaf_bwd = narrow.n_leaf_bwd;
    P.fwd_task_wide.swap(wide.fwd);
    P.bwd_task_wide.swap(wide.bwd);
    P.nsplit_rows = std::max(narrow.nsplit_rows, wide.nsplit_rows);
    return P;
}

void fill_block_plan_values(BlockPlan &P, const int *colptr, const int *rowidx, const double *val, int threads) {
    const int nb = P.nblocks;
    P.mfwd.assign((size_t)P.moff[nb], 0.0);
    P.mbwd.assign((size_t)P.moff[nb], 0.0);
    const int hw = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    const int nt = std::max(1, std::min(hw, threads > 0 ? threads : hw));
    // Dense inverses per block. Blocks are independent: the few large ones
    // (top-of-tree separator chains, which would otherwise bound the time)
    // run one after another with all threads inside their products, the
    // rest are shared out over the threads, heaviest first.
    struct Work {
        std::vector<int> pos;
        std::vector<double> Lbb, Pb, Linv, Q;
    };
    auto compute = [&](int b, int inner, Work &w) {
        const int a = P.start[b], e = P.start[b + 1];
        const int m = e - a;
        const int R = P.rptr[b + 1] - P.rptr[b];
        const int *rl = P.rlist.data() + P.rptr[b];
        if (w.pos.empty()) w.pos.assign((size_t)P.n, -1);
        for (int q = 0; q < R; ++q) w.pos[rl[q]] = q;
        // Column-major L_BB (m x m) and L_RB (R x m); X := L_BB^{-T} (so
        // L_BB^{-1}[rho][c] = X[c][rho]); then C = -L_RB L_BB^{-1}, the
        // lower part of M_B.
        w.Lbb.assign((size_t)m * m, 0.0);
        w.Pb.assign((size_t)R * m, 0.0);
        for (int c = a; c < e; ++c)
            for (int p = colptr[c]; p < colptr[c + 1]; ++p) {
                const int i = rowidx[p];
                if (i < e) w.Lbb[(size_t)(c - a) * m + (i - a)] = val[p];
                else w.Pb[(size_t)(c - a) * R + w.pos[i]] = val[p];
            }
        for (int q = 0; q < R; ++q) w.pos[rl[q]] = -1;
        w.Linv.assign((size_t)m * m, 0.0);
        for (int c = 0; c < m; ++c) w.Linv[(size_t)c * m + c] = 1.0;
        dense::lower_inverse_trans(m, w.Lbb.data(), m, w.Linv.data(), m, inner);
        w.Q.assign((size_t)R * m, 0.0);
        if (R > 0) dense::gemm_lower_inverse(R, m, w.Pb.data(), R, w.Linv.data(), m, w.Q.data(), R, inner);
        const int rows = m + R;
        double *F = P.mfwd.data() + P.moff[b];
        double *Bk = P.mbwd.data() + P.moff[b];
        const double *Linv = w.Linv.data(), *Q = w.Q.data();
        // Both layouts written in kTile x kTile tiles, so that the
        // column-major one is not written with a stride per entry.
        constexpr int kTile = 32;
        const int tiles = (rows + kTile - 1) / kTile;
        dense::parallel_parts(inner, inner, [&](int part) {
            const int t0 = dense::part_bound(tiles, inner, part, 1), t1 = dense::part_bound(tiles, inner, part + 1, 1);
            for (int r0 = t0 * kTile; r0 < std::min(rows, t1 * kTile); r0 += kTile) {
                const int r1 = std::min(rows, r0 + kTile);
                for (int c0 = 0; c0 <= std::min(r1 - 1, m - 1); c0 += kTile)
                    for (int rho = r0; rho < r1; ++rho) {
                        const int cmax = std::min(rho < m ? rho : m - 1, c0 + kTile - 1);
                        for (int c = c0; c <= cmax; ++c) {
                            const double v = rho < m ? Linv[(size_t)rho * m + c] : Q[(size_t)c * R + (rho - m)];
                            F[fwd_off(rows, c, rho)] = v;
                            Bk[bwd_off(m, rho, c)] = v;
                        }
                    }
            }
        });
    };
    std::vector<int> by_cost(nb);
    std::iota(by_cost.begin(), by_cost.end(), 0);
    auto cost = [&](int b) {
        const double m = P.start[b + 1] - P.start[b], R = P.rptr[b + 1] - P.rptr[b];
        return m * m * m / 3.0 + R * m * m;
    };
    std::stable_sort(by_cost.begin(), by_cost.end(), [&](int x, int y) { return cost(x) > cost(y); });
    constexpr double kBigBlock = 4.0e6;   // multiply-adds worth splitting over threads
    int nbig = 0;
    if (nt > 1)
        while (nbig < nb && cost(by_cost[nbig]) >= kBigBlock) ++nbig;
    {
        Work w;
        for (int t = 0; t < nbig; ++t) compute(by_cost[t], nt, w);
    }
    std::atomic<int> next_block(nbig);
    auto dense_work = [&]() {
        Work w;
        for (int t = next_block.fetch_add(1); t < nb; t = next_block.fetch_add(1)) compute(by_cost[t], 1, w);
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < nt; ++t) pool.emplace_back(dense_work);
    dense_work();
    for (auto &t : pool) t.join();
}

BlockPlan build_block_plan(int n, const int *colptr, const int *rowidx, const double *val,
                           const BlockPlanOptions &opt) {
    BlockPlan P = build_block_plan_structure(n, colptr, rowidx, opt);
    fill_block_plan_values(P, colptr, rowidx, val, opt.threads);
    return P;
}

}  // namespace cholespy_amd
