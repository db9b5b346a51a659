// CPU emulation of the block passes' task lists (cholespy_amd/csrc/device/
// block_plan.hpp), used by tests/test_plan_emulator.py.
//
// Runs the forward and then the backward tasks of one list one at a time in
// ticket order. Every value a task publishes starts as NaN (the sentinel), and
// the emulation fails when a task reads a value that no earlier ticket wrote:
// on the GPU that is a wait that never ends, or a schedule that can deadlock,
// since workgroups take tickets in order. It also fails when:
//   - a split row is written twice before its group is combined;
//   - the combining split (split nsplit - 1) is not the last of its group;
//   - a split row is not reset by its combine;
//   - the solution differs from a sequential solve with L.
// For the narrow list it checks the leaf launches' assumptions as well:
//   - the leaf tasks come first (forward) and last (backward);
//   - they fit the leaf kernels' shapes;
//   - their rows receive no partial sums;
//   - no gather task (kind 1) of the list belongs to a leaf block.
//
// Usage: plan_emulator FILE narrow|wide [KEY=VALUE ...]
// FILE holds L in compressed-column form (rows ascending, diagonal first,
// parent(j) > j): int64 n, int64 nnz, int32 colptr[n + 1], int32 rowidx[nnz],
// float64 val[nnz]. KEYs are BlockPlanOptions fields (see parse_option), and
// reverse=1, which runs the forward tasks in reverse ticket order: a planted
// scheduling mistake the emulation must report.
#include "block_plan.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <set>
#include <string>
#include <vector>

using namespace cholespy_amd;

namespace {

constexpr int kRhs = 2;

int64_t fwd_off(int64_t rows, int64_t c, int64_t rho) { return c * rows - c * (c - 1) / 2 + (rho - c); }
int64_t bwd_off(int64_t m, int64_t rho, int64_t c) {
    return rho < m ? rho * (rho + 1) / 2 + c : m * (m + 1) / 2 + (rho - m) * m + c;
}

bool parse_option(BlockPlanOptions &o, const std::string &kv) {
    const size_t eq = kv.find('=');
    if (eq == std::string::npos) return false;
    const std::string key = kv.substr(0, eq);
    const int v = std::atoi(kv.c_str() + eq + 1);
    struct Field {
        const char *name;
        int BlockPlanOptions::*ptr;
    };
    static const Field fields[] = {
        {"max_block", &BlockPlanOptions::max_block},
        {"leaf_max", &BlockPlanOptions::leaf_max},
        {"fwd_rows", &BlockPlanOptions::fwd_rows},
        {"bwd_cols", &BlockPlanOptions::bwd_cols},
        {"split_gather", &BlockPlanOptions::split_gather},
        {"fwd_rows_big", &BlockPlanOptions::fwd_rows_big},
        {"bwd_cols_big", &BlockPlanOptions::bwd_cols_big},
        {"split_entries", &BlockPlanOptions::split_entries},
        {"max_split", &BlockPlanOptions::max_split},
        {"level_tasks", &BlockPlanOptions::level_tasks},
        {"min_split_entries", &BlockPlanOptions::min_split_entries},
        {"wide_fwd_rows", &BlockPlanOptions::wide_fwd_rows},
        {"wide_fwd_rows_big", &BlockPlanOptions::wide_fwd_rows_big},
        {"wide_split_gather", &BlockPlanOptions::wide_split_gather},
        {"wide_level_tasks", &BlockPlanOptions::wide_level_tasks},
        {"wide_min_split_entries", &BlockPlanOptions::wide_min_split_entries},
        {"wide_fine_bwd_cols_big", &BlockPlanOptions::wide_fine_bwd_cols_big},
    };
    for (const Field &f : fields)
        if (key == f.name) {
            o.*(f.ptr) = v;
            return true;
        }
    if (key == "wide_list") {
        o.wide_list = v != 0;
        return true;
    }
    if (key == "critical_order") {
        o.critical_order = v != 0;
        return true;
    }
    return false;
}

struct Failure {
    std::string what;
};

[[noreturn]] void fail(const std::string &what) { throw Failure{what}; }

double need(double v, const char *what, size_t tk) {
    if (std::isnan(v)) fail(std::string("task ") + std::to_string(tk) + " reads an unwritten " + what);
    return v;
}

// Split bookkeeping shared by both passes: results of the splits of one
// group (indexed by its first split row) and the splits seen so far.
struct Splits {
    std::vector<double> rows;   // nsplit_rows x kRhs, NaN at rest
    std::vector<int> seen;
    explicit Splits(int nrows) : rows((size_t)std::max(nrows, 1) * kRhs, NAN), seen((size_t)std::max(nrows, 1), 0) {}
    double &at(int row, int t) { return rows[(size_t)row * kRhs + t]; }
    void put(int row, int t, double v, size_t tk) {
        double &e = at(row, t);
        if (!std::isnan(e)) fail("task " + std::to_string(tk) + " overwrites a split row before its combine");
        e = v;
    }
    // Counts split sidx of the group at sbase; true for the combining split,
    // which must come after all the others.
    bool arrive(int sbase, int sidx, int nsplit, size_t tk) {
        if (sidx == nsplit - 1) {
            if (seen[sbase] != nsplit - 1) fail("task " + std::to_string(tk) + " combines before all splits ran");
            seen[sbase] = 0;
            return true;
        }
        ++seen[sbase];
        return false;
    }
    void check_reset(const char *pass) {
        for (double e : rows)
            if (!std::isnan(e)) fail(std::string(pass) + " pass leaves a split row set");
    }
};

int run(const char *path, bool wide, const BlockPlanOptions &opt, bool reverse) {
    FILE *f = std::fopen(path, "rb");
    if (!f) fail("cannot open the factor file");
    int64_t hdr[2];
    if (std::fread(hdr, sizeof(int64_t), 2, f) != 2) fail("short factor file");
    const int n = (int)hdr[0];
    const int64_t nnz = hdr[1];
    std::vector<int> cp((size_t)n + 1), ri((size_t)nnz);
    std::vector<double> lv((size_t)nnz);
    if (std::fread(cp.data(), sizeof(int), cp.size(), f) != cp.size() ||
        std::fread(ri.data(), sizeof(int), ri.size(), f) != ri.size() ||
        std::fread(lv.data(), sizeof(double), lv.size(), f) != lv.size())
        fail("short factor file");
    std::fclose(f);

    const BlockPlan P = build_block_plan(n, cp.data(), ri.data(), lv.data(), opt);
    const std::vector<int> &ftask = wide ? P.fwd_task_wide : P.fwd_task;
    const std::vector<int> &btask = wide ? P.bwd_task_wide : P.bwd_task;
    const size_t nf = ftask.size() / kTaskInts, nbt = btask.size() / kTaskInts;

    // Right-hand sides (in the factor's order) and the sequential solution.
    std::mt19937 gen(7);
    std::uniform_real_distribution<double> U(-1.0, 1.0);
    std::vector<double> b((size_t)n * kRhs), ref;
    for (double &e : b) e = U(gen);
    ref = b;
    for (int t = 0; t < kRhs; ++t) {
        for (int j = 0; j < n; ++j) {
            const double yj = ref[(size_t)j * kRhs + t] / lv[cp[j]];
            ref[(size_t)j * kRhs + t] = yj;
            for (int p = cp[j] + 1; p < cp[j + 1]; ++p) ref[(size_t)ri[p] * kRhs + t] -= lv[p] * yj;
        }
        for (int j = n - 1; j >= 0; --j) {
            double s = ref[(size_t)j * kRhs + t];
            for (int p = cp[j] + 1; p < cp[j + 1]; ++p) s -= lv[p] * ref[(size_t)ri[p] * kRhs + t];
            ref[(size_t)j * kRhs + t] = s / lv[cp[j]];
        }
    }

    // Leaf launches (narrow list): the first n_leaf_fwd forward and the last
    // n_leaf_bwd backward tasks.
    std::set<int> leaf_blocks;
    if (!wide) {
        if (P.n_leaf_fwd > (int)nf || P.n_leaf_bwd > (int)nbt) fail("leaf counts exceed the task lists");
        for (int t = 0; t < P.n_leaf_fwd; ++t) {
            const int *T = &ftask[(size_t)t * kTaskInts];
            const int blk = T[0], m = P.start[blk + 1] - P.start[blk], R = P.rptr[blk + 1] - P.rptr[blk];
            if (T[3] != 0 || P.level_fwd[blk] != 0 || m > kLeafMaxM || m + R > kLeafMaxRows || T[2] - T[1] > 64)
                fail("leaf forward task " + std::to_string(t) + " does not fit the leaf kernel");
            for (int i = P.start[blk]; i < P.start[blk + 1]; ++i)
                if (P.contrib_ptr[i + 1] != P.contrib_ptr[i]) fail("a leaf block's row receives partial sums");
            leaf_blocks.insert(blk);
        }
        std::set<int> bwd_leaf;
        for (size_t t = nbt - (size_t)P.n_leaf_bwd; t < nbt; ++t) {
            const int *T = &btask[t * kTaskInts];
            if (T[3] != 0 || T[2] - T[1] > 32) fail("leaf backward task does not fit the leaf kernel");
            bwd_leaf.insert(T[0]);
        }
        if (bwd_leaf != leaf_blocks) fail("forward and backward leaf tasks cover different blocks");
        for (size_t t = 0; t < nf; ++t)
            if (ftask[t * kTaskInts + 3] == 1 && leaf_blocks.count(ftask[t * kTaskInts])) fail("gather task on a leaf block");
    }

    // Forward pass.
    const int nslots = P.rptr[P.nblocks];
    std::vector<double> y((size_t)n * kRhs, NAN), z((size_t)n * kRhs, NAN), rbuf((size_t)n * kRhs, NAN);
    std::vector<double> part((size_t)std::max(nslots, 1) * kRhs, NAN);
    Splits sp(P.nsplit_rows);
    std::vector<double> r;
    int nsplit_tasks = 0;
    for (size_t tk = 0; tk < nf; ++tk) {
        const int *T = &ftask[(reverse ? nf - 1 - tk : tk) * kTaskInts];
        const int blk = T[0], first = T[1], last = T[2], kind = T[3], lo = T[4], hi = T[5], sbase = T[6];
        const int nsp = split_count(T[7]), sidx = split_index(T[7]);
        if (kind < 0 || kind > 3) fail("unknown forward task kind");
        const int a0 = P.start[blk], m = P.start[blk + 1] - a0, r0 = P.rptr[blk];
        const int rows = m + P.rptr[blk + 1] - r0;
        const double *M = P.mfwd.data() + P.moff[blk];
        auto gather = [&](int i, int t) {
            double s = b[(size_t)(a0 + i) * kRhs + t];
            for (int p = P.contrib_ptr[a0 + i]; p < P.contrib_ptr[a0 + i + 1]; ++p)
                s += need(part[(size_t)P.contrib_slot[p] * kRhs + t], "partial slot", tk);
            return s;
        };
        if (kind == 1) {
            for (int i = first; i < last; ++i)
                for (int t = 0; t < kRhs; ++t) rbuf[(size_t)(a0 + i) * kRhs + t] = gather(i, t);
            continue;
        }
        // The rows of r_B the task stages (and so waits on): all m for kind 0
        // (gathered), its reduction range from rbuf for kinds 2 and 3.
        const int clo = kind == 3 ? lo : 0, chi = kind == 3 ? hi : m;
        r.assign((size_t)m * kRhs, NAN);
        for (int c = clo; c < chi; ++c)
            for (int t = 0; t < kRhs; ++t)
                r[(size_t)c * kRhs + t] =
                    kind == 0 ? gather(c, t) : need(rbuf[(size_t)(a0 + c) * kRhs + t], "rbuf row", tk);
        auto finish = [&](int rho, int t, double v) {
            if (rho < m)
                y[(size_t)(a0 + rho) * kRhs + t] = v;
            else
                part[(size_t)(r0 + rho - m) * kRhs + t] = v;
        };
        for (int rho = first; rho < last; ++rho)
            for (int t = 0; t < kRhs; ++t) {
                double s = 0.0;
                const int cend = std::min(rho < m ? rho + 1 : m, chi);
                for (int c = clo; c < cend; ++c) s += M[fwd_off(rows, c, rho)] * r[(size_t)c * kRhs + t];
                if (kind == 3)
                    sp.put(sbase + sidx * (last - first) + rho - first, t, s, tk);
                else
                    finish(rho, t, s);
            }
        if (kind == 3) {
            ++nsplit_tasks;
            if (sp.arrive(sbase, sidx, nsp, tk))
                for (int i = 0; i < last - first; ++i)
                    for (int t = 0; t < kRhs; ++t) {
                        double s = 0.0;
                        for (int q = 0; q < nsp; ++q) {
                            double &e = sp.at(sbase + q * (last - first) + i, t);
                            s += need(e, "split row", tk);
                            e = NAN;
                        }
                        finish(first + i, t, s);
                    }
        }
    }
    sp.check_reset("forward");

    // Backward pass.
    std::vector<double> v;
    for (size_t tk = 0; tk < nbt; ++tk) {
        const int *T = &btask[tk * kTaskInts];
        const int blk = T[0], first = T[1], last = T[2], kind = T[3], lo = T[4], hi = T[5], sbase = T[6];
        const int nsp = split_count(T[7]), sidx = split_index(T[7]);
        if (kind != 0 && kind != 1) fail("unknown backward task kind");
        const int a0 = P.start[blk], m = P.start[blk + 1] - a0, r0 = P.rptr[blk];
        const int rows = m + P.rptr[blk + 1] - r0;
        const double *M = P.mbwd.data() + P.moff[blk];
        const int hlo = kind == 1 ? lo : 0, hhi = kind == 1 ? hi : rows;
        v.assign((size_t)rows * kRhs, NAN);
        for (int rho = std::max(first, hlo); rho < hhi; ++rho)
            for (int t = 0; t < kRhs; ++t)
                v[(size_t)rho * kRhs + t] =
                    rho < m ? need(y[(size_t)(a0 + rho) * kRhs + t], "y row", tk)
                            : need(z[(size_t)P.rlist[r0 + rho - m] * kRhs + t], "z row", tk);
        for (int c = first; c < last; ++c)
            for (int t = 0; t < kRhs; ++t) {
                double s = 0.0;
                for (int rho = std::max(hlo, c); rho < hhi; ++rho) s += M[bwd_off(m, rho, c)] * v[(size_t)rho * kRhs + t];
                if (kind == 1)
                    sp.put(sbase + sidx * (last - first) + c - first, t, s, tk);
                else
                    z[(size_t)(a0 + c) * kRhs + t] = s;
            }
        if (kind == 1 && sp.arrive(sbase, sidx, nsp, tk))
            for (int i = 0; i < last - first; ++i)
                for (int t = 0; t < kRhs; ++t) {
                    double s = 0.0;
                    for (int q = 0; q < nsp; ++q) {
                        double &e = sp.at(sbase + q * (last - first) + i, t);
                        s += need(e, "split row", tk);
                        e = NAN;
                    }
                    z[(size_t)(a0 + first + i) * kRhs + t] = s;
                }
    }
    sp.check_reset("backward");

    double err = 0.0, scale = 0.0;
    for (size_t i = 0; i < z.size(); ++i) {
        err = std::max(err, std::fabs(need(z[i], "solution entry", nbt) - ref[i]));
        scale = std::max(scale, std::fabs(ref[i]));
    }
    if (!(err <= 1e-10 * scale)) fail("solution differs from the sequential solve: " + std::to_string(err / scale));
    std::printf("ok n %d forward tasks %zu (split %d) backward tasks %zu leaf %d/%d rel err %.2e\n", n, nf,
                nsplit_tasks, nbt, wide ? 0 : P.n_leaf_fwd, wide ? 0 : P.n_leaf_bwd, err / scale);
    return 0;
}

}  // namespace

int main(int argc, char **argv) {
    if (argc < 3 || (std::strcmp(argv[2], "narrow") != 0 && std::strcmp(argv[2], "wide") != 0)) {
        std::fprintf(stderr, "usage: %s FILE narrow|wide [KEY=VALUE ...]\n", argv[0]);
        return 2;
    }
    BlockPlanOptions opt;
    opt.threads = 1;
    bool reverse = false;
    for (int i = 3; i < argc; ++i)
        if (std::strcmp(argv[i], "reverse=1") == 0) {
            reverse = true;
        } else if (!parse_option(opt, argv[i])) {
            std::fprintf(stderr, "unknown option %s\n", argv[i]);
            return 2;
        }
    try {
        return run(argv[1], std::strcmp(argv[2], "wide") == 0, opt, reverse);
    } catch (const Failure &e) {
        std::printf("FAIL %s\n", e.what.c_str());
        return 1;
    } catch (const std::exception &e) {
        std::printf("FAIL %s\n", e.what());
        return 1;
    }
}
