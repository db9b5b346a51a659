#include "supernodal.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <functional>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>

#include "dense.hpp"
#include "factor.hpp"

namespace cholespy_amd {

namespace {

using dense::parallel_parts;

constexpr int kNB = 64;             // block columns of the partial factorisation
constexpr int kParallelRows = 768;  // trailing sizes worth splitting over threads

struct NotPositiveDefinite {};

// Partial Cholesky of a frontal matrix F (N x N lower, leading dimension N):
// the first ns columns are factored, the trailing N - ns block becomes the
// update matrix F22 - L21 L21'. Blocked right-looking: unblocked factor of
// a kNB diagonal block, TRSM of the rows below it, SYRK of the trailing part.
void partial_factor(double *F, int N, int ns, int threads) {
    const dense::GemmFn gemm = dense::gemm_nt();
    for (int j0 = 0; j0 < ns; j0 += kNB) {
        const int jb = std::min(kNB, ns - j0);
        // Diagonal block (already updated by earlier block columns).
        for (int j = j0; j < j0 + jb; ++j) {
            double *cj = F + (size_t)j * N;
            const double d2 = cj[j];
            if (!(d2 > 0.0) || !std::isfinite(d2)) throw NotPositiveDefinite{};
            const double d = std::sqrt(d2);
            cj[j] = d;
            for (int i = j + 1; i < j0 + jb; ++i) cj[i] /= d;
            for (int c = j + 1; c < j0 + jb; ++c) {
                const double lcj = cj[c];
                double *cc = F + (size_t)c * N;
                for (int i = c; i < j0 + jb; ++i) cc[i] -= cj[i] * lcj;
            }
        }
        const int r0 = j0 + jb, m = N - r0;
        if (m <= 0) continue;
        // Rows below: X := X L_bb^{-T}, split over row ranges.
        const int rparts = (m >= kParallelRows && threads > 1) ? threads : 1;
        parallel_parts(rparts, threads, [&](int part) {
            const int a = r0 + (int)((int64_t)m * part / rparts), b = r0 + (int)((int64_t)m * (part + 1) / rparts);
            for (int j = j0; j < j0 + jb; ++j) {
                double *xj = F + (size_t)j * N;
                for (int p = j0; p < j; ++p) {
                    const double l = F[j + (size_t)p * N];
                    const double *xp = F + (size_t)p * N;
                    for (int i = a; i < b; ++i) xj[i] -= l * xp[i];
                }
                const double d = F[j + (size_t)j * N];
                for (int i = a; i < b; ++i) xj[i] /= d;
            }
        });
        // Trailing update: F[r0:, r0:] -= P P^T, P = F[r0:, j0:j0+jb]; column
        // ranges of about equal lower-triangle area per thread.
        const double *P = F + r0 + (size_t)j0 * N;
        double *T = F + r0 + (size_t)r0 * N;
        const int cparts = (m >= kParallelRows && threads > 1) ? threads : 1;
        std::vector<int> cut((size_t)cparts + 1, 0);
        for (int q = 1; q < cparts; ++q) {
            // area of columns [0, c) of an m x m lower triangle = c m - c^2 / 2
            const double target = (double)q / cparts * 0.5 * (double)m * (double)m;
            const double c = (double)m - std::sqrt(std::max(0.0, (double)m * m - 2.0 * target));
            cut[q] = std::min(m, std::max(cut[q - 1], ((int)c + dense::kNR - 1) / dense::kNR * dense::kNR));
        }
        cut[cparts] = m;
        parallel_parts(cparts, threads, [&](int part) {
            const int ca = cut[part], cb = cut[part + 1];
            if (cb <= ca) return;
            gemm(true, cb - ca, cb - ca, jb, P + ca, N, P + ca, N, T + ca + (size_t)ca * N, N);
            if (cb < m) gemm(false, m - cb, cb - ca, jb, P + cb, N, P + ca, N, T + cb + (size_t)ca * N, N);
        });
    }
}

}  // namespace

void supernodal_numeric(int n, const std::vector<int> &colptr, const std::vector<int> &rowidx,
                        const std::vector<double> &val, const std::vector<int> &parent,
                        const std::vector<int64_t> &count, Factor &f, int threads) {
    if (threads <= 0) threads = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    // 1. Fundamental supernodes (j + 1 joins j when j is its only child and
    //    the column counts differ by one).
    std::vector<int> nchild(n, 0);
    for (int j = 0; j < n; ++j)
        if (parent[j] >= 0) nchild[parent[j]]++;
    std::vector<int> sfirst;
    for (int j = 0; j < n; ++j)
        if (j == 0 || !(parent[j - 1] == j && nchild[j] == 1 && count[j - 1] == count[j] + 1)) sfirst.push_back(j);
    const int ns_total = (int)sfirst.size();
    sfirst.push_back(n);
    std::vector<int> sn_of(n);
    for (int s = 0; s < ns_total; ++s)
        for (int j = sfirst[s]; j < sfirst[s + 1]; ++j) sn_of[j] = s;
    std::vector<int> sparent(ns_total, -1), nkids(ns_total, 0);
    std::vector<std::vector<int>> kids(ns_total);
    for (int s = 0; s < ns_total; ++s) {
        const int pj = parent[sfirst[s + 1] - 1];
        if (pj >= 0) {
            sparent[s] = sn_of[pj];
            kids[sparent[s]].push_back(s);
            nkids[sparent[s]]++;
        }
    }
    // 2. Row structure of every front: own columns, the permuted matrix below
    //    them and the children's update rows (supernodes are in postorder, so
    //    children come first).
    std::vector<std::vector<int>> srows(ns_total);
    {
        std::vector<int> mark(n, -1);
        for (int s = 0; s < ns_total; ++s) {
            const int fc = sfirst[s], lc = sfirst[s + 1];
            std::vector<int> &r = srows[s];
            for (int j = fc; j < lc; ++j) {
                r.push_back(j);
                mark[j] = s;
            }
            for (int j = fc; j < lc; ++j)
                for (int p = colptr[j]; p < colptr[j + 1]; ++p) {
                    const int i = rowidx[p];
                    if (i >= lc && mark[i] != s) {
                        mark[i] = s;
                        r.push_back(i);
                    }
                }
            for (int c : kids[s]) {
                const std::vector<int> &cr = srows[c];
                const int cw = sfirst[c + 1] - sfirst[c];
                for (size_t q = (size_t)cw; q < cr.size(); ++q) {
                    const int i = cr[q];
                    if (mark[i] != s) {
                        mark[i] = s;
                        r.push_back(i);
                    }
                }
            }
            std::sort(r.begin() + (lc - fc), r.end());
            if ((int64_t)r.size() != count[fc])
                throw std::runtime_error("supernodal factorisation: front structure does not match the column counts");
        }
    }
    // 3. Numeric: fronts in dependency order on worker threads.
    std::vector<std::vector<double>> update(ns_total);
    std::vector<std::atomic<int>> pending(ns_total);
    for (int s = 0; s < ns_total; ++s) pending[s].store(nkids[s]);
    std::mutex mu;
    std::condition_variable cv;
    std::deque<int> ready;
    for (int s = 0; s < ns_total; ++s)
        if (nkids[s] == 0) ready.push_back(s);
    int done = 0, active = 0;
    bool failed = false, not_pd = false;
    std::string error;
    const int workers = std::max(1, threads);

    auto process = [&](int s, std::vector<int> &pos, int inner_threads) {
        const int fc = sfirst[s], lc = sfirst[s + 1], w = lc - fc;
        const std::vector<int> &r = srows[s];
        const int N = (int)r.size();
        for (int q = 0; q < N; ++q) pos[r[q]] = q;
        std::vector<double> F((size_t)N * N, 0.0);
        for (int j = fc; j < lc; ++j) {
            double *col = F.data() + (size_t)(j - fc) * N;
            for (int p = colptr[j]; p < colptr[j + 1]; ++p) col[pos[rowidx[p]]] += val[p];
        }
        for (int c : kids[s]) {
            const std::vector<int> &cr = srows[c];
            const int cw = sfirst[c + 1] - sfirst[c];
            const int M = (int)cr.size() - cw;
            const double *U = update[c].data();
            for (int y = 0; y < M; ++y) {
                double *col = F.data() + (size_t)pos[cr[cw + y]] * N;
                const double *uy = U + (size_t)y * M;
                for (int x = y; x < M; ++x) col[pos[cr[cw + x]]] += uy[x];
            }
            std::vector<double>().swap(update[c]);
        }
        partial_factor(F.data(), N, w, inner_threads);
        for (int j = fc; j < lc; ++j) {
            const int64_t dst = f.colptr[j];
            const int q0 = j - fc;
            for (int q = q0; q < N; ++q) {
                f.rowidx[(size_t)(dst + q - q0)] = r[q];
                f.val[(size_t)(dst + q - q0)] = F[(size_t)q0 * N + q];
            }
        }
        const int M = N - w;
        if (sparent[s] >= 0 && M > 0) {
            std::vector<double> U((size_t)M * M);
            for (int y = 0; y < M; ++y)
                std::memcpy(U.data() + (size_t)y * M + y, F.data() + (size_t)(w + y) * N + w + y,
                            sizeof(double) * (size_t)(M - y));
            update[s].swap(U);
        }
        for (int q = 0; q < N; ++q) pos[r[q]] = -1;
    };

    auto worker = [&]() {
        std::vector<int> pos(n, -1);
        while (true) {
            int s;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv.wait(lk, [&] { return failed || !ready.empty() || done == ns_total; });
                if (failed || (ready.empty() && done == ns_total)) return;
                s = ready.front();
                ready.pop_front();
                ++active;
            }
            // Large fronts get the idle threads for their dense kernels.
            const int N = (int)srows[s].size();
            int inner = 1;
            if (N >= kParallelRows) {
                std::lock_guard<std::mutex> lk(mu);
                inner = std::max(1, workers - active + 1);
            }
            try {
                process(s, pos, inner);
            } catch (const NotPositiveDefinite &) {
                std::lock_guard<std::mutex> lk(mu);
                failed = true;
                not_pd = true;
                cv.notify_all();
                return;
            } catch (const std::exception &e) {
                std::lock_guard<std::mutex> lk(mu);
                failed = true;
                error = e.what();
                cv.notify_all();
                return;
            }
            {
                std::lock_guard<std::mutex> lk(mu);
                --active;
                ++done;
                const int p = sparent[s];
                if (p >= 0 && pending[p].fetch_sub(1) == 1) ready.push_back(p);
            }
            cv.notify_all();
        }
    };
    if (workers == 1) {
        worker();
    } else {
        std::vector<std::thread> pool;
        for (int t = 0; t < workers; ++t) pool.emplace_back(worker);
        for (auto &t : pool) t.join();
    }
    if (not_pd) throw std::invalid_argument("Matrix is not positive definite!");
    if (failed) throw std::runtime_error(error.empty() ? "supernodal factorisation failed" : error);
}

}  // namespace cholespy_amd
