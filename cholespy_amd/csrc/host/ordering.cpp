#include "ordering.hpp"

#include <algorithm>
#include <cstdint>
#include <cstdlib>

namespace cholespy_amd {

namespace {

class Dissector {
public:
    Dissector(const Graph &g, const NDOptions &opt)
        : g_(g), opt_(opt), region_(g.n, 0), level_(g.n, -1), seen_(g.n, 0),
          side_(g.n, 0), local_(g.n, -1) {
        perm_.reserve(g.n);
    }

    std::vector<int> run() {
        std::vector<int> all(g_.n);
        for (int v = 0; v < g_.n; ++v) all[v] = v;
        order_set(all, 0);
        return std::move(perm_);
    }

private:
    const Graph &g_;
    NDOptions opt_;
    std::vector<int> region_;   // label of the vertex set a vertex currently belongs to
    std::vector<int> level_;    // BFS level scratch, -1 when unvisited
    std::vector<int> seen_;     // stamp for component search
    std::vector<char> side_;    // 0 = A, 1 = separator, 2 = B during one split
    std::vector<int> local_;    // vertex -> position inside a leaf
    std::vector<int> perm_;
    int next_region_ = 1;
    int stamp_ = 0;

    int restricted_degree(int v, int rid) const {
        int d = 0;
        for (int p = g_.xadj[v]; p < g_.xadj[v + 1]; ++p) d += region_[g_.adj[p]] == rid;
        return d;
    }

    // Rooted level structure of the component containing `root` inside region
    // `rid`. Fills `order` with vertices in BFS order and `levptr` with level
    // boundaries; leaves level_[] set for the visited vertices.
    int bfs(int root, int rid, std::vector<int> &order, std::vector<int> &levptr) {
        order.clear();
        levptr.clear();
        order.push_back(root);
        level_[root] = 0;
        levptr.push_back(0);
        size_t head = 0;
        int cur = 0;
        while (head < order.size()) {
            const int v = order[head];
            if (level_[v] != cur) {
                levptr.push_back((int)head);
                cur = level_[v];
            }
            ++head;
            for (int p = g_.xadj[v]; p < g_.xadj[v + 1]; ++p) {
                const int u = g_.adj[p];
                if (region_[u] == rid && level_[u] < 0) {
                    level_[u] = cur + 1;
                    order.push_back(u);
                }
            }
        }
        levptr.push_back((int)order.size());
        return (int)levptr.size() - 1;
    }

    void clear_levels(const std::vector<int> &order) {
        for (int v : order) level_[v] = -1;
    }

    void order_set(std::vector<int> &verts, int depth) {
        if (verts.empty()) return;
        const int rid = next_region_++;
        for (int v : verts) region_[v] = rid;
        // Split into connected components (all of them before any recursion,
        // which reuses the stamp array); each is dissected on its own.
        const int stamp = ++stamp_;
        std::vector<std::vector<int>> comps;
        for (int s : verts) {
            if (seen_[s] == stamp) continue;
            std::vector<int> comp{s};
            seen_[s] = stamp;
            for (size_t h = 0; h < comp.size(); ++h) {
                const int v = comp[h];
                for (int p = g_.xadj[v]; p < g_.xadj[v + 1]; ++p) {
                    const int u = g_.adj[p];
                    if (region_[u] == rid && seen_[u] != stamp) {
                        seen_[u] = stamp;
                        comp.push_back(u);
                    }
                }
            }
            comps.push_back(std::move(comp));
        }
        for (auto &comp : comps) dissect(comp, rid, depth);
    }

    void dissect(std::vector<int> &comp, int rid, int depth) {
        const int N = (int)comp.size();
        if (N <= opt_.leaf_size || depth > 400) {
            leaf(comp, rid);
            return;
        }
        // Pseudo-peripheral root: repeat BFS from a minimum-degree vertex of
        // the deepest level while the eccentricity keeps growing.
        int root = comp[0], best_deg = restricted_degree(root, rid);
        for (int v : comp) {
            const int d = restricted_degree(v, rid);
            if (d < best_deg) { best_deg = d; root = v; }
        }
        std::vector<int> order, levptr;
        int h = bfs(root, rid, order, levptr);
        for (int it = 0; it < 8; ++it) {
            int cand = -1, cand_deg = 1 << 30;
            for (int p = levptr[h - 1]; p < levptr[h]; ++p) {
                const int d = restricted_degree(order[p], rid);
                if (d < cand_deg) { cand_deg = d; cand = order[p]; }
            }
            clear_levels(order);
            std::vector<int> o2, l2;
            const int h2 = bfs(cand, rid, o2, l2);
            if (h2 > h) {
                order.swap(o2);
                levptr.swap(l2);
                h = h2;
            } else {
                clear_levels(o2);
                bfs(root, rid, order, levptr);
                break;
            }
            root = cand;
        }
        if (h < 3) {   // compact component: no useful separator
            clear_levels(order);
            leaf(comp, rid);
            return;
        }
        // Separator level: smallest level leaving both sides balanced enough.
        int m_best = -1;
        long best_sep = 0, best_gap = 0;
        for (int m = 1; m <= h - 2; ++m) {
            const long before = levptr[m], sep = levptr[m + 1] - levptr[m];
            const long after = N - before - sep;
            if (std::min(before, after) < (long)(opt_.min_balance * N)) continue;
            const long gap = std::labs(before - after);
            if (m_best < 0 || sep < best_sep || (sep == best_sep && gap < best_gap)) {
                m_best = m; best_sep = sep; best_gap = gap;
            }
        }
        if (m_best < 0) {
            for (int m = 1; m <= h - 2; ++m) {
                const long before = levptr[m], sep = levptr[m + 1] - levptr[m];
                const long gap = std::labs(before - (N - before - sep));
                if (m_best < 0 || gap < best_gap) { m_best = m; best_gap = gap; }
            }
        }
        const int m = m_best;
        for (int p = 0; p < N; ++p) {
            const int v = order[p];
            side_[v] = level_[v] < m ? 0 : (level_[v] == m ? 1 : 2);
        }
        clear_levels(order);
        // Trim to a minimal separator: a separator vertex with no neighbour on
        // one side can join the other side without connecting A and B.
        std::vector<int> A, B, S;
        for (int p = levptr[m]; p < levptr[m + 1]; ++p) {
            const int s = order[p];
            bool touches_b = false;
            for (int q = g_.xadj[s]; q < g_.xadj[s + 1]; ++q) {
                const int u = g_.adj[q];
                if (region_[u] == rid && side_[u] == 2) { touches_b = true; break; }
            }
            if (!touches_b) side_[s] = 0;
        }
        for (int p = levptr[m]; p < levptr[m + 1]; ++p) {
            const int s = order[p];
            if (side_[s] != 1) continue;
            bool touches_a = false;
            for (int q = g_.xadj[s]; q < g_.xadj[s + 1]; ++q) {
                const int u = g_.adj[q];
                if (region_[u] == rid && side_[u] == 0) { touches_a = true; break; }
            }
            if (!touches_a) side_[s] = 2;
        }
        for (int v : order) {
            if (side_[v] == 0) A.push_back(v);
            else if (side_[v] == 2) B.push_back(v);
            else S.push_back(v);
        }
        if (A.empty() || B.empty()) {   // degenerate split: order as a leaf
            leaf(comp, rid);
            return;
        }
        // Mark the separator out of both halves before recursing.
        const int sep_rid = next_region_++;
        for (int v : S) region_[v] = sep_rid;
        order_set(A, depth + 1);
        order_set(B, depth + 1);
        for (int v : S) perm_.push_back(v);
    }

    // Exact minimum degree on the induced subgraph (bitset elimination graph)
    // for leaves of up to 64 vertices; larger compact pieces keep BFS order.
    void leaf(std::vector<int> &comp, int rid) {
        const int m = (int)comp.size();
        if (m > 64) {
            for (int v : comp) perm_.push_back(v);
            return;
        }
        for (int i = 0; i < m; ++i) local_[comp[i]] = i;
        std::vector<uint64_t> nb(m, 0);
        for (int i = 0; i < m; ++i) {
            const int v = comp[i];
            for (int p = g_.xadj[v]; p < g_.xadj[v + 1]; ++p) {
                const int u = g_.adj[p];
                if (region_[u] == rid && local_[u] >= 0 && u != v) nb[i] |= uint64_t(1) << local_[u];
            }
        }
        uint64_t alive = (m == 64) ? ~uint64_t(0) : ((uint64_t(1) << m) - 1);
        for (int step = 0; step < m; ++step) {
            int best = -1, best_deg = 1 << 30;
            for (int i = 0; i < m; ++i) {
                if (!((alive >> i) & 1)) continue;
                const int d = __builtin_popcountll(nb[i] & alive);
                if (d < best_deg) { best_deg = d; best = i; }
            }
            const uint64_t bit = uint64_t(1) << best;
            const uint64_t clique = nb[best] & alive & ~bit;
            for (int j = 0; j < m; ++j)
                if ((clique >> j) & 1) nb[j] |= clique & ~(uint64_t(1) << j);
            alive &= ~bit;
            perm_.push_back(comp[best]);
        }
        for (int v : comp) local_[v] = -1;
    }
};

}  // namespace

std::vector<int> nested_dissection(const Graph &g, const NDOptions &opt) {
    Dissector d(g, opt);
    return d.run();
}

}  // namespace cholespy_amd
