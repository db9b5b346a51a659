#include "factor.hpp"

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <stdexcept>

#include "ordering.hpp"
#include "supernodal.hpp"

namespace cholespy_amd {

namespace {

// Upper-triangular compressed columns of P A P' (column k holds rows i <= k),
// built from the lower triangle of A and pinv (pinv[old] = new).
struct UpperCSC {
    int n = 0;
    std::vector<int> colptr, rowidx;
    std::vector<double> val;
};

// a_to_c (optional) receives, per entry of A, its position in the result.
UpperCSC permute_upper(const LowerCSC &A, const std::vector<int> &pinv, std::vector<int> *a_to_c = nullptr) {
    const int n = A.n;
    UpperCSC C;
    C.n = n;
    C.colptr.assign((size_t)n + 1, 0);
    for (int j = 0; j < n; ++j)
        for (int p = A.colptr[j]; p < A.colptr[j + 1]; ++p) {
            const int a = pinv[A.rowidx[p]], b = pinv[j];
            C.colptr[(size_t)std::max(a, b) + 1]++;
        }
    for (int k = 0; k < n; ++k) C.colptr[k + 1] += C.colptr[k];
    C.rowidx.resize((size_t)C.colptr[n]);
    C.val.resize((size_t)C.colptr[n]);
    if (a_to_c) a_to_c->assign((size_t)A.nnz(), -1);
    std::vector<int> next(C.colptr.begin(), C.colptr.end() - 1);
    for (int j = 0; j < n; ++j)
        for (int p = A.colptr[j]; p < A.colptr[j + 1]; ++p) {
            const int a = pinv[A.rowidx[p]], b = pinv[j];
            const int col = std::max(a, b), row = std::min(a, b);
            const int dst = next[col]++;
            C.rowidx[dst] = row;
            C.val[dst] = A.val[p];
            if (a_to_c) (*a_to_c)[(size_t)p] = dst;
        }
    return C;
}

// Elimination tree with path compression through a virtual-ancestor array.
std::vector<int> etree_upper(const UpperCSC &C) {
    const int n = C.n;
    std::vector<int> parent(n, -1), anc(n, -1);
    for (int k = 0; k < n; ++k) {
        for (int p = C.colptr[k]; p < C.colptr[k + 1]; ++p) {
            int i = C.rowidx[p];
            while (i != -1 && i < k) {
                const int up = anc[i];
                anc[i] = k;
                if (up == -1) parent[i] = k;
                i = up;
            }
        }
    }
    return parent;
}

// Depth-first postorder of the forest; children are visited in increasing
// index order. Returns post[k] = node visited k-th.
std::vector<int> postorder(const std::vector<int> &parent) {
    const int n = (int)parent.size();
    std::vector<int> first_child(n, -1), next_sib(n, -1);
    for (int j = n - 1; j >= 0; --j) {
        const int p = parent[j];
        if (p >= 0) {
            next_sib[j] = first_child[p];
            first_child[p] = j;
        }
    }
    std::vector<int> post;
    post.reserve(n);
    std::vector<int> stack;
    for (int r = 0; r < n; ++r) {
        if (parent[r] != -1) continue;
        stack.push_back(r);
        while (!stack.empty()) {
            const int v = stack.back();
            const int c = first_child[v];
            if (c != -1) {
                first_child[v] = next_sib[c];   // consume the child
                stack.push_back(c);
            } else {
                post.push_back(v);
                stack.pop_back();
            }
        }
    }
    return post;
}

// Pattern of row k of L (excluding the diagonal): the union of the etree
// paths from each i in C(:, k), i < k, up to k. Written to buf[top..n) in an
// order where every node precedes its etree ancestors; returns top.
inline int row_pattern(const UpperCSC &C, int k, const std::vector<int> &parent,
                       std::vector<int> &mark, std::vector<int> &path, std::vector<int> &buf) {
    const int n = C.n;
    int top = n;
    mark[k] = k;
    for (int p = C.colptr[k]; p < C.colptr[k + 1]; ++p) {
        int i = C.rowidx[p];
        if (i >= k) continue;
        int len = 0;
        while (mark[i] != k) {
            path[len++] = i;
            mark[i] = k;
            i = parent[i];
        }
        while (len > 0) buf[--top] = path[--len];
    }
    return top;
}

}  // namespace

SymbolicFactor symbolic_factor(const LowerCSC &A, const FactorOptions &opt) {
    const int n = A.n;
    SymbolicFactor S;
    S.n = n;
    S.threads = opt.threads;

    // 1. Fill-reducing ordering on the adjacency graph of A.
    std::vector<int> perm0(n);
    if (opt.natural) {
        for (int i = 0; i < n; ++i) perm0[i] = i;
    } else {
        Graph g;
        g.n = n;
        g.xadj.assign((size_t)n + 1, 0);
        for (int j = 0; j < n; ++j)
            for (int p = A.colptr[j]; p < A.colptr[j + 1]; ++p) {
                const int i = A.rowidx[p];
                if (i == j) continue;
                g.xadj[(size_t)i + 1]++;
                g.xadj[(size_t)j + 1]++;
            }
        for (int v = 0; v < n; ++v) g.xadj[v + 1] += g.xadj[v];
        g.adj.resize((size_t)g.xadj[n]);
        std::vector<int> next(g.xadj.begin(), g.xadj.end() - 1);
        for (int j = 0; j < n; ++j)
            for (int p = A.colptr[j]; p < A.colptr[j + 1]; ++p) {
                const int i = A.rowidx[p];
                if (i == j) continue;
                g.adj[next[i]++] = j;
                g.adj[next[j]++] = i;
            }
        NDOptions nd;
        nd.leaf_size = opt.nd_leaf_size;
        nd.min_balance = opt.nd_min_balance;
        perm0 = nested_dissection(g, nd);
    }

    // 2. Postorder the elimination tree of the reordered matrix.
    std::vector<int> pinv(n);
    for (int k = 0; k < n; ++k) pinv[perm0[k]] = k;
    {
        UpperCSC C0 = permute_upper(A, pinv);
        std::vector<int> par0 = etree_upper(C0);
        std::vector<int> post = postorder(par0);
        S.perm.resize(n);
        for (int k = 0; k < n; ++k) S.perm[k] = perm0[post[k]];
    }
    for (int k = 0; k < n; ++k) pinv[S.perm[k]] = k;
    UpperCSC C = permute_upper(A, pinv, &S.a_to_c);
    S.parent = etree_upper(C);

    // 3. Column counts from the row patterns (row subtrees).
    std::vector<int> mark(n, -1), path(n), buf(n);
    S.count.assign(n, 1);
    for (int k = 0; k < n; ++k) {
        const int top = row_pattern(C, k, S.parent, mark, path, buf);
        for (int t = top; t < n; ++t) S.count[buf[t]]++;
    }
    S.colptr.assign((size_t)n + 1, 0);
    int64_t total = 0;
    for (int j = 0; j < n; ++j) {
        if (total > std::numeric_limits<int>::max())
            throw std::runtime_error("Cholesky factor too large for 32-bit indices.");
        S.colptr[j] = (int)total;
        total += S.count[j];
    }
    if (total > std::numeric_limits<int>::max())
        throw std::runtime_error("Cholesky factor too large for 32-bit indices.");
    S.colptr[n] = (int)total;

    FactorOptions::Method method = opt.method;
    if (method == FactorOptions::Method::Auto) {
        const char *e = std::getenv("CHOLESPY_FACTOR");
        if (e && std::strcmp(e, "simplicial") == 0)
            method = FactorOptions::Method::Simplicial;
        else if (e && std::strcmp(e, "supernodal") == 0)
            method = FactorOptions::Method::Supernodal;
        else
            method = total >= opt.supernodal_min_nnz ? FactorOptions::Method::Supernodal
                                                     : FactorOptions::Method::Simplicial;
    }
    S.supernodal = method == FactorOptions::Method::Supernodal;
    if (S.supernodal) {
        // The lower CSC of P A P' (the transpose of C) for the multifrontal phase.
        S.l_colptr.assign((size_t)n + 1, 0);
        S.l_rowidx.resize((size_t)C.colptr[n]);
        S.c_to_l.resize((size_t)C.colptr[n]);
        for (int k = 0; k < n; ++k)
            for (int p = C.colptr[k]; p < C.colptr[k + 1]; ++p) S.l_colptr[(size_t)C.rowidx[p] + 1]++;
        for (int j = 0; j < n; ++j) S.l_colptr[j + 1] += S.l_colptr[j];
        std::vector<int> next(S.l_colptr.begin(), S.l_colptr.end() - 1);
        for (int k = 0; k < n; ++k)
            for (int p = C.colptr[k]; p < C.colptr[k + 1]; ++p) {
                const int q = next[C.rowidx[p]]++;
                S.l_rowidx[q] = k;
                S.c_to_l[p] = q;
            }
    }
    S.c_colptr = std::move(C.colptr);
    S.c_rowidx = std::move(C.rowidx);
    return S;
}

Factor numeric_factor(const SymbolicFactor &S, const std::vector<double> &a_val) {
    const int n = S.n;
    if (a_val.size() != S.a_to_c.size())
        throw std::invalid_argument("numeric_factor: values do not match the symbolic factor.");
    Factor f;
    f.n = n;
    f.perm = S.perm;
    f.parent = S.parent;
    f.colptr = S.colptr;
    const int64_t total = S.colptr[n];
    f.rowidx.resize((size_t)total);
    f.val.resize((size_t)total);
    const size_t nc = (size_t)S.c_colptr[n];

    if (S.supernodal) {
        // Multifrontal numeric factorisation on the lower CSC of P A P'.
        std::vector<double> lv(nc);
        for (size_t p = 0; p < a_val.size(); ++p) lv[(size_t)S.c_to_l[(size_t)S.a_to_c[p]]] = a_val[p];
        supernodal_numeric(n, S.l_colptr, S.l_rowidx, lv, S.parent, S.count, f, S.threads);
        return f;
    }

    // Up-looking numeric factorisation: row k of L solves a sparse
    // triangular system against the already computed columns.
    UpperCSC C;
    C.n = n;
    C.colptr = S.c_colptr;
    C.rowidx = S.c_rowidx;
    C.val.resize(nc);
    for (size_t p = 0; p < a_val.size(); ++p) C.val[(size_t)S.a_to_c[p]] = a_val[p];
    std::vector<int> mark(n, -1), path(n), buf(n);
    std::vector<int> fill(f.colptr.begin(), f.colptr.end() - 1);   // next free slot per column
    std::vector<double> x(n, 0.0);
    for (int k = 0; k < n; ++k) {
        const int top = row_pattern(C, k, f.parent, mark, path, buf);
        for (int p = C.colptr[k]; p < C.colptr[k + 1]; ++p) x[C.rowidx[p]] += C.val[p];
        double d = x[k];
        x[k] = 0.0;
        for (int t = top; t < n; ++t) {
            const int j = buf[t];
            const double lkj = x[j] / f.val[f.colptr[j]];
            x[j] = 0.0;
            for (int p = f.colptr[j] + 1; p < fill[j]; ++p) x[f.rowidx[p]] -= f.val[p] * lkj;
            d -= lkj * lkj;
            const int q = fill[j]++;
            f.rowidx[q] = k;
            f.val[q] = lkj;
        }
        if (!(d > 0.0) || !std::isfinite(d))
            throw std::invalid_argument("Matrix is not positive definite!");
        const int q = fill[k]++;
        f.rowidx[q] = k;
        f.val[q] = std::sqrt(d);
    }
    return f;
}

Factor cholesky_factor(const LowerCSC &A, const FactorOptions &opt) {
    return numeric_factor(symbolic_factor(A, opt), A.val);
}

void factor_rows(const Factor &f, std::vector<int> &rowptr, std::vector<int> &colidx,
                 std::vector<double> &val) {
    const int n = f.n;
    rowptr.assign((size_t)n + 1, 0);
    for (int64_t p = 0; p < f.nnz(); ++p) rowptr[(size_t)f.rowidx[p] + 1]++;
    for (int i = 0; i < n; ++i) rowptr[i + 1] += rowptr[i];
    colidx.resize((size_t)f.nnz());
    val.resize((size_t)f.nnz());
    std::vector<int> next(rowptr.begin(), rowptr.end() - 1);
    for (int j = 0; j < n; ++j)
        for (int p = f.colptr[j]; p < f.colptr[j + 1]; ++p) {
            const int dst = next[f.rowidx[p]]++;
            colidx[dst] = j;
            val[dst] = f.val[p];
        }
}

void factor_solve_colmajor(const Factor &f, int k, double *bx) {
    const int n = f.n;
    std::vector<double> y(n);
    for (int t = 0; t < k; ++t) {
        double *col = bx + (size_t)t * n;
        for (int i = 0; i < n; ++i) y[i] = col[f.perm[i]];
        for (int j = 0; j < n; ++j) {
            const int p0 = f.colptr[j];
            const double yj = y[j] / f.val[p0];
            y[j] = yj;
            for (int p = p0 + 1; p < f.colptr[j + 1]; ++p) y[f.rowidx[p]] -= f.val[p] * yj;
        }
        for (int j = n - 1; j >= 0; --j) {
            const int p0 = f.colptr[j];
            double s = y[j];
            for (int p = p0 + 1; p < f.colptr[j + 1]; ++p) s -= f.val[p] * y[f.rowidx[p]];
            y[j] = s / f.val[p0];
        }
        for (int i = 0; i < n; ++i) col[f.perm[i]] = y[i];
    }
}

}  // namespace cholespy_amd
