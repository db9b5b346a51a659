#include "sparse_input.hpp"

#include <algorithm>
#include <stdexcept>
#include <string>

namespace cholespy_amd {

namespace {

// One stored entry of A restricted to the lower triangle.
struct Entry {
    int row, col;
    double v;
    int64_t src;   // position in the caller's x
};

void check_index(int64_t idx, int n, const char *what) {
    if (idx < 0 || idx >= n)
        throw std::invalid_argument(std::string("Sparse matrix: ") + what +
                                    " index out of range [0, n_rows).");
}

void check_pointer(const int *ptr, int n, int64_t nnz, const char *what) {
    if (ptr[0] != 0 || (int64_t)ptr[n] != nnz)
        throw std::invalid_argument(std::string("Sparse matrix: ") + what +
                                    " pointer array must start at 0 and end at nnz.");
    for (int i = 0; i < n; ++i)
        if (ptr[i + 1] < ptr[i])
            throw std::invalid_argument(std::string("Sparse matrix: ") + what +
                                        " pointer array must be non-decreasing.");
}

}  // namespace

LowerCSC normalize_input(int n, int64_t nnz, const int *ii, const int *jj,
                         const double *x, MatrixType type, std::vector<int64_t> *entry_dst) {
    if (n < 0) throw std::invalid_argument("Sparse matrix: n_rows must be non-negative.");
    std::vector<Entry> lower;
    lower.reserve((size_t)nnz);

    auto push = [&](int r, int c, double v, int64_t e) {
        if (r >= c) lower.push_back({r, c, v, e});   // stype = -1: only row >= col is read
    };

    if (type == MatrixType::COO) {
        for (int64_t e = 0; e < nnz; ++e) {
            check_index(ii[e], n, "row");
            check_index(jj[e], n, "column");
            push(ii[e], jj[e], x[e], e);
        }
    } else if (type == MatrixType::CSR) {
        check_pointer(ii, n, nnz, "row");
        for (int r = 0; r < n; ++r)
            for (int p = ii[r]; p < ii[r + 1]; ++p) {
                check_index(jj[p], n, "column");
                push(r, jj[p], x[p], p);
            }
    } else {
        check_pointer(ii, n, nnz, "column");
        for (int c = 0; c < n; ++c)
            for (int p = ii[c]; p < ii[c + 1]; ++p) {
                check_index(jj[p], n, "row");
                push(jj[p], c, x[p], p);
            }
    }

    // Counting sort by column (stable), then order rows inside each column.
    LowerCSC out;
    out.n = n;
    out.colptr.assign((size_t)n + 1, 0);
    for (const Entry &e : lower) out.colptr[(size_t)e.col + 1]++;
    for (int c = 0; c < n; ++c) out.colptr[c + 1] += out.colptr[c];

    std::vector<int> rows(lower.size());
    std::vector<double> vals(lower.size());
    std::vector<int64_t> srcs(entry_dst ? lower.size() : 0);
    {
        std::vector<int> next(out.colptr.begin(), out.colptr.end() - 1);
        for (const Entry &e : lower) {
            int dst = next[e.col]++;
            rows[dst] = e.row;
            vals[dst] = e.v;
            if (entry_dst) srcs[dst] = e.src;
        }
    }
    if (entry_dst) entry_dst->assign((size_t)nnz, -1);

    out.rowidx.reserve(lower.size());
    out.val.reserve(lower.size());
    std::vector<int> order;
    int64_t written = 0;
    for (int c = 0; c < n; ++c) {
        const int beg = out.colptr[c], end = out.colptr[c + 1];
        order.resize((size_t)(end - beg));
        for (int p = beg; p < end; ++p) order[p - beg] = p;
        // Stable so that duplicates are summed in input order.
        std::stable_sort(order.begin(), order.end(),
                         [&](int a, int b) { return rows[a] < rows[b]; });
        out.colptr[c] = (int)written;
        size_t q = 0;
        while (q < order.size()) {
            const int r = rows[order[q]];
            double s = vals[order[q]];
            if (entry_dst) (*entry_dst)[(size_t)srcs[order[q]]] = written;
            ++q;
            while (q < order.size() && rows[order[q]] == r) {
                if (entry_dst) (*entry_dst)[(size_t)srcs[order[q]]] = written;
                s += vals[order[q++]];
            }
            out.rowidx.push_back(r);
            out.val.push_back(s);
            ++written;
        }
    }
    out.colptr[n] = (int)written;
    return out;
}

void refill_values(const std::vector<int64_t> &entry_dst, const double *x, std::vector<double> &val) {
    std::vector<char> seen(val.size(), 0);
    for (size_t e = 0; e < entry_dst.size(); ++e) {
        const int64_t q = entry_dst[e];
        if (q < 0) continue;
        if (seen[(size_t)q]) {
            val[(size_t)q] += x[e];
        } else {
            val[(size_t)q] = x[e];
            seen[(size_t)q] = 1;
        }
    }
}

}  // namespace cholespy_amd
