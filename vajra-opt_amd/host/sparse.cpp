// Sparse matrix operations on the PDHG path (semantics of src/linalg/sparse.cpp:13-238 of
// the reference; see include/indus/sparse.hpp).
#include "indus/sparse.hpp"

#include <algorithm>
#include <cmath>
#include <utility>

#include "indus/tolerances.hpp"

namespace indus::la {

SparseMatrixCSC::SparseMatrixCSC(int rows, int cols) : m(rows), n(cols), col_ptr(static_cast<size_t>(cols) + 1, 0) {}
SparseMatrixCSR::SparseMatrixCSR(int rows, int cols) : m(rows), n(cols), row_ptr(static_cast<size_t>(rows) + 1, 0) {}

void SparseMatrixCSC::clear() noexcept {
    m = n = 0;
    col_ptr.clear();
    row_idx.clear();
    values.clear();
}

void SparseMatrixCSC::resize(int rows, int cols) {
    m = rows;
    n = cols;
    col_ptr.assign(static_cast<size_t>(cols) + 1, 0);
    row_idx.clear();
    values.clear();
}

View<int> SparseMatrixCSC::col_rows(int c) const {
    if (c < 0 || c >= n || col_ptr.empty()) return {};
    return {row_idx.data() + col_ptr[c], static_cast<size_t>(col_ptr[c + 1] - col_ptr[c])};
}

View<double> SparseMatrixCSC::col_vals(int c) const {
    if (c < 0 || c >= n || col_ptr.empty()) return {};
    return {values.data() + col_ptr[c], static_cast<size_t>(col_ptr[c + 1] - col_ptr[c])};
}

double SparseMatrixCSC::coeff(int r, int c) const {
    const View<int> rows = col_rows(c);
    const int* it = std::lower_bound(rows.begin(), rows.end(), r);
    if (it != rows.end() && *it == r) return col_vals(c)[static_cast<size_t>(it - rows.begin())];
    return 0.0;
}

// Keep an entry unless it is a finite |v| <= 1e-11 (NaN and Inf are kept, sparse.cpp:59-60).
static bool keep_entry(const Triplet& t, int rows, int cols) {
    if (t.row < 0 || t.row >= rows || t.col < 0 || t.col >= cols) return false;
    return std::isnan(t.value) || std::isinf(t.value) || std::fabs(t.value) > tol::kZeroDrop;
}

void SparseMatrixCSC::set_from_triplets(int rows, int cols, const std::vector<Triplet>& triplets, bool sum_duplicates) {
    resize(rows, cols);
    // bucket by column, then sort each column by row (stable for equal rows) and merge
    std::vector<int64_t> count(static_cast<size_t>(cols) + 1, 0);
    for (const Triplet& t : triplets)
        if (keep_entry(t, rows, cols)) ++count[static_cast<size_t>(t.col) + 1];
    for (int j = 0; j < cols; ++j) count[j + 1] += count[j];
    std::vector<std::pair<int, double>> bucket(static_cast<size_t>(count[cols]));
    std::vector<int64_t> cursor(count.begin(), count.end() - 1);
    for (const Triplet& t : triplets)
        if (keep_entry(t, rows, cols)) bucket[static_cast<size_t>(cursor[t.col]++)] = {t.row, t.value};
    row_idx.reserve(bucket.size());
    values.reserve(bucket.size());
    for (int j = 0; j < cols; ++j) {
        col_ptr[j] = static_cast<int64_t>(row_idx.size());
        auto b = bucket.begin() + count[j], e = bucket.begin() + count[j + 1];
        std::stable_sort(b, e, [](const auto& a, const auto& z) { return a.first < z.first; });
        for (auto it = b; it != e; ++it) {
            if (sum_duplicates && static_cast<int64_t>(row_idx.size()) > col_ptr[j] && row_idx.back() == it->first) {
                values.back() += it->second;
            } else {
                row_idx.push_back(it->first);
                values.push_back(it->second);
            }
        }
    }
    col_ptr[cols] = static_cast<int64_t>(row_idx.size());
}

SparseMatrixCSC SparseMatrixCSC::from_triplets(int rows, int cols, const std::vector<Triplet>& triplets,
                                               bool sum_duplicates) {
    SparseMatrixCSC a;
    a.set_from_triplets(rows, cols, triplets, sum_duplicates);
    return a;
}

SparseMatrixCSR SparseMatrixCSC::to_csr() const {
    SparseMatrixCSR csr(m, n);
    if (nnz() == 0) return csr;
    for (int r : row_idx) ++csr.row_ptr[static_cast<size_t>(r) + 1];
    for (int i = 0; i < m; ++i) csr.row_ptr[i + 1] += csr.row_ptr[i];
    csr.col_idx.resize(static_cast<size_t>(nnz()));
    csr.values.resize(static_cast<size_t>(nnz()));
    std::vector<int64_t> cursor(csr.row_ptr.begin(), csr.row_ptr.end() - 1);
    for (int j = 0; j < n; ++j)
        for (int64_t k = col_ptr[j]; k < col_ptr[j + 1]; ++k) {
            const int64_t d = cursor[row_idx[k]]++;
            csr.col_idx[d] = j;
            csr.values[d] = values[k];
        }
    return csr;
}

static void apply_beta(std::vector<double>& y, double beta) {
    if (beta == 0.0) std::fill(y.begin(), y.end(), 0.0);
    else if (beta != 1.0)
        for (double& v : y) v *= beta;
}

void SparseMatrixCSC::multiply(View<double> x, std::vector<double>& y, double alpha, double beta) const {
    y.resize(static_cast<size_t>(m));
    apply_beta(y, beta);
    for (int j = 0; j < n; ++j) {
        const double xj = x[j];
        if (std::fabs(xj) <= tol::kZeroDrop) continue;  // sparse.cpp:207
        const double axj = alpha * xj;
        for (int64_t k = col_ptr[j]; k < col_ptr[j + 1]; ++k) y[row_idx[k]] += axj * values[k];
    }
}

void SparseMatrixCSC::multiply_transpose(View<double> x, std::vector<double>& y, double alpha, double beta) const {
    y.resize(static_cast<size_t>(n));
    apply_beta(y, beta);
    for (int j = 0; j < n; ++j) {
        double dot = 0.0;
        for (int64_t k = col_ptr[j]; k < col_ptr[j + 1]; ++k) dot += values[k] * x[row_idx[k]];
        y[j] += alpha * dot;
    }
}

double SparseMatrixCSC::norm_inf() const noexcept {
    std::vector<double> rows(static_cast<size_t>(m), 0.0);
    for (size_t k = 0; k < values.size(); ++k) rows[row_idx[k]] += std::fabs(values[k]);
    double mx = 0.0;
    for (double v : rows) mx = std::max(mx, v);
    return mx;
}

}  // namespace indus::la
