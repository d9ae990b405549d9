// Sparse matrices on the PDHG path (restates include/indus/sparse.hpp:20-112 of the
// reference in C++17: the reference uses std::span, here a small read-only View with the same
// use pattern). Semantics kept exactly (src/linalg/sparse.cpp): triplet assembly sorts each
// column and drops |v| <= 1e-11; CSC multiply is a column scatter that skips |x_j| <= 1e-11;
// multiply_transpose is a per-column gather; to_csr keeps columns ascending in each row.
#pragma once
#include <cstddef>
#include <cstdint>
#include <vector>

namespace indus::la {

template <class T>
struct View {
    const T* ptr = nullptr;
    size_t len = 0;
    View() = default;
    View(const T* p, size_t n) : ptr(p), len(n) {}
    View(const std::vector<T>& v) : ptr(v.data()), len(v.size()) {}  // NOLINT(implicit)
    [[nodiscard]] size_t size() const noexcept { return len; }
    [[nodiscard]] bool empty() const noexcept { return len == 0; }
    const T& operator[](size_t i) const { return ptr[i]; }
    const T* begin() const { return ptr; }
    const T* end() const { return ptr + len; }
};

struct Triplet {
    int row = 0;
    int col = 0;
    double value = 0.0;
};

struct SparseMatrixCSR;

struct SparseMatrixCSC {
    int m = 0;
    int n = 0;
    std::vector<int64_t> col_ptr;  // n + 1
    std::vector<int> row_idx;      // nnz, ascending within a column
    std::vector<double> values;    // nnz

    SparseMatrixCSC() = default;
    SparseMatrixCSC(int rows, int cols);

    [[nodiscard]] int64_t nnz() const noexcept { return static_cast<int64_t>(values.size()); }
    [[nodiscard]] double coeff(int r, int c) const;
    [[nodiscard]] View<int> col_rows(int c) const;
    [[nodiscard]] View<double> col_vals(int c) const;

    void clear() noexcept;
    void resize(int rows, int cols);
    void set_from_triplets(int rows, int cols, const std::vector<Triplet>& triplets, bool sum_duplicates = true);
    static SparseMatrixCSC from_triplets(int rows, int cols, const std::vector<Triplet>& triplets,
                                         bool sum_duplicates = true);
    [[nodiscard]] SparseMatrixCSR to_csr() const;

    // y = alpha*A*x + beta*y
    void multiply(View<double> x, std::vector<double>& y, double alpha = 1.0, double beta = 0.0) const;
    // y = alpha*A^T*x + beta*y
    void multiply_transpose(View<double> x, std::vector<double>& y, double alpha = 1.0, double beta = 0.0) const;

    [[nodiscard]] double norm_inf() const noexcept;
};

struct SparseMatrixCSR {
    int m = 0;
    int n = 0;
    std::vector<int64_t> row_ptr;  // m + 1
    std::vector<int> col_idx;      // nnz, ascending within a row
    std::vector<double> values;

    SparseMatrixCSR() = default;
    SparseMatrixCSR(int rows, int cols);
    [[nodiscard]] int64_t nnz() const noexcept { return static_cast<int64_t>(values.size()); }
};

}  // namespace indus::la
