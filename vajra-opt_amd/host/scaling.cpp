// Ruiz equilibration (semantics of src/linalg/scaling.cpp:7-177 of the reference).
#include "indus/scaling.hpp"

#include <algorithm>
#include <cmath>

#include "indus/tolerances.hpp"

namespace indus::la {

ScalingFactors RuizScaler::compute_and_scale(SparseMatrixCSC& A, std::vector<double>& c, std::vector<double>& row_lower,
                                             std::vector<double>& row_upper, std::vector<double>& col_lower,
                                             std::vector<double>& col_upper, int max_iters, double tolerance) {
    ScalingFactors sf;
    const int m = A.m, n = A.n;
    if (m <= 0 || n <= 0) return sf;
    sf.row_scale.assign(m, 1.0);
    sf.col_scale.assign(n, 1.0);
    std::vector<double> rk(m, 1.0), ck(n, 1.0), rinf(m), cinf(n);
    auto step = [](double norm) { return norm > tol::kZeroDrop ? 1.0 / std::sqrt(norm) : 1.0; };
    for (int it = 0; it < max_iters; ++it) {
        std::fill(rinf.begin(), rinf.end(), 0.0);
        for (int j = 0; j < n; ++j) {
            double cm = 0.0;
            for (int64_t k = A.col_ptr[j]; k < A.col_ptr[j + 1]; ++k) {
                const double a = std::fabs(A.values[k]);
                rinf[A.row_idx[k]] = std::max(rinf[A.row_idx[k]], a);
                cm = std::max(cm, a);
            }
            cinf[j] = cm;
        }
        double err = 0.0;
        for (double v : rinf)
            if (v > tol::kZeroDrop) err = std::max(err, std::fabs(1.0 - v));
        double cerr = 0.0;
        for (double v : cinf)
            if (v > tol::kZeroDrop) cerr = std::max(cerr, std::fabs(1.0 - v));
        if (err <= tolerance && cerr <= tolerance) break;
        for (int i = 0; i < m; ++i) {
            rk[i] = step(rinf[i]);
            sf.row_scale[i] *= rk[i];
        }
        for (int j = 0; j < n; ++j) {
            ck[j] = step(cinf[j]);
            sf.col_scale[j] *= ck[j];
        }
        for (int j = 0; j < n; ++j)
            for (int64_t k = A.col_ptr[j]; k < A.col_ptr[j + 1]; ++k) A.values[k] *= rk[A.row_idx[k]] * ck[j];
    }
    sf.inv_row_scale.resize(m);
    sf.inv_col_scale.resize(n);
    for (int i = 0; i < m; ++i) sf.inv_row_scale[i] = 1.0 / sf.row_scale[i];
    for (int j = 0; j < n; ++j) sf.inv_col_scale[j] = 1.0 / sf.col_scale[j];
    if (c.size() == static_cast<size_t>(n))
        for (int j = 0; j < n; ++j) c[j] *= sf.col_scale[j];
    for (int i = 0; i < m && row_lower.size() == static_cast<size_t>(m); ++i)
        if (row_lower[i] > -1e20) row_lower[i] *= sf.row_scale[i];
    for (int i = 0; i < m && row_upper.size() == static_cast<size_t>(m); ++i)
        if (row_upper[i] < 1e20) row_upper[i] *= sf.row_scale[i];
    for (int j = 0; j < n && col_lower.size() == static_cast<size_t>(n); ++j)
        if (col_lower[j] > -1e20) col_lower[j] *= sf.inv_col_scale[j];
    for (int j = 0; j < n && col_upper.size() == static_cast<size_t>(n); ++j)
        if (col_upper[j] < 1e20) col_upper[j] *= sf.inv_col_scale[j];
    sf.is_scaled = true;
    return sf;
}

void RuizScaler::unscale_primal(const ScalingFactors& sf, std::vector<double>& x) {
    if (!sf.is_scaled) return;
    for (size_t j = 0; j < std::min(x.size(), sf.col_scale.size()); ++j) x[j] *= sf.col_scale[j];
}

void RuizScaler::unscale_dual(const ScalingFactors& sf, std::vector<double>& y) {
    if (!sf.is_scaled) return;
    for (size_t i = 0; i < std::min(y.size(), sf.row_scale.size()); ++i) y[i] *= sf.row_scale[i];
}

void RuizScaler::unscale_reduced_costs(const ScalingFactors& sf, std::vector<double>& d) {
    if (!sf.is_scaled) return;
    for (size_t j = 0; j < std::min(d.size(), sf.inv_col_scale.size()); ++j) d[j] *= sf.inv_col_scale[j];
}

}  // namespace indus::la
