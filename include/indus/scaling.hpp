// Ruiz equilibration on the host (restates src/linalg/scaling.hpp:11-43 and
// scaling.cpp:7-177 of the reference); used by the CPU PDHG.
#pragma once
#include <vector>

#include "indus/sparse.hpp"

namespace indus::la {

struct ScalingFactors {
    std::vector<double> row_scale;
    std::vector<double> col_scale;
    std::vector<double> inv_row_scale;
    std::vector<double> inv_col_scale;
    bool is_scaled = false;
};

class RuizScaler {
public:
    static ScalingFactors compute_and_scale(SparseMatrixCSC& A, std::vector<double>& c, std::vector<double>& row_lower,
                                            std::vector<double>& row_upper, std::vector<double>& col_lower,
                                            std::vector<double>& col_upper, int max_iters = 10,
                                            double tolerance = 0.05);
    static void unscale_primal(const ScalingFactors& sf, std::vector<double>& x);
    static void unscale_dual(const ScalingFactors& sf, std::vector<double>& y);
    static void unscale_reduced_costs(const ScalingFactors& sf, std::vector<double>& d);
};

}  // namespace indus::la
