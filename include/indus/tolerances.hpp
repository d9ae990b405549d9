// Tolerances used on the PDHG path (include/indus/tolerances.hpp:7-55 of the reference).
#pragma once

namespace indus::tol {

inline constexpr double kPrimalFeasibility = 1e-7;
inline constexpr double kDualFeasibility = 1e-7;
inline constexpr double kZeroDrop = 1e-11;
inline constexpr int kMaxRuizIterations = 10;
inline constexpr double kRuizTolerance = 0.05;

}  // namespace indus::tol
