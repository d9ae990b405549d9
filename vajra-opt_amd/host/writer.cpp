// Solution writers. The .sol and JSON layouts are wire formats read by the reference's file
// verifier (src/engine/verifier.cpp:660-668) and its Python audit, so field names, order and
// value formatting (17 significant digits) follow src/io/writer.cpp:85-264 of the reference.
#include <cstdio>
#include <fstream>
#include <iomanip>
#include <stdexcept>
#include <string>

#include "indus/io.hpp"

namespace indus::io {

namespace {

const char* kProducer = "VAJRA-OPT MI355X engine";
const char* kCommit = "vajra-opt-amd";

const char* status_name(SolveStatus s) {
    // the writer's own table (writer.cpp:11-26): UNSUPPORTED prints as UNKNOWN
    return s == SolveStatus::kUnsupported ? "UNKNOWN" : to_string(s);
}

const char* basis_name(BasisStatus b) {
    switch (b) {
        case BasisStatus::kBasic: return "BASIC";
        case BasisStatus::kAtLower: return "AT_LOWER";
        case BasisStatus::kAtUpper: return "AT_UPPER";
        case BasisStatus::kFixed: return "FIXED";
        case BasisStatus::kNonbasicFree: return "FREE";
        default: return "UNKNOWN";
    }
}

const char* class_name(const Model& m) {
    switch (m.classify()) {
        case ProblemClass::kQp: return "QP";
        case ProblemClass::kMilp: return "MILP";
        case ProblemClass::kMiqp: return "MIQP";
        default: return "LP";
    }
}

std::string esc(const std::string& s) {
    std::string o;
    o.reserve(s.size() + 2);
    for (unsigned char ch : s) {
        if (ch == '"') o += "\\\"";
        else if (ch == '\\') o += "\\\\";
        else if (ch == '\n') o += "\\n";
        else if (ch == '\r') o += "\\r";
        else if (ch == '\t') o += "\\t";
        else if (ch < 0x20u) {
            char b[8];
            std::snprintf(b, sizeof b, "\\u%04x", ch);
            o += b;
        } else o += static_cast<char>(ch);
    }
    return o;
}

std::string nm(const std::vector<std::string>& names, int k, char p) {
    if (static_cast<size_t>(k) < names.size() && !names[static_cast<size_t>(k)].empty()) return names[static_cast<size_t>(k)];
    return std::string(1, p) + std::to_string(k);
}

double at(const std::vector<double>& v, int k) { return static_cast<size_t>(k) < v.size() ? v[static_cast<size_t>(k)] : 0.0; }

const char* bs(const std::vector<BasisStatus>& v, int k) {
    return static_cast<size_t>(k) < v.size() ? basis_name(v[static_cast<size_t>(k)]) : "UNKNOWN";
}

const char* tf(bool b) { return b ? "true" : "false"; }

std::ofstream open_out(const std::string& path, const char* what) {
    std::ofstream f(path);
    if (!f.is_open()) throw std::runtime_error(std::string("Could not open ") + what + " file for writing: " + path);
    f << std::setprecision(17);
    return f;
}

}  // namespace

void write_solution(const Solution& s, const Model& m, const std::string& path) {
    std::ofstream f = open_out(path, "solution");
    const auto& q = s.quality;
    f << "# Solution produced by " << kProducer << "\n"
      << "# Model: " << (m.name.empty() ? "unnamed" : m.name) << "\n"
      << "# Problem Class: " << class_name(m) << "\n"
      << "# Rows: " << m.num_rows << " Cols: " << m.num_cols << " Nonzeros: " << m.A.nnz() << "\n"
      << "# Algorithm: " << s.algorithm_used << "\n"
      << "# Git Commit: " << kCommit << "\n"
      << "# Status: " << status_name(s.status) << "\n"
      << "# Message: " << s.status_message << "\n"
      << "# Termination Reason: " << (s.termination_reason.empty() ? s.status_message : s.termination_reason) << "\n"
      << "# Objective: " << s.objective_value << "\n"
      << "# Best Dual Bound: " << s.best_dual_bound << "\n"
      << "# Has Incumbent: " << tf(s.has_incumbent) << "\n"
      << "# Absolute Gap: " << s.absolute_gap << "\n"
      << "# Relative Gap: " << s.relative_gap << "\n"
      << "# Iterations: " << s.iterations << "\n"
      << "# Nodes: " << s.nodes << "\n"
      << "# Open Nodes: " << s.open_nodes << "\n"
      << "# Search Completed: " << tf(s.search_completed) << "\n"
      << "# Solve Time: " << s.solve_time_seconds << " s\n"
      << "# Primal Feasible: " << tf(q.is_primal_feasible) << "\n"
      << "# Dual Feasible: " << tf(q.is_dual_feasible) << "\n"
      << "# Max Primal Violation: " << q.max_primal_violation << "\n"
      << "# Max Dual Violation: " << q.max_dual_violation << "\n"
      << "# Max Stationarity Residual: " << q.max_stationarity_residual << "\n"
      << "# Max Complementarity Violation: " << q.max_complementarity_violation << "\n\n";
    f << "# Columns (Variables)\n# Name  Value  ReducedCost  Status\n";
    for (int j = 0; j < m.num_cols; ++j)
        f << nm(m.col_names, j, 'c') << " " << at(s.col_value, j) << " " << at(s.col_dual, j) << " "
          << bs(s.col_basis_status, j) << "\n";
    f << "\n# Rows (Constraints)\n# Name  Activity  DualMultiplier  Status\n";
    for (int i = 0; i < m.num_rows; ++i)
        f << nm(m.row_names, i, 'r') << " " << at(s.row_value, i) << " " << at(s.row_dual, i) << " "
          << bs(s.row_basis_status, i) << "\n";
    if (s.certificate_type != "none" && !s.certificate_vector.empty()) {
        f << "\n# Certificate: " << s.certificate_type << "\n";
        for (double v : s.certificate_vector) f << v << " ";
        f << "\n";
    }
}

void write_json(const Solution& s, const Model& m, const std::string& path) {
    std::ofstream f = open_out(path, "JSON telemetry");
    const auto& q = s.quality;
    const std::string pc = class_name(m);
    const char* verif = (s.status == SolveStatus::kOptimal && q.is_primal_feasible && (pc != "QP" || q.is_stationary))
                            ? "PASSED"
                            : (s.status == SolveStatus::kOptimal ? "FAILED" : "NOT_APPLICABLE");
    const std::string reason = s.termination_reason.empty() ? s.status_message : s.termination_reason;
    f << "{\n"
      << "  \"problem_class\": \"" << pc << "\",\n"
      << "  \"model_name\": \"" << esc(m.name.empty() ? "unnamed" : m.name) << "\",\n"
      << "  \"num_rows\": " << m.num_rows << ",\n"
      << "  \"num_cols\": " << m.num_cols << ",\n"
      << "  \"num_nonzeros\": " << m.A.nnz() << ",\n"
      << "  \"algorithm\": \"" << esc(s.algorithm_used) << "\",\n"
      << "  \"git_commit\": \"" << esc(kCommit) << "\",\n"
      << "  \"status\": \"" << status_name(s.status) << "\",\n"
      << "  \"verification_status\": \"" << verif << "\",\n"
      << "  \"status_message\": \"" << esc(s.status_message) << "\",\n"
      << "  \"termination_reason\": \"" << esc(reason) << "\",\n"
      << "  \"objective_value\": " << s.objective_value << ",\n"
      << "  \"max_primal_violation\": " << q.max_primal_violation << ",\n"
      << "  \"max_stationarity_residual\": " << q.max_stationarity_residual << ",\n"
      << "  \"max_complementarity_violation\": " << q.max_complementarity_violation << ",\n"
      << "  \"best_dual_bound\": " << s.best_dual_bound << ",\n"
      << "  \"has_incumbent\": " << tf(s.has_incumbent) << ",\n"
      << "  \"absolute_gap\": " << s.absolute_gap << ",\n"
      << "  \"relative_gap\": " << s.relative_gap << ",\n"
      << "  \"iterations\": " << s.iterations << ",\n"
      << "  \"nodes\": " << s.nodes << ",\n"
      << "  \"open_nodes\": " << s.open_nodes << ",\n"
      << "  \"search_completed\": " << tf(s.search_completed) << ",\n"
      << "  \"solve_time_seconds\": " << s.solve_time_seconds << ",\n"
      << "  \"quality\": {\n"
      << "    \"is_primal_feasible\": " << tf(q.is_primal_feasible) << ",\n"
      << "    \"is_dual_feasible\": " << tf(q.is_dual_feasible) << ",\n"
      << "    \"is_stationary\": " << tf(q.is_stationary) << ",\n"
      << "    \"max_primal_violation\": " << q.max_primal_violation << ",\n"
      << "    \"max_dual_violation\": " << q.max_dual_violation << ",\n"
      << "    \"max_stationarity_residual\": " << q.max_stationarity_residual << ",\n"
      << "    \"max_complementarity_violation\": " << q.max_complementarity_violation << ",\n"
      << "    \"duality_gap\": " << q.duality_gap << "\n"
      << "  },\n"
      << "  \"presolve\": {\n"
      << "    \"presolved_rows\": " << s.presolve_num_rows << ",\n"
      << "    \"presolved_cols\": " << s.presolve_num_cols << ",\n"
      << "    \"total_reductions\": " << s.presolve_total_reductions << ",\n"
      << "    \"doubleton_reductions\": " << s.presolve_doubleton_reductions << "\n"
      << "  },\n"
      << "  \"certificate\": {\n"
      << "    \"type\": \"" << esc(s.certificate_type) << "\",\n"
      << "    \"values\": [";
    for (size_t k = 0; k < s.certificate_vector.size(); ++k) f << (k ? ", " : "") << s.certificate_vector[k];
    f << "]\n  },\n  \"variables\": [\n";
    for (int j = 0; j < m.num_cols; ++j)
        f << "    {\"name\": \"" << esc(nm(m.col_names, j, 'c')) << "\", \"value\": " << at(s.col_value, j)
          << ", \"reduced_cost\": " << at(s.col_dual, j) << ", \"status\": \"" << bs(s.col_basis_status, j) << "\"}"
          << (j + 1 < m.num_cols ? "," : "") << "\n";
    f << "  ],\n  \"constraints\": [\n";
    for (int i = 0; i < m.num_rows; ++i)
        f << "    {\"name\": \"" << esc(nm(m.row_names, i, 'r')) << "\", \"activity\": " << at(s.row_value, i)
          << ", \"dual\": " << at(s.row_dual, i) << ", \"status\": \"" << bs(s.row_basis_status, i) << "\"}"
          << (i + 1 < m.num_rows ? "," : "") << "\n";
    f << "  ]\n}\n";
}

}  // namespace indus::io
