// Options store (src/engine/options.cpp:6-89 of the reference).
#include "indus/options.hpp"

namespace indus {

Options::Options() {
    set("time_limit", time_limit);
    set("iteration_limit", iteration_limit);
    set("node_limit", node_limit);
    set("mip_relative_gap", mip_relative_gap);
    set("mip_absolute_gap", mip_absolute_gap);
    set("integer_tolerance", integer_tolerance);
    set("branching", branching_strategy);
    set("node_selection", node_selection);
    set("enable_presolve", enable_presolve);
    set("enable_scaling", enable_scaling);
    set("enable_root_cuts", enable_root_cuts);
    set("use_gpu", use_gpu);
    set("algorithm", algorithm);
}

void Options::set(const std::string& key, const std::string& value) {
    values_[key] = value;
    if (key == "algorithm") algorithm = value;
    else if (key == "branching") branching_strategy = value;
    else if (key == "node_selection") node_selection = value;
}

void Options::set(const std::string& key, const char* value) { set(key, std::string(value ? value : "")); }

void Options::set(const std::string& key, double value) {
    values_[key] = value;
    if (key == "time_limit") time_limit = value;
    else if (key == "mip_relative_gap" || key == "mip_gap") mip_relative_gap = value;
    else if (key == "mip_absolute_gap") mip_absolute_gap = value;
    else if (key == "integer_tolerance" || key == "integer_tol") integer_tolerance = value;
}

void Options::set(const std::string& key, int64_t value) {
    values_[key] = value;
    if (key == "iteration_limit") iteration_limit = value;
    else if (key == "node_limit") node_limit = value;
}

void Options::set(const std::string& key, bool value) {
    values_[key] = value;
    if (key == "enable_presolve") enable_presolve = value;
    else if (key == "enable_scaling") enable_scaling = value;
    else if (key == "enable_root_cuts") enable_root_cuts = value;
    else if (key == "use_gpu") use_gpu = value;
}

std::string Options::get_string(const std::string& key, const std::string& default_val) const {
    auto it = values_.find(key);
    if (it != values_.end())
        if (const auto* s = std::get_if<std::string>(&it->second)) return *s;
    return default_val;
}

double Options::get_double(const std::string& key, double default_val) const {
    auto it = values_.find(key);
    if (it == values_.end()) return default_val;
    if (const auto* d = std::get_if<double>(&it->second)) return *d;
    if (const auto* i = std::get_if<int64_t>(&it->second)) return static_cast<double>(*i);
    return default_val;
}

int64_t Options::get_int(const std::string& key, int64_t default_val) const {
    auto it = values_.find(key);
    if (it == values_.end()) return default_val;
    if (const auto* i = std::get_if<int64_t>(&it->second)) return *i;
    if (const auto* d = std::get_if<double>(&it->second)) return static_cast<int64_t>(*d);
    return default_val;
}

bool Options::get_bool(const std::string& key, bool default_val) const {
    auto it = values_.find(key);
    if (it != values_.end())
        if (const auto* b = std::get_if<bool>(&it->second)) return *b;
    return default_val;
}

}  // namespace indus
