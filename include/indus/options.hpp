// Solver options (restates include/indus/options.hpp:10-41 and src/engine/options.cpp:6-89
// of the reference: typed shortcut members plus a string-keyed variant store).
//
// One deliberate difference: set(key, const char*) exists here. In the reference a string
// literal binds to set(key, bool) (pointer-to-bool is a standard conversion and beats the
// user-defined conversion to std::string), so set("algorithm", "pdhg_cpu") silently stores a
// bool and leaves `algorithm` at "auto". With this overload the call means what it says.
//
// PDHG keys read on this path: "tolerance", "gap_tolerance", "check_frequency", "verbose"
// (reference, pdhg.cpp:62-67) and the engine's own: "pdhg_mode" ("restarted" | "reference"),
// "pdhg_precision" ("fp64" | "fp32" | "bf16" | "fp16"), "pdhg_graph" (bool), "allow_cpu_fallback"
// (bool: with no device, run the CPU PDHG instead of throwing).
#pragma once
#include <cstdint>
#include <string>
#include <unordered_map>
#include <variant>

namespace indus {

class Options {
public:
    Options();

    void set(const std::string& key, const std::string& value);
    void set(const std::string& key, const char* value);
    void set(const std::string& key, double value);
    void set(const std::string& key, int64_t value);
    void set(const std::string& key, int value) { set(key, static_cast<int64_t>(value)); }
    void set(const std::string& key, bool value);

    [[nodiscard]] std::string get_string(const std::string& key, const std::string& default_val = "") const;
    [[nodiscard]] double get_double(const std::string& key, double default_val = 0.0) const;
    [[nodiscard]] int64_t get_int(const std::string& key, int64_t default_val = 0) const;
    [[nodiscard]] bool get_bool(const std::string& key, bool default_val = false) const;
    [[nodiscard]] bool has(const std::string& key) const { return values_.count(key) != 0; }

    double time_limit = 1e20;
    int64_t iteration_limit = 1000000;
    int64_t node_limit = 500000;
    double mip_relative_gap = 1e-4;
    double mip_absolute_gap = 1e-6;
    double integer_tolerance = 1e-5;
    std::string branching_strategy = "most_fractional";
    std::string node_selection = "best_bound";
    bool enable_presolve = true;
    bool enable_scaling = true;
    bool enable_root_cuts = false;
    bool use_gpu = false;
    std::string algorithm = "auto";

private:
    std::unordered_map<std::string, std::variant<std::string, double, int64_t, bool>> values_;
};

}  // namespace indus
