// Test infrastructure: a fixed script of Options::set calls and a text dump of every typed
// member and of get_string / get_double / get_int / get_bool on the keys the script touches.
// Compiled against this build's include/indus/options.hpp (tests/cpp/test_indus_api.cpp) and
// against the reference's (oracle/ref_shim.cpp), so the two dumps can be compared text for
// text. Only the reference's overloads are used (std::string, double, int64_t, bool).
#pragma once
#include <cstdint>
#include <cstdio>
#include <string>

namespace options_probe {

template <class O>
void dump(const O& o, const char* tag, std::string& out) {
    char buf[512];
    std::snprintf(buf, sizeof buf, "[%s] members %a %lld %lld %a %a %a %s %s %d %d %d %d %s\n", tag, o.time_limit,
                  static_cast<long long>(o.iteration_limit), static_cast<long long>(o.node_limit), o.mip_relative_gap,
                  o.mip_absolute_gap, o.integer_tolerance, o.branching_strategy.c_str(), o.node_selection.c_str(),
                  o.enable_presolve ? 1 : 0, o.enable_scaling ? 1 : 0, o.enable_root_cuts ? 1 : 0, o.use_gpu ? 1 : 0,
                  o.algorithm.c_str());
    out += buf;
    const char* keys[] = {"time_limit", "iteration_limit", "node_limit", "mip_relative_gap", "mip_gap",
                          "mip_absolute_gap", "integer_tolerance", "integer_tol", "branching", "node_selection",
                          "enable_presolve", "enable_scaling", "enable_root_cuts", "use_gpu", "algorithm",
                          "tolerance", "check_frequency", "verbose", "custom", "missing"};
    for (const char* k : keys) {
        std::snprintf(buf, sizeof buf, "[%s] %s: s=%s d=%a i=%lld b=%d\n", tag, k, o.get_string(k, "<none>").c_str(),
                      o.get_double(k, -1.5), static_cast<long long>(o.get_int(k, -7)), o.get_bool(k, true) ? 1 : 0);
        out += buf;
    }
}

template <class O>
std::string run() {
    std::string out;
    O o;
    dump(o, "defaults", out);
    o.set("time_limit", 12.5);
    o.set("iteration_limit", static_cast<int64_t>(777));
    o.set("node_limit", 3.75);                          // a double under an integer member's key
    o.set("mip_gap", 0.25);                             // alias of mip_relative_gap
    o.set("integer_tol", 1e-3);                         // alias of integer_tolerance
    o.set("mip_absolute_gap", static_cast<int64_t>(2)); // an integer under a double member's key
    o.set("branching", std::string("pseudo_cost"));
    o.set("node_selection", std::string("depth_first"));
    o.set("algorithm", std::string("pdhg_cpu"));
    o.set("enable_presolve", false);
    o.set("use_gpu", true);
    o.set("enable_scaling", std::string("no"));         // a string under a bool member's key
    o.set("tolerance", 1e-8);
    o.set("check_frequency", static_cast<int64_t>(64));
    o.set("verbose", true);
    o.set("custom", std::string("value"));
    dump(o, "after", out);
    o.set("time_limit", true);                          // re-typed key: stored, member kept
    o.set("iteration_limit", 2.9);
    o.set("mip_relative_gap", static_cast<int64_t>(5));
    dump(o, "retyped", out);
    return out;
}

}  // namespace options_probe
