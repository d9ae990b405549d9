// Options store: the typed members and the string-keyed variant store of the reference's
// Options (include/indus/options.hpp:10-41, behaviour of src/engine/options.cpp:6-89).
//
// Every typed member is reachable through a table of key -> member bindings. A value set
// under a key always lands in the store; it also updates the member bound to that key when
// the value's type is the member's type (a value of another type is stored only). The
// constructor seeds the store with each member's default under its primary key; aliases
// ("mip_gap", "integer_tol") update the member but are not seeded.
#include "indus/options.hpp"

#include <cstring>

namespace indus {

namespace {

using MemberRef = std::variant<std::string Options::*, double Options::*, int64_t Options::*, bool Options::*>;

struct Binding {
    const char* key;
    MemberRef member;
    bool primary;   // seeded by the constructor
};

const Binding kBindings[] = {
    {"time_limit", &Options::time_limit, true},
    {"iteration_limit", &Options::iteration_limit, true},
    {"node_limit", &Options::node_limit, true},
    {"mip_relative_gap", &Options::mip_relative_gap, true},
    {"mip_gap", &Options::mip_relative_gap, false},
    {"mip_absolute_gap", &Options::mip_absolute_gap, true},
    {"integer_tolerance", &Options::integer_tolerance, true},
    {"integer_tol", &Options::integer_tolerance, false},
    {"branching", &Options::branching_strategy, true},
    {"node_selection", &Options::node_selection, true},
    {"enable_presolve", &Options::enable_presolve, true},
    {"enable_scaling", &Options::enable_scaling, true},
    {"enable_root_cuts", &Options::enable_root_cuts, true},
    {"use_gpu", &Options::use_gpu, true},
    {"algorithm", &Options::algorithm, true},
};

// Write `value` into the member bound to `key`, if the binding's member has type T.
template <class T>
void update_bound_member(Options& o, const std::string& key, const T& value) {
    for (const Binding& b : kBindings) {
        if (key != b.key) continue;
        if (const auto* m = std::get_if<T Options::*>(&b.member)) o.**m = value;
        return;
    }
}

// A stored value read as the requested arithmetic type (integers and doubles convert into
// each other; anything else is absent).
template <class T, class Store>
bool numeric_value(const Store& values, const std::string& key, T& out) {
    auto it = values.find(key);
    if (it == values.end()) return false;
    if (const auto* d = std::get_if<double>(&it->second)) {
        out = static_cast<T>(*d);
        return true;
    }
    if (const auto* i = std::get_if<int64_t>(&it->second)) {
        out = static_cast<T>(*i);
        return true;
    }
    return false;
}

}  // namespace

Options::Options() {
    for (const Binding& b : kBindings)
        if (b.primary) std::visit([&](auto member) { set(b.key, this->*member); }, b.member);
}

void Options::set(const std::string& key, const std::string& value) {
    values_[key] = value;
    update_bound_member(*this, key, value);
}

void Options::set(const std::string& key, const char* value) { set(key, std::string(value ? value : "")); }

void Options::set(const std::string& key, double value) {
    values_[key] = value;
    update_bound_member(*this, key, value);
}

void Options::set(const std::string& key, int64_t value) {
    values_[key] = value;
    update_bound_member(*this, key, value);
}

void Options::set(const std::string& key, bool value) {
    values_[key] = value;
    update_bound_member(*this, key, value);
}

std::string Options::get_string(const std::string& key, const std::string& default_val) const {
    auto it = values_.find(key);
    if (it == values_.end()) return default_val;
    const auto* s = std::get_if<std::string>(&it->second);
    return s ? *s : default_val;
}

double Options::get_double(const std::string& key, double default_val) const {
    double v = default_val;
    return numeric_value(values_, key, v) ? v : default_val;
}

int64_t Options::get_int(const std::string& key, int64_t default_val) const {
    int64_t v = default_val;
    return numeric_value(values_, key, v) ? v : default_val;
}

bool Options::get_bool(const std::string& key, bool default_val) const {
    auto it = values_.find(key);
    if (it == values_.end()) return default_val;
    const auto* b = std::get_if<bool>(&it->second);
    return b ? *b : default_val;
}

}  // namespace indus
