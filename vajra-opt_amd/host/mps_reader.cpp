// MPS reader with the observable behaviour of src/io/mps_reader.cpp:48-476 of the reference
// (quirks listed in indus/io.hpp). The file is read whole, split into lines and tokens in
// place; entries are collected as triplets and assembled once, as the reference assembles
// them (sum duplicates, drop |v| <= 1e-11).
#include <algorithm>
#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "indus/io.hpp"

namespace indus::io {

namespace {

enum class Sec { kNone, kName, kObjsense, kRows, kColumns, kRhs, kRanges, kBounds, kQuad, kEnd };

// Number in Fortran or C notation; false when the token is not a number (the reference
// swallows std::stod's exception and skips the entry).
bool to_number(const std::string& tok, double& out) {
    std::string s = tok;
    for (char& ch : s)
        if (ch == 'D' || ch == 'd') ch = 'E';
    // std::stod semantics: leading blanks skipped, longest valid prefix, failure if none
    const char* b = s.c_str();
    char* e = nullptr;
    errno = 0;
    const double v = std::strtod(b, &e);
    if (e == b) return false;
    if (errno == ERANGE) return false;  // std::stod throws out_of_range here
    out = v;
    return true;
}

void split(const std::string& line, std::vector<std::string>& tok) {
    tok.clear();
    size_t i = 0, L = line.size();
    while (i < L) {
        while (i < L && std::isspace(static_cast<unsigned char>(line[i]))) ++i;
        size_t j = i;
        while (j < L && !std::isspace(static_cast<unsigned char>(line[j]))) ++j;
        if (j > i) tok.emplace_back(line, i, j - i);
        i = j;
    }
}

std::string upper(std::string s) {
    for (char& ch : s) ch = static_cast<char>(std::toupper(static_cast<unsigned char>(ch)));
    return s;
}

}  // namespace

Model read_mps(const std::string& filepath) {
    std::ifstream in(filepath, std::ios::binary);
    if (!in.is_open()) throw std::runtime_error("Could not open MPS file: " + filepath);

    Model model;
    model.source_path = filepath;
    model.sense = ObjSense::kMinimize;
    Sec sec = Sec::kNone;
    std::string obj_row;
    std::vector<std::string> rnames, cnames;
    std::vector<char> rtype;
    std::unordered_map<std::string, int> rindex, cindex;
    std::vector<la::Triplet> trips, qtrips;
    std::unordered_map<int, double> rhs, rng, lo_set, up_set;
    std::unordered_map<int, VarType> vtype;
    bool int_block = false;
    std::vector<std::string> tok;
    std::string line;

    // a pair (name, value) of a COLUMNS, RHS or RANGES line
    auto column_entry = [&](int j, const std::string& r, const std::string& vtok) {
        double v;
        if (!to_number(vtok, v)) return;
        if (r == obj_row) {
            if (static_cast<size_t>(j) >= model.c.size()) model.c.resize(static_cast<size_t>(j) + 1, 0.0);
            model.c[static_cast<size_t>(j)] += v;
            return;
        }
        auto it = rindex.find(r);
        if (it != rindex.end()) trips.push_back({it->second, j, v});
    };
    auto row_entries = [&](std::unordered_map<int, double>& dst) {
        const size_t o = (tok.size() == 3 || tok.size() == 5) ? 1 : 0;
        for (size_t q = o; q + 1 < tok.size() && q <= o + 2; q += 2) {
            double v;
            if (!to_number(tok[q + 1], v)) continue;
            auto it = rindex.find(tok[q]);
            if (it != rindex.end()) dst[it->second] = v;
        }
    };

    while (std::getline(in, line)) {
        if (line.empty() || line[0] == '*') continue;
        if (!std::isspace(static_cast<unsigned char>(line[0]))) {
            split(line, tok);
            if (tok.empty()) continue;
            const std::string h = upper(tok[0]);
            bool header = true;
            if (h == "NAME") {
                sec = Sec::kName;
                model.name = tok.size() > 1 ? tok[1] : "unnamed";
            } else if (h == "OBJSENSE") sec = Sec::kObjsense;
            else if (h == "ROWS") sec = Sec::kRows;
            else if (h == "COLUMNS") sec = Sec::kColumns;
            else if (h == "RHS") sec = Sec::kRhs;
            else if (h == "RANGES") sec = Sec::kRanges;
            else if (h == "BOUNDS") sec = Sec::kBounds;
            else if (h == "QUADOBJ" || h == "QMATRIX") sec = Sec::kQuad;
            else if (h == "ENDATA") break;
            else header = false;
            if (header) continue;
        }
        split(line, tok);
        if (tok.empty()) continue;
        switch (sec) {
            case Sec::kObjsense: {
                const std::string s = upper(tok[0]);
                if (s == "MAX" || s == "MAXIMIZE") model.sense = ObjSense::kMaximize;
                else if (s == "MIN" || s == "MINIMIZE") model.sense = ObjSense::kMinimize;
                break;
            }
            case Sec::kRows: {
                if (tok.size() < 2) break;
                const char t = static_cast<char>(std::toupper(static_cast<unsigned char>(tok[0][0])));
                if (t == 'N') {
                    if (obj_row.empty()) obj_row = tok[1];
                } else if ((t == 'E' || t == 'L' || t == 'G') && !rindex.count(tok[1])) {
                    rindex.emplace(tok[1], static_cast<int>(rnames.size()));
                    rnames.push_back(tok[1]);
                    rtype.push_back(t);
                }
                break;
            }
            case Sec::kColumns: {
                if (tok.size() < 3) break;
                if (tok[1] == "'MARKER'" || tok[1] == "MARKER") {
                    if (tok[2].find("INTORG") != std::string::npos) int_block = true;
                    else if (tok[2].find("INTEND") != std::string::npos) int_block = false;
                    break;
                }
                int j;
                auto it = cindex.find(tok[0]);
                if (it == cindex.end()) {
                    j = static_cast<int>(cnames.size());
                    cindex.emplace(tok[0], j);
                    cnames.push_back(tok[0]);
                    vtype[j] = int_block ? VarType::kInteger : VarType::kContinuous;
                } else {
                    j = it->second;
                    if (int_block) vtype[j] = VarType::kInteger;
                }
                column_entry(j, tok[1], tok[2]);
                if (tok.size() >= 5) column_entry(j, tok[3], tok[4]);
                break;
            }
            case Sec::kRhs: row_entries(rhs); break;
            case Sec::kRanges: row_entries(rng); break;
            case Sec::kBounds: {
                if (tok.size() < 2) break;
                const std::string bt = upper(tok[0]);
                std::string cname;
                double v = 0.0;
                if (tok.size() == 2) {
                    cname = tok[1];
                } else if (tok.size() == 3) {
                    if (cindex.count(tok[1])) {
                        cname = tok[1];
                        if (!to_number(tok[2], v)) v = 0.0;
                    } else {
                        cname = tok[2];
                    }
                } else {
                    cname = tok[2];
                    if (!to_number(tok[3], v)) v = 0.0;
                }
                auto it = cindex.find(cname);
                if (it == cindex.end()) break;
                const int j = it->second;
                if (bt == "LO") lo_set[j] = v;
                else if (bt == "UP") {
                    up_set[j] = v;
                    if (v < 0.0 && !lo_set.count(j)) lo_set[j] = -1e20;
                } else if (bt == "FX") lo_set[j] = up_set[j] = v;
                else if (bt == "FR") { lo_set[j] = -1e20; up_set[j] = 1e20; }
                else if (bt == "MI") lo_set[j] = -1e20;
                else if (bt == "PL") up_set[j] = 1e20;
                else if (bt == "BV") { lo_set[j] = 0.0; up_set[j] = 1.0; vtype[j] = VarType::kInteger; }
                else if (bt == "UI") { up_set[j] = v; vtype[j] = VarType::kInteger; }
                else if (bt == "LI") { lo_set[j] = v; vtype[j] = VarType::kInteger; }
                break;
            }
            case Sec::kQuad: {
                if (tok.size() < 3) break;
                auto a = cindex.find(tok[0]), b = cindex.find(tok[1]);
                double v;
                if (a == cindex.end() || b == cindex.end() || !to_number(tok[2], v)) break;
                int r = a->second, c = b->second;
                if (r < c) std::swap(r, c);  // lower triangle
                qtrips.push_back({r, c, v});
                break;
            }
            default: break;
        }
    }

    const int m = static_cast<int>(rnames.size()), n = static_cast<int>(cnames.size());
    model.num_rows = m;
    model.num_cols = n;
    model.row_names = std::move(rnames);
    model.col_names = std::move(cnames);
    if (static_cast<int>(model.c.size()) < n) model.c.resize(static_cast<size_t>(n), 0.0);
    model.row_lower.assign(static_cast<size_t>(m), -1e20);
    model.row_upper.assign(static_cast<size_t>(m), 1e20);
    for (int i = 0; i < m; ++i) {
        auto bi = rhs.find(i);
        const double b = bi == rhs.end() ? 0.0 : bi->second;
        auto ri = rng.find(i);
        const bool has_r = ri != rng.end();
        const double r = has_r ? ri->second : 0.0;
        double& lo = model.row_lower[static_cast<size_t>(i)];
        double& hi = model.row_upper[static_cast<size_t>(i)];
        switch (rtype[static_cast<size_t>(i)]) {
            case 'E':
                if (!has_r) lo = hi = b;
                else if (r > 0.0) { lo = b; hi = b + r; }
                else { lo = b + r; hi = b; }
                break;
            case 'L': lo = has_r ? b - std::abs(r) : -1e20; hi = b; break;
            default: lo = b; hi = has_r ? b + std::abs(r) : 1e20; break;
        }
    }
    model.col_lower.assign(static_cast<size_t>(n), 0.0);
    model.col_upper.assign(static_cast<size_t>(n), 1e20);
    model.col_type.assign(static_cast<size_t>(n), VarType::kContinuous);
    for (const auto& [j, v] : lo_set) model.col_lower[static_cast<size_t>(j)] = v;
    for (const auto& [j, v] : up_set) model.col_upper[static_cast<size_t>(j)] = v;
    for (const auto& [j, t] : vtype) model.col_type[static_cast<size_t>(j)] = t;
    model.A.set_from_triplets(m, n, trips, true);
    if (!qtrips.empty()) model.Q.set_from_triplets(n, n, qtrips, true);
    return model;
}

}  // namespace indus::io
