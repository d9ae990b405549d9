// CPLEX-style LP file reader with the observable behaviour of src/io/lp_reader.cpp:223-793
// of the reference (same grammar, same defaults, same handling of odd input), so that a
// file reads into the same Model:
//   * sections MINIMIZE/MAXIMIZE (MIN, MINIMUM, MAX, MAXIMUM), SUBJECT TO / SUCH THAT / ST /
//     S.T., BOUNDS, GENERALS (GEN, INTEGER(S), INT), BINARY (BINARIES, BIN), END; case-blind;
//   * "\" starts a comment to the end of the line;
//   * an objective label "name:" becomes the model name; a quadratic part "[ ... ] / 2" is
//     read into Q (lower triangle, coefficients as written);
//   * rows "name: expr op rhs" or "lhs op expr op rhs"; unnamed rows are R<k>; < and > mean
//     <= and >=; an operator without a number leaves the bound infinite;
//   * bounds "x >= a", "x <= b", "x = v", "x free", "a <= x <= b" (INF/INFINITY allowed);
//     default bounds 0 <= x < +inf; binaries get [0, 1] unless bounded already.
#include <algorithm>
#include <cctype>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "indus/io.hpp"

namespace indus::io {

namespace {

enum class Tk { kIdent, kNum, kColon, kPlus, kMinus, kLe, kGe, kEq, kLt, kGt, kLBr, kRBr, kStar, kSlash, kCaret, kEof };

struct Tok {
    Tk t = Tk::kEof;
    std::string s;
    double v = 0.0;
};

std::string up(std::string s) {
    for (char& ch : s) ch = static_cast<char>(std::toupper(static_cast<unsigned char>(ch)));
    return s;
}

bool is_inf_word(const std::string& s) {
    const std::string u = up(s);
    return u == "INF" || u == "INFINITY" || u == "+INF" || u == "-INF";
}

bool is_delim(char c) {
    return std::isspace(static_cast<unsigned char>(c)) || c == ':' || c == '+' || c == '-' || c == '<' || c == '>' ||
           c == '=' || c == '[' || c == ']' || c == '*' || c == '/' || c == '^' || c == '\\';
}

std::vector<Tok> lex(const std::string& src) {
    std::vector<Tok> out;
    size_t p = 0;
    const size_t L = src.size();
    auto two = [&](char next) { return p + 1 < L && src[p + 1] == next; };
    while (true) {
        while (p < L) {   // blanks and comments
            if (std::isspace(static_cast<unsigned char>(src[p]))) ++p;
            else if (src[p] == '\\') {
                while (p < L && src[p] != '\n') ++p;
            } else break;
        }
        if (p >= L) break;
        const char c = src[p];
        Tok t;
        switch (c) {
            case ':': t.t = Tk::kColon; ++p; break;
            case '+': t.t = Tk::kPlus; ++p; break;
            case '-': t.t = Tk::kMinus; ++p; break;
            case '[': t.t = Tk::kLBr; ++p; break;
            case ']': t.t = Tk::kRBr; ++p; break;
            case '*': t.t = Tk::kStar; ++p; break;
            case '/': t.t = Tk::kSlash; ++p; break;
            case '^': t.t = Tk::kCaret; ++p; break;
            case '<': t.t = two('=') ? Tk::kLe : Tk::kLt; p += two('=') ? 2 : 1; break;
            case '>': t.t = two('=') ? Tk::kGe : Tk::kGt; p += two('=') ? 2 : 1; break;
            case '=':
                if (two('<')) { t.t = Tk::kLe; p += 2; }
                else if (two('>')) { t.t = Tk::kGe; p += 2; }
                else if (two('=')) { t.t = Tk::kEq; p += 2; }
                else { t.t = Tk::kEq; ++p; }
                break;
            default: {
                const bool num = std::isdigit(static_cast<unsigned char>(c)) ||
                                 (c == '.' && p + 1 < L && std::isdigit(static_cast<unsigned char>(src[p + 1])));
                const size_t b = p;
                if (num) {
                    bool dot = false, ex = false;
                    while (p < L) {
                        const char d = src[p];
                        if (std::isdigit(static_cast<unsigned char>(d))) ++p;
                        else if (d == '.' && !dot && !ex) { dot = true; ++p; }
                        else if ((d == 'e' || d == 'E') && !ex) {
                            ex = true;
                            ++p;
                            if (p < L && (src[p] == '+' || src[p] == '-')) ++p;
                        } else break;
                    }
                    t.t = Tk::kNum;
                    t.s = src.substr(b, p - b);
                    try {
                        t.v = std::stod(t.s);
                    } catch (...) {
                        t.v = 0.0;
                    }
                } else {
                    while (p < L && !is_delim(src[p])) ++p;
                    t.t = Tk::kIdent;
                    t.s = src.substr(b, p - b);
                }
            }
        }
        out.push_back(std::move(t));
    }
    out.push_back(Tok{});
    return out;
}

enum class Sec { kStart, kObj, kRows, kBounds, kGen, kBin, kEnd };

bool is_rel(Tk t) { return t == Tk::kLe || t == Tk::kLt || t == Tk::kGe || t == Tk::kGt || t == Tk::kEq; }
bool is_le(Tk t) { return t == Tk::kLe || t == Tk::kLt; }
bool is_ge(Tk t) { return t == Tk::kGe || t == Tk::kGt; }

class Parser {
public:
    explicit Parser(std::vector<Tok> toks) : tk_(std::move(toks)) {}

    Model run(const std::string& path) {
        M_.source_path = path;
        M_.sense = ObjSense::kMinimize;
        while (at().t != Tk::kEof) {
            Sec next = sec_;
            if (header(at(), next)) {
                ++k_;
                if (next == Sec::kRows && at().t == Tk::kIdent && up(at().s) == "TO") ++k_;
                sec_ = next;
                if (sec_ == Sec::kEnd) break;
                continue;
            }
            switch (sec_) {
                case Sec::kStart: ++k_; break;
                case Sec::kObj: objective(); break;
                case Sec::kRows: row(); break;
                case Sec::kBounds: bound(); break;
                case Sec::kGen: integer(false); break;
                case Sec::kBin: integer(true); break;
                case Sec::kEnd: k_ = tk_.size() - 1; break;
            }
        }
        return finish();
    }

private:
    const Tok& at(size_t o = 0) const { return k_ + o < tk_.size() ? tk_[k_ + o] : tk_.back(); }
    const Tok& take() {
        const Tok& t = at();
        if (k_ < tk_.size()) ++k_;
        return t;
    }
    bool header(const Tok& t, Sec& next) {
        if (t.t != Tk::kIdent) return false;
        const std::string u = up(t.s);
        if (u == "MAXIMIZE" || u == "MAXIMUM" || u == "MAX") { M_.sense = ObjSense::kMaximize; next = Sec::kObj; return true; }
        if (u == "MINIMIZE" || u == "MINIMUM" || u == "MIN") { M_.sense = ObjSense::kMinimize; next = Sec::kObj; return true; }
        if (u == "SUBJECT" || u == "SUCH" || u == "ST" || u == "S.T.") { next = Sec::kRows; return true; }
        if (u == "BOUNDS" || u == "BOUND") { next = Sec::kBounds; return true; }
        if (u == "GENERALS" || u == "GENERAL" || u == "GEN" || u == "INTEGERS" || u == "INTEGER" || u == "INT") {
            next = Sec::kGen;
            return true;
        }
        if (u == "BINARY" || u == "BINARIES" || u == "BIN") { next = Sec::kBin; return true; }
        if (u == "END") { next = Sec::kEnd; return true; }
        return false;
    }
    bool at_header() {
        Sec s;
        const ObjSense keep = M_.sense;   // peeking must not change the sense
        const bool h = header(at(), s);
        M_.sense = keep;
        return h;
    }
    int col(const std::string& name) {
        auto it = cidx_.find(name);
        if (it != cidx_.end()) return it->second;
        const int j = static_cast<int>(cnames_.size());
        cnames_.push_back(name);
        cidx_.emplace(name, j);
        return j;
    }
    // optional sign token; returns -1 or +1
    double sign() {
        if (at().t == Tk::kPlus) { ++k_; return 1.0; }
        if (at().t == Tk::kMinus) { ++k_; return -1.0; }
        return 1.0;
    }

    void objective() {
        if (at().t == Tk::kIdent && at(1).t == Tk::kColon) {
            M_.name = at().s;
            k_ += 2;
        }
        while (at().t != Tk::kEof && !at_header()) {
            if (at().t == Tk::kLBr) {
                quadratic();
                continue;
            }
            const double s = sign();
            double a = 1.0;
            if (at().t == Tk::kNum) a = take().v;
            if (at().t == Tk::kIdent) {
                if (at_header()) break;
                const int j = col(take().s);
                if (static_cast<size_t>(j) >= M_.c.size()) M_.c.resize(static_cast<size_t>(j) + 1, 0.0);
                M_.c[static_cast<size_t>(j)] += s * a;
            } else if (at().t != Tk::kEof) {
                ++k_;
            }
        }
    }

    void quadratic() {
        ++k_;  // '['
        double s = 1.0;
        while (at().t != Tk::kRBr && at().t != Tk::kEof) {
            if (at().t == Tk::kPlus || at().t == Tk::kMinus) {
                s = take().t == Tk::kPlus ? 1.0 : -1.0;
                continue;
            }
            double a = 1.0;
            if (at().t == Tk::kNum) a = take().v;
            if (at().t != Tk::kIdent) {
                ++k_;
                continue;
            }
            const int j1 = col(take().s);
            if (at().t == Tk::kCaret) {
                ++k_;
                if (at().t == Tk::kNum) ++k_;
                q_.push_back({j1, j1, s * a});
            } else if (at().t == Tk::kStar) {
                ++k_;
                if (at().t == Tk::kIdent) {
                    const int j2 = col(take().s);
                    q_.push_back({std::max(j1, j2), std::min(j1, j2), s * a});
                }
            } else {
                q_.push_back({j1, j1, s * a});
            }
        }
        if (at().t == Tk::kRBr) ++k_;
        if (at().t == Tk::kSlash) {
            ++k_;
            if (at().t == Tk::kNum) ++k_;
        }
    }

    void row() {
        std::string name;
        if (at().t == Tk::kIdent && at(1).t == Tk::kColon) {
            name = at().s;
            k_ += 2;
        } else {
            name = "R" + std::to_string(rnames_.size());
        }
        // optional left-hand side "[+-] number op"
        bool has_lhs = false;
        double lhs = -1e20;
        Tk lop = Tk::kLe;
        const size_t o = (at().t == Tk::kPlus || at().t == Tk::kMinus) ? 1 : 0;
        if (at(o).t == Tk::kNum && is_rel(at(o + 1).t)) {
            const double s = o ? (take().t == Tk::kMinus ? -1.0 : 1.0) : 1.0;
            lhs = s * take().v;
            lop = take().t;
            has_lhs = true;
        }
        const int i = static_cast<int>(rnames_.size());
        rnames_.push_back(name);
        while (at().t != Tk::kEof && !at_header() && !is_rel(at().t)) {
            const double s = sign();
            double a = 1.0;
            if (at().t == Tk::kNum) a = take().v;
            if (at().t != Tk::kIdent) break;
            trips_.push_back({i, col(take().s), s * a});
        }
        double rhs = 1e20;
        Tk rop = Tk::kLe;
        if (is_rel(at().t)) {
            rop = take().t;
            const double s = sign();
            if (at().t == Tk::kNum) rhs = s * take().v;
        }
        double lo = -1e20, hi = 1e20;
        if (has_lhs) {
            if (is_le(lop)) lo = lhs;
            else if (is_ge(lop)) hi = lhs;
            if (is_le(rop)) hi = rhs;
            else if (is_ge(rop)) lo = rhs;
        } else if (rop == Tk::kEq) {
            lo = hi = rhs;
        } else if (is_le(rop)) {
            hi = rhs;
        } else if (is_ge(rop)) {
            lo = rhs;
        }
        rlo_.push_back(lo);
        rhi_.push_back(hi);
    }

    // number or +-infinity after an optional sign that was already read as `s`
    bool value(double s, double& v) {
        if (at().t == Tk::kNum) {
            v = s * take().v;
            return true;
        }
        if (at().t == Tk::kIdent && is_inf_word(at().s)) {
            ++k_;
            v = s < 0.0 ? -1e20 : 1e20;
            return true;
        }
        return false;
    }

    void bound() {
        if (at_header()) return;
        const size_t o = (at().t == Tk::kPlus || at().t == Tk::kMinus) ? 1 : 0;
        if (at(o).t == Tk::kNum || (at(o).t == Tk::kIdent && is_inf_word(at(o).s))) {
            // "a op x [op b]"
            const double s0 = o ? (take().t == Tk::kMinus ? -1.0 : 1.0) : 1.0;
            double a;
            if (at().t == Tk::kNum) a = s0 * take().v;
            else {
                const std::string w = take().s;
                a = (s0 < 0.0 || w[0] == '-') ? -1e20 : 1e20;
            }
            const Tk op1 = take().t;
            if (at().t != Tk::kIdent) return;
            const int j = col(take().s);
            if (is_le(op1)) lo_[j] = a;
            else if (is_ge(op1)) hi_[j] = a;
            if (is_le(at().t) || is_ge(at().t)) {
                const Tk op2 = take().t;
                double b;
                if (value(sign(), b)) (is_le(op2) ? hi_ : lo_)[j] = b;
            }
            return;
        }
        if (at().t != Tk::kIdent) {
            ++k_;
            return;
        }
        const int j = col(take().s);
        if (at().t == Tk::kIdent && up(at().s) == "FREE") {
            ++k_;
            lo_[j] = -1e20;
            hi_[j] = 1e20;
        } else if (is_ge(at().t)) {
            ++k_;
            double b;
            if (value(sign(), b)) lo_[j] = b;
        } else if (is_le(at().t)) {
            ++k_;
            double b;
            if (value(sign(), b)) hi_[j] = b;
        } else if (at().t == Tk::kEq) {
            ++k_;
            const double s = sign();
            if (at().t == Tk::kNum) lo_[j] = hi_[j] = s * take().v;
        }
    }

    void integer(bool binary) {
        if (at_header()) return;
        if (at().t != Tk::kIdent) {
            ++k_;
            return;
        }
        const int j = col(take().s);
        type_[j] = VarType::kInteger;
        if (binary) {
            if (!lo_.count(j)) lo_[j] = 0.0;
            if (!hi_.count(j)) hi_[j] = 1.0;
        }
    }

    Model finish() {
        const int m = static_cast<int>(rnames_.size()), n = static_cast<int>(cnames_.size());
        M_.num_rows = m;
        M_.num_cols = n;
        M_.row_names = std::move(rnames_);
        M_.col_names = std::move(cnames_);
        if (static_cast<int>(M_.c.size()) < n) M_.c.resize(static_cast<size_t>(n), 0.0);
        M_.row_lower = std::move(rlo_);
        M_.row_upper = std::move(rhi_);
        M_.col_lower.assign(static_cast<size_t>(n), 0.0);
        M_.col_upper.assign(static_cast<size_t>(n), 1e20);
        M_.col_type.assign(static_cast<size_t>(n), VarType::kContinuous);
        for (const auto& [j, v] : lo_) M_.col_lower[static_cast<size_t>(j)] = v;
        for (const auto& [j, v] : hi_) M_.col_upper[static_cast<size_t>(j)] = v;
        for (const auto& [j, t] : type_) M_.col_type[static_cast<size_t>(j)] = t;
        M_.A.set_from_triplets(m, n, trips_, true);
        if (!q_.empty()) M_.Q.set_from_triplets(n, n, q_, true);
        return std::move(M_);
    }

    std::vector<Tok> tk_;
    size_t k_ = 0;
    Sec sec_ = Sec::kStart;
    Model M_;
    std::vector<std::string> cnames_, rnames_;
    std::unordered_map<std::string, int> cidx_;
    std::vector<la::Triplet> trips_, q_;
    std::vector<double> rlo_, rhi_;
    std::unordered_map<int, double> lo_, hi_;
    std::unordered_map<int, VarType> type_;
};

}  // namespace

Model read_lp(const std::string& filepath) {
    std::ifstream in(filepath, std::ios::binary);
    if (!in.is_open()) throw std::runtime_error("Could not open LP file: " + filepath);
    std::stringstream ss;
    ss << in.rdbuf();
    return Parser(lex(ss.str())).run(filepath);
}

}  // namespace indus::io
