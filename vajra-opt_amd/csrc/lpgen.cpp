// Seeded LP generators (host code). Declared in include/vajra_lpgen.h.
#include "vajra_lpgen.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

namespace {

struct SplitMix64 {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
};

// xoshiro256** 1.0 (Blackman & Vigna).
struct Xoshiro {
    uint64_t s[4];
    explicit Xoshiro(uint64_t seed) {
        SplitMix64 sm{seed};
        for (auto& w : s) w = sm.next();
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        const uint64_t result = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return result;
    }
    double uniform() { return static_cast<double>(next() >> 11) * 0x1.0p-53; }
    double uniform(double a, double b) { return a + (b - a) * uniform(); }
    // integer on [0, bound) by Lemire's multiply-shift with rejection
    uint32_t below(uint32_t bound) {
        uint64_t x = next() >> 32;
        uint64_t mres = x * bound;
        uint32_t l = static_cast<uint32_t>(mres);
        if (l < bound) {
            const uint32_t t = static_cast<uint32_t>(-bound) % bound;
            while (l < t) {
                x = next() >> 32;
                mres = x * bound;
                l = static_cast<uint32_t>(mres);
            }
        }
        return static_cast<uint32_t>(mres >> 32);
    }
    double normal() {
        // Box-Muller, one value per call (the second is discarded so the stream
        // position depends only on the number of calls).
        double u1 = uniform();
        while (u1 <= 0.0) u1 = uniform();
        const double u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

// Floyd's algorithm: k distinct values from [0, m), returned ascending.
void floyd_sample(Xoshiro& rng, int m, int k, int32_t* out) {
    int cnt = 0;
    for (int j = m - k; j < m; ++j) {
        const int t = static_cast<int>(rng.below(static_cast<uint32_t>(j + 1)));
        bool seen = false;
        for (int q = 0; q < cnt; ++q)
            if (out[q] == t) { seen = true; break; }
        out[cnt++] = seen ? j : t;
    }
    std::sort(out, out + k);
}

// b = A x0 (row sums in ascending column order), c = A^T y + s.
void form_rhs(int m, int n, const int64_t* col_ptr, const int32_t* row_idx, const double* vals,
              const std::vector<double>& x0, double* row_lower, double* row_upper) {
    std::vector<double> b(static_cast<size_t>(m), 0.0);
    for (int j = 0; j < n; ++j)
        for (int64_t p = col_ptr[j]; p < col_ptr[j + 1]; ++p)
            b[static_cast<size_t>(row_idx[p])] += vals[p] * x0[static_cast<size_t>(j)];
    for (int i = 0; i < m; ++i) {
        row_lower[i] = b[static_cast<size_t>(i)];
        row_upper[i] = b[static_cast<size_t>(i)];
    }
}

}  // namespace

extern "C" int vj_gen_random_lp(uint64_t seed, int m, int n, int k, double scaled_frac,
                                double log10_range, int64_t* col_ptr, int32_t* row_idx,
                                double* vals, double* c, double* col_lower, double* col_upper,
                                double* row_lower, double* row_upper, double* x0_out,
                                double* y0_out) {
    if (m <= 0 || n <= 0 || k <= 0 || k > m) return -1;
    Xoshiro rng(seed);
    col_ptr[0] = 0;
    for (int j = 0; j < n; ++j) {
        int32_t* rows = row_idx + static_cast<int64_t>(j) * k;
        floyd_sample(rng, m, k, rows);
        double* v = vals + static_cast<int64_t>(j) * k;
        for (int q = 0; q < k; ++q) v[q] = rng.normal();
        col_ptr[j + 1] = static_cast<int64_t>(j + 1) * k;
    }
    if (scaled_frac > 0.0) {
        for (int j = 0; j < n; ++j) {
            const bool pick = rng.uniform() < scaled_frac;
            const double e = rng.uniform(-log10_range, log10_range);
            if (!pick) continue;
            const double s = std::pow(10.0, e);
            double* v = vals + static_cast<int64_t>(j) * k;
            for (int q = 0; q < k; ++q) v[q] *= s;
        }
    }
    std::vector<double> x0(static_cast<size_t>(n)), y0(static_cast<size_t>(m));
    for (auto& v : x0) v = rng.uniform();
    for (auto& v : y0) v = rng.normal();
    form_rhs(m, n, col_ptr, row_idx, vals, x0, row_lower, row_upper);
    for (int j = 0; j < n; ++j) {
        double dot = 0.0;
        for (int64_t p = col_ptr[j]; p < col_ptr[j + 1]; ++p)
            dot += vals[p] * y0[static_cast<size_t>(row_idx[p])];
        c[j] = dot + rng.uniform();
        col_lower[j] = 0.0;
        col_upper[j] = 10.0;
    }
    if (x0_out) std::copy(x0.begin(), x0.end(), x0_out);
    if (y0_out) std::copy(y0.begin(), y0.end(), y0_out);
    return 0;
}

extern "C" int vj_gen_staircase_lp(uint64_t seed, int periods, int rows_pp, int cols_pp, int k,
                                   int inv, int64_t* col_ptr, int32_t* row_idx, double* vals,
                                   double* c, double* col_lower, double* col_upper,
                                   double* row_lower, double* row_upper, double* x0_out) {
    if (periods <= 0 || rows_pp <= 0 || cols_pp <= 0 || k <= 0 || k > rows_pp || inv < 0 ||
        inv > cols_pp || inv > rows_pp)
        return -1;
    Xoshiro rng(seed);
    const int m = periods * rows_pp, n = periods * cols_pp;
    std::vector<int32_t> rows(static_cast<size_t>(k));
    int64_t p = 0;
    col_ptr[0] = 0;
    for (int t = 0; t < periods; ++t) {
        for (int q = 0; q < cols_pp; ++q) {
            floyd_sample(rng, rows_pp, k, rows.data());
            for (int r = 0; r < k; ++r) {
                row_idx[p] = t * rows_pp + rows[static_cast<size_t>(r)];
                vals[p] = rng.normal();
                ++p;
            }
            const int inv_idx = q - (cols_pp - inv);
            if (inv_idx >= 0 && t + 1 < periods) {
                // identity coupling into the next period: strictly below this block's rows,
                // so the column stays sorted ascending
                row_idx[p] = (t + 1) * rows_pp + inv_idx;
                vals[p] = 1.0;
                ++p;
            }
            col_ptr[t * cols_pp + q + 1] = p;
        }
    }
    std::vector<double> x0(static_cast<size_t>(n));
    for (int j = 0; j < n; ++j) {
        c[j] = rng.uniform(1.0, 10.0);
        col_lower[j] = 0.0;
        col_upper[j] = rng.uniform(5.0, 50.0);
        x0[static_cast<size_t>(j)] = rng.uniform(0.0, 5.0);
    }
    form_rhs(m, n, col_ptr, row_idx, vals, x0, row_lower, row_upper);
    if (x0_out) std::copy(x0.begin(), x0.end(), x0_out);
    return 0;
}

extern "C" int vj_gen_uniform_ints(uint64_t seed, int count, int lo, int hi, int32_t* out) {
    if (count < 0 || hi < lo) return -1;
    Xoshiro rng(seed);
    const uint32_t span = static_cast<uint32_t>(hi - lo) + 1u;
    for (int i = 0; i < count; ++i) out[i] = lo + static_cast<int32_t>(rng.below(span));
    return 0;
}
