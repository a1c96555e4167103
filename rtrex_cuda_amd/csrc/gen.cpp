// Seeded synthetic graph generators for the benchmark workloads.
//
// Host C++ (built into librtrex_gen.so).  Every generator is deterministic for
// a given seed: the random stream is xoshiro256** seeded through splitmix64
// (both published by Blackman & Vigna).  Outputs are simple undirected edge
// lists (u < v, no self-loops, no duplicates), sorted by (u, v), with uint32
// vertex ids -- the shape the Escape loader consumes after the SNAP
// conversion rules (drop self-loops, dedupe, sorted pairs; RTRex_CUDA
// snap_to_rtr.py:81-82,112-142).  Optional float32 edge weights are produced
// as a sidecar; the decomposition path never reads them (the reference's
// loader parses exactly two integers per line, GraphIO.cpp:40-41).
//
// Models:
//   chung_lu : expected-degree power law w_i ~ (i + i0)^(-1/(gamma-1)) scaled to
//              the requested mean and capped at max_deg; m distinct edges are
//              drawn with endpoints proportional to w ("fast Chung-Lu").
//   planted  : communities of uniform random size in [cmin, cmax] over a random
//              vertex permutation; each intra pair present with probability p;
//              plus a Chung-Lu background of m_bg edges over all vertices.
//   rmat     : R-MAT(a, b, c, d) with ef * 2^scale samples, symmetrised,
//              deduplicated, self-loops dropped, vertex ids randomly permuted.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <unordered_set>
#include <vector>

namespace {

struct Rng {
    uint64_t s[4];
    static uint64_t splitmix(uint64_t& x) {
        uint64_t z = (x += 0x9e3779b97f4a7c15ULL);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
        return z ^ (z >> 31);
    }
    explicit Rng(uint64_t seed) {
        uint64_t x = seed;
        for (auto& w : s) w = splitmix(x);
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        const uint64_t r = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    double uniform() { return (next() >> 11) * 0x1.0p-53; }  // [0,1)
    uint64_t below(uint64_t k) { return k ? next() % k : 0; }
    double normal() {  // Box-Muller
        double u1 = 1.0 - uniform(), u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

inline uint64_t key(uint32_t a, uint32_t b) {
    if (a > b) std::swap(a, b);
    return (uint64_t(a) << 32) | b;
}

// LSD radix sort of 64-bit keys on the low `bits` bits.
void radix_sort(std::vector<uint64_t>& k, int bits) {
    std::vector<uint64_t> tmp(k.size());
    for (int sh = 0; sh < bits; sh += 11) {
        size_t cnt[2049] = {0};
        for (uint64_t x : k) cnt[((x >> sh) & 2047) + 1]++;
        for (int i = 0; i < 2048; ++i) cnt[i + 1] += cnt[i];
        for (uint64_t x : k) tmp[cnt[(x >> sh) & 2047]++] = x;
        k.swap(tmp);
    }
}

int bits_for(uint64_t n) {
    int b = 1;
    while ((1ULL << b) < n) ++b;
    return b;
}

// Sort + dedupe keys into (u, v) arrays.
int64_t emit(std::vector<uint64_t>& keys, uint64_t n, uint32_t* u, uint32_t* v,
             int64_t cap) {
    radix_sort(keys, 32 + bits_for(n + 1));
    keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
    if ((int64_t)keys.size() > cap) return -(int64_t)keys.size();
    for (size_t i = 0; i < keys.size(); ++i) {
        u[i] = uint32_t(keys[i] >> 32);
        v[i] = uint32_t(keys[i]);
    }
    return (int64_t)keys.size();
}

// Expected-degree weights for Chung-Lu: power law with exponent gamma, mean
// `mean`, capped at max_deg (rescaled after capping so the mean holds).
std::vector<double> cl_weights(uint64_t n, double gamma, double mean,
                               double max_deg) {
    std::vector<double> w(n);
    const double ex = -1.0 / (gamma - 1.0);
    for (uint64_t i = 0; i < n; ++i) w[i] = std::pow(double(i) + 1.0, ex);
    double scale = 1.0;
    for (int it = 0; it < 60; ++it) {
        double s = 0;
        for (uint64_t i = 0; i < n; ++i) s += std::min(w[i] * scale, max_deg);
        double cur = s / double(n);
        if (std::fabs(cur - mean) < 1e-9 * mean) break;
        scale *= mean / cur;
    }
    for (auto& x : w) x = std::min(x * scale, max_deg);
    return w;
}

struct Sampler {  // inverse-CDF sampling proportional to weights
    std::vector<double> cdf;
    explicit Sampler(const std::vector<double>& w) : cdf(w.size()) {
        double s = 0;
        for (size_t i = 0; i < w.size(); ++i) cdf[i] = (s += w[i]);
        for (auto& c : cdf) c /= s;
    }
    uint32_t draw(Rng& r) const {
        double x = r.uniform();
        size_t i = std::upper_bound(cdf.begin(), cdf.end(), x) - cdf.begin();
        return uint32_t(std::min(i, cdf.size() - 1));
    }
};

// Draws m_target distinct non-loop edges with endpoints ~ w (over ids map[i]).
// With hard_cap > 0 a draw is rejected when an endpoint already has that many
// edges from this call, so realised degrees never exceed the cap.
void chung_lu_into(std::unordered_set<uint64_t>& seen,
                   std::vector<uint64_t>& keys, const std::vector<double>& w,
                   const std::vector<uint32_t>* map, uint64_t m_target, Rng& r,
                   uint32_t hard_cap = 0) {
    Sampler smp(w);
    const uint64_t n = w.size();
    std::vector<uint32_t> dcount(hard_cap ? n : 0, 0);
    uint64_t added = 0, tries = 0, limit = 50 * m_target + 1000;
    const uint64_t max_pairs = n * (n - 1) / 2;
    if (m_target > max_pairs) m_target = max_pairs;
    while (added < m_target && tries < limit) {
        ++tries;
        uint32_t a = smp.draw(r), b = smp.draw(r);
        if (a == b) continue;
        if (hard_cap && (dcount[a] >= hard_cap || dcount[b] >= hard_cap))
            continue;
        const uint32_t ia = a, ib = b;
        if (map) {
            a = (*map)[a];
            b = (*map)[b];
        }
        uint64_t k = key(a, b);
        if (seen.insert(k).second) {
            keys.push_back(k);
            ++added;
            if (hard_cap) {
                dcount[ia]++;
                dcount[ib]++;
            }
        }
    }
}

std::vector<uint32_t> permutation(uint64_t n, Rng& r) {
    std::vector<uint32_t> p(n);
    std::iota(p.begin(), p.end(), 0u);
    for (uint64_t i = n; i > 1; --i) std::swap(p[i - 1], p[r.below(i)]);
    return p;
}

}  // namespace

extern "C" {

// Returns the number of edges written (>= 0), or -(needed) if cap is too small.
int64_t rtrex_gen_chung_lu(uint64_t n, uint64_t m, double gamma,
                           double max_deg, uint64_t seed, uint32_t* u,
                           uint32_t* v, int64_t cap) {
    if (n < 2) return 0;
    Rng r(seed);
    auto w = cl_weights(n, gamma, 2.0 * double(m) / double(n), max_deg);
    std::unordered_set<uint64_t> seen;
    seen.reserve(m * 2);
    std::vector<uint64_t> keys;
    keys.reserve(m);
    chung_lu_into(seen, keys, w, nullptr, m, r, uint32_t(max_deg));
    return emit(keys, n, u, v, cap);
}

int64_t rtrex_gen_planted(uint64_t n, uint64_t n_comm, uint32_t cmin,
                          uint32_t cmax, double p_in, double bg_gamma,
                          uint64_t m_bg, double bg_max_deg, uint64_t seed,
                          uint32_t* u, uint32_t* v, int64_t cap) {
    if (n < 2) return 0;
    Rng r(seed);
    auto perm = permutation(n, r);
    std::unordered_set<uint64_t> seen;
    std::vector<uint64_t> keys;
    uint64_t pos = 0;
    // n_comm == 0: partition every vertex into communities.
    for (uint64_t c = 0; (n_comm == 0 || c < n_comm) && pos < n; ++c) {
        uint64_t s = cmin + r.below(uint64_t(cmax) - cmin + 1);
        if (pos + s > n) s = n - pos;
        for (uint64_t i = 0; i < s; ++i)
            for (uint64_t j = i + 1; j < s; ++j)
                if (r.uniform() < p_in) {
                    uint64_t k = key(perm[pos + i], perm[pos + j]);
                    if (seen.insert(k).second) keys.push_back(k);
                }
        pos += s;
    }
    if (m_bg) {
        auto w = cl_weights(n, bg_gamma, 2.0 * double(m_bg) / double(n),
                            bg_max_deg);
        chung_lu_into(seen, keys, w, &perm, m_bg, r);
    }
    return emit(keys, n, u, v, cap);
}

int64_t rtrex_gen_rmat(uint32_t scale, uint32_t ef, double a, double b,
                       double c, uint64_t seed, uint32_t* u, uint32_t* v,
                       int64_t cap) {
    const uint64_t n = 1ULL << scale;
    const uint64_t samples = uint64_t(ef) << scale;
    Rng r(seed);
    auto perm = permutation(n, r);
    std::vector<uint64_t> keys;
    keys.reserve(samples);
    const double ab = a + b, abc = a + b + c;
    for (uint64_t e = 0; e < samples; ++e) {
        uint64_t x = 0, y = 0;
        for (uint32_t lvl = 0; lvl < scale; ++lvl) {
            double q = r.uniform();
            uint64_t bx = 0, by = 0;
            if (q < a) {
            } else if (q < ab) {
                by = 1;
            } else if (q < abc) {
                bx = 1;
            } else {
                bx = 1;
                by = 1;
            }
            x = (x << 1) | bx;
            y = (y << 1) | by;
        }
        if (x == y) continue;
        keys.push_back(key(perm[x], perm[y]));
    }
    return emit(keys, n, u, v, cap);
}

// kind 0: uniform in (0,1];  kind 1: log-normal(0,1) clipped to (0,10].
void rtrex_gen_weights(uint64_t m, int kind, uint64_t seed, float* w) {
    Rng r(seed ^ 0x5eedf00dULL);
    for (uint64_t i = 0; i < m; ++i) {
        if (kind == 0) {
            w[i] = float(1.0 - r.uniform());
            if (w[i] <= 0.0f) w[i] = 1e-7f;
        } else {
            double x = std::exp(r.normal());
            if (x > 10.0) x = 10.0;
            if (x <= 0.0) x = 1e-7;
            w[i] = float(x);
        }
    }
}

}  // extern "C"
