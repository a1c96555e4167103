// Order-independent certification of cleaning decisions.
//
// The reference's support of edge f at any moment is its count-phase value S0
// (bit-identical here) minus the weights of f's dead triangles, subtracted one
// by one in the order its LIFO cascade happened to kill them
// (Decomposition.h:217-223).  Any decision "support < eps/(d_u d_v)" the
// reference takes is therefore a property of (S0, D, order) with D the
// multiset of dead-triangle weights.  A decision is certain when it comes out
// the same for EVERY order of D:
//   1. E = S0 - sum(D) is formed exactly enough in double-double; every order
//      rounds |D| times by at most ulp(S0)/2 each (running values never exceed
//      S0), so |fold - E| <= |D| ulp(S0)/2 decides most cases;
//   2. otherwise, for |D| <= kMaxPerm every distinct order is evaluated with
//      the reference's own double subtractions.
// An undecided case is reported (ambiguous_clean) and resolved by this
// library's own value.
//
// Monotonicity (fl(a - t) <= fl(b - t) for a <= b, and fl(a - t) <= a for
// t > 0) makes "removed" = "some prefix of the kills drops below the
// threshold" = "the whole current set D does", which is what is certified.
#pragma once

#include "dev.hpp"
#include "internal.hpp"

namespace rtrex {

constexpr int kMaxPerm = 8;

enum : int { CERT_KEEP = 0, CERT_REMOVE = 1, CERT_UNDECIDED = 2 };

__device__ __forceinline__ double ulp_of(double x) {
    if (!(x > 0)) return 0.0;
    return ldexp(1.0, ilogb(x) - 52);
}

// Cheap test valid for any history of at most k subtractions from s0.  The
// reference's value lies within k ulp(s0)/2 of the exact difference (its
// subtractions round to nearest; running values never exceed s0).  This
// library's value v may come from FP64 atomics (global atomic add, built with
// -munsafe-fp-atomics), whose rounding is not relied upon: each is counted as
// off by up to a whole ulp(s0), so v lies within k ulp(s0) of the exact
// difference and the two within 3k/2 ulp(s0) of each other.
__device__ __forceinline__ double cheap_bound(uint32_t k, double s0) {
    return 1.5 * (double)k * ulp_of(s0) + 0x1.0p-1060;
}

__device__ __forceinline__ void two_sum(double a, double b, double& s,
                                        double& e) {
    s = a + b;
    double bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}

// Decision from E = s0 - sum(D) held as hi + lo (hi + lo_exact equal to E
// exactly, lo's own additions rounded; at most nd + 16 of them, each off by
// at most 2^-53 |lo| <= 2^-53 (nd + 16) ulp(s0) / 2), with nd = |D|: KEEP or
// REMOVE when every order of the subtractions lands on that side of thr, else
// CERT_UNDECIDED (the caller may enumerate the orders).
__device__ __forceinline__ int dd_decide(double hi, double lo, uint32_t nd, double s0,
                                         double thr) {
    const double u = ulp_of(s0);
    const double n = (double)nd;
    const double half =
        0.5 * u * (n * (1.0 + 0x1.0p-40) + (n + 16.0) * (n + 16.0) * 0x1.0p-52) + 0x1.0p-1060;
    double d, de;
    two_sum(hi, -thr, d, de);
    d += de + lo;
    if (d >= half) return CERT_KEEP;
    if (d < -half) return CERT_REMOVE;
    return CERT_UNDECIDED;
}

// Every order of N items as rows of indices (lexicographic), built at
// compile time for the unrolled enumeration below.
template <int N>
struct PermTable {
    static constexpr int rows() {
        int f = 1;
        for (int i = 2; i <= N; ++i) f *= i;
        return f;
    }
    uint8_t p[rows()][N];
    constexpr PermTable() : p{} {
        uint8_t a[N] = {};
        for (int i = 0; i < N; ++i) a[i] = (uint8_t)i;
        for (int r = 0; r < rows(); ++r) {
            for (int i = 0; i < N; ++i) p[r][i] = a[i];
            int i = N - 2;  // next lexicographic permutation of a
            while (i >= 0 && a[i] > a[i + 1]) --i;
            if (i < 0) break;
            int j = N - 1;
            while (a[j] < a[i]) --j;
            const uint8_t x = a[i];
            a[i] = a[j];
            a[j] = x;
            for (int l = i + 1, h = N - 1; l < h; ++l, --h) {
                const uint8_t y = a[l];
                a[l] = a[h];
                a[h] = y;
            }
        }
    }
};

// All N! orders of t[0..N) with static register indices (repeated values
// give repeated orders; at most 120 of them).
template <int N>
__device__ __forceinline__ int enumerate_unrolled(double s0, double thr, const double* src) {
    constexpr PermTable<N> P{};
    double t[N];
#pragma unroll
    for (int i = 0; i < N; ++i) t[i] = src[i];
    bool seen_lt = false, seen_ge = false;
#pragma unroll
    for (int r = 0; r < PermTable<N>::rows(); ++r) {
        double v = s0;
#pragma unroll
        for (int i = 0; i < N; ++i) v -= t[P.p[r][i]];
        if (v < thr) seen_lt = true;
        else seen_ge = true;
    }
    return seen_lt && seen_ge ? CERT_UNDECIDED : seen_lt ? CERT_REMOVE : CERT_KEEP;
}

// Combines two outcomes over disjoint sets of orders.
__device__ __forceinline__ int join_outcomes(int a, int b) {
    if (a < 0) return b;
    return a == b ? a : CERT_UNDECIDED;
}

// x[i] for a run-time i < N, by selects (no run-time register index).
template <int N>
__device__ __forceinline__ double pick(const double* x, int i) {
    double r = x[0];
#pragma unroll
    for (int j = 1; j < N; ++j) r = i == j ? x[j] : r;
    return r;
}

// Next lexicographic permutation of a[0..k) (ascending start); false at end.
__device__ __forceinline__ bool next_perm(double* a, int k) {
    int i = k - 2;
    while (i >= 0 && !(a[i] < a[i + 1])) --i;
    if (i < 0) return false;
    int j = k - 1;
    while (!(a[i] < a[j])) --j;
    double t = a[i];
    a[i] = a[j];
    a[j] = t;
    for (int l = i + 1, r = k - 1; l < r; ++l, --r) {
        t = a[l];
        a[l] = a[r];
        a[r] = t;
    }
    return true;
}

// Outcome over every order of subtracting src[0..nd) (nd <= kMaxPerm) from s0
// with the reference's plain double subtractions.  FAST: six to eight values
// go through run-time loops over the first values subtracted (the running
// value after them is the reference's own), down to five left, whose 120
// orders run unrolled -- every index static.  Otherwise six or more values
// are sorted and stepped through their distinct orders, in few registers (for
// kernels that stream, where these cases are rare).
template <bool FAST>
__device__ inline int enumerate_orders(double s0, double thr, const double* src, int nd) {
    switch (nd) {
        case 0: return s0 < thr ? CERT_REMOVE : CERT_KEEP;
        case 1: return enumerate_unrolled<1>(s0, thr, src);
        case 2: return enumerate_unrolled<2>(s0, thr, src);
        case 3: return enumerate_unrolled<3>(s0, thr, src);
        case 4: return enumerate_unrolled<4>(s0, thr, src);
        case 5: return enumerate_unrolled<5>(s0, thr, src);
        default: break;
    }
    double t[kMaxPerm];
#pragma unroll
    for (int i = 0; i < kMaxPerm; ++i) t[i] = i < nd ? src[i] : 0.0;
    if (FAST) {
        int out = -1;
        for (int a = 0; a < nd && out != CERT_UNDECIDED; ++a) {
            double q[7];  // the nd - 1 values after the first
#pragma unroll
            for (int j = 0; j < 7; ++j) q[j] = j < a ? t[j] : t[j + 1];
            const double sa = s0 - pick<8>(t, a);
            if (nd == 6) {
                out = join_outcomes(out, enumerate_unrolled<5>(sa, thr, q));
                continue;
            }
            for (int b = 0; b < nd - 1 && out != CERT_UNDECIDED; ++b) {
                double w[6];
#pragma unroll
                for (int j = 0; j < 6; ++j) w[j] = j < b ? q[j] : q[j + 1];
                const double sb = sa - pick<7>(q, b);
                if (nd == 7) {
                    out = join_outcomes(out, enumerate_unrolled<5>(sb, thr, w));
                    continue;
                }
                for (int c = 0; c < 6 && out != CERT_UNDECIDED; ++c) {  // nd == 8
                    double x[5];
#pragma unroll
                    for (int j = 0; j < 5; ++j) x[j] = j < c ? w[j] : w[j + 1];
                    out = join_outcomes(out, enumerate_unrolled<5>(sb - pick<6>(w, c), thr, x));
                }
            }
        }
        return out;
    }
    for (int i = 1; i < nd; ++i)  // insertion sort ascending
        for (int j = i; j > 0 && t[j] < t[j - 1]; --j) {
            const double x = t[j];
            t[j] = t[j - 1];
            t[j - 1] = x;
        }
    bool seen_lt = false, seen_ge = false;
    do {
        double v = s0;
        for (int i = 0; i < nd; ++i) v -= t[i];
        if (v < thr) seen_lt = true;
        else seen_ge = true;
        if (seen_lt && seen_ge) return CERT_UNDECIDED;
    } while (next_perm(t, nd));
    return seen_lt ? CERT_REMOVE : CERT_KEEP;
}

// Accumulates one dead triangle of weight t into the certification state.
struct CertAcc {
    double hi, lo;
    double t8[kMaxPerm];
    int nd;
    __device__ explicit CertAcc(double s0) : hi(s0), lo(0.0), nd(0) {}
    __device__ void add(double t) {
        double s, e;
        two_sum(hi, -t, s, e);
        hi = s;
        lo += e;
        if (nd < kMaxPerm) t8[nd] = t;
        ++nd;
    }
    // Outcome over every order of the accumulated subtractions from s0.
    __device__ int finish(double s0, double thr) {
        const int c = dd_decide(hi, lo, (uint32_t)nd, s0, thr);
        if (c != CERT_UNDECIDED || nd > kMaxPerm) return c;
        return enumerate_orders<false>(s0, thr, t8, nd);
    }
};

__device__ __forceinline__ bool tri_dead(uint8_t sa, uint8_t sb, bool qdead) {
    bool dead = sa == ST_DEAD || sb == ST_DEAD;
    if (qdead) dead = dead || sa == ST_QUEUED || sb == ST_QUEUED;
    return dead;
}

// Certifies "support(f) < thr" for edge f = {x, y} with count-phase support
// s0.  A triangle {x, y, z} counts as dead when one of its other two edges is
// DEAD or QUEUED (qdead) / DEAD only (!qdead), or when z == extra_dead (the
// triangle shared with an edge being deleted right now).  Runs on one lane;
// enumerates f's triangles by merging the two sorted lists.
__device__ int certify(const GraphDev& g, uint32_t x, uint32_t y, double s0,
                       double thr, bool qdead, uint32_t extra_dead) {
    const uint32_t dx = g.deg[x], dy = g.deg[y];
    uint32_t a = g.off[x], ae = g.off[x + 1];
    uint32_t b = g.off[y], be = g.off[y + 1];
    CertAcc acc(s0);
    while (a < ae && b < be) {
        const uint32_t za = g.nb[a], zb = g.nb[b];
        if (za < zb) {
            ++a;
        } else if (zb < za) {
            ++b;
        } else {
            if (tri_dead(g.st[a], g.st[b], qdead) || za == extra_dead)
                acc.add(tri_w(dx, dy, g.deg[za]));
            ++a;
            ++b;
        }
    }
    return acc.finish(s0, thr);
}

// Same, reading edge f's stored triangle list; the triangle shared with edge
// extra_edge (being deleted right now) counts as dead.
__device__ int certify_terms(const GraphDev& g, const TermBuf& tb, uint32_t f,
                             double thr, bool qdead, uint32_t extra_edge) {
    const double s0 = g.sup0[f];
    const uint64_t k0 = tb.zoff[f], k1 = k0 + g.tcnt[f];
    CertAcc acc(s0);
    for (uint64_t k = k0; k < k1; ++k) {
        const Term q = tb.t[k];
        if (tri_dead(g.est[q.sa], g.est[q.sb], qdead) || q.sa == extra_edge || q.sb == extra_edge)
            acc.add(q.w);
    }
    return acc.finish(s0, thr);
}

// certify_terms by a whole wave: every lane calls it with the same arguments;
// s_t / s_n are kMaxPerm doubles and a counter in this wave's LDS.  The lanes
// walk the list U records per step (edge ids, then statuses, loaded before
// use), each sums its dead weights in double-double (exact partial sums) and
// gathers them into s_t while there are at most kMaxPerm; the partials are
// added pairwise across the wave (every lane ends with the same sum), and
// every lane decides from s0 + sum, enumerating the orders of the gathered
// weights within the error band.  A certain outcome is certify_terms'.
__device__ inline int certify_terms_wave(const GraphDev& g, const TermBuf& tb, uint32_t f,
                                         double thr, bool qdead, uint32_t extra_edge,
                                         double* s_t, uint32_t* s_n) {
    const int lane = lane_id();
    const double s0 = g.sup0[f];
    const uint64_t k0 = tb.zoff[f], k1 = k0 + g.tcnt[f];
    if (lane == 0) *s_n = 0;
    wave_sync();
    double hi = 0.0, lo = 0.0;
    uint32_t nd = 0;
    constexpr int U = 2;
    for (uint64_t k = k0 + lane; k < k1; k += (uint64_t)U * WAVE) {
        uint32_t fa[U], fb[U];
        double tw[U];
        bool dead[U];
#pragma unroll
        for (int r = 0; r < U; ++r) {
            const uint64_t q = k + (uint64_t)r * WAVE;
            const Term rec = q < k1 ? tb.t[q] : Term{0.0, 0u, 0u};
            fa[r] = rec.sa;
            fb[r] = rec.sb;
            tw[r] = rec.w;
        }
#pragma unroll
        for (int r = 0; r < U; ++r) {
            const uint64_t q = k + (uint64_t)r * WAVE;
            dead[r] = q < k1 && (tri_dead(g.est[fa[r]], g.est[fb[r]], qdead) ||
                                 fa[r] == extra_edge || fb[r] == extra_edge);
        }
#pragma unroll
        for (int r = 0; r < U; ++r) {
            if (!dead[r]) continue;
            double sn, e;
            two_sum(hi, -tw[r], sn, e);
            hi = sn;
            lo += e;
            ++nd;
            const uint32_t slot = atomicAdd(s_n, 1u);
            if (slot < (uint32_t)kMaxPerm) s_t[slot] = tw[r];
        }
    }
#pragma unroll
    for (int o = 1; o < WAVE; o <<= 1) {
        const double h2 = __shfl_xor(hi, o, WAVE);
        const double l2 = __shfl_xor(lo, o, WAVE);
        nd += __shfl_xor(nd, o, WAVE);
        double sn, e;
        two_sum(hi, h2, sn, e);
        hi = sn;
        lo = lo + l2 + e;
    }
    wave_sync();  // the gathered weights are visible to every lane
    double sn, e;
    two_sum(s0, hi, sn, e);
    int c = dd_decide(sn, lo + e, nd, s0, thr);
    if (c == CERT_UNDECIDED && nd <= (uint32_t)kMaxPerm)
        c = enumerate_orders<true>(s0, thr, s_t, (int)nd);
    wave_sync();  // s_t and s_n are read before the next call resets them
    return c;
}

// Either form, by what the graph keeps.  The extra dead triangle is named by
// its third vertex (intersection form) and by the edge being deleted (stored
// lists).
__device__ __forceinline__ int certify_edge(const GraphDev& g, const TermBuf& tb,
                                            uint32_t f, double thr, bool qdead,
                                            uint32_t extra_vertex, uint32_t extra_edge) {
    if (tb.t) return certify_terms(g, tb, f, thr, qdead, extra_edge);
    const uint32_t sl = g.eslot[f];
    return certify(g, g.esrc[f], g.nb[sl], g.sup0[f], thr, qdead, extra_vertex);
}

}  // namespace rtrex
