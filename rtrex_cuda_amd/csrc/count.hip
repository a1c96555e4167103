// Weighted triangle supports (the count phase).
//
// Reference: commonNbr (Decomposition.h:42-117) / triangleCountKernel +
// gpuCommonNbr (RTRex_CUDA/gpu_triangle_count.cu:15-140).
//
// The reference finds every triangle once, in lexicographic (a < b < c) order,
// and adds its weight t = 1/(d_a d_b d_c) to six edge slots and three vertices.
// Reading off that order per accumulator gives two fixed fold orders:
//   edge {x,y}:  left fold from 0 over z in N(x) ∩ N(y), ascending z;
//   vertex v:    left fold from 0 over y in N(v) ascending, then
//                z in N(v) ∩ N(y) with z > y ascending.
// Both are reproduced here exactly, so supports are bit-identical to the
// sequential reference, with no atomics on FP values:
//   k_count_edges<G>:    G lanes per undirected edge.  The shorter list is
//       streamed in chunks of G ascending ids, each lane probes the longer list
//       (binary search, window narrowed chunk by chunk), and the matches of a
//       chunk are folded in lane order (ballot + lane shuffle).
//   The matched weights of every edge are also stored, ascending z, in a term
//   buffer (capacity min(d_x, d_y) - 1 per edge), with the number of matches
//   below x and below y.  Vertex v's fold is then the concatenation, over y in
//   N(v) ascending, of the stored terms of edge {v, y} with z > y -- a
//   streaming read folded serially (k_vfold_lane / k_vfold_wave), instead of
//   a second round of intersections.
//   k_count_vertices<G>: fallback when the term buffer does not fit in HBM:
//       G lanes per vertex walking y in N(v) in order, each step a chunked
//       intersection restricted to ids > y.
// Edges and vertices are bucketed by work (shorter-list length, degree) so
// that G matches the work: 1, 4, 16 or 64 lanes.
#include <algorithm>

#include "dev.hpp"
#include "internal.hpp"

namespace rtrex {
namespace {

constexpr int T = 256;
constexpr int NB = 4;  // work buckets: <= 4, <= 16, <= 96, larger

__device__ __forceinline__ int bucket_of(uint32_t w) {
    return w <= 4u ? 0 : w <= 16u ? 1 : w <= 96u ? 2 : 3;
}

// A term written once here and read only by later phases: a streaming
// (nontemporal) store, so the term buffer does not push the lists the count
// kernels re-read out of the L2.
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ void store_term(Term* p, double w, uint32_t a, uint32_t b) {
    const uint64_t wb = (uint64_t)__double_as_longlong(w);
    const u32x4 v = {(uint32_t)wb, (uint32_t)(wb >> 32), a, b};
    __builtin_nontemporal_store(v, reinterpret_cast<u32x4*>(p));
}

// Hub classes (CountParams).  Edge {x, y} streams the list of
// lo = (d_x <= d_y ? x : y) (x < y) and probes the list of the other
// endpoint, its hub.  Edges whose hub degree is in (hub_min, hub_max] are
// counted per hub (k_count_hub*), the hub's list staged once in LDS; the rest
// per edge (k_count_edges).
constexpr int HUB_NL = 4, HUB_NG = 5;  // staging classes 64/512/4096/32768; G = 4..64
// Owned edges per unit, at most, in the large and huge classes.  A huge
// unit restages the tree and the filter of a hub of up to 32760 entries, so
// its units are larger (R-MAT scale 21 count: 123 -> 110 ms with 128 instead
// of 32); 32 keeps the few large hubs of planted graphs spread over the CUs.
constexpr uint32_t LARGE_UNIT = 32, HUGE_UNIT = 128;
// A unit's w field: d_hub in the low UNIT_DH_BITS bits (up to HUB_MAX_LIMIT),
// its owned-edge count above (at most d_hub, and at most 4096 below the huge
// class, whose units hold at most HUGE_UNIT).
constexpr uint32_t UNIT_DH_BITS = 18, UNIT_DH_MASK = (1u << UNIT_DH_BITS) - 1u;
static_assert(HUB_MAX_LIMIT <= UNIT_DH_MASK && HUGE_UNIT < (1u << (32 - UNIT_DH_BITS)) &&
                  4096u < (1u << (32 - UNIT_DH_BITS)),
              "a unit's degree and edge count fit its w field");
__host__ __device__ __forceinline__ bool hub_edge(uint32_t dhub, const CountParams& P) {
    return dhub > P.hub_min && dhub <= P.hub_max;
}
__host__ __device__ __forceinline__ int hub_lclass(uint32_t dhub, const CountParams& P) {
    return dhub <= P.l0_max ? 0 : dhub <= P.l1_max ? 1 : dhub <= P.l2_max ? 2 : 3;
}
// Medium and large hubs whose edges' shorter lists are short on average are
// counted by the streaming kernel (with a term buffer); long ones (over
// g_thr[3] elements on average, R-MAT's high-degree pairs) keep 64 lanes per
// edge.
__host__ __device__ __forceinline__ bool hub_stream_class(int lc, int gc) {
    return (lc == 1 || lc == 2) && gc < HUB_NG - 1;
}
// Lanes per edge from the mean shorter-list length of the hub's edges.
__host__ __device__ __forceinline__ int hub_gclass(uint64_t work, uint32_t edges,
                                                   const CountParams& P) {
    const uint64_t mean = (work + edges - 1) / edges;
    return mean <= P.g_thr[0] ? 0 : mean <= P.g_thr[1] ? 1 : mean <= P.g_thr[2] ? 2
           : mean <= P.g_thr[3] ? 3 : 4;
}

// Appends ids to per-bucket lists: work(e) = length of the shorter list for
// edges, the degree for vertices.
template <bool EDGES>
__global__ void __launch_bounds__(T) k_bucket(GraphDev g, uint32_t count,
                                              uint32_t* __restrict__ lists,
                                              uint32_t* __restrict__ cursor, CountParams P) {
    __shared__ uint32_t cnt[NB], base[NB];
    if (threadIdx.x < NB) cnt[threadIdx.x] = 0;
    __syncthreads();
    uint32_t i = blockIdx.x * T + threadIdx.x;
    int b = -1;
    uint32_t r = 0;
    if (i < count) {
        uint32_t w;
        if (EDGES) {
            uint32_t x = g.esrc[i], y = g.nb[g.eslot[i]];
            w = min(g.deg[x], g.deg[y]);
            if (hub_edge(max(g.deg[x], g.deg[y]), P)) w = 0xffffffffu;  // k_count_hub
        } else {
            w = g.deg[i];
        }
        if (w != 0xffffffffu) {
            b = bucket_of(w);
            r = atomicAdd(&cnt[b], 1u);
        }
    }
    __syncthreads();
    if (threadIdx.x < NB)
        base[threadIdx.x] = atomicAdd(&cursor[threadIdx.x], cnt[threadIdx.x]);
    __syncthreads();
    if (b >= 0) lists[(uint64_t)b * count + base[b] + r] = i;
}

// Folds, in lane order, the values t of the lanes of this lane's group whose
// hit flag is set.  All 64 lanes must call it.
__device__ __forceinline__ double readlane_f64(double x, int l) {
    const uint64_t b = __double_as_longlong(x);
    const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)b, l);
    const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(b >> 32), l);
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

template <int G>
__device__ __forceinline__ void group_fold(bool hit, double t, double& S,
                                           uint32_t& c) {
    if (G == 64) {
        // One group: the hit mask is wave-uniform, so each hit's value is
        // read with a scalar lane index (no LDS permute).
        uint64_t mask = ballot64(hit);
        c += (uint32_t)__popcll(mask);
        while (mask) {
            const int b = __ffsll((unsigned long long)mask) - 1;
            S += readlane_f64(t, b);
            mask &= mask - 1;
        }
        return;
    }
    const int lane = lane_id();
    const int gbase = lane & ~(G - 1);
    uint64_t bal = ballot64(hit);
    uint64_t mask = G == 64 ? bal : (bal >> gbase) & ((1ull << G) - 1);
    c += (uint32_t)__popcll(mask);
    while (__any(mask != 0)) {
        int b = mask ? __ffsll((unsigned long long)mask) - 1 : 0;
        double val = __shfl(t, mask ? gbase + b : lane, 64);
        if (mask) {
            S += val;
            mask &= mask - 1;
        }
    }
}
// One step of a chunked intersection: lanes of a group take G consecutive
// ids L[p .. p+G) of the streamed list and look each up in H[hlo, he).  On
// return hlo has advanced to the insertion point of the chunk's last id.
template <int G>
__device__ __forceinline__ bool probe_chunk(const uint32_t* __restrict__ nb,
                                            bool act, uint32_t p, uint32_t le,
                                            uint32_t& hlo, uint32_t he,
                                            uint32_t& z, uint32_t* hitpos = nullptr) {
    const int lane = lane_id();
    const int gl = lane & (G - 1);
    const int gbase = lane & ~(G - 1);
    bool in = act && p + gl < le;
    z = in ? nb[p + gl] : 0;
    uint32_t pos = in ? lower_bound(nb, hlo, he, z) : hlo;
    bool hit = in && pos < he && nb[pos] == z;
    if (hitpos) *hitpos = pos;
    // The group's last in-range lane holds the largest id of the chunk.
    uint32_t lastl = act ? min((uint32_t)(G - 1), le > p ? le - p - 1 : 0u) : 0u;
    uint32_t nh = __shfl(pos, gbase + lastl, 64);
    if (act) hlo = nh;
    return hit;
}


template <int G>
__global__ void __launch_bounds__(T) k_count_edges(GraphDev g,
                                                   const uint32_t* __restrict__ list,
                                                   uint32_t cnt,
                                                   TermBuf tb) {
    const int lane = lane_id();
    const int gl = lane & (G - 1);
    const int gbase = lane & ~(G - 1);
    const uint64_t below = ((1ull << gl) - 1ull);  // lanes of the group before me
    const uint32_t per_wave = WAVE / G;
    const uint32_t wave = (blockIdx.x * T + threadIdx.x) / WAVE;
    const uint32_t nwaves = gridDim.x * T / WAVE;
        for (uint32_t base = wave * per_wave; base < cnt; base += nwaves * per_wave) {
        const uint32_t idx = base + lane / G;
        bool act = idx < cnt;
        uint32_t e = act ? list[idx] : 0;
        uint32_t x = act ? g.esrc[e] : 0;
        uint32_t y = act ? g.nb[g.eslot[e]] : 0;
        uint32_t dx = act ? g.deg[x] : 0, dy = act ? g.deg[y] : 0;
        uint32_t lo = dx <= dy ? x : y, hi = dx <= dy ? y : x;
        uint32_t p = act ? g.off[lo] : 0, le = act ? g.off[lo + 1] : 0;
        uint32_t hlo = act ? g.off[hi] : 0, he = act ? g.off[hi + 1] : 0;
        const uint64_t wbase = (act && tb.t) ? tb.zoff[e] : 0;
        double S = 0.0;
        uint32_t c = 0;
        while (__any(act && p < le)) {
            bool go = act && p < le;
            uint32_t z, hp;
            bool hit = probe_chunk<G>(g.nb, go, p, le, hlo, he, z, &hp);
            double t = hit ? tri_w(dx, dy, g.deg[z]) : 0.0;
            if (tb.t) {
                const uint64_t hm = G == 64 ? ballot64(hit)
                                            : (ballot64(hit) >> gbase) & ((1ull << G) - 1);
                if (hit)  // edges {lo, z} and {hi, z}
                    store_term(tb.t + wbase + c + __popcll(hm & below), t, g.s2e[p + gl], g.s2e[hp]);
            }
            group_fold<G>(hit, t, S, c);
            if (go) p += G;
        }
        if (act && gl == 0) tb.rec[e] = SupRec{S, c, 0u};
    }
}

// Triangle incidences sum_e tcnt(e) (three per triangle): per-block partial
// sums, then one block adds them (no contended atomics).  The same stream
// spreads the counting kernels' per-edge records (support and count, one
// scattered 16-byte store per edge) into sup, sup0 (the count-phase value the
// clean starts from) and tcnt, in edge order.
constexpr int RED_BLOCKS = 1024;
__global__ void __launch_bounds__(T) k_sum_tcnt(const SupRec* __restrict__ rec,
                                                double* __restrict__ sup,
                                                double* __restrict__ sup0,
                                                uint32_t* __restrict__ tcnt, uint32_t m,
                                                unsigned long long* __restrict__ part) {
    __shared__ unsigned long long sw[T / WAVE];
    unsigned long long acc = 0;
    for (uint32_t e = blockIdx.x * T + threadIdx.x; e < m; e += gridDim.x * T) {
        const SupRec r = rec[e];
        acc += r.c;
        tcnt[e] = r.c;
        sup[e] = r.s;
        sup0[e] = r.s;
    }
    for (int d = 32; d >= 1; d >>= 1) acc += __shfl_xor(acc, d, 64);
    if (lane_id() == 0) sw[threadIdx.x / WAVE] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (int w = 0; w < T / WAVE; ++w) t += sw[w];
        part[blockIdx.x] = t;
    }
}

__global__ void __launch_bounds__(T) k_sum_final(const unsigned long long* __restrict__ part,
                                                 uint32_t nparts, unsigned long long* out) {
    __shared__ unsigned long long sw[T / WAVE];
    unsigned long long c = 0;
    for (uint32_t i = threadIdx.x; i < nparts; i += T) c += part[i];
    for (int d = 32; d >= 1; d >>= 1) c += __shfl_xor(c, d, 64);
    if (lane_id() == 0) sw[threadIdx.x / WAVE] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (int w = 0; w < T / WAVE; ++w) t += sw[w];
        *out = t;
    }
}

// ---------------------------------------------------------------- hub path
// Ownership flags: the hub-side slot of every hub-path edge; hwork[h] = the
// streamed work of hub h (sum of its owned edges' shorter lists, <= 32768^2).
__global__ void __launch_bounds__(T) k_own_flags(GraphDev g, uint32_t* __restrict__ flag,
                                                 uint32_t* __restrict__ hwork, CountParams P) {
    for (uint32_t e = blockIdx.x * T + threadIdx.x; e < g.m; e += gridDim.x * T) {
        const uint32_t sx = g.eslot[e], x = g.esrc[e], y = g.nb[sx];
        const uint32_t dx = g.deg[x], dy = g.deg[y];
        if (!hub_edge(max(dx, dy), P)) continue;
        flag[dx <= dy ? g.part[sx] : sx] = 1u;  // hub y (slot in N(y)) or hub x
        atomicAdd(&hwork[dx <= dy ? y : x], min(dx, dy));
    }
}

// own_idx[pos[j]] = local index of hub-side slot j in its hub's list;
// own_off[h] = pos[off[h]].
__global__ void __launch_bounds__(T) k_own_scatter(GraphDev g, const uint32_t* __restrict__ pos,
                                                   uint32_t* __restrict__ own_off,
                                                   uint32_t* __restrict__ own_idx, CountParams P) {
    for (uint32_t e = blockIdx.x * T + threadIdx.x; e < g.m; e += gridDim.x * T) {
        const uint32_t sx = g.eslot[e], x = g.esrc[e], y = g.nb[sx];
        const uint32_t dx = g.deg[x], dy = g.deg[y];
        if (!hub_edge(max(dx, dy), P)) continue;
        const uint32_t h = dx <= dy ? y : x;
        const uint32_t j = dx <= dy ? g.part[sx] : sx;
        own_idx[pos[j]] = j - g.off[h];
    }
    for (uint32_t v = blockIdx.x * T + threadIdx.x; v <= g.n; v += gridDim.x * T)
        own_off[v] = pos[g.off[v]];
}

// Work units of every hub: P parts of its owned edges, P from the streamed
// work (sum of the shorter lists), each part worth at least max(1024, d_h)
// elements so that staging the hub list is amortised.  COUNT: per-class
// totals; otherwise the units are written (cursor per class).
template <bool COUNT>
__global__ void __launch_bounds__(T) k_hub_units(GraphDev g, const uint32_t* __restrict__ own_off,
                                                 const uint32_t* __restrict__ hwork,
                                                 uint32_t* __restrict__ cls_cnt,
                                                 uint4* __restrict__ units,
                                                 const uint32_t* __restrict__ cls_base,
                                                 CountParams Pm, uint32_t* __restrict__ cls_max) {
    // One thread per hub; class totals (and, COUNT, largest hub degrees)
    // aggregated per block (LDS counters, one global atomic per class per
    // block).
    constexpr int NC = HUB_NL * HUB_NG;
    __shared__ uint32_t lc[NC], lb[NC], lm[NC];
    if (threadIdx.x < NC) {
        lc[threadIdx.x] = 0;
        lm[threadIdx.x] = 0;
    }
    __syncthreads();
    const uint32_t h = blockIdx.x * T + threadIdx.x;
    int c = -1;
    uint64_t P = 0;
    uint32_t at = 0;
    if (h < g.n) {
        const uint32_t oc = own_off[h + 1] - own_off[h];
        if (oc) {
            const uint32_t dh = g.deg[h];
            const uint64_t work = hwork[h];
            // Units of at least max(unit_min, d_h) elements (small hubs); the
            // medium and large classes restage their tree per unit:
            // max(2 d_h, 4 unit_min) (a unit feeds four waves); the huge class
            // unit_hash_mult * d_h, but at most LARGE_UNIT (huge: HUGE_UNIT)
            // owned edges apart so that hubs still spread out.
            // Classes whose shorter lists average over g_thr[3] elements keep
            // the lanes-per-edge kernels and their units (below); the others
            // of the medium and large classes stream (hub_stream_class).
            const int lc_ = hub_lclass(dh, Pm), gc_ = hub_gclass(work, oc, Pm);
            const bool stream = hub_stream_class(lc_, gc_);
            const uint64_t uw = lc_ == 0 ? max<uint64_t>(Pm.unit_min, dh)
                                : stream ? max<uint64_t>(2ull * dh, 4ull * Pm.unit_min)
                                : lc_ == 1 ? max<uint64_t>(Pm.unit_min, dh)
                                           : (uint64_t)Pm.unit_hash_mult * dh;
            P = (work + uw - 1) / uw;
            const uint32_t ucap = lc_ == 3 ? HUGE_UNIT : LARGE_UNIT;
            if (lc_ >= 2 && !stream) P = max<uint64_t>(P, (oc + ucap - 1) / ucap);
            P = max<uint64_t>(1, min<uint64_t>(P, min<uint64_t>(oc, 65535)));
            c = lc_ * HUB_NG + gc_;
            at = atomicAdd(&lc[c], (uint32_t)P);
            if (COUNT) atomicMax(&lm[c], dh);
        }
    }
    __syncthreads();
    if (threadIdx.x < NC && lc[threadIdx.x]) {
        lb[threadIdx.x] = atomicAdd(&cls_cnt[threadIdx.x], lc[threadIdx.x]);
        if (COUNT) atomicMax(&cls_max[threadIdx.x], lm[threadIdx.x]);
    }
    __syncthreads();
    if (!COUNT && c >= 0) {
        const uint32_t o0 = own_off[h], oc = own_off[h + 1] - o0;
        const uint32_t oh = g.off[h], dh = g.deg[h];
        for (uint64_t p = 0; p < P; ++p) {
            const uint32_t k0 = o0 + (uint32_t)((uint64_t)oc * p / P);
            const uint32_t k1 = o0 + (uint32_t)((uint64_t)oc * (p + 1) / P);
            units[cls_base[c] + lb[c] + at + p] =
                make_uint4(h, oh, k0, dh | ((k1 - k0) << UNIT_DH_BITS));
        }
    }
}

// Hub kernels: one unit (part of one hub's owned edges) at a time, per wave
// (small hubs) or per workgroup (larger ones).  The hub's sorted list is
// staged in LDS once per unit; its owned edges are counted G lanes per edge:
// the shorter list (with its slots' edge ids) is streamed ascending in chunks
// of G, every id is looked up in the staged list, and the matches of a chunk
// are folded in lane order.
constexpr int HT = 256;

// Hub-path work unit: part of one hub's owned edges.  x = hub, y = offset of
// its list, z = first own_idx entry, w = d_hub | (edge count << UNIT_DH_BITS).
__device__ __forceinline__ void unit_fields(const uint4& U, uint32_t& h, uint32_t& oh,
                                            uint32_t& k0, uint32_t& dh, uint32_t& cnt) {
    h = U.x;
    oh = U.y;
    k0 = U.z;
    dh = U.w & UNIT_DH_MASK;
    cnt = U.w >> UNIT_DH_BITS;
}

// Probe structure of a hub list staged in LDS: an 8-ary search tree.  The
// list is padded with UINT32_MAX (never equal to, never below, a vertex id)
// and stored whole (level 0, L = 8^(LV+1) entries); level l holds the last
// entry of every block of 8 of level l - 1; the top 8 pivots (the last entry
// of each eighth of the list) live in registers.  A probe counts the top
// pivots below z, then reads one 8-entry block per level (two 16-byte LDS
// reads) and ends on the only leaf block that can hold z: LV dependent LDS
// reads instead of log2(L) for a binary search, and no hash table to build.
struct Piv8 {
    uint32_t p[8];
};
constexpr uint32_t PAD_ID = 0xffffffffu;

__device__ __forceinline__ uint32_t below8(const uint4& a, const uint4& b, uint32_t z) {
    return (a.x < z) + (a.y < z) + (a.z < z) + (a.w < z) + (b.x < z) + (b.y < z) + (b.z < z) +
           (b.w < z);
}

// leaf: L entries, l1: L/8 pivots, l2: L/64 pivots; all 16-byte aligned.
// True and pos = index of z when z is in the list.
template <int LV>
__device__ __forceinline__ bool tree_probe(const uint32_t* leaf, const uint32_t* l1,
                                           const uint32_t* l2, const Piv8& top, uint32_t z,
                                           uint32_t& pos) {
    // Pivots below z, by a three-step binary search over the sorted top
    // pivots (the count of p[0..6] below z, at most 7); z above every id
    // when p[7] < z.
    // (Every pivot is chosen by selects on the comparison bits, so the
    // search stays branch-free.)
    const bool ok = !(top.p[7] < z);
    const bool c2 = top.p[3] < z;
    const bool c1 = (c2 ? top.p[5] : top.p[1]) < z;
    const uint32_t q1 = c2 ? top.p[6] : top.p[2];
    const uint32_t q0 = c2 ? top.p[4] : top.p[0];
    const bool c0 = (c1 ? q1 : q0) < z;
    uint32_t b = (c2 ? 4u : 0u) | (c1 ? 2u : 0u) | (c0 ? 1u : 0u);
    if (LV >= 3) {
        const uint4* q = reinterpret_cast<const uint4*>(l2) + 2 * b;
        b = 8 * b + min(below8(q[0], q[1], z), 7u);
    }
    if (LV >= 2) {
        const uint4* q = reinterpret_cast<const uint4*>(l1) + 2 * b;
        b = 8 * b + min(below8(q[0], q[1], z), 7u);
    }
    const uint4* q = reinterpret_cast<const uint4*>(leaf) + 2 * b;
    const uint4 a = q[0], c = q[1];
    // every comparison evaluated (a ?: chain lets the compiler read the block
    // entry by entry behind branches): the match's bit, then its position
    const uint32_t mk = (uint32_t)(a.x == z) | (uint32_t)(a.y == z) << 1 |
                        (uint32_t)(a.z == z) << 2 | (uint32_t)(a.w == z) << 3 |
                        (uint32_t)(c.x == z) << 4 | (uint32_t)(c.y == z) << 5 |
                        (uint32_t)(c.z == z) << 6 | (uint32_t)(c.w == z) << 7;
    pos = 8 * b + (mk ? (uint32_t)__ffs(mk) - 1u : 0u);
    return ok && mk != 0;
}

// The probe of a staged tree as a callable for hub_edges.
template <int LV>
struct TreeProbe {
    const uint32_t *leaf, *l1, *l2;
    Piv8 top;
    __device__ __forceinline__ bool operator()(uint32_t z, uint32_t& pos) const {
        return tree_probe<LV>(leaf, l1, l2, top, z, pos);
    }
};

// Entry i of the padded hub list (global loads; the pivots of the tree).
__device__ __forceinline__ uint32_t hub_id(const GraphDev& g, uint32_t oh, uint32_t dh,
                                           uint32_t i) {
    return i < dh ? g.nb[oh + i] : PAD_ID;
}

// Per-slot data of the hub list for the probe and the matches: the edge id
// of {hub, z}, d_z and z's list offset.  META: staged in LDS (sed, sdg, sof;
// small and medium hubs); otherwise read from the per-slot arrays in HBM at
// oh (large hubs, whose 4096-entry staging would leave no room for other
// workgroups).
template <bool META, bool HIT = META>
struct HubMeta {
    // HIT: the edge ids and degrees a match needs are in LDS (sed, sdg) even
    // when the owned edges' data is read from HBM.
    const uint32_t *sed, *sdg, *sof;
    const uint64_t* szo;
    uint32_t oh;
    __device__ __forceinline__ uint32_t ed(const GraphDev& g, uint32_t i) const {
        return META || HIT ? sed[i] : g.nbe[oh + i].y;
    }
    __device__ __forceinline__ uint2 dego(const GraphDev& g, uint32_t i) const {
        return META ? make_uint2(sdg[i], sof[i]) : g.ndo[oh + i];
    }
    __device__ __forceinline__ uint32_t deg(const GraphDev& g, uint32_t i) const {
        return META || HIT ? sdg[i] : g.ndo[oh + i].x;
    }
    // Term offset of the edge of hub slot i.
    __device__ __forceinline__ uint64_t zo(const TermBuf& tb, uint32_t i) const {
        return META ? szo[i] : tb.zsl[oh + i];
    }
};

// The owned edges [first, cnt) of a unit, stride apart, G lanes per edge:
// each group streams its edge's shorter list from HBM (R chunks of G ids per
// load, neighbour and edge id in one 8-byte load) and looks every id up in
// the hub's tree; own[] holds the owned slots (LDS or HBM).  Matches go to
// the term buffer in ascending z; the support is folded in lane order.
// sfold is this wave's 64-entry LDS slice.
// PRE: the owned edges' data is prefetched per owned position (pre: edge id,
// shorter-list length and offset; prez: term offset) instead of read through
// own[] and the hub's slots.
template <int G, class PR, class HM, bool PRE = false>
__device__ __forceinline__ void hub_edges(const GraphDev& g, const TermBuf& tb, uint32_t dh,
                                          const PR& probe,
                                          const HM& hm_, const uint32_t* own,
                                          double* sfold, uint32_t first, uint32_t cnt,
                                          uint32_t stride, const uint4* pre = nullptr,
                                          const uint64_t* prez = nullptr) {
    constexpr int R = 4;  // chunks of G ids per group and load round
    const int lane = lane_id();
    const int gl = lane & (G - 1);
    const int gbase = lane & ~(G - 1);
    const uint64_t below = ((1ull << gl) - 1ull);
    for (uint32_t base = first; base < cnt; base += stride) {
        const uint32_t idx = base + lane / G;
        const bool act = idx < cnt;
        uint32_t e;  // edge {hub, lo}
        uint2 yo;
        uint64_t wb0;
        if (PRE) {
            const uint4 q = pre[act ? idx : 0];
            e = q.x;
            yo = make_uint2(q.y, q.z);
            wb0 = prez[act ? idx : 0];
        } else {
            const uint32_t li = act ? own[idx] : 0;
            e = hm_.ed(g, li);
            yo = hm_.dego(g, li);
            wb0 = (act && tb.t) ? hm_.zo(tb, li) : 0;
        }
        const uint32_t dy = yo.x;
        uint32_t p = yo.y;
        const uint32_t le = act ? p + dy : p;
        const uint64_t wbase = (act && tb.t) ? wb0 : 0;
        double S = 0.0;
        uint32_t c = 0;
        while (__any(p < le)) {
            uint32_t zr[R], er[R];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const uint32_t q = p + r * G + gl;
                const uint2 a = q < le ? g.nbe[q] : make_uint2(0u, 0u);
                zr[r] = a.x;
                er[r] = a.y;  // edge {lo, z}
            }
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const bool go = p < le;
                const bool in = go && p + gl < le;
                const uint32_t z = zr[r];
                uint32_t pos = 0;
                const bool hit = in && probe(z, pos);
                const double t = hit ? tri_w(dh, dy, hm_.deg(g, pos)) : 0.0;
                const uint64_t hm = G == 64 ? ballot64(hit)
                                            : (ballot64(hit) >> gbase) & ((1ull << G) - 1);
                if (tb.t && hit)  // edges {lo, z} and {hub, z}
                    store_term(tb.t + wbase + c + __popcll(hm & below), t, er[r], hm_.ed(g, pos));
                c += (uint32_t)__popcll(hm);
                // Ordered fold: the group's match weights in lane order
                // (ascending z).
                if (G == 64) {
                    uint64_t mask = hm;
                    while (mask) {
                        const int b = __ffsll((unsigned long long)mask) - 1;
                        S += readlane_f64(t, b);
                        mask &= mask - 1;
                    }
                } else if (__any(hm != 0)) {
                    // the group's match weights compacted in lane order; its
                    // first lane adds them one after another
                    if (hit) sfold[gbase + __popcll(hm & below)] = t;
                    wave_sync();
                    if (gl == 0 && hm) {
                        const double* f = sfold + gbase;
                        const int nh = __popcll(hm);
                        int i = 0;
                        for (; i + 4 <= nh; i += 4) {
                            const double a0 = f[i], a1 = f[i + 1], a2 = f[i + 2], a3 = f[i + 3];
                            S += a0;
                            S += a1;
                            S += a2;
                            S += a3;
                        }
                        for (; i < nh; ++i) S += f[i];
                    }
                    wave_sync();
                }
                if (go) p += G;
            }
        }
        if (act && gl == 0) tb.rec[e] = SupRec{S, c, 0u};
    }
}

// Membership filter of a staged hub list: one bit per hash value of
// 2^FB, set for every id of the list (a multiplicative hash, so ids that
// share low bits, as R-MAT's do, still spread).  A clear bit proves z is not
// in the list; a set bit is confirmed by the tree probe.
constexpr uint32_t FILT_MUL = 0x9E3779B1u;
template <int FB>
__device__ __forceinline__ uint32_t filt_hash(uint32_t z) {
    return (z * FILT_MUL) >> (32 - FB);
}
template <int FB>
__device__ __forceinline__ void filt_set(uint32_t* filt, uint32_t z) {
    const uint32_t h = filt_hash<FB>(z);
    atomicOr(&filt[h >> 5], 1u << (h & 31u));
}
template <int FB>
__device__ __forceinline__ bool filt_test(const uint32_t* filt, uint32_t z) {
    const uint32_t h = filt_hash<FB>(z);
    return (filt[h >> 5] >> (h & 31u)) & 1u;
}
constexpr uint32_t RING = 128;  // filter-passing ids queued per wave
// Filter sizes of the 64-lane large- and huge-hub kernels (2^FB bits): at
// most 1/16 and 1/4 of the bits are set by the class's largest hub.  The huge
// kernel's 16 KB filter leaves room for three workgroups per CU; a 32 KB one
// (two) measured 13 % slower on R-MAT scale 20 and 21, an 8 KB one (more
// false positives) 3-5 % slower.
constexpr int LARGE_FB = 16, HUGE_FB = 17;
// Whether the large-hub kernel stages its slots' edge ids and degrees beside
// the tree: without a filter, where every id is probed.  With one, a match
// reads them from the hub's slots in HBM (L2-resident) and the smaller LDS
// footprint fits four workgroups per CU instead of two (R-MAT scale 20 count
// 54.5 -> 45.9 ms).
__host__ __device__ constexpr bool large_slot_meta(int fb) {
    return fb == 0;
}

// hub_edges for G = 64 (one owned edge per wave at a time; R-MAT's long
// shorter lists against hubs of thousands of entries, where most ids miss):
// each streamed id is first tested against the hub's filter, and only the
// ids that pass are queued, in stream order, in the wave's LDS ring (id and
// edge id {lo, z}); whenever 64 are queued they are probed in the tree
// together, so the probe, the match weight and the term store run on full
// waves of candidates instead of on every streamed id.  The queue keeps
// ascending z, so terms and the ordered fold are those of hub_edges.
template <int FB, class PR, class HM, bool PRE = false>
__device__ __forceinline__ void hub_edges_filt(const GraphDev& g, const TermBuf& tb, uint32_t dh,
                                               const PR& probe, const HM& hm_,
                                               const uint32_t* own, const uint32_t* filt,
                                               uint2* ring, uint32_t first, uint32_t cnt,
                                               uint32_t stride, const uint4* pre = nullptr,
                                               const uint64_t* prez = nullptr) {
    constexpr int R = 4;  // chunks of 64 ids per load round
    const int lane = lane_id();
    const uint64_t below = (1ull << lane) - 1ull;
    for (uint32_t idx = first; idx < cnt; idx += stride) {  // wave-uniform
        uint32_t e;  // edge {hub, lo}
        uint2 yo;
        uint64_t wb0;
        if (PRE) {
            const uint4 q = pre[idx];
            e = q.x;
            yo = make_uint2(q.y, q.z);
            wb0 = prez[idx];
        } else {
            const uint32_t li = own[idx];
            e = hm_.ed(g, li);
            yo = hm_.dego(g, li);
            wb0 = tb.t ? hm_.zo(tb, li) : 0;
        }
        const uint32_t dy = yo.x;
        uint32_t p = yo.y;
        const uint32_t le = p + dy;
        const uint64_t wbase = tb.t ? wb0 : 0;
        double S = 0.0;
        uint32_t c = 0, head = 0, tail = 0;  // ring positions (mod RING)
        // Probes the n <= 64 queued ids from head, one per lane.
        auto drain = [&](uint32_t n) {
            wave_sync();  // the ring entries written by other lanes
            const uint2 a = ring[(head + (uint32_t)lane) & (RING - 1)];
            const bool in = (uint32_t)lane < n;
            const uint32_t z = in ? a.x : 0u;
            uint32_t pos = 0;
            const bool hit = in && probe(z, pos);
            const double t = hit ? tri_w(dh, dy, hm_.deg(g, pos)) : 0.0;
            const uint64_t hm = ballot64(hit);
            if (tb.t && hit)  // edges {lo, z} and {hub, z}
                store_term(tb.t + wbase + c + __popcll(hm & below), t, a.y, hm_.ed(g, pos));
            c += (uint32_t)__popcll(hm);
            uint64_t mask = hm;  // ordered fold: ascending z
            while (mask) {
                const int b = __ffsll((unsigned long long)mask) - 1;
                S += readlane_f64(t, b);
                mask &= mask - 1;
            }
            head += n;
            wave_sync();  // these slots are written again only after the reads
        };
        while (p < le) {
            uint32_t zr[R], er[R];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const uint32_t q = p + r * WAVE + lane;
                const uint2 a = q < le ? g.nbe[q] : make_uint2(0u, 0u);
                zr[r] = a.x;
                er[r] = a.y;  // edge {lo, z}
            }
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const bool in = p + r * WAVE + lane < le;
                const bool pass = in && filt_test<FB>(filt, zr[r]);
                const uint64_t pm = ballot64(pass);
                if (pass)
                    ring[(tail + (uint32_t)__popcll(pm & below)) & (RING - 1)] =
                        make_uint2(zr[r], er[r]);
                tail += (uint32_t)__popcll(pm);
                if (tail - head >= WAVE) drain(WAVE);
            }
            p += R * WAVE;
        }
        if (tail != head) drain(tail - head);
        if (lane == 0) tb.rec[e] = SupRec{S, c, 0u};
    }
}

// Small hubs (d_hub <= 64): one wave per unit.  The wave stages its hub in
// its own LDS slice, one slot per lane in a single load round trip (the
// neighbours' degrees and offsets come precomputed per slot), so units of
// a block never wait on each other.  The tree is the list itself plus its 8
// block pivots, read from the lanes' registers.
constexpr int HW = 256;
template <int G>
__global__ void __launch_bounds__(HW) k_count_hub_w(GraphDev g, const uint4* __restrict__ units,
                                                    uint32_t nunits,
                                                    const uint32_t* __restrict__ own_idx,
                                                    TermBuf tb) {
    constexpr int NW = HW / WAVE;
    __shared__ __attribute__((aligned(16))) uint32_t s_nb[NW][WAVE];
    __shared__ uint32_t s_ed[NW][WAVE], s_dg[NW][WAVE], s_of[NW][WAVE], s_own[NW][WAVE];
    __shared__ uint64_t s_zo[NW][WAVE];
    __shared__ double s_fold[NW][WAVE];
    const int lane = lane_id();
    const int wid = threadIdx.x / WAVE;
    const HubMeta<true> meta{s_ed[wid], s_dg[wid], s_of[wid], s_zo[wid], 0u};
    for (uint32_t un = blockIdx.x * NW + wid; un < nunits; un += gridDim.x * NW) {
        uint32_t h, oh, k0, dh, cnt;
        unit_fields(units[un], h, oh, k0, dh, cnt);  // dh, cnt <= 64 in this class
        uint32_t z = PAD_ID;
        if (lane < dh) {
            const uint2 a = g.nbe[oh + lane];
            const uint2 b = g.ndo[oh + lane];
            z = a.x;
            s_ed[wid][lane] = a.y;
            s_dg[wid][lane] = b.x;
            s_of[wid][lane] = b.y;
            if (tb.t) s_zo[wid][lane] = tb.zsl[oh + lane];
        }
        s_nb[wid][lane] = z;
        Piv8 top;
#pragma unroll
        for (int i = 0; i < 8; ++i) top.p[i] = __builtin_amdgcn_readlane(z, 8 * i + 7);
        if (lane < cnt) s_own[wid][lane] = own_idx[k0 + lane];
        wave_sync();
        hub_edges<G>(g, tb, dh, TreeProbe<1>{s_nb[wid], nullptr, nullptr, top}, meta, s_own[wid],
                              s_fold[wid], 0, cnt, WAVE / G);
        wave_sync();  // the next unit restages the slice
    }
}

// Small hubs (d_hub <= 64) with a term buffer, batched: a wave takes a chunk
// of consecutive units (one hub each) and splits it into batches of whole
// units with at most 64 owned edges and 256 hub-list slots (each hub's list
// padded to a multiple of 8).  A batch stages all its hub lists in LDS at
// once and streams ALL its owned edges' shorter lists as one sequence, so the
// per-unit work of one hub per wave pass (staging, prefix scans, a partial
// batch of matches, a step half empty at the end of the hub) is paid once per
// batch, about six hubs.  Edge k of the batch covers elements [L_k, L_k + d_k)
// of the sequence; a lane finds its edge from the edge starts in the step
// (per-step tags in LDS, one ballot per 64-element part).
//
// Every lane probes its own edge's hub: the hub's list is an 8-ary tree whose
// eight top pivots are kept per hub in LDS, indexed by the hub's first block
// (the last entries of the hub's full blocks; UINT32_MAX from the hub's last
// block on, so a probe never leaves the hub's blocks), and the leaf is the
// hub's list itself.  The step
// after the current one has its edge lookup done and its (id, edge id) pairs
// loaded before the current step is probed, so the loads of one step are in
// flight while the previous one is worked on.
//
// Matches go to the wave's ring in LDS in stream order -- the edge id of
// {lo, z} from the streamed pair, the hub slot of z, the owned edge -- and are
// handled 64 at a time, one per lane: the weight 1/(d_lo d_hub d_z) (the
// product d_lo d_hub < 2^12 times d_z is exact in FP64, so the division sees
// the same value as the reference's integer product), the 16-byte term at the
// edge's term offset plus its earlier matches, and the support fold: the
// weights go to LDS in ring order and the last lane of every edge's run adds
// the run, one weight after another, to the edge's running sum (ascending z,
// the reference's order).  A batch's last matches are handled when it ends.
constexpr int SW = 64;
constexpr uint32_t S_HS = 256;   // hub-list slots per batch
constexpr uint32_t S_OE = 64;    // owned edges per batch
constexpr uint32_t S_Q = 128;    // pending-match ring per wave: < 64 left + 64 appended
constexpr uint32_t S_WORK = 1024; // streamed elements per chunk (wave task), about

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x, int lane) {
#pragma unroll
    for (int d = 1; d < WAVE; d <<= 1) {
        const uint32_t y = __shfl_up(x, (unsigned)d, 64);
        if (lane >= d) x += y;
    }
    return x;
}

// The leaf block of z in a hub list of at most 64 entries, given its top
// pivots pa = p0..p3, pb = p4..p7 (p7 and every pivot from the last block on
// are UINT32_MAX): the block is the count of p0..p6 below z, by three
// selects.  True and pos = index of z when z is in the list.
__device__ __forceinline__ bool probe64(const uint32_t* leaf, const uint4& pa, const uint4& pb,
                                        uint32_t z, uint32_t& pos) {
    const bool c2 = pa.w < z;
    const bool c1 = (c2 ? pb.y : pa.y) < z;
    const uint32_t q1 = c2 ? pb.z : pa.z;
    const uint32_t q0 = c2 ? pb.x : pa.x;
    const bool c0 = (c1 ? q1 : q0) < z;
    const uint32_t b = (c2 ? 4u : 0u) | (c1 ? 2u : 0u) | (c0 ? 1u : 0u);
    const uint4* q = reinterpret_cast<const uint4*>(leaf) + 2 * b;
    const uint4 a = q[0], c = q[1];
    const uint32_t mk = (uint32_t)(a.x == z) | (uint32_t)(a.y == z) << 1 |
                        (uint32_t)(a.z == z) << 2 | (uint32_t)(a.w == z) << 3 |
                        (uint32_t)(c.x == z) << 4 | (uint32_t)(c.y == z) << 5 |
                        (uint32_t)(c.z == z) << 6 | (uint32_t)(c.w == z) << 7;
    pos = 8 * b + (mk ? (uint32_t)__ffs(mk) - 1u : 0u);
    return mk != 0;
}

__global__ void __launch_bounds__(SW, 5) k_count_small(GraphDev g, const uint4* __restrict__ units,
                                                    uint32_t nunits,
                                                    const uint32_t* __restrict__ chunk,
                                                    uint32_t nchunks,
                                                    const uint32_t* __restrict__ own_idx,
                                                    TermBuf tb) {
    constexpr int NW = SW / WAVE;
    __shared__ __attribute__((aligned(16))) uint32_t s_nb[NW][S_HS];  // hub lists, padded
    __shared__ uint2 s_hd[NW][S_HS];      // hub slot: (edge id of {hub, z}, d_z)
    __shared__ uint2 s_ek[NW][S_OE];      // owned edge k: (list offset - L_k, hub base slot)
    __shared__ __attribute__((aligned(16))) uint4 s_pv[NW][S_HS / 8][2];  // top pivots per hub
    __shared__ uint32_t s_D[NW][S_OE];    // d_lo * d_hub
    __shared__ uint64_t s_zb[NW][S_OE];   // term offset of edge k
    __shared__ uint32_t s_cnt[NW][S_OE];  // matches so far
    __shared__ double s_S[NW][S_OE];      // support so far
    __shared__ double s_t[NW][WAVE];      // a batch's match weights, in ring order
    // Tagged marks: a batch's starts (owned edges at [0, 64), hub blocks at
    // [64, 96)), then each step's edge starts (positions in the step); one
    // tag per use, so a mark counts only in the use that wrote it.
    __shared__ uint32_t s_mark[NW][2 * WAVE];
    __shared__ uint2 s_q[NW][S_Q];        // pending matches: (edge id of {lo, z}, slot | k << 8)
    const int lane = lane_id();
    const int wid = threadIdx.x / WAVE;
    const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
    const uint64_t upto = below | (1ull << lane);
    uint32_t tag = 0;  // the marks start at 0
    s_mark[wid][lane] = 0;
    s_mark[wid][WAVE + lane] = 0;
    // The pending matches [qh, qh + nh) of the ring handled one per lane, the
    // support folded in ring order.
    auto handle = [&](uint32_t qh, uint32_t nh) {
        const bool act = (uint32_t)lane < nh;
        const uint2 q = s_q[wid][(qh + (uint32_t)lane) & (S_Q - 1)];
        const uint32_t hs = q.y & 0xffu, k = act ? q.y >> 8 : 0u;
        const uint32_t kup = __shfl_up(k, 1u, 64), kdn = __shfl_down(k, 1u, 64);
        const bool start = act && (lane == 0 || kup != k);
        const uint64_t sm = ballot64(start) | 1ull;
        const int r0 = 63 - __clzll((long long)(sm & upto));  // first lane of the run
        const bool last = act && ((uint32_t)lane + 1 == nh || kdn != k);
        uint32_t c0 = 0;
        if (act) {
            c0 = s_cnt[wid][k];
            const uint2 hd = s_hd[wid][hs];
            const double t = 1.0 / ((double)s_D[wid][k] * (double)hd.y);
            store_term(tb.t + s_zb[wid][k] + c0 + (uint32_t)(lane - r0), t, q.x, hd.x);
            s_t[wid][lane] = t;
        }
        wave_sync();
        if (last) {
            double S = s_S[wid][k];
            int i = r0;
            for (; i + 3 <= lane; i += 4) {
                double a[4];
#pragma unroll
                for (int j = 0; j < 4; ++j) a[j] = s_t[wid][i + j];
#pragma unroll
                for (int j = 0; j < 4; ++j) S += a[j];
            }
            for (; i <= lane; ++i) S += s_t[wid][i];
            s_S[wid][k] = S;
            s_cnt[wid][k] = c0 + (uint32_t)(lane - r0 + 1);
        }
        wave_sync();
    };
    // A batch's end: its last pending matches, then every owned edge's
    // (support, triangles) record.  Runs at the next batch's setup.
    bool pend = false, pko = false;
    uint32_t pqh = 0, pqt = 0, pve = 0;
    auto flush = [&]() {
        if (pqt != pqh) {
            wave_sync();
            handle(pqh, pqt - pqh);
        }
        if (pko) tb.rec[pve] = SupRec{s_S[wid][lane], s_cnt[wid][lane], 0u};
        wave_sync();
        pend = false;
    };
    // Chunks of about equal work (the plan's chunk offsets), round robin over
    // the launch, neighbouring chunks on one XCD (hubs of one community share
    // lists); the units of the next batch are loaded while the current one
    // streams.
    uint32_t ch = xcd_swizzle(blockIdx.x, gridDim.x) * NW + wid;
    if (ch >= nchunks) return;
    uint32_t u1 = chunk[ch + 1], s = chunk[ch];
    uint4 Un = (uint32_t)lane < u1 - s ? units[s + lane] : make_uint4(0u, 0u, 0u, 0u);
    for (;;) {
        {
            // Batch: units [s, s + nb) of the chunk [.., u1), the longest run
            // from s within both caps; lane j holds unit s + j.  (Wave-uniform
            // values are read with readlane so that they live in scalar
            // registers and the loops below stay uniform.)
            const uint4 U = (uint32_t)lane < u1 - s ? Un : make_uint4(0u, 0u, 0u, 0u);
            const uint32_t udh = U.w & UNIT_DH_MASK, ucnt = U.w >> UNIT_DH_BITS;
            const uint32_t upad = (udh + 7u) & ~7u;
            const uint32_t pc = wave_incl_scan(ucnt, lane), ph = wave_incl_scan(upad, lane);
            const bool fit = (uint32_t)lane < u1 - s && pc <= S_OE && ph <= S_HS;
            const uint32_t nb = (uint32_t)__popcll(ballot64(fit));
            const uint32_t noe = __builtin_amdgcn_readlane(pc, (int)nb - 1);
            const uint32_t nhs = __builtin_amdgcn_readlane(ph, (int)nb - 1);
            const bool inb = (uint32_t)lane < nb;
            const uint32_t ucb = pc - ucnt, uhb = ph - upad;  // unit's first edge, slot
            ++tag;
            if (inb) {
                s_mark[wid][ucb] = tag;
                s_mark[wid][WAVE + (uhb >> 3)] = tag;
            }
            wave_sync();
            const uint64_t kst = ballot64(s_mark[wid][lane] == tag);
            const uint64_t hst = ballot64(lane < 32 && s_mark[wid][WAVE + lane] == tag);
            // Hub slots, four rounds of 64: list entry, its edge id and degree.
            uint2 hA[4];
            uint32_t hD[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const uint32_t i = (uint32_t)lane + 64u * r;
                hA[r] = make_uint2(PAD_ID, 0u);
                hD[r] = 0u;
                if (64u * r < nhs) {
                    const uint32_t blk = min(i >> 3, 31u);
                    const int j = __popcll(hst & ((2ull << blk) - 1ull)) - 1;
                    const uint32_t oh = __shfl(U.y, j, 64), dh = __shfl(udh, j, 64);
                    const uint32_t t = i - __shfl(uhb, j, 64);
                    if (i < nhs && t < dh) {
                        hA[r] = g.nbe[oh + t];
                        hD[r] = g.ndo[oh + t].x;
                    }
                }
            }
            // Owned edge k = lane: its unit, hub-side slot and list data.
            const bool ko = (uint32_t)lane < noe;
            const int jk = __popcll(kst & upto) - 1;
            const uint32_t k0 = __shfl(U.z, jk, 64), cb = __shfl(ucb, jk, 64);
            const uint32_t koh = __shfl(U.y, jk, 64), kdh = __shfl(udh, jk, 64);
            const uint32_t khb = __shfl(uhb, jk, 64);
            const uint32_t li = ko ? own_idx[k0 + (uint32_t)lane - cb] : 0u;
            // The previous batch's last matches and results, while this
            // batch's staging loads are in flight (its LDS is overwritten
            // only below).
            if (pend) flush();
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const uint32_t i = (uint32_t)lane + 64u * r;
                if (64u * r < nhs) {
                    s_nb[wid][i] = hA[r].x;
                    s_hd[wid][i] = make_uint2(hA[r].y, hD[r]);
                }
            }
            // (loaded by every lane, slot 0 for lanes past the batch's edges,
            // so no load is left in flight on a path the compiler cannot
            // see through)
            const uint32_t olo = g.ndo[ko ? koh + li : 0u].y;
            const uint64_t zo = tb.zsl[ko ? koh + li : 0u];
            wave_sync();
            uint32_t dlo = 0, ve = 0;
            if (ko) {
                const uint2 hd = s_hd[wid][khb + li];  // slot li of the hub: lo
                ve = hd.x;
                dlo = hd.y;
            }
            if (inb) {  // the unit's hub pivots
                const uint32_t nblk = upad >> 3;
                uint32_t p[8];
#pragma unroll
                for (int b = 0; b < 8; ++b)
                    p[b] = (uint32_t)b + 1u < nblk ? s_nb[wid][uhb + 8u * b + 7u] : PAD_ID;
                s_pv[wid][uhb >> 3][0] = make_uint4(p[0], p[1], p[2], p[3]);
                s_pv[wid][uhb >> 3][1] = make_uint4(p[4], p[5], p[6], p[7]);
            }
            const uint32_t incl = wave_incl_scan(dlo, lane);
            const uint32_t Lk = incl - dlo;
            const uint32_t W = __builtin_amdgcn_readlane(incl, WAVE - 1);
            if (ko) {
                s_ek[wid][lane] = make_uint2(olo - Lk, khb);
                s_D[wid][lane] = dlo * kdh;
                s_zb[wid][lane] = zo;
                s_cnt[wid][lane] = 0;
                s_S[wid][lane] = 0.0;
            }
            wave_sync();

            // The next batch's units (loaded by every lane: the last unit
            // past the end, never used).
            uint32_t sn = s + nb;
            bool last_batch = false;
            if (sn >= u1) {
                ch += gridDim.x * NW;
                if (ch < nchunks) {
                    sn = chunk[ch];
                    u1 = chunk[ch + 1];
                } else {
                    last_batch = true;
                }
            }
            Un = units[min(sn + (uint32_t)lane, nunits - 1)];
            // Edge lookup and loads of the step at base: the step's edge
            // starts tagged in LDS, then per part the lane's edge (kb: the
            // edge of the element before the step; -1 past the end) and its
            // list's (id, edge id) pair.  The loads are issued by every lane
            // (slot 0 past the end; that pair is never used), so the wait
            // before a step's probe leaves the later steps' loads in flight.
            int kb = -1;
            struct Step {
                int k[2];
                uint2 z[2];
            };
            auto prep = [&](uint32_t base, Step& S) {
                ++tag;
                if (ko && Lk >= base && Lk - base < 2u * WAVE) s_mark[wid][Lk - base] = tag;
                wave_sync();
                uint32_t sq[2];
#pragma unroll
                for (int q = 0; q < 2; ++q) {
                    const bool st = s_mark[wid][q * WAVE + lane] == tag;
                    const uint64_t m = ballot64(st);
                    const int k = kb + (int)__popcll(m & below) + (st ? 1 : 0);
                    kb += (int)__popcll(m);
                    const bool aq = base + q * WAVE + (uint32_t)lane < W;  // then 0 <= k < noe
                    sq[q] = aq ? base + q * WAVE + (uint32_t)lane + s_ek[wid][k].x : 0u;
                    S.k[q] = aq ? k : -1;
                }
#pragma unroll
                for (int q = 0; q < 2; ++q) S.z[q] = g.nbe[sq[q]];
            };
            uint32_t qh = 0, qt = 0;
            auto probe = [&](const Step& S) {
#pragma unroll
                for (int hf = 0; hf < 2; ++hf) {
                    // Branch-free: every lane probes (a lane past the end
                    // probes edge 0's hub and is masked), so the step's
                    // loads are always waited for here.
                    const bool act = S.k[hf] >= 0;
                    const uint32_t hb = s_ek[wid][act ? S.k[hf] : 0].y;
                    const uint4 pa = s_pv[wid][hb >> 3][0], pb = s_pv[wid][hb >> 3][1];
                    uint32_t pos;
                    const bool hit = probe64(s_nb[wid] + hb, pa, pb, S.z[hf].x, pos) & act;
                    const uint64_t hm = ballot64(hit);
                    if (hit)
                        s_q[wid][(qt + (uint32_t)__popcll(hm & below)) & (S_Q - 1)] =
                            make_uint2(S.z[hf].y, (hb + pos) | ((uint32_t)S.k[hf] << 8));
                    qt += (uint32_t)__popcll(hm);
                    if (qt - qh >= (uint32_t)WAVE) {
                        wave_sync();
                        handle(qh, WAVE);
                        qh += WAVE;
                    }
                }
            };
            // Two step buffers used in turn (one step's loads in flight
            // while the other is probed; no register copies of loads in
            // flight).
            Step A, B;
            prep(0, A);
            for (uint32_t base = 0;;) {
                prep(base + 2 * WAVE, B);
                probe(A);
                base += 2 * WAVE;
                if (base >= W) break;
                prep(base + 2 * WAVE, A);
                probe(B);
                base += 2 * WAVE;
                if (base >= W) break;
            }
            pend = true;
            pko = ko;
            pve = ve;
            pqh = qh;
            pqt = qt;
            if (last_batch) break;
            s = sn;
        }
    }
    flush();
}

// Medium and large hubs (d_hub <= 512: LV = 2; <= 4096: LV = 3): one
// workgroup of NT lanes per unit, the hub's tree staged in LDS once and shared
// by its waves.  Medium hubs also stage the per-slot data and the owned slots;
// large ones read them from HBM (the hub's own slots: L2-resident).
template <int G, int LV, int NT>
__global__ void __launch_bounds__(NT) k_count_hub(GraphDev g, const uint4* __restrict__ units,
                                                  uint32_t nunits,
                                                  const uint32_t* __restrict__ own_idx,
                                                  TermBuf tb) {
    constexpr uint32_t L = LV == 2 ? 512 : 4096;
    constexpr bool META = LV == 2;  // per-slot data staged in LDS
    constexpr uint32_t ML = META ? L : 1;
    __shared__ __attribute__((aligned(16))) uint32_t s_leaf[L];
    __shared__ __attribute__((aligned(16))) uint32_t s_l1[L / 8];
    __shared__ __attribute__((aligned(16))) uint32_t s_l2[L / 64];
    __shared__ uint32_t s_ed[ML], s_dg[ML], s_of[ML], s_own[ML];
    __shared__ uint64_t s_zo[ML];
    __shared__ double sfold[NT];
    const int wid = threadIdx.x / WAVE;
    constexpr uint32_t per_wave = WAVE / G;
    for (uint32_t un = blockIdx.x; un < nunits; un += gridDim.x) {
        uint32_t h, oh, k0, dh, cnt;
        unit_fields(units[un], h, oh, k0, dh, cnt);
        // Leaf: the list, padded through the block after its last entry (a
        // probe above every id can land there); pivots for every position.
        const uint32_t nleaf = min(L, ((dh + 7) & ~7u) + 8);
        for (uint32_t i = threadIdx.x; i < nleaf; i += NT) {
            if (META && i < dh) {
                const uint2 a = g.nbe[oh + i];
                const uint2 b = g.ndo[oh + i];
                s_leaf[i] = a.x;
                s_ed[i] = a.y;
                s_dg[i] = b.x;
                s_of[i] = b.y;
                if (tb.t) s_zo[i] = tb.zsl[oh + i];
            } else {
                s_leaf[i] = hub_id(g, oh, dh, i);
            }
        }
        for (uint32_t i = threadIdx.x; i < L / 8; i += NT) s_l1[i] = hub_id(g, oh, dh, 8 * i + 7);
        if (LV == 3)
            for (uint32_t i = threadIdx.x; i < L / 64; i += NT)
                s_l2[i] = hub_id(g, oh, dh, 64 * i + 63);
        Piv8 top;
#pragma unroll
        for (int i = 0; i < 8; ++i) top.p[i] = hub_id(g, oh, dh, (L / 8) * i + L / 8 - 1);
        if (META)
            for (uint32_t i = threadIdx.x; i < cnt; i += NT) s_own[i] = own_idx[k0 + i];
        __syncthreads();
        const HubMeta<META> meta{s_ed, s_dg, s_of, s_zo, oh};
        hub_edges<G>(g, tb, dh, TreeProbe<LV>{s_leaf, s_l1, s_l2, top}, meta,
                               META ? s_own : own_idx + k0, sfold + wid * WAVE, wid * per_wave,
                               cnt, per_wave * (NT / WAVE));
        __syncthreads();  // the next unit restages the hub
    }
}

// Large hubs (degree in (l1_max, l2_max], l2_max <= 4096): one workgroup of HTL lanes per unit,
// the hub's 8-ary tree staged in LDS once and shared by its waves; the owned
// edges' own data is read from HBM (the hub's slots, L2-resident; fetching it
// during the staging measured slower, it delays the staging barrier).  The
// leaf lives in dynamic LDS of lcap entries, lcap >= the launch's largest hub
// degree + 8: a probe ends in the leaf block holding the hub's first padding
// entry at the latest.  FB = 0 (fewer than 64 lanes per edge): the edge id and
// degree of every hub slot are staged beside the leaf (two more arrays of
// lcap), so that a match needs no memory access.  FB > 0 (64 lanes): the ids
// are filtered first (hub_edges_filt) against a filter of 2^FB bits built
// while the hub is staged, and a match reads its slot's data from HBM.
constexpr int HTL = 512;
template <int G, int FB = 0>
__global__ void __launch_bounds__(HTL) k_count_hub_large(GraphDev g,
                                                         const uint4* __restrict__ units,
                                                         uint32_t nunits,
                                                         const uint32_t* __restrict__ own_idx,
                                                         TermBuf tb, uint32_t lcap) {
    static_assert(FB == 0 || G == WAVE, "the filtered path runs one edge per wave");
    constexpr uint32_t L = 4096;
    constexpr uint32_t NF = FB ? (1u << FB) / 32 : 1;
    constexpr bool SM = large_slot_meta(FB);  // slots' edge ids and degrees in LDS
    extern __shared__ __attribute__((aligned(16))) uint32_t s_dyn[];
    uint32_t* s_leaf = s_dyn;
    uint32_t* s_ed = s_dyn + lcap;
    uint32_t* s_dg = s_dyn + 2 * lcap;
    __shared__ __attribute__((aligned(16))) uint32_t s_l1[L / 8];
    __shared__ __attribute__((aligned(16))) uint32_t s_l2[L / 64];
    __shared__ double sfold[FB ? 1 : HTL];
    __shared__ __attribute__((aligned(16))) uint32_t s_filt[NF];
    __shared__ uint2 s_ring[FB ? (HTL / WAVE) * RING : 1];
    const int wid = threadIdx.x / WAVE;
    constexpr uint32_t per_wave = WAVE / G;
    for (uint32_t un = blockIdx.x; un < nunits; un += gridDim.x) {
        uint32_t h, oh, k0, dh, cnt;
        unit_fields(units[un], h, oh, k0, dh, cnt);
        if constexpr (FB > 0) {  // the previous unit's probes ended at the barrier below
            for (uint32_t i = threadIdx.x; i < NF; i += HTL) s_filt[i] = 0u;
            __syncthreads();
        }
        const uint32_t nleaf = ((dh + 7) & ~7u) + 8;  // <= lcap
        for (uint32_t i = threadIdx.x; i < nleaf; i += HTL) {
            uint32_t z = PAD_ID;
            if (i < dh) {
                if constexpr (SM) {
                    const uint2 a = g.nbe[oh + i];
                    z = a.x;
                    s_ed[i] = a.y;
                    s_dg[i] = g.ndo[oh + i].x;
                } else {
                    z = g.nb[oh + i];
                }
                if constexpr (FB > 0) filt_set<FB>(s_filt, z);
            }
            s_leaf[i] = z;
        }
        for (uint32_t i = threadIdx.x; i < L / 8; i += HTL) s_l1[i] = hub_id(g, oh, dh, 8 * i + 7);
        for (uint32_t i = threadIdx.x; i < L / 64; i += HTL) s_l2[i] = hub_id(g, oh, dh, 64 * i + 63);
        Piv8 top;
#pragma unroll
        for (int i = 0; i < 8; ++i) top.p[i] = hub_id(g, oh, dh, (L / 8) * i + L / 8 - 1);
        __syncthreads();
        const HubMeta<false, SM> meta{s_ed, s_dg, nullptr, nullptr, oh};
        if constexpr (FB > 0)
            hub_edges_filt<FB>(g, tb, dh, TreeProbe<3>{s_leaf, s_l1, s_l2, top}, meta,
                               own_idx + k0, s_filt, s_ring + wid * RING, wid, cnt,
                               HTL / WAVE);
        else
            hub_edges<G>(g, tb, dh, TreeProbe<3>{s_leaf, s_l1, s_l2, top}, meta, own_idx + k0,
                         sfold + wid * WAVE, wid * per_wave, cnt, per_wave * (HTL / WAVE));
        __syncthreads();  // the next unit restages the hub
    }
}

// Medium and large hubs with a term buffer (degree in (l0_max, l2_max], tree
// depth LV = 2 up to 512 entries, 3 up to 4096): one workgroup of four waves
// per unit (a hub and part of its owned edges).  The hub's list is staged in
// LDS once per unit as an 8-ary tree (leaf, every 8th and every 64th entry;
// the top pivots in registers) with the degree of every slot beside it.  The
// unit's owned edges are taken 32 at a time by the waves (an LDS counter);
// a wave streams its edges' shorter lists as one sequence, 128 elements per
// step with the next step's loads in flight, probes every id in the tree and
// handles the matches 64 at a time exactly like the small-hub kernel (the
// edge id of {hub, z} gathered from the hub's slots, L2-resident meanwhile).
constexpr int HS_T = 256;
constexpr uint32_t HS_E = 16;  // owned edges per wave batch

template <int LV>
__global__ void __launch_bounds__(HS_T) k_count_hub_stream(GraphDev g,
                                                           const uint4* __restrict__ units,
                                                           uint32_t nunits,
                                                           const uint32_t* __restrict__ own_idx,
                                                           TermBuf tb, uint32_t lcap) {
    constexpr uint32_t L = LV == 2 ? 512 : 4096;
    constexpr int NW = HS_T / WAVE;
    extern __shared__ __attribute__((aligned(16))) uint32_t s_dyn[];
    uint32_t* s_leaf = s_dyn;       // lcap entries: the list, padded
    uint32_t* s_dg = s_dyn + lcap;  // lcap entries: degree of each slot's vertex
    __shared__ __attribute__((aligned(16))) uint32_t s_l1[L / 8];
    __shared__ __attribute__((aligned(16))) uint32_t s_l2[L / 64];
    __shared__ uint32_t s_next;  // the unit's next batch
    __shared__ uint32_t s_ek[NW][HS_E];   // owned edge k: list offset - L_k
    __shared__ uint32_t s_D[NW][HS_E];    // d_lo * d_hub
    __shared__ uint64_t s_zb[NW][HS_E];   // term offset
    __shared__ uint32_t s_cnt[NW][HS_E];  // matches so far
    __shared__ double s_S[NW][HS_E];      // support so far
    __shared__ double s_t[NW][WAVE];
    __shared__ uint32_t s_mark[NW][2 * WAVE];
    __shared__ uint2 s_q[NW][S_Q];  // pending matches: (edge id of {lo, z}, hub slot | k << 16)
    const int lane = lane_id();
    const int wid = threadIdx.x / WAVE;
    const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
    const uint64_t upto = below | (1ull << lane);
    uint32_t tag = 0;
    s_mark[wid][lane] = 0;
    s_mark[wid][WAVE + lane] = 0;
    // (units of one hub are consecutive: one XCD stages the hub from its L2)
    for (uint32_t un = xcd_swizzle(blockIdx.x, gridDim.x); un < nunits; un += gridDim.x) {
        uint32_t h, oh, k0, dh, cnt;
        unit_fields(units[un], h, oh, k0, dh, cnt);
        const uint32_t nleaf = min(L, ((dh + 7) & ~7u) + 8);  // <= lcap
        for (uint32_t i = threadIdx.x; i < nleaf; i += HS_T) {
            uint32_t z = PAD_ID, d = 0;
            if (i < dh) {
                z = g.nb[oh + i];
                d = g.ndo[oh + i].x;
            }
            s_leaf[i] = z;
            s_dg[i] = d;
        }
        if (threadIdx.x == 0) s_next = 0;
        __syncthreads();
        // Pivot levels from the staged leaf (entries from nleaf on are
        // padding).
        auto leaf_at = [&](uint32_t i) { return i < nleaf ? s_leaf[i] : PAD_ID; };
        for (uint32_t i = threadIdx.x; i < L / 8; i += HS_T) s_l1[i] = leaf_at(8 * i + 7);
        if (LV == 3)
            for (uint32_t i = threadIdx.x; i < L / 64; i += HS_T) s_l2[i] = leaf_at(64 * i + 63);
        Piv8 top;
#pragma unroll
        for (int i = 0; i < 8; ++i) top.p[i] = leaf_at((L / 8) * i + L / 8 - 1);
        __syncthreads();
        const TreeProbe<LV> probe_tree{s_leaf, s_l1, s_l2, top};
        // The unit's owned edges split into nbat batches of at most HS_E
        // edges, at least one per wave (long lists: few edges per unit).
        const uint32_t nbat = min(cnt, max((uint32_t)NW, (cnt + HS_E - 1) / HS_E));
        for (;;) {
            uint32_t b = 0;
            if (lane == 0) b = atomicAdd(&s_next, 1u);
            b = __builtin_amdgcn_readfirstlane(b);
            if (b >= nbat) break;
            const uint32_t a0 = (uint32_t)((uint64_t)cnt * b / nbat);
            const uint32_t noe = (uint32_t)((uint64_t)cnt * (b + 1) / nbat) - a0;
            // Owned edge k = lane of the batch: its hub slot li, the shorter
            // list's offset and length, the term offset.
            const bool ko = (uint32_t)lane < noe;
            const uint32_t li = own_idx[k0 + a0 + (ko ? (uint32_t)lane : 0u)];
            const uint32_t olo = g.ndo[oh + li].y;
            const uint64_t zo = tb.zsl[oh + li];
            const uint32_t ve = g.nbe[oh + li].y;
            const uint32_t dlo = ko ? s_dg[li] : 0u;
            const uint32_t incl = wave_incl_scan(dlo, lane);
            const uint32_t Lk = incl - dlo;
            const uint32_t W = __builtin_amdgcn_readlane(incl, WAVE - 1);
            if (ko) {
                s_ek[wid][lane] = olo - Lk;
                s_D[wid][lane] = dlo * dh;
                s_zb[wid][lane] = zo;
                s_cnt[wid][lane] = 0;
                s_S[wid][lane] = 0.0;
            }
            wave_sync();
            auto handle = [&](uint32_t qh, uint32_t nh) {
                const bool act = (uint32_t)lane < nh;
                const uint2 q = s_q[wid][(qh + (uint32_t)lane) & (S_Q - 1)];
                const uint32_t hs = act ? q.y & 0xffffu : 0u, k = act ? q.y >> 16 : 0u;
                const uint32_t ez = g.nbe[oh + hs].y;  // edge {hub, z}
                const uint32_t kup = __shfl_up(k, 1u, 64), kdn = __shfl_down(k, 1u, 64);
                const bool start = act && (lane == 0 || kup != k);
                const uint64_t sm = ballot64(start) | 1ull;
                const int r0 = 63 - __clzll((long long)(sm & upto));
                const bool last = act && ((uint32_t)lane + 1 == nh || kdn != k);
                uint32_t c0 = 0;
                if (act) {
                    c0 = s_cnt[wid][k];
                    const double t = 1.0 / ((double)s_D[wid][k] * (double)s_dg[hs]);
                    store_term(tb.t + s_zb[wid][k] + c0 + (uint32_t)(lane - r0), t, q.x, ez);
                    s_t[wid][lane] = t;
                }
                wave_sync();
                if (last) {
                    double S = s_S[wid][k];
                    int i = r0;
                    for (; i + 3 <= lane; i += 4) {
                        double a[4];
#pragma unroll
                        for (int j = 0; j < 4; ++j) a[j] = s_t[wid][i + j];
#pragma unroll
                        for (int j = 0; j < 4; ++j) S += a[j];
                    }
                    for (; i <= lane; ++i) S += s_t[wid][i];
                    s_S[wid][k] = S;
                    s_cnt[wid][k] = c0 + (uint32_t)(lane - r0 + 1);
                }
                wave_sync();
            };
            int kb = -1;
            struct Step {
                int k[2];
                uint2 z[2];
            };
            auto prep = [&](uint32_t base, Step& S) {
                ++tag;
                if (ko && Lk >= base && Lk - base < 2u * WAVE) s_mark[wid][Lk - base] = tag;
                wave_sync();
                uint32_t sq[2];
#pragma unroll
                for (int q = 0; q < 2; ++q) {
                    const bool st = s_mark[wid][q * WAVE + lane] == tag;
                    const uint64_t m = ballot64(st);
                    const int k = kb + (int)__popcll(m & below) + (st ? 1 : 0);
                    kb += (int)__popcll(m);
                    const bool aq = base + q * WAVE + (uint32_t)lane < W;
                    sq[q] = aq ? base + q * WAVE + (uint32_t)lane + s_ek[wid][k] : 0u;
                    S.k[q] = aq ? k : -1;
                }
#pragma unroll
                for (int q = 0; q < 2; ++q) S.z[q] = g.nbe[sq[q]];
            };
            uint32_t qh = 0, qt = 0;
            auto probe = [&](const Step& S) {
#pragma unroll
                for (int hf = 0; hf < 2; ++hf) {
                    const bool act = S.k[hf] >= 0;
                    uint32_t pos;
                    const bool hit = probe_tree(S.z[hf].x, pos) & act;
                    const uint64_t hm = ballot64(hit);
                    if (hit)
                        s_q[wid][(qt + (uint32_t)__popcll(hm & below)) & (S_Q - 1)] =
                            make_uint2(S.z[hf].y, pos | ((uint32_t)S.k[hf] << 16));
                    qt += (uint32_t)__popcll(hm);
                    if (qt - qh >= (uint32_t)WAVE) {
                        wave_sync();
                        handle(qh, WAVE);
                        qh += WAVE;
                    }
                }
            };
            Step A, B;
            prep(0, A);
            for (uint32_t base = 0;;) {
                prep(base + 2 * WAVE, B);
                probe(A);
                base += 2 * WAVE;
                if (base >= W) break;
                prep(base + 2 * WAVE, A);
                probe(B);
                base += 2 * WAVE;
                if (base >= W) break;
            }
            if (qt != qh) {
                wave_sync();
                handle(qh, qt - qh);
            }
            if (ko) tb.rec[ve] = SupRec{s_S[wid][lane], s_cnt[wid][lane], 0u};
            wave_sync();
        }
        __syncthreads();  // the next unit restages the hub
    }
}

// Huge hubs (degree in (l2_max, HUB_MAX_LIMIT]): a four-level 8-ary tree whose
// upper levels (every 8th, 64th and 512th entry; 4096 + 512 + 64 ids, 18 KB)
// are staged in LDS and whose leaf is the hub's list itself in HBM: a probe
// descends in LDS, then reads one 8-entry block of the list (two aligned
// 16-byte loads; the hub's list is read by all of its units, L2-resident
// meanwhile) instead of the dozen dependent reads of a binary search.  The
// tree indexes the list from the 32-byte boundary at or before its start
// (sh = oh mod 8 entries earlier): those sh entries read as 0, so the sequence
// stays non-decreasing, and are never a match.
//
// Hubs above HUGE_TREE_MAX entries (R-MAT's largest) take a five-level tree
// over 262144 positions (MegaProbe): levels of every 64th, 512th and 4096th
// entry in the same LDS arrays, the top 8 in registers, then the 8 pivots of
// every 8th entry inside the chosen 64-entry stretch are read from the list
// itself (eight 4-byte loads, two lines), then the leaf block.
constexpr uint32_t HUGE_TREE_MAX = 32760;
static_assert(HUGE_TREE_MAX + 7 < 32768 && HUB_MAX_LIMIT + 7 < 262144,
              "a huge hub's list and its shift fit its tree");
static_assert(HUGE_UNIT <= HTL, "a unit's owned edges are prefetched one per thread");
__device__ __forceinline__ uint32_t huge_id(const GraphDev& g, uint32_t ob, uint32_t sh,
                                            uint32_t dh, uint32_t s) {
    return s < sh ? 0u : s < sh + dh ? g.nb[ob + s] : PAD_ID;
}
// Leaf block b of a huge hub's shifted list: positions 8b .. 8b + 7, read
// from memory as they are (the entries before sh belong to the previous list,
// those from sh + dh on to the next one); a match counts only inside the list,
// k in [kmin, kend).  True and pos = z's index in the list when found.
__device__ __forceinline__ bool huge_leaf(const uint4* blk, uint32_t sh, uint32_t dh, uint32_t b,
                                          uint32_t z, uint32_t& pos) {
    const uint4 a = blk[2 * b], c = blk[2 * b + 1];
    const uint32_t v[8] = {a.x, a.y, a.z, a.w, c.x, c.y, c.z, c.w};
    const uint32_t base = 8 * b;
    const uint32_t kmin = base >= sh ? 0u : sh - base;
    const uint32_t kend = sh + dh > base ? min(8u, sh + dh - base) : 0u;
    int j = -1;
#pragma unroll
    for (int k = 7; k >= 0; --k)
        if (v[k] == z && (uint32_t)k >= kmin && (uint32_t)k < kend) j = k;
    pos = base + (uint32_t)max(j, 0) - sh;
    return j >= 0;
}
struct HugeProbe {
    const uint4* blk;  // the list's 32-byte blocks, from its aligned base
    uint32_t sh, dh;
    const uint32_t *l1, *l2, *l3;
    Piv8 top;
    __device__ __forceinline__ bool operator()(uint32_t z, uint32_t& pos) const {
        uint32_t b = 0;
#pragma unroll
        for (int i = 0; i < 8; ++i) b += top.p[i] < z ? 1u : 0u;
        const bool ok = b < 8;
        b = min(b, 7u);
        const uint4* q = reinterpret_cast<const uint4*>(l3) + 2 * b;
        b = 8 * b + min(below8(q[0], q[1], z), 7u);
        q = reinterpret_cast<const uint4*>(l2) + 2 * b;
        b = 8 * b + min(below8(q[0], q[1], z), 7u);
        q = reinterpret_cast<const uint4*>(l1) + 2 * b;
        b = 8 * b + min(below8(q[0], q[1], z), 7u);
        return ok && huge_leaf(blk, sh, dh, b, z, pos);
    }
};

struct MegaProbe {
    const uint4* blk;    // the list's 32-byte blocks, from its aligned base
    const uint32_t* nb;  // the list's entries, from the same base
    uint32_t sh, dh;
    const uint32_t *l2, *l3, *l4;  // every 64th, 512th and 4096th position
    Piv8 top;                      // every 32768th
    __device__ __forceinline__ bool operator()(uint32_t z, uint32_t& pos) const {
        uint32_t b = 0;
#pragma unroll
        for (int i = 0; i < 8; ++i) b += top.p[i] < z ? 1u : 0u;
        const bool ok = b < 8;
        b = min(b, 7u);
        const uint4* q = reinterpret_cast<const uint4*>(l4) + 2 * b;
        b = 8 * b + min(below8(q[0], q[1], z), 7u);
        q = reinterpret_cast<const uint4*>(l3) + 2 * b;
        b = 8 * b + min(below8(q[0], q[1], z), 7u);
        q = reinterpret_cast<const uint4*>(l2) + 2 * b;
        b = 8 * b + min(below8(q[0], q[1], z), 7u);
        // positions 64 b + 8 k + 7 (never below sh <= 7): the list's entry,
        // PAD_ID past its end (the load clamped into the list)
        const uint32_t end = sh + dh;
        uint32_t nbelow = 0;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const uint32_t s = 64 * b + 8 * k + 7;
            const uint32_t x = nb[min(s, end - 1u)];
            nbelow += (s < end && x < z) ? 1u : 0u;
        }
        b = 8 * b + min(nbelow, 7u);
        return ok && huge_leaf(blk, sh, dh, b, z, pos);
    }
};

// FB > 0 (G = 64 only): the ids are filtered first (hub_edges_filt) against a
// filter of 2^FB bits, built from one pass over the hub's list that also
// fills the tree's lowest pivot level.
template <int G, int FB = 0>
__global__ void __launch_bounds__(HTL) k_count_hub_huge(GraphDev g,
                                                        const uint4* __restrict__ units,
                                                        uint32_t nunits,
                                                        const uint32_t* __restrict__ own_idx,
                                                        TermBuf tb, uint32_t tree_max) {
    static_assert(FB == 0 || G == WAVE, "the filtered path runs one edge per wave");
    constexpr uint32_t L = 32768;
    constexpr uint32_t NF = FB ? (1u << FB) / 32 : 1;
    __shared__ __attribute__((aligned(16))) uint32_t s_l1[L / 8];
    __shared__ __attribute__((aligned(16))) uint32_t s_l2[L / 64];
    __shared__ __attribute__((aligned(16))) uint32_t s_l3[L / 512];
    __shared__ double sfold[FB ? 1 : HTL];
    __shared__ uint4 s_pre[HUGE_UNIT];
    __shared__ uint64_t s_prez[HUGE_UNIT];
    __shared__ __attribute__((aligned(16))) uint32_t s_filt[NF];
    __shared__ uint2 s_ring[FB ? (HTL / WAVE) * RING : 1];
    const int wid = threadIdx.x / WAVE;
    constexpr uint32_t per_wave = WAVE / G;
    for (uint32_t un = blockIdx.x; un < nunits; un += gridDim.x) {
        uint32_t h, oh, k0, dh, cnt;
        unit_fields(units[un], h, oh, k0, dh, cnt);  // cnt <= HUGE_UNIT in this class
        if constexpr (FB > 0) {  // the previous unit's probes ended at the barrier below
            for (uint32_t i = threadIdx.x; i < NF; i += HTL) s_filt[i] = 0u;
            __syncthreads();
        }
        if (threadIdx.x < cnt) {
            const uint32_t li = own_idx[k0 + threadIdx.x];
            const uint2 yo = g.ndo[oh + li];
            s_pre[threadIdx.x] = make_uint4(g.nbe[oh + li].y, yo.x, yo.y, 0u);
            s_prez[threadIdx.x] = tb.t ? tb.zsl[oh + li] : 0;
        }
        const uint32_t sh = oh & 7u, ob = oh - sh;
        // the five-level tree (no filter) above tree_max <= HUGE_TREE_MAX
        const bool mega = dh > tree_max;
        Piv8 top;
        if (!mega) {  // dh + sh <= L
            if constexpr (FB > 0) {
                // every entry into the filter; position s = sh + j of the
                // shifted sequence is pivot s / 8 of the lowest level when
                // s % 8 == 7
                for (uint32_t j = threadIdx.x; j < dh; j += HTL) {
                    const uint32_t z = g.nb[oh + j];
                    filt_set<FB>(s_filt, z);
                    if (((sh + j) & 7u) == 7u) s_l1[(sh + j) >> 3] = z;
                }
                for (uint32_t i = ((sh + dh) >> 3) + threadIdx.x; i < L / 8; i += HTL)
                    s_l1[i] = PAD_ID;  // 8 i + 7 >= sh + dh: past the list
            } else {
                for (uint32_t i = threadIdx.x; i < L / 8; i += HTL)
                    s_l1[i] = huge_id(g, ob, sh, dh, 8 * i + 7);
            }
            for (uint32_t i = threadIdx.x; i < L / 64; i += HTL)
                s_l2[i] = huge_id(g, ob, sh, dh, 64 * i + 63);
            for (uint32_t i = threadIdx.x; i < L / 512; i += HTL)
                s_l3[i] = huge_id(g, ob, sh, dh, 512 * i + 511);
#pragma unroll
            for (int i = 0; i < 8; ++i) top.p[i] = huge_id(g, ob, sh, dh, (L / 8) * i + L / 8 - 1);
        } else {  // dh + sh <= 8 L: every 64th, 512th, 4096th, 32768th position
            for (uint32_t i = threadIdx.x; i < L / 8; i += HTL)
                s_l1[i] = huge_id(g, ob, sh, dh, 64 * i + 63);
            for (uint32_t i = threadIdx.x; i < L / 64; i += HTL)
                s_l2[i] = huge_id(g, ob, sh, dh, 512 * i + 511);
            for (uint32_t i = threadIdx.x; i < L / 512; i += HTL)
                s_l3[i] = huge_id(g, ob, sh, dh, 4096 * i + 4095);
#pragma unroll
            for (int i = 0; i < 8; ++i) top.p[i] = huge_id(g, ob, sh, dh, L * i + L - 1);
        }
        __syncthreads();
        const HubMeta<false, false> meta{nullptr, nullptr, nullptr, nullptr, oh};
        const uint4* blk = reinterpret_cast<const uint4*>(g.nb + ob);
        double* fold = FB ? nullptr : sfold + wid * WAVE;  // G < 64 only
        if (mega) {
            const MegaProbe pr{blk, g.nb + ob, sh, dh, s_l1, s_l2, s_l3, top};
            hub_edges<G, MegaProbe, HubMeta<false, false>, true>(
                g, tb, dh, pr, meta, nullptr, fold, wid * per_wave, cnt, per_wave * (HTL / WAVE),
                s_pre, s_prez);
        } else {
            const HugeProbe pr{blk, sh, dh, s_l1, s_l2, s_l3, top};
            if constexpr (FB > 0)
                hub_edges_filt<FB, HugeProbe, HubMeta<false, false>, true>(
                    g, tb, dh, pr, meta, nullptr, s_filt, s_ring + wid * RING, wid, cnt,
                    HTL / WAVE, s_pre, s_prez);
            else
                hub_edges<G, HugeProbe, HubMeta<false, false>, true>(
                    g, tb, dh, pr, meta, nullptr, fold, wid * per_wave, cnt,
                    per_wave * (HTL / WAVE), s_pre, s_prez);
        }
        __syncthreads();  // the next unit restages the hub
    }
}

// Segment of vertex v's fold contributed by its neighbour slot j: the stored
// terms of edge {v, y} with z > y.  The list is ascending in z, so the segment
// starts after the terms with z < y, found by binary search; the third vertex
// of a term is the endpoint of its edge sa = {lo, z} other than lo.
__device__ __forceinline__ void vseg(const GraphDev& g, const TermBuf& tb,
                                     uint32_t v, uint32_t j, uint64_t& start,
                                     uint32_t& len) {
    const uint32_t y = g.nb[j];
    const uint32_t e = g.s2e[j];
    const uint32_t x0 = min(v, y), y0 = max(v, y);
    const uint32_t lo = g.deg[x0] <= g.deg[y0] ? x0 : y0;
    const uint64_t k0 = tb.zoff[e];
    const uint32_t cnt = g.tcnt[e];
    uint32_t a = 0, b = cnt;
    while (a < b) {
        const uint32_t mid = (a + b) >> 1;
        const uint32_t f = tb.t[k0 + mid].sa;
        const uint32_t fx = g.esrc[f];
        const uint32_t z = fx == lo ? g.nb[g.eslot[f]] : fx;
        if (z < y) a = mid + 1;
        else b = mid;
    }
    start = k0 + a;
    len = cnt - a;
}

// Dead-triangle test of term k of the segment of slot j (PULL mode): the
// triangle is gone once any of its three edges is deleted.
__device__ __forceinline__ bool term_dead(const GraphDev& g, const TermBuf& tb,
                                          uint32_t j, uint64_t k) {
    const Term q = tb.t[k];
    return g.st[j] == ST_DEAD || g.est[q.sa] == ST_DEAD || g.est[q.sb] == ST_DEAD;
}

// One lane per (low-degree) vertex: a plain serial fold over its segments.
// PULL=false: the count-phase fold.  PULL=true: after a clean, the vertex
// support is the count-phase value minus its dead triangles, subtracted in the
// same fold order.
template <bool PULL>
__global__ void __launch_bounds__(T) k_vfold_lane(GraphDev g, TermBuf tb,
                                                  const uint32_t* __restrict__ list,
                                                  uint32_t cnt) {
    for (uint32_t i = blockIdx.x * T + threadIdx.x; i < cnt; i += gridDim.x * T) {
        const uint32_t v = list[i];
        double S = PULL ? g.pv0[v] : 0.0;
        uint32_t c = 0;
        for (uint32_t j = g.off[v]; j < g.off[v + 1]; ++j) {
            uint64_t st;
            uint32_t len;
            vseg(g, tb, v, j, st, len);
            for (uint32_t k = 0; k < len; ++k) {
                if (PULL) {
                    if (term_dead(g, tb, j, st + k)) S -= tb.t[st + k].w;
                } else {
                    S += tb.t[st + k].w;
                }
            }
            c += len;
        }
        g.pv[v] = S;
        if (!PULL) {
            g.pv0[v] = S;
            g.vtcnt[v] = c;
        }
    }
}


// One wave per vertex: 64 neighbour segments at a time are laid end to end
// (lane scan of their lengths), the wave gathers 64 consecutive terms of that
// stream per step and folds them serially, uniform across lanes.
constexpr int VT = 256;
template <bool PULL>
__global__ void __launch_bounds__(VT) k_vfold_wave(GraphDev g, TermBuf tb,
                                                   const uint32_t* __restrict__ list,
                                                   uint32_t cnt) {
    __shared__ uint64_t s_start[VT];
    __shared__ uint32_t s_pre[VT];
    __shared__ uint32_t s_slot[VT];
    const int lane = lane_id();
    const int wid = threadIdx.x / WAVE;
    uint64_t* sst = s_start + wid * WAVE;
    uint32_t* spre = s_pre + wid * WAVE;
    uint32_t* sslot = s_slot + wid * WAVE;
    const uint32_t wave = (blockIdx.x * VT + threadIdx.x) / WAVE;
    const uint32_t nwaves = gridDim.x * VT / WAVE;
    for (uint32_t i = wave; i < cnt; i += nwaves) {
        const uint32_t v = list[i];
        const uint32_t jb = g.off[v], je = g.off[v + 1];
        double S = PULL ? g.pv0[v] : 0.0;
        uint32_t c = 0;
        for (uint32_t j0 = jb; j0 < je; j0 += WAVE) {
            const uint32_t j = j0 + lane;
            uint64_t st = 0;
            uint32_t len = 0;
            if (j < je) vseg(g, tb, v, j, st, len);
            uint32_t inc = len;  // inclusive scan of segment lengths
            for (int d = 1; d < WAVE; d <<= 1) {
                uint32_t o = __shfl_up(inc, d, 64);
                if (lane >= d) inc += o;
            }
            sst[lane] = st;
            spre[lane] = inc - len;
            sslot[lane] = j;
            const uint32_t K = __shfl(inc, WAVE - 1, 64);
            __builtin_amdgcn_wave_barrier();
            for (uint32_t k0 = 0; k0 < K; k0 += WAVE) {
                const uint32_t k = k0 + lane;
                double val = 0.0;
                if (k < K) {
                    int lo = 0, hi = WAVE - 1;  // last segment with pre <= k
                    while (lo < hi) {
                        int mid = (lo + hi + 1) >> 1;
                        if (spre[mid] <= k) lo = mid;
                        else hi = mid - 1;
                    }
                    const uint64_t kk = sst[lo] + (k - spre[lo]);
                    val = tb.t[kk].w;
                    if (PULL && !term_dead(g, tb, sslot[lo], kk)) val = 0.0;
                }
                const uint32_t nk = min((uint32_t)WAVE, K - k0);
                if (PULL) {
                    for (uint32_t b2 = 0; b2 < nk; ++b2) S -= readlane_f64(val, (int)b2);
                } else {
                    for (uint32_t b2 = 0; b2 < nk; ++b2) S += readlane_f64(val, (int)b2);
                }
            }
            c += K;
            __builtin_amdgcn_wave_barrier();
        }
        if (lane == 0) {
            g.pv[v] = S;
            if (!PULL) {
                g.pv0[v] = S;
                g.vtcnt[v] = c;
            }
        }
    }
}

// Per edge: capacity of its term list, min(d_x, d_y) - 1.
__global__ void __launch_bounds__(T) k_term_caps(GraphDev g, uint32_t* __restrict__ cap) {
    for (uint32_t e = blockIdx.x * T + threadIdx.x; e < g.m; e += gridDim.x * T) {
        const uint32_t x = g.esrc[e], y = g.nb[g.eslot[e]];
        cap[e] = min(g.deg[x], g.deg[y]) - 1;
    }
}

// Term offset of every slot's edge (zsl[j] = zoff[s2e[j]]): the hub kernels
// read a hub's owned edges' offsets with its list, contiguously.
__global__ void __launch_bounds__(T) k_zsl(GraphDev g, const uint64_t* __restrict__ zoff,
                                           uint64_t* __restrict__ zsl) {
    for (uint64_t j = (uint64_t)blockIdx.x * T + threadIdx.x; j < g.nnz;
         j += (uint64_t)gridDim.x * T)
        zsl[j] = zoff[g.s2e[j]];
}

template <int G>
__global__ void __launch_bounds__(T) k_count_vertices(GraphDev g,
                                                      const uint32_t* __restrict__ list,
                                                      uint32_t cnt) {
    const int lane = lane_id();
    const int gl = lane & (G - 1);
    const uint32_t per_wave = WAVE / G;
    const uint32_t wave = (blockIdx.x * T + threadIdx.x) / WAVE;
    const uint32_t nwaves = gridDim.x * T / WAVE;
    for (uint32_t base = wave * per_wave; base < cnt; base += nwaves * per_wave) {
        const uint32_t idx = base + lane / G;
        bool act = idx < cnt;
        uint32_t v = act ? list[idx] : 0;
        uint32_t dv = act ? g.deg[v] : 0;
        uint32_t j = act ? g.off[v] : 0, jend = act ? g.off[v + 1] : 0;
        double S = 0.0;
        uint32_t c = 0;
        bool have = false;
        uint32_t p = 0, le = 0, hlo = 0, he = 0, dy = 0;
        while (__any(act)) {
            if (act && !have) {
                // Next y whose "ids > y" parts of N(v) and N(y) both exist.
                while (j < jend) {
                    uint32_t y = g.nb[j];
                    uint32_t a0 = j + 1, a1 = jend;
                    uint32_t b1 = g.off[y + 1];
                    uint32_t b0 = lower_bound(g.nb, g.off[y], b1, y);
                    ++j;
                    if (a0 < a1 && b0 < b1) {
                        dy = g.deg[y];
                        if (a1 - a0 <= b1 - b0) {
                            p = a0; le = a1; hlo = b0; he = b1;
                        } else {
                            p = b0; le = b1; hlo = a0; he = a1;
                        }
                        have = true;
                        break;
                    }
                }
                if (!have) act = false;
            }
            bool go = act && have;
            uint32_t z;
            bool hit = probe_chunk<G>(g.nb, go, p, le, hlo, he, z);
            double t = hit ? tri_w(dv, dy, g.deg[z]) : 0.0;
            group_fold<G>(hit, t, S, c);
            if (go) {
                p += G;
                if (p >= le) have = false;
            }
        }
        if (idx < cnt && gl == 0) {
            g.pv[v] = S;
            g.pv0[v] = S;
            g.vtcnt[v] = c;
        }
    }
}

unsigned grid_for(uint32_t items, int G) {
    uint64_t waves = ((uint64_t)items * G + WAVE - 1) / WAVE;
    uint64_t blocks = (waves * WAVE + T - 1) / T;
    if (blocks < 1) blocks = 1;
    if (blocks > 16384) blocks = 16384;
    return (unsigned)blocks;
}

template <int G, int L>
void launch_hub_gl(const GraphDev& g, const uint4* units, uint32_t nu, const uint32_t* own_idx,
                   const TermBuf& tb, hipStream_t s) {
    if (L == WAVE) {
        const uint32_t blocks = std::min<uint32_t>((nu + HW / WAVE - 1) / (HW / WAVE), 16384);
        k_count_hub_w<G><<<blocks, HW, 0, s>>>(g, units, nu, own_idx, tb);
    } else {
        k_count_hub<G, 2, HT><<<std::min<uint32_t>(nu, 4096), HT, 0, s>>>(g, units, nu, own_idx,
                                                                          tb);
    }
}

// Large hubs: 8 waves per unit share the tree; LDS sized for the class's
// largest hub.
template <int G>
void launch_hub_large(const GraphDev& g, const uint4* units, uint32_t nu, const uint32_t* own_idx,
                     const TermBuf& tb, uint32_t maxdh, hipStream_t s) {
    const uint32_t lcap = (std::min<uint32_t>(maxdh, 4096) + 8 + 15) & ~15u;
    constexpr int FB = G == WAVE ? LARGE_FB : 0;
    const uint64_t arrays = large_slot_meta(FB) ? 3 : 1;
    k_count_hub_large<G, FB><<<std::min<uint32_t>(nu, 4096), HTL, arrays * lcap * 4, s>>>(
        g, units, nu, own_idx, tb, lcap);
}

template <int G>
void launch_hub_huge(const GraphDev& g, const uint4* units, uint32_t nu, const uint32_t* own_idx,
                     const TermBuf& tb, uint32_t tree_max, hipStream_t s) {
    constexpr int FB = G == WAVE ? HUGE_FB : 0;
    k_count_hub_huge<G, FB><<<std::min<uint32_t>(nu, 8192), HTL, 0, s>>>(g, units, nu, own_idx,
                                                                        tb, tree_max);
}

template <int L>
void launch_hub_l(int gc, const GraphDev& g, const uint4* units, uint32_t nu,
                  const uint32_t* own_idx, const TermBuf& tb, hipStream_t s) {
    switch (gc) {
        case 0: launch_hub_gl<4, L>(g, units, nu, own_idx, tb, s); break;
        case 1: launch_hub_gl<8, L>(g, units, nu, own_idx, tb, s); break;
        case 2: launch_hub_gl<16, L>(g, units, nu, own_idx, tb, s); break;
        case 3: launch_hub_gl<32, L>(g, units, nu, own_idx, tb, s); break;
        default: launch_hub_gl<64, L>(g, units, nu, own_idx, tb, s); break;
    }
}

// Every hub-path unit class, one launch each; with a term buffer, the small
// hubs' classes (stored first) in one batched launch and the medium and large
// classes in one streaming launch each.
void hub_pass(rtrex_graph h, const TermBuf& tb) {
    uint64_t at = 0;
    int c0 = 0;
    if (tb.t) {
        uint64_t nu = 0;
        for (int c = 0; c < HUB_NG; ++c) nu += h->hub_ucnt[c];
        if (nu) {
            const uint32_t tasks = h->small_nchunks;
            const uint32_t blocks = std::min<uint32_t>((tasks + SW / WAVE - 1) / (SW / WAVE), 65536);
            k_count_small<<<blocks, SW, 0, h->stream>>>(h->g, h->hub_units, (uint32_t)nu,
                                                        h->small_chunk, tasks, h->own_idx, tb);
        }
        at = nu;
        c0 = HUB_NG;
    }
    for (int c = c0; c < HUB_NL * HUB_NG; ++c) {
        const int lc = c / HUB_NG, gc = c % HUB_NG;
        if (tb.t && hub_stream_class(lc, gc)) {
            // the class's streaming classes (consecutive) in one launch
            if (gc) continue;
            uint32_t n = 0, mx = 0;
            for (int cc = c; cc < c + HUB_NG && hub_stream_class(lc, cc % HUB_NG); ++cc) {
                n += h->hub_ucnt[cc];
                mx = std::max(mx, h->hub_umax[cc]);
            }
            if (n) {
                const uint4* u = h->hub_units + at;
                const uint32_t blocks = std::min<uint32_t>(n, 16384);
                if (lc == 1)
                    k_count_hub_stream<2><<<blocks, HS_T, 2 * 512 * 4, h->stream>>>(
                        h->g, u, n, h->own_idx, tb, 512);
                else {
                    const uint32_t lcap = (std::min<uint32_t>(mx, 4096) + 8 + 15) & ~15u;
                    k_count_hub_stream<3><<<blocks, HS_T, 2ull * lcap * 4, h->stream>>>(
                        h->g, u, n, h->own_idx, tb, lcap);
                }
            }
            at += n;
            continue;
        }
        const uint32_t nu = h->hub_ucnt[c];
        if (!nu) continue;
        const uint4* u = h->hub_units + at;
        if (lc == 0) launch_hub_l<WAVE>(gc, h->g, u, nu, h->own_idx, tb, h->stream);
        else if (lc == 1) launch_hub_l<512>(gc, h->g, u, nu, h->own_idx, tb, h->stream);
        else if (lc == 3) {
            const uint32_t tl3 = std::min(h->cparams.l3_max, HUGE_TREE_MAX);
            switch (gc) {
                case 0: launch_hub_huge<4>(h->g, u, nu, h->own_idx, tb, tl3, h->stream); break;
                case 1: launch_hub_huge<8>(h->g, u, nu, h->own_idx, tb, tl3, h->stream); break;
                case 2: launch_hub_huge<16>(h->g, u, nu, h->own_idx, tb, tl3, h->stream); break;
                case 3: launch_hub_huge<32>(h->g, u, nu, h->own_idx, tb, tl3, h->stream); break;
                default: launch_hub_huge<64>(h->g, u, nu, h->own_idx, tb, tl3, h->stream); break;
            }
        } else {
            const uint32_t mx = h->hub_umax[c];
            switch (gc) {
                case 0: launch_hub_large<4>(h->g, u, nu, h->own_idx, tb, mx, h->stream); break;
                case 1: launch_hub_large<8>(h->g, u, nu, h->own_idx, tb, mx, h->stream); break;
                case 2: launch_hub_large<16>(h->g, u, nu, h->own_idx, tb, mx, h->stream); break;
                case 3: launch_hub_large<32>(h->g, u, nu, h->own_idx, tb, mx, h->stream); break;
                default: launch_hub_large<64>(h->g, u, nu, h->own_idx, tb, mx, h->stream); break;
            }
        }
        at += nu;
    }
}

// Processing order of the small-hub units (results do not depend on it): the
// units sorted by a MinHash of their hub's closed neighbourhood, min over
// z in N(h) + {h} of a mixed 32-bit hash of z.  Two hubs share that value with
// probability the Jaccard similarity of their neighbourhoods, so hubs with
// common neighbours tend to run near each other (one XCD, close in time) and
// re-read the same lists from its L2.  Measured on C2 against the units'
// build order: 1.6 % faster; against the hub id's hash alone: 0-4 %.  key = unit index << 32 | the hash's low 20
// bits, sorted on those bits (stable: ties keep their order).  The groups
// come in pseudo-random order; ordered by the whole hash (whose small values
// come from large neighbourhoods) or grouped by its top bits (minima are
// small, so the top bits merge many groups) measured 7-8 % slower.
__device__ __forceinline__ uint32_t mix32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}
__global__ void __launch_bounds__(T) k_small_keys(GraphDev g, const uint4* __restrict__ units,
                                                  uint32_t nu, uint64_t* __restrict__ keys) {
    for (uint32_t u = blockIdx.x * T + threadIdx.x; u < nu; u += gridDim.x * T) {
        const uint4 U = units[u];
        const uint32_t dh = U.w & UNIT_DH_MASK;
        uint32_t m = mix32(U.x);
        for (uint32_t i = 0; i < dh; ++i) m = min(m, mix32(g.nb[U.y + i]));
        keys[u] = ((uint64_t)u << 32) | (m & 0xfffffu);
    }
}
__global__ void __launch_bounds__(T) k_small_permute(const uint64_t* __restrict__ keys,
                                                     const uint4* __restrict__ in, uint32_t nu,
                                                     uint4* __restrict__ out) {
    for (uint32_t i = blockIdx.x * T + threadIdx.x; i < nu; i += gridDim.x * T)
        out[i] = in[(uint32_t)(keys[i] >> 32)];
}

// Streamed work of every small-hub unit: the sum of its owned edges' shorter
// lists (the degrees of the hub's neighbours at the owned slots).
__global__ void __launch_bounds__(T) k_small_work(GraphDev g, const uint4* __restrict__ units,
                                                  uint32_t nu, const uint32_t* __restrict__ own_idx,
                                                  uint32_t* __restrict__ w) {
    for (uint32_t u = blockIdx.x * T + threadIdx.x; u < nu; u += gridDim.x * T) {
        const uint4 U = units[u];
        const uint32_t cnt = U.w >> UNIT_DH_BITS;
        uint32_t sum = 0;
        for (uint32_t i = 0; i < cnt; ++i) sum += g.ndo[U.y + own_idx[U.z + i]].x;
        w[u] = sum;
    }
}

// Chunk c takes the units that start in the work window [c wt, (c + 1) wt)
// (P: exclusive prefix sums of the unit work): chunk[c] = the first unit with
// P >= c wt; chunk[nchunks] = nu.  A window no unit starts in is empty.
__global__ void __launch_bounds__(T) k_small_chunks(const uint64_t* __restrict__ P, uint32_t nu,
                                                    uint32_t wt, uint32_t nchunks,
                                                    uint32_t* __restrict__ chunk) {
    for (uint32_t u = blockIdx.x * T + threadIdx.x; u < nu; u += gridDim.x * T) {
        const uint64_t c0 = u ? P[u - 1] / wt + 1 : 0, c1 = P[u] / wt;
        for (uint64_t c = c0; c <= c1; ++c) chunk[c] = u;
        if (u == nu - 1)
            for (uint64_t c = c1 + 1; c <= nchunks; ++c) chunk[c] = nu;
    }
}

// Hub plan (per graph and parameters): ownership lists and work units.  The
// arrays are sized for any parameters (own_idx, hub_units: at most m
// entries, every unit has an owned edge) and refilled when they change.
rtrex_status prepare_hubs(rtrex_graph h) {
    const CountParams P = h->cparams;
    GraphDev& g = h->g;
    hipStream_t s = h->stream;
    Scratch sc;
    uint32_t *flag, *pos, *ccnt, *cbase, *hwork;
    RT_TRY(sc.get(&flag, g.nnz));
    RT_TRY(sc.get(&hwork, g.n));
    RT_HIP(hipMemsetAsync(hwork, 0, (uint64_t)g.n * 4, s));
    RT_TRY(sc.get(&pos, g.nnz + 1));
    RT_TRY(sc.get(&ccnt, 2 * HUB_NL * HUB_NG));  // unit counts, then largest degrees
    RT_TRY(sc.get(&cbase, HUB_NL * HUB_NG));
    RT_HIP(hipMemsetAsync(flag, 0, g.nnz * 4, s));
    const unsigned be = blocks_for(g.m, T, 8192);
    if (g.m) k_own_flags<<<be, T, 0, s>>>(g, flag, hwork, P);
    RT_TRY(exclusive_scan_u32_32(flag, pos, g.nnz, s));
    uint32_t nown = 0;
    RT_HIP(hipMemcpyAsync(&nown, pos + g.nnz, 4, hipMemcpyDeviceToHost, s));
    RT_HIP(hipStreamSynchronize(s));
    if (!h->own_off) {
        RT_TRY(dalloc(h, &h->own_off, (uint64_t)g.n + 1));
        RT_TRY(dalloc(h, &h->own_idx, std::max<uint64_t>(g.m, 1)));
        RT_TRY(dalloc(h, &h->hub_units, std::max<uint64_t>(g.m, 1)));
    }
    k_own_scatter<<<blocks_for(std::max<uint64_t>(g.m, g.n + 1ull), T, 8192), T, 0, s>>>(
        g, pos, h->own_off, h->own_idx, P);
    RT_HIP(hipMemsetAsync(ccnt, 0, 2 * HUB_NL * HUB_NG * 4, s));
    const unsigned bv = (unsigned)std::max<uint64_t>(1, ((uint64_t)g.n + T - 1) / T);  // one thread per hub
    if (nown)
        k_hub_units<true><<<bv, T, 0, s>>>(g, h->own_off, hwork, ccnt, nullptr, nullptr, P,
                                            ccnt + HUB_NL * HUB_NG);
    RT_HIP(hipGetLastError());
    RT_HIP(hipMemcpyAsync(h->hub_ucnt, ccnt, sizeof(h->hub_ucnt), hipMemcpyDeviceToHost, s));
    RT_HIP(hipMemcpyAsync(h->hub_umax, ccnt + HUB_NL * HUB_NG, sizeof(h->hub_umax),
                          hipMemcpyDeviceToHost, s));
    RT_HIP(hipStreamSynchronize(s));
    uint32_t base[HUB_NL * HUB_NG];
    uint64_t total = 0;
    for (int c = 0; c < HUB_NL * HUB_NG; ++c) {
        base[c] = (uint32_t)total;
        total += h->hub_ucnt[c];
    }
    if (total) {
        RT_HIP(hipMemcpyAsync(cbase, base, sizeof(base), hipMemcpyHostToDevice, s));
        RT_HIP(hipMemsetAsync(ccnt, 0, HUB_NL * HUB_NG * 4, s));
        k_hub_units<false><<<bv, T, 0, s>>>(g, h->own_off, hwork, ccnt, h->hub_units, cbase, P,
                                             nullptr);
        RT_HIP(hipGetLastError());
    }
    // Small-hub chunks (the classes below l0_max come first in hub_units).
    h->small_nchunks = 0;
    uint32_t nsmall = 0;
    for (int c = 0; c < HUB_NG; ++c) nsmall += h->hub_ucnt[c];
    if (nsmall) {
        {
            uint64_t *keys, *tmp;
            uint4* cpy;
            RT_TRY(sc.get(&keys, nsmall));
            RT_TRY(sc.get(&tmp, nsmall));
            RT_TRY(sc.get(&cpy, nsmall));
            k_small_keys<<<blocks_for(nsmall, T, 4096), T, 0, s>>>(g, h->hub_units, nsmall, keys);
            RT_TRY(radix_sort_u64(keys, tmp, nsmall, 20, s));
            RT_HIP(hipMemcpyAsync(cpy, h->hub_units, (uint64_t)nsmall * sizeof(uint4),
                                  hipMemcpyDeviceToDevice, s));
            k_small_permute<<<blocks_for(nsmall, T, 4096), T, 0, s>>>(keys, cpy, nsmall,
                                                                      h->hub_units);
            RT_HIP(hipGetLastError());
        }
        uint32_t* w;
        uint64_t* pre;
        RT_TRY(sc.get(&w, nsmall));
        RT_TRY(sc.get(&pre, nsmall + 1));
        k_small_work<<<blocks_for(nsmall, T, 4096), T, 0, s>>>(g, h->hub_units, nsmall,
                                                                h->own_idx, w);
        RT_TRY(exclusive_scan_u32(w, pre, nsmall, s));
        uint64_t last = 0;
        RT_HIP(hipMemcpyAsync(&last, pre + nsmall - 1, 8, hipMemcpyDeviceToHost, s));
        RT_HIP(hipStreamSynchronize(s));
        const uint64_t nch = last / S_WORK + 1;  // windows up to the last unit's start
        if (nch + 1 > h->small_chunk_cap) {
            h->small_chunk_cap = 2 * (nch + 1);
            RT_TRY(dalloc(h, &h->small_chunk, h->small_chunk_cap));
        }
        k_small_chunks<<<blocks_for(nsmall, T, 4096), T, 0, s>>>(pre, nsmall, S_WORK,
                                                                  (uint32_t)nch, h->small_chunk);
        RT_HIP(hipGetLastError());
        h->small_nchunks = (uint32_t)nch;
    }
    RT_HIP(hipStreamSynchronize(s));
    return RTREX_OK;
}

}  // namespace

// Term buffer of `total` 16-byte records when it fits in 60 % of the free HBM
// (the rest stays for the phase scratch and other work).
constexpr uint64_t TERM_BOUND_DIRECT = 4ull << 30;  // bytes

bool try_alloc_terms(rtrex_graph h, uint64_t total) {
    const uint64_t bytes = (total + 1) * sizeof(Term);
    if (h->slab && h->slab_top + bytes + 256 <= h->slab_cap) {
        h->zt = static_cast<Term*>(slab_take(h, bytes));
        return true;
    }
    size_t freeb = 0, totalb = 0;
    if (hipMemGetInfo(&freeb, &totalb) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    freeb += dev_cached_bytes();  // the cache gives its blocks back on demand
    if (bytes >= freeb / 5 * 3) return false;
    if (void* q = dev_alloc(bytes, nullptr)) {
        h->allocs.push_back(q);
        h->zt = static_cast<Term*>(q);
        return true;
    }
    return false;
}

// The count plan: work buckets and hub units (rebuilt when the parameters
// change) and the term buffer (depends on the graph only, built once).
rtrex_status prepare_count(rtrex_graph h) {
    GraphDev& g = h->g;
    hipStream_t s = h->stream;
    if (!h->plan_valid) {
        if (!h->cnt_lists)
            RT_TRY(dalloc(h, &h->cnt_lists, (uint64_t)NB * (g.m + g.n) + 2 * NB));
        uint32_t* elist = h->cnt_lists;
        uint32_t* vlist = elist + (uint64_t)NB * g.m;
        uint32_t* cur = vlist + (uint64_t)NB * g.n;
        RT_HIP(hipMemsetAsync(cur, 0, 2 * NB * sizeof(uint32_t), s));
        if (g.m) k_bucket<true><<<blocks_for(g.m, T), T, 0, s>>>(g, g.m, elist, cur, h->cparams);
        if (g.n)
            k_bucket<false><<<blocks_for(g.n, T), T, 0, s>>>(g, g.n, vlist, cur + NB, h->cparams);
        RT_HIP(hipGetLastError());
        RT_HIP(hipMemcpyAsync(h->cnt_counts, cur, sizeof(h->cnt_counts), hipMemcpyDeviceToHost, s));
        RT_TRY(prepare_hubs(h));
        h->plan_valid = true;
    }
    if (h->zoff) return RTREX_OK;
    RT_TRY(dalloc(h, &h->red, RED_BLOCKS));
    RT_TRY(dalloc(h, &h->srec, std::max<uint64_t>(g.m, 1)));
    RT_TRY(dalloc(h, &h->zsl, std::max<uint64_t>(g.nnz, 1)));
    // term buffer: capacity min(d_x, d_y) - 1 per edge
    RT_TRY(dalloc(h, &h->zoff, (uint64_t)g.m + 1));
    {
        Scratch sc;
        uint32_t* cap;
        RT_TRY(sc.get(&cap, g.m));
        if (g.m) k_term_caps<<<blocks_for(g.m, T, 8192), T, 0, s>>>(g, cap);
        RT_TRY(exclusive_scan_u32(cap, h->zoff, g.m, s));
    }
    if (g.nnz) k_zsl<<<blocks_for(g.nnz, T, 8192), T, 0, s>>>(g, h->zoff, h->zsl);
    RT_HIP(hipGetLastError());
    uint64_t total = 0;
    RT_HIP(hipMemcpyAsync(&total, h->zoff + g.m, 8, hipMemcpyDeviceToHost, s));
    RT_HIP(hipStreamSynchronize(s));
    h->zt = nullptr;
    h->terms_exact_pending = false;
    // The bound sizes the buffer directly when that costs no large
    // allocation: carved from a server job's range, small, or a block the
    // device cache already holds.  Otherwise a counting pass sizes it
    // exactly, which costs far less than mapping and releasing a large block
    // (R-MAT scale 20: a 110 GB bound for 20 GB of terms; about 3 s per
    // handle for the bound's block, 70 ms for the pass).
    const uint64_t bound_bytes = (total + 1) * sizeof(Term);
    const bool cheap = h->slab ? h->slab_top + bound_bytes + 256 <= h->slab_cap
                               : bound_bytes <= TERM_BOUND_DIRECT || dev_cache_fits(bound_bytes);
    if (cheap && try_alloc_terms(h, total)) return RTREX_OK;
    h->terms_exact_pending = true;
    return RTREX_OK;
}

// Launches the per-vertex folds over the stored triangle lists: the count
// fold (pull = false) or the post-clean pull (pull = true).
rtrex_status vertex_folds(rtrex_graph h, bool pull) {
    GraphDev& g = h->g;
    hipStream_t s = h->stream;
    const TermBuf tb = h->terms();
    const uint32_t* vlist = h->cnt_lists + (uint64_t)NB * g.m;
    const uint32_t* vc = h->cnt_counts + NB;
    for (int b = 0; b < 4; ++b) {
        if (!vc[b]) continue;
        const uint32_t* lst = vlist + (uint64_t)b * g.n;
        if (b < 2) {
            const unsigned grid = blocks_for(vc[b], T, 8192);
            if (pull) k_vfold_lane<true><<<grid, T, 0, s>>>(g, tb, lst, vc[b]);
            else k_vfold_lane<false><<<grid, T, 0, s>>>(g, tb, lst, vc[b]);
        } else {
            const unsigned grid =
                (unsigned)std::min<uint64_t>(((uint64_t)vc[b] * WAVE + VT - 1) / VT, 8192);
            if (pull) k_vfold_wave<true><<<grid, VT, 0, s>>>(g, tb, lst, vc[b]);
            else k_vfold_wave<false><<<grid, VT, 0, s>>>(g, tb, lst, vc[b]);
        }
    }
    RT_HIP(hipGetLastError());
    return RTREX_OK;
}

// Per-vertex supports over stored triangle lists are computed on demand: the
// reference only reads them for extraction candidates (Decomposition.h:856),
// so extraction asks for its component vertices and rtrex_get_supports for
// all.  The value is the count-phase fold, then (after a clean) minus the dead
// triangles in the same order -- the same as computing it eagerly.
rtrex_status ensure_vertex_supports(rtrex_graph h, const uint32_t* list, uint32_t cnt) {
    if (!h->zt || h->pv_full) return RTREX_OK;  // intersection paths keep them
    GraphDev& g = h->g;
    hipStream_t s = h->stream;
    const TermBuf tb = h->terms();
    const bool cleaned = h->phase >= PH_CLEANED;
    if (!list) {
        RT_TRY(vertex_folds(h, false));
        if (cleaned) RT_TRY(vertex_folds(h, true));
        h->pv_full = true;
    } else if (cnt) {
        // One wave per candidate: its slots' segment bounds (binary
        // searches) are found in parallel rather than one after another.
        const unsigned grid =
            (unsigned)std::min<uint64_t>(((uint64_t)cnt * WAVE + VT - 1) / VT, 8192);
        k_vfold_wave<false><<<grid, VT, 0, s>>>(g, tb, list, cnt);
        if (cleaned) k_vfold_wave<true><<<grid, VT, 0, s>>>(g, tb, list, cnt);
    }
    RT_HIP(hipGetLastError());
    return RTREX_OK;
}

// Clean state after a count: every slot and edge ALIVE, no edge touched or
// claimed, the clean's counters zero.  16 status bytes per thread step.
__global__ void __launch_bounds__(T) k_reset_clean(GraphDev g, Counters* ctr) {
    const uint64_t stride = (uint64_t)gridDim.x * T;
    const uint64_t t0 = (uint64_t)blockIdx.x * T + threadIdx.x;
    static_assert(ST_ALIVE == 0, "status bytes are cleared to zero");
    for (uint64_t i = t0; i < (g.nnz + 15) / 16; i += stride) {
        if (16 * i + 16 <= g.nnz && ((uintptr_t)g.st & 15) == 0)
            reinterpret_cast<uint4*>(g.st)[i] = make_uint4(0u, 0u, 0u, 0u);
        else
            for (uint64_t j = 16 * i; j < min<uint64_t>(16 * i + 16, g.nnz); ++j) g.st[j] = 0;
    }
    for (uint64_t e = t0; e < g.m; e += stride) {
        g.est[e] = ST_ALIVE;
        g.touched[e] = 0;
        g.claim[e] = 0;
    }
    if (t0 == 0) {
        ctr->tri_removed = 0;
        ctr->edges_removed = 0;
        ctr->ambiguous_clean = 0;
        ctr->ambiguous_ratio = 0;
        ctr->seeds = 0;
        ctr->n_undecided = 0;
        ctr->replayed_edges = 0;
    }
}

rtrex_status run_count(rtrex_graph h) {
    if (h->phase < PH_BUILT) return fail(RTREX_ESTATE, "graph not built");
    GraphDev& g = h->g;
    hipStream_t s = h->stream;
    RT_TRY(prepare_count(h));
    const uint32_t* counts = h->cnt_counts;
    uint32_t* elist = h->cnt_lists;
    uint32_t* vlist = elist + (uint64_t)NB * g.m;
    auto ep = [&](int b) { return elist + (uint64_t)b * g.m; };
    auto edge_pass = [&](const TermBuf& tb) {
        if (counts[0]) k_count_edges<1><<<grid_for(counts[0], 1), T, 0, s>>>(g, ep(0), counts[0], tb);
        if (counts[1]) k_count_edges<4><<<grid_for(counts[1], 4), T, 0, s>>>(g, ep(1), counts[1], tb);
        if (counts[2]) k_count_edges<16><<<grid_for(counts[2], 16), T, 0, s>>>(g, ep(2), counts[2], tb);
        if (counts[3]) k_count_edges<64><<<grid_for(counts[3], 64), T, 0, s>>>(g, ep(3), counts[3], tb);
        hub_pass(h, tb);
        k_sum_tcnt<<<RED_BLOCKS, T, 0, s>>>(h->srec, g.sup, g.sup0, g.tcnt, g.m, h->red);
        k_sum_final<<<1, T, 0, s>>>(h->red, RED_BLOCKS, &h->ctr->tri_total);
    };
    if (h->terms_exact_pending) {
        // Counting pass, exact offsets from the triangle counts, then the
        // buffer is filled by the timed pass below (once per graph).
        edge_pass(TermBuf{nullptr, h->zoff, h->zsl, h->srec});
        RT_HIP(hipGetLastError());
        RT_TRY(exclusive_scan_u32(g.tcnt, h->zoff, g.m, s));
        if (g.nnz) k_zsl<<<blocks_for(g.nnz, T, 8192), T, 0, s>>>(g, h->zoff, h->zsl);
        RT_HIP(hipGetLastError());
        uint64_t total = 0;
        RT_HIP(hipMemcpyAsync(&total, h->zoff + g.m, 8, hipMemcpyDeviceToHost, s));
        RT_HIP(hipStreamSynchronize(s));
        h->terms_exact_pending = false;
        try_alloc_terms(h, total);
    }
    const TermBuf tb = h->terms();

    RT_HIP(hipEventRecord(h->ev[2], s));
    edge_pass(tb);
    RT_HIP(hipEventRecord(h->ev[3], s));
    auto vp = [&](int b) { return vlist + (uint64_t)b * g.n; };
    const uint32_t* vc = counts + NB;
    h->pv_full = false;
    if (tb.t) {
        // on demand (ensure_vertex_supports)
    } else {
        if (vc[0]) k_count_vertices<1><<<grid_for(vc[0], 1), T, 0, s>>>(g, vp(0), vc[0]);
        if (vc[1]) k_count_vertices<4><<<grid_for(vc[1], 4), T, 0, s>>>(g, vp(1), vc[1]);
        if (vc[2]) k_count_vertices<16><<<grid_for(vc[2], 16), T, 0, s>>>(g, vp(2), vc[2]);
        if (vc[3]) k_count_vertices<64><<<grid_for(vc[3], 64), T, 0, s>>>(g, vp(3), vc[3]);
    }
    RT_HIP(hipGetLastError());
    RT_HIP(hipEventRecord(h->ev[4], s));
    // Fresh clean state for the supports just computed (one launch).
    k_reset_clean<<<blocks_for(std::max<uint64_t>(g.nnz / 16, 1), T, 4096), T, 0, s>>>(g, h->ctr);
    RT_HIP(hipGetLastError());
    // No wait here: the next phase's launches queue behind these; the phase
    // time is read from the events later (resolve_timings).
    h->tm_count_pending = true;
    h->audit = AuditHost();
    h->n_alive = g.m;
    h->phase = PH_COUNTED;
    return RTREX_OK;
}

rtrex_status resolve_timings(rtrex_graph h) {
    if (h->tm_count_pending) {
        RT_HIP(hipEventSynchronize(h->ev[4]));
        float a = 0, b = 0;
        RT_HIP(hipEventElapsedTime(&a, h->ev[2], h->ev[3]));
        RT_HIP(hipEventElapsedTime(&b, h->ev[3], h->ev[4]));
        h->tm.count_edge_ms = a;
        h->tm.count_vertex_ms = b;
        h->tm.count_ms = a + b;
        h->tm_count_pending = false;
    }
    if (h->tm_clean_pending) {
        RT_HIP(hipEventSynchronize(h->ev[6]));
        float ms = 0;
        RT_HIP(hipEventElapsedTime(&ms, h->ev[5], h->ev[6]));
        h->tm.clean_ms = ms;
        h->tm_clean_pending = false;
    }
    if (h->tm_extract_pending) {
        RT_HIP(hipEventSynchronize(h->ev[8]));
        float ms = 0;
        RT_HIP(hipEventElapsedTime(&ms, h->ev[7], h->ev[8]));
        h->tm.extract_ms = ms;
        h->tm_extract_pending = false;
    }
    return RTREX_OK;
}

}  // namespace rtrex
