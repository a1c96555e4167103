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
//   k_count_vertices<G>: G lanes per vertex walking y in N(v) in order, each
//       step the same chunked intersection restricted to ids > y.
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

// Appends ids to per-bucket lists: work(e) = length of the shorter list for
// edges, the degree for vertices.
template <bool EDGES>
__global__ void __launch_bounds__(T) k_bucket(GraphDev g, uint32_t count,
                                              uint32_t* __restrict__ lists,
                                              uint32_t* __restrict__ cursor) {
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
        } else {
            w = g.deg[i];
        }
        b = bucket_of(w);
        r = atomicAdd(&cnt[b], 1u);
    }
    __syncthreads();
    if (threadIdx.x < NB)
        base[threadIdx.x] = atomicAdd(&cursor[threadIdx.x], cnt[threadIdx.x]);
    __syncthreads();
    if (b >= 0) lists[(uint64_t)b * count + base[b] + r] = i;
}

// Folds, in lane order, the values t of the lanes of this lane's group whose
// hit flag is set.  All 64 lanes must call it.
template <int G>
__device__ __forceinline__ void group_fold(bool hit, double t, double& S,
                                           uint32_t& c) {
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
                                            uint32_t& z) {
    const int lane = lane_id();
    const int gl = lane & (G - 1);
    const int gbase = lane & ~(G - 1);
    bool in = act && p + gl < le;
    z = in ? nb[p + gl] : 0;
    uint32_t pos = in ? lower_bound(nb, hlo, he, z) : hlo;
    bool hit = in && pos < he && nb[pos] == z;
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
                                                   unsigned long long* tri3) {
    const int lane = lane_id();
    const int gl = lane & (G - 1);
    const uint32_t per_wave = WAVE / G;
    const uint32_t wave = (blockIdx.x * T + threadIdx.x) / WAVE;
    const uint32_t nwaves = gridDim.x * T / WAVE;
    unsigned long long mytri = 0;
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
        double S = 0.0;
        uint32_t c = 0;
        while (__any(act && p < le)) {
            bool go = act && p < le;
            uint32_t z;
            bool hit = probe_chunk<G>(g.nb, go, p, le, hlo, he, z);
            double t = hit ? tri_w(dx, dy, g.deg[z]) : 0.0;
            group_fold<G>(hit, t, S, c);
            if (go) p += G;
        }
        if (act && gl == 0) {
            g.sup[e] = S;
            g.sup0[e] = S;
            g.tcnt[e] = c;
            mytri += c;
        }
    }
    // Wave-reduce the triangle incidences, one atomic per wave.
    for (int d = 32; d >= 1; d >>= 1) mytri += __shfl_xor(mytri, d, 64);
    if (lane == 0 && mytri) atomicAdd(tri3, mytri);
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

}  // namespace

rtrex_status run_count(rtrex_graph h) {
    if (h->phase < PH_BUILT) return fail(RTREX_ESTATE, "graph not built");
    GraphDev& g = h->g;
    hipStream_t s = h->stream;
    Scratch sc;
    uint32_t *elist, *vlist, *cur;
    RT_TRY(sc.get(&elist, (uint64_t)NB * g.m));
    RT_TRY(sc.get(&vlist, (uint64_t)NB * g.n));
    RT_TRY(sc.get(&cur, 2 * NB));
    RT_HIP(hipMemsetAsync(cur, 0, 2 * NB * sizeof(uint32_t), s));
    RT_HIP(hipMemsetAsync(&h->ctr->tri_total, 0, sizeof(unsigned long long), s));
    if (g.m) k_bucket<true><<<blocks_for(g.m, T), T, 0, s>>>(g, g.m, elist, cur);
    if (g.n) k_bucket<false><<<blocks_for(g.n, T), T, 0, s>>>(g, g.n, vlist, cur + NB);
    RT_HIP(hipGetLastError());
    uint32_t counts[2 * NB];
    RT_HIP(hipMemcpyAsync(counts, cur, sizeof(counts), hipMemcpyDeviceToHost, s));
    RT_HIP(hipStreamSynchronize(s));

    RT_HIP(hipEventRecord(h->ev[2], s));
    unsigned long long* tri3 = &h->ctr->tri_total;
    auto ep = [&](int b) { return elist + (uint64_t)b * g.m; };
    if (counts[0]) k_count_edges<1><<<grid_for(counts[0], 1), T, 0, s>>>(g, ep(0), counts[0], tri3);
    if (counts[1]) k_count_edges<4><<<grid_for(counts[1], 4), T, 0, s>>>(g, ep(1), counts[1], tri3);
    if (counts[2]) k_count_edges<16><<<grid_for(counts[2], 16), T, 0, s>>>(g, ep(2), counts[2], tri3);
    if (counts[3]) k_count_edges<64><<<grid_for(counts[3], 64), T, 0, s>>>(g, ep(3), counts[3], tri3);
    RT_HIP(hipEventRecord(h->ev[3], s));
    auto vp = [&](int b) { return vlist + (uint64_t)b * g.n; };
    const uint32_t* vc = counts + NB;
    if (vc[0]) k_count_vertices<1><<<grid_for(vc[0], 1), T, 0, s>>>(g, vp(0), vc[0]);
    if (vc[1]) k_count_vertices<4><<<grid_for(vc[1], 4), T, 0, s>>>(g, vp(1), vc[1]);
    if (vc[2]) k_count_vertices<16><<<grid_for(vc[2], 16), T, 0, s>>>(g, vp(2), vc[2]);
    if (vc[3]) k_count_vertices<64><<<grid_for(vc[3], 64), T, 0, s>>>(g, vp(3), vc[3]);
    RT_HIP(hipGetLastError());
    RT_HIP(hipEventRecord(h->ev[4], s));
    RT_HIP(hipEventSynchronize(h->ev[4]));
    float a = 0, b = 0;
    RT_HIP(hipEventElapsedTime(&a, h->ev[2], h->ev[3]));
    RT_HIP(hipEventElapsedTime(&b, h->ev[3], h->ev[4]));
    h->tm.count_edge_ms = a;
    h->tm.count_vertex_ms = b;
    h->tm.count_ms = a + b;
    // Fresh clean state for the supports just computed.
    RT_HIP(hipMemsetAsync(g.st, ST_ALIVE, g.nnz, s));
    RT_HIP(hipMemsetAsync(g.touched, 0, g.m, s));
    RT_HIP(hipMemsetAsync(g.claim, 0, (uint64_t)g.m * 4, s));
    RT_HIP(hipMemsetAsync(&h->ctr->tri_removed, 0, 5 * sizeof(unsigned long long), s));
    RT_HIP(hipStreamSynchronize(s));
    h->audit = AuditHost();
    h->phase = PH_COUNTED;
    return RTREX_OK;
}

}  // namespace rtrex
