// Cascading clean: frontier peel to convergence.
//
// Reference: initialClean + deleteAndClean (Decomposition.h:147-460);
// gpuInitialClean / runBulkCleanLoop / markLowWeightEdgesKernel /
// processStackEdgesKernel (RTRex_CUDA/gpu_clean.cu:29-258).
//
// The reference pops one edge at a time from a LIFO stack, removes every live
// triangle through it (subtracting its weight from the three edges and three
// vertices) and pushes any other edge whose support falls below
// eps * 1/(d_u d_v).  Removal is monotone, so the survivors are the least
// fixpoint of "remove while below threshold" and the work can run in rounds
// over a frontier:
//   round r:  every frontier edge enumerates its live triangles (both other
//             edges not deleted); a triangle with several frontier edges is
//             removed once, by its smallest edge id; its weight is subtracted
//             (FP64 atomics) from its three edges and three vertices;
//   decision: an alive edge whose value is below its threshold by more than
//             the cheap bound (audit.hpp) is claimed (CAS) for the next
//             frontier at once; one within the bound goes to a near list and
//             is certified at the end of the round against every order the
//             reference could have subtracted in (certify()).
// There is no round cap (the reference's CUDA port stops at 100/50 rounds).
#include <algorithm>

#include "audit.hpp"
#include "dev.hpp"
#include "internal.hpp"

namespace rtrex {
namespace {

constexpr int T = 256;

// Initial frontier: alive edges below threshold (Decomposition.h:378-405).
// The count-phase supports are bit-identical to the reference's, so these
// decisions are the reference's exactly.
__global__ void __launch_bounds__(T) k_seed_frontier(GraphDev g, double eps,
                                                     uint32_t* __restrict__ F,
                                                     unsigned* nF) {
    for (uint32_t e = blockIdx.x * T + threadIdx.x; e < g.m;
         e += gridDim.x * T) {
        const uint32_t sl = g.eslot[e];
        if (g.st[sl] != ST_ALIVE) continue;
        const uint32_t x = g.esrc[e], y = g.nb[sl];
        if (g.sup[e] < edge_thr(eps, g.deg[x], g.deg[y])) {
            g.st[sl] = ST_QUEUED;
            g.st[g.part[sl]] = ST_QUEUED;
            F[atomicAdd(nF, 1u)] = e;
        }
    }
}

// Explicit deletions (rtrex_delete_and_clean): the listed edges, which must
// be alive, become the frontier.
__global__ void __launch_bounds__(T) k_seed_pairs(GraphDev g,
                                                  const uint32_t* __restrict__ pr,
                                                  uint32_t k,
                                                  uint32_t* __restrict__ F,
                                                  unsigned* nF, unsigned* err) {
    for (uint32_t q = blockIdx.x * T + threadIdx.x; q < k; q += gridDim.x * T) {
        uint32_t a = pr[2 * q], b = pr[2 * q + 1];
        if (a >= g.n || b >= g.n) {
            atomicOr(err, ERR_RANGE);
            continue;
        }
        int64_t sl = find_slot(g.nb, g.off[a], g.off[a + 1], b);
        if (sl < 0) {
            atomicOr(err, ERR_RANGE);
            continue;
        }
        const uint32_t e = g.s2e[sl];
        if (g.st[sl] != ST_ALIVE || atomicCAS(&g.claim[e], 0u, 2u) != 0u) {
            atomicOr(err, ERR_UNSORTED_OR_DUP);
            continue;
        }
        g.st[sl] = ST_QUEUED;
        g.st[g.part[sl]] = ST_QUEUED;
        F[atomicAdd(nF, 1u)] = e;
    }
}

__device__ __forceinline__ void claim_edge(const GraphDev& g, uint32_t f,
                                           uint32_t* Fn, unsigned* nFn) {
    if (atomicCAS(&g.claim[f], 0u, 1u) == 0u) {
        const uint32_t sl = g.eslot[f];
        g.st[sl] = ST_PENDING;
        g.st[g.part[sl]] = ST_PENDING;
        Fn[atomicAdd(nFn, 1u)] = f;
    }
}

struct PeelArgs {
    GraphDev g;
    double eps;
    const uint32_t* F;
    uint32_t nF;
    uint32_t* Fn;
    unsigned* nFn;
    uint32_t* near;
    unsigned* nNear;
    uint32_t* nearst;
    uint32_t stamp;  // this round's near-list stamp
    Counters* ctr;
};

// One peel round with G lanes per frontier edge.
template <int G>
__global__ void __launch_bounds__(T) k_peel(PeelArgs A) {
    const GraphDev& g = A.g;
    const int lane = lane_id();
    const int gl = lane & (G - 1);
    const int gbase = lane & ~(G - 1);
    const uint32_t per_wave = WAVE / G;
    const uint32_t wave = (blockIdx.x * T + threadIdx.x) / WAVE;
    const uint32_t nwaves = gridDim.x * T / WAVE;
    unsigned long long removed = 0;
    for (uint32_t base = wave * per_wave; base < A.nF; base += nwaves * per_wave) {
        const uint32_t idx = base + lane / G;
        bool act = idx < A.nF;
        const uint32_t e = act ? A.F[idx] : 0;
        const uint32_t x = act ? g.esrc[e] : 0;
        const uint32_t y = act ? g.nb[g.eslot[e]] : 0;
        const uint32_t dx = act ? g.deg[x] : 0, dy = act ? g.deg[y] : 0;
        const uint32_t lo = dx < dy ? x : y, hi = dx < dy ? y : x;
        const uint32_t dlo = dx < dy ? dx : dy, dhi = dx < dy ? dy : dx;
        uint32_t p = act ? g.off[lo] : 0;
        const uint32_t le = act ? g.off[lo + 1] : 0;
        uint32_t hlo = act ? g.off[hi] : 0;
        const uint32_t he = act ? g.off[hi + 1] : 0;
        while (__any(act && p < le)) {
            const bool go = act && p < le;
            const bool in = go && p + gl < le;
            const uint32_t j = p + gl;
            const uint8_t sj = in ? g.st[j] : ST_DEAD;
            const uint32_t w = in ? g.nb[j] : 0;
            uint32_t pos = in ? lower_bound(g.nb, hlo, he, w) : hlo;
            const uint32_t lastl = go ? min((uint32_t)(G - 1), le - p - 1) : 0u;
            const uint32_t nh = __shfl(pos, gbase + lastl, 64);
            bool tri = in && sj != ST_DEAD && pos < he && g.nb[pos] == w;
            uint8_t sh = tri ? g.st[pos] : ST_DEAD;
            tri = tri && sh != ST_DEAD;
            if (tri) {
                const uint32_t fj = g.s2e[j], fh = g.s2e[pos];
                const bool owned = !(sj == ST_QUEUED && fj < e) &&
                                   !(sh == ST_QUEUED && fh < e);
                if (owned) {
                    const uint32_t dw = g.deg[w];
                    const double t = tri_w(dx, dy, dw);
                    atomicAdd(&g.sup[e], -t);
                    atomicAdd(&g.pv[x], -t);
                    atomicAdd(&g.pv[y], -t);
                    atomicAdd(&g.pv[w], -t);
                    removed++;
                    // The two other edges: (lo, w) at slot j, (hi, w) at pos.
#pragma unroll
                    for (int side = 0; side < 2; ++side) {
                        const uint32_t f = side ? fh : fj;
                        const uint8_t sf = side ? sh : sj;
                        const double old = atomicAdd(&g.sup[f], -t);
                        if (sf == ST_QUEUED) continue;
                        g.touched[f] = 1;
                        if (sf != ST_ALIVE) continue;
                        const double nv = old - t;
                        const double thr = edge_thr(A.eps, side ? dhi : dlo, dw);
                        const double cb = cheap_bound(g.tcnt[f], g.sup0[f]);
                        if (nv < thr - cb) {
                            claim_edge(g, f, A.Fn, A.nFn);
                        } else if (nv < thr + cb &&
                                   atomicExch(&A.nearst[f], A.stamp) != A.stamp) {
                            A.near[atomicAdd(A.nNear, 1u)] = f;
                        }
                    }
                }
            }
            if (go) {
                hlo = nh;
                p += G;
            }
        }
    }
    for (int d = 32; d >= 1; d >>= 1) removed += __shfl_xor(removed, d, 64);
    if (lane == 0 && removed) atomicAdd(&A.ctr->tri_removed, removed);
}

// Certifies the near-threshold edges of the round (statuses: the frontier is
// QUEUED, so its triangles count as dead).
__global__ void __launch_bounds__(T) k_near_check(PeelArgs A, uint32_t nNear) {
    const GraphDev& g = A.g;
    unsigned long long amb = 0;
    for (uint32_t i = blockIdx.x * T + threadIdx.x; i < nNear; i += gridDim.x * T) {
        const uint32_t f = A.near[i];
        const uint32_t sl = g.eslot[f];
        if (g.st[sl] != ST_ALIVE) continue;  // already claimed
        const uint32_t x = g.esrc[f], y = g.nb[sl];
        const double thr = edge_thr(A.eps, g.deg[x], g.deg[y]);
        int c = certify(g, x, y, g.sup0[f], thr, true, 0xffffffffu);
        if (c == CERT_UNDECIDED) {
            amb++;
            c = g.sup[f] < thr ? CERT_REMOVE : CERT_KEEP;
        }
        if (c == CERT_REMOVE) claim_edge(g, f, A.Fn, A.nFn);
    }
    for (int d = 32; d >= 1; d >>= 1) amb += __shfl_xor(amb, d, 64);
    if (lane_id() == 0 && amb) atomicAdd(&A.ctr->ambiguous_clean, amb);
}

// End of a round: the frontier is deleted, the claimed edges are queued.
__global__ void __launch_bounds__(T) k_end_round(GraphDev g,
                                                 const uint32_t* __restrict__ F,
                                                 uint32_t nF,
                                                 const uint32_t* __restrict__ Fn,
                                                 uint32_t nFn, Counters* ctr) {
    for (uint32_t i = blockIdx.x * T + threadIdx.x; i < max(nF, nFn);
         i += gridDim.x * T) {
        if (i < nF) {
            const uint32_t sl = g.eslot[F[i]];
            g.st[sl] = ST_DEAD;
            g.st[g.part[sl]] = ST_DEAD;
        }
        if (i < nFn) {
            const uint32_t sl = g.eslot[Fn[i]];
            g.st[sl] = ST_QUEUED;
            g.st[g.part[sl]] = ST_QUEUED;
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0)
        atomicAdd(&ctr->edges_removed, (unsigned long long)nF);
}

// After convergence, a safety net: every alive edge that lost triangles and
// is not clearly above its threshold is certified once more.
__global__ void __launch_bounds__(T) k_audit_survivors(GraphDev g, double eps,
                                                       Counters* ctr) {
    unsigned long long amb = 0;
    for (uint32_t e = blockIdx.x * T + threadIdx.x; e < g.m;
         e += gridDim.x * T) {
        if (!g.touched[e]) continue;
        const uint32_t sl = g.eslot[e];
        if (g.st[sl] != ST_ALIVE) continue;
        const uint32_t x = g.esrc[e], y = g.nb[sl];
        const double thr = edge_thr(eps, g.deg[x], g.deg[y]);
        if (g.sup[e] >= thr + cheap_bound(g.tcnt[e], g.sup0[e])) continue;
        if (certify(g, x, y, g.sup0[e], thr, false, 0xffffffffu) != CERT_KEEP) amb++;
    }
    for (int d = 32; d >= 1; d >>= 1) amb += __shfl_xor(amb, d, 64);
    if (lane_id() == 0 && amb) atomicAdd(&ctr->ambiguous_clean, amb);
}

rtrex_status peel(rtrex_graph h, double eps, uint32_t* F, uint32_t* Fn,
                  uint32_t nF) {
    GraphDev& g = h->g;
    hipStream_t s = h->stream;
    Scratch sc;
    uint32_t* near;
    RT_TRY(sc.get(&near, g.m));
    if (!h->nearst) {
        RT_TRY(dalloc(h, &h->nearst, g.m));
        RT_HIP(hipMemsetAsync(h->nearst, 0, (uint64_t)g.m * 4, s));
    }
    unsigned* cnts = &h->ctr->frontier_next;  // [0] next frontier, [3] near
    PeelArgs A;
    A.g = g;
    A.eps = eps;
    A.nFn = cnts;
    A.near = near;
    A.nNear = &h->ctr->changed;
    A.nearst = h->nearst;
    A.ctr = h->ctr;
    while (nF) {
        RT_HIP(hipMemsetAsync(A.nFn, 0, sizeof(unsigned), s));
        RT_HIP(hipMemsetAsync(A.nNear, 0, sizeof(unsigned), s));
        A.F = F;
        A.nF = nF;
        A.Fn = Fn;
        A.stamp = ++h->round_stamp;
        uint64_t waves = ((uint64_t)nF * 16 + WAVE - 1) / WAVE;
        unsigned grid = (unsigned)std::min<uint64_t>((waves * WAVE + T - 1) / T, 8192);
        k_peel<16><<<grid, T, 0, s>>>(A);
        uint32_t nNear = 0;
        RT_HIP(hipMemcpyAsync(&nNear, A.nNear, sizeof(unsigned), hipMemcpyDeviceToHost, s));
        RT_HIP(hipStreamSynchronize(s));
        if (nNear)
            k_near_check<<<blocks_for(nNear, T, 4096), T, 0, s>>>(A, nNear);
        uint32_t nFn = 0;
        RT_HIP(hipMemcpyAsync(&nFn, A.nFn, sizeof(unsigned), hipMemcpyDeviceToHost, s));
        RT_HIP(hipStreamSynchronize(s));
        k_end_round<<<blocks_for(std::max(nF, nFn), T, 4096), T, 0, s>>>(
            g, F, nF, Fn, nFn, h->ctr);
        RT_HIP(hipGetLastError());
        h->audit.clean_rounds++;
        std::swap(F, Fn);
        nF = nFn;
    }
    k_audit_survivors<<<blocks_for(g.m, T, 4096), T, 0, s>>>(g, eps, h->ctr);
    RT_HIP(hipGetLastError());
    RT_HIP(hipStreamSynchronize(s));
    return RTREX_OK;
}

}  // namespace

rtrex_status run_initial_clean(rtrex_graph h, double eps) {
    if (h->phase < PH_COUNTED) return fail(RTREX_ESTATE, "count has not run");
    if (h->phase >= PH_CLEANED)
        return fail(RTREX_ESTATE, "initial clean already ran; run count again");
    if (!(eps >= 0)) return fail(RTREX_EINVAL, "eps must be >= 0");
    GraphDev& g = h->g;
    hipStream_t s = h->stream;
    Scratch sc;
    uint32_t *F, *Fn;
    RT_TRY(sc.get(&F, g.m));
    RT_TRY(sc.get(&Fn, g.m));
    RT_HIP(hipEventRecord(h->ev[5], s));
    unsigned* cnt = &h->ctr->frontier_next;
    RT_HIP(hipMemsetAsync(cnt, 0, sizeof(unsigned), s));
    k_seed_frontier<<<blocks_for(g.m, T, 8192), T, 0, s>>>(g, eps, F, cnt);
    RT_HIP(hipGetLastError());
    uint32_t nF = 0;
    RT_HIP(hipMemcpyAsync(&nF, cnt, sizeof(unsigned), hipMemcpyDeviceToHost, s));
    RT_HIP(hipStreamSynchronize(s));
    RT_TRY(peel(h, eps, F, Fn, nF));
    RT_HIP(hipEventRecord(h->ev[6], s));
    RT_HIP(hipEventSynchronize(h->ev[6]));
    float ms = 0;
    RT_HIP(hipEventElapsedTime(&ms, h->ev[5], h->ev[6]));
    h->tm.clean_ms = ms;
    h->eps = eps;
    h->phase = PH_CLEANED;
    return RTREX_OK;
}

rtrex_status run_delete_and_clean(rtrex_graph h, const int64_t* pairs,
                                  int64_t k, double eps) {
    if (h->phase < PH_COUNTED) return fail(RTREX_ESTATE, "count has not run");
    if (h->phase >= PH_EXTRACTED)
        return fail(RTREX_ESTATE, "graph already extracted; run count again");
    if (k < 0 || (k && !pairs)) return fail(RTREX_EINVAL, "bad pair list");
    if (k == 0) return RTREX_OK;
    GraphDev& g = h->g;
    hipStream_t s = h->stream;
    std::vector<uint32_t> pr(2 * k);
    for (int64_t i = 0; i < 2 * k; ++i) {
        if (pairs[i] < 0 || pairs[i] >= (int64_t)g.n)
            return fail(RTREX_EINVAL, "pair vertex out of range");
        pr[i] = (uint32_t)pairs[i];
    }
    Scratch sc;
    uint32_t *F, *Fn, *dpr;
    RT_TRY(sc.get(&F, g.m));
    RT_TRY(sc.get(&Fn, g.m));
    RT_TRY(sc.get(&dpr, 2 * k));
    RT_HIP(hipMemcpyAsync(dpr, pr.data(), pr.size() * 4, hipMemcpyHostToDevice, s));
    unsigned* cnt = &h->ctr->frontier_next;
    RT_HIP(hipMemsetAsync(cnt, 0, sizeof(unsigned), s));
    RT_HIP(hipMemsetAsync(&h->ctr->error, 0, sizeof(unsigned), s));
    k_seed_pairs<<<blocks_for(k, T, 4096), T, 0, s>>>(g, dpr, (uint32_t)k, F, cnt,
                                                      &h->ctr->error);
    RT_HIP(hipGetLastError());
    uint32_t nF = 0;
    unsigned err = 0;
    RT_HIP(hipMemcpyAsync(&nF, cnt, sizeof(unsigned), hipMemcpyDeviceToHost, s));
    RT_HIP(hipMemcpyAsync(&err, &h->ctr->error, sizeof(unsigned),
                          hipMemcpyDeviceToHost, s));
    RT_HIP(hipStreamSynchronize(s));
    if (err)
        return fail(RTREX_EINVAL,
                    "delete_and_clean: a pair is not an alive edge of the "
                    "graph, or is listed twice (state is undefined now)");
    RT_TRY(peel(h, eps, F, Fn, nF));
    return RTREX_OK;
}

}  // namespace rtrex
