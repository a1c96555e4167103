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

// Status change of edge e during a peel.  With stored triangle lists the
// peel reads the per-edge status only, so only est is written and the slot
// copies are brought in line once the clean ends (sync_slot_status); the
// intersection form reads the slot copies and keeps all three in step.
__device__ __forceinline__ void mark(const GraphDev& g, uint32_t e, uint8_t c, bool terms) {
    if (terms) g.est[e] = c;
    else set_status(g, e, c);
}

// Initial frontier: alive edges below threshold (Decomposition.h:378-405).
// The count-phase supports are bit-identical to the reference's, so these
// decisions are the reference's exactly.
__device__ __forceinline__ int work_bucket(uint32_t w) {
    return w <= 4u ? 0 : w <= 16u ? 1 : w <= 96u ? 2 : 3;
}

// When bk is given (stored triangle lists), the frontier is also written
// split by work bucket of tcnt (bk: 4 lists of capacity m, counts in bkc),
// and the edges that stay alive with triangles go to the alive buckets ak
// (same layout, counts in akc), the work list of a pull round.  At eps = 0.3 nearly every
// edge starts in the frontier, so appends are aggregated per workgroup (LDS
// counters, one global atomic per list per tile of SEED_K * T edges) rather
// than per wave.
constexpr int SEED_K = 8;
// The seeding workgroup is 1024 threads (tiles of 8192 edges): a quarter of
// the same-address global atomics of 256-thread workgroups at the same
// occupancy (about 18 % faster on a 2.1M-edge graph).
constexpr int SEED_T = 1024;
__global__ void __launch_bounds__(SEED_T) k_seed_frontier(GraphDev g, double eps,
                                                          uint32_t* __restrict__ F,
                                                          unsigned* nF,
                                                          uint32_t* __restrict__ bk,
                                                          unsigned* bkc,
                                                          uint32_t* __restrict__ ak,
                                                          unsigned* akc) {
    // counters: 0-3 frontier buckets, 4 frontier, 5-8 alive buckets
    __shared__ uint32_t lcnt[9], lbase[9];
    const bool terms = bk != nullptr;
    for (uint64_t tile = (uint64_t)blockIdx.x * SEED_K * SEED_T; tile < g.m;
         tile += (uint64_t)gridDim.x * SEED_K * SEED_T) {
        if (threadIdx.x < 9) lcnt[threadIdx.x] = 0;
        __syncthreads();
        int qb[SEED_K];
        uint32_t rf[SEED_K], rb[SEED_K];
#pragma unroll
        for (int k = 0; k < SEED_K; ++k) {
            const uint64_t e = tile + (uint64_t)k * SEED_T + threadIdx.x;
            qb[k] = -1;
            if (e < g.m && g.est[e] == ST_ALIVE) {
                const uint32_t tc = terms ? g.tcnt[e] : 0u;
                if (g.sup[e] < eps * g.ew[e]) {
                    mark(g, (uint32_t)e, ST_QUEUED, terms);
                    qb[k] = terms ? work_bucket(tc) : 0;
                    rf[k] = atomicAdd(&lcnt[4], 1u);
                    if (terms) rb[k] = atomicAdd(&lcnt[qb[k]], 1u);
                } else if (terms && tc) {
                    qb[k] = 8 + work_bucket(tc);  // alive bucket
                    rb[k] = atomicAdd(&lcnt[5 + work_bucket(tc)], 1u);
                }
            }
        }
        __syncthreads();
        if (threadIdx.x < 9 && lcnt[threadIdx.x]) {
            const int i = threadIdx.x;
            unsigned* c = i == 4 ? nF : i < 4 ? bkc + i : akc + (i - 5);
            if (i == 4 || terms) lbase[i] = atomicAdd(c, lcnt[i]);
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < SEED_K; ++k) {
            if (qb[k] < 0) continue;
            const uint32_t e = (uint32_t)(tile + (uint64_t)k * SEED_T + threadIdx.x);
            if (qb[k] >= 8) {
                const int b = qb[k] - 8;
                ak[(uint64_t)b * g.m + lbase[5 + b] + rb[k]] = e;
                continue;
            }
            F[lbase[4] + rf[k]] = e;
            if (terms) bk[(uint64_t)qb[k] * g.m + lbase[qb[k]] + rb[k]] = e;
        }
        __syncthreads();  // lcnt / lbase reused by the next tile
    }
}

// Explicit deletions (rtrex_delete_and_clean).  First the edge of every
// pair and whether it is alive, with no write to the graph (the host checks
// them); then the listed edges, all alive and distinct, become the frontier.
constexpr uint32_t PAIR_NOT_EDGE = 0xffffffffu, PAIR_NOT_ALIVE = 0x80000000u;
__global__ void __launch_bounds__(T) k_pair_edges(GraphDev g, const uint32_t* __restrict__ pr,
                                                  uint32_t k, uint32_t* __restrict__ eid) {
    for (uint32_t q = blockIdx.x * T + threadIdx.x; q < k; q += gridDim.x * T) {
        const uint32_t a = pr[2 * q], b = pr[2 * q + 1];
        const int64_t sl = a < g.n && b < g.n ? find_slot(g.nb, g.off[a], g.off[a + 1], b) : -1;
        if (sl < 0) {
            eid[q] = PAIR_NOT_EDGE;
            continue;
        }
        const uint32_t e = g.s2e[sl];
        eid[q] = e | (g.est[e] != ST_ALIVE ? PAIR_NOT_ALIVE : 0u);
    }
}

__global__ void __launch_bounds__(T) k_seed_edges(GraphDev g, const uint32_t* __restrict__ eid,
                                                  uint32_t k, uint32_t* __restrict__ F,
                                                  bool terms) {
    for (uint32_t q = blockIdx.x * T + threadIdx.x; q < k; q += gridDim.x * T) {
        const uint32_t e = eid[q];
        g.claim[e] = 2u;
        mark(g, e, ST_QUEUED, terms);
        F[q] = e;
    }
}

__device__ __forceinline__ void claim_edge(const GraphDev& g, uint32_t f,
                                           uint32_t* Fn, unsigned* nFn, bool terms) {
    if (atomicCAS(&g.claim[f], 0u, 1u) == 0u) {
        mark(g, f, ST_PENDING, terms);
        Fn[wave_append(nFn)] = f;
    }
}

// Splits a frontier into four work buckets (shorter-list length <= 4, 16,
// 96, larger) so each edge gets 1, 4, 16 or 64 lanes.
__global__ void __launch_bounds__(T) k_frontier_buckets(GraphDev g, bool terms,
                                                        const uint32_t* __restrict__ F,
                                                        uint32_t nF,
                                                        uint32_t* __restrict__ out,
                                                        unsigned* cur) {
    for (uint32_t i = blockIdx.x * T + threadIdx.x; i < nF; i += gridDim.x * T) {
        const uint32_t e = F[i];
        const uint32_t w = terms ? g.tcnt[e]
                                 : min(g.deg[g.esrc[e]], g.deg[g.nb[g.eslot[e]]]);
        const int b = work_bucket(w);
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (b == k) out[(uint64_t)k * nF + wave_append(&cur[k])] = e;
    }
}

struct PeelArgs {
    GraphDev g;
    double eps;
    const uint32_t* F;
    uint32_t nF;               // frontier length, or an upper bound when nFd is set
    const unsigned* nFd;       // device-side length (no host round trip), or null
    uint32_t* Fn;
    unsigned* nFn;
    uint32_t* near;
    unsigned* nNear;
    uint32_t* nearst;
    uint32_t stamp;  // this round's near-list stamp
    Counters* ctr;
    TermBuf tb;
    uint32_t* undec;     // edges no bound could decide (initial clean), or null
    uint32_t undec_cap;  // capacity; ctr->n_undecided counts every attempt
};

// Records an edge whose decision no error bound settles (for the exact
// replay); past the capacity only the count grows.
__device__ __forceinline__ void note_undecided(uint32_t* list, uint32_t cap, Counters* ctr,
                                               uint32_t e) {
    if (!list) return;
    const unsigned i = atomicAdd(&ctr->n_undecided, 1u);
    if (i < cap) list[i] = e;
}

// Shared tail of both peel kernels for one owned triangle {x, y, z}: the two
// other edges f (slot sf_slot, status sf) lose t; an alive one that falls
// below its threshold is claimed, or near-listed when within the cheap bound.
__device__ __forceinline__ void peel_side(const PeelArgs& A, uint32_t f, uint8_t sf,
                                          double t) {
    const GraphDev& g = A.g;
    if (sf != ST_ALIVE) return;  // queued or claimed edges die anyway
    const double old = atomicAdd(&g.sup[f], -t);
    g.touched[f] = 1;
    const double nv = old - t;
    const double thr = A.eps * g.ew[f];
    const double cb = cheap_bound(g.tcnt[f], g.sup0[f]);
    if (nv < thr - cb) {
        claim_edge(g, f, A.Fn, A.nFn, A.tb.t != nullptr);
    } else if (nv < thr + cb && atomicExch(&A.nearst[f], A.stamp) != A.stamp) {
        A.near[wave_append(A.nNear)] = f;
    }
}

// One "push" peel round over stored triangle lists, G lanes per frontier edge:
// no intersections, each lane takes one (weight, edge, edge) record.  Vertex
// supports and the live-triangle total are not touched here: both are pulled
// once the clean converges.
template <int G>
__global__ void __launch_bounds__(T) k_peel_terms(PeelArgs A) {
    const GraphDev& g = A.g;
    const TermBuf& tb = A.tb;
    const int lane = lane_id();
    const int gl = lane & (G - 1);
    const uint32_t per_wave = WAVE / G;
    const uint32_t wave = (blockIdx.x * T + threadIdx.x) / WAVE;
    const uint32_t nwaves = gridDim.x * T / WAVE;
    const uint32_t nF = A.nFd ? *A.nFd : A.nF;
    for (uint32_t base = wave * per_wave; base < nF; base += nwaves * per_wave) {
        const uint32_t idx = base + lane / G;
        const bool act = idx < nF;
        const uint32_t e = act ? A.F[idx] : 0;
        const uint64_t k0 = act ? tb.zoff[e] : 0;
        const uint64_t k1 = act ? k0 + g.tcnt[e] : 0;
        uint64_t k = k0 + gl;
        while (__any(act && k - gl < k1)) {
            const bool go = act && k - gl < k1;
            const bool in = go && k < k1;
            if (in) {
                const Term q = tb.t[k];
                const uint32_t fj = q.sa, fh = q.sb;
                const uint8_t sj = g.est[fj], sh = g.est[fh];
                // Only a triangle with an alive edge needs work: that edge
                // loses t, once, by the smallest frontier edge id.
                if (sj != ST_DEAD && sh != ST_DEAD && (sj == ST_ALIVE || sh == ST_ALIVE) &&
                    !(sj == ST_QUEUED && fj < e) && !(sh == ST_QUEUED && fh < e)) {
                    const double t = q.w;
                    peel_side(A, fj, sj, t);
                    peel_side(A, fh, sh, t);
                }
            }
            k += G;
        }
    }
}

// Alive edges split by work bucket of tcnt (lists of capacity m, counts in
// cur[0..3]), block-aggregated appends.
__global__ void __launch_bounds__(T) k_alive_buckets(GraphDev g, uint32_t* __restrict__ bk,
                                                     unsigned* cur) {
    __shared__ uint32_t lcnt[4], lbase[4];
    for (uint64_t tile = (uint64_t)blockIdx.x * SEED_K * T; tile < g.m;
         tile += (uint64_t)gridDim.x * SEED_K * T) {
        if (threadIdx.x < 4) lcnt[threadIdx.x] = 0;
        __syncthreads();
        int qb[SEED_K];
        uint32_t rb[SEED_K];
#pragma unroll
        for (int k = 0; k < SEED_K; ++k) {
            const uint64_t e = tile + (uint64_t)k * T + threadIdx.x;
            qb[k] = -1;
            if (e < g.m && g.est[e] == ST_ALIVE && g.tcnt[e]) {
                qb[k] = work_bucket(g.tcnt[e]);
                rb[k] = atomicAdd(&lcnt[qb[k]], 1u);
            }
        }
        __syncthreads();
        if (threadIdx.x < 4 && lcnt[threadIdx.x])
            lbase[threadIdx.x] = atomicAdd(&cur[threadIdx.x], lcnt[threadIdx.x]);
        __syncthreads();
#pragma unroll
        for (int k = 0; k < SEED_K; ++k)
            if (qb[k] >= 0)
                bk[(uint64_t)qb[k] * g.m + lbase[qb[k]] + rb[k]] =
                    (uint32_t)(tile + (uint64_t)k * T + threadIdx.x);
        __syncthreads();
    }
}

// One "pull" peel round: every alive edge subtracts the triangles that die
// this round (one other edge in the frontier, none already deleted) and
// decides for itself.  Chosen when the frontier is larger than the alive
// set, e.g. the first round, where nearly every edge is below its threshold.
// No atomics on supports and no ownership rule: each triangle is seen by each
// of its alive edges exactly once.  G lanes per edge (bucketed by tcnt) sum
// the dying weights, in any order: the result is S0 minus the dead terms with
// at most tcnt roundings of size <= ulp(S0)/2 in all, inside the cheap bound,
// and every decision within that bound is certified from the terms.
// Blocks blk of nblk of one bucket's launch range.
template <int G>
__device__ __forceinline__ void pull_body(const PeelArgs& A, const uint32_t* __restrict__ list,
                                          uint32_t cnt, uint32_t blk, uint32_t nblk) {
    const GraphDev& g = A.g;
    const TermBuf& tb = A.tb;
    const int lane = lane_id();
    const int gl = lane & (G - 1);
    const uint32_t per_wave = WAVE / G;
    const uint32_t wave = (blk * T + threadIdx.x) / WAVE;
    const uint32_t nwaves = nblk * T / WAVE;
    for (uint32_t base = wave * per_wave; base < cnt; base += nwaves * per_wave) {
        const uint32_t idx = base + lane / G;
        const bool in = idx < cnt;
        const uint32_t e = in ? list[idx] : 0;
        // An item claimed by the last round's near check is queued by now.
        const uint8_t se = in ? g.est[e] : (uint8_t)ST_DEAD;
        const uint64_t z = in ? tb.zoff[e] : 0;
        const uint32_t tc = in ? g.tcnt[e] : 0u;
        // the decision's inputs, loaded with the list bounds (they depend on
        // e only) rather than after the group's sum
        const double sup_e = in ? g.sup[e] : 0.0, ew_e = in ? g.ew[e] : 0.0;
        const double sup0_e = in ? g.sup0[e] : 0.0;
        const bool act = se == ST_ALIVE;
        const uint64_t k0 = z;
        const uint64_t k1 = act ? z + tc : z;
        double d = 0.0;
        bool any = false;
        // U terms per lane per step: all their edge ids, then all their
        // statuses, are loaded before any is used (two round trips per step).
        constexpr int U = 4;
        for (uint64_t k = k0 + gl; k < k1; k += (uint64_t)U * G) {
            uint32_t fa[U], fb[U];
            double tw[U];
            uint8_t sa_[U], sb_[U];
#pragma unroll
            for (int r = 0; r < U; ++r) {
                const uint64_t q = k + (uint64_t)r * G;
                const Term rec = q < k1 ? tb.t[q] : Term{0.0, 0u, 0u};
                fa[r] = rec.sa;
                fb[r] = rec.sb;
                tw[r] = rec.w;
            }
#pragma unroll
            for (int r = 0; r < U; ++r) {
                const uint64_t q = k + (uint64_t)r * G;
                sa_[r] = q < k1 ? g.est[fa[r]] : (uint8_t)ST_DEAD;
                sb_[r] = q < k1 ? g.est[fb[r]] : (uint8_t)ST_DEAD;
            }
#pragma unroll
            for (int r = 0; r < U; ++r) {
                const uint8_t a = sa_[r], b = sb_[r];
                if ((a == ST_QUEUED || b == ST_QUEUED) && a != ST_DEAD && b != ST_DEAD) {
                    d += tw[r];
                    any = true;
                }
            }
        }
#pragma unroll
        for (int o = 1; o < G; o <<= 1) d += __shfl_xor(d, o, 64);
        const uint64_t ab = ballot64(any);
        const bool gany = G == 64 ? ab != 0 : ((ab >> (lane & ~(G - 1))) & ((1ull << G) - 1)) != 0;
        if (!act || gl != 0 || !gany) continue;
        const double v = sup_e - d;
        g.sup[e] = v;
        g.touched[e] = 1;
        const double thr = A.eps * ew_e;
        const double cb = cheap_bound(tc, sup0_e);
        if (v < thr - cb) {
            // This group is the only writer of claim[e] in a pull round;
            // k_collect_claims appends the claimed edges to the next
            // frontier (a per-wave append here would make one same-address
            // atomic per wave, nearly every wave, in the first round).
            g.claim[e] = 1u;
        } else if (v < thr + cb && atomicExch(&A.nearst[e], A.stamp) != A.stamp) {
            A.near[wave_append(A.nNear)] = e;
        }
    }
}

// The four alive buckets in one launch: blocks [bb[b], bb[b + 1]) take
// bucket b of ak (1, 4, 16 or 64 lanes per edge; counts in acnt[b]).
struct BucketBlocks {
    uint32_t bb[5];
};
__global__ void __launch_bounds__(T) k_peel_pull(PeelArgs A, const uint32_t* __restrict__ ak,
                                                 const unsigned* __restrict__ acnt,
                                                 BucketBlocks R) {
    const uint32_t x = blockIdx.x;
    const int b = x < R.bb[1] ? 0 : x < R.bb[2] ? 1 : x < R.bb[3] ? 2 : 3;
    const uint32_t* L = ak + (uint64_t)b * A.g.m;
    const uint32_t blk = x - R.bb[b], nblk = R.bb[b + 1] - R.bb[b];
    switch (b) {
        case 0: pull_body<1>(A, L, acnt[0], blk, nblk); break;
        case 1: pull_body<4>(A, L, acnt[1], blk, nblk); break;
        case 2: pull_body<16>(A, L, acnt[2], blk, nblk); break;
        default: pull_body<64>(A, L, acnt[3], blk, nblk); break;
    }
}

// After the pull kernels of a round, the items of the alive buckets ak
// (counts ac) that are still ALIVE split two ways: the claimed ones (claim set) go to the next
// frontier as PENDING, the rest to ak2, the next pull round's buckets (same
// bucket, counts ac2).  An edge k_near_check claims later in the round stays in ak2; the
// pull kernels skip items that are no longer ALIVE and the next collection
// drops them.  One global atomic per list per tile of SEED_K * T items.
__global__ void __launch_bounds__(T) k_collect_claims(GraphDev g, const uint32_t* __restrict__ ak,
                                                      const unsigned* __restrict__ ac,
                                                      uint32_t* __restrict__ Fn,
                                                      unsigned* nFn,
                                                      uint32_t* __restrict__ ak2,
                                                      unsigned* ac2) {
    __shared__ uint32_t lcnt[5], lbase[5];  // 0-3 buckets of ak2, 4 next frontier
    const uint32_t c[4] = {ac[0], ac[1], ac[2], ac[3]};
    const uint32_t na = c[0] + c[1] + c[2] + c[3];
    for (uint32_t tile = blockIdx.x * SEED_K * T; tile < na; tile += gridDim.x * SEED_K * T) {
        if (threadIdx.x < 5) lcnt[threadIdx.x] = 0;
        __syncthreads();
        int qb[SEED_K];
        uint32_t r[SEED_K], ev[SEED_K];
#pragma unroll
        for (int k = 0; k < SEED_K; ++k) {
            const uint32_t i = tile + k * T + threadIdx.x;
            qb[k] = -1;
            if (i < na) {
                uint32_t j = i;
                int b = 0;
                while (b < 3 && j >= c[b]) j -= c[b++];
                ev[k] = ak[(uint64_t)b * g.m + j];
                if (g.est[ev[k]] == ST_ALIVE) {
                    qb[k] = g.claim[ev[k]] ? 4 : b;
                    r[k] = atomicAdd(&lcnt[qb[k]], 1u);
                }
            }
        }
        __syncthreads();
        if (threadIdx.x < 5 && lcnt[threadIdx.x])
            lbase[threadIdx.x] =
                atomicAdd(threadIdx.x == 4 ? nFn : &ac2[threadIdx.x], lcnt[threadIdx.x]);
        __syncthreads();
#pragma unroll
        for (int k = 0; k < SEED_K; ++k) {
            if (qb[k] < 0) continue;
            if (qb[k] == 4) {
                g.est[ev[k]] = ST_PENDING;  // pull rounds run on stored triangle lists only
                Fn[lbase[4] + r[k]] = ev[k];
            } else {
                ak2[(uint64_t)qb[k] * g.m + lbase[qb[k]] + r[k]] = ev[k];
            }
        }
        __syncthreads();  // lcnt / lbase are reset for the next tile
    }
}

// Slot statuses from the edge statuses, after a peel on stored triangle
// lists (which writes est only).
// Four slots per step (one 16-byte load of edge ids, one 4-byte store) when
// the arrays allow it.
__device__ __forceinline__ void sync_slot_status(const GraphDev& g, uint64_t t0,
                                                 uint64_t stride) {
    const bool vec = (((uintptr_t)g.s2e & 15) | ((uintptr_t)g.st & 3)) == 0;
    const uint64_t n4 = vec ? g.nnz / 4 : 0;
    const uint4* __restrict__ e4 = reinterpret_cast<const uint4*>(g.s2e);
    uint32_t* __restrict__ st4 = reinterpret_cast<uint32_t*>(g.st);
    for (uint64_t q = t0; q < n4; q += stride) {
        const uint4 e = e4[q];
        st4[q] = (uint32_t)g.est[e.x] | (uint32_t)g.est[e.y] << 8 | (uint32_t)g.est[e.z] << 16 |
                 (uint32_t)g.est[e.w] << 24;
    }
    for (uint64_t j = 4 * n4 + t0; j < g.nnz; j += stride) g.st[j] = g.est[g.s2e[j]];
}
__global__ void __launch_bounds__(T) k_slot_status(GraphDev g) {
    sync_slot_status(g, (uint64_t)blockIdx.x * T + threadIdx.x, (uint64_t)gridDim.x * T);
}

// Triangles removed by the clean: the count phase's total minus the live ones
// (incidences are three per triangle).
__global__ void k_tri_removed(const unsigned long long* __restrict__ inc, Counters* ctr) {
    if (threadIdx.x == 0) ctr->tri_removed = ctr->tri_total / 3 - *inc / 3;
}

// Triangles with all three edges alive, counted once per alive edge.
__global__ void __launch_bounds__(T) k_live_incidences(GraphDev g, TermBuf tb,
                                                       unsigned long long* out) {
    unsigned long long c = 0;
    for (uint32_t e = blockIdx.x * T + threadIdx.x; e < g.m; e += gridDim.x * T) {
        if (g.est[e] != ST_ALIVE) continue;
        const uint64_t k0 = tb.zoff[e], k1 = k0 + g.tcnt[e];
        for (uint64_t k = k0; k < k1; ++k)
            c += g.est[tb.t[k].sa] == ST_ALIVE && g.est[tb.t[k].sb] == ST_ALIVE;
    }
    for (int d = 32; d >= 1; d >>= 1) c += __shfl_xor(c, d, 64);
    if (lane_id() == 0 && c) atomicAdd(out, c);
}

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
    const uint32_t nF = A.nFd ? *A.nFd : A.nF;
    for (uint32_t base = wave * per_wave; base < nF; base += nwaves * per_wave) {
        const uint32_t idx = base + lane / G;
        bool act = idx < nF;
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
            double t = 0.0;
            if (tri) {
                const uint32_t fj = g.s2e[j], fh = g.s2e[pos];
                const bool owned = !(sj == ST_QUEUED && fj < e) &&
                                   !(sh == ST_QUEUED && fh < e);
                if (owned) {
                    const uint32_t dw = g.deg[w];
                    t = tri_w(dx, dy, dw);
                    atomicAdd(&g.pv[w], -t);
                    removed++;
                    // The two other edges: (lo, w) at slot j, (hi, w) at pos.
                    // Edges already queued or claimed die anyway: their
                    // supports are not maintained (set to 0 at deletion).
#pragma unroll
                    for (int side = 0; side < 2; ++side) {
                        const uint32_t f = side ? fh : fj;
                        const uint8_t sf = side ? sh : sj;
                        if (sf != ST_ALIVE) continue;
                        const double old = atomicAdd(&g.sup[f], -t);
                        g.touched[f] = 1;
                        const double nv = old - t;
                        const double thr = edge_thr(A.eps, side ? dhi : dlo, dw);
                        const double cb = cheap_bound(g.tcnt[f], g.sup0[f]);
                        if (nv < thr - cb) {
                            claim_edge(g, f, A.Fn, A.nFn, false);  // intersection form
                        } else if (nv < thr + cb &&
                                   atomicExch(&A.nearst[f], A.stamp) != A.stamp) {
                            A.near[wave_append(A.nNear)] = f;
                        }
                    }
                }
            }
            // x and y lose every owned triangle of the chunk: one subtraction
            // of the group's sum each.
            double ts = t;
#pragma unroll
            for (int d = 1; d < G; d <<= 1) ts += __shfl_xor(ts, d, 64);
            if (go && gl == 0 && ts != 0.0) {
                atomicAdd(&g.pv[x], -ts);
                atomicAdd(&g.pv[y], -ts);
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
__global__ void __launch_bounds__(T) k_near_check(PeelArgs A) {
    const GraphDev& g = A.g;
    const uint32_t nNear = *A.nNear;
    unsigned long long amb = 0;
    for (uint32_t i = blockIdx.x * T + threadIdx.x; i < nNear; i += gridDim.x * T) {
        const uint32_t f = A.near[i];
        if (g.est[f] != ST_ALIVE) continue;  // already claimed
        const double thr = A.eps * g.ew[f];
        int c = certify_edge(g, A.tb, f, thr, true, 0xffffffffu, 0xffffffffu);
        if (c == CERT_UNDECIDED) {
            amb++;
            note_undecided(A.undec, A.undec_cap, A.ctr, f);
            c = g.sup[f] < thr ? CERT_REMOVE : CERT_KEEP;
        }
        if (c == CERT_REMOVE) claim_edge(g, f, A.Fn, A.nFn, A.tb.t != nullptr);
    }
    for (int d = 32; d >= 1; d >>= 1) amb += __shfl_xor(amb, d, 64);
    if (lane_id() == 0 && amb) atomicAdd(&A.ctr->ambiguous_clean, amb);
}

// The same over stored triangle lists, G lanes per near edge: the lanes of a
// group walk the edge's list U records per step (all edge ids, then all
// statuses, loaded before any is used), each lane sums the weights of its dead
// triangles in double-double (exact partial sums) and appends them to the
// group's LDS slots while there are at most kMaxPerm, the group adds the
// partials pairwise, and the group's first lane decides from s0 + sum; within
// the error band, it enumerates the orders of the gathered weights.
template <int G>
__global__ void __launch_bounds__(T) k_near_check_terms(PeelArgs A) {
    __shared__ double s_t[T / G][kMaxPerm];
    __shared__ uint32_t s_n[T / G];
    const GraphDev& g = A.g;
    const TermBuf& tb = A.tb;
    const uint32_t nNear = *A.nNear;
    const int lane = lane_id();
    const int gl = lane & (G - 1);
    const int grp = (int)threadIdx.x / G;
    const uint32_t per_wave = WAVE / G;
    const uint32_t wave = (blockIdx.x * T + threadIdx.x) / WAVE;
    const uint32_t nwaves = gridDim.x * T / WAVE;
    unsigned long long amb = 0;
    for (uint32_t base = wave * per_wave; base < nNear; base += nwaves * per_wave) {
        const uint32_t idx = base + lane / G;
        const bool in = idx < nNear;
        const uint32_t f = in ? A.near[idx] : 0;
        const bool act = in && g.est[f] == ST_ALIVE;  // else already claimed
        const uint64_t k0 = act ? tb.zoff[f] : 0;
        const uint64_t k1 = act ? k0 + g.tcnt[f] : 0;
        if (gl == 0) s_n[grp] = 0;
        wave_sync();
        double hi = 0.0, lo = 0.0;
        uint32_t nd = 0;
        constexpr int U = 4;
        for (uint64_t k = k0 + gl; k < k1; k += (uint64_t)U * G) {
            uint32_t fa[U], fb[U];
            double tw[U];
            uint8_t sa_[U], sb_[U];
#pragma unroll
            for (int r = 0; r < U; ++r) {
                const uint64_t q = k + (uint64_t)r * G;
                const Term rec = q < k1 ? tb.t[q] : Term{0.0, 0u, 0u};
                fa[r] = rec.sa;
                fb[r] = rec.sb;
                tw[r] = rec.w;
            }
#pragma unroll
            for (int r = 0; r < U; ++r) {
                const uint64_t q = k + (uint64_t)r * G;
                sa_[r] = q < k1 ? g.est[fa[r]] : (uint8_t)ST_ALIVE;
                sb_[r] = q < k1 ? g.est[fb[r]] : (uint8_t)ST_ALIVE;
            }
#pragma unroll
            for (int r = 0; r < U; ++r) {
                if (tri_dead(sa_[r], sb_[r], true)) {
                    double sn, e;
                    two_sum(hi, -tw[r], sn, e);
                    hi = sn;
                    lo += e;
                    ++nd;
                    const uint32_t slot = atomicAdd(&s_n[grp], 1u);
                    if (slot < (uint32_t)kMaxPerm) s_t[grp][slot] = tw[r];
                }
            }
        }
#pragma unroll
        for (int o = 1; o < G; o <<= 1) {
            const double h2 = __shfl_xor(hi, o, 64);
            const double l2 = __shfl_xor(lo, o, 64);
            nd += __shfl_xor(nd, o, 64);
            double sn, e;
            two_sum(hi, h2, sn, e);
            hi = sn;
            lo = lo + l2 + e;
        }
        wave_sync();  // the group's gathered weights are visible to its first lane
        if (!act || gl != 0) continue;
        const double s0 = g.sup0[f];
        const double thr = A.eps * g.ew[f];
        double sn, e;
        two_sum(s0, hi, sn, e);
        int c = dd_decide(sn, lo + e, nd, s0, thr);
        if (c == CERT_UNDECIDED && nd <= (uint32_t)kMaxPerm)
            c = enumerate_orders<true>(s0, thr, s_t[grp], (int)nd);
        if (c == CERT_UNDECIDED) {
            amb++;
            note_undecided(A.undec, A.undec_cap, A.ctr, f);
            c = g.sup[f] < thr ? CERT_REMOVE : CERT_KEEP;
        }
        if (c == CERT_REMOVE) claim_edge(g, f, A.Fn, A.nFn, true);
    }
    for (int d = 32; d >= 1; d >>= 1) amb += __shfl_xor(amb, d, 64);
    if (lane == 0 && amb) atomicAdd(&A.ctr->ambiguous_clean, amb);
}

// End of a round: the frontier is deleted, the claimed edges are queued.
// It also clears the counters the next round appends to (zero2: the other
// parity's next-frontier and near-list counts) and, after a pull round, the
// counts of the alive lists just consumed (zero4), which become the next
// round's output lists: no clearing launch per round.  The next frontier's
// length goes to the host mailbox at once, so the host sizes the next round
// while this kernel still runs.
__global__ void __launch_bounds__(T) k_end_round(GraphDev g,
                                                 const uint32_t* __restrict__ F,
                                                 uint32_t nF,
                                                 const uint32_t* __restrict__ Fn,
                                                 const unsigned* __restrict__ nFnp,
                                                 Counters* ctr, bool terms,
                                                 unsigned* zero2, unsigned* zero4,
                                                 uint64_t* box, uint32_t seq) {
    const uint32_t nFn = *nFnp;
    if (blockIdx.x == 0 && threadIdx.x < 4) {
        if (threadIdx.x == 0) publish(box, seq, nFn);  // final: every appender has ended
        if (threadIdx.x < 2) zero2[threadIdx.x] = 0;
        if (zero4) zero4[threadIdx.x] = 0;
    }
    for (uint32_t i = blockIdx.x * T + threadIdx.x; i < max(nF, nFn);
         i += gridDim.x * T) {
        if (i < nF) {
            const uint32_t f = F[i];
            mark(g, f, ST_DEAD, terms);
            g.sup[f] = 0.0;
        }
        if (i < nFn) mark(g, Fn[i], ST_QUEUED, terms);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0)
        atomicAdd(&ctr->edges_removed, (unsigned long long)nF);
}

// After convergence, one pass over the edges:
//  - a safety net: every alive edge that lost triangles and is not clearly
//    above its threshold is certified once more;
//  - xact (when given): a support the fast peel never decremented is the
//    count-phase value, which is the reference's (no triangle of the edge
//    died there either);
//  - inc (when given, stored lists): triangles with all three edges alive,
//    counted once per alive edge.
__global__ void __launch_bounds__(T) k_post_peel(GraphDev g, TermBuf tb, double eps,
                                                 Counters* ctr, uint32_t* undec,
                                                 uint32_t undec_cap, uint8_t* __restrict__ xact,
                                                 unsigned long long* inc) {
    unsigned long long amb = 0, live = 0;
    for (uint32_t e = blockIdx.x * T + threadIdx.x; e < g.m;
         e += gridDim.x * T) {
        const uint8_t touched = g.touched[e];
        if (xact) xact[e] = touched ? 0 : 1;
        if (g.est[e] != ST_ALIVE) continue;
        if (inc) {
            const uint64_t k0 = tb.zoff[e], k1 = k0 + g.tcnt[e];
            for (uint64_t k = k0; k < k1; ++k) {
                const Term q = tb.t[k];
                live += g.est[q.sa] == ST_ALIVE && g.est[q.sb] == ST_ALIVE;
            }
        }
        if (!touched) continue;
        const double thr = eps * g.ew[e];
        if (g.sup[e] >= thr + cheap_bound(g.tcnt[e], g.sup0[e])) continue;
        if (certify_edge(g, tb, e, thr, false, 0xffffffffu, 0xffffffffu) != CERT_KEEP) {
            amb++;
            note_undecided(undec, undec_cap, ctr, e);
        }
    }
    for (int d = 32; d >= 1; d >>= 1) {
        amb += __shfl_xor(amb, d, 64);
        live += __shfl_xor(live, d, 64);
    }
    if (lane_id() == 0 && amb) atomicAdd(&ctr->ambiguous_clean, amb);
    if (lane_id() == 0 && live) atomicAdd(inc, live);
}

// Packed survivor bitmap (bit e = edge e alive), kept after the initial
// clean; with slots, the slot statuses are brought in line in the same launch.
__global__ void __launch_bounds__(T) k_survivor_bits(GraphDev g, uint32_t* bits, bool slots) {
    if (slots) sync_slot_status(g, (uint64_t)blockIdx.x * T + threadIdx.x, (uint64_t)gridDim.x * T);
    const uint32_t nw = (g.m + 31) / 32;
    const bool vec = ((uintptr_t)g.est & 15) == 0;
    static_assert(ST_ALIVE == 0, "alive status bytes are zero");
    for (uint32_t w = blockIdx.x * T + threadIdx.x; w < nw; w += gridDim.x * T) {
        uint32_t b = 0;
        if (vec && (uint64_t)w * 32 + 32 <= g.m) {  // the word's 32 status bytes in two loads
            const uint4* q = reinterpret_cast<const uint4*>(g.est) + 2ull * w;
            const uint4 lo = q[0], hi = q[1];
            const uint32_t x[8] = {lo.x, lo.y, lo.z, lo.w, hi.x, hi.y, hi.z, hi.w};
#pragma unroll
            for (int k = 0; k < 8; ++k)
#pragma unroll
                for (int i = 0; i < 4; ++i)
                    b |= (uint32_t)(((x[k] >> (8 * i)) & 0xffu) == 0u) << (4 * k + i);
        } else {
            for (uint32_t i = 0; i < 32; ++i) {
                const uint32_t e = w * 32 + i;
                if (e < g.m && g.est[e] == ST_ALIVE) b |= 1u << i;
            }
        }
        bits[w] = b;
    }
}

// ------------------------------------------------------------ exact replay
// When a cleaning decision is undecided (its value lies within the error
// bound of the reference's own subtraction order), the clean re-runs, from
// the count-phase supports, every triangle-connected region holding such an
// edge in the reference's own order: the initial pushes of the region's
// below-threshold edges in ascending edge id (initialClean,
// Decomposition.h:378-405), then LIFO pops (deleteAndClean,
// Decomposition.h:147-283) that visit the popped edge's triangles in
// ascending third vertex, subtract the weight from both other edges and push
// (lower, w) before (higher, w), lower being the popped pair's endpoint of
// smaller degree, its second on a tie.  Edges of different regions share no
// triangle, so a region's subtraction sequence -- the global one restricted
// to it -- and thus its FP64 values and decisions are the reference's.  One
// wave per region; the terms of a pop are handled 64 at a time (they touch
// distinct edges), pushes placed in the reference's order.
__device__ __forceinline__ uint32_t rg_root(uint32_t* parent, uint32_t x) {
    uint32_t p = parent[x];
    while (p != x) {
        const uint32_t gp = parent[p];
        if (gp != p) parent[x] = gp;
        x = p;
        p = gp;
    }
    return x;
}

__device__ __forceinline__ void rg_union(uint32_t* parent, uint32_t a, uint32_t b) {
    a = rg_root(parent, a);
    b = rg_root(parent, b);
    while (a != b) {
        if (a > b) {
            const uint32_t t = a;
            a = b;
            b = t;
        }
        const uint32_t old = atomicCAS(&parent[b], b, a);
        if (old == b) return;
        b = rg_root(parent, old);
        a = rg_root(parent, a);
    }
}

__global__ void __launch_bounds__(T) k_rg_init(uint32_t m, uint32_t* parent) {
    for (uint32_t e = blockIdx.x * T + threadIdx.x; e < m; e += gridDim.x * T) parent[e] = e;
}

// Unions the three edges of every triangle (once, from its smallest edge id).
__global__ void __launch_bounds__(T) k_rg_hook(GraphDev g, TermBuf tb, uint32_t* parent) {
    for (uint32_t e = blockIdx.x * T + threadIdx.x; e < g.m; e += gridDim.x * T) {
        const uint64_t k0 = tb.zoff[e], k1 = k0 + g.tcnt[e];
        for (uint64_t k = k0; k < k1; ++k) {
            const uint32_t a = tb.t[k].sa, b = tb.t[k].sb;
            if (e < a && e < b) {
                rg_union(parent, e, a);
                rg_union(parent, e, b);
            }
        }
    }
}

// parent[] := root, writing roots only (no path-halving races).
__global__ void __launch_bounds__(T) k_rg_compress(uint32_t m, uint32_t* parent) {
    for (uint32_t e = blockIdx.x * T + threadIdx.x; e < m; e += gridDim.x * T) {
        uint32_t r = e, p = parent[e];
        while (p != r) {
            r = p;
            p = parent[r];
        }
        parent[e] = r;
    }
}

__global__ void __launch_bounds__(T) k_rg_flag(const uint32_t* __restrict__ undec, uint32_t n,
                                               const uint32_t* __restrict__ parent,
                                               uint8_t* __restrict__ flag, bool all, uint32_t m) {
    if (all) {
        for (uint32_t e = blockIdx.x * T + threadIdx.x; e < m; e += gridDim.x * T) flag[e] = 1;
        return;
    }
    for (uint32_t i = blockIdx.x * T + threadIdx.x; i < n; i += gridDim.x * T)
        flag[parent[undec[i]]] = 1;
}

// Keys (root, edge id) of the edges of flagged regions; region edges are
// reset to their state before the clean.
__global__ void __launch_bounds__(T) k_rg_keys(GraphDev g, const uint32_t* __restrict__ parent,
                                               const uint8_t* __restrict__ flag, int Bi,
                                               uint64_t* __restrict__ keys, unsigned* nk,
                                               uint8_t* __restrict__ xact) {
    for (uint32_t e = blockIdx.x * T + threadIdx.x; e < g.m; e += gridDim.x * T) {
        const uint32_t r = parent[e];
        if (!flag[r]) continue;
        keys[wave_append(nk)] = ((uint64_t)r << Bi) | e;
        if (xact) xact[e] = 1;  // replayed: the reference's own values
        set_status(g, e, ST_ALIVE);
        g.sup[e] = g.sup0[e];
        g.touched[e] = 0;
        g.claim[e] = 0;
    }
}

// ---- regions of the undecided edges only: a breadth-first search over stored
// triangles from them (the visited set is the union of their regions), then
// the same union-find restricted to the visited edges.
__device__ __forceinline__ bool first_visit(uint8_t* vis, uint32_t e) {
    unsigned* w = reinterpret_cast<unsigned*>(vis + (e & ~3u));
    const unsigned bit = 1u << (8 * (e & 3u));
    return (atomicOr(w, bit) & bit) == 0u;
}

// Queue counters c: [0, 1) the round's frontier, [2] the tail.
__global__ void __launch_bounds__(T) k_bfs_seed(const uint32_t* __restrict__ undec, uint32_t n,
                                                uint8_t* vis, uint32_t* __restrict__ q,
                                                unsigned* c) {
    for (uint32_t i = blockIdx.x * T + threadIdx.x; i < n; i += gridDim.x * T) {
        const uint32_t e = undec[i];
        if (first_visit(vis, e)) q[wave_append(&c[2])] = e;
    }
}

__global__ void k_bfs_begin(unsigned* c) {
    if (threadIdx.x == 0) {
        c[0] = c[1];
        c[1] = c[2];
    }
}

// One round: every frontier edge's triangles; an edge seen for the first
// time joins the queue (the next frontier).
__global__ void __launch_bounds__(T) k_bfs_expand(GraphDev g, TermBuf tb, uint8_t* vis,
                                                  uint32_t* __restrict__ q, unsigned* c) {
    const uint32_t h0 = c[0], h1 = c[1];  // set by k_bfs_begin; only c[2] moves here
    for (uint32_t i = h0 + blockIdx.x * T + threadIdx.x; i < h1; i += gridDim.x * T) {
        const uint32_t e = q[i];
        const uint64_t k0 = tb.zoff[e], k1 = k0 + g.tcnt[e];
        for (uint64_t k = k0; k < k1; ++k) {
            const uint32_t a = tb.t[k].sa, b = tb.t[k].sb;
            if (first_visit(vis, a)) q[wave_append(&c[2])] = a;
            if (first_visit(vis, b)) q[wave_append(&c[2])] = b;
        }
    }
}

__global__ void __launch_bounds__(T) k_rgl_init(const uint32_t* __restrict__ q, uint32_t nv,
                                                uint32_t* parent) {
    for (uint32_t i = blockIdx.x * T + threadIdx.x; i < nv; i += gridDim.x * T) parent[q[i]] = q[i];
}

__global__ void __launch_bounds__(T) k_rgl_hook(GraphDev g, TermBuf tb,
                                                const uint32_t* __restrict__ q, uint32_t nv,
                                                uint32_t* parent) {
    for (uint32_t i = blockIdx.x * T + threadIdx.x; i < nv; i += gridDim.x * T) {
        const uint32_t e = q[i];
        const uint64_t k0 = tb.zoff[e], k1 = k0 + g.tcnt[e];
        for (uint64_t k = k0; k < k1; ++k) {
            const uint32_t a = tb.t[k].sa, b = tb.t[k].sb;
            if (e < a && e < b) {
                rg_union(parent, e, a);
                rg_union(parent, e, b);
            }
        }
    }
}

__global__ void __launch_bounds__(T) k_rgl_compress(const uint32_t* __restrict__ q, uint32_t nv,
                                                    uint32_t* parent) {
    for (uint32_t i = blockIdx.x * T + threadIdx.x; i < nv; i += gridDim.x * T) {
        const uint32_t e = q[i];
        uint32_t r = e, p = parent[e];
        while (p != r) {
            r = p;
            p = parent[r];
        }
        parent[e] = r;
    }
}

// Keys (root, edge id) of the visited edges, which are reset to their state
// before the clean.
__global__ void __launch_bounds__(T) k_rgl_keys(GraphDev g, const uint32_t* __restrict__ parent,
                                                const uint32_t* __restrict__ q, uint32_t nv,
                                                int Bi, uint64_t* __restrict__ keys,
                                                uint8_t* __restrict__ xact) {
    for (uint32_t i = blockIdx.x * T + threadIdx.x; i < nv; i += gridDim.x * T) {
        const uint32_t e = q[i];
        keys[i] = ((uint64_t)parent[e] << Bi) | e;
        if (xact) xact[e] = 1;  // replayed: the reference's own values
        set_status(g, e, ST_ALIVE);
        g.sup[e] = g.sup0[e];
        g.touched[e] = 0;
        g.claim[e] = 0;
    }
}

__global__ void __launch_bounds__(T) k_rg_layout(const uint64_t* __restrict__ keys, uint32_t nk,
                                                 int Bi, uint32_t* __restrict__ redge,
                                                 uint32_t* __restrict__ rflag) {
    const uint64_t mask = (1ull << Bi) - 1;
    for (uint32_t i = blockIdx.x * T + threadIdx.x; i < nk; i += gridDim.x * T) {
        redge[i] = (uint32_t)(keys[i] & mask);
        rflag[i] = (i == 0 || (keys[i] >> Bi) != (keys[i - 1] >> Bi)) ? 1u : 0u;
    }
}

__global__ void __launch_bounds__(T) k_rg_starts(const uint32_t* __restrict__ rflag,
                                                 const uint32_t* __restrict__ rpos, uint32_t nk,
                                                 uint32_t* __restrict__ rstart) {
    for (uint32_t i = blockIdx.x * T + threadIdx.x; i < nk; i += gridDim.x * T)
        if (rflag[i]) rstart[rpos[i]] = i;
    if (blockIdx.x == 0 && threadIdx.x == 0) rstart[rpos[nk]] = nk;
}

__device__ __forceinline__ uint64_t lanes_lt() {
    const int l = lane_id();
    return l ? (~0ull >> (64 - l)) : 0ull;
}

struct ReplayArgs {
    GraphDev g;
    TermBuf tb;
    double eps;
    const uint32_t* redge;   // region edges, ascending id within a region
    const uint32_t* rstart;  // nreg + 1 boundaries
    uint32_t nreg;
    uint64_t* stk;           // per region: (edge << 32) | first endpoint of the pair
    unsigned* next;          // work queue head
    // Explicit deletions since the initial clean (rtrex_delete_and_clean), in
    // call order: entry j deletes edge dl_e[j] given as the pair (dl_a[j], .),
    // in call dl_c[j].  Region r's entries, ascending j, are
    // dl_idx[dl_start[r] .. dl_end[r]) (no log: null).
    const uint32_t *dl_e = nullptr, *dl_a = nullptr, *dl_c = nullptr;
    const uint32_t *dl_idx = nullptr, *dl_start = nullptr, *dl_end = nullptr;
};

constexpr int RT = 64;  // one wave per workgroup
// One chunk of a popped edge's triangles (lane k of the chunk holds record
// q of its list, in = within the list; the statuses, supports and thresholds
// of the record's two edges loaded by the caller): the alive triangles lose
// their weight from both other edges, which are pushed in (lower, w) then
// (higher, w) order when they fall below their threshold.  Returns the
// number of pushes; *last is the entry on top of the stack afterwards when
// there were any.
__device__ __forceinline__ uint32_t replay_chunk(const ReplayArgs& A, uint64_t* stk, uint32_t top,
                                                 bool in, const Term& q, bool swap,
                                                 uint32_t lower, uint32_t higher, uint8_t sa,
                                                 uint8_t sb, double sup_a, double sup_b,
                                                 double ew_a, double ew_b, uint64_t* last) {
    const GraphDev& g = A.g;
    const int lane = lane_id();
    const uint32_t jl = swap ? q.sb : q.sa, jh = swap ? q.sa : q.sb;
    const uint8_t sl = swap ? sb : sa, sh = swap ? sa : sb;
    const double sup_l = swap ? sup_b : sup_a, sup_h = swap ? sup_a : sup_b;
    const double ew_l = swap ? ew_b : ew_a, ew_h = swap ? ew_a : ew_b;
    const bool alive = in && sl != ST_DEAD && sh != ST_DEAD;
    bool pl = false, ph = false;
    if (alive) {
        const double t = q.w;
        const double vl = sup_l - t, vh = sup_h - t;
        g.sup[jl] = vl;
        g.sup[jh] = vh;
        g.touched[jl] = 1;
        g.touched[jh] = 1;
        pl = sl != ST_QUEUED && vl < A.eps * ew_l;
        ph = sh != ST_QUEUED && vh < A.eps * ew_h;
    }
    const uint32_t np = (uint32_t)pl + (uint32_t)ph;
    uint32_t pre = np;  // exclusive prefix of pushes over lanes
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t o = __shfl_up(pre, d, 64);
        if (lane >= d) pre += o;
    }
    const uint32_t tot = __shfl(pre, 63, 64);
    pre -= np;
    // Edge statuses only: the slot copies are brought in line once the
    // clean ends (sync_slot_status).
    const uint64_t el = ((uint64_t)jl << 32) | lower, eh = ((uint64_t)jh << 32) | higher;
    if (pl) {
        stk[top + pre] = el;
        g.est[jl] = ST_QUEUED;
    }
    if (ph) {
        stk[top + pre + (uint32_t)pl] = eh;
        g.est[jh] = ST_QUEUED;
    }
    if (tot) {  // the new top of the stack, from the highest pushing lane
        const uint64_t pb = ballot64(np != 0);
        const int hl = 63 - __clzll((long long)pb);
        *last = __shfl(ph ? eh : el, hl, 64);
    }
    return tot;
}

// LIFO pops until the region's stack (top entries) is empty.
__device__ __forceinline__ void replay_pops(const ReplayArgs& A, uint64_t* stk, uint32_t top) {
    const GraphDev& g = A.g;
    const TermBuf& tb = A.tb;
    const int lane = lane_id();
    // The entry on top of the stack when the last pop pushed (registers;
    // otherwise it is read back).
    bool have = false;
    uint64_t next = 0;
    while (top > 0) {
        --top;
        const uint64_t ent = have ? next : stk[top];
        have = false;
        const uint32_t e = (uint32_t)(ent >> 32), a = (uint32_t)ent;
        // Three waves of loads per pop: what depends on e (and a); the
        // first chunk's records and e's other endpoint; the degrees and
        // both record edges' statuses, supports and thresholds.  Lanes
        // past the list take a blank record (term_or_blank), so that no
        // load waits on a branch.
        const uint32_t x = g.esrc[e], sle = g.eslot[e], tc = g.tcnt[e];
        const uint64_t k0 = tb.zoff[e];
        const uint32_t da = g.deg[a];
        const bool in0 = (uint32_t)lane < tc;
        const Term q0 = term_or_blank(tb.t, k0 + lane, in0);
        const uint32_t y = g.nb[sle];
        const uint32_t b = a == x ? y : x;
        const uint32_t db = g.deg[b];
        const uint8_t sa0 = g.est[q0.sa], sb0 = g.est[q0.sb];
        const double supa0 = g.sup[q0.sa], supb0 = g.sup[q0.sb];
        const double ewa0 = g.ew[q0.sa], ewb0 = g.ew[q0.sb];
        const uint32_t lower = da < db ? a : b, higher = da < db ? b : a;
        // stored terms: sa = {lo, z}, sb = {hi, z}, lo = d_x <= d_y ? x : y
        const uint32_t lo = (a == x ? da : db) <= (a == x ? db : da) ? x : y;
        const bool swap = lower != lo;
        uint64_t last = 0;
        uint32_t pushed = replay_chunk(A, stk, top, in0, q0, swap, lower, higher, sa0, sb0,
                                       supa0, supb0, ewa0, ewb0, &last);
        top += pushed;
        bool any = pushed != 0;
        const uint64_t k1 = k0 + tc;
        for (uint64_t p = k0 + 64; p < k1; p += 64) {  // longer lists
            const uint64_t k = p + lane;
            const bool in = k < k1;
            const Term q = term_or_blank(tb.t, k, in);
            const uint32_t pushed2 = replay_chunk(
                A, stk, top, in, q, swap, lower, higher, g.est[q.sa], g.est[q.sb],
                g.sup[q.sa], g.sup[q.sb], g.ew[q.sa], g.ew[q.sb], &last);
            top += pushed2;
            any = any || pushed2 != 0;
        }
        if (lane == 0) {
            g.est[e] = ST_DEAD;
            g.sup[e] = 0.0;
        }
        have = any;
        next = last;
        __syncthreads();
    }
}

__global__ void __launch_bounds__(RT) k_replay(ReplayArgs A) {
    const GraphDev& g = A.g;
    const int lane = lane_id();
    for (;;) {
        uint32_t r = 0;
        if (lane == 0) r = atomicAdd(A.next, 1u);
        r = __shfl(r, 0, 64);
        if (r >= A.nreg) break;
        const uint32_t b0 = A.rstart[r], b1 = A.rstart[r + 1];
        uint64_t* stk = A.stk + b0;
        uint32_t top = 0;
        // initial pushes: (u, v), u < v, in ascending edge id
        for (uint32_t p = b0; p < b1; p += 64) {
            const uint32_t i = p + lane;
            const uint32_t e = i < b1 ? A.redge[i] : 0;
            const bool push = i < b1 && g.sup0[e] < A.eps * g.ew[e];
            const uint64_t bal = ballot64(push);
            if (push) {
                stk[top + __popcll(bal & lanes_lt())] = ((uint64_t)e << 32) | g.esrc[e];
                g.est[e] = ST_QUEUED;
            }
            top += (uint32_t)__popcll(bal);
        }
        __syncthreads();
        replay_pops(A, stk, top);
        if (!A.dl_idx) continue;
        // Then every later deletion call's pairs of this region, pushed in
        // the caller's order (deleteAndClean pops the last one first), each
        // call cascaded to convergence before the next.
        const uint32_t l0 = A.dl_start[r], l1 = A.dl_end[r];
        uint32_t p = l0;
        while (p < l1) {
            const uint32_t call = A.dl_c[A.dl_idx[p]];
            top = 0;
            for (;;) {
                const uint32_t i = p + lane;
                const uint32_t j = i < l1 ? A.dl_idx[i] : 0;
                const bool mine = i < l1 && A.dl_c[j] == call;
                const uint32_t e = mine ? A.dl_e[j] : 0;
                const bool push = mine && g.est[e] == ST_ALIVE;
                const uint64_t bal = ballot64(push);
                if (push) {
                    stk[top + __popcll(bal & lanes_lt())] = ((uint64_t)e << 32) | A.dl_a[j];
                    g.est[e] = ST_QUEUED;
                }
                top += (uint32_t)__popcll(bal);
                const uint64_t mm = ballot64(mine);
                p += (uint32_t)__popcll(mm);
                if (mm != ~0ull) break;  // this call's entries end in this chunk
            }
            __syncthreads();
            replay_pops(A, stk, top);
        }
    }
}

__global__ void __launch_bounds__(T) k_count_alive(GraphDev g, unsigned long long* out) {
    unsigned long long c = 0;
    for (uint32_t e = blockIdx.x * T + threadIdx.x; e < g.m; e += gridDim.x * T)
        c += g.est[e] == ST_ALIVE;
    for (int d = 32; d >= 1; d >>= 1) c += __shfl_xor(c, d, 64);
    if (lane_id() == 0 && c) atomicAdd(out, c);
}

int bits_of(uint64_t n) {
    int b = 1;
    while ((1ull << b) < n) ++b;
    return b;
}

// ---- deletions since the initial clean, per replayed region
constexpr uint32_t NO_REGION = 0xffffffffu;

__global__ void __launch_bounds__(T) k_fill_u32(uint32_t* p, uint64_t n, uint32_t v) {
    for (uint64_t i = blockIdx.x * (uint64_t)T + threadIdx.x; i < n; i += (uint64_t)gridDim.x * T)
        p[i] = v;
}

// rid[e] = region of every region edge (position i of the sorted region
// edges belongs to region rpos[i] + rflag[i] - 1).
__global__ void __launch_bounds__(T) k_rg_rid(const uint32_t* __restrict__ redge,
                                              const uint32_t* __restrict__ rflag,
                                              const uint32_t* __restrict__ rpos, uint32_t nk,
                                              uint32_t* __restrict__ rid) {
    for (uint32_t i = blockIdx.x * T + threadIdx.x; i < nk; i += gridDim.x * T)
        rid[redge[i]] = rpos[i] + rflag[i] - 1u;
}

// Keys (region, entry) of the logged deletions inside replayed regions.
__global__ void __launch_bounds__(T) k_log_keys(const uint32_t* __restrict__ dl_e, uint32_t nl,
                                                const uint32_t* __restrict__ rid, int Bj,
                                                uint64_t* __restrict__ keys, unsigned* cnt) {
    for (uint32_t j = blockIdx.x * T + threadIdx.x; j < nl; j += gridDim.x * T) {
        const uint32_t r = rid[dl_e[j]];
        if (r != NO_REGION) keys[wave_append(cnt)] = ((uint64_t)r << Bj) | j;
    }
}

__global__ void __launch_bounds__(T) k_log_layout(const uint64_t* __restrict__ keys, uint32_t n,
                                                  int Bj, uint32_t* __restrict__ idx,
                                                  uint32_t* __restrict__ start,
                                                  uint32_t* __restrict__ end) {
    for (uint32_t i = blockIdx.x * T + threadIdx.x; i < n; i += gridDim.x * T) {
        const uint64_t k = keys[i];
        const uint32_t r = (uint32_t)(k >> Bj);
        idx[i] = (uint32_t)(k & ((1ull << Bj) - 1));
        if (i == 0 || (uint32_t)(keys[i - 1] >> Bj) != r) start[r] = i;
        if (i + 1 == n || (uint32_t)(keys[i + 1] >> Bj) != r) end[r] = i + 1;
    }
}

// Gives the replay the logged deletions of its regions, in call order.
rtrex_status replay_log(rtrex_graph h, Scratch& sc, const uint32_t* redge, const uint32_t* rflag,
                        const uint32_t* rpos, uint32_t nk, uint32_t nreg, ReplayArgs& A) {
    GraphDev& g = h->g;
    hipStream_t s = h->stream;
    const uint32_t nl = (uint32_t)h->dlog_e.size();
    uint32_t *rid, *de, *da, *dc, *idx, *st, *en;
    uint64_t *keys, *tmp;
    unsigned* cnt;
    RT_TRY(sc.get(&rid, g.m));
    RT_TRY(sc.get(&de, nl));
    RT_TRY(sc.get(&da, nl));
    RT_TRY(sc.get(&dc, nl));
    RT_TRY(sc.get(&keys, nl));
    RT_TRY(sc.get(&tmp, nl));
    RT_TRY(sc.get(&idx, nl));
    RT_TRY(sc.get(&st, nreg));
    RT_TRY(sc.get(&en, nreg));
    RT_TRY(sc.get(&cnt, 1));
    RT_HIP(hipMemcpyAsync(de, h->dlog_e.data(), nl * 4ull, hipMemcpyHostToDevice, s));
    RT_HIP(hipMemcpyAsync(da, h->dlog_a.data(), nl * 4ull, hipMemcpyHostToDevice, s));
    RT_HIP(hipMemcpyAsync(dc, h->dlog_c.data(), nl * 4ull, hipMemcpyHostToDevice, s));
    k_fill_u32<<<blocks_for(g.m, T, 4096), T, 0, s>>>(rid, g.m, NO_REGION);
    k_fill_u32<<<blocks_for(nreg, T, 4096), T, 0, s>>>(st, nreg, 0u);
    k_fill_u32<<<blocks_for(nreg, T, 4096), T, 0, s>>>(en, nreg, 0u);
    k_fill_u32<<<1, T, 0, s>>>(cnt, 1, 0u);
    k_rg_rid<<<blocks_for(nk, T, 8192), T, 0, s>>>(redge, rflag, rpos, nk, rid);
    const int Bj = bits_of((uint64_t)nl + 1), Br = bits_of((uint64_t)nreg + 1);
    k_log_keys<<<blocks_for(nl, T, 4096), T, 0, s>>>(de, nl, rid, Bj, keys, cnt);
    RT_HIP(hipGetLastError());
    uint32_t n = 0;
    RT_HIP(hipMemcpyAsync(&n, cnt, 4, hipMemcpyDeviceToHost, s));
    RT_HIP(hipStreamSynchronize(s));
    if (n) {
        RT_TRY(radix_sort_u64(keys, tmp, n, Bj + Br, s));
        k_log_layout<<<blocks_for(n, T, 4096), T, 0, s>>>(keys, n, Bj, idx, st, en);
        RT_HIP(hipGetLastError());
    }
    A.dl_e = de;
    A.dl_a = da;
    A.dl_c = dc;
    A.dl_idx = idx;
    A.dl_start = st;
    A.dl_end = en;
    return RTREX_OK;
}

// Re-runs the regions of the undecided edges exactly (see above).
rtrex_status exact_replay(rtrex_graph h, double eps, const uint32_t* undec, uint32_t cap,
                          unsigned nund) {
    GraphDev& g = h->g;
    hipStream_t s = h->stream;
    if (!nund || !h->zt) return RTREX_OK;
    Scratch sc;
    uint32_t *parent, *redge, *rflag, *rpos, *rstart;
    uint64_t *keys, *tmp, *stk;
    uint8_t* flag;
    unsigned long long* cnt;
    RT_TRY(sc.get(&parent, g.m));
    RT_TRY(sc.get(&flag, (g.m + 3) & ~3ull));
    RT_TRY(sc.get(&cnt, 2));
    const unsigned be = blocks_for(g.m, T, 8192);
    const bool all = nund > cap;
    const int Bi = bits_of((uint64_t)g.m + 1);
    RT_TRY(sc.get(&keys, g.m));
    RT_HIP(hipMemsetAsync(cnt, 0, 16, s));  // [0] region edges (all path), [1] alive edges
    uint32_t nk = 0;
    if (all) {  // every region: a union-find over all stored triangles
        k_rg_init<<<be, T, 0, s>>>(g.m, parent);
        k_rg_hook<<<be, T, 0, s>>>(g, h->terms(), parent);
        k_rg_compress<<<be, T, 0, s>>>(g.m, parent);
        RT_HIP(hipMemsetAsync(flag, 0, g.m, s));
        k_rg_flag<<<be, T, 0, s>>>(undec, nund, parent, flag, true, g.m);
        k_rg_keys<<<be, T, 0, s>>>(g, parent, flag, Bi, keys, (unsigned*)cnt, h->xact);
        RT_HIP(hipGetLastError());
        RT_HIP(hipMemcpyAsync(&nk, cnt, 4, hipMemcpyDeviceToHost, s));
        RT_HIP(hipStreamSynchronize(s));
    } else {
        // The regions of the undecided edges: breadth-first over stored
        // triangles, eight rounds per host round trip (rounds past the end
        // find an empty frontier), then the union-find over the visited
        // edges only.
        uint32_t* q;
        unsigned* c;
        RT_TRY(sc.get(&q, g.m));
        RT_TRY(sc.get(&c, 4));
        RT_HIP(hipMemsetAsync(flag, 0, (g.m + 3) & ~3ull, s));
        RT_HIP(hipMemsetAsync(c, 0, 16, s));
        k_bfs_seed<<<blocks_for(nund, T, 4096), T, 0, s>>>(undec, nund, flag, q, c);
        const TermBuf tb = h->terms();
        unsigned hc[3] = {0, 0, 0};
        for (;;) {
            for (int r = 0; r < 8; ++r) {
                k_bfs_begin<<<1, WAVE, 0, s>>>(c);
                k_bfs_expand<<<1024, T, 0, s>>>(g, tb, flag, q, c);
            }
            RT_HIP(hipGetLastError());
            RT_HIP(hipMemcpyAsync(hc, c, 12, hipMemcpyDeviceToHost, s));
            RT_HIP(hipStreamSynchronize(s));
            if (hc[1] == hc[2]) break;  // the last round added no edge
        }
        nk = hc[2];
        const unsigned bv = blocks_for(nk, T, 8192);
        if (nk) {
            k_rgl_init<<<bv, T, 0, s>>>(q, nk, parent);
            k_rgl_hook<<<bv, T, 0, s>>>(g, tb, q, nk, parent);
            k_rgl_compress<<<bv, T, 0, s>>>(q, nk, parent);
            k_rgl_keys<<<bv, T, 0, s>>>(g, parent, q, nk, Bi, keys, h->xact);
            RT_HIP(hipGetLastError());
        }
    }
    if (nk) {
        RT_TRY(sc.get(&tmp, nk));
        RT_TRY(radix_sort_u64(keys, tmp, nk, 2 * Bi, s));
        RT_TRY(sc.get(&redge, nk));
        RT_TRY(sc.get(&rflag, nk));
        RT_TRY(sc.get(&rpos, (uint64_t)nk + 1));
        const unsigned bk = blocks_for(nk, T, 8192);
        k_rg_layout<<<bk, T, 0, s>>>(keys, nk, Bi, redge, rflag);
        RT_TRY(exclusive_scan_u32_32(rflag, rpos, nk, s));
        uint32_t nreg = 0;
        RT_HIP(hipMemcpyAsync(&nreg, rpos + nk, 4, hipMemcpyDeviceToHost, s));
        RT_HIP(hipStreamSynchronize(s));
        RT_TRY(sc.get(&rstart, (uint64_t)nreg + 1));
        k_rg_starts<<<bk, T, 0, s>>>(rflag, rpos, nk, rstart);
        RT_TRY(sc.get(&stk, nk));
        RT_HIP(hipMemsetAsync(&h->ctr->work_next, 0, 4, s));
        ReplayArgs A{g, h->terms(), eps, redge, rstart, nreg, stk, &h->ctr->work_next};
        if (!h->dlog_e.empty()) RT_TRY(replay_log(h, sc, redge, rflag, rpos, nk, nreg, A));
        k_replay<<<std::min<uint32_t>(nreg, 8192), RT, 0, s>>>(A);
        RT_HIP(hipGetLastError());
    }
    // Every undecided decision now follows the reference's order.
    k_count_alive<<<be, T, 0, s>>>(g, cnt + 1);
    unsigned long long alive = 0;
    RT_HIP(hipMemcpyAsync(&alive, cnt + 1, 8, hipMemcpyDeviceToHost, s));
    RT_HIP(hipStreamSynchronize(s));
    const unsigned long long removed = g.m - alive, replayed = nk, zero = 0;
    RT_HIP(hipMemcpyAsync(&h->ctr->edges_removed, &removed, 8, hipMemcpyHostToDevice, s));
    RT_HIP(hipMemcpyAsync(&h->ctr->replayed_edges, &replayed, 8, hipMemcpyHostToDevice, s));
    RT_HIP(hipMemcpyAsync(&h->ctr->ambiguous_clean, &zero, 8, hipMemcpyHostToDevice, s));
    RT_HIP(hipStreamSynchronize(s));
    h->n_alive = alive;
    return RTREX_OK;
}

// Blocks for `items` work items at G lanes each (capped; kernels stride).
// Near-list checks: grid-stride over the near list (its length is on the
// device), NEAR_G lanes per edge with stored triangle lists, one otherwise.
constexpr unsigned NEAR_BLOCKS = 256;
constexpr int NEAR_G = 16;

unsigned grid_items(uint32_t items, int G) {
    const uint64_t waves = ((uint64_t)items * G + WAVE - 1) / WAVE;
    return (unsigned)std::min<uint64_t>(std::max<uint64_t>(1, (waves * WAVE + T - 1) / T), 8192);
}

// Work lists of a clean, all of capacity 4 x m (four buckets): the first
// frontier split by bucket (bk), and the two alive-bucket lists pull rounds
// alternate between (ak[0] walked, ak[1] written).  Their counts, the
// frontier's and the rounds' counters live in one block (cnt, CNT_WORDS
// words) cleared by one memset before the clean starts.
enum : uint32_t {
    C_FRONT = 0,   // first frontier length (seeding)
    C_BK = 4,      // 4 bucket counts of bk
    C_AK0 = 8,     // 4 bucket counts of ak[0] as passed in
    C_AK1 = 12,    // ... of ak[1]
    C_ROUND = 16,  // [2p, 2p + 1]: next-frontier and near counts, round parity p
    C_INC = 20,    // live-incidence total (64 bits)
    C_FBK = 24,    // 4 bucket counts of a later frontier's split
    CNT_WORDS = 28,
};
struct CleanLists {
    uint32_t* bk = nullptr;
    uint32_t* ak[2] = {nullptr, nullptr};
    uint32_t* cnt = nullptr;
    unsigned* akc[2] = {nullptr, nullptr};
};

// The lists are allocated with stored triangle lists only (pull rounds and
// the first frontier's split need them); bk when seeding writes it.
rtrex_status clean_lists(rtrex_graph h, Scratch& sc, bool with_bk, CleanLists& L) {
    const uint64_t m = h->g.m;
    if (h->zt) {
        if (with_bk) RT_TRY(sc.get(&L.bk, 4 * m));
        RT_TRY(sc.get(&L.ak[0], 4 * m));
        RT_TRY(sc.get(&L.ak[1], 4 * m));
    }
    RT_TRY(sc.get(&L.cnt, CNT_WORDS));
    L.akc[0] = L.cnt + C_AK0;
    L.akc[1] = L.cnt + C_AK1;
    RT_HIP(hipMemsetAsync(L.cnt, 0, CNT_WORDS * sizeof(uint32_t), h->stream));
    return RTREX_OK;
}

// Frontier rounds to convergence.  first_bk: the first frontier is also
// split by work bucket in L.bk (seeding).  With stored triangle lists, the
// alive buckets a pull round walks are L.ak[0], valid on entry when ak_valid
// (seeding wrote them), carried from one pull round to the next by
// k_collect_claims, and rebuilt by a scan of every edge after a push round.
// Per round: one launch per kernel, the counters the kernels append to
// cleared by the round before (k_end_round), and one host round trip (the
// next frontier's length, read after the round ends); every launch in between
// takes its item count from device memory and is sized from a host-side upper
// bound.  L.cnt must be zero on entry but for the seeding's counts.
rtrex_status peel(rtrex_graph h, double eps, uint32_t* F, uint32_t* Fn,
                  uint32_t nF, CleanLists L, bool ak_valid, bool first_bk,
                  uint32_t* undec = nullptr, uint32_t undec_cap = 0) {
    GraphDev& g = h->g;
    hipStream_t s = h->stream;
    Scratch sc;
    uint32_t* near;
    RT_TRY(sc.get(&near, g.m));
    if (!h->nearst) {
        RT_TRY(dalloc(h, &h->nearst, g.m));
        RT_HIP(hipMemsetAsync(h->nearst, 0, (uint64_t)g.m * 4, s));
    }
    RT_TRY(box_open(h));
    uint32_t* rc = L.cnt + C_ROUND;
    unsigned long long* inc = reinterpret_cast<unsigned long long*>(L.cnt + C_INC);
    PeelArgs A{};
    A.g = g;
    A.eps = eps;
    A.near = near;
    A.nearst = h->nearst;
    A.ctr = h->ctr;
    A.tb = h->terms();
    A.undec = undec;
    A.undec_cap = undec_cap;
    uint32_t* fbk;  // a later frontier split by bucket (4 x m ids)
    RT_TRY(sc.get(&fbk, 4ull * g.m));
    unsigned* bcur = L.cnt + C_FBK;
    for (uint32_t round = 0; nF; ++round) {
        unsigned* rcur = rc + 2 * (round & 1);
        A.nFn = rcur;
        A.nNear = rcur + 1;
        A.Fn = Fn;
        A.stamp = ++h->round_stamp;
        const bool terms = A.tb.t != nullptr;
        const bool pull = terms && (uint64_t)nF > h->n_alive;
        if (pull) {
            first_bk = false;  // pull rounds walk the alive edges instead
            A.F = F;
            A.nF = nF;
            A.nFd = nullptr;
            unsigned* acur = L.akc[0];
            if (!ak_valid)  // its counts were cleared by the round before (k_end_round)
                k_alive_buckets<<<blocks_for((g.m + SEED_K - 1) / SEED_K, T, 4096), T, 0, s>>>(
                    g, L.ak[0], acur);
            if (h->n_alive) {
                // one launch over the four buckets, each sized for every
                // alive edge (the bucket counts stay on the device)
                BucketBlocks R{};
                for (int b = 0; b < 4; ++b)
                    R.bb[b + 1] = R.bb[b] + grid_items((uint32_t)h->n_alive, 1 << (2 * b));
                k_peel_pull<<<R.bb[4], T, 0, s>>>(A, L.ak[0], acur, R);
                k_collect_claims<<<blocks_for((h->n_alive + SEED_K - 1) / SEED_K, T, 4096), T, 0,
                                   s>>>(g, L.ak[0], acur, Fn, A.nFn, L.ak[1], L.akc[1]);
            }
        } else {
            // Work buckets of the frontier (1, 4, 16 or 64 lanes per edge):
            // split while it was seeded (first round), split here when large,
            // otherwise one list at 16 lanes.
            const uint32_t* lists[4] = {nullptr, nullptr, nullptr, nullptr};
            const unsigned* cnts[4] = {nullptr, nullptr, nullptr, nullptr};
            if (first_bk) {
                for (int b = 0; b < 4; ++b) {
                    lists[b] = L.bk + (uint64_t)b * g.m;
                    cnts[b] = L.cnt + C_BK + b;
                }
                first_bk = false;
            } else if (nF > 4096) {
                RT_HIP(hipMemsetAsync(bcur, 0, 4 * sizeof(unsigned), s));
                k_frontier_buckets<<<blocks_for(nF, T, 8192), T, 0, s>>>(g, A.tb.t != nullptr,
                                                                         F, nF, fbk, bcur);
                for (int b = 0; b < 4; ++b) {
                    lists[b] = fbk + (uint64_t)b * nF;
                    cnts[b] = bcur + b;
                }
            } else {
                lists[2] = F;
            }
            for (int b = 0; b < 4; ++b) {
                if (!lists[b]) continue;
                const int G = b == 0 ? 1 : b == 1 ? 4 : b == 2 ? 16 : 64;
                A.F = lists[b];
                A.nF = nF;
                A.nFd = cnts[b];
                const unsigned grid = grid_items(nF, G);
                if (A.tb.t) {
                    if (G == 1) k_peel_terms<1><<<grid, T, 0, s>>>(A);
                    else if (G == 4) k_peel_terms<4><<<grid, T, 0, s>>>(A);
                    else if (G == 16) k_peel_terms<16><<<grid, T, 0, s>>>(A);
                    else k_peel_terms<64><<<grid, T, 0, s>>>(A);
                } else {
                    if (G == 1) k_peel<1><<<grid, T, 0, s>>>(A);
                    else if (G == 4) k_peel<4><<<grid, T, 0, s>>>(A);
                    else if (G == 16) k_peel<16><<<grid, T, 0, s>>>(A);
                    else k_peel<64><<<grid, T, 0, s>>>(A);
                }
            }
            A.F = F;
            A.nF = nF;
            A.nFd = nullptr;
        }
        // Near-threshold edges (a few hundred at most in practice): certified
        // from their terms; count from device memory.
        if (terms) k_near_check_terms<NEAR_G><<<NEAR_BLOCKS, T, 0, s>>>(A);
        else k_near_check<<<NEAR_BLOCKS, T, 0, s>>>(A);
        // The frontier dies, the claimed edges are queued.  Both lists are
        // bounded by the edges alive at the start of the round.  The next
        // round's counters are cleared here, and so are the counts of
        // L.ak[0]: after a pull round it is the list just walked (the next
        // round's output list, after the swap); after a push round it is
        // stale, and a later pull round rebuilds it from a zero count (the
        // rebuild appends; on top of the old counts every alive edge would be
        // listed twice and a claimed edge queued twice).
        k_end_round<<<blocks_for(std::max<uint64_t>(nF, h->n_alive), T, 4096), T, 0, s>>>(
            g, F, nF, Fn, A.nFn, h->ctr, terms, rc + 2 * ((round + 1) & 1),
            terms ? L.akc[0] : nullptr, h->box_dev, ++h->box_seq);
        if (pull) {  // ak[1] (now ak[0]): the next pull round's work list
            std::swap(L.ak[0], L.ak[1]);
            std::swap(L.akc[0], L.akc[1]);
        }
        ak_valid = pull;
        RT_HIP(hipGetLastError());
        uint32_t nFn = 0;
        RT_TRY(box_wait(h, 0, h->box_seq, &nFn));
        h->audit.clean_rounds++;
        h->n_alive -= nFn;
        std::swap(F, Fn);
        nF = nFn;
    }
    // The safety-net certification, the exact-support flags and the live
    // triangles (the peel kernels do not count them) in one pass.
    if (undec && !h->xact) RT_TRY(dalloc(h, &h->xact, g.m));
    k_post_peel<<<blocks_for(g.m, T, 4096), T, 0, s>>>(g, A.tb, eps, h->ctr, undec, undec_cap,
                                                       undec ? h->xact : nullptr,
                                                       A.tb.t ? inc : nullptr);
    if (A.tb.t) k_tri_removed<<<1, WAVE, 0, s>>>(inc, h->ctr);
    RT_HIP(hipGetLastError());
    if (undec) {  // the only wait after the last round: were any decisions open?
        unsigned nund = 0;
        k_publish<<<1, WAVE, 0, s>>>(h->box_dev, ++h->box_seq, &h->ctr->n_undecided);
        RT_HIP(hipGetLastError());
        RT_TRY(box_wait(h, 0, h->box_seq, &nund));
        if (nund && h->zt) {
            RT_TRY(exact_replay(h, eps, undec, undec_cap, nund));
            // live triangles again, after the replayed regions
            RT_HIP(hipMemsetAsync(inc, 0, 8, s));
            k_live_incidences<<<blocks_for(g.m, T, 4096), T, 0, s>>>(g, A.tb, inc);
            k_tri_removed<<<1, WAVE, 0, s>>>(inc, h->ctr);
            RT_HIP(hipGetLastError());
        }
    }
    h->pv_full = false;  // vertex supports are pulled again on demand
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
    const bool terms = h->zt != nullptr;
    RT_HIP(hipEventRecord(h->ev[5], s));
    CleanLists L;
    RT_TRY(clean_lists(h, sc, terms, L));
    unsigned* cnt = L.cnt + C_FRONT;
    k_seed_frontier<<<blocks_for((g.m + SEED_K - 1) / SEED_K, SEED_T, 4096), SEED_T, 0, s>>>(
        g, eps, F, cnt, L.bk, L.cnt + C_BK, terms ? L.ak[0] : nullptr, L.akc[0]);
    RT_TRY(box_open(h));
    k_publish<<<1, WAVE, 0, s>>>(h->box_dev, ++h->box_seq, cnt);
    RT_HIP(hipGetLastError());
    uint32_t nF = 0;
    RT_TRY(box_wait(h, 0, h->box_seq, &nF));
    h->n_alive = g.m - nF;
    uint32_t* undec = nullptr;
    RT_TRY(sc.get(&undec, g.m));
    // Regions an earlier extraction could not decide are re-run as well.
    if (terms && (h->force_all || h->replay_all || !h->force_replay.empty())) {
        const bool all = h->force_all || h->replay_all;
        const uint32_t nf = all ? g.m + 1 : (uint32_t)h->force_replay.size();
        if (!all)
            RT_HIP(hipMemcpyAsync(undec, h->force_replay.data(), nf * 4ull, hipMemcpyHostToDevice,
                                  s));
        RT_HIP(hipMemcpyAsync(&h->ctr->n_undecided, &nf, 4, hipMemcpyHostToDevice, s));
        RT_HIP(hipStreamSynchronize(s));
    }
    h->dlog_e.clear();  // a new clean: no deletion since
    h->dlog_a.clear();
    h->dlog_b.clear();
    h->dlog_c.clear();
    h->dlog_calls = 0;
    RT_TRY(peel(h, eps, F, Fn, nF, L, terms, terms, terms ? undec : nullptr, g.m));
    if (!h->surv) RT_TRY(dalloc(h, &h->surv, (g.m + 31) / 32));
    k_survivor_bits<<<blocks_for(std::max<uint64_t>(terms ? g.nnz / 4 : 0, (g.m + 31) / 32), T,
                                 8192),
                      T, 0, s>>>(g, h->surv, terms);
    RT_HIP(hipGetLastError());
    RT_HIP(hipEventRecord(h->ev[6], s));
    h->tm_clean_pending = true;  // read later (resolve_timings), no wait here
    h->eps = eps;
    h->phase = PH_CLEANED;
    return RTREX_OK;
}

rtrex_status run_delete_and_clean(rtrex_graph h, const int64_t* pairs,
                                  int64_t k, double eps) {
    if (h->phase < PH_CLEANED) return fail(RTREX_ESTATE, "initial clean has not run");
    if (h->phase >= PH_EXTRACTED)
        return fail(RTREX_ESTATE, "graph already extracted; run count again");
    if (eps != h->eps) return fail(RTREX_EINVAL, "delete_and_clean needs the eps of the clean");
    if (k < 0 || (k && !pairs)) return fail(RTREX_EINVAL, "bad pair list");
    if (k == 0) return RTREX_OK;
    GraphDev& g = h->g;
    hipStream_t s = h->stream;
    if ((uint64_t)k > g.m) return fail(RTREX_EINVAL, "more pairs than edges");
    std::vector<uint32_t> pr(2 * k);
    for (int64_t i = 0; i < 2 * k; ++i) {
        if (pairs[i] < 0 || pairs[i] >= (int64_t)g.n)
            return fail(RTREX_EINVAL, "pair vertex out of range");
        pr[i] = (uint32_t)pairs[i];
    }
    Scratch sc;
    uint32_t *F, *Fn, *dpr, *eid;
    const bool terms = h->zt != nullptr;
    RT_TRY(sc.get(&F, g.m));
    RT_TRY(sc.get(&Fn, g.m));
    RT_TRY(sc.get(&dpr, 2 * k));
    RT_TRY(sc.get(&eid, k));
    // Every pair checked before anything changes: an edge of the graph,
    // alive, listed once (the state is untouched when one is not).
    RT_HIP(hipMemcpyAsync(dpr, pr.data(), pr.size() * 4, hipMemcpyHostToDevice, s));
    k_pair_edges<<<blocks_for(k, T, 4096), T, 0, s>>>(g, dpr, (uint32_t)k, eid);
    RT_HIP(hipGetLastError());
    std::vector<uint32_t> he(k);
    RT_HIP(hipMemcpyAsync(he.data(), eid, k * 4ull, hipMemcpyDeviceToHost, s));
    RT_HIP(hipStreamSynchronize(s));
    for (int64_t q = 0; q < k; ++q) {
        if (he[q] == PAIR_NOT_EDGE)
            return fail(RTREX_EINVAL, "pair " + std::to_string(q) + " is not an edge of the graph");
        if (he[q] & PAIR_NOT_ALIVE)
            return fail(RTREX_EINVAL, "pair " + std::to_string(q) + " is not an alive edge");
    }
    std::vector<uint32_t> sorted(he);
    std::sort(sorted.begin(), sorted.end());
    if (std::adjacent_find(sorted.begin(), sorted.end()) != sorted.end())
        return fail(RTREX_EINVAL, "an edge is listed twice");
    CleanLists L;
    RT_TRY(clean_lists(h, sc, false, L));
    k_seed_edges<<<blocks_for(k, T, 4096), T, 0, s>>>(g, eid, (uint32_t)k, F, terms);
    RT_HIP(hipGetLastError());
    for (int64_t q = 0; q < k; ++q) {
        h->dlog_e.push_back(he[q]);
        h->dlog_a.push_back(pr[2 * q]);
        h->dlog_b.push_back(pr[2 * q + 1]);
        h->dlog_c.push_back(h->dlog_calls);
    }
    h->dlog_calls++;
    if (h->xact) RT_HIP(hipMemsetAsync(h->xact, 0, g.m, s));  // supports no longer the reference's
    h->n_alive -= (uint64_t)k;
    uint32_t* undec = nullptr;
    if (terms) {
        RT_TRY(sc.get(&undec, g.m));
        RT_HIP(hipMemsetAsync(&h->ctr->n_undecided, 0, 4, s));
        if (h->replay_all) {  // every deleted edge's region re-run
            RT_HIP(hipMemcpyAsync(undec, eid, k * 4ull, hipMemcpyDeviceToDevice, s));
            const uint32_t nf = (uint32_t)k;
            RT_HIP(hipMemcpyAsync(&h->ctr->n_undecided, &nf, 4, hipMemcpyHostToDevice, s));
            RT_HIP(hipStreamSynchronize(s));
        }
    }
    RT_TRY(peel(h, eps, F, Fn, (uint32_t)k, L, false, false, undec, g.m));
    if (terms) k_slot_status<<<blocks_for(g.nnz / 4 + 1, T, 8192), T, 0, s>>>(g);
    RT_HIP(hipGetLastError());
    RT_HIP(hipStreamSynchronize(s));
    return RTREX_OK;
}

}  // namespace rtrex
