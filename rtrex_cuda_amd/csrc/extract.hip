// Greedy extraction (seed loop) on the device.
//
// Reference: disjointExtract (Decomposition.h:538-940), populateStats
// (:473-491); CUDA port: disjointExtractCUDA (RTRex_CUDA/main.cu:130-338),
// gpuExtractInner / extractInnerKernel / gpuMarkClusterEdges
// (RTRex_CUDA/gpu_extract.cu:9-317).
//
// Independence.  A seed's cluster is itself, its neighbours over alive edges
// and two-hop vertices reached through alive triangles; deleting the cluster's
// edges and cascading only touches alive edges reachable through alive
// triangles.  So every connected component of the alive graph (after the
// initial clean) evolves on its own, its seeds taken in the global
// (degree, id) order restricted to it.  Components are extracted in parallel,
// one wave each, pulled from a work queue in one persistent launch; there is
// no host round trip per seed (the CUDA port makes ~10 launches and O(n + m)
// PCIe traffic per seed).
//
// Per seed u (all 64 lanes of the wave):
//   1. alive neighbours A of u join the cluster (slot order);
//   2. the reference's inner loop (:728-805) in its exact order: pairs (v, w)
//      with w at or after v in A, both passing the degree filter d > d_u/eps;
//      internal += 1/(d_u d_v d_w); for x in alive N(lower) with (higher, x)
//      alive and x not in {lower, higher, u}: if x is a neighbour of u the
//      weight joins `internal` when x > v, w, else it accumulates into the
//      two-hop weight of x.  Lanes probe 64 ids of N(lower) at a time; the
//      `internal` fold and each candidate's fold keep the reference's order
//      (lane order inside a chunk, chunks and pairs in sequence);
//   3. candidates ascending by id, potential = fold of their weights,
//      std::sort-identical ordering by weight (introsort.hpp), greedy prefix
//      with the strict '>' ratio rule (:847-864);
//   4. every alive edge at a cluster vertex is queued and the deletion
//      cascades (deleteAndClean semantics, LIFO) inside the wave.
// Per-vertex supports read in step 3 went through the initial clean's atomic
// subtractions, so they may differ from the reference's by a few ulps; each
// ratio decision is audited against that bound (ambiguous_ratio).
#include <algorithm>
#include <mutex>
#include <unordered_set>

#include "audit.hpp"
#include "dev.hpp"
#include "internal.hpp"
#include "introsort.hpp"
#include "uset.hpp"

namespace rtrex {
namespace {

constexpr int T = 256;
constexpr int XT = 64;  // one wave per workgroup
constexpr uint32_t NONE = 0xffffffffu;
constexpr uint32_t LDS_CAND = 1024;
constexpr uint32_t XSTAGE = 2048;  // longest N(hi) the inner loop stages in LDS
// Clusters of up to US_LDS members emulate the set's order in LDS, when the
// set has at most US_LDS_BKT buckets at that size.
constexpr uint32_t US_LDS = 128, US_LDS_BKT = 512;
constexpr int CWG = 8;             // waves per handed-over cascade (k_cascade_wg)

using isort::Rec;

// ---------------------------------------------------------------- setup
__global__ void __launch_bounds__(T) k_seed_keys(const GraphDev g,
                                                 const uint32_t* __restrict__ vseg,
                                                 int Bd, int Bi,
                                                 uint64_t* __restrict__ keys) {
    for (uint32_t v = blockIdx.x * T + threadIdx.x; v < g.n; v += gridDim.x * T) {
        uint64_t sg = vseg ? vseg[v] : 0;
        keys[v] = (sg << (Bd + Bi)) | ((uint64_t)g.deg[v] << Bi) | v;
    }
}

__global__ void __launch_bounds__(T) k_ranks(const uint64_t* __restrict__ keys,
                                             uint32_t n, int Bi,
                                             uint32_t* __restrict__ order,
                                             uint32_t* __restrict__ rank) {
    const uint64_t mask = (1ull << Bi) - 1;
    for (uint32_t r = blockIdx.x * T + threadIdx.x; r < n; r += gridDim.x * T) {
        uint32_t v = (uint32_t)(keys[r] & mask);
        order[r] = v;
        rank[v] = r;
    }
}

__device__ __forceinline__ uint32_t uf_find(uint32_t* parent, uint32_t x) {
    uint32_t p = parent[x];
    while (p != x) {
        uint32_t gp = parent[p];
        if (gp != p) parent[x] = gp;  // path halving
        x = p;
        p = gp;
    }
    return x;
}

// Also clears what the extraction's later kernels count or stamp into (one
// launch instead of a memset each): the neighbour and candidate stamps mk, cm
// (n each), the largest-component slot big, and the counters nk (vertices in
// components), n_undecided, work_next and error.
__global__ void __launch_bounds__(T) k_uf_init(uint32_t n, uint32_t* parent,
                                               uint32_t* csize, uint32_t* cseed,
                                               uint32_t* __restrict__ mk,
                                               uint32_t* __restrict__ cm, uint32_t* big,
                                               unsigned* nk, Counters* ctr) {
    for (uint32_t v = blockIdx.x * T + threadIdx.x; v < n; v += gridDim.x * T) {
        parent[v] = v;
        csize[v] = 0;
        cseed[v] = NONE;
        mk[v] = 0;
        cm[v] = 0;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        *big = 0;
        *nk = 0;
        ctr->n_undecided = 0;
        ctr->work_next = 0;
        ctr->error = 0;
    }
}

// Lock-free union of the endpoints of every alive edge (smaller root wins).
__global__ void __launch_bounds__(T) k_uf_hook(const GraphDev g, uint32_t* parent) {
    for (uint32_t e = blockIdx.x * T + threadIdx.x; e < g.m; e += gridDim.x * T) {
        const uint32_t sl = g.eslot[e];
        if (g.st[sl] != ST_ALIVE) continue;
        uint32_t a = uf_find(parent, g.esrc[e]);
        uint32_t b = uf_find(parent, g.nb[sl]);
        while (a != b) {
            if (a > b) {
                uint32_t t = a;
                a = b;
                b = t;
            }
            uint32_t old = atomicCAS(&parent[b], b, a);
            if (old == b) break;
            b = uf_find(parent, old);
            a = uf_find(parent, a);
        }
    }
}

// Every parent[] entry set to its root.  The walk up does not shorten paths:
// a halving write (parent[x] = grandparent) racing with another thread's
// parent[x] = root could leave x pointing at a non-root, and the component
// keys below read parent[] as the root.  Only roots are written here.
__global__ void __launch_bounds__(T) k_uf_compress(uint32_t n, uint32_t* parent,
                                                   uint32_t* csize) {
    for (uint32_t v = blockIdx.x * T + threadIdx.x; v < n; v += gridDim.x * T) {
        uint32_t r = v, p = parent[v];
        while (p != r) {
            r = p;
            p = parent[r];
        }
        parent[v] = r;
        atomicAdd(&csize[r], 1u);
    }
}

// Vertices of size-1 components have no alive edge: each is a singleton
// cluster seeded by itself.  The rest are keyed (root, seed rank).
__global__ void __launch_bounds__(T) k_comp_keys(uint32_t n,
                                                 const uint32_t* __restrict__ parent,
                                                 const uint32_t* __restrict__ csize,
                                                 const uint32_t* __restrict__ rank,
                                                 uint32_t* __restrict__ cseed,
                                                 uint32_t* __restrict__ mrank,
                                                 uint64_t* __restrict__ keys,
                                                 unsigned* nkeys) {
    for (uint32_t v = blockIdx.x * T + threadIdx.x; v < n; v += gridDim.x * T) {
        uint32_t r = parent[v];
        if (csize[r] < 2) {
            cseed[v] = v;
            mrank[v] = 0;
        } else {
            keys[wave_append(nkeys)] = ((uint64_t)r << 32) | rank[v];
        }
    }
}

__global__ void __launch_bounds__(T) k_comp_layout(const uint64_t* __restrict__ keys,
                                                   uint32_t nk,
                                                   const uint32_t* __restrict__ order,
                                                   const GraphDev g,
                                                   uint32_t* __restrict__ cv,
                                                   uint32_t* __restrict__ vdeg,
                                                   uint32_t* __restrict__ flag) {
    for (uint32_t i = blockIdx.x * T + threadIdx.x; i < nk; i += gridDim.x * T) {
        uint32_t v = order[(uint32_t)(keys[i] & 0xffffffffu)];
        cv[i] = v;
        vdeg[i] = g.deg[v];
        flag[i] = (i == 0 || (keys[i] >> 32) != (keys[i - 1] >> 32)) ? 1u : 0u;
    }
}

// Scalars read back at the end of an extraction, through the host mailbox
// (slots 1-5): component count, largest component, undecided cascade
// decisions, error flag, clusters (absent sources read as 0).
__global__ void k_final_scalars(const uint32_t* ncomp, const uint32_t* big,
                                const unsigned* nund, const unsigned* err,
                                const uint32_t* total, uint64_t* box, uint32_t seq) {
    if (threadIdx.x == 0) {
        publish(box + 1, seq, ncomp ? *ncomp : 0u);
        publish(box + 2, seq, big ? *big : 0u);
        publish(box + 3, seq, nund ? *nund : 0u);
        publish(box + 4, seq, *err);
        publish(box + 5, seq, *total);
    }
}

// Per component c < ncomp: its unordered_set bucket count bcnt[size] (0 past
// ncomp, up to bound), and the largest size.
__global__ void __launch_bounds__(T) k_comp_hsize(const uint32_t* __restrict__ cstart,
                                                  const uint32_t* __restrict__ ncomp_d,
                                                  uint32_t bound,
                                                  const uint32_t* __restrict__ bcnt,
                                                  uint32_t* __restrict__ hsz, uint32_t* big) {
    const uint32_t ncomp = *ncomp_d;
    uint32_t mx = 0;
    for (uint32_t c = blockIdx.x * T + threadIdx.x; c < bound; c += gridDim.x * T) {
        uint32_t v = 0;
        if (c < ncomp) {
            const uint32_t len = cstart[c + 1] - cstart[c];
            v = bcnt[len];
            mx = max(mx, len);
        }
        hsz[c] = v;
    }
    for (int d = 32; d >= 1; d >>= 1) mx = max(mx, (uint32_t)__shfl_xor(mx, d, 64));
    if (lane_id() == 0 && mx) atomicMax(big, mx);
}

__global__ void __launch_bounds__(T) k_comp_starts(const uint32_t* __restrict__ flag,
                                                   const uint32_t* __restrict__ fpos,
                                                   uint32_t nk,
                                                   uint32_t* __restrict__ cstart) {
    for (uint32_t i = blockIdx.x * T + threadIdx.x; i < nk; i += gridDim.x * T)
        if (flag[i]) cstart[fpos[i]] = i;
    if (blockIdx.x == 0 && threadIdx.x == 0) cstart[fpos[nk]] = nk;
}

// Layouts of up to SB_MAX component keys run in one workgroup of SB_T
// threads (k_comp_small), bitonic sort in LDS included; larger ones take the
// multi-pass radix sort and the per-element kernels above.
constexpr int SB_T = 1024;
constexpr uint32_t SB_MAX = 8192;

// Exclusive scan over the workgroup of one value per thread (SB_T threads):
// returns this thread's prefix, *total the sum.  part: SB_T / WAVE entries.
template <class V>
__device__ V block_excl_scan(V x, V* part, V* total) {
    const int lane = lane_id(), w = threadIdx.x / WAVE;
    V incl = x;
    for (int d = 1; d < WAVE; d <<= 1) {
        const V o = __shfl_up(incl, d, 64);
        if (lane >= d) incl += o;
    }
    if (lane == WAVE - 1) part[w] = incl;
    __syncthreads();
    if (threadIdx.x == 0) {
        V acc = 0;
        for (int i = 0; i < SB_T / WAVE; ++i) {
            const V t = part[i];
            part[i] = acc;
            acc += t;
        }
        *total = acc;
    }
    __syncthreads();
    return part[w] + incl - x;
}

// The alive components' layout for up to SB_MAX keys (root << 32 | seed
// rank) in one workgroup, in place of the sort, layout, start, bucket-size
// kernels and their three scans: cv (vertices grouped by component, seed
// order inside), cstart (ncomp + 1 boundaries), *ncomp, dpre (nk + 1 prefix
// of degrees), hoff (nk + 1 prefix of the components' set bucket counts
// bcnt[size], 0 past ncomp), *big (largest component).  Thread t takes keys
// [8t, 8t + 8).
constexpr uint32_t CS_E = SB_MAX / SB_T;
__global__ void __launch_bounds__(SB_T) k_comp_small(const uint64_t* __restrict__ keys,
                                                     uint32_t nk,
                                                     const uint32_t* __restrict__ order,
                                                     const GraphDev g,
                                                     const uint32_t* __restrict__ bcnt,
                                                     uint32_t* __restrict__ cv,
                                                     uint32_t* __restrict__ cstart,
                                                     uint32_t* __restrict__ ncomp,
                                                     uint64_t* __restrict__ dpre,
                                                     uint32_t* __restrict__ hoff,
                                                     uint32_t* __restrict__ big) {
    __shared__ uint64_t a[SB_MAX];
    __shared__ uint32_t s_start[SB_MAX / 2 + 1];
    __shared__ uint32_t p32[SB_T / WAVE], t32;
    __shared__ uint64_t p64[SB_T / WAVE], t64;
    uint32_t np = 1;
    while (np < nk) np <<= 1;
    for (uint32_t i = threadIdx.x; i < np; i += SB_T) a[i] = i < nk ? keys[i] : ~0ull;
    __syncthreads();
    for (uint32_t k = 2; k <= np; k <<= 1) {
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t i = threadIdx.x; i < np; i += SB_T) {
                const uint32_t l = i ^ j;
                if (l > i) {
                    const uint64_t ai = a[i], al = a[l];
                    if ((ai > al) == ((i & k) == 0)) {
                        a[i] = al;
                        a[l] = ai;
                    }
                }
            }
            __syncthreads();
        }
    }
    const uint32_t i0 = threadIdx.x * CS_E;
    uint32_t fl = 0, fmask = 0;
    uint64_t dg = 0;
    uint32_t d[CS_E];
#pragma unroll
    for (uint32_t k = 0; k < CS_E; ++k) {
        const uint32_t i = i0 + k;
        d[k] = 0;
        if (i < nk) {
            const uint32_t v = order[(uint32_t)(a[i] & 0xffffffffu)];
            cv[i] = v;
            d[k] = g.deg[v];
            if (i == 0 || (a[i] >> 32) != (a[i - 1] >> 32)) {
                fmask |= 1u << k;
                ++fl;
            }
        }
        dg += d[k];
    }
    uint32_t f = block_excl_scan<uint32_t>(fl, p32, &t32);
    uint64_t dp = block_excl_scan<uint64_t>(dg, p64, &t64);
    const uint32_t nc = t32;
#pragma unroll
    for (uint32_t k = 0; k < CS_E; ++k) {
        const uint32_t i = i0 + k;
        if (i < nk) {
            dpre[i] = dp;
            if (fmask >> k & 1u) {
                s_start[f] = i;
                cstart[f] = i;
                ++f;
            }
        }
        dp += d[k];
    }
    if (threadIdx.x == 0) {
        s_start[nc] = nk;
        cstart[nc] = nk;
        *ncomp = nc;
        dpre[nk] = t64;
    }
    __syncthreads();
    uint32_t hs = 0, mx = 0, hv[CS_E];
#pragma unroll
    for (uint32_t k = 0; k < CS_E; ++k) {
        const uint32_t c = i0 + k;
        hv[k] = 0;
        if (c < nc) {
            const uint32_t len = s_start[c + 1] - s_start[c];
            hv[k] = bcnt[len];
            mx = max(mx, len);
        }
        hs += hv[k];
    }
    uint32_t h = block_excl_scan<uint32_t>(hs, p32, &t32);
#pragma unroll
    for (uint32_t k = 0; k < CS_E; ++k) {
        const uint32_t c = i0 + k;
        if (c < nk) hoff[c] = h;
        h += hv[k];
    }
    if (threadIdx.x == 0) hoff[nk] = t32;
    for (int dd = 32; dd >= 1; dd >>= 1) mx = max(mx, (uint32_t)__shfl_xor(mx, dd, 64));
    if (lane_id() == 0 && mx) atomicMax(big, mx);
}

// ---------------------------------------------------------------- extraction
struct XArgs {
    GraphDev g;
    double eps;
    const uint32_t* cv;      // component-grouped vertices, seed order inside
    const uint32_t* cstart;  // ncomp + 1 boundaries into cv
    const uint64_t* dpre;    // prefix of degrees over cv (scratch placement)
    const uint32_t* ncomp;   // number of components (device)
    uint32_t* cseed;         // seed of each vertex's cluster, NONE if not yet
    uint32_t* mrank;         // position at which the vertex joined its cluster
    uint32_t* mk;            // neighbour stamp (seed + 1)
    uint32_t* cm;            // candidate stamp (seed + 1)
    double* tw;              // two-hop weight per candidate vertex
    uint32_t* mem;           // [start, start+len): u, then alive neighbours
    uint32_t* cand;          // [2 start, ...): candidates, padded for sorting
    Rec* grec;               // [start, ...): candidate records (large case)
    double* gpv;             // [start, ...): prefetched supports (large case)
    uint64_t* wl;            // [dpre[start], ...): deletion worklist (LIFO):
                             // (edge << 32) | first endpoint of the pair
    TermBuf tb;              // stored triangle lists (tb.t null: intersect)
    Counters* ctr;
    const uint8_t* xact;     // per edge: post-clean support is the reference's bits
    const uint32_t* bcnt;    // unordered_set bucket count after k insertions
    bool uset_lds;           // bcnt[min(US_LDS, largest cluster)] <= US_LDS_BKT
    const uint32_t* hoff;    // per component: start of its bucket scratch
    uint32_t* hnext;         // [start, ...): set node links
    uint32_t* hbkt;          // [hoff[c], ...): set buckets
    uint32_t* undec;         // edges no bound decided (for an exact re-run)
    uint32_t undec_cap;
    // Cascades handed to the workgroup kernel (k_cascade_wg): a cascade whose
    // stack holds cascade_min entries or more between two pops stops; its
    // component is listed in dlist (count *ndefer) with the stack length
    // (dtop) and the next seed (resume), and the component waits.  clist
    // (count *nclist), when given, is the list of components this launch
    // resumes instead of every component.  cascade_min 0: no hand-over.
    uint32_t cascade_min = 0;
    uint32_t* resume = nullptr;
    uint32_t* dtop = nullptr;
    uint32_t* dlist = nullptr;
    unsigned* ndefer = nullptr;
    const uint32_t* clist = nullptr;
    const unsigned* nclist = nullptr;
};

// Cascade decision for alive edge f whose support just became v (threshold
// thr): certain outside the cheap bound, otherwise (need set) certified
// against every subtraction order (audit.hpp) and settled by settle().
// An edge whose post-clean support is the reference's own value (xact) has,
// in the reference's cascade order, the reference's value here too: its
// comparison is taken as is.
__device__ __forceinline__ bool decide_cheap_v(bool exact, uint32_t tcnt, double sup0, double v,
                                               double thr, bool& need) {
    if (exact) return v < thr;
    const double cb = cheap_bound(tcnt, sup0);
    if (v < thr - cb) return true;
    if (v >= thr + cb) return false;
    need = true;
    return false;
}
__device__ bool decide_cheap(const XArgs& A, uint32_t f, double v, double thr, bool& need) {
    const GraphDev& g = A.g;
    return decide_cheap_v(A.xact && A.xact[f], g.tcnt[f], g.sup0[f], v, thr, need);
}

// Outcome c of a certification; an undecided one is recorded for an exact
// re-run of its region and takes this library's comparison.
__device__ bool settle(const XArgs& A, uint32_t f, double v, double thr, int c,
                       unsigned long long& amb) {
    if (c == CERT_UNDECIDED) {
        amb++;
        if (A.undec) {
            const unsigned i = atomicAdd(&A.ctr->n_undecided, 1u);
            if (i < A.undec_cap) A.undec[i] = f;
        }
        return v < thr;
    }
    return c == CERT_REMOVE;
}

__device__ __forceinline__ uint64_t lanes_below() {
    const int l = lane_id();
    return l ? (~0ull >> (64 - l)) : 0ull;
}

// Ascending bitonic sort of a[0, np) by the wave (np a power of two).
__device__ void wave_bitonic(uint32_t* a, uint32_t np) {
    const int lane = lane_id();
    for (uint32_t k = 2; k <= np; k <<= 1) {
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t i = lane; i < np; i += XT) {
                uint32_t l = i ^ j;
                if (l > i) {
                    uint32_t ai = a[i], al = a[l];
                    bool up = (i & k) == 0;
                    if ((ai > al) == up) {
                        a[i] = al;
                        a[l] = ai;
                    }
                }
            }
            __syncthreads();
        }
    }
}

// Status change of edge e whose two slots (sl, its mirror sp) are known.
__device__ __forceinline__ void set_status_at(const GraphDev& g, uint32_t e, uint32_t sl,
                                              uint32_t sp, uint8_t c) {
    g.est[e] = c;
    g.st[sl] = c;
    g.st[sp] = c;
}

// The cascade of one cluster (deleteAndClean, Decomposition.h:147-283) over
// stored triangle lists: the LIFO worklist wl[ebase, ebase + top) popped
// edge by edge, in the reference's order.  Per pop the loads come in five
// waves (the popped edge's data; its other endpoint, mirror slot and first
// chunk of triangle records, lanes past the list with a blank record so that
// no load waits on a branch; the degrees and the records' edges' statuses, supports
// and decision inputs; the third vertices and mirror slots; the third
// vertices' supports), and the entry on top of the stack is carried in
// registers when the pop pushed.  Returns 0, or the stack length when the
// cascade is handed to the workgroup kernel (the stack is wl[ebase, +top)
// between pops).
__device__ uint32_t cascade_terms(const XArgs& A, uint64_t ebase, uint32_t top,
                                  unsigned long long& amb_c, unsigned long long& n_tri,
                                  unsigned long long& n_edges, double* s_ct, uint32_t* s_cn) {
    const GraphDev& g = A.g;
    const TermBuf& tb = A.tb;
    const int lane = lane_id();
    bool have = false;
    uint64_t next = 0;
    while (top > 0) {
        if (A.cascade_min && top >= A.cascade_min) return top;
        --top;
        const uint64_t ent = have ? next : A.wl[ebase + top];
        have = false;
        const uint32_t e = (uint32_t)(ent >> 32), pa = (uint32_t)ent;
        const uint32_t x = g.esrc[e], sle = g.eslot[e], tc = g.tcnt[e];
        const uint64_t c0 = tb.zoff[e];
        const uint32_t dpa = g.deg[pa];
        const uint32_t y = g.nb[sle];
        const uint32_t spe = g.part[sle];
        double px = g.pv[x];
        const uint32_t pb = pa == x ? y : x;
        const uint32_t dpb = g.deg[pb];
        double py = g.pv[y];
        const uint32_t dx = pa == x ? dpa : dpb, dy = pa == x ? dpb : dpa;
        // the stored lists' lower side: d_x <= d_y ? x : y; the reference's
        // lower endpoint of the pair (pa, pb): smaller degree, pb on a tie
        const uint32_t tlo = dx <= dy ? x : y;
        const uint32_t rlo = dpa < dpb ? pa : pb;
        const uint32_t rhi = rlo == pa ? pb : pa;
        const bool swp = rlo != tlo;
        uint64_t last = 0;
        bool any = false;
        const uint64_t c1 = c0 + tc;
        for (uint64_t p = c0; p < c1; p += XT) {
            const uint64_t kk = p + lane;
            const bool in = kk < c1;
            const Term q = term_or_blank(tb.t, kk, in);
            const uint32_t fj = q.sa, fh = q.sb;
            const uint8_t sj = g.est[fj], sh = g.est[fh];
            const uint32_t aj = g.esrc[fj], slj = g.eslot[fj], slh = g.eslot[fh];
            const double supj = g.sup[fj], suph = g.sup[fh];
            const double ewj = g.ew[fj], ewh = g.ew[fh];
            const bool xj = A.xact && A.xact[fj], xh = A.xact && A.xact[fh];
            const uint32_t tcj = g.tcnt[fj], tch = g.tcnt[fh];
            const double s0j = g.sup0[fj], s0h = g.sup0[fh];
            const uint32_t wj = g.nb[slj];
            const uint32_t spj = g.part[slj], sph = g.part[slh];
            const bool hit = in && sj != ST_DEAD && sh != ST_DEAD;
            const uint32_t w = aj == tlo ? wj : aj;  // third vertex
            const double pvw = g.pv[w];
            const double t = q.w;
            uint64_t hb = ballot64(hit);
            n_tri += (lane == 0) ? (unsigned long long)__popcll(hb) : 0ull;
            while (hb) {  // the popped edge's endpoints, in w order
                const int bl = __ffsll((unsigned long long)hb) - 1;
                const double tv = __shfl(t, bl, 64);
                px -= tv;
                py -= tv;
                hb &= hb - 1;
            }
            // fj is the edge (tlo, w), fh the other side's; the reference's
            // (lower, w) is fj unless its lower endpoint is the other one
            const uint32_t jl = swp ? fh : fj, jh = swp ? fj : fh;
            const uint8_t sl = swp ? sh : sj, shh = swp ? sj : sh;
            const double supl = swp ? suph : supj, suph2 = swp ? supj : suph;
            const double ewl = swp ? ewh : ewj, ewh2 = swp ? ewj : ewh;
            const bool xl = swp ? xh : xj, xh2 = swp ? xj : xh;
            const uint32_t tcl = swp ? tch : tcj, tch2 = swp ? tcj : tch;
            const double s0l = swp ? s0h : s0j, s0h2 = swp ? s0j : s0h;
            bool push1 = false, push2 = false, need1 = false, need2 = false;
            double v1 = 0.0, v2 = 0.0, thr1 = 0.0, thr2 = 0.0;
            if (hit) {
                g.pv[w] = pvw - t;
                // (lower, w) then (higher, w), as Decomposition.h:238-264.
                // Queued edges die anyway; their supports are not kept.
                if (sl == ST_ALIVE) {
                    v1 = supl - t;
                    g.sup[jl] = v1;
                    g.touched[jl] = 1;
                    thr1 = A.eps * ewl;
                    push1 = decide_cheap_v(xl, tcl, s0l, v1, thr1, need1);
                }
                if (shh == ST_ALIVE) {
                    v2 = suph2 - t;
                    g.sup[jh] = v2;
                    g.touched[jh] = 1;
                    thr2 = A.eps * ewh2;
                    push2 = decide_cheap_v(xh2, tch2, s0h2, v2, thr2, need2);
                }
            }
            // The decisions the cheap bound leaves open read statuses only
            // (no support written above), so they may follow all of the
            // chunk's updates: one edge at a time by the whole wave.
#pragma unroll
            for (int side = 0; side < 2; ++side) {
                uint64_t nb = ballot64(side == 0 ? need1 : need2);
                while (nb) {  // wave-uniform
                    const int bl = __ffsll((unsigned long long)nb) - 1;
                    nb &= nb - 1;
                    const uint32_t f = __shfl(side == 0 ? jl : jh, bl, 64);
                    const double th = __shfl(side == 0 ? thr1 : thr2, bl, 64);
                    const int cc = certify_terms_wave(g, tb, f, th, false, e, s_ct, s_cn);
                    if (lane == bl) {
                        if (side == 0) push1 = settle(A, f, v1, th, cc, amb_c);
                        else push2 = settle(A, f, v2, th, cc, amb_c);
                    }
                }
            }
            const uint32_t np2 = (uint32_t)push1 + (uint32_t)push2;
            uint32_t pre = np2;  // exclusive prefix of pushes over lanes
            for (int d = 1; d < 64; d <<= 1) {
                const uint32_t o = __shfl_up(pre, d, 64);
                if (lane >= d) pre += o;
            }
            const uint32_t tot = __shfl(pre, 63, 64);
            pre -= np2;
            const uint64_t e1 = ((uint64_t)jl << 32) | rlo, e2 = ((uint64_t)jh << 32) | rhi;
            // slots of jl and jh: (slj, spj) belong to fj, (slh, sph) to fh
            if (push1) {
                A.wl[ebase + top + pre] = e1;
                set_status_at(g, jl, swp ? slh : slj, swp ? sph : spj, ST_QUEUED);
            }
            if (push2) {
                A.wl[ebase + top + pre + (uint32_t)push1] = e2;
                set_status_at(g, jh, swp ? slj : slh, swp ? spj : sph, ST_QUEUED);
            }
            if (tot) {  // the new top of the stack, from the highest pushing lane
                const uint64_t pbm = ballot64(np2 != 0);
                const int hl = 63 - __clzll((long long)pbm);
                last = __shfl(push2 ? e2 : e1, hl, 64);
                any = true;
            }
            top += tot;
            __syncthreads();
        }
        if (lane == 0) {
            g.pv[x] = px;
            g.pv[y] = py;
            set_status_at(g, e, sle, spe, ST_DEAD);
            g.sup[e] = 0.0;
        }
        n_edges++;
        have = any;
        next = last;
        __syncthreads();
    }
    return 0;
}

// The inner loop of seed u (Decomposition.h:728-805): every pair of alive
// neighbours nbr[a], nbr[b] (b >= a, both of degree at most lim = d_u / eps)
// adds the triangle (u, v, w) to internal and enumerates the triangles
// (v, w, x) over the lower-degree side's list, searched in the other side's;
// x a neighbour of u (mk[x] == stamp) adds to internal when it is the largest
// of v, w, x, any other x is a two-hop candidate: appended to cand (in
// discovery order) the first time (cm[x] = stamp, tw[x] = its weight), its
// weight added to tw[x] after that.  Called by the whole wave; returns the
// number of candidates and adds to internal in the reference's order.
__device__ uint32_t seed_pairs(const GraphDev& g, uint32_t u, uint32_t du, double lim,
                               uint32_t stamp, const uint32_t* nbr, uint32_t nA,
                               const uint32_t* mk, uint32_t* cm, double* tw, uint32_t* cand,
                               uint32_t* s_hn, uint8_t* s_ha, double& internal) {
    const int lane = lane_id();
    uint32_t nc = 0;
    // Up to 64 alive neighbours: ids, degrees and list bounds one per
    // lane, broadcast per pair (else read per pair).
    const bool regs = nA <= (uint32_t)XT;
    uint32_t r_v = 0, r_d = 0, r_o0 = 0, r_o1 = 0;
    if (regs && (uint32_t)lane < nA) {
        r_v = nbr[lane];
        r_d = g.deg[r_v];
        r_o0 = g.off[r_v];
        r_o1 = g.off[r_v + 1];
    }
    for (uint32_t a = 0; a < nA; ++a) {
        const uint32_t v = regs ? __shfl(r_v, (int)a, 64) : nbr[a];
        const uint32_t dv = regs ? __shfl(r_d, (int)a, 64) : g.deg[v];
        if ((double)dv > lim) continue;
        const uint32_t v0 = regs ? __shfl(r_o0, (int)a, 64) : g.off[v];
        const uint32_t v1 = regs ? __shfl(r_o1, (int)a, 64) : g.off[v + 1];
        for (uint32_t b = a; b < nA; ++b) {
            const uint32_t w = regs ? __shfl(r_v, (int)b, 64) : nbr[b];
            const uint32_t dw = regs ? __shfl(r_d, (int)b, 64) : g.deg[w];
            if ((double)dw > lim) continue;
            const uint32_t w0 = regs ? __shfl(r_o0, (int)b, 64) : g.off[w];
            const uint32_t w1 = regs ? __shfl(r_o1, (int)b, 64) : g.off[w + 1];
            internal += tri_w(du, dv, dw);
            const uint32_t lo = dv < dw ? v : w, hi = dv < dw ? w : v;
            uint32_t p = dv < dw ? v0 : w0;
            const uint32_t le = dv < dw ? v1 : w1;
            uint32_t hlo = dv < dw ? w0 : v0;
            const uint32_t he = dv < dw ? w1 : v1;
            // N(hi) and its alive flags staged in LDS when short
            // enough (the first chunk of N(lo) loaded alongside), so
            // that each lane's search reads LDS; else searched in HBM.
            const uint32_t dh = he - hlo;
            const bool staged = dh <= XSTAGE;
            uint32_t x0 = 0, d0 = 0;
            uint8_t st0 = ST_DEAD;
            if (staged) {
                if (p + lane < le) {
                    x0 = g.nb[p + lane];
                    st0 = g.st[p + lane];
                    d0 = g.ndo[p + lane].x;
                }
                for (uint32_t i = lane; i < dh; i += XT) {
                    s_hn[i] = g.nb[hlo + i];
                    s_ha[i] = g.st[hlo + i] != ST_DEAD;
                }
                wave_sync();
            }
            for (bool first = true; p < le; p += XT, first = false) {
                const uint32_t l = p + lane;
                const bool in = l < le;
                const uint32_t x = !in ? 0 : staged && first ? x0 : g.nb[l];
                const uint8_t sx = !in ? (uint8_t)ST_DEAD : staged && first ? st0 : g.st[l];
                // d_x from the slot's copy (loaded with x, not after it)
                const uint32_t dx = !in ? 0u : staged && first ? d0 : g.ndo[l].x;
                const bool ok = in && sx != ST_DEAD && x != lo && x != hi && x != u;
                bool hit;
                if (staged) {
                    uint32_t a0 = 0, a1 = dh;  // lower bound of x in the staged list
                    while (a0 < a1) {
                        const uint32_t mid = (a0 + a1) >> 1;
                        if (s_hn[mid] < x) a0 = mid + 1;
                        else a1 = mid;
                    }
                    hit = ok && a0 < dh && s_hn[a0] == x && s_ha[a0];
                } else {
                    const uint32_t pos = in ? lower_bound(g.nb, hlo, he, x) : hlo;
                    const uint32_t nh = __shfl(pos, (int)min(63u, le - p - 1), 64);
                    hit = ok && pos < he && g.nb[pos] == x && g.st[pos] != ST_DEAD;
                    hlo = nh;
                }
                // x's marks and weight in one round of loads
                uint32_t mkx = 0, cmx = 0;
                double twx = 0.0;
                if (hit) {
                    mkx = mk[x];
                    cmx = cm[x];
                    twx = tw[x];
                }
                const double t = hit ? tri_w(dx, dv, dw) : 0.0;
                const bool isn = hit && mkx == stamp;
                uint64_t ib = ballot64(isn && x > v && x > w);
                while (ib) {  // wave-uniform
                    const int bl = __ffsll((unsigned long long)ib) - 1;
                    internal += __shfl(t, bl, 64);
                    ib &= ib - 1;
                }
                const bool ct = hit && !isn;
                const bool fresh = ct && cmx != stamp;
                const uint64_t fb = ballot64(fresh);
                if (fresh) {
                    cm[x] = stamp;
                    tw[x] = t;
                    cand[nc + __popcll(fb & lanes_below())] = x;
                } else if (ct) {
                    tw[x] = twx + t;  // x is one lane's alone within a pair
                }
                nc += (uint32_t)__popcll(fb);
            }
            __syncthreads();  // candidate weights of this pair are final
        }
    }
    return nc;
}

// The nc candidates of a seed (cand, discovery order) in ascending id, the
// std::set order of Decomposition.h:811-827: rec[i] = (x, tw[x]); returns
// their weights' sum in that order (potentialWgt).  keys: room for nc
// rounded up to a power of two (LDS, or cand itself).  Called by the wave.
__device__ double seed_candidates(const uint32_t* cand, uint32_t nc, uint32_t* keys,
                                  const double* tw, Rec* rec) {
    const int lane = lane_id();
    uint32_t np = 1;
    while (np < nc) np <<= 1;
    for (uint32_t i = lane; i < np; i += XT) keys[i] = i < nc ? cand[i] : NONE;
    __syncthreads();
    if (nc > 1) wave_bitonic(keys, np);
    double potential = 0.0;
    for (uint32_t p = 0; p < nc; p += XT) {  // ascending x (Decomposition.h:815-827)
        const uint32_t i = p + lane;
        const bool in = i < nc;
        const uint32_t x = in ? keys[i] : 0;
        const double wx = in ? tw[x] : 0.0;
        if (in) {
            rec[i].vertex = x;
            rec[i].wgt = wx;
        }
        const uint32_t cnt = min(64u, nc - p);
        for (uint32_t b2 = 0; b2 < cnt; ++b2) potential += __shfl(wx, (int)b2, 64);
    }
    return potential;
}

template <bool TERMS>
__global__ void __launch_bounds__(XT) k_extract(XArgs A) {
    __shared__ Rec srec[LDS_CAND];
    __shared__ double spv[LDS_CAND];
    __shared__ double sbd[LDS_CAND];
    __shared__ uint32_t skey[LDS_CAND];
    __shared__ int s_maxind;
    __shared__ double s_ct[kMaxPerm];  // certification: gathered dead weights
    __shared__ uint32_t s_cn;
    __shared__ uint32_t s_hn[XSTAGE];  // inner loop: the pair's N(hi) ...
    __shared__ uint8_t s_ha[XSTAGE];   // ... and its alive flags
    // the cluster's set order: members, iteration order, links, buckets,
    // bucket counts after k insertions
    __shared__ uint32_t s_uin[US_LDS], s_uord[US_LDS], s_unxt[US_LDS], s_ubkt[US_LDS_BKT];
    __shared__ uint32_t s_ubc[US_LDS + 1];
    const GraphDev& g = A.g;
    const int lane = lane_id();
    unsigned long long n_seeds = 0, n_tri = 0, n_edges = 0, amb_c = 0, amb_r = 0;
    const uint32_t nwork = A.clist ? *A.nclist : *A.ncomp;

    while (true) {
        uint32_t i = 0;
        if (lane == 0) i = atomicAdd(&A.ctr->work_next, 1u);
        i = __shfl(i, 0, 64);
        if (i >= nwork) break;
        const uint32_t c = A.clist ? A.clist[i] : i;
        const uint32_t start = A.cstart[c], len = A.cstart[c + 1] - start;
        const uint64_t ebase = A.dpre[start];

        for (uint32_t r = A.resume ? A.resume[c] : 0; r < len; ++r) {
            const uint32_t u = A.cv[start + r];
            if (A.cseed[u] != NONE) continue;
            const uint32_t du = g.deg[u];
            const double lim = (double)du / A.eps;
            const uint32_t stamp = u + 1;
            n_seeds++;

            // 1. alive neighbours (Decomposition.h:620-631)
            uint32_t nA = 0;
            const uint32_t ub = g.off[u], ue = g.off[u + 1];
            for (uint32_t p = ub; p < ue; p += XT) {
                const uint32_t j = p + lane;
                const bool alive = j < ue && g.st[j] != ST_DEAD;
                const uint32_t v = j < ue ? g.nb[j] : 0;
                const uint64_t bal = ballot64(alive);
                if (alive) {
                    const uint32_t r = nA + (uint32_t)__popcll(bal & lanes_below());
                    A.mem[start + 1 + r] = v;
                    A.mk[v] = stamp;
                    A.cseed[v] = u;
                    A.mrank[v] = 1 + r;
                }
                nA += (uint32_t)__popcll(bal);
            }
            if (lane == 0) {
                A.mem[start] = u;
                A.cseed[u] = u;
                A.mrank[u] = 0;
            }
            __syncthreads();
            if (nA == 0) continue;  // singleton cluster {u}

            // 2. inner loop (Decomposition.h:728-805)
            double internal = 0.0;
            const uint64_t cb = 2ull * start;
            const uint32_t nc = seed_pairs(g, u, du, lim, stamp, A.mem + start + 1, nA, A.mk, A.cm,
                                           A.tw, A.cand + cb, s_hn, s_ha, internal);

            // 3. candidates ascending (std::set), potential, sort, prefix
            uint32_t np = 1;
            while (np < nc) np <<= 1;
            const bool small = np <= LDS_CAND;
            Rec* rec = small ? srec : A.grec + start;
            double* pvv = small ? spv : A.gpv + start;
            double* bdd = sbd;  // per-candidate support bounds (LDS case only)
            const double potential =
                seed_candidates(A.cand + cb, nc, small ? skey : A.cand + cb, A.tw, rec);
            __syncthreads();
            if (lane == 0) isort::sort(rec, rec + nc);
            __syncthreads();
            // supports of the candidates in sorted order
            for (uint32_t i = lane; i < nc; i += XT) {
                const uint32_t x = rec[i].vertex;
                pvv[i] = g.pv[x];
                if (small) bdd[i] = cheap_bound(2 * g.vtcnt[x], g.pv0[x]);
            }
            __syncthreads();
            if (lane == 0) {
                const double D0 = internal + potential;
                double ratio = internal / D0;
                double eratio = 0.0;
                int maxind = -1;
                double num = 0.0, den = 0.0, eden = 0.0;
                const bool audit = ratio == ratio;  // NaN ratio: every '>' is false
                for (uint32_t i = 0; i < nc; ++i) {
                    num += rec[i].wgt;
                    den += pvv[i];
                    const double bi = small ? bdd[i]
                                            : cheap_bound(2 * g.vtcnt[rec[i].vertex],
                                                           g.pv0[rec[i].vertex]);
                    eden += bi + 0x1.0p-52 * fabs(den);
                    const double D = D0 + den;
                    const double nr = (internal + num) / D;
                    if (audit) {
                        const double enr =
                            1.01 * nr * ((eden + 0x1.0p-52 * D) / D) + 0x1.0p-51 * nr;
                        if (fabs(nr - ratio) <= enr + eratio) amb_r++;
                        if (nr > ratio) eratio = enr;
                    }
                    if (nr > ratio) {
                        ratio = nr;
                        maxind = (int)i;
                    }
                }
                s_maxind = maxind;
            }
            __syncthreads();
            const int maxind = s_maxind;
            for (int i = lane; i <= maxind; i += XT) {
                A.cseed[rec[i].vertex] = u;
                A.mrank[rec[i].vertex] = 1 + nA + (uint32_t)i;
            }
            __syncthreads();

            // 4. queue every alive edge of the cluster, its vertices in the
            // iteration order of the reference's unordered_set (filled as the
            // cluster grew: seed, alive neighbours, chosen candidates), their
            // slots in order (Decomposition.h:897-914)
            // Up to US_LDS members: the set is emulated in LDS (its bucket
            // counts staged first), so lane 0's dependent steps read LDS.
            uint32_t top = 0;
            const uint32_t nmem = 1 + nA + (uint32_t)(maxind + 1);
            const bool ulds = A.uset_lds && nmem <= US_LDS;
            uint32_t* ins = ulds ? s_uin : A.cand + cb;  // members in insertion order
            uint32_t* ord = ulds ? s_uord : A.cand + cb + len;  // ... in iteration order
            if (ulds)
                for (uint32_t q = lane; q <= nmem; q += XT) s_ubc[q] = A.bcnt[q];
            for (uint32_t q = lane; q < nmem; q += XT)
                ins[q] = q < 1 + nA ? A.mem[start + q] : rec[q - 1 - nA].vertex;
            __syncthreads();
            if (lane == 0) {
                if (ulds) uset_order(ins, nmem, s_ubc, s_unxt, s_ubkt, ord);
                else uset_order(ins, nmem, A.bcnt, A.hnext + start, A.hbkt + A.hoff[c], ord);
            }
            __syncthreads();
            // Up to 64 members: ids and list bounds one per lane, read once.
            // A member's statuses are read after the earlier members' writes
            // (an edge queued from one end is not queued again from the
            // other); the queued edge's slots are j and its mirror part[j].
            const bool mregs = nmem <= (uint32_t)XT;
            uint32_t m_z = 0, m_b = 0, m_e = 0;
            if (mregs && (uint32_t)lane < nmem) {
                m_z = ord[lane];
                m_b = g.off[m_z];
                m_e = g.off[m_z + 1];
            }
            for (uint32_t q = 0; q < nmem; ++q) {
                const uint32_t z = mregs ? __shfl(m_z, (int)q, 64) : ord[q];
                const uint32_t zb = mregs ? __shfl(m_b, (int)q, 64) : g.off[z];
                const uint32_t ze = mregs ? __shfl(m_e, (int)q, 64) : g.off[z + 1];
                for (uint32_t p = zb; p < ze; p += XT) {
                    const uint32_t j = p + lane;
                    const bool in = j < ze;
                    const uint8_t sj = in ? g.st[j] : (uint8_t)ST_DEAD;
                    const uint32_t f = in ? g.s2e[j] : 0u;
                    const uint32_t pj = in ? g.part[j] : 0u;
                    const bool al = sj == ST_ALIVE;
                    const uint64_t bal = ballot64(al);
                    if (al) {
                        A.wl[ebase + top + __popcll(bal & lanes_below())] =
                            ((uint64_t)f << 32) | z;  // the pair (z, nbor)
                        set_status_at(g, f, j, pj, ST_QUEUED);
                    }
                    top += (uint32_t)__popcll(bal);
                }
                __syncthreads();
            }

            // ... and cascade (deleteAndClean, Decomposition.h:147-283)
            if constexpr (TERMS) {
                const uint32_t left = cascade_terms(A, ebase, top, amb_c, n_tri, n_edges, s_ct,
                                                    &s_cn);
                if (left) {  // the workgroup kernel finishes it; seeds resume at r + 1
                    if (lane == 0) {
                        A.resume[c] = r + 1;
                        A.dtop[c] = left;
                        A.dlist[atomicAdd(A.ndefer, 1u)] = c;
                    }
                    break;
                }
            } else {
            while (top > 0) {
                --top;
                const uint64_t ent = A.wl[ebase + top];
                const uint32_t e = (uint32_t)(ent >> 32), pa = (uint32_t)ent;
                const uint32_t sl = g.eslot[e];
                const uint32_t x = g.esrc[e], y = g.nb[sl];
                const uint32_t dx = g.deg[x], dy = g.deg[y];
                const uint32_t lo = dx < dy ? x : y, hi = dx < dy ? y : x;
                // The reference's lower endpoint of the popped pair (pa, pb):
                // the one of smaller degree, pb on a tie.  (lower, w) is
                // checked and pushed before (higher, w).
                const uint32_t pb = pa == x ? y : x;
                const uint32_t rlo = g.deg[pa] < g.deg[pb] ? pa : pb;
                const uint32_t rhi = rlo == pa ? pb : pa;
                // Triangles of e: from its stored list (ascending third
                // vertex, the order the reference visits them), or by
                // intersecting N(lo) with N(hi).  Either way, per chunk of 64:
                // slot ja of (lo, w) and jb of (hi, w), both not deleted.
                constexpr bool terms = false;
                const uint32_t tlo = terms ? (dx <= dy ? x : y) : lo;
                const uint64_t c0 = terms ? A.tb.zoff[e] : g.off[lo];
                const uint64_t c1 = terms ? c0 + g.tcnt[e] : g.off[lo + 1];
                uint32_t hlo = g.off[hi];
                const uint32_t he = g.off[hi + 1];
                // The endpoints' supports lose the triangles' weights in w
                // order on lane 0, in registers until the pop ends (no other
                // lane and no other wave writes them meanwhile).
                double px = 0.0, py = 0.0;
                if (lane == 0) {
                    px = g.pv[x];
                    py = g.pv[y];
                }
                for (uint64_t p = c0; p < c1; p += XT) {
                    const uint64_t kk = p + lane;
                    const bool in = kk < c1;
                    uint32_t fj = 0, fh = 0, w = 0;
                    bool hit = false;
                    double t = 0.0;
                    if (terms) {
                        if (in) {
                            const Term q = A.tb.t[kk];
                            fj = q.sa;
                            fh = q.sb;
                            hit = g.est[fj] != ST_DEAD && g.est[fh] != ST_DEAD;
                            if (hit) {
                                const uint32_t a0 = g.esrc[fj];  // third vertex
                                w = a0 == tlo ? g.nb[g.eslot[fj]] : a0;
                                t = q.w;
                            }
                        }
                    } else {
                        const uint32_t ja = (uint32_t)kk;
                        w = in ? g.nb[ja] : 0;
                        const uint32_t jb = in ? lower_bound(g.nb, hlo, he, w) : hlo;
                        const uint32_t nh = __shfl(jb, (int)min<uint64_t>(63, c1 - p - 1), 64);
                        hit = in && g.st[ja] != ST_DEAD && jb < he && g.nb[jb] == w &&
                              g.st[jb] != ST_DEAD;
                        if (hit) {
                            fj = g.s2e[ja];
                            fh = g.s2e[jb];
                            t = tri_w(dx, dy, g.deg[w]);
                        }
                        hlo = nh;
                    }
                    const uint8_t sj = hit ? g.est[fj] : ST_DEAD;
                    const uint8_t sh = hit ? g.est[fh] : ST_DEAD;
                    uint64_t hb = ballot64(hit);
                    n_tri += (lane == 0) ? (unsigned long long)__popcll(hb) : 0ull;
                    while (hb) {  // the popped edge's endpoints, in w order
                        const int bl = __ffsll((unsigned long long)hb) - 1;
                        const double tv = __shfl(t, bl, 64);
                        if (lane == 0) {
                            px -= tv;
                            py -= tv;
                        }
                        hb &= hb - 1;
                    }
                    // fj is the edge (L, w) of our lower side L, fh (H, w);
                    // the reference's (lower, w) is fj unless its lower
                    // endpoint is H (degree tie resolved by the pair order).
                    const bool swp = rlo != (terms ? tlo : lo);
                    const uint32_t jl = swp ? fh : fj, jh = swp ? fj : fh;
                    const uint8_t sl = swp ? sh : sj, shh = swp ? sj : sh;
                    bool push1 = false, push2 = false, need1 = false, need2 = false;
                    double v1 = 0.0, v2 = 0.0, thr1 = 0.0, thr2 = 0.0;
                    if (hit) {
                        g.pv[w] -= t;
                        // (lower, w) then (higher, w), as Decomposition.h:238-264.
                        // Queued edges die anyway; their supports are not kept.
                        if (sl == ST_ALIVE) {
                            v1 = g.sup[jl] - t;
                            g.sup[jl] = v1;
                            g.touched[jl] = 1;
                            thr1 = A.eps * g.ew[jl];
                            push1 = decide_cheap(A, jl, v1, thr1, need1);
                        }
                        if (shh == ST_ALIVE) {
                            v2 = g.sup[jh] - t;
                            g.sup[jh] = v2;
                            g.touched[jh] = 1;
                            thr2 = A.eps * g.ew[jh];
                            push2 = decide_cheap(A, jh, v2, thr2, need2);
                        }
                    }
                    // The decisions the cheap bound leaves open read statuses
                    // only (no support written above), so they may follow all
                    // of the chunk's updates: with stored lists one edge at a
                    // time by the whole wave, else per lane.
                    if (terms) {
#pragma unroll
                        for (int side = 0; side < 2; ++side) {
                            uint64_t nb = ballot64(side == 0 ? need1 : need2);
                            while (nb) {  // wave-uniform
                                const int bl = __ffsll((unsigned long long)nb) - 1;
                                nb &= nb - 1;
                                const uint32_t f = __shfl(side == 0 ? jl : jh, bl, 64);
                                const double th = __shfl(side == 0 ? thr1 : thr2, bl, 64);
                                const int cc = certify_terms_wave(g, A.tb, f, th, false, e, s_ct,
                                                                  &s_cn);
                                if (lane == bl) {
                                    if (side == 0) push1 = settle(A, f, v1, th, cc, amb_c);
                                    else push2 = settle(A, f, v2, th, cc, amb_c);
                                }
                            }
                        }
                    } else {
                        if (need1)
                            push1 = settle(A, jl, v1, thr1,
                                           certify_edge(g, A.tb, jl, thr1, false, rhi, e), amb_c);
                        if (need2)
                            push2 = settle(A, jh, v2, thr2,
                                           certify_edge(g, A.tb, jh, thr2, false, rlo, e), amb_c);
                    }
                    const uint32_t np2 = (uint32_t)push1 + (uint32_t)push2;
                    // exclusive prefix of pushes over lanes
                    uint32_t pre = 0;
                    {
                        uint32_t sc = np2;
                        for (int d = 1; d < 64; d <<= 1) {
                            uint32_t o = __shfl_up(sc, d, 64);
                            if (lane >= d) sc += o;
                        }
                        pre = sc - np2;
                    }
                    const uint32_t tot = __shfl(pre + np2, 63, 64);
                    if (push1) {
                        A.wl[ebase + top + pre] = ((uint64_t)jl << 32) | rlo;
                        set_status(g, jl, ST_QUEUED);
                    }
                    if (push2) {
                        A.wl[ebase + top + pre + (uint32_t)push1] = ((uint64_t)jh << 32) | rhi;
                        set_status(g, jh, ST_QUEUED);
                    }
                    top += tot;
                    __syncthreads();
                }
                if (lane == 0) {
                    g.pv[x] = px;
                    g.pv[y] = py;
                    set_status(g, e, ST_DEAD);
                    g.sup[e] = 0.0;
                }
                n_edges++;
                __syncthreads();
            }
            }
            __syncthreads();
        }
    }
    if (lane == 0) {
        atomicAdd(&A.ctr->seeds, n_seeds);
        atomicAdd(&A.ctr->tri_removed, n_tri);
        atomicAdd(&A.ctr->edges_removed, n_edges);
    }
    for (int d = 32; d >= 1; d >>= 1) {
        amb_c += __shfl_xor(amb_c, d, 64);
        amb_r += __shfl_xor(amb_r, d, 64);
    }
    if (lane == 0) {
        if (amb_c) atomicAdd(&A.ctr->ambiguous_clean, amb_c);
        if (amb_r) atomicAdd(&A.ctr->ambiguous_ratio, amb_r);
    }
}

// The cascades k_extract handed over (one workgroup of NW waves per listed
// component): the same LIFO pops over stored triangle lists, each pop's list
// split into rounds of NW chunks of 64 records, one chunk per wave.  Within
// a pop the records are distinct triangles {x, y, w}: their third vertices
// and other edges are distinct, no edge but the popped one dies before the
// pop ends, and a certification counts DEAD edges only, so the chunks of a
// round are independent; their pushes go on the stack in chunk order (an
// exclusive prefix of the waves' push counts), which is the reference's
// order.  The popped edge's endpoints lose the sum of the pop's live
// triangle weights in one subtraction (partial sums per lane, then per wave
// in wave order) instead of one weight at a time: vertex supports are only
// compared within the error bound that covers any subtraction order
// (ambiguous_ratio).  Thread 0 of block 0 also clears the work queue head
// and next_count, the count the next k_extract launch appends to.
template <int NW>
__global__ void __launch_bounds__(NW * WAVE) k_cascade_wg(XArgs A, const uint32_t* __restrict__ list,
                                                           unsigned* next_count) {
    __shared__ double s_ct[NW][kMaxPerm];
    __shared__ uint32_t s_cn[NW];
    __shared__ uint32_t s_tot[2][NW];
    __shared__ double s_sum[NW];
    const GraphDev& g = A.g;
    const TermBuf& tb = A.tb;
    const int lane = lane_id();
    const int wid = (int)threadIdx.x / WAVE;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        *next_count = 0;
        A.ctr->work_next = 0;
    }
    const uint32_t c = list[blockIdx.x];
    const uint64_t ebase = A.dpre[A.cstart[c]];
    uint32_t top = A.dtop[c];
    unsigned long long amb_c = 0, n_tri = 0, n_edges = 0;
    int par = 0;
    while (top > 0) {
        --top;
        const uint64_t ent = A.wl[ebase + top];
        const uint32_t e = (uint32_t)(ent >> 32), pa = (uint32_t)ent;
        const uint32_t x = g.esrc[e], sle = g.eslot[e], tc = g.tcnt[e];
        const uint64_t c0 = tb.zoff[e], c1 = c0 + tc;
        const uint32_t y = g.nb[sle];
        const uint32_t pb = pa == x ? y : x;
        const uint32_t dpa = g.deg[pa], dpb = g.deg[pb];
        const uint32_t dx = pa == x ? dpa : dpb, dy = pa == x ? dpb : dpa;
        const uint32_t tlo = dx <= dy ? x : y;
        const uint32_t rlo = dpa < dpb ? pa : pb;
        const uint32_t rhi = rlo == pa ? pb : pa;
        const bool swp = rlo != tlo;
        double hsum = 0.0;  // this lane's live triangle weights, for the endpoints
        const uint32_t nch = (tc + WAVE - 1) / WAVE;
        for (uint32_t k = 0; k < nch; k += NW) {  // rounds: the same count in every wave
            const uint64_t kk = c0 + (uint64_t)(k + wid) * WAVE + lane;
            const bool in = kk < c1;
            const Term q = term_or_blank(tb.t, kk, in);
            const uint32_t fj = q.sa, fh = q.sb;
            const uint8_t sj = g.est[fj], sh = g.est[fh];
            const uint32_t aj = g.esrc[fj], slj = g.eslot[fj], slh = g.eslot[fh];
            const double supj = g.sup[fj], suph = g.sup[fh];
            const double ewj = g.ew[fj], ewh = g.ew[fh];
            const bool xj = A.xact && A.xact[fj], xh = A.xact && A.xact[fh];
            const uint32_t tcj = g.tcnt[fj], tch = g.tcnt[fh];
            const double s0j = g.sup0[fj], s0h = g.sup0[fh];
            const uint32_t wj = g.nb[slj];
            const uint32_t spj = g.part[slj], sph = g.part[slh];
            const bool hit = in && sj != ST_DEAD && sh != ST_DEAD;
            const uint32_t w = aj == tlo ? wj : aj;  // third vertex
            const double t = q.w;
            n_tri += lane == 0 ? (unsigned long long)__popcll(ballot64(hit)) : 0ull;
            const uint32_t jl = swp ? fh : fj, jh = swp ? fj : fh;
            const uint8_t sl = swp ? sh : sj, shh = swp ? sj : sh;
            const double supl = swp ? suph : supj, suph2 = swp ? supj : suph;
            const double ewl = swp ? ewh : ewj, ewh2 = swp ? ewj : ewh;
            const bool xl = swp ? xh : xj, xh2 = swp ? xj : xh;
            const uint32_t tcl = swp ? tch : tcj, tch2 = swp ? tcj : tch;
            const double s0l = swp ? s0h : s0j, s0h2 = swp ? s0j : s0h;
            bool push1 = false, push2 = false, need1 = false, need2 = false;
            double v1 = 0.0, v2 = 0.0, thr1 = 0.0, thr2 = 0.0;
            if (hit) {
                hsum += t;
                g.pv[w] -= t;
                if (sl == ST_ALIVE) {
                    v1 = supl - t;
                    g.sup[jl] = v1;
                    g.touched[jl] = 1;
                    thr1 = A.eps * ewl;
                    push1 = decide_cheap_v(xl, tcl, s0l, v1, thr1, need1);
                }
                if (shh == ST_ALIVE) {
                    v2 = suph2 - t;
                    g.sup[jh] = v2;
                    g.touched[jh] = 1;
                    thr2 = A.eps * ewh2;
                    push2 = decide_cheap_v(xh2, tch2, s0h2, v2, thr2, need2);
                }
            }
#pragma unroll
            for (int side = 0; side < 2; ++side) {
                uint64_t nb = ballot64(side == 0 ? need1 : need2);
                while (nb) {  // wave-uniform
                    const int bl = __ffsll((unsigned long long)nb) - 1;
                    nb &= nb - 1;
                    const uint32_t f = __shfl(side == 0 ? jl : jh, bl, 64);
                    const double th = __shfl(side == 0 ? thr1 : thr2, bl, 64);
                    const int cc = certify_terms_wave(g, tb, f, th, false, e, s_ct[wid], &s_cn[wid]);
                    if (lane == bl) {
                        if (side == 0) push1 = settle(A, f, v1, th, cc, amb_c);
                        else push2 = settle(A, f, v2, th, cc, amb_c);
                    }
                }
            }
            const uint32_t np2 = (uint32_t)push1 + (uint32_t)push2;
            uint32_t pre = np2;  // inclusive prefix of pushes over the wave's lanes
            for (int d = 1; d < 64; d <<= 1) {
                const uint32_t o = __shfl_up(pre, d, 64);
                if (lane >= d) pre += o;
            }
            if (lane == 63) s_tot[par][wid] = pre;
            pre -= np2;
            __syncthreads();
            uint32_t base = 0, rtot = 0;
#pragma unroll
            for (int w2 = 0; w2 < NW; ++w2) {
                const uint32_t tw = s_tot[par][w2];
                base += w2 < wid ? tw : 0u;
                rtot += tw;
            }
            par ^= 1;
            const uint64_t e1 = ((uint64_t)jl << 32) | rlo, e2 = ((uint64_t)jh << 32) | rhi;
            if (push1) {
                A.wl[ebase + top + base + pre] = e1;
                set_status_at(g, jl, swp ? slh : slj, swp ? sph : spj, ST_QUEUED);
            }
            if (push2) {
                A.wl[ebase + top + base + pre + (uint32_t)push1] = e2;
                set_status_at(g, jh, swp ? slj : slh, swp ? spj : sph, ST_QUEUED);
            }
            top += rtot;
        }
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) hsum += __shfl_xor(hsum, d, 64);
        if (lane == 0) s_sum[wid] = hsum;
        __syncthreads();
        if (threadIdx.x == 0) {
            double tot = 0.0;
            for (int w2 = 0; w2 < NW; ++w2) tot += s_sum[w2];
            g.pv[x] -= tot;
            g.pv[y] -= tot;
            set_status_at(g, e, sle, g.part[sle], ST_DEAD);
            g.sup[e] = 0.0;
            n_edges++;
        }
        __syncthreads();  // the stack, statuses and supports for the next pop
    }
    for (int d = 32; d >= 1; d >>= 1) amb_c += __shfl_xor(amb_c, d, 64);
    if (lane == 0) {
        if (n_tri) atomicAdd(&A.ctr->tri_removed, n_tri);
        if (n_edges) atomicAdd(&A.ctr->edges_removed, n_edges);
        if (amb_c) atomicAdd(&A.ctr->ambiguous_clean, amb_c);
    }
}

// ---------------------------------------------------------------- per seed
// One seed's inner loop on the current state (rtrex_extract_inner): the alive
// neighbours of u marked with stamp in mk (Decomposition.h:620-631), the
// pairs (seed_pairs), the candidates in ascending id with their weights
// (seed_candidates) into rec, and out = {internal, potential}, nout = the
// candidate count.  One wave.
struct InnerArgs {
    GraphDev g;
    uint32_t u, stamp;
    double eps;
    uint32_t *mk, *cm;  // n each, stamps (never reset: stamps grow per call)
    double* tw;         // n
    uint32_t* nbr;      // deg(u): the alive neighbours
    uint32_t* cand;     // 2 n: candidates, then sort keys when LDS is short
    Rec* rec;           // n: the result
    double* out;        // internal, potential
    uint32_t* nout;
};
__global__ void __launch_bounds__(XT) k_extract_inner(InnerArgs A) {
    __shared__ Rec srec[LDS_CAND];
    __shared__ uint32_t skey[LDS_CAND];
    __shared__ uint32_t s_hn[XSTAGE];
    __shared__ uint8_t s_ha[XSTAGE];
    const GraphDev& g = A.g;
    const int lane = lane_id();
    const uint32_t u = A.u, du = g.deg[u];
    uint32_t nA = 0;
    for (uint32_t p = g.off[u]; p < g.off[u + 1]; p += XT) {
        const uint32_t j = p + lane;
        const bool alive = j < g.off[u + 1] && g.st[j] != ST_DEAD;
        const uint32_t v = alive ? g.nb[j] : 0u;
        const uint64_t bal = ballot64(alive);
        if (alive) {
            A.nbr[nA + (uint32_t)__popcll(bal & lanes_below())] = v;
            A.mk[v] = A.stamp;
        }
        nA += (uint32_t)__popcll(bal);
    }
    __syncthreads();
    double internal = 0.0;
    const uint32_t nc = seed_pairs(g, u, du, (double)du / A.eps, A.stamp, A.nbr, nA, A.mk, A.cm,
                                   A.tw, A.cand, s_hn, s_ha, internal);
    uint32_t np = 1;
    while (np < nc) np <<= 1;
    const bool small = np <= LDS_CAND;
    Rec* rec = small ? srec : A.rec;
    const double potential = seed_candidates(A.cand, nc, small ? skey : A.cand, A.tw, rec);
    __syncthreads();
    if (small)
        for (uint32_t i = lane; i < nc; i += XT) A.rec[i] = srec[i];
    if (lane == 0) {
        A.out[0] = internal;
        A.out[1] = potential;
        A.nout[0] = nc;
    }
}

// The pairs deleting a cluster's edges (rtrex_cluster_edges): members in the
// given order, each one's alive slots in order, an edge listed from its first
// member only (pos: member index per vertex, NONE elsewhere) -- the push order
// of Decomposition.h:897-914, where the first listing marks both slots 'S'.
// One wave; out[2i], out[2i + 1] = (member, neighbour) for i < cap; *nout =
// the number of pairs.
__global__ void __launch_bounds__(XT) k_cluster_edges(GraphDev g, const uint32_t* __restrict__ mem,
                                                      uint32_t nm, const uint32_t* __restrict__ pos,
                                                      uint32_t* __restrict__ out, uint64_t cap,
                                                      unsigned long long* nout) {
    const int lane = lane_id();
    uint64_t c = 0;
    for (uint32_t q = 0; q < nm; ++q) {
        const uint32_t z = mem[q];
        const uint32_t zb = g.off[z], ze = g.off[z + 1];
        for (uint32_t p = zb; p < ze; p += XT) {
            const uint32_t j = p + lane;
            const bool in = j < ze;
            const uint32_t y = in ? g.nb[j] : 0u;
            const bool take = in && g.st[j] == ST_ALIVE && !(pos[y] < q);
            const uint64_t bal = ballot64(take);
            const uint64_t i = c + (uint64_t)__popcll(bal & lanes_below());
            if (take && i < cap) {
                out[2 * i] = z;
                out[2 * i + 1] = y;
            }
            c += (uint64_t)__popcll(bal);
        }
    }
    if (lane == 0) *nout = c;
}

__global__ void __launch_bounds__(T) k_set_pos(const uint32_t* __restrict__ mem, uint32_t nm,
                                               uint32_t* __restrict__ pos, bool clear) {
    for (uint32_t q = blockIdx.x * T + threadIdx.x; q < nm; q += gridDim.x * T)
        pos[mem[q]] = clear ? NONE : q;
}

}  // namespace
}  // namespace rtrex

namespace rtrex {
namespace {

// Cluster ids: seeds numbered in seed order (Decomposition.h:612,892-893).
__global__ void __launch_bounds__(T) k_seed_flags(uint32_t n,
                                                  const uint32_t* __restrict__ order,
                                                  const uint32_t* __restrict__ cseed,
                                                  uint32_t* __restrict__ flag) {
    for (uint32_t r = blockIdx.x * T + threadIdx.x; r < n; r += gridDim.x * T) {
        const uint32_t v = order[r];
        flag[r] = cseed[v] == v ? 1u : 0u;
    }
}

__global__ void __launch_bounds__(T) k_gather_u32(const uint32_t* __restrict__ src,
                                                  const uint32_t* __restrict__ idx, uint32_t k,
                                                  uint32_t* __restrict__ out) {
    for (uint32_t i = blockIdx.x * T + threadIdx.x; i < k; i += gridDim.x * T) out[i] = src[idx[i]];
}

__global__ void __launch_bounds__(T) k_cluster_ids(uint32_t n,
                                                   const uint32_t* __restrict__ cseed,
                                                   const uint32_t* __restrict__ rank,
                                                   const uint32_t* __restrict__ cidr,
                                                   uint32_t* __restrict__ cid,
                                                   unsigned* err) {
    for (uint32_t v = blockIdx.x * T + threadIdx.x; v < n; v += gridDim.x * T) {
        const uint32_t sd = cseed[v];
        if (sd >= n) {  // every vertex must have been assigned a cluster
            cid[v] = NONE;
            atomicOr(err, ERR_SCRATCH);
            continue;
        }
        cid[v] = cidr[rank[sd]];
    }
}

// populateStats over the original graph (Decomposition.h:473-491).
__global__ void __launch_bounds__(T) k_cluster_stats(const GraphDev g,
                                                     const uint32_t* __restrict__ cid,
                                                     unsigned long long* nv,
                                                     unsigned long long* ne,
                                                     unsigned long long* cut) {
    for (uint32_t z = blockIdx.x * T + threadIdx.x; z < g.n; z += gridDim.x * T) {
        const uint32_t c = cid[z];
        unsigned long long in = 0, out = 0;
        for (uint32_t i = g.off[z]; i < g.off[z + 1]; ++i) {
            if (cid[g.nb[i]] == c) in++;
            else out++;
        }
        atomicAdd(&nv[c], 1ull);
        if (in) atomicAdd(&ne[c], in);
        if (out) atomicAdd(&cut[c], out);
    }
}

// ---- connectedComp over the original graph (Decomposition.h:1023-1068)
__global__ void __launch_bounds__(T) k_cc_init(uint32_t n, uint32_t* parent) {
    for (uint32_t v = blockIdx.x * T + threadIdx.x; v < n; v += gridDim.x * T) parent[v] = v;
}

// Every edge joins its endpoints' trees, the larger root under the smaller,
// so each component's root is its smallest vertex.
__global__ void __launch_bounds__(T) k_cc_hook(const GraphDev g, uint32_t* parent) {
    for (uint32_t e = blockIdx.x * T + threadIdx.x; e < g.m; e += gridDim.x * T) {
        uint32_t a = uf_find(parent, g.esrc[e]);
        uint32_t b = uf_find(parent, g.nb[g.eslot[e]]);
        while (a != b) {
            if (a > b) {
                const uint32_t t = a;
                a = b;
                b = t;
            }
            const uint32_t old = atomicCAS(&parent[b], b, a);
            if (old == b) break;
            b = uf_find(parent, old);
            a = uf_find(parent, a);
        }
    }
}

// parent[] := root (roots only written, as in k_uf_compress); isroot flags.
__global__ void __launch_bounds__(T) k_cc_roots(uint32_t n, uint32_t* parent,
                                               uint32_t* __restrict__ isroot) {
    for (uint32_t v = blockIdx.x * T + threadIdx.x; v < n; v += gridDim.x * T) {
        uint32_t r = v, p = parent[v];
        while (p != r) {
            r = p;
            p = parent[r];
        }
        parent[v] = r;
        isroot[v] = r == v ? 1u : 0u;
    }
}

// The reference numbers components as its BFS meets them: by smallest
// vertex, i.e. the rank of the root among the roots.
__global__ void __launch_bounds__(T) k_cc_label(uint32_t n, const uint32_t* __restrict__ parent,
                                               const uint32_t* __restrict__ rk,
                                               int64_t* __restrict__ out) {
    for (uint32_t v = blockIdx.x * T + threadIdx.x; v < n; v += gridDim.x * T)
        out[v] = rk[parent[v]];
}

int bits_for(uint64_t n) {
    int b = 1;
    while ((1ull << b) < n) ++b;
    return b;
}

}  // namespace

// unordered_set<int64> bucket counts after k insertions, measured once per
// process from the library (the iteration order the extraction emulates
// depends on them) and grown on demand by further insertions, with the
// largest ratio bcnt[k] / k over k >= 2.
struct BucketTable {
    std::mutex mu;
    std::unordered_set<int64_t> probe;
    std::vector<uint32_t> bcnt;
    uint32_t ratio = 1;
};
BucketTable& bucket_table() {
    static BucketTable t;
    return t;
}

// The handle's copy of bcnt[0..upto] (its own buffer, stable while an
// upload reads it), its device copy, and the ratio (which bounds the set
// scratch of components).
rtrex_status ensure_bucket_table(rtrex_graph h, uint64_t upto) {
    if (h->bcnt.size() < upto + 1) {
        BucketTable& t = bucket_table();
        std::lock_guard<std::mutex> lk(t.mu);
        if (t.bcnt.empty()) t.bcnt.push_back((uint32_t)t.probe.bucket_count());
        for (uint64_t k = t.bcnt.size(); k <= upto; ++k) {
            t.probe.insert((int64_t)k);
            const uint32_t b = (uint32_t)t.probe.bucket_count();
            t.bcnt.push_back(b);
            if (k >= 2) t.ratio = std::max<uint32_t>(t.ratio, (uint32_t)((b + k - 1) / k));
        }
        h->bcnt.assign(t.bcnt.begin(), t.bcnt.begin() + (upto + 1));
        h->bcnt_ratio = t.ratio;
    }
    if (!h->bcnt_dev) {
        RT_TRY(dalloc(h, &h->bcnt_dev, (uint64_t)h->g.n + 1));
        h->bcnt_dev_len = 0;
    }
    if (h->bcnt_dev_len < upto + 1) {
        // the host table stays unchanged until a later extraction, which
        // begins after this one has synchronised
        RT_HIP(hipMemcpyAsync(h->bcnt_dev, h->bcnt.data(), (upto + 1) * 4,
                              hipMemcpyHostToDevice, h->stream));
        h->bcnt_dev_len = upto + 1;
    }
    return RTREX_OK;
}

rtrex_status run_extract(rtrex_graph h, double eps) {
    if (h->phase < PH_CLEANED)
        return fail(RTREX_ESTATE, "extract needs count and initial clean first");
    if (h->phase >= PH_EXTRACTED)
        return fail(RTREX_ESTATE, "graph already extracted; run count again");
    if (!(eps > 0)) return fail(RTREX_EINVAL, "eps must be > 0");
    GraphDev& g = h->g;
    hipStream_t s = h->stream;
    const uint32_t n = g.n;
    RT_HIP(hipEventRecord(h->ev[7], s));
    Scratch sc;
    uint64_t *keys, *tmp;
    uint32_t *parent, *csize, *vseg = nullptr;
    RT_TRY(sc.get(&keys, n));
    RT_TRY(sc.get(&tmp, n));
    RT_TRY(sc.get(&parent, n));
    RT_TRY(sc.get(&csize, n));
    const int Bi = bits_for((uint64_t)n + 1);
    int Bs = 0;
    if (!h->seg_vtx.empty() && !h->seed_order) {
        const size_t ns = h->seg_vtx.size() - 1;
        Bs = bits_for(ns + 1);
        if (!h->vseg_d) {  // each vertex's segment, kept for the per-graph ids
            std::vector<uint32_t> seg(n);
            for (size_t q = 0; q < ns; ++q)
                for (int64_t v = h->seg_vtx[q]; v < h->seg_vtx[q + 1]; ++v) seg[v] = (uint32_t)q;
            RT_TRY(dalloc(h, &h->vseg_d, n));
            RT_HIP(hipMemcpyAsync(h->vseg_d, seg.data(), (uint64_t)n * 4, hipMemcpyHostToDevice,
                                  s));
            RT_HIP(hipStreamSynchronize(s));
        }
        vseg = h->vseg_d;
    }
    if (Bs + 2 * Bi > 64) return fail(RTREX_EINVAL, "too many vertices/segments");
    const unsigned bn = blocks_for(n, T, 8192);
    // seed order: (segment, degree, id) -- pairCompareSecond (Graph.cpp:40-45);
    // static for the graph, so computed once per handle.
    if (!h->seed_order) {
        RT_TRY(dalloc(h, &h->seed_order, n));
        RT_TRY(dalloc(h, &h->seed_rank, n));
        k_seed_keys<<<bn, T, 0, s>>>(g, vseg, Bi, Bi, keys);
        RT_HIP(hipGetLastError());
        RT_TRY(radix_sort_u64(keys, tmp, n, Bs + 2 * Bi, s));
        k_ranks<<<bn, T, 0, s>>>(keys, n, Bi, h->seed_order, h->seed_rank);
    }
    uint32_t* order = h->seed_order;
    uint32_t* rank = h->seed_rank;
    // alive components
    uint32_t *mk, *cm, *big_d;
    RT_TRY(sc.get(&mk, n));
    RT_TRY(sc.get(&cm, n));
    RT_TRY(sc.get(&big_d, 1));
    unsigned* nk_d = &h->ctr->frontier_next;
    k_uf_init<<<bn, T, 0, s>>>(n, parent, csize, h->cseed, mk, cm, big_d, nk_d, h->ctr);
    k_uf_hook<<<blocks_for(g.m, T, 8192), T, 0, s>>>(g, parent);
    k_uf_compress<<<bn, T, 0, s>>>(n, parent, csize);
    if (!h->mrank) RT_TRY(dalloc(h, &h->mrank, n));
    k_comp_keys<<<bn, T, 0, s>>>(n, parent, csize, rank, h->cseed, h->mrank, keys, nk_d);
    RT_TRY(box_open(h));
    k_publish<<<1, WAVE, 0, s>>>(h->box_dev, ++h->box_seq, nk_d);
    RT_HIP(hipGetLastError());
    uint32_t nk = 0;
    RT_TRY(box_wait(h, 0, h->box_seq, &nk));
    h->audit.components = 0;
    h->audit.largest_component = nk ? 0 : (n ? 1 : 0);
    uint32_t* x_undec_list = nullptr;  // device list of undecided cascade decisions
    const uint32_t* ncomp_d = nullptr;  // component count (device)
    if (nk) {
        uint32_t *cv, *cstart, *fpos, *hoff, *hnext, *hbkt;
        uint64_t* dpre;
        RT_TRY(sc.get(&cv, nk));
        RT_TRY(sc.get(&fpos, (uint64_t)nk + 1));
        RT_TRY(sc.get(&cstart, (uint64_t)nk + 1));
        RT_TRY(sc.get(&dpre, (uint64_t)nk + 1));
        // Bucket counts of the reference's unordered_set (measured from the
        // library) for up to nk elements, and each component's share of the
        // set scratch: hoff = prefix of bcnt[size] (at most ratio * nk).
        RT_TRY(ensure_bucket_table(h, nk));
        RT_TRY(sc.get(&hoff, (uint64_t)nk + 1));
        RT_TRY(sc.get(&hnext, nk));
        RT_TRY(sc.get(&hbkt, (uint64_t)h->bcnt_ratio * nk + 1));
        // The component count stays on the device (fpos[nk]); everything
        // below is sized from nk (components have two vertices or more).
        ncomp_d = fpos + nk;
        if (nk <= SB_MAX) {  // one workgroup: sort, layout and the three scans
            k_comp_small<<<1, SB_T, 0, s>>>(keys, nk, order, g, h->bcnt_dev, cv, cstart,
                                            fpos + nk, dpre, hoff, big_d);
            RT_HIP(hipGetLastError());
        } else {
            RT_TRY(radix_sort_u64(keys, tmp, nk, 32 + Bi, s));
            uint32_t *vdeg, *flag, *hsz;
            RT_TRY(sc.get(&vdeg, nk));
            RT_TRY(sc.get(&flag, nk));
            RT_TRY(sc.get(&hsz, nk));
            const unsigned bk = blocks_for(nk, T, 8192);
            k_comp_layout<<<bk, T, 0, s>>>(keys, nk, order, g, cv, vdeg, flag);
            RT_HIP(hipGetLastError());
            RT_TRY(exclusive_scan_u32_32(flag, fpos, nk, s));
            k_comp_starts<<<bk, T, 0, s>>>(flag, fpos, nk, cstart);
            RT_HIP(hipGetLastError());
            RT_TRY(exclusive_scan_u32(vdeg, dpre, nk, s));
            k_comp_hsize<<<bk, T, 0, s>>>(cstart, ncomp_d, nk, h->bcnt_dev, hsz, big_d);
            RT_HIP(hipGetLastError());
            RT_TRY(exclusive_scan_u32_32(hsz, hoff, nk, s));
        }
        RT_TRY(ensure_vertex_supports(h, cv, nk));  // candidates' supports
        uint32_t *mem, *cand, *undec;
        uint64_t* wl;
        double *tw, *gpv;
        Rec* grec;
        RT_TRY(sc.get(&mem, n));
        RT_TRY(sc.get(&cand, 2ull * n));
        RT_TRY(sc.get(&tw, n));
        RT_TRY(sc.get(&gpv, n));
        RT_TRY(sc.get(&grec, n));
        RT_TRY(sc.get(&wl, g.nnz));
        RT_TRY(sc.get(&undec, g.m));
        XArgs a;
        a.g = g;
        a.eps = eps;
        a.cv = cv;
        a.cstart = cstart;
        a.dpre = dpre;
        a.ncomp = ncomp_d;
        a.cseed = h->cseed;
        a.mrank = h->mrank;
        a.mk = mk;
        a.cm = cm;
        a.tw = tw;
        a.mem = mem;
        a.cand = cand;
        a.grec = grec;
        a.gpv = gpv;
        a.wl = wl;
        a.tb = h->terms();
        a.ctr = h->ctr;
        a.xact = h->xact;
        a.bcnt = h->bcnt_dev;
        a.uset_lds = h->bcnt[std::min<uint64_t>(US_LDS, nk)] <= US_LDS_BKT;
        a.hoff = hoff;
        a.hnext = hnext;
        a.hbkt = hbkt;
        a.undec = undec;
        a.undec_cap = g.m;
        const unsigned grid = std::min<uint32_t>(nk / 2, 8192);
        // Long cascades go to a workgroup each (k_cascade_wg): listed by
        // k_extract, finished, then their components' seed loops resumed,
        // until no launch hands one over.  Components have two vertices or
        // more, so at most nk / 2 are listed per launch.
        const bool handoff = h->cascade_min && a.tb.t;
        const uint32_t dcap = nk / 2 + 1;
        uint32_t *resume = nullptr, *dtop = nullptr, *dlist = nullptr;
        unsigned* dcount = nullptr;
        if (handoff) {
            RT_TRY(sc.get(&resume, nk));
            RT_TRY(sc.get(&dtop, nk));
            RT_TRY(sc.get(&dlist, 2ull * dcap));
            RT_TRY(sc.get(&dcount, 2));
            RT_HIP(hipMemsetAsync(resume, 0, (uint64_t)nk * 4, s));
            RT_HIP(hipMemsetAsync(dcount, 0, 2 * sizeof(unsigned), s));
            a.cascade_min = h->cascade_min;
            a.resume = resume;
            a.dtop = dtop;
            a.dlist = dlist;
            a.ndefer = dcount;
        }
        if (a.tb.t) k_extract<true><<<grid, XT, 0, s>>>(a);
        else k_extract<false><<<grid, XT, 0, s>>>(a);
        RT_HIP(hipGetLastError());
        for (int cur = 0; handoff; cur ^= 1) {
            k_publish<<<1, WAVE, 0, s>>>(h->box_dev, ++h->box_seq, dcount + cur);
            RT_HIP(hipGetLastError());
            uint32_t nd = 0;
            RT_TRY(box_wait(h, 0, h->box_seq, &nd));
            if (!nd) break;
            if (nd > dcap) return fail(RTREX_EDEVICE, "cascade hand-over list overflow");
            const uint32_t* lst = dlist + (uint64_t)cur * dcap;
            k_cascade_wg<CWG><<<nd, CWG * WAVE, 0, s>>>(a, lst, dcount + (cur ^ 1));
            XArgs b = a;
            b.clist = lst;
            b.nclist = dcount + cur;
            b.dlist = dlist + (uint64_t)(cur ^ 1) * dcap;
            b.ndefer = dcount + (cur ^ 1);
            k_extract<true><<<std::min<uint32_t>(nd, grid), XT, 0, s>>>(b);
            RT_HIP(hipGetLastError());
        }
        // Undecided cascade decisions (inexact post-clean supports): kept for
        // an exact re-run of their regions (rtrex_extract); their count is
        // read with the phase's last results below.
        x_undec_list = undec;
    }
    // cluster ids in seed order; they stay in HBM (copy_cluster_map)
    uint32_t *flag2, *cidr;
    RT_TRY(sc.get(&flag2, n));
    RT_TRY(sc.get(&cidr, (uint64_t)n + 1));
    if (!h->cid) RT_TRY(dalloc(h, &h->cid, n));
    k_seed_flags<<<bn, T, 0, s>>>(n, order, h->cseed, flag2);
    RT_TRY(exclusive_scan_u32_32(flag2, cidr, n, s));
    k_cluster_ids<<<bn, T, 0, s>>>(n, h->cseed, rank, cidr, h->cid, &h->ctr->error);
    RT_HIP(hipGetLastError());
    const size_t ns = h->seg_vtx.empty() ? 0 : h->seg_vtx.size() - 1;
    uint32_t* bounds_d = nullptr;
    if (ns) {
        // Ranks of segment q start at seg_vtx[q]: its clusters are
        // [cidr[seg_vtx[q]], cidr[seg_vtx[q + 1]]).
        if (!h->seg_pos_d) {
            std::vector<uint32_t> sp(h->seg_vtx.begin(), h->seg_vtx.end());
            RT_TRY(dalloc(h, &h->seg_pos_d, ns + 1));
            RT_HIP(hipMemcpyAsync(h->seg_pos_d, sp.data(), (ns + 1) * 4, hipMemcpyHostToDevice, s));
            RT_HIP(hipStreamSynchronize(s));
        }
        if (!h->seg_bounds_d) RT_TRY(dalloc(h, &h->seg_bounds_d, ns + 1));
        bounds_d = h->seg_bounds_d;
        k_gather_u32<<<blocks_for(ns + 1, T, 1024), T, 0, s>>>(cidr, h->seg_pos_d,
                                                             (uint32_t)(ns + 1), bounds_d);
        RT_HIP(hipGetLastError());
    }
    // the phase's scalar results, published to the host mailbox
    RT_TRY(box_open(h));
    const uint32_t fseq = ++h->box_seq;
    k_final_scalars<<<1, WAVE, 0, s>>>(ncomp_d, ncomp_d ? big_d : nullptr,
                                       x_undec_list ? &h->ctr->n_undecided : nullptr,
                                       &h->ctr->error, cidr + n, h->box_dev, fseq);
    RT_HIP(hipGetLastError());
    if (ns && !h->sb_pin) {  // page-locked: a pageable landing takes the slow staged path
        h->sb_pin = static_cast<uint32_t*>(pin_alloc((ns + 1) * 4));
        if (!h->sb_pin) return fail(RTREX_ENOMEM, "page-locked staging allocation failed");
    }
    if (ns)
        RT_HIP(hipMemcpyAsync(h->sb_pin, bounds_d, (ns + 1) * 4, hipMemcpyDeviceToHost, s));
    RT_HIP(hipEventRecord(h->ev[8], s));
    uint32_t hf[5] = {0, 0, 0, 0, 0};
    for (int i = 0; i < 5; ++i) RT_TRY(box_wait(h, 1 + i, fseq, &hf[i]));
    if (ns) RT_HIP(hipEventSynchronize(h->ev[8]));  // the segment bounds
    h->seg_bounds.assign(h->sb_pin, h->sb_pin + (ns ? ns + 1 : 0));
    const uint32_t ncomp = hf[0], big = hf[1], nu = hf[2], cerr = hf[3], total = hf[4];
    if (cerr) return fail(RTREX_EDEVICE, "extraction left a vertex without a cluster");
    if (ncomp_d) {
        h->audit.components = ncomp;
        h->audit.largest_component = big;
    }
    h->x_undec.resize(std::min<uint32_t>(nu, g.m));
    h->x_undec_all = nu > g.m;
    if (nu) {
        RT_HIP(hipMemcpyAsync(h->x_undec.data(), x_undec_list, h->x_undec.size() * 4ull,
                              hipMemcpyDeviceToHost, s));
        RT_HIP(hipStreamSynchronize(s));
    }
    h->tm_extract_pending = true;  // read later (resolve_timings), no wait here
    h->n_clusters = total;
    h->seg_clusters.assign(ns, 0);
    for (size_t q = 0; q < ns; ++q)
        h->seg_clusters[q] = (int64_t)h->seg_bounds[q + 1] - h->seg_bounds[q];
    h->phase = PH_EXTRACTED;
    return RTREX_OK;
}

namespace {
// Cluster ids widened to int64 straight into page-locked host memory that
// the device can address (rtrex_host_alloc): no staging copy, no host loop.
__global__ void __launch_bounds__(T) k_cid_out(const uint32_t* __restrict__ cid, uint32_t n,
                                              int64_t* __restrict__ out) {
    for (uint32_t v = blockIdx.x * T + threadIdx.x; v < n; v += gridDim.x * T) out[v] = cid[v];
}
// The same for a batch: per-graph ids, each vertex's id less its graph's
// first cluster id (vseg: the vertex's graph, bounds: first ids).
__global__ void __launch_bounds__(T) k_cid_out_seg(const uint32_t* __restrict__ cid, uint32_t n,
                                                  const uint32_t* __restrict__ vseg,
                                                  const uint32_t* __restrict__ bounds,
                                                  int64_t* __restrict__ out) {
    for (uint32_t v = blockIdx.x * T + threadIdx.x; v < n; v += gridDim.x * T)
        out[v] = (int64_t)cid[v] - (int64_t)bounds[vseg[v]];
}
}  // namespace

bool mapped_host(const void* p, void** dptr) {
    hipPointerAttribute_t a;
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();  // ordinary pageable memory is not known to HIP
        return false;
    }
    if (a.type != hipMemoryTypeHost || !a.devicePointer) return false;
    *dptr = a.devicePointer;
    return true;
}

rtrex_status copy_cluster_map(rtrex_graph h, int64_t* out) {
    if (h->phase < PH_EXTRACTED || !h->cid) return fail(RTREX_ESTATE, "extract has not run");
    const uint32_t n = h->g.n;
    if (!n) return RTREX_OK;
    void* dout = nullptr;
    if ((h->seg_bounds.empty() || h->vseg_d) && mapped_host(out, &dout)) {
        if (h->seg_bounds.empty())
            k_cid_out<<<blocks_for(n, T, 4096), T, 0, h->stream>>>(h->cid, n,
                                                                   static_cast<int64_t*>(dout));
        else
            k_cid_out_seg<<<blocks_for(n, T, 4096), T, 0, h->stream>>>(
                h->cid, n, h->vseg_d, h->seg_bounds_d, static_cast<int64_t*>(dout));
        RT_HIP(hipGetLastError());
        RT_HIP(hipStreamSynchronize(h->stream));
        return RTREX_OK;
    }
    if (!h->cm_pin) h->cm_pin = static_cast<uint32_t*>(pin_alloc((uint64_t)n * 4));
    if (!h->cm_pin) return fail(RTREX_ENOMEM, "page-locked staging allocation failed");
    RT_HIP(hipMemcpyAsync(h->cm_pin, h->cid, (uint64_t)n * 4, hipMemcpyDeviceToHost, h->stream));
    RT_HIP(hipStreamSynchronize(h->stream));
    const uint32_t* cm = h->cm_pin;
    const size_t ns = h->seg_bounds.empty() ? 0 : h->seg_bounds.size() - 1;
    if (!ns) {
        host_parallel(n, 1u << 20, [cm, out](uint64_t a, uint64_t b) {
            for (uint64_t v = a; v < b; ++v) out[v] = cm[v];
        });
        return RTREX_OK;
    }
    // Per-graph ids: graph q's vertices [seg_vtx[q], seg_vtx[q+1]) hold ids
    // from seg_bounds[q] on.
    const std::vector<int64_t>& sv = h->seg_vtx;
    const std::vector<uint32_t>& sb = h->seg_bounds;
    host_parallel(n, 1u << 20, [cm, out, ns, &sv, &sb](uint64_t a, uint64_t b) {
        size_t q = (size_t)(std::upper_bound(sv.begin(), sv.begin() + ns + 1, (int64_t)a) -
                            sv.begin()) - 1;
        for (uint64_t v = a; v < b; ++v) {
            while ((int64_t)v >= sv[q + 1]) ++q;
            out[v] = (int64_t)cm[v] - (int64_t)sb[q];
        }
    });
    return RTREX_OK;
}

rtrex_status run_cluster_stats(rtrex_graph h, int64_t* nv, int64_t* ne,
                               int64_t* cut) {
    if (h->phase < PH_EXTRACTED || !h->cid)
        return fail(RTREX_ESTATE, "extract has not run");
    const uint64_t k = (uint64_t)h->n_clusters;
    hipStream_t s = h->stream;
    Scratch sc;
    unsigned long long *a, *b, *c;
    RT_TRY(sc.get(&a, k));
    RT_TRY(sc.get(&b, k));
    RT_TRY(sc.get(&c, k));
    RT_HIP(hipMemsetAsync(a, 0, k * 8, s));
    RT_HIP(hipMemsetAsync(b, 0, k * 8, s));
    RT_HIP(hipMemsetAsync(c, 0, k * 8, s));
    k_cluster_stats<<<blocks_for(h->g.n, T, 8192), T, 0, s>>>(h->g, h->cid, a, b, c);
    RT_HIP(hipGetLastError());
    std::vector<unsigned long long> ha(k), hb(k), hc(k);
    RT_HIP(hipMemcpyAsync(ha.data(), a, k * 8, hipMemcpyDeviceToHost, s));
    RT_HIP(hipMemcpyAsync(hb.data(), b, k * 8, hipMemcpyDeviceToHost, s));
    RT_HIP(hipMemcpyAsync(hc.data(), c, k * 8, hipMemcpyDeviceToHost, s));
    RT_HIP(hipStreamSynchronize(s));
    for (uint64_t i = 0; i < k; ++i) {
        if (nv) nv[i] = (int64_t)ha[i];
        if (ne) ne[i] = (int64_t)(hb[i] / 2);
        if (cut) cut[i] = (int64_t)hc[i];
    }
    return RTREX_OK;
}

rtrex_status run_connected_components(rtrex_graph h, int64_t* comp, int64_t* n_comp) {
    if (h->phase < PH_BUILT) return fail(RTREX_ESTATE, "graph not built");
    const GraphDev& g = h->g;
    const uint32_t n = g.n;
    hipStream_t s = h->stream;
    if (!n) {
        if (n_comp) *n_comp = 0;
        return RTREX_OK;
    }
    Scratch sc;
    uint32_t *parent, *isroot, *rk;
    int64_t* lab;
    RT_TRY(sc.get(&parent, n));
    RT_TRY(sc.get(&isroot, n));
    RT_TRY(sc.get(&rk, (uint64_t)n + 1));
    RT_TRY(sc.get(&lab, n));
    const unsigned bn = blocks_for(n, T, 8192);
    k_cc_init<<<bn, T, 0, s>>>(n, parent);
    k_cc_hook<<<blocks_for(g.m, T, 8192), T, 0, s>>>(g, parent);
    k_cc_roots<<<bn, T, 0, s>>>(n, parent, isroot);
    RT_HIP(hipGetLastError());
    RT_TRY(exclusive_scan_u32_32(isroot, rk, n, s));
    void* dout = nullptr;
    const bool direct = comp && mapped_host(comp, &dout);
    k_cc_label<<<bn, T, 0, s>>>(n, parent, rk, direct ? static_cast<int64_t*>(dout) : lab);
    RT_HIP(hipGetLastError());
    uint32_t nc = 0;
    RT_HIP(hipMemcpyAsync(&nc, rk + n, 4, hipMemcpyDeviceToHost, s));
    if (comp && !direct)
        RT_HIP(hipMemcpyAsync(comp, lab, (uint64_t)n * 8, hipMemcpyDeviceToHost, s));
    RT_HIP(hipStreamSynchronize(s));
    if (n_comp) *n_comp = nc;
    return RTREX_OK;
}

}  // namespace rtrex

// ------------------------------------------------- statistics, no allocation
// rtrex_cluster_stats_noalloc keeps the three per-cluster accumulators of
// k_cluster_stats in this array instead of the handle's scratch arena, so the
// call allocates nothing.  Each GPU holds its own copy for the life of the
// process (24 MiB); every call on one GPU shares it, and a call leaves its
// results in it, so each call clears what it uses first, and calls that
// overlap on one GPU race.
static constexpr uint64_t STATS_MAX_CLUSTERS = 1ull << 20;
static __device__ __align__(256) unsigned long long scratch[3 * STATS_MAX_CLUSTERS];

namespace rtrex {
namespace {

constexpr int MAX_DEVICES = 64;
std::mutex g_stats_scratch_mu;
unsigned long long* g_stats_scratch[MAX_DEVICES] = {};

// The device address of `scratch` on the current device (looked up once per
// device; host code never uses the array's name as an address).
rtrex_status stats_scratch(int device, unsigned long long** out) {
    if (device < 0 || device >= MAX_DEVICES)
        return fail(RTREX_EINVAL, "device index beyond the allocation-free path's table");
    std::lock_guard<std::mutex> lk(g_stats_scratch_mu);
    if (!g_stats_scratch[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return hip_fail(e, "hipGetSymbolAddress(scratch)", __FILE__,
                                                   __LINE__);
        g_stats_scratch[device] = static_cast<unsigned long long*>(p);
    }
    *out = g_stats_scratch[device];
    return RTREX_OK;
}

__global__ void __launch_bounds__(T) k_zero_u64(unsigned long long* p, uint64_t count) {
    for (uint64_t i = blockIdx.x * (uint64_t)T + threadIdx.x; i < count;
         i += (uint64_t)gridDim.x * T)
        p[i] = 0ull;
}

// Internal edges were counted from both endpoints (populateStats halves them,
// Decomposition.h:489).
__global__ void __launch_bounds__(T) k_halve_u64(unsigned long long* p, uint64_t count) {
    for (uint64_t i = blockIdx.x * (uint64_t)T + threadIdx.x; i < count;
         i += (uint64_t)gridDim.x * T)
        p[i] >>= 1;
}

}  // namespace

rtrex_status run_cluster_stats_noalloc(rtrex_graph h, int64_t* nv, int64_t* ne, int64_t* cut) {
    if (h->phase < PH_EXTRACTED || !h->cid)
        return fail(RTREX_ESTATE, "extract has not run");
    const uint64_t k = (uint64_t)h->n_clusters;
    if (k > STATS_MAX_CLUSTERS)
        return fail(RTREX_EINVAL, "the graph has " + std::to_string(k) +
                                      " clusters; rtrex_cluster_stats_noalloc holds at most " +
                                      std::to_string(STATS_MAX_CLUSTERS) +
                                      " (rtrex_cluster_stats has no limit)");
    if (k == 0) return RTREX_OK;
    unsigned long long* base = nullptr;
    RT_TRY(stats_scratch(h->device, &base));
    unsigned long long *a = base, *b = base + STATS_MAX_CLUSTERS, *c = base + 2 * STATS_MAX_CLUSTERS;
    hipStream_t s = h->stream;
    // the last call's results are still there: clear the three ranges used
    k_zero_u64<<<blocks_for(k, T, 2048), T, 0, s>>>(a, k);
    k_zero_u64<<<blocks_for(k, T, 2048), T, 0, s>>>(b, k);
    k_zero_u64<<<blocks_for(k, T, 2048), T, 0, s>>>(c, k);
    k_cluster_stats<<<blocks_for(h->g.n, T, 8192), T, 0, s>>>(h->g, h->cid, a, b, c);
    k_halve_u64<<<blocks_for(k, T, 2048), T, 0, s>>>(b, k);
    RT_HIP(hipGetLastError());
    // counts are below 2^63: the u64 words are the int64 results
    if (nv) RT_HIP(hipMemcpyAsync(nv, a, k * 8, hipMemcpyDeviceToHost, s));
    if (ne) RT_HIP(hipMemcpyAsync(ne, b, k * 8, hipMemcpyDeviceToHost, s));
    if (cut) RT_HIP(hipMemcpyAsync(cut, c, k * 8, hipMemcpyDeviceToHost, s));
    RT_HIP(hipStreamSynchronize(s));
    return RTREX_OK;
}

// One seed's inner loop (gpuExtractInner, gpu_extract.cuh:28-37; the
// sequential Decomposition.h:728-827) on the handle's current state.
rtrex_status run_extract_inner(rtrex_graph h, int64_t u, double eps, int64_t* cand,
                               double* wgt, int64_t cap, int64_t* ncand, double* internal,
                               double* potential) {
    if (h->phase < PH_CLEANED || h->phase >= PH_EXTRACTED)
        return fail(RTREX_ESTATE, "extract_inner needs a cleaned, not yet extracted graph");
    if (!(eps > 0)) return fail(RTREX_EINVAL, "eps must be > 0");
    GraphDev& g = h->g;
    if (u < 0 || u >= (int64_t)g.n) return fail(RTREX_EINVAL, "seed out of range");
    if (cap < 0 || (cap && (!cand || !wgt)) || !ncand || !internal || !potential)
        return fail(RTREX_EINVAL, "bad output arguments");
    hipStream_t s = h->stream;
    const uint32_t n = g.n;
    if (!h->xi_mk) {  // stamps start at 1: zeroed arrays hold none
        RT_TRY(dalloc(h, &h->xi_mk, n));
        RT_TRY(dalloc(h, &h->xi_cm, n));
        RT_TRY(dalloc(h, &h->xi_tw, n));
        h->xi_stamp = 0xffffffffu;
    }
    if (h->xi_stamp == 0xffffffffu) {
        RT_HIP(hipMemsetAsync(h->xi_mk, 0, (uint64_t)n * 4, s));
        RT_HIP(hipMemsetAsync(h->xi_cm, 0, (uint64_t)n * 4, s));
        h->xi_stamp = 0;
    }
    Scratch sc;
    InnerArgs A;
    A.g = g;
    A.u = (uint32_t)u;
    A.stamp = ++h->xi_stamp;
    A.eps = eps;
    A.mk = h->xi_mk;
    A.cm = h->xi_cm;
    A.tw = h->xi_tw;
    RT_TRY(sc.get(&A.nbr, (uint64_t)n));
    RT_TRY(sc.get(&A.cand, 2ull * n));
    RT_TRY(sc.get(&A.rec, (uint64_t)n));
    RT_TRY(sc.get(&A.out, 2));
    RT_TRY(sc.get(&A.nout, 1));
    k_extract_inner<<<1, XT, 0, s>>>(A);
    RT_HIP(hipGetLastError());
    double o[2];
    uint32_t nc = 0;
    RT_HIP(hipMemcpyAsync(o, A.out, sizeof o, hipMemcpyDeviceToHost, s));
    RT_HIP(hipMemcpyAsync(&nc, A.nout, 4, hipMemcpyDeviceToHost, s));
    RT_HIP(hipStreamSynchronize(s));
    const uint64_t k = std::min<uint64_t>(nc, (uint64_t)cap);
    if (k) {
        std::vector<Rec> r(k);
        RT_HIP(hipMemcpyAsync(r.data(), A.rec, k * sizeof(Rec), hipMemcpyDeviceToHost, s));
        RT_HIP(hipStreamSynchronize(s));
        for (uint64_t i = 0; i < k; ++i) {
            cand[i] = r[i].vertex;
            wgt[i] = r[i].wgt;
        }
    }
    *ncand = nc;
    *internal = o[0];
    *potential = o[1];
    return RTREX_OK;
}

// The pairs that delete a cluster's alive edges (gpuMarkClusterEdges,
// gpu_extract.cuh:53-59; the sequential Decomposition.h:897-914), statuses
// untouched: the caller hands them to rtrex_delete_and_clean.
rtrex_status run_cluster_edges(rtrex_graph h, const int64_t* verts, int64_t nverts,
                               int64_t* pairs, int64_t cap, int64_t* npairs) {
    if (h->phase < PH_CLEANED || h->phase >= PH_EXTRACTED)
        return fail(RTREX_ESTATE, "cluster_edges needs a cleaned, not yet extracted graph");
    GraphDev& g = h->g;
    if (nverts < 0 || (nverts && !verts) || cap < 0 || (cap && !pairs) || !npairs)
        return fail(RTREX_EINVAL, "bad arguments");
    if ((uint64_t)nverts > g.n) return fail(RTREX_EINVAL, "more vertices than the graph has");
    std::vector<uint32_t> mv(nverts);
    for (int64_t i = 0; i < nverts; ++i) {
        if (verts[i] < 0 || verts[i] >= (int64_t)g.n)
            return fail(RTREX_EINVAL, "vertex out of range");
        mv[i] = (uint32_t)verts[i];
    }
    {
        std::vector<uint32_t> sorted(mv);
        std::sort(sorted.begin(), sorted.end());
        if (std::adjacent_find(sorted.begin(), sorted.end()) != sorted.end())
            return fail(RTREX_EINVAL, "a vertex is listed twice");
    }
    *npairs = 0;
    if (!nverts) return RTREX_OK;
    hipStream_t s = h->stream;
    const uint32_t n = g.n;
    if (!h->xi_pos) {
        RT_TRY(dalloc(h, &h->xi_pos, n));
        RT_HIP(hipMemsetAsync(h->xi_pos, 0xff, (uint64_t)n * 4, s));  // NONE
    }
    Scratch sc;
    uint32_t *dm, *dout;
    unsigned long long* dn;
    const uint64_t k = std::min<uint64_t>((uint64_t)cap, g.m);  // at most m pairs exist
    RT_TRY(sc.get(&dm, (uint64_t)nverts));
    RT_TRY(sc.get(&dout, 2 * std::max<uint64_t>(k, 1)));
    RT_TRY(sc.get(&dn, 1));
    RT_HIP(hipMemcpyAsync(dm, mv.data(), mv.size() * 4, hipMemcpyHostToDevice, s));
    const unsigned bm = blocks_for((uint64_t)nverts, T, 1024);
    k_set_pos<<<bm, T, 0, s>>>(dm, (uint32_t)nverts, h->xi_pos, false);
    k_cluster_edges<<<1, XT, 0, s>>>(g, dm, (uint32_t)nverts, h->xi_pos, dout, k, dn);
    k_set_pos<<<bm, T, 0, s>>>(dm, (uint32_t)nverts, h->xi_pos, true);
    RT_HIP(hipGetLastError());
    unsigned long long total = 0;
    RT_HIP(hipMemcpyAsync(&total, dn, 8, hipMemcpyDeviceToHost, s));
    RT_HIP(hipStreamSynchronize(s));
    const uint64_t w = std::min<uint64_t>(total, k);
    if (w) {
        std::vector<uint32_t> hp(2 * w);
        RT_HIP(hipMemcpyAsync(hp.data(), dout, 2 * w * 4, hipMemcpyDeviceToHost, s));
        RT_HIP(hipStreamSynchronize(s));
        for (uint64_t i = 0; i < 2 * w; ++i) pairs[i] = hp[i];
    }
    *npairs = (int64_t)total;
    return RTREX_OK;
}

}  // namespace rtrex
