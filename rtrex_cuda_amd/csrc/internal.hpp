// Internal definitions shared by the HIP translation units of librtrex_hip.so.
//
// Device data layout of one resident graph (all arrays in HBM, 32-bit ids):
//   off[n+1]   u32  CSR offsets (reference layout, lists sorted by id)
//   nb[nnz]    u32  neighbour ids
//   deg[n]     u32  original degree (never changes; every weight uses it)
//   part[nnz]  u32  slot of the reverse copy of each directed slot
//   s2e[nnz]   u32  undirected edge id of each slot; edge ids are the ranks of
//                   the canonical slots (the copy stored in the lower-id
//                   endpoint's list), so edge order = lexicographic (x<y)
//   esrc[m], eslot[m]  u32  lower-id endpoint and canonical slot per edge
//   st[nnz]    u8   per-slot status, both copies always equal
//   est[m]     u8   the same status per undirected edge
//   sup[m]     f64  current support of each undirected edge
//   sup0[m]    f64  support after the count phase (audit bound)
//   tcnt[m]    u32  triangles on the edge after the count phase
//   pv[n]      f64  current per-vertex support; pv0[n] after count
//   touched[m] u8   edge lost at least one triangle since the count phase
//   ew[m]      f64  1/(d_x d_y) of each edge (threshold = eps * ew)
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <algorithm>
#include <thread>
#include <vector>

#include "../../include/rtrex.h"

namespace rtrex {

// f(a, b) over contiguous ranges covering [0, n), on up to 8 host threads
// (one range when n < 2 * min_range): host loops over per-vertex results.
template <class F>
void host_parallel(uint64_t n, uint64_t min_range, F&& f) {
    const unsigned hw = std::max(1u, std::min(8u, std::thread::hardware_concurrency()));
    const unsigned nt = (unsigned)std::min<uint64_t>(hw, n / std::max<uint64_t>(1, min_range));
    if (nt <= 1) {
        f(uint64_t(0), n);
        return;
    }
    const uint64_t per = (n + nt - 1) / nt;
    std::vector<std::thread> th;
    for (unsigned t = 1; t < nt; ++t) {
        const uint64_t a = std::min<uint64_t>(n, t * per), b = std::min<uint64_t>(n, a + per);
        if (a < b) th.emplace_back([&f, a, b] { f(a, b); });
    }
    f(uint64_t(0), std::min<uint64_t>(n, per));
    for (auto& x : th) x.join();
}

// Per-slot status codes.  The reference's characters map as Y=ALIVE,
// S=QUEUED, N=DEAD; PENDING is an edge claimed for the next peel round.
enum : uint8_t { ST_ALIVE = 0, ST_QUEUED = 1, ST_DEAD = 2, ST_PENDING = 3 };

struct GraphDev {
    uint32_t* off = nullptr;
    uint32_t* nb = nullptr;
    uint32_t* deg = nullptr;
    uint32_t* part = nullptr;
    uint32_t* s2e = nullptr;
    uint2* nbe = nullptr;  // per slot: (nb, s2e) side by side, one 8-byte load
    uint2* ndo = nullptr;  // per slot: (deg, off) of the neighbour nb
    uint32_t* esrc = nullptr;
    uint32_t* eslot = nullptr;
    uint8_t* st = nullptr;
    uint8_t* est = nullptr;  // per-edge copy of the status (coalesced scans)
    double* sup = nullptr;
    double* sup0 = nullptr;
    uint32_t* tcnt = nullptr;
    double* pv = nullptr;
    double* pv0 = nullptr;
    uint32_t* vtcnt = nullptr;
    uint8_t* touched = nullptr;
    uint32_t* claim = nullptr;
    double* ew = nullptr;  // per edge: 1 / (d_x d_y), the reference's edgeWt
    uint32_t n = 0, m = 0;
    uint64_t nnz = 0;
};

// Counters written by kernels (one device block per graph).
struct Counters {
    unsigned long long tri_total;      // triangles (count phase)
    unsigned long long tri_removed;    // triangles removed by cleaning
    unsigned long long edges_removed;  // edges removed by cleaning
    unsigned long long ambiguous_clean;
    unsigned long long ambiguous_ratio;
    unsigned long long seeds;
    unsigned int frontier_next;        // size of the next peel frontier
    unsigned int error;                // validation / capacity flags
    unsigned int work_next;            // extraction work queue head
    unsigned int changed;              // union-find progress flag
    unsigned int n_undecided;          // edges the clean could not decide
    unsigned int pad;
    unsigned long long replayed_edges; // edges re-run by the exact replay
};

enum : uint32_t {
    ERR_SELF_LOOP = 1u,
    ERR_UNSORTED_OR_DUP = 2u,
    ERR_RANGE = 4u,
    ERR_ASYMMETRIC = 8u,
    ERR_SCRATCH = 16u,
};

// Triangle list of every edge, written by the count phase: for edge e with
// lower-degree endpoint lo (ties: the lower id) and the other endpoint hi,
// terms [zoff[e], zoff[e] + tcnt[e]) hold, ascending in the third vertex z,
// one 16-byte record each (one vector load or store): the weight
// 1/(d_x d_y d_z) and the edge ids of {lo, z} (sa) and {hi, z} (sb).
struct __attribute__((aligned(16))) Term {
    double w;
    uint32_t sa, sb;
};
// A count kernel's result for one edge: its support and triangle count in
// one 16-byte record, so that each edge costs one scattered store; a
// streaming pass then fills sup, sup0 and tcnt in edge order.
struct __attribute__((aligned(16))) SupRec {
    double s;
    uint32_t c, pad;
};
// Record k of the term buffer for lane-uniform code: every lane loads (k is
// an in-range index when in is set, and entry 0 otherwise: the buffer always
// holds at least one entry), but a lane past its list gets the record of edge
// 0 with weight 0, never the entry it read.  Entries past an edge's matches
// are capacity, not records, and may hold anything (entry 0 included, when
// edge 0 has no triangle).
__device__ __forceinline__ Term term_or_blank(const Term* t, uint64_t k, bool in) {
    const Term q = t[in ? k : 0];
    return in ? q : Term{0.0, 0u, 0u};
}

struct TermBuf {
    Term* t;  // null when the buffer did not fit (intersection fallbacks)
    const uint64_t* zoff;
    const uint64_t* zsl = nullptr;  // per slot: zoff of the slot's edge
    SupRec* rec = nullptr;          // count kernels' per-edge results
};

struct Timings {
    double build_ms = 0, count_ms = 0, clean_ms = 0, extract_ms = 0;
    double count_edge_ms = 0, count_vertex_ms = 0;
};

struct AuditHost {
    int64_t clean_rounds = 0, edges_removed = 0, triangles_removed = 0;
    int64_t ambiguous_clean = 0, ambiguous_ratio = 0, seeds = 0;
    int64_t components = 0, largest_component = 0;
};

// Largest hub degree the count kernels accept (the largest hubs' tree spans
// 262144 entries from the 8-entry boundary at or before the list's start).
constexpr uint32_t HUB_MAX_LIMIT = 262136;

// Count-plan parameters (the autotuner's search space).  An edge whose
// longer-list endpoint (its hub) has degree in (hub_min, hub_max] is counted
// per hub, the hub list staged as an 8-ary pivot tree: whole in LDS for
// degrees up to 64 (l0_max), 512 (l1_max) and 4096 (l2_max), its upper levels
// in LDS and its leaf blocks read from HBM above l2_max; the rest per edge.
// A hub's edges get G = 4, 8, 16, 32 or 64 lanes each by their mean
// shorter-list length against g_thr[0..3].  Work units carry at least
// unit_min streamed elements (classes above l1_max: unit_hash_mult * d_h).
struct CountParams {
    uint32_t hub_min = 0, hub_max = HUB_MAX_LIMIT;
    uint32_t l0_max = 64, l1_max = 512, l2_max = 4096;
    uint32_t l3_max = 32760;  // above: the huge-hub tree of five levels
    uint32_t g_thr[4] = {3, 6, 24, 80};
    uint32_t unit_min = 1024, unit_hash_mult = 8;
};

// Per-handle scratch arena (see Scratch below).
struct Arena {
    char* base = nullptr;
    uint64_t cap = 0, top = 0, vtop = 0, demand = 0;
    bool owned = true;  // false: carved from the handle's slab
};

enum Phase { PH_NONE = 0, PH_BUILT = 1, PH_COUNTED = 2, PH_CLEANED = 3, PH_EXTRACTED = 4 };

// ---- process-wide caches (memcache.hip) -------------------------------------
constexpr int NEV = 9;  // timing events per handle
struct StreamSet {
    hipStream_t s = nullptr;
    hipEvent_t ev[NEV] = {};
};
// Device memory of the current device (null and *err on failure); blocks
// come back to a cache that later requests reuse.
void* dev_alloc(uint64_t bytes, hipError_t* err);
void dev_free(void* p);
void dev_trim();  // every cached block of the current device released
uint64_t dev_cached_bytes();
bool dev_cache_fits(uint64_t bytes);  // a cached block would serve dev_alloc(bytes)
// Page-locked host memory (null on failure), cached the same way.
void* pin_alloc(uint64_t bytes);
void pin_free(void* p);
// A stream of the current device with its events, from the pool (false: none
// pooled); given back once drained.
bool stream_take(StreamSet* out);
void stream_give(const StreamSet& s);

}  // namespace rtrex

// The opaque handle of the C ABI.
struct rtrex_graph_s {
    int device = 0;
    hipStream_t stream = nullptr;
    rtrex::GraphDev g;
    rtrex::Counters* ctr = nullptr;  // device
    std::vector<void*> allocs;       // owned device allocations
    rtrex::Phase phase = rtrex::PH_NONE;
    double eps = 0;
    // Extraction results (device): seed vertex of each vertex's cluster,
    // and the cluster id of every vertex.
    uint32_t* cseed = nullptr;
    uint32_t* cid = nullptr;  // cluster id per vertex (global numbering)
    uint32_t* mrank = nullptr;  // position at which each vertex joined its cluster
    // Per-seed calls (rtrex_extract_inner, rtrex_cluster_edges): neighbour and
    // candidate stamps (xi_stamp per call), candidate weights, member index.
    uint32_t *xi_mk = nullptr, *xi_cm = nullptr, *xi_pos = nullptr;
    double* xi_tw = nullptr;
    uint32_t xi_stamp = 0;
    uint8_t* xact = nullptr;    // per edge: post-clean support is the reference's bits
    std::vector<uint32_t> bcnt;     // unordered_set bucket count after k insertions
    uint32_t* bcnt_dev = nullptr;   // device copy of bcnt[0, bcnt_dev_len), capacity n + 1
    uint64_t bcnt_dev_len = 0;
    uint32_t bcnt_ratio = 0;        // max over k >= 2 of ceil(bcnt[k] / k)
    std::vector<uint32_t> x_undec;  // edges the last extraction could not decide
    bool x_undec_all = false;       // ... more than the list holds
    // Deletions since the initial clean (rtrex_delete_and_clean), in call
    // order: per entry the edge, the first endpoint of the pair as given
    // (deleteAndClean's "lower" on a degree tie is the second), the call.
    std::vector<uint32_t> dlog_e, dlog_a, dlog_b, dlog_c;
    uint32_t dlog_calls = 0;
    bool replay_all = false;  // every clean decision re-run in the reference's order (checking)
    // Extraction: a cascade whose stack reaches this many entries is finished
    // by a workgroup of several waves instead of one wave (0: never).
    uint32_t cascade_min = 4096;
    std::vector<uint32_t> force_replay;  // edges whose regions the clean re-runs
    bool force_all = false;
    double force_eps = -1;  // the eps force_replay was found at
    int64_t n_clusters = -1;
    std::vector<int64_t> seg_clusters;  // per segment (batch)
    // Cluster ids stay in HBM (cid); copy_cluster_map brings them to the host
    // on request through this pinned staging buffer (n entries, lazily).
    uint32_t* cm_pin = nullptr;
    uint32_t* sb_pin = nullptr;  // batch: page-locked landing of seg_bounds
    void* up_pin = nullptr;  // page-locked staging of an edge-list upload (pin cache)
    std::vector<uint32_t> seg_bounds;  // batch: first cluster id of each segment, and the total
    uint32_t* seg_pos_d = nullptr;     // batch: seg_vtx on the device (u32)
    uint32_t* vseg_d = nullptr;        // batch: each vertex's segment (graph)
    uint32_t* seg_bounds_d = nullptr;  // batch: seg_bounds on the device
    // Batch layout (one graph = one segment): vertex segment boundaries.
    std::vector<int64_t> seg_vtx;  // empty for a single graph
    uint32_t* nearst = nullptr;  // per edge: last peel round it was near-threshold
    uint32_t* surv = nullptr;    // packed survivors of the last initial clean
    uint64_t n_alive = 0;        // alive edges (host bookkeeping of the peel)
    // Count-phase plan (static per graph): work-bucket lists, their sizes and
    // the term buffer (per-edge matched weights, ascending z).
    uint32_t* cnt_lists = nullptr;
    uint32_t *seed_order = nullptr, *seed_rank = nullptr;  // (segment, degree, id)
    uint32_t cnt_counts[8] = {};
    // Hub plan: edges owned by their longer-list endpoint ("hub"), listed per
    // hub (own_off[n+1], own_idx = local slot index in the hub's list), and
    // work units (hub | part << 32 | parts << 48) per (staging size, lanes)
    // class, hub_ucnt[] units each, concatenated in hub_units.
    uint32_t* own_off = nullptr;
    uint32_t* own_idx = nullptr;
    uint4* hub_units = nullptr;
    // Small-hub units split into chunks of about equal streamed work: chunk c
    // is units [small_chunk[c], small_chunk[c + 1]).
    uint32_t* small_chunk = nullptr;
    uint64_t small_chunk_cap = 0;
    uint32_t small_nchunks = 0;
    uint32_t hub_ucnt[20] = {};
    uint32_t hub_umax[20] = {};  // largest hub degree per class
    unsigned long long* red = nullptr;  // per-block partial sums (count phase)
    rtrex::CountParams cparams;
    bool plan_valid = false;  // buckets and hub units built for cparams
    uint64_t* zoff = nullptr;
    uint64_t* zsl = nullptr;         // zoff per slot (hub kernels stage it with the hub list)
    rtrex::SupRec* srec = nullptr;   // count kernels' per-edge results
    rtrex::Term* zt = nullptr;  // term buffer (TermBuf::t)
    bool terms_exact_pending = false;  // size the buffer after a counting pass
    bool pv_full = false;              // per-vertex supports valid everywhere
    uint32_t round_stamp = 0;
    rtrex::Timings tm;
    rtrex::AuditHost audit;
    rtrex::Arena arena;  // phase scratch
    // Slab mode (serving): every device allocation of the handle is carved
    // from [slab, slab + slab_cap) of a pool the caller owns; nothing is freed
    // per allocation.  The stream may belong to the caller as well.
    char* slab = nullptr;
    uint64_t slab_cap = 0, slab_top = 0;
    bool own_stream = true;
    rtrex::TermBuf terms() const { return {zt, zoff, zsl, srec}; }
    // Events: 0-1 build, 2-4 count, 5-6 clean, 7-8 extract.  The count and
    // clean times are read once something waits for the stream anyway
    // (resolve_timings), so those phases end without a host round trip.
    hipEvent_t ev[rtrex::NEV] = {};
    bool tm_count_pending = false, tm_clean_pending = false, tm_extract_pending = false;
    // Host mailbox: pinned words a kernel publishes small results into
    // (value | sequence << 32), polled by the host in place of a copy and a
    // stream synchronisation (box: host address, box_dev: device address).
    uint64_t* box = nullptr;
    uint64_t* box_dev = nullptr;
    uint32_t box_seq = 0;
};

namespace rtrex {

// Mailbox slots (rtrex_graph_s::box).
constexpr int BOX_SLOTS = 8;
// Takes a mailbox for the handle on first use (box_close gives it back; the
// handle's stream must have drained).
rtrex_status box_open(rtrex_graph h);
void box_close(rtrex_graph h);
// Waits until slot holds sequence number seq and returns its value; fails if
// the stream reports an error, or drains without the value arriving.
rtrex_status box_wait(rtrex_graph h, int slot, uint32_t seq, uint32_t* out);

// Bytes taken from the handle's slab, 256-byte aligned; null if it is full.
inline void* slab_take(rtrex_graph h, uint64_t bytes) {
    if (!h->slab) return nullptr;
    bytes = (bytes + 255) & ~255ull;
    if (h->slab_top + bytes > h->slab_cap) return nullptr;
    void* q = h->slab + h->slab_top;
    h->slab_top += bytes;
    return q;
}

// Device bytes a graph of n vertices and m undirected edges holds while
// resident (every per-graph array, count plans included) and the largest
// phase scratch; the stored triangle lists are 16 bytes per (term + 1) on top.
uint64_t resident_bytes(uint64_t n, uint64_t m);
uint64_t scratch_bytes(uint64_t n, uint64_t m);

// Graph from an undirected edge list on `device`.  With `stream` the handle
// uses (and does not own) that stream; with `slab` every device allocation
// is carved from [slab, slab + slab_cap) (hipMalloc only past its end).
rtrex_status create_edges_in(const uint32_t* u, const uint32_t* v, int64_t n, int64_t m,
                             int device, hipStream_t stream, char* slab, uint64_t slab_cap,
                             rtrex_graph* out);
// Concatenated per-graph edge lists (local ids) -> one disjoint union.
rtrex_status union_edges(int64_t n_graphs, const int64_t* vtx, const int64_t* eoff,
                         const uint32_t* u, const uint32_t* v, std::vector<int64_t>& vbase,
                         std::vector<uint32_t>& U, std::vector<uint32_t>& V);

// Sets the status of undirected edge e in its per-edge byte and both slots.
__device__ __forceinline__ void set_status(const GraphDev& g, uint32_t e, uint8_t c) {
    g.est[e] = c;
    const uint32_t sl = g.eslot[e];
    g.st[sl] = c;
    g.st[g.part[sl]] = c;
}

// ---- error reporting (api.cpp) -------------------------------------------
void set_error(const std::string& msg);
rtrex_status fail(rtrex_status code, const std::string& msg);
rtrex_status hip_fail(hipError_t e, const char* what, const char* file,
                      int line);

#define RT_HIP(call)                                                       \
    do {                                                                   \
        hipError_t e_ = (call);                                            \
        if (e_ != hipSuccess)                                              \
            return ::rtrex::hip_fail(e_, #call, __FILE__, __LINE__);       \
    } while (0)

#define RT_TRY(call)                                                       \
    do {                                                                   \
        rtrex_status s_ = (call);                                          \
        if (s_ != RTREX_OK) return s_;                                     \
    } while (0)

// ---- device allocation owned by a handle ---------------------------------
template <class T>
rtrex_status dalloc(rtrex_graph h, T** p, uint64_t count) {
    void* q = nullptr;
    uint64_t bytes = count ? count * sizeof(T) : sizeof(T);
    if ((q = slab_take(h, bytes))) {
        *p = static_cast<T*>(q);
        return RTREX_OK;
    }
    hipError_t e = hipSuccess;
    q = dev_alloc(bytes, &e);
    if (!q)
        return fail(RTREX_ENOMEM, "device allocation of " + std::to_string(bytes) +
                                      " bytes failed: " + hipGetErrorString(e));
    h->allocs.push_back(q);
    *p = static_cast<T*>(q);
    return RTREX_OK;
}

// Per-handle scratch arena: one device buffer reused by every phase call
// (stack discipline: a Scratch releases what it took when it goes out of
// scope).  Work on a handle is ordered on its one stream, so memory released
// by a finished phase can be handed to the next one without a sync.  A
// request that does not fit is served by hipMalloc for that call and raises
// `demand`; the arena grows to the demand at the start of the next call.
// Arena of the handle whose API call is running on this thread (or null).
extern thread_local Arena* tl_arena;
// The stream of the API call in progress on this thread (set with tl_arena).
extern thread_local hipStream_t tl_stream;

// Scratch that lives for one call.
struct Scratch {
    std::vector<void*> ptrs;
    Arena* a = tl_arena;
    hipStream_t stream = tl_stream;
    uint64_t mark = a ? a->top : 0, vmark = a ? a->vtop : 0;
    ~Scratch() {
        // Kernels queued on this buffer may still run (callers do not wait
        // for their last launches): buffers of their own are freed only once
        // the call's stream has drained (the device, outside an API call).
        // Arena scratch is reused in stream order.
        if (!ptrs.empty()) {
            if (stream) (void)hipStreamSynchronize(stream);
            else (void)hipDeviceSynchronize();
        }
        for (void* p : ptrs) dev_free(p);
        if (a) {
            a->top = mark;
            a->vtop = vmark;
        }
    }
    template <class T>
    rtrex_status get(T** p, uint64_t count) {
        void* q = nullptr;
        uint64_t bytes = count ? count * sizeof(T) : sizeof(T);
        if (a) {
            bytes = (bytes + 255) & ~255ull;
            a->vtop += bytes;
            a->demand = std::max(a->demand, a->vtop);
            if (a->top + bytes <= a->cap) {
                *p = reinterpret_cast<T*>(a->base + a->top);
                a->top += bytes;
                return RTREX_OK;
            }
        }
        hipError_t e = hipSuccess;
        q = dev_alloc(bytes, &e);
        if (!q)
            return fail(RTREX_ENOMEM, "scratch allocation of " + std::to_string(bytes) +
                                          " bytes failed");
        ptrs.push_back(q);
        *p = static_cast<T*>(q);
        return RTREX_OK;
    }
};

// ---- primitives (sort.hip) ------------------------------------------------
// In-place LSD radix sort of u64 keys on bits [0, bits); tmp has n entries.
rtrex_status radix_sort_u64(uint64_t* keys, uint64_t* tmp, uint64_t n,
                            int bits, hipStream_t s);
// Exclusive scan of u32 counts into u64 (out has n+1 entries, out[n] = sum).
rtrex_status exclusive_scan_u32(const uint32_t* in, uint64_t* out, uint64_t n,
                                hipStream_t s);
// Same, u32 output (caller guarantees the sum fits in 32 bits).
rtrex_status exclusive_scan_u32_32(const uint32_t* in, uint32_t* out,
                                   uint64_t n, hipStream_t s);

// ---- phases ---------------------------------------------------------------
rtrex_status build_from_csr(rtrex_graph h, const int64_t* off,
                            const int64_t* nb, int64_t n, int64_t nnz);
rtrex_status build_from_edges(rtrex_graph h, const uint32_t* u,
                              const uint32_t* v, int64_t n, int64_t m);
// Edge-list text parsed on the device (ingest.hip): load_text builds the
// graph of an Escape or SNAP file; convert_text writes the Escape file of a
// SNAP file (snap_to_rtr.py's conversion).
rtrex_status load_text(rtrex_graph h, const char* path, int format);
rtrex_status convert_text(rtrex_graph h, const char* in, const char* out, int64_t* n,
                          int64_t* m);
// The same from an edge list already in device memory.
rtrex_status build_from_device_edges(rtrex_graph h, const uint32_t* du, const uint32_t* dv,
                                     int64_t n, int64_t m);
rtrex_status run_count(rtrex_graph h);
// Phase times whose events are still pending (waits for them).
rtrex_status resolve_timings(rtrex_graph h);
rtrex_status run_initial_clean(rtrex_graph h, double eps);
rtrex_status run_delete_and_clean(rtrex_graph h, const int64_t* pairs,
                                  int64_t k, double eps);
rtrex_status run_extract(rtrex_graph h, double eps);
rtrex_status run_extract_inner(rtrex_graph h, int64_t u, double eps, int64_t* cand,
                               double* wgt, int64_t cap, int64_t* ncand, double* internal,
                               double* potential);
rtrex_status run_cluster_edges(rtrex_graph h, const int64_t* verts, int64_t nverts,
                               int64_t* pairs, int64_t cap, int64_t* npairs);
// The cluster map of the last extraction, as int64 per vertex (per-graph ids
// on a batch handle).
rtrex_status copy_cluster_map(rtrex_graph h, int64_t* out);
// True when p is page-locked host memory the device can address; *dptr is
// its device address.
bool mapped_host(const void* p, void** dptr);
rtrex_status run_cluster_stats(rtrex_graph h, int64_t* nv, int64_t* ne,
                               int64_t* cut);
rtrex_status run_connected_components(rtrex_graph h, int64_t* comp, int64_t* n_comp);
rtrex_status run_cluster_stats_noalloc(rtrex_graph h, int64_t* nv, int64_t* ne,
                                       int64_t* cut);
// Per-vertex supports of the listed vertices (all when list is null) on
// demand, when stored triangle lists are kept (count.hip).
rtrex_status ensure_vertex_supports(rtrex_graph h, const uint32_t* list, uint32_t cnt);

inline unsigned blocks_for(uint64_t work, unsigned per_block,
                           unsigned cap = 1u << 20) {
    uint64_t b = (work + per_block - 1) / per_block;
    if (b < 1) b = 1;
    if (b > cap) b = cap;
    return (unsigned)b;
}

}  // namespace rtrex
