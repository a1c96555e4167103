// Device-resident graph: CSR upload or on-device CSR construction, partner
// slots and undirected edge numbering.
//
// Reference counterparts: transferGraphToDevice/allocateDeviceState
// (RTRex_CUDA/gpu_graph.cu:46-135), makeCSR + sortById (Graph.cpp:344-384,
// 201-205) and getPartnerMap (Graph.cpp:320-340).
//
// Edge-list build: degrees by atomics, offsets by an exclusive scan, both
// directions of every edge scattered into their lists, then each list sorted
// on its own (by a thread, a wave or a wave in LDS by length; lists over 4,096
// entries together by the LSD radix sort) -- the order makeCSR's std::sort of
// (src, dst) pairs gives, without sorting all 2m pairs.  Partner slots are found by a binary
// search of src in dst's sorted list (one dependent search per slot, coalesced
// outer loop); edge ids are the ranks of the canonical slots (src < dst).
#include <algorithm>

#include <cstring>
#include <thread>
#include <vector>

#include "dev.hpp"
#include "internal.hpp"

namespace rtrex {
namespace {

constexpr int T = 256;

// src of every slot: the vertex whose [off[v], off[v+1]) range holds it.
__global__ void __launch_bounds__(T) k_slot_src(const uint32_t* __restrict__ off,
                                                uint32_t n, uint64_t nnz,
                                                uint32_t* __restrict__ src) {
    for (uint64_t i = (uint64_t)blockIdx.x * T + threadIdx.x; i < nnz;
         i += (uint64_t)gridDim.x * T) {
        // Largest v with off[v] <= i (lists may be empty).
        uint32_t lo = 0, hi = n;  // answer in [lo, hi)
        while (hi - lo > 1) {
            uint32_t mid = lo + ((hi - lo) >> 1);
            if (off[mid] <= i) lo = mid;
            else hi = mid;
        }
        src[i] = lo;
    }
}

// Degrees and the CSR validity checks of the uploaded graph: ids in range,
// no self-loops, each list strictly increasing (sorted, no duplicates).
__global__ void __launch_bounds__(T) k_check_csr(
    const uint32_t* __restrict__ off, const uint32_t* __restrict__ nb,
    const uint32_t* __restrict__ src, uint32_t n, uint64_t nnz,
    uint32_t* __restrict__ deg, unsigned* err) {
    for (uint64_t i = (uint64_t)blockIdx.x * T + threadIdx.x; i < nnz;
         i += (uint64_t)gridDim.x * T) {
        uint32_t s = src[i], x = nb[i];
        unsigned e = 0;
        if (x >= n) e |= ERR_RANGE;
        if (x == s) e |= ERR_SELF_LOOP;
        if (i + 1 < off[s + 1] && nb[i + 1] <= x) e |= ERR_UNSORTED_OR_DUP;
        if (e) atomicOr(err, e);
    }
    for (uint64_t v = (uint64_t)blockIdx.x * T + threadIdx.x; v < n;
         v += (uint64_t)gridDim.x * T)
        deg[v] = off[v + 1] - off[v];
}

__global__ void __launch_bounds__(T) k_partner(const uint32_t* __restrict__ off,
                                               const uint32_t* __restrict__ nb,
                                               const uint32_t* __restrict__ src,
                                               uint64_t nnz,
                                               uint32_t* __restrict__ part,
                                               uint32_t* __restrict__ canon,
                                               unsigned* err) {
    for (uint64_t i = (uint64_t)blockIdx.x * T + threadIdx.x; i < nnz;
         i += (uint64_t)gridDim.x * T) {
        uint32_t s = src[i], x = nb[i];
        int64_t p = find_slot(nb, off[x], off[x + 1], s);
        if (p < 0) {
            atomicOr(err, ERR_ASYMMETRIC);
            p = i;
        }
        part[i] = (uint32_t)p;
        canon[i] = x > s ? 1u : 0u;
    }
}

__global__ void __launch_bounds__(T) k_edge_ids(const uint32_t* __restrict__ nb,
                                                const uint32_t* __restrict__ src,
                                                const uint32_t* __restrict__ canon,
                                                const uint32_t* __restrict__ rank,
                                                const uint32_t* __restrict__ deg,
                                                uint64_t nnz,
                                                uint32_t* __restrict__ s2e,
                                                uint32_t* __restrict__ esrc,
                                                uint32_t* __restrict__ eslot,
                                                double* __restrict__ ew) {
    for (uint64_t i = (uint64_t)blockIdx.x * T + threadIdx.x; i < nnz;
         i += (uint64_t)gridDim.x * T) {
        if (canon[i]) {
            uint32_t e = rank[i];
            const uint32_t x = src[i], y = nb[i];
            s2e[i] = e;
            esrc[e] = x;
            eslot[e] = (uint32_t)i;
            // edgeWt = 1/double(d_u * d_v), the product in int64
            // (Escape/Decomposition.h:395); depends on the degrees only.
            ew[e] = 1.0 / (double)((uint64_t)deg[x] * (uint64_t)deg[y]);
        }
    }
}

__global__ void __launch_bounds__(T) k_edge_ids_mirror(
    const uint32_t* __restrict__ part, const uint32_t* __restrict__ canon,
    const uint32_t* __restrict__ rank, uint64_t nnz, uint32_t* __restrict__ s2e) {
    for (uint64_t i = (uint64_t)blockIdx.x * T + threadIdx.x; i < nnz;
         i += (uint64_t)gridDim.x * T)
        if (!canon[i]) s2e[i] = rank[part[i]];
}

// Per slot, side by side: the neighbour id and the edge id (nbe), the
// neighbour's degree and list offset (ndo).  The counting kernels stream nbe
// for every element of a shorter list (one 8-byte load touches half the cache
// lines of two 4-byte loads from separate arrays) and stage a hub's nbe and
// ndo in one load round trip, with no dependent gather of degrees/offsets.
__global__ void __launch_bounds__(T) k_pack_slots(const uint32_t* __restrict__ nb,
                                                  const uint32_t* __restrict__ s2e,
                                                  const uint32_t* __restrict__ deg,
                                                  const uint32_t* __restrict__ off, uint64_t nnz,
                                                  uint2* __restrict__ nbe,
                                                  uint2* __restrict__ ndo) {
    for (uint64_t i = (uint64_t)blockIdx.x * T + threadIdx.x; i < nnz;
         i += (uint64_t)gridDim.x * T) {
        const uint32_t z = nb[i];
        nbe[i] = make_uint2(z, s2e[i]);
        ndo[i] = make_uint2(deg[z], off[z]);
    }
}

// ---- edge-list build: the CSR without a global sort -- degrees, offsets, each
// edge scattered into both endpoints' lists, then every list sorted on its
// own.  An out-of-range id is flagged and the edge taken as (0, 0), a
// self-loop flagged (the build then fails in finish()), so every slot holds
// ids in range.
__device__ __forceinline__ void edge_ends(const uint32_t* __restrict__ u,
                                          const uint32_t* __restrict__ v, uint64_t i,
                                          uint32_t n, uint32_t& a, uint32_t& b, unsigned& e) {
    a = u[i];
    b = v[i];
    e = 0;
    if (a >= n || b >= n) {
        e |= ERR_RANGE;
        a = b = 0;
    }
    if (a == b) e |= ERR_SELF_LOOP;
}

__global__ void __launch_bounds__(T) k_edge_degrees(const uint32_t* __restrict__ u,
                                                    const uint32_t* __restrict__ v, uint64_t m,
                                                    uint32_t n, uint32_t* __restrict__ deg,
                                                    unsigned* err) {
    for (uint64_t i = (uint64_t)blockIdx.x * T + threadIdx.x; i < m;
         i += (uint64_t)gridDim.x * T) {
        uint32_t a, b;
        unsigned e;
        edge_ends(u, v, i, n, a, b, e);
        if (e) atomicOr(err, e);
        atomicAdd(&deg[a], 1u);
        atomicAdd(&deg[b], 1u);
    }
}

// cur: per-vertex fill counts, zero on entry.  Lists come out unordered.
__global__ void __launch_bounds__(T) k_edge_fill(const uint32_t* __restrict__ u,
                                                 const uint32_t* __restrict__ v, uint64_t m,
                                                 uint32_t n, const uint32_t* __restrict__ off,
                                                 uint32_t* __restrict__ cur,
                                                 uint32_t* __restrict__ nb,
                                                 uint32_t* __restrict__ src) {
    for (uint64_t i = (uint64_t)blockIdx.x * T + threadIdx.x; i < m;
         i += (uint64_t)gridDim.x * T) {
        uint32_t a, b;
        unsigned e;
        edge_ends(u, v, i, n, a, b, e);
        const uint32_t pa = off[a] + atomicAdd(&cur[a], 1u);
        const uint32_t pb = off[b] + atomicAdd(&cur[b], 1u);
        nb[pa] = b;
        src[pa] = a;
        nb[pb] = a;
        src[pb] = b;
    }
}

// Every list sorted ascending, each list by the cheapest unit for its
// length: up to 16 entries by one thread (a bitonic network in registers),
// up to 64 by one wave (a bitonic network across the lanes), up to SL_LDS by
// one wave in LDS; longer lists are listed in big (count *nbig) for the radix
// sort.  A repeated neighbour is an error (duplicate edge).
__device__ __forceinline__ void cswap(uint32_t& a, uint32_t& b) {
    const uint32_t lo = min(a, b), hi = max(a, b);
    a = lo;
    b = hi;
}

constexpr uint32_t SL_REG = 16;
__global__ void __launch_bounds__(T) k_sort_short(const uint32_t* __restrict__ off, uint32_t n,
                                                  uint32_t* __restrict__ nb, unsigned* err) {
    for (uint32_t x = blockIdx.x * T + threadIdx.x; x < n; x += gridDim.x * T) {
        const uint32_t b0 = off[x], d = off[x + 1] - b0;
        if (d < 2 || d > SL_REG) continue;
        uint32_t r[SL_REG];
#pragma unroll
        for (uint32_t i = 0; i < SL_REG; ++i) r[i] = i < d ? nb[b0 + i] : 0xffffffffu;
#pragma unroll
        for (uint32_t k = 2; k <= SL_REG; k <<= 1) {
#pragma unroll
            for (uint32_t j = k >> 1; j > 0; j >>= 1) {
#pragma unroll
                for (uint32_t i = 0; i < SL_REG; ++i) {
                    const uint32_t l = i ^ j;
                    if (l > i) {
                        if ((i & k) == 0) cswap(r[i], r[l]);
                        else cswap(r[l], r[i]);
                    }
                }
            }
        }
        bool dup = false;
#pragma unroll
        for (uint32_t i = 0; i < SL_REG; ++i) {
            if (i < d) nb[b0 + i] = r[i];
            if (i + 1 < d) dup |= r[i] == r[i + 1];
        }
        if (dup) atomicOr(err, ERR_UNSORTED_OR_DUP);
    }
}

constexpr int SL_W = 4;
constexpr uint32_t SL_LDS = 4096;
__global__ void __launch_bounds__(SL_W * WAVE) k_sort_long(const uint32_t* __restrict__ off,
                                                           uint32_t n, uint32_t* __restrict__ nb,
                                                           uint32_t* __restrict__ big,
                                                           unsigned* nbig, unsigned* err) {
    __shared__ uint32_t s_a[SL_W][SL_LDS];
    const int lane = lane_id();
    const uint32_t wave = (blockIdx.x * SL_W * WAVE + threadIdx.x) / WAVE;
    const uint32_t nwaves = gridDim.x * SL_W;
    uint32_t* a = s_a[threadIdx.x / WAVE];
    // 64 vertices per step; the lists longer than SL_REG among them in turn
    for (uint32_t c = wave * WAVE; c < n; c += nwaves * WAVE) {
        const uint32_t xl = c + lane;
        const uint32_t o0 = xl < n ? off[xl] : 0, o1 = xl < n ? off[xl + 1] : 0;
        uint64_t todo = ballot64(o1 - o0 > SL_REG);
        while (todo) {  // wave-uniform
            const int q = __ffsll((unsigned long long)todo) - 1;
            todo &= todo - 1;
            const uint32_t x = c + (uint32_t)q;
            const uint32_t b0 = __shfl(o0, q, WAVE), d = __shfl(o1, q, WAVE) - b0;
            if (d <= (uint32_t)WAVE) {
                uint32_t y = (uint32_t)lane < d ? nb[b0 + lane] : 0xffffffffu;
#pragma unroll
                for (int k = 2; k <= WAVE; k <<= 1) {
#pragma unroll
                    for (int j = k >> 1; j > 0; j >>= 1) {
                        const uint32_t o = __shfl_xor(y, j, WAVE);
                        const bool asc = (lane & k) == 0, low = (lane & j) == 0;
                        y = low == asc ? min(y, o) : max(y, o);
                    }
                }
                const uint32_t nx = __shfl_down(y, 1, WAVE);
                if ((uint32_t)lane + 1 < d && y == nx) atomicOr(err, ERR_UNSORTED_OR_DUP);
                if ((uint32_t)lane < d) nb[b0 + lane] = y;
            } else if (d <= SL_LDS) {
                uint32_t np = 2 * WAVE;
                while (np < d) np <<= 1;
                for (uint32_t i = lane; i < np; i += WAVE) a[i] = i < d ? nb[b0 + i] : 0xffffffffu;
                wave_sync();
                for (uint32_t k = 2; k <= np; k <<= 1) {
                    for (uint32_t j = k >> 1; j > 0; j >>= 1) {
                        for (uint32_t i = lane; i < np; i += WAVE) {
                            const uint32_t l = i ^ j;
                            if (l > i) {
                                const uint32_t ai = a[i], al = a[l];
                                if ((ai > al) == ((i & k) == 0)) {
                                    a[i] = al;
                                    a[l] = ai;
                                }
                            }
                        }
                        wave_sync();
                    }
                }
                for (uint32_t i = lane; i < d; i += WAVE) {
                    const uint32_t y = a[i];
                    if (i + 1 < d && a[i + 1] == y) atomicOr(err, ERR_UNSORTED_OR_DUP);
                    nb[b0 + i] = y;
                }
                wave_sync();  // a[] is reused by the wave's next list
            } else if (lane == 0) {
                big[atomicAdd(nbig, 1u)] = x;
            }
        }
    }
}

// Lists longer than SL_LDS: keys (list index << B | neighbour) gathered at
// pos[i] for big list i, radix-sorted, written back; one workgroup per list.
__global__ void __launch_bounds__(T) k_big_keys(const uint32_t* __restrict__ off,
                                                const uint32_t* __restrict__ big,
                                                const uint64_t* __restrict__ pos, int B,
                                                const uint32_t* __restrict__ nb,
                                                uint64_t* __restrict__ keys) {
    const uint32_t i = blockIdx.x, x = big[i];
    const uint32_t b0 = off[x], d = off[x + 1] - b0;
    for (uint32_t j = threadIdx.x; j < d; j += T) keys[pos[i] + j] = ((uint64_t)i << B) | nb[b0 + j];
}

__global__ void __launch_bounds__(T) k_big_back(const uint32_t* __restrict__ off,
                                                const uint32_t* __restrict__ big,
                                                const uint64_t* __restrict__ pos, int B,
                                                const uint64_t* __restrict__ keys,
                                                uint32_t* __restrict__ nb, unsigned* err) {
    const uint32_t i = blockIdx.x, x = big[i];
    const uint32_t b0 = off[x], d = off[x + 1] - b0;
    const uint64_t mask = (1ull << B) - 1;
    for (uint32_t j = threadIdx.x; j < d; j += T) {
        const uint64_t k = keys[pos[i] + j];
        if (j + 1 < d && keys[pos[i] + j + 1] == k) atomicOr(err, ERR_UNSORTED_OR_DUP);
        nb[b0 + j] = (uint32_t)(k & mask);
    }
}

__global__ void k_big_bounds(const uint32_t* __restrict__ off, const uint32_t* __restrict__ big,
                             uint32_t nbig, uint32_t* __restrict__ len) {
    for (uint32_t i = blockIdx.x * T + threadIdx.x; i < nbig; i += gridDim.x * T)
        len[i] = off[big[i] + 1] - off[big[i]];
}

int bits_for(uint64_t n) {
    int b = 1;
    while ((1ull << b) < n) ++b;
    return b;
}

std::string describe(unsigned err) {
    std::string s;
    if (err & ERR_RANGE) s += " vertex id out of range;";
    if (err & ERR_SELF_LOOP) s += " self-loop;";
    if (err & ERR_UNSORTED_OR_DUP) s += " unsorted list or duplicate edge;";
    if (err & ERR_ASYMMETRIC) s += " not symmetric (missing reverse slot);";
    return s;
}

rtrex_status alloc_state(rtrex_graph h, uint32_t n, uint64_t nnz) {
    GraphDev& g = h->g;
    const uint64_t m = nnz / 2;
    g.n = n;
    g.nnz = nnz;
    g.m = (uint32_t)m;
    RT_TRY(dalloc(h, &g.off, (uint64_t)n + 1));
    // 8 entries of padding: a hub-list block read (32 aligned bytes) may run
    // past the last list
    RT_TRY(dalloc(h, &g.nb, nnz + 8));
    RT_TRY(dalloc(h, &g.deg, n));
    RT_TRY(dalloc(h, &g.part, nnz));
    RT_TRY(dalloc(h, &g.s2e, nnz));
    RT_TRY(dalloc(h, &g.nbe, nnz));
    RT_TRY(dalloc(h, &g.ndo, nnz));
    RT_TRY(dalloc(h, &g.esrc, m));
    RT_TRY(dalloc(h, &g.eslot, m));
    RT_TRY(dalloc(h, &g.st, nnz));
    RT_TRY(dalloc(h, &g.est, m));
    RT_TRY(dalloc(h, &g.sup, m));
    RT_TRY(dalloc(h, &g.sup0, m));
    RT_TRY(dalloc(h, &g.tcnt, m));
    RT_TRY(dalloc(h, &g.pv, n));
    RT_TRY(dalloc(h, &g.pv0, n));
    RT_TRY(dalloc(h, &g.vtcnt, n));
    RT_TRY(dalloc(h, &g.touched, m));
    RT_TRY(dalloc(h, &g.claim, m));
    RT_TRY(dalloc(h, &g.ew, m));
    RT_TRY(dalloc(h, &h->cseed, n));
    RT_TRY(dalloc(h, &h->ctr, 1));
    RT_HIP(hipMemsetAsync(h->ctr, 0, sizeof(Counters), h->stream));
    return RTREX_OK;
}

// Shared tail of both builds: partner slots, edge ids, state reset.
rtrex_status finish(rtrex_graph h, uint32_t* src) {
    GraphDev& g = h->g;
    hipStream_t s = h->stream;
    const uint64_t nnz = g.nnz;
    unsigned* err = &h->ctr->error;
    Scratch sc;
    uint32_t *canon, *rank;
    RT_TRY(sc.get(&canon, nnz));
    RT_TRY(sc.get(&rank, nnz + 1));
    const unsigned bs = blocks_for(nnz, T, 65536);
    k_partner<<<bs, T, 0, s>>>(g.off, g.nb, src, nnz, g.part, canon, err);
    RT_TRY(exclusive_scan_u32_32(canon, rank, nnz, s));
    unsigned hostErr = 0;
    uint32_t ncanon = 0;
    RT_HIP(hipMemcpyAsync(&hostErr, err, sizeof(unsigned),
                          hipMemcpyDeviceToHost, s));
    RT_HIP(hipMemcpyAsync(&ncanon, rank + nnz, sizeof(uint32_t),
                          hipMemcpyDeviceToHost, s));
    RT_HIP(hipStreamSynchronize(s));
    if (hostErr)
        return fail(RTREX_EINVAL, "graph is not a simple undirected CSR:" +
                                      describe(hostErr));
    if ((uint64_t)ncanon * 2 != nnz)
        return fail(RTREX_EINVAL, "graph is not symmetric");
    k_edge_ids<<<bs, T, 0, s>>>(g.nb, src, canon, rank, g.deg, nnz, g.s2e, g.esrc,
                                g.eslot, g.ew);
    k_edge_ids_mirror<<<bs, T, 0, s>>>(g.part, canon, rank, nnz, g.s2e);
    k_pack_slots<<<bs, T, 0, s>>>(g.nb, g.s2e, g.deg, g.off, nnz, g.nbe, g.ndo);
    RT_HIP(hipGetLastError());
    RT_HIP(hipMemsetAsync(g.st, ST_ALIVE, nnz, s));
    RT_HIP(hipMemsetAsync(g.est, ST_ALIVE, g.m, s));
    RT_HIP(hipMemsetAsync(g.touched, 0, g.m, s));
    RT_HIP(hipMemsetAsync(g.claim, 0, (uint64_t)g.m * 4, s));
    RT_HIP(hipStreamSynchronize(s));
    h->phase = PH_BUILT;
    return RTREX_OK;
}

}  // namespace

rtrex_status build_from_csr(rtrex_graph h, const int64_t* off,
                            const int64_t* nb, int64_t n, int64_t nnz) {
    if (n < 0 || nnz < 0 || n >= (int64_t)UINT32_MAX ||
        nnz >= (int64_t)UINT32_MAX)
        return fail(RTREX_EINVAL, "n and nnz must be in [0, 2^32-1)");
    if (nnz % 2) return fail(RTREX_EINVAL, "nnz must be even (undirected)");
    if (!off || (nnz && !nb)) return fail(RTREX_EINVAL, "null CSR array");
    if (off[0] != 0 || off[n] != nnz)
        return fail(RTREX_EINVAL, "offsets must start at 0 and end at nnz");
    std::vector<uint32_t> off32(n + 1), nb32(nnz);
    for (int64_t v = 0; v <= n; ++v) {
        if (v && off[v] < off[v - 1])
            return fail(RTREX_EINVAL, "offsets must be non-decreasing");
        off32[v] = (uint32_t)off[v];
    }
    for (int64_t i = 0; i < nnz; ++i) {
        if (nb[i] < 0 || nb[i] >= n)
            return fail(RTREX_EINVAL, "neighbour id out of range");
        nb32[i] = (uint32_t)nb[i];
    }
    RT_TRY(alloc_state(h, (uint32_t)n, (uint64_t)nnz));
    GraphDev& g = h->g;
    hipStream_t s = h->stream;
    RT_HIP(hipEventRecord(h->ev[0], s));
    RT_HIP(hipMemcpyAsync(g.off, off32.data(), (n + 1) * sizeof(uint32_t),
                          hipMemcpyHostToDevice, s));
    if (nnz)
        RT_HIP(hipMemcpyAsync(g.nb, nb32.data(), nnz * sizeof(uint32_t),
                              hipMemcpyHostToDevice, s));
    Scratch sc;
    uint32_t* src;
    RT_TRY(sc.get(&src, nnz));
    const unsigned bs = blocks_for(std::max<uint64_t>(nnz, n), T, 65536);
    k_slot_src<<<bs, T, 0, s>>>(g.off, g.n, nnz, src);
    k_check_csr<<<bs, T, 0, s>>>(g.off, g.nb, src, g.n, nnz, g.deg,
                                 &h->ctr->error);
    RT_HIP(hipGetLastError());
    RT_TRY(finish(h, src));
    RT_HIP(hipEventRecord(h->ev[1], s));
    RT_HIP(hipEventSynchronize(h->ev[1]));
    float ms = 0;
    RT_HIP(hipEventElapsedTime(&ms, h->ev[0], h->ev[1]));
    h->tm.build_ms = ms;
    return RTREX_OK;
}

// The CSR of an edge list already on the device (du, dv: m entries each;
// the build reads them before any of its own scratch is taken).
rtrex_status build_from_device_edges(rtrex_graph h, const uint32_t* du, const uint32_t* dv,
                                     int64_t n, int64_t m) {
    if (n < 0 || m < 0 || n >= (int64_t)UINT32_MAX || 2 * m >= (int64_t)UINT32_MAX)
        return fail(RTREX_EINVAL, "n and 2m must be in [0, 2^32-1)");
    const uint64_t nnz = 2 * (uint64_t)m;
    if (h->g.nb == nullptr) RT_TRY(alloc_state(h, (uint32_t)n, nnz));
    GraphDev& g = h->g;
    hipStream_t s = h->stream;
    Scratch sc;
    uint32_t *src, *cur, *big;
    RT_TRY(sc.get(&src, nnz));
    RT_TRY(sc.get(&cur, (uint64_t)g.n + 2));
    RT_TRY(sc.get(&big, (uint64_t)g.n + 1));
    unsigned* nbig = cur + g.n;  // one counter after the fill counts
    const unsigned bm = blocks_for(std::max<uint64_t>(m, 1), T, 65536);
    unsigned* err = &h->ctr->error;
    RT_HIP(hipMemsetAsync(g.deg, 0, (uint64_t)g.n * 4, s));
    RT_HIP(hipMemsetAsync(cur, 0, ((uint64_t)g.n + 1) * 4, s));
    if (m) k_edge_degrees<<<bm, T, 0, s>>>(du, dv, m, g.n, g.deg, err);
    RT_HIP(hipGetLastError());
    RT_TRY(exclusive_scan_u32_32(g.deg, g.off, g.n, s));
    if (m) {
        k_edge_fill<<<bm, T, 0, s>>>(du, dv, m, g.n, g.off, cur, g.nb, src);
        k_sort_short<<<blocks_for(g.n, T, 8192), T, 0, s>>>(g.off, g.n, g.nb, err);
        k_sort_long<<<blocks_for(g.n, SL_W * WAVE, 4096), SL_W * WAVE, 0, s>>>(g.off, g.n, g.nb,
                                                                             big, nbig, err);
        RT_HIP(hipGetLastError());
        uint32_t nb_host = 0;
        RT_HIP(hipMemcpyAsync(&nb_host, nbig, 4, hipMemcpyDeviceToHost, s));
        RT_HIP(hipStreamSynchronize(s));
        if (nb_host) {  // lists longer than SL_LDS: one radix sort over all of them
            uint32_t* len;
            uint64_t* pos;
            RT_TRY(sc.get(&len, nb_host));
            RT_TRY(sc.get(&pos, (uint64_t)nb_host + 1));
            k_big_bounds<<<blocks_for(nb_host, T, 1024), T, 0, s>>>(g.off, big, nb_host, len);
            RT_HIP(hipGetLastError());
            RT_TRY(exclusive_scan_u32(len, pos, nb_host, s));
            uint64_t total = 0;
            RT_HIP(hipMemcpyAsync(&total, pos + nb_host, 8, hipMemcpyDeviceToHost, s));
            RT_HIP(hipStreamSynchronize(s));
            uint64_t *keys, *tmp;
            RT_TRY(sc.get(&keys, total));
            RT_TRY(sc.get(&tmp, total));
            const int B = bits_for((uint64_t)n + 1);
            k_big_keys<<<nb_host, T, 0, s>>>(g.off, big, pos, B, g.nb, keys);
            RT_HIP(hipGetLastError());
            RT_TRY(radix_sort_u64(keys, tmp, total, bits_for(nb_host) + B, s));
            k_big_back<<<nb_host, T, 0, s>>>(g.off, big, pos, B, keys, g.nb, err);
            RT_HIP(hipGetLastError());
        }
    }
    return finish(h, src);  // ends with the stream drained
}

rtrex_status build_from_edges(rtrex_graph h, const uint32_t* u,
                              const uint32_t* v, int64_t n, int64_t m) {
    if (n < 0 || m < 0 || n >= (int64_t)UINT32_MAX ||
        2 * m >= (int64_t)UINT32_MAX)
        return fail(RTREX_EINVAL, "n and 2m must be in [0, 2^32-1)");
    if (m && (!u || !v)) return fail(RTREX_EINVAL, "null edge array");
    const uint64_t nnz = 2 * (uint64_t)m;
    RT_TRY(alloc_state(h, (uint32_t)n, nnz));
    hipStream_t s = h->stream;
    Scratch sc;
    uint32_t *du, *dv;
    RT_TRY(sc.get(&du, m));
    RT_TRY(sc.get(&dv, m));
    RT_HIP(hipEventRecord(h->ev[0], s));
    if (m) {
        // Through page-locked staging (from the pin cache, released when the
        // build has drained): the host copies in parallel, the DMA engine
        // runs at link speed instead of the runtime's pageable path.
        const uint64_t bytes = (uint64_t)m * 4;
        if (!h->up_pin) h->up_pin = pin_alloc(2 * bytes);
        if (h->up_pin) {
            // Slices of both arrays, one per host thread; the caller copies
            // slice 0 and then, in slice order, waits for each thread and
            // queues its slice's two transfers, so the DMA of the first
            // slices overlaps the copies of the later ones.
            uint32_t* su = static_cast<uint32_t*>(h->up_pin);
            uint32_t* sv = su + m;
            const unsigned hw = std::max(1u, std::min(8u, std::thread::hardware_concurrency()));
            const uint64_t ns = std::min<uint64_t>(hw, (uint64_t)m / (256u << 10) + 1);
            const uint64_t per = ((uint64_t)m + ns - 1) / ns;
            auto slice = [&](uint64_t t) {
                const uint64_t a = std::min<uint64_t>(m, t * per), b = std::min<uint64_t>(m, a + per);
                if (a < b) {
                    std::memcpy(su + a, u + a, (b - a) * 4);
                    std::memcpy(sv + a, v + a, (b - a) * 4);
                }
            };
            auto queue = [&](uint64_t t) -> hipError_t {
                const uint64_t a = std::min<uint64_t>(m, t * per), b = std::min<uint64_t>(m, a + per);
                if (a >= b) return hipSuccess;
                hipError_t e = hipMemcpyAsync(du + a, su + a, (b - a) * 4, hipMemcpyHostToDevice, s);
                if (e == hipSuccess)
                    e = hipMemcpyAsync(dv + a, sv + a, (b - a) * 4, hipMemcpyHostToDevice, s);
                return e;
            };
            std::vector<std::thread> th;
            for (uint64_t t = 1; t < ns; ++t) th.emplace_back(slice, t);
            slice(0);
            hipError_t e = queue(0);
            for (uint64_t t = 1; t < ns; ++t) {
                th[t - 1].join();
                if (e == hipSuccess) e = queue(t);
            }
            RT_HIP(e);
        } else {
            RT_HIP(hipMemcpyAsync(du, u, bytes, hipMemcpyHostToDevice, s));
            RT_HIP(hipMemcpyAsync(dv, v, bytes, hipMemcpyHostToDevice, s));
        }
    }
    const rtrex_status fs = build_from_device_edges(h, du, dv, n, m);
    if (fs != RTREX_OK) (void)hipStreamSynchronize(s);
    pin_free(h->up_pin);
    h->up_pin = nullptr;
    RT_TRY(fs);
    RT_HIP(hipEventRecord(h->ev[1], s));
    RT_HIP(hipEventSynchronize(h->ev[1]));
    float ms = 0;
    RT_HIP(hipEventElapsedTime(&ms, h->ev[0], h->ev[1]));
    h->tm.build_ms = ms;
    return RTREX_OK;
}

}  // namespace rtrex
