// Device primitives: exclusive scans and an LSD radix sort of u64 keys.
//
// Radix sort: 4-bit digits, one pass per digit.  A tile is 256 threads x 16
// keys held in blocked order (thread t owns keys 16t..16t+15 of its tile), so
// per-tile ranks come from one 4096-entry block scan in digit-major order and
// the sort is stable.  Per pass: tile histograms -> one exclusive scan of the
// digit-major histogram table -> stable scatter.  Keys carry their payload in
// the low bits (e.g. src<<32|dst), so no separate value array is needed.
#include "internal.hpp"

namespace rtrex {
namespace {

constexpr int SCAN_T = 256;
constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_TILE = SCAN_T * SCAN_ITEMS;

// Inclusive scan of v across a 256-thread block; returns the inclusive value,
// *total receives the block sum.
template <class T>
__device__ T block_inclusive_scan(T v, T* total) {
    __shared__ T wsum[SCAN_T / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        T o = __shfl_up(v, d, 64);
        if (lane >= d) v += o;
    }
    if (lane == 63) wsum[wid] = v;
    __syncthreads();
    T pre = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < SCAN_T / 64; ++w) {
        if (w < wid) pre += wsum[w];
        tot += wsum[w];
    }
    __syncthreads();
    *total = tot;
    return v + pre;
}

template <class TI>
__global__ void __launch_bounds__(SCAN_T) k_block_sums(const TI* in, uint64_t n,
                                                       uint64_t* bsum) {
    uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE;
    uint64_t s = 0;
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        uint64_t k = base + (uint64_t)i * SCAN_T + threadIdx.x;
        if (k < n) s += in[k];
    }
    uint64_t tot;
    block_inclusive_scan<uint64_t>(s, &tot);
    if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

// Exclusive scan of bsum[0..nb) in place, single block; bsum[nb] = total.
__global__ void __launch_bounds__(SCAN_T) k_scan_sums(uint64_t* bsum,
                                                      uint64_t nb) {
    uint64_t carry = 0;
    for (uint64_t base = 0; base < nb; base += SCAN_T) {
        uint64_t k = base + threadIdx.x;
        uint64_t v = k < nb ? bsum[k] : 0;
        uint64_t tot;
        uint64_t inc = block_inclusive_scan<uint64_t>(v, &tot);
        if (k < nb) bsum[k] = carry + inc - v;
        carry += tot;
        __syncthreads();
    }
    if (threadIdx.x == 0) bsum[nb] = carry;
}

template <class TI, class TO>
__global__ void __launch_bounds__(SCAN_T) k_block_scan(const TI* in, uint64_t n,
                                                       const uint64_t* bsum,
                                                       TO* out) {
    // Blocked arrangement: thread t owns items t*ITEMS .. t*ITEMS+ITEMS-1.
    uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE +
                    (uint64_t)threadIdx.x * SCAN_ITEMS;
    uint64_t v[SCAN_ITEMS];
    uint64_t s = 0;
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        uint64_t k = base + i;
        v[i] = k < n ? (uint64_t)in[k] : 0;
        s += v[i];
    }
    uint64_t tot;
    uint64_t inc = block_inclusive_scan<uint64_t>(s, &tot);
    uint64_t run = bsum[blockIdx.x] + inc - s;
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        uint64_t k = base + i;
        if (k < n) out[k] = (TO)run;
        run += v[i];
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == SCAN_T - 1)
        out[n] = (TO)bsum[gridDim.x];
}

// Small inputs (n <= SCAN_SMALL): the whole exclusive scan in one workgroup,
// tile after tile with a carry -- one launch instead of three.
constexpr uint64_t SCAN_SMALL = 4 * SCAN_TILE;
template <class TI, class TO>
__global__ void __launch_bounds__(SCAN_T) k_scan_small(const TI* in, uint64_t n, TO* out) {
    uint64_t carry = 0;
    for (uint64_t t0 = 0; t0 < n; t0 += SCAN_TILE) {
        const uint64_t base = t0 + (uint64_t)threadIdx.x * SCAN_ITEMS;
        uint64_t v[SCAN_ITEMS];
        uint64_t s = 0;
        for (int i = 0; i < SCAN_ITEMS; ++i) {
            const uint64_t k = base + i;
            v[i] = k < n ? (uint64_t)in[k] : 0;
            s += v[i];
        }
        uint64_t tot;
        const uint64_t inc = block_inclusive_scan<uint64_t>(s, &tot);
        uint64_t run = carry + inc - s;
        for (int i = 0; i < SCAN_ITEMS; ++i) {
            const uint64_t k = base + i;
            if (k < n) out[k] = (TO)run;
            run += v[i];
        }
        carry += tot;
    }
    if (threadIdx.x == 0) out[n] = (TO)carry;
}

inline uint64_t scan_blocks(uint64_t n) {
    uint64_t nb = (n + SCAN_TILE - 1) / SCAN_TILE;
    return nb ? nb : 1;
}

// Enqueues the scan; bsum must hold scan_blocks(n) + 1 entries.
template <class TI, class TO>
void scan_enqueue(const TI* in, TO* out, uint64_t n, uint64_t* bsum,
                  hipStream_t s) {
    const uint64_t nb = scan_blocks(n);
    k_block_sums<TI><<<(unsigned)nb, SCAN_T, 0, s>>>(in, n, bsum);
    k_scan_sums<<<1, SCAN_T, 0, s>>>(bsum, nb);
    k_block_scan<TI, TO><<<(unsigned)nb, SCAN_T, 0, s>>>(in, n, bsum, out);
}

template <class TI, class TO>
rtrex_status exclusive_scan_impl(const TI* in, TO* out, uint64_t n,
                                 hipStream_t s) {
    if (n <= SCAN_SMALL) {
        k_scan_small<TI, TO><<<1, SCAN_T, 0, s>>>(in, n, out);
        RT_HIP(hipGetLastError());
        return RTREX_OK;
    }
    Scratch sc;
    uint64_t* bsum;
    RT_TRY(sc.get(&bsum, scan_blocks(n) + 1));
    scan_enqueue<TI, TO>(in, out, n, bsum, s);
    RT_HIP(hipGetLastError());
    // Arena scratch is reused in stream order; hipMalloc'ed scratch is freed
    // on return, after the work that reads it.
    if (!tl_arena) RT_HIP(hipStreamSynchronize(s));
    return RTREX_OK;
}

// ---------------------------------------------------------------- radix sort
constexpr int RS_T = 256;
constexpr int RS_ITEMS = 16;
constexpr int RS_TILE = RS_T * RS_ITEMS;
constexpr int RS_DIG = 16;

__global__ void __launch_bounds__(RS_T) k_rs_hist(const uint64_t* keys,
                                                  uint64_t n, int shift,
                                                  uint32_t* hist,
                                                  uint32_t ntiles) {
    __shared__ uint32_t h[RS_DIG];
    if (threadIdx.x < RS_DIG) h[threadIdx.x] = 0;
    __syncthreads();
    uint64_t base = (uint64_t)blockIdx.x * RS_TILE;
    for (int i = 0; i < RS_ITEMS; ++i) {
        uint64_t k = base + (uint64_t)i * RS_T + threadIdx.x;
        if (k < n) atomicAdd(&h[(keys[k] >> shift) & 15], 1u);
    }
    __syncthreads();
    if (threadIdx.x < RS_DIG)
        hist[(uint64_t)threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

__global__ void __launch_bounds__(RS_T) k_rs_scatter(const uint64_t* keys,
                                                     uint64_t* out, uint64_t n,
                                                     int shift,
                                                     const uint64_t* goff,
                                                     uint32_t ntiles) {
    __shared__ uint32_t sc[RS_TILE];  // digit-major counts, then ranks
    const int t = threadIdx.x;
    uint64_t base = (uint64_t)blockIdx.x * RS_TILE + (uint64_t)t * RS_ITEMS;
    uint64_t k[RS_ITEMS];
    uint32_t dg[RS_ITEMS];  // digit per key; RS_DIG marks a padding slot
#pragma unroll
    for (int i = 0; i < RS_ITEMS; ++i) {
        uint64_t idx = base + i;
        k[i] = idx < n ? keys[idx] : 0;
        dg[i] = idx < n ? (uint32_t)((k[i] >> shift) & 15) : RS_DIG;
    }
#pragma unroll
    for (int d = 0; d < RS_DIG; ++d) {
        uint32_t c = 0;
#pragma unroll
        for (int i = 0; i < RS_ITEMS; ++i) c += dg[i] == (uint32_t)d;
        sc[d * RS_T + t] = c;
    }
    __syncthreads();
    // Exclusive scan over the 4096 digit-major entries; thread t owns
    // entries 16t..16t+15.
    uint32_t loc[RS_ITEMS];
    uint32_t s = 0;
#pragma unroll
    for (int i = 0; i < RS_ITEMS; ++i) {
        loc[i] = s;
        s += sc[t * RS_ITEMS + i];
    }
    uint32_t tot;
    uint32_t inc = block_inclusive_scan<uint32_t>(s, &tot);
    uint32_t pre = inc - s;
    __syncthreads();
#pragma unroll
    for (int i = 0; i < RS_ITEMS; ++i) sc[t * RS_ITEMS + i] = pre + loc[i];
    __syncthreads();
    // sc[d*256+t] - sc[d*256] = keys of digit d held by threads before t.
#pragma unroll
    for (int i = 0; i < RS_ITEMS; ++i) {
        const uint32_t d = dg[i];
        if (d < RS_DIG) {
            uint32_t r = 0;
#pragma unroll
            for (int j = 0; j < i; ++j) r += dg[j] == d;
            uint64_t pos = goff[(uint64_t)d * ntiles + blockIdx.x] +
                           (sc[d * RS_T + t] - sc[d * RS_T]) + r;
            out[pos] = k[i];
        }
    }
}

}  // namespace

rtrex_status exclusive_scan_u32(const uint32_t* in, uint64_t* out, uint64_t n,
                                hipStream_t s) {
    return exclusive_scan_impl<uint32_t, uint64_t>(in, out, n, s);
}

rtrex_status exclusive_scan_u32_32(const uint32_t* in, uint32_t* out,
                                   uint64_t n, hipStream_t s) {
    return exclusive_scan_impl<uint32_t, uint32_t>(in, out, n, s);
}

rtrex_status radix_sort_u64(uint64_t* keys, uint64_t* tmp, uint64_t n,
                            int bits, hipStream_t s) {
    if (n <= 1) return RTREX_OK;
    const uint64_t ntiles64 = (n + RS_TILE - 1) / RS_TILE;
    if (ntiles64 > 0xffffffffull / RS_DIG)
        return fail(RTREX_EINVAL, "radix sort input too large");
    const uint32_t ntiles = (uint32_t)ntiles64;
    Scratch sc;
    uint32_t* hist;
    uint64_t* goff;
    uint64_t* bsum;
    const uint64_t nh = (uint64_t)RS_DIG * ntiles;
    RT_TRY(sc.get(&hist, nh));
    RT_TRY(sc.get(&goff, nh + 1));
    RT_TRY(sc.get(&bsum, scan_blocks(nh) + 1));
    uint64_t* a = keys;
    uint64_t* b = tmp;
    int passes = (bits + 3) / 4;
    for (int p = 0; p < passes; ++p) {
        int shift = 4 * p;
        k_rs_hist<<<ntiles, RS_T, 0, s>>>(a, n, shift, hist, ntiles);
        scan_enqueue<uint32_t, uint64_t>(hist, goff, nh, bsum, s);
        k_rs_scatter<<<ntiles, RS_T, 0, s>>>(a, b, n, shift, goff, ntiles);
        RT_HIP(hipGetLastError());
        uint64_t* t = a;
        a = b;
        b = t;
    }
    if (a != keys)
        RT_HIP(hipMemcpyAsync(keys, a, n * sizeof(uint64_t),
                              hipMemcpyDeviceToDevice, s));
    if (!tl_arena) RT_HIP(hipStreamSynchronize(s));
    return RTREX_OK;
}

}  // namespace rtrex
