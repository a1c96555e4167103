// Device-side helpers shared by the kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace rtrex {

constexpr int WAVE = 64;

// Position of x in the sorted range a[lo, hi), or -1.  Same contract as the
// reference's getEdgeBinary (Graph.cpp:130-149), on 32-bit slots.
__device__ __forceinline__ int64_t find_slot(const uint32_t* __restrict__ a,
                                             uint32_t lo, uint32_t hi,
                                             uint32_t x) {
    while (lo < hi) {
        uint32_t mid = lo + ((hi - lo) >> 1);
        uint32_t y = a[mid];
        if (y == x) return mid;
        if (y < x) lo = mid + 1;
        else hi = mid;
    }
    return -1;
}

// First position in a[lo, hi) with a[pos] >= x.
__device__ __forceinline__ uint32_t lower_bound(const uint32_t* __restrict__ a,
                                                uint32_t lo, uint32_t hi,
                                                uint32_t x) {
    while (lo < hi) {
        uint32_t mid = lo + ((hi - lo) >> 1);
        if (a[mid] < x) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// Triangle weight 1/(d_a d_b d_c): the degree product is formed exactly in 64
// bits and converted once (Decomposition.h:97,213,760,785).
__device__ __forceinline__ double tri_w(uint32_t a, uint32_t b, uint32_t c) {
    return 1.0 / (double)((uint64_t)a * (uint64_t)b * (uint64_t)c);
}

// Cleaning threshold eps * 1/(d_a d_b) (Decomposition.h:239-254,395-397).
__device__ __forceinline__ double edge_thr(double eps, uint32_t a, uint32_t b) {
    double ew = 1.0 / (double)((uint64_t)a * (uint64_t)b);
    return eps * ew;
}

// The hardware deals a launch's workgroups to the chip's 8 XCDs (each with
// its own L2) round robin by workgroup id.  This maps workgroup b of an
// n-workgroup launch to a work index such that the workgroups of one XCD take
// one contiguous range of the work, in order: neighbouring work items (which
// share data) then run on the same XCD and share its L2.
constexpr uint32_t N_XCD = 8;
__device__ __forceinline__ uint32_t xcd_swizzle(uint32_t b, uint32_t n) {
    const uint32_t x = b % N_XCD, i = b / N_XCD;
    const uint32_t per = n / N_XCD, rem = n % N_XCD;
    return x * per + min(x, rem) + i;
}

__device__ __forceinline__ uint64_t ballot64(bool p) {
    return __ballot(p);
}

__device__ __forceinline__ int lane_id() { return threadIdx.x & (WAVE - 1); }

// Orders this wave's LDS writes before its later LDS reads by other lanes
// (LDS operations of one wave complete in order; the fences keep the
// compiler from moving accesses across).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// A small result for the host: value and sequence number in one 64-bit
// system-scope store to pinned host memory (ordered after this thread's
// earlier writes).  Called by one lane.
__device__ __forceinline__ void publish(uint64_t* box, uint32_t seq, uint32_t v) {
    __hip_atomic_store(box, ((uint64_t)seq << 32) | v, __ATOMIC_RELEASE,
                       __HIP_MEMORY_SCOPE_SYSTEM);
}

// Slot for one item per calling lane in a list whose length is *ctr: one
// atomic per wave (the lowest active lane reserves for all active lanes).
// Callable from divergent code; every active lane must call it.
__device__ __forceinline__ uint32_t wave_append(unsigned* ctr) {
    const uint64_t act = __ballot(1);
    const int lane = lane_id();
    const int leader = __ffsll((unsigned long long)act) - 1;
    const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(ctr, (unsigned)__popcll(act));
    base = __shfl(base, leader, 64);
    return base + (uint32_t)__popcll(act & below);
}

// Publishes *src (final: earlier kernels of the stream have ended).
__global__ void k_publish(uint64_t* box, uint32_t seq, const unsigned* src);

}  // namespace rtrex
