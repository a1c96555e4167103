// Host builds of device-side routines, for the CPU test suite: the candidate
// introsort (introsort.hpp) is compared permutation-for-permutation with the
// standard library's std::sort under the same comparison, and the
// unordered_set iteration order (uset.hpp) with std::unordered_set itself.
#include <algorithm>
#include <cstdint>
#include <unordered_set>
#include <vector>

#include "introsort.hpp"
#include "uset.hpp"

using rtrex::isort::Rec;

extern "C" {

// Sorts (ids[i], w[i]) pairs by weight, greater first, with the device
// routine; writes the resulting id order back into ids.
void rtrex_hostcheck_introsort(uint32_t* ids, const double* w, int64_t n) {
    std::vector<Rec> a(n);
    for (int64_t i = 0; i < n; ++i) a[i] = Rec{ids[i], w[ids[i]]};
    rtrex::isort::sort(a.data(), a.data() + n);
    for (int64_t i = 0; i < n; ++i) ids[i] = a[i].vertex;
}

// Same input through std::sort with the reference's comparison
// (ClusterStructures.h:44-47).
void rtrex_hostcheck_stdsort(uint32_t* ids, const double* w, int64_t n) {
    std::vector<Rec> a(n);
    for (int64_t i = 0; i < n; ++i) a[i] = Rec{ids[i], w[ids[i]]};
    std::sort(a.begin(), a.end(), [](Rec x, Rec y) { return x.wgt > y.wgt; });
    for (int64_t i = 0; i < n; ++i) ids[i] = a[i].vertex;
}

// Bucket count of a std::unordered_set<int64_t> after each of k = 0..kmax
// single insertions of distinct keys into an empty set.
void rtrex_hostcheck_bucket_counts(uint32_t kmax, uint32_t* bcnt) {
    std::unordered_set<int64_t> s;
    bcnt[0] = (uint32_t)s.bucket_count();
    for (uint32_t k = 1; k <= kmax; ++k) {
        s.insert((int64_t)k * 7919);
        bcnt[k] = (uint32_t)s.bucket_count();
    }
}

// Iteration order of keys[0..n) inserted in order: the emulation (emu) and
// std::unordered_set itself (ref).
void rtrex_hostcheck_uset(const uint32_t* keys, uint32_t n, uint32_t* emu, uint32_t* ref) {
    std::vector<uint32_t> bcnt(n + 1);
    rtrex_hostcheck_bucket_counts(n, bcnt.data());
    std::vector<uint32_t> nxt(n + 1), bkt(bcnt[n] + 1);
    rtrex::uset_order(keys, n, bcnt.data(), nxt.data(), bkt.data(), emu);
    std::unordered_set<int64_t> s;
    for (uint32_t i = 0; i < n; ++i) s.insert(keys[i]);
    uint32_t k = 0;
    for (int64_t x : s) ref[k++] = (uint32_t)x;
}

}  // extern "C"
