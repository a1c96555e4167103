// Host builds of device-side routines, for the CPU test suite: the candidate
// introsort (introsort.hpp) is compared permutation-for-permutation with the
// standard library's std::sort under the same comparison.
#include <algorithm>
#include <cstdint>
#include <vector>

#include "introsort.hpp"

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

}  // extern "C"
