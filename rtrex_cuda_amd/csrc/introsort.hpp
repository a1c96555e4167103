// Introsort with the exact element movements of GCC libstdc++'s std::sort
// (the SGI STL design: median-of-three pivot moved to the front, unguarded
// Hoare partition, recursion on the right part with a 2*floor(log2 n) depth
// budget and heap sort when it runs out, chunks of <= 16 left for one final
// insertion sort).  The reference orders extraction candidates with
// std::sort(candidates, ..., wgtCompareDecreasing) (Decomposition.h:829,
// ClusterStructures.h:44-47), which is not stable; equal-weight candidates end
// up in the order this procedure produces, and the greedy prefix depends on
// it.  Reproducing the procedure step for step reproduces that order.
//
// The routine is written once for host and device: the device copy runs on
// one lane over an LDS or global array; the host copy is checked against
// std::sort by the test suite.
#pragma once

#include <cstdint>

#if defined(__HIPCC__)
#define RT_HD __host__ __device__
#else
#define RT_HD
#endif

namespace rtrex {
namespace isort {

// Candidate record: the key the reference compares is the weight, greater
// first.
struct Rec {
    uint32_t vertex;
    double wgt;
};

RT_HD inline bool before(const Rec& a, const Rec& b) { return a.wgt > b.wgt; }

RT_HD inline void swap_rec(Rec* a, Rec* b) {
    Rec t = *a;
    *a = *b;
    *b = t;
}

RT_HD inline int floor_log2(int64_t n) {
    int r = 0;
    while (n > 1) {
        n >>= 1;
        ++r;
    }
    return r;
}

RT_HD inline void move_median_to_first(Rec* result, Rec* a, Rec* b, Rec* c) {
    if (before(*a, *b)) {
        if (before(*b, *c)) swap_rec(result, b);
        else if (before(*a, *c)) swap_rec(result, c);
        else swap_rec(result, a);
    } else if (before(*a, *c)) {
        swap_rec(result, a);
    } else if (before(*b, *c)) {
        swap_rec(result, c);
    } else {
        swap_rec(result, b);
    }
}

RT_HD inline Rec* unguarded_partition(Rec* first, Rec* last, Rec* pivot) {
    while (true) {
        while (before(*first, *pivot)) ++first;
        --last;
        while (before(*pivot, *last)) --last;
        if (!(first < last)) return first;
        swap_rec(first, last);
        ++first;
    }
}

RT_HD inline void push_heap(Rec* first, int64_t hole, int64_t top, Rec value) {
    int64_t parent = (hole - 1) / 2;
    while (hole > top && before(first[parent], value)) {
        first[hole] = first[parent];
        hole = parent;
        parent = (hole - 1) / 2;
    }
    first[hole] = value;
}

RT_HD inline void adjust_heap(Rec* first, int64_t hole, int64_t len, Rec value) {
    const int64_t top = hole;
    int64_t child = hole;
    while (child < (len - 1) / 2) {
        child = 2 * (child + 1);
        if (before(first[child], first[child - 1])) child--;
        first[hole] = first[child];
        hole = child;
    }
    if ((len & 1) == 0 && child == (len - 2) / 2) {
        child = 2 * (child + 1);
        first[hole] = first[child - 1];
        hole = child - 1;
    }
    push_heap(first, hole, top, value);
}

RT_HD inline void heap_sort(Rec* first, Rec* last) {
    const int64_t len = last - first;
    if (len >= 2) {  // make_heap
        int64_t parent = (len - 2) / 2;
        while (true) {
            Rec value = first[parent];
            adjust_heap(first, parent, len, value);
            if (parent == 0) break;
            parent--;
        }
    }
    while (last - first > 1) {  // sort_heap: pop_heap(first, last-1, last-1)
        --last;
        Rec value = *last;
        *last = *first;
        adjust_heap(first, 0, last - first, value);
    }
}

RT_HD inline void unguarded_linear_insert(Rec* last) {
    Rec val = *last;
    Rec* next = last - 1;
    while (before(val, *next)) {
        *last = *next;
        last = next;
        --next;
    }
    *last = val;
}

RT_HD inline void insertion_sort(Rec* first, Rec* last) {
    if (first == last) return;
    for (Rec* i = first + 1; i != last; ++i) {
        if (before(*i, *first)) {
            Rec val = *i;
            for (Rec* p = i; p != first; --p) *p = *(p - 1);  // move_backward
            *first = val;
        } else {
            unguarded_linear_insert(i);
        }
    }
}

constexpr int64_t kThreshold = 16;

// The recursion of the reference's introsort loop, as an explicit stack of
// [first, last) ranges: the right part is sorted before the loop continues on
// the left part, exactly as the recursive call precedes `last = cut`.
RT_HD inline void sort(Rec* first, Rec* last) {
    if (last - first <= 1) return;
    struct Frame {
        Rec* first;
        Rec* last;
        int depth;
    };
    Frame stack[64];
    int sp = 0;
    stack[sp++] = {first, last, 2 * floor_log2(last - first)};
    while (sp) {
        Frame f = stack[--sp];
        Rec* lo = f.first;
        Rec* hi = f.last;
        int depth = f.depth;
        while (hi - lo > kThreshold) {
            if (depth == 0) {
                heap_sort(lo, hi);
                break;
            }
            --depth;
            Rec* mid = lo + (hi - lo) / 2;
            move_median_to_first(lo, lo + 1, mid, hi - 1);
            Rec* cut = unguarded_partition(lo + 1, hi, lo);
            // Recursion on [cut, hi) happens first and completes before the
            // loop resumes on [lo, cut): push the resumption, then the call.
            stack[sp++] = {lo, cut, depth};
            stack[sp++] = {cut, hi, depth};
            break;
        }
    }
    // final insertion sort
    if (last - first > kThreshold) {
        insertion_sort(first, first + kThreshold);
        for (Rec* i = first + kThreshold; i != last; ++i) unguarded_linear_insert(i);
    } else {
        insertion_sort(first, last);
    }
}

}  // namespace isort
}  // namespace rtrex
