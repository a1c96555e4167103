// Iteration order of a std::unordered_set<int64_t> (libstdc++, as the
// reference's Cluster::vertices, ClusterStructures.h:27-33) filled by inserting
// distinct keys one at a time into an empty set.
//
// The order follows from the container's node list (libstdc++ hashtable.h:
// _M_insert_bucket_begin, _M_rehash_aux for unique keys): the hash of an
// integer is itself, a key goes to bucket key % bucket_count; a node entering
// an empty bucket goes to the front of the list (and the bucket that led the
// list is re-pointed at it), otherwise right after its bucket's predecessor;
// a rehash walks the list in order re-inserting each node the same way.  When
// rehashes happen depends only on the element count, so the bucket count
// after each insertion, bcnt[k] for k elements, is measured once from the
// library itself (bucket_counts) and passed in.
//
// Host and device code; one thread runs it.
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#define RT_USET_HD __host__ __device__
#else
#define RT_USET_HD
#endif

namespace rtrex {

constexpr uint32_t US_NIL = 0xffffffffu;     // end of list
constexpr uint32_t US_EMPTY = 0xfffffffeu;   // empty bucket
constexpr uint32_t US_BEFORE = 0xfffffffdu;  // bucket led by the list head

// keys[0..n) in insertion order -> out[0..n) in iteration order.  Scratch:
// nxt[n], bkt[bcnt[n]] (bcnt is non-decreasing).
RT_USET_HD inline void uset_order(const uint32_t* keys, uint32_t n, const uint32_t* bcnt,
                                           uint32_t* nxt, uint32_t* bkt, uint32_t* out) {
    uint32_t head = US_NIL, nb = 0;
    auto succ = [&](uint32_t at) -> uint32_t { return at == US_BEFORE ? head : nxt[at]; };
    auto set_succ = [&](uint32_t at, uint32_t v) {
        if (at == US_BEFORE) head = v;
        else nxt[at] = v;
    };
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t want = bcnt[i + 1];
        if (want != nb) {  // rehash (before inserting the new node)
            for (uint32_t b = 0; b < want; ++b) bkt[b] = US_EMPTY;
            uint32_t p = head, bbegin = 0;
            head = US_NIL;
            while (p != US_NIL) {
                const uint32_t next = nxt[p];
                const uint32_t b = keys[p] % want;
                if (bkt[b] == US_EMPTY) {
                    nxt[p] = head;
                    head = p;
                    bkt[b] = US_BEFORE;
                    if (nxt[p] != US_NIL) bkt[bbegin] = p;
                    bbegin = b;
                } else {
                    nxt[p] = succ(bkt[b]);
                    set_succ(bkt[b], p);
                }
                p = next;
            }
            nb = want;
        }
        const uint32_t b = keys[i] % nb;
        if (bkt[b] != US_EMPTY) {
            nxt[i] = succ(bkt[b]);
            set_succ(bkt[b], i);
        } else {
            nxt[i] = head;
            head = i;
            if (nxt[i] != US_NIL) bkt[keys[nxt[i]] % nb] = i;
            bkt[b] = US_BEFORE;
        }
    }
    uint32_t k = 0;
    for (uint32_t p = head; p != US_NIL; p = nxt[p]) out[k++] = keys[p];
}

}  // namespace rtrex
