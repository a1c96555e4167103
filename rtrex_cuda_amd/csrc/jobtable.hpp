// Job id table of the serving loop (host code, no device calls), kept apart
// so that the CPU test suite can drive it from many threads.
//
// Every member is called with the owner's mutex held.  A waiter claims a job
// once (a second waiter on the same id is refused), waits on the owner's
// condition variable -- which releases the mutex, so submitters may insert
// meanwhile and the map may rehash -- and then erases the job BY KEY: no
// iterator is held across the wait.
#pragma once

#include <cstdint>
#include <unordered_map>

namespace rtrex {

template <class J>  // J has: int64_t id; bool claimed;
struct JobTable {
    std::unordered_map<int64_t, J*> map;
    int64_t next_id = 1;

    int64_t add(J* j) {
        j->id = next_id++;
        map[j->id] = j;
        return j->id;
    }

    enum Claim { CLAIMED, UNKNOWN, TAKEN };

    // Marks the job as being waited on; *out receives it when CLAIMED.
    Claim claim(int64_t id, J** out) {
        auto it = map.find(id);
        if (it == map.end()) return UNKNOWN;
        if (it->second->claimed) return TAKEN;
        it->second->claimed = true;
        *out = it->second;
        return CLAIMED;
    }

    // Removes a claimed job after its wait (the caller then owns it).
    void release(int64_t id) { map.erase(id); }
};

}  // namespace rtrex
