// Host builds of device-side routines, for the CPU test suite: the candidate
// introsort (introsort.hpp) is compared permutation-for-permutation with the
// standard library's std::sort under the same comparison, and the
// unordered_set iteration order (uset.hpp) with std::unordered_set itself;
// the serving loop's job table (jobtable.hpp) is driven from many threads.
#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstdint>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_set>
#include <vector>

#include "introsort.hpp"
#include "jobtable.hpp"
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

// The serving loop's wait/submit pattern on the job table: `submitters`
// threads add `per` jobs each (the map rehashes many times), a completer
// marks them done, and every id is waited on by two threads at once, one of
// which must be refused.  Waits release the mutex, as the server's do.
// Returns the number of violations (0 expected): a job waited on successfully
// other than exactly once, or a table not empty at the end.
int64_t rtrex_hostcheck_jobtable(int submitters, int per, int waiters) {
    struct J {
        int64_t id = 0;
        bool claimed = false, done = false;
        int waited = 0;
    };
    std::mutex mu;
    std::condition_variable cv;
    rtrex::JobTable<J> tab;
    std::deque<int64_t> todo_done, todo_wait;
    std::vector<J*> all;
    std::atomic<int64_t> bad{0};
    const int64_t total = (int64_t)submitters * per;
    int64_t waits_left = 2 * total;
    std::vector<std::thread> ts;
    for (int s = 0; s < submitters; ++s)
        ts.emplace_back([&] {
            for (int i = 0; i < per; ++i) {
                J* j = new J();
                std::lock_guard<std::mutex> lk(mu);
                const int64_t id = tab.add(j);
                all.push_back(j);
                todo_done.push_back(id);
                todo_wait.push_back(id);
                todo_wait.push_back(id);
                cv.notify_all();
            }
        });
    ts.emplace_back([&] {  // completer
        std::unique_lock<std::mutex> lk(mu);
        for (int64_t k = 0; k < total; ++k) {
            cv.wait(lk, [&] { return !todo_done.empty(); });
            const int64_t id = todo_done.front();
            todo_done.pop_front();
            for (J* j : all)
                if (j->id == id) j->done = true;
            cv.notify_all();
        }
    });
    for (int w = 0; w < waiters; ++w)
        ts.emplace_back([&] {
            std::unique_lock<std::mutex> lk(mu);
            for (;;) {
                cv.wait(lk, [&] { return !todo_wait.empty() || waits_left == 0; });
                if (waits_left == 0) return;
                const int64_t id = todo_wait.front();
                todo_wait.pop_front();
                waits_left--;
                J* j = nullptr;
                const auto c = tab.claim(id, &j);
                if (c != rtrex::JobTable<J>::CLAIMED) continue;  // the other waiter has it
                cv.wait(lk, [&] { return j->done; });
                tab.release(id);
                j->waited++;
                cv.notify_all();
            }
        });
    for (auto& t : ts) t.join();
    for (J* j : all) {
        if (j->waited != 1) bad++;
        delete j;
    }
    if (!tab.map.empty()) bad++;
    return bad.load();
}

}  // extern "C"
