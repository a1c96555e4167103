// Process-wide caches of device memory, page-locked host memory and streams
// (host code).
//
// A graph handle holds a few dozen device arrays, a scratch arena and a
// triangle-list buffer, one HIP stream with its timing events, and a
// page-locked staging buffer.  Allocating and freeing them per handle costs
// far more than a decomposition (hipFree of a large buffer waits for the
// device and unmaps its pages; stream creation takes milliseconds), and a
// serving loop or a batch of small graphs pays it per graph.  So released
// blocks go back to a per-device cache and are handed out again: device
// blocks by best fit (a cached block at most a quarter larger than the
// request), page-locked blocks likewise, streams with their events as they
// are.  Every block comes back only after its handle's stream has drained.
// The device cache holds at most CACHE_CAP bytes per device, and a request
// that hipMalloc cannot satisfy first returns every cached block of the
// device to the runtime and tries again.
#include <map>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "internal.hpp"

namespace rtrex {
namespace {

constexpr int MAXDEV = 64;
constexpr uint64_t CACHE_CAP = 48ull << 30;  // per device

struct Block {
    uint64_t bytes;
    int dev;
};

struct Caches {
    std::mutex mu;
    std::multimap<uint64_t, void*> dev_free[MAXDEV];
    uint64_t dev_cached[MAXDEV] = {};
    std::unordered_map<void*, Block> dev_live;  // every block dev_alloc handed out
    std::multimap<uint64_t, void*> pin_free;
    std::unordered_map<void*, uint64_t> pin_live;
    std::vector<StreamSet> streams[MAXDEV];
};

Caches& caches() {
    static Caches* c = new Caches();  // never destroyed: blocks outlive static teardown
    return *c;
}

// Sizes are rounded so that repeated requests of one size class meet.
uint64_t round_bytes(uint64_t b) {
    if (b < 4096) return 4096;
    if (b < (2ull << 20)) {
        uint64_t p = 4096;
        while (p < b) p <<= 1;
        return p;
    }
    return (b + (2ull << 20) - 1) & ~((2ull << 20) - 1);
}

int current_device() {
    int d = 0;
    if (hipGetDevice(&d) != hipSuccess) {
        (void)hipGetLastError();
        d = 0;
    }
    return d < 0 || d >= MAXDEV ? 0 : d;
}

}  // namespace

void* dev_alloc(uint64_t bytes, hipError_t* err) {
    const uint64_t want = round_bytes(bytes);
    const int d = current_device();
    Caches& c = caches();
    {
        std::lock_guard<std::mutex> lk(c.mu);
        auto& fr = c.dev_free[d];
        auto it = fr.lower_bound(want);
        if (it != fr.end() && it->first <= want + want / 4) {
            void* p = it->second;
            c.dev_cached[d] -= it->first;
            c.dev_live[p] = Block{it->first, d};
            fr.erase(it);
            if (err) *err = hipSuccess;
            return p;
        }
    }
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, want);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        dev_trim();
        e = hipMalloc(&p, want);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            if (err) *err = e;
            return nullptr;
        }
    }
    std::lock_guard<std::mutex> lk(c.mu);
    c.dev_live[p] = Block{want, d};
    if (err) *err = hipSuccess;
    return p;
}

void dev_free(void* p) {
    if (!p) return;
    Caches& c = caches();
    std::unique_lock<std::mutex> lk(c.mu);
    auto it = c.dev_live.find(p);
    if (it == c.dev_live.end()) {  // not one of ours
        lk.unlock();
        (void)hipFree(p);
        return;
    }
    const Block b = it->second;
    c.dev_live.erase(it);
    if (c.dev_cached[b.dev] + b.bytes <= CACHE_CAP) {
        c.dev_free[b.dev].emplace(b.bytes, p);
        c.dev_cached[b.dev] += b.bytes;
        return;
    }
    lk.unlock();
    (void)hipFree(p);
}

void dev_trim() {
    const int d = current_device();
    Caches& c = caches();
    std::vector<void*> drop;
    {
        std::lock_guard<std::mutex> lk(c.mu);
        for (auto& kv : c.dev_free[d]) drop.push_back(kv.second);
        c.dev_free[d].clear();
        c.dev_cached[d] = 0;
    }
    for (void* p : drop) (void)hipFree(p);
}

bool dev_cache_fits(uint64_t bytes) {
    const uint64_t want = round_bytes(bytes);
    const int d = current_device();
    Caches& c = caches();
    std::lock_guard<std::mutex> lk(c.mu);
    auto it = c.dev_free[d].lower_bound(want);
    return it != c.dev_free[d].end() && it->first <= want + want / 4;
}

uint64_t dev_cached_bytes() {
    const int d = current_device();
    Caches& c = caches();
    std::lock_guard<std::mutex> lk(c.mu);
    return c.dev_cached[d];
}

void* pin_alloc(uint64_t bytes) {
    const uint64_t want = round_bytes(bytes);
    Caches& c = caches();
    {
        std::lock_guard<std::mutex> lk(c.mu);
        auto it = c.pin_free.lower_bound(want);
        if (it != c.pin_free.end() && it->first <= want + want / 4) {
            void* p = it->second;
            c.pin_live[p] = it->first;
            c.pin_free.erase(it);
            return p;
        }
    }
    void* p = nullptr;
    if (hipHostMalloc(&p, want, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    std::lock_guard<std::mutex> lk(c.mu);
    c.pin_live[p] = want;
    return p;
}

void pin_free(void* p) {
    if (!p) return;
    Caches& c = caches();
    std::lock_guard<std::mutex> lk(c.mu);
    auto it = c.pin_live.find(p);
    if (it == c.pin_live.end()) return;
    c.pin_free.emplace(it->second, p);
    c.pin_live.erase(it);
}

bool stream_take(StreamSet* out) {
    const int d = current_device();
    Caches& c = caches();
    std::lock_guard<std::mutex> lk(c.mu);
    if (c.streams[d].empty()) return false;
    *out = c.streams[d].back();
    c.streams[d].pop_back();
    return true;
}

void stream_give(const StreamSet& s) {
    const int d = current_device();
    Caches& c = caches();
    std::lock_guard<std::mutex> lk(c.mu);
    c.streams[d].push_back(s);
}

}  // namespace rtrex
