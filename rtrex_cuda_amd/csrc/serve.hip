// Serving: a request loop over decomposition jobs with HBM-footprint
// admission and continuous batching (no reference counterpart: the
// reference decomposes one graph per process, RTRex_CUDA/main.cu:343-391).
//
// Workers (host threads, one HIP stream each) take jobs from one FIFO queue.
// A worker forms a launch from the head of the queue: the head job alone, or,
// when the head is small, the head and the small jobs queued right behind it
// (same eps, up to an edge and a graph cap) as one disjoint-union batch.  The
// launch is admitted when its footprint -- resident arrays, the upper bound of
// its stored triangle lists (at most half the pool) and its phase
// scratch -- fits in one free range of the server's device pool; every
// allocation of the launch is carved from that range, so a job costs no
// hipMalloc/hipFree (except exactly sized triangle lists that outgrow the
// range, allocated outside the pool).  Launches of different workers
// run concurrently on their streams, and jobs that arrive while others run
// join the next launch that a worker forms.  A launch that does not fit while
// nothing else runs is run outside the pool.  A batch is the batched path:
// per-graph results equal running each graph alone.
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>

#include "internal.hpp"
#include "jobtable.hpp"

using namespace rtrex;

namespace {

struct Job {
    int64_t id = 0;
    std::unique_ptr<uint32_t[]> u, v;  // the submitted edges (m each), copied at submission
    int64_t n = 0, m = 0;
    double eps = 0;
    uint64_t cap = 0;  // upper bound of stored triangle-list records
    std::vector<int64_t> cmap;
    int64_t k = 0;
    rtrex_status st = RTREX_OK;
    std::string err;
    rtrex_job_info info{};
    bool done = false;
    bool claimed = false;  // a waiter holds it (see JobTable)
};

// First-fit ranges of the device pool (offset -> length), coalesced on free.
struct Pool {
    char* base = nullptr;
    uint64_t cap = 0;
    std::map<uint64_t, uint64_t> fr;
    bool take(uint64_t bytes, uint64_t* off) {
        for (auto it = fr.begin(); it != fr.end(); ++it) {
            if (it->second < bytes) continue;
            *off = it->first;
            const uint64_t rest = it->second - bytes;
            const uint64_t o2 = it->first + bytes;
            fr.erase(it);
            if (rest) fr[o2] = rest;
            return true;
        }
        return false;
    }
    void give(uint64_t off, uint64_t bytes) {
        auto it = fr.emplace(off, bytes).first;
        auto nx = std::next(it);
        if (nx != fr.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            fr.erase(nx);
        }
        if (it != fr.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                fr.erase(it);
            }
        }
    }
};

// Rounded to 4 KiB so that every range the pool hands out stays aligned
// (device arrays, atomics included, are carved from the start of a range).
uint64_t launch_bytes(uint64_t n, uint64_t m, uint64_t cap) {
    const uint64_t b = resident_bytes(n, m) + scratch_bytes(n, m) + 16 * (cap + 1) + (64u << 10);
    return (b + 4095) & ~4095ull;
}

}  // namespace

struct rtrex_server_s {
    rtrex_server_config cfg{};
    Pool pool;
    std::vector<hipStream_t> streams;
    std::vector<std::thread> workers;
    std::mutex mu;
    std::condition_variable cv;
    std::deque<Job*> queue;
    std::deque<Job*> intake;  // submitted, footprint not yet computed
    int intake_busy = 0;      // jobs an intake thread is sizing
    std::condition_variable cv_in;
    std::vector<std::thread> intake_threads;
    JobTable<Job> jobs;
    int running = 0;
    bool stop = false;
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    double now() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
};

namespace {

void fail_job(Job* j, rtrex_status st) {
    j->st = st;
    j->err = rtrex_last_error();
}

void run_single(rtrex_server s, Job* j, hipStream_t st, char* slab, uint64_t bytes) {
    rtrex_graph g = nullptr;
    rtrex_status r = create_edges_in(j->u.get(), j->v.get(), j->n, j->m, s->cfg.device, st,
                                     slab, bytes, &g);
    if (r == RTREX_OK) {
        j->cmap.resize((size_t)j->n);
        r = rtrex_decompose(g, j->eps, j->cmap.data(), &j->k);
    }
    if (r != RTREX_OK) fail_job(j, r);
    rtrex_graph_destroy(g);
}

void run_batch(rtrex_server s, std::vector<Job*>& js, hipStream_t st, char* slab,
               uint64_t bytes) {
    const int64_t B = (int64_t)js.size();
    std::vector<int64_t> vtx(B), eoff(B + 1, 0);
    for (int64_t i = 0; i < B; ++i) {
        vtx[i] = js[i]->n;
        eoff[i + 1] = eoff[i] + js[i]->m;
    }
    std::vector<uint32_t> u(eoff[B]), v(eoff[B]);
    for (int64_t i = 0; i < B; ++i) {
        std::copy(js[i]->u.get(), js[i]->u.get() + js[i]->m, u.begin() + eoff[i]);
        std::copy(js[i]->v.get(), js[i]->v.get() + js[i]->m, v.begin() + eoff[i]);
    }
    std::vector<int64_t> vbase;
    std::vector<uint32_t> U, V;
    rtrex_graph g = nullptr;
    rtrex_status r = union_edges(B, vtx.data(), eoff.data(), u.data(), v.data(), vbase, U, V);
    if (r == RTREX_OK)
        r = create_edges_in(U.data(), V.data(), vbase.back(), (int64_t)U.size(), s->cfg.device,
                            st, slab, bytes, &g);
    std::vector<int64_t> cm, kc(B);
    if (r == RTREX_OK) {
        g->seg_vtx = vbase;
        cm.resize((size_t)vbase.back());
        r = rtrex_decompose(g, js[0]->eps, cm.data(), nullptr);
    }
    if (r == RTREX_OK) r = rtrex_batch_cluster_counts(g, kc.data(), B);
    rtrex_graph_destroy(g);
    if (r != RTREX_OK) {
        // One bad graph fails the union: run each job alone to isolate it.
        // The batch's slab range is free again and is reused in turn.
        for (Job* j : js) run_single(s, j, st, slab, bytes);
        return;
    }
    for (int64_t i = 0; i < B; ++i) {
        js[i]->cmap.assign(cm.begin() + vbase[i], cm.begin() + vbase[i + 1]);
        js[i]->k = kc[i];
    }
}

// Upper bound of the stored triangle lists: sum over edges of
// min(d_x, d_y) - 1 (the count phase sizes its term buffer the same way).
uint64_t term_cap(const Job* j) {
    std::vector<uint32_t> deg((size_t)j->n, 0);
    for (int64_t i = 0; i < j->m; ++i) {
        deg[j->u[i]]++;
        deg[j->v[i]]++;
    }
    uint64_t cap = 0;
    for (int64_t i = 0; i < j->m; ++i) {
        const uint32_t d = std::min(deg[j->u[i]], deg[j->v[i]]);
        cap += d ? d - 1 : 0;
    }
    return cap;
}

// Moves submitted jobs to the run queue once their footprint is known, off
// the submitting thread.  Several intake threads size jobs side by side, so a
// job enters the queue when it is sized: a large graph (tens of milliseconds
// of degree counting) does not hold up the small jobs submitted after it.
void intake_main(rtrex_server s) {
    std::unique_lock<std::mutex> lk(s->mu);
    for (;;) {
        s->cv_in.wait(lk, [&] { return s->stop || !s->intake.empty(); });
        if (s->intake.empty()) {
            if (s->stop) return;
            continue;
        }
        Job* j = s->intake.front();
        s->intake.pop_front();
        s->intake_busy++;
        lk.unlock();
        j->cap = term_cap(j);
        lk.lock();
        s->intake_busy--;
        s->queue.push_back(j);
        s->cv.notify_all();
    }
}

void worker_main(rtrex_server s, int w) {
    (void)hipSetDevice(s->cfg.device);
    hipStream_t st = s->streams[w];
    std::unique_lock<std::mutex> lk(s->mu);
    // The launch starting at queue position p: that job, plus the small jobs
    // right behind it.  The stored triangle lists count at most a quarter of
    // the pool: a job whose upper bound is larger (skewed graphs: the bound can
    // be several times the exact size) sizes its lists exactly after a
    // counting pass and takes them from its range when they fit, else from
    // the device cache -- so it runs beside other jobs instead of waiting for
    // the whole server to drain.
    auto form = [&](size_t p, std::vector<Job*>& take) {
        take.assign(1, s->queue[p]);
        const Job* h0 = take[0];
        uint64_t N = h0->n, M = h0->m, Cp = h0->cap;
        if ((int64_t)h0->m <= s->cfg.small_job_edges) {
            for (size_t i = p + 1; i < s->queue.size(); ++i) {
                const Job* j = s->queue[i];
                if ((int64_t)j->m > s->cfg.small_job_edges || j->eps != h0->eps) break;
                if ((int64_t)(M + j->m) > s->cfg.batch_max_edges ||
                    (int64_t)take.size() >= s->cfg.batch_max_graphs)
                    break;
                take.push_back(s->queue[i]);
                N += j->n;
                M += j->m;
                Cp += j->cap;
            }
        }
        return launch_bytes(N, M, std::min<uint64_t>(Cp, s->pool.cap / 4 / sizeof(Term)));
    };
    // Reserved workers: with three workers or more, worker 0 takes only
    // launches headed by a small job, and with four or more, worker 1 only
    // jobs of at most 32 times that many edges -- each the first such job in
    // the queue -- so that small and medium jobs always find a worker while
    // large graphs hold the others for hundreds of milliseconds.
    const int64_t lim = w == 0 && s->cfg.workers >= 3   ? s->cfg.small_job_edges
                        : w == 1 && s->cfg.workers >= 4 ? 32 * s->cfg.small_job_edges
                                                        : INT64_MAX;
    const bool small_only = lim != INT64_MAX;
    auto first_small = [&]() -> size_t {
        const size_t look = std::min<size_t>(s->queue.size(), 256);
        for (size_t p = 0; p < look; ++p)
            if ((int64_t)s->queue[p]->m <= lim) return p;
        return SIZE_MAX;
    };
    std::vector<Job*> take;
    for (;;) {
        s->cv.wait(lk, [&] {
            return s->stop || (small_only ? first_small() != SIZE_MAX : !s->queue.empty());
        });
        const size_t p0 = small_only ? first_small() : (s->queue.empty() ? SIZE_MAX : 0);
        if (p0 == SIZE_MAX) {
            if (s->stop && s->queue.empty() && s->intake.empty() && s->intake_busy == 0) return;
            if (s->stop) s->cv.wait(lk);  // jobs still in the intake or for other workers
            continue;
        }
        uint64_t bytes = form(p0, take);
        size_t at = p0;
        uint64_t off = 0;
        bool pooled = s->pool.take(bytes, &off);
        if (!pooled && s->running > 0) {
            // Admission: the head waits for a running launch to free memory.
            // Meanwhile a small launch further back (at most 1/16 of the pool)
            // that fits now may go first; the head keeps its place and is
            // tried first whenever a launch ends.
            const size_t look = std::min<size_t>(s->queue.size(), 64);
            for (size_t p = p0 + 1; p < look && !pooled; ++p) {
                const uint64_t b = form(p, take);
                if (b > s->pool.cap / 16) continue;
                if (s->pool.take(b, &off)) {
                    pooled = true;
                    bytes = b;
                    at = p;
                }
            }
            if (!pooled) {
                s->cv.wait(lk);
                continue;
            }
        }
        s->queue.erase(s->queue.begin() + at, s->queue.begin() + at + take.size());
        s->running++;
        const double t_start = s->now();
        lk.unlock();
        char* slab = pooled ? s->pool.base + off : nullptr;
        if (take.size() == 1) run_single(s, take[0], st, slab, bytes);
        else run_batch(s, take, st, slab, bytes);
        (void)hipStreamSynchronize(st);
        const double t_end = s->now();
        lk.lock();
        if (pooled) s->pool.give(off, bytes);
        s->running--;
        for (Job* j : take) {
            j->info.start_s = t_start;
            j->info.end_s = t_end;
            j->info.batch_graphs = (int64_t)take.size();
            j->info.bytes = (int64_t)bytes;
            j->info.worker = w;
            j->info.pooled = pooled ? 1 : 0;
            j->done = true;
        }
        s->cv.notify_all();
    }
}

}  // namespace

extern "C" {

rtrex_status rtrex_server_start(const rtrex_server_config* cfg, rtrex_server* out) {
    if (!cfg || !out) return fail(RTREX_EINVAL, "null server config or handle");
    *out = nullptr;
    if (cfg->workers < 1 || cfg->workers > 16)
        return fail(RTREX_EINVAL, "workers must be in [1, 16]");
    if (cfg->hbm_budget_bytes <= 0) return fail(RTREX_EINVAL, "hbm_budget_bytes must be > 0");
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || cfg->device < 0 || cfg->device >= count) {
        (void)hipGetLastError();
        return fail(RTREX_EDEVICE, "no such HIP device");
    }
    hipDeviceProp_t prop;
    RT_HIP(hipGetDeviceProperties(&prop, cfg->device));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(RTREX_EDEVICE, std::string("device is ") + prop.gcnArchName +
                                       ", this build targets gfx950 only");
    RT_HIP(hipSetDevice(cfg->device));
    auto* s = new rtrex_server_s();
    s->cfg = *cfg;
    if (s->cfg.small_job_edges <= 0) s->cfg.small_job_edges = 1 << 16;
    if (s->cfg.batch_max_edges <= 0) s->cfg.batch_max_edges = 1 << 22;
    if (s->cfg.batch_max_graphs <= 0) s->cfg.batch_max_graphs = 1024;
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, (size_t)cfg->hbm_budget_bytes);
    if (e != hipSuccess) {  // blocks cached by earlier handles go back first
        (void)hipGetLastError();
        dev_trim();
        e = hipMalloc(&q, (size_t)cfg->hbm_budget_bytes);
    }
    if (e != hipSuccess) {
        (void)hipGetLastError();
        delete s;
        return fail(RTREX_ENOMEM, "server pool hipMalloc failed");
    }
    s->pool.base = static_cast<char*>(q);
    s->pool.cap = (uint64_t)cfg->hbm_budget_bytes;
    s->pool.fr[0] = s->pool.cap;
    s->streams.resize(cfg->workers, nullptr);
    for (auto& st : s->streams) {
        e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
        if (e != hipSuccess) {
            for (auto& x : s->streams)
                if (x) (void)hipStreamDestroy(x);
            (void)hipFree(q);
            delete s;
            return hip_fail(e, "server stream creation", __FILE__, __LINE__);
        }
    }
    for (int t = 0; t < 4; ++t) s->intake_threads.emplace_back(intake_main, s);
    for (int w = 0; w < cfg->workers; ++w) s->workers.emplace_back(worker_main, s, w);
    *out = s;
    return RTREX_OK;
}

rtrex_status rtrex_server_submit(rtrex_server s, const uint32_t* u, const uint32_t* v,
                                 int64_t n, int64_t m, double eps, int64_t* job) {
    if (!s || !job) return fail(RTREX_EINVAL, "null server or job pointer");
    if (n < 0 || m < 0 || n >= (int64_t)UINT32_MAX || 2 * m >= (int64_t)UINT32_MAX)
        return fail(RTREX_EINVAL, "n and 2m must be in [0, 2^32-1)");
    if (m && (!u || !v)) return fail(RTREX_EINVAL, "null edge array");
    if (!(eps > 0)) return fail(RTREX_EINVAL, "eps must be > 0");
    // The edges are copied (the caller may reuse its arrays) and checked in
    // the same pass, on several host threads for a large graph.
    std::unique_ptr<Job> jp(new Job());
    jp->u.reset(new uint32_t[std::max<int64_t>(m, 1)]);
    jp->v.reset(new uint32_t[std::max<int64_t>(m, 1)]);
    std::atomic<bool> bad{false};
    uint32_t* du = jp->u.get();
    uint32_t* dv = jp->v.get();
    host_parallel((uint64_t)m, 1u << 20, [&](uint64_t a, uint64_t b) {
        uint32_t mx = 0;
        for (uint64_t i = a; i < b; ++i) {
            du[i] = u[i];
            dv[i] = v[i];
            mx = std::max(mx, std::max(u[i], v[i]));
        }
        if (b > a && (int64_t)mx >= n) bad = true;
    });
    if (bad) return fail(RTREX_EINVAL, "edge endpoint >= n");
    Job* j = jp.release();
    j->n = n;
    j->m = m;
    j->eps = eps;
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) {
        delete j;
        return fail(RTREX_ESTATE, "server is stopping");
    }
    j->info.submit_s = s->now();
    s->jobs.add(j);
    s->intake.push_back(j);
    *job = j->id;
    s->cv_in.notify_one();
    return RTREX_OK;
}

rtrex_status rtrex_server_wait(rtrex_server s, int64_t job, int64_t* cluster_map,
                               int64_t* n_clusters, rtrex_job_info* info) {
    if (!s) return fail(RTREX_EINVAL, "null server");
    Job* j = nullptr;
    {
        std::unique_lock<std::mutex> lk(s->mu);
        switch (s->jobs.claim(job, &j)) {
            case JobTable<Job>::UNKNOWN: return fail(RTREX_EINVAL, "unknown job id");
            case JobTable<Job>::TAKEN:
                return fail(RTREX_ESTATE, "another thread is already waiting on this job");
            case JobTable<Job>::CLAIMED: break;
        }
        s->cv.wait(lk, [&] { return j->done; });
        s->jobs.release(job);
    }
    const rtrex_status st = j->st;
    if (st != RTREX_OK) {
        set_error(j->err);
    } else {
        if (cluster_map) std::copy(j->cmap.begin(), j->cmap.end(), cluster_map);
        if (n_clusters) *n_clusters = j->k;
    }
    if (info) *info = j->info;
    delete j;
    return st;
}

double rtrex_server_clock(rtrex_server s) { return s ? s->now() : 0.0; }

rtrex_status rtrex_server_stop(rtrex_server s) {
    if (!s) return RTREX_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        s->cv.notify_all();
        s->cv_in.notify_all();
    }
    for (auto& t : s->intake_threads) t.join();
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->cv.notify_all();
    }
    for (auto& t : s->workers) t.join();
    (void)hipSetDevice(s->cfg.device);
    for (auto& st : s->streams) (void)hipStreamDestroy(st);
    (void)hipFree(s->pool.base);
    for (auto& kv : s->jobs.map) delete kv.second;
    delete s;
    return RTREX_OK;
}

}  // extern "C"
