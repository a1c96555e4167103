// C ABI of librtrex_hip.so (declared in include/rtrex.h).
//
// Every entry point validates its arguments, maps HIP failures to status codes
// and records a message for rtrex_last_error(); nothing exits the process
// (the reference's CUDA_CHECK prints and calls exit(1), gpu_common.cuh:31-39).
// The library runs only on a gfx950 device; there is no CPU fallback.
#include <algorithm>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "dev.hpp"
#include "internal.hpp"

namespace rtrex {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

thread_local Arena* tl_arena = nullptr;
thread_local hipStream_t tl_stream = nullptr;

namespace {
// Makes the handle's scratch arena (and its stream) current for one API
// call, growing the arena first (no scratch is outstanding between calls)
// when the last call asked for more than it holds.  If the larger buffer
// cannot be had, calls keep allocating per request.
struct ArenaScope {
    Arena* prev = tl_arena;
    hipStream_t prev_stream = tl_stream;
    explicit ArenaScope(rtrex_graph h) {
        tl_stream = h->stream;
        Arena& a = h->arena;
        if (a.top == 0 && a.demand > a.cap) {
            (void)hipStreamSynchronize(h->stream);
            if (a.base && a.owned) dev_free(a.base);
            a.base = nullptr;
            a.cap = 0;
            a.owned = true;
            const uint64_t want = a.demand + a.demand / 4;
            if (void* q = dev_alloc(want, nullptr)) {
                a.base = static_cast<char*>(q);
                a.cap = want;
            }
        }
        tl_arena = &a;
    }
    ~ArenaScope() {
        tl_arena = prev;
        tl_stream = prev_stream;
    }
};

// The arena sized up front for the largest phase scratch of a graph this
// size (from the cache), so that not even the first calls on a new handle
// allocate scratch per request.
void presize_arena(rtrex_graph h, int64_t n, int64_t m) {
    if (n < 0 || m < 0 || h->arena.base) return;
    const uint64_t want = scratch_bytes((uint64_t)n, (uint64_t)m);
    if (void* q = dev_alloc(want, nullptr)) {
        h->arena.base = static_cast<char*>(q);
        h->arena.cap = want;
        h->arena.owned = true;
    }
}
}  // namespace

__global__ void k_publish(uint64_t* box, uint32_t seq, const unsigned* src) {
    if (threadIdx.x == 0) publish(box, seq, *src);
}

// Mailboxes come from a process-wide pool of pinned, device-mapped chunks
// (never returned to the system): serving opens one handle per job, and a
// pinned allocation per handle would cost more than the waits it saves.
namespace {
struct BoxPool {
    std::mutex mu;
    std::vector<std::pair<uint64_t*, uint64_t*>> free;  // (host, device)
};
BoxPool& box_pool() {
    static BoxPool p;
    return p;
}
constexpr int BOX_CHUNK = 256;  // mailboxes per pinned chunk
}  // namespace

rtrex_status box_open(rtrex_graph h) {
    if (h->box) return RTREX_OK;
    BoxPool& bp = box_pool();
    std::lock_guard<std::mutex> lk(bp.mu);
    if (bp.free.empty()) {
        void* p = nullptr;
        RT_HIP(hipHostMalloc(&p, (uint64_t)BOX_CHUNK * BOX_SLOTS * sizeof(uint64_t),
                             hipHostMallocMapped | hipHostMallocCoherent |
                                 hipHostMallocPortable));
        void* d = nullptr;
        const hipError_t e = hipHostGetDevicePointer(&d, p, 0);
        if (e != hipSuccess) {
            (void)hipHostFree(p);
            return hip_fail(e, "hipHostGetDevicePointer", __FILE__, __LINE__);
        }
        for (int i = BOX_CHUNK - 1; i >= 0; --i)
            bp.free.emplace_back(static_cast<uint64_t*>(p) + (uint64_t)i * BOX_SLOTS,
                                 static_cast<uint64_t*>(d) + (uint64_t)i * BOX_SLOTS);
    }
    h->box = bp.free.back().first;
    h->box_dev = bp.free.back().second;
    bp.free.pop_back();
    // sequence numbers start at 1 for every handle; the last owner's stream
    // was drained when it was destroyed
    for (int i = 0; i < BOX_SLOTS; ++i) __atomic_store_n(h->box + i, 0ull, __ATOMIC_RELEASE);
    h->box_seq = 0;
    return RTREX_OK;
}

void box_close(rtrex_graph h) {
    if (!h->box) return;
    BoxPool& bp = box_pool();
    std::lock_guard<std::mutex> lk(bp.mu);
    bp.free.emplace_back(h->box, h->box_dev);
    h->box = h->box_dev = nullptr;
}

rtrex_status box_wait(rtrex_graph h, int slot, uint32_t seq, uint32_t* out) {
    const uint64_t* b = h->box + slot;
    for (uint32_t it = 1;; ++it) {
        const uint64_t v = __atomic_load_n(b, __ATOMIC_ACQUIRE);
        if ((uint32_t)(v >> 32) == seq) {
            *out = (uint32_t)v;
            return RTREX_OK;
        }
        if ((it & 255) == 0) {  // every few microseconds: is the stream still healthy?
            const hipError_t e = hipStreamQuery(h->stream);
            if (e == hipSuccess) {
                const uint64_t w = __atomic_load_n(b, __ATOMIC_ACQUIRE);
                if ((uint32_t)(w >> 32) == seq) {
                    *out = (uint32_t)w;
                    return RTREX_OK;
                }
                return fail(RTREX_EDEVICE, "a kernel ended without publishing its result");
            }
            if (e != hipErrorNotReady) return hip_fail(e, "stream query", __FILE__, __LINE__);
        }
        __builtin_ia32_pause();
    }
}

rtrex_status fail(rtrex_status code, const std::string& msg) {
    set_error(msg);
    return code;
}

rtrex_status hip_fail(hipError_t e, const char* what, const char* file,
                      int line) {
    std::string m = std::string("HIP error ") + hipGetErrorString(e) + " in " +
                    what + " (" + file + ":" + std::to_string(line) + ")";
    (void)hipGetLastError();
    return fail(e == hipErrorOutOfMemory ? RTREX_ENOMEM : RTREX_EDEVICE, m);
}

namespace {

rtrex_status open_device(int device) {
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count == 0)
        return fail(RTREX_EDEVICE, "no HIP device available (this library "
                                   "runs only on MI355X / gfx950)");
    if (device < 0 || device >= count)
        return fail(RTREX_EINVAL, "device index out of range");
    hipDeviceProp_t prop;
    RT_HIP(hipGetDeviceProperties(&prop, device));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(RTREX_EDEVICE, std::string("device is ") + prop.gcnArchName +
                                       ", this build targets gfx950 only");
    RT_HIP(hipSetDevice(device));
    return RTREX_OK;
}

rtrex_status new_handle(int device, rtrex_graph* out, hipStream_t stream = nullptr) {
    if (!out) return fail(RTREX_EINVAL, "null output handle");
    *out = nullptr;
    RT_TRY(open_device(device));
    rtrex_graph h = new rtrex_graph_s();
    h->device = device;
    // Checking mode: every cleaning decision re-run in the reference's own
    // order (regions of all removed edges), whatever the certification says.
    if (const char* ra = std::getenv("RTREX_REPLAY_ALL")) h->replay_all = std::atoi(ra) != 0;
    if (const char* cw = std::getenv("RTREX_CASCADE_WG_MIN"))
        h->cascade_min = (uint32_t)std::strtoul(cw, nullptr, 10);
    hipError_t e = hipSuccess;
    StreamSet ss;
    if (stream) {
        h->stream = stream;
        h->own_stream = false;
        for (auto& ev : h->ev)
            if (e == hipSuccess) e = hipEventCreate(&ev);
    } else if (stream_take(&ss)) {  // a drained stream of an earlier handle
        h->stream = ss.s;
        for (int i = 0; i < NEV; ++i) h->ev[i] = ss.ev[i];
    } else {
        e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
        for (auto& ev : h->ev)
            if (e == hipSuccess) e = hipEventCreate(&ev);
    }
    if (e != hipSuccess) {
        rtrex_graph_destroy(h);
        return hip_fail(e, "stream/event creation", __FILE__, __LINE__);
    }
    *out = h;
    return RTREX_OK;
}

rtrex_status check(rtrex_graph g) {
    if (!g) return fail(RTREX_EINVAL, "null graph handle");
    RT_HIP(hipSetDevice(g->device));
    return RTREX_OK;
}

// Whole-file Escape parser; the rules of loadGraph_Escape (GraphIO.cpp:19-74).
struct EscapeFile {
    std::vector<char> buf;
    size_t pos = 0;
    bool next_line(const char** b, const char** e) {
        while (pos < buf.size()) {
            const char* s = buf.data() + pos;
            const char* nl = static_cast<const char*>(
                std::memchr(s, '\n', buf.size() - pos));
            const char* t = nl ? nl : buf.data() + buf.size();
            pos = (size_t)(t - buf.data()) + (nl ? 1 : 0);
            if (s[0] == '#') continue;  // comment line (GraphIO.cpp:34-35)
            const char* q = s;
            while (q < t && std::isspace((unsigned char)*q)) ++q;
            if (q == t) continue;  // blank line (GraphIO.cpp:37-38)
            *b = q;
            *e = t;
            return true;
        }
        return false;
    }
};

bool two_ints(const char* b, const char* e, int64_t* x, int64_t* y) {
    std::string line(b, e);
    char* p = nullptr;
    errno = 0;
    long long a = std::strtoll(line.c_str(), &p, 10);
    if (p == line.c_str() || errno) return false;
    char* q = nullptr;
    long long c = std::strtoll(p, &q, 10);
    if (q == p || errno) return false;
    *x = a;
    *y = c;
    return true;
}

rtrex_status read_file(const char* path, EscapeFile* f) {
    if (!path) return fail(RTREX_EINVAL, "null path");
    FILE* fp = std::fopen(path, "rb");
    if (!fp) return fail(RTREX_EIO, std::string("could not open file ") + path);
    std::fseek(fp, 0, SEEK_END);
    long sz = std::ftell(fp);
    std::fseek(fp, 0, SEEK_SET);
    f->buf.resize(sz > 0 ? (size_t)sz : 0);
    size_t got = sz > 0 ? std::fread(f->buf.data(), 1, (size_t)sz, fp) : 0;
    std::fclose(fp);
    if ((long)got != sz) return fail(RTREX_EIO, "short read");
    return RTREX_OK;
}

}  // namespace
}  // namespace rtrex

namespace rtrex {
// Every array a resident graph allocates (alloc_state, the count plans, the
// clean's near stamps and survivor bits, the seed order and cluster ids),
// each rounded up to 256 bytes.
uint64_t resident_bytes(uint64_t n, uint64_t m) {
    const uint64_t nnz = 2 * m;
    return 37 * nnz + 107 * m + 72 * n + 96 * 256 + 8192 + 64;
}
// The largest phase scratch: the build (edge upload, radix keys, partner
// ranks), the hub plan, the clean (frontiers and buckets) or the extraction
// (component layout, candidates, worklist), plus sort/scan block sums.
uint64_t scratch_bytes(uint64_t n, uint64_t m) {
    const uint64_t nnz = 2 * m;
    const uint64_t build = 28 * nnz + 8 * m;
    const uint64_t hubs = 8 * nnz + 4 * n;
    const uint64_t clean = 128 * m;  // frontiers, buckets, alive lists, undecided list, exact replay
    const uint64_t extract = 160 * n + 4 * nnz;
    return std::max(std::max(build, hubs), std::max(clean, extract)) + nnz / 16 +
           (4ull << 20) + 64 * 256;
}
}  // namespace rtrex

namespace rtrex {

rtrex_status create_edges_in(const uint32_t* u, const uint32_t* v, int64_t n, int64_t m,
                             int device, hipStream_t stream, char* slab, uint64_t slab_cap,
                             rtrex_graph* out) {
    rtrex_graph h = nullptr;
    RT_TRY(new_handle(device, &h, stream));
    rtrex_status st = RTREX_OK;
    if (slab) {
        const uint64_t pad = (256 - ((uintptr_t)slab & 255)) & 255;  // keep 256-byte alignment
        h->slab = slab + pad;
        h->slab_cap = slab_cap > pad ? slab_cap - pad : 0;
        // The phase scratch (the build's included) is carved up front.
        const uint64_t sb = (n >= 0 && m >= 0) ? scratch_bytes((uint64_t)n, (uint64_t)m) : 0;
        if (void* q = slab_take(h, sb)) {
            h->arena.base = static_cast<char*>(q);
            h->arena.cap = (sb + 255) & ~255ull;
            h->arena.owned = false;
        }
        ArenaScope scope_(h);
        st = build_from_edges(h, u, v, n, m);
    } else {
        presize_arena(h, n, m);
        ArenaScope scope_(h);
        st = build_from_edges(h, u, v, n, m);
    }
    if (st != RTREX_OK) {
        std::string msg = g_last_error;
        rtrex_graph_destroy(h);
        set_error(msg);
        return st;
    }
    *out = h;
    return RTREX_OK;
}

// One disjoint union: graph i's vertices are shifted by the vertices of the
// graphs before it.  Triangles, components and seeds never cross graphs, so
// every phase runs over all graphs per launch.
rtrex_status union_edges(int64_t n_graphs, const int64_t* vtx, const int64_t* eoff,
                         const uint32_t* u, const uint32_t* v, std::vector<int64_t>& vbase,
                         std::vector<uint32_t>& U, std::vector<uint32_t>& V) {
    if (n_graphs <= 0 || !vtx || !eoff) return fail(RTREX_EINVAL, "bad batch description");
    vbase.assign(n_graphs + 1, 0);
    for (int64_t i = 0; i < n_graphs; ++i) {
        if (vtx[i] < 0 || eoff[i + 1] < eoff[i]) return fail(RTREX_EINVAL, "bad batch sizes");
        vbase[i + 1] = vbase[i] + vtx[i];
    }
    const int64_t N = vbase[n_graphs], M = eoff[n_graphs] - eoff[0];
    if (N >= (int64_t)UINT32_MAX) return fail(RTREX_EINVAL, "batch too large");
    if (M && (!u || !v)) return fail(RTREX_EINVAL, "null edge array");
    U.resize(M);
    V.resize(M);
    for (int64_t i = 0; i < n_graphs; ++i)
        for (int64_t k = eoff[i]; k < eoff[i + 1]; ++k) {
            if ((int64_t)u[k] >= vtx[i] || (int64_t)v[k] >= vtx[i])
                return fail(RTREX_EINVAL, "edge endpoint outside its graph");
            U[k - eoff[0]] = (uint32_t)(u[k] + vbase[i]);
            V[k - eoff[0]] = (uint32_t)(v[k] + vbase[i]);
        }
    return RTREX_OK;
}

}  // namespace rtrex

using namespace rtrex;

extern "C" {

const char* rtrex_last_error(void) { return g_last_error.c_str(); }

const char* rtrex_version(void) {
    return "rtrex-mi355x 1.0 (gfx950; count/clean/extract HIP kernels)";
}

rtrex_status rtrex_escape_header(const char* path, int64_t* n, int64_t* m) {
    EscapeFile f;
    RT_TRY(read_file(path, &f));
    const char *b, *e;
    int64_t x, y;
    if (!f.next_line(&b, &e) || !two_ints(b, e, &x, &y) || x < 0 || y < 0)
        return fail(RTREX_EIO, "missing or malformed header line 'n m'");
    if (n) *n = x;
    if (m) *m = y;
    return RTREX_OK;
}

rtrex_status rtrex_escape_read(const char* path, int64_t m, uint32_t* u,
                               uint32_t* v) {
    EscapeFile f;
    RT_TRY(read_file(path, &f));
    const char *b, *e;
    int64_t x, y;
    if (!f.next_line(&b, &e) || !two_ints(b, e, &x, &y))
        return fail(RTREX_EIO, "missing header");
    if (m != y) return fail(RTREX_EINVAL, "m does not match the header");
    int64_t i = 0;
    while (i < m && f.next_line(&b, &e)) {
        if (!two_ints(b, e, &x, &y))
            return fail(RTREX_EIO, "edge line " + std::to_string(i) +
                                       " does not start with two integers");
        if (x < 0 || y < 0 || x > UINT32_MAX || y > UINT32_MAX)
            return fail(RTREX_EINVAL, "vertex id outside uint32");
        u[i] = (uint32_t)x;
        v[i] = (uint32_t)y;
        ++i;
    }
    if (i < m)  // GraphIO.cpp:67-71
        return fail(RTREX_EIO, "expected " + std::to_string(m) + " edges, only got " +
                                   std::to_string(i));
    return RTREX_OK;
}

rtrex_status rtrex_graph_create_csr(const int64_t* offsets,
                                    const int64_t* nbors, int64_t n,
                                    int64_t nnz, int device, rtrex_graph* out) {
    rtrex_graph h = nullptr;
    RT_TRY(new_handle(device, &h));
    presize_arena(h, n, nnz / 2);
    rtrex_status s;
    {
        ArenaScope scope_(h);
        s = build_from_csr(h, offsets, nbors, n, nnz);
    }
    if (s != RTREX_OK) {
        std::string msg = g_last_error;
        rtrex_graph_destroy(h);
        set_error(msg);
        return s;
    }
    *out = h;
    return RTREX_OK;
}

rtrex_status rtrex_graph_create_edges(const uint32_t* u, const uint32_t* v,
                                      int64_t n, int64_t m, int device,
                                      rtrex_graph* out) {
    rtrex_graph h = nullptr;
    RT_TRY(new_handle(device, &h));
    presize_arena(h, n, m);
    rtrex_status s;
    {
        ArenaScope scope_(h);
        s = build_from_edges(h, u, v, n, m);
    }
    if (s != RTREX_OK) {
        std::string msg = g_last_error;
        rtrex_graph_destroy(h);
        set_error(msg);
        return s;
    }
    *out = h;
    return RTREX_OK;
}

rtrex_status rtrex_graph_load(const char* path, int format, int device, rtrex_graph* out) {
    rtrex_graph h = nullptr;
    RT_TRY(new_handle(device, &h));
    rtrex_status s;
    {
        ArenaScope scope_(h);
        s = load_text(h, path, format);
    }
    if (s != RTREX_OK) {
        std::string msg = g_last_error;
        rtrex_graph_destroy(h);
        set_error(msg);
        return s;
    }
    *out = h;
    return RTREX_OK;
}

rtrex_status rtrex_snap_convert(const char* in_path, const char* out_path, int device,
                                int64_t* n, int64_t* m) {
    if (!in_path || !out_path) return fail(RTREX_EINVAL, "null path");
    rtrex_graph h = nullptr;  // a handle for its stream and scratch only
    RT_TRY(new_handle(device, &h));
    rtrex_status s;
    {
        ArenaScope scope_(h);
        s = convert_text(h, in_path, out_path, n, m);
    }
    std::string msg = g_last_error;
    rtrex_graph_destroy(h);
    if (s != RTREX_OK) set_error(msg);
    return s;
}

void rtrex_graph_destroy(rtrex_graph g) {
    if (!g) return;
    (void)hipSetDevice(g->device);
    if (g->stream) (void)hipStreamSynchronize(g->stream);
    // Everything goes back to the process-wide caches (the stream drained).
    for (void* p : g->allocs) dev_free(p);
    if (g->arena.base && g->arena.owned) dev_free(g->arena.base);
    pin_free(g->cm_pin);
    pin_free(g->sb_pin);
    pin_free(g->up_pin);
    box_close(g);  // after the stream drained above
    bool pooled = false;
    if (g->stream && g->own_stream) {
        bool all = true;
        for (auto& ev : g->ev) all = all && ev;
        if (all && hipStreamQuery(g->stream) == hipSuccess) {
            StreamSet ss;
            ss.s = g->stream;
            for (int i = 0; i < NEV; ++i) ss.ev[i] = g->ev[i];
            stream_give(ss);
            pooled = true;
        }
        (void)hipGetLastError();
    }
    if (!pooled) {
        for (auto& ev : g->ev)
            if (ev) (void)hipEventDestroy(ev);
        if (g->stream && g->own_stream) (void)hipStreamDestroy(g->stream);
    }
    delete g;
}

rtrex_status rtrex_graph_info(rtrex_graph g, int64_t* n, int64_t* nnz) {
    RT_TRY(check(g));
    if (n) *n = g->g.n;
    if (nnz) *nnz = (int64_t)g->g.nnz;
    return RTREX_OK;
}

rtrex_status rtrex_graph_get_csr(rtrex_graph g, int64_t* offsets,
                                 int64_t* nbors) {
    RT_TRY(check(g));
    std::vector<uint32_t> o(g->g.n + 1ull), b(g->g.nnz);
    RT_HIP(hipMemcpyAsync(o.data(), g->g.off, o.size() * 4, hipMemcpyDeviceToHost, g->stream));
    if (g->g.nnz)
        RT_HIP(hipMemcpyAsync(b.data(), g->g.nb, b.size() * 4, hipMemcpyDeviceToHost, g->stream));
    RT_HIP(hipStreamSynchronize(g->stream));
    if (offsets) std::copy(o.begin(), o.end(), offsets);
    if (nbors) std::copy(b.begin(), b.end(), nbors);
    return RTREX_OK;
}


int64_t rtrex_graph_footprint_bytes(int64_t n, int64_t nnz) {
    if (n < 0 || nnz < 0) return -1;
    return (int64_t)(resident_bytes((uint64_t)n, (uint64_t)nnz / 2) +
                     scratch_bytes((uint64_t)n, (uint64_t)nnz / 2));
}

rtrex_status rtrex_count(rtrex_graph g, int64_t* total) {
    RT_TRY(check(g));
    ArenaScope scope_(g);
    RT_TRY(run_count(g));
    unsigned long long t3 = 0;
    RT_HIP(hipMemcpyAsync(&t3, &g->ctr->tri_total, 8, hipMemcpyDeviceToHost, g->stream));
    RT_HIP(hipStreamSynchronize(g->stream));
    if (total) *total = (int64_t)(t3 / 3);
    return RTREX_OK;
}

rtrex_status rtrex_get_supports(rtrex_graph g, double* perEdge,
                                double* perVertex, int64_t* total) {
    RT_TRY(check(g));
    ArenaScope scope_(g);
    if (g->phase < PH_COUNTED) return fail(RTREX_ESTATE, "count has not run");
    if (perVertex) RT_TRY(ensure_vertex_supports(g, nullptr, 0));
    hipStream_t s = g->stream;
    if (perEdge && g->g.nnz) {
        std::vector<double> sup(g->g.m);
        std::vector<uint32_t> s2e(g->g.nnz);
        RT_HIP(hipMemcpyAsync(sup.data(), g->g.sup, sup.size() * 8, hipMemcpyDeviceToHost, s));
        RT_HIP(hipMemcpyAsync(s2e.data(), g->g.s2e, s2e.size() * 4, hipMemcpyDeviceToHost, s));
        RT_HIP(hipStreamSynchronize(s));
        for (uint64_t i = 0; i < g->g.nnz; ++i) perEdge[i] = sup[s2e[i]];
    }
    if (perVertex && g->g.n)
        RT_HIP(hipMemcpyAsync(perVertex, g->g.pv, g->g.n * 8ull, hipMemcpyDeviceToHost, s));
    unsigned long long c[2] = {0, 0};
    RT_HIP(hipMemcpyAsync(c, &g->ctr->tri_total, 16, hipMemcpyDeviceToHost, s));
    RT_HIP(hipStreamSynchronize(s));
    if (total) *total = (int64_t)(c[0] / 3 - c[1]);
    return RTREX_OK;
}

rtrex_status rtrex_initial_clean(rtrex_graph g, double eps) {
    RT_TRY(check(g));
    ArenaScope scope_(g);
    return run_initial_clean(g, eps);
}

rtrex_status rtrex_get_edge_status(rtrex_graph g, uint8_t* status) {
    RT_TRY(check(g));
    if (!status) return fail(RTREX_EINVAL, "null status array");
    std::vector<uint8_t> st(g->g.nnz);
    if (g->g.nnz)
        RT_HIP(hipMemcpyAsync(st.data(), g->g.st, st.size(), hipMemcpyDeviceToHost, g->stream));
    RT_HIP(hipStreamSynchronize(g->stream));
    for (uint64_t i = 0; i < g->g.nnz; ++i)
        status[i] = st[i] == ST_ALIVE ? 'Y' : st[i] == ST_DEAD ? 'N' : 'S';
    return RTREX_OK;
}

rtrex_status rtrex_host_alloc(size_t bytes, void** out) {
    if (!out) return fail(RTREX_EINVAL, "null output pointer");
    *out = nullptr;
    if (!bytes) return RTREX_OK;
    RT_HIP(hipHostMalloc(out, bytes, hipHostMallocMapped | hipHostMallocPortable));
    return RTREX_OK;
}

void rtrex_host_free(void* p) {
    if (p) (void)hipHostFree(p);
}

namespace {
// The survivor bitmap written by the device straight into page-locked host
// memory (rtrex_host_alloc), as the cluster ids are: no copy-engine transfer.
constexpr int WT = 256;
__global__ void __launch_bounds__(WT) k_words_out(const uint32_t* __restrict__ src, uint64_t n,
                                                  uint32_t* __restrict__ dst) {
    const uint64_t stride = (uint64_t)gridDim.x * WT;
    for (uint64_t i = (uint64_t)blockIdx.x * WT + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}
}  // namespace

rtrex_status rtrex_get_survivors(rtrex_graph g, uint32_t* bits, int64_t nwords) {
    RT_TRY(check(g));
    if (!g->surv) return fail(RTREX_ESTATE, "initial clean has not run");
    const int64_t need = ((int64_t)g->g.m + 31) / 32;
    if (!bits || nwords < need) return fail(RTREX_EINVAL, "survivor buffer too small");
    void* dbits = nullptr;
    if (need && mapped_host(bits, &dbits)) {
        k_words_out<<<blocks_for((uint64_t)need, WT, 1024), WT, 0, g->stream>>>(
            g->surv, (uint64_t)need, static_cast<uint32_t*>(dbits));
        RT_HIP(hipGetLastError());
    } else if (need) {
        RT_HIP(hipMemcpyAsync(bits, g->surv, need * 4, hipMemcpyDeviceToHost, g->stream));
    }
    RT_HIP(hipStreamSynchronize(g->stream));
    return RTREX_OK;
}

rtrex_status rtrex_delete_and_clean(rtrex_graph g, const int64_t* pairs,
                                    int64_t k, double eps) {
    RT_TRY(check(g));
    ArenaScope scope_(g);
    return run_delete_and_clean(g, pairs, k, eps);
}

rtrex_status rtrex_extract_inner(rtrex_graph g, int64_t u, double eps, int64_t* candidates,
                                 double* cand_wgt, int64_t max_cand, int64_t* n_cand,
                                 double* internal, double* potential) {
    RT_TRY(check(g));
    ArenaScope scope_(g);
    return run_extract_inner(g, u, eps, candidates, cand_wgt, max_cand, n_cand, internal,
                             potential);
}

rtrex_status rtrex_cluster_edges(rtrex_graph g, const int64_t* verts, int64_t n_verts,
                                 int64_t* pairs, int64_t max_pairs, int64_t* n_pairs) {
    RT_TRY(check(g));
    ArenaScope scope_(g);
    return run_cluster_edges(g, verts, n_verts, pairs, max_pairs, n_pairs);
}

namespace {
// Extraction whose every decision is the reference's.  A cascade decision the
// certified post-clean supports cannot settle is recorded; the clean then
// runs again with those edges' triangle-connected regions replayed in the
// reference's order (their supports become the reference's bits) and the
// extraction runs again -- at most three times.  The regions are remembered
// for later decompositions at the same eps.
rtrex_status extract_exact(rtrex_graph g, double eps) {
    if (g->force_eps != eps) {
        g->force_replay.clear();
        g->force_all = false;
        g->force_eps = eps;
    }
    RT_TRY(run_extract(g, eps));
    for (int tries = 0; tries < 3 && g->zt && (!g->x_undec.empty() || g->x_undec_all); ++tries) {
        if (g->x_undec_all) g->force_all = true;
        g->force_replay.insert(g->force_replay.end(), g->x_undec.begin(), g->x_undec.end());
        std::sort(g->force_replay.begin(), g->force_replay.end());
        g->force_replay.erase(std::unique(g->force_replay.begin(), g->force_replay.end()),
                              g->force_replay.end());
        if (g->force_replay.size() > g->g.m) g->force_all = true;
        // The deletions made since the initial clean are made again, call by
        // call, after the clean re-runs (it clears them).
        const std::vector<uint32_t> la = g->dlog_a, lb = g->dlog_b, lc = g->dlog_c;
        RT_TRY(run_count(g));
        RT_TRY(run_initial_clean(g, eps));
        for (size_t i = 0; i < lc.size();) {
            std::vector<int64_t> pr;
            size_t j = i;
            for (; j < lc.size() && lc[j] == lc[i]; ++j) {
                pr.push_back(la[j]);
                pr.push_back(lb[j]);
            }
            RT_TRY(run_delete_and_clean(g, pr.data(), (int64_t)(j - i), eps));
            i = j;
        }
        RT_TRY(run_extract(g, eps));
    }
    return RTREX_OK;
}
}  // namespace

rtrex_status rtrex_extract(rtrex_graph g, double eps, int64_t* cluster_map,
                           int64_t* n_clusters) {
    RT_TRY(check(g));
    ArenaScope scope_(g);
    if (g->phase >= PH_CLEANED && g->eps != eps)
        return fail(RTREX_EINVAL, "extract needs the eps of the initial clean");
    RT_TRY(extract_exact(g, eps));
    if (cluster_map) RT_TRY(copy_cluster_map(g, cluster_map));
    if (n_clusters) *n_clusters = g->n_clusters;
    return RTREX_OK;
}

rtrex_status rtrex_decompose(rtrex_graph g, double eps, int64_t* cluster_map,
                             int64_t* n_clusters) {
    RT_TRY(check(g));
    ArenaScope scope_(g);
    if (g->force_eps != eps) {  // regions remembered for another eps
        g->force_replay.clear();
        g->force_all = false;
        g->force_eps = eps;
    }
    RT_TRY(run_count(g));
    RT_TRY(run_initial_clean(g, eps));
    RT_TRY(extract_exact(g, eps));
    if (cluster_map) RT_TRY(copy_cluster_map(g, cluster_map));
    if (n_clusters) *n_clusters = g->n_clusters;
    return RTREX_OK;
}

namespace {
rtrex_count_params to_abi(const CountParams& c) {
    rtrex_count_params p;
    p.hub_min = (int32_t)c.hub_min;
    p.hub_max = (int32_t)c.hub_max;
    p.l0_max = (int32_t)c.l0_max;
    p.l1_max = (int32_t)c.l1_max;
    for (int i = 0; i < 4; ++i) p.g_thr[i] = (int32_t)c.g_thr[i];
    p.unit_min = (int32_t)c.unit_min;
    p.unit_hash_mult = (int32_t)c.unit_hash_mult;
    p.l2_max = (int32_t)c.l2_max;
    p.l3_max = (int32_t)c.l3_max;
    return p;
}

rtrex_status from_abi(const rtrex_count_params& p, CountParams* c) {
    if (p.hub_min < 0 || p.hub_min >= p.hub_max || p.hub_max > (int32_t)HUB_MAX_LIMIT)
        return fail(RTREX_EINVAL,
                    "need 0 <= hub_min < hub_max <= " + std::to_string(HUB_MAX_LIMIT));
    if (p.l0_max < 0 || p.l0_max > 64 || p.l1_max < p.l0_max || p.l1_max > 512)
        return fail(RTREX_EINVAL, "need 0 <= l0_max <= 64 and l0_max <= l1_max <= 512");
    if (p.l2_max < p.l1_max || p.l2_max > 4096)
        return fail(RTREX_EINVAL, "need l1_max <= l2_max <= 4096");
    if (p.l3_max < p.l2_max || p.l3_max > 32760)
        return fail(RTREX_EINVAL, "need l2_max <= l3_max <= 32760");
    for (int i = 0; i < 4; ++i)
        if (p.g_thr[i] < 0 || (i && p.g_thr[i] < p.g_thr[i - 1]))
            return fail(RTREX_EINVAL, "g_thr must be non-negative and non-decreasing");
    if (p.unit_min < 1 || p.unit_hash_mult < 1)
        return fail(RTREX_EINVAL, "unit_min and unit_hash_mult must be >= 1");
    c->hub_min = (uint32_t)p.hub_min;
    c->hub_max = (uint32_t)p.hub_max;
    c->l0_max = (uint32_t)p.l0_max;
    c->l1_max = (uint32_t)p.l1_max;
    for (int i = 0; i < 4; ++i) c->g_thr[i] = (uint32_t)p.g_thr[i];
    c->unit_min = (uint32_t)p.unit_min;
    c->unit_hash_mult = (uint32_t)p.unit_hash_mult;
    c->l2_max = (uint32_t)p.l2_max;
    c->l3_max = (uint32_t)p.l3_max;
    return RTREX_OK;
}

// Best (smallest) count-phase device time of `repeats` runs with the
// handle's current parameters.
rtrex_status time_count(rtrex_graph g, int repeats, double* ms) {
    double best = 1e300;
    for (int r = 0; r < repeats; ++r) {
        RT_TRY(run_count(g));
        RT_TRY(resolve_timings(g));
        best = std::min(best, (double)g->tm.count_edge_ms);
    }
    *ms = best;
    return RTREX_OK;
}
}  // namespace

rtrex_status rtrex_get_count_params(rtrex_graph g, rtrex_count_params* out) {
    RT_TRY(check(g));
    if (!out) return fail(RTREX_EINVAL, "null output");
    *out = to_abi(g->cparams);
    return RTREX_OK;
}

rtrex_status rtrex_set_count_params(rtrex_graph g, const rtrex_count_params* p) {
    RT_TRY(check(g));
    if (!p) return fail(RTREX_EINVAL, "null parameters");
    CountParams c;
    RT_TRY(from_abi(*p, &c));
    g->cparams = c;
    g->plan_valid = false;
    return RTREX_OK;
}

rtrex_status rtrex_autotune_count(rtrex_graph g, int repeats, rtrex_count_params* best,
                                  double* best_ms, double* start_ms) {
    RT_TRY(check(g));
    ArenaScope scope_(g);
    if (g->phase < PH_BUILT) return fail(RTREX_ESTATE, "graph not built");
    if (repeats < 1) repeats = 1;
    CountParams cur = g->cparams;
    g->plan_valid = false;
    double t_cur = 0;
    RT_TRY(time_count(g, repeats, &t_cur));
    const double t_start = t_cur;
    auto trial = [&](const CountParams& c, double* t) -> rtrex_status {
        g->cparams = c;
        g->plan_valid = false;
        return time_count(g, repeats, t);
    };
    // Coordinate search: each parameter over its candidates, the others held.
    for (int coord = 0; coord < 6; ++coord) {
        std::vector<CountParams> cand;
        CountParams c = cur;
        switch (coord) {
            case 0:
                for (uint32_t v : {0u, 4u, 8u, 16u, 32u, 64u}) { c.hub_min = v; cand.push_back(c); }
                break;
            case 1:
                for (auto lv : {std::pair<uint32_t, uint32_t>{64, 512}, {32, 512}, {64, 256},
                                {32, 256}}) {
                    c.l0_max = lv.first;
                    c.l1_max = lv.second;
                    cand.push_back(c);
                }
                break;
            case 2:
                for (uint32_t num : {1u, 2u, 4u}) {  // thresholds x 0.5, 1, 2
                    const uint32_t base[4] = {6, 12, 48, 160};
                    for (int i = 0; i < 4; ++i) c.g_thr[i] = base[i] * num / 2;
                    cand.push_back(c);
                }
                break;
            case 3:
                for (uint32_t v : {512u, 1024u, 4096u}) { c.unit_min = v; cand.push_back(c); }
                break;
            case 4:
                for (uint32_t v : {4u, 8u, 16u}) { c.unit_hash_mult = v; cand.push_back(c); }
                break;
            default:
                for (uint32_t v : {1024u, 4096u, 32760u, HUB_MAX_LIMIT}) {
                    c.hub_max = v;
                    if (c.hub_min < v) cand.push_back(c);
                }
                break;
        }
        for (const CountParams& k : cand) {
            double t = 0;
            RT_TRY(trial(k, &t));
            if (t < t_cur) {
                t_cur = t;
                cur = k;
            }
        }
    }
    g->cparams = cur;
    g->plan_valid = false;
    if (best) *best = to_abi(cur);
    if (best_ms) *best_ms = t_cur;
    if (start_ms) *start_ms = t_start;
    return RTREX_OK;
}

rtrex_status rtrex_get_cluster_map(rtrex_graph g, int64_t* cluster_map) {
    RT_TRY(check(g));
    if (!cluster_map) return fail(RTREX_EINVAL, "null cluster map");
    return copy_cluster_map(g, cluster_map);
}

rtrex_status rtrex_get_member_order(rtrex_graph g, int64_t* rank) {
    RT_TRY(check(g));
    if (g->phase < PH_EXTRACTED || !g->mrank)
        return fail(RTREX_ESTATE, "extract has not run");
    if (!rank) return fail(RTREX_EINVAL, "null rank array");
    std::vector<uint32_t> r(g->g.n);
    RT_HIP(hipMemcpyAsync(r.data(), g->mrank, (uint64_t)g->g.n * 4, hipMemcpyDeviceToHost,
                          g->stream));
    RT_HIP(hipStreamSynchronize(g->stream));
    for (uint32_t v = 0; v < g->g.n; ++v) rank[v] = r[v];
    return RTREX_OK;
}

rtrex_status rtrex_cluster_stats(rtrex_graph g, int64_t* n_vertices,
                                 int64_t* n_edges, int64_t* cut) {
    RT_TRY(check(g));
    ArenaScope scope_(g);
    return run_cluster_stats(g, n_vertices, n_edges, cut);
}

rtrex_status rtrex_connected_components(rtrex_graph g, int64_t* comp, int64_t* n_comp) {
    RT_TRY(check(g));
    ArenaScope scope_(g);
    return run_connected_components(g, comp, n_comp);
}

rtrex_status rtrex_cluster_stats_noalloc(rtrex_graph g, int64_t* n_vertices,
                                         int64_t* n_edges, int64_t* cut) {
    RT_TRY(check(g));
    return run_cluster_stats_noalloc(g, n_vertices, n_edges, cut);
}

rtrex_status rtrex_get_audit(rtrex_graph g, rtrex_audit* out) {
    RT_TRY(check(g));
    if (!out) return fail(RTREX_EINVAL, "null output");
    Counters c;
    RT_HIP(hipMemcpyAsync(&c, g->ctr, sizeof(c), hipMemcpyDeviceToHost, g->stream));
    RT_HIP(hipStreamSynchronize(g->stream));
    out->clean_rounds = g->audit.clean_rounds;
    out->edges_removed = (int64_t)c.edges_removed;
    out->triangles_removed = (int64_t)c.tri_removed;
    out->ambiguous_clean = (int64_t)c.ambiguous_clean;
    out->ambiguous_ratio = (int64_t)c.ambiguous_ratio;
    out->seeds = g->phase >= PH_EXTRACTED ? g->n_clusters : 0;
    out->components = g->audit.components;
    out->largest_component = g->audit.largest_component;
    out->replayed_edges = (int64_t)c.replayed_edges;
    return RTREX_OK;
}

rtrex_status rtrex_get_timings(rtrex_graph g, rtrex_timings* out) {
    RT_TRY(check(g));
    if (!out) return fail(RTREX_EINVAL, "null output");
    RT_TRY(resolve_timings(g));
    out->build_ms = g->tm.build_ms;
    out->count_ms = g->tm.count_ms;
    out->clean_ms = g->tm.clean_ms;
    out->extract_ms = g->tm.extract_ms;
    out->count_edge_kernel_ms = g->tm.count_edge_ms;
    out->count_vertex_kernel_ms = g->tm.count_vertex_ms;
    return RTREX_OK;
}


rtrex_status rtrex_batch_create(int64_t n_graphs, const int64_t* vtx,
                                const int64_t* eoff, const uint32_t* u,
                                const uint32_t* v, int device, rtrex_graph* out) {
    if (!out) return fail(RTREX_EINVAL, "null output handle");
    std::vector<int64_t> vbase;
    std::vector<uint32_t> U, V;
    RT_TRY(union_edges(n_graphs, vtx, eoff, u, v, vbase, U, V));
    rtrex_graph h = nullptr;
    RT_TRY(create_edges_in(U.data(), V.data(), vbase.back(), (int64_t)U.size(), device, nullptr,
                           nullptr, 0, &h));
    h->seg_vtx = vbase;
    *out = h;
    return RTREX_OK;
}

rtrex_status rtrex_batch_cluster_counts(rtrex_graph g, int64_t* n_clusters,
                                        int64_t n_graphs) {
    RT_TRY(check(g));
    if (g->seg_vtx.empty()) return fail(RTREX_EINVAL, "not a batch handle");
    if (g->phase < PH_EXTRACTED) return fail(RTREX_ESTATE, "extract has not run");
    if (!n_clusters || n_graphs != (int64_t)g->seg_clusters.size())
        return fail(RTREX_EINVAL, "n_graphs does not match the batch");
    std::copy(g->seg_clusters.begin(), g->seg_clusters.end(), n_clusters);
    return RTREX_OK;
}

rtrex_status rtrex_batch_decompose(int64_t n_graphs, const int64_t* vtx,
                                   const int64_t* eoff, const uint32_t* u,
                                   const uint32_t* v, double eps, int device,
                                   int64_t* cluster_map, int64_t* n_clusters,
                                   rtrex_timings* timings, rtrex_audit* audit) {
    rtrex_graph h = nullptr;
    RT_TRY(rtrex_batch_create(n_graphs, vtx, eoff, u, v, device, &h));
    struct Guard {
        rtrex_graph h;
        ~Guard() {
            std::string msg = g_last_error;
            rtrex_graph_destroy(h);
            set_error(msg);
        }
    } guard{h};
    RT_TRY(rtrex_decompose(h, eps, cluster_map, nullptr));
    if (n_clusters) RT_TRY(rtrex_batch_cluster_counts(h, n_clusters, n_graphs));
    if (timings) RT_TRY(rtrex_get_timings(h, timings));
    if (audit) RT_TRY(rtrex_get_audit(h, audit));
    return RTREX_OK;
}

}  // extern "C"
