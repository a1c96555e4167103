// Edge-list text parsed on the device: the reference's Escape loader and the
// SNAP conversion of its snap_to_rtr.py, with the CSR then built on the
// device from the parsed edges (no host edge list).
//
// Reference: loadGraph_Escape (RTRex_Sequential/GraphIO.cpp:19-88) and
// RTRex_CUDA/snap_to_rtr.py:47-142.
//
// The file is read once into page-locked memory and copied to HBM.  Lines are
// found with two passes over the bytes: every thread counts the newlines of
// its 256-byte chunk, a scan of the counts gives each chunk its first line
// number, and the second pass writes the line starts.  Then one thread per
// line classifies and parses it:
//   Escape (GraphIO.cpp:33-63): '#' in the first column and blank lines are
//     skipped; otherwise two integers as sscanf("%ld%ld") reads them.  The
//     first such line is the header "n m"; the next m are the edges (a line
//     that does not start with two integers is an error, as is a file with
//     fewer edge lines).
//   SNAP (snap_to_rtr.py:63-84): the line stripped; empty lines and lines
//     starting with '#' or '%' skipped; fewer than two whitespace-separated
//     tokens, or a token that is not an integer (Python's int(): sign,
//     digits, single underscores between digits), skipped; self-loops
//     dropped.  Then the ids are remapped to their rank among the distinct
//     ids (snap_to_rtr.py:100-116 remaps by sorted order when the ids are not
//     already 0..n-1, where the rank is the id itself), each undirected pair
//     kept once and the edges sorted (:118-141) -- radix sorts of the ids and
//     of the (min, max) pairs.
// The converted graph can also be written as an Escape file, formatted on the
// device (digit counts, a scan of the line lengths, one thread per edge).
#include <cinttypes>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "dev.hpp"
#include "internal.hpp"

namespace rtrex {
namespace {

constexpr int T = 256;
constexpr uint32_t CH = 256;  // bytes per thread in the line scans

enum : uint8_t { L_SKIP = 0, L_EDGE = 1, L_BAD = 2, L_BIG = 3 };

__global__ void __launch_bounds__(T) k_nl_count(const uint8_t* __restrict__ text, uint64_t size,
                                                uint64_t nchunks, uint32_t* __restrict__ cnt) {
    for (uint64_t c = blockIdx.x * (uint64_t)T + threadIdx.x; c < nchunks;
         c += (uint64_t)gridDim.x * T) {
        const uint64_t a = c * CH, b = min<uint64_t>(size, a + CH);
        uint32_t k = 0;
        if (b - a == CH) {  // whole chunk: 16-byte loads
            const uint4* q = reinterpret_cast<const uint4*>(text + a);
#pragma unroll 4
            for (int i = 0; i < (int)(CH / 16); ++i) {
                const uint4 w = q[i];
                const uint32_t ws[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
                for (int j = 0; j < 4; ++j)
#pragma unroll
                    for (int t = 0; t < 4; ++t) k += ((ws[j] >> (8 * t)) & 0xffu) == '\n';
            }
        } else {
            for (uint64_t i = a; i < b; ++i) k += text[i] == '\n';
        }
        cnt[c] = k;
    }
}

// starts[0] = 0; the line after the k-th newline starts past it.
__global__ void __launch_bounds__(T) k_line_starts(const uint8_t* __restrict__ text, uint64_t size,
                                                   uint64_t nchunks,
                                                   const uint64_t* __restrict__ base,
                                                   uint64_t* __restrict__ starts) {
    for (uint64_t c = blockIdx.x * (uint64_t)T + threadIdx.x; c < nchunks;
         c += (uint64_t)gridDim.x * T) {
        const uint64_t a = c * CH, b = min<uint64_t>(size, a + CH);
        uint64_t k = base[c] + 1;
        for (uint64_t i = a; i < b; ++i)
            if (text[i] == '\n') starts[k++] = i + 1;
        if (c == 0) starts[0] = 0;
    }
}

__device__ __forceinline__ bool c_space(uint8_t c) {  // C isspace
    return c == ' ' || (c >= '\t' && c <= '\r');
}
__device__ __forceinline__ bool py_space(uint8_t c) {  // Python str.split, ASCII
    return c_space(c) || (c >= 0x1c && c <= 0x1f);
}

// The parse keeps no boolean across a loop whose lanes leave at different
// trips: what a loop finds is carried in positions and counts (integers), and
// judged after the loop.  (A flag kept across such a loop exit was read wrong
// for lanes that had left early when other lanes of the wave still iterated.)

// Value of the decimal digits t[a, b) (at most 19 of them, so no wrap).
__device__ __forceinline__ uint64_t digits_value(const uint8_t* t, uint64_t a, uint64_t b) {
    uint64_t v = 0;
    for (uint64_t i = a; i < b; ++i) v = v * 10 + (uint64_t)(t[i] - '0');
    return v;
}

// strtoll-style integer at text[*p, e): optional sign, at least one digit.
// Returns 0 when there is none, 1 for a value, 2 when it does not fit int64.
__device__ __forceinline__ int c_int(const uint8_t* t, uint64_t* p, uint64_t e, int64_t* out) {
    uint64_t i = *p;
    while (i < e && c_space(t[i])) ++i;
    uint64_t neg = 0;
    if (i < e && (t[i] == '+' || t[i] == '-')) {
        neg = t[i] == '-';
        ++i;
    }
    const uint64_t d0 = i;
    while (i < e && t[i] >= '0' && t[i] <= '9') ++i;
    const uint64_t nd = i - d0;
    if (nd == 0) return 0;
    uint64_t z = d0;  // leading zeros do not count towards the width
    while (z + 1 < i && t[z] == '0') ++z;
    if (i - z > 19) return 2;
    const uint64_t v = digits_value(t, z, i);
    if (v > 0x7fffffffffffffffull) return 2;
    *out = neg ? -(int64_t)v : (int64_t)v;
    *p = i;
    return 1;
}

// Python int() of the token t[a, b): [+-]? digit (_? digit)*.  0 when it is
// not an integer, 1 for a value, 2 when it does not fit int64.
__device__ __forceinline__ int py_int(const uint8_t* t, uint64_t a, uint64_t b, int64_t* out) {
    uint64_t neg = 0;
    if (a < b && (t[a] == '+' || t[a] == '-')) {
        neg = t[a] == '-';
        ++a;
    }
    if (a >= b) return 0;
    // counts: digits, underscores not between two digits, other characters
    uint64_t nd = 0, badu = 0, other = 0;
    for (uint64_t i = a; i < b; ++i) {
        const uint8_t c = t[i];
        const uint64_t dig = c >= '0' && c <= '9';
        const uint64_t und = c == '_';
        nd += dig;
        other += 1 - dig - und;
        if (und) {
            const uint64_t before = i > a && t[i - 1] >= '0' && t[i - 1] <= '9';
            const uint64_t after = i + 1 < b && t[i + 1] >= '0' && t[i + 1] <= '9';
            badu += 1 - (before & after);
        }
    }
    if (nd == 0 || other || badu) return 0;
    // the value, skipping underscores and leading zeros
    uint64_t v = 0, width = 0;
    for (uint64_t i = a; i < b; ++i) {
        const uint8_t c = t[i];
        if (c == '_') continue;
        const uint64_t d = c - '0';
        width += (width > 0) | (d != 0);
        if (width <= 19) v = v * 10 + d;
    }
    if (width > 19 || v > 0x7fffffffffffffffull) return 2;
    *out = neg ? -(int64_t)v : (int64_t)v;
    return 1;
}

template <bool SNAP>
__global__ void __launch_bounds__(T) k_parse(const uint8_t* __restrict__ text,
                                             const uint64_t* __restrict__ starts, uint64_t nl,
                                             uint64_t size, uint8_t* __restrict__ kind,
                                             int64_t* __restrict__ A, int64_t* __restrict__ Bv) {
    for (uint64_t l = blockIdx.x * (uint64_t)T + threadIdx.x; l < nl;
         l += (uint64_t)gridDim.x * T) {
        const uint64_t s = starts[l];
        uint64_t e = l + 1 < nl ? starts[l + 1] - 1 : size;  // without its newline
        if (e > size) e = size;
        uint8_t k = L_SKIP;
        int64_t a = 0, b = 0;
        if (!SNAP) {
            uint64_t i = s;
            while (i < e && c_space(text[i])) ++i;  // i == e: a blank line
            if (i < e && text[s] != '#') {
                uint64_t p = s;
                const int ra = c_int(text, &p, e, &a);
                const int rb = ra ? c_int(text, &p, e, &b) : 0;
                k = ra == 0 || rb == 0 ? L_BAD : (ra == 2 || rb == 2 ? L_BIG : L_EDGE);
            }
        } else {
            uint64_t i = s;
            while (i < e && py_space(text[i])) ++i;
            if (i < e && text[i] != '#' && text[i] != '%') {
                const uint64_t a0 = i;
                while (i < e && !py_space(text[i])) ++i;
                const uint64_t a1 = i;
                while (i < e && py_space(text[i])) ++i;
                const uint64_t b0 = i;
                while (i < e && !py_space(text[i])) ++i;
                const uint64_t b1 = i;
                if (b0 < b1) {
                    const int ra = py_int(text, a0, a1, &a);
                    const int rb = py_int(text, b0, b1, &b);
                    if (ra && rb)
                        k = ra == 2 || rb == 2 ? L_BIG : (a == b ? L_SKIP : L_EDGE);  // self-loops dropped
                }
            }
        }
        kind[l] = k;
        A[l] = a;
        Bv[l] = b;
    }
}

__global__ void __launch_bounds__(T) k_first_line(const uint8_t* __restrict__ kind, uint64_t nl,
                                                  unsigned long long* first) {
    for (uint64_t l = blockIdx.x * (uint64_t)T + threadIdx.x; l < nl;
         l += (uint64_t)gridDim.x * T)
        if (kind[l] != L_SKIP) atomicMin(first, (unsigned long long)l);
}

// Flags of the lines that count as edges: Escape, every non-skipped line
// after the header; SNAP, every parsed edge.
__global__ void __launch_bounds__(T) k_edge_flags(const uint8_t* __restrict__ kind, uint64_t nl,
                                                  uint64_t after, bool snap,
                                                  uint32_t* __restrict__ f,
                                                  unsigned long long* big) {
    for (uint64_t l = blockIdx.x * (uint64_t)T + threadIdx.x; l < nl;
         l += (uint64_t)gridDim.x * T) {
        const uint8_t k = kind[l];
        f[l] = snap ? k == L_EDGE : (l > after && k != L_SKIP);
        if (snap && k == L_BIG) atomicOr(big, 1ull);  // an id beyond 64 bits
    }
}

// Escape: the first m edge lines into (u, v); err: [0] first bad line's
// ordinal, [1] first ordinal whose ids do not fit in 32 bits.
__global__ void __launch_bounds__(T) k_take_escape(const uint8_t* __restrict__ kind,
                                                   const int64_t* __restrict__ A,
                                                   const int64_t* __restrict__ Bv,
                                                   const uint32_t* __restrict__ f,
                                                   const uint64_t* __restrict__ ord, uint64_t nl,
                                                   uint64_t m, uint32_t* __restrict__ u,
                                                   uint32_t* __restrict__ v,
                                                   unsigned long long* err) {
    for (uint64_t l = blockIdx.x * (uint64_t)T + threadIdx.x; l < nl;
         l += (uint64_t)gridDim.x * T) {
        if (!f[l]) continue;
        const uint64_t k = ord[l];
        if (k >= m) continue;  // lines beyond m are ignored
        if (kind[l] != L_EDGE) {
            atomicMin(&err[0], (unsigned long long)k);
            continue;
        }
        const int64_t a = A[l], b = Bv[l];
        if (a < 0 || b < 0 || a > 0xffffffffll || b > 0xffffffffll) {
            atomicMin(&err[1], (unsigned long long)k);
            continue;
        }
        u[k] = (uint32_t)a;
        v[k] = (uint32_t)b;
    }
}

// SNAP: the parsed edges, compacted; both endpoints as order-preserving
// unsigned keys (sign bit flipped) for the id sort.
__global__ void __launch_bounds__(T) k_take_snap(const int64_t* __restrict__ A,
                                                 const int64_t* __restrict__ Bv,
                                                 const uint32_t* __restrict__ f,
                                                 const uint64_t* __restrict__ ord, uint64_t nl,
                                                 uint64_t* __restrict__ ea, uint64_t* __restrict__ eb,
                                                 uint64_t* __restrict__ ids, uint64_t ne) {
    for (uint64_t l = blockIdx.x * (uint64_t)T + threadIdx.x; l < nl;
         l += (uint64_t)gridDim.x * T) {
        if (!f[l]) continue;
        const uint64_t k = ord[l];
        const uint64_t a = (uint64_t)A[l] ^ 0x8000000000000000ull;
        const uint64_t b = (uint64_t)Bv[l] ^ 0x8000000000000000ull;
        ea[k] = a;
        eb[k] = b;
        ids[k] = a;
        ids[ne + k] = b;
    }
}

__global__ void __launch_bounds__(T) k_uniq_flags(const uint64_t* __restrict__ x, uint64_t n,
                                                  uint32_t* __restrict__ f) {
    for (uint64_t i = blockIdx.x * (uint64_t)T + threadIdx.x; i < n; i += (uint64_t)gridDim.x * T)
        f[i] = i == 0 || x[i] != x[i - 1];
}

__global__ void __launch_bounds__(T) k_uniq_take(const uint64_t* __restrict__ x, uint64_t n,
                                                 const uint32_t* __restrict__ f,
                                                 const uint64_t* __restrict__ ord,
                                                 uint64_t* __restrict__ out) {
    for (uint64_t i = blockIdx.x * (uint64_t)T + threadIdx.x; i < n; i += (uint64_t)gridDim.x * T)
        if (f[i]) out[ord[i]] = x[i];
}

__device__ __forceinline__ uint32_t rank_of(const uint64_t* __restrict__ u, uint64_t n,
                                            uint64_t key) {
    uint64_t lo = 0, hi = n;
    while (lo < hi) {
        const uint64_t mid = (lo + hi) >> 1;
        if (u[mid] < key) lo = mid + 1;
        else hi = mid;
    }
    return (uint32_t)lo;
}

// Each pair as its ranks (min, max) in one key.
__global__ void __launch_bounds__(T) k_pair_keys(const uint64_t* __restrict__ ea,
                                                 const uint64_t* __restrict__ eb, uint64_t ne,
                                                 const uint64_t* __restrict__ uniq, uint64_t nu,
                                                 int B, uint64_t* __restrict__ keys) {
    for (uint64_t i = blockIdx.x * (uint64_t)T + threadIdx.x; i < ne;
         i += (uint64_t)gridDim.x * T) {
        const uint32_t a = rank_of(uniq, nu, ea[i]), b = rank_of(uniq, nu, eb[i]);
        keys[i] = a < b ? ((uint64_t)a << B) | b : ((uint64_t)b << B) | a;
    }
}

__global__ void __launch_bounds__(T) k_pair_take(const uint64_t* __restrict__ keys, uint64_t n,
                                                 const uint32_t* __restrict__ f,
                                                 const uint64_t* __restrict__ ord, int B,
                                                 uint32_t* __restrict__ u,
                                                 uint32_t* __restrict__ v) {
    for (uint64_t i = blockIdx.x * (uint64_t)T + threadIdx.x; i < n; i += (uint64_t)gridDim.x * T)
        if (f[i]) {
            u[ord[i]] = (uint32_t)(keys[i] >> B);
            v[ord[i]] = (uint32_t)(keys[i] & ((1ull << B) - 1));
        }
}

__device__ __forceinline__ uint32_t ndigits(uint32_t x) {
    uint32_t d = 1;
    while (x >= 10) {
        x /= 10;
        ++d;
    }
    return d;
}

__global__ void __launch_bounds__(T) k_fmt_len(const uint32_t* __restrict__ u,
                                               const uint32_t* __restrict__ v, uint64_t m,
                                               uint32_t* __restrict__ len) {
    for (uint64_t i = blockIdx.x * (uint64_t)T + threadIdx.x; i < m; i += (uint64_t)gridDim.x * T)
        len[i] = ndigits(u[i]) + ndigits(v[i]) + 2;
}

__device__ __forceinline__ void put_u32(uint8_t* o, uint32_t x, uint32_t nd) {
    for (uint32_t k = nd; k-- > 0;) {
        o[k] = (uint8_t)('0' + x % 10);
        x /= 10;
    }
}

// "u v\n" per edge at its offset.
__global__ void __launch_bounds__(T) k_fmt(const uint32_t* __restrict__ u,
                                           const uint32_t* __restrict__ v, uint64_t m,
                                           const uint64_t* __restrict__ off,
                                           uint8_t* __restrict__ out) {
    for (uint64_t i = blockIdx.x * (uint64_t)T + threadIdx.x; i < m; i += (uint64_t)gridDim.x * T) {
        uint8_t* o = out + off[i];
        const uint32_t du = ndigits(u[i]), dv = ndigits(v[i]);
        put_u32(o, u[i], du);
        o[du] = ' ';
        put_u32(o + du + 1, v[i], dv);
        o[du + dv + 1] = '\n';
    }
}

unsigned grid(uint64_t n) { return blocks_for(std::max<uint64_t>(n, 1), T, 65536); }

// The whole file into page-locked memory (read by several threads), then to
// the device; size is the file's, the device copy is zero-padded to CH.
rtrex_status read_to_device(const char* path, Scratch& sc, hipStream_t s, uint8_t** dtext,
                            uint64_t* size) {
    if (!path) return fail(RTREX_EINVAL, "null path");
    FILE* f = std::fopen(path, "rb");
    if (!f) return fail(RTREX_EIO, std::string("could not open file ") + path);
    std::fseek(f, 0, SEEK_END);
    const long sz = std::ftell(f);
    std::fclose(f);
    if (sz < 0) return fail(RTREX_EIO, std::string("cannot size ") + path);
    *size = (uint64_t)sz;
    const uint64_t padded = ((uint64_t)sz + CH) / CH * CH;
    RT_TRY(sc.get(dtext, padded));
    RT_HIP(hipMemsetAsync(*dtext + (uint64_t)sz, 0, padded - (uint64_t)sz, s));
    if (!sz) return RTREX_OK;
    void* pin = pin_alloc((uint64_t)sz);
    if (!pin) return fail(RTREX_ENOMEM, "page-locked staging allocation failed");
    const unsigned nt = (unsigned)std::min<uint64_t>(
        std::max(1u, std::min(8u, std::thread::hardware_concurrency())), (uint64_t)sz / (8 << 20) + 1);
    const uint64_t per = ((uint64_t)sz + nt - 1) / nt;
    std::vector<std::thread> th;
    std::vector<int> ok(nt, 0);
    for (unsigned t = 0; t < nt; ++t)
        th.emplace_back([&, t] {
            const uint64_t a = std::min<uint64_t>((uint64_t)sz, t * per);
            const uint64_t b = std::min<uint64_t>((uint64_t)sz, a + per);
            FILE* g = std::fopen(path, "rb");
            if (!g) return;
            std::fseek(g, (long)a, SEEK_SET);
            ok[t] = std::fread(static_cast<char*>(pin) + a, 1, b - a, g) == b - a;
            std::fclose(g);
        });
    for (auto& x : th) x.join();
    for (int o : ok)
        if (!o) {
            pin_free(pin);
            return fail(RTREX_EIO, std::string("short read of ") + path);
        }
    const hipError_t e = hipMemcpyAsync(*dtext, pin, (uint64_t)sz, hipMemcpyHostToDevice, s);
    (void)hipStreamSynchronize(s);
    pin_free(pin);
    if (e != hipSuccess) return hip_fail(e, "text upload", __FILE__, __LINE__);
    return RTREX_OK;
}

// Parsed edges on the device: (du, dv) and their count and vertex count.
struct Parsed {
    uint32_t *u = nullptr, *v = nullptr;
    int64_t n = 0, m = 0;
};

rtrex_status parse_device(rtrex_graph h, const char* path, int format, Scratch& sc, Parsed* P) {
    hipStream_t s = h->stream;
    const bool snap = format == RTREX_FORMAT_SNAP;
    if (format != RTREX_FORMAT_ESCAPE && !snap) return fail(RTREX_EINVAL, "unknown text format");
    uint8_t* text = nullptr;
    uint64_t size = 0;
    RT_TRY(read_to_device(path, sc, s, &text, &size));
    const uint64_t nch = (size + CH - 1) / CH + 1;
    uint32_t* cnt;
    uint64_t* base;
    RT_TRY(sc.get(&cnt, nch));
    RT_TRY(sc.get(&base, nch + 1));
    k_nl_count<<<grid(nch), T, 0, s>>>(text, size, nch, cnt);
    RT_TRY(exclusive_scan_u32(cnt, base, nch, s));
    uint64_t nnl = 0;
    RT_HIP(hipMemcpyAsync(&nnl, base + nch, 8, hipMemcpyDeviceToHost, s));
    RT_HIP(hipStreamSynchronize(s));
    const uint64_t nl = nnl + 1;  // the text after the last newline is a line too
    uint64_t* starts;
    uint8_t* kind;
    int64_t *A, *Bv;
    uint32_t* f;
    uint64_t* ord;
    unsigned long long* red;
    RT_TRY(sc.get(&starts, nl));
    RT_TRY(sc.get(&kind, nl));
    RT_TRY(sc.get(&A, nl));
    RT_TRY(sc.get(&Bv, nl));
    RT_TRY(sc.get(&f, nl));
    RT_TRY(sc.get(&ord, nl + 1));
    RT_TRY(sc.get(&red, 2));
    k_line_starts<<<grid(nch), T, 0, s>>>(text, size, nch, base, starts);
    if (snap) k_parse<true><<<grid(nl), T, 0, s>>>(text, starts, nl, size, kind, A, Bv);
    else k_parse<false><<<grid(nl), T, 0, s>>>(text, starts, nl, size, kind, A, Bv);
    RT_HIP(hipGetLastError());
    if (!snap) {
        const unsigned long long none = ~0ull;
        unsigned long long hdr = none;
        RT_HIP(hipMemcpyAsync(red, &none, 8, hipMemcpyHostToDevice, s));
        k_first_line<<<grid(nl), T, 0, s>>>(kind, nl, red);
        RT_HIP(hipMemcpyAsync(&hdr, red, 8, hipMemcpyDeviceToHost, s));
        RT_HIP(hipStreamSynchronize(s));
        int64_t hv[2] = {0, 0};
        uint8_t hk = L_BAD;
        if (hdr != none) {
            RT_HIP(hipMemcpyAsync(&hk, kind + hdr, 1, hipMemcpyDeviceToHost, s));
            RT_HIP(hipMemcpyAsync(&hv[0], A + hdr, 8, hipMemcpyDeviceToHost, s));
            RT_HIP(hipMemcpyAsync(&hv[1], Bv + hdr, 8, hipMemcpyDeviceToHost, s));
            RT_HIP(hipStreamSynchronize(s));
        }
        if (hdr == none || hk != L_EDGE || hv[0] < 0 || hv[1] < 0)
            return fail(RTREX_EIO, "missing or malformed header line 'n m'");
        P->n = hv[0];
        P->m = hv[1];
        if (P->n >= (int64_t)UINT32_MAX || 2 * P->m >= (int64_t)UINT32_MAX)
            return fail(RTREX_EINVAL, "n and 2m must be in [0, 2^32-1)");
        k_edge_flags<<<grid(nl), T, 0, s>>>(kind, nl, hdr, false, f, red);
        RT_TRY(exclusive_scan_u32(f, ord, nl, s));
        uint64_t have = 0;
        RT_HIP(hipMemcpyAsync(&have, ord + nl, 8, hipMemcpyDeviceToHost, s));
        RT_TRY(sc.get(&P->u, (uint64_t)P->m));
        RT_TRY(sc.get(&P->v, (uint64_t)P->m));
        const unsigned long long e2[2] = {none, none};
        RT_HIP(hipMemcpyAsync(red, e2, 16, hipMemcpyHostToDevice, s));
        k_take_escape<<<grid(nl), T, 0, s>>>(kind, A, Bv, f, ord, nl, (uint64_t)P->m, P->u, P->v,
                                             red);
        RT_HIP(hipGetLastError());
        unsigned long long err[2];
        RT_HIP(hipMemcpyAsync(err, red, 16, hipMemcpyDeviceToHost, s));
        RT_HIP(hipStreamSynchronize(s));
        if (have < (uint64_t)P->m)  // GraphIO.cpp:67-71
            return fail(RTREX_EIO, "expected " + std::to_string(P->m) + " edges, only got " +
                                       std::to_string(have));
        if (err[0] != none)
            return fail(RTREX_EIO, "edge line " + std::to_string(err[0]) +
                                       " does not start with two integers");
        if (err[1] != none) return fail(RTREX_EINVAL, "vertex id outside uint32");
        return RTREX_OK;
    }
    // SNAP conversion
    RT_HIP(hipMemsetAsync(red, 0, 8, s));
    k_edge_flags<<<grid(nl), T, 0, s>>>(kind, nl, 0, true, f, red);
    RT_TRY(exclusive_scan_u32(f, ord, nl, s));
    uint64_t ne = 0;
    unsigned long long big = 0;
    RT_HIP(hipMemcpyAsync(&ne, ord + nl, 8, hipMemcpyDeviceToHost, s));
    RT_HIP(hipMemcpyAsync(&big, red, 8, hipMemcpyDeviceToHost, s));
    RT_HIP(hipStreamSynchronize(s));
    if (big) return fail(RTREX_EINVAL, "a vertex id does not fit in 64 bits");
    if (ne == 0) return fail(RTREX_EINVAL, "no edge in the file");
    uint64_t *ea, *eb, *ids, *tmp, *uq, *uord, *keys;
    uint32_t* uf;
    RT_TRY(sc.get(&ea, ne));
    RT_TRY(sc.get(&eb, ne));
    RT_TRY(sc.get(&ids, 2 * ne));
    RT_TRY(sc.get(&tmp, 2 * ne));
    k_take_snap<<<grid(nl), T, 0, s>>>(A, Bv, f, ord, nl, ea, eb, ids, ne);
    RT_TRY(radix_sort_u64(ids, tmp, 2 * ne, 64, s));
    RT_TRY(sc.get(&uf, 2 * ne));
    RT_TRY(sc.get(&uord, 2 * ne + 1));
    k_uniq_flags<<<grid(2 * ne), T, 0, s>>>(ids, 2 * ne, uf);
    RT_TRY(exclusive_scan_u32(uf, uord, 2 * ne, s));
    uint64_t nu = 0;
    RT_HIP(hipMemcpyAsync(&nu, uord + 2 * ne, 8, hipMemcpyDeviceToHost, s));
    RT_HIP(hipStreamSynchronize(s));
    if (nu >= UINT32_MAX) return fail(RTREX_EINVAL, "more than 2^32-1 distinct vertex ids");
    RT_TRY(sc.get(&uq, nu));
    k_uniq_take<<<grid(2 * ne), T, 0, s>>>(ids, 2 * ne, uf, uord, uq);
    int B = 1;
    while ((1ull << B) < nu) ++B;
    RT_TRY(sc.get(&keys, ne));
    k_pair_keys<<<grid(ne), T, 0, s>>>(ea, eb, ne, uq, nu, B, keys);
    RT_TRY(radix_sort_u64(keys, tmp, ne, 2 * B, s));
    k_uniq_flags<<<grid(ne), T, 0, s>>>(keys, ne, uf);
    RT_TRY(exclusive_scan_u32(uf, uord, ne, s));
    uint64_t m = 0;
    RT_HIP(hipMemcpyAsync(&m, uord + ne, 8, hipMemcpyDeviceToHost, s));
    RT_HIP(hipStreamSynchronize(s));
    if (2 * m >= UINT32_MAX) return fail(RTREX_EINVAL, "2m must be below 2^32-1");
    RT_TRY(sc.get(&P->u, m));
    RT_TRY(sc.get(&P->v, m));
    k_pair_take<<<grid(ne), T, 0, s>>>(keys, ne, uf, uord, B, P->u, P->v);
    RT_HIP(hipGetLastError());
    P->n = (int64_t)nu;
    P->m = (int64_t)m;
    return RTREX_OK;
}

rtrex_status write_escape_device(rtrex_graph h, const Parsed& P, const char* out, Scratch& sc) {
    hipStream_t s = h->stream;
    FILE* f = std::fopen(out, "wb");
    if (!f) return fail(RTREX_EIO, std::string("cannot write ") + out);
    std::fprintf(f, "%" PRId64 " %" PRId64 "\n", P.n, P.m);
    rtrex_status st = RTREX_OK;
    if (P.m) {
        uint32_t* len;
        uint64_t* off;
        uint8_t* buf;
        st = sc.get(&len, (uint64_t)P.m);
        if (st == RTREX_OK) st = sc.get(&off, (uint64_t)P.m + 1);
        uint64_t bytes = 0;
        if (st == RTREX_OK) {
            k_fmt_len<<<grid(P.m), T, 0, s>>>(P.u, P.v, (uint64_t)P.m, len);
            st = exclusive_scan_u32(len, off, (uint64_t)P.m, s);
        }
        if (st == RTREX_OK && hipMemcpyAsync(&bytes, off + P.m, 8, hipMemcpyDeviceToHost, s) ==
                                  hipSuccess &&
            hipStreamSynchronize(s) == hipSuccess) {
            st = sc.get(&buf, bytes);
            void* pin = st == RTREX_OK ? pin_alloc(bytes) : nullptr;
            if (st == RTREX_OK && !pin) st = fail(RTREX_ENOMEM, "page-locked staging failed");
            if (st == RTREX_OK) {
                k_fmt<<<grid(P.m), T, 0, s>>>(P.u, P.v, (uint64_t)P.m, off, buf);
                if (hipMemcpyAsync(pin, buf, bytes, hipMemcpyDeviceToHost, s) != hipSuccess ||
                    hipStreamSynchronize(s) != hipSuccess)
                    st = fail(RTREX_EDEVICE, "formatted edges could not be copied back");
                else if (std::fwrite(pin, 1, bytes, f) != bytes)
                    st = fail(RTREX_EIO, std::string("short write to ") + out);
                pin_free(pin);
            }
        } else if (st == RTREX_OK) {
            st = fail(RTREX_EDEVICE, "line lengths could not be read back");
        }
    }
    if (std::fclose(f) != 0 && st == RTREX_OK) st = fail(RTREX_EIO, "close failed");
    return st;
}

}  // namespace

rtrex_status load_text(rtrex_graph h, const char* path, int format) {
    hipStream_t s = h->stream;
    Scratch sc;
    Parsed P;
    RT_HIP(hipEventRecord(h->ev[0], s));
    RT_TRY(parse_device(h, path, format, sc, &P));
    RT_TRY(build_from_device_edges(h, P.u, P.v, P.n, P.m));
    RT_HIP(hipEventRecord(h->ev[1], s));
    RT_HIP(hipEventSynchronize(h->ev[1]));
    float ms = 0;
    RT_HIP(hipEventElapsedTime(&ms, h->ev[0], h->ev[1]));
    h->tm.build_ms = ms;
    return RTREX_OK;
}

rtrex_status convert_text(rtrex_graph h, const char* in, const char* out, int64_t* n,
                          int64_t* m) {
    Scratch sc;
    Parsed P;
    RT_TRY(parse_device(h, in, RTREX_FORMAT_SNAP, sc, &P));
    if (n) *n = P.n;
    if (m) *m = P.m;
    if (out) RT_TRY(write_escape_device(h, P, out, sc));
    return RTREX_OK;
}

}  // namespace rtrex
