// CPU oracle for the RTRExtractor decomposition path.
//
// TEST INFRASTRUCTURE ONLY (see rtrex_oracle.h): a project-written C++
// restatement of the reference's sequential algorithm, used as the checker by
// tests/, smoke() and bench.py's cpu_baseline leg.  The product never calls it.
//
// Where the reference's results depend on container semantics, the same
// libstdc++ containers are used so that the behaviour is identical:
//   - the deletion worklist is LIFO (std::stack in Decomposition.h:147-283);
//   - cluster vertex sets are std::unordered_set<int64_t> filled in the same
//     insertion sequence (ClusterStructures.h:27-33), so their iteration order,
//     which decides the extraction push order (Decomposition.h:897-914), is
//     the reference's;
//   - two-hop candidates are visited in ascending id (std::set,
//     Decomposition.h:615,815) and ordered with std::sort under the same
//     strict "weight greater" comparison (ClusterStructures.h:44-47,
//     Decomposition.h:829).
// Arithmetic: every triangle weight is 1.0/double(int64 product of the three
// ORIGINAL degrees) and every edge weight 1.0/double(int64 product of two),
// accumulated by plain left folds in the reference's loop order.

#include "rtrex_oracle.h"

#include <algorithm>
#include <chrono>
#include <cstring>
#include <queue>
#include <set>
#include <unordered_set>
#include <utility>
#include <vector>

namespace {

using i64 = int64_t;

struct Graph {
    i64 n;
    const i64* off;
    const i64* nb;
    i64 deg(i64 v) const { return off[v + 1] - off[v]; }

    // Slot of b inside a's sorted list, or -1 (Graph.cpp:130-149).
    i64 slot(i64 a, i64 b) const {
        if (a >= n) return -1;
        i64 lo = off[a], hi = off[a + 1] - 1;
        while (lo <= hi) {
            i64 mid = (lo + hi) / 2;
            if (nb[mid] == b) return mid;
            if (nb[mid] > b) hi = mid - 1;
            else lo = mid + 1;
        }
        return -1;
    }
};

// partner[i] = slot of the reverse edge (Graph.cpp:320-340).
std::vector<i64> partner_slots(const Graph& g) {
    std::vector<i64> p(g.off[g.n]);
    for (i64 u = 0; u < g.n; ++u)
        for (i64 i = g.off[u]; i < g.off[u + 1]; ++i) p[i] = g.slot(g.nb[i], u);
    return p;
}

inline double tri_weight(i64 da, i64 db, i64 dc) {
    return 1 / double(da * db * dc);  // Decomposition.h:97,213,760,785
}
inline double edge_weight(i64 da, i64 db) {
    return 1 / double(da * db);  // Decomposition.h:239,253,395
}

struct State {
    const Graph& g;
    std::vector<i64> part;
    double* pe;  // perEdge, nnz
    double* pv;  // perVertex, n
    i64 total;
    double total_wgt;
    std::vector<char> st;  // 'Y' alive, 'N' deleted, 'S' on the worklist
    double eps;
    explicit State(const Graph& gg) : g(gg) {}
};

// commonNbr (Decomposition.h:42-117): triangles found once each, in
// lexicographic (u < v < w) order; the weight goes to both copies of all three
// edges and to all three vertices.
void count_supports(State& s) {
    const Graph& g = s.g;
    const i64 nnz = g.off[g.n];
    std::fill(s.pe, s.pe + nnz, 0.0);
    std::fill(s.pv, s.pv + g.n, 0.0);
    s.total = 0;
    s.total_wgt = 0;
    for (i64 u = 0; u < g.n; ++u) {
        const i64 du = g.deg(u);
        for (i64 i = g.off[u]; i < g.off[u + 1]; ++i) {
            const i64 v = g.nb[i];
            if (u > v) continue;
            const i64 dv = g.deg(v);
            const i64 lo = du < dv ? u : v;
            const i64 hi = du < dv ? v : u;
            for (i64 j = g.off[lo]; j < g.off[lo + 1]; ++j) {
                const i64 w = g.nb[j];
                if (w <= u || w <= v) continue;
                const i64 loc = g.slot(hi, w);
                if (loc < 0) continue;
                const double t = tri_weight(du, dv, g.deg(w));
                s.total++;
                s.total_wgt += t;
                s.pe[i] += t;
                s.pe[s.part[i]] += t;
                s.pe[j] += t;
                s.pe[s.part[j]] += t;
                s.pe[loc] += t;
                s.pe[s.part[loc]] += t;
                s.pv[u] += t;
                s.pv[v] += t;
                s.pv[w] += t;
            }
        }
    }
}

using Work = std::vector<std::pair<i64, i64>>;  // LIFO: push_back / back

inline void mark(State& s, i64 slot, char c) {
    s.st[slot] = c;
    s.st[s.part[slot]] = c;
}

// deleteAndClean (Decomposition.h:147-283).
void delete_and_clean(State& s, Work& todo) {
    const Graph& g = s.g;
    while (!todo.empty()) {
        const i64 u = todo.back().first, v = todo.back().second;
        todo.pop_back();
        const i64 luv = g.slot(u, v);
        if (luv < 0) continue;  // the reference only prints an error here
        const i64 puv = s.part[luv];
        const i64 du = g.deg(u), dv = g.deg(v);
        const i64 lo = du < dv ? u : v;
        const i64 hi = du < dv ? v : u;
        const i64 dlo = du < dv ? du : dv;
        const i64 dhi = du < dv ? dv : du;
        for (i64 j = g.off[lo]; j < g.off[lo + 1]; ++j) {
            if (s.st[j] == 'N') continue;
            const i64 w = g.nb[j];
            if (w == u || w == v) continue;
            const i64 lhw = g.slot(hi, w);
            if (lhw < 0 || s.st[lhw] == 'N') continue;
            const i64 pj = s.part[j], phw = s.part[lhw];
            const i64 dw = g.deg(w);
            const double t = tri_weight(du, dv, dw);
            s.pe[luv] -= t;
            s.pe[j] -= t;
            s.pe[lhw] -= t;
            s.pe[puv] -= t;
            s.pe[pj] -= t;
            s.pe[phw] -= t;
            s.pv[u] -= t;
            s.pv[v] -= t;
            s.pv[w] -= t;
            s.total--;
            s.total_wgt -= t;
            if (s.st[j] != 'S' && s.pe[j] < s.eps * edge_weight(dlo, dw)) {
                todo.emplace_back(lo, w);
                s.st[j] = 'S';
                s.st[pj] = 'S';
            }
            if (s.st[lhw] != 'S' && s.pe[lhw] < s.eps * edge_weight(dhi, dw)) {
                todo.emplace_back(hi, w);
                s.st[lhw] = 'S';
                s.st[phw] = 'S';
            }
        }
        s.st[luv] = 'N';
        s.st[puv] = 'N';
    }
}

// initialClean, sequential seeding branch (Decomposition.h:302-411).
void initial_clean(State& s) {
    const Graph& g = s.g;
    s.st.assign(g.off[g.n] + 1, 'Y');
    Work todo;
    for (i64 u = 0; u < g.n; ++u) {
        const i64 du = g.deg(u);
        for (i64 i = g.off[u]; i < g.off[u + 1]; ++i) {
            const i64 v = g.nb[i];
            if (u > v) continue;
            if (s.st[i] != 'S' && s.pe[i] < s.eps * edge_weight(du, g.deg(v))) {
                todo.emplace_back(u, v);
                mark(s, i, 'S');
            }
        }
    }
    delete_and_clean(s, todo);
}

struct Cluster {
    std::unordered_set<i64> vertices;  // ClusterStructures.h:27-33
    i64 nv = 0, ne = 0, cut = 0;
};

// populateStats over the original graph (Decomposition.h:473-491).
void cluster_stats(const Graph& g, Cluster& c) {
    c.nv = (i64)c.vertices.size();
    c.ne = 0;
    c.cut = 0;
    for (i64 z : c.vertices)
        for (i64 i = g.off[z]; i < g.off[z + 1]; ++i) {
            if (c.vertices.count(g.nb[i])) c.ne++;
            else c.cut++;
        }
    c.ne /= 2;
}

struct Cand {
    i64 vertex;
    double wgt;
};

// The deletion pairs of a cluster (Decomposition.h:897-914): members in the
// given order (the reference iterates its unordered_set), each one's 'Y'
// slots in order, both slots of a listed edge marked 'S'.
void cluster_pairs(State& s, const std::vector<i64>& members, Work& todo) {
    const Graph& g = s.g;
    for (i64 z : members)
        for (i64 j = g.off[z]; j < g.off[z + 1]; ++j) {
            if (s.st[j] != 'Y') continue;
            todo.emplace_back(z, g.nb[j]);
            mark(s, j, 'S');
        }
}

// The inner loop of one seed u (Decomposition.h:728-827, sequential branch):
// every pair of alive neighbours v, w (slot k >= j, both of degree at most
// d_u / eps) adds the triangle (u, v, w) to internal and enumerates the
// triangles (v, w, x) over the lower-degree side's list; x a neighbour of u
// (nbd) adds to internal when it is the largest of v, w, x, any other x is a
// two-hop candidate collecting the weight in triwgt.  cand receives the
// candidates in ascending id (std::set iteration) with their weights, and
// potential their sum in that order; triwgt is left zero.
void inner_loop(const State& s, i64 u, const std::vector<char>& nbd, std::vector<double>& triwgt,
                double& internal, std::vector<Cand>& cand, double& potential) {
    const Graph& g = s.g;
    const i64 du = g.deg(u);
    std::set<i64> two_hop;
    const double lim = (double)du / s.eps;
    for (i64 j = g.off[u]; j < g.off[u + 1]; ++j) {  // :728-805
        if (s.st[j] == 'N') continue;
        const i64 v = g.nb[j];
        const i64 dv = g.deg(v);
        if ((double)dv > lim) continue;
        for (i64 k = j; k < g.off[u + 1]; ++k) {
            if (s.st[k] == 'N') continue;
            const i64 w = g.nb[k];
            const i64 dw = g.deg(w);
            if ((double)dw > lim) continue;
            // The reference re-finds w in u's list (:754-756); with the
            // sorted simple CSR that is slot k itself, already alive.
            const i64 loc = g.slot(u, w);
            if (loc < 0 || s.st[loc] == 'N') continue;
            internal += tri_weight(du, dv, dw);
            const i64 lo = dv < dw ? v : w;
            const i64 hi = dv < dw ? w : v;
            for (i64 l = g.off[lo]; l < g.off[lo + 1]; ++l) {
                if (s.st[l] == 'N') continue;
                const i64 x = g.nb[l];
                if (x == lo || x == hi || x == u) continue;
                const i64 lhx = g.slot(hi, x);
                if (lhx < 0 || s.st[lhx] == 'N') continue;
                const double t = tri_weight(g.deg(x), dv, dw);
                if (nbd[x]) {
                    if (x > v && x > w) internal += t;
                    continue;
                }
                two_hop.insert(x);
                triwgt[x] += t;
            }
        }
    }
    cand.clear();  // :811-827
    cand.reserve(two_hop.size());
    potential = 0;
    for (i64 x : two_hop) {
        cand.push_back({x, triwgt[x]});
        potential += triwgt[x];
        triwgt[x] = 0;
    }
}

// disjointExtract, sequential inner-loop branch (Decomposition.h:538-940),
// entered after count and initial clean.
std::vector<Cluster> extract(State& s, i64* cluster_map, i64* member_rank) {
    const Graph& g = s.g;
    const i64 n = g.n;
    std::vector<std::pair<i64, i64>> order(n);  // (vertex, degree)
    for (i64 i = 0; i < n; ++i) order[i] = {i, g.deg(i)};
    // pairCompareSecond (Graph.cpp:40-45): degree, then id.  A total order.
    std::sort(order.begin(), order.end(),
              [](const std::pair<i64, i64>& a, const std::pair<i64, i64>& b) {
                  if (a.second != b.second) return a.second < b.second;
                  return a.first < b.first;
              });
    std::vector<double> triwgt(n, 0.0);
    std::vector<char> nbd(n, 0), clustered(n, 0);
    std::fill(cluster_map, cluster_map + n, -1);
    std::vector<Cluster> out;
    i64 nclus = 0;

    for (i64 r = 0; r < n; ++r) {
        const i64 u = order[r].first;
        if (clustered[u]) continue;
        Cluster c;
        c.vertices.insert(u);
        cluster_map[u] = nclus;
        clustered[u] = 1;
        i64 joined = 0;  // order in which vertices join the cluster
        if (member_rank) member_rank[u] = joined;
        double internal = 0;

        for (i64 j = g.off[u]; j < g.off[u + 1]; ++j) {  // :620-631
            if (s.st[j] == 'N') continue;
            const i64 v = g.nb[j];
            c.vertices.insert(v);
            if (member_rank) member_rank[v] = ++joined;
            nbd[v] = 1;
            cluster_map[v] = nclus;
            clustered[v] = 1;
        }
        std::vector<Cand> cand;
        double potential = 0;
        inner_loop(s, u, nbd, triwgt, internal, cand, potential);
        std::sort(cand.begin(), cand.end(),
                  [](Cand a, Cand b) { return a.wgt > b.wgt; });

        double ratio = internal / (internal + potential);  // :847-864
        int maxind = -1;
        double num = 0, den = 0;
        const i64 size = (i64)cand.size();
        for (i64 idx = 0; idx < size; ++idx) {
            num += cand[idx].wgt;
            den += s.pv[cand[idx].vertex];
            const double nr = (internal + num) / (internal + potential + den);
            if (nr > ratio) {
                ratio = nr;
                maxind = (int)idx;
            }
        }
        for (i64 idx = 0; idx <= maxind; ++idx) {  // :870-876
            const i64 x = cand[idx].vertex;
            c.vertices.insert(x);
            if (member_rank) member_rank[x] = ++joined;
            cluster_map[x] = nclus;
            clustered[x] = 1;
        }
        for (i64 j = g.off[u]; j < g.off[u + 1]; ++j) nbd[g.nb[j]] = 0;

        cluster_stats(g, c);
        out.push_back(c);
        nclus++;

        Work todo;  // :897-918
        cluster_pairs(s, std::vector<i64>(c.vertices.begin(), c.vertices.end()), todo);
        delete_and_clean(s, todo);
    }
    return out;
}

bool valid(i64 n, const i64* off, const i64* nb) {
    if (n < 0 || !off || (off[n] > 0 && !nb)) return false;
    if (off[0] != 0) return false;
    for (i64 v = 0; v < n; ++v)
        if (off[v + 1] < off[v]) return false;
    for (i64 i = 0; i < off[n]; ++i)
        if (nb[i] < 0 || nb[i] >= n) return false;
    return true;
}

double since(std::chrono::steady_clock::time_point a) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - a)
        .count();
}

}  // namespace

extern "C" {

int oracle_make_csr(int64_t n, int64_t nnz, const int64_t* src,
                    const int64_t* dst, int64_t* offsets, int64_t* nbors) {
    std::vector<std::pair<i64, i64>> e(nnz);
    for (i64 i = 0; i < nnz; ++i) {
        if (src[i] < 0 || src[i] >= n || dst[i] < 0 || dst[i] >= n) return 1;
        e[i] = {src[i], dst[i]};
    }
    std::sort(e.begin(), e.end());  // Graph.cpp:352-362, then :381
    i64 cv = 0;
    for (i64 i = 0; i < nnz; ++i)
        while (cv <= e[i].first) offsets[cv++] = i;  // Graph.cpp:369-377
    while (cv <= n) offsets[cv++] = nnz;
    for (i64 i = 0; i < nnz; ++i) nbors[i] = e[i].second;
    return 0;
}

int oracle_count(int64_t n, const int64_t* offsets, const int64_t* nbors,
                 int64_t* total, double* perEdge, double* perVertex) {
    if (!valid(n, offsets, nbors)) return 1;
    Graph g{n, offsets, nbors};
    State s(g);
    s.part = partner_slots(g);
    for (i64 p : s.part)
        if (p < 0) return 2;  // not undirected (Graph.cpp:332-336)
    s.pe = perEdge;
    s.pv = perVertex;
    count_supports(s);
    *total = s.total;
    return 0;
}

int oracle_initial_clean(int64_t n, const int64_t* offsets,
                         const int64_t* nbors, double eps, double* perEdge,
                         double* perVertex, int64_t* total, uint8_t* status) {
    if (!valid(n, offsets, nbors)) return 1;
    Graph g{n, offsets, nbors};
    State s(g);
    s.part = partner_slots(g);
    s.pe = perEdge;
    s.pv = perVertex;
    s.total = *total;
    s.eps = eps;
    initial_clean(s);
    *total = s.total;
    const i64 nnz = offsets[n];
    for (i64 i = 0; i < nnz; ++i) status[i] = (uint8_t)s.st[i];
    return 0;
}

int oracle_delete_and_clean(int64_t n, const int64_t* offsets,
                            const int64_t* nbors, double eps,
                            const int64_t* pairs, int64_t k, double* perEdge,
                            double* perVertex, int64_t* total,
                            uint8_t* status) {
    if (!valid(n, offsets, nbors)) return 1;
    Graph g{n, offsets, nbors};
    State s(g);
    s.part = partner_slots(g);
    s.pe = perEdge;
    s.pv = perVertex;
    s.total = *total;
    s.eps = eps;
    const i64 nnz = offsets[n];
    s.st.assign(status, status + nnz);
    Work todo;
    for (i64 q = 0; q < k; ++q) {
        const i64 a = pairs[2 * q], b = pairs[2 * q + 1];
        const i64 sl = g.slot(a, b);
        if (sl < 0 || s.st[sl] != 'Y') return 3;
        todo.emplace_back(a, b);
        mark(s, sl, 'S');
    }
    delete_and_clean(s, todo);
    *total = s.total;
    for (i64 i = 0; i < nnz; ++i) status[i] = (uint8_t)s.st[i];
    return 0;
}

// disjointExtract's seed loop (Decomposition.h:538-940) entered from a given
// state (supports and statuses after initialClean and any deleteAndClean
// calls), as disjointExtract continues after its own initialClean (:548-556).
int oracle_extract_state(int64_t n, const int64_t* offsets, const int64_t* nbors,
                         double eps, double* perEdge, double* perVertex,
                         int64_t* total, const uint8_t* status, int64_t* cluster_map,
                         int64_t* n_clusters) {
    if (!valid(n, offsets, nbors)) return 1;
    Graph g{n, offsets, nbors};
    State s(g);
    s.part = partner_slots(g);
    s.pe = perEdge;
    s.pv = perVertex;
    s.total = *total;
    s.eps = eps;
    const i64 nnz = offsets[n];
    s.st.assign(status, status + nnz);
    s.st.push_back('Y');
    std::vector<Cluster> dec = extract(s, cluster_map, nullptr);
    *n_clusters = (i64)dec.size();
    *total = s.total;
    return 0;
}

int oracle_extract_inner(int64_t n, const int64_t* offsets, const int64_t* nbors,
                         const uint8_t* status, int64_t u, double eps, int64_t* cand,
                         double* wgt, int64_t cap, int64_t* ncand, double* internal,
                         double* potential) {
    if (!valid(n, offsets, nbors) || u < 0 || u >= n || cap < 0) return 1;
    Graph g{n, offsets, nbors};
    State s(g);
    s.eps = eps;
    s.st.assign(status, status + offsets[n]);
    std::vector<char> nbd(n, 0);
    for (i64 j = g.off[u]; j < g.off[u + 1]; ++j)  // :620-631
        if (s.st[j] != 'N') nbd[g.nb[j]] = 1;
    std::vector<double> triwgt(n, 0.0);
    std::vector<Cand> c;
    double in = 0, pot = 0;
    inner_loop(s, u, nbd, triwgt, in, c, pot);
    for (i64 i = 0; i < (i64)c.size() && i < cap; ++i) {
        cand[i] = c[i].vertex;
        wgt[i] = c[i].wgt;
    }
    *ncand = (i64)c.size();
    *internal = in;
    *potential = pot;
    return 0;
}

int oracle_cluster_edges(int64_t n, const int64_t* offsets, const int64_t* nbors,
                         const uint8_t* status, const int64_t* verts, int64_t nverts,
                         int64_t* pairs, int64_t cap, int64_t* npairs) {
    if (!valid(n, offsets, nbors) || nverts < 0 || cap < 0) return 1;
    Graph g{n, offsets, nbors};
    State s(g);
    s.part = partner_slots(g);
    s.st.assign(status, status + offsets[n]);
    std::vector<i64> members(verts, verts + nverts);
    for (i64 z : members)
        if (z < 0 || z >= n) return 1;
    Work todo;
    cluster_pairs(s, members, todo);
    for (i64 i = 0; i < (i64)todo.size() && i < cap; ++i) {
        pairs[2 * i] = todo[i].first;
        pairs[2 * i + 1] = todo[i].second;
    }
    *npairs = (i64)todo.size();
    return 0;
}

int oracle_decompose(int64_t n, const int64_t* offsets, const int64_t* nbors,
                     double eps, int64_t* cluster_map, int64_t* n_clusters,
                     uint8_t* status_after_clean, int64_t* flat,
                     int64_t flat_cap, int64_t* flat_len, double* t,
                     int64_t* member_rank) {
    if (!valid(n, offsets, nbors)) return 1;
    Graph g{n, offsets, nbors};
    State s(g);
    const i64 nnz = offsets[n];
    std::vector<double> pe(nnz + 1), pv(n + 1);
    s.pe = pe.data();
    s.pv = pv.data();
    s.eps = eps;
    auto t0 = std::chrono::steady_clock::now();
    s.part = partner_slots(g);
    for (i64 p : s.part)
        if (p < 0) return 2;
    count_supports(s);
    double tc = since(t0);
    auto t1 = std::chrono::steady_clock::now();
    initial_clean(s);
    double tcl = since(t1);
    if (status_after_clean)
        for (i64 i = 0; i < nnz; ++i) status_after_clean[i] = (uint8_t)s.st[i];
    auto t2 = std::chrono::steady_clock::now();
    std::vector<Cluster> dec = extract(s, cluster_map, member_rank);
    double te = since(t2);
    *n_clusters = (i64)dec.size();
    if (t) {
        t[0] = tc;
        t[1] = tcl;
        t[2] = te;
    }
    if (flat) {
        i64 p = 0;
        for (const Cluster& c : dec) {
            if (p + 3 + c.nv > flat_cap) return 4;
            flat[p++] = c.nv;
            flat[p++] = c.ne;
            flat[p++] = c.cut;
            for (i64 v : c.vertices) flat[p++] = v;
        }
        if (flat_len) *flat_len = p;
    }
    return 0;
}

int oracle_connected_components(int64_t n, const int64_t* offsets,
                                const int64_t* nbors, int64_t* comp,
                                int64_t* n_comp) {
    if (!valid(n, offsets, nbors)) return 1;
    std::fill(comp, comp + n, -1);
    i64 c = 0;
    std::queue<i64> q;
    for (i64 u = 0; u < n; ++u) {  // Decomposition.h:1033-1064
        if (comp[u] >= 0) continue;
        comp[u] = c;
        q.push(u);
        while (!q.empty()) {
            i64 x = q.front();
            q.pop();
            for (i64 j = offsets[x]; j < offsets[x + 1]; ++j)
                if (comp[nbors[j]] < 0) {
                    comp[nbors[j]] = c;
                    q.push(nbors[j]);
                }
        }
        c++;
    }
    *n_comp = c;
    return 0;
}

}  // extern "C"
