// Output files of the reference's driver, byte for byte (host code only).
//
// Reference: printDecomposition (RTRex_Sequential/clustering/RTRex.cpp:19-118),
// expandClusters (Decomposition.h:943-1008) and printScores
// (RTRex.cpp:120-135), populateStats (Escape/Decomposition.h:473-491), connectedComp
// (Decomposition.h:1023-1068) and the survivor dump of initialClean
// (Decomposition.h:419-452).
//
// The byte order of a decomposition file follows two library behaviours of
// the reference: the vertices of a cluster are printed in the iteration
// order of a std::unordered_set<int64_t>, which depends only on the sequence
// of insertions, and the clusters are ordered by std::sort (not stable) on
// size.  Both are reproduced by doing the same operations: each cluster's set
// is filled in the order its vertices joined it (the member order recorded by
// the extraction, or BFS discovery order for components), and an index array
// is sorted with the same comparison -- the sort takes the same decisions,
// so it yields the same permutation.  Copies of a set keep its order.
#include <algorithm>
#include <cinttypes>
#include <cstdio>
#include <map>
#include <queue>
#include <string>
#include <thread>
#include <unordered_set>
#include <vector>

#include "internal.hpp"

using namespace rtrex;

namespace {

struct Clus {
    std::unordered_set<int64_t> vertices;
    int64_t nVertices = 0, nEdges = 0, cut = 0;
};

// populateStats over the original graph.
void stats(const int64_t* off, const int64_t* nb, Clus& c) {
    c.nVertices = (int64_t)c.vertices.size();
    c.nEdges = c.cut = 0;
    for (int64_t u : c.vertices)
        for (int64_t i = off[u]; i < off[u + 1]; ++i) {
            if (c.vertices.count(nb[i])) c.nEdges++;
            else c.cut++;
        }
    c.nEdges /= 2;
}

struct File {
    FILE* f = nullptr;
    explicit File(const std::string& p) : f(std::fopen(p.c_str(), "w")) {}
    ~File() {
        if (f) std::fclose(f);
    }
};

rtrex_status write_all(const std::string& name, int64_t n, const int64_t* off, const int64_t* nb,
                       std::vector<Clus>& D) {
    File f(name + "-decomposition.txt"), fs(name + "-stats.txt"), fcut(name + "-cut.txt"),
        ffreq(name + "-freq.txt");
    if (!f.f || !fs.f || !fcut.f || !ffreq.f)
        return fail(RTREX_EIO, "cannot write the output files of " + name);
    std::vector<int64_t> idx(D.size());
    for (size_t i = 0; i < D.size(); ++i) idx[i] = (int64_t)i;
    // clusterCompareDecreasing (ClusterStructures.h:51-54)
    std::sort(idx.begin(), idx.end(),
              [&](int64_t a, int64_t b) { return D[a].nVertices > D[b].nVertices; });
    std::vector<int64_t> freq((size_t)n + 1, 0);
    int64_t maxsize = 0;
    for (int64_t ci : idx) {
        const Clus& c = D[ci];
        for (int64_t v : c.vertices) std::fprintf(f.f, "%" PRId64 " ", v);
        std::fprintf(f.f, "\n");
        maxsize = std::max(maxsize, c.nVertices);
        freq[c.nVertices]++;
        if (c.nVertices != 1) {
            const double density =
                2 * (c.nEdges) / ((double)(c.nVertices) * (double)(c.nVertices - 1));
            std::fprintf(fs.f, "%" PRId64 " %.2f\n", c.nVertices, density);
        }
        for (int64_t vert : c.vertices)
            for (int64_t j = off[vert]; j < off[vert + 1]; ++j) {
                const int64_t nbr = nb[j];
                if (nbr < vert) continue;
                if (!c.vertices.count(nbr))
                    std::fprintf(fcut.f, "%" PRId64 " %" PRId64 "\n", vert, nbr);
            }
    }
    for (int64_t i = 0; i < maxsize + 1; ++i)
        std::fprintf(ffreq.f, "Size = %" PRId64 ", freq = %" PRId64 ", total vertices = %" PRId64 "\n",
                     i, freq[i], i * freq[i]);
    return RTREX_OK;
}

rtrex_status check_csr(int64_t n, const int64_t* off, const int64_t* nb) {
    if (n < 0 || !off || (off[n] && !nb)) return fail(RTREX_EINVAL, "bad CSR");
    return RTREX_OK;
}

// Members of every cluster, each set filled in the order its vertices
// joined the cluster (member_rank), stats over the original graph.
rtrex_status build_clusters(int64_t n, const int64_t* offsets, const int64_t* nbors,
                            const int64_t* cluster_of, const int64_t* member_rank,
                            int64_t n_clusters, std::vector<Clus>& D) {
    if (n && (!cluster_of || !member_rank)) return fail(RTREX_EINVAL, "null cluster arrays");
    if (n_clusters < 0 || n_clusters > n) return fail(RTREX_EINVAL, "bad cluster count");
    std::vector<int64_t> cnt((size_t)n_clusters + 1, 0);
    for (int64_t v = 0; v < n; ++v) {
        const int64_t c = cluster_of[v];
        if (c < 0 || c >= n_clusters) return fail(RTREX_EINVAL, "cluster id out of range");
        cnt[c + 1]++;
    }
    for (int64_t c = 0; c < n_clusters; ++c) cnt[c + 1] += cnt[c];
    std::vector<int64_t> mem((size_t)n, -1);
    for (int64_t v = 0; v < n; ++v) {
        const int64_t c = cluster_of[v], r = member_rank[v];
        if (r < 0 || r >= cnt[c + 1] - cnt[c] || mem[cnt[c] + r] != -1)
            return fail(RTREX_EINVAL, "member ranks are not a permutation within each cluster");
        mem[cnt[c] + r] = v;
    }
    D.assign((size_t)n_clusters, Clus());
    for (int64_t c = 0; c < n_clusters; ++c) {
        for (int64_t i = cnt[c]; i < cnt[c + 1]; ++i) D[c].vertices.insert(mem[i]);
        stats(offsets, nbors, D[c]);
    }
    return RTREX_OK;
}

// The cluster a singleton v joins (Decomposition.h:970-990): neighbours'
// clusters counted in slot order; a cluster becomes the best when its count
// exceeds the best count so far, which starts at the threshold.  cnt is a
// zeroed per-thread counter array, left zeroed.
void best_cluster(int64_t v, const int64_t* off, const int64_t* nb, const int64_t* cluster_of,
                  int32_t threshold, std::vector<int64_t>& cnt, int64_t* best, int64_t* top) {
    int64_t b = -1, mx = threshold;
    for (int64_t j = off[v]; j < off[v + 1]; ++j) {
        const int64_t c = cluster_of[nb[j]];
        if (c < 0) continue;
        if (++cnt[c] > mx) {
            b = c;
            mx = cnt[c];
        }
    }
    for (int64_t j = off[v]; j < off[v + 1]; ++j) {
        const int64_t c = cluster_of[nb[j]];
        if (c >= 0) cnt[c] = 0;
    }
    *best = b;
    *top = mx;
}

}  // namespace

extern "C" {

rtrex_status rtrex_write_decomposition(const char* name, int64_t n, const int64_t* offsets,
                                       const int64_t* nbors, const int64_t* cluster_of,
                                       const int64_t* member_rank, int64_t n_clusters) {
    if (!name) return fail(RTREX_EINVAL, "null name");
    RT_TRY(check_csr(n, offsets, nbors));
    std::vector<Clus> D;
    RT_TRY(build_clusters(n, offsets, nbors, cluster_of, member_rank, n_clusters, D));
    return write_all(name, n, offsets, nbors, D);
}

rtrex_status rtrex_write_expanded(const char* name, int64_t n, const int64_t* offsets,
                                  const int64_t* nbors, const int64_t* cluster_of,
                                  const int64_t* member_rank, int64_t n_clusters,
                                  int32_t threshold, int64_t* expanded_of, double* score) {
    if (!name) return fail(RTREX_EINVAL, "null name");
    if (threshold < 0) return fail(RTREX_EINVAL, "threshold must be >= 0");
    RT_TRY(check_csr(n, offsets, nbors));
    std::vector<Clus> D;
    RT_TRY(build_clusters(n, offsets, nbors, cluster_of, member_rank, n_clusters, D));
    // Every singleton's neighbouring cluster, from the extraction's map (which
    // the moves below never change): in parallel over vertices.
    std::vector<int64_t> best((size_t)n, -1), top((size_t)n, 0);
    std::vector<char> single((size_t)n, 0);
    for (const Clus& c : D)
        if (c.nVertices == 1) single[(size_t)*c.vertices.cbegin()] = 1;
    const unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; ++t)
        th.emplace_back([&, t] {
            std::vector<int64_t> cnt((size_t)n_clusters, 0);
            for (int64_t v = t; v < n; v += nt)
                if (single[v]) best_cluster(v, offsets, nbors, cluster_of, threshold, cnt,
                                            &best[v], &top[v]);
        });
    for (auto& x : th) x.join();
    // expandClusters' loop (Decomposition.h:958-1001): singleton clusters in
    // id order, each one's first vertex in its set's iteration order moves to
    // its best cluster; stats recomputed at the end.
    std::map<int64_t, double> scores;
    const int64_t size = (int64_t)D.size();
    for (int64_t i = 0; i < size; ++i) {
        if (D[i].nVertices > 1 || D[i].vertices.empty()) continue;
        const int64_t v = *D[i].vertices.cbegin();
        if (best[v] < 0) continue;
        D[best[v]].vertices.insert(v);
        D[i].vertices.erase(v);
        scores[v] = (double)top[v] / (double)(offsets[v + 1] - offsets[v]);
    }
    for (Clus& c : D) stats(offsets, nbors, c);
    if (expanded_of) {
        for (int64_t c = 0; c < size; ++c)
            for (int64_t v : D[c].vertices) expanded_of[v] = c;
    }
    if (score) {
        std::fill(score, score + n, -1.0);
        for (const auto& kv : scores) score[kv.first] = kv.second;
    }
    RT_TRY(write_all(name, n, offsets, nbors, D));
    // printScores (RTRex.cpp:120-135): moved vertices ascending
    File fsc(std::string(name) + "-scores.txt");
    if (!fsc.f) return fail(RTREX_EIO, std::string("cannot write ") + name + "-scores.txt");
    std::fprintf(fsc.f, "Vertex,Score\n");
    for (const auto& kv : scores) std::fprintf(fsc.f, "%" PRId64 ",%f\n", kv.first, kv.second);
    return RTREX_OK;
}

rtrex_status rtrex_write_components(const char* name, int64_t n, const int64_t* offsets,
                                    const int64_t* nbors) {
    if (!name) return fail(RTREX_EINVAL, "null name");
    RT_TRY(check_csr(n, offsets, nbors));
    // connectedComp: BFS from every unvisited vertex in id order; members in
    // discovery order.
    std::vector<char> seen((size_t)n, 0);
    std::vector<Clus> D;
    std::queue<int64_t> q;
    for (int64_t u = 0; u < n; ++u) {
        if (seen[u]) continue;
        Clus c;
        c.vertices.insert(u);
        seen[u] = 1;
        q.push(u);
        while (!q.empty()) {
            const int64_t x = q.front();
            q.pop();
            for (int64_t j = offsets[x]; j < offsets[x + 1]; ++j) {
                const int64_t y = nbors[j];
                if (y < 0 || y >= n) return fail(RTREX_EINVAL, "neighbour id out of range");
                if (!seen[y]) {
                    q.push(y);
                    seen[y] = 1;
                    c.vertices.insert(y);
                }
            }
        }
        stats(offsets, nbors, c);
        D.push_back(std::move(c));
    }
    return write_all(name, n, offsets, nbors, D);
}

rtrex_status rtrex_write_survivors(const char* path, int64_t n, const int64_t* offsets,
                                   const int64_t* nbors, const uint8_t* status) {
    if (!path) return fail(RTREX_EINVAL, "null path");
    RT_TRY(check_csr(n, offsets, nbors));
    File f(path);
    if (!f.f) return fail(RTREX_EIO, std::string("cannot write ") + path);
    // (u, v) with u < v in CSR order is already (u, v)-sorted.
    for (int64_t u = 0; u < n; ++u)
        for (int64_t j = offsets[u]; j < offsets[u + 1]; ++j)
            if (status[j] == 'Y' && u < nbors[j])
                std::fprintf(f.f, "%" PRId64 " %" PRId64 "\n", u, nbors[j]);
    return RTREX_OK;
}

}  // extern "C"
