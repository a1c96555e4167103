// ref_dump: a thin driver around the REFERENCE implementation, used only to
// produce golden vectors and the reference CPU timing.  It is test
// infrastructure: nothing in the product links or runs it.
//
// It is compiled (by oracle/Makefile) against the reference sources where
// they lie under /root/reference/RTRex_Sequential (Graph.cpp, GraphIO.cpp and
// the header-only Escape/Decomposition.h); no reference source is copied into
// this repository.  The driver itself is project code: it calls the
// reference's phase functions one after another and writes their raw outputs.
//
//   usage: ref_dump <escape-graph.txt> <eps> <outdir> [phases]
//     phases: "all" (default) | "noextract"
//
// Outputs (raw little-endian arrays, one file each, in <outdir>):
//   csr_offsets.i64 csr_nbors.i64           makeCSR           Graph.cpp:344-384
//   count_total.i64 count_perEdge.f64
//   count_perVertex.f64                     commonNbr         Decomposition.h:42-117
//   clean_status.u8 clean_perEdge.f64
//   clean_perVertex.f64 clean_total.i64     initialClean      Decomposition.h:302-460
//   clusters.i64   [k, then per cluster: nVertices nEdges cut v_0..v_{s-1}]
//                  vertices in the reference's unordered_set iteration order
//   cluster_map.i64                         disjointExtract   Decomposition.h:538-940
//   times.txt      wall seconds per phase (load+csr, count, clean, extract)
//
// The reference prints progress text and writes amazon_cleaned_edges.txt into
// the current directory (Decomposition.h:441); run this with cwd = a scratch
// directory.  Its stdout is sent to /dev/null so only our own messages remain
// on stderr.

#include "Escape/GraphIO.h"
#include "Escape/Graph.h"
#include "Escape/Decomposition.h"

#include <chrono>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

using namespace Escape;

static void put(const std::string& dir, const char* name, const void* p,
                size_t bytes)
{
    std::string path = dir + "/" + name;
    FILE* f = fopen(path.c_str(), "wb");
    if (!f) {
        fprintf(stderr, "ref_dump: cannot write %s\n", path.c_str());
        exit(2);
    }
    if (bytes && fwrite(p, 1, bytes, f) != bytes) {
        fprintf(stderr, "ref_dump: short write %s\n", path.c_str());
        exit(2);
    }
    fclose(f);
}

static double secs(std::chrono::steady_clock::time_point a,
                   std::chrono::steady_clock::time_point b)
{
    return std::chrono::duration<double>(b - a).count();
}

int main(int argc, char** argv)
{
    if (argc < 4) {
        fprintf(stderr, "usage: %s graph.txt eps outdir [all|noextract]\n",
                argv[0]);
        return 1;
    }
    const char* path = argv[1];
    double eps = atof(argv[2]);
    std::string out(argv[3]);
    bool do_extract = !(argc > 4 && strcmp(argv[4], "noextract") == 0);

    if (!freopen("/dev/null", "w", stdout)) {
        fprintf(stderr, "ref_dump: cannot silence stdout\n");
        return 2;
    }

    auto t0 = std::chrono::steady_clock::now();
    Graph g;
    if (loadGraph(path, g, 1, IOFormat::escape)) {
        fprintf(stderr, "ref_dump: loadGraph failed\n");
        return 3;
    }
    CGraph cg = makeCSR(g);
    auto t1 = std::chrono::steady_clock::now();
    const int64_t n = cg.nVertices, nnz = cg.nEdges;
    put(out, "csr_offsets.i64", cg.offsets, sizeof(int64_t) * (n + 1));
    put(out, "csr_nbors.i64", cg.nbors, sizeof(int64_t) * nnz);

    // Phase 1: weighted triangle supports.
    auto t2 = std::chrono::steady_clock::now();
    WeightedTriangleInfo tri = commonNbr(&cg);
    auto t3 = std::chrono::steady_clock::now();
    put(out, "count_total.i64", &tri.total, sizeof(int64_t));
    put(out, "count_perEdge.f64", tri.perEdge, sizeof(double) * nnz);
    put(out, "count_perVertex.f64", tri.perVertex, sizeof(double) * n);

    // Phase 2: initial clean to convergence (sequential seeding, mode 's').
    auto t4 = std::chrono::steady_clock::now();
    char* status = initialClean(&cg, &tri, eps, false);
    auto t5 = std::chrono::steady_clock::now();
    put(out, "clean_status.u8", status, nnz);
    put(out, "clean_perEdge.f64", tri.perEdge, sizeof(double) * nnz);
    put(out, "clean_perVertex.f64", tri.perVertex, sizeof(double) * n);
    put(out, "clean_total.i64", &tri.total, sizeof(int64_t));

    double t_extract = 0.0;
    if (do_extract) {
        // Phase 3: the full driver call (it recomputes count and clean
        // internally, exactly as RTRex.cpp:193 does).
        std::map<VertexIdx, VertexIdx> cmap;
        auto t6 = std::chrono::steady_clock::now();
        std::vector<Cluster> dec = disjointExtract(&cg, eps, cmap, false);
        auto t7 = std::chrono::steady_clock::now();
        t_extract = secs(t6, t7);
        std::vector<int64_t> flat;
        flat.push_back((int64_t)dec.size());
        for (const Cluster& c : dec) {
            flat.push_back(c.nVertices);
            flat.push_back(c.nEdges);
            flat.push_back(c.cut);
            for (VertexIdx v : c.vertices) flat.push_back(v);
        }
        put(out, "clusters.i64", flat.data(), sizeof(int64_t) * flat.size());
        std::vector<int64_t> cm(n, -1);
        for (auto& kv : cmap) cm[kv.first] = kv.second;
        put(out, "cluster_map.i64", cm.data(), sizeof(int64_t) * n);
    }

    std::string tpath = out + "/times.txt";
    FILE* ft = fopen(tpath.c_str(), "w");
    if (ft) {
        fprintf(ft, "load_csr %.6f\ncount %.6f\nclean %.6f\nfull_pipeline %.6f\n",
                secs(t0, t1), secs(t2, t3), secs(t4, t5), t_extract);
        fclose(ft);
    }
    return 0;
}
