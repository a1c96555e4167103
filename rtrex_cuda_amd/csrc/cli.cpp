// rtrex: the command-line driver of the decomposition path, on the C ABI.
//
// Mirrors the reference's drivers:
//   rtrex <graph_file> <output_name> <eps> [mode]        RTRex.cpp:141-206
//   rtrex <graph_file> <output_name> <eps> <gpu_device>  main.cu:343-391
// A numeric fourth argument selects a GPU the way main.cu does and writes the
// "RTRex-CUDA-<name>" files only; otherwise the mode letter is RTRex.cpp's
// (s simple, p parallel: the same device path here; e expand, followed by
// the threshold, RTRex.cpp:166-174) and the files are those RTRex.cpp writes:
//   RTRex-<name>-{decomposition,stats,cut,freq}.txt
//   RTRex-expanded-<name>-{decomposition,stats,cut,freq,scores}.txt (mode e)
//   connected-components-<name>-{decomposition,stats,cut,freq}.txt
//   amazon_cleaned_edges.txt (the survivor dump of initialClean)
// All of them byte-identical to the reference's, written in the working
// directory.  --device N picks the GPU in either form.
#include <cctype>
#include <chrono>
#include <cinttypes>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/rtrex.h"

namespace {

int die(const char* what) {
    std::fprintf(stderr, "rtrex: %s: %s\n", what, rtrex_last_error());
    return 1;
}

double secs_since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

bool numeric(const char* s) {
    if (!*s) return false;
    for (; *s; ++s)
        if (!std::isdigit((unsigned char)*s)) return false;
    return true;
}

}  // namespace

int main(int argc, char** argv) {
    std::vector<const char*> pos;
    int device = 0;
    for (int i = 1; i < argc; ++i) {
        if (!std::strcmp(argv[i], "--device") && i + 1 < argc) device = std::atoi(argv[++i]);
        else pos.push_back(argv[i]);
    }
    if (pos.size() < 3) {
        std::printf("Usage: %s <graph_file> <output_name> <eps> [mode s|p|e [threshold] | "
                    "gpu_device] [--device N]\n",
                    argv[0]);
        return 1;
    }
    const std::string path = pos[0], name = pos[1];
    const double eps = std::atof(pos[2]);
    bool cuda_names = false;
    char mode = 's';
    if (pos.size() > 3) {
        if (numeric(pos[3])) {
            cuda_names = true;
            device = std::atoi(pos[3]);
        } else {
            mode = pos[3][0];
        }
    }
    std::printf("eps = %f\n", eps);
    if (!cuda_names) std::printf("mode = %c\n", mode);
    int threshold = 0;
    if (mode == 'e') {
        if (pos.size() <= 4) {
            std::printf("Error: Missing threshold value\n");
            return 1;
        }
        threshold = std::atoi(pos[4]);
        std::printf("threshold = %i\n", threshold);
    }
    if (mode != 's' && mode != 'p' && mode != 'e') {
        std::fprintf(stderr, "rtrex: unknown mode %c\n", mode);
        return 1;
    }

    int64_t n = 0, m = 0;
    if (rtrex_escape_header(path.c_str(), &n, &m) != RTREX_OK) return die("load");
    std::vector<uint32_t> u((size_t)m), v((size_t)m);
    if (rtrex_escape_read(path.c_str(), m, u.data(), v.data()) != RTREX_OK) return die("load");
    std::printf("Loaded graph\n");

    auto t0 = std::chrono::steady_clock::now();
    rtrex_graph g = nullptr;
    if (rtrex_graph_create_edges(u.data(), v.data(), n, m, device, &g) != RTREX_OK)
        return die("graph");
    int64_t nn = 0, nnz = 0;
    rtrex_graph_info(g, &nn, &nnz);
    std::printf("Vertices = %" PRId64 "\n", nn);
    std::printf("Edges = %" PRId64 "\n", nnz / 2);

    int64_t triangles = 0;
    if (rtrex_count(g, &triangles) != RTREX_OK) return die("count");
    std::printf("Total triangle count = %" PRId64 "\n", triangles);
    if (rtrex_initial_clean(g, eps) != RTREX_OK) return die("clean");
    std::vector<int64_t> off((size_t)nn + 1), nb((size_t)nnz);
    std::vector<uint8_t> st((size_t)nnz);
    if (rtrex_graph_get_csr(g, off.data(), nb.data()) != RTREX_OK) return die("csr");
    if (rtrex_get_edge_status(g, st.data()) != RTREX_OK) return die("status");
    int64_t alive = 0;
    for (uint8_t c : st) alive += c == 'Y';
    std::printf("Edges remaining after initialClean (Undirected): %" PRId64 "\n", alive / 2);
    if (!cuda_names &&
        rtrex_write_survivors("amazon_cleaned_edges.txt", nn, off.data(), nb.data(),
                              st.data()) != RTREX_OK)
        return die("survivors");

    std::vector<int64_t> cluster_of((size_t)nn), rank((size_t)nn);
    int64_t k = 0;
    if (rtrex_extract(g, eps, cluster_of.data(), &k) != RTREX_OK) return die("extract");
    if (rtrex_get_member_order(g, rank.data()) != RTREX_OK) return die("member order");
    std::printf("Total clusters extracted: %" PRId64 "\n", k);
    rtrex_timings tm;
    rtrex_get_timings(g, &tm);
    std::printf("Device time (ms): build %.3f count %.3f clean %.3f extract %.3f\n", tm.build_ms,
                tm.count_ms, tm.clean_ms, tm.extract_ms);
    std::printf("Time from load to clusters: %.6fs\n", secs_since(t0));
    rtrex_graph_destroy(g);

    const std::string pre = cuda_names ? "RTRex-CUDA-" + name : "RTRex-" + name;
    if (rtrex_write_decomposition(pre.c_str(), nn, off.data(), nb.data(), cluster_of.data(),
                                  rank.data(), k) != RTREX_OK)
        return die("write decomposition");
    if (mode == 'e' &&
        rtrex_write_expanded(("RTRex-expanded-" + name).c_str(), nn, off.data(), nb.data(),
                             cluster_of.data(), rank.data(), k, threshold, nullptr,
                             nullptr) != RTREX_OK)
        return die("write expanded");
    if (!cuda_names &&
        rtrex_write_components(("connected-components-" + name).c_str(), nn, off.data(),
                               nb.data()) != RTREX_OK)
        return die("write components");
    std::printf("Done.\n");
    return 0;
}
