/* rtrex.h -- C ABI of the MI355X (gfx950) RTRExtractor decomposition path.
 *
 * Drop-in boundary for the reference's GPU layer (RTRex_CUDA/gpu_*.cuh) and its
 * CPU twin (RTRex_Sequential/Escape/Decomposition.h): graph load -> CSR ->
 * weighted triangle supports -> cascading clean -> greedy extraction.
 *
 * Conventions
 *   - C linkage, plain pointers and sizes; no C++ or torch types.
 *   - Every call returns an rtrex_status; nothing calls exit().  The message of
 *     the last failure on the calling thread is in rtrex_last_error().
 *   - Host arrays are caller-allocated; the library copies in and out.
 *   - A graph handle owns its device memory and one HIP stream.  Different
 *     handles may be used from different host threads; one handle must not be
 *     used by two threads at once (same as the reference's DeviceGraph).
 *   - CSR layout is the reference's (Escape/Graph.h:64-106): offsets int64[n+1],
 *     nbors int64[nnz], nnz = 2m directed slots, lists sorted by id, the graph
 *     simple and undirected.  Slot indices in outputs refer to that layout.
 *   - Edge status bytes use the reference's characters: 'Y' alive, 'N' deleted,
 *     'S' queued for deletion (Decomposition.h:127).
 *   - Limits: nnz < 2^32 and n < 2^32 (device ids and slots are 32-bit).
 */
#ifndef RTREX_H
#define RTREX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    RTREX_OK = 0,
    RTREX_EINVAL = 1,   /* bad argument or malformed graph                   */
    RTREX_EDEVICE = 2,  /* HIP runtime error or no usable gfx950 device      */
    RTREX_ENOMEM = 3,   /* device or host allocation failed                  */
    RTREX_EIO = 4,      /* file could not be read / is truncated             */
    RTREX_ESTATE = 5,   /* phase called out of order                         */
    RTREX_EBUSY = 6     /* server/batch queue full                           */
} rtrex_status;

typedef struct rtrex_graph_s* rtrex_graph;

/* Message of the last failure on this thread ("" if none). */
const char* rtrex_last_error(void);

/* Library build string (kernels, target). */
const char* rtrex_version(void);

/* ------------------------------------------------------------------------ */
/* Graph load.  Replaces loadGraph/loadGraph_Escape (GraphIO.cpp:19-88):    */
/* header "n m", '#' comments and blank lines skipped, the first two        */
/* integers of each line are the edge (extra columns such as weights are    */
/* ignored, GraphIO.cpp:40-41), fewer than m edge lines is RTREX_EIO        */
/* (GraphIO.cpp:67-71).  Edges beyond m are ignored.  Returns the edges as  */
/* written (one per line, not mirrored).                                     */
rtrex_status rtrex_escape_header(const char* path, int64_t* n, int64_t* m);
rtrex_status rtrex_escape_read(const char* path, int64_t m, uint32_t* u,
                               uint32_t* v);

/* The same rules parsed on the device, and the reference's SNAP conversion
 * (RTRex_CUDA/snap_to_rtr.py:47-142): '#'/'%' comment lines and blank lines
 * skipped, lines without two integer tokens skipped, self-loops dropped, ids
 * remapped to 0..n-1 by sorted order when they are not already, each
 * undirected pair kept once, edges sorted. */
enum { RTREX_FORMAT_ESCAPE = 0, RTREX_FORMAT_SNAP = 1 };

/* Reads an edge-list text file (format RTREX_FORMAT_ESCAPE or _SNAP), parses
 * it on the device and builds the graph there from the parsed edges, with no
 * edge list on the host: loadGraph_Escape + makeCSR (GraphIO.cpp:19-88,
 * Graph.cpp:344-384), or snap_to_rtr.py then both.  Errors as
 * rtrex_escape_read and rtrex_graph_create_edges. */
rtrex_status rtrex_graph_load(const char* path, int format, int device, rtrex_graph* out);

/* snap_to_rtr.py's conversion on the device: the SNAP file in_path written
 * as the Escape file out_path ("n m", then "u v" per edge, u < v, sorted);
 * *n and *m (nullable) receive its vertex and edge counts. */
rtrex_status rtrex_snap_convert(const char* in_path, const char* out_path, int device,
                                int64_t* n, int64_t* m);

/* ------------------------------------------------------------------------ */
/* Graph lifecycle.                                                          */

/* Host CSR -> device.  Replaces transferGraphToDevice + allocateDeviceState
 * (gpu_graph.cuh:47-57; gpu_graph.cu:46-135).  The partner slot of every edge
 * (CGraph::getPartnerMap, Graph.cpp:320-340) is derived on the device.  The
 * CSR must be simple, symmetric and sorted (RTREX_EINVAL otherwise). */
rtrex_status rtrex_graph_create_csr(const int64_t* offsets,
                                    const int64_t* nbors, int64_t n,
                                    int64_t nnz, int device, rtrex_graph* out);

/* Undirected edge list (each edge once, any order) -> device CSR.  Replaces
 * the Escape loader's mirroring (GraphIO.cpp:54-62) + makeCSR
 * (Graph.cpp:344-384) + getPartnerMap: the CSR is built on the device with a
 * radix sort.  Self-loops, duplicate edges and ids >= n are RTREX_EINVAL. */
rtrex_status rtrex_graph_create_edges(const uint32_t* u, const uint32_t* v,
                                      int64_t n, int64_t m, int device,
                                      rtrex_graph* out);

/* Releases all device memory of the handle.  Replaces freeDeviceGraph
 * (gpu_graph.cuh:62).  NULL is accepted. */
void rtrex_graph_destroy(rtrex_graph g);

rtrex_status rtrex_graph_info(rtrex_graph g, int64_t* n, int64_t* nnz);

/* Copies the device CSR back in the reference layout (for makeCSR parity). */
rtrex_status rtrex_graph_get_csr(rtrex_graph g, int64_t* offsets,
                                 int64_t* nbors);

/* Device bytes a graph of this size holds while resident (all phases). */
int64_t rtrex_graph_footprint_bytes(int64_t n, int64_t nnz);

/* ------------------------------------------------------------------------ */
/* Phases (each also callable alone, in this order).                         */

/* Weighted triangle supports.  Replaces gpuCommonNbr
 * (gpu_triangle_count.cuh:23) / commonNbr (Decomposition.h:42-117).  The
 * per-edge and per-vertex FP64 sums are accumulated in the reference's order
 * (ascending third vertex per edge; lexicographic per vertex), so they are
 * bit-identical to the sequential reference.  *total = triangle count. */
rtrex_status rtrex_count(rtrex_graph g, int64_t* total);

/* Current supports, per directed slot (both copies of an edge are equal) and
 * per vertex.  Replaces syncWeightsDeviceToHost (gpu_graph.cuh:67-71).  Any
 * pointer may be NULL. */
rtrex_status rtrex_get_supports(rtrex_graph g, double* perEdge,
                                double* perVertex, int64_t* total);

/* Cascading clean to convergence.  Replaces gpuInitialClean
 * (gpu_clean.cuh:30) / initialClean (Decomposition.h:302-460): every edge
 * with support < eps * 1/(d_u d_v) is deleted, with cascades, until none
 * remains.  There is no round cap. */
rtrex_status rtrex_initial_clean(rtrex_graph g, double eps);

/* Status byte per directed slot ('Y','N','S').  Replaces
 * syncEdgeStatusDeviceToHost (gpu_graph.cuh:76). */
rtrex_status rtrex_get_edge_status(rtrex_graph g, uint8_t* status);

/* Page-locked host memory that the device can address, for result arrays
 * (cluster maps, survivor words): a single graph's cluster map given in it is
 * written by the device directly, with no staging copy.  No reference
 * counterpart (the reference's results are ordinary host arrays,
 * gpu_graph.cuh:67-81).  rtrex_host_free(NULL) is a no-op. */
rtrex_status rtrex_host_alloc(size_t bytes, void** out);
void rtrex_host_free(void* p);

/* Survivors of the last rtrex_initial_clean, packed: bit (e % 32) of
 * bits[e / 32] is set when undirected edge e is alive, edges numbered in
 * (u < v) lexicographic order -- the order of the reference's sorted survivor
 * dump (Decomposition.h:419-452).  Stays valid through extraction.  nwords
 * must be >= (nnz/2 + 31) / 32. */
rtrex_status rtrex_get_survivors(rtrex_graph g, uint32_t* bits, int64_t nwords);

/* Deletes the k undirected edges pairs[2k] (each must be alive) and cleans to
 * convergence.  Replaces gpuDeleteAndClean (gpu_clean.cuh:44-45) /
 * deleteAndClean (Decomposition.h:147-283). */
rtrex_status rtrex_delete_and_clean(rtrex_graph g, const int64_t* pairs,
                                    int64_t k, double eps);

/* One seed's inner loop on the current state (after rtrex_initial_clean and
 * any rtrex_delete_and_clean calls, before rtrex_extract): the two-hop
 * candidates of seed u in ascending id with their triangle weights (the
 * first max_cand written; *n_cand = how many there are), the internal and
 * the potential weight.  Replaces gpuExtractInner (gpu_extract.cuh:28-37),
 * with the sequential reference's loop (Decomposition.h:728-827): pairs of
 * alive neighbours v, w with slot(w) >= slot(v), both of degree at most
 * d_u / eps; weights added in that loop's order.  A seed without alive
 * neighbours returns no candidates and zero weights.  The statuses are left
 * unchanged. */
rtrex_status rtrex_extract_inner(rtrex_graph g, int64_t u, double eps, int64_t* candidates,
                                 double* cand_wgt, int64_t max_cand, int64_t* n_cand,
                                 double* internal, double* potential);

/* The (member, neighbour) pairs that delete a cluster's alive edges, for
 * rtrex_delete_and_clean: members in the order given (the reference iterates
 * its unordered_set), each member's alive slots in order, an edge listed from
 * its first member only (Decomposition.h:897-914).  pairs[2 * max_pairs]
 * receives the first max_pairs; *n_pairs = how many there are.  Replaces
 * gpuMarkClusterEdges (gpu_extract.cuh:53-59) without marking: the statuses
 * are left unchanged, rtrex_delete_and_clean marks and cascades. */
rtrex_status rtrex_cluster_edges(rtrex_graph g, const int64_t* verts, int64_t n_verts,
                                 int64_t* pairs, int64_t max_pairs, int64_t* n_pairs);

/* Greedy extraction after rtrex_initial_clean.  Replaces disjointExtractCUDA
 * (RTRex_CUDA/main.cu:130-338) / disjointExtract (Decomposition.h:538-940).
 * cluster_map[n] receives the reference's cluster id of every vertex (ids in
 * the order the sequential reference creates them); *n_clusters the count.
 * Every decision is certified against the reference's own arithmetic or
 * re-run in its order; a ratio-prefix decision within the per-vertex
 * supports' error bound cannot be re-run (the reference subtracts from a
 * vertex's support in its global stack order) and is counted in
 * rtrex_audit.ambiguous_ratio: when that count is 0 the clusters are provably
 * the reference's. */
rtrex_status rtrex_extract(rtrex_graph g, double eps, int64_t* cluster_map,
                           int64_t* n_clusters);

/* count + initial clean + extract in one call (RTRex.cpp:193).  For both
 * calls a null cluster_map leaves the map in device memory only (no host
 * copy); rtrex_get_cluster_map fetches it later. */
rtrex_status rtrex_decompose(rtrex_graph g, double eps, int64_t* cluster_map,
                             int64_t* n_clusters);

/* The cluster map of the last extraction (n entries, per-graph ids on a
 * batch handle): the cluster_map output of rtrex_extract, fetched on demand.
 * No reference counterpart (the reference holds its clusters in host memory,
 * Decomposition.h:538-543). */
rtrex_status rtrex_get_cluster_map(rtrex_graph g, int64_t* cluster_map);

/* Position at which every vertex joined its cluster in the last extraction:
 * 0 for the seed, then its alive neighbours in slot order, then the chosen
 * candidates in sorted order (Decomposition.h:610-631,869-875).  With
 * rtrex_write_decomposition this reproduces the reference's file order. */
rtrex_status rtrex_get_member_order(rtrex_graph g, int64_t* rank);

/* Per-cluster statistics over the original graph, populateStats
 * (Decomposition.h:473-491): arrays of length n_clusters, indexed by id. */
rtrex_status rtrex_cluster_stats(rtrex_graph g, int64_t* n_vertices,
                                 int64_t* n_edges, int64_t* cut);

/* rtrex_cluster_stats without any allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (serving loops,
 * harnesses that forbid hipMalloc).  The same results, with the per-cluster
 * accumulators in a device array of the library (24 MiB per GPU, held until
 * the process ends) instead of the handle's scratch.  At most 1,048,576
 * clusters (RTREX_EINVAL beyond).  Calls on one GPU share that array: calls
 * that overlap on one GPU race. */
rtrex_status rtrex_cluster_stats_noalloc(rtrex_graph g, int64_t* n_vertices,
                                         int64_t* n_edges, int64_t* cut);

/* Connected components of the original graph.  Replaces connectedComp
 * (Decomposition.h:1023-1068), which RTRex.cpp:203-204 runs after the
 * extraction: comp[n] (nullable) receives each vertex's component, numbered
 * as the reference's BFS finds them (by smallest vertex), *n_comp their
 * number.  A union-find on the device; callable once the graph is built. */
rtrex_status rtrex_connected_components(rtrex_graph g, int64_t* comp, int64_t* n_comp);

/* Decision audit.  Every threshold or ratio decision whose margin is within
 * the FP64 error bound of the reference's own summation is counted as
 * ambiguous; with zero ambiguous decisions the survivor and cluster sets are
 * provably those of the sequential reference.  A cleaning decision that the
 * bound cannot settle makes the clean re-run the triangle-connected regions
 * holding such edges in the reference's own stack order (exact FP64, so those
 * decisions are the reference's): they are counted in replayed_edges, not in
 * ambiguous_clean. */
typedef struct {
    int64_t clean_rounds;         /* frontier rounds, all clean calls        */
    int64_t edges_removed;        /* undirected edges deleted by cleaning    */
    int64_t triangles_removed;
    int64_t ambiguous_clean;      /* near-threshold removal/survival         */
    int64_t ambiguous_ratio;      /* near-tie ratio-prefix decisions         */
    int64_t seeds;                /* clusters formed                          */
    int64_t components;           /* alive components at extraction start    */
    int64_t largest_component;    /* vertices in the largest one             */
    int64_t replayed_edges;       /* edges of the regions the clean re-ran   */
                                  /* in the reference's exact stack order    */
} rtrex_audit;
rtrex_status rtrex_get_audit(rtrex_graph g, rtrex_audit* out);

/* Device time of the last run of each phase, milliseconds (HIP events). */
typedef struct {
    double build_ms, count_ms, clean_ms, extract_ms;
    double count_edge_kernel_ms;  /* per-edge support kernel(s) alone        */
    double count_vertex_kernel_ms;/* per-vertex fold kernel(s) alone         */
} rtrex_timings;
rtrex_status rtrex_get_timings(rtrex_graph g, rtrex_timings* out);

/* ------------------------------------------------------------------------ */
/* Count-plan parameters and the autotuner (no reference counterpart).        */
/* An edge whose longer-list endpoint (its hub) has degree in                 */
/* (hub_min, hub_max] is counted per hub with the hub's list staged in LDS:   */
/* a sorted list of up to 64 entries (degree <= l0_max), 512 (<= l1_max) or   */
/* 4096 (<= l2_max) with an 8-ary pivot tree; above l2_max (hub_max <=       */
/* 262136) the tree's upper levels are staged and its leaf is the list in     */
/* device memory; other edges are counted per edge.                            */
/* With stored triangle lists, hubs up to l0_max are counted in batches of    */
/* whole hubs (units of at least unit_min streamed elements, grouped into     */
/* chunks of about 1,024), and hubs up to l2_max whose edges' shorter lists   */
/* average at most g_thr[3] elements by a streaming kernel (units of          */
/* max(2 * hub degree, 4 * unit_min) elements); other hubs' edges get 4, 8,   */
/* 16, 32 or 64 lanes each by their mean shorter-list length against          */
/* g_thr[0..3] (units: unit_min, or unit_hash_mult * hub degree above l1_max).*/
/* Results never depend on the parameters, only the time does.                */
typedef struct {
    int32_t hub_min, hub_max, l0_max, l1_max;
    int32_t g_thr[4];
    int32_t unit_min, unit_hash_mult;
    int32_t l2_max; /* l1_max <= l2_max <= 4096 */
    int32_t l3_max; /* l2_max <= l3_max <= 32760: hubs above take a tree of */
                    /* five levels instead of four (up to 262136 entries)   */
} rtrex_count_params;

rtrex_status rtrex_get_count_params(rtrex_graph g, rtrex_count_params* out);

/* Validated; the plan is rebuilt at the next count. */
rtrex_status rtrex_set_count_params(rtrex_graph g, const rtrex_count_params* p);

/* Times the count phase (device time of its kernels, best of `repeats`) over
 * a coordinate search of the parameters above, starting from the current
 * ones, and keeps the fastest setting on the handle.  *best receives it,
 * *best_ms its time and *start_ms the starting setting's time (nullable).
 * Leaves the handle counted (as after rtrex_count). */
rtrex_status rtrex_autotune_count(rtrex_graph g, int repeats,
                                  rtrex_count_params* best, double* best_ms,
                                  double* start_ms);

/* ------------------------------------------------------------------------ */
/* Batched decomposition (no reference counterpart: the reference runs one   */
/* graph per process).  n_graphs independent graphs given as concatenated    */
/* edge lists with LOCAL ids: graph i has vtx[i] vertices and edges          */
/* [eoff[i], eoff[i+1]) of u/v.  All graphs are resident at once as one      */
/* disjoint union and every phase runs over all of them per launch.          */
/* Results equal running each graph alone.                                   */

/* Builds the resident batch.  The handle works with every phase call above;
 * rtrex_decompose / rtrex_extract then return PER-GRAPH cluster ids (graph i
 * occupies cluster_map[sum(vtx[0..i)) ...]). */
rtrex_status rtrex_batch_create(int64_t n_graphs, const int64_t* vtx,
                                const int64_t* eoff, const uint32_t* u,
                                const uint32_t* v, int device,
                                rtrex_graph* out);

/* Clusters per graph of the last extraction of a batch handle. */
rtrex_status rtrex_batch_cluster_counts(rtrex_graph g, int64_t* n_clusters,
                                        int64_t n_graphs);

/* One-shot: rtrex_batch_create + rtrex_decompose + counts + destroy. */
rtrex_status rtrex_batch_decompose(int64_t n_graphs, const int64_t* vtx,
                                   const int64_t* eoff, const uint32_t* u,
                                   const uint32_t* v, double eps, int device,
                                   int64_t* cluster_map, int64_t* n_clusters,
                                   rtrex_timings* timings, rtrex_audit* audit);

/* ------------------------------------------------------------------------ */
/* Output files, byte-identical to the reference driver's (host code; no GPU  */
/* needed).  CSR in the reference layout.                                      */

/* <name>-decomposition.txt, -stats.txt, -cut.txt, -freq.txt as
 * printDecomposition (RTRex.cpp:19-118) writes them for the clusters given by
 * cluster_of[n] (ids 0..n_clusters-1) with member_rank[n] the order in which
 * vertices joined each cluster (rtrex_get_member_order). */
rtrex_status rtrex_write_decomposition(const char* name, int64_t n,
                                       const int64_t* offsets,
                                       const int64_t* nbors,
                                       const int64_t* cluster_of,
                                       const int64_t* member_rank,
                                       int64_t n_clusters);

/* expandClusters (Decomposition.h:943-1008) on the clusters given as for
 * rtrex_write_decomposition, then the same four files for the expanded
 * clusters plus <name>-scores.txt (printScores, RTRex.cpp:120-135): each
 * singleton cluster's vertex joins the neighbouring cluster holding the most
 * of its neighbours, when that count exceeds threshold (the reference
 * driver's mode e, RTRex.cpp:166-174,196-201, writes them with the name
 * "RTRex-expanded-<name>").  expanded_of[n] (nullable) receives the cluster
 * of every vertex after expansion; score[n] (nullable) the moved vertices'
 * score (neighbours in the cluster / degree), -1 for the others. */
rtrex_status rtrex_write_expanded(const char* name, int64_t n, const int64_t* offsets,
                                  const int64_t* nbors, const int64_t* cluster_of,
                                  const int64_t* member_rank, int64_t n_clusters,
                                  int32_t threshold, int64_t* expanded_of, double* score);

/* The same four files for the connected components, connectedComp
 * (Decomposition.h:1023-1068), as RTRex.cpp:203-204 writes them. */
rtrex_status rtrex_write_components(const char* name, int64_t n,
                                    const int64_t* offsets,
                                    const int64_t* nbors);

/* The survivor dump of initialClean (Decomposition.h:419-452): "u v" per
 * surviving edge, u < v, sorted; status[nnz] as rtrex_get_edge_status. */
rtrex_status rtrex_write_survivors(const char* path, int64_t n,
                                   const int64_t* offsets, const int64_t* nbors,
                                   const uint8_t* status);

/* ------------------------------------------------------------------------ */
/* Serving (no reference counterpart).  A request loop: jobs are queued,      */
/* admitted by HBM footprint into one device pool of hbm_budget_bytes, run    */
/* by `workers` host threads on their own HIP streams (launches overlap on    */
/* the GPU), and small jobs waiting behind each other are batched into one    */
/* launch (continuous batching; per-graph results equal running alone).       */
typedef struct rtrex_server_s* rtrex_server;

typedef struct {
    int device;
    int workers;               /* concurrent launches, 1..16                  */
    int64_t hbm_budget_bytes;  /* device pool shared by admitted launches     */
    int64_t small_job_edges;   /* jobs up to this many edges may be batched   */
    int64_t batch_max_edges;   /* edge cap of one batched launch              */
    int64_t batch_max_graphs;  /* graph cap of one batched launch             */
} rtrex_server_config;         /* zero caps select defaults (64Ki, 4Mi, 1024) */

typedef struct {
    double submit_s, start_s, end_s; /* server clock (rtrex_server_clock)     */
    int64_t batch_graphs;            /* graphs in the launch that ran the job */
    int64_t bytes;                   /* HBM reserved for that launch          */
    int32_t worker;
    int32_t pooled;                  /* 0: ran outside the pool (too large)   */
} rtrex_job_info;

rtrex_status rtrex_server_start(const rtrex_server_config* cfg, rtrex_server* out);

/* Queues one decomposition of an undirected edge list (each edge once, ids <
 * n); the edges are copied.  *job receives the job id. */
rtrex_status rtrex_server_submit(rtrex_server s, const uint32_t* u,
                                 const uint32_t* v, int64_t n, int64_t m,
                                 double eps, int64_t* job);

/* Blocks until the job is done, then returns its status (the job's own error
 * in rtrex_last_error) and, on success, the cluster ids (cluster_map[n],
 * nullable) and count.  The job id is released. */
rtrex_status rtrex_server_wait(rtrex_server s, int64_t job, int64_t* cluster_map,
                               int64_t* n_clusters, rtrex_job_info* info);

/* Seconds since rtrex_server_start (the clock of rtrex_job_info). */
double rtrex_server_clock(rtrex_server s);

/* Runs every queued job, joins the workers and frees the pool. */
rtrex_status rtrex_server_stop(rtrex_server s);

#ifdef __cplusplus
}
#endif
#endif /* RTREX_H */
