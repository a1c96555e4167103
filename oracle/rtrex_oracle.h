/* CPU oracle for the RTRExtractor decomposition path (count -> clean -> extract).
 *
 * TEST INFRASTRUCTURE ONLY.  This is a project-written C++ restatement of the
 * reference's sequential algorithm (RTRex_Sequential, mode 's').  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it, and only as
 * the checker.  The product (rtrex_cuda_amd) never links, loads or calls it.
 *
 * Every function cites the reference file:line it restates.  The restatement
 * is pinned by tests/test_oracle.py against: the golden vectors produced by the
 * reference itself (oracle/_ref/ref_dump built from the reference sources),
 * brute-force triangle enumeration, K_n closed forms and the order-independent
 * fixpoint of the clean phase.
 *
 * All arrays use the reference's CSR layout: offsets int64[n+1], nbors
 * int64[nnz] with nnz = 2m directed slots, each adjacency list sorted by id.
 */
#ifndef RTREX_ORACLE_H
#define RTREX_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* COO (both directions already present, as the Escape loader emits them,
 * GraphIO.cpp:54-62) -> CSR sorted by (src, dst) (Graph.cpp:344-384). */
int oracle_make_csr(int64_t n, int64_t nnz, const int64_t* src,
                    const int64_t* dst, int64_t* offsets, int64_t* nbors);

/* Weighted triangle supports, commonNbr (Decomposition.h:42-117). */
int oracle_count(int64_t n, const int64_t* offsets, const int64_t* nbors,
                 int64_t* total, double* perEdge, double* perVertex);

/* initialClean with sequential seeding (Decomposition.h:302-460) applied to
 * supports produced by oracle_count.  perEdge/perVertex/total are updated in
 * place; status receives 'Y' or 'N' per directed slot. */
int oracle_initial_clean(int64_t n, const int64_t* offsets,
                         const int64_t* nbors, double eps, double* perEdge,
                         double* perVertex, int64_t* total, uint8_t* status);

/* deleteAndClean (Decomposition.h:147-283) on an existing state: the k
 * undirected edges in pairs[2k] (each must be 'Y') are pushed in order, marked
 * 'S' and then cascaded to convergence. */
int oracle_delete_and_clean(int64_t n, const int64_t* offsets,
                            const int64_t* nbors, double eps,
                            const int64_t* pairs, int64_t k, double* perEdge,
                            double* perVertex, int64_t* total,
                            uint8_t* status);

/* disjointExtract's seed loop alone (Decomposition.h:538-940), from a state
 * after oracle_initial_clean and any oracle_delete_and_clean calls (the
 * supports are updated in place by the extraction's cascades). */
int oracle_extract_state(int64_t n, const int64_t* offsets, const int64_t* nbors,
                         double eps, double* perEdge, double* perVertex,
                         int64_t* total, const uint8_t* status, int64_t* cluster_map,
                         int64_t* n_clusters);

/* The inner loop of one seed u (Decomposition.h:728-827) on the given slot
 * statuses ('N' deleted): the two-hop candidates in ascending id with their
 * triangle weights (the first cap of them written; *ncand = how many),
 * internal and potential weight.  The same function runs inside
 * oracle_extract_state and oracle_decompose. */
int oracle_extract_inner(int64_t n, const int64_t* offsets, const int64_t* nbors,
                         const uint8_t* status, int64_t u, double eps, int64_t* cand,
                         double* wgt, int64_t cap, int64_t* ncand, double* internal,
                         double* potential);

/* The deletion pairs of a cluster (Decomposition.h:897-914) on the given slot
 * statuses: members in the given order, each one's 'Y' slots in order, an
 * edge listed once (its slots marked 'S' on a copy of the statuses).  The
 * same function builds the seed loop's worklists. */
int oracle_cluster_edges(int64_t n, const int64_t* offsets, const int64_t* nbors,
                         const uint8_t* status, const int64_t* verts, int64_t nverts,
                         int64_t* pairs, int64_t cap, int64_t* npairs);

/* Full disjointExtract (Decomposition.h:538-940): count, clean, then greedy
 * extraction.  Outputs: cluster_map[n] (cluster id per vertex, ids in
 * extraction order), *n_clusters, survivors status[nnz] after the initial
 * clean (may be NULL), and phase wall times in seconds t[3] (may be NULL).
 * If flat != NULL it receives, per cluster in id order: nVertices, nEdges,
 * cut, then the vertices in the reference's iteration order; flat_cap bounds
 * it (n_clusters*3 + n suffices) and *flat_len receives the length used.
 * member_rank[n] (may be NULL) receives the position at which each vertex
 * joined its cluster (seed 0, alive neighbours, then chosen candidates,
 * Decomposition.h:610-631,869-875). */
int oracle_decompose(int64_t n, const int64_t* offsets, const int64_t* nbors,
                     double eps, int64_t* cluster_map, int64_t* n_clusters,
                     uint8_t* status_after_clean, int64_t* flat,
                     int64_t flat_cap, int64_t* flat_len, double* t,
                     int64_t* member_rank);

/* connectedComp (Decomposition.h:1023-1068): component id per vertex in BFS
 * discovery order; returns the number of components in *n_comp. */
int oracle_connected_components(int64_t n, const int64_t* offsets,
                                const int64_t* nbors, int64_t* comp,
                                int64_t* n_comp);

#ifdef __cplusplus
}
#endif
#endif
