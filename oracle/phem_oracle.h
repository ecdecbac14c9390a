/*
 * phem_oracle.h -- CPU restatement of PHEM's MST / connected-components path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only the test suite,
 * the smoke check and the CPU-baseline leg of the benchmark may load it.  The
 * product (the HIP library) never links, loads or calls anything here.
 *
 * Conventions shared with the product:
 *   - an edge is (u[i], v[i], w[i]) with id i (its index in the input arrays);
 *   - an edge with u == v (self-loop) or with an endpoint >= n is not part of
 *     the graph and is ignored;
 *   - w is a 32-bit pattern.  wmode 0: unsigned integer weight (the reference's
 *     `typedef uint32_t weight`, MST/src/constants.h:9).  wmode 1: IEEE float32,
 *     mapped to an unsigned order key by or_wkey();
 *   - the total order on edges is key(e) = (wkey(w_e) << 32) | id_e, i.e. weight
 *     first, edge id second (MST/src/utils.cu:801-802, MST/src/bench/serialMST.C:67-68;
 *     the GPU path's (weight, position) order with stable compaction,
 *     MST/src/kernels.cu:24-70,641-646, is the same order);
 *   - component labels are the minimum vertex id of the component
 *     (hooking is always larger root -> smaller root, CC/host_spanning_tree.cu:31-38).
 */
#ifndef PHEM_ORACLE_H
#define PHEM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_NO_EDGE 0xFFFFFFFFu

/* Order-preserving map of a weight pattern to an unsigned key. */
uint32_t or_wkey(uint32_t w, int wmode);

/* Sequential Kruskal over key order with a disjoint-set forest.
 * Writes the MSF edge ids in ascending id order to out_ids (capacity >= n) and,
 * if out_labels != NULL, min-id component labels.  Returns the MSF size. */
int64_t or_kruskal(uint32_t n, uint64_t m, const uint32_t *u, const uint32_t *v,
                   const uint32_t *w, int wmode, uint32_t *out_ids,
                   uint32_t *out_labels);

/* Canonical Boruvka restatement of the reference round loop
 * (MST/src/kernels.cu:507-689): per round, every component picks its minimum-key
 * incident alive edge; a mutual pick keeps the lower id as root
 * (kernels.cu:73-134); all vertices are shortcut to their root (kernels.cu:254-271);
 * edges are relabelled and self-loops dropped (kernels.cu:273-321,641-649).
 * log_m[r] = alive edges at the start of round r, log_c[r] = components with at
 * least one alive edge in round r (both may be NULL; capacity max_rounds).
 * Returns the MSF size (ids ascending in out_ids), or -1 if max_rounds is hit. */
int64_t or_boruvka(uint32_t n, uint64_t m, const uint32_t *u, const uint32_t *v,
                   const uint32_t *w, int wmode, uint32_t *out_ids,
                   uint32_t *out_labels, uint64_t *log_m, uint64_t *log_c,
                   int max_rounds, int *rounds_out);
/* State of or_boruvka's round loop after `rounds` rounds: forest edges chosen
 * so far (count returned) and min-id labels of the components so far. */
int64_t or_boruvka_rounds(uint32_t n, uint64_t m, const uint32_t *u,
                          const uint32_t *v, const uint32_t *w, int wmode,
                          int rounds, uint32_t *out_ids, uint32_t *out_labels);

/* Union-find labels with the reference hook rule (larger root -> smaller root,
 * CC/host_spanning_tree.cu:10-42), processed sequentially in edge order.
 * If forest != NULL (n slots), forest[root] = id of the edge that hooked root,
 * OR_NO_EDGE otherwise (the sequential schedule of CC/ExternalSpanningTree.cu:46-59). */
void or_uf_labels(uint32_t n, uint64_t m, const uint32_t *u, const uint32_t *v,
                  uint32_t *labels, uint32_t *forest);

/* Breadth-first-search labels: vertices scanned in increasing id, so each BFS
 * tree is labelled with its smallest vertex.  Independent of union-find. */
void or_bfs_labels(uint32_t n, uint64_t m, const uint32_t *u, const uint32_t *v,
                   uint32_t *labels);

/* Spanning-forest invariants: every id < m names a valid edge, no id repeats,
 * the edges are acyclic, |F| = n - (number of components), and the forest's
 * components equal `labels` (min-id).  Returns 0 if all hold, else a code > 0. */
int or_check_forest(uint32_t n, uint64_t m, const uint32_t *u, const uint32_t *v,
                    const uint32_t *ids, uint64_t count, const uint32_t *labels);

/* Exact total weight of the edges `ids`, summed in ascending id order.
 * wmode 1: returns sum of w * 2^24 as an integer and sets *exact = 0 if some
 * weight is negative or not a multiple of 2^-24.  wmode 0: integer sum. */
uint64_t or_weight_sum(uint64_t count, const uint32_t *ids, const uint32_t *w,
                       int wmode, int *exact);

/* ---- parallel CPU baseline (OpenMP), restating the reference CPU paths ---- */

/* Filter-Kruskal with deterministic reservations: the PBBS mst()
 * (MST/src/utils.cu:794-841, reserve/commit :757-792,
 * MST/src/common/speculative_for.h:51-113, MST/src/algorithm/union_find.h:32-83).
 * Same output contract as or_kruskal. threads <= 0: OpenMP default. */
int64_t or_par_mst(uint32_t n, uint64_t m, const uint32_t *u, const uint32_t *v,
                   const uint32_t *w, int wmode, uint32_t *out_ids,
                   int threads);

/* Asynchronous union-find over all edges in parallel, then min-id labels
 * (SimpleUnionAsync, CC/host_spanning_tree.cu:67-86). forest as or_uf_labels
 * but holding a race-dependent choice of hooking edge. */
void or_par_cc(uint32_t n, uint64_t m, const uint32_t *u, const uint32_t *v,
               uint32_t *labels, uint32_t *forest, int threads);

#ifdef __cplusplus
}
#endif
#endif
