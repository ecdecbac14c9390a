/*
 * phem_dropin.h -- C++ entry points with the reference's names, argument
 * meaning and error behaviour, implemented over the C ABI of phem.h, so that
 * the reference's callers (its main programs and loaders) link against this
 * library in place of its CUDA sources:
 *
 *   MST/src/kernels.cu, MST/src/new_boruvka.cu   -> boruvka_for_edge_array,
 *                                                   device_boruvka,
 *                                                   new_boruvka
 *   MST/src/utils.cu (file reading)               -> getGraphInfo,
 *                                                   mmapPopulateEdgeArray
 *   CC/ExternalSpanningTree.cu                    -> SpanningTree
 *   CC/GPUSpanningTree.cu                         -> gpu_spanning_tree
 *
 * The CPU path (CC/host_spanning_tree.cu with its ParlayLib) stays the
 * caller's: SpanningTree calls the caller's host_spanning_tree() on the CPU
 * share of the edges while the GPU batches run.  As in the reference, a HIP
 * failure in these entry points prints a message and exits the process
 * (CC/include/cuda_utility.cuh:7-17); the C ABI itself never exits.
 */
#ifndef PHEM_DROPIN_H
#define PHEM_DROPIN_H

#include <hip/hip_runtime.h>
#include <thrust/device_ptr.h>

#include <cstdint>
#include <fstream>
#include <limits>
#include <string>

/* ---- MST types (MST/src/constants.h:8-24) ------------------------------- */
typedef uint32_t vertex;
typedef uint32_t weight;
typedef uint64_t large_vertex;
const weight sentinel = std::numeric_limits<weight>::max();
const vertex sentinel_vertex = std::numeric_limits<vertex>::max();
struct edge {
  vertex u;
  vertex v;
  weight w;
};
struct outgoing_edge_id {
  vertex e_id;
  weight w;
};

/* Minimum spanning forest of host edges (MST/src/new_boruvka.cu:40-179).
 * Weights are unsigned integers (the reference's `weight`).  Writes the
 * forest's edges (u, v, w) in ascending input order to
 * final_msf[number_of_filled_edges ...] (the caller provides num_nodes
 * slots) and adds their count to number_of_filled_edges.  `result` is
 * accepted and unused, as in the reference's plain path. */
void boruvka_for_edge_array(edge *edges, vertex num_nodes, vertex num_edges,
                            edge *final_msf, vertex &number_of_filled_edges,
                            std::string result = "res.txt");

/* Vertex count and entry count from a MatrixMarket header
 * (MST/src/utils.cu:160-179); leaves the stream after the header. */
void getGraphInfo(std::ifstream &file, vertex &num_nodes,
                  large_vertex &num_edges_large);

/* Reads num_edges edges of a MatrixMarket file into edges (MST/src/utils.cu:
 * 896-1036): 0-based, u <= v, integer weights or the generated weight
 * ((u-1)(v-1) mod num_nodes) + 1.  Prints and returns on an unreadable file,
 * as the reference does. */
void mmapPopulateEdgeArray(const std::string &filename, edge *edges,
                           vertex num_edges, vertex num_nodes,
                           bool generate_random_weights = false);

/* The reference's MST driver (MST/src/new_boruvka.cu:330-510): reads a
 * MatrixMarket file and computes its minimum spanning forest.
 *  - default: one solve of the whole graph (boruvka_for_edge_array);
 *  - use_streamline: chunks of chunk_size edges (or num_chunks_ideal
 *    chunks) folded into a forest kept on the device, F = MSF(F u chunk)
 *    (MST/src/kernels.cu:849-1154);
 *  - use_cpu_gpu_streamline: the same, with part of the last chunk first
 *    reduced to its own forest on the CPU (phem_cpu_msf) while the GPU folds
 *    in the others (MST/src/kernels.cu:1155-1826): the part the cpu_share
 *    knob gives (phem_tuning; none by default, the measured best here);
 *  - use_malloc_managed (debug or not): the unified-memory mode
 *    (MST/src/new_boruvka.cu:183-323) as a host-resident graph read by the
 *    device in chunks sized to its free memory (phem_msf_host_chunks);
 *    prints "Time to execute T microseconds", "total weight W" and "number
 *    of filled edges K";
 *  - use_cpu_only (not in debug mode, as in the reference): the forest is
 *    computed on the host by the function registered with phem_set_cpu_mst
 *    (the caller's ParlayLib cpu_compute_pbbs, MST/src/utils.cu:887, behind
 *    an adapter), or by the library's host solver phem_cpu_msf when none is
 *    registered; prints "pbbs execution time T microseconds";
 *  - debug: the streamline modes run 10 times, printing "i th execution
 *    time T" (microseconds) and "Average exec time T"
 *    (MST/src/new_boruvka.cu:356-384, MST/src/kernels.cu:878-1138);
 *  - debug && save_weighted_graph: writes weighted.mtx and returns.
 * Dispatch order as in the reference (MST/src/new_boruvka.cu:356-506).
 * Prints "Total weight W" (the exact forest weight).  When result_filename is
 * not "res.txt" the forest is written there as "u+1 v+1 w" lines. */
void new_boruvka(std::string filename, std::string result_filename, bool debug,
                 vertex chunk_size, int num_chunks_ideal,
                 bool generate_random_weights = false,
                 bool use_malloc_managed = false, bool use_streamline = false,
                 bool use_cpu_only = false, bool use_cpu_gpu_streamline = false,
                 bool save_weighted_graph = false);

/* The CPU minimum spanning forest new_boruvka(use_cpu_only) calls: the n,
 * m edges (u[i], v[i], w[i]) with unsigned weights; writes the forest's edge
 * ids (any order; n slots) to ids_out and their number to *count; returns 0
 * on success.  No counterpart in the reference, which calls its ParlayLib
 * cpu_compute_pbbs directly (MST/src/new_boruvka.cu:468-480): this library
 * cannot name that function's ParlayLib argument type, so the caller
 * registers a plain-array adapter around it.  NULL restores the library's
 * host solver (phem_cpu_msf). */
typedef int (*phem_cpu_mst_fn)(uint32_t n, uint64_t m, const uint32_t *u,
                               const uint32_t *v, const uint32_t *w,
                               uint32_t *ids_out, uint64_t *count);
extern "C" void phem_set_cpu_mst(phem_cpu_mst_fn fn);

/* Minimum spanning forest of device edges (MST/src/kernels.cu:507-689).
 * Reads d_edges[0 .. num_edges); appends the forest's edges to
 * d_final_msf + number_of_filled_edges; writes min-id component labels to
 * d_representatives; sets num_edges to 0 (every edge consumed).  The other
 * buffers (d_iter_edges, d_msf, d_new_representatives,
 * d_iter_representatives, d_outgoing_edges) are accepted for signature
 * compatibility and left untouched: the solver needs about 32 B per edge of
 * workspace, more than those buffers hold, so its workspace is allocated and
 * cached by the library (grown on demand, one per device) until
 * phem_dropin_release_workspace frees it. */
void device_boruvka(edge *d_edges, edge *d_final_msf, edge *d_iter_edges,
                    edge *d_msf, vertex *d_representatives,
                    vertex *d_new_representatives,
                    vertex *d_iter_representatives,
                    outgoing_edge_id *d_outgoing_edges, vertex &num_nodes,
                    vertex &num_edges, vertex &number_of_filled_edges,
                    hipStream_t kernel_stream = 0);

/* Frees the workspace device_boruvka cached on the current device (no
 * counterpart in the reference, whose device_boruvka works in the caller's
 * buffers, MST/src/kernels.cu:507-518).  The next device_boruvka allocates
 * it again. */
extern "C" void phem_dropin_release_workspace(void);

/* The reference's full parameter list (its four thrust::device_ptr views of
 * the buffers above are ignored). */
void device_boruvka(edge *d_edges, edge *d_final_msf, edge *d_iter_edges,
                    edge *d_msf, vertex *d_representatives,
                    vertex *d_new_representatives,
                    vertex *d_iter_representatives,
                    outgoing_edge_id *d_outgoing_edges, vertex &num_nodes,
                    vertex &num_edges, vertex &number_of_filled_edges,
                    thrust::device_ptr<vertex> d_new_representatives_ptr,
                    thrust::device_ptr<vertex> d_representatives_ptr,
                    thrust::device_ptr<vertex> d_iter_representatives_ptr,
                    thrust::device_ptr<edge> d_final_msf_ptr,
                    hipStream_t kernel_stream = 0);

/* ---- CC types (CC/include/hbcg_utils.cuh:4-42) -------------------------- */
struct graph_data_host; /* the caller's (ParlayLib sequences) */
struct graph_data {
  uint64_t *edges;
  uint64_t *edges2;
  int *temp_label;
  int *label;
  int *V;
  long *E;
  long *size;
  long *size2;
  uint64_t *T1edges;
  uint64_t *T2edges;
  int *first_occ;
  int *last_occ;
  int *parent;
  int *w1;
  int *w2;
  int *w1_copy;
  int *w2_copy;
  uint64_t *SpanningForest;
  int *low;
  int *high;
  bool *fg1;
};

/* Batch size and GPU fraction of the edges (CC/ExternalSpanningTree.cu:7-8);
 * defined by the caller's main program (the library holds weak defaults:
 * 2^20 and 1.0). */
extern int Batch_Size;
extern float GPU_share;

/* The caller's CPU union-find (CC/host_spanning_tree.cu:89-92), found by
 * symbol when the caller's program defines it. */
int host_spanning_tree(struct graph_data_host *);

/* Registers the CPU union-find explicitly (no counterpart in the reference;
 * for callers that load this library and their CPU path as separate shared
 * objects, where the symbol above is not visible to the library).  NULL
 * restores the symbol lookup. */
extern "C" void phem_set_host_spanning_tree(
    int (*fn)(struct graph_data_host *));

/* Batched union-find over the GPU share of the edges, double-buffered on two
 * streams, overlapped with host_spanning_tree on the CPU share
 * (CC/ExternalSpanningTree.cu:133-218).  h_input->edges: host records
 * (pinned for overlap); d_input: caller-allocated device buffers (edges and
 * edges2 of Batch_Size records; temp_label = parent, label, T1edges of V).
 * Returns the elapsed milliseconds.  Unhooked T1edges slots hold all ones. */
float SpanningTree(struct graph_data *h_input, struct graph_data *d_input,
                   struct graph_data_host *h_input_host);

/* Merges the GPU forest d_input->T1edges with the vert candidate edges in
 * extraedges (the CPU forest; all ones or INT_MAX mark empty slots) into
 * d_input->T2edges, leaving labels (minimum vertex ids) in d_input->label
 * (CC/GPUSpanningTree.cu:152-198).  Returns the elapsed milliseconds. */
float gpu_spanning_tree(struct graph_data *d_input, uint64_t *extraedges,
                        int vert);

#endif /* PHEM_DROPIN_H */
