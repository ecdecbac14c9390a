/*
 * phem.h -- C ABI of the MI355X (gfx950) implementation of PHEM's graph hot
 * path: Boruvka minimum spanning forest and union-find connected components /
 * spanning forest, single-graph, streaming and batched.
 *
 * Conventions
 *  - Plain pointers and sizes only.  Pointers named d_* are device (HBM)
 *    pointers; h_* are host pointers.  `stream` is a hipStream_t (NULL = the
 *    null stream).
 *  - Every entry point returns PHEM_OK (0) or a negative PHEM_E* code; none
 *    throws or exits.  phem_last_error() gives a message for the last failure
 *    on the calling thread.
 *  - An MST edge i is (u[i], v[i], w[i]), its id is i.  Edges with u == v or
 *    with an endpoint >= n are not part of the graph and are ignored.
 *    Duplicate edges are kept.
 *  - Weights are 32-bit patterns compared through an order key:
 *    PHEM_W_U32 = unsigned integers (the reference's `typedef uint32_t
 *    weight`, MST/src/constants.h:9); PHEM_W_F32 = IEEE float32 in the
 *    IEEE total order.  Edges are ordered by (weight, id): this makes the
 *    minimum spanning forest unique and equal to the reference's
 *    (MST/src/utils.cu:801-802; the GPU path's (weight, position) order with
 *    stable compaction, MST/src/kernels.cu:24-70,641-646, is the same order).
 *  - Component labels are the minimum vertex id of each component (the
 *    reference hooks larger root -> smaller root,
 *    CC/ExternalSpanningTree.cu:52-53, CC/host_spanning_tree.cu:31-38).
 *  - Edge ids are 32-bit: m <= PHEM_MAX_EDGES.  n <= PHEM_MAX_VERTICES.
 */
#ifndef PHEM_H
#define PHEM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PHEM_OK 0
#define PHEM_EINVAL (-1)    /* bad argument */
#define PHEM_ENOMEM (-2)    /* workspace too small / allocation failed */
#define PHEM_EHIP (-3)      /* HIP runtime error */
#define PHEM_ENOCONV (-4)   /* round cap reached */
#define PHEM_ENODEV (-5)    /* no usable gfx950 device */
#define PHEM_EBUSY (-6)     /* admission refused (serving) */
#define PHEM_ENOTFOUND (-7) /* unknown request / handle */

#define PHEM_W_U32 0
#define PHEM_W_F32 1

#define PHEM_MAX_EDGES 0xFFFFFFFEull
#define PHEM_MAX_VERTICES 0xFFFFFFFEu
#define PHEM_NO_EDGE 0xFFFFFFFFFFFFFFFFull /* empty spanning-forest slot */
#define PHEM_MAX_ROUNDS 64

/* Process-wide tunables (the autotuner's knobs).  Built-in defaults, the
 * environment variables PHEM_BLOCKS_PER_CU, PHEM_MST_LIGHT_FACTOR and
 * PHEM_CC_SAMPLE_FACTOR at first use, then phem_set_tuning.  Set them between
 * solves, not while one runs.  Results do not depend on them: the forest,
 * labels and weights are the same for every setting.  The kernel shapes
 * select among kernels compiled for each allowed value (the reference fixes
 * its block size at compile time, BLOCKSIZE 1024, MST/src/utils.cu:12, and
 * tb_size 1024, CC/ExternalSpanningTree.cu:18); workspace sizes do not
 * depend on them. */
typedef struct {
  uint32_t blocks_per_cu;  /* resident 256-thread blocks per CU of the looping
                              kernels (1..64; 0 in phem_set_tuning: unchanged;
                              512-thread blocks count twice) */
  uint32_t reserved;
  double mst_light_factor; /* default light_factor (NaN: unchanged) */
  double cc_sample_factor; /* CC sampling phase: the first f*n records,
                              then a second giant after 3*f*n records
                              (default 1.0; <= 0: one phase; NaN: unchanged) */
  /* kernel shapes; 0 (NaN for mst_split_ratio) in phem_set_tuning: unchanged */
  uint32_t mst_block;       /* threads per block of the MST input passes
                               (light and heavy filter): 256, 512 (default) */
  uint32_t mst_ept;         /* their edges per thread: 4 (default), 8, 16;
                               mst_block x mst_ept <= 4096; graphs of fewer
                               than 2^22 edges use at most 2048 edges per
                               tile (per thread: 2048 / mst_block) */
  uint32_t mst_ept_compact; /* edges per thread of the compaction rounds of
                               graphs with m >= 2^mst_compact_log2: 8, 16
                               (default; smaller graphs use 8) */
  uint32_t mst_compact_log2; /* that edge-count threshold (1..40, default 22;
                                below 22 it acts as 22) */
  double mst_split_ratio;   /* the light/heavy split runs when
                               m > ratio * light_factor * n (default 1.5;
                               > 1) */
  uint32_t cc_block;        /* threads per block of the unfiltered CC union:
                               256 (default), 512 */
  uint32_t cc_ept;          /* records per thread of the unfiltered union:
                               8, 16 (default), 32 */
  uint32_t cc_ept_filter;   /* records per thread of the filtered unions:
                               4, 8 (default), 16 */
  uint32_t cc_block_filter; /* threads per block of the filtered unions, whose
                               LDS summary of the giant bitmap grows with the
                               block (16 KB per 256 threads): 256, 512, 1024
                               (default; cc_ept_filter 4 or 8) */
  double cpu_share;         /* phem_msf_host_chunks with PHEM_CPU_EDGES_AUTO:
                               the host reduces cpu_share x (edges before
                               the last chunk) edges of the last chunk
                               (>= 0; default 0: on MI355X every host share
                               measured slowed the solve; NaN: unchanged) */
  double mst_light_factor_skewed; /* light factor of solves that leave it to
                               the library when their threshold sample
                               finds degree skew (hubs repeat as sampled
                               endpoints; default 0.625; 0: the one
                               default factor for every graph; NaN:
                               unchanged) */
} phem_tuning;
int phem_get_tuning(phem_tuning *t);
int phem_set_tuning(const phem_tuning *t);

/* Library / device information. */
const char *phem_version(void);
const char *phem_last_error(void);
/* 0 if a gfx950 device is visible and the kernels can launch on it. */
int phem_device_check(int *device_count);

/* Per-kernel timing with HIP events recorded around every launch on the
 * launch's own stream (no counterpart in the reference, which times whole
 * phases with cudaEvents, CC/ExternalSpanningTree.cu:147-151).  Off by
 * default.  phem_prof_count() waits for the recorded events and returns the
 * number of distinct kernels seen since the last reset. */
int phem_prof_enable(int on);
/* Records only the launches of the kernel called `name` (as reported by
 * phem_prof_get); NULL or "" records every kernel.  Keeps the timing of one
 * kernel from adding event records around all the others. */
int phem_prof_only(const char *name);
int phem_prof_reset(void);
int phem_prof_count(void);
int phem_prof_get(int i, char *name, int name_cap, uint64_t *launches,
                  double *total_ms);

/* ---------------------------------------------------------------------------
 * Boruvka minimum spanning forest (one graph)
 *
 * Replaces the plain GPU path boruvka_for_edge_array / device_boruvka
 * (MST/src/new_boruvka.cu:40-179, MST/src/kernels.cu:507-689).
 * ------------------------------------------------------------------------- */

typedef struct phem_mst_stats {
  uint32_t rounds;        /* Boruvka rounds executed (all phases) */
  uint32_t mst_count;     /* edges in the forest */
  uint64_t weight_fx;     /* exact total weight: U32 sum of w; F32 sum of w*2^24 */
  int32_t weight_exact;   /* F32: 0 if a weight is negative or off the 2^-24 grid */
  uint32_t light_rounds;  /* rounds of the light phase (all rounds: no split) */
  uint64_t sum_m;         /* sum over rounds of alive edges */
  uint64_t sum_c;         /* sum over rounds of active components */
  uint64_t valid_edges;   /* input edges that are part of the graph */
  uint64_t tau;           /* light/heavy key threshold (all ones: no split) */
  uint64_t light_edges;   /* edges of the light phase */
  uint64_t heavy_edges;   /* alive edges entering the heavy rounds */
  uint64_t giant_edges;   /* of those: one per component adjacent to the giant
                             component (its lightest such edge), for the
                             components not hooked to the giant at once
                             because a lighter edge or a smaller id decides */
  uint64_t m_log[PHEM_MAX_ROUNDS]; /* alive edges at the start of round r */
  uint64_t c_log[PHEM_MAX_ROUNDS]; /* components with >= 1 alive edge in round r */
} phem_mst_stats;

/* Solver options (NULL = defaults).
 * light_factor: the solve first runs Boruvka rounds over the ~light_factor*n
 *   lightest edges (threshold from a sample of the keys), then drops every
 *   heavier edge whose endpoints are already connected and runs rounds over
 *   the rest (filter-Kruskal, as the CPU path does, MST/src/utils.cu:794-841).
 *   This split runs when m > 1.5 * light_factor * n; otherwise, and for
 *   <= 0, the solve runs the canonical Boruvka round loop over all edges
 *   (MST/src/kernels.cu:507-689), whose per-round log is the reference's.
 *   NaN, like NULL options: the mst_light_factor knob (phem_set_tuning;
 *   1.0 unless set or given by the environment variable
 *   PHEM_MST_LIGHT_FACTOR), or the mst_light_factor_skewed knob (0.625)
 *   when the key sample's endpoints repeat far more often than in a
 *   uniform graph (hubs).  Whether the split runs follows from the first.
 * flags: PHEM_MST_MAXIMUM or 0. */
typedef struct phem_mst_opts {
  double light_factor;
  uint32_t flags; /* PHEM_MST_* bits */
  uint32_t reserved;
} phem_mst_opts;

/* Maximum spanning forest: edges ordered by (weight, id) descending, the
 * order of the reference's populate_max_weight / populate_eid_max
 * (MST/src/kernels.cu:437-502: atomicMax on the weight, then on the id). */
#define PHEM_MST_MAXIMUM 1u

/* Bytes of device workspace phem_mst_solve needs for (n, m). */
int phem_mst_workspace_bytes(uint32_t n, uint64_t m, uint64_t *bytes);

/* Minimum spanning forest of (d_u, d_v, d_w).  Writes the forest's edge ids in
 * ascending order to d_mst_ids (capacity >= n), and min-id labels to d_labels
 * (n entries; may be NULL).  d_ws: device workspace of >= the queried bytes.
 * Blocks the calling thread until the result is in HBM (the round loop reads
 * one counter per round); stats (host, may be NULL) receives the round log. */
int phem_mst_solve(uint32_t n, uint64_t m, const uint32_t *d_u,
                   const uint32_t *d_v, const uint32_t *d_w, int wmode,
                   void *d_ws, uint64_t ws_bytes, uint32_t *d_mst_ids,
                   uint32_t *d_labels, const phem_mst_opts *opts,
                   phem_mst_stats *stats, void *stream);

/* Same as phem_mst_solve over the reference's array-of-records edge layout:
 * d_edges holds m records {u, v, w} of three 32-bit words
 * (struct edge, MST/src/constants.h:15-19).  Same workspace size. */
int phem_mst_solve_aos(uint32_t n, uint64_t m, const uint32_t *d_edges,
                       int wmode, void *d_ws, uint64_t ws_bytes,
                       uint32_t *d_mst_ids, uint32_t *d_labels,
                       const phem_mst_opts *opts, phem_mst_stats *stats,
                       void *stream);

/* phem_mst_solve with no allocation of any kind (no counterpart in the
 * reference: for callers that cannot allocate, such as code under HIP graph
 * capture, tight serving loops and harnesses that forbid hipMalloc).  Same
 * arguments and results as phem_mst_solve, without d_ws / ws_bytes: the
 * workspace is a static device array of 64 MiB (67108864 B) per device,
 * held from the first call on that device until the process ends.  Graphs
 * whose phem_mst_workspace_bytes exceeds it (for instance n = 2^17,
 * m = 2^21; n = 2^17, m = 2^20 fits) return PHEM_ENOMEM.  Every call on one
 * device uses the same array: calls that overlap on one device (other
 * streams or threads) race and must be serialised by the caller. */
int phem_mst_solve_noalloc(uint32_t n, uint64_t m, const uint32_t *d_u,
                           const uint32_t *d_v, const uint32_t *d_w, int wmode,
                           uint32_t *d_mst_ids, uint32_t *d_labels,
                           const phem_mst_opts *opts, phem_mst_stats *stats,
                           void *stream);

/* A whole phem_mst_solve recorded once into a HIP graph and replayed.
 * Replaces the host-driven round loop of device_boruvka
 * (MST/src/kernels.cu:526-689: a count_if / remove_if and a host read of
 * num_edges every round, :577-649) by a fixed number of rounds per phase
 * that end on the device: a round that starts with no alive edge does
 * nothing, and no counter is read until phem_mst_graph_stats.
 * phem_mst_graph_create records the solve of (d_u, d_v, d_w) with these
 * arguments (the same as phem_mst_solve's, the pointers fixed); a replay
 * solves whatever those buffers hold at launch time (n and m fixed).
 * rounds: Boruvka rounds per phase (the canonical rounds after round 0; with
 *   the light/heavy split, the light phase and the heavy phase each);
 *   0 = one eager solve of the current inputs picks them (its rounds per
 *   phase + 1) and writes its result to the outputs.
 * phem_mst_graph_launch enqueues one replay on `stream` and returns.
 * phem_mst_graph_stats waits for the last replay on `stream`, fills stats
 *   (may be NULL) as phem_mst_solve would, and returns PHEM_ENOCONV when a
 *   phase still had alive edges after its recorded rounds: the outputs are
 *   then incomplete (record again with more rounds, or call phem_mst_solve).
 * The graph holds no device memory of the solve: d_ws and the outputs stay
 * the caller's and must outlive it.  Replays on one workspace must not
 * overlap. */
typedef struct phem_mst_graph phem_mst_graph_t;
int phem_mst_graph_create(uint32_t n, uint64_t m, const uint32_t *d_u,
                          const uint32_t *d_v, const uint32_t *d_w, int wmode,
                          void *d_ws, uint64_t ws_bytes, uint32_t *d_mst_ids,
                          uint32_t *d_labels, const phem_mst_opts *opts,
                          int rounds, void *stream, phem_mst_graph_t **out);
int phem_mst_graph_launch(phem_mst_graph_t *g, void *stream);
int phem_mst_graph_stats(phem_mst_graph_t *g, phem_mst_stats *stats,
                         void *stream);
int phem_mst_graph_destroy(phem_mst_graph_t *g);

/* d_out[k] = record d_ids[k] of the 3-word {u, v, w} records in d_src, for
 * k < count: materialises forest edges as the reference's `edge` records
 * (the MSF copy-out of device_boruvka, MST/src/kernels.cu:590-596). */
int phem_gather_edges_aos(const uint32_t *d_src, const uint32_t *d_ids,
                          uint64_t count, uint32_t *d_out, void *stream);

/* ---------------------------------------------------------------------------
 * Union-find connected components / spanning forest
 *
 * Edges are packed uint64 (u << 32) | v (CC/graph.cu:74-77).  As in the
 * reference kernel (CC/ExternalSpanningTree.cu:77-86) only records with u < v
 * are processed; records with v >= n are ignored.  The forest holds, for each
 * vertex that was hooked as a root, the packed edge that hooked it
 * (T1edges, CC/ExternalSpanningTree.cu:52-55), else PHEM_NO_EDGE; which edge is
 * chosen depends on the schedule, as in the reference.
 * ------------------------------------------------------------------------- */

/* parent[i] = i, forest[i] = PHEM_NO_EDGE (init_parent_label_T1edges,
 * CC/ExternalSpanningTree.cu:108-116).  d_forest may be NULL. */
int phem_cc_init(uint32_t n, uint32_t *d_parent, uint64_t *d_forest,
                 void *stream);
/* Union all m edges into the parent forest (union_find_gpu_COO,
 * CC/ExternalSpanningTree.cu:77-86).  Asynchronous on `stream`.  The parent
 * array stays valid between calls, so batches can be applied one after the
 * other (CC/ExternalSpanningTree.cu:170-204). */
int phem_cc_union(uint32_t n, const uint64_t *d_edges, uint64_t m,
                  uint32_t *d_parent, uint64_t *d_forest, void *stream);
/* labels[i] = root of i = minimum vertex id of i's component (`cc`,
 * CC/GPUSpanningTree.cu:122-128).  Asynchronous. */
int phem_cc_labels(uint32_t n, const uint32_t *d_parent, uint32_t *d_labels,
                   void *stream);
/* init + union + labels.  Asynchronous. */
int phem_cc_solve(uint32_t n, const uint64_t *d_edges, uint64_t m,
                  uint32_t *d_parent, uint64_t *d_forest, uint32_t *d_labels,
                  void *stream);
/* Merge the device forest d_forest (n slots) with n extra candidate edges
 * (the CPU-side forest) into a second forest d_forest2 and recompute parent
 * and labels (gpu_spanning_tree, CC/GPUSpanningTree.cu:74-96,152-158,169-198).
 * Asynchronous. */
int phem_cc_merge_forests(uint32_t n, const uint64_t *d_forest,
                          const uint64_t *d_extra, uint32_t *d_parent,
                          uint64_t *d_forest2, uint32_t *d_labels,
                          void *stream);

/* ---------------------------------------------------------------------------
 * Streaming spanning forest / connected components with device-resident
 * state (the reference's chunked MSF, MSF(chunk u previous MSF),
 * MST/src/kernels.cu:942-1102, and batch-by-batch union-find,
 * CC/ExternalSpanningTree.cu:170-204).  After each batch the forest is the
 * minimum spanning forest of every edge inserted so far, ordered by
 * (weight, global id) with global ids in insertion order; labels are the
 * minimum vertex id of each component.  The handle owns its device memory
 * (forest and parent array stay in HBM between batches).
 * ------------------------------------------------------------------------- */

typedef struct phem_stream phem_stream_t;

/* Device bytes phem_stream_create allocates for n vertices and batches of up
 * to max_batch edges (no counterpart in the reference; the server admits
 * resident streams by it). */
int phem_stream_bytes(uint32_t n, uint64_t max_batch, uint64_t *bytes);
/* Device memory for n vertices and batches of up to max_batch edges
 * (n + max_batch <= PHEM_MAX_EDGES), on the current device. */
int phem_stream_create(uint32_t n, uint64_t max_batch, int wmode,
                       phem_stream_t **out);
int phem_stream_destroy(phem_stream_t *s);
/* Inserts mb edges (device arrays); returns once forest and parent array are
 * updated.  On an error before the batch's union-find ran, the stream is as
 * before the call; on a later error (a HIP failure while the result is read)
 * the union-find state already holds the batch, so the stream refuses every
 * further phem_stream_apply and phem_stream_save with PHEM_EINVAL. */
int phem_stream_apply(phem_stream_t *s, const uint32_t *d_u,
                      const uint32_t *d_v, const uint32_t *d_w, uint64_t mb,
                      void *stream);
/* Device pointers to the current forest: *count edges (u, v, w, global id),
 * ascending global id.  Valid until the next phem_stream_apply. */
int phem_stream_forest(const phem_stream_t *s, uint64_t *count,
                       const uint32_t **d_u, const uint32_t **d_v,
                       const uint32_t **d_w, const uint64_t **d_gid);
/* Copies the current forest into caller buffers (NULL entries skipped). */
int phem_stream_copy_forest(const phem_stream_t *s, uint32_t *d_u,
                            uint32_t *d_v, uint32_t *d_w, uint64_t *d_gid,
                            void *stream);
/* Statistics of the last solve (its weight_fx is the forest's total weight),
 * edges inserted and batches applied so far. */
int phem_stream_stats(const phem_stream_t *s, phem_mst_stats *stats,
                      uint64_t *edges_inserted, uint64_t *batches);
/* Connected-component labels of every edge inserted so far. */
int phem_stream_labels(const phem_stream_t *s, uint32_t *d_labels,
                       void *stream);
/* Union-find spanning forest (packed hooking records, n slots). */
int phem_stream_cc_forest(const phem_stream_t *s, const uint64_t **d_forest);
/* Copies those n slots to d_out (the T1edges array that the reference's
 * SpanningTree fills, CC/ExternalSpanningTree.cu:46-59). */
int phem_stream_copy_cc_forest(const phem_stream_t *s, uint64_t *d_out,
                               void *stream);
/* Host snapshot of a stream, for checkpoint and resume.  No counterpart in
 * the reference, which keeps no state between runs (its only persistence is
 * the weighted graph dump of new_boruvka, MST/src/new_boruvka.cu:399-410).
 * phem_stream_save writes the forest, the union-find state, the insertion
 * count and the last statistics into `host` (any host memory of at least
 * phem_stream_snapshot_bytes bytes; the size grows with the forest) and
 * returns when the copy is complete.  phem_stream_restore creates a stream
 * on the current device with that state and room for batches of max_batch
 * edges; later batches then give exactly what the saved stream would have
 * given.  A buffer that is not a snapshot, is shorter than its header says,
 * or holds a value the kernels could not index with, returns PHEM_EINVAL.
 * The format is this library's in-memory layout (host byte order, the stats
 * struct as compiled): restore with the same build that saved. */
int phem_stream_snapshot_bytes(const phem_stream_t *s, uint64_t *bytes);
int phem_stream_save(const phem_stream_t *s, void *host, uint64_t bytes,
                     void *stream);
int phem_stream_restore(const void *host, uint64_t bytes, uint64_t max_batch,
                        phem_stream_t **out);

/* ---------------------------------------------------------------------------
 * Batched solve: k independent graphs in one pass (one disjoint union).
 * Graph g has h_voff[g+1] - h_voff[g] vertices and edges
 * [h_eoff[g], h_eoff[g+1]) of the concatenated arrays d_u, d_v, d_w, with
 * vertex ids local to the graph (0-based).  Results are per graph and local:
 * forest edge ids (local, ascending) at d_ids[h_voff[g] ...] (d_counts[g] of
 * them), min-id labels at d_labels[h_voff[g] ...] (may be NULL), exact
 * forest weights d_weights[g] with d_inexact[g] != 0 where a float weight was
 * off the 2^-24 grid (both may be NULL).  h_voff / h_eoff: k + 1 host
 * offsets starting at 0.  Blocks until done, like phem_mst_solve.
 * ------------------------------------------------------------------------- */
int phem_batch_workspace_bytes(uint64_t n_total, uint64_t m_total, uint32_t k,
                               uint64_t *bytes);
int phem_batch_solve(uint32_t k, const uint64_t *h_voff,
                     const uint64_t *h_eoff, const uint32_t *d_u,
                     const uint32_t *d_v, const uint32_t *d_w, int wmode,
                     void *d_ws, uint64_t ws_bytes, uint32_t *d_ids,
                     uint32_t *d_counts, uint32_t *d_labels,
                     uint64_t *d_weights, uint64_t *d_inexact,
                     const phem_mst_opts *opts, phem_mst_stats *stats,
                     void *stream);
/* The same with graph g's edges in its own device arrays h_pu[g], h_pv[g],
 * h_pw[g] (host tables of k device pointers; h_eoff gives the edge counts). */
int phem_batch_solve_ptrs(uint32_t k, const uint64_t *h_voff,
                          const uint64_t *h_eoff, const uint32_t *const *h_pu,
                          const uint32_t *const *h_pv,
                          const uint32_t *const *h_pw, int wmode, void *d_ws,
                          uint64_t ws_bytes, uint32_t *d_ids,
                          uint32_t *d_counts, uint32_t *d_labels,
                          uint64_t *d_weights, uint64_t *d_inexact,
                          const phem_mst_opts *opts, phem_mst_stats *stats,
                          void *stream);

/* A batched solve recorded once into a HIP graph and replayed (the batched
 * counterpart of phem_mst_graph_*; no counterpart in the reference, whose
 * round loop is host-driven, MST/src/kernels.cu:526-689).
 * phem_batch_graph_create records a batch of up to k_cap - 1 graphs with at
 * most n_cap vertices and m_cap edges in total, edges given per graph as in
 * phem_batch_solve_ptrs, with `rounds` Boruvka rounds per phase (0: 12) and
 * no counter read.  d_ws (>= phem_batch_workspace_bytes(n_cap, m_cap,
 * k_cap)) and the outputs (d_ids, d_labels: n_cap slots; d_counts,
 * d_weights, d_inexact: k_cap) stay the caller's and must outlive the graph.
 * phem_batch_graph_launch copies a batch's offsets and pointer tables (as
 * for phem_batch_solve_ptrs: k <= k_cap - 1 graphs that fit the
 * capacities) and enqueues one replay on `stream`; the batch is padded with
 * one graph of no valid edge that takes the remaining vertices and edges.
 * Graph g's results are where phem_batch_solve_ptrs puts them.  Launches
 * must not overlap (replays share the workspace); the edge arrays must stay
 * valid until the replay is done.  phem_batch_graph_stats waits for the last
 * replay, fills stats for the whole padded solve, and returns PHEM_ENOCONV
 * when a phase needed more rounds than recorded: the results are then
 * incomplete and the batch must be solved another way (phem_batch_solve). */
typedef struct phem_batch_graph phem_batch_graph_t;
int phem_batch_graph_create(uint32_t k_cap, uint64_t n_cap, uint64_t m_cap,
                            int wmode, void *d_ws, uint64_t ws_bytes,
                            uint32_t *d_ids, uint32_t *d_counts,
                            uint32_t *d_labels, uint64_t *d_weights,
                            uint64_t *d_inexact, const phem_mst_opts *opts,
                            int rounds, phem_batch_graph_t **out);
int phem_batch_graph_launch(phem_batch_graph_t *g, uint32_t k,
                            const uint64_t *h_voff, const uint64_t *h_eoff,
                            const uint32_t *const *h_pu,
                            const uint32_t *const *h_pv,
                            const uint32_t *const *h_pw, void *stream);
int phem_batch_graph_stats(phem_batch_graph_t *g, phem_mst_stats *stats,
                           void *stream);
int phem_batch_graph_destroy(phem_batch_graph_t *g);

/* ---------------------------------------------------------------------------
 * Serving: a request loop for MST + CC queries on independent graphs.
 * No counterpart in the reference (its drivers solve one graph per process
 * run, MST/main.cu:5-30, CC/main.cu); the batched solve above does the work.
 *
 * Requests are queued as they arrive.  Each of `slots` workers (one host
 * thread and one HIP stream each, its own device memory of mem_bytes / slots)
 * takes the queued requests in arrival order while their combined device
 * footprint fits its memory and there are at most max_graphs of them, and
 * runs them as one batch; a request that arrives while batches are running
 * joins the next batch a worker forms, so work from several batches overlaps
 * on the device.  Results go to the request's own host buffers.
 * A request larger than one slot can hold is refused with PHEM_EBUSY.
 * The server also keeps resident streams (phem_server_stream_open): their
 * forest, parent array and union-find state stay in HBM between the
 * streaming edge batches that clients submit to them (PHEM_REQ_STREAM
 * requests), which queue with the graph requests and run on the same
 * workers, each batch on its own.
 * Times are seconds on phem_server_clock() (a monotonic clock).
 * ------------------------------------------------------------------------- */
typedef struct phem_server phem_server;

typedef struct {
  uint64_t mem_bytes;       /* device memory for all slots; 0: 1/4 of free */
  uint64_t max_batch_edges; /* edges per batch; 0: 2^27 */
  uint32_t slots;           /* batches in flight; 0: 2 */
  uint32_t max_graphs;      /* graphs per batch; 0: 4096 */
  uint32_t window_us;       /* an idle worker that finds fewer than max_graphs
                               requests waits up to this long for more */
  int32_t wmode;            /* PHEM_W_U32 or PHEM_W_F32 for every request */
  uint32_t record_rounds;   /* > 0: batches replay a batched solve recorded
                               (phem_batch_graph_*) for the batch's capacity
                               bucket: a bucket's first batch runs eagerly,
                               later ones replay a recording with the rounds
                               per phase it needed plus two, at most this
                               many; a batch that needs more falls back to
                               the eager solve (and the bucket is recorded
                               again with more rounds).  0: every batch is
                               solved eagerly */
  uint32_t stream_mem_mib;  /* device memory of the resident streams, MiB
                               (0: as much as mem_bytes) */
} phem_server_opts;

#define PHEM_REQ_DEVICE 1u /* u, v, w are device pointers */
#define PHEM_REQ_STREAM 2u /* a batch of m edges (the stream's vertex ids) for
                              resident stream stream_id; n is ignored and
                              forest_ids must be NULL */

typedef struct {
  uint32_t n;
  uint32_t flags;
  uint64_t m;
  const uint32_t *u, *v, *w; /* edges with local ids; valid until done */
  uint32_t *forest_ids;      /* host, n slots: forest edge ids ascending
                                (NULL: not returned) */
  uint32_t *labels;          /* host, n slots: min-id labels (or NULL) */
  double arrival;            /* < 0: the time of submission */
  uint64_t stream_id;        /* PHEM_REQ_STREAM: the resident stream */
} phem_request;

typedef struct {
  int32_t status;        /* PHEM_OK or the error of its batch */
  uint32_t forest_count; /* edges in forest_ids (a stream batch: edges of the
                            stream's forest after the batch) */
  uint64_t weight_fx;    /* exact forest weight (x 2^24 for float weights) */
  int32_t weight_exact;
  uint32_t batch_graphs; /* graphs in the batch the request ran in */
  double arrival, start, done;
  char error[160];       /* the batch's error message ("" when status is 0);
                            phem_last_error() of the client thread does not
                            hold it, the batch ran on a worker thread */
} phem_result;

typedef struct {
  uint64_t submitted, completed, refused, batches, edges;
  double first_arrival, last_done;          /* of completed requests */
  double lat_p50, lat_p99, lat_max, lat_mean; /* done - arrival, seconds */
  double mean_batch_graphs;
  uint64_t slot_bytes, slot_edges;          /* capacity of one slot */
  uint64_t recorded_batches;  /* batches that replayed a recorded solve */
  uint64_t record_fallbacks;  /* of those, solved again eagerly (more rounds
                                 needed than recorded) */
  uint64_t recordings;        /* recorded solves created (capture + instantiate) */
  /* stream batches (counted in completed, edges and the latencies above
   * too), their own latency percentiles, and the resident streams */
  uint64_t stream_batches, stream_edges;
  double stream_lat_p50, stream_lat_p99;
  uint64_t streams_open, stream_bytes; /* device bytes the open streams hold */
} phem_server_stats;

double phem_server_clock(void);
int phem_server_create(const phem_server_opts *opts, phem_server **out);
int phem_server_destroy(phem_server *srv);
int phem_server_submit(phem_server *srv, const phem_request *req,
                       uint64_t *ticket);
/* Blocks until the request is done, fills *res and forgets the ticket.  One
 * waiter per ticket: a second wait, or a poll, while one waits returns
 * PHEM_EBUSY. */
int phem_server_wait(phem_server *srv, uint64_t ticket, phem_result *res);
/* *done = 1 and *res filled (ticket forgotten) if done, else *done = 0
 * (PHEM_EBUSY while a phem_server_wait waits for the ticket). */
int phem_server_poll(phem_server *srv, uint64_t ticket, phem_result *res,
                     int *done);
/* Blocks until every submitted request is done. */
int phem_server_drain(phem_server *srv);
int phem_server_get_stats(phem_server *srv, phem_server_stats *st);
int phem_server_reset_stats(phem_server *srv);
/* A resident stream of n vertices taking batches of up to max_batch edges
 * (a phem_stream owned by the server, on the server's device).  Admitted by
 * its device footprint (phem_stream_bytes): beyond the stream memory still
 * free (opts.stream_mem_mib) it is refused with PHEM_EBUSY.  A stream's
 * batches are applied in submission order, whichever workers take them; a
 * batch's result holds the forest's edge count and exact weight after it,
 * and its labels (if requested) the stream's min-id labels after it.  No
 * counterpart in the reference, whose streams live in one process run
 * (MST/src/kernels.cu:942-1102, CC/ExternalSpanningTree.cu:170-204). */
int phem_server_stream_open(phem_server *srv, uint32_t n, uint64_t max_batch,
                            uint64_t *stream_id);
/* Frees the stream; PHEM_EBUSY while batches of it are queued or running,
 * PHEM_ENOTFOUND for an unknown id. */
int phem_server_stream_close(phem_server *srv, uint64_t stream_id);

/* ---------------------------------------------------------------------------
 * Host-resident edge lists streamed through the device in chunks.
 * ------------------------------------------------------------------------- */

typedef struct {
  double seconds;      /* the whole call */
  double cpu_seconds;  /* the host's share (0 without one) */
  uint64_t chunks;     /* chunks folded in on the device */
  uint64_t cpu_edges;  /* edges the host reduced */
  uint64_t cpu_kept;   /* edges of their forest */
  uint64_t reserved[3];
} phem_chunks_stats;

#define PHEM_CPU_EDGES_AUTO 0xFFFFFFFFFFFFFFFFull

/* Minimum spanning forest of the host edges (u[i], v[i], w[i]), i < m, in
 * chunks of chunk_edges folded one by one into a forest kept in HBM,
 * F = MSF(F u chunk) (the reference's streamline path, MST/src/kernels.cu:
 * 849-1154, whose chunks are the contiguous slices of the edge array).  The
 * next chunk is copied to the device while the current one is folded in.
 * cpu_edges: the first cpu_edges edges of the last chunk are reduced on the
 * host (phem_cpu_msf) while the device folds in the other chunks, and the
 * last chunk carries their forest instead (the heterogeneous streamline,
 * MST/src/kernels.cu:1155-1826, which gives the host at most 10^8 edges);
 * PHEM_CPU_EDGES_AUTO: cpu_share (phem_tuning) x the edges before the last
 * chunk; 0, or a single chunk: none.  ids_out (n host slots) receives the
 * forest's edge ids ascending, *count their number: the forest of the whole
 * list under the (weight, id) order whatever the chunking.  Runs on the
 * current device on streams of its own and returns when done.  stats may be
 * NULL. */
int phem_msf_host_chunks(uint32_t n, uint64_t m, const uint32_t *u,
                         const uint32_t *v, const uint32_t *w, int wmode,
                         uint64_t chunk_edges, uint64_t cpu_edges,
                         uint32_t *ids_out, uint64_t *count,
                         phem_chunks_stats *stats);

/* ---------------------------------------------------------------------------
 * Host backend: the same results computed on the CPU.
 * ------------------------------------------------------------------------- */

/* Minimum spanning forest on the host: Kruskal over (weight, id) with a
 * multithreaded stable radix sort.  The CPU pre-filter of the heterogeneous
 * chunked solve (MST/src/kernels.cu:1503-1788, where the reference calls its
 * ParlayLib mst() on the last chunk).  ids_out (n slots) receives the forest's
 * edge ids ascending, *count their number.  threads <= 0: all host threads. */
int phem_cpu_msf(uint32_t n, uint64_t m, const uint32_t *u, const uint32_t *v,
                 const uint32_t *w, int wmode, uint32_t *ids_out,
                 uint64_t *count, int threads);

/* Connected components on the host: lock-free union-find over host threads
 * (the same hook rule as the device path), labels = minimum vertex id.  Edges
 * with u == v or an endpoint >= n are ignored.  threads <= 0: all host
 * threads. */
int phem_cpu_cc(uint32_t n, uint64_t m, const uint32_t *u, const uint32_t *v,
                uint32_t *labels, int threads);

/* ---------------------------------------------------------------------------
 * Graph files (host memory).  The readers fill a phem_graph_file whose arrays
 * the library allocates; phem_graph_file_free releases them.
 * ------------------------------------------------------------------------- */
typedef struct {
  uint32_t n;           /* vertices */
  int32_t wmode;        /* order of w: PHEM_W_U32 or PHEM_W_F32 */
  uint64_t m;           /* edges (u, v, w forms) or records (edges form) */
  uint32_t *u, *v, *w;  /* MST forms: m edges, 0-based */
  uint64_t *edges;      /* CC forms: m records (u << 32) | v with u < v */
} phem_graph_file;

/* MatrixMarket coordinate file (MST/src/utils.cu:160-179, 896-1036): comment
 * lines, "rows cols entries", then "u v [w]" lines with 1-based ids; n = rows,
 * at most `entries` edges, stored 0-based with u <= v.  Integer weights are
 * read (a missing weight reads as 0), or with generate_weights set to
 * w = ((u-1)(v-1) mod n) + 1 in 32-bit unsigned arithmetic (utils.cu:1002).
 * Fills u, v, w (PHEM_W_U32). */
int phem_read_mtx(const char *path, int generate_weights, phem_graph_file *g);
/* HAST edge list (.edges / .eg2 / .txt, CC/graph.cu:45-90): "V E" then E
 * directed pairs "u v"; keeps the pairs with u < v.  Fills edges. */
int phem_read_edge_list(const char *path, phem_graph_file *g);
/* ECL binary CSR (.egr, CC/graph.cu:92-121, CC/README.md:56-66): int64 V,
 * int64 E, int64 offsets[V+1], int32 neighbours[E]; keeps u < neighbour.
 * Fills edges. */
int phem_read_ecl(const char *path, phem_graph_file *g);
/* PBBS weighted edge array (MST/src/common/graphIO.h:169-188): the word
 * "WeightedEdgeArray" then triples "u v w", w read as a double and stored as
 * float32; n = largest id + 1.  Fills u, v, w (PHEM_W_F32). */
int phem_read_pbbs_weighted(const char *path, phem_graph_file *g);
/* Writes u, v, w as the reference's weighted MatrixMarket dump
 * (MST/src/new_boruvka.cu:399-408): both directions, 1-based. */
int phem_write_mtx_weighted(const char *path, const phem_graph_file *g);
void phem_graph_file_free(phem_graph_file *g);

/* ---------------------------------------------------------------------------
 * Synthetic inputs on the device (counter-based splitmix64 streams; the same
 * bytes as the host generator on every host).  Edges are stored with u < v.
 * ------------------------------------------------------------------------- */

/* m uniform edges over [0, n); ids id0 .. id0+m-1 index the streams.
 * positive_w: weights in (0,1] instead of [0,1). */
int phem_gen_uniform(uint32_t n, uint64_t m, uint64_t seed, uint64_t id0,
                     int positive_w, uint32_t *d_u, uint32_t *d_v,
                     uint32_t *d_w, void *stream);
/* m RMAT edges over 2^scale vertices with cumulative 32-bit thresholds
 * t1 = a*2^32, t2 = (a+b)*2^32, t3 = (a+b+c)*2^32. */
int phem_gen_rmat(uint32_t scale, uint64_t m, uint64_t seed, uint32_t t1,
                  uint32_t t2, uint32_t t3, int positive_w, uint32_t *d_u,
                  uint32_t *d_v, uint32_t *d_w, void *stream);
/* Pack (u, v) into (u << 32) | v records. */
int phem_pack_edges(const uint32_t *d_u, const uint32_t *d_v, uint64_t m,
                    uint64_t *d_edges, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* PHEM_H */
