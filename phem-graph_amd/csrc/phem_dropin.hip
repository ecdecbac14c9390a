// phem_dropin.hip -- the reference's C++ entry points over the C ABI.
#include <chrono>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <mutex>
#include <algorithm>
#include <string>
#include <thread>
#include <vector>

#include "phem.h"
#include "phem_dev.h"
#include "phem_dropin.h"
#include "phem_internal.h"

namespace phem {

constexpr int DBLK = 256;

// out[k] = record ids[k] of the 3-word records in src
__global__ __launch_bounds__(DBLK) void k_gather_aos(
    const uint32_t *__restrict__ src, const uint32_t *__restrict__ ids,
    uint64_t count, uint32_t *__restrict__ out) {
  for (uint64_t k = (uint64_t)blockIdx.x * DBLK + threadIdx.x; k < count;
       k += (uint64_t)gridDim.x * DBLK) {
    const uint64_t e = ids[k];
    out[3 * k] = src[3 * e];
    out[3 * k + 1] = src[3 * e + 1];
    out[3 * k + 2] = src[3 * e + 2];
  }
}

namespace {

// The reference's error convention for these entry points
// (CC/include/cuda_utility.cuh:7-17): report and exit.
void die_hip(hipError_t e, const char *what) {
  if (e == hipSuccess) return;
  std::fprintf(stderr, "HIP error in %s: %s\n", what, hipGetErrorString(e));
  std::exit(EXIT_FAILURE);
}
void die_phem(int rc, const char *what) {
  if (rc == PHEM_OK) return;
  std::fprintf(stderr, "%s failed (%d): %s\n", what, rc, phem_last_error());
  std::exit(EXIT_FAILURE);
}

// Grow-only solver workspace per device, kept for the life of the process.
struct CachedWs {
  void *p = nullptr;
  size_t cap = 0;
};
std::mutex g_ws_mu;
CachedWs g_ws[64];

void *cached_workspace(size_t bytes) {
  int dev = 0;
  die_hip(hipGetDevice(&dev), "hipGetDevice");
  if (dev < 0 || dev >= 64) dev = 0;
  std::lock_guard<std::mutex> lk(g_ws_mu);
  CachedWs &w = g_ws[dev];
  if (w.cap < bytes) {
    if (w.p) die_hip(hipFree(w.p), "hipFree");
    w.p = nullptr;
    die_hip(hipMalloc(&w.p, bytes), "hipMalloc (MST workspace)");
    w.cap = bytes;
  }
  return w.p;
}

void release_workspace() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) {
    (void)hipGetLastError();
    return;
  }
  if (dev < 0 || dev >= 64) dev = 0;
  std::lock_guard<std::mutex> lk(g_ws_mu);
  CachedWs &w = g_ws[dev];
  if (w.p) (void)hipFree(w.p);
  w.p = nullptr;
  w.cap = 0;
}

}  // namespace
}  // namespace phem

using namespace phem;

extern "C" void phem_dropin_release_workspace(void) { release_workspace(); }

extern "C" int phem_gather_edges_aos(const uint32_t *d_src, const uint32_t *d_ids,
                                     uint64_t count, uint32_t *d_out,
                                     void *stream) {
  if (count && (!d_src || !d_ids || !d_out))
    return set_err(PHEM_EINVAL, "NULL array");
  if (!count) return PHEM_OK;
  {
    PHEM_PROF("k_gather_aos", (hipStream_t)stream);
    hipLaunchKernelGGL(k_gather_aos, dim3(grid_for(count, DBLK)), dim3(DBLK),
                       0, (hipStream_t)stream, d_src, d_ids, count, d_out);
  }
  PHEM_LAUNCHED("k_gather_aos");
  return PHEM_OK;
}

// ---- MST --------------------------------------------------------------------

void device_boruvka(edge *d_edges, edge *d_final_msf, edge * /*d_iter_edges*/,
                    edge * /*d_msf*/, vertex *d_representatives,
                    vertex * /*d_new_representatives*/,
                    vertex * /*d_iter_representatives*/,
                    outgoing_edge_id * /*d_outgoing_edges*/, vertex &num_nodes,
                    vertex &num_edges, vertex &number_of_filled_edges,
                    hipStream_t kernel_stream) {
  const uint32_t n = num_nodes;
  const uint64_t m = num_edges;
  uint64_t wsb = 0;
  die_phem(phem_mst_workspace_bytes(n, m, &wsb), "phem_mst_workspace_bytes");
  const uint64_t ids_off = align_up(wsb, 256);
  char *ws = (char *)cached_workspace(ids_off + 4ull * (n ? n : 1));
  uint32_t *ids = (uint32_t *)(ws + ids_off);
  phem_mst_stats st;
  die_phem(phem_mst_solve_aos(n, m, (const uint32_t *)d_edges, PHEM_W_U32, ws,
                              wsb, ids, d_representatives, nullptr, &st,
                              kernel_stream),
           "phem_mst_solve_aos");
  if ((uint64_t)number_of_filled_edges + st.mst_count > (uint64_t)n) {
    std::fprintf(stderr,
                 "device_boruvka: %u forest edges do not fit after %u filled "
                 "slots of %u\n",
                 st.mst_count, number_of_filled_edges, n);
    std::exit(EXIT_FAILURE);
  }
  die_phem(phem_gather_edges_aos((const uint32_t *)d_edges, ids, st.mst_count,
                                 (uint32_t *)(d_final_msf + number_of_filled_edges),
                                 kernel_stream),
           "phem_gather_edges_aos");
  die_hip(hipStreamSynchronize(kernel_stream), "hipStreamSynchronize");
  number_of_filled_edges += st.mst_count;
  num_edges = 0;
}

void device_boruvka(edge *d_edges, edge *d_final_msf, edge *d_iter_edges,
                    edge *d_msf, vertex *d_representatives,
                    vertex *d_new_representatives,
                    vertex *d_iter_representatives,
                    outgoing_edge_id *d_outgoing_edges, vertex &num_nodes,
                    vertex &num_edges, vertex &number_of_filled_edges,
                    thrust::device_ptr<vertex>, thrust::device_ptr<vertex>,
                    thrust::device_ptr<vertex>, thrust::device_ptr<edge>,
                    hipStream_t kernel_stream) {
  device_boruvka(d_edges, d_final_msf, d_iter_edges, d_msf, d_representatives,
                 d_new_representatives, d_iter_representatives,
                 d_outgoing_edges, num_nodes, num_edges, number_of_filled_edges,
                 kernel_stream);
}

void boruvka_for_edge_array(edge *edges, vertex num_nodes, vertex num_edges,
                            edge *final_msf, vertex &number_of_filled_edges,
                            std::string /*result*/) {
  const uint32_t n = num_nodes;
  edge *d_edges = nullptr, *d_out = nullptr;
  vertex *d_rep = nullptr;
  die_hip(hipMalloc(&d_edges, sizeof(edge) * (num_edges ? num_edges : 1)),
          "hipMalloc");
  die_hip(hipMalloc(&d_out, sizeof(edge) * (n ? n : 1)), "hipMalloc");
  die_hip(hipMalloc(&d_rep, sizeof(vertex) * (n ? n : 1)), "hipMalloc");
  if (num_edges)
    die_hip(hipMemcpy(d_edges, edges, sizeof(edge) * num_edges,
                      hipMemcpyHostToDevice),
            "hipMemcpy H2D");
  vertex nn = num_nodes, mm = num_edges, filled = 0;
  device_boruvka(d_edges, d_out, nullptr, nullptr, d_rep, nullptr, nullptr,
                 nullptr, nn, mm, filled, 0);
  if ((uint64_t)number_of_filled_edges + filled > (uint64_t)n) {
    std::fprintf(stderr, "boruvka_for_edge_array: forest does not fit\n");
    std::exit(EXIT_FAILURE);
  }
  if (filled)
    die_hip(hipMemcpy(final_msf + number_of_filled_edges, d_out,
                      sizeof(edge) * filled, hipMemcpyDeviceToHost),
            "hipMemcpy D2H");
  number_of_filled_edges += filled;
  die_hip(hipFree(d_edges), "hipFree");
  die_hip(hipFree(d_out), "hipFree");
  die_hip(hipFree(d_rep), "hipFree");
}

// ---- MST driver: MatrixMarket file -> forest --------------------------------

void getGraphInfo(std::ifstream &file, vertex &num_nodes,
                  large_vertex &num_edges_large) {
  std::string line;
  while (std::getline(file, line))
    if (line.empty() || line[0] != '%') break;
  unsigned long long rows = 0, cols = 0, entries = 0;
  if (std::sscanf(line.c_str(), "%llu %llu %llu", &rows, &cols, &entries) != 3)
    rows = entries = 0;
  num_nodes = (vertex)rows;
  num_edges_large = (large_vertex)entries;
}

void mmapPopulateEdgeArray(const std::string &filename, edge *edges,
                           vertex num_edges, vertex /*num_nodes*/,
                           bool generate_random_weights) {
  phem_graph_file g;
  int rc = phem_read_mtx(filename.c_str(), generate_random_weights ? 1 : 0, &g);
  if (rc != PHEM_OK) {
    std::fprintf(stderr, "Error reading %s: %s\n", filename.c_str(),
                 phem_last_error());
    return;
  }
  const uint64_t k = g.m < num_edges ? g.m : num_edges;
  for (uint64_t i = 0; i < k; i++) edges[i] = edge{g.u[i], g.v[i], g.w[i]};
  phem_graph_file_free(&g);
}

namespace {

// Forest as "u+1 v+1 w" lines (the streamline path's result file,
// MST/src/kernels.cu:1104-1117) and its exact weight.
uint64_t report_forest(const edge *f, uint64_t k, const std::string &result) {
  uint64_t w = 0;
  for (uint64_t i = 0; i < k; i++) w += f[i].w;
  if (result != "res.txt") {
    FILE *o = std::fopen(result.c_str(), "w");
    if (!o) {
      std::fprintf(stderr, "cannot write %s\n", result.c_str());
      std::exit(EXIT_FAILURE);
    }
    for (uint64_t i = 0; i < k; i++)
      std::fprintf(o, "%llu %llu %u\n", (unsigned long long)f[i].u + 1,
                   (unsigned long long)f[i].v + 1, f[i].w);
    std::fclose(o);
  }
  return w;
}

}  // namespace

namespace {

// The caller's CPU MST for use_cpu_only (phem_set_cpu_mst); NULL: the
// library's host solver.
phem_cpu_mst_fn g_cpu_mst = nullptr;

// The streamline paths: chunks of cs edges folded into a forest kept on the
// device, F = MSF(F u chunk) (MST/src/kernels.cu:849-1154); with `hetero`
// the last chunk is first reduced on the CPU while the GPU folds in the
// others (MST/src/kernels.cu:1155-1826), by the measured rule of
// PHEM_CPU_EDGES_AUTO instead of the reference's fixed 10^8 edges.  Returns
// the forest in ascending edge id and the elapsed microseconds of the solve
// (reading excluded).
std::vector<edge> streamline_forest(const phem_graph_file &g, uint64_t cs,
                                    bool hetero, long long *usec) {
  const auto t0 = std::chrono::steady_clock::now();
  std::vector<uint32_t> ids(g.n ? g.n : 1);
  uint64_t fc = 0;
  phem_chunks_stats st{};
  die_phem(phem_msf_host_chunks(g.n, g.m, g.u, g.v, g.w, PHEM_W_U32, cs,
                                hetero ? PHEM_CPU_EDGES_AUTO : 0, ids.data(),
                                &fc, &st),
           "phem_msf_host_chunks");
  if (hetero && st.cpu_edges)
    std::printf("%llu cut down to %llu\n", (unsigned long long)st.cpu_edges,
                (unsigned long long)st.cpu_kept);
  std::vector<edge> forest(fc);
  for (uint64_t i = 0; i < fc; i++)
    forest[i] = edge{g.u[ids[i]], g.v[ids[i]], g.w[ids[i]]};
  *usec = std::chrono::duration_cast<std::chrono::microseconds>(
              std::chrono::steady_clock::now() - t0).count();
  return forest;
}

// Chunk size of the managed mode: the largest of m, m/2, m/4, ... whose
// phem_msf_host_chunks state (two forest|chunk edge buffers, the solve
// workspace, per-vertex arrays, two staged chunks) takes at most 3/4 of the
// device's free memory.
uint64_t managed_chunk(uint32_t n, uint64_t m) {
  size_t free_b = 0, total_b = 0;
  if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) {
    (void)hipGetLastError();
    return m ? m : 1;
  }
  uint64_t c = m ? m : 1;
  for (; c > (1u << 16); c = (c + 1) / 2) {
    uint64_t ws = 0;
    if (phem_mst_workspace_bytes(n, (uint64_t)n + c, &ws) != PHEM_OK) continue;
    const uint64_t need = 40 * ((uint64_t)n + c) + ws + 20ull * n + 32 * c;
    if (need <= free_b / 4 * 3) break;
  }
  return c;
}

}  // namespace

extern "C" void phem_set_cpu_mst(phem_cpu_mst_fn fn) { g_cpu_mst = fn; }

void new_boruvka(std::string filename, std::string result_filename, bool debug,
                 vertex chunk_size, int num_chunks_ideal,
                 bool generate_random_weights, bool use_malloc_managed,
                 bool use_streamline, bool use_cpu_only,
                 bool use_cpu_gpu_streamline, bool save_weighted_graph) {
  std::printf("new boruvka %s\n", filename.c_str());
  phem_graph_file g;
  die_phem(phem_read_mtx(filename.c_str(), generate_random_weights ? 1 : 0, &g),
           "phem_read_mtx");
  const uint32_t n = g.n;
  const uint64_t m = g.m;
  std::printf("Total num edges %llu\nTotal num nodes %u\n", (unsigned long long)m, n);
  if (num_chunks_ideal > 0)  // MST/src/new_boruvka.cu:326-354
    chunk_size = (vertex)((float)m / ((float)num_chunks_ideal + 0.4f));
  // Dispatch in the reference's order (MST/src/new_boruvka.cu:356-506):
  // debug runs the streamline paths 10 times and reports each time; it
  // checks save_weighted_graph and never use_cpu_only; otherwise
  // use_cpu_only is checked after the streamline and managed modes.
  if (use_streamline || use_cpu_gpu_streamline) {
    const uint64_t cs = chunk_size ? chunk_size : (m ? m : 1);
    std::printf("streamline chunk boruvka, chunk size %llu\n", (unsigned long long)cs);
    const int iters = debug ? 10 : 1;  // execution_time_iterations, :358
    std::vector<edge> forest;
    long long total = 0;
    for (int it = 0; it < iters; it++) {
      long long us = 0;
      forest = streamline_forest(g, cs, use_cpu_gpu_streamline, &us);
      total += us;
      if (debug) std::printf("%d th execution time %lld\n", it, us);
    }
    if (debug) {
      std::printf("Average exec time %lld\n", total / iters);
      std::printf("debug %d\n", (int)debug);
    }
    const uint64_t w = report_forest(forest.data(), forest.size(), result_filename);
    std::printf("final weight by streamline approach %llu\n", (unsigned long long)w);
    phem_graph_file_free(&g);
    return;
  }
  if (use_malloc_managed) {
    // The reference's unified-memory path (MST/src/new_boruvka.cu:183-323,
    // reached at :378-383 and :453-458, debug or not) keeps the graph in
    // memory the host filled and the device pages in on first touch.  Here
    // the graph stays in the host memory the reader filled and the device
    // reads it in chunks sized to its free memory (one chunk when the graph
    // fits, as it does up to billions of edges in 288 GB): explicit copies
    // instead of page faults, which need XNACK.
    int dev = 0, attr = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&attr, hipDeviceAttributeManagedMemory, dev);
    const uint64_t cs = managed_chunk(n, m);
    std::printf("malloc managed (device managed-memory attribute %d): host-resident "
                "graph in chunks of %llu edges\n", attr, (unsigned long long)cs);
    long long us = 0;
    std::vector<edge> forest = streamline_forest(g, cs, false, &us);
    std::printf("Time to execute %lld microseconds\n", us);
    const uint64_t w = report_forest(forest.data(), forest.size(), result_filename);
    std::printf("total weight %llu\n", (unsigned long long)w);
    std::printf("number of filled edges %llu\n", (unsigned long long)forest.size());
    phem_graph_file_free(&g);
    return;
  }
  if (debug && save_weighted_graph) {
    die_phem(phem_write_mtx_weighted("weighted.mtx", &g), "phem_write_mtx_weighted");
    phem_graph_file_free(&g);
    return;
  }
  if (!debug && use_cpu_only) {
    // the CPU path (MST/src/new_boruvka.cu:468-492): the caller's solver
    // registered with phem_set_cpu_mst (the ParlayLib cpu_compute_pbbs
    // behind a plain-array adapter), else the library's host solver
    std::vector<uint32_t> ids(n ? n : 1);
    uint64_t cnt = 0;
    const auto t0 = std::chrono::steady_clock::now();
    if (g_cpu_mst) {
      if (g_cpu_mst(n, m, g.u, g.v, g.w, ids.data(), &cnt) != 0 || cnt > n) {
        std::fprintf(stderr, "new_boruvka: the registered CPU MST failed\n");
        std::exit(EXIT_FAILURE);
      }
      std::sort(ids.begin(), ids.begin() + cnt);
    } else {
      die_phem(phem_cpu_msf(n, m, g.u, g.v, g.w, PHEM_W_U32, ids.data(), &cnt, 0),
               "phem_cpu_msf");
    }
    const long long us = std::chrono::duration_cast<std::chrono::microseconds>(
                             std::chrono::steady_clock::now() - t0).count();
    std::vector<edge> forest(cnt);
    for (uint64_t i = 0; i < cnt; i++) {
      const uint32_t e = ids[i];
      forest[i] = edge{g.u[e], g.v[e], g.w[e]};
    }
    const uint64_t w = report_forest(forest.data(), cnt, result_filename);
    std::printf("pbbs execution time %lld microseconds\n", us);
    std::printf("Total weight %llu\n", (unsigned long long)w);
    phem_graph_file_free(&g);
    return;
  }
  if (m > PHEM_MAX_EDGES) {
    std::fprintf(stderr, "new_boruvka: %llu edges beyond the 32-bit id range\n",
                 (unsigned long long)m);
    std::exit(EXIT_FAILURE);
  }
  std::vector<edge> edges(m ? m : 1);
  for (uint64_t i = 0; i < m; i++) edges[i] = edge{g.u[i], g.v[i], g.w[i]};
  std::vector<edge> forest(n ? n : 1);
  vertex filled = 0;
  std::printf("monolithic GPU approach\n");
  boruvka_for_edge_array(edges.data(), n, (vertex)m, forest.data(), filled,
                         result_filename);
  const uint64_t w = report_forest(forest.data(), filled, result_filename);
  std::printf("Total weight %llu\n", (unsigned long long)w);
  phem_graph_file_free(&g);
}

// ---- CC ---------------------------------------------------------------------

// Defaults for programs that do not define the reference's globals; the
// caller's definitions take precedence.
__attribute__((weak)) int Batch_Size = 1 << 20;
__attribute__((weak)) float GPU_share = 1.0f;
// The caller's CPU path; absent when the whole graph goes to the GPU.
__attribute__((weak)) int host_spanning_tree(struct graph_data_host *);

static int (*g_host_spanning_tree)(struct graph_data_host *) = nullptr;

extern "C" void phem_set_host_spanning_tree(
    int (*fn)(struct graph_data_host *)) {
  g_host_spanning_tree = fn;
}

float SpanningTree(struct graph_data *h_input, struct graph_data *d_input,
                   struct graph_data_host *h_input_host) {
  const uint32_t V = (uint32_t)h_input->V[0];
  const long E = h_input->E[0];
  uint32_t *parent = (uint32_t *)d_input->temp_label;
  die_phem(phem_cc_init(V, parent, d_input->T1edges, nullptr), "phem_cc_init");
  die_phem(phem_cc_labels(V, parent, (uint32_t *)d_input->label, nullptr),
           "phem_cc_labels");
  die_hip(hipDeviceSynchronize(), "hipDeviceSynchronize");

  hipStream_t s1, s2;
  die_hip(hipStreamCreate(&s1), "hipStreamCreate");
  die_hip(hipStreamCreate(&s2), "hipStreamCreate");
  hipEvent_t start, stop, d1, d2;
  die_hip(hipEventCreate(&start), "hipEventCreate");
  die_hip(hipEventCreate(&stop), "hipEventCreate");
  die_hip(hipEventCreate(&d1), "hipEventCreate");
  die_hip(hipEventCreate(&d2), "hipEventCreate");

  const long batch = Batch_Size > 0 ? Batch_Size : 1;
  // GPU share of the records, computed in double precision (a float product
  // loses low bits above 2^24 records)
  long edges_for_gpu = GPU_share >= 1.0f ? E : (long)((double)GPU_share * (double)E);
  if (edges_for_gpu < 0) edges_for_gpu = 0;
  if (edges_for_gpu > E) edges_for_gpu = E;

  die_hip(hipEventRecord(start, nullptr), "hipEventRecord");
  long k = 0;
  for (long i = 0; i < edges_for_gpu; i += batch, k++) {
    const long len = edges_for_gpu - i < batch ? edges_for_gpu - i : batch;
    // alternate buffers and streams: batch k + 2 reuses batch k's buffer
    // after it on the same stream
    uint64_t *buf = (k & 1) ? d_input->edges2 : d_input->edges;
    hipStream_t s = (k & 1) ? s2 : s1;
    die_hip(hipMemcpyAsync(buf, h_input->edges + i, sizeof(uint64_t) * len,
                           hipMemcpyHostToDevice, s),
            "hipMemcpyAsync");
    die_phem(phem_cc_union(V, buf, (uint64_t)len, parent, d_input->T1edges, s),
             "phem_cc_union");
  }
  if (edges_for_gpu < E) {
    int (*cpu)(struct graph_data_host *) =
        g_host_spanning_tree ? g_host_spanning_tree : host_spanning_tree;
    if (!cpu) {
      std::fprintf(stderr,
                   "SpanningTree: GPU_share < 1 but no host_spanning_tree is "
                   "linked\n");
      std::exit(EXIT_FAILURE);
    }
    cpu(h_input_host);
  }
  die_hip(hipEventRecord(d1, s1), "hipEventRecord");
  die_hip(hipEventRecord(d2, s2), "hipEventRecord");
  die_hip(hipStreamWaitEvent(nullptr, d1, 0), "hipStreamWaitEvent");
  die_hip(hipStreamWaitEvent(nullptr, d2, 0), "hipStreamWaitEvent");
  die_hip(hipEventRecord(stop, nullptr), "hipEventRecord");
  die_hip(hipEventSynchronize(stop), "hipEventSynchronize");
  float ms = 0.f;
  die_hip(hipEventElapsedTime(&ms, start, stop), "hipEventElapsedTime");
  die_hip(hipStreamDestroy(s1), "hipStreamDestroy");
  die_hip(hipStreamDestroy(s2), "hipStreamDestroy");
  die_hip(hipEventDestroy(start), "hipEventDestroy");
  die_hip(hipEventDestroy(stop), "hipEventDestroy");
  die_hip(hipEventDestroy(d1), "hipEventDestroy");
  die_hip(hipEventDestroy(d2), "hipEventDestroy");
  return ms;
}

float gpu_spanning_tree(struct graph_data *d_input, uint64_t *extraedges,
                        int vert) {
  hipEvent_t start, stop;
  die_hip(hipEventCreate(&start), "hipEventCreate");
  die_hip(hipEventCreate(&stop), "hipEventCreate");
  die_hip(hipEventRecord(start, nullptr), "hipEventRecord");
  die_phem(phem_cc_merge_forests((uint32_t)vert, d_input->T1edges, extraedges,
                                 (uint32_t *)d_input->temp_label,
                                 d_input->T2edges, (uint32_t *)d_input->label,
                                 nullptr),
           "phem_cc_merge_forests");
  die_hip(hipEventRecord(stop, nullptr), "hipEventRecord");
  die_hip(hipEventSynchronize(stop), "hipEventSynchronize");
  float ms = 0.f;
  die_hip(hipEventElapsedTime(&ms, start, stop), "hipEventElapsedTime");
  die_hip(hipEventDestroy(start), "hipEventDestroy");
  die_hip(hipEventDestroy(stop), "hipEventDestroy");
  return ms;
}
