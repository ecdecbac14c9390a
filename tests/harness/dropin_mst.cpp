// Test harness: C entry points around the C++ drop-in MST functions of
// libphem.so (boruvka_for_edge_array, device_boruvka), for ctypes tests.
#include <cstring>
#include <thrust/device_ptr.h>

#include "phem.h"
#include "phem_dropin.h"

extern "C" int harness_boruvka_for_edge_array(const uint32_t *h_aos, uint32_t n,
                                              uint32_t m, uint32_t *out_aos,
                                              uint32_t *filled_io) {
  vertex filled = *filled_io;
  boruvka_for_edge_array((edge *)h_aos, n, m, (edge *)out_aos, filled);
  *filled_io = filled;
  return 0;
}

// device_boruvka with the reference's full parameter list on device copies.
extern "C" int harness_device_boruvka(const uint32_t *h_aos, uint32_t n,
                                      uint32_t m, uint32_t *out_aos,
                                      uint32_t *labels, uint32_t *filled_io,
                                      uint32_t *m_after) {
  edge *d_edges, *d_final, *d_iter, *d_msf;
  vertex *d_rep, *d_new, *d_it;
  outgoing_edge_id *d_out;
  const size_t nn = n ? n : 1, mm = m ? m : 1;
  if (hipMalloc(&d_edges, sizeof(edge) * mm) || hipMalloc(&d_final, sizeof(edge) * nn) ||
      hipMalloc(&d_iter, sizeof(edge) * mm) || hipMalloc(&d_msf, sizeof(edge) * nn) ||
      hipMalloc(&d_rep, 4 * nn) || hipMalloc(&d_new, 4 * nn) || hipMalloc(&d_it, 4 * nn) ||
      hipMalloc(&d_out, 8 * nn))
    return -1;
  if (m && hipMemcpy(d_edges, h_aos, sizeof(edge) * m, hipMemcpyHostToDevice)) return -1;
  vertex vn = n, vm = m, filled = 0;
  device_boruvka(d_edges, d_final, d_iter, d_msf, d_rep, d_new, d_it, d_out, vn, vm,
                 filled, thrust::device_ptr<vertex>(d_new),
                 thrust::device_ptr<vertex>(d_rep), thrust::device_ptr<vertex>(d_it),
                 thrust::device_ptr<edge>(d_final), 0);
  if (filled && hipMemcpy(out_aos, d_final, sizeof(edge) * filled, hipMemcpyDeviceToHost))
    return -1;
  if (n && hipMemcpy(labels, d_rep, 4 * n, hipMemcpyDeviceToHost)) return -1;
  *filled_io = filled;
  *m_after = vm;
  (void)hipFree(d_edges); (void)hipFree(d_final); (void)hipFree(d_iter); (void)hipFree(d_msf);
  (void)hipFree(d_rep); (void)hipFree(d_new); (void)hipFree(d_it); (void)hipFree(d_out);
  return 0;
}

// new_boruvka on a MatrixMarket file; mode 0 plain, 1 streamline,
// 2 heterogeneous streamline, 3 managed, 4 cpu only.
extern "C" int harness_new_boruvka(const char *file, const char *result,
                                   uint32_t chunk_size, int num_chunks,
                                   int generate_weights, int mode, int debug) {
  new_boruvka(file, result, debug != 0, chunk_size, num_chunks,
              generate_weights != 0, mode == 3, mode == 1, mode == 4, mode == 2,
              false);
  return 0;
}

// A caller's CPU MST registered with phem_set_cpu_mst: the library's host
// solver, ids handed back in descending order, calls counted.
static int g_cpu_calls = 0;
static int cpu_mst_reversed(uint32_t n, uint64_t m, const uint32_t *u,
                            const uint32_t *v, const uint32_t *w, uint32_t *ids,
                            uint64_t *count) {
  g_cpu_calls++;
  int rc = phem_cpu_msf(n, m, u, v, w, PHEM_W_U32, ids, count, 2);
  for (uint64_t i = 0, j = *count; i + 1 < j; i++, j--) {
    uint32_t t = ids[i];
    ids[i] = ids[j - 1];
    ids[j - 1] = t;
  }
  return rc;
}
extern "C" void harness_register_cpu_mst(int on) {
  phem_set_cpu_mst(on ? cpu_mst_reversed : nullptr);
}
extern "C" int harness_cpu_mst_calls() { return g_cpu_calls; }
extern "C" void harness_release_workspace() { phem_dropin_release_workspace(); }

extern "C" int harness_graph_info(const char *file, uint32_t *n, uint64_t *m) {
  std::ifstream f(file);
  vertex nn = 0;
  large_vertex mm = 0;
  getGraphInfo(f, nn, mm);
  *n = nn;
  *m = mm;
  return 0;
}

extern "C" int harness_mmap_populate(const char *file, uint32_t *out_aos, uint32_t m,
                                     uint32_t n, int generate_weights) {
  mmapPopulateEdgeArray(file, (edge *)out_aos, m, n, generate_weights != 0);
  return 0;
}
