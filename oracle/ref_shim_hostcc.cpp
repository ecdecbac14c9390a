// C-ABI driver for the reference's HAST CPU union-find
// (CC/host_spanning_tree.cu:67-92 via the graph_data_host struct of
// CC/include/hbcg_utils.cuh:4-18).
// TEST INFRASTRUCTURE ONLY: built into oracle/_ref/ from the reference sources
// where they lie; used to validate the oracle.  Not part of the product.
#include <climits>
#include <cstdint>
#include "include/hbcg_utils.cuh"
#include "include/host_spanning_tree.cuh"

extern "C" __attribute__((visibility("default"))) int ref_host_cc(int n, long m, const uint64_t *edges,
                           uint32_t *labels, uint64_t *sptree_out) {
  graph_data_host h;
  int V = n;
  long E = m, es = m;
  h.V = &V;
  h.E = &E;
  h.edges_size = &es;
  h.edges = parlay::sequence<uint64_t>(edges, edges + m);
  h.label = parlay::sequence<int>::from_function(n, [](size_t i) { return (int)i; });
  h.temp_label = parlay::sequence<int>::from_function(n, [](size_t i) { return (int)i; });
  h.sptree = parlay::sequence<uint64_t>(n, (uint64_t)INT_MAX);
  int rc = host_spanning_tree(&h);
  // the reference leaves the labels pass commented out
  // (CC/host_spanning_tree.cu:79-82): resolve roots through the parent array
  for (int i = 0; i < n; i++) {
    int x = i;
    while (h.temp_label[x] != x) x = h.temp_label[x];
    labels[i] = (uint32_t)x;
  }
  if (sptree_out)
    for (int i = 0; i < n; i++) sptree_out[i] = h.sptree[i];
  return rc;
}
