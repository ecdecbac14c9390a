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

// The reference's CPU path timed alone: host_spanning_tree on a prepared
// graph_data_host (the parlay sequences are built before the clock starts,
// as CC/main.cu builds them before calling SpanningTree).  The reference
// leaves its label pass out (CC/host_spanning_tree.cu:79-82), so the timed
// region is the union-find over all m records.  Returns seconds.
#include <chrono>
extern "C" __attribute__((visibility("default"))) double ref_host_cc_time(
    int n, long m, const uint64_t *edges) {
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
  const auto t0 = std::chrono::steady_clock::now();
  host_spanning_tree(&h);
  const auto t1 = std::chrono::steady_clock::now();
  return std::chrono::duration<double>(t1 - t0).count();
}

// Number of worker threads of the reference's ParlayLib scheduler (set by
// PARLAY_NUM_THREADS before its first use).
extern "C" __attribute__((visibility("default"))) int ref_host_cc_workers() {
  return (int)parlay::num_workers();
}
