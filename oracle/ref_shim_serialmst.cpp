// C-ABI driver for the reference's serial filter-Kruskal MST
// (MST/src/bench/serialMST.C:62-101, types MST/src/bench/MST.h:5-9).
// TEST INFRASTRUCTURE ONLY: built into oracle/_ref/ from the reference sources
// where they lie; used to validate the oracle.  Not part of the product.
#include <cstdint>
#include <cstring>
#include "bench/MST.h"

extern "C" __attribute__((visibility("default"))) long ref_serial_mst(int n, long m, const uint32_t *u,
                               const uint32_t *v, const float *w,
                               uint32_t *out_ids) {
  parlay::sequence<wghEdge<vertexId, edgeWeight>> E(m);
  for (long i = 0; i < m; i++)
    E[i] = wghEdge<vertexId, edgeWeight>((int)u[i], (int)v[i], w[i]);
  wghEdgeArray<vertexId, edgeWeight> G(std::move(E), n);
  parlay::sequence<edgeId> r = mst(G);
  for (size_t i = 0; i < r.size(); i++) out_ids[i] = r[i];
  return (long)r.size();
}
