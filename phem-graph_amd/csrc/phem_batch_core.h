// phem_batch_core.h -- the batched solve as called by the C ABI and by the
// serving loop.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "phem.h"

namespace phem {

// One batch of k graphs.  Edges come concatenated (d_u / d_v / d_w, graph g
// at h_eoff[g]) or, when h_pu is set, as host tables of k device pointers.
struct BatchArgs {
  uint32_t k;
  const uint64_t *h_voff, *h_eoff;
  const uint32_t *d_u, *d_v, *d_w;
  const uint32_t *const *h_pu, *const *h_pv, *const *h_pw;
  int wmode;
  void *d_ws;
  uint64_t ws_bytes;
  uint32_t *d_ids, *d_counts, *d_labels;
  uint64_t *d_weights, *d_inexact;
  const phem_mst_opts *opts;
  phem_mst_stats *stats;
};

int batch_workspace_bytes(uint64_t n_total, uint64_t m_total, uint32_t k,
                          uint64_t *bytes);
int batch_solve(const BatchArgs &a, hipStream_t s);

}  // namespace phem
