// phem_batch.hip -- many independent graphs in one solve.
//
// k graphs are solved as one disjoint union: graph g's vertex x is global
// vertex voff[g] + x and its edge j is global edge eoff[g] + j.  The minimum
// spanning forest of a disjoint union is the union of the forests, and the
// (weight, global id) order restricted to one graph is its (weight, local id)
// order, so one solve gives every graph's exact forest; a graph whose
// components are finished simply has no alive edges or active components in
// later rounds.  Labels stay minimum vertex ids (voff[g] + local minimum).
//
// The graphs come either concatenated (one u / v / w array, graph g at
// eoff[g]) or as k separate device arrays (a pointer table per field); the
// rebase pass reads either and writes the global endpoint arrays the solve
// runs on, so scattered requests are gathered in the same single pass.
#include <cstring>

#include "phem_batch_core.h"
#include "phem_dev.h"
#include "phem_internal.h"

namespace phem {

constexpr int BBLK = 256;

__device__ __forceinline__ uint64_t ceil_div_u64(uint64_t a, uint64_t b) {
  return (a + b - 1) / b;
}

__global__ __launch_bounds__(BBLK) void k_tile_map(const ull *__restrict__ off,
                                                   uint32_t k, uint64_t total,
                                                   uint32_t *__restrict__ map) {
  const uint64_t tiles = ceil_div_u64(total, MT);
  for (uint64_t t = (uint64_t)blockIdx.x * BBLK + threadIdx.x; t < tiles;
       t += (uint64_t)gridDim.x * BBLK)
    map[t] = seg_of(off, k, t * MT);
}

// Global endpoints (a local id outside its graph maps to an invalid vertex)
// and, from pointer tables, the gathered weights.  Consecutive threads take
// consecutive edges, so the reads of one graph's arrays are coalesced.  A
// block takes RB_PT * BBLK consecutive edges at a time and looks their graph
// up once when they all lie in one graph.
constexpr int RB_PT = 4;
template <bool TABLE>
__global__ __launch_bounds__(BBLK) void k_batch_rebase(
    const uint32_t *__restrict__ u, const uint32_t *__restrict__ v,
    const uint64_t *__restrict__ pu, const uint64_t *__restrict__ pv,
    const uint64_t *__restrict__ pw, uint64_t m, const ull *__restrict__ voff,
    const ull *__restrict__ eoff, const uint32_t *__restrict__ emap, uint32_t k,
    uint32_t *__restrict__ gu, uint32_t *__restrict__ gv,
    uint32_t *__restrict__ gw) {
  constexpr uint64_t T = (uint64_t)RB_PT * BBLK;
  for (uint64_t t0 = (uint64_t)blockIdx.x * T; t0 < m; t0 += (uint64_t)gridDim.x * T) {
    const uint64_t t1 = t0 + T < m ? t0 + T : m;
    const uint32_t g0 = graph_of(emap, eoff, k, t0);
    const bool one = g0 == graph_of(emap, eoff, k, t1 - 1);
#pragma unroll
    for (int r = 0; r < RB_PT; r++) {
      const uint64_t i = t0 + (uint64_t)r * BBLK + threadIdx.x;
      if (i < t1) {
        const uint32_t g = one ? g0 : graph_of(emap, eoff, k, i);
        const ull base = voff[g], ng = voff[g + 1] - base;
        uint32_t a, b;
        if (TABLE) {
          const uint64_t j = i - eoff[g];
          a = ((const uint32_t *)pu[g])[j];
          b = ((const uint32_t *)pv[g])[j];
          gw[i] = ((const uint32_t *)pw[g])[j];
        } else {
          a = u[i];
          b = v[i];
        }
        gu[i] = a < ng ? (uint32_t)(base + a) : U32_NONE;
        gv[i] = b < ng ? (uint32_t)(base + b) : U32_NONE;
      }
    }
  }
}

// per-graph forest sizes: ids is ascending, graph g owns [eoff[g], eoff[g+1])
__global__ __launch_bounds__(BBLK) void k_batch_counts(
    const uint32_t *__restrict__ ids, uint64_t count,
    const ull *__restrict__ eoff, uint32_t k, uint32_t *__restrict__ first,
    uint32_t *__restrict__ counts) {
  for (uint64_t g = (uint64_t)blockIdx.x * BBLK + threadIdx.x; g < k;
       g += (uint64_t)gridDim.x * BBLK) {
    auto lower = [&](ull x) {  // first index with ids[i] >= x
      uint64_t lo = 0, hi = count;
      while (lo < hi) {
        const uint64_t mid = (lo + hi) >> 1;
        if (ids[mid] < x) lo = mid + 1; else hi = mid;
      }
      return lo;
    };
    const uint64_t a = lower(eoff[g]), b = lower(eoff[g + 1]);
    first[g] = (uint32_t)a;
    counts[g] = (uint32_t)(b - a);
  }
}

// forest edge i (global id) -> local id at out[voff[g] + (i - first[g])]
__global__ __launch_bounds__(BBLK) void k_batch_split_ids(
    const uint32_t *__restrict__ ids, uint64_t count,
    const ull *__restrict__ voff, const ull *__restrict__ eoff,
    const uint32_t *__restrict__ emap, uint32_t k,
    const uint32_t *__restrict__ first, uint32_t *__restrict__ out) {
  for (uint64_t i = (uint64_t)blockIdx.x * BBLK + threadIdx.x; i < count;
       i += (uint64_t)gridDim.x * BBLK) {
    const uint32_t e = ids[i];
    const uint32_t g = graph_of(emap, eoff, k, e);
    out[voff[g] + (i - first[g])] = (uint32_t)(e - eoff[g]);
  }
}

__global__ __launch_bounds__(BBLK) void k_batch_split_labels(
    const uint32_t *__restrict__ glab, uint64_t n_total,
    const ull *__restrict__ voff, const uint32_t *__restrict__ vmap, uint32_t k,
    uint32_t *__restrict__ out) {
  for (uint64_t x = (uint64_t)blockIdx.x * BBLK + threadIdx.x; x < n_total;
       x += (uint64_t)gridDim.x * BBLK) {
    const uint32_t g = graph_of(vmap, voff, k, x);
    out[x] = (uint32_t)(glab[x] - voff[g]);
  }
}

// Per-graph exact forest weight.  Each thread sums WCH consecutive forest
// edges (ids ascending), flushing at a graph change; then the wave combines
// the lanes' last partial sums by graph (the lanes' graphs are non-decreasing
// across the wave) so that one atomic per graph run of a wave is left.
constexpr int WCH = 8;
__global__ __launch_bounds__(BBLK) void k_batch_weights(
    const uint32_t *__restrict__ ids, uint64_t count,
    const ull *__restrict__ eoff, const uint32_t *__restrict__ emap, uint32_t k,
    const uint32_t *__restrict__ w, int wmode, ull *__restrict__ weights,
    ull *__restrict__ inexact) {
  const uint64_t nthreads = ceil_div_u64(count, WCH);
  const uint64_t stride = (uint64_t)gridDim.x * BBLK;
  // every lane of a wave runs the same number of iterations (the shuffles)
  const uint64_t t0 = (uint64_t)blockIdx.x * BBLK + threadIdx.x;
  const uint64_t iters = ceil_div_u64(nthreads, stride);
  for (uint64_t it = 0; it < iters; it++) {
    const uint64_t t = t0 + it * stride;
    uint32_t g = 0xFFFFFFFFu;
    ull sum = 0, bad = 0;
    if (t < nthreads) {
      const uint64_t a = t * WCH, b = a + WCH < count ? a + WCH : count;
      // the chunk's ids, then its weights, are loaded together
      uint32_t e[WCH], we[WCH];
#pragma unroll
      for (int j = 0; j < WCH; j++) e[j] = ids[a + j < b ? a + j : a];
#pragma unroll
      for (int j = 0; j < WCH; j++) we[j] = w[e[j]];
      g = graph_of(emap, eoff, k, e[0]);
#pragma unroll
      for (int j = 0; j < WCH; j++) {
        if (a + j < b) {
          while (g + 1 < k && eoff[g + 1] <= e[j]) {
            if (sum) atomicAdd(&weights[g], sum);
            if (bad) atomicAdd(&inexact[g], bad);
            sum = bad = 0;
            g++;
          }
          ull fx;
          if (weight_fx(we[j], wmode, &fx)) sum += fx; else bad++;
        }
      }
    }
    // segmented inclusive scan over runs of equal g (lane order = id order)
    const uint32_t lane = lane_id();
#pragma unroll
    for (int d = 1; d < WAVE; d <<= 1) {
      const uint32_t og = __shfl_up(g, d);
      const ull os = __shfl_up(sum, d), ob = __shfl_up(bad, d);
      if (lane >= (uint32_t)d && og == g) {
        sum += os;
        bad += ob;
      }
    }
    const uint32_t ng = __shfl_down(g, 1);
    const bool last = lane == WAVE - 1 || ng != g;
    if (last && g != 0xFFFFFFFFu) {
      if (sum) atomicAdd(&weights[g], sum);
      if (bad) atomicAdd(&inexact[g], bad);
    }
  }
}

__global__ __launch_bounds__(BBLK) void k_batch_zero(ull *a, uint64_t cnt) {
  for (uint64_t i = (uint64_t)blockIdx.x * BBLK + threadIdx.x; i < cnt;
       i += (uint64_t)gridDim.x * BBLK)
    a[i] = 0;
}

struct BatchLayout {
  uint64_t mst_ws, mst_bytes, gu, gv, gw, ids, glab, voff, eoff, ptrs, first,
      emap, vmap, groot, taus, total;
};

static BatchLayout batch_layout(uint64_t n_total, uint64_t m_total, uint32_t k,
                                uint64_t mst_bytes) {
  BatchLayout L;
  uint64_t o = 0;
  auto take = [&](uint64_t bytes) {
    uint64_t at = o;
    o = align_up(o + (bytes ? bytes : 1), 256);
    return at;
  };
  L.mst_bytes = mst_bytes;
  L.mst_ws = take(mst_bytes);
  L.gu = take(4 * m_total);
  L.gv = take(4 * m_total);
  L.gw = take(4 * m_total);
  L.ids = take(4 * n_total);
  L.glab = take(4 * n_total);
  L.voff = take(8ull * (k + 1));
  L.eoff = take(8ull * (k + 1));
  L.ptrs = take(24ull * k);
  L.first = take(4ull * k);
  L.emap = take(4 * (ceil_div(m_total, MT) + 1));
  L.vmap = take(4 * (ceil_div(n_total, MT) + 1));
  L.groot = take(4ull * k);
  L.taus = take(8ull * k);
  L.total = o;
  return L;
}

int batch_workspace_bytes(uint64_t n_total, uint64_t m_total, uint32_t k,
                          uint64_t *bytes) {
  if (!bytes) return set_err(PHEM_EINVAL, "bytes is NULL");
  if (n_total > PHEM_MAX_VERTICES || m_total > PHEM_MAX_EDGES || k == 0)
    return set_err(PHEM_EINVAL, "batch of %u graphs, %llu vertices, %llu edges "
                   "beyond limits", k, (unsigned long long)n_total,
                   (unsigned long long)m_total);
  uint64_t mb = 0;
  int rc = phem_mst_workspace_bytes((uint32_t)n_total, m_total, &mb);
  if (rc) return rc;
  *bytes = batch_layout(n_total, m_total, k, mb).total;
  return PHEM_OK;
}

int batch_solve(const BatchArgs &A, hipStream_t s) {
  const uint32_t k = A.k;
  const uint64_t *h_voff = A.h_voff, *h_eoff = A.h_eoff;
  if (k == 0 || !h_voff || !h_eoff)
    return set_err(PHEM_EINVAL, "empty batch or NULL offsets");
  if (h_voff[0] != 0 || h_eoff[0] != 0)
    return set_err(PHEM_EINVAL, "offsets must start at 0");
  for (uint32_t g = 0; g < k; g++)
    if (h_voff[g + 1] < h_voff[g] || h_eoff[g + 1] < h_eoff[g])
      return set_err(PHEM_EINVAL, "offsets of graph %u decrease", g);
  const uint64_t n_total = h_voff[k], m_total = h_eoff[k];
  uint64_t need = 0;
  int rc = batch_workspace_bytes(n_total, m_total, k, &need);
  if (rc) return rc;
  if (!A.d_ws || A.ws_bytes < need)
    return set_err(PHEM_ENOMEM, "batch workspace %llu B < required %llu B",
                   (unsigned long long)A.ws_bytes, (unsigned long long)need);
  const bool table = A.h_pu != nullptr;
  if (table) {
    if (!A.h_pv || !A.h_pw)
      return set_err(PHEM_EINVAL, "pointer tables must all be given");
    for (uint32_t g = 0; g < k; g++)
      if (h_eoff[g + 1] > h_eoff[g] && (!A.h_pu[g] || !A.h_pv[g] || !A.h_pw[g]))
        return set_err(PHEM_EINVAL, "graph %u has edges but a NULL array", g);
  } else if (m_total && (!A.d_u || !A.d_v || !A.d_w)) {
    return set_err(PHEM_EINVAL, "NULL edge array");
  }
  if ((n_total && !A.d_ids) || !A.d_counts)
    return set_err(PHEM_EINVAL, "NULL output array");
  if ((A.d_weights != nullptr) != (A.d_inexact != nullptr))
    return set_err(PHEM_EINVAL, "weights and inexact go together");
  uint64_t mb = 0;
  rc = phem_mst_workspace_bytes((uint32_t)n_total, m_total, &mb);
  if (rc) return rc;
  const BatchLayout L = batch_layout(n_total, m_total, k, mb);
  char *ws = (char *)A.d_ws;
  uint32_t *gu = (uint32_t *)(ws + L.gu), *gv = (uint32_t *)(ws + L.gv);
  uint32_t *ids = (uint32_t *)(ws + L.ids), *glab = (uint32_t *)(ws + L.glab);
  ull *voff = (ull *)(ws + L.voff), *eoff = (ull *)(ws + L.eoff);
  uint64_t *ptrs = (uint64_t *)(ws + L.ptrs);
  uint32_t *first = (uint32_t *)(ws + L.first);
  uint32_t *emap = (uint32_t *)(ws + L.emap), *vmap = (uint32_t *)(ws + L.vmap);
  const uint32_t *gw = table ? (const uint32_t *)(ws + L.gw) : A.d_w;
  // offsets (and pointer tables) go up from a pinned buffer of the shared
  // pool, held until the solve below has synchronized, so the copies are
  // asynchronous
  const uint64_t words = 2ull * (k + 1) + (table ? 3ull * k : 0ull);
  PinnedLease stage(8 * words, s);
  if (!stage.get()) return set_err(PHEM_ENOMEM, "pinned offset staging");
  uint64_t *hstage = (uint64_t *)stage.get();
  std::memcpy(hstage, h_voff, 8ull * (k + 1));
  std::memcpy(hstage + (k + 1), h_eoff, 8ull * (k + 1));
  PHEM_HIP(hipMemcpyAsync(voff, hstage, 8ull * (k + 1), hipMemcpyHostToDevice, s));
  PHEM_HIP(hipMemcpyAsync(eoff, hstage + (k + 1), 8ull * (k + 1),
                          hipMemcpyHostToDevice, s));
  if (table) {
    uint64_t *t = hstage + 2ull * (k + 1);
    for (uint32_t g = 0; g < k; g++) {
      t[g] = (uint64_t)(uintptr_t)A.h_pu[g];
      t[k + g] = (uint64_t)(uintptr_t)A.h_pv[g];
      t[2ull * k + g] = (uint64_t)(uintptr_t)A.h_pw[g];
    }
    PHEM_HIP(hipMemcpyAsync(ptrs, t, 24ull * k, hipMemcpyHostToDevice, s));
  }
  {
    PHEM_PROF("k_tile_map", s);
    if (m_total)
      hipLaunchKernelGGL(k_tile_map, dim3(grid_for(ceil_div(m_total, MT), BBLK)),
                         dim3(BBLK), 0, s, eoff, k, m_total, emap);
    if (n_total)
      hipLaunchKernelGGL(k_tile_map, dim3(grid_for(ceil_div(n_total, MT), BBLK)),
                         dim3(BBLK), 0, s, voff, k, n_total, vmap);
  }
  PHEM_LAUNCHED("k_tile_map");
  if (m_total) {
    PHEM_PROF("k_batch_rebase", s);
    if (table)
      hipLaunchKernelGGL(k_batch_rebase<true>, dim3(grid_for(m_total, RB_PT * BBLK)),
                         dim3(BBLK), 0, s, nullptr, nullptr, ptrs, ptrs + k,
                         ptrs + 2ull * k, m_total, voff, eoff, emap, k, gu, gv,
                         (uint32_t *)(ws + L.gw));
    else
      hipLaunchKernelGGL(k_batch_rebase<false>, dim3(grid_for(m_total, RB_PT * BBLK)),
                         dim3(BBLK), 0, s, A.d_u, A.d_v, nullptr, nullptr,
                         nullptr, m_total, voff, eoff, emap, k, gu, gv, nullptr);
  }
  PHEM_LAUNCHED("k_batch_rebase");
  phem_mst_stats st;
  const GraphGroups groups{k,    voff, eoff, vmap, emap, (ull *)(ws + L.taus),
                           (uint32_t *)(ws + L.groot)};
  rc = mst_solve_groups((uint32_t)n_total, m_total, gu, gv, gw, A.wmode,
                        ws + L.mst_ws, L.mst_bytes, ids, glab, A.opts, &st, s,
                        groups);
  if (rc) return rc;
  stage.done();  // the solve's counter reads synchronized the stream
  {
    PHEM_PROF("k_batch_counts", s);
    hipLaunchKernelGGL(k_batch_counts, dim3(grid_for(k, BBLK)), dim3(BBLK), 0, s,
                       ids, (uint64_t)st.mst_count, eoff, k, first, A.d_counts);
  }
  PHEM_LAUNCHED("k_batch_counts");
  if (st.mst_count) {
    PHEM_PROF("k_batch_split_ids", s);
    hipLaunchKernelGGL(k_batch_split_ids, dim3(grid_for(st.mst_count, BBLK)),
                       dim3(BBLK), 0, s, ids, (uint64_t)st.mst_count, voff,
                       eoff, emap, k, first, A.d_ids);
  }
  PHEM_LAUNCHED("k_batch_split_ids");
  if (A.d_labels && n_total) {
    PHEM_PROF("k_batch_split_labels", s);
    hipLaunchKernelGGL(k_batch_split_labels, dim3(grid_for(n_total, BBLK)),
                       dim3(BBLK), 0, s, glab, n_total, voff, vmap, k,
                       A.d_labels);
  }
  PHEM_LAUNCHED("k_batch_split_labels");
  if (A.d_weights) {
    {
      PHEM_PROF("k_batch_zero", s);
      hipLaunchKernelGGL(k_batch_zero, dim3(grid_for(k, BBLK)), dim3(BBLK), 0, s,
                         (ull *)A.d_weights, (uint64_t)k);
      hipLaunchKernelGGL(k_batch_zero, dim3(grid_for(k, BBLK)), dim3(BBLK), 0, s,
                         (ull *)A.d_inexact, (uint64_t)k);
    }
    PHEM_LAUNCHED("k_batch_zero");
    if (st.mst_count) {
      PHEM_PROF("k_batch_weights", s);
      hipLaunchKernelGGL(k_batch_weights,
                         dim3(grid_for(ceil_div(st.mst_count, WCH), BBLK)),
                         dim3(BBLK), 0, s, ids, (uint64_t)st.mst_count, eoff,
                         emap, k, gw, A.wmode, (ull *)A.d_weights,
                         (ull *)A.d_inexact);
    }
    PHEM_LAUNCHED("k_batch_weights");
  }
  if (A.stats) *A.stats = st;
  return PHEM_OK;
}

}  // namespace phem

using namespace phem;

extern "C" int phem_batch_workspace_bytes(uint64_t n_total, uint64_t m_total,
                                          uint32_t k, uint64_t *bytes) {
  return batch_workspace_bytes(n_total, m_total, k, bytes);
}

extern "C" int phem_batch_solve(uint32_t k, const uint64_t *h_voff,
                                const uint64_t *h_eoff, const uint32_t *d_u,
                                const uint32_t *d_v, const uint32_t *d_w,
                                int wmode, void *d_ws, uint64_t ws_bytes,
                                uint32_t *d_ids, uint32_t *d_counts,
                                uint32_t *d_labels, uint64_t *d_weights,
                                uint64_t *d_inexact, const phem_mst_opts *opts,
                                phem_mst_stats *stats, void *stream) {
  BatchArgs A{};
  A.k = k;
  A.h_voff = h_voff;
  A.h_eoff = h_eoff;
  A.d_u = d_u;
  A.d_v = d_v;
  A.d_w = d_w;
  A.wmode = wmode;
  A.d_ws = d_ws;
  A.ws_bytes = ws_bytes;
  A.d_ids = d_ids;
  A.d_counts = d_counts;
  A.d_labels = d_labels;
  A.d_weights = d_weights;
  A.d_inexact = d_inexact;
  A.opts = opts;
  A.stats = stats;
  return batch_solve(A, (hipStream_t)stream);
}

extern "C" int phem_batch_solve_ptrs(
    uint32_t k, const uint64_t *h_voff, const uint64_t *h_eoff,
    const uint32_t *const *h_pu, const uint32_t *const *h_pv,
    const uint32_t *const *h_pw, int wmode, void *d_ws, uint64_t ws_bytes,
    uint32_t *d_ids, uint32_t *d_counts, uint32_t *d_labels,
    uint64_t *d_weights, uint64_t *d_inexact, const phem_mst_opts *opts,
    phem_mst_stats *stats, void *stream) {
  if (!h_pu) return set_err(PHEM_EINVAL, "pointer tables are NULL");
  BatchArgs A{};
  A.k = k;
  A.h_voff = h_voff;
  A.h_eoff = h_eoff;
  A.h_pu = h_pu;
  A.h_pv = h_pv;
  A.h_pw = h_pw;
  A.wmode = wmode;
  A.d_ws = d_ws;
  A.ws_bytes = ws_bytes;
  A.d_ids = d_ids;
  A.d_counts = d_counts;
  A.d_labels = d_labels;
  A.d_weights = d_weights;
  A.d_inexact = d_inexact;
  A.opts = opts;
  A.stats = stats;
  return batch_solve(A, (hipStream_t)stream);
}
