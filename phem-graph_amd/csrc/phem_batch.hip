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
#include <new>

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
          const uint32_t *pug = (const uint32_t *)pu[g];
          if (!pug) {  // a recorded batch's padding graph: no valid edge
            gu[i] = gv[i] = U32_NONE;
            gw[i] = 0;
            continue;
          }
          const uint64_t j = i - eoff[g];
          a = pug[j];
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
// count: the forest size, or *d_count when d_count is set (a recorded solve
// leaves it on the device; count is then only an upper bound).
__global__ __launch_bounds__(BBLK) void k_batch_counts(
    const uint32_t *__restrict__ ids, uint64_t count, const ull *d_count,
    const ull *__restrict__ eoff, uint32_t k, uint32_t *__restrict__ first,
    uint32_t *__restrict__ counts) {
  if (d_count) count = *d_count;
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
    const uint32_t *__restrict__ ids, uint64_t count, const ull *d_count,
    const ull *__restrict__ voff, const ull *__restrict__ eoff,
    const uint32_t *__restrict__ emap, uint32_t k,
    const uint32_t *__restrict__ first, uint32_t *__restrict__ out) {
  if (d_count) count = *d_count;
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
    const uint32_t *__restrict__ ids, uint64_t count, const ull *d_count,
    const ull *__restrict__ eoff, const uint32_t *__restrict__ emap, uint32_t k,
    const uint32_t *__restrict__ w, int wmode, ull *__restrict__ weights,
    ull *__restrict__ inexact) {
  if (d_count) count = *d_count;
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

// Workspace pointers of a batch (the layout of batch_layout).
struct BatchDev {
  uint32_t *gu, *gv, *gw_own, *ids, *glab, *first, *emap, *vmap, *groot;
  ull *voff, *eoff, *taus;
  uint64_t *ptrs;
  char *mst_ws;
  uint64_t mst_bytes;
};

static BatchDev batch_dev(char *ws, const BatchLayout &L) {
  BatchDev D;
  D.gu = (uint32_t *)(ws + L.gu);
  D.gv = (uint32_t *)(ws + L.gv);
  D.gw_own = (uint32_t *)(ws + L.gw);
  D.ids = (uint32_t *)(ws + L.ids);
  D.glab = (uint32_t *)(ws + L.glab);
  D.first = (uint32_t *)(ws + L.first);
  D.emap = (uint32_t *)(ws + L.emap);
  D.vmap = (uint32_t *)(ws + L.vmap);
  D.groot = (uint32_t *)(ws + L.groot);
  D.voff = (ull *)(ws + L.voff);
  D.eoff = (ull *)(ws + L.eoff);
  D.taus = (ull *)(ws + L.taus);
  D.ptrs = (uint64_t *)(ws + L.ptrs);
  D.mst_ws = ws + L.mst_ws;
  D.mst_bytes = L.mst_bytes;
  return D;
}

// The device work of a batched solve whose offsets (and pointer tables) are
// already in the workspace: tile maps, rebase, the solve over the disjoint
// union, then the per-graph split of the results.  fr: the fixed-round
// schedule of a recording (the forest size then stays on the device and the
// result kernels read it there; their grids come from upper bounds).
static int batch_device(const BatchDev &D, uint32_t k, uint64_t n_total,
                        uint64_t m_total, bool table, const uint32_t *d_u,
                        const uint32_t *d_v, const uint32_t *d_w, int wmode,
                        uint32_t *d_ids, uint32_t *d_counts, uint32_t *d_labels,
                        uint64_t *d_weights, uint64_t *d_inexact,
                        const phem_mst_opts *opts, phem_mst_stats *stats,
                        hipStream_t s, const FixedRounds *fr) {
  const uint32_t *gw = table ? D.gw_own : d_w;
  {
    PHEM_PROF("k_tile_map", s);
    if (m_total)
      hipLaunchKernelGGL(k_tile_map, dim3(grid_for(ceil_div(m_total, MT), BBLK)),
                         dim3(BBLK), 0, s, (const ull *)D.eoff, k, m_total, D.emap);
    if (n_total)
      hipLaunchKernelGGL(k_tile_map, dim3(grid_for(ceil_div(n_total, MT), BBLK)),
                         dim3(BBLK), 0, s, (const ull *)D.voff, k, n_total, D.vmap);
  }
  PHEM_LAUNCHED("k_tile_map");
  if (m_total) {
    PHEM_PROF("k_batch_rebase", s);
    if (table)
      hipLaunchKernelGGL(k_batch_rebase<true>, dim3(grid_for(m_total, RB_PT * BBLK)),
                         dim3(BBLK), 0, s, nullptr, nullptr, D.ptrs, D.ptrs + k,
                         D.ptrs + 2ull * k, m_total, (const ull *)D.voff,
                         (const ull *)D.eoff, D.emap, k, D.gu, D.gv, D.gw_own);
    else
      hipLaunchKernelGGL(k_batch_rebase<false>, dim3(grid_for(m_total, RB_PT * BBLK)),
                         dim3(BBLK), 0, s, d_u, d_v, nullptr, nullptr, nullptr,
                         m_total, (const ull *)D.voff, (const ull *)D.eoff, D.emap,
                         k, D.gu, D.gv, nullptr);
  }
  PHEM_LAUNCHED("k_batch_rebase");
  phem_mst_stats st{};
  const GraphGroups groups{k, D.voff, D.eoff, D.vmap, D.emap, D.taus, D.groot};
  int rc = mst_solve_groups((uint32_t)n_total, m_total, D.gu, D.gv, gw, wmode,
                            D.mst_ws, D.mst_bytes, D.ids, D.glab, opts, &st, s,
                            groups, fr);
  if (rc) return rc;
  // forest size: from the stats, or (recorded) on the device with n_total
  // as its bound
  const ull *d_count = fr ? solve_count_word(D.mst_ws, (uint32_t)n_total, m_total)
                          : nullptr;
  const uint64_t count = fr ? n_total : (uint64_t)st.mst_count;
  {
    PHEM_PROF("k_batch_counts", s);
    hipLaunchKernelGGL(k_batch_counts, dim3(grid_for(k, BBLK)), dim3(BBLK), 0, s,
                       D.ids, count, d_count, (const ull *)D.eoff, k, D.first,
                       d_counts);
  }
  PHEM_LAUNCHED("k_batch_counts");
  if (count) {
    PHEM_PROF("k_batch_split_ids", s);
    hipLaunchKernelGGL(k_batch_split_ids, dim3(grid_for(count, BBLK)),
                       dim3(BBLK), 0, s, D.ids, count, d_count,
                       (const ull *)D.voff, (const ull *)D.eoff, D.emap, k,
                       D.first, d_ids);
  }
  PHEM_LAUNCHED("k_batch_split_ids");
  if (d_labels && n_total) {
    PHEM_PROF("k_batch_split_labels", s);
    hipLaunchKernelGGL(k_batch_split_labels, dim3(grid_for(n_total, BBLK)),
                       dim3(BBLK), 0, s, D.glab, n_total, (const ull *)D.voff,
                       D.vmap, k, d_labels);
  }
  PHEM_LAUNCHED("k_batch_split_labels");
  if (d_weights) {
    {
      PHEM_PROF("k_batch_zero", s);
      hipLaunchKernelGGL(k_batch_zero, dim3(grid_for(k, BBLK)), dim3(BBLK), 0, s,
                         (ull *)d_weights, (uint64_t)k);
      hipLaunchKernelGGL(k_batch_zero, dim3(grid_for(k, BBLK)), dim3(BBLK), 0, s,
                         (ull *)d_inexact, (uint64_t)k);
    }
    PHEM_LAUNCHED("k_batch_zero");
    if (count) {
      PHEM_PROF("k_batch_weights", s);
      hipLaunchKernelGGL(k_batch_weights,
                         dim3(grid_for(ceil_div(count, WCH), BBLK)), dim3(BBLK),
                         0, s, D.ids, count, d_count, (const ull *)D.eoff,
                         D.emap, k, gw, wmode, (ull *)d_weights,
                         (ull *)d_inexact);
    }
    PHEM_LAUNCHED("k_batch_weights");
  }
  if (stats && !fr) *stats = st;
  return PHEM_OK;
}

// Offsets checks shared by the eager and the recorded batch.
static int check_offsets(uint32_t k, const uint64_t *h_voff,
                         const uint64_t *h_eoff) {
  if (k == 0 || !h_voff || !h_eoff)
    return set_err(PHEM_EINVAL, "empty batch or NULL offsets");
  if (h_voff[0] != 0 || h_eoff[0] != 0)
    return set_err(PHEM_EINVAL, "offsets must start at 0");
  for (uint32_t g = 0; g < k; g++)
    if (h_voff[g + 1] < h_voff[g] || h_eoff[g + 1] < h_eoff[g])
      return set_err(PHEM_EINVAL, "offsets of graph %u decrease", g);
  return PHEM_OK;
}

static int check_tables(uint32_t k, const uint64_t *h_eoff,
                        const uint32_t *const *h_pu, const uint32_t *const *h_pv,
                        const uint32_t *const *h_pw) {
  if (!h_pu || !h_pv || !h_pw)
    return set_err(PHEM_EINVAL, "pointer tables must all be given");
  for (uint32_t g = 0; g < k; g++)
    if (h_eoff[g + 1] > h_eoff[g] && (!h_pu[g] || !h_pv[g] || !h_pw[g]))
      return set_err(PHEM_EINVAL, "graph %u has edges but a NULL array", g);
  return PHEM_OK;
}

int batch_solve(const BatchArgs &A, hipStream_t s) {
  const uint32_t k = A.k;
  const uint64_t *h_voff = A.h_voff, *h_eoff = A.h_eoff;
  int rc = check_offsets(k, h_voff, h_eoff);
  if (rc) return rc;
  const uint64_t n_total = h_voff[k], m_total = h_eoff[k];
  uint64_t need = 0;
  rc = batch_workspace_bytes(n_total, m_total, k, &need);
  if (rc) return rc;
  if (!A.d_ws || A.ws_bytes < need)
    return set_err(PHEM_ENOMEM, "batch workspace %llu B < required %llu B",
                   (unsigned long long)A.ws_bytes, (unsigned long long)need);
  const bool table = A.h_pu != nullptr;
  if (table) {
    if ((rc = check_tables(k, h_eoff, A.h_pu, A.h_pv, A.h_pw))) return rc;
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
  const BatchDev D = batch_dev((char *)A.d_ws, L);
  // offsets (and pointer tables) go up from a pinned buffer of the shared
  // pool, held until the solve below has synchronized, so the copies are
  // asynchronous
  const uint64_t words = 2ull * (k + 1) + (table ? 3ull * k : 0ull);
  PinnedLease stage(8 * words, s);
  if (!stage.get()) return set_err(PHEM_ENOMEM, "pinned offset staging");
  uint64_t *hstage = (uint64_t *)stage.get();
  std::memcpy(hstage, h_voff, 8ull * (k + 1));
  std::memcpy(hstage + (k + 1), h_eoff, 8ull * (k + 1));
  PHEM_HIP(hipMemcpyAsync(D.voff, hstage, 8ull * (k + 1), hipMemcpyHostToDevice, s));
  PHEM_HIP(hipMemcpyAsync(D.eoff, hstage + (k + 1), 8ull * (k + 1),
                          hipMemcpyHostToDevice, s));
  if (table) {
    uint64_t *t = hstage + 2ull * (k + 1);
    for (uint32_t g = 0; g < k; g++) {
      t[g] = (uint64_t)(uintptr_t)A.h_pu[g];
      t[k + g] = (uint64_t)(uintptr_t)A.h_pv[g];
      t[2ull * k + g] = (uint64_t)(uintptr_t)A.h_pw[g];
    }
    PHEM_HIP(hipMemcpyAsync(D.ptrs, t, 24ull * k, hipMemcpyHostToDevice, s));
  }
  rc = batch_device(D, k, n_total, m_total, table, A.d_u, A.d_v, A.d_w, A.wmode,
                    A.d_ids, A.d_counts, A.d_labels, A.d_weights, A.d_inexact,
                    A.opts, A.stats, s, nullptr);
  if (rc) return rc;
  stage.done();  // the solve's counter reads synchronized the stream
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

// ---------------------------------------------------------------------------
// A batched solve recorded once into a HIP graph for capacities (k_cap
// graphs, n_cap vertices, m_cap edges) and replayed for any batch that fits:
// the batch is padded with one graph of no valid edge (a NULL pointer
// table entry, which the rebase turns into invalid endpoints) that takes
// the remaining vertices and edges, and by empty graphs after it.  Offsets
// and pointer tables are copied from pinned staging before each replay.

struct phem_batch_graph {
  uint32_t k_cap = 0;
  uint64_t n_cap = 0, m_cap = 0;
  bool split = false;
  bool launched = false;
  phem::FixedRounds fr{0, 0};
  phem::BatchDev D{};
  uint64_t *h_stage = nullptr;  // pinned: voff, eoff (k_cap + 1 each), 3 tables
  hipEvent_t staged = nullptr;  // the last launch's staging copies are done
  unsigned long long *d_sums = nullptr, *h_sums = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

static void batch_graph_free(phem_batch_graph *g) {
  if (!g) return;
  if (g->staged) (void)hipEventSynchronize(g->staged);
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  if (g->staged) (void)hipEventDestroy(g->staged);
  (void)hipFree(g->d_sums);
  (void)hipHostFree(g->h_sums);
  (void)hipHostFree(g->h_stage);
  delete g;
}

extern "C" int phem_batch_graph_create(uint32_t k_cap, uint64_t n_cap,
                                       uint64_t m_cap, int wmode, void *d_ws,
                                       uint64_t ws_bytes, uint32_t *d_ids,
                                       uint32_t *d_counts, uint32_t *d_labels,
                                       uint64_t *d_weights, uint64_t *d_inexact,
                                       const phem_mst_opts *opts, int rounds,
                                       phem_batch_graph_t **out) {
  if (!out) return set_err(PHEM_EINVAL, "out is NULL");
  *out = nullptr;
  if (k_cap < 2)
    return set_err(PHEM_EINVAL, "k_cap %u < 2 (one graph slot pads the batch)", k_cap);
  if (rounds < 0) return set_err(PHEM_EINVAL, "rounds %d < 0", rounds);
  if (!rounds) rounds = 12;
  uint64_t need = 0;
  int rc = batch_workspace_bytes(n_cap, m_cap, k_cap, &need);
  if (rc) return rc;
  if (!d_ws || ws_bytes < need)
    return set_err(PHEM_ENOMEM, "batch workspace %llu B < required %llu B",
                   (unsigned long long)ws_bytes, (unsigned long long)need);
  if ((n_cap && !d_ids) || !d_counts)
    return set_err(PHEM_EINVAL, "NULL output array");
  if ((d_weights != nullptr) != (d_inexact != nullptr))
    return set_err(PHEM_EINVAL, "weights and inexact go together");
  if (wmode != PHEM_W_U32 && wmode != PHEM_W_F32)
    return set_err(PHEM_EINVAL, "unknown weight mode %d", wmode);
  const bool split = mst_uses_split((uint32_t)n_cap, m_cap, opts);
  const FixedRounds fr{rounds, split ? rounds : 0};
  if ((split ? fr.light + fr.heavy : 1 + fr.light) > PHEM_MAX_ROUNDS)
    return set_err(PHEM_EINVAL, "%d rounds per phase exceed the %d-round limit",
                   rounds, PHEM_MAX_ROUNDS);
  phem_batch_graph *g = new (std::nothrow) phem_batch_graph();
  if (!g) return set_err(PHEM_ENOMEM, "host allocation failed");
  g->k_cap = k_cap;
  g->n_cap = n_cap;
  g->m_cap = m_cap;
  g->split = split;
  g->fr = fr;
  uint64_t mb = 0;
  (void)phem_mst_workspace_bytes((uint32_t)n_cap, m_cap, &mb);
  g->D = batch_dev((char *)d_ws, batch_layout(n_cap, m_cap, k_cap, mb));
  const size_t stage_bytes = 8ull * (2ull * (k_cap + 1) + 3ull * k_cap);
  if (hipHostMalloc((void **)&g->h_stage, stage_bytes) != hipSuccess ||
      hipEventCreateWithFlags(&g->staged, hipEventDisableTiming) != hipSuccess ||
      hipMalloc((void **)&g->d_sums, sizeof(ull) * fixed_sums_words()) != hipSuccess ||
      hipHostMalloc((void **)&g->h_sums, sizeof(ull) * fixed_sums_words()) != hipSuccess) {
    (void)hipGetLastError();
    batch_graph_free(g);
    return set_err(PHEM_ENOMEM, "recorded batch buffers");
  }
  // record on a stream of our own; the launch sizing reads the device's
  // CU count before recording starts
  (void)resident_blocks();
  hipStream_t cs = nullptr;
  if (hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess) {
    batch_graph_free(g);
    return set_err(PHEM_EHIP, "hipStreamCreateWithFlags failed");
  }
  if (hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) != hipSuccess) {
    rc = set_err(PHEM_EHIP, "hipStreamBeginCapture failed");
  } else {
    rc = batch_device(g->D, k_cap, n_cap, m_cap, true, nullptr, nullptr, nullptr,
                      wmode, d_ids, d_counts, d_labels, d_weights, d_inexact, opts,
                      nullptr, cs, &g->fr);
    const hipError_t e = hipStreamEndCapture(cs, &g->graph);
    if (!rc && e != hipSuccess)
      rc = set_err(PHEM_EHIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
  }
  (void)hipStreamDestroy(cs);
  if (!rc && hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0) !=
                 hipSuccess)
    rc = set_err(PHEM_EHIP, "hipGraphInstantiate failed");
  if (rc) {
    batch_graph_free(g);
    return rc;
  }
  *out = g;
  return PHEM_OK;
}

extern "C" int phem_batch_graph_launch(phem_batch_graph_t *g, uint32_t k,
                                       const uint64_t *h_voff,
                                       const uint64_t *h_eoff,
                                       const uint32_t *const *h_pu,
                                       const uint32_t *const *h_pv,
                                       const uint32_t *const *h_pw,
                                       void *stream) {
  if (!g) return set_err(PHEM_EINVAL, "graph handle is NULL");
  int rc = check_offsets(k, h_voff, h_eoff);
  if (rc) return rc;
  if (k >= g->k_cap || h_voff[k] > g->n_cap || h_eoff[k] > g->m_cap)
    return set_err(PHEM_EINVAL,
                   "batch of %u graphs, %llu vertices, %llu edges exceeds the "
                   "recorded %u, %llu, %llu (one graph slot pads)", k,
                   (unsigned long long)h_voff[k], (unsigned long long)h_eoff[k],
                   g->k_cap - 1, (unsigned long long)g->n_cap,
                   (unsigned long long)g->m_cap);
  if ((rc = check_tables(k, h_eoff, h_pu, h_pv, h_pw))) return rc;
  hipStream_t s = (hipStream_t)stream;
  // the previous launch's copies must have read the staging first
  if (g->launched) PHEM_HIP(hipEventSynchronize(g->staged));
  const uint32_t K = g->k_cap;
  uint64_t *vo = g->h_stage, *eo = vo + (K + 1), *t = eo + (K + 1);
  for (uint32_t i = 0; i <= k; i++) {
    vo[i] = h_voff[i];
    eo[i] = h_eoff[i];
  }
  for (uint32_t i = k + 1; i <= K; i++) {  // graph k pads, the rest are empty
    vo[i] = g->n_cap;
    eo[i] = g->m_cap;
  }
  for (uint32_t i = 0; i < K; i++) {
    const bool real = i < k;
    t[i] = real ? (uint64_t)(uintptr_t)h_pu[i] : 0;
    t[K + i] = real ? (uint64_t)(uintptr_t)h_pv[i] : 0;
    t[2ull * K + i] = real ? (uint64_t)(uintptr_t)h_pw[i] : 0;
  }
  PHEM_HIP(hipMemcpyAsync(g->D.voff, vo, 8ull * (K + 1), hipMemcpyHostToDevice, s));
  PHEM_HIP(hipMemcpyAsync(g->D.eoff, eo, 8ull * (K + 1), hipMemcpyHostToDevice, s));
  PHEM_HIP(hipMemcpyAsync(g->D.ptrs, t, 24ull * K, hipMemcpyHostToDevice, s));
  PHEM_HIP(hipEventRecord(g->staged, s));
  PHEM_HIP(hipGraphLaunch(g->exec, s));
  g->launched = true;
  return PHEM_OK;
}

extern "C" int phem_batch_graph_stats(phem_batch_graph_t *g,
                                      phem_mst_stats *stats, void *stream) {
  if (!g) return set_err(PHEM_EINVAL, "graph handle is NULL");
  if (!g->launched) return set_err(PHEM_EINVAL, "graph not launched yet");
  return fixed_round_stats(g->D.mst_ws, (uint32_t)g->n_cap, g->m_cap, g->split,
                           g->fr, g->d_sums, g->h_sums, stats,
                           (hipStream_t)stream);
}

extern "C" int phem_batch_graph_destroy(phem_batch_graph_t *g) {
  batch_graph_free(g);
  return PHEM_OK;
}
