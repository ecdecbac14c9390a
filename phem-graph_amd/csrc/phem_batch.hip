// phem_batch.hip -- many independent graphs in one solve.
//
// k graphs are solved as one disjoint union: graph g's vertex x is global
// vertex voff[g] + x and its edge j is global edge eoff[g] + j.  The minimum
// spanning forest of a disjoint union is the union of the forests, and the
// (weight, global id) order restricted to one graph is its (weight, local id)
// order, so one solve gives every graph's exact forest; a graph whose
// components are finished simply has no alive edges or active components in
// later rounds.  Labels stay minimum vertex ids (voff[g] + local minimum).
#include <cstring>

#include "phem_dev.h"
#include "phem_internal.h"

namespace phem {

constexpr int BBLK = 256;

// index g with off[g] <= x < off[g + 1] (off ascending, off[0] = 0)
__device__ __forceinline__ uint64_t ceil_div_u64(uint64_t a, uint64_t b) {
  return (a + b - 1) / b;
}

__device__ __forceinline__ uint32_t seg_of(const ull *__restrict__ off,
                                           uint32_t k, ull x) {
  uint32_t lo = 0, hi = k - 1;
  while (lo < hi) {
    const uint32_t mid = (lo + hi + 1) >> 1;
    if (off[mid] <= x) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// global endpoints; a local id outside its graph maps to an invalid vertex
__global__ __launch_bounds__(BBLK) void k_batch_rebase(
    const uint32_t *__restrict__ u, const uint32_t *__restrict__ v,
    uint64_t m, const ull *__restrict__ voff, const ull *__restrict__ eoff,
    uint32_t k, uint32_t *__restrict__ gu, uint32_t *__restrict__ gv) {
  for (uint64_t i = (uint64_t)blockIdx.x * BBLK + threadIdx.x; i < m;
       i += (uint64_t)gridDim.x * BBLK) {
    const uint32_t g = seg_of(eoff, k, i);
    const ull base = voff[g], ng = voff[g + 1] - base;
    const uint32_t a = u[i], b = v[i];
    gu[i] = a < ng ? (uint32_t)(base + a) : U32_NONE;
    gv[i] = b < ng ? (uint32_t)(base + b) : U32_NONE;
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
    const ull *__restrict__ voff, const ull *__restrict__ eoff, uint32_t k,
    const uint32_t *__restrict__ first, uint32_t *__restrict__ out) {
  for (uint64_t i = (uint64_t)blockIdx.x * BBLK + threadIdx.x; i < count;
       i += (uint64_t)gridDim.x * BBLK) {
    const uint32_t e = ids[i];
    const uint32_t g = seg_of(eoff, k, e);
    out[voff[g] + (i - first[g])] = (uint32_t)(e - eoff[g]);
  }
}

__global__ __launch_bounds__(BBLK) void k_batch_split_labels(
    const uint32_t *__restrict__ glab, uint64_t n_total,
    const ull *__restrict__ voff, uint32_t k, uint32_t *__restrict__ out) {
  for (uint64_t x = (uint64_t)blockIdx.x * BBLK + threadIdx.x; x < n_total;
       x += (uint64_t)gridDim.x * BBLK) {
    const uint32_t g = seg_of(voff, k, x);
    out[x] = (uint32_t)(glab[x] - voff[g]);
  }
}

// per-graph exact forest weight: each thread sums a run of WCH consecutive
// forest edges (ids ascending, so a run rarely crosses a graph) and flushes
// one atomic per graph it touched
constexpr int WCH = 256;
__global__ __launch_bounds__(BBLK) void k_batch_weights(
    const uint32_t *__restrict__ ids, uint64_t count,
    const ull *__restrict__ eoff, uint32_t k, const uint32_t *__restrict__ w,
    int wmode, ull *__restrict__ weights, ull *__restrict__ inexact) {
  const uint64_t nthreads = ceil_div_u64(count, WCH);
  for (uint64_t t = (uint64_t)blockIdx.x * BBLK + threadIdx.x; t < nthreads;
       t += (uint64_t)gridDim.x * BBLK) {
    const uint64_t a = t * WCH, b = a + WCH < count ? a + WCH : count;
    uint32_t g = seg_of(eoff, k, ids[a]);
    ull sum = 0, bad = 0;
    for (uint64_t i = a; i < b; i++) {
      const uint32_t e = ids[i];
      while (g + 1 < k && eoff[g + 1] <= e) {
        if (sum) atomicAdd(&weights[g], sum);
        if (bad) atomicAdd(&inexact[g], bad);
        sum = bad = 0;
        g++;
      }
      ull fx;
      if (weight_fx(w[e], wmode, &fx)) sum += fx; else bad++;
    }
    if (sum) atomicAdd(&weights[g], sum);
    if (bad) atomicAdd(&inexact[g], bad);
  }
}

__global__ __launch_bounds__(BBLK) void k_batch_zero(ull *a, uint64_t cnt) {
  for (uint64_t i = (uint64_t)blockIdx.x * BBLK + threadIdx.x; i < cnt;
       i += (uint64_t)gridDim.x * BBLK)
    a[i] = 0;
}

struct BatchLayout {
  uint64_t mst_ws, mst_bytes, gu, gv, ids, glab, voff, eoff, first, total;
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
  L.ids = take(4 * n_total);
  L.glab = take(4 * n_total);
  L.voff = take(8ull * (k + 1));
  L.eoff = take(8ull * (k + 1));
  L.first = take(4ull * k);
  L.total = o;
  return L;
}

}  // namespace phem

using namespace phem;

extern "C" int phem_batch_workspace_bytes(uint64_t n_total, uint64_t m_total,
                                          uint32_t k, uint64_t *bytes) {
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

extern "C" int phem_batch_solve(uint32_t k, const uint64_t *h_voff,
                                const uint64_t *h_eoff, const uint32_t *d_u,
                                const uint32_t *d_v, const uint32_t *d_w,
                                int wmode, void *d_ws, uint64_t ws_bytes,
                                uint32_t *d_ids, uint32_t *d_counts,
                                uint32_t *d_labels, uint64_t *d_weights,
                                uint64_t *d_inexact, const phem_mst_opts *opts,
                                phem_mst_stats *stats, void *stream) {
  if (k == 0 || !h_voff || !h_eoff)
    return set_err(PHEM_EINVAL, "empty batch or NULL offsets");
  if (h_voff[0] != 0 || h_eoff[0] != 0)
    return set_err(PHEM_EINVAL, "offsets must start at 0");
  for (uint32_t g = 0; g < k; g++)
    if (h_voff[g + 1] < h_voff[g] || h_eoff[g + 1] < h_eoff[g])
      return set_err(PHEM_EINVAL, "offsets of graph %u decrease", g);
  const uint64_t n_total = h_voff[k], m_total = h_eoff[k];
  uint64_t need = 0;
  int rc = phem_batch_workspace_bytes(n_total, m_total, k, &need);
  if (rc) return rc;
  if (!d_ws || ws_bytes < need)
    return set_err(PHEM_ENOMEM, "batch workspace %llu B < required %llu B",
                   (unsigned long long)ws_bytes, (unsigned long long)need);
  if ((m_total && (!d_u || !d_v || !d_w)) || (n_total && !d_ids) || !d_counts)
    return set_err(PHEM_EINVAL, "NULL array");
  uint64_t mb = 0;
  rc = phem_mst_workspace_bytes((uint32_t)n_total, m_total, &mb);
  if (rc) return rc;
  const BatchLayout L = batch_layout(n_total, m_total, k, mb);
  hipStream_t s = (hipStream_t)stream;
  char *ws = (char *)d_ws;
  uint32_t *gu = (uint32_t *)(ws + L.gu), *gv = (uint32_t *)(ws + L.gv);
  uint32_t *ids = (uint32_t *)(ws + L.ids), *glab = (uint32_t *)(ws + L.glab);
  ull *voff = (ull *)(ws + L.voff), *eoff = (ull *)(ws + L.eoff);
  uint32_t *first = (uint32_t *)(ws + L.first);
  PHEM_HIP(hipMemcpyAsync(voff, h_voff, 8ull * (k + 1), hipMemcpyHostToDevice, s));
  PHEM_HIP(hipMemcpyAsync(eoff, h_eoff, 8ull * (k + 1), hipMemcpyHostToDevice, s));
  if (m_total) {
    PHEM_PROF("k_batch_rebase", s);
    hipLaunchKernelGGL(k_batch_rebase, dim3(grid_for(m_total, BBLK)), dim3(BBLK),
                       0, s, d_u, d_v, m_total, voff, eoff, k, gu, gv);
  }
  PHEM_LAUNCHED("k_batch_rebase");
  phem_mst_stats st;
  rc = phem_mst_solve((uint32_t)n_total, m_total, gu, gv, d_w, wmode,
                      ws + L.mst_ws, L.mst_bytes, ids, glab, opts, &st, stream);
  if (rc) return rc;
  {
    PHEM_PROF("k_batch_counts", s);
    hipLaunchKernelGGL(k_batch_counts, dim3(grid_for(k, BBLK)), dim3(BBLK), 0, s,
                       ids, (uint64_t)st.mst_count, eoff, k, first, d_counts);
  }
  PHEM_LAUNCHED("k_batch_counts");
  if (st.mst_count) {
    PHEM_PROF("k_batch_split_ids", s);
    hipLaunchKernelGGL(k_batch_split_ids, dim3(grid_for(st.mst_count, BBLK)),
                       dim3(BBLK), 0, s, ids, (uint64_t)st.mst_count, voff,
                       eoff, k, first, d_ids);
  }
  PHEM_LAUNCHED("k_batch_split_ids");
  if (d_labels && n_total) {
    PHEM_PROF("k_batch_split_labels", s);
    hipLaunchKernelGGL(k_batch_split_labels, dim3(grid_for(n_total, BBLK)),
                       dim3(BBLK), 0, s, glab, n_total, voff, k, d_labels);
  }
  PHEM_LAUNCHED("k_batch_split_labels");
  if (d_weights || d_inexact) {
    if (!d_weights || !d_inexact)
      return set_err(PHEM_EINVAL, "weights and inexact go together");
    {
      PHEM_PROF("k_batch_zero", s);
      hipLaunchKernelGGL(k_batch_zero, dim3(grid_for(k, BBLK)), dim3(BBLK), 0, s,
                         (ull *)d_weights, (uint64_t)k);
      hipLaunchKernelGGL(k_batch_zero, dim3(grid_for(k, BBLK)), dim3(BBLK), 0, s,
                         (ull *)d_inexact, (uint64_t)k);
    }
    PHEM_LAUNCHED("k_batch_zero");
    if (st.mst_count) {
      PHEM_PROF("k_batch_weights", s);
      hipLaunchKernelGGL(k_batch_weights,
                         dim3(grid_for(ceil_div(st.mst_count, WCH), BBLK)),
                         dim3(BBLK), 0, s, ids, (uint64_t)st.mst_count, eoff, k,
                         d_w, wmode, (ull *)d_weights, (ull *)d_inexact);
    }
    PHEM_LAUNCHED("k_batch_weights");
  }
  if (stats) *stats = st;
  return PHEM_OK;
}
