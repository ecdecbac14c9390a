// phem_cc.hip -- union-find connected components / spanning forest for gfx950.
//
// Semantics follow the reference HAST GPU path: records are packed
// (u << 32) | v and only u < v is processed (union_find_gpu_COO,
// CC/ExternalSpanningTree.cu:77-86); the larger root is hooked under the
// smaller by compare-and-swap and the hooking edge is stored at the hooked
// root (union_async, CC/ExternalSpanningTree.cu:46-59); labels are roots, i.e.
// minimum vertex ids (cc, CC/GPUSpanningTree.cu:122-128).  Finds use path
// halving instead of the reference's two-pass compression
// (custom_compress, CC/ExternalSpanningTree.cu:27-43): both only ever move a
// pointer to an ancestor.
#include <cstdlib>

#include "phem_dev.h"
#include "phem_internal.h"
#include "phem_split.h"

namespace phem {

constexpr int CBLK = 256;  // threads per block of the small passes
// The unions run with B threads per block and E records per thread per tile
// (the cc_block / cc_ept knobs for the unfiltered union, cc_block_filter /
// cc_ept_filter for the filtered ones).  Built-in shapes: 256 threads and 16
// records per thread in the unfiltered union (every record a union: more
// records in flight per thread measured faster, 0.72 ms against 0.80 ms for
// the first n records at n = 2^24); 1024 threads and 8 records per thread in
// the filtered passes (8 rather than 16: fewer registers and more resident
// waves, 2.09 ms against 2.46 ms at 256 threads; 1024 threads: a 64 KB
// summary at 8 waves per SIMD, 1.79 ms).

// giant-summary words in LDS of the filtered unions: 16 KB per 256 threads
// (8 resident blocks of 256 threads, 4 of 512 or 2 of 1024 use 128 KB of
// the 160 KB)
template <int B>
constexpr uint32_t gsum_cc_words() {
  return B >= 1024 ? 16384u : B >= 512 ? 8192u : 4096u;
}

__global__ __launch_bounds__(CBLK) void k_cc_init(uint32_t n, uint32_t *par,
                                                  ull *forest) {
  for (uint64_t i = (uint64_t)blockIdx.x * CBLK + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * CBLK) {
    par[i] = (uint32_t)i;
    if (forest) forest[i] = KEY_NONE;
  }
}

// CHECK (and check at run time): device-scope read of a root before its
// compare-and-swap (uf_hook_roots)
template <bool CHECK>
__device__ __forceinline__ void cc_edge(ull e, uint32_t n, uint32_t *par,
                                        ull *forest, bool check = true) {
  const uint32_t a = (uint32_t)(e >> 32), b = (uint32_t)e;
  if (a < b && b < n) {
    uint32_t h = uf_union<CHECK>(par, a, b, check);
    if (forest && h != U32_NONE) forest[h] = e;
  }
}

// FILTER: edges whose endpoints are both in the giant component of the
// sampling phase (bitmap `giant`) are skipped without touching par.
// gcount: how many of the GSAMP sampled vertices the giant holds.
template <bool FILTER, int B, int CEPT>
__global__ __launch_bounds__(B) void k_cc_union(const ull *__restrict__ edges,
                                                uint64_t m, uint32_t n,
                                                uint32_t *par, ull *forest,
                                                const uint32_t *__restrict__ giant,
                                                const uint32_t *gsum,
                                                const uint32_t *gcount) {
  __shared__ uint32_t s_sum[FILTER ? gsum_cc_words<B>() : 1];
  uint32_t covered = 0;
  if (FILTER) covered = gsum_load<B, gsum_cc_words<B>()>(s_sum, gsum, n);
  // Checked hooks (uf_hook_roots) where many records become unions: the
  // unfiltered union, which builds trees from single vertices in a CC
  // solve's sample phase (0.70 -> 0.66 ms at RMAT n = 2^24, m = 2^28;
  // 2.1-6.1 -> 0.84-0.91 ms with the vertex ids permuted; streaming
  // batches on a resident parent array unchanged), and filtered passes whose
  // giant holds under 7/8 of the sampled vertices, so that by 1 - (7/8)^2
  // about a quarter or more of the records (with even degrees) are not
  // skipped: the 4096^2 grid's pass keeps all its 10 M records (1.36 ->
  // 0.79 ms), a uniform graph's first pass a third.  Behind a giant holding
  // nearly all of the sample (RMAT: 4-9 % of the records kept) the hooks are
  // few and the check measured 0.09 ms slower.
  const bool check = !FILTER || *gcount < GSAMP * 7 / 8;
  constexpr int CTILE = B * CEPT;
  for (uint64_t tile = blockIdx.x; tile * CTILE < m; tile += gridDim.x) {
    const uint64_t base = tile * CTILE + threadIdx.x;
    ull e[CEPT];
#pragma unroll
    for (int j = 0; j < CEPT; j++) {
      uint64_t i = base + (uint64_t)j * B;
      e[j] = i < m ? __builtin_nontemporal_load(&edges[i]) : KEY_NONE;
    }
    bool ok[CEPT];
#pragma unroll
    for (int j = 0; j < CEPT; j++) {
      const uint32_t a = (uint32_t)(e[j] >> 32), b = (uint32_t)e[j];
      ok[j] = a < b && b < n;
      // b's test only when a is in the giant component (measured faster
      // than issuing both bitmap reads together: 2.09 against 2.22 ms of
      // filtered unions at RMAT n = 2^24, m = 2^28)
      if (FILTER && ok[j])
        ok[j] = !(giant_test(a, s_sum, covered, giant) &&
                  giant_test(b, s_sum, covered, giant));
    }
    // Fast path: the parents of both endpoints of all edges are loaded
    // together (independent loads in flight); an edge whose endpoints share a
    // parent is inside one tree already.  A parent read here may be older than
    // the current one, but every value it ever held is an ancestor, so equal
    // parents still prove one tree.
    uint32_t pa[CEPT], pb[CEPT];
#pragma unroll
    for (int j = 0; j < CEPT; j++) {
      const uint32_t a = (uint32_t)(e[j] >> 32), b = (uint32_t)e[j];
      pa[j] = ok[j] ? par[a] : 0u;
      pb[j] = ok[j] ? par[b] : 0u;
    }
#pragma unroll
    for (int j = 0; j < CEPT; j++)
      if (pa[j] != pb[j]) cc_edge<true>(e[j], n, par, forest, check);
  }
}

__global__ __launch_bounds__(CBLK) void k_cc_labels(uint32_t n,
                                                    const uint32_t *__restrict__ par,
                                                    uint32_t *__restrict__ labels) {
  for (uint64_t i = (uint64_t)blockIdx.x * CBLK + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * CBLK)
    labels[i] = uf_root(par, (uint32_t)i);
}

// Union over the 2n candidates forest[0..n) ++ extra[0..n)
// (union_find_gpu, CC/GPUSpanningTree.cu:74-96).
__global__ __launch_bounds__(CBLK) void k_cc_merge(uint32_t n,
                                                   const ull *__restrict__ f1,
                                                   const ull *__restrict__ extra,
                                                   uint32_t *par, ull *f2) {
  const uint64_t tot = 2ull * n;
  for (uint64_t i = (uint64_t)blockIdx.x * CBLK + threadIdx.x; i < tot;
       i += (uint64_t)gridDim.x * CBLK) {
    ull e = i < n ? f1[i] : extra[i - n];
    if (e != KEY_NONE) cc_edge<true>(e, n, par, f2);  // a fresh parent array
  }
}

}  // namespace phem

namespace phem {

// Launches the union of records [0, m) of `edges` with the knobs' shape.
template <bool FILTER>
static void launch_cc_union(const ull *edges, uint64_t m, uint32_t n,
                            uint32_t *par, ull *forest, const uint32_t *giant,
                            const uint32_t *gsum, const uint32_t *gcount,
                            hipStream_t s) {
  const KernelShape k = kernel_shape();
  const uint32_t B = FILTER ? k.cc_block_filter : k.cc_block;
  const uint32_t E = FILTER ? k.cc_ept_filter : k.cc_ept;
  const unsigned grid = grid_for_threads(m, (uint64_t)B * E, B);
#define PHEM_CC_UNION(BB, EE)                                                  \
  if (B == BB && E == EE) {                                                    \
    hipLaunchKernelGGL((k_cc_union<FILTER, BB, EE>), dim3(grid), dim3(BB), 0, \
                       s, edges, m, n, par, forest, giant, gsum, gcount);      \
    return;                                                                    \
  }
  if constexpr (FILTER) {
    PHEM_CC_UNION(256, 4) PHEM_CC_UNION(256, 8) PHEM_CC_UNION(256, 16)
    PHEM_CC_UNION(512, 4) PHEM_CC_UNION(512, 8) PHEM_CC_UNION(512, 16)
    PHEM_CC_UNION(1024, 4) PHEM_CC_UNION(1024, 8)
  } else {
    PHEM_CC_UNION(256, 8) PHEM_CC_UNION(256, 16) PHEM_CC_UNION(256, 32)
    PHEM_CC_UNION(512, 8) PHEM_CC_UNION(512, 16) PHEM_CC_UNION(512, 32)
  }
#undef PHEM_CC_UNION
  // phem_set_tuning admits no other shape; 256 threads as a backstop
  hipLaunchKernelGGL((k_cc_union<FILTER, 256, FILTER ? 8 : 16>),
                     dim3(grid_for(m, 256ull * (FILTER ? 8 : 16))), dim3(256),
                     0, s, edges, m, n, par, forest, giant, gsum, gcount);
}

}  // namespace phem

using namespace phem;

extern "C" int phem_cc_init(uint32_t n, uint32_t *d_parent, uint64_t *d_forest,
                            void *stream) {
  if (n > PHEM_MAX_VERTICES) return set_err(PHEM_EINVAL, "n=%u too large", n);
  if (n && !d_parent) return set_err(PHEM_EINVAL, "d_parent is NULL");
  if (!n) return PHEM_OK;
  {
    PHEM_PROF("k_cc_init", (hipStream_t)stream);
    hipLaunchKernelGGL(k_cc_init, dim3(grid_for(n, CBLK)), dim3(CBLK), 0,
                       (hipStream_t)stream, n, d_parent, (ull *)d_forest);
  }
  PHEM_LAUNCHED("k_cc_init");
  return PHEM_OK;
}

extern "C" int phem_cc_union(uint32_t n, const uint64_t *d_edges, uint64_t m,
                             uint32_t *d_parent, uint64_t *d_forest,
                             void *stream) {
  if (n > PHEM_MAX_VERTICES) return set_err(PHEM_EINVAL, "n=%u too large", n);
  if (m && (!d_edges || !d_parent))
    return set_err(PHEM_EINVAL, "edges or parent is NULL");
  if (!m || !n) return PHEM_OK;
  {
    PHEM_PROF("k_cc_union", (hipStream_t)stream);
    launch_cc_union<false>((const ull *)d_edges, m, n, d_parent, (ull *)d_forest,
                           nullptr, nullptr, nullptr, (hipStream_t)stream);
  }
  PHEM_LAUNCHED("k_cc_union");
  return PHEM_OK;
}

extern "C" int phem_cc_labels(uint32_t n, const uint32_t *d_parent,
                              uint32_t *d_labels, void *stream) {
  if (n && (!d_parent || !d_labels))
    return set_err(PHEM_EINVAL, "parent or labels is NULL");
  if (!n) return PHEM_OK;
  {
    PHEM_PROF("k_cc_labels", (hipStream_t)stream);
    hipLaunchKernelGGL(k_cc_labels, dim3(grid_for(n, CBLK)), dim3(CBLK), 0,
                       (hipStream_t)stream, n, d_parent, d_labels);
  }
  PHEM_LAUNCHED("k_cc_labels");
  return PHEM_OK;
}

// Sampling phases when m is well above n (the sampling idea of Afforest,
// Sutton et al., IPDPS 2018): union the first m1 = f*n records (f: the
// cc_sample_factor knob), flatten the parent forest and mark its giant
// component; union the records up to 3*m1, skipping those with both
// endpoints in the giant component; mark the giant again, now that it has
// grown; union the remaining records the same way.  d_labels doubles as
// scratch for the bitmap and the giant root before receiving the labels.
extern "C" int phem_cc_solve(uint32_t n, const uint64_t *d_edges, uint64_t m,
                             uint32_t *d_parent, uint64_t *d_forest,
                             uint32_t *d_labels, void *stream) {
  int rc = phem_cc_init(n, d_parent, d_forest, stream);
  if (rc) return rc;
  const double f = cc_default_sample_factor();
  const uint64_t m1 = f > 0 ? (uint64_t)(f * (double)n) : m;
  if (!d_labels || m1 >= m || n < 64) {
    rc = phem_cc_union(n, d_edges, m, d_parent, d_forest, stream);
    if (rc) return rc;
    return d_labels ? phem_cc_labels(n, d_parent, d_labels, stream) : PHEM_OK;
  }
  hipStream_t s = (hipStream_t)stream;
  uint32_t *giant_root = d_labels;       // 1 word
  uint32_t *giant_count = d_labels + 1;  // 1 word
  uint32_t *bits = d_labels + 2;         // ceil(n / 32) words
  // summary: gsum_words <= ceil(n / 1024) words; all fit in n for n >= 64
  uint32_t *gsum = bits + (n + 31) / 32;
  rc = phem_cc_union(n, d_edges, m1, d_parent, d_forest, stream);
  if (rc) return rc;
  // flatten, pick the giant root, giant bitmap and its LDS summary
  auto mark = [&]() -> int {
    {
      PHEM_PROF("k_flatten", s);
      hipLaunchKernelGGL(k_flatten, dim3(grid_for(n, SPLIT_BLK)),
                         dim3(SPLIT_BLK), 0, s, n, d_parent);
    }
    PHEM_LAUNCHED("k_flatten");
    {
      PHEM_PROF("k_pick_giant", s);
      hipLaunchKernelGGL(k_pick_giant, dim3(1), dim3(SBLK), 0, s, n,
                         (const uint32_t *)d_parent, giant_root, giant_count);
    }
    PHEM_LAUNCHED("k_pick_giant");
    {
      PHEM_PROF("k_giant_bits", s);
      hipLaunchKernelGGL(k_giant_bits, dim3(grid_for(n, SPLIT_BLK)),
                         dim3(SPLIT_BLK), 0, s, n, (const uint32_t *)d_parent,
                         (const uint32_t *)giant_root, bits);
    }
    PHEM_LAUNCHED("k_giant_bits");
    {
      PHEM_PROF("k_giant_summary", s);
      hipLaunchKernelGGL(k_giant_summary,
                         dim3(grid_for(32ull * gsum_words(n, GSUM_MAX), SPLIT_BLK)),
                         dim3(SPLIT_BLK), 0, s, n, (const uint32_t *)bits, gsum);
    }
    PHEM_LAUNCHED("k_giant_summary");
    return PHEM_OK;
  };
  // records [a0, a1), skipping those inside the marked giant component
  auto filtered = [&](uint64_t a0, uint64_t a1) -> int {
    if (a1 <= a0) return PHEM_OK;
    {
      PHEM_PROF("k_cc_union_filtered", s);
      launch_cc_union<true>((const ull *)d_edges + a0, a1 - a0, n, d_parent,
                            (ull *)d_forest, bits, gsum, giant_count, s);
    }
    PHEM_LAUNCHED("k_cc_union_filtered");
    return PHEM_OK;
  };
  const uint64_t m2 = m1 < m / 3 ? 3 * m1 : m;
  if ((rc = mark()) || (rc = filtered(m1, m2))) return rc;
  if (m2 < m && ((rc = mark()) || (rc = filtered(m2, m)))) return rc;
  return phem_cc_labels(n, d_parent, d_labels, stream);
}

extern "C" int phem_cc_merge_forests(uint32_t n, const uint64_t *d_forest,
                                     const uint64_t *d_extra,
                                     uint32_t *d_parent, uint64_t *d_forest2,
                                     uint32_t *d_labels, void *stream) {
  if (n && (!d_forest || !d_extra || !d_parent))
    return set_err(PHEM_EINVAL, "forest, extra or parent is NULL");
  int rc = phem_cc_init(n, d_parent, d_forest2, stream);
  if (rc || !n) return rc;
  {
    PHEM_PROF("k_cc_merge", (hipStream_t)stream);
    hipLaunchKernelGGL(k_cc_merge, dim3(grid_for(2ull * n, CBLK)), dim3(CBLK), 0,
                       (hipStream_t)stream, n, (const ull *)d_forest,
                       (const ull *)d_extra, d_parent, (ull *)d_forest2);
  }
  PHEM_LAUNCHED("k_cc_merge");
  return d_labels ? phem_cc_labels(n, d_parent, d_labels, stream) : PHEM_OK;
}
