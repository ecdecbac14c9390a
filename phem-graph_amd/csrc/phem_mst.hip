// phem_mst.hip -- Boruvka minimum spanning forest for gfx950.
//
// Round structure (the reference's device_boruvka round loop,
// MST/src/kernels.cu:507-689, re-planned for HBM traffic):
//
//   round 0   k_sweep0     : every input edge (u, v, w; id implicit) offers
//                            key = (wkey(w) << 32) | id to best[u] and best[v]
//                            (populate_min_weight + populate_eid,
//                            kernels.cu:24-70, as one 64-bit atomicMin).
//   round r   k_hook       : every active component c with best[c] = key of
//                            edge e marks e in the forest bitmap and unions
//                            the endpoints of e (new_grafting_edges,
//                            kernels.cu:73-134); the larger root is hooked
//                            under the smaller, so every root is the minimum
//                            vertex of its component.
//             k_compress   : every component active in round r is pointed at
//                            its root (new_shortcutting_step,
//                            kernels.cu:254-271) and the roots form the next
//                            round's component list.
//             k_compact    : one pass over the alive edges relabels both
//                            endpoints to roots, drops edges inside one
//                            component, writes the survivors (u', v', key)
//                            through an LDS staging tile and offers each
//                            survivor's key to best[] for round r + 1
//                            (relabelling_step + filtering_step + the two
//                            remove_if passes, kernels.cu:273-321,641-649,
//                            fused with the next round's sweep).
//
// The packed key makes every component's choice unique, so the forest is the
// unique minimum spanning forest under (weight, id) order; compaction need not
// be stable because ids travel with the edges.
#include <cstring>

#include "phem_dev.h"
#include "phem_internal.h"

namespace phem {

constexpr int BLK = 256;       // threads per block (4 waves)
constexpr int EPT = 8;         // edges per thread per tile
constexpr int TILE = BLK * EPT;
constexpr int WPT = 8;         // bitmap words per thread
constexpr int WPB = BLK * WPT; // bitmap words per block
constexpr int IPT = 4;         // components per thread in hook / compress

// counter slots (ull) in the workspace
constexpr int C_VALID = 0;   // valid input edges (m_0)
constexpr int C_WSUM = 1;    // exact weight sum
constexpr int C_INEXACT = 2; // weights off the 2^-24 grid (F32)
constexpr int C_MSTCNT = 3;  // forest size
constexpr int C_ROUND = 8;   // + 4 r: [E(r), A(r), L(r), -]
constexpr int NCTR = C_ROUND + 4 * (PHEM_MAX_ROUNDS + 1);

__host__ __device__ constexpr int cE(int r) { return C_ROUND + 4 * r + 0; }
__host__ __device__ constexpr int cA(int r) { return C_ROUND + 4 * r + 1; }
__host__ __device__ constexpr int cL(int r) { return C_ROUND + 4 * r + 2; }

// ---------------------------------------------------------------------------
// Input edge sources: structure-of-arrays (u[], v[], w[]) or the reference's
// array of `edge {u, v, w}` records (MST/src/constants.h:15-19).

struct SrcSoA {
  const uint32_t *u, *v, *w;
  __device__ __forceinline__ uint32_t U(uint64_t i) const { return u[i]; }
  __device__ __forceinline__ uint32_t V(uint64_t i) const { return v[i]; }
  __device__ __forceinline__ uint32_t W(uint64_t i) const { return w[i]; }
  __device__ __forceinline__ void stream_load(uint64_t i, uint32_t &a,
                                              uint32_t &b, uint32_t &c) const {
    a = __builtin_nontemporal_load(&u[i]);
    b = __builtin_nontemporal_load(&v[i]);
    c = __builtin_nontemporal_load(&w[i]);
  }
};

struct SrcAoS {
  const uint32_t *e;  // 3 words per edge
  __device__ __forceinline__ uint32_t U(uint64_t i) const { return e[3 * i]; }
  __device__ __forceinline__ uint32_t V(uint64_t i) const { return e[3 * i + 1]; }
  __device__ __forceinline__ uint32_t W(uint64_t i) const { return e[3 * i + 2]; }
  __device__ __forceinline__ void stream_load(uint64_t i, uint32_t &a,
                                              uint32_t &b, uint32_t &c) const {
    a = __builtin_nontemporal_load(&e[3 * i]);
    b = __builtin_nontemporal_load(&e[3 * i + 1]);
    c = __builtin_nontemporal_load(&e[3 * i + 2]);
  }
};

// ---------------------------------------------------------------------------

__global__ __launch_bounds__(BLK) void k_mst_init(uint32_t n, uint32_t *par,
                                                  ull *best, uint64_t words,
                                                  uint32_t *bitmap, ull *ctr) {
  uint64_t lim = n;
  if (words > lim) lim = words;
  if ((uint64_t)NCTR > lim) lim = NCTR;
  for (uint64_t i = (uint64_t)blockIdx.x * BLK + threadIdx.x; i < lim;
       i += (uint64_t)gridDim.x * BLK) {
    if (i < n) {
      par[i] = (uint32_t)i;
      best[i] = KEY_NONE;
    }
    if (i < words) bitmap[i] = 0u;
    if (i < (uint64_t)NCTR) ctr[i] = 0ull;
  }
}

// Round-0 sweep over the input edge arrays.
template <class Src>
__global__ __launch_bounds__(BLK) void k_sweep0(Src src, uint64_t m, uint32_t n,
                                                int wmode, ull *best) {
  for (uint64_t tile = blockIdx.x; tile * TILE < m; tile += gridDim.x) {
    const uint64_t base = tile * TILE + threadIdx.x;
    uint32_t a[EPT], b[EPT], ww[EPT];
#pragma unroll
    for (int j = 0; j < EPT; j++) {
      uint64_t i = base + (uint64_t)j * BLK;
      if (i < m) {
        src.stream_load(i, a[j], b[j], ww[j]);
      } else {
        a[j] = b[j] = U32_NONE;
        ww[j] = 0;
      }
    }
#pragma unroll
    for (int j = 0; j < EPT; j++) {
      uint64_t i = base + (uint64_t)j * BLK;
      if (a[j] < n && b[j] < n && a[j] != b[j]) {
        uint64_t key = ((uint64_t)wkey(ww[j], wmode) << 32) | i;
        min_store(best, a[j], key);
        min_store(best, b[j], key);
      }
    }
  }
}

// Hook: components of `list` (or 0..n_ident-1 when list is NULL).
template <class Src>
__global__ __launch_bounds__(BLK) void k_hook(
    const uint32_t *__restrict__ list, const ull *list_count, uint32_t n_ident,
    ull *best, Src src, uint32_t *par, uint32_t *bitmap, uint32_t *act,
    ull *act_count) {
  const uint64_t cnt = list ? *list_count : (uint64_t)n_ident;
  for (uint64_t base = (uint64_t)blockIdx.x * (BLK * IPT); base < cnt;
       base += (uint64_t)gridDim.x * (BLK * IPT)) {
    bool active[IPT];
    uint32_t c[IPT];
    uint64_t slot[IPT];
#pragma unroll
    for (int j = 0; j < IPT; j++) {
      const uint64_t idx = base + (uint64_t)j * BLK + threadIdx.x;
      active[j] = false;
      c[j] = 0;
      if (idx < cnt) {
        c[j] = list ? list[idx] : (uint32_t)idx;
        ull key = best[c[j]];
        if (key != KEY_NONE) {
          active[j] = true;
          best[c[j]] = KEY_NONE;  // reset for the next round's sweep
          uint32_t e = (uint32_t)key;
          atomicOr(&bitmap[e >> 5], 1u << (e & 31));
          uf_union(par, src.U(e), src.V(e));
        }
      }
    }
    block_append<BLK, IPT>(active, slot, act_count);
#pragma unroll
    for (int j = 0; j < IPT; j++)
      if (active[j]) act[slot[j]] = c[j];
  }
}

// Point every active component at its root; collect the roots.
__global__ __launch_bounds__(BLK) void k_compress(const uint32_t *__restrict__ act,
                                                  const ull *act_count,
                                                  uint32_t *par, uint32_t *next,
                                                  ull *next_count) {
  const uint64_t cnt = *act_count;
  for (uint64_t base = (uint64_t)blockIdx.x * (BLK * IPT); base < cnt;
       base += (uint64_t)gridDim.x * (BLK * IPT)) {
    bool root[IPT];
    uint32_t c[IPT];
    uint64_t slot[IPT];
#pragma unroll
    for (int j = 0; j < IPT; j++) {
      const uint64_t idx = base + (uint64_t)j * BLK + threadIdx.x;
      root[j] = false;
      c[j] = 0;
      if (idx < cnt) {
        c[j] = act[idx];
        uint32_t r = uf_root(par, c[j]);
        if (r != c[j]) par[c[j]] = r;
        root[j] = (r == c[j]);
      }
    }
    block_append<BLK, IPT>(root, slot, next_count);
#pragma unroll
    for (int j = 0; j < IPT; j++)
      if (root[j]) next[slot[j]] = c[j];
  }
}

// Relabel + drop intra-component edges + compact + next-round sweep.
// R0: the input is the original (u, v, w) arrays, keys are formed from w and
// the index, and invalid edges are dropped.
template <bool R0, class Src>
__global__ __launch_bounds__(BLK) void k_compact(
    Src src, const uint32_t *__restrict__ ia, const uint32_t *__restrict__ ib,
    const ull *__restrict__ ik, const ull *in_count, uint64_t m0, uint32_t n, int wmode,
    const uint32_t *__restrict__ par, uint32_t *__restrict__ oa,
    uint32_t *__restrict__ ob, ull *__restrict__ ok, ull *out_count,
    ull *valid_count, ull *best) {
  __shared__ uint32_t s_a[TILE];
  __shared__ uint32_t s_b[TILE];
  __shared__ ull s_k[TILE];
  __shared__ uint32_t s_cnt, s_valid;
  __shared__ ull s_base;
  const uint64_t mr = R0 ? m0 : *in_count;
  const uint32_t lane = lane_id();
  for (uint64_t tile = blockIdx.x; tile * TILE < mr; tile += gridDim.x) {
    if (threadIdx.x == 0) {
      s_cnt = 0;
      s_valid = 0;
    }
    __syncthreads();
    const uint64_t base = tile * TILE + threadIdx.x;
    uint32_t a[EPT], b[EPT];
    ull key[EPT];
    bool val[EPT];
#pragma unroll
    for (int j = 0; j < EPT; j++) {
      uint64_t i = base + (uint64_t)j * BLK;
      val[j] = false;
      a[j] = b[j] = 0;
      key[j] = KEY_NONE;
      if (i < mr) {
        if (R0) {
          uint32_t wv;
          src.stream_load(i, a[j], b[j], wv);
          key[j] = ((ull)wkey(wv, wmode) << 32) | i;
          val[j] = a[j] < n && b[j] < n && a[j] != b[j];
        } else {
          a[j] = __builtin_nontemporal_load(&ia[i]);
          b[j] = __builtin_nontemporal_load(&ib[i]);
          key[j] = __builtin_nontemporal_load(&ik[i]);
          val[j] = true;
        }
      }
    }
#pragma unroll
    for (int j = 0; j < EPT; j++) {
      if (val[j]) {
        a[j] = par[a[j]];
        b[j] = par[b[j]];
      }
    }
#pragma unroll
    for (int j = 0; j < EPT; j++) {
      const bool keep = val[j] && a[j] != b[j];
      const uint64_t mask = __ballot(keep);
      const uint64_t vmask = R0 ? __ballot(val[j]) : 0ull;
      uint32_t wbase = 0;
      if (lane == 0) {
        if (mask) wbase = atomicAdd(&s_cnt, (uint32_t)__popcll(mask));
        if (R0 && vmask) atomicAdd(&s_valid, (uint32_t)__popcll(vmask));
      }
      wbase = __shfl(wbase, 0);
      if (keep) {
        uint32_t pos = wbase + lanes_below(mask);
        s_a[pos] = a[j];
        s_b[pos] = b[j];
        s_k[pos] = key[j];
        min_store(best, a[j], key[j]);
        min_store(best, b[j], key[j]);
      }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      s_base = s_cnt ? atomicAdd(out_count, (ull)s_cnt) : 0ull;
      if (R0 && s_valid) atomicAdd(valid_count, (ull)s_valid);
    }
    __syncthreads();
    const uint32_t cnt = s_cnt;
    const ull ob_base = s_base;
    for (uint32_t k = threadIdx.x; k < cnt; k += BLK) {
      oa[ob_base + k] = s_a[k];
      ob[ob_base + k] = s_b[k];
      ok[ob_base + k] = s_k[k];
    }
    __syncthreads();
  }
}

__global__ __launch_bounds__(BLK) void k_labels(uint32_t n,
                                                const uint32_t *__restrict__ par,
                                                uint32_t *__restrict__ labels) {
  for (uint64_t i = (uint64_t)blockIdx.x * BLK + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * BLK)
    labels[i] = uf_root(par, (uint32_t)i);
}

// ---- block scan helpers (BLK threads, wave64) ------------------------------

__device__ __forceinline__ ull wave_incl_scan(ull x) {
  const uint32_t lane = lane_id();
#pragma unroll
  for (int d = 1; d < WAVE; d <<= 1) {
    ull y = __shfl_up(x, d);
    if (lane >= (uint32_t)d) x += y;
  }
  return x;
}

// Exclusive scan over the block; *total receives the block sum.
__device__ __forceinline__ ull block_excl_scan(ull x, ull *total) {
  __shared__ ull s_w[BLK / WAVE];
  const uint32_t lane = lane_id(), wv = threadIdx.x / WAVE;
  ull inc = wave_incl_scan(x);
  if (lane == WAVE - 1) s_w[wv] = inc;
  __syncthreads();
  ull off = 0, tot = 0;
#pragma unroll
  for (int k = 0; k < BLK / WAVE; k++) {
    if ((uint32_t)k < wv) off += s_w[k];
    tot += s_w[k];
  }
  __syncthreads();
  *total = tot;
  return off + inc - x;
}

__global__ __launch_bounds__(BLK) void k_bm_count(const uint32_t *__restrict__ bm,
                                                  uint64_t words, ull *blk) {
  const uint64_t w0 = (uint64_t)blockIdx.x * WPB + (uint64_t)threadIdx.x * WPT;
  ull c = 0;
#pragma unroll
  for (int k = 0; k < WPT; k++)
    if (w0 + k < words) c += __popc(bm[w0 + k]);
  ull tot;
  block_excl_scan(c, &tot);
  if (threadIdx.x == 0) blk[blockIdx.x] = tot;
}

// Exclusive scan of blk[0..nb) by one block; total -> *sum.
__global__ __launch_bounds__(BLK) void k_scan_blocks(ull *blk, uint64_t nb,
                                                     ull *sum) {
  ull carry = 0;
  for (uint64_t base = 0; base < nb; base += BLK) {
    uint64_t i = base + threadIdx.x;
    ull x = i < nb ? blk[i] : 0ull;
    ull tot;
    ull ex = block_excl_scan(x, &tot);
    if (i < nb) blk[i] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0) *sum = carry;
}

__device__ __forceinline__ bool weight_fx(uint32_t bits, int wmode, ull *out) {
  if (wmode == 0) {
    *out = bits;
    return true;
  }
  float f = __builtin_bit_cast(float, bits);
  if (!(f >= 0.0f)) return false;
  float s = f * 16777216.0f;  // exact scaling by 2^24
  if (!(s < 1.8446744e19f)) return false;
  if (truncf(s) != s) return false;
  *out = (ull)s;
  return true;
}

// Bitmap -> ascending ids; exact weight sum of the forest.
template <class Src>
__global__ __launch_bounds__(BLK) void k_bm_emit(const uint32_t *__restrict__ bm,
                                                 uint64_t words,
                                                 const ull *__restrict__ blk,
                                                 uint32_t *__restrict__ ids,
                                                 Src src, int wmode, ull *wsum,
                                                 ull *inexact) {
  const uint64_t w0 = (uint64_t)blockIdx.x * WPB + (uint64_t)threadIdx.x * WPT;
  uint32_t word[WPT];
  ull c = 0;
#pragma unroll
  for (int k = 0; k < WPT; k++) {
    word[k] = (w0 + k < words) ? bm[w0 + k] : 0u;
    c += __popc(word[k]);
  }
  ull tot;
  ull pos = blk[blockIdx.x] + block_excl_scan(c, &tot);
  ull s = 0, bad = 0;
#pragma unroll
  for (int k = 0; k < WPT; k++) {
    uint32_t x = word[k];
    while (x) {
      uint32_t bit = __ffs(x) - 1;
      x &= x - 1;
      uint32_t e = (uint32_t)((w0 + k) * 32 + bit);
      ids[pos++] = e;
      ull fx;
      if (weight_fx(src.W(e), wmode, &fx)) s += fx; else bad++;
    }
  }
  ull ts, tb;
  block_excl_scan(s, &ts);
  block_excl_scan(bad, &tb);
  if (threadIdx.x == 0) {
    if (ts) atomicAdd(wsum, ts);
    if (tb) atomicAdd(inexact, tb);
  }
}

// ---------------------------------------------------------------------------
// workspace layout

struct MstLayout {
  uint64_t par, best, listL, listA, ea_a, ea_b, ea_k, eb_a, eb_b, eb_k, bitmap,
      ctr, blk, total;
  uint64_t words, nb;
};

static MstLayout mst_layout(uint32_t n, uint64_t m) {
  MstLayout L;
  const uint64_t A = 256;
  uint64_t o = 0;
  auto take = [&](uint64_t bytes) {
    uint64_t at = o;
    o = align_up(o + (bytes ? bytes : 1), A);
    return at;
  };
  L.words = ceil_div(m ? m : 1, 32);
  L.nb = ceil_div(L.words, WPB);
  L.par = take(4ull * n);
  L.best = take(8ull * n);
  L.listL = take(4ull * n);
  L.listA = take(4ull * n);
  L.ea_a = take(4ull * m);
  L.ea_b = take(4ull * m);
  L.ea_k = take(8ull * m);
  L.eb_a = take(4ull * m);
  L.eb_b = take(4ull * m);
  L.eb_k = take(8ull * m);
  L.bitmap = take(4ull * L.words);
  L.ctr = take(8ull * NCTR);
  L.blk = take(8ull * L.nb);
  L.total = o;
  return L;
}

}  // namespace phem

using namespace phem;

extern "C" int phem_mst_workspace_bytes(uint32_t n, uint64_t m,
                                        uint64_t *bytes) {
  if (!bytes) return set_err(PHEM_EINVAL, "bytes is NULL");
  if (m > PHEM_MAX_EDGES || n > PHEM_MAX_VERTICES)
    return set_err(PHEM_EINVAL, "n=%u m=%llu beyond limits", n,
                   (unsigned long long)m);
  *bytes = mst_layout(n, m).total;
  return PHEM_OK;
}

namespace phem {

template <class Src>
static int mst_solve_impl(uint32_t n, uint64_t m, Src src, int wmode,
                          void *d_ws, uint64_t ws_bytes, uint32_t *d_mst_ids,
                          uint32_t *d_labels, phem_mst_stats *stats,
                          hipStream_t s) {
  if (m > PHEM_MAX_EDGES || n > PHEM_MAX_VERTICES)
    return set_err(PHEM_EINVAL, "n=%u m=%llu beyond limits", n,
                   (unsigned long long)m);
  if (wmode != PHEM_W_U32 && wmode != PHEM_W_F32)
    return set_err(PHEM_EINVAL, "unknown weight mode %d", wmode);
  if (n && !d_mst_ids) return set_err(PHEM_EINVAL, "d_mst_ids is NULL");
  const MstLayout L = mst_layout(n, m);
  if (!d_ws || ws_bytes < L.total)
    return set_err(PHEM_ENOMEM, "workspace %llu B < required %llu B",
                   (unsigned long long)ws_bytes, (unsigned long long)L.total);
  char *ws = (char *)d_ws;
  uint32_t *par = (uint32_t *)(ws + L.par);
  ull *best = (ull *)(ws + L.best);
  uint32_t *listL = (uint32_t *)(ws + L.listL);
  uint32_t *listA = (uint32_t *)(ws + L.listA);
  uint32_t *Ea[2] = {(uint32_t *)(ws + L.ea_a), (uint32_t *)(ws + L.eb_a)};
  uint32_t *Eb[2] = {(uint32_t *)(ws + L.ea_b), (uint32_t *)(ws + L.eb_b)};
  ull *Ek[2] = {(ull *)(ws + L.ea_k), (ull *)(ws + L.eb_k)};
  uint32_t *bitmap = (uint32_t *)(ws + L.bitmap);
  ull *ctr = (ull *)(ws + L.ctr);
  ull *blk = (ull *)(ws + L.blk);

  if (stats) std::memset(stats, 0, sizeof(*stats));

  uint64_t initw = n;
  if (L.words > initw) initw = L.words;
  if ((uint64_t)NCTR > initw) initw = NCTR;
  {
    PHEM_PROF("k_mst_init", s);
    hipLaunchKernelGGL(k_mst_init, dim3(grid_for(initw, BLK)), dim3(BLK), 0, s,
                       n, par, best, L.words, bitmap, ctr);
  }
  PHEM_LAUNCHED("k_mst_init");

  int rounds = 0;
  if (m > 0 && n > 1) {
    // round 0: sweep, hook and compress over all vertices, compact the input
    {
      PHEM_PROF("k_sweep0", s);
      hipLaunchKernelGGL(k_sweep0<Src>, dim3(grid_for(m, TILE)), dim3(BLK), 0,
                         s, src, m, n, wmode, best);
    }
    PHEM_LAUNCHED("k_sweep0");
    {
      PHEM_PROF("k_hook", s);
      hipLaunchKernelGGL(k_hook<Src>, dim3(grid_for(n, BLK * IPT)), dim3(BLK),
                         0, s, (const uint32_t *)nullptr, (const ull *)nullptr,
                         n, best, src, par, bitmap, listA, ctr + cA(0));
    }
    PHEM_LAUNCHED("k_hook");
    {
      PHEM_PROF("k_compress", s);
      hipLaunchKernelGGL(k_compress, dim3(grid_for(n, BLK * IPT)), dim3(BLK), 0,
                         s, listA, ctr + cA(0), par, listL, ctr + cL(1));
    }
    PHEM_LAUNCHED("k_compress");
    {
      PHEM_PROF("k_compact<R0>", s);
      hipLaunchKernelGGL((k_compact<true, Src>), dim3(grid_for(m, TILE)),
                         dim3(BLK), 0, s, src, (const uint32_t *)nullptr,
                         (const uint32_t *)nullptr, (const ull *)nullptr,
                         (const ull *)nullptr, m, n, wmode, par, Ea[0], Eb[0],
                         Ek[0], ctr + cE(1), ctr + C_VALID, best);
    }
    PHEM_LAUNCHED("k_compact<R0>");
    ull h[16];
    PHEM_HIP(hipMemcpyAsync(h, ctr, 16 * sizeof(ull), hipMemcpyDeviceToHost, s));
    PHEM_HIP(hipStreamSynchronize(s));
    uint64_t m_r = h[cE(1)], list_ub = h[cL(1)];
    if (stats) {
      stats->m_log[0] = h[C_VALID];
      stats->c_log[0] = h[cA(0)];
    }
    // a round counts when it starts with at least one alive edge
    rounds = h[C_VALID] ? 1 : 0;
    int cur = 0;
    while (m_r > 0) {
      if (rounds >= PHEM_MAX_ROUNDS)
        return set_err(PHEM_ENOCONV, "no convergence after %d rounds", rounds);
      const int r = rounds;
      {
        PHEM_PROF("k_hook", s);
        hipLaunchKernelGGL(k_hook<Src>, dim3(grid_for(list_ub, BLK * IPT)),
                           dim3(BLK), 0, s, listL, ctr + cL(r), 0u, best, src,
                           par, bitmap, listA, ctr + cA(r));
      }
      PHEM_LAUNCHED("k_hook");
      {
        PHEM_PROF("k_compress", s);
        hipLaunchKernelGGL(k_compress, dim3(grid_for(list_ub, BLK * IPT)),
                           dim3(BLK), 0, s, listA, ctr + cA(r), par, listL,
                           ctr + cL(r + 1));
      }
      PHEM_LAUNCHED("k_compress");
      {
        PHEM_PROF("k_compact", s);
        hipLaunchKernelGGL((k_compact<false, Src>), dim3(grid_for(m_r, TILE)),
                           dim3(BLK), 0, s, src, Ea[cur], Eb[cur], Ek[cur],
                           ctr + cE(r), 0ull, n, wmode, par, Ea[cur ^ 1],
                           Eb[cur ^ 1], Ek[cur ^ 1], ctr + cE(r + 1),
                           (ull *)nullptr, best);
      }
      PHEM_LAUNCHED("k_compact");
      ull hr[8];
      PHEM_HIP(hipMemcpyAsync(hr, ctr + cE(r), 8 * sizeof(ull),
                              hipMemcpyDeviceToHost, s));
      PHEM_HIP(hipStreamSynchronize(s));
      if (stats) {
        stats->m_log[r] = hr[0];
        stats->c_log[r] = hr[1];
      }
      m_r = hr[4];      // E(r + 1)
      list_ub = hr[6];  // L(r + 1)
      cur ^= 1;
      rounds++;
    }
  }

  if (d_labels && n) {
    {
      PHEM_PROF("k_labels", s);
      hipLaunchKernelGGL(k_labels, dim3(grid_for(n, BLK)), dim3(BLK), 0, s, n,
                         par, d_labels);
    }
    PHEM_LAUNCHED("k_labels");
  }
  {
    PHEM_PROF("k_bm_count", s);
    hipLaunchKernelGGL(k_bm_count, dim3((unsigned)L.nb), dim3(BLK), 0, s,
                       bitmap, L.words, blk);
  }
  PHEM_LAUNCHED("k_bm_count");
  {
    PHEM_PROF("k_scan_blocks", s);
    hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(BLK), 0, s, blk, L.nb,
                       ctr + C_MSTCNT);
  }
  PHEM_LAUNCHED("k_scan_blocks");
  {
    PHEM_PROF("k_bm_emit", s);
    hipLaunchKernelGGL(k_bm_emit<Src>, dim3((unsigned)L.nb), dim3(BLK), 0, s,
                       bitmap, L.words, blk, d_mst_ids, src, wmode,
                       ctr + C_WSUM, ctr + C_INEXACT);
  }
  PHEM_LAUNCHED("k_bm_emit");
  ull tail[4];
  PHEM_HIP(hipMemcpyAsync(tail, ctr, 4 * sizeof(ull), hipMemcpyDeviceToHost, s));
  PHEM_HIP(hipStreamSynchronize(s));
  if (stats) {
    stats->rounds = (uint32_t)rounds;
    stats->mst_count = (uint32_t)tail[C_MSTCNT];
    stats->weight_fx = tail[C_WSUM];
    stats->weight_exact = tail[C_INEXACT] == 0;
    for (int r = 0; r < rounds; r++) {
      stats->sum_m += stats->m_log[r];
      stats->sum_c += stats->c_log[r];
    }
  }
  return PHEM_OK;
}

}  // namespace phem

extern "C" int phem_mst_solve(uint32_t n, uint64_t m, const uint32_t *d_u,
                              const uint32_t *d_v, const uint32_t *d_w,
                              int wmode, void *d_ws, uint64_t ws_bytes,
                              uint32_t *d_mst_ids, uint32_t *d_labels,
                              phem_mst_stats *stats, void *stream) {
  if (m && (!d_u || !d_v || !d_w))
    return set_err(PHEM_EINVAL, "edge arrays are NULL");
  return mst_solve_impl(n, m, SrcSoA{d_u, d_v, d_w}, wmode, d_ws, ws_bytes,
                        d_mst_ids, d_labels, stats, (hipStream_t)stream);
}

extern "C" int phem_mst_solve_aos(uint32_t n, uint64_t m,
                                  const uint32_t *d_edges, int wmode,
                                  void *d_ws, uint64_t ws_bytes,
                                  uint32_t *d_mst_ids, uint32_t *d_labels,
                                  phem_mst_stats *stats, void *stream) {
  if (m && !d_edges) return set_err(PHEM_EINVAL, "edge array is NULL");
  return mst_solve_impl(n, m, SrcAoS{d_edges}, wmode, d_ws, ws_bytes, d_mst_ids,
                        d_labels, stats, (hipStream_t)stream);
}
