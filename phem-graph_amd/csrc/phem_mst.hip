// phem_mst.hip -- Boruvka minimum spanning forest for gfx950.
//
// Round structure (the reference's device_boruvka round loop,
// MST/src/kernels.cu:507-689, re-planned for HBM traffic):
//
//   k_hook     : every active component c with best[c] = key of edge e marks
//                e in the forest bitmap and unions the endpoints of e
//                (new_grafting_edges, kernels.cu:73-134); the larger root is
//                hooked under the smaller, so every root is the minimum
//                vertex of its component.
//   k_compress : every component active in the round is pointed at its root
//                (new_shortcutting_step, kernels.cu:254-271) and the roots
//                form the next round's component list.
//   k_pass     : one pass over the alive edges relabels both endpoints to
//                roots, drops edges inside one component, writes the
//                survivors (u', v', key) to the tile's slots of the output
//                segment (one atomic per tile) and offers each survivor's
//                key = (wkey(w) << 32) | id to best[] of both components
//                for the next hook (relabelling_step +
//                filtering_step + the two remove_if passes,
//                kernels.cu:273-321,641-649, fused with the next round's
//                populate_min_weight / populate_eid sweep, kernels.cu:24-70).
//
// Two schedules:
//   canonical  k_sweep0 over the input, then rounds over all edges; the
//              per-round log equals the reference round loop's.
//   split      (default) filter-Kruskal in front of the rounds, as the CPU
//              path does (MST/src/utils.cu:794-841): a sample of the keys
//              fixes a threshold tau; rounds run over the edges with key <
//              tau; the parent forest is flattened and its giant component
//              marked in a bitmap; one pass keeps the heavier edges whose
//              endpoints are still in different components (an edge between
//              a component and the giant only offers its key to that
//              component's minimum: one edge per component comes back, and
//              a component whose lightest edge it is gets hooked to the
//              giant at once); rounds finish those.  It runs when the heavy
//              edges are at least half as many as the light ones; a batch's
//              graphs (disjoint union) each get their own threshold and
//              giant component.
// A whole solve can also be recorded into a HIP graph with a fixed number of
// rounds per phase and no counter read (phem_mst_graph_*).
// The packed key makes every component's choice unique, so both give the
// unique minimum spanning forest under (weight, id) order; compaction need
// not be stable because ids travel with the edges.
//
// Every appended list (alive edges, component lists) is segmented: tile v of
// a producer appends to segment v mod KSEG, whose counter has its own cache
// line, so no global atomic is shared by all tiles.
#include <atomic>
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <initializer_list>
#include <memory>
#include <mutex>
#include <new>
#include <type_traits>
#include <utility>

#include "phem_dev.h"
#include "phem_internal.h"
#include "phem_split.h"

namespace phem {

constexpr int BLK = 256;       // threads per block (4 waves)
constexpr int EPT = 8;         // edges per thread per tile (input passes)
constexpr int TILE = BLK * EPT;
// The input passes (light and heavy filter) run with mst_block threads and
// mst_ept edges per thread (512 x 4 built in); the compaction rounds of
// graphs with m >= 2^mst_compact_log2 with mst_ept_compact edges per thread
// (16 built in), smaller graphs with EPT.  Segment capacities are sized for
// the largest tile any shape can use on the graph: 4096 edges for graphs of
// at least 2^22 edges (WIDE_LOG2), 2048 below, where a larger shape runs
// with its block size and 2048 edges per tile.  So the workspace does not
// depend on the knobs, and small graphs keep small segments.
constexpr int MAX_EDGE_TILE = 4096;  // 256 x 16 or 512 x 8 edges
constexpr int WIDE_LOG2 = 22;
static inline uint64_t max_edge_tile(uint64_t m) {
  return m >= (1ull << WIDE_LOG2) ? MAX_EDGE_TILE : TILE;
}
constexpr int WPT = 8;         // bitmap words per thread
constexpr int WPB = BLK * WPT; // bitmap words per block
constexpr int IPT = 4;         // vertices or components per thread (compress,
                               // list collection, giant emission)
constexpr int LTILE = BLK * IPT;
constexpr int HPT = 1;         // components per thread in the hook
constexpr int HTILE = BLK * HPT;
constexpr int SPEC = 3;  // Boruvka rounds launched per counter read
constexpr uint32_t GSUM_HEAVY = 8192;  // giant-summary words in LDS (32 KB)

// Counter area (ull), every counter on its own 128-byte line (CSTR ull):
// scalars, then four blocks of KSEG partial sums (valid edges, weight sum,
// inexact weights, edges emitted by k_giant_emit), then per round three segmented counters: E(r) alive edges
// at the start of round r, A(r) components active in round r, L(r) component
// list for round r.
constexpr int C_MSTCNT = 0 * CSTR;  // forest size
constexpr int C_TAU = 1 * CSTR;     // light / heavy key threshold
constexpr int C_GIANT = 2 * CSTR;   // giant-component root (32-bit)
constexpr int C_SUMS = 3 * CSTR;    // k_seg_sums output (up to 16 values)
constexpr int C_GATE = 4 * CSTR;    // fixed rounds: light phase not converged
constexpr int C_PARTS = 5 * CSTR;
constexpr int C_ROUND = C_PARTS + 4 * KSEG * CSTR;
constexpr int RSTRIDE = 3 * KSEG * CSTR;
constexpr int NCTR = C_ROUND + RSTRIDE * (PHEM_MAX_ROUNDS + 1);

__host__ __device__ constexpr int cE(int r) { return C_ROUND + RSTRIDE * r; }
__host__ __device__ constexpr int cA(int r) { return cE(r) + KSEG * CSTR; }
__host__ __device__ constexpr int cL(int r) { return cE(r) + 2 * KSEG * CSTR; }
// which: 0 valid edges, 1 weight sum, 2 inexact weights, 3 giant edges
__host__ __device__ constexpr int cPart(int which) {
  return C_PARTS + which * KSEG * CSTR;
}

__device__ __forceinline__ uint64_t ceil_div_dev(uint64_t a, uint64_t b) {
  return (a + b - 1) / b;
}
__device__ __forceinline__ uint64_t min_u64(uint64_t a, uint64_t b) {
  return a < b ? a : b;
}

// ---------------------------------------------------------------------------
// Input edge sources: structure-of-arrays (u[], v[], w[]) or the reference's
// array of `edge {u, v, w}` records (MST/src/constants.h:15-19).  load4 reads
// the four consecutive edges i4 .. i4+3 below `end` (with 128-bit loads when
// the arrays are 16-byte aligned and all four are in range).

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned long long u64x2 __attribute__((ext_vector_type(2)));

struct SrcSoA {
  const uint32_t *u, *v, *w;
  bool vec;
  __device__ __forceinline__ uint32_t U(uint64_t i) const { return u[i]; }
  __device__ __forceinline__ uint32_t V(uint64_t i) const { return v[i]; }
  __device__ __forceinline__ uint32_t W(uint64_t i) const { return w[i]; }
  // the four edges i4 .. i4+3, all in range, arrays 16-byte aligned
  __device__ __forceinline__ void load4_full(uint64_t i4, uint32_t *a,
                                             uint32_t *b, uint32_t *c) const {
    u32x4 x = __builtin_nontemporal_load((const u32x4 *)(u + i4));
    u32x4 y = __builtin_nontemporal_load((const u32x4 *)(v + i4));
    u32x4 z = __builtin_nontemporal_load((const u32x4 *)(w + i4));
    a[0] = x.x; a[1] = x.y; a[2] = x.z; a[3] = x.w;
    b[0] = y.x; b[1] = y.y; b[2] = y.z; b[3] = y.w;
    c[0] = z.x; c[1] = z.y; c[2] = z.z; c[3] = z.w;
  }
  __device__ __forceinline__ void load4(uint64_t i4, uint64_t end, uint32_t *a,
                                        uint32_t *b, uint32_t *c) const {
    if (vec && i4 + 4 <= end) {
      u32x4 x = __builtin_nontemporal_load((const u32x4 *)(u + i4));
      u32x4 y = __builtin_nontemporal_load((const u32x4 *)(v + i4));
      u32x4 z = __builtin_nontemporal_load((const u32x4 *)(w + i4));
      a[0] = x.x; a[1] = x.y; a[2] = x.z; a[3] = x.w;
      b[0] = y.x; b[1] = y.y; b[2] = y.z; b[3] = y.w;
      c[0] = z.x; c[1] = z.y; c[2] = z.z; c[3] = z.w;
    } else {
#pragma unroll
      for (int q = 0; q < 4; q++) {
        if (i4 + q < end) {
          a[q] = __builtin_nontemporal_load(&u[i4 + q]);
          b[q] = __builtin_nontemporal_load(&v[i4 + q]);
          c[q] = __builtin_nontemporal_load(&w[i4 + q]);
        } else {
          a[q] = b[q] = U32_NONE;
          c[q] = 0;
        }
      }
    }
  }
};

struct SrcAoS {
  const uint32_t *e;  // 3 words per edge
  bool vec;
  __device__ __forceinline__ uint32_t U(uint64_t i) const { return e[3 * i]; }
  __device__ __forceinline__ uint32_t V(uint64_t i) const { return e[3 * i + 1]; }
  __device__ __forceinline__ uint32_t W(uint64_t i) const { return e[3 * i + 2]; }
  // the four edges i4 .. i4+3, all in range, array 16-byte aligned
  __device__ __forceinline__ void load4_full(uint64_t i4, uint32_t *a,
                                             uint32_t *b, uint32_t *c) const {
    const u32x4 *p = (const u32x4 *)(e + 3 * i4);
    u32x4 x = __builtin_nontemporal_load(p);
    u32x4 y = __builtin_nontemporal_load(p + 1);
    u32x4 z = __builtin_nontemporal_load(p + 2);
    a[0] = x.x; b[0] = x.y; c[0] = x.z; a[1] = x.w;
    b[1] = y.x; c[1] = y.y; a[2] = y.z; b[2] = y.w;
    c[2] = z.x; a[3] = z.y; b[3] = z.z; c[3] = z.w;
  }
  __device__ __forceinline__ void load4(uint64_t i4, uint64_t end, uint32_t *a,
                                        uint32_t *b, uint32_t *c) const {
    if (vec && i4 + 4 <= end) {
      const u32x4 *p = (const u32x4 *)(e + 3 * i4);
      u32x4 x = __builtin_nontemporal_load(p);
      u32x4 y = __builtin_nontemporal_load(p + 1);
      u32x4 z = __builtin_nontemporal_load(p + 2);
      a[0] = x.x; b[0] = x.y; c[0] = x.z; a[1] = x.w;
      b[1] = y.x; c[1] = y.y; a[2] = y.z; b[2] = y.w;
      c[2] = z.x; a[3] = z.y; b[3] = z.z; c[3] = z.w;
    } else {
#pragma unroll
      for (int q = 0; q < 4; q++) {
        if (i4 + q < end) {
          a[q] = __builtin_nontemporal_load(&e[3 * (i4 + q)]);
          b[q] = __builtin_nontemporal_load(&e[3 * (i4 + q) + 1]);
          c[q] = __builtin_nontemporal_load(&e[3 * (i4 + q) + 2]);
        } else {
          a[q] = b[q] = U32_NONE;
          c[q] = 0;
        }
      }
    }
  }
};

// ---------------------------------------------------------------------------

__global__ __launch_bounds__(BLK) void k_mst_init(uint32_t n, uint32_t *par,
                                                  ull *best, ull *gmin,
                                                  uint64_t words,
                                                  uint32_t *bitmap, ull *ctr) {
  uint64_t lim = n;
  if (words > lim) lim = words;
  if ((uint64_t)NCTR > lim) lim = NCTR;
  for (uint64_t i = (uint64_t)blockIdx.x * BLK + threadIdx.x; i < lim;
       i += (uint64_t)gridDim.x * BLK) {
    if (i < n) {
      par[i] = (uint32_t)i;
      best[i] = KEY_NONE;
      if (gmin) gmin[i] = KEY_NONE;
    }
    if (i < words) bitmap[i] = 0u;
    if (i < (uint64_t)NCTR) ctr[i] = 0ull;
  }
}

// Round-0 sweep over the input edges (canonical schedule).
template <class Src>
__global__ __launch_bounds__(BLK) void k_sweep0(Src src, uint64_t m, uint32_t n,
                                                int ord, ull *best) {
  for (uint64_t tile = blockIdx.x; tile * TILE < m; tile += gridDim.x) {
    uint32_t a[EPT], b[EPT], ww[EPT];
#pragma unroll
    for (int g = 0; g < EPT / 4; g++)
      src.load4(tile * TILE + ((uint64_t)g * BLK + threadIdx.x) * 4, m,
                a + 4 * g, b + 4 * g, ww + 4 * g);
#pragma unroll
    for (int j = 0; j < EPT; j++) {
      const uint64_t i =
          tile * TILE + ((uint64_t)(j / 4) * BLK + threadIdx.x) * 4 + (j % 4);
      if (a[j] < n && b[j] < n && a[j] != b[j]) {
        uint64_t key = make_key(ww[j], i, ord);
        min_store(best, a[j], key);
        min_store(best, b[j], key);
      }
    }
  }
}

// Hook: the components of the segmented list (or 0..n_ident-1 when the list
// counter is NULL).  Active components go to the segmented list `act`.
template <class Src>
__global__ __launch_bounds__(BLK) void k_hook(
    const uint32_t *__restrict__ list, const ull *list_cnt, uint64_t list_cap,
    uint32_t n_ident, ull *best, Src src, int ord, uint32_t *par,
    uint32_t *bitmap, uint32_t *act, ull *act_cnt, uint64_t act_cap,
    ull *wsum_parts, ull *inexact_parts) {
  __shared__ SegView sv;
  if (list_cnt) seg_load<HTILE>(sv, list_cnt);
  // exact weight of the forest edges this thread hooks (each forest edge is
  // hooked exactly once: keys are unique and the chosen edges form a forest)
  ull wsum = 0, wbad = 0;
  const uint32_t ntiles_ident = (uint32_t)ceil_div_dev(n_ident, HTILE);
  for (uint32_t v = blockIdx.x;; v += gridDim.x) {
    uint64_t start;
    uint32_t len;
    if (list_cnt) {
      if (!seg_tile<HTILE>(sv, v, list_cap, &start, &len)) break;
    } else {
      if (v >= ntiles_ident) break;
      start = (uint64_t)v * HTILE;
      len = (uint32_t)min_u64(HTILE, (uint64_t)n_ident - start);
    }
    // Each step's loads (list, best, the chosen edges' endpoints, the bitmap
    // words, the endpoints' parents) are issued for all HPT components
    // before any is used.
    bool active[HPT];
    uint32_t c[HPT];
    ull key[HPT];
    uint64_t slot[HPT];
#pragma unroll
    for (int j = 0; j < HPT; j++) {
      const uint32_t idx = (uint32_t)j * BLK + threadIdx.x;
      c[j] = idx < len ? (list_cnt ? list[start + idx] : (uint32_t)(start + idx))
                       : U32_NONE;
    }
#pragma unroll
    for (int j = 0; j < HPT; j++) {
      key[j] = best[c[j] != U32_NONE ? c[j] : 0u];  // unconditional read
      if (c[j] == U32_NONE) key[j] = KEY_NONE;
    }
    uint32_t e[HPT], eu[HPT], ev[HPT], old[HPT];
#pragma unroll
    for (int j = 0; j < HPT; j++) {
      active[j] = key[j] != KEY_NONE;
      e[j] = active[j] ? key_id(key[j], ord) : 0u;
      if (active[j]) best[c[j]] = KEY_NONE;  // reset for the next round's sweep
    }
#pragma unroll
    for (int j = 0; j < HPT; j++) {
      eu[j] = src.U(e[j]);  // e = 0 for inactive entries
      ev[j] = src.V(e[j]);
    }
    // a component whose edge was already marked (both sides chose the same
    // edge) adds no weight: only the atomicOr that sets the bit does
#pragma unroll
    for (int j = 0; j < HPT; j++)
      old[j] = active[j] ? atomicOr(&bitmap[e[j] >> 5], 1u << (e[j] & 31)) : ~0u;
#pragma unroll
    for (int j = 0; j < HPT; j++) {
      if (!((old[j] >> (e[j] & 31)) & 1u)) {
        ull fx;
        if (weight_fx(key_weight(key[j], ord), ord & ORD_F32, &fx)) wsum += fx;
        else wbad++;
      }
    }
    uint32_t pu[HPT], pv[HPT];
#pragma unroll
    for (int j = 0; j < HPT; j++) {
      pu[j] = par_load(&par[active[j] ? eu[j] : 0u]);
      pv[j] = par_load(&par[active[j] ? ev[j] : 0u]);
    }
#pragma unroll
    for (int j = 0; j < HPT; j++)
      if (active[j]) uf_union_from(par, eu[j], pu[j], ev[j], pv[j]);
    const uint32_t seg = v % KSEG;
    block_append<BLK, HPT>(active, slot, act_cnt + seg * CSTR,
                           (uint64_t)seg * act_cap);
#pragma unroll
    for (int j = 0; j < HPT; j++)
      if (active[j]) act[slot[j]] = c[j];
  }
  __shared__ ull s_w[2][BLK / WAVE];
#pragma unroll
  for (int d = 1; d < WAVE; d <<= 1) {
    wsum += __shfl_xor(wsum, d);
    wbad += __shfl_xor(wbad, d);
  }
  if (lane_id() == 0) {
    s_w[0][threadIdx.x / WAVE] = wsum;
    s_w[1][threadIdx.x / WAVE] = wbad;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    ull a = 0, b = 0;
    for (int w = 0; w < BLK / WAVE; w++) {
      a += s_w[0][w];
      b += s_w[1][w];
    }
    const uint32_t seg = blockIdx.x % KSEG;
    if (a) atomicAdd(wsum_parts + seg * CSTR, a);
    if (b) atomicAdd(inexact_parts + seg * CSTR, b);
  }
}

// Clears the KSEG counters of a segmented list (one thread per counter).
__global__ __launch_bounds__(KSEG) void k_zero_seg(ull *cnt) {
  cnt[threadIdx.x * CSTR] = 0ull;
}

// The vertices whose best[] holds an offer, into a segmented list: the first
// hook of a phase then runs over active components only (a third or less of
// the vertices at the start of either phase) instead of all n.
__global__ __launch_bounds__(BLK) void k_collect_active(
    uint32_t n, const ull *__restrict__ best, uint32_t *__restrict__ list,
    ull *list_cnt, uint64_t list_cap) {
  const uint32_t ntiles = (uint32_t)ceil_div_dev(n, LTILE);
  for (uint32_t v = blockIdx.x; v < ntiles; v += gridDim.x) {
    bool on[IPT];
    uint64_t slot[IPT];
#pragma unroll
    for (int j = 0; j < IPT; j++) {
      const uint64_t x = (uint64_t)v * LTILE + (uint64_t)j * BLK + threadIdx.x;
      on[j] = x < n && best[x] != KEY_NONE;
    }
    const uint32_t seg = v % KSEG;
    block_append<BLK, IPT>(on, slot, list_cnt + seg * CSTR,
                           (uint64_t)seg * list_cap);
#pragma unroll
    for (int j = 0; j < IPT; j++)
      if (on[j]) list[slot[j]] = (uint32_t)((uint64_t)v * LTILE + (uint64_t)j * BLK + threadIdx.x);
  }
}

// Point every active component at its root; collect the roots.
__global__ __launch_bounds__(BLK) void k_compress(
    const uint32_t *__restrict__ act, const ull *act_cnt, uint64_t act_cap,
    uint32_t *par, uint32_t *next, ull *next_cnt, uint64_t next_cap) {
  __shared__ SegView sv;
  seg_load<LTILE>(sv, act_cnt);
  for (uint32_t v = blockIdx.x;; v += gridDim.x) {
    uint64_t start;
    uint32_t len;
    if (!seg_tile<LTILE>(sv, v, act_cap, &start, &len)) break;
    bool root[IPT];
    uint32_t c[IPT];
    uint64_t slot[IPT];
#pragma unroll
    for (int j = 0; j < IPT; j++) {
      const uint32_t idx = (uint32_t)j * BLK + threadIdx.x;
      root[j] = false;
      c[j] = 0;
      if (idx < len) {
        c[j] = act[start + idx];
        uint32_t r = uf_root(par, c[j]);
        if (r != c[j]) par[c[j]] = r;
        root[j] = (r == c[j]);
      }
    }
    const uint32_t seg = v % KSEG;
    block_append<BLK, IPT>(root, slot, next_cnt + seg * CSTR,
                           (uint64_t)seg * next_cap);
#pragma unroll
    for (int j = 0; j < IPT; j++)
      if (root[j]) next[slot[j]] = c[j];
  }
}

// Edge passes.  Every pass reads a stream of edges, decides which survive,
// writes the survivors (component u, component v, key) straight from
// registers to the tile's slots of its output segment (one atomic per tile,
// on the segment's counter) and offers each survivor's key to best[] of both
// components.
//   P_ALL   input edges, relabelled through par (canonical round 0)
//   P_LIGHT input edges with key < tau, not relabelled (par is the identity)
//   P_HEAVY input edges with key >= tau whose endpoints are not both in the
//           giant component (bitmap) and have different roots (par flattened)
//   P_EDGES segmented alive edges of the previous round, relabelled
enum PassMode : int { P_ALL = 0, P_LIGHT = 1, P_HEAVY = 2, P_EDGES = 3 };

// Light thresholds per graph of a batch (taus[g] for the edges of graph g),
// or taus == NULL: one threshold for all edges.
struct EdgeTaus {
  const ull *taus, *eoff;
  const uint32_t *emap;
  uint32_t k;
};

template <int MODE, class Src, int NE_ = EPT, int B_ = BLK>
__global__ __launch_bounds__(B_) void k_pass(
    Src src, const uint32_t *__restrict__ ia, const uint32_t *__restrict__ ib,
    const ull *__restrict__ ik, const ull *in_cnt, uint64_t in_cap,
    uint64_t m0, uint32_t n, int ord, const uint32_t *__restrict__ par,
    const ull *tau_ptr, const uint32_t *__restrict__ giant,
    uint32_t *__restrict__ oa, uint32_t *__restrict__ ob,
    ull *__restrict__ ok, ull *out_cnt, uint64_t out_cap, ull *valid_parts,
    ull *best, ull *gmin, const uint32_t *gsum, EdgeTaus et, int two_step) {
  // tile shape: the compaction rounds of large graphs measured faster with
  // 16 edges per thread (1.43 against 1.52 ms of rounds at n = 2^24,
  // m = 2^28), the input passes with EPT (8; 16 made k_light and k_heavy
  // 0.15-0.4 ms slower)
  constexpr int NE = NE_;
  constexpr int KT = B_ * NE;
  __shared__ SegView sv;
  __shared__ uint32_t s_sum[MODE == P_HEAVY ? GSUM_HEAVY : 1];
  constexpr bool INPUT = MODE != P_EDGES;
  if (!INPUT) seg_load<KT>(sv, in_cnt);
  uint32_t covered = 0;
  if (MODE == P_HEAVY) covered = gsum_load<B_, GSUM_HEAVY>(s_sum, gsum, n);
  const ull tau = (MODE == P_LIGHT || MODE == P_HEAVY) ? *tau_ptr : 0ull;
  const uint32_t ntiles_in = INPUT ? (uint32_t)ceil_div_dev(m0, KT) : 0u;
  uint32_t nval = 0;  // valid input edges seen by this thread
  for (uint32_t v = blockIdx.x;; v += gridDim.x) {
    uint64_t start;
    uint32_t len;
    if (INPUT) {
      if (v >= ntiles_in) break;
      start = (uint64_t)v * KT;
      len = (uint32_t)min_u64(KT, m0 - start);
    } else {
      if (!seg_tile<KT>(sv, v, in_cap, &start, &len)) break;
    }
    const uint64_t end = start + len;
    uint32_t a[NE], b[NE];
    ull key[NE];
    bool val[NE];
    // A full tile (block-uniform) issues all its loads unconditionally, so
    // they are in flight together; a partial tile takes the guarded path.
    const bool full = len == (uint32_t)KT && (!INPUT || src.vec);
    if (INPUT) {
      uint32_t wv[NE];
      if (full) {
#pragma unroll
        for (int g = 0; g < NE / 4; g++)
          src.load4_full(start + ((uint64_t)g * B_ + threadIdx.x) * 4, a + 4 * g,
                         b + 4 * g, wv + 4 * g);
      } else {
#pragma unroll
        for (int g = 0; g < NE / 4; g++)
          src.load4(start + ((uint64_t)g * B_ + threadIdx.x) * 4, end, a + 4 * g,
                    b + 4 * g, wv + 4 * g);
      }
#pragma unroll
      for (int j = 0; j < NE; j++) {
        const uint64_t i = start + ((uint64_t)(j / 4) * B_ + threadIdx.x) * 4 + (j % 4);
        key[j] = make_key(wv[j], i, ord);
        val[j] = (a[j] < n) & (b[j] < n) & (a[j] != b[j]);
      }
    } else if (full) {
#pragma unroll
      for (int g = 0; g < NE / 4; g++) {
        const uint64_t i4 = start + ((uint64_t)g * B_ + threadIdx.x) * 4;
        u32x4 x = __builtin_nontemporal_load((const u32x4 *)(ia + i4));
        u32x4 y = __builtin_nontemporal_load((const u32x4 *)(ib + i4));
        u64x2 k0 = __builtin_nontemporal_load((const u64x2 *)(ik + i4));
        u64x2 k1 = __builtin_nontemporal_load((const u64x2 *)(ik + i4 + 2));
        a[4 * g] = x.x; a[4 * g + 1] = x.y; a[4 * g + 2] = x.z; a[4 * g + 3] = x.w;
        b[4 * g] = y.x; b[4 * g + 1] = y.y; b[4 * g + 2] = y.z; b[4 * g + 3] = y.w;
        key[4 * g] = k0.x; key[4 * g + 1] = k0.y;
        key[4 * g + 2] = k1.x; key[4 * g + 3] = k1.y;
      }
#pragma unroll
      for (int j = 0; j < NE; j++) val[j] = true;
    } else {
#pragma unroll
      for (int j = 0; j < NE; j++) {
        const uint64_t i = start + ((uint64_t)(j / 4) * B_ + threadIdx.x) * 4 + (j % 4);
        const bool in = i < end;
        a[j] = in ? ia[i] : 0u;
        b[j] = in ? ib[i] : 0u;
        key[j] = in ? ik[i] : KEY_NONE;
        val[j] = in;
      }
    }
    if (INPUT) {
#pragma unroll
      for (int j = 0; j < NE; j++) nval += val[j];
    }
    // the threshold of edge j (its graph's, for a batch): one lookup per
    // tile when the whole tile lies in one graph (graph ids grow with the
    // edge index, so the first and last edge decide)
    ull ttau = tau;
    bool one_graph = true;
    if ((MODE == P_LIGHT || MODE == P_HEAVY) && et.taus) {
      const uint32_t g0 = graph_of(et.emap, et.eoff, et.k, start);
      one_graph = g0 == graph_of(et.emap, et.eoff, et.k, end - 1);
      ttau = et.taus[g0];
    }
    auto tau_of = [&](int j) -> ull {
      if (one_graph) return ttau;
      const uint64_t i = start + ((uint64_t)(j / 4) * B_ + threadIdx.x) * 4 + (j % 4);
      return et.taus[graph_of(et.emap, et.eoff, et.k, i < end ? i : start)];
    };
    if (MODE == P_LIGHT) {
#pragma unroll
      for (int j = 0; j < NE; j++) val[j] = val[j] && key[j] < tau_of(j);
    }
    if (MODE == P_HEAVY) {
      // both endpoints in the giant component: inside one component.  One
      // endpoint x in it: of the parallel edges between x's component and
      // the giant only the lightest can be in the forest (cycle property),
      // so the edge is only offered to gmin[x]; k_giant_fold moves each
      // vertex's minimum to its root and k_giant_emit turns each
      // component's minimum into one alive edge.  The bitmap words of the
      // endpoints the LDS summary leaves open are read together.
      bool ga[NE], gb[NE];
      uint32_t wa[NE], wb[NE];
#pragma unroll
      for (int j = 0; j < NE; j++) {
        val[j] = val[j] && key[j] >= tau_of(j);
        ga[j] = val[j] && giant_summary_hit(a[j], s_sum, covered);
        gb[j] = val[j] && giant_summary_hit(b[j], s_sum, covered);
      }
      // unconditional reads (word 0 stands in for settled endpoints)
#pragma unroll
      for (int j = 0; j < NE; j++) {
        wa[j] = giant[val[j] && !ga[j] ? a[j] >> 5 : 0u];
        wb[j] = giant[val[j] && !gb[j] ? b[j] >> 5 : 0u];
      }
      bool one[NE];
      uint32_t x[NE];
#pragma unroll
      for (int j = 0; j < NE; j++) {
        ga[j] = val[j] && (ga[j] || ((wa[j] >> (a[j] & 31)) & 1u));
        gb[j] = val[j] && (gb[j] || ((wb[j] >> (b[j] & 31)) & 1u));
        one[j] = val[j] && ga[j] != gb[j];
        x[j] = ga[j] ? b[j] : a[j];
        val[j] = val[j] && !(ga[j] || gb[j]);
      }
      min_store_n<NE, true>(gmin, x, key, one);
    }
    if (MODE == P_EDGES) {
      // alive edges of late rounds name a few hot components; one edge's
      // lookups at a time measured faster here than the batched form below
#pragma unroll
      for (int j = 0; j < NE; j++) {
        if (val[j]) {
          a[j] = par[a[j]];
          b[j] = par[b[j]];
          if (two_step) {  // see run_rounds: one more step reaches the root
            a[j] = par[a[j]];
            b[j] = par[b[j]];
          }
        }
      }
    } else if (MODE != P_LIGHT) {
      uint32_t pa[NE], pb[NE];
#pragma unroll
      for (int j = 0; j < NE; j++) {
        pa[j] = par[val[j] ? a[j] : 0u];  // unconditional, as above
        pb[j] = par[val[j] ? b[j] : 0u];
      }
#pragma unroll
      for (int j = 0; j < NE; j++) {
        a[j] = pa[j];
        b[j] = pb[j];
      }
    }
    bool keep[NE];
#pragma unroll
    for (int j = 0; j < NE; j++) keep[j] = val[j] && a[j] != b[j];
    if (MODE == P_LIGHT || MODE == P_HEAVY) {
      // input edges spread their offers over many components: the
      // pre-check reads of the tile are issued together
      // (tie-checked: RMAT 2^24 / 2^28 with integer weights 1..16 took 5.26
      // ms in the light pass without, 1.11 ms with; float weights 1.08 ms
      // and 1.06 ms)
      min_store_ends<NE, true>(best, a, b, key, keep);
    } else {
      // relabelled edges concentrate on few components in late rounds: each
      // pre-check follows the thread's previous atomic, so a component's
      // lowered minimum filters the offers after it
#pragma unroll
      for (int j = 0; j < NE; j++) {
        if (keep[j]) {
          min_store(best, a[j], key[j]);
          min_store(best, b[j], key[j]);
        }
      }
    }
    // survivors go straight from registers to their slots in the tile's
    // output segment (one atomic per tile)
    const uint32_t seg = v % KSEG;
    uint64_t slot[NE];
    block_append<B_, NE>(keep, slot, out_cnt + seg * CSTR,
                           (uint64_t)seg * out_cap);
#pragma unroll
    for (int j = 0; j < NE; j++) {
      if (keep[j]) {
        oa[slot[j]] = a[j];
        ob[slot[j]] = b[j];
        ok[slot[j]] = key[j];
      }
    }
  }
  if (INPUT && valid_parts) {
    __shared__ ull s_w[B_ / WAVE];
    ull x = nval;
#pragma unroll
    for (int d = 1; d < WAVE; d <<= 1) x += __shfl_xor(x, d);
    if (lane_id() == 0) s_w[threadIdx.x / WAVE] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
      ull tot = 0;
      for (int w = 0; w < B_ / WAVE; w++) tot += s_w[w];
      if (tot) atomicAdd(valid_parts + (blockIdx.x % KSEG) * CSTR, tot);
    }
  }
}

// After the heavy pass: the minimum of every non-root vertex goes to its
// root (par is flat), and the vertex's own entry is cleared.  Only root
// entries are written by other threads, only non-root entries are read.
__global__ __launch_bounds__(BLK) void k_giant_fold(
    uint32_t n, const uint32_t *__restrict__ par, ull *__restrict__ gmin) {
  for (uint64_t x = (uint64_t)blockIdx.x * BLK + threadIdx.x; x < n;
       x += (uint64_t)gridDim.x * BLK) {
    const ull k = gmin[x];
    if (k == KEY_NONE) continue;
    const uint32_t r = par[x];
    if (r != x) {
      min_store(gmin, r, k);
      gmin[x] = KEY_NONE;
    }
  }
}

// After the heavy pass, gmin[x] holds the lightest edge between component x
// and the giant component G.  When nothing lighter reached best[x] in the
// heavy pass, that edge is x's Boruvka edge for the first heavy round; if
// also x > G, x is hooked under G here (union-find, forest bit, exact
// weight) and leaves the round (best[x] cleared; G stays the root, so no
// component list changes).  Otherwise the edge becomes one alive edge
// (x, G, key), offered to best[x].  Every key is offered to best[G] (one
// block minimum).  Tile v appends to segment v mod KSEG of the alive-edge
// list.  A hook here never meets a hook of the same edge by G in the round's
// k_hook: that one finds the forest bit set and adds no weight.  The hooked
// x are in no component list, so the round's compaction relabels through
// two parent steps (x -> G -> root).
struct GiantHook {
  uint32_t *par, *bitmap;
  int ord;
  ull *wsum_parts, *inexact_parts;
};

// Hooks root x > G under G with the edge of key k; adds its weight to wsum /
// wbad.  Within the emit kernel only x's own thread writes par[x], and G
// stays a root (every hook there puts a larger root under G), so a plain
// store is the union.
__device__ __forceinline__ void giant_hook(const GiantHook &gh, ull *best,
                                           uint32_t x, uint32_t G, ull k,
                                           ull &wsum, ull &wbad) {
  const uint32_t e = key_id(k, gh.ord);
  atomicOr(&gh.bitmap[e >> 5], 1u << (e & 31));
  ull fx;
  if (weight_fx(key_weight(k, gh.ord), gh.ord & ORD_F32, &fx)) wsum += fx;
  else wbad++;
  best[x] = KEY_NONE;
  par_store(&gh.par[x], G);
}

// Block sums of the edges appended to the alive list, and of the hooked
// edges' weights and inexact weights, into their KSEG-segment counters
// (every thread of the block calls it).
__device__ __forceinline__ void giant_emit_flush(ull emitted, ull wsum,
                                                 ull wbad, ull *parts,
                                                 const GiantHook &gh) {
  __shared__ ull s_w[3][BLK / WAVE];
#pragma unroll
  for (int d = 1; d < WAVE; d <<= 1) {
    emitted += __shfl_xor(emitted, d);
    wsum += __shfl_xor(wsum, d);
    wbad += __shfl_xor(wbad, d);
  }
  if (lane_id() == 0) {
    s_w[0][threadIdx.x / WAVE] = emitted;
    s_w[1][threadIdx.x / WAVE] = wsum;
    s_w[2][threadIdx.x / WAVE] = wbad;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    ull t[3] = {0, 0, 0};
    for (int w = 0; w < BLK / WAVE; w++)
      for (int q = 0; q < 3; q++) t[q] += s_w[q][w];
    const uint32_t seg = blockIdx.x % KSEG;
    if (t[0]) atomicAdd(parts + seg * CSTR, t[0]);
    if (t[1]) atomicAdd(gh.wsum_parts + seg * CSTR, t[1]);
    if (t[2]) atomicAdd(gh.inexact_parts + seg * CSTR, t[2]);
  }
}

__global__ __launch_bounds__(BLK) void k_giant_emit(
    uint32_t n, ull *__restrict__ gmin, const uint32_t *giant_root,
    uint32_t *__restrict__ oa, uint32_t *__restrict__ ob,
    ull *__restrict__ ok, ull *out_cnt, uint64_t out_cap, ull *best,
    ull *parts, GiantHook gh) {
  const uint32_t G = *giant_root;
  const uint32_t ntiles = (uint32_t)ceil_div_dev(n, LTILE);
  ull emitted = 0, wsum = 0, wbad = 0;  // by this thread
  ull gbest = KEY_NONE;  // lightest emitted key seen by this thread (for G)
  for (uint32_t v = blockIdx.x; v < ntiles; v += gridDim.x) {
    bool keep[IPT], on[IPT];
    ull key[IPT], bx[IPT];
    uint64_t slot[IPT];
#pragma unroll
    for (int j = 0; j < IPT; j++) {
      const uint64_t x = (uint64_t)v * LTILE + (uint64_t)j * BLK + threadIdx.x;
      key[j] = x < n ? gmin[x] : KEY_NONE;
      on[j] = key[j] != KEY_NONE;
    }
#pragma unroll
    for (int j = 0; j < IPT; j++) {
      const uint64_t x = (uint64_t)v * LTILE + (uint64_t)j * BLK + threadIdx.x;
      bx[j] = best[on[j] ? x : 0u];  // unconditional read
    }
#pragma unroll
    for (int j = 0; j < IPT; j++) {
      const uint64_t x = (uint64_t)v * LTILE + (uint64_t)j * BLK + threadIdx.x;
      const bool lightest = key[j] < bx[j];
      keep[j] = on[j] && !(lightest && x > G);
      if (on[j]) {
        gbest = key[j] < gbest ? key[j] : gbest;
        if (!keep[j]) {
          giant_hook(gh, best, (uint32_t)x, G, key[j], wsum, wbad);
        } else {
          emitted++;
          if (lightest) atomicMin(&best[x], key[j]);
        }
      }
    }
    const uint32_t seg = v % KSEG;
    block_append<BLK, IPT>(keep, slot, out_cnt + seg * CSTR,
                           (uint64_t)seg * out_cap);
#pragma unroll
    for (int j = 0; j < IPT; j++) {
      if (keep[j]) {
        const uint64_t x = (uint64_t)v * LTILE + (uint64_t)j * BLK + threadIdx.x;
        oa[slot[j]] = (uint32_t)x;
        ob[slot[j]] = G;
        ok[slot[j]] = key[j];
      }
    }
  }
  // one giant-side offer per block
  __shared__ ull s_g[BLK / WAVE];
#pragma unroll
  for (int d = 1; d < WAVE; d <<= 1) {
    const ull o = __shfl_xor(gbest, d);
    gbest = o < gbest ? o : gbest;
  }
  if (lane_id() == 0) s_g[threadIdx.x / WAVE] = gbest;
  __syncthreads();
  if (threadIdx.x == 0) {
    ull g = KEY_NONE;
    for (int w = 0; w < BLK / WAVE; w++) g = s_g[w] < g ? s_g[w] : g;
    if (g != KEY_NONE) min_store(best, G, g);
  }
  giant_emit_flush(emitted, wsum, wbad, parts, gh);
}

// Per graph g of a batch (vertices [voff[g], voff[g + 1])): the most frequent
// root among up to GSAMP_G evenly spaced vertices of the graph (after
// k_flatten), or U32_NONE for a graph without vertices.  One block per graph.
constexpr int GSAMP_G = 256;
__global__ __launch_bounds__(GSAMP_G) void k_pick_giant_groups(
    uint32_t k, const ull *__restrict__ voff, const uint32_t *par,
    uint32_t *__restrict__ groot) {
  __shared__ ull s[GSAMP_G];
  for (uint32_t g = blockIdx.x; g < k; g += gridDim.x) {
    const ull base = voff[g], ng = voff[g + 1] - base;
    const ull cnt = ng < (ull)GSAMP_G ? ng : (ull)GSAMP_G;
    const uint32_t j = threadIdx.x;  // one sample per thread
    s[j] = j < cnt ? (ull)par[base + (ull)j * ng / cnt] : ~0ull;
    __syncthreads();
    bitonic_sort_lds<GSAMP_G, GSAMP_G>(s);
    const ull r = longest_run<GSAMP_G, GSAMP_G>(s);
    if (threadIdx.x == 0) groot[g] = r == ~0ull ? U32_NONE : (uint32_t)r;
    __syncthreads();  // s is refilled for the next graph
  }
}

// bit x of `bits` = (par[x] == root of the giant of x's graph), for the
// graphs of a batch (par flattened).
__global__ __launch_bounds__(SPLIT_BLK) void k_giant_bits_groups(
    uint32_t n, const uint32_t *par, const ull *__restrict__ voff,
    const uint32_t *__restrict__ vmap, uint32_t k,
    const uint32_t *__restrict__ groot, uint32_t *bits) {
  const uint64_t words = ((uint64_t)n + 31) / 32;
  for (uint64_t base = (uint64_t)blockIdx.x * SPLIT_BLK; base < n;
       base += (uint64_t)gridDim.x * SPLIT_BLK) {
    const uint64_t i = base + threadIdx.x;
    const uint64_t mask =
        __ballot(i < n && par[i] == groot[graph_of(vmap, voff, k, i)]);
    if (lane_id() == 0) {
      const uint64_t w = i >> 5;  // i is a multiple of 64 for lane 0
      if (w < words) bits[w] = (uint32_t)mask;
      if (w + 1 < words) bits[w + 1] = (uint32_t)(mask >> 32);
    }
  }
}

// k_giant_emit for the graphs of a batch: component x's edge goes to the
// giant root of x's own graph (hooked at once, or kept as an alive edge, as
// above).  A tile inside one graph offers the block's minimum to that
// graph's giant once; a tile across graphs offers one key at a time
// (consecutive vertices share a graph, so each pre-check sees the thread's
// previous offer to that giant).
__global__ __launch_bounds__(BLK) void k_giant_emit_groups(
    uint32_t n, ull *__restrict__ gmin, const ull *__restrict__ voff,
    const uint32_t *__restrict__ vmap, uint32_t k,
    const uint32_t *__restrict__ groot, uint32_t *__restrict__ oa,
    uint32_t *__restrict__ ob, ull *__restrict__ ok, ull *out_cnt,
    uint64_t out_cap, ull *best, ull *parts, GiantHook gh) {
  __shared__ ull s_g[BLK / WAVE];
  const uint32_t ntiles = (uint32_t)ceil_div_dev(n, LTILE);
  ull emitted = 0, wsum = 0, wbad = 0;  // by this thread
  for (uint32_t v = blockIdx.x; v < ntiles; v += gridDim.x) {
    const uint64_t x0 = (uint64_t)v * LTILE;
    const uint64_t x1 = x0 + LTILE < n ? x0 + LTILE : n;
    const uint32_t g0 = graph_of(vmap, voff, k, x0);
    const bool one = g0 == graph_of(vmap, voff, k, x1 - 1);  // block-uniform
    bool keep[IPT], on[IPT];
    ull key[IPT], bx[IPT];
    uint32_t G[IPT];
    uint64_t slot[IPT];
#pragma unroll
    for (int j = 0; j < IPT; j++) {
      const uint64_t x = x0 + (uint64_t)j * BLK + threadIdx.x;
      key[j] = x < n ? gmin[x] : KEY_NONE;
      on[j] = key[j] != KEY_NONE;
      G[j] = on[j] ? groot[one ? g0 : graph_of(vmap, voff, k, x)] : 0u;
    }
#pragma unroll
    for (int j = 0; j < IPT; j++) {
      const uint64_t x = x0 + (uint64_t)j * BLK + threadIdx.x;
      bx[j] = best[on[j] ? x : 0u];  // unconditional read
    }
    ull gbest = KEY_NONE;  // this thread's lightest key for the tile's giant
#pragma unroll
    for (int j = 0; j < IPT; j++) {
      const uint64_t x = x0 + (uint64_t)j * BLK + threadIdx.x;
      const bool lightest = key[j] < bx[j];
      keep[j] = on[j] && !(lightest && x > G[j]);
      if (on[j]) {
        if (!keep[j]) {
          giant_hook(gh, best, (uint32_t)x, G[j], key[j], wsum, wbad);
        } else {
          emitted++;
          if (lightest) atomicMin(&best[x], key[j]);
        }
        if (one) gbest = key[j] < gbest ? key[j] : gbest;
        else min_store(best, G[j], key[j]);
      }
    }
    if (one) {
#pragma unroll
      for (int d = 1; d < WAVE; d <<= 1) {
        const ull o = __shfl_xor(gbest, d);
        gbest = o < gbest ? o : gbest;
      }
      if (lane_id() == 0) s_g[threadIdx.x / WAVE] = gbest;
      __syncthreads();
      if (threadIdx.x == 0) {
        ull g = KEY_NONE;
        for (int w = 0; w < BLK / WAVE; w++) g = s_g[w] < g ? s_g[w] : g;
        if (g != KEY_NONE) min_store(best, groot[g0], g);
      }
      __syncthreads();
    }
    const uint32_t seg = v % KSEG;
    block_append<BLK, IPT>(keep, slot, out_cnt + seg * CSTR,
                           (uint64_t)seg * out_cap);
#pragma unroll
    for (int j = 0; j < IPT; j++) {
      if (keep[j]) {
        const uint64_t x = x0 + (uint64_t)j * BLK + threadIdx.x;
        oa[slot[j]] = (uint32_t)x;
        ob[slot[j]] = G[j];
        ok[slot[j]] = key[j];
      }
    }
  }
  giant_emit_flush(emitted, wsum, wbad, parts, gh);
}

// tau = key of rank ~light_edges in a strided sample of TSAMP input keys
// (KEY_NONE: every edge is light).  The light phase only needs about
// light_edges edges; the results do not depend on tau.
//
// Degree skew picks light_edges (when light_skewed differs from it): the
// sampled edges' endpoints go into an LDS hash set, and an endpoint already
// there counts as a repeat.  Among k endpoints drawn uniformly from n
// vertices about k^2 / 2n repeat; a graph with hubs (RMAT, power laws)
// repeats many times more, and its light subgraph still has a giant
// component with fewer light edges (light_skewed), while a uniform graph's
// light subgraph needs an average degree above 1 (light_edges > n / 2) for
// one.  Skewed: at least 16 repeats and 6 times the uniform expectation.
constexpr int TSAMP = 1024;
constexpr int HSLOTS = 4096;  // hash slots for the 2 * TSAMP sampled endpoints

// Inserts vertex x into the LDS set hs of SLOTS (a power of two) slots;
// counts a repeat in *rep.
template <int SLOTS>
__device__ __forceinline__ void sample_set_insert(uint32_t *hs, uint32_t x,
                                                  uint32_t *rep) {
  static_assert((SLOTS & (SLOTS - 1)) == 0 && SLOTS > 1, "SLOTS: a power of two");
  // multiplicative hash: the product's top bits
  const uint32_t h = (x * 2654435761u) >> (32 - __builtin_ctz((unsigned)SLOTS));
  for (uint32_t p = 0; p < (uint32_t)SLOTS; p++) {
    const uint32_t slot = (h + p) & (SLOTS - 1);
    const uint32_t old = atomicCAS(&hs[slot], U32_NONE, x);
    if (old == U32_NONE) return;
    if (old == x) {
      atomicAdd(rep, 1u);
      return;
    }
  }
}

template <class Src>
__global__ __launch_bounds__(SBLK) void k_threshold(Src src, uint64_t m,
                                                    uint32_t n, int ord,
                                                    uint64_t light_edges,
                                                    uint64_t light_skewed,
                                                    ull *tau_out) {
  static_assert(HSLOTS >= 2 * TSAMP, "the endpoint set needs free slots");
  __shared__ ull s[TSAMP];
  __shared__ uint32_t hs[HSLOTS];
  __shared__ uint32_t s_rep, s_k;
  const bool detect = light_skewed != light_edges;
  if (detect) {
    for (int j = threadIdx.x; j < HSLOTS; j += SBLK) hs[j] = U32_NONE;
    if (threadIdx.x == 0) s_rep = s_k = 0;
    __syncthreads();
  }
  for (int j = threadIdx.x; j < TSAMP; j += SBLK) {
    const uint64_t i = (uint64_t)j * m / TSAMP;
    const uint32_t a = src.U(i), b = src.V(i);
    const bool valid = a < n && b < n && a != b;
    s[j] = valid ? make_key(src.W(i), i, ord) : KEY_NONE;
    if (detect && valid) {
      sample_set_insert<HSLOTS>(hs, a, &s_rep);
      sample_set_insert<HSLOTS>(hs, b, &s_rep);
      atomicAdd(&s_k, 2u);
    }
  }
  __syncthreads();
  bitonic_sort_lds<TSAMP>(s);
  if (threadIdx.x == 0) {
    uint64_t le = light_edges;
    if (detect) {
      const double k = (double)s_k;
      const double uniform = k * k / (2.0 * (double)n);
      if (s_rep >= 16 && (double)s_rep >= 6.0 * uniform) le = light_skewed;
    }
    ull r = (ull)((double)le * TSAMP / (double)m);
    *tau_out = (le >= m || r >= TSAMP) ? KEY_NONE : s[r];
  }
}

// Per graph g of a batch: taus[g] = the key of rank ~light_factor * n_g
// among up to TSAMP_G evenly spaced keys of the graph's edges (KEY_NONE:
// every edge of the graph is light).  One block per graph.
// Each graph's factor follows its own degree skew as in k_threshold (the
// sampled endpoints' repeats against k^2 / 2 n_g; skewed: at least 4
// repeats and 6 times that), when lf_skewed differs from light_factor.
constexpr int TSAMP_G = 256;
constexpr int HSLOTS_G = 1024;  // hash slots for the 2 * TSAMP_G endpoints
template <class Src>
__global__ __launch_bounds__(TSAMP_G) void k_threshold_groups(
    Src src, uint32_t k, const ull *__restrict__ voff,
    const ull *__restrict__ eoff, uint32_t n, int ord, double light_factor,
    double lf_skewed, ull *__restrict__ taus) {
  static_assert(HSLOTS_G >= 2 * TSAMP_G, "the endpoint set needs free slots");
  __shared__ ull s[TSAMP_G];
  __shared__ uint32_t hs[HSLOTS_G];
  __shared__ uint32_t s_rep, s_k;
  const bool detect = lf_skewed != light_factor;
  for (uint32_t g = blockIdx.x; g < k; g += gridDim.x) {
    const ull e0 = eoff[g], mg = eoff[g + 1] - e0;
    const ull cnt = mg < (ull)TSAMP_G ? mg : (ull)TSAMP_G;
    const uint32_t j = threadIdx.x;  // one sample per thread
    if (detect) {
      for (int x = j; x < HSLOTS_G; x += TSAMP_G) hs[x] = U32_NONE;
      if (j == 0) s_rep = s_k = 0;
      __syncthreads();
    }
    ull key = KEY_NONE;
    if (j < cnt) {
      const uint64_t i = e0 + (ull)j * mg / cnt;
      const uint32_t a = src.U(i), b = src.V(i);
      if (a < n && b < n && a != b) {
        key = make_key(src.W(i), i, ord);
        if (detect) {
          sample_set_insert<HSLOTS_G>(hs, a, &s_rep);
          sample_set_insert<HSLOTS_G>(hs, b, &s_rep);
          atomicAdd(&s_k, 2u);
        }
      }
    }
    s[j] = key;
    __syncthreads();
    bitonic_sort_lds<TSAMP_G, TSAMP_G>(s);
    if (j == 0) {
      const double ng = (double)(voff[g + 1] - voff[g]);
      double lf = light_factor;
      if (detect && ng > 0) {
        const double kk = (double)s_k;
        if (s_rep >= 4 && (double)s_rep >= 6.0 * kk * kk / (2.0 * ng)) lf = lf_skewed;
      }
      const double want = lf * ng;
      ull t = KEY_NONE;
      if (want < (double)mg) {
        const ull r = (ull)(want * (double)cnt / (double)mg);
        if (r < cnt) t = s[r];
      }
      taus[g] = t;
    }
    __syncthreads();  // s is refilled for the next graph
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

// Bitmap -> ascending ids (the exact weight is summed by k_hook).
__global__ __launch_bounds__(BLK) void k_bm_emit(const uint32_t *__restrict__ bm,
                                                 uint64_t words,
                                                 const ull *__restrict__ blk,
                                                 uint32_t *__restrict__ ids) {
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
#pragma unroll
  for (int k = 0; k < WPT; k++) {
    uint32_t x = word[k];
    while (x) {
      uint32_t bit = __ffs(x) - 1;
      x &= x - 1;
      ids[pos++] = (uint32_t)((w0 + k) * 32 + bit);
    }
  }
}

// Sums of counters for the host: value i is ctr[off[i]] (segs[i] == 1) or
// the sum of the KSEG segment counters at off[i].  One wave.
constexpr int MAX_SUMS = 16;  // values per k_seg_sums call (the C_SUMS slot)
struct SumReq {
  int n;
  int off[MAX_SUMS];
  int segs[MAX_SUMS];
};

__global__ __launch_bounds__(WAVE) void k_seg_sums(const ull *ctr, SumReq q,
                                                   ull *out) {
  const uint32_t lane = threadIdx.x;
  for (int i = 0; i < q.n; i++) {
    ull x = (int)lane < q.segs[i] ? ctr[q.off[i] + lane * CSTR] : 0ull;
#pragma unroll
    for (int d = 1; d < WAVE; d <<= 1) x += __shfl_xor(x, d);
    if (lane == 0) out[i] = x;
  }
}

// Fixed-round schedule (graph capture): before the heavy pass, a light phase
// that still has alive edges (E(r) > 0) turns the heavy pass and the giant
// emission into no-ops (tau = KEY_NONE selects no heavy edge, gmin stays
// empty), so their appends cannot land on top of the light phase's leftovers;
// the flag tells the host that the result is incomplete.
__global__ __launch_bounds__(WAVE) void k_phase_gate(const ull *e_cnt, ull *tau,
                                                     ull *gate) {
  const uint32_t lane = threadIdx.x;
  ull x = (int)lane < KSEG ? e_cnt[lane * CSTR] : 0ull;
#pragma unroll
  for (int d = 1; d < WAVE; d <<= 1) x += __shfl_xor(x, d);
  if (lane == 0 && x) {
    *tau = KEY_NONE;
    *gate = 1ull;
  }
}

// k_phase_gate for a batch's graphs, each with its own light threshold:
// every graph's threshold becomes all ones.
__global__ __launch_bounds__(BLK) void k_phase_gate_groups(const ull *e_cnt,
                                                           ull *taus,
                                                           uint32_t k,
                                                           ull *gate) {
  __shared__ ull s_x;
  if (threadIdx.x < WAVE) {
    const uint32_t lane = threadIdx.x;
    ull x = (int)lane < KSEG ? e_cnt[lane * CSTR] : 0ull;
#pragma unroll
    for (int d = 1; d < WAVE; d <<= 1) x += __shfl_xor(x, d);
    if (lane == 0) {
      s_x = x;
      if (x) *gate = 1ull;
    }
  }
  __syncthreads();
  if (s_x)
    for (uint32_t g = threadIdx.x; g < k; g += BLK) taus[g] = KEY_NONE;
}

// Counters of a fixed-round solve for the host, in one copy: forest size,
// tau, gate, the four partial sums, then E(r), A(r) for r = 0 .. r_end.
constexpr int FS_HEAD = 7;
__global__ __launch_bounds__(WAVE) void k_fixed_sums(const ull *ctr, int r_end,
                                                     ull *out) {
  const uint32_t lane = threadIdx.x;
  const int nv = FS_HEAD + 2 * (r_end + 1);
  for (int i = 0; i < nv; i++) {
    int off, segs = KSEG;
    if (i < 3) {
      off = i == 0 ? C_MSTCNT : i == 1 ? C_TAU : C_GATE;
      segs = 1;
    } else if (i < FS_HEAD) {
      off = cPart(i - 3);
    } else {
      const int r = (i - FS_HEAD) >> 1;
      off = (i - FS_HEAD) & 1 ? cA(r) : cE(r);
    }
    ull x = (int)lane < segs ? ctr[off + lane * CSTR] : 0ull;
#pragma unroll
    for (int d = 1; d < WAVE; d <<= 1) x += __shfl_xor(x, d);
    if (lane == 0) out[i] = x;
  }
}

// ---------------------------------------------------------------------------
// workspace layout

struct MstLayout {
  uint64_t par, best, gmin, gsum, listL, listA, ea_a, ea_b, ea_k, eb_a, eb_b, eb_k, bitmap,
      gbits, ctr, blk, total;
  uint64_t words, nb, ecap, lcap;
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
  // Segment capacities: a producer pass over at most ceil(x / T) + KSEG tiles
  // of T items sends at most ceil(ceil(x / T) / KSEG) + 1 tiles to one
  // segment, at most x / KSEG + 2 T items; the bound with the largest edge
  // tile of any shape (MAX_EDGE_TILE) covers every smaller tile too.
  // The heavy pass and k_giant_emit (LTILE vertices per tile) append to the
  // same list, so its segments hold both producers' worst case.
  // No list ever holds more than m + n alive edges (heavy survivors plus one
  // giant edge per component), so no segment does either: small graphs get
  // segments of m + n items instead of the tile-based bound.
  static_assert(MAX_EDGE_TILE >= 256 * 16 && MAX_EDGE_TILE >= 512 * 8 &&
                    TILE >= 512 * 4 && TILE >= 256 * 8,
                "segment capacity uses the largest edge tile");
  const uint64_t TE = max_edge_tile(m);
  L.ecap = (ceil_div(ceil_div(m ? m : 1, TE), KSEG) + 2) * TE +
           ceil_div(ceil_div(n ? n : 1, LTILE), KSEG) * LTILE;
  L.ecap = std::min<uint64_t>(L.ecap, align_up(m + n + 1, 4));
  // Component lists also come from the hook's smaller tiles (HTILE items):
  // ceil(ceil(n / HTILE) / KSEG) * HTILE <= n / KSEG + HTILE, below lcap.
  // A component list holds at most n entries.
  L.lcap = (ceil_div(ceil_div(n ? n : 1, LTILE), KSEG) + 2) * LTILE;
  L.lcap = std::min<uint64_t>(L.lcap, align_up((uint64_t)n + 1, 4));
  const uint64_t E = L.ecap * KSEG, LN = L.lcap * KSEG;
  L.par = take(4ull * n);
  L.best = take(8ull * n);
  L.gmin = take(8ull * n);
  L.gsum = take(4ull * GSUM_MAX);
  L.listL = take(4ull * LN);
  L.listA = take(4ull * LN);
  L.ea_a = take(4ull * E);
  L.ea_b = take(4ull * E);
  L.ea_k = take(8ull * E);
  L.eb_a = take(4ull * E);
  L.eb_b = take(4ull * E);
  L.eb_k = take(8ull * E);
  L.bitmap = take(4ull * L.words);
  L.gbits = take(4ull * ceil_div(n ? n : 1, 32));
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

static double mst_light_factor(const phem_mst_opts *opts) {
  return opts && !std::isnan(opts->light_factor) ? opts->light_factor
                                                 : mst_default_light_factor();
}

static uint64_t mst_light_edges(uint32_t n, uint64_t m, double light_factor) {
  return light_factor > 0 ? (uint64_t)(light_factor * (double)n) : m;
}

// The split pays for its second input pass and giant filter only when the
// heavy edges are at least half as many as the light ones (m > 1.5 x
// light_edges); sparser graphs run the canonical rounds.
static bool mst_split(uint32_t n, uint64_t m, double light_factor) {
  const uint64_t le = mst_light_edges(n, m, light_factor);
  return le < m && (double)m > kernel_shape().mst_split_ratio * (double)le;
}

// Launches k_pass<MODE> over the m input edges with the knobs' shape
// (mst_block x mst_ept) for structure-of-arrays inputs; the array-of-records
// layout keeps the built-in shape.
template <int MODE, class Src, class... A>
static void launch_input_pass(uint64_t m, hipStream_t s, A... args) {
  if constexpr (std::is_same<Src, SrcSoA>::value) {
    const KernelShape k = kernel_shape();
    const uint32_t B = k.mst_block;
    uint32_t E = k.mst_ept;
    if ((uint64_t)B * E > max_edge_tile(m)) E = (uint32_t)(max_edge_tile(m) / B);
#define PHEM_INPUT_PASS(BB, EE)                                                \
    if (B == BB && E == EE) {                                                  \
      hipLaunchKernelGGL((k_pass<MODE, Src, EE, BB>),                          \
                         dim3(grid_for_threads(m, (uint64_t)BB * EE, BB)),     \
                         dim3(BB), 0, s, args...);                             \
      return;                                                                  \
    }
    PHEM_INPUT_PASS(256, 4) PHEM_INPUT_PASS(256, 8) PHEM_INPUT_PASS(256, 16)
    PHEM_INPUT_PASS(512, 4) PHEM_INPUT_PASS(512, 8)
#undef PHEM_INPUT_PASS
  }
  hipLaunchKernelGGL((k_pass<MODE, Src, EPT, BLK>), dim3(grid_for(m, TILE)),
                     dim3(BLK), 0, s, args...);
}

// grp: the graphs of a batch (giant component per graph), or NULL.
// Rounds the last solve of each size class (log2 n, log2 m) needed in its
// first phase (canonical or light) and its heavy phase.  A solve launches
// that many rounds of a phase before the phase's first counter read (at
// most what one read reports), so a run of similar solves reads the
// counters once per phase; the results do not depend on it.
static std::atomic<uint8_t> g_phase_rounds[2][34][42];

static int size_log2(uint64_t x, int cap) {
  int l = 0;
  while (x > 1 && l < cap) {
    x >>= 1;
    l++;
  }
  return l;
}

// tail: work enqueued before the final counter read, or NULL.
// fr: fixed-round schedule with no counter read at all (stats unwritten).
// first_rounds: rounds of the first phase launched before its first counter
// read (a caller that expects that many, as a stream does from its previous
// batch; 0: the size class's last count, or the default batch).
template <class Src>
static int mst_solve_impl(uint32_t n, uint64_t m, Src src, int wmode,
                          void *d_ws, uint64_t ws_bytes, uint32_t *d_mst_ids,
                          uint32_t *d_labels, const phem_mst_opts *opts,
                          phem_mst_stats *stats, hipStream_t s,
                          const GraphGroups *grp = nullptr,
                          const MstTail *tail = nullptr,
                          const FixedRounds *fr = nullptr,
                          ull *host_sums = nullptr, int first_rounds = 0) {
  const bool groups = grp && grp->k > 1;
  if (m > PHEM_MAX_EDGES || n > PHEM_MAX_VERTICES)
    return set_err(PHEM_EINVAL, "n=%u m=%llu beyond limits", n,
                   (unsigned long long)m);
  if (wmode != PHEM_W_U32 && wmode != PHEM_W_F32)
    return set_err(PHEM_EINVAL, "unknown weight mode %d", wmode);
  const uint32_t flags = opts ? opts->flags : 0u;
  if (flags & ~(uint32_t)PHEM_MST_MAXIMUM)
    return set_err(PHEM_EINVAL, "unknown option flags 0x%x", flags);
  // key order passed to the kernels that build keys
  const int ord = wmode | ((flags & PHEM_MST_MAXIMUM) ? ORD_MAX : 0);
  if (n && !d_mst_ids) return set_err(PHEM_EINVAL, "d_mst_ids is NULL");
  const MstLayout L = mst_layout(n, m);
  const KernelShape shape = kernel_shape();
  // compaction rounds: 16 or 8 edges per thread
  const bool wide = m >= (1ull << std::max<uint32_t>(shape.mst_compact_log2,
                                                    WIDE_LOG2)) &&
                    shape.mst_ept_compact == 16;
  if (!d_ws || ws_bytes < L.total)
    return set_err(PHEM_ENOMEM, "workspace %llu B < required %llu B",
                   (unsigned long long)ws_bytes, (unsigned long long)L.total);
  const double light_factor = mst_light_factor(opts);
  const uint64_t light_edges = mst_light_edges(n, m, light_factor);
  const bool split = mst_split(n, m, light_factor);
  // a solve that leaves light_factor to the library takes the skewed-graph
  // factor when the threshold sample finds hubs (k_threshold)
  const double lf_skew = kernel_shape().mst_light_factor_skewed;
  const uint64_t light_skewed =
      (!opts || std::isnan(opts->light_factor)) && lf_skew > 0 && light_factor > 0
          ? mst_light_edges(n, m, lf_skew)
          : light_edges;

  const int lgn = size_log2(n, 33), lgm = size_log2(m, 41);
  int hint[2] = {first_rounds > 0 ? first_rounds
                                  : (int)g_phase_rounds[0][lgn][lgm].load(std::memory_order_relaxed),
                 (int)g_phase_rounds[1][lgn][lgm].load(std::memory_order_relaxed)};

  char *ws = (char *)d_ws;
  uint32_t *par = (uint32_t *)(ws + L.par);
  ull *best = (ull *)(ws + L.best);
  ull *gmin = (ull *)(ws + L.gmin);
  uint32_t *gsum = (uint32_t *)(ws + L.gsum);
  uint32_t *listL = (uint32_t *)(ws + L.listL);
  uint32_t *listA = (uint32_t *)(ws + L.listA);
  uint32_t *Ea[2] = {(uint32_t *)(ws + L.ea_a), (uint32_t *)(ws + L.eb_a)};
  uint32_t *Eb[2] = {(uint32_t *)(ws + L.ea_b), (uint32_t *)(ws + L.eb_b)};
  ull *Ek[2] = {(ull *)(ws + L.ea_k), (ull *)(ws + L.eb_k)};
  uint32_t *bitmap = (uint32_t *)(ws + L.bitmap);
  uint32_t *gbits = (uint32_t *)(ws + L.gbits);
  ull *ctr = (ull *)(ws + L.ctr);
  ull *blk = (ull *)(ws + L.blk);
  const uint32_t *nou = nullptr;
  const ull *noull = nullptr;

  if (stats) std::memset(stats, 0, sizeof(*stats));

  uint64_t initw = n;
  if (L.words > initw) initw = L.words;
  if ((uint64_t)NCTR > initw) initw = NCTR;
  {
    PHEM_PROF("k_mst_init", s);
    hipLaunchKernelGGL(k_mst_init, dim3(grid_for(initw, BLK)), dim3(BLK), 0, s,
                       n, par, best, split ? gmin : (ull *)nullptr, L.words,
                       bitmap, ctr);
  }
  PHEM_LAUNCHED("k_mst_init");

  // Counter values for the host: summed on the device, copied to host_sums
  // (the caller's host memory) or to a pinned buffer of the shared pool held
  // until the solve returns.
  std::unique_ptr<PinnedLease> lease;
  ull *hpin = host_sums;
  auto read_sum_req = [&](const SumReq &q, ull *out) -> int {
    if (!hpin) {
      lease.reset(new (std::nothrow) PinnedLease(MAX_SUMS * sizeof(ull), s));
      if (!lease || !lease->get())
        return set_err(PHEM_ENOMEM, "pinned counter buffer");
      hpin = (ull *)lease->get();
    }
    hipLaunchKernelGGL(k_seg_sums, dim3(1), dim3(WAVE), 0, s, ctr, q,
                       ctr + C_SUMS);
    PHEM_LAUNCHED("k_seg_sums");
    if (lease) lease->busy();
    PHEM_HIP(hipMemcpyAsync(hpin, ctr + C_SUMS, q.n * sizeof(ull),
                            hipMemcpyDeviceToHost, s));
    PHEM_HIP(hipStreamSynchronize(s));
    if (lease) lease->done();  // no copy into it is pending
    for (int i = 0; i < q.n; i++) out[i] = hpin[i];
    return PHEM_OK;
  };
  auto read_sums = [&](std::initializer_list<std::pair<int, int>> req,
                       ull *out) -> int {
    SumReq q{};
    for (auto &p : req) {
      q.off[q.n] = p.first;
      q.segs[q.n] = p.second;
      q.n++;
    }
    return read_sum_req(q, out);
  };
  ull hv[8];
  int r = 0;    // round index (a round = hook + compress + compact)
  int cur = 0;  // edge buffer holding the alive edges of round r
  uint64_t m_r = 0;

  // Rounds over the alive edges in buffer `cur` until none is left.  The
  // first hook of a phase runs over the vertices holding an offer (collected
  // from best[]); later hooks over the roots collected by the previous
  // compress.  On entry m_r may be an
  // upper bound of the alive edges (the producer's count not read yet): the
  // first counter read then reports E(r) with the rounds' counters, together
  // with the `extra` values (written to extra_out), so no read of its own is
  // needed between the producer and the first rounds.
  // giant_hooked: k_giant_emit hooked components outside the round's lists
  // (parent G, or an ancestor of G after path halving).  The round's compress
  // flattens every active component, G and the roots it hooks to included,
  // so the phase's first compaction relabels in two steps.
  auto run_rounds = [&](uint64_t list_ub, bool first_all,
                        std::initializer_list<std::pair<int, int>> extra,
                        ull *extra_out, int fixed, int phase,
                        bool giant_hooked = false) -> int {
    bool all = first_all;
    bool first = true;
    int phase_real = 0;  // rounds of this phase that had alive edges
    while (m_r > 0) {
      // Up to SPEC rounds per counter read (exactly `fixed` rounds and no
      // read for a recorded solve).  A round launched after the phase has
      // converged finds empty inputs and leaves its counters at zero; grids
      // come from upper bounds (alive edges and component lists only shrink
      // within a phase), and the kernels loop over the tiles the device
      // counters hold.
      // one read reports 2 batch + 2 values and, on a phase's first read,
      // the extra ones: at most MAX_SUMS
      const int read_cap = (MAX_SUMS - 2 - (first ? (int)extra.size() : 0)) / 2;
      int want = fixed ? fixed : SPEC;
      if (!fixed && first && hint[phase] > 0)
        want = hint[phase] < read_cap ? hint[phase] : read_cap;
      const int batch = PHEM_MAX_ROUNDS - r < want ? PHEM_MAX_ROUNDS - r : want;
      if (batch <= 0 || (fixed && batch < fixed))
        return set_err(PHEM_ENOCONV, "no convergence after %d rounds", r);
      const uint64_t list_bound = all ? n : list_ub;
      for (int k = 0; k < batch; k++) {
        const int rr = r + k;
        const int c = cur ^ (k & 1);
        if (all && k == 0) {
          // the phase's first component list: the vertices holding an offer
          // (its counters may hold a previous phase's roots)
          PHEM_PROF("k_collect_active", s);
          hipLaunchKernelGGL(k_zero_seg, dim3(1), dim3(KSEG), 0, s, ctr + cL(rr));
          hipLaunchKernelGGL(k_collect_active,
                             dim3(grid_for(ceil_div(n, LTILE), 1)), dim3(BLK), 0,
                             s, n, (const ull *)best, listL, ctr + cL(rr),
                             L.lcap);
        }
        PHEM_LAUNCHED("k_collect_active");
        const uint64_t hook_tiles = ceil_div(list_bound, HTILE) + KSEG;
        {
          PHEM_PROF("k_hook", s);
          hipLaunchKernelGGL(k_hook<Src>, dim3(grid_for(hook_tiles, 1)),
                             dim3(BLK), 0, s, listL, ctr + cL(rr), L.lcap, 0u,
                             best, src, ord, par, bitmap, listA, ctr + cA(rr),
                             L.lcap, ctr + cPart(1), ctr + cPart(2));
        }
        PHEM_LAUNCHED("k_hook");
        {
          PHEM_PROF("k_compress", s);
          hipLaunchKernelGGL(k_compress,
                             dim3(grid_for(ceil_div(list_bound, LTILE) + 2 * KSEG, 1)),
                             dim3(BLK), 0, s, listA, ctr + cA(rr), L.lcap, par,
                             listL, ctr + cL(rr + 1), L.lcap);
        }
        PHEM_LAUNCHED("k_compress");
        const int two_step = giant_hooked && first && k == 0;
        {
          PHEM_PROF("k_compact", s);
          if (wide)
            hipLaunchKernelGGL((k_pass<P_EDGES, Src, 16>),
                               dim3(grid_for(ceil_div(m_r, BLK * 16) + KSEG, 1)),
                               dim3(BLK), 0, s, src, Ea[c], Eb[c], Ek[c],
                               ctr + cE(rr), L.ecap, 0ull, n, ord, par, noull,
                               nou, Ea[c ^ 1], Eb[c ^ 1], Ek[c ^ 1],
                               ctr + cE(rr + 1), L.ecap, (ull *)nullptr, best,
                               (ull *)nullptr, nou, EdgeTaus{}, two_step);
          else
            hipLaunchKernelGGL((k_pass<P_EDGES, Src>),
                               dim3(grid_for(ceil_div(m_r, TILE) + KSEG, 1)),
                               dim3(BLK), 0, s, src, Ea[c], Eb[c], Ek[c],
                               ctr + cE(rr), L.ecap, 0ull, n, ord, par, noull,
                               nou, Ea[c ^ 1], Eb[c ^ 1], Ek[c ^ 1],
                               ctr + cE(rr + 1), L.ecap, (ull *)nullptr, best,
                               (ull *)nullptr, nou, EdgeTaus{}, two_step);
        }
        PHEM_LAUNCHED("k_compact");
      }
      if (fixed) {
        r += batch;
        cur ^= (batch & 1);
        break;
      }
      // E(r .. r+batch), A(r .. r+batch-1), L(r+batch)
      SumReq q{};
      for (int k = 0; k <= batch; k++) {
        q.off[q.n] = cE(r + k);
        q.segs[q.n++] = KSEG;
      }
      for (int k = 0; k < batch; k++) {
        q.off[q.n] = cA(r + k);
        q.segs[q.n++] = KSEG;
      }
      q.off[q.n] = cL(r + batch);
      q.segs[q.n++] = KSEG;
      const int nbase = q.n;
      if (first) {
        for (auto &p : extra) {
          q.off[q.n] = p.first;
          q.segs[q.n++] = p.second;
        }
      }
      ull got[MAX_SUMS];
      int rc = read_sum_req(q, got);
      if (rc) return rc;
      if (first) {
        for (int i = nbase; i < q.n; i++) extra_out[i - nbase] = got[i];
        first = false;
      }
      int real = 0;  // rounds that started with at least one alive edge
      while (real < batch && got[real] > 0) {
        if (stats) {
          stats->m_log[r + real] = got[real];
          stats->c_log[r + real] = got[batch + 1 + real];
        }
        real++;
      }
      phase_real += real;
      m_r = got[batch];        // E(r + batch); 0 when the phase converged
      list_ub = got[2 * batch + 1];
      cur ^= (real & 1);
      r += real;
      all = false;
      if (real < batch) m_r = 0;
    }
    if (!fixed && phase_real > 0)
      g_phase_rounds[phase][lgn][lgm].store((uint8_t)(phase_real < 255 ? phase_real : 255),
                                            std::memory_order_relaxed);
    return PHEM_OK;
  };

  if (m > 0 && n > 1 && !split) {
    // canonical Boruvka: round 0 sweeps the input, then hooks every vertex
    {
      PHEM_PROF("k_sweep0", s);
      hipLaunchKernelGGL(k_sweep0<Src>, dim3(grid_for(m, TILE)), dim3(BLK), 0,
                         s, src, m, n, ord, best);
    }
    PHEM_LAUNCHED("k_sweep0");
    const uint64_t t0 = ceil_div(n, LTILE);
    {
      PHEM_PROF("k_hook", s);
      hipLaunchKernelGGL(k_hook<Src>, dim3(grid_for(ceil_div(n, HTILE), 1)),
                         dim3(BLK), 0, s, listL, noull, L.lcap, n, best, src, ord,
                         par, bitmap,
                         listA, ctr + cA(0), L.lcap, ctr + cPart(1),
                         ctr + cPart(2));
    }
    PHEM_LAUNCHED("k_hook");
    {
      PHEM_PROF("k_compress", s);
      hipLaunchKernelGGL(k_compress, dim3(grid_for(t0 + KSEG, 1)), dim3(BLK), 0,
                         s, listA, ctr + cA(0), L.lcap, par, listL, ctr + cL(1),
                         L.lcap);
    }
    PHEM_LAUNCHED("k_compress");
    {
      PHEM_PROF("k_compact<R0>", s);
      hipLaunchKernelGGL((k_pass<P_ALL, Src>), dim3(grid_for(m, TILE)),
                         dim3(BLK), 0, s, src, nou, nou, noull, noull, 0ull, m,
                         n, ord, par, noull, nou, Ea[0], Eb[0], Ek[0],
                         ctr + cE(1), L.ecap, ctr + cPart(0), best,
                         (ull *)nullptr, nou, EdgeTaus{}, 0);
    }
    PHEM_LAUNCHED("k_compact<R0>");
    // rounds 1.. start at once from upper bounds (m alive edges, n roots);
    // round 0's counters arrive with their first counter read
    r = 1;
    m_r = m;
    cur = 0;
    int rc = run_rounds(n, false, {{cPart(0), KSEG}, {cA(0), KSEG}}, hv,
                        fr ? fr->light : 0, 0);
    if (rc) return rc;
    const uint64_t valid = fr ? 1 : hv[0];
    // a round counts when it starts with at least one alive edge
    if (!valid) r = 0;
    if (stats && !fr) {
      stats->m_log[0] = valid;
      stats->c_log[0] = hv[1];
      stats->light_edges = valid;
      stats->tau = KEY_NONE;
      stats->light_rounds = (uint32_t)r;
    }
  } else if (m > 0 && n > 1) {
    // light phase: edges with key < tau (about light_edges of them; for a
    // batch, each graph's own threshold gives it about light_factor * n_g)
    const EdgeTaus et = groups ? EdgeTaus{grp->taus, grp->eoff, grp->emap, grp->k}
                               : EdgeTaus{};
    {
      PHEM_PROF("k_threshold", s);
      if (groups)
        hipLaunchKernelGGL(k_threshold_groups<Src>, dim3(grid_for(grp->k, 1)),
                           dim3(TSAMP_G), 0, s, src, grp->k, grp->voff,
                           grp->eoff, n, ord, light_factor,
                           light_skewed != light_edges ? lf_skew : light_factor,
                           grp->taus);
      else
        hipLaunchKernelGGL(k_threshold<Src>, dim3(1), dim3(SBLK), 0, s, src, m,
                           n, ord, light_edges, light_skewed, ctr + C_TAU);
    }
    PHEM_LAUNCHED("k_threshold");
    {
      PHEM_PROF("k_light", s);
      launch_input_pass<P_LIGHT, Src>(m, s, src, nou, nou, noull, noull, 0ull, m,
                         n, ord, par, ctr + C_TAU, nou, Ea[0], Eb[0], Ek[0],
                         ctr + cE(0), L.ecap, ctr + cPart(0), best,
                         (ull *)nullptr, nou, et, 0);
    }
    PHEM_LAUNCHED("k_light");
    // the light edge count arrives with the first rounds' counters
    m_r = m;
    cur = 0;
    int rc = run_rounds(0, true, {{C_TAU, 1}, {cE(0), KSEG}}, hv,
                        fr ? fr->light : 0, 0);
    if (rc) return rc;
    if (fr) {
      if (groups)
        hipLaunchKernelGGL(k_phase_gate_groups, dim3(1), dim3(BLK), 0, s,
                           (const ull *)(ctr + cE(r)), grp->taus, grp->k,
                           ctr + C_GATE);
      else
        hipLaunchKernelGGL(k_phase_gate, dim3(1), dim3(WAVE), 0, s,
                           (const ull *)(ctr + cE(r)), ctr + C_TAU, ctr + C_GATE);
      PHEM_LAUNCHED("k_phase_gate");
    } else if (stats) {
      stats->tau = hv[0];
      stats->light_edges = hv[1];
      stats->light_rounds = (uint32_t)r;
    }
    // heavy phase: flatten, mark the giant component, filter the rest
    {
      PHEM_PROF("k_flatten", s);
      hipLaunchKernelGGL(k_flatten, dim3(grid_for(n, BLK)), dim3(BLK), 0, s, n,
                         par);
    }
    PHEM_LAUNCHED("k_flatten");
    if (groups) {
      {
        PHEM_PROF("k_pick_giant", s);
        hipLaunchKernelGGL(k_pick_giant_groups, dim3(grid_for(grp->k, 1)),
                           dim3(GSAMP_G), 0, s, grp->k, grp->voff, par,
                           grp->groot);
      }
      PHEM_LAUNCHED("k_pick_giant_groups");
      {
        PHEM_PROF("k_giant_bits", s);
        hipLaunchKernelGGL(k_giant_bits_groups, dim3(grid_for(n, SPLIT_BLK)),
                           dim3(SPLIT_BLK), 0, s, n, par, grp->voff, grp->vmap,
                           grp->k, (const uint32_t *)grp->groot, gbits);
      }
      PHEM_LAUNCHED("k_giant_bits_groups");
    } else {
      {
        PHEM_PROF("k_pick_giant", s);
        hipLaunchKernelGGL(k_pick_giant, dim3(1), dim3(SBLK), 0, s, n, par,
                           (uint32_t *)(ctr + C_GIANT), (uint32_t *)nullptr);
      }
      PHEM_LAUNCHED("k_pick_giant");
      {
        PHEM_PROF("k_giant_bits", s);
        hipLaunchKernelGGL(k_giant_bits, dim3(grid_for(n, BLK)), dim3(BLK), 0, s,
                           n, par, (const uint32_t *)(ctr + C_GIANT), gbits);
      }
      PHEM_LAUNCHED("k_giant_bits");
    }
    {
      PHEM_PROF("k_giant_summary", s);
      hipLaunchKernelGGL(k_giant_summary,
                         dim3(grid_for(32ull * gsum_words(n, GSUM_MAX), SPLIT_BLK)),
                         dim3(SPLIT_BLK), 0, s, n, (const uint32_t *)gbits, gsum);
    }
    PHEM_LAUNCHED("k_giant_summary");
    {
      PHEM_PROF("k_heavy", s);
      launch_input_pass<P_HEAVY, Src>(m, s, src, nou, nou, noull, noull, 0ull, m,
                         n, ord, par, ctr + C_TAU, gbits, Ea[cur], Eb[cur],
                         Ek[cur], ctr + cE(r), L.ecap, (ull *)nullptr, best,
                         gmin, gsum, et, 0);
    }
    PHEM_LAUNCHED("k_heavy");
    {
      PHEM_PROF("k_giant_fold", s);
      hipLaunchKernelGGL(k_giant_fold, dim3(grid_for(n, BLK)), dim3(BLK), 0, s,
                         n, par, gmin);
    }
    PHEM_LAUNCHED("k_giant_fold");
    const GiantHook gh{par, bitmap, ord, ctr + cPart(1), ctr + cPart(2)};
    {
      PHEM_PROF("k_giant_emit", s);
      if (groups)
        hipLaunchKernelGGL(k_giant_emit_groups,
                           dim3(grid_for(ceil_div(n, LTILE), 1)), dim3(BLK), 0,
                           s, n, gmin, grp->voff, grp->vmap, grp->k,
                           (const uint32_t *)grp->groot, Ea[cur], Eb[cur],
                           Ek[cur], ctr + cE(r), L.ecap, best, ctr + cPart(3),
                           gh);
      else
        hipLaunchKernelGGL(k_giant_emit, dim3(grid_for(ceil_div(n, LTILE), 1)),
                           dim3(BLK), 0, s, n, gmin,
                           (const uint32_t *)(ctr + C_GIANT), Ea[cur], Eb[cur],
                           Ek[cur], ctr + cE(r), L.ecap, best, ctr + cPart(3),
                           gh);
    }
    PHEM_LAUNCHED("k_giant_emit");
    // the heavy survivors (at most m + n) are counted with the first heavy
    // rounds' counters
    m_r = m + n;
    const int r_heavy = r;
    rc = run_rounds(0, true, {{cE(r_heavy), KSEG}, {cPart(3), KSEG}}, hv,
                    fr ? fr->heavy : 0, 1, true);
    if (rc) return rc;
    if (stats && !fr) {
      stats->heavy_edges = hv[0];
      stats->giant_edges = hv[1];
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
    hipLaunchKernelGGL(k_bm_emit, dim3((unsigned)L.nb), dim3(BLK), 0, s,
                       bitmap, L.words, blk, d_mst_ids);
  }
  PHEM_LAUNCHED("k_bm_emit");
  if (tail) {
    const int trc = (*tail)(ctr + C_MSTCNT);
    if (trc) return trc;
  }
  if (fr) return PHEM_OK;
  int rc = read_sums({{C_MSTCNT, 1}, {cPart(1), KSEG}, {cPart(2), KSEG},
                      {cPart(0), KSEG}}, hv);
  if (rc) return rc;
  if (stats) {
    stats->rounds = (uint32_t)r;
    stats->mst_count = (uint32_t)hv[0];
    stats->weight_fx = hv[1];
    stats->weight_exact = hv[2] == 0;
    for (int k = 0; k < r; k++) {
      stats->sum_m += stats->m_log[k];
      stats->sum_c += stats->c_log[k];
    }
    stats->valid_edges = hv[3];
  }
  return PHEM_OK;
}

static bool aligned16(const void *p) { return ((uintptr_t)p & 15u) == 0; }

int mst_solve_groups(uint32_t n, uint64_t m, const uint32_t *d_u,
                     const uint32_t *d_v, const uint32_t *d_w, int wmode,
                     void *d_ws, uint64_t ws_bytes, uint32_t *d_mst_ids,
                     uint32_t *d_labels, const phem_mst_opts *opts,
                     phem_mst_stats *stats, hipStream_t s,
                     const GraphGroups &groups, const FixedRounds *fr) {
  if (m && (!d_u || !d_v || !d_w))
    return set_err(PHEM_EINVAL, "edge arrays are NULL");
  if (groups.k > 1 && (!groups.voff || !groups.eoff || !groups.vmap ||
                       !groups.emap || !groups.taus || !groups.groot))
    return set_err(PHEM_EINVAL, "graph ranges are NULL");
  const bool vec = aligned16(d_u) && aligned16(d_v) && aligned16(d_w);
  return mst_solve_impl(n, m, SrcSoA{d_u, d_v, d_w, vec}, wmode, d_ws, ws_bytes,
                        d_mst_ids, d_labels, opts, fr ? nullptr : stats, s,
                        &groups, nullptr, fr);
}

const unsigned long long *solve_count_word(const void *d_ws, uint32_t n,
                                           uint64_t m) {
  return (const ull *)((const char *)d_ws + mst_layout(n, m).ctr) + C_MSTCNT;
}

bool mst_uses_split(uint32_t n, uint64_t m, const phem_mst_opts *opts) {
  return m > 0 && n > 1 && mst_split(n, m, mst_light_factor(opts));
}

int fixed_sums_words() { return FS_HEAD + 2 * (PHEM_MAX_ROUNDS + 1); }

int fixed_round_stats(const void *d_ws, uint32_t n, uint64_t m, bool split,
                      const FixedRounds &fr, ull *d_sums, ull *h_sums,
                      phem_mst_stats *stats, hipStream_t s) {
  const bool work = m > 0 && n > 1;
  const int r_end = work ? (split ? fr.light + fr.heavy : 1 + fr.light) : 0;
  const ull *ctr = (const ull *)((const char *)d_ws + mst_layout(n, m).ctr);
  const int nv = FS_HEAD + 2 * (r_end + 1);
  hipLaunchKernelGGL(k_fixed_sums, dim3(1), dim3(WAVE), 0, s, ctr, r_end,
                     d_sums);
  PHEM_LAUNCHED("k_fixed_sums");
  PHEM_HIP(hipMemcpyAsync(h_sums, d_sums, sizeof(ull) * nv,
                          hipMemcpyDeviceToHost, s));
  PHEM_HIP(hipStreamSynchronize(s));
  const ull *h = h_sums;
  auto E = [&](int r) { return h[FS_HEAD + 2 * r]; };
  auto A = [&](int r) { return h[FS_HEAD + 2 * r + 1]; };
  phem_mst_stats t;
  std::memset(&t, 0, sizeof(t));
  t.mst_count = (uint32_t)h[0];
  t.valid_edges = h[3];
  t.weight_fx = h[4];
  t.weight_exact = h[5] == 0;
  bool converged = true;
  if (work) {
    int r = 0;
    // the rounds that started with an alive edge are a prefix of each phase
    auto phase = [&](int k0, int k1) {
      for (int k = k0; k < k1 && E(k) > 0; k++, r++) {
        t.m_log[r] = E(k);
        t.c_log[r] = A(k);
      }
    };
    if (!split) {
      if (h[3]) {  // round 0 sweeps the input
        t.m_log[0] = h[3];
        t.c_log[0] = A(0);
        r = 1;
        phase(1, 1 + fr.light);
      }
      t.light_edges = h[3];
      t.tau = KEY_NONE;
      converged = E(1 + fr.light) == 0;
      t.light_rounds = (uint32_t)r;
    } else {
      phase(0, fr.light);
      t.light_rounds = (uint32_t)r;
      const int rh = fr.light;
      phase(rh, rh + fr.heavy);
      t.tau = h[1];
      t.light_edges = E(0);
      t.heavy_edges = E(rh);
      t.giant_edges = h[6];
      converged = h[2] == 0 && E(rh + fr.heavy) == 0;
    }
    t.rounds = (uint32_t)r;
    for (int k = 0; k < r; k++) {
      t.sum_m += t.m_log[k];
      t.sum_c += t.c_log[k];
    }
  }
  if (stats) *stats = t;
  if (!converged)
    return set_err(PHEM_ENOCONV,
                   "a phase needed more than its %d recorded rounds",
                   split ? (fr.light > fr.heavy ? fr.light : fr.heavy) : fr.light);
  return PHEM_OK;
}

int mst_solve_tail(uint32_t n, uint64_t m, const uint32_t *d_u,
                   const uint32_t *d_v, const uint32_t *d_w, int wmode,
                   void *d_ws, uint64_t ws_bytes, uint32_t *d_mst_ids,
                   uint32_t *d_labels, const phem_mst_opts *opts,
                   phem_mst_stats *stats, hipStream_t s, const MstTail &tail,
                   int first_rounds) {
  if (m && (!d_u || !d_v || !d_w))
    return set_err(PHEM_EINVAL, "edge arrays are NULL");
  const bool vec = aligned16(d_u) && aligned16(d_v) && aligned16(d_w);
  return mst_solve_impl(n, m, SrcSoA{d_u, d_v, d_w, vec}, wmode, d_ws, ws_bytes,
                        d_mst_ids, d_labels, opts, stats, s, nullptr, &tail,
                        nullptr, nullptr, first_rounds);
}

}  // namespace phem

extern "C" int phem_mst_solve(uint32_t n, uint64_t m, const uint32_t *d_u,
                              const uint32_t *d_v, const uint32_t *d_w,
                              int wmode, void *d_ws, uint64_t ws_bytes,
                              uint32_t *d_mst_ids, uint32_t *d_labels,
                              const phem_mst_opts *opts, phem_mst_stats *stats,
                              void *stream) {
  if (m && (!d_u || !d_v || !d_w))
    return set_err(PHEM_EINVAL, "edge arrays are NULL");
  const bool vec = aligned16(d_u) && aligned16(d_v) && aligned16(d_w);
  return mst_solve_impl(n, m, SrcSoA{d_u, d_v, d_w, vec}, wmode, d_ws,
                        ws_bytes, d_mst_ids, d_labels, opts, stats,
                        (hipStream_t)stream);
}

extern "C" int phem_mst_solve_aos(uint32_t n, uint64_t m,
                                  const uint32_t *d_edges, int wmode,
                                  void *d_ws, uint64_t ws_bytes,
                                  uint32_t *d_mst_ids, uint32_t *d_labels,
                                  const phem_mst_opts *opts,
                                  phem_mst_stats *stats, void *stream) {
  if (m && !d_edges) return set_err(PHEM_EINVAL, "edge array is NULL");
  return mst_solve_impl(n, m, SrcAoS{d_edges, aligned16(d_edges)}, wmode, d_ws,
                        ws_bytes, d_mst_ids, d_labels, opts, stats,
                        (hipStream_t)stream);
}

// ---------------------------------------------------------------------------
// A whole solve recorded once into a HIP graph and replayed.

struct phem_mst_graph {
  uint32_t n = 0;
  uint64_t m = 0;
  bool split = false;
  bool launched = false;
  phem::FixedRounds fr{0, 0};
  void *ws = nullptr;  // the caller's workspace (its counters give the stats)
  unsigned long long *d_sums = nullptr, *h_sums = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

static void mst_graph_free(phem_mst_graph *g) {
  if (!g) return;
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  (void)hipFree(g->d_sums);
  (void)hipHostFree(g->h_sums);
  delete g;
}

extern "C" int phem_mst_graph_create(uint32_t n, uint64_t m, const uint32_t *d_u,
                                     const uint32_t *d_v, const uint32_t *d_w,
                                     int wmode, void *d_ws, uint64_t ws_bytes,
                                     uint32_t *d_mst_ids, uint32_t *d_labels,
                                     const phem_mst_opts *opts, int rounds,
                                     void *stream, phem_mst_graph_t **out) {
  if (!out) return set_err(PHEM_EINVAL, "out is NULL");
  *out = nullptr;
  if (m && (!d_u || !d_v || !d_w))
    return set_err(PHEM_EINVAL, "edge arrays are NULL");
  if (rounds < 0) return set_err(PHEM_EINVAL, "rounds %d < 0", rounds);
  const bool work = m > 0 && n > 1;
  const bool split = work && mst_split(n, m, mst_light_factor(opts));
  FixedRounds fr{rounds, split ? rounds : 0};
  if (rounds == 0) {
    // one solve on the current inputs gives the rounds of each phase; the
    // recorded schedule keeps one spare round per phase
    phem_mst_stats st;
    const int rc = phem_mst_solve(n, m, d_u, d_v, d_w, wmode, d_ws, ws_bytes,
                                  d_mst_ids, d_labels, opts, &st, stream);
    if (rc) return rc;
    const int l = (int)st.light_rounds, h = (int)(st.rounds - st.light_rounds);
    fr.light = (split ? l : (l > 0 ? l - 1 : 0)) + 1;
    fr.heavy = split ? h + 1 : 0;
  }
  const int r_end = work ? (split ? fr.light + fr.heavy : 1 + fr.light) : 0;
  if (r_end > PHEM_MAX_ROUNDS)
    return set_err(PHEM_EINVAL, "%d rounds per phase exceed the %d-round limit",
                   rounds ? rounds : fr.light, PHEM_MAX_ROUNDS);
  phem_mst_graph *g = new (std::nothrow) phem_mst_graph();
  if (!g) return set_err(PHEM_ENOMEM, "host allocation failed");
  g->n = n;
  g->m = m;
  g->split = split;
  g->fr = fr;
  g->ws = d_ws;
  const size_t sums = sizeof(ull) * (FS_HEAD + 2 * (PHEM_MAX_ROUNDS + 1));
  if (hipMalloc((void **)&g->d_sums, sums) != hipSuccess ||
      hipHostMalloc((void **)&g->h_sums, sums) != hipSuccess) {
    (void)hipGetLastError();
    mst_graph_free(g);
    return set_err(PHEM_ENOMEM, "graph counter buffers");
  }
  // record on a stream of our own (the caller's may be the null stream);
  // the launch sizing reads the device's CU count before recording starts
  (void)resident_blocks();
  hipStream_t cs = nullptr;
  if (hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess) {
    mst_graph_free(g);
    return set_err(PHEM_EHIP, "hipStreamCreateWithFlags failed");
  }
  int rc = PHEM_OK;
  if (hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) != hipSuccess) {
    rc = set_err(PHEM_EHIP, "hipStreamBeginCapture failed");
  } else {
    const bool vec = aligned16(d_u) && aligned16(d_v) && aligned16(d_w);
    rc = mst_solve_impl(n, m, SrcSoA{d_u, d_v, d_w, vec}, wmode, d_ws,
                        ws_bytes, d_mst_ids, d_labels, opts, nullptr, cs,
                        nullptr, nullptr, &fr);
    const hipError_t e = hipStreamEndCapture(cs, &g->graph);
    if (!rc && e != hipSuccess)
      rc = set_err(PHEM_EHIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
  }
  (void)hipStreamDestroy(cs);
  if (!rc && hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0) !=
                 hipSuccess)
    rc = set_err(PHEM_EHIP, "hipGraphInstantiate failed");
  if (rc) {
    mst_graph_free(g);
    return rc;
  }
  *out = g;
  return PHEM_OK;
}

extern "C" int phem_mst_graph_launch(phem_mst_graph_t *g, void *stream) {
  if (!g) return set_err(PHEM_EINVAL, "graph handle is NULL");
  PHEM_HIP(hipGraphLaunch(g->exec, (hipStream_t)stream));
  g->launched = true;
  return PHEM_OK;
}

extern "C" int phem_mst_graph_stats(phem_mst_graph_t *g, phem_mst_stats *stats,
                                    void *stream) {
  if (!g) return set_err(PHEM_EINVAL, "graph handle is NULL");
  if (!g->launched) return set_err(PHEM_EINVAL, "graph not launched yet");
  return fixed_round_stats(g->ws, g->n, g->m, g->split, g->fr, g->d_sums,
                           g->h_sums, stats, (hipStream_t)stream);
}

extern "C" int phem_mst_graph_destroy(phem_mst_graph_t *g) {
  mst_graph_free(g);
  return PHEM_OK;
}

// ---------------------------------------------------------------------------
// phem_mst_solve without any allocation: the workspace is a file-scope
// device array of NOALLOC_BYTES, one per device that runs it, held until
// the process ends.  Its address on each device comes from
// hipGetSymbolAddress on the first call there (the array's name in host
// code is not a device address); kernels get that address, as they get the
// caller's workspace in phem_mst_solve.  The solve initialises every part of
// the workspace it reads (k_mst_init clears the counters, parents, minima
// and forest bitmap), so the previous call's contents do not matter; calls
// that overlap on one device share the array and race.  Counter values come
// back through host memory on the stack, not a pinned buffer.

namespace phem {

constexpr uint64_t NOALLOC_BYTES = 64ull << 20;
static_assert(NOALLOC_BYTES % 256 == 0, "scratch size");
__device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

namespace {
std::mutex g_scratch_mu;
void *g_scratch_ptr[64];  // device address of `scratch`, per device
}  // namespace

}  // namespace phem

extern "C" int phem_mst_solve_noalloc(uint32_t n, uint64_t m,
                                      const uint32_t *d_u, const uint32_t *d_v,
                                      const uint32_t *d_w, int wmode,
                                      uint32_t *d_mst_ids, uint32_t *d_labels,
                                      const phem_mst_opts *opts,
                                      phem_mst_stats *stats, void *stream) {
  if (m > PHEM_MAX_EDGES || n > PHEM_MAX_VERTICES)
    return set_err(PHEM_EINVAL, "n=%u m=%llu beyond limits", n,
                   (unsigned long long)m);
  const uint64_t need = mst_layout(n, m).total;
  if (need > NOALLOC_BYTES)
    return set_err(PHEM_ENOMEM,
                   "n=%u m=%llu needs %llu B of workspace, beyond the %llu B "
                   "static scratch of phem_mst_solve_noalloc (use "
                   "phem_mst_solve with a workspace)",
                   n, (unsigned long long)m, (unsigned long long)need,
                   (unsigned long long)NOALLOC_BYTES);
  if (m && (!d_u || !d_v || !d_w))
    return set_err(PHEM_EINVAL, "edge arrays are NULL");
  int dev = 0;
  PHEM_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) return set_err(PHEM_EINVAL, "device %d beyond 64", dev);
  void *ws = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
      void *p = nullptr;
      PHEM_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
      if (!p) return set_err(PHEM_EHIP, "scratch has no device address");
      g_scratch_ptr[dev] = p;
    }
    ws = g_scratch_ptr[dev];
  }
  ull sums[MAX_SUMS];
  const bool vec = aligned16(d_u) && aligned16(d_v) && aligned16(d_w);
  return mst_solve_impl(n, m, SrcSoA{d_u, d_v, d_w, vec}, wmode, ws,
                        NOALLOC_BYTES, d_mst_ids, d_labels, opts, stats,
                        (hipStream_t)stream, nullptr, nullptr, nullptr, sums);
}
