// phem_dev.h -- device helpers shared by the MST, CC and generator kernels
// (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace phem {

constexpr uint64_t KEY_NONE = 0xFFFFFFFFFFFFFFFFull;
constexpr uint32_t U32_NONE = 0xFFFFFFFFu;
constexpr int WAVE = 64;

typedef unsigned long long ull;

// Order key of a 32-bit weight pattern: identity for unsigned weights; for
// float32 the numeric order of the reference's comparator
// ((a.w < b.w) || (a.w == b.w && a.id < b.id), MST/src/utils.cu:801-802,
// MST/src/bench/serialMST.C:67-68): -0.0 and +0.0 are one weight (ties by
// id), negatives below positives.  NaN, which that comparator leaves
// unordered, is defined here as one weight above +inf.
__device__ __forceinline__ uint32_t wkey(uint32_t w, int mode) {
  if (!mode) return w;
  if ((w & 0x7FFFFFFFu) == 0) return 0x80000000u;           // +-0
  if ((w & 0x7FFFFFFFu) > 0x7F800000u) return 0xFFFFFFFFu;  // NaN
  return (w & 0x80000000u) ? ~w : (w | 0x80000000u);
}
__device__ __forceinline__ uint32_t wkey_inv(uint32_t k, int mode) {
  return mode ? ((k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k) : k;
}

// Order of the edges: bit 0 of `ord` = float32 weights, bit 1 = maximum
// spanning forest.  The minimum order is (wkey(w), id) ascending; the maximum
// order is (w, id) descending, encoded as (~wkey(w), 0xFFFFFFFE - id)
// ascending so that one minimum search serves both.  Ids are < 2^32 - 2, so
// no key of either order equals the all-ones "no edge" value.
constexpr int ORD_F32 = 1;
constexpr int ORD_MAX = 2;

__device__ __forceinline__ uint64_t make_key(uint32_t w, uint64_t id, int ord) {
  const uint32_t k = wkey(w, ord & ORD_F32);
  if (ord & ORD_MAX)
    return ((uint64_t)(~k) << 32) | (uint32_t)(0xFFFFFFFEu - (uint32_t)id);
  return ((uint64_t)k << 32) | (uint32_t)id;
}
__device__ __forceinline__ uint32_t key_id(uint64_t key, int ord) {
  return (ord & ORD_MAX) ? 0xFFFFFFFEu - (uint32_t)key : (uint32_t)key;
}
// the weight bit pattern a key was made from
__device__ __forceinline__ uint32_t key_weight(uint64_t key, int ord) {
  const uint32_t k = (uint32_t)(key >> 32);
  return wkey_inv((ord & ORD_MAX) ? ~k : k, ord & ORD_F32);
}

// ---- wave64 helpers --------------------------------------------------------

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// Number of set lanes below this lane in a 64-bit ballot.
__device__ __forceinline__ uint32_t lanes_below(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// Wave-aggregated append: every lane with `pred` gets a distinct slot of
// `*counter`; one atomic per wave.  Returns the slot (valid where pred).
__device__ __forceinline__ uint64_t wave_append(bool pred, ull *counter) {
  uint64_t mask = __ballot(pred);
  if (mask == 0) return 0;
  uint32_t leader = __ffsll((unsigned long long)mask) - 1;
  ull base = 0;
  if (lane_id() == leader) base = atomicAdd(counter, (ull)__popcll(mask));
  base = __shfl(base, leader);
  return base + lanes_below(mask);
}

// ---- segmented append buffers ---------------------------------------------
//
// A buffer of KSEG segments: segment k holds cnt[k * CSTR] elements starting
// at element k * cap.  A producer appends the output of its tile v to segment
// v mod KSEG with one atomic on that segment's counter.  The counters sit on
// separate 128-byte lines: atomics on one address serialize (about 12 ns each
// on MI355X), atomics spread over 64 lines do not.

constexpr int KSEG = 64;
constexpr int CSTR = 16;  // ull per counter slot (128 B)

// Per-block view of a segmented input cut into tiles of T elements.
struct SegView {
  uint32_t tpref[KSEG + 1];  // tiles before segment k
  uint32_t len[KSEG];
};

// All threads call; the first wave reads the KSEG counters.
template <int T>
__device__ __forceinline__ void seg_load(SegView &sv, const ull *cnt) {
  if (threadIdx.x < KSEG) {
    const uint32_t k = threadIdx.x;
    const uint32_t l = (uint32_t)cnt[k * CSTR];
    uint32_t x = (l + T - 1) / T;
#pragma unroll
    for (int d = 1; d < KSEG; d <<= 1) {
      uint32_t y = __shfl_up(x, d);
      if (k >= (uint32_t)d) x += y;
    }
    sv.len[k] = l;
    sv.tpref[k + 1] = x;
    if (k == 0) sv.tpref[0] = 0;
  }
  __syncthreads();
}

// Tile v of the segmented input: first element and length.
template <int T>
__device__ __forceinline__ bool seg_tile(const SegView &sv, uint32_t v,
                                         uint64_t cap, uint64_t *start,
                                         uint32_t *len) {
  if (v >= sv.tpref[KSEG]) return false;
  int lo = 0, hi = KSEG - 1;  // last k with tpref[k] <= v
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (sv.tpref[mid] <= v) lo = mid; else hi = mid - 1;
  }
  const uint32_t off = (v - sv.tpref[lo]) * T;
  *start = (uint64_t)lo * cap + off;
  const uint32_t rem = sv.len[lo] - off;
  *len = rem < (uint32_t)T ? rem : (uint32_t)T;
  return true;
}

// Block-aggregated append of IPT items per thread to one segment: one global
// atomic per call.  Every thread of the block must call it (uniform control
// flow).  slot[j] receives the element index of item j where pred[j] holds.
template <int BLOCK, int IPT>
__device__ __forceinline__ void block_append(const bool (&pred)[IPT],
                                             uint64_t (&slot)[IPT],
                                             ull *counter, uint64_t seg_base) {
  constexpr int NW = BLOCK / WAVE;
  __shared__ uint32_t s_c[IPT * NW];
  __shared__ ull s_base;
  const uint32_t lane = lane_id(), wv = threadIdx.x / WAVE;
  uint64_t mask[IPT];
#pragma unroll
  for (int j = 0; j < IPT; j++) {
    mask[j] = __ballot(pred[j]);
    if (lane == 0) s_c[j * NW + wv] = (uint32_t)__popcll(mask[j]);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (int k = 0; k < IPT * NW; k++) {
      uint32_t c = s_c[k];
      s_c[k] = t;
      t += c;
    }
    s_base = t ? atomicAdd(counter, (ull)t) : 0ull;
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < IPT; j++)
    slot[j] = seg_base + s_base + s_c[j * NW + wv] + lanes_below(mask[j]);
  __syncthreads();
}

// ---- ranges of concatenated graphs ----------------------------------------
//
// Graph g of a batch owns elements [off[g], off[g + 1]) of a concatenated
// vertex or edge range.

// index g with off[g] <= x < off[g + 1] (off ascending, off[0] = 0); empty
// ranges are skipped because the last such g is taken
__device__ __forceinline__ uint32_t seg_of(const ull *__restrict__ off,
                                           uint32_t k, ull x) {
  uint32_t lo = 0, hi = k - 1;
  while (lo < hi) {
    const uint32_t mid = (lo + hi + 1) >> 1;
    if (off[mid] <= x) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// Graph-of-tile maps: map[t] = the graph holding element t * MT of the
// concatenated range (one binary search per MT elements).  An element x of
// tile t then belongs to graph g = map[t] advanced while off[g + 1] <= x.
constexpr uint32_t MT = 256;

__device__ __forceinline__ uint32_t graph_of(const uint32_t *__restrict__ map,
                                             const ull *__restrict__ off,
                                             uint32_t k, uint64_t x) {
  uint32_t g = map[x / MT];
  while (g + 1 < k && off[g + 1] <= x) g++;
  return g;
}

// ---- minimum-key store ----------------------------------------------------

// best[c] = min(best[c], key).  The plain read filters out the atomics that
// cannot lower the value (best only decreases, so a stale read is an upper
// bound and only costs an unnecessary atomic).
__device__ __forceinline__ void min_store(ull *best, uint32_t c, uint64_t key) {
  if (key < best[c]) atomicMin(&best[c], (ull)key);
}

// min_store of key[j] to best[c[j]] for every j with on[j].  All pre-check
// reads are issued before the first compare, so a thread waits for one
// memory latency instead of N in a row; they are unconditional (entry 0
// stands in where on[j] is false) so that no branch splits them.  The
// pre-check reads only the high word of best[c]: best only decreases, so a
// stale high word is an upper bound of the current one and key >> 32 <= it
// whenever key < best[c].  best must have at least one entry.
//
// TIES: a key whose high word (the weight) equals the one read is sent to
// the atomic only after the whole entry is read again and found larger, a
// read issued only in waves where some lane has such a tie.  Otherwise every
// offer of an entry's own weight reaches it: with tied weights (a grid's
// integer weights 1..16) a component's many offers of its lightest weight
// queue on its one entry at the memory side, worst on hubs.  The heavy pass
// over the 4096^2 grid took 1.84 ms that way and takes 0.48 ms with the
// check; the light pass over RMAT n = 2^24, m = 2^28 with integer weights
// 1..16, 5.26 and 1.11 ms; with float weights the check is not measurable.
template <int N, bool TIES = false>
__device__ __forceinline__ void min_store_n(ull *best, const uint32_t (&c)[N],
                                            const ull (&key)[N],
                                            const bool (&on)[N]) {
  uint32_t hi[N];
#pragma unroll
  for (int j = 0; j < N; j++) hi[j] = (uint32_t)(best[on[j] ? c[j] : 0u] >> 32);
#pragma unroll
  for (int j = 0; j < N; j++) {
    const uint32_t kh = (uint32_t)(key[j] >> 32);
    if (on[j] && (TIES ? kh < hi[j] : kh <= hi[j])) atomicMin(&best[c[j]], key[j]);
    if (TIES && __ballot(on[j] && kh == hi[j]))
      if (on[j] && kh == hi[j] && key[j] < best[c[j]]) atomicMin(&best[c[j]], key[j]);
  }
}

// min_store_n to both endpoints a[j] and b[j] of every edge j with on[j]
// (TIES as there).
template <int N, bool TIES = false>
__device__ __forceinline__ void min_store_ends(ull *best, const uint32_t (&a)[N],
                                               const uint32_t (&b)[N],
                                               const ull (&key)[N],
                                               const bool (&on)[N]) {
  uint32_t ha[N], hb[N];
#pragma unroll
  for (int j = 0; j < N; j++) {
    ha[j] = (uint32_t)(best[on[j] ? a[j] : 0u] >> 32);
    hb[j] = (uint32_t)(best[on[j] ? b[j] : 0u] >> 32);
  }
#pragma unroll
  for (int j = 0; j < N; j++) {
    const uint32_t kh = (uint32_t)(key[j] >> 32);
    if (on[j]) {
      if (TIES ? kh < ha[j] : kh <= ha[j]) atomicMin(&best[a[j]], key[j]);
      if (TIES ? kh < hb[j] : kh <= hb[j]) atomicMin(&best[b[j]], key[j]);
    }
    if (TIES && __ballot(on[j] && (kh == ha[j] || kh == hb[j]))) {
      if (on[j] && kh == ha[j] && key[j] < best[a[j]]) atomicMin(&best[a[j]], key[j]);
      if (on[j] && kh == hb[j] && key[j] < best[b[j]]) atomicMin(&best[b[j]], key[j]);
    }
  }
}

// ---- union-find over a parent array where every pointer goes to a smaller
// id (roots are the minimum vertex of their tree) ---------------------------
//
// Reads are ordinary cached loads and may observe an older value of an entry
// (another XCD's L2 is not kept coherent within a kernel).  Every value an
// entry ever holds is an ancestor of the entry, so an older value still lies
// on a path to the true root; only device-scope accesses are authoritative:
// the compare-and-swap on a root (a failed one returns the current parent,
// which is strictly smaller) and the device-scope load that may precede it.
// Path-halving stores only write an ancestor into a non-root entry, which no
// compare-and-swap ever targets.

__device__ __forceinline__ uint32_t par_load(const uint32_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void par_store(uint32_t *p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// find with path halving
__device__ __forceinline__ uint32_t uf_find(uint32_t *par, uint32_t x) {
  uint32_t p = par_load(&par[x]);
  while (p != x) {
    uint32_t gp = par_load(&par[p]);
    if (gp == p) return p;
    par_store(&par[x], gp);
    x = gp;
    p = par_load(&par[x]);
  }
  return x;
}

// find with path halving from x whose parent p = par[x] is already loaded
__device__ __forceinline__ uint32_t uf_find_from(uint32_t *par, uint32_t x,
                                                 uint32_t p) {
  while (p != x) {
    uint32_t gp = par_load(&par[p]);
    if (gp == p) return p;
    par_store(&par[x], gp);
    x = gp;
    p = par_load(&par[x]);
  }
  return x;
}

// find without writes (for read-only walks)
__device__ __forceinline__ uint32_t uf_root(const uint32_t *par, uint32_t x) {
  uint32_t p = par[x];
  while (p != x) {
    x = p;
    p = par[x];
  }
  return x;
}

// Union of the trees whose roots were found as a and b: the larger root is
// hooked under the smaller by compare-and-swap (a failed swap means the root
// was hooked meanwhile; the walk goes on from its new root).  Returns the
// hooked root, or U32_NONE if the trees were already one.
//
// CHECK: read the entry at device scope first and swap only if it still
// holds the root.  A root that was hooked long ago can still read as a root
// from another XCD's L2, and every union that walks through it then fails a
// swap on that one entry, which the memory side performs one at a time.
// While a large component grows, its root changes many times (each merge
// with a component holding a smaller id), so the unions of a fresh parent
// array pile their failed swaps onto a few old roots: with random vertex ids
// (RMAT n = 2^24, first 2^24 records) 0.3-0.5 M failed swaps, 80 % of them
// on ten entries, and 2.1-6.1 ms instead of 0.84-0.91 ms with the check.
// Where the large component's root is fixed from the start (it has the
// smallest id) or the unions are few, the check costs about as much as it
// saves, so it is used where most records are unions: over a fresh parent
// array, and in CC passes that skip few records.  With CHECK, `check` can
// still turn it off at run time.
template <bool CHECK = false>
__device__ __forceinline__ uint32_t uf_hook_roots(uint32_t *par, uint32_t a,
                                                  uint32_t b, bool check = true) {
  while (a != b) {
    uint32_t hi = a > b ? a : b;
    uint32_t lo = a ^ b ^ hi;
    uint32_t old = hi;
    if (CHECK && check)
      old = __hip_atomic_load(&par[hi], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (old == hi) {
      old = atomicCAS(&par[hi], hi, lo);
      if (old == hi) return hi;
    }
    // hi was hooked meanwhile; `old` is its current parent
    uint32_t r = uf_find(par, old);
    if (hi == a) a = r; else b = r;
  }
  return U32_NONE;
}

// Union of the trees of a and b (uf_hook_roots of their roots).
template <bool CHECK = false>
__device__ __forceinline__ uint32_t uf_union(uint32_t *par, uint32_t a,
                                             uint32_t b, bool check = true) {
  return uf_hook_roots<CHECK>(par, uf_find(par, a), uf_find(par, b), check);
}

// uf_union of a and b whose parents pa = par[a], pb = par[b] are already
// loaded (callers load the parents of several edges together)
__device__ __forceinline__ uint32_t uf_union_from(uint32_t *par, uint32_t a,
                                                  uint32_t pa, uint32_t b,
                                                  uint32_t pb) {
  return uf_hook_roots(par, uf_find_from(par, a, pa), uf_find_from(par, b, pb));
}

// Exact weight of one edge as an integer: the unsigned weight itself
// (wmode 0), or w * 2^24 for a float32 weight that is non-negative and a
// multiple of 2^-24 (false otherwise).
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

// ---- counter-based random streams (splitmix64; Steele, Lea, Flood 2014) ----

constexpr uint64_t GOLDEN = 0x9E3779B97F4A7C15ull;

__host__ __device__ __forceinline__ uint64_t mix64(uint64_t x) {
  uint64_t z = x + GOLDEN;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__host__ __device__ __forceinline__ uint64_t stream_base(uint64_t seed,
                                                         uint64_t stream) {
  return mix64(seed * GOLDEN + stream);
}
__host__ __device__ __forceinline__ uint32_t mulhi_n(uint64_t r, uint32_t n) {
  return (uint32_t)(((r >> 32) * (uint64_t)n) >> 32);
}
// weight k * 2^-24 with k in [0, 2^24) (or [1, 2^24] if positive)
__host__ __device__ __forceinline__ uint32_t weight_bits(uint64_t r,
                                                         int positive) {
  uint32_t k = (uint32_t)(r >> 40) + (positive ? 1u : 0u);
  float f = (float)k * 5.9604644775390625e-08f;  // 2^-24
  return __builtin_bit_cast(uint32_t, f);
}

}  // namespace phem
