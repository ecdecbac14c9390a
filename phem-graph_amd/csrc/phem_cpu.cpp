// phem_cpu.cpp -- the library's host (CPU) backend: minimum spanning forest
// and connected components over the same argument conventions as the device
// entry points.
//
// The minimum spanning forest also pre-filters a chunk on the CPU while the
// GPU streams the other chunks (the reference's heterogeneous streamline, MST/src/kernels.cu:1503-1788: the
// last chunk is replaced by its own forest before the GPU folds it in, which
// keeps the result exact because an edge outside the forest of a subgraph is
// the heaviest edge of a cycle and stays outside every larger forest).
//
// Filter-Kruskal, the steps of the reference's PBBS mst() (MST/src/
// utils.cu:794-841): a threshold near the 5n/4-th smallest (weight, id) key
// from a sample; Kruskal over the edges below it; the edges at or above it
// filtered in parallel by one union-find lookup per endpoint; Kruskal over
// the survivors with the same union-find.  Kruskal here is one sweep over
// (key, position, endpoints) records in key order (sorted by a multithreaded
// stable radix sort, so equal weights keep position order); the sweeps touch
// about 5n/4 edges plus the survivors, the filter touches the rest once.
#include <algorithm>
#include <atomic>
#include <cstring>
#include <memory>
#include <new>
#include <thread>
#include <vector>

#include "phem.h"
#include "phem_internal.h"

namespace phem {
namespace {

// The device order key (phem_dev.h wkey): float32 -0.0 orders as +0.0 and
// every NaN as one value above +inf.
uint32_t order_key(uint32_t w, int wmode) {
  if (wmode == PHEM_W_U32) return w;
  if ((w & 0x7FFFFFFFu) == 0) return 0x80000000u;
  if ((w & 0x7FFFFFFFu) > 0x7F800000u) return 0xFFFFFFFFu;
  return (w & 0x80000000u) ? ~w : (w | 0x80000000u);
}

// Runs f(0) .. f(T-1) on T threads (f(0) on the caller's).
template <class F>
void run_threads(unsigned T, const F &f) {
  std::vector<std::thread> ts;
  for (unsigned t = 1; t < T; t++) ts.emplace_back(f, t);
  f(0);
  for (auto &t : ts) t.join();
}

// An edge as the sweeps read it: order key, input position, endpoints.
struct Rec {
  uint32_t key, pos, a, b;
};

// Stable LSD radix sort of records by key (T threads over contiguous
// blocks; a pass is skipped when every key agrees on its digit).
void radix_records(std::vector<Rec> &r, unsigned T) {
  const uint64_t m = r.size();
  if (m < (1u << 16)) T = 1;
  std::vector<Rec> tmp(m);
  std::vector<uint64_t> hist((size_t)T * 256);
  auto lo = [&](unsigned t) { return m * t / T; };
  for (int shift = 0; shift < 32; shift += 8) {
    std::fill(hist.begin(), hist.end(), 0);
    run_threads(T, [&](unsigned t) {
      uint64_t *h = &hist[(size_t)t * 256];
      for (uint64_t i = lo(t); i < lo(t + 1); i++) h[(r[i].key >> shift) & 255u]++;
    });
    // exclusive offsets: digit-major, then block order (stability)
    uint64_t run = 0;
    int digits = 0;
    for (int d = 0; d < 256; d++) {
      uint64_t dc = 0;
      for (unsigned t = 0; t < T; t++) {
        const uint64_t c = hist[(size_t)t * 256 + d];
        hist[(size_t)t * 256 + d] = run;
        run += c;
        dc += c;
      }
      digits += dc != 0;
    }
    if (digits <= 1) continue;
    run_threads(T, [&](unsigned t) {
      uint64_t *h = &hist[(size_t)t * 256];
      for (uint64_t i = lo(t); i < lo(t + 1); i++) tmp[h[(r[i].key >> shift) & 255u]++] = r[i];
    });
    r.swap(tmp);
  }
}

uint32_t find_root(uint32_t *par, uint32_t x) {
  while (par[x] != x) {
    par[x] = par[par[x]];
    x = par[x];
  }
  return x;
}

// Records of the edges i in [0, m) with keep(i), in position order: each
// thread gathers its block's records, then the blocks are concatenated.
template <class Keep>
void gather_records(uint64_t m, const uint32_t *key, const uint32_t *u,
                    const uint32_t *v, unsigned T, const Keep &keep,
                    std::vector<Rec> &out) {
  if (m < (1u << 16)) T = 1;
  std::vector<std::vector<Rec>> part(T);
  run_threads(T, [&](unsigned t) {
    std::vector<Rec> &o = part[t];
    for (uint64_t i = m * t / T; i < m * (t + 1) / T; i++)
      if (keep(i)) o.push_back(Rec{key[i], (uint32_t)i, u[i], v[i]});
  });
  std::vector<uint64_t> at(T + 1, 0);
  for (unsigned t = 0; t < T; t++) at[t + 1] = at[t] + part[t].size();
  out.resize(at[T]);
  run_threads(T, [&](unsigned t) {
    if (!part[t].empty())
      std::memcpy(out.data() + at[t], part[t].data(), sizeof(Rec) * part[t].size());
    std::vector<Rec>().swap(part[t]);
  });
}

// Kruskal sweep over records in key order: every edge that joins two trees
// is chosen (flag[pos] = 1).  The union-find entries of the records a few
// steps ahead are prefetched (the sweep is one dependent chain of reads).
uint64_t kruskal_sweep(const std::vector<Rec> &r, uint32_t n, uint32_t *par,
                       uint64_t chosen, uint8_t *flag) {
  constexpr uint64_t AHEAD = 16;
  const uint64_t k_end = r.size();
  for (uint64_t k = 0; k < k_end && chosen + 1 < (uint64_t)n; k++) {
    if (k + AHEAD < k_end) {
      const Rec &f = r[k + AHEAD];
      if (f.a < n) __builtin_prefetch(&par[f.a]);
      if (f.b < n) __builtin_prefetch(&par[f.b]);
    }
    const Rec &e = r[k];
    if (e.a == e.b || e.a >= n || e.b >= n) continue;
    const uint32_t ra = find_root(par, e.a), rb = find_root(par, e.b);
    if (ra == rb) continue;
    par[ra > rb ? ra : rb] = ra > rb ? rb : ra;
    flag[e.pos] = 1;
    chosen++;
  }
  return chosen;
}

}  // namespace
}  // namespace phem

using namespace phem;

extern "C" int phem_cpu_msf(uint32_t n, uint64_t m, const uint32_t *u,
                            const uint32_t *v, const uint32_t *w, int wmode,
                            uint32_t *ids_out, uint64_t *count, int threads) {
  if (!count) return set_err(PHEM_EINVAL, "count is NULL");
  if (m > PHEM_MAX_EDGES) return set_err(PHEM_EINVAL, "m=%llu beyond limits",
                                         (unsigned long long)m);
  if (wmode != PHEM_W_U32 && wmode != PHEM_W_F32)
    return set_err(PHEM_EINVAL, "unknown weight mode %d", wmode);
  if (m && (!u || !v || !w)) return set_err(PHEM_EINVAL, "NULL edge array");
  if (n && !ids_out) return set_err(PHEM_EINVAL, "ids_out is NULL");
  unsigned T = threads > 0 ? (unsigned)threads : std::thread::hardware_concurrency();
  T = std::max(1u, std::min(T, 64u));
  if (m < (1u << 16)) T = 1;
  // keys and flags are first touched by the threads that use them
  std::unique_ptr<uint32_t[]> key(new (std::nothrow) uint32_t[m ? m : 1]);
  std::unique_ptr<uint8_t[]> flag(new (std::nothrow) uint8_t[m ? m : 1]);
  if (!key || !flag) return set_err(PHEM_ENOMEM, "host allocation failed");
  run_threads(T, [&](unsigned t) {
    for (uint64_t i = m * t / T; i < m * (t + 1) / T; i++) {
      key[i] = order_key(w[i], wmode);
      flag[i] = 0;
    }
  });
  // threshold: the (5n/4 / m)-quantile of a strided sample of the keys;
  // every edge with key < thr precedes every other in (key, id) order
  const uint64_t k_light = std::min<uint64_t>(5ull * n / 4, m);
  uint64_t thr = (uint64_t)UINT32_MAX + 1;  // no filter: every edge is light
  if (m > 2 * k_light && k_light > 0) {
    const uint64_t S = std::min<uint64_t>(m, 1u << 16);
    std::vector<uint32_t> smp(S);
    for (uint64_t j = 0; j < S; j++) smp[j] = key[j * m / S];
    const uint64_t r = std::min<uint64_t>(S - 1, (k_light * S + m - 1) / m);
    std::nth_element(smp.begin(), smp.begin() + r, smp.end());
    thr = smp[r];
  }
  std::vector<uint32_t> par(n);
  for (uint32_t x = 0; x < n; x++) par[x] = x;
  std::vector<Rec> rec;
  const uint32_t *kp = key.get();
  gather_records(m, kp, u, v, T, [&](uint64_t i) { return kp[i] < thr; }, rec);
  radix_records(rec, T);
  uint64_t chosen = kruskal_sweep(rec, n, par.data(), 0, flag.get());
  if (thr <= UINT32_MAX && chosen + 1 < (uint64_t)n) {
    // flatten, then keep the heavy edges whose endpoints are in two trees
    for (uint32_t x = 0; x < n; x++) par[x] = find_root(par.data(), x);
    const uint32_t *fp = par.data();
    gather_records(m, kp, u, v, T, [&](uint64_t i) {
      const uint32_t a = u[i], b = v[i];
      return kp[i] >= thr && a < n && b < n && fp[a] != fp[b];
    }, rec);
    radix_records(rec, T);
    chosen = kruskal_sweep(rec, n, par.data(), chosen, flag.get());
  }
  // the chosen positions, ascending: per-block counts, then a parallel pack
  std::vector<uint64_t> at(T + 1, 0);
  run_threads(T, [&](unsigned t) {
    uint64_t c = 0;
    for (uint64_t i = m * t / T; i < m * (t + 1) / T; i++) c += flag[i];
    at[t + 1] = c;
  });
  for (unsigned t = 0; t < T; t++) at[t + 1] += at[t];
  run_threads(T, [&](unsigned t) {
    uint64_t o = at[t];
    for (uint64_t i = m * t / T; i < m * (t + 1) / T; i++)
      if (flag[i]) ids_out[o++] = (uint32_t)i;
  });
  *count = chosen;
  return PHEM_OK;
}

namespace phem {
namespace {

// Lock-free union-find on host threads: roots are the minimum vertex of
// their tree (the larger root is hooked under the smaller by CAS), finds
// halve paths with relaxed stores of ancestors.
uint32_t cpu_find(std::atomic<uint32_t> *par, uint32_t x) {
  uint32_t p = par[x].load(std::memory_order_relaxed);
  while (p != x) {
    const uint32_t gp = par[p].load(std::memory_order_relaxed);
    if (gp == p) return p;
    par[x].store(gp, std::memory_order_relaxed);
    x = gp;
    p = par[x].load(std::memory_order_relaxed);
  }
  return x;
}

void cpu_union(std::atomic<uint32_t> *par, uint32_t a, uint32_t b) {
  a = cpu_find(par, a);
  b = cpu_find(par, b);
  while (a != b) {
    uint32_t hi = a > b ? a : b, lo = a ^ b ^ hi;
    uint32_t expect = hi;
    if (par[hi].compare_exchange_strong(expect, lo, std::memory_order_acq_rel))
      return;
    const uint32_t r = cpu_find(par, expect);  // hi was hooked meanwhile
    if (hi == a) a = r; else b = r;
  }
}

}  // namespace
}  // namespace phem

extern "C" int phem_cpu_cc(uint32_t n, uint64_t m, const uint32_t *u,
                           const uint32_t *v, uint32_t *labels, int threads) {
  if (m && (!u || !v)) return set_err(PHEM_EINVAL, "NULL edge array");
  if (n && !labels) return set_err(PHEM_EINVAL, "labels is NULL");
  unsigned T = threads > 0 ? (unsigned)threads : std::thread::hardware_concurrency();
  T = std::max(1u, std::min(T, 64u));
  if (m < (1u << 16)) T = 1;
  std::vector<std::atomic<uint32_t>> par(n);
  for (uint32_t x = 0; x < n; x++) par[x].store(x, std::memory_order_relaxed);
  auto run = [&](unsigned t) {
    for (uint64_t i = m * t / T; i < m * (t + 1) / T; i++) {
      const uint32_t a = u[i], b = v[i];
      if (a != b && a < n && b < n) cpu_union(par.data(), a, b);
    }
  };
  std::vector<std::thread> ts;
  for (unsigned t = 1; t < T; t++) ts.emplace_back(run, t);
  run(0);
  for (auto &t : ts) t.join();
  // labels: the root of each vertex (the minimum id of its component)
  for (uint32_t x = 0; x < n; x++) {
    uint32_t r = x;
    while (par[r].load(std::memory_order_relaxed) != r)
      r = par[r].load(std::memory_order_relaxed);
    labels[x] = r;
  }
  return PHEM_OK;
}
