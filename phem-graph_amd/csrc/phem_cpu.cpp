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
// Kruskal: edge positions sorted stably by the 32-bit order key of the weight
// (four 8-bit LSD passes; each pass histograms, scans and scatters with
// several threads over contiguous blocks, so equal keys keep position order),
// then one union-find sweep in that order.
#include <algorithm>
#include <atomic>
#include <cstring>
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

// Stable LSD radix sort of positions 0..m-1 by key[pos].
void radix_positions(const std::vector<uint32_t> &key, std::vector<uint32_t> &pos,
                     unsigned T) {
  const uint64_t m = key.size();
  std::vector<uint32_t> tmp(m);
  pos.resize(m);
  for (uint64_t i = 0; i < m; i++) pos[i] = (uint32_t)i;
  std::vector<uint64_t> hist((size_t)T * 256);
  for (int shift = 0; shift < 32; shift += 8) {
    std::fill(hist.begin(), hist.end(), 0);
    auto block = [&](unsigned t, uint64_t *b, uint64_t *e) {
      *b = m * t / T;
      *e = m * (t + 1) / T;
    };
    auto count = [&](unsigned t) {
      uint64_t b, e;
      block(t, &b, &e);
      uint64_t *h = &hist[(size_t)t * 256];
      for (uint64_t i = b; i < e; i++) h[(key[pos[i]] >> shift) & 255u]++;
    };
    std::vector<std::thread> ts;
    for (unsigned t = 1; t < T; t++) ts.emplace_back(count, t);
    count(0);
    for (auto &t : ts) t.join();
    ts.clear();
    // exclusive offsets: digit-major, then block order (stability)
    uint64_t run = 0;
    for (int d = 0; d < 256; d++)
      for (unsigned t = 0; t < T; t++) {
        const uint64_t c = hist[(size_t)t * 256 + d];
        hist[(size_t)t * 256 + d] = run;
        run += c;
      }
    auto scatter = [&](unsigned t) {
      uint64_t b, e;
      block(t, &b, &e);
      uint64_t *h = &hist[(size_t)t * 256];
      for (uint64_t i = b; i < e; i++) {
        const uint32_t p = pos[i];
        tmp[h[(key[p] >> shift) & 255u]++] = p;
      }
    };
    for (unsigned t = 1; t < T; t++) ts.emplace_back(scatter, t);
    scatter(0);
    for (auto &t : ts) t.join();
    pos.swap(tmp);
  }
}

uint32_t find_root(std::vector<uint32_t> &par, uint32_t x) {
  while (par[x] != x) {
    par[x] = par[par[x]];
    x = par[x];
  }
  return x;
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
  std::vector<uint32_t> key(m), pos;
  for (uint64_t i = 0; i < m; i++) key[i] = order_key(w[i], wmode);
  radix_positions(key, pos, T);
  std::vector<uint32_t> par(n);
  for (uint32_t x = 0; x < n; x++) par[x] = x;
  std::vector<uint32_t> chosen;
  chosen.reserve(n ? n - 1 : 0);
  for (uint64_t k = 0; k < m && chosen.size() + 1 < (uint64_t)n; k++) {
    const uint32_t i = pos[k], a = u[i], b = v[i];
    if (a == b || a >= n || b >= n) continue;
    const uint32_t ra = find_root(par, a), rb = find_root(par, b);
    if (ra == rb) continue;
    par[ra > rb ? ra : rb] = ra > rb ? rb : ra;
    chosen.push_back(i);
  }
  std::sort(chosen.begin(), chosen.end());
  if (!chosen.empty()) std::memcpy(ids_out, chosen.data(), 4 * chosen.size());
  *count = chosen.size();
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
