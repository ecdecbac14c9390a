// phem_common.hip -- tuning knobs, profiler, version and device check.
#include <atomic>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <cmath>
#include <initializer_list>
#include <limits>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "phem_internal.h"

namespace phem {

// Process-wide tunables: built-in defaults, overridden by the environment at
// first use, then by phem_set_tuning.
namespace {
std::atomic<unsigned> g_blocks_per_cu{0};  // 0: not initialised
std::atomic<double> g_light_factor{std::numeric_limits<double>::quiet_NaN()};
std::atomic<double> g_sample_factor{std::numeric_limits<double>::quiet_NaN()};
std::once_flag g_tune_once;
// kernel shapes (built-in defaults: the shapes measured fastest at RMAT
// n = 2^24, m = 2^28; the filtered CC unions with 1024-thread blocks and a
// 64 KB LDS summary took 1.79 ms against 1.97 ms with 512 and 2.09 ms with
// 256 threads; the MST input passes in 512-thread blocks with 4 edges per
// thread, 7 waves per SIMD, took the step from 8.71 to 8.67 ms against 256 x
// 8 at 4 waves, mean of 3, and the served mix from 15.3-15.6 to 15.9-17.6 G
// edges/s)
std::atomic<uint32_t> g_mst_block{512}, g_mst_ept{4}, g_mst_ept_compact{16},
    g_mst_compact_log2{22}, g_cc_block{256}, g_cc_ept{16}, g_cc_ept_filter{8},
    g_cc_block_filter{1024};
std::atomic<double> g_mst_split_ratio{1.5};
// host share of the chunked solve: none by default.  Measured on MI355X
// with 16 host threads (RMAT 2^24 / 2^28 in chunks of 2^26; RMAT 2^22 / 2^24
// plus 2^28 uniform edges in chunks of 2^24): the device path runs at
// 2.9-3.7 G edges/s including the host-to-device copies, the host reduces
// 0.02-0.1 G edges/s, and every host share tried (2^20 edges up to the
// whole last chunk) made the solve slower.
std::atomic<double> g_cpu_share{0.0};
// light factor of degree-skewed graphs (measured on MI355X, MST ms at light
// factor 0.5 / 0.625 / 1.0: RMAT 2^24 / 2^28 6.17 / 6.15 / 6.44, the same
// with permuted ids 6.59 / 6.64 / 7.11, RMAT 2^20 with 2^22 edges 0.84 /
// 0.87 / 0.91; uniform graphs need 1.0: 2^24 / 2^28 takes 77.8 / 47.5 /
// 12.2 ms)
std::atomic<double> g_lf_skewed{0.625};

bool one_of(uint32_t x, std::initializer_list<uint32_t> allowed) {
  for (uint32_t a : allowed)
    if (x == a) return true;
  return false;
}

void tuning_init() {
  std::call_once(g_tune_once, [] {
    unsigned per_cu = 32;  // measured best at RMAT 2^24 / 2^28 (phem_tune)
    if (const char *e = getenv("PHEM_BLOCKS_PER_CU")) {
      const int x = atoi(e);
      if (x > 0 && x <= 64) per_cu = (unsigned)x;
    }
    unsigned z = 0;
    g_blocks_per_cu.compare_exchange_strong(z, per_cu);
    double lf = 1.0, sf = 1.0;
    if (const char *e = getenv("PHEM_MST_LIGHT_FACTOR")) lf = atof(e);
    if (const char *e = getenv("PHEM_CC_SAMPLE_FACTOR")) sf = atof(e);
    // a value set before first use wins over the environment
    double cur = g_light_factor.load();
    if (std::isnan(cur)) g_light_factor.compare_exchange_strong(cur, lf);
    cur = g_sample_factor.load();
    if (std::isnan(cur)) g_sample_factor.compare_exchange_strong(cur, sf);
  });
}
}  // namespace

unsigned resident_blocks() {
  static std::atomic<int> cus_cache[64];
  tuning_init();
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) dev = 0;
  int cus = cus_cache[dev].load(std::memory_order_relaxed);
  if (!cus) {
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) !=
            hipSuccess || cus <= 0)
      cus = 256;
    cus_cache[dev].store(cus, std::memory_order_relaxed);
  }
  return (unsigned)cus * g_blocks_per_cu.load(std::memory_order_relaxed);
}

double mst_default_light_factor() {
  tuning_init();
  return g_light_factor.load(std::memory_order_relaxed);
}

double cc_default_sample_factor() {
  tuning_init();
  return g_sample_factor.load(std::memory_order_relaxed);
}

KernelShape kernel_shape() {
  KernelShape k;
  k.mst_block = g_mst_block.load(std::memory_order_relaxed);
  k.mst_ept = g_mst_ept.load(std::memory_order_relaxed);
  k.mst_ept_compact = g_mst_ept_compact.load(std::memory_order_relaxed);
  k.mst_compact_log2 = g_mst_compact_log2.load(std::memory_order_relaxed);
  k.mst_split_ratio = g_mst_split_ratio.load(std::memory_order_relaxed);
  k.cc_block = g_cc_block.load(std::memory_order_relaxed);
  k.cc_ept = g_cc_ept.load(std::memory_order_relaxed);
  k.cc_ept_filter = g_cc_ept_filter.load(std::memory_order_relaxed);
  k.cc_block_filter = g_cc_block_filter.load(std::memory_order_relaxed);
  k.cpu_share = g_cpu_share.load(std::memory_order_relaxed);
  k.mst_light_factor_skewed = g_lf_skewed.load(std::memory_order_relaxed);
  return k;
}

// ---- shared pinned host buffers ---------------------------------------------

namespace {
struct PinnedBuf {
  void *p;
  size_t cap;
};
std::mutex g_pin_mu;
std::vector<PinnedBuf> g_pin_free;  // buffers not in use
std::vector<PinnedBuf> g_pin_used;  // buffers handed out
}  // namespace

void *pinned_acquire(size_t bytes) {
  if (bytes < 4096) bytes = 4096;
  {
    std::lock_guard<std::mutex> lk(g_pin_mu);
    // the smallest free buffer that is large enough
    size_t best = g_pin_free.size();
    for (size_t i = 0; i < g_pin_free.size(); i++)
      if (g_pin_free[i].cap >= bytes &&
          (best == g_pin_free.size() || g_pin_free[i].cap < g_pin_free[best].cap))
        best = i;
    if (best < g_pin_free.size()) {
      PinnedBuf b = g_pin_free[best];
      g_pin_free.erase(g_pin_free.begin() + best);
      g_pin_used.push_back(b);
      return b.p;
    }
  }
  void *p = nullptr;
  if (hipHostMalloc(&p, bytes) != hipSuccess) {
    (void)hipGetLastError();
    return nullptr;
  }
  std::lock_guard<std::mutex> lk(g_pin_mu);
  g_pin_used.push_back(PinnedBuf{p, bytes});
  return p;
}

void pinned_release(void *p) {
  if (!p) return;
  std::lock_guard<std::mutex> lk(g_pin_mu);
  for (size_t i = 0; i < g_pin_used.size(); i++)
    if (g_pin_used[i].p == p) {
      g_pin_free.push_back(g_pin_used[i]);
      g_pin_used.erase(g_pin_used.begin() + i);
      return;
    }
}

// ---- per-kernel event profiler --------------------------------------------

namespace {
struct Pending {
  const char *name;
  hipEvent_t a, b;
};
struct Acc {
  std::string name;
  uint64_t launches = 0;
  double ms = 0.0;
};
std::atomic<bool> g_prof{false};
std::mutex g_prof_mu;
std::string g_prof_only;  // record only this kernel (empty: all); under g_prof_mu
std::vector<Pending> g_pending;
std::vector<std::pair<hipEvent_t, hipEvent_t>> g_free_events;
std::vector<Acc> g_acc;

void prof_resolve_locked() {
  for (const Pending &p : g_pending) {
    float ms = 0.f;
    if (hipEventSynchronize(p.b) == hipSuccess &&
        hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
      Acc *acc = nullptr;
      for (Acc &x : g_acc)
        if (x.name == p.name) acc = &x;
      if (!acc) {
        g_acc.push_back(Acc{p.name});
        acc = &g_acc.back();
      }
      acc->launches++;
      acc->ms += ms;
    }
    g_free_events.emplace_back(p.a, p.b);
  }
  g_pending.clear();
}
}  // namespace

ProfScope::ProfScope(const char *name, hipStream_t s) : name_(name), s_(s) {
  if (!g_prof.load(std::memory_order_relaxed)) return;
  // launches recorded into a graph are not timed
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(s, &cap) != hipSuccess ||
      cap != hipStreamCaptureStatusNone)
    return;
  std::lock_guard<std::mutex> lk(g_prof_mu);
  if (!g_prof_only.empty() && g_prof_only != name) return;
  if (!g_free_events.empty()) {
    a_ = g_free_events.back().first;
    b_ = g_free_events.back().second;
    g_free_events.pop_back();
  } else if (hipEventCreate(&a_) != hipSuccess ||
             hipEventCreate(&b_) != hipSuccess) {
    return;
  }
  on_ = hipEventRecord(a_, s_) == hipSuccess;
}

ProfScope::~ProfScope() {
  if (!on_) return;
  (void)hipEventRecord(b_, s_);
  std::lock_guard<std::mutex> lk(g_prof_mu);
  g_pending.push_back(Pending{name_, a_, b_});
}

}  // namespace phem

extern "C" int phem_prof_enable(int on) {
  phem::g_prof.store(on != 0);
  return PHEM_OK;
}

extern "C" int phem_prof_only(const char *name) {
  std::lock_guard<std::mutex> lk(phem::g_prof_mu);
  phem::g_prof_only = name ? name : "";
  return PHEM_OK;
}

extern "C" int phem_prof_reset(void) {
  std::lock_guard<std::mutex> lk(phem::g_prof_mu);
  phem::prof_resolve_locked();
  phem::g_acc.clear();
  return PHEM_OK;
}

extern "C" int phem_prof_count(void) {
  std::lock_guard<std::mutex> lk(phem::g_prof_mu);
  phem::prof_resolve_locked();
  return (int)phem::g_acc.size();
}

extern "C" int phem_prof_get(int i, char *name, int name_cap,
                             uint64_t *launches, double *total_ms) {
  std::lock_guard<std::mutex> lk(phem::g_prof_mu);
  if (i < 0 || i >= (int)phem::g_acc.size())
    return phem::set_err(PHEM_EINVAL, "profile entry %d out of range", i);
  const phem::Acc &a = phem::g_acc[i];
  if (name && name_cap > 0) {
    std::strncpy(name, a.name.c_str(), name_cap - 1);
    name[name_cap - 1] = 0;
  }
  if (launches) *launches = a.launches;
  if (total_ms) *total_ms = a.ms;
  return PHEM_OK;
}

extern "C" const char *phem_version(void) { return "phem-mi355x 0.1 (gfx950)"; }

extern "C" int phem_device_check(int *device_count) {
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (device_count) *device_count = (e == hipSuccess) ? n : 0;
  if (e != hipSuccess || n == 0)
    return phem::set_err(PHEM_ENODEV, "no HIP device visible (%s)",
                         hipGetErrorString(e));
  int dev = 0;
  PHEM_HIP(hipGetDevice(&dev));
  hipDeviceProp_t p;
  PHEM_HIP(hipGetDeviceProperties(&p, dev));
  if (std::strncmp(p.gcnArchName, "gfx950", 6) != 0)
    return phem::set_err(PHEM_ENODEV, "device %d is %s, library built for gfx950",
                         dev, p.gcnArchName);
  return PHEM_OK;
}

extern "C" int phem_get_tuning(phem_tuning *t) {
  if (!t) return phem::set_err(PHEM_EINVAL, "t is NULL");
  phem::tuning_init();
  t->blocks_per_cu = phem::g_blocks_per_cu.load();
  t->reserved = 0;
  t->mst_light_factor = phem::g_light_factor.load();
  t->cc_sample_factor = phem::g_sample_factor.load();
  const phem::KernelShape k = phem::kernel_shape();
  t->mst_block = k.mst_block;
  t->mst_ept = k.mst_ept;
  t->mst_ept_compact = k.mst_ept_compact;
  t->mst_compact_log2 = k.mst_compact_log2;
  t->mst_split_ratio = k.mst_split_ratio;
  t->cc_block = k.cc_block;
  t->cc_ept = k.cc_ept;
  t->cc_ept_filter = k.cc_ept_filter;
  t->cc_block_filter = k.cc_block_filter;
  t->cpu_share = k.cpu_share;
  t->mst_light_factor_skewed = k.mst_light_factor_skewed;
  return PHEM_OK;
}

extern "C" int phem_set_tuning(const phem_tuning *t) {
  if (!t) return phem::set_err(PHEM_EINVAL, "t is NULL");
  using phem::one_of;
  if (t->blocks_per_cu > 64)
    return phem::set_err(PHEM_EINVAL, "blocks_per_cu %u beyond 64", t->blocks_per_cu);
  if (t->mst_block && !one_of(t->mst_block, {256, 512}))
    return phem::set_err(PHEM_EINVAL, "mst_block %u: 256 or 512", t->mst_block);
  if (t->mst_ept && !one_of(t->mst_ept, {4, 8, 16}))
    return phem::set_err(PHEM_EINVAL, "mst_ept %u: 4, 8 or 16", t->mst_ept);
  {
    // at most 4096 edges per tile (the segment capacities' bound)
    const uint32_t b = t->mst_block ? t->mst_block : phem::g_mst_block.load();
    const uint32_t e = t->mst_ept ? t->mst_ept : phem::g_mst_ept.load();
    if (b * e > 4096)
      return phem::set_err(PHEM_EINVAL, "mst_block %u x mst_ept %u beyond 4096 edges",
                           b, e);
  }
  if (t->mst_ept_compact && !one_of(t->mst_ept_compact, {8, 16}))
    return phem::set_err(PHEM_EINVAL, "mst_ept_compact %u: 8 or 16",
                         t->mst_ept_compact);
  if (t->mst_compact_log2 > 40)
    return phem::set_err(PHEM_EINVAL, "mst_compact_log2 %u beyond 40",
                         t->mst_compact_log2);
  if (!std::isnan(t->mst_split_ratio) && !(t->mst_split_ratio > 1.0))
    return phem::set_err(PHEM_EINVAL, "mst_split_ratio must be above 1");
  if (t->cc_block && !one_of(t->cc_block, {256, 512}))
    return phem::set_err(PHEM_EINVAL, "cc_block %u: 256 or 512", t->cc_block);
  if (t->cc_ept && !one_of(t->cc_ept, {8, 16, 32}))
    return phem::set_err(PHEM_EINVAL, "cc_ept %u: 8, 16 or 32", t->cc_ept);
  if (t->cc_ept_filter && !one_of(t->cc_ept_filter, {4, 8, 16}))
    return phem::set_err(PHEM_EINVAL, "cc_ept_filter %u: 4, 8 or 16",
                         t->cc_ept_filter);
  if (t->cc_block_filter && !one_of(t->cc_block_filter, {256, 512, 1024}))
    return phem::set_err(PHEM_EINVAL, "cc_block_filter %u: 256, 512 or 1024",
                         t->cc_block_filter);
  {
    const uint32_t b = t->cc_block_filter ? t->cc_block_filter
                                          : phem::g_cc_block_filter.load();
    const uint32_t e = t->cc_ept_filter ? t->cc_ept_filter : phem::g_cc_ept_filter.load();
    if (b == 1024 && e > 8)
      return phem::set_err(PHEM_EINVAL, "cc_block_filter 1024 takes cc_ept_filter 4 or 8");
  }
  if (!std::isnan(t->cpu_share) && !(t->cpu_share >= 0.0 && t->cpu_share <= 1e6))
    return phem::set_err(PHEM_EINVAL, "cpu_share must be in [0, 1e6]");
  if (!std::isnan(t->mst_light_factor_skewed) &&
      !(t->mst_light_factor_skewed >= 0.0 && t->mst_light_factor_skewed <= 1e6))
    return phem::set_err(PHEM_EINVAL, "mst_light_factor_skewed must be in [0, 1e6]");
  phem::tuning_init();
  if (!std::isnan(t->cpu_share)) phem::g_cpu_share.store(t->cpu_share);
  if (!std::isnan(t->mst_light_factor_skewed))
    phem::g_lf_skewed.store(t->mst_light_factor_skewed);
  if (t->mst_block) phem::g_mst_block.store(t->mst_block);
  if (t->mst_ept) phem::g_mst_ept.store(t->mst_ept);
  if (t->mst_ept_compact) phem::g_mst_ept_compact.store(t->mst_ept_compact);
  if (t->mst_compact_log2) phem::g_mst_compact_log2.store(t->mst_compact_log2);
  if (!std::isnan(t->mst_split_ratio)) phem::g_mst_split_ratio.store(t->mst_split_ratio);
  if (t->cc_block) phem::g_cc_block.store(t->cc_block);
  if (t->cc_ept) phem::g_cc_ept.store(t->cc_ept);
  if (t->cc_ept_filter) phem::g_cc_ept_filter.store(t->cc_ept_filter);
  if (t->cc_block_filter) phem::g_cc_block_filter.store(t->cc_block_filter);
  if (t->blocks_per_cu) phem::g_blocks_per_cu.store(t->blocks_per_cu);
  if (!std::isnan(t->mst_light_factor)) phem::g_light_factor.store(t->mst_light_factor);
  if (!std::isnan(t->cc_sample_factor)) phem::g_sample_factor.store(t->cc_sample_factor);
  return PHEM_OK;
}
