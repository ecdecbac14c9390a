// phem_common.hip -- tuning knobs, profiler, version and device check.
#include <atomic>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <cmath>
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
  return PHEM_OK;
}

extern "C" int phem_set_tuning(const phem_tuning *t) {
  if (!t) return phem::set_err(PHEM_EINVAL, "t is NULL");
  if (t->blocks_per_cu > 64)
    return phem::set_err(PHEM_EINVAL, "blocks_per_cu %u beyond 64", t->blocks_per_cu);
  phem::tuning_init();
  if (t->blocks_per_cu) phem::g_blocks_per_cu.store(t->blocks_per_cu);
  if (!std::isnan(t->mst_light_factor)) phem::g_light_factor.store(t->mst_light_factor);
  if (!std::isnan(t->cc_sample_factor)) phem::g_sample_factor.store(t->cc_sample_factor);
  return PHEM_OK;
}
