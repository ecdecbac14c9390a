// phem_common.hip -- error reporting, version and device check.
#include <atomic>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "phem_internal.h"

namespace phem {

static thread_local char g_err[512] = "";

int set_err(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

unsigned resident_blocks() {
  static std::atomic<unsigned> cache[64];
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) dev = 0;
  unsigned v = cache[dev].load(std::memory_order_relaxed);
  if (v) return v;
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) !=
          hipSuccess || cus <= 0)
    cus = 256;
  unsigned per_cu = 16;
  if (const char *e = getenv("PHEM_BLOCKS_PER_CU")) {
    const int x = atoi(e);
    if (x > 0 && x <= 64) per_cu = (unsigned)x;
  }
  v = (unsigned)cus * per_cu;
  cache[dev].store(v, std::memory_order_relaxed);
  return v;
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
  std::lock_guard<std::mutex> lk(g_prof_mu);
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

extern "C" const char *phem_last_error(void) { return phem::g_err; }

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
