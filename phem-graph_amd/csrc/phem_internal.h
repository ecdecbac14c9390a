// phem_internal.h -- host-side helpers shared by the library's translation
// units (error reporting, launch sizing).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <functional>

#include "phem.h"

namespace phem {

// Records a formatted message for phem_last_error() and returns `code`.
int set_err(int code, const char *fmt, ...);

inline uint64_t ceil_div(uint64_t a, uint64_t b) { return (a + b - 1) / b; }
inline uint64_t align_up(uint64_t a, uint64_t b) { return (a + b - 1) / b * b; }

// Blocks launched for the library's looping kernels: compute units x the
// blocks_per_cu knob (32 by default; 256-thread blocks, the dispatcher
// backfills as blocks finish).
unsigned resident_blocks();

// Tunable defaults (phem_set_tuning, else the environment, else built-in).
double mst_default_light_factor();
double cc_default_sample_factor();

// Kernel shapes selected at launch (phem_tuning's mst_* and cc_* shapes).
struct KernelShape {
  uint32_t mst_block, mst_ept, mst_ept_compact, mst_compact_log2;
  double mst_split_ratio;
  uint32_t cc_block, cc_ept, cc_ept_filter, cc_block_filter;
  double cpu_share;
  double mst_light_factor_skewed;
};
KernelShape kernel_shape();

// Grid size for `work` items at `per_block` items per block, capped at the
// resident block count: kernels loop over their work, so blocks stay alive
// across many tiles instead of the dispatcher launching one block per tile.
inline unsigned grid_for(uint64_t work, uint64_t per_block) {
  const uint64_t g = ceil_div(work ? work : 1, per_block);
  const uint64_t cap = resident_blocks();
  return (unsigned)(g < cap ? g : cap);
}

// grid_for for blocks of `threads` threads: the resident cap counts
// 256-thread blocks, so larger blocks get proportionally fewer.
inline unsigned grid_for_threads(uint64_t work, uint64_t per_block,
                                 unsigned threads) {
  const uint64_t g = ceil_div(work ? work : 1, per_block);
  uint64_t cap = (uint64_t)resident_blocks() * 256 / threads;
  if (!cap) cap = 1;
  return (unsigned)(g < cap ? g : cap);
}

// A batch's graphs seen by the solve over their disjoint union (device
// arrays): graph g owns vertices [voff[g], voff[g + 1]) and edges
// [eoff[g], eoff[g + 1]); vmap / emap are the graph-of-tile maps of the two
// ranges.  The solve picks a light threshold per graph (taus, k keys) and a
// giant component per graph (groot, k roots): edges never join two graphs,
// so each graph is split and filtered as if it were solved alone.
struct GraphGroups {
  uint32_t k;
  const unsigned long long *voff, *eoff;
  const uint32_t *vmap, *emap;
  unsigned long long *taus;
  uint32_t *groot;
};

// Rounds launched per phase when a solve is recorded for replay: the first
// phase (canonical rounds after round 0, or the light phase) and the heavy
// phase.  Nothing is read back between rounds; a round launched after its
// phase converged finds no alive edge and does nothing.
struct FixedRounds {
  int light, heavy;
};

// phem_mst_solve over a batch's disjoint union (structure-of-arrays input).
// fr: the fixed-round schedule of a recorded solve (no counter is read, stats
// stay unwritten, the forest size stays in the device word solve_count_word
// names), or NULL.
int mst_solve_groups(uint32_t n, uint64_t m, const uint32_t *d_u,
                     const uint32_t *d_v, const uint32_t *d_w, int wmode,
                     void *d_ws, uint64_t ws_bytes, uint32_t *d_mst_ids,
                     uint32_t *d_labels, const phem_mst_opts *opts,
                     phem_mst_stats *stats, hipStream_t s,
                     const GraphGroups &groups,
                     const FixedRounds *fr = nullptr);

// Device addresses in a solve's workspace (d_ws of phem_mst_solve for n, m):
// the word holding the forest size after the solve, and the counter area
// that fixed_round_stats reads.
const unsigned long long *solve_count_word(const void *d_ws, uint32_t n,
                                           uint64_t m);

// Statistics of a fixed-round solve of (n, m) over workspace d_ws, with the
// schedule fr and light_factor its recording used: enqueues one summing
// kernel and one copy to h_sums (host, pinned or not) via d_sums (device,
// fixed_sums_words() words), waits for them on s, fills *stats as
// phem_mst_solve would and returns PHEM_ENOCONV when a phase still had alive
// edges after its recorded rounds.
int fixed_sums_words();
int fixed_round_stats(const void *d_ws, uint32_t n, uint64_t m, bool split,
                      const FixedRounds &fr, unsigned long long *d_sums,
                      unsigned long long *h_sums, phem_mst_stats *stats,
                      hipStream_t s);

// The light/heavy split decision of a solve of (n, m) with these options.
bool mst_uses_split(uint32_t n, uint64_t m, const phem_mst_opts *opts);

// Work a caller enqueues after the forest ids are written and before the
// solve's final counter read, so that read's wait covers it: d_count is the
// device word that holds the forest size (the ids are in d_mst_ids).
using MstTail = std::function<int(const unsigned long long *d_count)>;

// phem_mst_solve with a tail (structure-of-arrays input).  first_rounds:
// Boruvka rounds to launch before the first counter read (the rounds the
// caller expects; the results do not depend on it; 0: the default batch).
int mst_solve_tail(uint32_t n, uint64_t m, const uint32_t *d_u,
                   const uint32_t *d_v, const uint32_t *d_w, int wmode,
                   void *d_ws, uint64_t ws_bytes, uint32_t *d_mst_ids,
                   uint32_t *d_labels, const phem_mst_opts *opts,
                   phem_mst_stats *stats, hipStream_t s, const MstTail &tail,
                   int first_rounds = 0);

// Pinned host buffers shared by every thread of the process: a free list
// that short-lived threads (server workers) draw from and return to, so no
// thread keeps a buffer of its own after it ends.  Buffers are never freed;
// their number is bounded by the solves that ever ran at the same time.
void *pinned_acquire(size_t bytes);  // nullptr if hipHostMalloc fails
void pinned_release(void *p);

// A pinned buffer held for one scope.  `s` is the stream whose pending
// copies may still read or write it: the release waits for that stream
// unless done() marked it idle (after the caller's own synchronisation).
class PinnedLease {
 public:
  PinnedLease(size_t bytes, hipStream_t s)
      : p_(pinned_acquire(bytes)), s_(s) {}
  ~PinnedLease() {
    if (!p_) return;
    if (!idle_) (void)hipStreamSynchronize(s_);
    pinned_release(p_);
  }
  PinnedLease(const PinnedLease &) = delete;
  PinnedLease &operator=(const PinnedLease &) = delete;
  void *get() const { return p_; }
  void busy() { idle_ = false; }  // a copy into or out of it was enqueued
  void done() { idle_ = true; }   // the stream has been synchronised since

 private:
  void *p_;
  hipStream_t s_;
  bool idle_ = false;
};

// Records HIP events around the launches in its scope when profiling is on.
class ProfScope {
 public:
  ProfScope(const char *name, hipStream_t s);
  ~ProfScope();
  ProfScope(const ProfScope &) = delete;
  ProfScope &operator=(const ProfScope &) = delete;

 private:
  const char *name_;
  hipStream_t s_;
  hipEvent_t a_ = nullptr, b_ = nullptr;
  bool on_ = false;
};

}  // namespace phem

#define PHEM_PROF(name, s) phem::ProfScope phem_prof_scope_(name, s)

// A failed call also leaves its error as the thread's last HIP error; it is
// cleared here so that the next launch check does not report it again.
#define PHEM_HIP(x)                                                     \
  do {                                                                     \
    hipError_t e_ = (x);                                                   \
    if (e_ != hipSuccess) {                                                \
      (void)hipGetLastError();                                             \
      return phem::set_err(PHEM_EHIP, "%s failed: %s (%s:%d)", #x,         \
                           hipGetErrorString(e_), __FILE__, __LINE__);     \
    }                                                                      \
  } while (0)

#define PHEM_LAUNCHED(name)                                                \
  do {                                                                     \
    hipError_t e_ = hipGetLastError();                                     \
    if (e_ != hipSuccess)                                                  \
      return phem::set_err(PHEM_EHIP, "launch of %s failed: %s", name,     \
                           hipGetErrorString(e_));                         \
  } while (0)
