// phem_chunks.hip -- minimum spanning forest of a host-resident edge list
// streamed through the device in chunks, with an optional CPU share.
//
// F = MSF(F u chunk) over the chunks in input order, with F kept in HBM by a
// phem_stream (the reference's streamline path, MST/src/kernels.cu:849-1154).
// Host edges reach the device through a ring of pinned pieces: host threads
// copy a piece out of the caller's (pageable) arrays while the previous
// pieces travel to the device, and the next chunk is staged while the device
// folds in the current one.
//
// CPU share (the reference's heterogeneous streamline, MST/src/kernels.cu:
// 1155-1826): the first cpu_edges edges of the last chunk are reduced on the
// host to their own forest (phem_cpu_msf) while the device folds in the other
// chunks; the last chunk then carries that forest and the rest of the chunk.
// The result is exact: an edge outside the forest of a subgraph is the
// heaviest edge of a cycle and stays outside every larger forest.  Edge ids
// keep their input order through the filter (the forest's ids ascend, the
// unfiltered rest follows), so ties still break by input position.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "phem.h"
#include "phem_copypool.h"
#include "phem_internal.h"

namespace phem {
namespace {

constexpr uint64_t PIECE = 1ull << 22;  // elements (16 MB) per pinned piece
constexpr int RING = 4;                 // pinned pieces in flight

double now_s() {
  return std::chrono::duration<double>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

// Host-to-device staging through RING pinned pieces (from the process's
// pinned pool, returned when the call ends) on one copy stream.
struct Stager {
  hipStream_t cs = nullptr;
  uint32_t *pin[RING] = {};
  hipEvent_t ev[RING] = {};
  bool used[RING] = {};
  int next = 0;
  unsigned threads = 1;
  std::unique_ptr<CopyPool> pool;  // threads - 1 workers besides the caller

  int init(unsigned copy_threads) {
    threads = copy_threads ? copy_threads : 1;
    pool.reset(new CopyPool(threads - 1));
    if (hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess)
      return fail("copy stream");
    for (int i = 0; i < RING; i++) {
      pin[i] = (uint32_t *)pinned_acquire(4 * PIECE);
      if (!pin[i]) return set_err(PHEM_ENOMEM, "pinned staging");
      if (hipEventCreateWithFlags(&ev[i], hipEventDisableTiming) != hipSuccess) {
        ev[i] = nullptr;
        return fail("staging event");
      }
    }
    return PHEM_OK;
  }
  ~Stager() {
    if (cs) (void)hipStreamSynchronize(cs);
    for (int i = 0; i < RING; i++) {
      if (ev[i]) (void)hipEventDestroy(ev[i]);
      pinned_release(pin[i]);  // the copy stream was synchronised above
    }
    if (cs) (void)hipStreamDestroy(cs);
  }
  static int fail(const char *what) {
    const hipError_t e = hipGetLastError();
    return set_err(e == hipSuccess ? PHEM_ENOMEM : PHEM_EHIP, "%s: %s", what,
                   hipGetErrorString(e));
  }
  // Enqueues the copy of k elements of src to dst (device).
  int copy(uint32_t *dst, const uint32_t *src, uint64_t k) {
    for (uint64_t o = 0; o < k; o += PIECE) {
      const uint64_t len = std::min(PIECE, k - o);
      const int r = next;
      next = (next + 1) % RING;
      if (used[r] && hipEventSynchronize(ev[r]) != hipSuccess)
        return set_err(PHEM_EHIP, "staging copy failed");
      uint32_t *p = pin[r];
      const uint32_t *s = src + o;
      const uint64_t T = len >= (1u << 18) ? threads : 1;
      pool->run(T, [&](uint64_t t) {
        const uint64_t b = len * t / T, e = len * (t + 1) / T;
        std::memcpy(p + b, s + b, 4 * (e - b));
      });
      if (hipMemcpyAsync(dst + o, p, 4 * len, hipMemcpyHostToDevice, cs) != hipSuccess ||
          hipEventRecord(ev[r], cs) != hipSuccess)
        return fail("staging copy");
      used[r] = true;
    }
    return PHEM_OK;
  }
};

}  // namespace
}  // namespace phem

using namespace phem;

extern "C" int phem_msf_host_chunks(uint32_t n, uint64_t m, const uint32_t *u,
                                    const uint32_t *v, const uint32_t *w,
                                    int wmode, uint64_t chunk_edges,
                                    uint64_t cpu_edges, uint32_t *ids_out,
                                    uint64_t *count, phem_chunks_stats *stats) {
  const double t0 = now_s();
  if (!count) return set_err(PHEM_EINVAL, "count is NULL");
  *count = 0;
  if (stats) std::memset(stats, 0, sizeof(*stats));
  if (n < 1 || n > PHEM_MAX_VERTICES || m > PHEM_MAX_EDGES)
    return set_err(PHEM_EINVAL, "bad n=%u / m=%llu", n, (unsigned long long)m);
  if (wmode != PHEM_W_U32 && wmode != PHEM_W_F32)
    return set_err(PHEM_EINVAL, "unknown weight mode %d", wmode);
  if (m && (!u || !v || !w)) return set_err(PHEM_EINVAL, "NULL edge array");
  if (!ids_out) return set_err(PHEM_EINVAL, "ids_out is NULL");
  if (chunk_edges < 1) return set_err(PHEM_EINVAL, "chunk_edges is 0");
  if (chunk_edges > m) chunk_edges = m ? m : 1;
  if ((uint64_t)n + chunk_edges > PHEM_MAX_EDGES)
    return set_err(PHEM_EINVAL, "n + chunk_edges exceeds the edge-id range");
  const uint64_t nchunks = m ? (m + chunk_edges - 1) / chunk_edges : 0;
  const uint64_t last = nchunks ? (nchunks - 1) * chunk_edges : 0;
  // the CPU share needs a chunk for the device to fold in meanwhile
  uint64_t cpu_span = 0;
  if (nchunks > 1) {
    if (cpu_edges == PHEM_CPU_EDGES_AUTO)
      cpu_edges = (uint64_t)(kernel_shape().cpu_share * (double)(last));
    cpu_span = std::min(cpu_edges, m - last);
  }

  phem_stream_t *st = nullptr;
  int rc = phem_stream_create(n, chunk_edges, wmode, &st);
  if (rc) return rc;
  std::unique_ptr<phem_stream_t, int (*)(phem_stream_t *)> st_own(st, phem_stream_destroy);
  uint32_t *d_in[2] = {nullptr, nullptr};
  for (int b = 0; b < 2; b++)
    if (hipMalloc(&d_in[b], 12 * chunk_edges) != hipSuccess) {
      (void)hipGetLastError();
      if (d_in[0]) (void)hipFree(d_in[0]);
      return set_err(PHEM_ENOMEM, "device chunk buffers (%llu B)",
                     (unsigned long long)(24 * chunk_edges));
    }
  struct FreeIn {
    uint32_t **p;
    ~FreeIn() {
      for (int b = 0; b < 2; b++) (void)hipFree(p[b]);
    }
  } free_in{d_in};
  hipStream_t ks = nullptr;
  if (hipStreamCreateWithFlags(&ks, hipStreamNonBlocking) != hipSuccess)
    return Stager::fail("solve stream");
  struct FreeStream {
    hipStream_t s;
    ~FreeStream() {
      (void)hipStreamSynchronize(s);
      (void)hipStreamDestroy(s);
    }
  } free_ks{ks};
  hipEvent_t ready[2] = {nullptr, nullptr};
  for (int b = 0; b < 2; b++)
    if (hipEventCreateWithFlags(&ready[b], hipEventDisableTiming) != hipSuccess) {
      if (ready[0]) (void)hipEventDestroy(ready[0]);
      return Stager::fail("chunk event");
    }
  struct FreeEv {
    hipEvent_t *e;
    ~FreeEv() {
      for (int b = 0; b < 2; b++) (void)hipEventDestroy(e[b]);
    }
  } free_ev{ready};

  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  Stager stg;
  rc = stg.init(std::max(1u, std::min(8u, hw / 2)));
  if (rc) return rc;

  // CPU share: the forest of edges [last, last + cpu_span)
  std::vector<uint32_t> cpu_ids;
  uint64_t cpu_cnt = 0;
  int cpu_rc = PHEM_OK;
  double cpu_s = 0;
  // error messages are per thread: a worker's is carried back to the caller
  std::string cpu_err, next_err;
  std::thread cpu;
  if (cpu_span) {
    cpu_ids.resize(n);
    cpu = std::thread([&] {
      const double c0 = now_s();
      cpu_rc = phem_cpu_msf(n, cpu_span, u + last, v + last, w + last, wmode,
                            cpu_ids.data(), &cpu_cnt, (int)std::max(1u, hw - stg.threads));
      if (cpu_rc) cpu_err = phem_last_error();
      cpu_s = now_s() - c0;
    });
  }
  // the filtered last chunk: the forest of the CPU share (gathered on the
  // host), then the rest of the chunk straight from the caller's arrays
  std::vector<uint32_t> fl;
  uint64_t last_len = 0;

  // stages chunk c into device buffer c % 2 and records ready[c % 2]
  auto stage = [&](uint64_t c) -> int {
    uint32_t *d = d_in[c & 1];
    const uint64_t off = c * chunk_edges;
    const uint32_t *arr[3] = {u, v, w};
    const uint64_t k = std::min(chunk_edges, m - off);
    if (c + 1 == nchunks && cpu.joinable()) {
      cpu.join();
      if (cpu_rc) return set_err(cpu_rc, "host share: %s", cpu_err.c_str());
      const uint64_t rest = m - last - cpu_span;
      last_len = cpu_cnt + rest;
      fl.resize(cpu_cnt ? cpu_cnt : 1);
      for (int f = 0; f < 3; f++) {
        for (uint64_t i = 0; i < cpu_cnt; i++) fl[i] = arr[f][last + cpu_ids[i]];
        int r = stg.copy(d + f * chunk_edges, fl.data(), cpu_cnt);
        if (!r) r = stg.copy(d + f * chunk_edges + cpu_cnt, arr[f] + last + cpu_span, rest);
        if (r) return r;
      }
    } else {
      for (int f = 0; f < 3; f++) {
        const int r = stg.copy(d + f * chunk_edges, arr[f] + off, k);
        if (r) return r;
      }
    }
    if (hipEventRecord(ready[c & 1], stg.cs) != hipSuccess)
      return Stager::fail("chunk event");
    return PHEM_OK;
  };
  auto chunk_len = [&](uint64_t c) {
    if (c + 1 == nchunks && cpu_span) return last_len;
    return std::min(chunk_edges, m - c * chunk_edges);
  };

  if (nchunks) rc = stage(0);
  for (uint64_t c = 0; c < nchunks && rc == PHEM_OK; c++) {
    int next_rc = PHEM_OK;
    std::thread next;
    if (c + 1 < nchunks)
      next = std::thread([&, c] {
        next_rc = stage(c + 1);
        if (next_rc) next_err = phem_last_error();
      });
    if (hipStreamWaitEvent(ks, ready[c & 1], 0) != hipSuccess) {
      rc = Stager::fail("chunk wait");
    } else {
      uint32_t *d = d_in[c & 1];
      rc = phem_stream_apply(st, d, d + chunk_edges, d + 2 * chunk_edges,
                             chunk_len(c), ks);
    }
    if (next.joinable()) next.join();
    if (rc == PHEM_OK && next_rc)
      rc = set_err(next_rc, "staging chunk %llu: %s", (unsigned long long)(c + 1),
                   next_err.c_str());
  }
  if (cpu.joinable()) cpu.join();
  if (rc) return rc;

  // forest ids: global ids ascend with input positions
  uint64_t fc = 0;
  const uint64_t *d_gid = nullptr;
  rc = phem_stream_forest(st, &fc, nullptr, nullptr, nullptr, &d_gid);
  if (rc) return rc;
  std::vector<uint64_t> gid(fc);
  if (fc && hipMemcpy(gid.data(), d_gid, 8 * fc, hipMemcpyDeviceToHost) != hipSuccess)
    return Stager::fail("forest copy");
  for (uint64_t i = 0; i < fc; i++) {
    const uint64_t g = gid[i];
    uint64_t p = g;
    if (cpu_span && g >= last) {
      const uint64_t j = g - last;
      p = j < cpu_cnt ? last + cpu_ids[j] : last + cpu_span + (j - cpu_cnt);
    }
    ids_out[i] = (uint32_t)p;
  }
  *count = fc;
  if (stats) {
    stats->seconds = now_s() - t0;
    stats->cpu_seconds = cpu_s;
    stats->chunks = nchunks;
    stats->cpu_edges = cpu_span;
    stats->cpu_kept = cpu_cnt;
  }
  return PHEM_OK;
}
