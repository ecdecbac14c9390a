// phem_server.hip -- request loop with continuous batching (host code).
//
// Requests wait in one arrival-ordered queue.  Each slot owns a worker thread,
// a HIP stream, one block of device memory, pinned host staging buffers (the
// input one on the slot's first request from host memory) and a small pool
// of copy threads, so serving allocates nothing per request.  An idle worker
// takes the queued requests in order while the batch's device footprint (the
// batched solve's workspace, staged host inputs and outputs) fits its memory
// and the graph and edge caps hold, runs them as one batched solve and copies
// every result to its request's host buffers.  Several slots run at once, on
// their own streams, so the H2D staging and host-side scatter of one batch
// overlap the kernels of another; requests that arrive meanwhile form the
// next batch.
//
// With record_rounds set, a batch replays a batched solve recorded for its
// capacity bucket (graph count rounded up to a power of two, vertex and edge
// counts to a quarter octave) instead of launching the solve eagerly; each
// slot keeps its most recently used recordings.  A bucket's first batch runs
// eagerly and notes the rounds it needed; the recording is made when the
// bucket comes again, with those rounds plus two (at most record_rounds), so
// shapes seen once cost no recording and replays run no idle rounds beyond
// the margin.
//
// Resident streams (phem_server_stream_open) are phem_streams the server
// owns, admitted by their device footprint against a memory budget of their
// own.  A stream batch queues with the graph requests and runs alone on the
// worker that takes it.  A stream's batches carry sequence numbers, and a
// worker takes a stream batch only when its stream is idle and the batch
// before it has been applied (it passes over the others, which stay queued),
// so the batches of one stream are applied in submission order while batches
// of other streams and graph batches run on the other workers.  Host inputs
// are copied into
// pinned staging by the slot's copy threads in chunks, each chunk's
// host-to-device copy running while the next chunk is filled.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <functional>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "phem_batch_core.h"
#include "phem_copypool.h"
#include "phem_internal.h"

namespace phem {
namespace {

using Clock = std::chrono::steady_clock;

double now_s() {
  return std::chrono::duration<double>(Clock::now().time_since_epoch()).count();
}

struct Resident;

struct Req {
  phem_request q;
  phem_result res;
  uint64_t ticket;
  bool done = false;
  bool claimed = false;  // a phem_server_wait is waiting for it
  Resident *rs = nullptr;  // a stream batch: its stream and sequence number
  uint64_t seq = 0;
};

// A resident stream: the phem_stream, its footprint, and the sequence
// numbers of its batches (all under the server mutex).
struct Resident {
  phem_stream_t *s = nullptr;
  uint32_t n = 0;
  uint64_t max_batch = 0, bytes = 0;
  uint64_t submitted = 0;  // batches submitted: the next sequence number
  uint64_t applied = 0;    // batches finished (applied or failed)
  bool running = false;    // a worker is applying one of its batches
};

// A batch shape a recording is made for.
struct Bucket {
  uint32_t k_cap;
  uint64_t n_cap, m_cap;
  bool operator==(const Bucket &o) const {
    return k_cap == o.k_cap && n_cap == o.n_cap && m_cap == o.m_cap;
  }
};

// x rounded up to a quarter octave (at least lo): at most 25 % padding.
uint64_t quarter_octave(uint64_t x, uint64_t lo) {
  if (x < lo) x = lo;
  uint64_t e = 63 - __builtin_clzll(x);
  const uint64_t step = e >= 2 ? (1ull << (e - 2)) : 1;
  return (x + step - 1) / step * step;
}

Bucket bucket_of(uint32_t k, uint64_t n, uint64_t m) {
  uint32_t kc = 8;
  while (kc < k + 1) kc <<= 1;
  return Bucket{kc, quarter_octave(n, 4096), quarter_octave(m, 65536)};
}

struct Recorded {
  Bucket b;
  phem_batch_graph_t *g = nullptr;
  uint64_t used = 0;
};

constexpr size_t MAX_RECORDINGS = 16;  // per slot

// Rounds per phase a bucket's eager batch needed (the larger phase).
struct Seen {
  Bucket b;
  int rounds = 0;
  uint64_t used = 0;
};

constexpr size_t MAX_SEEN = 256;  // buckets remembered per slot

struct Slot {
  hipStream_t stream = nullptr;
  void *dev = nullptr;
  uint32_t *hin = nullptr;  // pinned: staged host edges (u | v | w)
  char *hout = nullptr;     // pinned: results on their way to the requests
  uint64_t hout_bytes = 0;
  std::thread th;
  std::unique_ptr<CopyPool> pool;
  std::vector<Recorded> rec;
  uint64_t tick = 0;
};

uint64_t a256(uint64_t x) { return align_up(x ? x : 1, 256); }

// Device bytes of the outputs and staging of a batch (capacities k, n; mh
// edges from host memory) after its workspace.
uint64_t io_bytes(uint64_t n, uint64_t mh, uint32_t k) {
  return 3 * a256(4 * mh) + 2 * a256(4 * n) + a256(4ull * k) + 2 * a256(8ull * k);
}

}  // namespace
}  // namespace phem

using namespace phem;

struct phem_server {
  int dev = 0;
  phem_server_opts o{};
  uint64_t slot_bytes = 0;
  std::vector<Slot> slots;
  std::mutex mu;
  // buckets seen by an eager batch, with the rounds per phase it needed
  // (any slot; under mu)
  std::vector<Seen> seen;
  uint64_t seen_tick = 0;
  std::condition_variable cv_work, cv_done;
  std::deque<Req *> queue;
  std::unordered_map<uint64_t, std::unique_ptr<Req>> reqs;
  uint64_t next_ticket = 1, in_flight = 0;
  bool stop = false;
  // resident streams (under mu)
  std::unordered_map<uint64_t, std::unique_ptr<Resident>> streams;
  uint64_t next_stream = 1, stream_mem = 0, stream_used = 0;
  // statistics since the last reset
  uint64_t submitted = 0, completed = 0, refused = 0, batches = 0, edges = 0,
           batch_graphs = 0, recorded_batches = 0, record_fallbacks = 0,
           recordings = 0;
  double first_arrival = 0, last_done = 0;
  std::vector<double> lat;
  uint64_t stream_batches = 0, stream_edges = 0;
  std::vector<double> stream_lat;

  // Device bytes of a batch of k graphs with n vertices, m edges of which
  // mh come from host memory, solved eagerly (exact sizes) or replaying the
  // recording of its capacity bucket.
  int footprint(uint64_t n, uint64_t m, uint64_t mh, uint32_t k,
                uint64_t *bytes) const {
    uint64_t ws = 0;
    int rc = batch_workspace_bytes(n, m, k, &ws);
    if (rc) return rc;
    *bytes = a256(ws) + io_bytes(n, mh, k);
    return PHEM_OK;
  }
  bool recorded_fits(uint64_t n, uint64_t m, uint64_t mh, uint32_t k) const {
    const Bucket b = bucket_of(k, n, m);
    if (b.n_cap > PHEM_MAX_VERTICES || b.m_cap > PHEM_MAX_EDGES) return false;
    uint64_t ws = 0;
    if (batch_workspace_bytes(b.n_cap, b.m_cap, b.k_cap, &ws) != PHEM_OK) return false;
    return a256(ws) + io_bytes(b.n_cap, mh, b.k_cap) <= slot_bytes;
  }
  bool fits(uint64_t n, uint64_t m, uint64_t mh, uint32_t k) const {
    if (k > o.max_graphs || m > o.max_batch_edges || n > o.max_batch_edges)
      return false;
    uint64_t b = 0;
    if (footprint(n, m, mh, k, &b) != PHEM_OK) return false;
    return b <= slot_bytes;
  }
  // Batches are formed so that they can replay their bucket's recording; a
  // request too large for that alone still runs, eagerly.
  bool admits(uint64_t n, uint64_t m, uint64_t mh, uint32_t k) const {
    if (!fits(n, m, mh, k)) return false;
    return !o.record_rounds || k == 1 || recorded_fits(n, m, mh, k);
  }
  // A batch of m edges for a stream of n vertices: the slot holds its labels
  // and, from host memory, its staged edges (within the pinned staging).
  bool stream_fits(uint32_t n, uint64_t mh) const {
    return mh <= o.max_batch_edges && a256(4ull * n) + 12 * mh <= slot_bytes;
  }
  void worker(int si);
  void run_batch(Slot &S, std::vector<Req *> &batch);
  void run_stream(Slot &S, Req *r);
  // Rounds per phase bucket b's eager batches needed (0: none seen yet), and
  // noting them (the least recently used bucket is forgotten beyond
  // MAX_SEEN).
  int seen_rounds(const Bucket &b) {
    std::lock_guard<std::mutex> lk(mu);
    for (auto &e : seen)
      if (e.b == b) {
        e.used = ++seen_tick;
        return e.rounds;
      }
    return 0;
  }
  void note_rounds(const Bucket &b, int rounds) {
    std::lock_guard<std::mutex> lk(mu);
    for (auto &e : seen)
      if (e.b == b) {
        e.rounds = rounds > e.rounds ? rounds : e.rounds;
        e.used = ++seen_tick;
        return;
      }
    if (seen.size() >= MAX_SEEN)
      seen.erase(std::min_element(seen.begin(), seen.end(), [](const Seen &x, const Seen &y) {
        return x.used < y.used;
      }));
    seen.push_back(Seen{b, rounds, ++seen_tick});
  }
  int stage_inputs(Slot &S, const std::vector<Req *> &batch,
                   const std::vector<uint64_t> &hoff, uint64_t mh,
                   uint32_t *d_su, uint32_t *d_sv, uint32_t *d_sw);
  phem_batch_graph_t *recording(Slot &S, const Bucket &b, int rounds, char *ws,
                                uint64_t ws_bytes, uint32_t *ids, uint32_t *cnt,
                                uint32_t *lab, uint64_t *wfx, uint64_t *bad);
};

void phem_server::worker(int si) {
  (void)hipSetDevice(dev);
  Slot &S = slots[si];
  std::vector<Req *> batch;
  for (;;) {
    batch.clear();
    {
      std::unique_lock<std::mutex> lk(mu);
      for (;;) {
        cv_work.wait(lk, [&] { return stop || !queue.empty(); });
        if (queue.empty()) return;  // stop requested and nothing left
        if (o.window_us && queue.size() < o.max_graphs && !stop) {
          const auto until = Clock::now() + std::chrono::microseconds(o.window_us);
          cv_work.wait_until(lk, until, [&] {
            return stop || queue.size() >= o.max_graphs;
          });
          if (queue.empty()) continue;  // another worker took them
        }
        // graph requests in arrival order while they fit; a stream batch
        // runs on its own, and only once its stream is idle and its
        // predecessor applied (others are passed over and stay queued)
        uint64_t n = 0, m = 0, mh = 0;
        for (auto it = queue.begin(); it != queue.end();) {
          Req *r = *it;
          if (r->rs) {
            if (batch.empty() && !r->rs->running && r->rs->applied == r->seq) {
              r->rs->running = true;
              batch.push_back(r);
              queue.erase(it);
              break;
            }
            ++it;
            continue;
          }
          const uint64_t rn = r->q.n, rm = r->q.m;
          const uint64_t rmh = (r->q.flags & PHEM_REQ_DEVICE) ? 0 : rm;
          if (!admits(n + rn, m + rm, mh + rmh, (uint32_t)batch.size() + 1)) break;
          n += rn;
          m += rm;
          mh += rmh;
          batch.push_back(r);
          it = queue.erase(it);
        }
        if (!batch.empty()) break;
        // only stream batches waiting for their streams: sleep until a
        // submission or a finished stream batch
        cv_work.wait(lk);
      }
      in_flight += batch.size();
    }
    if (batch.size() == 1 && batch[0]->rs)
      run_stream(S, batch[0]);
    else if (!batch.empty())
      run_batch(S, batch);
  }
}

// One batch of a resident stream: staged to the slot's memory if it comes
// from host memory, applied after the stream's previous batch, then the
// forest's size and weight (and labels) returned.
void phem_server::run_stream(Slot &S, Req *r) {
  Resident *rs = r->rs;  // taken only when idle with its predecessor applied
  const double t0 = now_s();
  const phem_request &q = r->q;
  const uint64_t m = q.m;
  const bool host = !(q.flags & PHEM_REQ_DEVICE);
  // the slot's memory: labels (n words), then the staged u, v, w (m words
  // each) of a batch from host memory (stream_fits checked both at submit)
  uint32_t *d_lab = (uint32_t *)S.dev;
  uint32_t *d_st = (uint32_t *)((char *)S.dev + a256(4ull * rs->n));
  int rc = PHEM_OK;
  const uint32_t *du = q.u, *dv = q.v, *dw = q.w;
  if (host && m) {
    std::vector<Req *> one{r};
    std::vector<uint64_t> hoff{0, m};
    rc = stage_inputs(S, one, hoff, m, d_st, d_st + m, d_st + 2 * m);
    du = d_st;
    dv = d_st + m;
    dw = d_st + 2 * m;
  }
  phem_mst_stats st{};
  if (rc == PHEM_OK) rc = phem_stream_apply(rs->s, du, dv, dw, m, S.stream);
  if (rc == PHEM_OK) rc = phem_stream_stats(rs->s, &st, nullptr, nullptr);
  if (rc == PHEM_OK && q.labels) {
    rc = phem_stream_labels(rs->s, d_lab, S.stream);
    if (rc == PHEM_OK) {
      // through the pinned result buffer when it holds them
      const bool pinned = rs->n <= o.max_batch_edges;
      uint32_t *h_lab = pinned ? (uint32_t *)(S.hout + S.hout_bytes - a256(4 * o.max_batch_edges))
                               : q.labels;
      hipError_t e = hipMemcpyAsync(h_lab, d_lab, 4ull * rs->n,
                                    hipMemcpyDeviceToHost, S.stream);
      if (e == hipSuccess) e = hipStreamSynchronize(S.stream);
      if (e != hipSuccess)
        rc = set_err(PHEM_EHIP, "label copy: %s", hipGetErrorString(e));
      else if (pinned)
        std::memcpy(q.labels, h_lab, 4ull * rs->n);
    }
  }
  if (rc != PHEM_OK) (void)hipStreamSynchronize(S.stream);
  const double t1 = now_s();
  char err[sizeof(((phem_result *)nullptr)->error)] = "";
  if (rc != PHEM_OK) {
    std::strncpy(err, phem_last_error(), sizeof(err) - 1);
    err[sizeof(err) - 1] = 0;
  }
  std::lock_guard<std::mutex> lk(mu);
  r->res.status = rc;
  std::memcpy(r->res.error, err, sizeof(err));
  r->res.forest_count = rc == PHEM_OK ? (uint32_t)st.mst_count : 0;
  r->res.weight_fx = rc == PHEM_OK ? st.weight_fx : 0;
  r->res.weight_exact = rc == PHEM_OK && st.weight_exact;
  r->res.batch_graphs = 1;
  r->res.start = t0;
  r->res.done = t1;
  r->done = true;
  if (rc == PHEM_OK) {
    if (completed == 0 || r->res.arrival < first_arrival)
      first_arrival = r->res.arrival;
    completed++;
    edges += m;
    lat.push_back(t1 - r->res.arrival);
    stream_batches++;
    stream_edges += m;
    stream_lat.push_back(t1 - r->res.arrival);
    last_done = std::max(last_done, t1);
  }
  rs->applied++;
  rs->running = false;
  in_flight -= 1;
  cv_done.notify_all();
  cv_work.notify_all();  // the stream's next batch may run now
}

// Host edges of the batch -> pinned staging -> device, in chunks of the
// staging arrays: the slot's copy threads fill chunk c + 1 while chunk c's
// host-to-device copy runs.  Staging index i of a field belongs to the host
// request g with hoff[g] <= i < hoff[g + 1].
int phem_server::stage_inputs(Slot &S, const std::vector<Req *> &batch,
                              const std::vector<uint64_t> &hoff, uint64_t mh,
                              uint32_t *d_su, uint32_t *d_sv, uint32_t *d_sw) {
  if (!mh) return PHEM_OK;
  if (!S.hin) {
    hipError_t e = hipHostMalloc((void **)&S.hin, 12 * o.max_batch_edges);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      S.hin = nullptr;
      return set_err(PHEM_ENOMEM, "pinned input staging: %s", hipGetErrorString(e));
    }
  }
  const uint32_t k = (uint32_t)batch.size();
  constexpr uint64_t CHUNK = 4ull << 20;  // elements (16 MB) per chunk
  constexpr uint64_t PIECE = 1ull << 18;  // elements per copy-thread task
  uint32_t *hst[3] = {S.hin, S.hin + mh, S.hin + 2 * mh};
  uint32_t *dst[3] = {d_su, d_sv, d_sw};
  for (int f = 0; f < 3; f++) {
    for (uint64_t c0 = 0; c0 < mh; c0 += CHUNK) {
      const uint64_t c1 = std::min(mh, c0 + CHUNK);
      const uint64_t pieces = (c1 - c0 + PIECE - 1) / PIECE;
      S.pool->run(pieces, [&](uint64_t p) {
        uint64_t i = c0 + p * PIECE;
        const uint64_t end = std::min(c1, i + PIECE);
        // the host request holding staging index i
        uint32_t g = (uint32_t)(std::upper_bound(hoff.begin(), hoff.begin() + k + 1, i) -
                                hoff.begin()) - 1;
        while (i < end) {
          while (hoff[g + 1] <= i) g++;
          const phem_request &q = batch[g]->q;
          const uint32_t *src = f == 0 ? q.u : f == 1 ? q.v : q.w;
          const uint64_t j = i - hoff[g];
          const uint64_t len = std::min(end, hoff[g + 1]) - i;
          std::memcpy(hst[f] + i, src + j, 4 * len);
          i += len;
        }
      });
      const hipError_t e = hipMemcpyAsync(dst[f] + c0, hst[f] + c0, 4 * (c1 - c0),
                                          hipMemcpyHostToDevice, S.stream);
      if (e != hipSuccess) {
        (void)hipGetLastError();
        return set_err(PHEM_EHIP, "staging copy: %s", hipGetErrorString(e));
      }
    }
  }
  return PHEM_OK;
}

// Rounds per phase of a solve's larger phase, from its statistics.
int phase_rounds(const phem_mst_stats &st) {
  const int light = (int)st.light_rounds, heavy = (int)st.rounds - light;
  return light > heavy ? light : heavy;
}

// The slot's recording for bucket b (made on first use with `rounds` per
// phase; the least recently used one is dropped beyond MAX_RECORDINGS), or
// NULL if recording failed.
phem_batch_graph_t *phem_server::recording(Slot &S, const Bucket &b, int rounds,
                                           char *ws,
                                           uint64_t ws_bytes, uint32_t *ids,
                                           uint32_t *cnt, uint32_t *lab,
                                           uint64_t *wfx, uint64_t *bad) {
  S.tick++;
  for (auto &r : S.rec)
    if (r.b == b) {
      r.used = S.tick;
      return r.g;
    }
  if (S.rec.size() >= MAX_RECORDINGS) {
    auto lru = std::min_element(S.rec.begin(), S.rec.end(),
                                [](const Recorded &x, const Recorded &y) {
                                  return x.used < y.used;
                                });
    (void)hipStreamSynchronize(S.stream);  // no replay of it is in flight
    phem_batch_graph_destroy(lru->g);
    S.rec.erase(lru);
  }
  phem_batch_graph_t *g = nullptr;
  if (phem_batch_graph_create(b.k_cap, b.n_cap, b.m_cap, o.wmode, ws, ws_bytes,
                              ids, cnt, lab, wfx, bad, nullptr, rounds, &g) != PHEM_OK)
    return nullptr;
  S.rec.push_back(Recorded{b, g, S.tick});
  {
    std::lock_guard<std::mutex> lk(mu);
    recordings++;
  }
  return g;
}

void phem_server::run_batch(Slot &S, std::vector<Req *> &batch) {
  const double t0 = now_s();
  const uint32_t k = (uint32_t)batch.size();
  std::vector<uint64_t> voff(k + 1, 0), eoff(k + 1, 0), hoff(k + 1, 0);
  for (uint32_t g = 0; g < k; g++) {
    const phem_request &q = batch[g]->q;
    voff[g + 1] = voff[g] + q.n;
    eoff[g + 1] = eoff[g] + q.m;
    hoff[g + 1] = hoff[g] + ((q.flags & PHEM_REQ_DEVICE) ? 0 : q.m);
  }
  const uint64_t n = voff[k], m = eoff[k], mh = hoff[k];
  // a recorded replay when the batch's bucket fits the slot, else eager
  const bool rec = o.record_rounds && recorded_fits(n, m, mh, k);
  const Bucket b = rec ? bucket_of(k, n, m) : Bucket{k, n, m};
  uint64_t ws = 0;
  int rc = batch_workspace_bytes(b.n_cap, b.m_cap, b.k_cap, &ws);
  // carve the slot's device memory
  char *base = (char *)S.dev;
  uint64_t off = 0;
  auto take = [&](uint64_t bytes) {
    char *p = base + off;
    off += a256(bytes);
    return p;
  };
  char *d_ws = take(ws);
  uint32_t *d_ids = (uint32_t *)take(4 * b.n_cap), *d_lab = (uint32_t *)take(4 * b.n_cap);
  uint32_t *d_cnt = (uint32_t *)take(4ull * b.k_cap);
  uint64_t *d_wfx = (uint64_t *)take(8ull * b.k_cap), *d_bad = (uint64_t *)take(8ull * b.k_cap);
  uint32_t *d_su = (uint32_t *)take(4 * mh), *d_sv = (uint32_t *)take(4 * mh),
           *d_sw = (uint32_t *)take(4 * mh);
  if (rc == PHEM_OK) rc = stage_inputs(S, batch, hoff, mh, d_su, d_sv, d_sw);
  std::vector<const uint32_t *> pu(k), pv(k), pw(k);
  for (uint32_t g = 0; g < k; g++) {
    const phem_request &q = batch[g]->q;
    if (q.flags & PHEM_REQ_DEVICE) {
      pu[g] = q.u;
      pv[g] = q.v;
      pw[g] = q.w;
    } else {
      pu[g] = d_su + hoff[g];
      pv[g] = d_sv + hoff[g];
      pw[g] = d_sw + hoff[g];
    }
  }
  bool replayed = false, fell_back = false;
  // a bucket seen before replays its recording; its first batch runs eagerly
  const int need = rec ? seen_rounds(b) : 0;
  if (rc == PHEM_OK && need > 0) {
    phem_mst_stats st{};
    const int rounds = need + 2 < (int)o.record_rounds ? need + 2 : (int)o.record_rounds;
    phem_batch_graph_t *g =
        recording(S, b, rounds, d_ws, ws, d_ids, d_cnt, d_lab, d_wfx, d_bad);
    if (g) {
      rc = phem_batch_graph_launch(g, k, voff.data(), eoff.data(), pu.data(),
                                   pv.data(), pw.data(), S.stream);
      if (rc == PHEM_OK) rc = phem_batch_graph_stats(g, &st, S.stream);
      replayed = true;
      if (rc == PHEM_ENOCONV) {
        // more rounds than recorded: solve eagerly, and record the bucket
        // again next time with the rounds this batch needs
        fell_back = true;
        rc = PHEM_OK;
        (void)hipStreamSynchronize(S.stream);
        for (auto it = S.rec.begin(); it != S.rec.end(); ++it)
          if (it->b == b) {
            phem_batch_graph_destroy(it->g);
            S.rec.erase(it);
            break;
          }
      }
    }
  }
  if (rc == PHEM_OK && (!replayed || fell_back)) {
    phem_mst_stats st{};
    BatchArgs A{};
    A.k = k;
    A.h_voff = voff.data();
    A.h_eoff = eoff.data();
    A.h_pu = pu.data();
    A.h_pv = pv.data();
    A.h_pw = pw.data();
    A.wmode = o.wmode;
    A.d_ws = d_ws;
    A.ws_bytes = ws;
    A.d_ids = d_ids;
    A.d_counts = d_cnt;
    A.d_labels = d_lab;
    A.d_weights = d_wfx;
    A.d_inexact = d_bad;
    A.stats = &st;
    rc = batch_solve(A, S.stream);
    if (rc == PHEM_OK && rec) note_rounds(b, phase_rounds(st));
  }
  // results: one copy per field into the pinned buffer, then to each request
  char *h = S.hout;
  uint32_t *h_cnt = (uint32_t *)h;
  uint64_t *h_wfx = (uint64_t *)(h + a256(4ull * o.max_graphs));
  uint64_t *h_bad = (uint64_t *)((char *)h_wfx + a256(8ull * o.max_graphs));
  uint32_t *h_ids = (uint32_t *)((char *)h_bad + a256(8ull * o.max_graphs));
  uint32_t *h_lab = (uint32_t *)((char *)h_ids + a256(4 * o.max_batch_edges));
  if (rc == PHEM_OK) {
    hipError_t e = hipMemcpyAsync(h_cnt, d_cnt, 4ull * k, hipMemcpyDeviceToHost, S.stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(h_wfx, d_wfx, 8ull * k, hipMemcpyDeviceToHost, S.stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(h_bad, d_bad, 8ull * k, hipMemcpyDeviceToHost, S.stream);
    if (e == hipSuccess && n)
      e = hipMemcpyAsync(h_ids, d_ids, 4 * n, hipMemcpyDeviceToHost, S.stream);
    if (e == hipSuccess && n)
      e = hipMemcpyAsync(h_lab, d_lab, 4 * n, hipMemcpyDeviceToHost, S.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(S.stream);
    if (e != hipSuccess) rc = set_err(PHEM_EHIP, "result copy: %s", hipGetErrorString(e));
  }
  if (rc == PHEM_OK) {
    auto scatter = [&](uint32_t g0, uint32_t g1) {
      for (uint32_t g = g0; g < g1; g++) {
        const phem_request &q = batch[g]->q;
        if (q.forest_ids && h_cnt[g])
          std::memcpy(q.forest_ids, h_ids + voff[g], 4ull * h_cnt[g]);
        if (q.labels && q.n) std::memcpy(q.labels, h_lab + voff[g], 4ull * q.n);
      }
    };
    // the slot's copy threads for large batches, contiguous runs of graphs
    // with about the same vertex count each
    if (n < (2u << 20) || k == 1) {
      scatter(0, k);
    } else {
      constexpr uint32_t PARTS = 8;
      std::vector<uint32_t> cut(PARTS + 1, k);
      cut[0] = 0;
      uint32_t g = 0;
      for (uint32_t t = 1; t < PARTS; t++) {
        const uint64_t target = n * t / PARTS;
        while (g < k && voff[g] < target) g++;
        cut[t] = g;
      }
      S.pool->run(PARTS, [&](uint64_t t) {
        if (cut[t + 1] > cut[t]) scatter(cut[t], cut[t + 1]);
      });
    }
  } else {
    (void)hipStreamSynchronize(S.stream);
  }
  const double t1 = now_s();
  // the failure's message lives in this worker thread's error slot: each
  // request of the batch carries a copy to its client
  char err[sizeof(((phem_result *)nullptr)->error)] = "";
  if (rc != PHEM_OK) {
    std::strncpy(err, phem_last_error(), sizeof(err) - 1);
    err[sizeof(err) - 1] = 0;
  }
  std::lock_guard<std::mutex> lk(mu);
  for (uint32_t g = 0; g < k; g++) {
    Req *r = batch[g];
    r->res.status = rc;
    std::memcpy(r->res.error, err, sizeof(err));
    r->res.forest_count = rc == PHEM_OK ? h_cnt[g] : 0;
    r->res.weight_fx = rc == PHEM_OK ? h_wfx[g] : 0;
    r->res.weight_exact = rc == PHEM_OK && h_bad[g] == 0;
    r->res.batch_graphs = k;
    r->res.start = t0;
    r->res.done = t1;
    r->done = true;
    if (rc == PHEM_OK) {
      if (completed == 0 || r->res.arrival < first_arrival)
        first_arrival = r->res.arrival;
      completed++;
      edges += r->q.m;
      lat.push_back(t1 - r->res.arrival);
    }
  }
  if (rc == PHEM_OK) {
    last_done = std::max(last_done, t1);
    batches++;
    batch_graphs += k;
    if (replayed) recorded_batches++;
    if (fell_back) record_fallbacks++;
  }
  in_flight -= k;
  cv_done.notify_all();
}

extern "C" double phem_server_clock(void) { return now_s(); }

extern "C" int phem_server_create(const phem_server_opts *opts,
                                  phem_server **out) {
  if (!out) return set_err(PHEM_EINVAL, "out is NULL");
  *out = nullptr;
  phem_server_opts o{};
  if (opts) o = *opts;
  if (o.wmode != PHEM_W_U32 && o.wmode != PHEM_W_F32)
    return set_err(PHEM_EINVAL, "unknown weight mode %d", o.wmode);
  if (!o.slots) o.slots = 2;
  if (o.slots > 16) return set_err(PHEM_EINVAL, "at most 16 slots");
  if (!o.max_graphs) o.max_graphs = 4096;
  if (!o.max_batch_edges) o.max_batch_edges = 1ull << 27;
  if (o.max_batch_edges > PHEM_MAX_EDGES)
    return set_err(PHEM_EINVAL, "max_batch_edges beyond %llu",
                   (unsigned long long)PHEM_MAX_EDGES);
  int dev = 0;
  PHEM_HIP(hipGetDevice(&dev));
  if (!o.mem_bytes) {
    size_t fr = 0, tot = 0;
    PHEM_HIP(hipMemGetInfo(&fr, &tot));
    o.mem_bytes = fr / 4;
  }
  std::unique_ptr<phem_server> S(new phem_server());
  S->dev = dev;
  S->o = o;
  S->slot_bytes = o.mem_bytes / o.slots / 256 * 256;
  S->stream_mem = o.stream_mem_mib ? (uint64_t)o.stream_mem_mib << 20 : o.mem_bytes;
  S->slots.resize(o.slots);
  // copy threads per slot: the box's share of host threads over the slots
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  const unsigned per_slot = std::max(1u, std::min(8u, std::min(hw, 16u) / o.slots)) - 1;
  for (auto &sl : S->slots) {
    sl.pool.reset(new CopyPool(per_slot));
    hipError_t e = hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc(&sl.dev, S->slot_bytes);
    sl.hout_bytes = a256(4ull * o.max_graphs) + 2 * a256(8ull * o.max_graphs) +
                    2 * a256(4 * o.max_batch_edges);
    if (e == hipSuccess) e = hipHostMalloc((void **)&sl.hout, sl.hout_bytes);
    if (e != hipSuccess) {
      (void)hipGetLastError();  // not left for the next launch check
      int rc = set_err(PHEM_ENOMEM, "slot allocation (%llu device bytes): %s",
                       (unsigned long long)S->slot_bytes, hipGetErrorString(e));
      phem_server_destroy(S.release());
      return rc;
    }
  }
  for (int i = 0; i < (int)o.slots; i++)
    S->slots[i].th = std::thread(&phem_server::worker, S.get(), i);
  *out = S.release();
  return PHEM_OK;
}

extern "C" int phem_server_destroy(phem_server *srv) {
  if (!srv) return PHEM_OK;
  {
    std::lock_guard<std::mutex> lk(srv->mu);
    srv->stop = true;
  }
  srv->cv_work.notify_all();
  for (auto &sl : srv->slots)
    if (sl.th.joinable()) sl.th.join();
  for (auto &sl : srv->slots) {
    if (sl.stream) (void)hipStreamSynchronize(sl.stream);
    for (auto &r : sl.rec) phem_batch_graph_destroy(r.g);
    sl.rec.clear();
    sl.pool.reset();
    if (sl.dev) (void)hipFree(sl.dev);
    if (sl.hin) (void)hipHostFree(sl.hin);
    if (sl.hout) (void)hipHostFree(sl.hout);
    if (sl.stream) (void)hipStreamDestroy(sl.stream);
  }
  for (auto &e : srv->streams) phem_stream_destroy(e.second->s);
  srv->streams.clear();
  delete srv;
  return PHEM_OK;
}

extern "C" int phem_server_submit(phem_server *srv, const phem_request *req,
                                  uint64_t *ticket) {
  if (!srv || !req || !ticket) return set_err(PHEM_EINVAL, "NULL argument");
  if (req->m && (!req->u || !req->v || !req->w))
    return set_err(PHEM_EINVAL, "request has edges but NULL arrays");
  const double t = req->arrival >= 0 ? req->arrival : now_s();
  const uint64_t mh = (req->flags & PHEM_REQ_DEVICE) ? 0 : req->m;
  std::unique_ptr<Req> r(new Req());
  r->q = *req;
  r->res = phem_result{};
  r->res.arrival = t;
  std::lock_guard<std::mutex> lk(srv->mu);
  if (srv->stop) return set_err(PHEM_EINVAL, "server is shutting down");
  if (req->flags & PHEM_REQ_STREAM) {
    auto it = srv->streams.find(req->stream_id);
    if (it == srv->streams.end())
      return set_err(PHEM_ENOTFOUND, "unknown stream %llu",
                     (unsigned long long)req->stream_id);
    Resident *rs = it->second.get();
    if (req->forest_ids)
      return set_err(PHEM_EINVAL, "a stream batch returns no forest ids");
    if (req->m > rs->max_batch)
      return set_err(PHEM_EINVAL, "batch of %llu edges exceeds the stream's max_batch %llu",
                     (unsigned long long)req->m, (unsigned long long)rs->max_batch);
    if (!srv->stream_fits(rs->n, mh)) {
      srv->refused++;
      return set_err(PHEM_EBUSY, "stream batch of %llu host edges exceeds one slot",
                     (unsigned long long)mh);
    }
    r->rs = rs;
    r->seq = rs->submitted++;
    r->q.n = rs->n;
  } else if (!srv->fits(req->n, req->m, mh, 1)) {
    srv->refused++;
    return set_err(PHEM_EBUSY,
                   "request of %u vertices, %llu edges exceeds one slot "
                   "(%llu device bytes, %llu edges)", req->n,
                   (unsigned long long)req->m,
                   (unsigned long long)srv->slot_bytes,
                   (unsigned long long)srv->o.max_batch_edges);
  }
  r->ticket = srv->next_ticket++;
  *ticket = r->ticket;
  srv->submitted++;
  srv->queue.push_back(r.get());
  srv->reqs.emplace(r->ticket, std::move(r));
  srv->cv_work.notify_one();
  return PHEM_OK;
}

extern "C" int phem_server_wait(phem_server *srv, uint64_t ticket,
                                phem_result *res) {
  if (!srv) return set_err(PHEM_EINVAL, "NULL server");
  std::unique_lock<std::mutex> lk(srv->mu);
  auto it = srv->reqs.find(ticket);
  if (it == srv->reqs.end())
    return set_err(PHEM_ENOTFOUND, "unknown ticket %llu", (unsigned long long)ticket);
  Req *r = it->second.get();
  if (r->claimed)
    return set_err(PHEM_EBUSY, "ticket %llu is already being waited for",
                   (unsigned long long)ticket);
  // claimed: a poll or second wait meanwhile gets PHEM_EBUSY and leaves it
  r->claimed = true;
  srv->cv_done.wait(lk, [&] { return r->done; });
  if (res) *res = r->res;
  srv->reqs.erase(ticket);  // by key: other submissions may have rehashed
  return PHEM_OK;
}

extern "C" int phem_server_poll(phem_server *srv, uint64_t ticket,
                                phem_result *res, int *done) {
  if (!srv || !done) return set_err(PHEM_EINVAL, "NULL argument");
  std::lock_guard<std::mutex> lk(srv->mu);
  auto it = srv->reqs.find(ticket);
  if (it == srv->reqs.end())
    return set_err(PHEM_ENOTFOUND, "unknown ticket %llu", (unsigned long long)ticket);
  if (it->second->claimed)
    return set_err(PHEM_EBUSY, "ticket %llu is being waited for",
                   (unsigned long long)ticket);
  *done = it->second->done ? 1 : 0;
  if (*done) {
    if (res) *res = it->second->res;
    srv->reqs.erase(it);
  }
  return PHEM_OK;
}

extern "C" int phem_server_drain(phem_server *srv) {
  if (!srv) return set_err(PHEM_EINVAL, "NULL server");
  std::unique_lock<std::mutex> lk(srv->mu);
  srv->cv_done.wait(lk, [&] { return srv->queue.empty() && srv->in_flight == 0; });
  return PHEM_OK;
}

extern "C" int phem_server_get_stats(phem_server *srv, phem_server_stats *st) {
  if (!srv || !st) return set_err(PHEM_EINVAL, "NULL argument");
  std::lock_guard<std::mutex> lk(srv->mu);
  phem_server_stats s{};
  s.submitted = srv->submitted;
  s.completed = srv->completed;
  s.refused = srv->refused;
  s.batches = srv->batches;
  s.edges = srv->edges;
  s.first_arrival = srv->first_arrival;
  s.last_done = srv->last_done;
  if (!srv->lat.empty()) {
    std::vector<double> l = srv->lat;
    std::sort(l.begin(), l.end());
    // nearest-rank percentiles
    auto pct = [&](double p) {
      size_t i = (size_t)std::ceil(p * l.size());
      return l[i ? i - 1 : 0];
    };
    s.lat_p50 = pct(0.50);
    s.lat_p99 = pct(0.99);
    s.lat_max = l.back();
    double sum = 0;
    for (double x : l) sum += x;
    s.lat_mean = sum / l.size();
  }
  s.mean_batch_graphs = srv->batches ? (double)srv->batch_graphs / srv->batches : 0;
  s.slot_bytes = srv->slot_bytes;
  s.slot_edges = srv->o.max_batch_edges;
  s.recorded_batches = srv->recorded_batches;
  s.record_fallbacks = srv->record_fallbacks;
  s.recordings = srv->recordings;
  s.stream_batches = srv->stream_batches;
  s.stream_edges = srv->stream_edges;
  if (!srv->stream_lat.empty()) {
    std::vector<double> l = srv->stream_lat;
    std::sort(l.begin(), l.end());
    auto pct = [&](double p) {
      size_t i = (size_t)std::ceil(p * l.size());
      return l[i ? i - 1 : 0];
    };
    s.stream_lat_p50 = pct(0.50);
    s.stream_lat_p99 = pct(0.99);
  }
  s.streams_open = srv->streams.size();
  s.stream_bytes = srv->stream_used;
  *st = s;
  return PHEM_OK;
}

extern "C" int phem_server_reset_stats(phem_server *srv) {
  if (!srv) return set_err(PHEM_EINVAL, "NULL server");
  std::lock_guard<std::mutex> lk(srv->mu);
  srv->submitted = srv->completed = srv->refused = srv->batches = srv->edges = 0;
  srv->batch_graphs = srv->recorded_batches = srv->record_fallbacks = 0;
  srv->recordings = 0;
  srv->first_arrival = srv->last_done = 0;
  srv->lat.clear();
  srv->stream_batches = srv->stream_edges = 0;
  srv->stream_lat.clear();
  return PHEM_OK;
}

extern "C" int phem_server_stream_open(phem_server *srv, uint32_t n,
                                       uint64_t max_batch,
                                       uint64_t *stream_id) {
  if (!srv || !stream_id) return set_err(PHEM_EINVAL, "NULL argument");
  *stream_id = 0;
  uint64_t bytes = 0;
  int rc = phem_stream_bytes(n, max_batch, &bytes);
  if (rc) return rc;
  {
    // admission: reserve the footprint before allocating
    std::lock_guard<std::mutex> lk(srv->mu);
    if (srv->stop) return set_err(PHEM_EINVAL, "server is shutting down");
    if (srv->stream_used + bytes > srv->stream_mem) {
      srv->refused++;
      return set_err(PHEM_EBUSY,
                     "stream of %u vertices needs %llu device bytes, %llu of %llu free",
                     n, (unsigned long long)bytes,
                     (unsigned long long)(srv->stream_mem - srv->stream_used),
                     (unsigned long long)srv->stream_mem);
    }
    srv->stream_used += bytes;
  }
  // the stream's memory on the server's device (the caller's device is
  // restored afterwards)
  std::unique_ptr<Resident> rs(new (std::nothrow) Resident());
  int dev = 0;
  if (!rs) {
    rc = set_err(PHEM_ENOMEM, "host allocation failed");
  } else if (hipGetDevice(&dev) != hipSuccess ||
             (dev != srv->dev && hipSetDevice(srv->dev) != hipSuccess)) {
    (void)hipGetLastError();
    rc = set_err(PHEM_EHIP, "cannot select the server's device %d", srv->dev);
  } else {
    rc = phem_stream_create(n, max_batch, srv->o.wmode, &rs->s);
    if (dev != srv->dev) (void)hipSetDevice(dev);
  }
  std::lock_guard<std::mutex> lk(srv->mu);
  if (rc) {
    srv->stream_used -= bytes;
    return rc;
  }
  rs->n = n;
  rs->max_batch = max_batch;
  rs->bytes = bytes;
  *stream_id = srv->next_stream++;
  srv->streams.emplace(*stream_id, std::move(rs));
  return PHEM_OK;
}

extern "C" int phem_server_stream_close(phem_server *srv, uint64_t stream_id) {
  if (!srv) return set_err(PHEM_EINVAL, "NULL server");
  std::unique_ptr<Resident> rs;
  {
    std::lock_guard<std::mutex> lk(srv->mu);
    auto it = srv->streams.find(stream_id);
    if (it == srv->streams.end())
      return set_err(PHEM_ENOTFOUND, "unknown stream %llu",
                     (unsigned long long)stream_id);
    if (it->second->applied != it->second->submitted)
      return set_err(PHEM_EBUSY, "stream %llu has batches queued or running",
                     (unsigned long long)stream_id);
    rs = std::move(it->second);
    srv->streams.erase(it);
    srv->stream_used -= rs->bytes;
  }
  return phem_stream_destroy(rs->s);
}
