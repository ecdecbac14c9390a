// phem_stream.hip -- streaming minimum spanning forest and connected
// components with device-resident state.
//
// The forest after batch t is F_t = MSF(F_{t-1} u B_t), which equals the
// minimum spanning forest of every edge inserted so far (cycle property): an
// edge left out of F_{t-1} is the heaviest edge of a cycle of edges inserted
// by then, and stays excluded.  Edges carry a global id (insertion order);
// the solve orders ties by position in [F_{t-1} | B_t], and F_{t-1} is kept
// sorted by global id ahead of the batch, so position order is global-id
// order.  This is the chunked MSF of the reference's streamline path
// (d_edges + chunk_size <- previous MSF, MST/src/kernels.cu:942-1102,
// 949-953) with the forest kept in HBM between batches.  Connected components
// follow the batch-by-batch union-find of the reference's SpanningTree
// (CC/ExternalSpanningTree.cu:170-204) on a parent array kept in HBM.
#include <cstring>
#include <new>

#include "phem_dev.h"
#include "phem_internal.h"

struct phem_stream {
  uint32_t n = 0;
  uint64_t max_batch = 0, cap = 0;
  int wmode = 0;
  int device = 0;
  // two [forest | batch] buffers; `cur` holds the forest in [0, f)
  uint32_t *u[2] = {nullptr, nullptr}, *v[2] = {nullptr, nullptr},
           *w[2] = {nullptr, nullptr};
  unsigned long long *gid[2] = {nullptr, nullptr};
  int cur = 0;
  uint64_t f = 0;
  uint64_t next_gid = 0;
  uint64_t batches = 0;
  void *ws = nullptr;
  uint64_t ws_bytes = 0;
  uint32_t *ids = nullptr;      // forest positions of the last solve
  uint32_t *labels = nullptr;   // min-id labels of the last solve
  uint32_t *parent = nullptr;   // union-find parent (connected components)
  unsigned long long *cc_forest = nullptr;
  unsigned long long *packed = nullptr;  // batch records (u << 32) | v
  phem_mst_stats last{};
  // an apply failed after the batch's union-find was enqueued: the
  // union-find state holds the batch, the forest does not
  bool broken = false;
};

namespace phem {

constexpr int SBLK_S = 256;

__global__ __launch_bounds__(SBLK_S) void k_stream_stage(
    const uint32_t *__restrict__ bu, const uint32_t *__restrict__ bv,
    const uint32_t *__restrict__ bw, uint64_t mb, uint64_t gid0,
    uint32_t *__restrict__ u, uint32_t *__restrict__ v, uint32_t *__restrict__ w,
    ull *__restrict__ gid, ull *__restrict__ packed) {
  for (uint64_t i = (uint64_t)blockIdx.x * SBLK_S + threadIdx.x; i < mb;
       i += (uint64_t)gridDim.x * SBLK_S) {
    const uint32_t a = bu[i], b = bv[i];
    u[i] = a;
    v[i] = b;
    w[i] = bw[i];
    gid[i] = gid0 + i;
    const uint32_t lo = a < b ? a : b, hi = a < b ? b : a;
    packed[i] = ((ull)lo << 32) | hi;
  }
}

// count: the device word holding the forest size (at most `bound`).
__global__ __launch_bounds__(SBLK_S) void k_stream_keep(
    const uint32_t *__restrict__ ids, const ull *__restrict__ count_p,
    const uint32_t *__restrict__ su, const uint32_t *__restrict__ sv,
    const uint32_t *__restrict__ sw, const ull *__restrict__ sg,
    uint32_t *__restrict__ du, uint32_t *__restrict__ dv,
    uint32_t *__restrict__ dw, ull *__restrict__ dg) {
  const uint64_t count = *count_p;
  for (uint64_t k = (uint64_t)blockIdx.x * SBLK_S + threadIdx.x; k < count;
       k += (uint64_t)gridDim.x * SBLK_S) {
    const uint32_t p = ids[k];
    du[k] = su[p];
    dv[k] = sv[p];
    dw[k] = sw[p];
    dg[k] = sg[p];
  }
}

static void stream_free(phem_stream *s) {
  for (int b = 0; b < 2; b++) {
    (void)hipFree(s->u[b]);
    (void)hipFree(s->v[b]);
    (void)hipFree(s->w[b]);
    (void)hipFree(s->gid[b]);
  }
  (void)hipFree(s->ws);
  (void)hipFree(s->ids);
  (void)hipFree(s->labels);
  (void)hipFree(s->parent);
  (void)hipFree(s->cc_forest);
  (void)hipFree(s->packed);
  delete s;
}

}  // namespace phem

using namespace phem;

extern "C" int phem_stream_bytes(uint32_t n, uint64_t max_batch,
                                 uint64_t *bytes) {
  if (!bytes) return set_err(PHEM_EINVAL, "bytes is NULL");
  *bytes = 0;
  if (n < 1 || n > PHEM_MAX_VERTICES || max_batch < 1)
    return set_err(PHEM_EINVAL, "bad n=%u / max_batch=%llu", n,
                   (unsigned long long)max_batch);
  const uint64_t cap = (uint64_t)n + max_batch;
  if (cap > PHEM_MAX_EDGES)
    return set_err(PHEM_EINVAL, "n + max_batch exceeds the edge-id range");
  uint64_t ws = 0;
  int rc = phem_mst_workspace_bytes(n, cap, &ws);
  if (rc) return rc;
  // the allocations of phem_stream_create: two [forest | batch] buffers
  // (u, v, w, global id), the solve's workspace, forest positions, labels,
  // parent, union-find forest and the batch's packed records
  *bytes = 2 * 20 * cap + ws + 12ull * n + 8ull * n + 8 * max_batch;
  return PHEM_OK;
}

extern "C" int phem_stream_create(uint32_t n, uint64_t max_batch, int wmode,
                                  phem_stream_t **out) {
  if (!out) return set_err(PHEM_EINVAL, "out is NULL");
  *out = nullptr;
  if (n < 1 || n > PHEM_MAX_VERTICES || max_batch < 1)
    return set_err(PHEM_EINVAL, "bad n=%u / max_batch=%llu", n,
                   (unsigned long long)max_batch);
  if (wmode != PHEM_W_U32 && wmode != PHEM_W_F32)
    return set_err(PHEM_EINVAL, "unknown weight mode %d", wmode);
  const uint64_t cap = (uint64_t)n + max_batch;
  if (cap > PHEM_MAX_EDGES)
    return set_err(PHEM_EINVAL, "n + max_batch exceeds the edge-id range");
  phem_stream *s = new (std::nothrow) phem_stream();
  if (!s) return set_err(PHEM_ENOMEM, "host allocation failed");
  s->n = n;
  s->max_batch = max_batch;
  s->cap = cap;
  s->wmode = wmode;
  (void)hipGetDevice(&s->device);
  int rc = phem_mst_workspace_bytes(n, cap, &s->ws_bytes);
  if (rc) {
    stream_free(s);
    return rc;
  }
  bool ok = true;
  for (int b = 0; b < 2 && ok; b++) {
    ok = hipMalloc(&s->u[b], 4 * cap) == hipSuccess &&
         hipMalloc(&s->v[b], 4 * cap) == hipSuccess &&
         hipMalloc(&s->w[b], 4 * cap) == hipSuccess &&
         hipMalloc(&s->gid[b], 8 * cap) == hipSuccess;
  }
  ok = ok && hipMalloc(&s->ws, s->ws_bytes) == hipSuccess &&
       hipMalloc(&s->ids, 4ull * n) == hipSuccess &&
       hipMalloc(&s->labels, 4ull * n) == hipSuccess &&
       hipMalloc(&s->parent, 4ull * n) == hipSuccess &&
       hipMalloc(&s->cc_forest, 8ull * n) == hipSuccess &&
       hipMalloc(&s->packed, 8 * max_batch) == hipSuccess;
  if (!ok) {
    (void)hipGetLastError();  // not left for the next launch check
    stream_free(s);
    return set_err(PHEM_ENOMEM, "device allocation for the stream failed");
  }
  rc = phem_cc_init(n, s->parent, (uint64_t *)s->cc_forest, nullptr);
  if (rc == PHEM_OK && hipDeviceSynchronize() != hipSuccess)
    rc = set_err(PHEM_EHIP, "stream initialisation failed");
  if (rc) {
    stream_free(s);
    return rc;
  }
  *out = s;
  return PHEM_OK;
}

extern "C" int phem_stream_destroy(phem_stream_t *s) {
  if (!s) return PHEM_OK;
  stream_free(s);
  return PHEM_OK;
}

extern "C" int phem_stream_apply(phem_stream_t *s, const uint32_t *d_u,
                                 const uint32_t *d_v, const uint32_t *d_w,
                                 uint64_t mb, void *stream) {
  if (!s) return set_err(PHEM_EINVAL, "stream handle is NULL");
  if (s->broken)
    return set_err(PHEM_EINVAL, "stream is unusable after a failed apply");
  if (mb > s->max_batch)
    return set_err(PHEM_EINVAL, "batch of %llu edges exceeds max_batch %llu",
                   (unsigned long long)mb, (unsigned long long)s->max_batch);
  if (mb && (!d_u || !d_v || !d_w)) return set_err(PHEM_EINVAL, "NULL batch");
  if (s->next_gid + mb > 0xFFFFFFFFFFFFFFFFull - 1)
    return set_err(PHEM_EINVAL, "global edge ids exhausted");
  hipStream_t st = (hipStream_t)stream;
  const int c = s->cur;
  if (mb) {
    PHEM_PROF("k_stream_stage", st);
    hipLaunchKernelGGL(k_stream_stage, dim3(grid_for(mb, SBLK_S)), dim3(SBLK_S), 0,
                       st, d_u, d_v, d_w, mb, (ull)s->next_gid, s->u[c] + s->f,
                       s->v[c] + s->f, s->w[c] + s->f, s->gid[c] + s->f,
                       s->packed);
  }
  PHEM_LAUNCHED("k_stream_stage");
  int rc;
  // The new forest (at most n - 1 edges) is gathered before the solve's
  // final counter read, so one wait covers the whole batch.
  const uint64_t total = s->f + mb;
  const uint64_t bound = s->n ? s->n - 1 : 0;
  const MstTail keep = [&](const ull *d_count) -> int {
    if (!bound) return PHEM_OK;
    PHEM_PROF("k_stream_keep", st);
    hipLaunchKernelGGL(k_stream_keep, dim3(grid_for(bound, SBLK_S)),
                       dim3(SBLK_S), 0, st, s->ids, d_count, s->u[c], s->v[c],
                       s->w[c], s->gid[c], s->u[c ^ 1], s->v[c ^ 1],
                       s->w[c ^ 1], s->gid[c ^ 1]);
    PHEM_LAUNCHED("k_stream_keep");
    return PHEM_OK;
  };
  // The batch's union-find is enqueued after the solve's kernels and before
  // its final counter read (in the tail), so a solve that fails before that
  // read leaves parent and union-find forest as they were.
  bool unioned = false;
  const MstTail tail = [&](const ull *d_count) -> int {
    int trc = keep(d_count);
    if (trc || !mb) return trc;
    unioned = true;
    return phem_cc_union(s->n, (const uint64_t *)s->packed, mb, s->parent,
                         (uint64_t *)s->cc_forest, stream);
  };
  // consecutive batches of a stream take about as many rounds: the previous
  // canonical solve's rounds after round 0 are launched before the first
  // counter read, which then usually finds the solve converged
  const int hint = s->last.rounds > 1 && s->last.tau == PHEM_NO_EDGE
                       ? (int)s->last.rounds - 1 : 0;
  phem_mst_stats last{};
  rc = mst_solve_tail(s->n, total, s->u[c], s->v[c], s->w[c], s->wmode, s->ws,
                      s->ws_bytes, s->ids, s->labels, nullptr, &last, st,
                      tail, hint);
  if (rc) {
    // once the union is enqueued the two states disagree: refuse later use
    if (unioned) s->broken = true;
    return rc;
  }
  s->last = last;
  const uint64_t count = s->last.mst_count;
  PHEM_HIP(hipStreamSynchronize(st));
  s->cur ^= 1;
  s->f = count;
  s->next_gid += mb;
  s->batches++;
  return PHEM_OK;
}

extern "C" int phem_stream_forest(const phem_stream_t *s, uint64_t *count,
                                  const uint32_t **d_u, const uint32_t **d_v,
                                  const uint32_t **d_w,
                                  const uint64_t **d_gid) {
  if (!s) return set_err(PHEM_EINVAL, "stream handle is NULL");
  if (count) *count = s->f;
  if (d_u) *d_u = s->u[s->cur];
  if (d_v) *d_v = s->v[s->cur];
  if (d_w) *d_w = s->w[s->cur];
  if (d_gid) *d_gid = (const uint64_t *)s->gid[s->cur];
  return PHEM_OK;
}

extern "C" int phem_stream_stats(const phem_stream_t *s, phem_mst_stats *stats,
                                 uint64_t *edges_inserted, uint64_t *batches) {
  if (!s) return set_err(PHEM_EINVAL, "stream handle is NULL");
  if (stats) *stats = s->last;
  if (edges_inserted) *edges_inserted = s->next_gid;
  if (batches) *batches = s->batches;
  return PHEM_OK;
}

extern "C" int phem_stream_labels(const phem_stream_t *s, uint32_t *d_labels,
                                  void *stream) {
  if (!s) return set_err(PHEM_EINVAL, "stream handle is NULL");
  return phem_cc_labels(s->n, s->parent, d_labels, stream);
}

extern "C" int phem_stream_cc_forest(const phem_stream_t *s,
                                     const uint64_t **d_forest) {
  if (!s) return set_err(PHEM_EINVAL, "stream handle is NULL");
  if (d_forest) *d_forest = (const uint64_t *)s->cc_forest;
  return PHEM_OK;
}

extern "C" int phem_stream_copy_cc_forest(const phem_stream_t *s,
                                          uint64_t *d_out, void *stream) {
  if (!s) return set_err(PHEM_EINVAL, "stream handle is NULL");
  if (!d_out) return set_err(PHEM_EINVAL, "d_out is NULL");
  if (s->n)
    PHEM_HIP(hipMemcpyAsync(d_out, s->cc_forest, 8ull * s->n,
                            hipMemcpyDeviceToDevice, (hipStream_t)stream));
  return PHEM_OK;
}

extern "C" int phem_stream_copy_forest(const phem_stream_t *s, uint32_t *d_u,
                                       uint32_t *d_v, uint32_t *d_w,
                                       uint64_t *d_gid, void *stream) {
  if (!s) return set_err(PHEM_EINVAL, "stream handle is NULL");
  hipStream_t st = (hipStream_t)stream;
  const int c = s->cur;
  if (s->f) {
    if (d_u) PHEM_HIP(hipMemcpyAsync(d_u, s->u[c], 4 * s->f, hipMemcpyDeviceToDevice, st));
    if (d_v) PHEM_HIP(hipMemcpyAsync(d_v, s->v[c], 4 * s->f, hipMemcpyDeviceToDevice, st));
    if (d_w) PHEM_HIP(hipMemcpyAsync(d_w, s->w[c], 4 * s->f, hipMemcpyDeviceToDevice, st));
    if (d_gid)
      PHEM_HIP(hipMemcpyAsync(d_gid, s->gid[c], 8 * s->f, hipMemcpyDeviceToDevice, st));
  }
  return PHEM_OK;
}

// ---------------------------------------------------------------------------
// Host snapshot of a stream: a header, then the forest (u, v, w as 32-bit
// words, global ids as 64-bit words, f entries each), the union-find parent
// array (n words) and its forest (n 64-bit records).

namespace phem {

struct StreamSnapHeader {
  char magic[8];
  uint32_t n;
  int32_t wmode;
  uint64_t f, next_gid, batches;
  phem_mst_stats last;
};
constexpr char SNAP_MAGIC[8] = {'P', 'H', 'E', 'M', 'S', 'T', 'R', '1'};

static uint64_t snap_bytes(uint64_t n, uint64_t f) {
  return sizeof(StreamSnapHeader) + 20ull * f + 12ull * n;
}

}  // namespace phem

extern "C" int phem_stream_snapshot_bytes(const phem_stream_t *s,
                                          uint64_t *bytes) {
  if (!s || !bytes) return set_err(PHEM_EINVAL, "stream handle or bytes is NULL");
  *bytes = snap_bytes(s->n, s->f);
  return PHEM_OK;
}

extern "C" int phem_stream_save(const phem_stream_t *s, void *host,
                                uint64_t bytes, void *stream) {
  if (!s || !host) return set_err(PHEM_EINVAL, "stream handle or buffer is NULL");
  if (s->broken)
    return set_err(PHEM_EINVAL, "stream is unusable after a failed apply");
  const uint64_t need = snap_bytes(s->n, s->f);
  if (bytes < need)
    return set_err(PHEM_EINVAL, "snapshot buffer of %llu B < %llu B",
                   (unsigned long long)bytes, (unsigned long long)need);
  StreamSnapHeader h;
  std::memset(&h, 0, sizeof(h));
  std::memcpy(h.magic, SNAP_MAGIC, 8);
  h.n = s->n;
  h.wmode = s->wmode;
  h.f = s->f;
  h.next_gid = s->next_gid;
  h.batches = s->batches;
  h.last = s->last;
  char *p = (char *)host;
  std::memcpy(p, &h, sizeof(h));
  p += sizeof(h);
  hipStream_t st = (hipStream_t)stream;
  const int c = s->cur;
  const uint64_t f = s->f, n = s->n;
  if (f) {
    PHEM_HIP(hipMemcpyAsync(p, s->u[c], 4 * f, hipMemcpyDeviceToHost, st));
    PHEM_HIP(hipMemcpyAsync(p + 4 * f, s->v[c], 4 * f, hipMemcpyDeviceToHost, st));
    PHEM_HIP(hipMemcpyAsync(p + 8 * f, s->w[c], 4 * f, hipMemcpyDeviceToHost, st));
    PHEM_HIP(hipMemcpyAsync(p + 12 * f, s->gid[c], 8 * f, hipMemcpyDeviceToHost, st));
  }
  p += 20 * f;
  PHEM_HIP(hipMemcpyAsync(p, s->parent, 4 * n, hipMemcpyDeviceToHost, st));
  PHEM_HIP(hipMemcpyAsync(p + 4 * n, s->cc_forest, 8 * n, hipMemcpyDeviceToHost, st));
  PHEM_HIP(hipStreamSynchronize(st));
  return PHEM_OK;
}

extern "C" int phem_stream_restore(const void *host, uint64_t bytes,
                                   uint64_t max_batch, phem_stream_t **out) {
  if (!out) return set_err(PHEM_EINVAL, "out is NULL");
  *out = nullptr;
  if (!host || bytes < sizeof(StreamSnapHeader))
    return set_err(PHEM_EINVAL, "snapshot is NULL or shorter than its header");
  StreamSnapHeader h;
  std::memcpy(&h, host, sizeof(h));
  if (std::memcmp(h.magic, SNAP_MAGIC, 8) != 0)
    return set_err(PHEM_EINVAL, "not a stream snapshot");
  if (h.n < 1 || h.f >= h.n || bytes < snap_bytes(h.n, h.f))
    return set_err(PHEM_EINVAL, "snapshot of n=%u, %llu forest edges is inconsistent "
                   "with its %llu bytes", h.n, (unsigned long long)h.f,
                   (unsigned long long)bytes);
  // the device kernels index with these values: check them before upload
  const char *p = (const char *)host + sizeof(h);
  const uint64_t f = h.f, n = h.n;
  {
    const uint32_t *su = (const uint32_t *)p, *sv = su + f;
    const char *gp = p + 12 * f;
    // the forest is kept in ascending global-id order (the tie order of
    // later batches depends on it): ids must strictly increase
    uint64_t prev = 0;
    for (uint64_t i = 0; i < f; i++) {
      uint64_t g;
      std::memcpy(&g, gp + 8 * i, 8);
      if (su[i] >= n || sv[i] >= n || su[i] == sv[i] || g >= h.next_gid ||
          (i > 0 && g <= prev))
        return set_err(PHEM_EINVAL, "snapshot forest edge %llu is invalid",
                       (unsigned long long)i);
      prev = g;
    }
    const char *pp = p + 20 * f, *fp = pp + 4 * n;
    for (uint64_t x = 0; x < n; x++) {
      uint32_t par;
      uint64_t rec;
      std::memcpy(&par, pp + 4 * x, 4);
      std::memcpy(&rec, fp + 8 * x, 8);
      const uint64_t a = rec >> 32, b = rec & 0xFFFFFFFFull;
      if (par > x || (rec != ~0ull && !(a < b && b < n)))
        return set_err(PHEM_EINVAL, "snapshot union-find entry %llu is invalid",
                       (unsigned long long)x);
    }
  }
  phem_stream_t *s = nullptr;
  int rc = phem_stream_create(h.n, max_batch, h.wmode, &s);
  if (rc) return rc;
  hipError_t e = hipSuccess;
  if (f) {
    e = hipMemcpy(s->u[0], p, 4 * f, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(s->v[0], p + 4 * f, 4 * f, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(s->w[0], p + 8 * f, 4 * f, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(s->gid[0], p + 12 * f, 8 * f, hipMemcpyHostToDevice);
  }
  p += 20 * f;
  if (e == hipSuccess) e = hipMemcpy(s->parent, p, 4 * n, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(s->cc_forest, p + 4 * n, 8 * n, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    phem_stream_destroy(s);
    return set_err(PHEM_EHIP, "snapshot upload: %s", hipGetErrorString(e));
  }
  s->cur = 0;
  s->f = f;
  s->next_gid = h.next_gid;
  s->batches = h.batches;
  s->last = h.last;
  *out = s;
  return PHEM_OK;
}
