/*
 * phem_oracle.c -- CPU restatement of PHEM's MST / connected-components path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Loaded by the test suite, the smoke
 * check and the benchmark's CPU-baseline leg; never by the product library.
 * Conventions are stated in phem_oracle.h.
 */
#include "phem_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define KEY_NONE 0xFFFFFFFFFFFFFFFFull

uint32_t or_wkey(uint32_t w, int wmode) {
  if (wmode == 0) return w;
  /* Numeric order of the reference's float comparator, (a.w < b.w) ||
   * (a.w == b.w && a.id < b.id) (MST/src/utils.cu:801-802,
   * MST/src/bench/serialMST.C:67-68): -0.0 equals +0.0 (ties by id);
   * negative values reversed and placed below the positive ones.  NaN,
   * unordered there, is defined as one value above +inf. */
  if ((w & 0x7FFFFFFFu) == 0) return 0x80000000u;
  if ((w & 0x7FFFFFFFu) > 0x7F800000u) return 0xFFFFFFFFu;
  return (w & 0x80000000u) ? ~w : (w | 0x80000000u);
}

static inline int edge_valid(uint32_t n, uint32_t a, uint32_t b) {
  return a < n && b < n && a != b;
}

static inline uint64_t edge_key(const uint32_t *w, int wmode, uint64_t id) {
  return ((uint64_t)or_wkey(w[id], wmode) << 32) | id;
}

/* ------------------------------------------------------------------------ */
/* sequential disjoint-set forest (union by size, path halving)              */

typedef struct {
  uint32_t *p;
  uint32_t *sz;
} dsu_t;

static int dsu_init(dsu_t *d, uint32_t n) {
  d->p = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
  d->sz = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
  if (!d->p || !d->sz) return -1;
  for (uint32_t i = 0; i < n; i++) {
    d->p[i] = i;
    d->sz[i] = 1;
  }
  return 0;
}
static void dsu_free(dsu_t *d) {
  free(d->p);
  free(d->sz);
}
static inline uint32_t dsu_find(dsu_t *d, uint32_t x) {
  while (d->p[x] != x) {
    d->p[x] = d->p[d->p[x]];
    x = d->p[x];
  }
  return x;
}
static inline int dsu_union(dsu_t *d, uint32_t a, uint32_t b) {
  a = dsu_find(d, a);
  b = dsu_find(d, b);
  if (a == b) return 0;
  if (d->sz[a] < d->sz[b]) {
    uint32_t t = a;
    a = b;
    b = t;
  }
  d->p[b] = a;
  d->sz[a] += d->sz[b];
  return 1;
}
/* min-id labels from a DSU */
static void dsu_labels(dsu_t *d, uint32_t n, uint32_t *labels) {
  uint32_t *mn = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
  for (uint32_t i = 0; i < n; i++) mn[i] = OR_NO_EDGE;
  for (uint32_t i = 0; i < n; i++) {
    uint32_t r = dsu_find(d, i);
    if (i < mn[r]) mn[r] = i;
  }
  for (uint32_t i = 0; i < n; i++) labels[i] = mn[dsu_find(d, i)];
  free(mn);
}

/* ------------------------------------------------------------------------ */
/* stable LSD radix sort of edge ids by their 32-bit weight key               */
/* Input ids ascending => output in (wkey, id) order.                         */

static void sort_ids_by_wkey(uint32_t *ids, uint64_t cnt, const uint32_t *w,
                             int wmode) {
  uint32_t *tmp = (uint32_t *)malloc(sizeof(uint32_t) * (cnt ? cnt : 1));
  uint64_t *hist = (uint64_t *)malloc(sizeof(uint64_t) * 65536);
  for (int pass = 0; pass < 2; pass++) {
    int sh = pass * 16;
    memset(hist, 0, sizeof(uint64_t) * 65536);
    for (uint64_t i = 0; i < cnt; i++)
      hist[(or_wkey(w[ids[i]], wmode) >> sh) & 0xFFFF]++;
    uint64_t s = 0;
    for (int b = 0; b < 65536; b++) {
      uint64_t c = hist[b];
      hist[b] = s;
      s += c;
    }
    for (uint64_t i = 0; i < cnt; i++) {
      uint32_t id = ids[i];
      tmp[hist[(or_wkey(w[id], wmode) >> sh) & 0xFFFF]++] = id;
    }
    memcpy(ids, tmp, sizeof(uint32_t) * cnt);
  }
  free(tmp);
  free(hist);
}

static int64_t flags_to_ids(const uint8_t *flag, uint64_t m, uint32_t *out) {
  int64_t c = 0;
  for (uint64_t i = 0; i < m; i++)
    if (flag[i]) out[c++] = (uint32_t)i;
  return c;
}

/* ------------------------------------------------------------------------ */

int64_t or_kruskal(uint32_t n, uint64_t m, const uint32_t *u, const uint32_t *v,
                   const uint32_t *w, int wmode, uint32_t *out_ids,
                   uint32_t *out_labels) {
  uint32_t *ids = (uint32_t *)malloc(sizeof(uint32_t) * (m ? m : 1));
  uint8_t *flag = (uint8_t *)calloc(m ? m : 1, 1);
  dsu_t d;
  if (!ids || !flag || dsu_init(&d, n)) return -1;
  uint64_t cnt = 0;
  for (uint64_t i = 0; i < m; i++)
    if (edge_valid(n, u[i], v[i])) ids[cnt++] = (uint32_t)i;
  sort_ids_by_wkey(ids, cnt, w, wmode);
  for (uint64_t k = 0; k < cnt; k++) {
    uint32_t e = ids[k];
    if (dsu_union(&d, u[e], v[e])) flag[e] = 1;
  }
  int64_t c = flags_to_ids(flag, m, out_ids);
  if (out_labels) dsu_labels(&d, n, out_labels);
  dsu_free(&d);
  free(ids);
  free(flag);
  return c;
}

/* ------------------------------------------------------------------------ */

/* partial: after max_rounds rounds, return the forest edges chosen so far and
 * the components so far instead of failing. */
static int64_t boruvka_impl(uint32_t n, uint64_t m, const uint32_t *u,
                            const uint32_t *v, const uint32_t *w, int wmode,
                            uint32_t *out_ids, uint32_t *out_labels,
                            uint64_t *log_m, uint64_t *log_c, int max_rounds,
                            int *rounds_out, int partial) {
  /* alive edge list: current component endpoints + original id */
  uint32_t *cu = (uint32_t *)malloc(sizeof(uint32_t) * (m ? m : 1));
  uint32_t *cv = (uint32_t *)malloc(sizeof(uint32_t) * (m ? m : 1));
  uint32_t *eid = (uint32_t *)malloc(sizeof(uint32_t) * (m ? m : 1));
  uint32_t *rep = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
  uint32_t *nrep = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
  uint64_t *best = (uint64_t *)malloc(sizeof(uint64_t) * (n ? n : 1));
  uint8_t *flag = (uint8_t *)calloc(m ? m : 1, 1);
  if (!cu || !cv || !eid || !rep || !nrep || !best || !flag) return -1;

  uint64_t alive = 0;
  for (uint64_t i = 0; i < m; i++)
    if (edge_valid(n, u[i], v[i])) {
      cu[alive] = u[i];
      cv[alive] = v[i];
      eid[alive] = (uint32_t)i;
      alive++;
    }
  for (uint32_t i = 0; i < n; i++) rep[i] = i; /* populate_representative_array */

  int r = 0;
  int64_t result = 0;
  while (alive > 0) {
    if (r >= max_rounds) {
      result = partial ? 0 : -1;
      break;
    }
    /* reset_outgoing_edge_array_to_sentinel + populate_min_weight/eid */
    for (uint32_t i = 0; i < n; i++) best[i] = KEY_NONE;
    for (uint64_t k = 0; k < alive; k++) {
      uint64_t key = edge_key(w, wmode, eid[k]);
      if (key < best[cu[k]]) best[cu[k]] = key;
      if (key < best[cv[k]]) best[cv[k]] = key;
    }
    uint64_t comps = 0;
    for (uint32_t i = 0; i < n; i++) comps += (best[i] != KEY_NONE);
    if (log_m) log_m[r] = alive;
    if (log_c) log_c[r] = comps;

    /* new_grafting_edges: orient the best edge, break mutual picks by lower id.
     * Component ids are vertex ids; best[c] of a non-root c is KEY_NONE. */
    for (uint32_t i = 0; i < n; i++) nrep[i] = rep[i];
    for (uint32_t c = 0; c < n; c++) {
      if (best[c] == KEY_NONE) continue;
      uint32_t e = (uint32_t)(best[c] & 0xFFFFFFFFu);
      uint32_t a = rep[u[e]], b = rep[v[e]]; /* current components of e */
      uint32_t d = (a == c) ? b : a;
      int same = (best[d] == best[c]);
      if (same && c < d) continue; /* c stays root */
      nrep[c] = d;
      flag[e] = 1;
    }
    /* new_shortcutting_step: every vertex to its root */
    for (uint32_t i = 0; i < n; i++) {
      uint32_t x = nrep[i];
      while (nrep[x] != x) x = nrep[x];
      rep[i] = x;
    }
    /* copy back so that roots are fixed points */
    for (uint32_t i = 0; i < n; i++) nrep[i] = rep[i];
    /* relabelling_step + remove_if(u == v) */
    uint64_t k2 = 0;
    for (uint64_t k = 0; k < alive; k++) {
      uint32_t a = rep[cu[k]], b = rep[cv[k]];
      if (a != b) {
        cu[k2] = a;
        cv[k2] = b;
        eid[k2] = eid[k];
        k2++;
      }
    }
    alive = k2;
    r++;
  }
  if (rounds_out) *rounds_out = r;
  if (result == 0) {
    result = flags_to_ids(flag, m, out_ids);
    if (out_labels) {
      /* canonicalise roots to the minimum vertex id */
      for (uint32_t i = 0; i < n; i++) nrep[i] = OR_NO_EDGE;
      for (uint32_t i = 0; i < n; i++)
        if (i < nrep[rep[i]]) nrep[rep[i]] = i;
      for (uint32_t i = 0; i < n; i++) out_labels[i] = nrep[rep[i]];
    }
  }
  free(cu);
  free(cv);
  free(eid);
  free(rep);
  free(nrep);
  free(best);
  free(flag);
  return result;
}

int64_t or_boruvka(uint32_t n, uint64_t m, const uint32_t *u, const uint32_t *v,
                   const uint32_t *w, int wmode, uint32_t *out_ids,
                   uint32_t *out_labels, uint64_t *log_m, uint64_t *log_c,
                   int max_rounds, int *rounds_out) {
  return boruvka_impl(n, m, u, v, w, wmode, out_ids, out_labels, log_m, log_c,
                      max_rounds, rounds_out, 0);
}

/* The state of the same round loop after `rounds` rounds (fewer if it ends
 * earlier): the forest edges chosen so far (ascending ids, returned count)
 * and min-id labels of the components so far. */
int64_t or_boruvka_rounds(uint32_t n, uint64_t m, const uint32_t *u,
                          const uint32_t *v, const uint32_t *w, int wmode,
                          int rounds, uint32_t *out_ids, uint32_t *out_labels) {
  return boruvka_impl(n, m, u, v, w, wmode, out_ids, out_labels, NULL, NULL,
                      rounds, NULL, 1);
}

/* ------------------------------------------------------------------------ */
/* union-find with the reference hook rule                                   */

/* find_compress (CC/host_spanning_tree.cu:10-21): find the root, then point
 * every vertex on the path whose parent is larger than the root at the root. */
static inline uint32_t find_compress_seq(uint32_t i, uint32_t *p) {
  uint32_t j = i;
  while (p[j] != j) j = p[j];
  uint32_t t;
  while ((t = p[i]) > j) {
    p[i] = j;
    i = t;
  }
  return j;
}

void or_uf_labels(uint32_t n, uint64_t m, const uint32_t *u, const uint32_t *v,
                  uint32_t *labels, uint32_t *forest) {
  uint32_t *p = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
  for (uint32_t i = 0; i < n; i++) p[i] = i;
  if (forest)
    for (uint32_t i = 0; i < n; i++) forest[i] = OR_NO_EDGE;
  for (uint64_t e = 0; e < m; e++) {
    if (!edge_valid(n, u[e], v[e])) continue;
    uint32_t a = find_compress_seq(u[e], p), b = find_compress_seq(v[e], p);
    if (a == b) continue;
    uint32_t hi = a > b ? a : b, lo = a > b ? b : a;
    p[hi] = lo; /* larger root hooks under the smaller */
    if (forest) forest[hi] = (uint32_t)e;
  }
  for (uint32_t i = 0; i < n; i++) labels[i] = find_compress_seq(i, p);
  free(p);
}

void or_bfs_labels(uint32_t n, uint64_t m, const uint32_t *u, const uint32_t *v,
                   uint32_t *labels) {
  uint64_t *off = (uint64_t *)calloc((uint64_t)n + 1, sizeof(uint64_t));
  uint64_t cnt = 0;
  for (uint64_t e = 0; e < m; e++)
    if (edge_valid(n, u[e], v[e])) {
      off[u[e] + 1]++;
      off[v[e] + 1]++;
      cnt += 2;
    }
  for (uint32_t i = 0; i < n; i++) off[i + 1] += off[i];
  uint32_t *adj = (uint32_t *)malloc(sizeof(uint32_t) * (cnt ? cnt : 1));
  uint64_t *pos = (uint64_t *)malloc(sizeof(uint64_t) * ((uint64_t)n + 1));
  memcpy(pos, off, sizeof(uint64_t) * ((uint64_t)n + 1));
  for (uint64_t e = 0; e < m; e++)
    if (edge_valid(n, u[e], v[e])) {
      adj[pos[u[e]]++] = v[e];
      adj[pos[v[e]]++] = u[e];
    }
  uint32_t *queue = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
  for (uint32_t i = 0; i < n; i++) labels[i] = OR_NO_EDGE;
  for (uint32_t s = 0; s < n; s++) {
    if (labels[s] != OR_NO_EDGE) continue;
    uint32_t qh = 0, qt = 0;
    queue[qt++] = s;
    labels[s] = s;
    while (qh < qt) {
      uint32_t x = queue[qh++];
      for (uint64_t k = off[x]; k < off[x + 1]; k++) {
        uint32_t y = adj[k];
        if (labels[y] == OR_NO_EDGE) {
          labels[y] = s;
          queue[qt++] = y;
        }
      }
    }
  }
  free(off);
  free(adj);
  free(pos);
  free(queue);
}

int or_check_forest(uint32_t n, uint64_t m, const uint32_t *u, const uint32_t *v,
                    const uint32_t *ids, uint64_t count, const uint32_t *labels) {
  int rc = 0;
  uint8_t *seen = (uint8_t *)calloc(m ? m : 1, 1);
  dsu_t d;
  if (!seen || dsu_init(&d, n)) return 99;
  for (uint64_t k = 0; k < count && !rc; k++) {
    uint32_t e = ids[k];
    if (e >= m) rc = 1;                             /* id out of range */
    else if (seen[e]) rc = 2;                       /* repeated id */
    else if (!edge_valid(n, u[e], v[e])) rc = 6;    /* not a graph edge */
    else {
      seen[e] = 1;
      if (!dsu_union(&d, u[e], v[e])) rc = 3;       /* cycle */
    }
  }
  if (!rc) {
    uint64_t comps = 0;
    for (uint32_t i = 0; i < n; i++) comps += (labels[i] == i);
    if (count != (uint64_t)n - comps) rc = 4;       /* |F| != n - c */
  }
  if (!rc) {
    uint32_t *fl = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
    dsu_labels(&d, n, fl);
    for (uint32_t i = 0; i < n && !rc; i++)
      if (fl[i] != labels[i]) rc = 5;               /* partition differs */
    free(fl);
  }
  dsu_free(&d);
  free(seen);
  return rc;
}

uint64_t or_weight_sum(uint64_t count, const uint32_t *ids, const uint32_t *w,
                       int wmode, int *exact) {
  uint64_t s = 0;
  int ok = 1;
  for (uint64_t k = 0; k < count; k++) {
    uint32_t bits = w[ids[k]];
    if (wmode == 0) {
      s += bits;
    } else {
      float f;
      memcpy(&f, &bits, 4);
      double d = (double)f * 16777216.0;
      if (!(d >= 0.0) || d != floor(d) || d >= 9.0e18) {
        ok = 0;
        continue;
      }
      s += (uint64_t)d;
    }
  }
  if (exact) *exact = ok;
  return s;
}

/* ------------------------------------------------------------------------ */
/* Parallel CPU baseline                                                     */

static inline uint32_t ld32(const uint32_t *p) {
  return __atomic_load_n(p, __ATOMIC_RELAXED);
}
static inline void st32(uint32_t *p, uint32_t x) {
  __atomic_store_n(p, x, __ATOMIC_RELAXED);
}
static inline int cas32(uint32_t *p, uint32_t o, uint32_t x) {
  return __atomic_compare_exchange_n(p, &o, x, 0, __ATOMIC_RELAXED,
                                     __ATOMIC_RELAXED);
}
static inline void write_min32(uint32_t *p, uint32_t x) {
  uint32_t c = ld32(p);
  while (x < c && !__atomic_compare_exchange_n(p, &c, x, 1, __ATOMIC_RELAXED,
                                               __ATOMIC_RELAXED)) {
  }
}

static int nthreads_of(int threads) {
#ifdef _OPENMP
  return threads > 0 ? threads : omp_get_max_threads();
#else
  (void)threads;
  return 1;
#endif
}

/* splitmix64 finaliser (Steele, Lea, Flood 2014) used for sampling */
static inline uint64_t mix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

/* parallel stable pack of indices [0, cnt) where keep(i) holds */
typedef int (*keep_fn)(uint64_t i, const void *ctx);

static uint64_t par_pack(uint64_t cnt, const uint32_t *in, uint32_t *out,
                         keep_fn keep, const void *ctx, int T) {
  uint64_t *tc = (uint64_t *)calloc((size_t)T + 1, sizeof(uint64_t));
#pragma omp parallel num_threads(T)
  {
#ifdef _OPENMP
    int t = omp_get_thread_num();
#else
    int t = 0;
#endif
    uint64_t lo = cnt * (uint64_t)t / (uint64_t)T,
             hi = cnt * (uint64_t)(t + 1) / (uint64_t)T, c = 0;
    for (uint64_t i = lo; i < hi; i++) c += keep(in[i], ctx) ? 1 : 0;
    tc[t + 1] = c;
#pragma omp barrier
#pragma omp single
    for (int k = 0; k < T; k++) tc[k + 1] += tc[k];
    uint64_t o = tc[t];
    for (uint64_t i = lo; i < hi; i++)
      if (keep(in[i], ctx)) out[o++] = in[i];
  }
  uint64_t total = tc[T];
  free(tc);
  return total;
}

/* parallel stable LSD radix sort of ids by their weight key (2 x 16 bits) */
static void par_sort_ids(uint32_t *ids, uint64_t cnt, const uint32_t *w,
                         int wmode, int T) {
  if (cnt < 2) return;
  uint32_t *tmp = (uint32_t *)malloc(sizeof(uint32_t) * cnt);
  uint64_t *h = (uint64_t *)malloc(sizeof(uint64_t) * 65536 * (size_t)T);
  for (int pass = 0; pass < 2; pass++) {
    int sh = pass * 16;
    memset(h, 0, sizeof(uint64_t) * 65536 * (size_t)T);
#pragma omp parallel num_threads(T)
    {
#ifdef _OPENMP
      int t = omp_get_thread_num();
#else
      int t = 0;
#endif
      uint64_t lo = cnt * (uint64_t)t / (uint64_t)T,
               hi = cnt * (uint64_t)(t + 1) / (uint64_t)T;
      uint64_t *ht = h + (size_t)t * 65536;
      for (uint64_t i = lo; i < hi; i++)
        ht[(or_wkey(w[ids[i]], wmode) >> sh) & 0xFFFF]++;
#pragma omp barrier
#pragma omp single
      {
        uint64_t s = 0;
        for (int b = 0; b < 65536; b++)
          for (int k = 0; k < T; k++) {
            uint64_t c = h[(size_t)k * 65536 + b];
            h[(size_t)k * 65536 + b] = s;
            s += c;
          }
      }
      for (uint64_t i = lo; i < hi; i++) {
        uint32_t id = ids[i];
        tmp[ht[(or_wkey(w[id], wmode) >> sh) & 0xFFFF]++] = id;
      }
    }
    memcpy(ids, tmp, sizeof(uint32_t) * cnt);
  }
  free(tmp);
  free(h);
}

/* PBBS union-find: parents[i] == ROOT marks a root; find shortcuts the path
 * (MST/src/algorithm/union_find.h:44-57); link(u, v) sets parents[u] = v. */
#define UF_ROOT 0xFFFFFFFFu
static inline uint32_t uf_find(uint32_t *par, uint32_t i) {
  uint32_t p = ld32(&par[i]);
  if (p == UF_ROOT) return i;
  uint32_t gp = ld32(&par[p]);
  if (gp == UF_ROOT) return p;
  do {
    st32(&par[i], gp);
    i = p;
    p = gp;
    gp = ld32(&par[p]);
  } while (gp != UF_ROOT);
  return p;
}

typedef struct {
  const uint32_t *u, *v, *w;
  int wmode;
  uint64_t kth_key;
  uint32_t *par;
  uint32_t n;
} mst_ctx_t;

static int keep_below_kth(uint64_t id, const void *c) {
  const mst_ctx_t *x = (const mst_ctx_t *)c;
  return edge_key(x->w, x->wmode, id) < x->kth_key;
}
static int keep_rest(uint64_t id, const void *c) {
  const mst_ctx_t *x = (const mst_ctx_t *)c;
  return edge_key(x->w, x->wmode, id) >= x->kth_key &&
         uf_find(x->par, x->u[id]) != uf_find(x->par, x->v[id]);
}
static int keep_valid(uint64_t id, const void *c) {
  const mst_ctx_t *x = (const mst_ctx_t *)c;
  return edge_valid(x->n, x->u[id], x->v[id]);
}

/* speculative_for over the sorted sequence E[0..cnt) with the UnionFindStep
 * reserve/commit (MST/src/utils.cu:768-791) and the round-size adaptation of
 * MST/src/common/speculative_for.h:51-113 (granularity 5). */
static void spec_for_uf(const uint32_t *E, uint64_t cnt, const uint32_t *u,
                        const uint32_t *v, uint32_t *par, uint32_t *R,
                        uint8_t *inST, int T) {
  if (cnt == 0) return;
  const int64_t gran = 5;
  int64_t maxRound = (int64_t)cnt / gran + 1;
  int64_t curRound = maxRound / 4;
  if (curRound < 1) curRound = 1;
  uint64_t *I = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)maxRound);
  uint64_t *hold = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)maxRound);
  uint32_t *cu = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)maxRound);
  uint32_t *cv = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)maxRound);
  uint8_t *keep = (uint8_t *)malloc((size_t)maxRound);
  int64_t nkeep = 0; /* failed iterations carried to the next round */
  uint64_t done = 0; /* iterations finished */
  while (done < cnt) {
    int64_t size = curRound;
    if ((uint64_t)size > cnt - done) size = (int64_t)(cnt - done);
    /* carried iterations first, then fresh ones: [done + nkeep, ...) */
#pragma omp parallel for num_threads(T) schedule(static)
    for (int64_t i = 0; i < size; i++) {
      uint64_t it = (i < nkeep) ? hold[i] : done + (uint64_t)i;
      I[i] = it;
      uint32_t e = E[it];
      uint32_t a = uf_find(par, u[e]), b = uf_find(par, v[e]);
      cu[i] = a;
      cv[i] = b;
      if (a != b) { /* reserve both endpoints for the smallest iteration */
        write_min32(&R[b], (uint32_t)it);
        write_min32(&R[a], (uint32_t)it);
        keep[i] = 1;
      } else {
        keep[i] = 0;
      }
    }
#pragma omp parallel for num_threads(T) schedule(static)
    for (int64_t i = 0; i < size; i++) {
      if (!keep[i]) continue;
      uint32_t it = (uint32_t)I[i];
      uint32_t a = cu[i], b = cv[i];
      if (ld32(&R[b]) == it) { /* commit: link b under a */
        if (ld32(&R[a]) == it) st32(&R[a], UF_ROOT);
        st32(&par[b], a);
        inST[E[it]] = 1;
        keep[i] = 0;
      } else if (ld32(&R[a]) == it) { /* link a under b */
        st32(&par[a], b);
        inST[E[it]] = 1;
        keep[i] = 0;
      }
    }
    int64_t k = 0;
    for (int64_t i = 0; i < size; i++)
      if (keep[i]) hold[k++] = I[i];
    nkeep = k;
    done += (uint64_t)(size - k);
    if ((double)k / (double)size > 0.2) {
      int64_t a2 = curRound / 2, b2 = maxRound / 64 + 1;
      if (b2 < k) b2 = k;
      curRound = a2 > b2 ? a2 : b2;
    } else if ((double)k / (double)size < 0.1) {
      curRound = curRound * 2 < maxRound ? curRound * 2 : maxRound;
    }
  }
  free(I);
  free(hold);
  free(cu);
  free(cv);
  free(keep);
}

int64_t or_par_mst(uint32_t n, uint64_t m, const uint32_t *u, const uint32_t *v,
                   const uint32_t *w, int wmode, uint32_t *out_ids,
                   int threads) {
  int T = nthreads_of(threads);
  uint32_t *all = (uint32_t *)malloc(sizeof(uint32_t) * (m ? m : 1));
  uint32_t *E1 = (uint32_t *)malloc(sizeof(uint32_t) * (m ? m : 1));
  uint32_t *par = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
  uint32_t *R = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
  uint8_t *inST = (uint8_t *)calloc(m ? m : 1, 1);
  if (!all || !E1 || !par || !R || !inST) return -1;
#pragma omp parallel for num_threads(T) schedule(static)
  for (int64_t i = 0; i < (int64_t)m; i++) all[i] = (uint32_t)i;
#pragma omp parallel for num_threads(T) schedule(static)
  for (int64_t i = 0; i < (int64_t)n; i++) {
    par[i] = UF_ROOT;
    R[i] = UF_ROOT;
  }
  mst_ctx_t ctx = {u, v, w, wmode, 0, par, n};
  /* edges with an endpoint out of range are not part of the graph */
  uint64_t mv = par_pack(m, all, E1, keep_valid, &ctx, T);
  memcpy(all, E1, sizeof(uint32_t) * mv);

  /* approximate k-th smallest by sampling (MST/src/algorithm/kth_smallest.h:28-38) */
  uint64_t k = (uint64_t)n * 5 / 4;
  if (k > mv) k = mv;
  uint64_t ns = mv ? (uint64_t)(mv / sqrt((double)mv)) : 0;
  uint64_t kth_key = KEY_NONE;
  if (ns > 0 && k < mv) {
    uint32_t *smp = (uint32_t *)malloc(sizeof(uint32_t) * ns);
    for (uint64_t i = 0; i < ns; i++) smp[i] = all[mix64(i) % mv];
    /* sort sample ids by full key (wkey, id): insertion via radix on id then wkey */
    /* sample ids are not in id order; sort by id first (stable radix on id) */
    uint32_t *tmp = (uint32_t *)malloc(sizeof(uint32_t) * ns);
    for (int pass = 0; pass < 2; pass++) {
      uint64_t h[65537];
      memset(h, 0, sizeof(h));
      for (uint64_t i = 0; i < ns; i++) h[((smp[i] >> (16 * pass)) & 0xFFFF) + 1]++;
      for (int b = 0; b < 65536; b++) h[b + 1] += h[b];
      for (uint64_t i = 0; i < ns; i++) tmp[h[(smp[i] >> (16 * pass)) & 0xFFFF]++] = smp[i];
      memcpy(smp, tmp, sizeof(uint32_t) * ns);
    }
    sort_ids_by_wkey(smp, ns, w, wmode);
    uint64_t at = k * ns / mv;
    if (at >= ns) at = ns - 1;
    kth_key = edge_key(w, wmode, smp[at]);
    free(smp);
    free(tmp);
  }
  ctx.kth_key = kth_key;

  /* prefix below the k-th key: filter, sort, speculative union-find */
  uint64_t c1 = par_pack(mv, all, E1, keep_below_kth, &ctx, T);
  par_sort_ids(E1, c1, w, wmode, T);
  spec_for_uf(E1, c1, u, v, par, R, inST, T);

  /* the rest: drop edges already inside one tree, sort, union-find */
  uint64_t c2 = par_pack(mv, all, E1, keep_rest, &ctx, T);
  par_sort_ids(E1, c2, w, wmode, T);
  spec_for_uf(E1, c2, u, v, par, R, inST, T);

  int64_t c = flags_to_ids(inST, m, out_ids);
  free(all);
  free(E1);
  free(par);
  free(R);
  free(inST);
  return c;
}

/* unite_impl (CC/host_spanning_tree.cu:24-42) with relaxed atomics */
static inline uint32_t find_compress_par(uint32_t i, uint32_t *p) {
  uint32_t j = i;
  if (ld32(&p[j]) == j) return j;
  do {
    j = ld32(&p[j]);
  } while (ld32(&p[j]) != j);
  uint32_t t;
  while ((t = ld32(&p[i])) > j) {
    st32(&p[i], j);
    i = t;
  }
  return j;
}

void or_par_cc(uint32_t n, uint64_t m, const uint32_t *u, const uint32_t *v,
               uint32_t *labels, uint32_t *forest, int threads) {
  int T = nthreads_of(threads);
  uint32_t *p = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
#pragma omp parallel for num_threads(T) schedule(static)
  for (int64_t i = 0; i < (int64_t)n; i++) {
    p[i] = (uint32_t)i;
    if (forest) forest[i] = OR_NO_EDGE;
  }
#pragma omp parallel for num_threads(T) schedule(dynamic, 4096)
  for (int64_t e = 0; e < (int64_t)m; e++) {
    uint32_t a = u[e], b = v[e];
    if (!edge_valid(n, a, b)) continue;
    while (a != b) {
      a = find_compress_par(a, p);
      b = find_compress_par(b, p);
      if (a > b && ld32(&p[a]) == a && cas32(&p[a], a, b)) {
        if (forest) forest[a] = (uint32_t)e;
        break;
      } else if (b > a && ld32(&p[b]) == b && cas32(&p[b], b, a)) {
        if (forest) forest[b] = (uint32_t)e;
        break;
      }
    }
  }
#pragma omp parallel for num_threads(T) schedule(static)
  for (int64_t i = 0; i < (int64_t)n; i++)
    labels[i] = find_compress_par((uint32_t)i, p);
  free(p);
}
