// phem_io.cpp -- graph file readers (host code): MatrixMarket coordinate
// files, HAST edge lists, ECL binary CSR and PBBS weighted edge arrays.
//
// Files are memory-mapped and parsed by several threads in two passes: each
// thread counts the records in its byte range (cut at line or token starts),
// a prefix sum gives every range its first record index, and the second pass
// writes the records in file order.  Arrays are allocated with malloc and
// released by phem_graph_file_free.
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "phem.h"
#include "phem_internal.h"

namespace phem {
namespace {

struct Mapped {
  const char *p = nullptr;
  size_t size = 0;
  int fd = -1;
  ~Mapped() {
    if (p && size) munmap((void *)p, size);
    if (fd >= 0) close(fd);
  }
};

int map_file(const char *path, Mapped *mp) {
  if (!path) return set_err(PHEM_EINVAL, "path is NULL");
  mp->fd = open(path, O_RDONLY);
  if (mp->fd < 0)
    return set_err(PHEM_ENOTFOUND, "cannot open %s: %s", path, strerror(errno));
  struct stat sb;
  if (fstat(mp->fd, &sb) != 0)
    return set_err(PHEM_EINVAL, "cannot stat %s: %s", path, strerror(errno));
  mp->size = (size_t)sb.st_size;
  if (mp->size == 0) return PHEM_OK;
  void *q = mmap(nullptr, mp->size, PROT_READ, MAP_PRIVATE, mp->fd, 0);
  if (q == MAP_FAILED) {
    mp->size = 0;
    return set_err(PHEM_ENOMEM, "cannot map %s: %s", path, strerror(errno));
  }
  mp->p = (const char *)q;
  return PHEM_OK;
}

unsigned io_threads() {
  const unsigned h = std::thread::hardware_concurrency();
  return h == 0 ? 1u : (h > 16 ? 16u : h);
}

template <class F>
void run_ranges(unsigned T, F f) {
  if (T <= 1) {
    f(0u);
    return;
  }
  std::vector<std::thread> ts;
  for (unsigned t = 0; t < T; t++) ts.emplace_back(f, t);
  for (auto &t : ts) t.join();
}

bool is_space(char c) { return c == ' ' || c == '\t' || c == '\r' || c == '\n'; }
bool is_digit(char c) { return c >= '0' && c <= '9'; }

// Range cuts at line starts (cut[i] is the first byte of a line).
std::vector<const char *> line_cuts(const char *b, const char *e, unsigned T) {
  std::vector<const char *> cuts{b};
  for (unsigned t = 1; t < T; t++) {
    const char *c = b + (size_t)(e - b) * t / T;
    if (c <= cuts.back()) continue;
    while (c < e && c[-1] != '\n') c++;
    if (c < e && c > cuts.back()) cuts.push_back(c);
  }
  cuts.push_back(e);
  return cuts;
}

// Range cuts at token starts (a non-space byte after a space byte).
std::vector<const char *> token_cuts(const char *b, const char *e, unsigned T) {
  std::vector<const char *> cuts{b};
  for (unsigned t = 1; t < T; t++) {
    const char *c = b + (size_t)(e - b) * t / T;
    if (c <= cuts.back()) continue;
    while (c < e && !(is_space(c[-1]) && !is_space(*c))) c++;
    if (c < e && c > cuts.back()) cuts.push_back(c);
  }
  cuts.push_back(e);
  return cuts;
}

// unsigned decimal digits, accumulated modulo 2^32 as a 32-bit counter would
const char *parse_u32(const char *p, const char *e, uint32_t *out, bool *any) {
  uint32_t x = 0;
  *any = false;
  while (p < e && is_digit(*p)) {
    x = x * 10u + (uint32_t)(*p - '0');
    p++;
    *any = true;
  }
  *out = x;
  return p;
}

const char *skip_blank(const char *p, const char *e) {
  while (p < e && (*p == ' ' || *p == '\t')) p++;
  return p;
}

// A data line of a MatrixMarket body: "u v [w]" (1-based), or nothing.
struct MtxLine {
  uint32_t u, v, w;
  bool ok;
};

MtxLine parse_mtx_line(const char *p, const char *e) {
  MtxLine r{0, 0, 0, false};
  p = skip_blank(p, e);
  if (p >= e || *p == '\n' || *p == '\r' || *p == '%') return r;
  bool any;
  p = parse_u32(p, e, &r.u, &any);
  if (!any) return r;
  p = skip_blank(p, e);
  p = parse_u32(p, e, &r.v, &any);
  if (!any) return r;
  p = skip_blank(p, e);
  parse_u32(p, e, &r.w, &any);  // absent or non-digit weight reads as 0
  r.ok = true;
  return r;
}

const char *line_end(const char *p, const char *e) {
  const char *q = (const char *)memchr(p, '\n', (size_t)(e - p));
  return q ? q : e;
}

void free_graph(phem_graph_file *g) {
  if (!g) return;
  free(g->u);
  free(g->v);
  free(g->w);
  free(g->edges);
  g->u = g->v = g->w = nullptr;
  g->edges = nullptr;
}

}  // namespace
}  // namespace phem

using namespace phem;

extern "C" void phem_graph_file_free(phem_graph_file *g) {
  free_graph(g);
  if (g) {
    g->n = 0;
    g->m = 0;
  }
}

extern "C" int phem_read_mtx(const char *path, int generate_weights,
                             phem_graph_file *g) {
  if (!g) return set_err(PHEM_EINVAL, "g is NULL");
  std::memset(g, 0, sizeof(*g));
  Mapped mp;
  int rc = map_file(path, &mp);
  if (rc) return rc;
  const char *p = mp.p, *e = mp.p + mp.size;
  // header: comment lines, then "rows cols entries"
  while (p < e && *p == '%') p = line_end(p, e) + (line_end(p, e) < e);
  uint64_t dims[3] = {0, 0, 0};
  {
    const char *le = line_end(p, e);
    const char *q = p;
    for (int k = 0; k < 3; k++) {
      while (q < le && !is_digit(*q)) q++;
      if (q >= le) return set_err(PHEM_EINVAL, "%s: no dimension line", path);
      uint64_t x = 0;
      while (q < le && is_digit(*q)) x = x * 10 + (uint64_t)(*q++ - '0');
      dims[k] = x;
    }
    p = le < e ? le + 1 : e;
  }
  if (dims[0] > PHEM_MAX_VERTICES || dims[2] > PHEM_MAX_EDGES)
    return set_err(PHEM_EINVAL, "%s: %llu vertices / %llu entries beyond limits",
                   path, (unsigned long long)dims[0], (unsigned long long)dims[2]);
  const uint32_t n = (uint32_t)dims[0];
  if (generate_weights && n == 0)
    return set_err(PHEM_EINVAL, "%s: generated weights need n > 0", path);
  const uint64_t want = dims[2];
  const unsigned T = io_threads();
  const std::vector<const char *> cuts = line_cuts(p, e, T);
  const unsigned R = (unsigned)cuts.size() - 1;
  std::vector<uint64_t> cnt(R + 1, 0);
  run_ranges(R, [&](unsigned t) {
    uint64_t c = 0;
    for (const char *q = cuts[t]; q < cuts[t + 1];) {
      const char *le = line_end(q, cuts[t + 1]);
      if (parse_mtx_line(q, le).ok) c++;
      q = le + 1;
    }
    cnt[t + 1] = c;
  });
  for (unsigned t = 0; t < R; t++) cnt[t + 1] += cnt[t];
  const uint64_t m = std::min<uint64_t>(cnt[R], want);
  g->n = n;
  g->m = m;
  g->wmode = PHEM_W_U32;
  g->u = (uint32_t *)malloc(4 * (m ? m : 1));
  g->v = (uint32_t *)malloc(4 * (m ? m : 1));
  g->w = (uint32_t *)malloc(4 * (m ? m : 1));
  if (!g->u || !g->v || !g->w) {
    free_graph(g);
    return set_err(PHEM_ENOMEM, "%s: no host memory for %llu edges", path,
                   (unsigned long long)m);
  }
  run_ranges(R, [&](unsigned t) {
    uint64_t i = cnt[t];
    for (const char *q = cuts[t]; q < cuts[t + 1] && i < m;) {
      const char *le = line_end(q, cuts[t + 1]);
      const MtxLine L = parse_mtx_line(q, le);
      if (L.ok) {
        // 1-based ids; stored 0-based with u <= v (MST/src/utils.cu:1015-1022)
        const uint32_t a = L.u - 1u, b = L.v - 1u;
        g->u[i] = L.u > L.v ? b : a;
        g->v[i] = L.u > L.v ? a : b;
        // generated weight: ((u - 1)(v - 1) mod n) + 1 in 32-bit unsigned
        // arithmetic (MST/src/utils.cu:1002)
        g->w[i] = generate_weights ? (a * b) % n + 1u : L.w;
        i++;
      }
      q = le + 1;
    }
  });
  return PHEM_OK;
}

extern "C" int phem_read_edge_list(const char *path, phem_graph_file *g) {
  if (!g) return set_err(PHEM_EINVAL, "g is NULL");
  std::memset(g, 0, sizeof(*g));
  Mapped mp;
  int rc = map_file(path, &mp);
  if (rc) return rc;
  const char *b = mp.p, *e = mp.p + mp.size;
  // tokens: V E, then E pairs "u v"; each undirected edge is listed in both
  // directions and only u < v is kept (CC/graph.cu:45-90)
  const unsigned T = io_threads();
  const std::vector<const char *> cuts = token_cuts(b, e, T);
  const unsigned R = (unsigned)cuts.size() - 1;
  std::vector<uint64_t> cnt(R + 1, 0);
  run_ranges(R, [&](unsigned t) {
    uint64_t c = 0;
    for (const char *q = cuts[t]; q < cuts[t + 1];) {
      while (q < cuts[t + 1] && is_space(*q)) q++;
      if (q >= cuts[t + 1]) break;
      c++;
      while (q < cuts[t + 1] && !is_space(*q)) q++;
    }
    cnt[t + 1] = c;
  });
  for (unsigned t = 0; t < R; t++) cnt[t + 1] += cnt[t];
  const uint64_t ntok = cnt[R];
  if (ntok < 2) return set_err(PHEM_EINVAL, "%s: missing 'V E' header", path);
  std::vector<uint64_t> tok(ntok);
  bool bad = false;
  run_ranges(R, [&](unsigned t) {
    uint64_t k = cnt[t];
    for (const char *q = cuts[t]; q < cuts[t + 1];) {
      while (q < cuts[t + 1] && is_space(*q)) q++;
      if (q >= cuts[t + 1]) break;
      uint64_t x = 0;
      bool dig = true;
      while (q < cuts[t + 1] && !is_space(*q)) {
        if (!is_digit(*q)) dig = false;
        x = x * 10 + (uint64_t)(*q - '0');
        q++;
      }
      if (!dig) bad = true;
      tok[k++] = x;
    }
  });
  if (bad) return set_err(PHEM_EINVAL, "%s: non-numeric token", path);
  const uint64_t V = tok[0], E = tok[1];
  if (V > PHEM_MAX_VERTICES)
    return set_err(PHEM_EINVAL, "%s: %llu vertices beyond limits", path,
                   (unsigned long long)V);
  const uint64_t pairs = std::min<uint64_t>(E, (ntok - 2) / 2);
  uint64_t kept = 0;
  for (uint64_t i = 0; i < pairs; i++) kept += tok[2 + 2 * i] < tok[3 + 2 * i];
  g->n = (uint32_t)V;
  g->m = kept;
  g->edges = (uint64_t *)malloc(8 * (kept ? kept : 1));
  if (!g->edges) return set_err(PHEM_ENOMEM, "%s: no host memory", path);
  uint64_t j = 0;
  for (uint64_t i = 0; i < pairs; i++) {
    const uint64_t a = tok[2 + 2 * i], c = tok[3 + 2 * i];
    if (a < c) g->edges[j++] = ((uint64_t)(uint32_t)a << 32) | (uint32_t)c;
  }
  return PHEM_OK;
}

extern "C" int phem_read_ecl(const char *path, phem_graph_file *g) {
  if (!g) return set_err(PHEM_EINVAL, "g is NULL");
  std::memset(g, 0, sizeof(*g));
  Mapped mp;
  int rc = map_file(path, &mp);
  if (rc) return rc;
  // int64 V, int64 E, int64 offsets[V + 1], int32 neighbours[E]
  // (CC/graph.cu:92-121)
  if (mp.size < 16) return set_err(PHEM_EINVAL, "%s: short ECL header", path);
  int64_t V, E;
  std::memcpy(&V, mp.p, 8);
  std::memcpy(&E, mp.p + 8, 8);
  // sizes compared without multiplying E, so no E wraps to a matching size
  const uint64_t head = 16 + 8 * ((uint64_t)V + 1);  // V <= 2^32 here
  if (V < 0 || E < 0 || (uint64_t)V > PHEM_MAX_VERTICES || mp.size < head ||
      (mp.size - head) % 4 != 0 || (mp.size - head) / 4 != (uint64_t)E)
    return set_err(PHEM_EINVAL, "%s: ECL sizes (V=%lld, E=%lld) do not match "
                   "the file size %zu", path, (long long)V, (long long)E, mp.size);
  const char *offp = mp.p + 16, *nbp = offp + 8 * ((uint64_t)V + 1);
  auto off = [&](int64_t i) {
    int64_t x;
    std::memcpy(&x, offp + 8 * i, 8);
    return x;
  };
  auto nb = [&](int64_t j) {
    int32_t x;
    std::memcpy(&x, nbp + 4 * j, 4);
    return x;
  };
  for (int64_t i = 0; i <= V; i++) {
    const int64_t o = off(i);
    if (o < 0 || o > E || (i > 0 && o < off(i - 1)))
      return set_err(PHEM_EINVAL, "%s: bad CSR offset at vertex %lld", path,
                     (long long)i);
  }
  const unsigned T = io_threads();
  std::vector<uint64_t> cnt(T + 1, 0);
  run_ranges(T, [&](unsigned t) {
    uint64_t c = 0;
    for (int64_t i = V * t / T; i < V * (t + 1) / T; i++)
      for (int64_t j = off(i); j < off(i + 1); j++) c += (int64_t)nb(j) > i;
    cnt[t + 1] = c;
  });
  for (unsigned t = 0; t < T; t++) cnt[t + 1] += cnt[t];
  g->n = (uint32_t)V;
  g->m = cnt[T];
  g->edges = (uint64_t *)malloc(8 * (g->m ? g->m : 1));
  if (!g->edges) return set_err(PHEM_ENOMEM, "%s: no host memory", path);
  run_ranges(T, [&](unsigned t) {
    uint64_t k = cnt[t];
    for (int64_t i = V * t / T; i < V * (t + 1) / T; i++)
      for (int64_t j = off(i); j < off(i + 1); j++) {
        const int32_t x = nb(j);
        if ((int64_t)x > i) g->edges[k++] = ((uint64_t)i << 32) | (uint32_t)x;
      }
  });
  return PHEM_OK;
}

extern "C" int phem_read_pbbs_weighted(const char *path, phem_graph_file *g) {
  if (!g) return set_err(PHEM_EINVAL, "g is NULL");
  std::memset(g, 0, sizeof(*g));
  Mapped mp;
  int rc = map_file(path, &mp);
  if (rc) return rc;
  // "WeightedEdgeArray" then whitespace-separated triples "u v w"; weights
  // read as double and stored as float32 (common/graphIO.h:169-188 with the
  // float Weight of MST/src/bench/MST.h:7); n = max id + 1
  const char *b = mp.p, *e = mp.p + mp.size;
  static const char kHdr[] = "WeightedEdgeArray";
  while (b < e && is_space(*b)) b++;
  const size_t hl = sizeof(kHdr) - 1;
  if ((size_t)(e - b) < hl || std::memcmp(b, kHdr, hl) != 0 ||
      (b + hl < e && !is_space(b[hl])))
    return set_err(PHEM_EINVAL, "%s: missing WeightedEdgeArray header", path);
  b += hl;
  const unsigned T = io_threads();
  const std::vector<const char *> cuts = token_cuts(b, e, T);
  const unsigned R = (unsigned)cuts.size() - 1;
  std::vector<uint64_t> cnt(R + 1, 0);
  run_ranges(R, [&](unsigned t) {
    uint64_t c = 0;
    for (const char *q = cuts[t]; q < cuts[t + 1];) {
      while (q < cuts[t + 1] && is_space(*q)) q++;
      if (q >= cuts[t + 1]) break;
      c++;
      while (q < cuts[t + 1] && !is_space(*q)) q++;
    }
    cnt[t + 1] = c;
  });
  for (unsigned t = 0; t < R; t++) cnt[t + 1] += cnt[t];
  const uint64_t m = cnt[R] / 3;
  if (m > PHEM_MAX_EDGES) return set_err(PHEM_EINVAL, "%s: too many edges", path);
  g->m = m;
  g->wmode = PHEM_W_F32;
  g->u = (uint32_t *)malloc(4 * (m ? m : 1));
  g->v = (uint32_t *)malloc(4 * (m ? m : 1));
  g->w = (uint32_t *)malloc(4 * (m ? m : 1));
  if (!g->u || !g->v || !g->w) {
    free_graph(g);
    return set_err(PHEM_ENOMEM, "%s: no host memory", path);
  }
  std::vector<uint32_t> vmax(R, 0);
  bool bad = false;
  run_ranges(R, [&](unsigned t) {
    uint64_t k = cnt[t];
    uint32_t mx = 0;
    char buf[64];
    for (const char *q = cuts[t]; q < cuts[t + 1];) {
      while (q < cuts[t + 1] && is_space(*q)) q++;
      if (q >= cuts[t + 1]) break;
      const char *s = q;
      while (q < cuts[t + 1] && !is_space(*q)) q++;
      const uint64_t i = k / 3;
      const int field = (int)(k % 3);
      k++;
      if (i >= m) continue;
      const size_t len = std::min<size_t>((size_t)(q - s), sizeof(buf) - 1);
      std::memcpy(buf, s, len);
      buf[len] = 0;
      char *endp = nullptr;
      if (field < 2) {
        const long long x = std::strtoll(buf, &endp, 10);
        if (*endp || x < 0 || x > (long long)PHEM_MAX_VERTICES) {
          bad = true;
          continue;
        }
        (field == 0 ? g->u : g->v)[i] = (uint32_t)x;
        mx = std::max(mx, (uint32_t)x);
      } else {
        const float f = (float)std::strtod(buf, &endp);
        if (*endp) bad = true;
        std::memcpy(&g->w[i], &f, 4);
      }
    }
    vmax[t] = mx;
  });
  if (bad) {
    free_graph(g);
    return set_err(PHEM_EINVAL, "%s: malformed edge record", path);
  }
  uint32_t mx = 0;
  for (uint32_t x : vmax) mx = std::max(mx, x);
  g->n = m ? mx + 1 : 0;
  return PHEM_OK;
}

extern "C" int phem_write_mtx_weighted(const char *path, const phem_graph_file *g) {
  if (!path || !g) return set_err(PHEM_EINVAL, "NULL argument");
  if (g->m && (!g->u || !g->v || !g->w))
    return set_err(PHEM_EINVAL, "graph has no (u, v, w) arrays");
  FILE *f = std::fopen(path, "w");
  if (!f) return set_err(PHEM_EINVAL, "cannot create %s: %s", path, strerror(errno));
  // the reference's weighted dump: both directions of every edge, 1-based
  // (MST/src/new_boruvka.cu:399-408 with utils.cu:742-745's header)
  std::string out =
      "%%MatrixMarket matrix coordinate integer general\n"
      "%-------------------------------------------------------------------------------\n"
      "%-------------------------------------------------------------------------------\n";
  out += std::to_string(g->n) + " " + std::to_string(g->n) + " " +
         std::to_string(2 * g->m) + "\n";
  char line[96];
  for (uint64_t i = 0; i < g->m; i++) {
    const unsigned long long a = (unsigned long long)g->u[i] + 1,
                             b = (unsigned long long)g->v[i] + 1;
    int k = std::snprintf(line, sizeof(line), "%llu %llu %u\n%llu %llu %u\n", a, b,
                          g->w[i], b, a, g->w[i]);
    out.append(line, (size_t)k);
    if (out.size() > (8u << 20)) {
      std::fwrite(out.data(), 1, out.size(), f);
      out.clear();
    }
  }
  std::fwrite(out.data(), 1, out.size(), f);
  const bool err = std::ferror(f);
  std::fclose(f);
  return err ? set_err(PHEM_EINVAL, "write to %s failed", path) : PHEM_OK;
}
