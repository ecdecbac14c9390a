// Host-only checks of the library's host code (CPU backend, graph-file
// readers and writer, error reporting), built with address / undefined /
// thread sanitizers by tests/harness/Makefile and run by
// tests/test_sanitize.py.  Each case is compared with the oracle's
// sequential Kruskal and BFS labels.  Exit status 0 = every check passed.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "phem.h"
#include "phem_oracle.h"

static int failures = 0;
#define CHECK(c)                                                          \
  do {                                                                    \
    if (!(c)) {                                                           \
      std::fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); \
      failures++;                                                         \
    }                                                                     \
  } while (0)

static uint64_t rng_state = 0x9E3779B97F4A7C15ull;
static uint64_t next_u64() {  // splitmix64
  uint64_t z = (rng_state += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

static void graph_case(uint32_t n, uint64_t m, int wmode, int threads) {
  std::vector<uint32_t> u(m), v(m), w(m);
  for (uint64_t i = 0; i < m; i++) {
    u[i] = (uint32_t)(next_u64() % (n + 2));  // a few out of range
    v[i] = (uint32_t)(next_u64() % (n + 2));
    const uint32_t k = (uint32_t)(next_u64() % 8);  // ties
    if (wmode == PHEM_W_F32) {
      float f = 0.125f * (float)k;
      std::memcpy(&w[i], &f, 4);
    } else {
      w[i] = k;
    }
  }
  std::vector<uint32_t> ids(n ? n : 1), want(n ? n : 1), lab(n ? n : 1),
      wlab(n ? n : 1);
  uint64_t cnt = 0;
  CHECK(phem_cpu_msf(n, m, u.data(), v.data(), w.data(), wmode, ids.data(), &cnt,
                     threads) == PHEM_OK);
  const int64_t kc = or_kruskal(n, m, u.data(), v.data(), w.data(), wmode,
                                want.data(), wlab.data());
  CHECK(kc >= 0 && (uint64_t)kc == cnt);
  CHECK(kc >= 0 && std::memcmp(ids.data(), want.data(), 4 * (size_t)kc) == 0);
  CHECK(phem_cpu_cc(n, m, u.data(), v.data(), lab.data(), threads) == PHEM_OK);
  or_bfs_labels(n, m, u.data(), v.data(), wlab.data());
  CHECK(std::memcmp(lab.data(), wlab.data(), 4 * (size_t)n) == 0);
}

static std::string tmp_path(const char *name) {
  const char *d = std::getenv("TMPDIR");
  return std::string(d && *d ? d : "/tmp") + "/phem_sanitize_" + name;
}

static void file_cases() {
  // MatrixMarket with weights, a comment, a self-loop and 1-based ids
  const std::string mtx = tmp_path("a.mtx");
  FILE *f = std::fopen(mtx.c_str(), "w");
  CHECK(f != nullptr);
  if (!f) return;
  std::fprintf(f, "%%%%MatrixMarket matrix coordinate integer general\n%% note\n"
                  "5 5 6\n1 2 7\n2 3 1\n3 3 4\n4 5 2\n5 1 9\n2 4 3\n");
  std::fclose(f);
  phem_graph_file g{};
  CHECK(phem_read_mtx(mtx.c_str(), 0, &g) == PHEM_OK);
  CHECK(g.n == 5 && g.m >= 5);
  if (g.m) {
    std::vector<uint32_t> ids(g.n);
    uint64_t cnt = 0;
    CHECK(phem_cpu_msf(g.n, g.m, g.u, g.v, g.w, g.wmode, ids.data(), &cnt, 2) ==
          PHEM_OK);
    CHECK(cnt == 4);
  }
  // written back and read again: same vertex count
  const std::string out = tmp_path("b.mtx");
  CHECK(phem_write_mtx_weighted(out.c_str(), &g) == PHEM_OK);
  phem_graph_file h{};
  CHECK(phem_read_mtx(out.c_str(), 0, &h) == PHEM_OK);
  CHECK(h.n == g.n);
  phem_graph_file_free(&h);
  phem_graph_file_free(&g);
  // HAST edge list: pairs in both directions, only u < v kept
  const std::string el = tmp_path("c.edges");
  f = std::fopen(el.c_str(), "w");
  CHECK(f != nullptr);
  if (!f) return;
  std::fprintf(f, "4 6\n0 1\n1 0\n1 2\n2 1\n2 3\n3 2\n");
  std::fclose(f);
  CHECK(phem_read_edge_list(el.c_str(), &g) == PHEM_OK);
  CHECK(g.n == 4 && g.m == 3);
  phem_graph_file_free(&g);
  // errors: a missing file and NULL arguments return codes, never crash
  CHECK(phem_read_mtx(tmp_path("missing.mtx").c_str(), 0, &g) != PHEM_OK);
  CHECK(std::strlen(phem_last_error()) > 0);
  CHECK(phem_read_mtx(nullptr, 0, &g) != PHEM_OK);
  uint64_t cnt = 0;
  CHECK(phem_cpu_msf(3, 1, nullptr, nullptr, nullptr, PHEM_W_U32, nullptr, &cnt, 1) !=
        PHEM_OK);
  std::remove(mtx.c_str());
  std::remove(out.c_str());
  std::remove(el.c_str());
}

int main() {
  for (int t = 0; t < 40; t++) {
    const uint32_t n = 1 + (uint32_t)(next_u64() % 300);
    const uint64_t m = next_u64() % 2000;
    graph_case(n, m, t % 2 ? PHEM_W_F32 : PHEM_W_U32, 1 + t % 4);
  }
  graph_case(20000, 200000, PHEM_W_F32, 8);
  graph_case(0, 0, PHEM_W_F32, 2);
  graph_case(1, 10, PHEM_W_U32, 2);
  file_cases();
  if (failures) std::fprintf(stderr, "%d checks failed\n", failures);
  else std::printf("sanitize_host: all checks passed\n");
  return failures ? 1 : 0;
}
