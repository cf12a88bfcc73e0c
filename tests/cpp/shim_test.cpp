// C++ call-site test of include/xb/zblossom.hpp, written the way the
// reference's engine tests drive zblossom::HGraph / XBlossomGpuEngine10
// (apps/xb/test/xblsm_gpu/xblsm_gpu_10_test.cu:57-73): load a CSR from the two
// text files, build the engine, run, check.  Here the check is the closed form
// of BFS on an R x C grid (level = Manhattan distance, parent = lowest-id
// neighbour one step closer), bit-exact, in both engine modes; and a
// maximum matching of the same grid through FindMaximumMatch.
//
// Usage: shim_test <tmpdir>      exit 0 = pass
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "xb/zblossom.hpp"

static int fails = 0;
#define EXPECT(c)                                                  \
  do {                                                             \
    if (!(c)) {                                                    \
      std::fprintf(stderr, "FAIL %s:%d: %s\n", __FILE__, __LINE__, #c); \
      ++fails;                                                     \
    }                                                              \
  } while (0)

int main(int argc, char** argv) {
  const std::string dir = argc > 1 ? argv[1] : "/tmp";
  const int R = 37, C = 53, n = R * C;
  std::vector<int32_t> ro(1, 0), ci;
  for (int v = 0; v < n; ++v) {
    int r = v / C, c = v % C;
    if (r > 0) ci.push_back(v - C);
    if (c > 0) ci.push_back(v - 1);
    if (c < C - 1) ci.push_back(v + 1);
    if (r < R - 1) ci.push_back(v + C);
    ro.push_back((int32_t)ci.size());
  }
  const std::string rp = dir + "/grid_row_offsets.txt", cp = dir + "/grid_col_indices.txt";
  {
    FILE* f = std::fopen(rp.c_str(), "w");
    for (int32_t x : ro) std::fprintf(f, "%d\n", x);
    std::fclose(f);
    f = std::fopen(cp.c_str(), "w");
    for (int32_t x : ci) std::fprintf(f, "%d\n", x);
    std::fclose(f);
  }
  zblossom::HGraph graph(rp, cp);
  EXPECT(graph.nnodes == (size_t)n);
  EXPECT(graph.nedges == ci.size() / 2);

  zblossom::XBlossomGpuEngine10 engine(graph, 80);
  std::vector<int32_t> sources = {0, n - 1, (R / 2) * C + C / 2, 5, n - 7};
  const xb_tune_report rep = engine.Autotune(sources);
  EXPECT(rep.trials >= 20 && rep.best_ms_per_source <= rep.default_ms_per_source);
  for (int mode : {XB_MODE_BATCHED, XB_MODE_PLAIN}) {
    engine.Options().mode = mode;
    std::vector<int32_t> lv, pa;
    engine.MultiSourceBFS(sources, lv, pa);
    for (size_t s = 0; s < sources.size(); ++s) {
      const int rs = sources[s] / C, cs = sources[s] % C;
      for (int v = 0; v < n; ++v) {
        const int r = v / C, c = v % C;
        const int lev = std::abs(r - rs) + std::abs(c - cs);
        int par = v;
        if (v != sources[s]) {
          par = INT32_MAX;
          if (r > rs) par = std::min(par, v - C);
          if (c > cs) par = std::min(par, v - 1);
          if (c < cs) par = std::min(par, v + 1);
          if (r < rs) par = std::min(par, v + C);
        }
        EXPECT(lv[s * n + v] == lev);
        EXPECT(pa[s * n + v] == par);
        if (fails > 10) return 1;
      }
    }
  }
  bool threw = false;
  try {
    std::vector<int32_t> lv, pa;
    engine.MultiSourceBFS({n}, lv, pa);
  } catch (const zblossom::XbException& e) {
    threw = e.code() == XB_ERR_SOURCE_RANGE;
  }
  EXPECT(threw);
  // FindMaximumMatch as the reference runner calls it (run_xb_pp_by_dataset.cu:
  // 42-44): initial matching all nnodes.  The grid is bipartite with an odd
  // vertex count, so a maximum matching leaves exactly one vertex exposed.
  const zblossom::Matching m = engine.FindMaximumMatch(zblossom::Matching(n, n));
  int exposed = 0;
  for (int v = 0; v < n; ++v) {
    if (m[v] == n) {
      ++exposed;
      continue;
    }
    const int d = std::abs(m[v] - v);
    EXPECT(m[v] >= 0 && m[v] < n && m[m[v]] == v && (d == C || (d == 1 && m[v] / C == v / C)));
  }
  EXPECT(exposed == 1);
  EXPECT(engine.LastMatchStats().blossom_edges == 0);
  EXPECT(engine.LastMatchStats().matched_vertices == n - 1);
  EXPECT(engine.FindMaximumMatch() == m);  // deterministic
  if (fails) return 1;
  std::printf("shim_test ok\n");
  return 0;
}
