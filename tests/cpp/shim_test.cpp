// C++ call-site test of include/xb/zblossom.hpp, written the way the
// reference's engine tests drive zblossom::HGraph / XBlossomGpuEngine10
// (apps/xb/test/xblsm_gpu/xblsm_gpu_10_test.cu:57-73): load a CSR from the two
// text files, build the engine, run, check.  Here the check is the closed form
// of BFS on an R x C grid (level = Manhattan distance, parent = lowest-id
// neighbour one step closer), bit-exact, in both engine modes; a maximum
// matching of the same grid through FindMaximumMatch; the engine calls of the
// reference's matching runner (constructor forms, timing summary, counters);
// and a blossom case with a known answer.
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

  // The reference runner's engine calls (run_xb_pp_by_dataset.cu:21-69, here
  // written as our own loop): host_graph_t and Matching at global scope, the
  // one-argument constructor, FindMaximumMatch, PrintTimingSummary and every
  // counter getter, over a few rounds with a fresh engine per round.  The
  // graph is a random one (3,000 vertices, 4,500 LCG-drawn edge pairs) whose
  // search needs many phases, blossoms and reused trees; the expected
  // counters are the CPU restatement's for the same graph (maximum matching
  // 1,390 pairs, also Edmonds' size).
  std::vector<int32_t> rro(1, 0), rci;
  {
    const int rn = 3000;
    std::vector<std::vector<int32_t>> adj(rn);
    uint64_t x = 12345;
    auto next = [&x]() {
      x = x * 6364136223846793005ull + 1442695040888963407ull;
      return (int32_t)((x >> 33) % 3000);
    };
    for (int i = 0; i < 4500; ++i) {
      const int32_t u = next(), v = next();
      if (u == v) continue;
      adj[u].push_back(v);
      adj[v].push_back(u);
    }
    for (auto& a : adj) {
      std::sort(a.begin(), a.end());
      a.erase(std::unique(a.begin(), a.end()), a.end());
      rci.insert(rci.end(), a.begin(), a.end());
      rro.push_back((int32_t)rci.size());
    }
    const std::string rr = dir + "/rnd_row_offsets.txt", rc = dir + "/rnd_col_indices.txt";
    FILE* f = std::fopen(rr.c_str(), "w");
    for (int32_t v : rro) std::fprintf(f, "%d\n", v);
    std::fclose(f);
    f = std::fopen(rc.c_str(), "w");
    for (int32_t v : rci) std::fprintf(f, "%d\n", v);
    std::fclose(f);
    host_graph_t g2(rr, rc);
    size_t max_degree = 0;
    const auto& offsets = g2.GetRowOffsetsVector();
    for (size_t i = 0; i < g2.nnodes; ++i) max_degree = std::max(max_degree, (size_t)(offsets[i + 1] - offsets[i]));
    EXPECT(max_degree > 0 && g2.nedges == rci.size() / 2);
    Matching first;
    for (int round = 0; round < 3; ++round) {
      Matching initial_matching(g2.nnodes, (int32_t)g2.nnodes);
      zblossom::XBlossomGpuEngine10 xb_engine(g2);
      Matching actual = xb_engine.FindMaximumMatch(initial_matching);
      xb_engine.PrintTimingSummary();
      if (round == 0) first = actual;
      EXPECT(actual == first);  // identical every round
      size_t pairs = 0;
      for (int32_t v = 0; v < rn; ++v) {
        const int32_t w = actual[v];
        if (w == rn) continue;
        EXPECT(w >= 0 && w < rn && actual[w] == v &&
               std::binary_search(rci.begin() + rro[v], rci.begin() + rro[v + 1], w));
        pairs += (v < w);
      }
      EXPECT(pairs == 1390);
      EXPECT(xb_engine.NumBlossoms() == 100);
      EXPECT(xb_engine.NumAugmentingPathEdgesProcessed() == 45561);
      EXPECT(xb_engine.NumExpandEdgesProcessed() == 13559);
      EXPECT(xb_engine.NumBlossomEdgesProcessed() == 2556);
      EXPECT(xb_engine.NumEdgesProcessed() == 45561 + 13559 + 2556);
      EXPECT(xb_engine.NumReusedTreeNodes() == 6602);
      EXPECT(xb_engine.NumResetTreeNodes() == 32398);  // + reused = 3,000 x 13 phases
      EXPECT(xb_engine.LastMatchStats().phases == 13);
      double a = -1, e = -1, b = -1, t = -1;
      xb_engine.GetTimingSummary(a, e, b, t);
      EXPECT(a > 0 && e > 0 && b > 0 && t == a + e + b);
      // A second call on the same engine adds to the counters, as the
      // reference's members do.
      if (round == 2) {
        xb_engine.FindMaximumMatch(initial_matching);
        EXPECT(xb_engine.NumEdgesProcessed() == 2 * (45561 + 13559 + 2556));
        EXPECT(xb_engine.NumBlossoms() == 200);
      }
    }
    // Regrowing the forest every phase: the same maximum size, no reuse.
    zblossom::XBlossomGpuEngine10 regrow(g2, 80);
    regrow.RegrowForest(true);
    const Matching rm = regrow.FindMaximumMatch();
    size_t rpairs = 0;
    for (int32_t v = 0; v < rn; ++v) rpairs += (rm[v] != rn && v < rm[v]);
    EXPECT(rpairs == 1390 && regrow.NumReusedTreeNodes() == 0 && regrow.LastMatchStats().phases == 18);
  }

  // Blossoms: triangle {0, 1, 2} with 1 = 2 matched and a pendant 3 on 1.  The
  // blossom {0, 1, 2} makes 1 outer; its edge to root 3 closes the augmenting
  // path 3-1=2-0: mates [2, 3, 0, 1], one blossom.
  {
    zblossom::HGraph tri(std::vector<int32_t>{0, 2, 5, 7, 8}, std::vector<int32_t>{1, 2, 0, 2, 3, 0, 1, 1});
    zblossom::XBlossomGpuEngine10 eng(tri);
    const Matching got = eng.FindMaximumMatch(Matching{4, 2, 1, 4});
    EXPECT((got == Matching{2, 3, 0, 1}));
    EXPECT(eng.NumBlossoms() == 1);
    EXPECT(eng.NumBlossomEdgesProcessed() > 0);
  }
  if (fails) return 1;
  std::printf("shim_test ok\n");
  return 0;
}
