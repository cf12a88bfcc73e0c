// graphgen.cpp -- deterministic synthetic CSR graphs and source samplers.
//
// Host-only C++ (no GPU).  These make the synthetic inputs the multi-source
// BFS engine is measured and tested on: R-MAT graphs (the recursion of
// apps/ligra/utils/rMatGraph.C:33-71, here with a counter-based splitmix64
// hash of (seed, draw, bit) instead of Ligra's hashInt, so every draw is
// independent of thread count), Chung-Lu power-law graphs, and seeded source
// samplers (uniform over non-isolated vertices, or degree-weighted; the
// reference samples unseeded, scripts/2-datasets/src_nodes/
// generate_src_lists.py:128-160).
//
// Every CSR produced here has ascending adjacency lists, no self-loops and no
// duplicate edges, which is the form the engine's lowest-id parent rule needs.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <thread>
#include <vector>

namespace {

inline uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

inline uint64_t hash3(uint64_t seed, uint64_t a, uint64_t b) {
  return splitmix64(splitmix64(splitmix64(seed) ^ a) ^ (b * 0xD1B54A32D192ED03ull));
}

inline double unit(uint64_t h) { return (double)(h >> 11) * (1.0 / 9007199254740992.0); }

int resolve_threads(int t) {
  if (t > 0) return t;
  unsigned hc = std::thread::hardware_concurrency();
  if (hc == 0) hc = 1;
  return (int)std::min<unsigned>(hc, 16u);
}

void parallel_for(int64_t n, int nthreads, const std::function<void(int64_t, int64_t)>& fn) {
  if (n <= 0) return;
  nthreads = (int)std::max<int64_t>(1, std::min<int64_t>(nthreads, n));
  if (nthreads == 1) {
    fn(0, n);
    return;
  }
  std::vector<std::thread> th;
  int64_t chunk = (n + nthreads - 1) / nthreads;
  for (int t = 0; t < nthreads; ++t) {
    int64_t b = t * chunk, e = std::min<int64_t>(n, b + chunk);
    if (b >= e) break;
    th.emplace_back([=, &fn] { fn(b, e); });
  }
  for (auto& x : th) x.join();
}

}  // namespace

extern "C" {

typedef struct {
  int32_t n;
  int64_t m;
  int32_t* rowoff;  // n+1
  int32_t* col;     // m
} xbgen_csr;

void xbgen_free(xbgen_csr* g) {
  if (!g) return;
  free(g->rowoff);
  free(g->col);
  g->rowoff = nullptr;
  g->col = nullptr;
  g->n = 0;
  g->m = 0;
}

// Builds a CSR from k directed pairs (src[i] -> dst[i]); drops self-loops and
// duplicates and sorts each row.  Returns 0 on success.
static int build_csr(int32_t n, const int32_t* src, const int32_t* dst, int64_t k, int nthreads,
                     xbgen_csr* out) {
  std::vector<std::atomic<int64_t>> cnt((size_t)n);
  for (int32_t v = 0; v < n; ++v) cnt[v].store(0, std::memory_order_relaxed);
  parallel_for(k, nthreads, [&](int64_t b, int64_t e) {
    for (int64_t i = b; i < e; ++i)
      if (src[i] != dst[i]) cnt[src[i]].fetch_add(1, std::memory_order_relaxed);
  });
  std::vector<int64_t> off((size_t)n + 1, 0);
  for (int32_t v = 0; v < n; ++v) off[v + 1] = off[v] + cnt[v].load(std::memory_order_relaxed);
  int64_t tot = off[n];
  std::vector<int32_t> tmp((size_t)std::max<int64_t>(tot, 1));
  {
    std::vector<std::atomic<int64_t>> pos((size_t)n);
    for (int32_t v = 0; v < n; ++v) pos[v].store(off[v], std::memory_order_relaxed);
    parallel_for(k, nthreads, [&](int64_t b, int64_t e) {
      for (int64_t i = b; i < e; ++i) {
        if (src[i] == dst[i]) continue;
        int64_t p = pos[src[i]].fetch_add(1, std::memory_order_relaxed);
        tmp[p] = dst[i];
      }
    });
  }
  std::vector<int64_t> ndeg((size_t)n, 0);
  parallel_for(n, nthreads, [&](int64_t b, int64_t e) {
    for (int64_t v = b; v < e; ++v) {
      int32_t* s = tmp.data() + off[v];
      int64_t len = off[v + 1] - off[v];
      std::sort(s, s + len);
      int64_t u = std::unique(s, s + len) - s;
      ndeg[v] = u;
    }
  });
  int64_t m = 0;
  for (int32_t v = 0; v < n; ++v) m += ndeg[v];
  if (m > INT32_MAX) return -2;
  out->n = n;
  out->m = m;
  out->rowoff = (int32_t*)malloc(sizeof(int32_t) * ((size_t)n + 1));
  out->col = (int32_t*)malloc(sizeof(int32_t) * (size_t)std::max<int64_t>(m, 1));
  if (!out->rowoff || !out->col) return -3;
  out->rowoff[0] = 0;
  for (int32_t v = 0; v < n; ++v) out->rowoff[v + 1] = (int32_t)(out->rowoff[v] + ndeg[v]);
  parallel_for(n, nthreads, [&](int64_t b, int64_t e) {
    for (int64_t v = b; v < e; ++v)
      std::memcpy(out->col + out->rowoff[v], tmp.data() + off[v], sizeof(int32_t) * (size_t)ndeg[v]);
  });
  return 0;
}

// R-MAT: 2^scale vertices, ndraws edge draws.  symmetric != 0 adds the
// reverse of every draw.  Quadrant probabilities a, b, c, d = 1-a-b-c.
int xbgen_rmat_capped(int32_t scale, int32_t n_cap, int64_t ndraws, double a, double b, double c,
                      uint64_t seed, int32_t symmetric, int32_t nthreads, xbgen_csr* out);

int xbgen_rmat(int32_t scale, int64_t ndraws, double a, double b, double c, uint64_t seed,
               int32_t symmetric, int32_t nthreads, xbgen_csr* out) {
  if (scale < 1 || scale > 30) return -1;
  return xbgen_rmat_capped(scale, (int32_t)1 << scale, ndraws, a, b, c, seed, symmetric, nthreads, out);
}

// R-MAT on 2^scale ids restricted to the first n_cap: a draw with an endpoint
// at or above n_cap is dropped (so n need not be a power of two).
int xbgen_rmat_capped(int32_t scale, int32_t n_cap, int64_t ndraws, double a, double b, double c,
                      uint64_t seed, int32_t symmetric, int32_t nthreads, xbgen_csr* out) {
  if (!out || scale < 1 || scale > 30 || ndraws < 0 || a + b + c > 1.0 + 1e-12) return -1;
  if (n_cap < 1 || (int64_t)n_cap > ((int64_t)1 << scale)) return -1;
  nthreads = resolve_threads(nthreads);
  const int32_t n = n_cap;
  const double ab = a + b, abc = a + b + c;
  const int64_t k = symmetric ? 2 * ndraws : ndraws;
  std::vector<int32_t> src((size_t)std::max<int64_t>(k, 1)), dst((size_t)std::max<int64_t>(k, 1));
  parallel_for(ndraws, nthreads, [&](int64_t lo, int64_t hi) {
    for (int64_t i = lo; i < hi; ++i) {
      uint32_t u = 0, v = 0;
      for (int bit = scale - 1; bit >= 0; --bit) {
        double r = unit(hash3(seed, (uint64_t)i, (uint64_t)bit));
        if (r < a) {
        } else if (r < ab) {
          v |= 1u << bit;
        } else if (r < abc) {
          u |= 1u << bit;
        } else {
          u |= 1u << bit;
          v |= 1u << bit;
        }
      }
      // Out-of-range draws become self-loops of vertex 0, which the CSR build drops.
      if ((int64_t)u >= n || (int64_t)v >= n) u = v = 0;
      src[i] = (int32_t)u;
      dst[i] = (int32_t)v;
      if (symmetric) {
        src[ndraws + i] = (int32_t)v;
        dst[ndraws + i] = (int32_t)u;
      }
    }
  });
  return build_csr(n, src.data(), dst.data(), k, nthreads, out);
}

// Chung-Lu: vertex i has weight w_i proportional to (i + i0)^(-1/(gamma-1)),
// with i0 chosen so that max weight / mean weight = max_deg / mean_deg.
// ndraws = n*mean_deg/2 * overdraw undirected pairs with both endpoints drawn
// by weight (Vose alias tables), symmetrised.  permute != 0 relabels vertices
// by a seeded random permutation so that id carries no degree information.
int xbgen_chung_lu(int32_t n, double gamma, double mean_deg, double max_deg, double overdraw,
                   uint64_t seed, int32_t permute, int32_t nthreads, xbgen_csr* out) {
  if (!out || n < 2 || gamma <= 1.0 || mean_deg <= 0 || max_deg < mean_deg) return -1;
  nthreads = resolve_threads(nthreads);
  const double beta = 1.0 / (gamma - 1.0);
  const double target = max_deg / mean_deg;
  auto ratio = [&](double i0) {
    std::vector<double> part((size_t)nthreads, 0.0);
    int64_t chunk = ((int64_t)n + nthreads - 1) / nthreads;
    std::vector<std::thread> th;
    for (int t = 0; t < nthreads; ++t)
      th.emplace_back([&, t] {
        int64_t b = t * chunk, e = std::min<int64_t>(n, b + chunk);
        double s = 0;
        for (int64_t i = b; i < e; ++i) s += std::pow((double)i + i0, -beta);
        part[t] = s;
      });
    for (auto& x : th) x.join();
    double s = 0;
    for (double p : part) s += p;
    return std::pow(i0, -beta) / (s / n);
  };
  double lo = 1e-6, hi = (double)n;
  for (int it = 0; it < 60; ++it) {
    double mid = std::sqrt(lo * hi);
    if (ratio(mid) > target)
      lo = mid;
    else
      hi = mid;
  }
  const double i0 = std::sqrt(lo * hi);
  std::vector<double> w((size_t)n);
  double wsum = 0;
  for (int32_t i = 0; i < n; ++i) {
    w[i] = std::pow((double)i + i0, -beta);
    wsum += w[i];
  }
  // Vose alias table.
  std::vector<double> prob((size_t)n);
  std::vector<int32_t> alias((size_t)n, 0);
  {
    std::vector<double> p((size_t)n);
    std::vector<int32_t> small, large;
    small.reserve(n);
    large.reserve(n);
    for (int32_t i = 0; i < n; ++i) {
      p[i] = w[i] * n / wsum;
      (p[i] < 1.0 ? small : large).push_back(i);
    }
    while (!small.empty() && !large.empty()) {
      int32_t s = small.back();
      small.pop_back();
      int32_t l = large.back();
      large.pop_back();
      prob[s] = p[s];
      alias[s] = l;
      p[l] = (p[l] + p[s]) - 1.0;
      (p[l] < 1.0 ? small : large).push_back(l);
    }
    for (int32_t l : large) prob[l] = 1.0;
    for (int32_t s : small) prob[s] = 1.0;
  }
  std::vector<int32_t> perm((size_t)n);
  for (int32_t i = 0; i < n; ++i) perm[i] = i;
  if (permute) {
    for (int32_t i = n - 1; i > 0; --i) {
      uint64_t h = hash3(seed ^ 0x5EED5EEDull, 0xFFFFFFFFull, (uint64_t)i);
      int32_t j = (int32_t)(h % (uint64_t)(i + 1));
      std::swap(perm[i], perm[j]);
    }
  }
  const int64_t ndraws = (int64_t)std::llround((double)n * mean_deg / 2.0 * overdraw);
  std::vector<int32_t> src((size_t)2 * ndraws), dst((size_t)2 * ndraws);
  parallel_for(ndraws, nthreads, [&](int64_t lo2, int64_t hi2) {
    for (int64_t i = lo2; i < hi2; ++i) {
      int32_t e[2];
      for (int k = 0; k < 2; ++k) {
        uint64_t h = hash3(seed, (uint64_t)i, (uint64_t)k);
        int32_t col = (int32_t)((h >> 32) % (uint64_t)n);
        double r = (double)(h & 0xFFFFFFFFull) * (1.0 / 4294967296.0);
        e[k] = perm[r < prob[col] ? col : alias[col]];
      }
      src[i] = e[0];
      dst[i] = e[1];
      src[ndraws + i] = e[1];
      dst[ndraws + i] = e[0];
    }
  });
  return build_csr(n, src.data(), dst.data(), 2 * ndraws, nthreads, out);
}

// Samples S distinct sources.  mode 0: uniform over vertices with out-degree
// > 0; mode 1: proportional to out-degree.  Returns the number written (less
// than S only if fewer eligible vertices exist).
int32_t xbgen_sample_sources(int32_t n, const int32_t* rowoff, int32_t S, uint64_t seed, int32_t mode,
                             int32_t* out) {
  if (n <= 0 || S <= 0 || !rowoff || !out) return 0;
  std::vector<int32_t> cand;
  cand.reserve(n);
  for (int32_t v = 0; v < n; ++v)
    if (rowoff[v + 1] > rowoff[v]) cand.push_back(v);
  if (cand.empty()) return 0;
  int32_t want = (int32_t)std::min<int64_t>(S, (int64_t)cand.size());
  if (mode == 0) {
    // Partial Fisher-Yates over the candidate list.
    for (int32_t i = 0; i < want; ++i) {
      uint64_t h = hash3(seed, 0xC0FFEEull, (uint64_t)i);
      int32_t j = i + (int32_t)(h % (uint64_t)(cand.size() - i));
      std::swap(cand[i], cand[j]);
      out[i] = cand[i];
    }
    return want;
  }
  // Degree-weighted without replacement: inverse CDF draws, rejecting repeats.
  std::vector<int64_t> cdf(cand.size());
  int64_t acc = 0;
  for (size_t i = 0; i < cand.size(); ++i) {
    acc += rowoff[cand[i] + 1] - rowoff[cand[i]];
    cdf[i] = acc;
  }
  std::vector<uint8_t> taken(cand.size(), 0);
  int32_t got = 0;
  for (uint64_t draw = 0; got < want; ++draw) {
    uint64_t h = hash3(seed, 0xDE6ull, draw);
    int64_t r = (int64_t)(h % (uint64_t)acc);
    size_t i = (size_t)(std::upper_bound(cdf.begin(), cdf.end(), r) - cdf.begin());
    if (taken[i]) continue;
    taken[i] = 1;
    out[got++] = cand[i];
  }
  return got;
}

// FNV-1a 64 over (n, m, rowoff[0..n], col[0..m-1]) -- pins generated graphs.
uint64_t xbgen_checksum(int32_t n, const int32_t* rowoff, int64_t m, const int32_t* col) {
  uint64_t h = 0xCBF29CE484222325ull;
  auto mix = [&](const void* p, size_t len) {
    const uint8_t* b = (const uint8_t*)p;
    for (size_t i = 0; i < len; ++i) {
      h ^= b[i];
      h *= 0x100000001B3ull;
    }
  };
  mix(&n, sizeof(n));
  mix(&m, sizeof(m));
  mix(rowoff, sizeof(int32_t) * ((size_t)n + 1));
  mix(col, sizeof(int32_t) * (size_t)m);
  return h;
}

}  // extern "C"
