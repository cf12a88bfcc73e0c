// xb_host.cpp -- host side of the MI355X multi-source BFS engine: CSR ingest
// and validation, device graph, the plain per-source and batched bit-parallel
// level loops, and the extern "C" boundary declared in include/xb/xb_capi.h.
//
// Reference counterparts (paths relative to the reference repository):
//   CSR ingest           apps/xb/src/include/graph/hgraph_gpu.h:37-48,
//                        apps/xb/src/include/common/utils_gpu.h:9-25
//   device graph         apps/xb/src/include/graph/dgraph.h:50-63
//   source seeding       apps/xb/src/xblsm_gpu/xblossom_gpu_engine_10.cu:83-185
//   level loop           apps/xb/src/xblsm_gpu/xblossom_gpu_engine_10.cu:72-81
//   expand step          apps/xb/src/xblsm_gpu/xblossom_gpu_engine_10.cu:270-359
//   push/pull switch     apps/ligra/ligra/ligra.h:236-277 (|F|+outdeg(F) > m/20)
//
// Unlike the reference's level loop (about ten blocking scalar reads per
// level), each level here costs one 8-byte device-to-host read of the packed
// (edges, count) queue counter.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cerrno>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <unordered_map>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "../../include/xb/xb_capi.h"
#include "xb_internal.hpp"

// The light emit of a tile in one launch instead of one launch per word group:
// the launch boundaries between groups (a grid's tail each) are gone
// (LiveJournal-shaped graph, 16-word tile: emit 18.4 -> 16.2 ms per step,
// 3 alternating rounds on one box).
#ifndef XB_EMIT_ONE_LAUNCH
#define XB_EMIT_ONE_LAUNCH 1
#endif
// A pull level's three in-degree buckets and its hub pieces in one launch,
// hub pieces and heavy buckets first (LiveJournal-shaped pull 11.3 -> 10.7 ms
// per step, 3 alternating rounds on one box).
#ifndef XB_PULL_FUSED
#define XB_PULL_FUSED 1
#endif
// A level leaves the words of its frontier set instead of clearing them (a
// launch per level over the queue's words, 0.73 ms per LiveJournal-shaped
// step); the next tile clears both frontier arrays with two memsets.  Left
// words are harmless: a word left in front[cur] at level d + 2 holds the
// level-d frontier bits of a vertex u, and every out-neighbour of u has those
// sources in its seen words since level d, so the pull's need masks and the
// push's seen filter drop them.  The words of every vertex that gains bits
// are overwritten, so the push, which reads the words of the queue's
// vertices only, sees exact frontier words for the new bits.
#ifndef XB_LAZY_FRONT_CLEAR
#define XB_LAZY_FRONT_CLEAR 1
#endif

namespace {

thread_local std::string g_last_error;

struct XbError {
  int32_t code;
  std::string msg;
};

[[noreturn]] void fail(int32_t code, const std::string& msg) { throw XbError{code, msg}; }

#define HIPCHK(expr)                                                                             \
  do {                                                                                           \
    hipError_t e_ = (expr);                                                                      \
    if (e_ != hipSuccess)                                                                        \
      fail(e_ == hipErrorOutOfMemory ? XB_ERR_NOMEM : XB_ERR_DEVICE,                             \
           std::string(#expr) + ": " + hipGetErrorString(e_));                                   \
  } while (0)

template <class F>
int32_t guarded(F&& f) {
  try {
    f();
    return XB_OK;
  } catch (const XbError& e) {
    g_last_error = e.msg;
    return e.code;
  } catch (const std::bad_alloc&) {
    g_last_error = "host allocation failed";
    return XB_ERR_NOMEM;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return XB_ERR_DEVICE;
  }
}

template <class T>
T* dalloc(size_t count) {
  void* p = nullptr;
  if (count == 0) count = 1;
  HIPCHK(hipMalloc(&p, count * sizeof(T)));
  return (T*)p;
}

void dfree(void* p) {
  if (p) (void)hipFree(p);
}

// Copies on the owning (non-blocking) stream and waits, so later work on that
// stream is ordered after the copy (a plain hipMemcpy/hipMemset runs on the
// null stream, which non-blocking streams do not wait for).
void h2d(void* dst, const void* src, size_t bytes, hipStream_t st) {
  if (!bytes) return;
  HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st));
  HIPCHK(hipStreamSynchronize(st));
}

void d2h(void* dst, const void* src, size_t bytes, hipStream_t st) {
  if (!bytes) return;
  HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
}

int resolve_threads() {
  unsigned hc = std::thread::hardware_concurrency();
  return (int)std::max(1u, std::min(hc, 16u));
}

template <class F>
void parallel_rows(int64_t n, F&& fn) {
  int T = resolve_threads();
  if (n < 65536) T = 1;
  std::vector<std::thread> th;
  int64_t chunk = (n + T - 1) / T;
  for (int t = 0; t < T; ++t) {
    int64_t b = t * chunk, e = std::min<int64_t>(n, b + chunk);
    if (b >= e) break;
    th.emplace_back([=, &fn] { fn(b, e); });
  }
  for (auto& x : th) x.join();
}

// Pinned host scratch for the per-level counter read.
struct Pinned {
  unsigned long long* p = nullptr;
  Pinned() { HIPCHK(hipHostMalloc((void**)&p, 64, hipHostMallocDefault)); }
  ~Pinned() {
    if (p) (void)hipHostFree(p);
  }
};

// Device state of one engine instance.  The stream and the pinned counter slot
// belong to the owner (the graph's default workspace uses the graph's; each
// serving slot has its own), so instances on different streams run at once.
struct Workspace {
  hipStream_t st = nullptr;
  unsigned long long* hctr = nullptr;  // pinned host slot for per-level counter reads
  int T = 0;
  int32_t n = 0;
  xb::BPState bp{};
  bool bp_dirty = false;
  // Frontier words left set by the last tile (levels do not clear the words of
  // their frontier): cleared at the start of the next tile.
  size_t front_dirty_words = 0;
  xb::Queue q[2]{};
  unsigned long long* ctrs = nullptr;  // q[0].ctr, q[1].ctr
  uint32_t* tcount = nullptr;
  int32_t* d_src = nullptr;
  int64_t src_cap = 0;
  int32_t* side = nullptr;
  int64_t side_cap = 0;
  int32_t* stage = nullptr;  // host-output staging: levels rows then parent rows
  int64_t stage_cap = 0;
  int32_t* q0_host = nullptr;  // pinned upload buffer for the first frontier
  int64_t q0_cap = 0;

  void free_bp() {
    dfree(bp.seen);
    dfree(bp.front[0]);
    dfree(bp.front[1]);
    dfree(bp.nextw);
    dfree(bp.planes);
    dfree(bp.planes_hi);
    dfree(bp.tflag);
    dfree(bp.tlist);
    bp = xb::BPState{};
    T = 0;
  }
  void free_all() {
    free_bp();
    for (auto& x : q) {
      dfree(x.v);
      dfree(x.eo);
    }
    dfree(ctrs);
    dfree(tcount);
    dfree(d_src);
    dfree(side);
    dfree(stage);
    if (q0_host) (void)hipHostFree(q0_host);
    hipStream_t keep_st = st;
    unsigned long long* keep_hctr = hctr;
    *this = Workspace{};
    st = keep_st;
    hctr = keep_hctr;
  }
};

}  // namespace

struct xb_graph {
  int32_t n = 0;
  int64_t m = 0;
  uint32_t flags = 0;
  std::vector<int32_t> h_off;  // out-edge offsets, original ids (host copy)
  int32_t* d_off = nullptr;    // original-id CSR: plain engine, traversed-edge counts
  int32_t* d_col = nullptr;
  int32_t* d_oin_off = nullptr;  // original-order in-offsets (the emit); == d_off when symmetric
  int32_t* d_oin_col = nullptr;  // original-order in-lists with engine-id entries (the emit)
  int32_t* d_oin_colo = nullptr; // the same in-lists with original ids; == d_col when symmetric
  // Bit-parallel engine graph on engine ids (descending out-degree order).
  std::vector<int32_t> h_new_of;    // original id -> engine id
  std::vector<int32_t> h_ein_off;   // engine in-edge offsets (layout of the pull and emit)
  int32_t* d_new_of = nullptr;
  int32_t* d_orig_of = nullptr;
  int32_t* d_eoff = nullptr;        // engine out-CSR (push); == the in-CSR when symmetric
  int32_t* d_ecol = nullptr;
  int32_t* d_ein_off = nullptr;     // engine in-CSR, each list in ascending ORIGINAL id
  int32_t* d_ein_col = nullptr;     // in-neighbours as engine ids
  int32_t* d_ein_col_o = nullptr;   // the same in-neighbours as original ids
  hipStream_t stream = nullptr;
  Pinned pinned;
  // Pinned upload buffer of the allocation-free entry point: a tile's first
  // frontier (distinct sources and their degree prefix) and its source ids,
  // at most 3 x 1,024 entries.  Allocated with the graph.
  int32_t* q0_pinned = nullptr;
  // In-degree layout for the pull and emit, rebuilt when a cut changes.
  int32_t cut_small = -1, cut_medium = -1, cut_hub = -1, hub_chunk = -1;
  // Bumped on every rebuild of the layout (d_perm, d_hub_chunks and d_hidx are
  // freed and reallocated then).  A batch or server holding those pointers
  // records the generation it was built with and refuses to run on another:
  // equal cut values after a change and a change back do not mean the same
  // device arrays.
  uint64_t layout_gen = 0;
  int32_t* d_perm = nullptr;          // engine ids with in-degree > 0, bucket-major, ascending per bucket
  int32_t bucket_end[3] = {0, 0, 0};  // perm ranges of the small / medium / large buckets
  // A bucket whose perm range is consecutive ids (degree-ordered ids of a
  // symmetric graph) is addressed as index + shift, without the list.
  bool bucket_contig[3] = {false, false, false};
  int32_t bucket_shift[3] = {0, 0, 0};
  int64_t bucket_edges[3] = {0, 0, 0};
  xb::HubChunk* d_hub_chunks = nullptr;
  int32_t n_hub_chunks = 0;
  int32_t n_hubs = 0;
  int64_t hub_edges = 0;
  int32_t* d_hidx = nullptr;   // engine id -> row of a heavy vertex (in-degree > cut_medium) in hpar, else -1
  int32_t n_heavy = 0;
  Workspace ws;
  // Second workspace of the batched engine, with its own stream: a run of
  // several tiles alternates the two, so one tile's level loop overlaps the
  // previous tile's emit.  Created by the first such run (ensure_ws2).
  hipStream_t stream2 = nullptr;
  std::unique_ptr<Pinned> pinned2;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  Workspace ws2;
  // Forest-mode state (about 90 B per vertex), kept from the first matching
  // run until the graph is destroyed.
  std::shared_ptr<void> match_state;

  int32_t outdeg(int32_t v) const { return h_off[v + 1] - h_off[v]; }

  ~xb_graph() {
    ws.free_all();
    ws2.free_all();
    if (ev_fork) (void)hipEventDestroy(ev_fork);
    if (ev_join) (void)hipEventDestroy(ev_join);
    if (stream2) (void)hipStreamDestroy(stream2);
    if (d_eoff != d_ein_off) dfree(d_eoff);
    if (d_ecol != d_ein_col) dfree(d_ecol);
    dfree(d_ein_off);
    dfree(d_ein_col);
    dfree(d_ein_col_o);
    dfree(d_new_of);
    dfree(d_orig_of);
    if (d_oin_off != d_off) dfree(d_oin_off);
    if (d_oin_colo != d_col) dfree(d_oin_colo);
    dfree(d_oin_col);
    dfree(d_off);
    dfree(d_col);
    dfree(d_perm);
    dfree(d_hub_chunks);
    dfree(d_hidx);
    if (q0_pinned) (void)hipHostFree(q0_pinned);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

// Builds the pull layout for the default cut-offs (defined with the layout
// code below), so that a fresh graph can run the allocation-free entry point.
void init_default_layout(xb_graph* g);

void ensure_device() {
  int cnt = 0;
  hipError_t e = hipGetDeviceCount(&cnt);
  if (e != hipSuccess || cnt <= 0) fail(XB_ERR_NO_DEVICE, "no HIP device visible");
}

// Validates and (if needed) sorts; returns the possibly sorted column array.
void validate_csr(const int32_t* off, int64_t n, const int32_t* col, int64_t m, std::vector<int32_t>& col_out) {
  if (n < 0) fail(XB_ERR_INVALID_ARG, "n_plus_1 must be >= 1");
  if (m < 0 || m > INT32_MAX) fail(XB_ERR_INVALID_ARG, "m out of range");
  if (n > INT32_MAX - 1) fail(XB_ERR_INVALID_ARG, "n out of range");
  if (off[0] != 0) fail(XB_ERR_BAD_CSR, "row_offsets[0] != 0");
  for (int64_t v = 0; v < n; ++v)
    if (off[v + 1] < off[v]) fail(XB_ERR_BAD_CSR, "row_offsets not monotone at row " + std::to_string(v));
  if (off[n] != m) fail(XB_ERR_BAD_CSR, "row_offsets[n] != number of column indices");
  std::vector<char> bad((size_t)resolve_threads() + 1, 0), unsorted((size_t)resolve_threads() + 1, 0);
  std::vector<int64_t> first_bad((size_t)resolve_threads() + 1, -1);
  col_out.assign(col, col + m);
  // Range check and sortedness check in parallel.
  {
    int T = resolve_threads();
    if (n < 65536) T = 1;
    std::vector<std::thread> th;
    int64_t chunk = (n + T - 1) / T;
    for (int t = 0; t < T; ++t) {
      int64_t b = t * chunk, e = std::min<int64_t>(n, b + chunk);
      if (b >= e) break;
      th.emplace_back([&, t, b, e] {
        for (int64_t v = b; v < e; ++v) {
          for (int32_t j = off[v]; j < off[v + 1]; ++j) {
            int32_t u = col[j];
            if (u < 0 || u >= n) {
              bad[t] = 1;
              if (first_bad[t] < 0) first_bad[t] = j;
            }
            if (j > off[v] && col[j - 1] > u) unsorted[t] = 1;
          }
        }
      });
    }
    for (auto& x : th) x.join();
  }
  for (size_t t = 0; t < bad.size(); ++t)
    if (bad[t])
      fail(XB_ERR_BAD_CSR, "column index out of range at position " + std::to_string(first_bad[t]));
  bool any_unsorted = false;
  for (char c : unsorted) any_unsorted |= (c != 0);
  if (any_unsorted) {
    parallel_rows(n, [&](int64_t b, int64_t e) {
      for (int64_t v = b; v < e; ++v) std::sort(col_out.begin() + off[v], col_out.begin() + off[v + 1]);
    });
  }
}

// Engine ids: vertices in descending out-degree order, ties by ascending
// original id (a counting sort on the degree).  The pull and the emit gather
// per-vertex state of in-neighbours, and an in-neighbour u is gathered once
// per out-edge of u, so this order packs the state that most gathers hit into
// the lowest ids, where it stays cached.  In-lists keep ascending ORIGINAL id
// order and carry the original ids beside the engine ids, so "first provider
// in list order" remains the lowest original id.
void build_engine_graph(xb_graph* g, const int32_t* off, const int32_t* col, const int32_t* in_off,
                        const int32_t* in_col, bool directed) {
  const int32_t n = g->n;
  const int64_t m = g->m;
  int32_t maxd = 0;
  for (int32_t v = 0; v < n; ++v) maxd = std::max(maxd, off[v + 1] - off[v]);
  std::vector<int64_t> start((size_t)maxd + 2, 0);
  for (int32_t v = 0; v < n; ++v) start[(size_t)(maxd - (off[v + 1] - off[v])) + 1]++;
  for (size_t d = 1; d < start.size(); ++d) start[d] += start[d - 1];
  std::vector<int32_t> orig_of((size_t)n);
  for (int32_t v = 0; v < n; ++v) orig_of[(size_t)start[(size_t)(maxd - (off[v + 1] - off[v]))]++] = v;
  g->h_new_of.assign((size_t)n, 0);
  for (int32_t i = 0; i < n; ++i) g->h_new_of[(size_t)orig_of[i]] = i;
  const std::vector<int32_t>& new_of = g->h_new_of;
  // Engine CSR from an original-id CSR: row i = the list of orig_of[i], in its
  // original order, as engine ids (and as original ids when `orig` is given).
  auto relabel = [&](const int32_t* ro, const int32_t* co, std::vector<int32_t>& eoff, std::vector<int32_t>& ecol,
                     std::vector<int32_t>* orig) {
    eoff.assign((size_t)n + 1, 0);
    for (int32_t i = 0; i < n; ++i) eoff[i + 1] = eoff[i] + (ro[orig_of[i] + 1] - ro[orig_of[i]]);
    ecol.assign((size_t)std::max<int64_t>(m, 1), 0);
    if (orig) orig->assign((size_t)std::max<int64_t>(m, 1), 0);
    parallel_rows(n, [&](int64_t b, int64_t e) {
      for (int64_t i = b; i < e; ++i) {
        const int32_t o = orig_of[(size_t)i];
        int32_t p = eoff[(size_t)i];
        for (int32_t j = ro[o]; j < ro[o + 1]; ++j, ++p) {
          ecol[(size_t)p] = new_of[(size_t)co[j]];
          if (orig) (*orig)[(size_t)p] = co[j];
        }
      }
    });
  };
  std::vector<int32_t> ein_col, ein_col_o;
  relabel(in_off, in_col, g->h_ein_off, ein_col, &ein_col_o);
  g->d_new_of = dalloc<int32_t>((size_t)n);
  g->d_orig_of = dalloc<int32_t>((size_t)n);
  h2d(g->d_orig_of, orig_of.data(), sizeof(int32_t) * (size_t)n, g->stream);
  g->d_ein_off = dalloc<int32_t>((size_t)n + 1);
  g->d_ein_col = dalloc<int32_t>((size_t)m);
  g->d_ein_col_o = dalloc<int32_t>((size_t)m);
  h2d(g->d_new_of, new_of.data(), sizeof(int32_t) * (size_t)n, g->stream);
  h2d(g->d_ein_off, g->h_ein_off.data(), sizeof(int32_t) * ((size_t)n + 1), g->stream);
  h2d(g->d_ein_col, ein_col.data(), sizeof(int32_t) * (size_t)m, g->stream);
  h2d(g->d_ein_col_o, ein_col_o.data(), sizeof(int32_t) * (size_t)m, g->stream);
  {
    // The emit's in-lists: original vertex and list order, engine-id entries.
    std::vector<int32_t> oin((size_t)std::max<int64_t>(m, 1));
    parallel_rows(m, [&](int64_t b, int64_t e) {
      for (int64_t j = b; j < e; ++j) oin[(size_t)j] = new_of[(size_t)in_col[j]];
    });
    g->d_oin_col = dalloc<int32_t>((size_t)m);
    h2d(g->d_oin_col, oin.data(), sizeof(int32_t) * (size_t)m, g->stream);
  }
  if (directed) {
    std::vector<int32_t> eoff, ecol;
    relabel(off, col, eoff, ecol, nullptr);
    g->d_eoff = dalloc<int32_t>((size_t)n + 1);
    g->d_ecol = dalloc<int32_t>((size_t)m);
    h2d(g->d_eoff, eoff.data(), sizeof(int32_t) * ((size_t)n + 1), g->stream);
    h2d(g->d_ecol, ecol.data(), sizeof(int32_t) * (size_t)m, g->stream);
  } else {
    g->d_eoff = g->d_ein_off;
    g->d_ecol = g->d_ein_col;
  }
}

xb_graph* build_graph(const int32_t* off, int64_t n_plus_1, const int32_t* col, int64_t m, uint32_t flags) {
  if (!off || n_plus_1 < 1 || (m > 0 && !col)) fail(XB_ERR_INVALID_ARG, "null or empty CSR arrays");
  if (flags & ~XB_GRAPH_DIRECTED) fail(XB_ERR_INVALID_ARG, "unknown graph flag");
  const int64_t n = n_plus_1 - 1;
  std::vector<int32_t> colv;
  validate_csr(off, n, col, m, colv);
  ensure_device();
  std::unique_ptr<xb_graph> g(new xb_graph());
  g->n = (int32_t)n;
  g->m = m;
  g->flags = flags;
  g->h_off.assign(off, off + n + 1);
  HIPCHK(hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking));
  g->ws.st = g->stream;
  g->ws.hctr = g->pinned.p;
  g->d_off = dalloc<int32_t>((size_t)n + 1);
  g->d_col = dalloc<int32_t>((size_t)m);
  h2d(g->d_off, off, sizeof(int32_t) * ((size_t)n + 1), g->stream);
  h2d(g->d_col, colv.data(), sizeof(int32_t) * (size_t)m, g->stream);
  std::vector<int32_t> ioff, icol;
  const int32_t* in_off = off;
  const int32_t* in_col = colv.data();
  if (flags & XB_GRAPH_DIRECTED) {
    // Transpose by counting sort on the destination; sources are visited in
    // ascending order, so every in-list comes out ascending.
    ioff.assign((size_t)n + 1, 0);
    for (int64_t j = 0; j < m; ++j) ioff[(size_t)colv[j] + 1]++;
    for (int64_t v = 0; v < n; ++v) ioff[v + 1] += ioff[v];
    std::vector<int32_t> pos(ioff.begin(), ioff.end() - 1);
    icol.assign((size_t)std::max<int64_t>(m, 1), 0);
    for (int64_t u = 0; u < n; ++u)
      for (int32_t j = off[u]; j < off[u + 1]; ++j) icol[pos[colv[j]]++] = (int32_t)u;
    in_off = ioff.data();
    in_col = icol.data();
    g->d_oin_off = dalloc<int32_t>((size_t)n + 1);
    g->d_oin_colo = dalloc<int32_t>((size_t)m);
    h2d(g->d_oin_off, ioff.data(), sizeof(int32_t) * ((size_t)n + 1), g->stream);
    h2d(g->d_oin_colo, icol.data(), sizeof(int32_t) * (size_t)m, g->stream);
  } else {
    g->d_oin_off = g->d_off;
    g->d_oin_colo = g->d_col;
  }
  build_engine_graph(g.get(), off, colv.data(), in_off, in_col, (flags & XB_GRAPH_DIRECTED) != 0);
  init_default_layout(g.get());
  HIPCHK(hipHostMalloc((void**)&g->q0_pinned, sizeof(int32_t) * 3 * 1024, hipHostMallocDefault));
  return g.release();
}

// Reads whitespace-separated int32 values (ReadFileIntoHVector's format,
// apps/xb/src/include/common/utils_gpu.h:9-25), but fails on a missing file,
// a non-integer token or an out-of-range value instead of stopping.  The
// buffer is split at whitespace into one chunk per thread; each chunk is
// parsed with the same rules and the earliest failure in file order wins.
bool is_ws(char c) { return c == ' ' || c == '\n' || c == '\r' || c == '\t'; }

std::vector<int32_t> read_ints(const char* path) {
  FILE* f = std::fopen(path, "rb");
  if (!f) fail(XB_ERR_IO, std::string("cannot open ") + path + ": " + std::strerror(errno));
  std::fseek(f, 0, SEEK_END);
  long sz = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  std::string buf((size_t)std::max(sz, 0L), '\0');
  size_t rd = sz > 0 ? std::fread(&buf[0], 1, (size_t)sz, f) : 0;
  std::fclose(f);
  if ((long)rd != sz) fail(XB_ERR_IO, std::string("short read on ") + path);
  const char* base = buf.data();
  const size_t len = buf.size();
  int T = (len < ((size_t)1 << 22)) ? 1 : resolve_threads();
  std::vector<size_t> cut((size_t)T + 1, len);
  cut[0] = 0;
  for (int t = 1; t < T; ++t) {
    size_t c = std::max(cut[t - 1], len / T * t);
    while (c < len && !is_ws(base[c])) ++c;  // never split a token
    cut[t] = c;
  }
  std::vector<std::vector<int32_t>> part((size_t)T);
  std::vector<std::string> err((size_t)T);
  auto parse = [&](int t) {
    const char* p = base + cut[t];
    const char* end = base + cut[t + 1];
    std::vector<int32_t>& out = part[(size_t)t];
    out.reserve((size_t)(end - p) / 4 + 1);
    while (p < end) {
      while (p < end && is_ws(*p)) ++p;
      if (p >= end) break;
      bool neg = false;
      if (*p == '-' || *p == '+') {
        neg = (*p == '-');
        ++p;
      }
      if (p >= end || *p < '0' || *p > '9') {
        err[(size_t)t] = std::string("non-integer token in ") + path;
        return;
      }
      int64_t v = 0;
      while (p < end && *p >= '0' && *p <= '9') {
        v = v * 10 + (*p - '0');
        if (v > (int64_t)INT32_MAX + 1) {
          err[(size_t)t] = std::string("integer overflow in ") + path;
          return;
        }
        ++p;
      }
      if (p < end && !is_ws(*p)) {
        err[(size_t)t] = std::string("non-integer token in ") + path;
        return;
      }
      v = neg ? -v : v;
      if (v > INT32_MAX) {
        err[(size_t)t] = std::string("integer overflow in ") + path;
        return;
      }
      out.push_back((int32_t)v);
    }
  };
  if (T == 1) {
    parse(0);
  } else {
    std::vector<std::thread> th;
    for (int t = 0; t < T; ++t) th.emplace_back(parse, t);
    for (auto& x : th) x.join();
  }
  for (int t = 0; t < T; ++t)
    if (!err[(size_t)t].empty()) fail(XB_ERR_IO, err[(size_t)t]);
  size_t total = 0;
  for (auto& v : part) total += v.size();
  std::vector<int32_t> out;
  out.reserve(total);
  for (auto& v : part) out.insert(out.end(), v.begin(), v.end());
  return out;
}

// ---------------------------------------------------------------------------
// Binary CSR files
//   char magic[8] "XBCSR01" (NUL-padded); uint32 flags (XB_GRAPH_*); uint32 0;
//   int64 n_plus_1; int64 m; uint64 checksum; int32 row_offsets[n_plus_1];
//   int32 col_indices[m]; all little-endian.  The checksum is a 64-bit
//   FNV-1a over the 32-bit words of row_offsets then col_indices.
// ---------------------------------------------------------------------------
struct CsrFileHeader {
  char magic[8];
  uint32_t flags;
  uint32_t zero;
  int64_t n_plus_1;
  int64_t m;
  uint64_t checksum;
};
static_assert(sizeof(CsrFileHeader) == 40, "CSR file header layout");
const char kCsrMagic[8] = {'X', 'B', 'C', 'S', 'R', '0', '1', 0};

uint64_t csr_words_hash(uint64_t h, const int32_t* w, int64_t count) {
  for (int64_t i = 0; i < count; ++i) {
    h ^= (uint32_t)w[i];
    h *= 0x100000001b3ull;
  }
  return h;
}

uint64_t csr_checksum(const int32_t* off, int64_t n_plus_1, const int32_t* col, int64_t m) {
  return csr_words_hash(csr_words_hash(0xcbf29ce484222325ull, off, n_plus_1), col, m);
}

void write_csr_file(const char* path, const int32_t* off, int64_t n_plus_1, const int32_t* col, int64_t m,
                    uint32_t flags) {
  CsrFileHeader h{};
  std::memcpy(h.magic, kCsrMagic, 8);
  h.flags = flags;
  h.n_plus_1 = n_plus_1;
  h.m = m;
  h.checksum = csr_checksum(off, n_plus_1, col, m);
  FILE* f = std::fopen(path, "wb");
  if (!f) fail(XB_ERR_IO, std::string("cannot create ") + path + ": " + std::strerror(errno));
  bool ok = std::fwrite(&h, sizeof(h), 1, f) == 1;
  ok = ok && std::fwrite(off, sizeof(int32_t), (size_t)n_plus_1, f) == (size_t)n_plus_1;
  if (m > 0) ok = ok && std::fwrite(col, sizeof(int32_t), (size_t)m, f) == (size_t)m;
  ok = (std::fclose(f) == 0) && ok;
  if (!ok) fail(XB_ERR_IO, std::string("write failed on ") + path);
}

CsrFileHeader read_csr_header(FILE* f, const char* path) {
  CsrFileHeader h{};
  if (std::fread(&h, sizeof(h), 1, f) != 1) fail(XB_ERR_IO, std::string("truncated header in ") + path);
  if (std::memcmp(h.magic, kCsrMagic, 8) != 0) fail(XB_ERR_IO, std::string("not a binary CSR file: ") + path);
  if (h.flags & ~XB_GRAPH_DIRECTED) fail(XB_ERR_IO, std::string("unknown graph flags in ") + path);
  if (h.n_plus_1 < 1 || h.n_plus_1 - 1 > INT32_MAX - 1 || h.m < 0 || h.m > INT32_MAX)
    fail(XB_ERR_IO, std::string("n or m out of range in ") + path);
  return h;
}

// Reads the arrays into caller buffers of exactly n_plus_1 and m entries and
// checks the checksum (structure is validated where the graph is built).
void read_csr_file(const char* path, int32_t* off, int64_t n_plus_1, int32_t* col, int64_t m, uint32_t* flags) {
  FILE* f = std::fopen(path, "rb");
  if (!f) fail(XB_ERR_IO, std::string("cannot open ") + path + ": " + std::strerror(errno));
  struct Closer {
    FILE* f;
    ~Closer() { std::fclose(f); }
  } closer{f};
  const CsrFileHeader h = read_csr_header(f, path);
  if (h.n_plus_1 != n_plus_1 || h.m != m) fail(XB_ERR_INVALID_ARG, "buffer sizes do not match the file header");
  if (std::fread(off, sizeof(int32_t), (size_t)n_plus_1, f) != (size_t)n_plus_1 ||
      (m > 0 && std::fread(col, sizeof(int32_t), (size_t)m, f) != (size_t)m))
    fail(XB_ERR_IO, std::string("truncated arrays in ") + path);
  if (csr_checksum(off, n_plus_1, col, m) != h.checksum) fail(XB_ERR_IO, std::string("checksum mismatch in ") + path);
  if (flags) *flags = h.flags;
}

// ---------------------------------------------------------------------------
// Batched-engine options
// ---------------------------------------------------------------------------
struct BPOpts {
  int T;
  int alpha;
  int cut_small;
  int cut_medium;
  int cut_hub;
  int hub_chunk;
};

// ---------------------------------------------------------------------------
// In-degree layout: bucket order for the pull, hub pieces, heavy emit list
// ---------------------------------------------------------------------------
void ensure_layout(xb_graph* g, const BPOpts& o) {
  const auto& io = g->h_ein_off;
  const int32_t n = g->n;
  if (o.cut_small != g->cut_small || o.cut_medium != g->cut_medium || o.cut_hub != g->cut_hub ||
      o.hub_chunk != g->hub_chunk) {
    auto cls_of = [&](int32_t d) { return d <= o.cut_small ? 0 : (d <= o.cut_medium ? 1 : (d <= o.cut_hub ? 2 : 3)); };
    std::vector<int32_t> perm;
    perm.reserve((size_t)n);
    for (int c = 0; c < 3; ++c) {
      int64_t edges = 0;
      for (int32_t v = 0; v < n; ++v) {
        const int32_t d = io[v + 1] - io[v];
        if (d > 0 && cls_of(d) == c) {
          perm.push_back(v);
          edges += d;
        }
      }
      const int32_t first = c == 0 ? 0 : g->bucket_end[c - 1];
      g->bucket_end[c] = (int32_t)perm.size();
      g->bucket_edges[c] = edges;
      bool contig = true;
      for (int32_t i = first + 1; i < g->bucket_end[c] && contig; ++i) contig = perm[(size_t)i] == perm[(size_t)i - 1] + 1;
      g->bucket_contig[c] = contig;
      g->bucket_shift[c] = (first < g->bucket_end[c]) ? perm[(size_t)first] - first : 0;
    }
    std::vector<xb::HubChunk> chunks;
    int32_t hubs = 0;
    int64_t hub_edges = 0;
    for (int32_t v = 0; v < n; ++v) {
      const int32_t d = io[v + 1] - io[v];
      if (d > 0 && cls_of(d) == 3) {
        ++hubs;
        hub_edges += d;
        for (int32_t jb = io[v]; jb < io[v + 1]; jb += o.hub_chunk)
          chunks.push_back(xb::HubChunk{v, jb, std::min(jb + o.hub_chunk, io[v + 1]), 0});
      }
    }
    dfree(g->d_perm);
    dfree(g->d_hub_chunks);
    g->d_perm = nullptr;
    g->d_hub_chunks = nullptr;
    g->d_perm = dalloc<int32_t>(perm.size());
    if (!perm.empty())
      h2d(g->d_perm, perm.data(), sizeof(int32_t) * perm.size(), g->stream);
    g->d_hub_chunks = dalloc<xb::HubChunk>(chunks.size());
    if (!chunks.empty())
      h2d(g->d_hub_chunks, chunks.data(), sizeof(xb::HubChunk) * chunks.size(), g->stream);
    g->n_hub_chunks = (int32_t)chunks.size();
    g->n_hubs = hubs;
    g->hub_edges = hub_edges;
    // Heavy vertices (bucket 2 and hubs): the pull and the push record their
    // parents in the heavy-parent table, row hidx[v].
    std::vector<int32_t> hidx((size_t)n, -1);
    int32_t n_heavy = 0;
    for (int32_t v = 0; v < n; ++v)
      if (io[v + 1] - io[v] > o.cut_medium) hidx[v] = n_heavy++;
    dfree(g->d_hidx);
    g->d_hidx = nullptr;
    g->d_hidx = dalloc<int32_t>((size_t)n);
    h2d(g->d_hidx, hidx.data(), sizeof(int32_t) * (size_t)n, g->stream);
    g->n_heavy = n_heavy;
    g->cut_small = o.cut_small;
    g->cut_medium = o.cut_medium;
    g->cut_hub = o.cut_hub;
    g->hub_chunk = o.hub_chunk;
    g->layout_gen++;
  }
}

// ---------------------------------------------------------------------------
// Workspace
// ---------------------------------------------------------------------------
void ensure_queues(const xb_graph* g, Workspace& w) {
  if (w.ctrs) return;
  const size_t n = (size_t)g->n + 1;
  w.n = g->n;
  w.ctrs = dalloc<unsigned long long>(2);
  for (int i = 0; i < 2; ++i) {
    w.q[i].v = dalloc<int32_t>(n);
    w.q[i].eo = dalloc<int32_t>(n);
    w.q[i].ctr = w.ctrs + i;
  }
  w.tcount = dalloc<uint32_t>(1);
}

// Bit-parallel state for tiles of up to T words per vertex (a tile with fewer
// words uses the leading n x words entries of each array).
void ensure_bp(const xb_graph* g, Workspace& w, int T) {
  if (w.T >= T) return;
  w.free_bp();
  const size_t nw = (size_t)g->n * (size_t)T + 1;
  w.bp.seen = dalloc<uint64_t>(nw);
  w.bp.front[0] = dalloc<uint64_t>(nw);
  w.bp.front[1] = dalloc<uint64_t>(nw);
  w.bp.nextw = dalloc<uint64_t>(nw);
  w.bp.planes = dalloc<uint64_t>(nw * 4);
  w.bp.planes_hi = nullptr;  // allocated when a tile first reaches level 15
  w.bp.tflag = dalloc<uint32_t>((size_t)g->n + 1);
  w.bp.tlist = dalloc<int32_t>((size_t)g->n + 1);
  w.bp.tcount = w.tcount;
  w.bp.n = g->n;
  w.T = T;
  w.bp_dirty = true;
}

void ensure_src(Workspace& w, int64_t cnt) {
  if (w.src_cap >= cnt) return;
  dfree(w.d_src);
  w.d_src = nullptr;
  w.d_src = dalloc<int32_t>((size_t)cnt);
  w.src_cap = cnt;
}

// Heavy-parent table: n_heavy x T x 64 entries.
void ensure_hpar(const xb_graph* g, Workspace& w, int T) {
  const int64_t need = (int64_t)std::max(g->n_heavy, 1) * T * 64;
  if (w.side_cap < need) {
    dfree(w.side);
    w.side = nullptr;
    w.side = dalloc<int32_t>((size_t)need);
    w.side_cap = need;
  }
  w.bp.hpar = (uint32_t*)w.side;
  w.bp.hidx = g->d_hidx;
  w.bp.new_of = g->d_new_of;
  w.bp.orig_of = g->d_orig_of;
  w.bp.in_col_o = g->d_ein_col_o;
}

int32_t* ensure_stage(Workspace& w, int64_t ints) {
  if (w.stage_cap < ints) {
    dfree(w.stage);
    w.stage = nullptr;
    w.stage = dalloc<int32_t>((size_t)ints);
    w.stage_cap = ints;
  }
  return w.stage;
}

int32_t* ensure_q0_host(Workspace& w, int64_t ints) {
  if (w.q0_cap < ints) {
    if (w.q0_host) (void)hipHostFree(w.q0_host);
    w.q0_host = nullptr;
    HIPCHK(hipHostMalloc((void**)&w.q0_host, sizeof(int32_t) * (size_t)ints, hipHostMallocDefault));
    w.q0_cap = ints;
  }
  return w.q0_host;
}

// ---------------------------------------------------------------------------
// Per-class event timing and algorithmic-byte accounting
// ---------------------------------------------------------------------------
struct Profiler {
  bool on = false;
  hipStream_t st = nullptr;
  struct Mark {
    int cat;
    hipEvent_t a, b;
  };
  std::vector<Mark> marks;
  std::vector<hipEvent_t> pool;
  hipEvent_t open_ev = nullptr;
  int open_cat = -1;
  double bytes[XB_K_CLASSES] = {};
  int64_t launches[XB_K_CLASSES] = {};

  hipEvent_t get() {
    if (!pool.empty()) {
      hipEvent_t e = pool.back();
      pool.pop_back();
      return e;
    }
    hipEvent_t e;
    HIPCHK(hipEventCreate(&e));
    return e;
  }
  void begin(int cat) {
    launches[cat]++;
    if (!on) return;
    open_ev = get();
    open_cat = cat;
    HIPCHK(hipEventRecord(open_ev, st));
  }
  void end() {
    if (!on || open_cat < 0) return;
    hipEvent_t e = get();
    HIPCHK(hipEventRecord(e, st));
    marks.push_back(Mark{open_cat, open_ev, e});
    open_cat = -1;
  }
  void add(int cat, double b) { bytes[cat] += b; }
  // Takes over another profiler's marks and counts (the second workspace's
  // stream of a pipelined run), so one collect() reports both.
  void absorb(Profiler& o) {
    marks.insert(marks.end(), o.marks.begin(), o.marks.end());
    o.marks.clear();
    for (int c = 0; c < XB_K_CLASSES; ++c) {
      bytes[c] += o.bytes[c];
      launches[c] += o.launches[c];
      o.bytes[c] = 0;
      o.launches[c] = 0;
    }
  }
  void collect(xb_run_stats* s) {
    for (auto& mk : marks) {
      float t = 0;
      HIPCHK(hipEventSynchronize(mk.b));
      HIPCHK(hipEventElapsedTime(&t, mk.a, mk.b));
      if (s) s->kernel_ms[mk.cat] += t;
      pool.push_back(mk.a);
      pool.push_back(mk.b);
    }
    marks.clear();
    if (s)
      for (int c = 0; c < XB_K_CLASSES; ++c) {
        s->kernel_bytes[c] = bytes[c];
        s->kernel_launches[c] = launches[c];
      }
  }
  ~Profiler() {
    for (auto& mk : marks) {
      (void)hipEventDestroy(mk.a);
      (void)hipEventDestroy(mk.b);
    }
    for (auto e : pool) (void)hipEventDestroy(e);
  }
};

// Reads the packed (edges << 32 | count) counter of a queue.
void read_ctr(Workspace& w, const unsigned long long* dctr, int32_t* qn, int32_t* E) {
  HIPCHK(hipMemcpyAsync(w.hctr, dctr, sizeof(unsigned long long), hipMemcpyDeviceToHost, w.st));
  HIPCHK(hipStreamSynchronize(w.st));
  const unsigned long long v = *w.hctr;
  *qn = (int32_t)(uint32_t)(v & 0xFFFFFFFFull);
  *E = (int32_t)(uint32_t)(v >> 32);
}

// ---------------------------------------------------------------------------
// Plain per-source engine
// ---------------------------------------------------------------------------
void run_plain(const xb_graph* g, Workspace& w, const int32_t* sources, int32_t S, int32_t* lvl_rows,
               int32_t* par_rows, bool dev_out, Profiler& prof, xb_run_stats* st) {
  ensure_queues(g, w);
  const int32_t n = g->n;
  int32_t* stage = nullptr;
  if (!dev_out || !lvl_rows) stage = ensure_stage(w, 2 * (int64_t)n);
  for (int32_t s = 0; s < S; ++s) {
    const int32_t src = sources[s];
    int32_t* lv = (dev_out && lvl_rows) ? lvl_rows + (int64_t)s * n : stage;
    int32_t* pa = dev_out ? par_rows + (int64_t)s * n : stage + n;
    prof.begin(XB_K_PLAIN_INIT);
    xb::launch_plain_init(w.st, lv, pa, n);
    xb::launch_plain_set_source(w.st, lv, pa, src, w.q[0]);
    prof.end();
    prof.add(XB_K_PLAIN_INIT, 8.0 * n);
    int32_t qn = 1, E = g->outdeg(src);
    int cur = 0, d = 0;
    while (qn > 0) {
      HIPCHK(hipMemsetAsync(w.q[cur ^ 1].ctr, 0, sizeof(unsigned long long), w.st));
      prof.begin(XB_K_PLAIN_PUSH);
      xb::launch_plain_push(w.st, g->d_off, g->d_col, w.q[cur], qn, E, lv, pa, d, w.q[cur ^ 1]);
      prof.end();
      const int32_t E_prev = E;
      read_ctr(w, w.q[cur ^ 1].ctr, &qn, &E);
      // col + level gather per frontier edge; level + parent + queue write per new vertex.
      prof.add(XB_K_PLAIN_PUSH, 8.0 * E_prev + 16.0 * qn);
      cur ^= 1;
      if (qn > 0) ++d;
    }
    if (st) {
      st->max_level = std::max(st->max_level, d);
      st->push_levels += d + 1;
    }
    if (!dev_out) {
      if (lvl_rows)
        HIPCHK(hipMemcpyAsync(lvl_rows + (int64_t)s * n, stage, sizeof(int32_t) * (size_t)n, hipMemcpyDeviceToHost,
                              w.st));
      HIPCHK(hipMemcpyAsync(par_rows + (int64_t)s * n, stage + n, sizeof(int32_t) * (size_t)n,
                            hipMemcpyDeviceToHost, w.st));
      HIPCHK(hipStreamSynchronize(w.st));
    }
  }
}

// ---------------------------------------------------------------------------
// Batched bit-parallel engine
// ---------------------------------------------------------------------------

// Returns false when the tile got deeper than the 8-bit level code allows.
bool run_bp_tile(const xb_graph* g, Workspace& w, const BPOpts& o, const int32_t* src, int32_t cnt,
                 int32_t* lvl_rows, int32_t* par_rows, const xb::RowMap& rows, Profiler& prof, xb_run_stats* st) {
  const int T = o.T;
  const int32_t n = g->n;
  const size_t nw = (size_t)n * (size_t)T;
  xb::BPState& s = w.bp;
  prof.begin(XB_K_MEMSET);
  if (w.bp_dirty) {
    // Whole capacity: later tiles may use more words than this one.
    const size_t cap = (size_t)n * (size_t)w.T;
    HIPCHK(hipMemsetAsync(s.front[0], 0, sizeof(uint64_t) * cap, w.st));
    HIPCHK(hipMemsetAsync(s.front[1], 0, sizeof(uint64_t) * cap, w.st));
    HIPCHK(hipMemsetAsync(s.nextw, 0, sizeof(uint64_t) * cap, w.st));
    HIPCHK(hipMemsetAsync(s.tflag, 0, sizeof(uint32_t) * ((size_t)n + 1), w.st));
    prof.add(XB_K_MEMSET, 24.0 * cap + 4.0 * n);
    w.bp_dirty = false;
    w.front_dirty_words = 0;
  }
  if (w.front_dirty_words > 0) {
    HIPCHK(hipMemsetAsync(s.front[0], 0, sizeof(uint64_t) * w.front_dirty_words, w.st));
    HIPCHK(hipMemsetAsync(s.front[1], 0, sizeof(uint64_t) * w.front_dirty_words, w.st));
    prof.add(XB_K_MEMSET, 16.0 * w.front_dirty_words);
    w.front_dirty_words = 0;
  }
  // Planes 0-3 (codes below 16) are cleared here; planes 4-7 only before the
  // first level whose code needs them (level 15), which most tiles never reach.
  const size_t half = (size_t)T * (size_t)n * 4;
  HIPCHK(hipMemsetAsync(s.seen, 0, sizeof(uint64_t) * nw, w.st));
  HIPCHK(hipMemsetAsync(s.planes, 0, sizeof(uint64_t) * half, w.st));
  if (g->n_heavy > 0)
    HIPCHK(hipMemsetAsync(s.hpar, 0xFF, sizeof(uint32_t) * (size_t)g->n_heavy * T * 64, w.st));
  prof.end();
  prof.add(XB_K_MEMSET, 8.0 * nw + 8.0 * half + 256.0 * T * g->n_heavy);
  bool hi_planes_clear = false;
  // First frontier: the distinct source vertices with their degree prefix.
  std::vector<int32_t> uniq(src, src + cnt);
  std::sort(uniq.begin(), uniq.end());
  uniq.erase(std::unique(uniq.begin(), uniq.end()), uniq.end());
  const int32_t qn0 = (int32_t)uniq.size();
  HIPCHK(hipStreamSynchronize(w.st));  // earlier uploads from the pinned buffer are done
  int32_t* q0 = ensure_q0_host(w, 2 * (int64_t)qn0 + cnt);
  int64_t E0 = 0;
  for (int32_t i = 0; i < qn0; ++i) {
    q0[i] = g->h_new_of[(size_t)uniq[i]];  // the frontier holds engine ids
    q0[qn0 + i] = (int32_t)E0;
    E0 += g->outdeg(uniq[i]);
  }
  std::memcpy(q0 + 2 * qn0, src, sizeof(int32_t) * (size_t)cnt);
  HIPCHK(hipMemcpyAsync(w.q[0].v, q0, sizeof(int32_t) * (size_t)qn0, hipMemcpyHostToDevice, w.st));
  HIPCHK(hipMemcpyAsync(w.q[0].eo, q0 + qn0, sizeof(int32_t) * (size_t)qn0, hipMemcpyHostToDevice, w.st));
  HIPCHK(hipMemcpyAsync(w.d_src, q0 + 2 * qn0, sizeof(int32_t) * (size_t)cnt, hipMemcpyHostToDevice, w.st));
  prof.begin(XB_K_SEED);
  xb::launch_bp_seed(w.st, T, s, w.d_src, cnt);
  prof.end();
  prof.add(XB_K_SEED, 28.0 * cnt);

  int32_t qn = qn0, E = (int32_t)E0;
  int cur = 0, d = 0;
  int pushes = 0, pulls = 0;
  const int64_t thresh = g->m / std::max(o.alpha, 1);
  bool overflow = false;
  while (qn > 0) {
    const int32_t code = d + 2;  // level d+1
    if (code > 255) {
      overflow = true;
      break;
    }
    if (code >= 16 && !hi_planes_clear) {
      if (!s.planes_hi) s.planes_hi = dalloc<uint64_t>((size_t)g->n * (size_t)w.T * 4 + 4);
      prof.begin(XB_K_MEMSET);
      HIPCHK(hipMemsetAsync(s.planes_hi, 0, sizeof(uint64_t) * half, w.st));
      prof.end();
      prof.add(XB_K_MEMSET, 8.0 * half);
      hi_planes_clear = true;
    }
    xb::Queue& qc = w.q[cur];
    xb::Queue& qx = w.q[cur ^ 1];
    HIPCHK(hipMemsetAsync(qx.ctr, 0, sizeof(unsigned long long), w.st));
    HIPCHK(hipMemsetAsync(w.tcount, 0, sizeof(uint32_t), w.st));
    if ((int64_t)qn + E > thresh && XB_PULL_FUSED) {
      xb::PullPlan plan;
      int32_t begin = 0;
      double bytes = 0;
      for (int c = 0; c < 3; ++c) {
        const int32_t end = g->bucket_end[c];
        plan.perm[c] = g->bucket_contig[c] ? nullptr : g->d_perm;
        plan.shift[c] = g->bucket_shift[c];
        plan.begin[c] = begin;
        plan.end[c] = end;
        if (end > begin) bytes += 4.0 * (end - begin + 1) + 4.0 * g->bucket_edges[c] + 24.0 * T * (end - begin);
        begin = end;
      }
      prof.begin(XB_K_PULL);
      xb::launch_bp_pull_fused(w.st, T, s, cur, code, cnt, g->d_ein_off, g->d_ein_col, g->d_eoff, plan, qx,
                               g->d_hub_chunks, g->n_hub_chunks);
      prof.end();
      // offsets + in-edges + (front, seen r/w) per vertex and word, buckets and hub pieces
      prof.add(XB_K_PULL, bytes + (g->n_hub_chunks > 0 ? 4.0 * g->hub_edges + 24.0 * T * g->n_hubs : 0.0));
      if (g->n_hub_chunks > 0) {
        prof.begin(XB_K_COMMIT);
        xb::launch_bp_commit(w.st, T, s, cur, code, g->d_eoff, qx, g->n_hubs);
        prof.end();
      }
      ++pulls;
    } else if ((int64_t)qn + E > thresh) {
      int32_t begin = 0;
      for (int c = 0; c < 3; ++c) {
        const int32_t end = g->bucket_end[c];
        if (end > begin) {
          prof.begin(XB_K_PULL);
          xb::launch_bp_pull_bucket(w.st, T, c, s, cur, code, cnt, g->d_ein_off, g->d_ein_col, g->d_eoff,
                                    g->bucket_contig[c] ? nullptr : g->d_perm, g->bucket_shift[c], begin, end, qx);
          prof.end();
          // offsets + in-edges + (front, seen r/w) per vertex and word
          prof.add(XB_K_PULL, 4.0 * (end - begin + 1) + 4.0 * g->bucket_edges[c] + 24.0 * T * (end - begin));
        }
        begin = end;
      }
      if (g->n_hub_chunks > 0) {
        prof.begin(XB_K_PULL_HUB);
        xb::launch_bp_pull_hub(w.st, T, s, cur, cnt, g->d_ein_col, g->d_hub_chunks, g->n_hub_chunks);
        prof.end();
        prof.add(XB_K_PULL_HUB, 4.0 * g->hub_edges + 24.0 * T * g->n_hubs);
        prof.begin(XB_K_COMMIT);
        xb::launch_bp_commit(w.st, T, s, cur, code, g->d_eoff, qx, g->n_hubs);
        prof.end();
      }
      ++pulls;
    } else {
      prof.begin(XB_K_PUSH);
      xb::launch_bp_push(w.st, T, s, cur, g->d_eoff, g->d_ecol, qc, qn, E);
      prof.end();
      // queue + frontier words of the queue, then col + seen words per edge
      prof.add(XB_K_PUSH, (8.0 + 8.0 * T) * qn + (4.0 + 8.0 * T) * E);
      prof.begin(XB_K_COMMIT);
      xb::launch_bp_commit(w.st, T, s, cur, code, g->d_eoff, qx, (int32_t)std::min<int64_t>(E, n));
      prof.end();
      ++pushes;
    }
    if (!XB_LAZY_FRONT_CLEAR) {
      prof.begin(XB_K_CLEAR);
      xb::launch_bp_clear(w.st, T, s, cur, qc, qn);
      prof.end();
      prof.add(XB_K_CLEAR, (4.0 + 8.0 * T) * qn);
    }
    read_ctr(w, qx.ctr, &qn, &E);
    cur ^= 1;
    if (qn > 0) ++d;
  }
  if (XB_LAZY_FRONT_CLEAR) w.front_dirty_words = std::max(w.front_dirty_words, nw);
  if (overflow) {
    w.bp_dirty = true;
    return false;
  }
  // Emit: K planes cover codes up to d+1.
  int K = 0;
  while ((1 << K) <= d + 1) ++K;
  const int words = (cnt + 63) / 64;
  // Heavy vertices need no pass of their own: the pull and the push recorded
  // the parent of every bit they discovered in hpar, and the emit copies it.
  // Light vertices, word groups of emit_pass_words(T) words of 64 rows each.
  const int W = xb::emit_pass_words(T);
#if XB_EMIT_ONE_LAUNCH
  {
    // One launch: each workgroup runs every word group of its chunk in turn.
    const int ngrp = (words + W - 1) / W;
    prof.begin(XB_K_EMIT);
    xb::launch_bp_emit(w.st, T, s, 0, ngrp, K, cnt, w.d_src, g->d_oin_off, g->d_oin_col, g->d_oin_colo, n,
                       o.cut_medium, lvl_rows, par_rows, rows, (g->flags & XB_GRAPH_DIRECTED) == 0);
    prof.end();
    prof.add(XB_K_EMIT, (lvl_rows ? 8.0 : 4.0) * cnt * (double)n);
  }
#else
  for (int grp = 0; grp * W < words; ++grp) {
    prof.begin(XB_K_EMIT);
    xb::launch_bp_emit(w.st, T, s, grp, 1, K, cnt, w.d_src, g->d_oin_off, g->d_oin_col, g->d_oin_colo, n,
                       o.cut_medium, lvl_rows, par_rows, rows, (g->flags & XB_GRAPH_DIRECTED) == 0);
    prof.end();
    const int rows_g = std::min(64 * W, cnt - 64 * W * grp);
    // level + parent int32 per (source, vertex) of this group's rows
    prof.add(XB_K_EMIT, (lvl_rows ? 8.0 : 4.0) * rows_g * (double)n);
  }
#endif
  if (st) {
    st->max_level = std::max(st->max_level, d);
    st->push_levels += pushes;
    st->pull_levels += pulls;
    st->tiles += 1;
  }
  return true;
}

// Words per vertex for a tile of cnt sources: the smallest power of two with
// 64 * words >= cnt, at most Tmax.  State traffic scales with the words, so a
// short tile does not pay for the empty ones.
// Widest tile: 16 words = 1,024 sources (words_per_tile 1, 2, 4, 8 or 16).
constexpr int kMaxTileWords = 16;

int tile_words(int32_t cnt, int Tmax) {
  int t = 1;
  while (t * 64 < cnt && t < Tmax) t *= 2;
  return t;
}

// Whether a second workspace for tiles of T words fits: it is already there,
// or twice its size (tile state 64 B per vertex and word, queues, heavy
// parents) is free.  A run that does not fit keeps to one stream.
bool ws2_fits(const xb_graph* g, int T) {
  if (g->stream2 && g->ws2.T >= T) return true;
  const double need = 64.0 * T * g->n + 24.0 * g->n + 256.0 * T * std::max(g->n_heavy, 1);
  size_t fr = 0, tot = 0;
  HIPCHK(hipMemGetInfo(&fr, &tot));
  return 2.0 * need <= (double)fr;
}

// The second workspace and its stream (see xb_graph::ws2).
void ensure_ws2(xb_graph* g) {
  if (g->stream2) return;
  HIPCHK(hipStreamCreateWithFlags(&g->stream2, hipStreamNonBlocking));
  g->pinned2.reset(new Pinned());
  HIPCHK(hipEventCreateWithFlags(&g->ev_fork, hipEventDisableTiming));
  HIPCHK(hipEventCreateWithFlags(&g->ev_join, hipEventDisableTiming));
  g->ws2.st = g->stream2;
  g->ws2.hctr = g->pinned2->p;
}

// Tiles of up to 64*T sources, one after another.  With a second workspace
// (w2, prof2 on its stream) and device outputs, tile k runs on workspace
// k % 2: a tile's level loop, which waits on the host for every level's
// frontier count, then overlaps the emit its predecessor left queued on the
// other stream, and the gather-bound traversal shares the device with the
// write-bound emit.  The second stream starts after the caller's earlier work
// on the first, and the first waits for the second at the end, so the run is
// ordered on w.st like a one-stream run.
void run_batched(xb_graph* g, Workspace& w, const BPOpts& o, const int32_t* sources, int32_t S, int32_t* lvl,
                 int32_t* par, bool dev_out, Profiler& prof, xb_run_stats* st, Workspace* w2 = nullptr,
                 Profiler* prof2 = nullptr) {
  const int32_t tile = 64 * o.T;
  const int Tneed = tile_words(std::min(S, tile), o.T);  // the widest tile of this run
  const bool pipe = w2 != nullptr && prof2 != nullptr && dev_out && S > tile;
  Workspace* ws[2] = {&w, pipe ? w2 : &w};
  Profiler* pr[2] = {&prof, pipe ? prof2 : &prof};
  ensure_layout(g, o);
  for (int i = 0; i < (pipe ? 2 : 1); ++i) {
    ensure_queues(g, *ws[i]);
    ensure_bp(g, *ws[i], Tneed);
    ensure_hpar(g, *ws[i], Tneed);
    ensure_src(*ws[i], tile);
  }
  const int32_t n = g->n;
  int32_t* stage = nullptr;
  if (!dev_out) stage = ensure_stage(w, 2 * (int64_t)tile * n);
  if (pipe) {
    HIPCHK(hipEventRecord(g->ev_fork, w.st));
    HIPCHK(hipStreamWaitEvent(w2->st, g->ev_fork, 0));
  }
  for (int32_t base = 0, k = 0; base < S; base += tile, ++k) {
    Workspace& wk = *ws[k & 1];
    Profiler& pk = *pr[k & 1];
    const int32_t cnt = std::min(tile, S - base);
    int32_t* lr = dev_out ? (lvl ? lvl + (int64_t)base * n : nullptr) : (lvl ? stage : nullptr);
    int32_t* pr_rows = dev_out ? par + (int64_t)base * n : stage + (int64_t)tile * n;
    BPOpts ot = o;
    ot.T = tile_words(cnt, o.T);
    bool ok = run_bp_tile(g, wk, ot, sources + base, cnt, lr, pr_rows, xb::RowMap{}, pk, st);
    if (!ok) {
      // Deeper than 254 levels: these sources take the plain path.
      if (st) st->fallback_sources += cnt;
      run_plain(g, wk, sources + base, cnt, lvl ? lvl + (int64_t)base * n : nullptr, par + (int64_t)base * n,
                dev_out, pk, st);
      continue;
    }
    if (!dev_out) {
      if (lvl)
        HIPCHK(hipMemcpyAsync(lvl + (int64_t)base * n, stage, sizeof(int32_t) * (size_t)cnt * n,
                              hipMemcpyDeviceToHost, w.st));
      HIPCHK(hipMemcpyAsync(par + (int64_t)base * n, stage + (int64_t)tile * n, sizeof(int32_t) * (size_t)cnt * n,
                            hipMemcpyDeviceToHost, w.st));
      HIPCHK(hipStreamSynchronize(w.st));
    }
  }
  if (pipe) {
    HIPCHK(hipEventRecord(g->ev_join, w2->st));
    HIPCHK(hipStreamWaitEvent(w.st, g->ev_join, 0));
  }
}

xb_msbfs_opts parse_opts(const xb_msbfs_opts* opts) {
  xb_msbfs_opts o;
  xb_opts_default(&o);
  if (opts) {
    o.mode = opts->mode;
    o.flags = opts->flags;
    if (opts->words_per_tile) o.words_per_tile = opts->words_per_tile;
    if (opts->alpha) o.alpha = opts->alpha;
    if (opts->pull_cut_small) o.pull_cut_small = opts->pull_cut_small;
    if (opts->pull_cut_medium) o.pull_cut_medium = opts->pull_cut_medium;
    if (opts->pull_cut_hub) o.pull_cut_hub = opts->pull_cut_hub;
    if (opts->hub_chunk) o.hub_chunk = opts->hub_chunk;
  }
  if (o.mode < XB_MODE_AUTO || o.mode > XB_MODE_BATCHED) fail(XB_ERR_INVALID_ARG, "unknown mode");
  const int T = o.words_per_tile;
  if (T < 1 || T > kMaxTileWords || (T & (T - 1)) != 0)
    fail(XB_ERR_INVALID_ARG, "words_per_tile must be 1, 2, 4, 8 or 16");
  if (o.alpha < 1 || o.pull_cut_small < 1 || o.pull_cut_medium < o.pull_cut_small ||
      o.pull_cut_hub < o.pull_cut_medium || o.hub_chunk < 1)
    fail(XB_ERR_INVALID_ARG, "bad tunable");
  return o;
}

BPOpts bp_opts(const xb_msbfs_opts& o) {
  return BPOpts{o.words_per_tile, o.alpha, o.pull_cut_small, o.pull_cut_medium, o.pull_cut_hub, o.hub_chunk};
}

void init_default_layout(xb_graph* g) {
  xb_msbfs_opts o;
  xb_opts_default(&o);
  ensure_layout(g, bp_opts(o));
}

void check_sources(const xb_graph* g, const int32_t* sources, int32_t S) {
  for (int32_t i = 0; i < S; ++i)
    if (sources[i] < 0 || sources[i] >= g->n)
      fail(XB_ERR_SOURCE_RANGE, "source " + std::to_string(i) + " = " + std::to_string(sources[i]) +
                                    " outside [0, " + std::to_string(g->n) + ")");
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

int32_t xb_opts_default(xb_msbfs_opts* o) {
  if (!o) {
    g_last_error = "null opts";
    return XB_ERR_INVALID_ARG;
  }
  std::memset(o, 0, sizeof(*o));
  o->mode = XB_MODE_AUTO;
  o->words_per_tile = 8;
  o->alpha = 20;
  o->pull_cut_small = 8;
  o->pull_cut_medium = 64;
  o->pull_cut_hub = 1024;
  o->hub_chunk = 1024;
  return XB_OK;
}

int32_t xb_graph_create_csr(const int32_t* row_offsets, int64_t n_plus_1, const int32_t* col_indices, int64_t m,
                            uint32_t flags, xb_graph** out) {
  return guarded([&] {
    if (!out) fail(XB_ERR_INVALID_ARG, "null out handle");
    *out = build_graph(row_offsets, n_plus_1, col_indices, m, flags);
  });
}

int32_t xb_graph_load_text(const char* row_path, const char* col_path, uint32_t flags, xb_graph** out) {
  return guarded([&] {
    if (!out || !row_path || !col_path) fail(XB_ERR_INVALID_ARG, "null argument");
    std::vector<int32_t> off = read_ints(row_path);
    std::vector<int32_t> col = read_ints(col_path);
    if (off.empty()) fail(XB_ERR_IO, std::string("no row offsets in ") + row_path);
    *out = build_graph(off.data(), (int64_t)off.size(), col.data(), (int64_t)col.size(), flags);
  });
}

int32_t xb_csr_save(const char* path, const int32_t* row_offsets, int64_t n_plus_1, const int32_t* col_indices,
                    int64_t m, uint32_t flags) {
  return guarded([&] {
    if (!path || !row_offsets || n_plus_1 < 1 || m < 0 || (m > 0 && !col_indices))
      fail(XB_ERR_INVALID_ARG, "null or empty CSR arrays");
    if (flags & ~XB_GRAPH_DIRECTED) fail(XB_ERR_INVALID_ARG, "unknown graph flag");
    write_csr_file(path, row_offsets, n_plus_1, col_indices, m, flags);
  });
}

int32_t xb_csr_load_header(const char* path, int64_t* n_plus_1, int64_t* m, uint32_t* flags) {
  return guarded([&] {
    if (!path) fail(XB_ERR_INVALID_ARG, "null path");
    FILE* f = std::fopen(path, "rb");
    if (!f) fail(XB_ERR_IO, std::string("cannot open ") + path + ": " + std::strerror(errno));
    struct Closer {
      FILE* f;
      ~Closer() { std::fclose(f); }
    } closer{f};
    const CsrFileHeader h = read_csr_header(f, path);
    if (n_plus_1) *n_plus_1 = h.n_plus_1;
    if (m) *m = h.m;
    if (flags) *flags = h.flags;
  });
}

int32_t xb_csr_load(const char* path, int32_t* row_offsets, int64_t n_plus_1, int32_t* col_indices, int64_t m) {
  return guarded([&] {
    if (!path || !row_offsets || (m > 0 && !col_indices)) fail(XB_ERR_INVALID_ARG, "null argument");
    read_csr_file(path, row_offsets, n_plus_1, col_indices, m, nullptr);
  });
}

int32_t xb_graph_load_binary(const char* path, xb_graph** out) {
  return guarded([&] {
    if (!path || !out) fail(XB_ERR_INVALID_ARG, "null argument");
    int64_t np1 = 0, m = 0;
    {
      FILE* f = std::fopen(path, "rb");
      if (!f) fail(XB_ERR_IO, std::string("cannot open ") + path + ": " + std::strerror(errno));
      struct Closer {
        FILE* f;
        ~Closer() { std::fclose(f); }
      } closer{f};
      const CsrFileHeader h = read_csr_header(f, path);
      np1 = h.n_plus_1;
      m = h.m;
    }
    std::vector<int32_t> off((size_t)np1), col((size_t)std::max<int64_t>(m, 1));
    uint32_t flags = 0;
    read_csr_file(path, off.data(), np1, col.data(), m, &flags);
    *out = build_graph(off.data(), np1, col.data(), m, flags);
  });
}

int32_t xb_graph_save_binary(const xb_graph* g, const char* path) {
  return guarded([&] {
    if (!g || !path) fail(XB_ERR_INVALID_ARG, "null argument");
    std::vector<int32_t> col((size_t)std::max<int64_t>(g->m, 1));
    d2h(col.data(), g->d_col, sizeof(int32_t) * (size_t)g->m, g->stream);
    write_csr_file(path, g->h_off.data(), (int64_t)g->h_off.size(), col.data(), g->m, g->flags);
  });
}

int32_t xb_graph_info(const xb_graph* g, int64_t* n, int64_t* m, uint32_t* flags) {
  return guarded([&] {
    if (!g) fail(XB_ERR_INVALID_ARG, "null graph");
    if (n) *n = g->n;
    if (m) *m = g->m;
    if (flags) *flags = g->flags;
  });
}

int32_t xb_graph_destroy(xb_graph* g) {
  return guarded([&] {
    if (!g) return;
    if (g->stream) (void)hipStreamSynchronize(g->stream);
    delete g;
  });
}

int32_t xb_msbfs_run(xb_graph* g, const int32_t* sources, int32_t S, int32_t* levels, int32_t* parents,
                     const xb_msbfs_opts* opts, xb_run_stats* stats) {
  return guarded([&] {
    if (!g) fail(XB_ERR_INVALID_ARG, "null graph");
    if (S < 0) fail(XB_ERR_INVALID_ARG, "S < 0");
    if (stats) std::memset(stats, 0, sizeof(*stats));
    if (S == 0) return;
    if (!sources || !parents) fail(XB_ERR_INVALID_ARG, "null sources or parents");
    const xb_msbfs_opts o = parse_opts(opts);
    if (!levels && !(o.flags & XB_NO_LEVELS)) fail(XB_ERR_INVALID_ARG, "null levels (set XB_NO_LEVELS)");
    check_sources(g, sources, S);
    const bool dev_out = (o.flags & XB_OUT_DEVICE) != 0;
    Profiler prof;
    prof.on = (o.flags & XB_PROFILE) != 0;
    prof.st = g->stream;
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    HIPCHK(hipEventRecord(e0, g->stream));
    if (o.mode == XB_MODE_PLAIN) {
      run_plain(g, g->ws, sources, S, levels, parents, dev_out, prof, stats);
    } else {
      const BPOpts bo = bp_opts(o);
      Profiler prof2;
      prof2.on = prof.on;
      Workspace* w2 = nullptr;
      if (!(o.flags & XB_NO_PIPELINE) && dev_out && S > 64 * bo.T && ws2_fits(g, bo.T)) {
        ensure_ws2(g);
        w2 = &g->ws2;
        prof2.st = g->stream2;
      }
      run_batched(g, g->ws, bo, sources, S, levels, parents, dev_out, prof, stats, w2, &prof2);
      prof.absorb(prof2);
    }
    HIPCHK(hipEventRecord(e1, g->stream));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (stats) {
      prof.collect(stats);
      stats->ms_total = ms;
    }
  });
}

int32_t xb_count_traversed(xb_graph* g, const int32_t* levels_dev, int32_t S, int64_t* edges_out) {
  return guarded([&] {
    if (!g || !levels_dev || !edges_out || S < 0) fail(XB_ERR_INVALID_ARG, "bad argument");
    unsigned long long* d = dalloc<unsigned long long>(1);
    HIPCHK(hipMemsetAsync(d, 0, sizeof(unsigned long long), g->stream));
    if (S > 0 && g->n > 0) xb::launch_count_traversed(g->stream, g->d_off, levels_dev, S, g->n, d);
    unsigned long long h = 0;
    HIPCHK(hipMemcpyAsync(&h, d, sizeof(h), hipMemcpyDeviceToHost, g->stream));
    HIPCHK(hipStreamSynchronize(g->stream));
    dfree(d);
    *edges_out = (int64_t)h;
  });
}

int32_t xb_test_edge_map(const int32_t* degrees, int32_t k, int32_t* owner_out, int32_t* offset_out,
                         int64_t* total_out) {
  return guarded([&] {
    if (!degrees || k < 0 || !total_out) fail(XB_ERR_INVALID_ARG, "bad argument");
    int64_t total = 0;
    for (int32_t i = 0; i < k; ++i) {
      if (degrees[i] < 0) fail(XB_ERR_INVALID_ARG, "negative degree");
      total += degrees[i];
    }
    if (total > INT32_MAX) fail(XB_ERR_INVALID_ARG, "total tasklets exceed int32");
    *total_out = total;
    if (k == 0 || total == 0) return;
    if (!owner_out || !offset_out) fail(XB_ERR_INVALID_ARG, "null outputs");
    ensure_device();
    hipStream_t st;
    HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    int32_t* d_deg = dalloc<int32_t>((size_t)k);
    int32_t* d_v = dalloc<int32_t>((size_t)k);
    int32_t* d_eo = dalloc<int32_t>((size_t)k);
    unsigned long long* d_ctr = dalloc<unsigned long long>(1);
    int32_t* d_owner = dalloc<int32_t>((size_t)total);
    int32_t* d_off = dalloc<int32_t>((size_t)total);
    h2d(d_deg, degrees, sizeof(int32_t) * (size_t)k, st);
    HIPCHK(hipMemsetAsync(d_ctr, 0, sizeof(unsigned long long), st));
    xb::Queue q{d_v, d_eo, d_ctr};
    xb::launch_test_build_queue(st, d_deg, k, q);
    xb::launch_test_edge_map(st, q, k, (int32_t)total, d_owner, d_off);
    HIPCHK(hipStreamSynchronize(st));
    unsigned long long ctr = 0;
    d2h(&ctr, d_ctr, sizeof(ctr), st);
    d2h(owner_out, d_owner, sizeof(int32_t) * (size_t)total, st);
    d2h(offset_out, d_off, sizeof(int32_t) * (size_t)total, st);
    for (void* p : {(void*)d_deg, (void*)d_v, (void*)d_eo, (void*)d_ctr, (void*)d_owner, (void*)d_off}) dfree(p);
    (void)hipStreamDestroy(st);
    if ((int64_t)(ctr & 0xFFFFFFFFull) != k || (int64_t)(ctr >> 32) != total)
      fail(XB_ERR_DEVICE, "queue counter does not match the input");
  });
}

int32_t xb_device_checksum(const int32_t* a_dev, int64_t count, uint64_t* out) {
  return guarded([&] {
    if ((!a_dev && count > 0) || count < 0 || !out) fail(XB_ERR_INVALID_ARG, "bad argument");
    ensure_device();
    unsigned long long* d = dalloc<unsigned long long>(1);
    hipStream_t st;
    HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    HIPCHK(hipMemsetAsync(d, 0, sizeof(unsigned long long), st));
    if (count > 0) xb::launch_checksum(st, a_dev, count, d);
    unsigned long long h = 0;
    HIPCHK(hipMemcpyAsync(&h, d, sizeof(h), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    (void)hipStreamDestroy(st);
    dfree(d);
    *out = (uint64_t)h;
  });
}

int32_t xb_device_alloc(size_t bytes, void** ptr) {
  return guarded([&] {
    if (!ptr) fail(XB_ERR_INVALID_ARG, "null ptr");
    ensure_device();
    *ptr = nullptr;
    HIPCHK(hipMalloc(ptr, bytes ? bytes : 1));
  });
}

int32_t xb_device_free(void* ptr) {
  return guarded([&] {
    if (ptr) HIPCHK(hipFree(ptr));
  });
}

int32_t xb_memcpy_d2h(void* dst, const void* src, size_t bytes) {
  return guarded([&] {
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    HIPCHK(hipDeviceSynchronize());
  });
}

int32_t xb_memcpy_h2d(void* dst, const void* src, size_t bytes) {
  return guarded([&] {
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    HIPCHK(hipDeviceSynchronize());
  });
}

int32_t xb_last_error(char* buf, size_t len) {
  if (!buf || len == 0) return XB_ERR_INVALID_ARG;
  std::snprintf(buf, len, "%s", g_last_error.c_str());
  return XB_OK;
}

}  // extern "C"

// ===========================================================================
// Allocation-free runs: the workspace lives in a static device array
// ===========================================================================
// One copy of `scratch` per GPU (each device the library runs on holds its own
// at its own address until the process ends).  Host code never uses the
// array's name as an address: it is a host-side stand-in.  The device address
// comes from hipGetSymbolAddress, once per device, with that device current.
static __device__ __align__(256) unsigned char scratch[XB_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kNoAllocMaxDevices = 64;

struct NoAllocDevice {
  std::mutex mu;         // serializes the calls on one device (they share its array)
  void* ptr = nullptr;   // device address of that device's copy of `scratch`
};
NoAllocDevice g_noalloc[kNoAllocMaxDevices];

// The current device's scratch address and lock, set up on first use.
NoAllocDevice& noalloc_device() {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kNoAllocMaxDevices) fail(XB_ERR_UNSUPPORTED, "device index beyond the scratch table");
  NoAllocDevice& d = g_noalloc[dev];
  std::lock_guard<std::mutex> lk(d.mu);
  if (!d.ptr) {
    void* p = nullptr;
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) fail(XB_ERR_DEVICE, "hipGetSymbolAddress returned a null scratch address");
    d.ptr = p;
  }
  return d;
}

// Carves a workspace out of `base` (null: only sizes it).  T = 0: the plain
// engine's queues only; T > 0: also the bit-parallel tile state for tiles of
// up to T words.  Every piece is 256-byte aligned.  Returns the bytes used.
size_t carve_workspace(const xb_graph* g, int T, char* base, Workspace* w) {
  size_t used = 0;
  auto take = [&](size_t bytes) -> void* {
    void* p = base ? base + used : nullptr;
    used += (bytes + 255) & ~(size_t)255;
    return p;
  };
  const size_t n1 = (size_t)g->n + 1;
  Workspace ws;
  ws.ctrs = (unsigned long long*)take(2 * sizeof(unsigned long long));
  ws.tcount = (uint32_t*)take(sizeof(uint32_t));
  for (int i = 0; i < 2; ++i) {
    ws.q[i].v = (int32_t*)take(sizeof(int32_t) * n1);
    ws.q[i].eo = (int32_t*)take(sizeof(int32_t) * n1);
    ws.q[i].ctr = ws.ctrs ? ws.ctrs + i : nullptr;
  }
  if (T > 0) {
    const size_t nw = (size_t)g->n * (size_t)T + 1;
    ws.bp.seen = (uint64_t*)take(sizeof(uint64_t) * nw);
    ws.bp.front[0] = (uint64_t*)take(sizeof(uint64_t) * nw);
    ws.bp.front[1] = (uint64_t*)take(sizeof(uint64_t) * nw);
    ws.bp.nextw = (uint64_t*)take(sizeof(uint64_t) * nw);
    ws.bp.planes = (uint64_t*)take(sizeof(uint64_t) * nw * 4);
    ws.bp.planes_hi = (uint64_t*)take(sizeof(uint64_t) * nw * 4);
    ws.bp.tflag = (uint32_t*)take(sizeof(uint32_t) * n1);
    ws.bp.tlist = (int32_t*)take(sizeof(int32_t) * n1);
    ws.bp.tcount = ws.tcount;
    ws.bp.n = g->n;
    ws.side_cap = (int64_t)std::max(g->n_heavy, 1) * T * 64;
    ws.side = (int32_t*)take(sizeof(int32_t) * (size_t)ws.side_cap);
    ws.src_cap = 64 * (int64_t)T;
    ws.d_src = (int32_t*)take(sizeof(int32_t) * (size_t)ws.src_cap);
    ws.T = T;
    ws.bp_dirty = true;  // the array keeps the last call's contents: clear all tile state first
  }
  if (w) {
    ws.n = g->n;
    ws.st = g->stream;
    ws.hctr = g->pinned.p;
    ws.q0_host = g->q0_pinned;
    ws.q0_cap = 3 * 1024;
    *w = ws;
  }
  return used;
}

void run_noalloc(xb_graph* g, const int32_t* sources, int32_t S, int32_t* levels, int32_t* parents,
                 const xb_msbfs_opts* opts, xb_run_stats* stats) {
  if (!g) fail(XB_ERR_INVALID_ARG, "null graph");
  if (S < 0) fail(XB_ERR_INVALID_ARG, "S < 0");
  if (stats) std::memset(stats, 0, sizeof(*stats));
  if (S == 0) return;
  if (!sources || !levels || !parents) fail(XB_ERR_INVALID_ARG, "null sources, levels or parents");
  const xb_msbfs_opts o = parse_opts(opts);
  if (o.flags & (XB_NO_LEVELS | XB_PROFILE))
    fail(XB_ERR_UNSUPPORTED, "the allocation-free run writes both device tables and takes no XB_PROFILE");
  check_sources(g, sources, S);
  BPOpts bo = bp_opts(o);
  if (o.mode != XB_MODE_PLAIN &&
      (bo.cut_small != g->cut_small || bo.cut_medium != g->cut_medium || bo.cut_hub != g->cut_hub ||
       bo.hub_chunk != g->hub_chunk))
    fail(XB_ERR_UNSUPPORTED, "the graph's pull layout was built for other cut-offs; run xb_msbfs_run once with "
                             "these options to rebuild it");
  // Widest tile (at most the requested words) whose state fits the scratch.
  int T = 0;
  if (o.mode != XB_MODE_PLAIN)
    for (int t = bo.T; t >= 1 && T == 0; t /= 2)
      if (carve_workspace(g, t, nullptr, nullptr) <= XB_NOALLOC_SCRATCH_BYTES) T = t;
  const bool plain = o.mode == XB_MODE_PLAIN || (T == 0 && o.mode == XB_MODE_AUTO);
  if (!plain && T == 0)
    fail(XB_ERR_BUDGET, "graph too large for the allocation-free scratch: one 64-source tile needs " +
                            std::to_string(carve_workspace(g, 1, nullptr, nullptr)) + " bytes, the scratch holds " +
                            std::to_string((size_t)XB_NOALLOC_SCRATCH_BYTES));
  if (plain && carve_workspace(g, 0, nullptr, nullptr) > XB_NOALLOC_SCRATCH_BYTES)
    fail(XB_ERR_BUDGET, "graph too large for the allocation-free scratch: the plain engine's queues need " +
                            std::to_string(carve_workspace(g, 0, nullptr, nullptr)) + " bytes, the scratch holds " +
                            std::to_string((size_t)XB_NOALLOC_SCRATCH_BYTES));
  NoAllocDevice& dev = noalloc_device();
  std::lock_guard<std::mutex> lk(dev.mu);
  Workspace w;
  carve_workspace(g, plain ? 0 : T, (char*)dev.ptr, &w);
  Profiler prof;  // off: byte and launch counts only, no events
  prof.st = g->stream;
  const auto t0 = std::chrono::steady_clock::now();
  if (plain) {
    run_plain(g, w, sources, S, levels, parents, true, prof, stats);
  } else {
    bo.T = T;
    run_batched(g, w, bo, sources, S, levels, parents, true, prof, stats);
  }
  HIPCHK(hipStreamSynchronize(g->stream));
  if (stats) {
    prof.collect(stats);
    stats->ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
}

}  // namespace

extern "C" int32_t xb_msbfs_run_noalloc(xb_graph* g, const int32_t* sources, int32_t S, int32_t* levels_dev,
                                        int32_t* parents_dev, const xb_msbfs_opts* opts, xb_run_stats* stats) {
  return guarded([&] { run_noalloc(g, sources, S, levels_dev, parents_dev, opts, stats); });
}

// ===========================================================================
// Batched multi-graph runs: one persistent, hipGraph-captured launch
// ===========================================================================
// Fewest workgroups a re-split of a batch's grid leaves a team (the first
// plan gives each team at least 8, before anything is measured).
#ifndef XB_REBALANCE_FLOOR
#define XB_REBALANCE_FLOOR 4
#endif

// A team of one tile leaves its emit to standalone launches captured after
// the persistent kernel (two words per pass in 512-thread workgroups).
#ifndef XB_BATCH_EMIT_AFTER
#define XB_BATCH_EMIT_AFTER 1
#endif
#ifndef XB_BATCH_EMIT_BRANCHES
#define XB_BATCH_EMIT_BRANCHES 4
#endif

// The persistent kernel instantiates its tile loop for 1, 2, 4, 8 and 16
// words, as the host-driven engine does.
constexpr int kBatchMaxWords = 16;

struct xb_batch {
  hipStream_t stream = nullptr;
  int T = 8;
  int alpha = 20;
  std::vector<xb_graph*> graphs;
  std::vector<uint64_t> layout_gen;  // each graph's layout generation when the batch was built
  std::vector<void*> allocs;  // device memory owned by the batch
  xb::Team* d_teams = nullptr;
  xb::GraphView* d_views = nullptr;
  xb::TileJob* d_jobs = nullptr;
  xb::TeamCtl* d_ctl = nullptr;
  unsigned* d_err = nullptr;
  std::vector<xb::TileJob> h_jobs;
  std::vector<xb::Team> h_teams;
  std::vector<int> team_words;  // words per vertex the team's workspace holds
  std::vector<int> team_floor, team_cap;  // workgroups a team keeps at least / can use at most
  bool rebalance = true;                  // re-split the grid from each run's team clocks
  std::vector<std::vector<int32_t>> job_sources;  // host copy of each tile's sources (plain re-runs)
  int32_t n_teams = 0;
  int grid = 0;
  bool profile = false;  // XB_PROFILE: phase clocks of the persistent kernel go into the stats
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  // Branches of the captured graph that run the emits after the kernel.
  std::vector<hipStream_t> branch_streams;
  hipEvent_t fork = nullptr;
  std::vector<hipEvent_t> join;

  template <class U>
  U* alloc(size_t count) {
    U* p = dalloc<U>(count);
    allocs.push_back(p);
    return p;
  }
  ~xb_batch() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (stream) (void)hipStreamSynchronize(stream);
    for (hipEvent_t e : join) (void)hipEventDestroy(e);
    if (fork) (void)hipEventDestroy(fork);
    for (hipStream_t x : branch_streams) (void)hipStreamDestroy(x);
    for (void* p : allocs) dfree(p);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

// Clean-state invariant of team workspaces: fronts, next words and touched
// flags are zero between tiles.  Re-established after an aborted launch.
void batch_clear_state(xb_batch* b) {
  for (size_t i = 0; i < b->h_teams.size(); ++i) {
    const xb::Team& tm = b->h_teams[i];
    const xb_graph* g = b->graphs[b->h_jobs[tm.job_begin].graph];
    const size_t nw = (size_t)g->n * b->team_words[i] + 1;
    HIPCHK(hipMemsetAsync(tm.s.front[0], 0, sizeof(uint64_t) * nw, b->stream));
    HIPCHK(hipMemsetAsync(tm.s.front[1], 0, sizeof(uint64_t) * nw, b->stream));
    HIPCHK(hipMemsetAsync(tm.s.nextw, 0, sizeof(uint64_t) * nw, b->stream));
    HIPCHK(hipMemsetAsync(tm.s.tflag, 0, sizeof(uint32_t) * ((size_t)g->n + 1), b->stream));
  }
  HIPCHK(hipStreamSynchronize(b->stream));
}

xb_batch* build_batch(xb_graph* const* graphs, int32_t G, const int32_t* const* sources, const int32_t* S,
                      int32_t* const* levels_dev, int32_t* const* parents_dev, const xb_msbfs_opts& o) {
  if (G <= 0 || !graphs || !sources || !S || !parents_dev) fail(XB_ERR_INVALID_ARG, "empty or null batch");
  const bool no_levels = (o.flags & XB_NO_LEVELS) != 0;
  if (!levels_dev && !no_levels) fail(XB_ERR_INVALID_ARG, "null levels (set XB_NO_LEVELS)");
  for (int32_t i = 0; i < G; ++i) {
    if (!graphs[i]) fail(XB_ERR_INVALID_ARG, "null graph in batch");
    if (S[i] < 0) fail(XB_ERR_INVALID_ARG, "negative source count");
    if (S[i] > 0 && (!sources[i] || !parents_dev[i] || (!no_levels && !levels_dev[i])))
      fail(XB_ERR_INVALID_ARG, "null sources or outputs for graph " + std::to_string(i));
    check_sources(graphs[i], sources[i], S[i]);
  }
  std::unique_ptr<xb_batch> b(new xb_batch());
  const BPOpts bo = bp_opts(o);
  b->T = std::min(bo.T, kBatchMaxWords);
  b->alpha = bo.alpha;
  b->profile = (o.flags & XB_PROFILE) != 0;
  b->rebalance = (o.flags & XB_FIXED_TEAMS) == 0;
  HIPCHK(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
  const int T = b->T;
  const int32_t tile = 64 * T;

  // Graph views (one per graph of the batch).
  std::vector<xb::GraphView> views((size_t)G);
  for (int32_t i = 0; i < G; ++i) {
    xb_graph* g = graphs[i];
    b->graphs.push_back(g);
    ensure_layout(g, bo);
    b->layout_gen.push_back(g->layout_gen);
    xb::GraphView& v = views[i];
    v.off = g->d_eoff;
    v.col = g->d_ecol;
    v.in_off = g->d_ein_off;
    v.in_col = g->d_ein_col;
    v.in_col_o = g->d_ein_col_o;
    v.new_of = g->d_new_of;
    v.orig_of = g->d_orig_of;
    v.oin_off = g->d_oin_off;
    v.oin_col = g->d_oin_col;
    v.oin_colo = g->d_oin_colo;
    v.perm = g->d_perm;
    for (int c = 0; c < 3; ++c) {
      v.bucket_end[c] = g->bucket_end[c];
      v.bucket_perm[c] = g->bucket_contig[c] ? nullptr : g->d_perm;
      v.bucket_shift[c] = g->bucket_shift[c];
    }
    v.n_hub_chunks = g->n_hub_chunks;
    v.hubs = g->d_hub_chunks;
    v.hidx = g->d_hidx;
    v.n_heavy = g->n_heavy;
    v.n = g->n;
    v.cut_medium = bo.cut_medium;
    v.pad = 0;
    v.m = g->m;
  }

  // Tiles: sources and first frontiers, packed into one device buffer.
  std::vector<int32_t> packed;
  struct TileHost {
    int32_t graph, base, cnt, words;
    size_t src_off, q0_off;
    int32_t qn0, E0;
  };
  std::vector<TileHost> tiles;
  for (int32_t i = 0; i < G; ++i) {
    for (int32_t base = 0; base < S[i]; base += tile) {
      const int32_t cnt = std::min(tile, S[i] - base);
      TileHost th{i, base, cnt, tile_words(cnt, T), packed.size(), 0, 0, 0};
      packed.insert(packed.end(), sources[i] + base, sources[i] + base + cnt);
      std::vector<int32_t> uniq(sources[i] + base, sources[i] + base + cnt);
      std::sort(uniq.begin(), uniq.end());
      uniq.erase(std::unique(uniq.begin(), uniq.end()), uniq.end());
      th.q0_off = packed.size();
      th.qn0 = (int32_t)uniq.size();
      int64_t E = 0;
      std::vector<int32_t> eo(uniq.size());
      for (size_t k = 0; k < uniq.size(); ++k) {
        eo[k] = (int32_t)E;
        E += graphs[i]->outdeg(uniq[k]);
      }
      th.E0 = (int32_t)E;
      for (int32_t& u : uniq) u = graphs[i]->h_new_of[(size_t)u];  // the frontier holds engine ids
      packed.insert(packed.end(), uniq.begin(), uniq.end());
      packed.insert(packed.end(), eo.begin(), eo.end());
      tiles.push_back(th);
      b->job_sources.emplace_back(sources[i] + base, sources[i] + base + cnt);
    }
  }
  int32_t* d_packed = b->alloc<int32_t>(packed.size());
  if (!packed.empty())
    h2d(d_packed, packed.data(), sizeof(int32_t) * packed.size(), b->stream);

  // Teams: one per graph with tiles.  A team's wall time is its cost over its
  // workgroups, so workgroups go in proportion to an estimate of the bytes
  // its tiles move (the launch ends with the slowest team):
  //   ~3 pull levels  x (CSR + one frontier word per in-edge and word),
  //   ~8 levels       x 24 B of state per vertex and word,
  //   the output rows, and the emit's plane and in-edge reads.
  hipDeviceProp_t prop;
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  HIPCHK(hipGetDeviceProperties(&prop, dev));
  const int per_cu = std::max(1, xb::persistent_blocks_per_cu());
  const int budget = per_cu * prop.multiProcessorCount;
  const double out_b = no_levels ? 4.0 : 8.0;
  std::vector<int32_t> team_graph;
  std::vector<double> cost;
  for (int32_t i = 0; i < G; ++i) {
    if (S[i] == 0) continue;
    const double n = graphs[i]->n, m = (double)graphs[i]->m;
    double c = 0;
    for (const TileHost& th : tiles)
      if (th.graph == i) {
        const double w = th.words;
        c += 3.0 * (4.0 * m + 8.0 * w * m) + 8.0 * 24.0 * w * n + out_b * th.cnt * n + w * (64.0 * n + 16.0 * m);
      }
    team_graph.push_back(i);
    cost.push_back(c);
  }
  const int nt = (int)team_graph.size();
  if (nt == 0) fail(XB_ERR_INVALID_ARG, "batch has no sources");
  if (nt > budget) fail(XB_ERR_UNSUPPORTED, "more graphs in one batch than co-resident workgroups");
  // Workgroups a team can keep busy: one per 4096 vertices or edges.
  auto useful = [&](int t) {
    const xb_graph* g = graphs[team_graph[t]];
    return (int)std::max<int64_t>(1, (g->m + (int64_t)g->n) / 4096);
  };
  // A floor of min(useful, 8) keeps small graphs off a single workgroup's
  // latency; the rest is shared in proportion to cost.
  std::vector<int> blocks((size_t)nt);
  int used = 0;
  for (int t = 0; t < nt; ++t) {
    blocks[t] = std::min(useful(t), std::max(1, std::min(8, budget / nt)));
    used += blocks[t];
    b->team_floor.push_back(std::min(blocks[t], XB_REBALANCE_FLOOR));
    b->team_cap.push_back(useful(t));
  }
  for (int round = 0; round < 4 && used < budget; ++round) {
    double open_cost = 0;
    for (int t = 0; t < nt; ++t)
      if (blocks[t] < useful(t)) open_cost += cost[t];
    if (open_cost <= 0) break;
    const int spare = budget - used;
    for (int t = 0; t < nt; ++t) {
      if (blocks[t] >= useful(t)) continue;
      const int want = (int)(spare * cost[t] / open_cost);
      const int add = std::min(want, useful(t) - blocks[t]);
      blocks[t] += add;
      used += add;
    }
  }
  // Rounding leftovers: one at a time to the team with the highest cost per workgroup.
  while (used < budget) {
    int best = -1;
    double best_load = -1;
    for (int t = 0; t < nt; ++t)
      if (blocks[t] < useful(t) && cost[t] / blocks[t] > best_load) {
        best_load = cost[t] / blocks[t];
        best = t;
      }
    if (best < 0) break;
    ++blocks[best];
    ++used;
  }

  b->d_ctl = b->alloc<xb::TeamCtl>((size_t)nt);
  b->d_err = b->alloc<unsigned>(1);
  int first = 0;
  for (int t = 0; t < nt; ++t) {
    const int32_t gi = team_graph[t];
    xb_graph* g = graphs[gi];
    xb::Team tm{};
    tm.first_block = first;
    tm.nblocks = blocks[t];
    first += blocks[t];
    tm.job_begin = (int32_t)b->h_jobs.size();
    int tw = 1;  // widest tile of the team
    for (const TileHost& th : tiles) {
      if (th.graph != gi) continue;
      xb::TileJob j{};
      j.graph = gi;
      j.cnt = th.cnt;
      j.src = d_packed + th.src_off;
      j.q0v = d_packed + th.q0_off;
      j.q0eo = d_packed + th.q0_off + th.qn0;
      j.qn0 = th.qn0;
      j.E0 = th.E0;
      j.lvl_out = no_levels ? nullptr : levels_dev[gi] + (int64_t)th.base * g->n;
      j.par_out = parents_dev[gi] + (int64_t)th.base * g->n;
      j.status = -1;
      j.levels = -1;
      j.words = th.words;
      j.pad_ = 0;
      tw = std::max(tw, th.words);
      b->h_jobs.push_back(j);
    }
    tm.job_end = (int32_t)b->h_jobs.size();
    tm.emit_after =
        (XB_BATCH_EMIT_AFTER && tm.job_end - tm.job_begin == 1 && (o.flags & XB_BATCH_EMIT_INSIDE) == 0) ? 1 : 0;
    tm.pad_ = 0;
    const size_t nw = (size_t)g->n * tw + 1;
    tm.s.seen = b->alloc<uint64_t>(nw);
    tm.s.front[0] = b->alloc<uint64_t>(nw);
    tm.s.front[1] = b->alloc<uint64_t>(nw);
    tm.s.nextw = b->alloc<uint64_t>(nw);
    tm.s.planes = b->alloc<uint64_t>(nw * 8);
    tm.s.n = g->n;
    tm.s.tflag = b->alloc<uint32_t>((size_t)g->n + 1);
    tm.s.tlist = b->alloc<int32_t>((size_t)g->n + 1);
    tm.s.hpar = b->alloc<uint32_t>((size_t)std::max(g->n_heavy, 1) * tw * 64);
    tm.s.tcount = nullptr;  // set per level from the control block
    for (int k = 0; k < 2; ++k) {
      tm.qv[k] = b->alloc<int32_t>((size_t)g->n + 1);
      tm.qeo[k] = b->alloc<int32_t>((size_t)g->n + 1);
    }
    tm.ctl = b->d_ctl + t;
    b->h_teams.push_back(tm);
    b->team_words.push_back(tw);
  }
  // Job order follows the teams; remap the host-side source copies to match.
  {
    std::vector<std::vector<int32_t>> srcs;
    for (int t = 0; t < nt; ++t)
      for (size_t k = 0; k < tiles.size(); ++k)
        if (tiles[k].graph == team_graph[t]) srcs.push_back(b->job_sources[k]);
    b->job_sources.swap(srcs);
  }
  b->n_teams = nt;
  b->grid = first;
  b->d_teams = b->alloc<xb::Team>((size_t)nt);
  b->d_views = b->alloc<xb::GraphView>((size_t)G);
  b->d_jobs = b->alloc<xb::TileJob>(b->h_jobs.size());
  h2d(b->d_teams, b->h_teams.data(), sizeof(xb::Team) * nt, b->stream);
  h2d(b->d_views, views.data(), sizeof(xb::GraphView) * G, b->stream);
  h2d(b->d_jobs, b->h_jobs.data(), sizeof(xb::TileJob) * b->h_jobs.size(), b->stream);
  batch_clear_state(b.get());

  // Capture the launch (control reset + persistent kernel) once, followed by
  // the emits of the one-tile teams: per team one launch for codes below 16
  // and one for deeper tiles, each gated on the tile's status and depth as
  // the kernel leaves them in the job table.  The emits go to up to
  // XB_BATCH_EMIT_BRANCHES branches of the graph (forked after the kernel,
  // joined at its end), largest tile first to the least loaded branch, so
  // the emits of small graphs run beside each other.
  HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  xb::launch_bp_persistent(b->stream, b->grid, b->d_teams, b->n_teams, b->d_views, b->d_jobs,
                           (int32_t)b->h_jobs.size(), b->d_ctl, b->alpha, b->d_err);
  std::vector<int> emit_teams;
  for (int t = 0; t < nt; ++t)
    if (b->h_teams[t].emit_after) emit_teams.push_back(t);
  auto emit_cells = [&](int t) {
    const xb::TileJob& j = b->h_jobs[b->h_teams[t].job_begin];
    return (double)j.cnt * (double)graphs[j.graph]->n;
  };
  std::stable_sort(emit_teams.begin(), emit_teams.end(), [&](int a, int c) { return emit_cells(a) > emit_cells(c); });
  const int nbr = std::max(1, std::min<int>(XB_BATCH_EMIT_BRANCHES, (int)emit_teams.size()));
  std::vector<hipStream_t> br((size_t)nbr, b->stream);
  if (nbr > 1) {
    HIPCHK(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming));
    HIPCHK(hipEventRecord(b->fork, b->stream));
    for (int i = 1; i < nbr; ++i) {
      hipStream_t sx = nullptr;
      HIPCHK(hipStreamCreateWithFlags(&sx, hipStreamNonBlocking));
      b->branch_streams.push_back(sx);
      br[(size_t)i] = sx;
      HIPCHK(hipStreamWaitEvent(sx, b->fork, 0));
    }
  }
  std::vector<double> load((size_t)nbr, 0.0);
  for (int t : emit_teams) {
    const int bi = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    load[(size_t)bi] += emit_cells(t);
    const xb::Team& tm = b->h_teams[t];
    const xb::TileJob& j = b->h_jobs[tm.job_begin];
    const xb_graph* g = graphs[j.graph];
    xb::BPState s = tm.s;
    s.hidx = g->d_hidx;
    s.new_of = g->d_new_of;
    s.orig_of = g->d_orig_of;
    s.in_col_o = g->d_ein_col_o;
    s.planes_hi = s.planes + (size_t)g->n * j.words * 4;  // as the kernel's tile set-up places them
    const int W = xb::emit_pass_words(j.words);
    const int ngrp = (j.words + W - 1) / W;
    const int32_t* gate = &b->d_jobs[tm.job_begin].status;  // {status, levels}
    const bool sym = (g->flags & XB_GRAPH_DIRECTED) == 0;
    // Codes up to 15 (deepest level 14) in 4 planes, deeper tiles in 8.
    xb::launch_bp_emit(br[(size_t)bi], j.words, s, 0, ngrp, 4, j.cnt, j.src, g->d_oin_off, g->d_oin_col,
                       g->d_oin_colo, g->n, bo.cut_medium, j.lvl_out, j.par_out, xb::RowMap{}, sym, gate, 0, 14);
    xb::launch_bp_emit(br[(size_t)bi], j.words, s, 0, ngrp, 8, j.cnt, j.src, g->d_oin_off, g->d_oin_col,
                       g->d_oin_colo, g->n, bo.cut_medium, j.lvl_out, j.par_out, xb::RowMap{}, sym, gate, 15,
                       1 << 30);
  }
  for (int i = 1; i < nbr; ++i) {
    hipEvent_t ev = nullptr;
    HIPCHK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    b->join.push_back(ev);
    HIPCHK(hipEventRecord(ev, br[(size_t)i]));
    HIPCHK(hipStreamWaitEvent(b->stream, ev, 0));
  }
  HIPCHK(hipStreamEndCapture(b->stream, &b->graph));
  HIPCHK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  return b.release();
}

// Re-splits the grid between the teams for the next replay.  A team's
// measured work is its span (the sum of its phase clocks) times its
// workgroups; the target gives every team workgroups in proportion to its
// work, within its floor and cap, and the new split goes half-way from the
// current one towards it (team spans are not linear in the workgroup count,
// so a few replays converge instead of one jumping).  The grid total stays
// the captured launch's; results do not depend on the split.
void rebalance_teams(xb_batch* b, const std::vector<xb::TeamCtl>& ctl) {
  const int nt = b->n_teams;
  if (nt < 2) return;
  std::vector<double> w((size_t)nt);
  double W = 0;
  for (int t = 0; t < nt; ++t) {
    double ticks = 0;
    for (int c = 0; c < XB_K_CLASSES; ++c) ticks += (double)ctl[t].tphase[c];
    w[t] = ticks * b->h_teams[t].nblocks;
    W += w[t];
  }
  if (W <= 0) return;
  const int total = b->grid;
  auto lo = [&](int t) { return b->team_floor[t]; };
  auto hi = [&](int t) { return b->team_cap[t]; };
  // Water-filling of the target.
  std::vector<int> tgt((size_t)nt);
  int used = 0;
  for (int t = 0; t < nt; ++t) used += (tgt[t] = lo(t));
  for (int round = 0; round < 8 && used < total; ++round) {
    double open = 0;
    for (int t = 0; t < nt; ++t)
      if (tgt[t] < hi(t)) open += w[t];
    if (open <= 0) break;
    const int spare = total - used;
    int added = 0;
    for (int t = 0; t < nt; ++t) {
      if (tgt[t] >= hi(t)) continue;
      const int add = std::min(hi(t) - tgt[t], (int)(spare * w[t] / open));
      tgt[t] += add;
      added += add;
    }
    used += added;
    if (added == 0) break;
  }
  // Half-way, then exact total: extra workgroups go to the highest work per
  // workgroup, surplus ones leave the lowest.
  std::vector<int> nb((size_t)nt);
  int sum = 0;
  for (int t = 0; t < nt; ++t) {
    nb[t] = std::min(hi(t), std::max(lo(t), (b->h_teams[t].nblocks + tgt[t] + 1) / 2));
    sum += nb[t];
  }
  while (sum != total) {
    int pick = -1;
    double best = 0;
    for (int t = 0; t < nt; ++t) {
      const double load = w[t] / nb[t];
      if (sum < total && nb[t] < hi(t) && (pick < 0 || load > best)) pick = t, best = load;
      if (sum > total && nb[t] > lo(t) && (pick < 0 || load < best)) pick = t, best = load;
    }
    if (pick < 0) return;  // no admissible move: keep the current split
    nb[pick] += sum < total ? 1 : -1;
    sum += sum < total ? 1 : -1;
  }
  bool same = true;
  for (int t = 0; t < nt; ++t) same = same && nb[t] == b->h_teams[t].nblocks;
  if (same) return;
  int first = 0;
  for (int t = 0; t < nt; ++t) {
    b->h_teams[t].first_block = first;
    b->h_teams[t].nblocks = nb[t];
    first += nb[t];
  }
  h2d(b->d_teams, b->h_teams.data(), sizeof(xb::Team) * nt, b->stream);
}

void run_batch(xb_batch* b, xb_run_stats* st) {
  // The captured launch holds the graphs' layout arrays: any rebuild since the
  // batch was built (even one that restored the same cut-offs) freed them.
  for (size_t i = 0; i < b->graphs.size(); ++i)
    if (b->graphs[i]->layout_gen != b->layout_gen[i])
      fail(XB_ERR_INVALID_ARG, "a graph of the batch was re-tuned after the batch was created");
  if (st) std::memset(st, 0, sizeof(*st));
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  HIPCHK(hipEventRecord(e0, b->stream));
  HIPCHK(hipGraphLaunch(b->exec, b->stream));
  HIPCHK(hipEventRecord(e1, b->stream));
  HIPCHK(hipEventSynchronize(e1));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  unsigned err = 0;
  d2h(&err, b->d_err, sizeof(err), b->stream);
  std::vector<xb::TileJob> jobs(b->h_jobs.size());
  d2h(jobs.data(), b->d_jobs, sizeof(xb::TileJob) * jobs.size(), b->stream);
  if (err) {
    batch_clear_state(b);
    fail(XB_ERR_DEVICE, "persistent batch aborted: a team barrier timed out (workgroups not co-resident)");
  }
  double bytes = 0;
  int32_t max_level = 0;
  for (size_t k = 0; k < jobs.size(); ++k) {
    const xb::TileJob& j = jobs[k];
    xb_graph* g = b->graphs[j.graph];
    if (j.status == 1) {
      // Deeper than the level code: the plain path writes this tile's rows.
      Profiler prof;
      prof.st = g->stream;
      run_plain(g, g->ws, b->job_sources[k].data(), j.cnt, j.lvl_out, j.par_out, true, prof, st);
      if (st) st->fallback_sources += j.cnt;
      continue;
    }
    if (j.status != 0) fail(XB_ERR_DEVICE, "a batch tile did not complete");
    max_level = std::max(max_level, j.levels);
    // Algorithmic bytes of the tile: per level one CSR stream and the state
    // streams (24 B per word per vertex), plus the output tables.
    const double L = j.levels + 1.0, W = (j.cnt + 63) / 64, n = g->n, m = (double)g->m;
    bytes += L * (4.0 * (n + 1) + 4.0 * m) + L * 24.0 * W * n + (j.lvl_out ? 8.0 : 4.0) * j.cnt * n;
  }
  std::vector<xb::TeamCtl> ctl;
  if ((st && b->profile) || b->rebalance) {
    ctl.resize((size_t)b->n_teams);
    d2h(ctl.data(), b->d_ctl, sizeof(xb::TeamCtl) * ctl.size(), b->stream);
  }
  if (st && b->profile) {
    // Phase clocks: each team's critical-path time per phase, weighted by the
    // team's share of the grid, so the classes add up to about the wall time
    // (with the split this run used).
    int dev = 0, khz = 0;
    HIPCHK(hipGetDevice(&dev));
    HIPCHK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
    if (khz > 0)
      for (int32_t t = 0; t < b->n_teams; ++t) {
        const double share = (double)b->h_teams[t].nblocks / b->grid;
        for (int c = 0; c < XB_K_CLASSES; ++c)
          if (c != XB_K_PERSISTENT) st->kernel_ms[c] += (double)ctl[t].tphase[c] / khz * share;
      }
  }
  if (b->rebalance) rebalance_teams(b, ctl);
  if (st) {
    st->ms_total = ms;
    st->kernel_ms[XB_K_PERSISTENT] = ms;
    st->kernel_launches[XB_K_PERSISTENT] = 1;
    st->kernel_bytes[XB_K_PERSISTENT] = bytes;
    st->max_level = std::max(st->max_level, max_level);
    st->tiles = (int32_t)jobs.size();
  }
}

}  // namespace

extern "C" {

int32_t xb_batch_create(xb_graph* const* graphs, int32_t G, const int32_t* const* sources, const int32_t* S,
                        int32_t* const* levels_dev, int32_t* const* parents_dev, const xb_msbfs_opts* opts,
                        xb_batch** out) {
  return guarded([&] {
    if (!out) fail(XB_ERR_INVALID_ARG, "null out handle");
    *out = build_batch(graphs, G, sources, S, levels_dev, parents_dev, parse_opts(opts));
  });
}

int32_t xb_batch_run(xb_batch* b, xb_run_stats* stats) {
  return guarded([&] {
    if (!b) fail(XB_ERR_INVALID_ARG, "null batch");
    run_batch(b, stats);
  });
}

int32_t xb_batch_destroy(xb_batch* b) {
  return guarded([&] { delete b; });
}

int32_t xb_msbfs_run_batched(xb_graph* const* graphs, int32_t G, const int32_t* const* sources, const int32_t* S,
                             int32_t* const* levels, int32_t* const* parents, const xb_msbfs_opts* opts,
                             xb_run_stats* stats) {
  return guarded([&] {
    const xb_msbfs_opts o = parse_opts(opts);
    if (G <= 0 || !graphs || !S || !parents) fail(XB_ERR_INVALID_ARG, "empty or null batch");
    const bool no_levels = (o.flags & XB_NO_LEVELS) != 0;
    if (o.flags & XB_OUT_DEVICE) {
      std::unique_ptr<xb_batch> b(build_batch(graphs, G, sources, S, levels, parents, o));
      run_batch(b.get(), stats);
      return;
    }
    // Host outputs: stage the tables on the device, then copy back.
    std::vector<int32_t*> dl((size_t)G, nullptr), dp((size_t)G, nullptr);
    std::vector<void*> owned;
    struct Free {
      std::vector<void*>& v;
      ~Free() {
        for (void* p : v) dfree(p);
      }
    } freer{owned};
    for (int32_t i = 0; i < G; ++i) {
      if (!graphs[i]) fail(XB_ERR_INVALID_ARG, "null graph in batch");
      const size_t cells = (size_t)std::max(S[i], 0) * (size_t)graphs[i]->n;
      dp[i] = dalloc<int32_t>(cells);
      owned.push_back(dp[i]);
      if (!no_levels) {
        dl[i] = dalloc<int32_t>(cells);
        owned.push_back(dl[i]);
      }
    }
    {
      std::unique_ptr<xb_batch> b(build_batch(graphs, G, sources, S, no_levels ? nullptr : dl.data(), dp.data(), o));
      run_batch(b.get(), stats);
    }
    for (int32_t i = 0; i < G; ++i) {
      const size_t cells = (size_t)std::max(S[i], 0) * (size_t)graphs[i]->n;
      if (cells == 0) continue;
      HIPCHK(hipDeviceSynchronize());
      HIPCHK(hipMemcpy(parents[i], dp[i], sizeof(int32_t) * cells, hipMemcpyDeviceToHost));
      if (!no_levels) HIPCHK(hipMemcpy(levels[i], dl[i], sizeof(int32_t) * cells, hipMemcpyDeviceToHost));
    }
  });
}

}  // extern "C"

// ===========================================================================
// Forest mode: maximum matching by alternating forests
// ===========================================================================
namespace {

// Device arrays of one matching run (one entry per vertex unless noted).
struct MatchState {
  int32_t n = 0;
  int32_t* mate = nullptr;
  int32_t* root = nullptr;
  int32_t* lvl = nullptr;
  int32_t* par = nullptr;
  int32_t* best = nullptr;
  int32_t* acc = nullptr;
  int32_t* roots = nullptr;  // the phase's even roots (exposed vertices) not yet taken
  int32_t* roots2 = nullptr; // compaction target for roots
  int32_t* pairs = nullptr;  // the round's taken tree pairs, by lower root
  int32_t* tlist = nullptr;  // first claims of the level
  int32_t* cursor = nullptr; // level-0 scan position in each root's adjacency
  int32_t* fp = nullptr;     // outer vertex -> a vertex above it on its path (FIRST by following outer ones)
  unsigned long long* claim = nullptr;
  unsigned long long* bridge = nullptr;
  unsigned long long* elab = nullptr;  // blossom edge label (x << 32 | y) of a vertex made outer, else ~0
  unsigned long long* lab = nullptr;   // blossom step: lowest (candidate << 1 | side) marking a vertex, else ~0
  int32_t* marked = nullptr;           // blossom step: vertices marked this step
  uint8_t* used = nullptr;
  int4* pool = nullptr;                // the flip's pending second calls
  uint32_t pool_cap = 0;
  // Blossom candidates (grown on demand).
  unsigned long long* cand = nullptr;
  unsigned long long* cand_sorted = nullptr;
  int32_t* join_of = nullptr;  // each sorted candidate's join
  uint32_t* formed = nullptr;  // candidates that labelled a vertex
  void* sort_tmp = nullptr;
  size_t sort_bytes = 0;
  int32_t cand_cap = 0;
  // Counters.  Per phase: [0] same-tree even-even edges met by the scans,
  // [2] low word: taken pairs, [3] low word: first claims, [4] low word: roots
  // kept by a compaction, [5] low word: candidates, [8] low word: vertices
  // marked by a blossom step, [9] low word: flip pool entries.  Per run: [1] matched count /
  // validation failures, [6] entries read by the level-0 scans, [7] error
  // bits, [10] blossoms formed, [11] vertices made outer by them.
  unsigned long long* ctl = nullptr;

  explicit MatchState(int32_t n_) : n(n_) {
    const size_t N = (size_t)n;
    mate = dalloc<int32_t>(N);
    root = dalloc<int32_t>(N);
    lvl = dalloc<int32_t>(N);
    par = dalloc<int32_t>(N);
    best = dalloc<int32_t>(N);
    acc = dalloc<int32_t>(N);
    roots = dalloc<int32_t>(N);
    roots2 = dalloc<int32_t>(N);
    pairs = dalloc<int32_t>(N);
    tlist = dalloc<int32_t>(N);
    cursor = dalloc<int32_t>(N);
    fp = dalloc<int32_t>(N);
    claim = dalloc<unsigned long long>(N);
    bridge = dalloc<unsigned long long>(N);
    elab = dalloc<unsigned long long>(N);
    lab = dalloc<unsigned long long>(N);
    marked = dalloc<int32_t>(N);
    used = dalloc<uint8_t>(N);
    // Each blossom label expands at most once per flip: n pending calls cover a round.
    pool_cap = (uint32_t)N + 64;
    pool = dalloc<int4>(pool_cap);
    ctl = dalloc<unsigned long long>(16);
  }
  ~MatchState() {
    for (void* p : {(void*)mate, (void*)root, (void*)lvl, (void*)par, (void*)best, (void*)acc, (void*)roots,
                    (void*)roots2, (void*)cursor, (void*)fp, (void*)pairs, (void*)tlist, (void*)claim,
                    (void*)bridge, (void*)elab, (void*)lab, (void*)marked, (void*)used, (void*)pool, (void*)ctl,
                    (void*)cand, (void*)cand_sorted, (void*)join_of, (void*)formed, sort_tmp})
      dfree(p);
  }
  uint32_t* npairs() { return (uint32_t*)(ctl + 2); }
  uint32_t* tcount() { return (uint32_t*)(ctl + 3); }
  uint32_t* kept() { return (uint32_t*)(ctl + 4); }
  uint32_t* ncand() { return (uint32_t*)(ctl + 5); }
  unsigned* err() { return (unsigned*)(ctl + 7); }
  uint32_t* nmarked() { return (uint32_t*)(ctl + 8); }
  uint32_t* pool_ctr() { return (uint32_t*)(ctl + 9); }

  void ensure_cand(int32_t C) {
    if (C <= cand_cap) return;
    const int32_t cap = std::max(C, std::max(2 * cand_cap, 4096));
    for (void* p : {(void*)cand, (void*)cand_sorted, (void*)join_of, (void*)formed, sort_tmp}) dfree(p);
    cand = cand_sorted = nullptr;
    join_of = nullptr;
    formed = nullptr;
    sort_tmp = nullptr;
    cand = dalloc<unsigned long long>((size_t)cap);
    cand_sorted = dalloc<unsigned long long>((size_t)cap);
    join_of = dalloc<int32_t>((size_t)cap);
    formed = dalloc<uint32_t>((size_t)cap);
    sort_bytes = xb::sort_u64_temp_bytes(cap);
    sort_tmp = dalloc<uint8_t>(sort_bytes);
    cand_cap = cap;
  }
};

// Copies the first `count` (<= 8) counters into the pinned slot.
const unsigned long long* read_ctl(Workspace& w, const unsigned long long* d, int count) {
  HIPCHK(hipMemcpyAsync(w.hctr, d, sizeof(unsigned long long) * (size_t)count, hipMemcpyDeviceToHost, w.st));
  HIPCHK(hipStreamSynchronize(w.st));
  return w.hctr;
}

unsigned long long read_u64(Workspace& w, const unsigned long long* d) {
  HIPCHK(hipMemcpyAsync(w.hctr, d, sizeof(unsigned long long), hipMemcpyDeviceToHost, w.st));
  HIPCHK(hipStreamSynchronize(w.st));
  return *w.hctr;
}

// Phase loop of FindMaximumMatch (max_match_engine.h:93-113) over the
// alternating-forest search (xblossom_gpu_engine_10.cu:72-81): grow the
// forest step by step; at every step, take rounds of tree-disjoint augmenting
// paths until none is left, and end the phase after the first step that
// yields any; stop when a whole phase (the forest grown until no vertex is
// added) yields none.  The first phase grows from the exposed vertices; with
// `reuse` every later phase keeps the trees whose root is still exposed
// (xblossom_gpu_engine_10.cu:119-175) and grows from all their outer
// vertices, else it regrows the forest from the exposed vertices.
void run_match(xb_graph* g, MatchState& ms, Profiler& prof, bool reuse, xb_match_stats* out) {
  Workspace& w = g->ws;
  ensure_queues(g, w);
  const int32_t n = g->n;
  const hipStream_t st = w.st;
  xb_match_stats s{};
  // Algorithmic bytes: per frontier edge col + level (+ root of an even
  // end) for the scans, col + level + the 8-byte claim for the claim pass;
  // per frontier vertex its queue entry, offsets and root; per committed
  // pair its list entry, mates, claims, the claimer's level and the six label
  // writes; per vertex the 41 bytes of forest state the reset writes (the
  // reuse set-up reads root, mate, the root's flag, level and offsets, and
  // clears 17 bytes of pairing state).
  constexpr double kScanEdge = 12, kClaimEdge = 16, kFrontVertex = 16, kCommitPair = 68, kResetVertex = 41,
                   kRootItem = 12, kReuseVertex = 38;
  HIPCHK(hipMemsetAsync(ms.ctl, 0, sizeof(unsigned long long) * 16, st));
  HIPCHK(hipMemsetAsync(ms.lab, 0xFF, sizeof(unsigned long long) * (size_t)n, st));
  const int32_t max_steps = 4 * n + 16;
  // Once a phase's level 0 has run to a round with no pair, no two exposed
  // vertices are adjacent, and that stays so: an augmentation matches both of
  // its exposed ends and exposes nothing.  Later phases skip the level-0
  // rounds, which could only scan the roots' lists and find nothing.
  bool roots_independent = false;
  // The phase's roots: the previous phase's untaken roots (kept by the
  // rounds' compaction) are exactly the exposed vertices of the next.
  int32_t nroots = 0;
  for (;;) {
    ++s.phases;
    HIPCHK(hipMemsetAsync(ms.ctl, 0, sizeof(unsigned long long) * 6, st));
    HIPCHK(hipMemsetAsync(w.q[0].ctr, 0, sizeof(unsigned long long), st));
    int32_t qn = 0, E = 0;
    bool roots_only;  // the frontier is the phase's roots, every tree one root
    if (reuse && s.phases > 1) {
      prof.begin(XB_K_MT_INIT);
      xb::launch_mt_reuse_init(st, n, ms.mate, g->d_off, ms.root, ms.lvl, ms.par, ms.fp, ms.elab, ms.claim, ms.used,
                               w.q[0], ms.ctl + 12);
      xb::launch_mt_reuse_clear(st, n, ms.best, ms.acc, ms.bridge, ms.used);
      prof.end();
      prof.add(XB_K_MT_INIT, kReuseVertex * n);
      read_ctr(w, w.q[0].ctr, &qn, &E);
      roots_only = false;
      if (nroots < 2) break;
    } else {
      prof.begin(XB_K_MT_INIT);
      xb::launch_mt_init(st, n, ms.mate, g->d_off, ms.root, ms.lvl, ms.par, ms.claim, ms.best, ms.acc, ms.bridge,
                         ms.used, ms.cursor, ms.fp, ms.elab, w.q[0]);
      prof.end();
      prof.add(XB_K_MT_INIT, (kResetVertex + 20) * n);
      s.reset_tree_vertices += n;
      read_ctr(w, w.q[0].ctr, &qn, &E);
      nroots = qn;
      roots_only = true;
      if (nroots < 2) break;  // at most one exposed vertex: no augmenting path exists
      HIPCHK(hipMemcpyAsync(ms.roots, w.q[0].v, sizeof(int32_t) * (size_t)nroots, hipMemcpyDeviceToDevice, st));
    }
    int cur = 0;
    int64_t found = 0;
    unsigned long long same_tree = 0;  // the phase's count of same-tree even-even edges so far
    while (qn > 0) {
      int64_t level_same_tree = 0;
      for (;;) {
        if (roots_only && roots_independent) break;
        if (roots_only) {
          prof.begin(XB_K_MT_SCAN);
          xb::launch_mt_scan0(st, ms.roots, nroots, g->d_off, g->d_col, ms.lvl, ms.used, ms.cursor, ms.best,
                              ms.ctl + 6);
          prof.end();
          // Cursor, offsets and best per root; col + level + taken flag per entry
          // read (the entry count is added at the end of the run).
          prof.add(XB_K_MT_SCAN, 16.0 * nroots);
        } else {
          prof.begin(XB_K_MT_TREE);
          xb::launch_mt_reset_best(st, ms.roots, nroots, ms.best);
          prof.end();
          prof.begin(XB_K_MT_SCAN);
          xb::launch_mt_scan(st, g->d_off, g->d_col, w.q[cur], qn, E, ms.root, ms.lvl, ms.used, ms.best, ms.ctl);
          prof.end();
          prof.add(XB_K_MT_SCAN, kScanEdge * E + kFrontVertex * qn);
          s.aug_path_edges += E;
        }
        // The round's counters: pairs, flip-stack pool, kept roots and the
        // next frontier queue (free since the last round's compaction).
        xb::launch_mt_zero_ctrs(st, ms.npairs(), ms.pool_ctr(), ms.kept(), w.q[cur ^ 1].ctr);
        prof.begin(XB_K_MT_TREE);
        xb::launch_mt_accept(st, ms.roots, nroots, ms.best, ms.acc, ms.pairs, ms.npairs(),
                             roots_only ? ms.bridge : nullptr);
        prof.end();
        prof.add(XB_K_MT_TREE, 2 * kRootItem * nroots);
        const unsigned long long* c = read_ctl(w, ms.ctl, 4);
        const int32_t np = (int32_t)(c[2] & 0xFFFFFFFFull);
        level_same_tree = (int64_t)(c[0] - same_tree);
        same_tree = c[0];
        if (np == 0) break;
        if (!roots_only) {
          prof.begin(XB_K_MT_SCAN);
          xb::launch_mt_bridge(st, g->d_off, g->d_col, w.q[cur], qn, E, ms.root, ms.lvl, ms.acc, ms.bridge);
          prof.end();
          prof.add(XB_K_MT_SCAN, kScanEdge * E + kFrontVertex * qn);
          s.aug_path_edges += E;
        }
        prof.begin(XB_K_MT_TREE);
        xb::launch_mt_flip(st, ms.pairs, ms.npairs(), np, ms.acc, ms.bridge, ms.par, ms.elab, ms.mate, ms.used,
                           xb::FlipStack{ms.pool, ms.pool_ctr(), ms.pool_cap}, max_steps, ms.err());
        // The next round sees only the frontier and roots of trees not taken.
        xb::launch_mt_compact(st, w.q[cur].v, qn, ms.root, ms.used, g->d_off, w.q[cur ^ 1], nullptr, nullptr);
        xb::launch_mt_compact(st, ms.roots, nroots, ms.root, ms.used, g->d_off, xb::Queue{}, ms.roots2,
                              ms.kept());
        prof.end();
        prof.add(XB_K_MT_TREE, 17.0 * qn + 9.0 * nroots);
        found += np;
        cur ^= 1;
        std::swap(ms.roots, ms.roots2);
        // Both counters in one round trip.
        HIPCHK(hipMemcpyAsync(w.hctr, w.q[cur].ctr, sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(w.hctr + 1, ms.ctl + 4, sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        qn = (int32_t)(uint32_t)(w.hctr[0] & 0xFFFFFFFFull);
        E = (int32_t)(uint32_t)(w.hctr[0] >> 32);
        nroots = (int32_t)(w.hctr[1] & 0xFFFFFFFFull);
      }
      if (roots_only) roots_independent = true;
      if (found) break;
      // Blossom candidates are the same-tree even-even edges the scan met;
      // take them before the claims grow the forest.
      const int32_t C = (int32_t)level_same_tree;
      if (C > 0) {
        ms.ensure_cand(C);
        HIPCHK(hipMemsetAsync(ms.ncand(), 0, sizeof(uint32_t), st));
        prof.begin(XB_K_MT_BLOSSOM);
        xb::launch_mt_collect(st, g->d_off, g->d_col, w.q[cur], qn, E, ms.root, ms.lvl, ms.cand, ms.ncand(),
                              (uint32_t)C, ms.err());
        prof.end();
        prof.add(XB_K_MT_BLOSSOM, kScanEdge * E + kFrontVertex * qn + 8.0 * C);
        s.blossom_scan_edges += E;
      }
      HIPCHK(hipMemsetAsync(ms.tcount(), 0, sizeof(uint32_t), st));
      HIPCHK(hipMemsetAsync(w.q[cur ^ 1].ctr, 0, sizeof(unsigned long long), st));
      prof.begin(XB_K_MT_EXPAND);
      xb::launch_mt_claim(st, g->d_off, g->d_col, w.q[cur], qn, E, ms.root, ms.lvl, ms.claim, ms.tlist,
                          ms.tcount());
      // First claims come one per frontier edge at most, and one per vertex.
      xb::launch_mt_commit(st, ms.tlist, ms.tcount(), std::min(E, n), ms.mate, g->d_off, ms.claim, ms.root, ms.lvl,
                           ms.par, ms.fp, w.q[cur ^ 1]);
      prof.end();
      prof.add(XB_K_MT_EXPAND, kClaimEdge * E + kFrontVertex * qn);
      s.expand_edges += E;
      if (C > 0) {
        // Per tree, in (root, edge position) order: odd vertices on each new
        // blossom become outer and join the next frontier.
        prof.begin(XB_K_MT_BLOSSOM);
        xb::sort_u64(ms.sort_tmp, ms.sort_bytes, ms.cand, ms.cand_sorted, C, st);
        HIPCHK(hipMemsetAsync(ms.nmarked(), 0, sizeof(uint32_t), st));
        HIPCHK(hipMemsetAsync(ms.formed, 0, sizeof(uint32_t) * (size_t)C, st));
        xb::launch_mt_blossom(st, g->d_off, g->d_col, n, ms.par, ms.lvl, ms.fp, ms.elab, ms.lab, ms.cand_sorted, C,
                              ms.join_of, ms.marked, ms.nmarked(), ms.formed, w.q[cur ^ 1], ms.ctl + 10, ms.err());
        prof.end();
        prof.add(XB_K_MT_BLOSSOM, 48.0 * C);
      }
      read_ctr(w, w.q[cur ^ 1].ctr, &qn, &E);
      prof.add(XB_K_MT_EXPAND, kCommitPair * qn);
      cur ^= 1;
      roots_only = false;
      ++s.levels;
    }
    s.augmentations += found;
    if (!found) {
      s.blossom_edges = (int64_t)read_u64(w, ms.ctl);
      break;
    }
  }
  HIPCHK(hipMemsetAsync(ms.ctl + 1, 0, sizeof(unsigned long long), st));
  xb::launch_mt_count(st, n, ms.mate, ms.ctl + 1);
  s.matched_vertices = (int64_t)read_u64(w, ms.ctl + 1);
  const unsigned long long* c = read_ctl(w, ms.ctl + 6, 8);  // [6] .. [13]
  const int64_t examined = (int64_t)c[0];
  const unsigned err = (unsigned)(c[1] & 0xFFFFFFFFull);
  s.blossoms = (int64_t)c[4];
  s.blossom_vertices = (int64_t)c[5];
  s.reused_tree_vertices += (int64_t)c[6];
  s.reset_tree_vertices += (int64_t)c[7];
  if (err)
    fail(XB_ERR_DEVICE, "matching: a device-side bound was exceeded (error bits " + std::to_string(err) + ")");
  s.aug_path_edges += examined;
  s.edges_scanned = s.aug_path_edges + s.expand_edges + s.blossom_scan_edges;
  prof.add(XB_K_MT_SCAN, 9.0 * examined);
  *out = s;
}

}  // namespace

extern "C" int32_t xb_match_run(xb_graph* g, const int32_t* init_mate, int32_t* mate_out, uint32_t flags,
                                xb_match_stats* stats) {
  return guarded([&] {
    if (!g || !mate_out) fail(XB_ERR_INVALID_ARG, "null graph or mate_out");
    if (g->flags & XB_GRAPH_DIRECTED) fail(XB_ERR_UNSUPPORTED, "matching needs a symmetric graph");
    if (flags & ~(uint32_t)(XB_OUT_DEVICE | XB_PROFILE | XB_MATCH_REGROW))
      fail(XB_ERR_INVALID_ARG, "unknown matching flags");
    if (stats) std::memset(stats, 0, sizeof(*stats));
    const int32_t n = g->n;
    if (n == 0) return;
    const bool dev = (flags & XB_OUT_DEVICE) != 0;
    const hipStream_t st = g->stream;
    if (!g->match_state)
      g->match_state = std::shared_ptr<void>(new MatchState(n), [](void* p) { delete (MatchState*)p; });
    MatchState& ms = *(MatchState*)g->match_state.get();
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    struct Ev {
      hipEvent_t a, b;
      ~Ev() {
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
      }
    } evs{e0, e1};
    HIPCHK(hipEventRecord(e0, st));
    if (init_mate) {
      HIPCHK(hipMemcpyAsync(ms.mate, init_mate, sizeof(int32_t) * (size_t)n,
                            dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, st));
      HIPCHK(hipMemsetAsync(ms.ctl + 1, 0, sizeof(unsigned long long), st));
      xb::launch_mt_validate(st, n, ms.mate, g->d_off, g->d_col, ms.ctl + 1);
      const unsigned long long bad = read_u64(g->ws, ms.ctl + 1);
      if (bad) fail(XB_ERR_INVALID_ARG, "init_mate is not a matching of the graph (" + std::to_string(bad) +
                                            " bad entries)");
    } else {
      HIPCHK(hipMemsetAsync(ms.mate, 0xFF, sizeof(int32_t) * (size_t)n, st));
    }
    xb_match_stats s{};
    Profiler prof;
    prof.on = (flags & XB_PROFILE) != 0;
    prof.st = st;
    run_match(g, ms, prof, (flags & XB_MATCH_REGROW) == 0, &s);
    HIPCHK(hipMemcpyAsync(mate_out, ms.mate, sizeof(int32_t) * (size_t)n,
                          dev ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, st));
    HIPCHK(hipEventRecord(e1, st));
    HIPCHK(hipEventSynchronize(e1));
    float ms_total = 0;
    HIPCHK(hipEventElapsedTime(&ms_total, e0, e1));
    s.ms_total = ms_total;
    xb_run_stats ks{};
    prof.collect(&ks);
    std::memcpy(s.kernel_ms, ks.kernel_ms, sizeof(s.kernel_ms));
    std::memcpy(s.kernel_bytes, ks.kernel_bytes, sizeof(s.kernel_bytes));
    std::memcpy(s.kernel_launches, ks.kernel_launches, sizeof(s.kernel_launches));
    if (stats) *stats = s;
  });
}

// ===========================================================================
// Autotuner: coordinate descent over the engine's tunables
// ===========================================================================
namespace {

struct TuneRun {
  xb_graph* g;
  const int32_t* sample;
  int32_t S;
  int32_t* dl = nullptr;
  int32_t* dp = nullptr;
  xb_tune_report* rep;

  // Milliseconds per source of the whole sample (best of two runs, after
  // one untimed run that also rebuilds the layout for new cut-offs), run as
  // xb_msbfs_run runs it: narrower tiles are several tiles on two streams.
  double measure(const xb_msbfs_opts& o) {
    const BPOpts bo = bp_opts(o);
    const int32_t cnt = S;
    Profiler prof, prof2;
    prof.st = g->stream;
    Workspace* w2 = nullptr;
    if (cnt > 64 * bo.T && ws2_fits(g, bo.T)) {
      ensure_ws2(g);
      w2 = &g->ws2;
      prof2.st = g->stream2;
    }
    run_batched(g, g->ws, bo, sample, cnt, dl, dp, true, prof, nullptr, w2, &prof2);
    hipEvent_t a, b;
    HIPCHK(hipEventCreate(&a));
    HIPCHK(hipEventCreate(&b));
    double best = 1e300;
    for (int rep_i = 0; rep_i < 2; ++rep_i) {
      HIPCHK(hipEventRecord(a, g->stream));
      run_batched(g, g->ws, bo, sample, cnt, dl, dp, true, prof, nullptr, w2, &prof2);
      HIPCHK(hipEventRecord(b, g->stream));
      HIPCHK(hipEventSynchronize(b));
      float ms = 0;
      HIPCHK(hipEventElapsedTime(&ms, a, b));
      best = std::min(best, (double)ms / cnt);
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (rep && rep->trials < XB_TUNE_MAX_TRIALS) {
      xb_tune_trial& t = rep->trial[rep->trials++];
      t.words_per_tile = o.words_per_tile;
      t.alpha = o.alpha;
      t.pull_cut_small = o.pull_cut_small;
      t.pull_cut_medium = o.pull_cut_medium;
      t.pull_cut_hub = o.pull_cut_hub;
      t.hub_chunk = o.hub_chunk;
      t.ms_per_source = best;
    }
    return best;
  }
};

bool valid_tunables(const xb_msbfs_opts& o) {
  return o.alpha >= 1 && o.pull_cut_small >= 1 && o.pull_cut_medium >= o.pull_cut_small &&
         o.pull_cut_hub >= o.pull_cut_medium && o.hub_chunk >= 1;
}

}  // namespace

extern "C" int32_t xb_autotune(xb_graph* g, const int32_t* sample, int32_t S, const xb_msbfs_opts* base,
                               xb_msbfs_opts* best, xb_tune_report* report) {
  return guarded([&] {
    if (!g || !sample || S < 1 || !best) fail(XB_ERR_INVALID_ARG, "bad argument");
    check_sources(g, sample, S);
    xb_msbfs_opts cur = parse_opts(base);
    cur.mode = XB_MODE_BATCHED;
    cur.flags = XB_OUT_DEVICE;
    if (report) std::memset(report, 0, sizeof(*report));
    TuneRun tr{g, sample, S, nullptr, nullptr, report};
    // One tile of the widest trial, unless its two scratch tables (8 bytes
    // per row and vertex) would take more than half the free device memory:
    // then fewer rows, in whole 64-source words, and wider trials run with
    // the rows there are (tile_words narrows them, so they cannot win).
    int32_t rows = std::min(S, 64 * kMaxTileWords);
    {
      size_t fr = 0, tot = 0;
      HIPCHK(hipMemGetInfo(&fr, &tot));
      while (rows > 64 && 8.0 * rows * (double)g->n > 0.5 * (double)fr) rows = std::max(64, rows / 2 / 64 * 64);
    }
    tr.S = rows;
    struct Free {
      TuneRun& t;
      ~Free() {
        dfree(t.dl);
        dfree(t.dp);
      }
    } freer{tr};
    tr.dl = dalloc<int32_t>((size_t)rows * (size_t)g->n);
    tr.dp = dalloc<int32_t>((size_t)rows * (size_t)g->n);
    double t_cur = tr.measure(cur);
    if (report) report->default_ms_per_source = t_cur;
    // Candidate values per tunable, tried in this order.
    struct Axis {
      int32_t xb_msbfs_opts::*field;
      std::vector<int32_t> values;
    };
    const std::vector<Axis> axes = {
        {&xb_msbfs_opts::words_per_tile, {1, 2, 4, 8, 16}},
        {&xb_msbfs_opts::alpha, {6, 10, 14, 20, 28, 40, 64}},
        {&xb_msbfs_opts::pull_cut_small, {4, 8, 16}},
        {&xb_msbfs_opts::pull_cut_medium, {32, 64, 128, 256}},
        {&xb_msbfs_opts::pull_cut_hub, {512, 1024, 2048, 4096}},
        {&xb_msbfs_opts::hub_chunk, {256, 512, 1024, 2048}},
    };
    int passes = 0;
    for (bool changed = true; changed && passes < 2; ++passes) {
      changed = false;
      for (const Axis& ax : axes)
        for (int32_t v : ax.values) {
          if (cur.*(ax.field) == v) continue;
          xb_msbfs_opts trial = cur;
          trial.*(ax.field) = v;
          if (!valid_tunables(trial)) continue;
          const double t = tr.measure(trial);
          if (t < 0.98 * t_cur) {
            cur = trial;
            t_cur = t;
            changed = true;
          }
        }
    }
    if (report) {
      report->passes = passes;
      report->best_ms_per_source = t_cur;
    }
    *best = cur;
    best->flags = base ? base->flags : 0u;
  });
}

// ===========================================================================
// Serving: continuous batching over one device-resident graph
// ===========================================================================
namespace {

using ServeClock = std::chrono::steady_clock;

// First-fit allocator over one device allocation (byte offsets, 256-aligned),
// with neighbour coalescing on release.
struct Arena {
  char* base = nullptr;
  int64_t size = 0, used = 0, peak = 0;
  std::map<int64_t, int64_t> free_;  // offset -> length

  static int64_t round(int64_t b) { return (b + 255) & ~(int64_t)255; }
  void init(int64_t bytes) {
    base = dalloc<char>((size_t)bytes);
    size = bytes;
    free_.clear();
    free_[0] = bytes;
  }
  int64_t alloc(int64_t bytes) {
    bytes = round(bytes);
    for (auto it = free_.begin(); it != free_.end(); ++it)
      if (it->second >= bytes) {
        const int64_t off = it->first, len = it->second;
        free_.erase(it);
        if (len > bytes) free_[off + bytes] = len - bytes;
        used += bytes;
        peak = std::max(peak, used);
        return off;
      }
    return -1;
  }
  void release(int64_t off, int64_t bytes) {
    bytes = round(bytes);
    used -= bytes;
    auto it = free_.emplace(off, bytes).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
};

struct ServeReq {
  uint64_t id = 0;
  std::vector<int32_t> src;
  int32_t S = 0;
  int64_t off = -1, bytes = 0;  // arena block: levels (S x n) then parents (S x n)
  int32_t next_row = 0;         // rows handed to tiles so far
  int32_t rows_done = 0;
  int32_t tiles = 0;
  int64_t edges = 0;
  int32_t status = 1;  // 1 while pending, then XB_OK or the failure's code
  std::string err;
  double t_submit = 0, t_admit = -1, t_start = -1, t_done = -1;
};

// Request rows [row0, row0 + rows) carried by consecutive rows of one tile.
struct TileSeg {
  ServeReq* req;
  int32_t row0, rows;
};

struct ServeSlot {
  Workspace ws;
  unsigned long long* hctr = nullptr;  // pinned, per-level counter reads
  int64_t* h_rows = nullptr;           // pinned: level row offsets, then parent row offsets
  int64_t* d_rows = nullptr;
  unsigned long long* h_edges = nullptr;  // pinned, one per segment
  unsigned long long* d_edges = nullptr;
  std::thread th;
};

}  // namespace

struct xb_server {
  xb_graph* g = nullptr;
  BPOpts o{};
  uint64_t layout_gen = 0;  // the graph's layout generation the slots were set up with
  int dev = 0;
  int32_t tile = 0;
  bool count_edges = true;
  int32_t max_queued = 0;
  Arena arena;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<ServeReq*> pending;  // arrival order; requests with rows not yet in a tile
  std::unordered_map<uint64_t, std::unique_ptr<ServeReq>> reqs;
  std::vector<std::unique_ptr<ServeSlot>> slots;
  uint64_t next_id = 1;
  bool stop = false;
  ServeClock::time_point t0;
  int64_t submitted = 0, completed = 0, failed = 0, released = 0, tiles = 0, tile_rows = 0;
  int32_t running = 0;

  double now_us() const {
    return std::chrono::duration<double, std::micro>(ServeClock::now() - t0).count();
  }
  ~xb_server() {
    for (auto& sl : slots) {
      sl->ws.free_all();
      dfree(sl->d_rows);
      dfree(sl->d_edges);
      if (sl->h_rows) (void)hipHostFree(sl->h_rows);
      if (sl->h_edges) (void)hipHostFree(sl->h_edges);
      if (sl->hctr) (void)hipHostFree(sl->hctr);
      if (sl->ws.st) (void)hipStreamDestroy(sl->ws.st);
    }
    dfree(arena.base);
  }
};

namespace {

// Under the lock: admits requests in arrival order while their tables fit the
// budget (the first that does not fit holds the rest back, so a large request
// is not starved) and packs their rows into one tile of up to `tile` sources.
bool form_tile(xb_server* sv, std::vector<TileSeg>& segs, std::vector<int32_t>& src) {
  segs.clear();
  src.clear();
  int32_t cap = sv->tile;
  for (auto it = sv->pending.begin(); it != sv->pending.end() && cap > 0;) {
    ServeReq* r = *it;
    if (r->off < 0) {
      const int64_t off = sv->arena.alloc(r->bytes);
      if (off < 0) break;
      r->off = off;
      r->t_admit = sv->now_us();
    }
    const int32_t k = std::min(cap, r->S - r->next_row);
    segs.push_back(TileSeg{r, r->next_row, k});
    src.insert(src.end(), r->src.begin() + r->next_row, r->src.begin() + r->next_row + k);
    if (r->t_start < 0) r->t_start = sv->now_us();
    r->tiles++;
    r->next_row += k;
    cap -= k;
    if (r->next_row == r->S)
      it = sv->pending.erase(it);
    else
      ++it;
  }
  return !segs.empty();
}

// One tile on a slot's stream: the bit-parallel engine with the tile's rows
// mapped into the requests' tables (rows deeper than the level code take the
// plain engine, row by row), then the traversed-edge count per segment.
void run_serve_tile(xb_server* sv, ServeSlot* sl, const std::vector<TileSeg>& segs,
                    const std::vector<int32_t>& src, Profiler& prof, std::vector<int64_t>& edges) {
  const xb_graph* g = sv->g;
  // The slots' tiles read the graph's layout arrays; a rebuild freed them.
  if (g->layout_gen != sv->layout_gen)
    fail(XB_ERR_INVALID_ARG, "the served graph was re-tuned while the server was running");
  const int64_t n = g->n;
  const int32_t cnt = (int32_t)src.size();
  int32_t* base = reinterpret_cast<int32_t*>(sv->arena.base);
  int32_t r = 0;
  for (const TileSeg& sg : segs) {
    const int64_t lvl0 = sg.req->off / (int64_t)sizeof(int32_t);
    const int64_t par0 = lvl0 + (int64_t)sg.req->S * n;
    for (int32_t i = 0; i < sg.rows; ++i, ++r) {
      sl->h_rows[r] = lvl0 + (int64_t)(sg.row0 + i) * n;
      sl->h_rows[sv->tile + r] = par0 + (int64_t)(sg.row0 + i) * n;
    }
  }
  HIPCHK(hipMemcpyAsync(sl->d_rows, sl->h_rows, sizeof(int64_t) * 2 * (size_t)sv->tile, hipMemcpyHostToDevice,
                        sl->ws.st));
  BPOpts ot = sv->o;
  ot.T = tile_words(cnt, sv->o.T);
  const xb::RowMap map{sl->d_rows, sl->d_rows + sv->tile};
  if (!run_bp_tile(g, sl->ws, ot, src.data(), cnt, base, base, map, prof, nullptr)) {
    for (int32_t i = 0; i < cnt; ++i)
      run_plain(g, sl->ws, &src[(size_t)i], 1, base + sl->h_rows[i], base + sl->h_rows[sv->tile + i], true, prof,
                nullptr);
  }
  if (sv->count_edges) {
    HIPCHK(hipMemsetAsync(sl->d_edges, 0, sizeof(unsigned long long) * segs.size(), sl->ws.st));
    int32_t r0 = 0;
    for (size_t k = 0; k < segs.size(); ++k) {
      xb::launch_count_traversed(sl->ws.st, g->d_off, base + sl->h_rows[r0], segs[k].rows, (int32_t)n,
                                 sl->d_edges + k);
      r0 += segs[k].rows;
    }
    HIPCHK(hipMemcpyAsync(sl->h_edges, sl->d_edges, sizeof(unsigned long long) * segs.size(), hipMemcpyDeviceToHost,
                          sl->ws.st));
  }
  HIPCHK(hipStreamSynchronize(sl->ws.st));
  for (size_t k = 0; k < segs.size(); ++k) edges[k] = sv->count_edges ? (int64_t)sl->h_edges[k] : -1;
}

void serve_worker(xb_server* sv, ServeSlot* sl) {
  (void)hipSetDevice(sv->dev);
  Profiler prof;
  prof.st = sl->ws.st;
  std::vector<TileSeg> segs;
  std::vector<int32_t> src;
  std::vector<int64_t> edges;
  for (;;) {
    {
      std::unique_lock<std::mutex> lk(sv->mu);
      for (;;) {
        if (sv->stop) return;
        if (form_tile(sv, segs, src)) break;
        sv->cv_work.wait(lk);
      }
      sv->running++;
    }
    int32_t code = XB_OK;
    std::string msg;
    edges.assign(segs.size(), -1);
    try {
      run_serve_tile(sv, sl, segs, src, prof, edges);
    } catch (const XbError& e) {
      code = e.code;
      msg = e.msg;
    } catch (const std::exception& e) {
      code = XB_ERR_DEVICE;
      msg = e.what();
    }
    {
      std::lock_guard<std::mutex> lk(sv->mu);
      const double t = sv->now_us();
      sv->running--;
      sv->tiles++;
      sv->tile_rows += (int64_t)src.size();
      for (size_t k = 0; k < segs.size(); ++k) {
        ServeReq* r = segs[k].req;
        r->rows_done += segs[k].rows;
        if (code != XB_OK && r->err.empty()) {
          r->err = msg;
          r->status = code;
        }
        if (edges[k] >= 0) r->edges += edges[k];
        if (r->rows_done == r->S) {
          if (r->status == 1) {
            r->status = XB_OK;
            sv->completed++;
          } else {
            sv->failed++;
          }
          r->t_done = t;
        }
      }
    }
    sv->cv_done.notify_all();
  }
}

ServeReq* find_req(xb_server* sv, uint64_t id) {
  auto it = sv->reqs.find(id);
  if (it == sv->reqs.end()) fail(XB_ERR_INVALID_ARG, "unknown request id " + std::to_string(id));
  return it->second.get();
}

}  // namespace

extern "C" {

int32_t xb_server_opts_default(xb_server_opts* o) {
  return guarded([&] {
    if (!o) fail(XB_ERR_INVALID_ARG, "null opts");
    std::memset(o, 0, sizeof(*o));
    o->slots = 4;
    o->max_queued = 65536;
    o->budget_bytes = 0;
    o->flags = XB_SERVE_COUNT_EDGES;
  });
}

int32_t xb_server_create(xb_graph* g, const xb_server_opts* opts, const xb_msbfs_opts* engine_opts,
                         xb_server** out) {
  return guarded([&] {
    if (!g || !out) fail(XB_ERR_INVALID_ARG, "null graph or out handle");
    xb_server_opts so;
    xb_server_opts_default(&so);
    if (opts) {
      if (opts->slots < 0 || opts->max_queued < 0 || opts->budget_bytes < 0)
        fail(XB_ERR_INVALID_ARG, "negative server option");
      if (opts->slots) so.slots = opts->slots;
      if (opts->max_queued) so.max_queued = opts->max_queued;
      so.budget_bytes = opts->budget_bytes;
      so.flags = opts->flags;
    }
    if (so.slots > 64) fail(XB_ERR_INVALID_ARG, "at most 64 slots");
    const xb_msbfs_opts eo = parse_opts(engine_opts);
    if (eo.mode == XB_MODE_PLAIN) fail(XB_ERR_UNSUPPORTED, "the server runs the bit-parallel engine");
    std::unique_ptr<xb_server> sv(new xb_server());
    sv->g = g;
    sv->o = bp_opts(eo);
    sv->tile = 64 * sv->o.T;
    sv->count_edges = (so.flags & XB_SERVE_COUNT_EDGES) != 0;
    sv->max_queued = so.max_queued;
    HIPCHK(hipGetDevice(&sv->dev));
    ensure_layout(g, sv->o);
    sv->layout_gen = g->layout_gen;
    // Slots: every buffer a tile needs is allocated here (freeing device
    // memory synchronizes the device, which would stall the other slots).
    for (int i = 0; i < so.slots; ++i) {
      std::unique_ptr<ServeSlot> sl(new ServeSlot());
      HIPCHK(hipStreamCreateWithFlags(&sl->ws.st, hipStreamNonBlocking));
      HIPCHK(hipHostMalloc((void**)&sl->hctr, 64, hipHostMallocDefault));
      sl->ws.hctr = sl->hctr;
      ensure_queues(g, sl->ws);
      ensure_bp(g, sl->ws, sv->o.T);
      ensure_hpar(g, sl->ws, sv->o.T);
      ensure_src(sl->ws, sv->tile);
      ensure_q0_host(sl->ws, 3 * (int64_t)sv->tile);
      sl->d_rows = dalloc<int64_t>(2 * (size_t)sv->tile);
      sl->d_edges = dalloc<unsigned long long>((size_t)sv->tile);
      HIPCHK(hipHostMalloc((void**)&sl->h_rows, sizeof(int64_t) * 2 * (size_t)sv->tile, hipHostMallocDefault));
      HIPCHK(hipHostMalloc((void**)&sl->h_edges, sizeof(unsigned long long) * (size_t)sv->tile,
                           hipHostMallocDefault));
      sv->slots.push_back(std::move(sl));
    }
    int64_t budget = so.budget_bytes;
    if (budget == 0) {
      size_t fr = 0, tot = 0;
      HIPCHK(hipMemGetInfo(&fr, &tot));
      budget = (int64_t)(fr / 8 * 7);
    }
    if (budget < 256) fail(XB_ERR_BUDGET, "serving budget below 256 bytes");
    sv->arena.init(budget);
    sv->t0 = ServeClock::now();
    for (auto& sl : sv->slots) sl->th = std::thread(serve_worker, sv.get(), sl.get());
    *out = sv.release();
  });
}

int32_t xb_server_submit(xb_server* sv, const int32_t* sources, int32_t S, uint64_t* id) {
  return guarded([&] {
    if (!sv || !sources || !id) fail(XB_ERR_INVALID_ARG, "null argument");
    if (S < 1) fail(XB_ERR_INVALID_ARG, "a request needs at least one source");
    check_sources(sv->g, sources, S);
    const int64_t bytes = 8 * (int64_t)S * sv->g->n;
    if (Arena::round(bytes) > sv->arena.size)
      fail(XB_ERR_BUDGET, "request tables need " + std::to_string(bytes) + " bytes, the serving budget is " +
                              std::to_string(sv->arena.size));
    std::unique_ptr<ServeReq> r(new ServeReq());
    r->src.assign(sources, sources + S);
    r->S = S;
    r->bytes = bytes;
    {
      std::lock_guard<std::mutex> lk(sv->mu);
      if ((int64_t)sv->reqs.size() >= sv->max_queued) fail(XB_ERR_INVALID_ARG, "serving queue is full");
      r->id = sv->next_id++;
      r->t_submit = sv->now_us();
      *id = r->id;
      sv->pending.push_back(r.get());
      sv->reqs.emplace(r->id, std::move(r));
      sv->submitted++;
    }
    sv->cv_work.notify_one();
  });
}

int32_t xb_server_wait(xb_server* sv, uint64_t id, int32_t timeout_ms, xb_request_info* info) {
  return guarded([&] {
    if (!sv) fail(XB_ERR_INVALID_ARG, "null server");
    std::unique_lock<std::mutex> lk(sv->mu);
    ServeReq* r = find_req(sv, id);
    auto done = [&] { return r->status != 1; };
    if (timeout_ms < 0)
      sv->cv_done.wait(lk, done);
    else if (!sv->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), done))
      fail(XB_ERR_TIMEOUT, "request " + std::to_string(id) + " not done within " + std::to_string(timeout_ms) + " ms");
    if (info) {
      std::memset(info, 0, sizeof(*info));
      info->status = r->status;
      info->S = r->S;
      info->n = sv->g->n;
      int32_t* base = reinterpret_cast<int32_t*>(sv->arena.base);
      info->levels = r->off >= 0 ? base + r->off / (int64_t)sizeof(int32_t) : nullptr;
      info->parents = r->off >= 0 ? info->levels + (int64_t)r->S * sv->g->n : nullptr;
      info->edges = sv->count_edges ? r->edges : -1;
      info->t_submit_us = r->t_submit;
      info->t_admit_us = r->t_admit;
      info->t_start_us = r->t_start;
      info->t_done_us = r->t_done;
      info->tiles = r->tiles;
    }
    if (r->status != XB_OK) fail(r->status, "request " + std::to_string(id) + " failed: " + r->err);
  });
}

int32_t xb_server_release(xb_server* sv, uint64_t id) {
  return guarded([&] {
    if (!sv) fail(XB_ERR_INVALID_ARG, "null server");
    {
      std::unique_lock<std::mutex> lk(sv->mu);
      ServeReq* r = find_req(sv, id);
      sv->cv_done.wait(lk, [&] { return r->status != 1; });
      if (r->off >= 0) sv->arena.release(r->off, r->bytes);
      sv->reqs.erase(id);
      sv->released++;
    }
    sv->cv_work.notify_all();
  });
}

int32_t xb_server_stats(xb_server* sv, xb_server_counters* st) {
  return guarded([&] {
    if (!sv || !st) fail(XB_ERR_INVALID_ARG, "null argument");
    std::lock_guard<std::mutex> lk(sv->mu);
    std::memset(st, 0, sizeof(*st));
    st->submitted = sv->submitted;
    st->completed = sv->completed;
    st->failed = sv->failed;
    st->released = sv->released;
    st->tiles = sv->tiles;
    st->tile_rows = sv->tile_rows;
    st->budget_bytes = sv->arena.size;
    st->used_bytes = sv->arena.used;
    st->peak_used_bytes = sv->arena.peak;
    st->queued = (int32_t)sv->pending.size();
    st->running_tiles = sv->running;
  });
}

int32_t xb_server_destroy(xb_server* sv) {
  return guarded([&] {
    if (!sv) return;
    {
      std::lock_guard<std::mutex> lk(sv->mu);
      sv->stop = true;
    }
    sv->cv_work.notify_all();
    for (auto& sl : sv->slots)
      if (sl->th.joinable()) sl->th.join();
    delete sv;
  });
}

}  // extern "C"
