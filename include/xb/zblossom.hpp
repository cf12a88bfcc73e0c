// zblossom.hpp -- header-only C++ shim over the C ABI (xb_capi.h) that keeps
// the reference's engine interface and gains MultiSourceBFS():
//
//   zblossom::HGraph(row_offsets_path, col_indices_path)
//       apps/xb/src/include/graph/hgraph_gpu.h:37-48 (host CSR, move-only;
//       nnodes = |rowOffsets| - 1, nedges = |colIndices| / 2 as at :44-45)
//   zblossom::XBlossomGpuEngine10(const HGraph&[, size_t path_table_buffer_ratio])
//       apps/xb/src/include/xbsm_gpu/xblossom_gpu_engine_10.h:23-25 (the engine
//       owns the device copy of the graph, as DGraph does in
//       apps/xb/src/include/graph/dgraph.h:50-63), with FindMaximumMatch
//       (max_match_engine.h:84-113) and the statistics and timing getters of
//       xblossom_gpu_engine_10.h:39-95 (PrintTimingSummary, GetTimingSummary,
//       NumEdgesProcessed, NumAugmentingPathEdgesProcessed,
//       NumExpandEdgesProcessed, NumBlossomEdgesProcessed, NumBlossoms,
//       NumReusedTreeNodes, NumResetTreeNodes)
//   Matching, host_graph_t  the global aliases of
//       apps/xb/src/include/utils/mm_types_gpu.h:9,20
//
// A call site that uses these (the engine calls of
// apps/xb/runner/xblsm_gpu/run_xb_pp_by_dataset.cu:21-69) compiles against this
// header; tests/cpp/shim_test.cpp is such a call site.  The runner's glog,
// gflags and timing helpers (common/*_guard.h) are command-line scaffolding
// outside the engine and are not provided.
//
// Status codes become exceptions (zblossom::XbException), replacing the
// reference's CUDA_CHECK terminate (apps/xb/src/include/utils/utils_gpu.h:54-63).
#pragma once

#include <climits>
#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "xb_capi.h"

namespace zblossom {

class XbException : public std::runtime_error {
 public:
  XbException(int32_t code, const std::string& msg) : std::runtime_error(msg), code_(code) {}
  int32_t code() const { return code_; }

 private:
  int32_t code_;
};

inline void XbCheck(int32_t rc) {
  if (rc != XB_OK) {
    char buf[1024];
    xb_last_error(buf, sizeof(buf));
    throw XbException(rc, std::string("xb error ") + std::to_string(rc) + ": " + buf);
  }
}

// Reads whitespace-separated int32 values, the file format of the
// reference's ReadFileIntoHVector (apps/xb/src/include/common/utils_gpu.h:9-25),
// but reports a missing file instead of returning an empty vector.
inline std::vector<int32_t> ReadIntFile(const std::string& path) {
  std::vector<int32_t> out;
  std::FILE* f = std::fopen(path.c_str(), "r");
  if (!f) throw XbException(XB_ERR_IO, "cannot open " + path);
  long long v;
  while (std::fscanf(f, "%lld", &v) == 1) {
    if (v < INT32_MIN || v > INT32_MAX) {
      std::fclose(f);
      throw XbException(XB_ERR_IO, "integer out of range in " + path);
    }
    out.push_back((int32_t)v);
  }
  std::fclose(f);
  return out;
}

// Host CSR graph (int32 rowOffsets / colIndices).
// One entry per vertex: its mate (apps/xb/src/include/utils/mm_types_gpu.h:9).
using Matching = std::vector<int32_t>;

class HGraph {
 public:
  HGraph() = default;
  HGraph(const std::string& row_offsets_path, const std::string& col_indices_path, bool directed = false)
      : row_offsets_(ReadIntFile(row_offsets_path)), col_indices_(ReadIntFile(col_indices_path)),
        directed_(directed) {
    if (row_offsets_.empty()) throw XbException(XB_ERR_IO, "no row offsets in " + row_offsets_path);
    Bind();
  }
  HGraph(std::vector<int32_t> row_offsets, std::vector<int32_t> col_indices, bool directed = false)
      : row_offsets_(std::move(row_offsets)), col_indices_(std::move(col_indices)), directed_(directed) {
    if (row_offsets_.empty()) throw XbException(XB_ERR_INVALID_ARG, "empty row offsets");
    Bind();
  }
  HGraph(const HGraph&) = delete;
  HGraph& operator=(const HGraph&) = delete;
  HGraph(HGraph&& o) noexcept { *this = std::move(o); }
  HGraph& operator=(HGraph&& o) noexcept {
    row_offsets_ = std::move(o.row_offsets_);
    col_indices_ = std::move(o.col_indices_);
    directed_ = o.directed_;
    Bind();
    return *this;
  }

  size_t nnodes = 0, nedges = 0;
  const int32_t* row_offsets = nullptr;
  const int32_t* col_indices = nullptr;

  const std::vector<int32_t>& GetRowOffsetsVector() const { return row_offsets_; }
  const std::vector<int32_t>& GetColIndicesVector() const { return col_indices_; }
  bool directed() const { return directed_; }

 private:
  void Bind() {
    nnodes = row_offsets_.empty() ? 0 : row_offsets_.size() - 1;
    nedges = directed_ ? col_indices_.size() : col_indices_.size() / 2;
    row_offsets = row_offsets_.data();
    col_indices = col_indices_.data();
  }
  std::vector<int32_t> row_offsets_, col_indices_;
  bool directed_ = false;
};

// Engine bound to one device-resident copy of a graph.
class XBlossomGpuEngine10 {
 public:
  static std::string EngineName() { return "XBlossomGpuEngine10(MI355X: Reuse+ParBlossom+MinKey)"; }

  // The reference's one-argument form (xblossom_gpu_engine_10.h:23) sizes its
  // path table by a default ratio; no path table exists here.
  explicit XBlossomGpuEngine10(const HGraph& g) : XBlossomGpuEngine10(g, 0) {}
  XBlossomGpuEngine10(const HGraph& g, size_t path_table_buffer_ratio)
      : n_(g.nnodes), path_table_buffer_ratio_(path_table_buffer_ratio) {
    XbCheck(xb_graph_create_csr(g.row_offsets, (int64_t)g.GetRowOffsetsVector().size(), g.col_indices,
                                (int64_t)g.GetColIndicesVector().size(),
                                g.directed() ? XB_GRAPH_DIRECTED : XB_GRAPH_SYMMETRIC, &graph_));
    XbCheck(xb_opts_default(&opts_));
  }
  ~XBlossomGpuEngine10() {
    if (graph_) xb_graph_destroy(graph_);
  }
  XBlossomGpuEngine10(const XBlossomGpuEngine10&) = delete;
  XBlossomGpuEngine10& operator=(const XBlossomGpuEngine10&) = delete;

  xb_msbfs_opts& Options() { return opts_; }
  const xb_run_stats& LastStats() const { return stats_; }
  size_t PathTableBufferRatio() const { return path_table_buffer_ratio_; }

  // Host tables, S x n row-major (row s = sources[s]).
  void MultiSourceBFS(const std::vector<int32_t>& sources, std::vector<int32_t>& levels,
                      std::vector<int32_t>& parents) {
    const size_t cells = sources.size() * n_;
    levels.assign(cells, 0);
    parents.assign(cells, 0);
    xb_msbfs_opts o = opts_;
    o.flags &= ~XB_OUT_DEVICE;
    XbCheck(xb_msbfs_run(graph_, sources.data(), (int32_t)sources.size(), levels.data(), parents.data(), &o,
                         &stats_));
  }

  // Tunes words per tile, alpha and the pull cut-offs on a sample of sources
  // (xb_autotune); later searches use the tuned options.
  xb_tune_report Autotune(const std::vector<int32_t>& sample) {
    xb_tune_report rep{};
    xb_msbfs_opts best{};
    XbCheck(xb_autotune(graph_, sample.data(), (int32_t)sample.size(), &opts_, &best, &rep));
    best.flags = opts_.flags;
    best.mode = opts_.mode;
    opts_ = best;
    return rep;
  }

  // Device tables (caller-allocated, S x n int32 each).
  void MultiSourceBFSDevice(const int32_t* sources, int32_t S, int32_t* levels_dev, int32_t* parents_dev) {
    xb_msbfs_opts o = opts_;
    o.flags |= XB_OUT_DEVICE;
    XbCheck(xb_msbfs_run(graph_, sources, S, levels_dev, parents_dev, &o, &stats_));
  }

  // FindMaximumMatch of apps/xb/src/include/xbsm_gpu/max_match_engine.h:84-113
  // (used by apps/xb/runner/xblsm_gpu/run_xb_pp_by_dataset.cu:42-44): entry v
  // is v's mate, nnodes for an exposed vertex, as in the reference GPU engine.
  // The result is a maximum matching (blossoms included).
  Matching FindMaximumMatch() { return FindMaximumMatch(Matching(n_, (int32_t)n_)); }

  // The forest keeps the trees of still-exposed roots between phases, as the
  // reference engine does; RegrowForest(true) resets it every phase instead.
  Matching FindMaximumMatch(const Matching& initial_matching) {
    if (initial_matching.size() != n_)
      throw XbException(XB_ERR_INVALID_ARG, "initial matching must have nnodes entries");
    Matching mate(initial_matching);
    for (auto& m : mate)
      if (m == (int32_t)n_) m = -1;
    const uint32_t flags = XB_PROFILE | (regrow_ ? XB_MATCH_REGROW : 0u);
    XbCheck(xb_match_run(graph_, mate.data(), mate.data(), flags, &match_stats_));
    for (auto& m : mate)
      if (m < 0) m = (int32_t)n_;
    // Counters and times add up over the engine's calls, as the reference's
    // members do (xblossom_gpu_engine_10.h:99-116).
    const xb_match_stats& s = match_stats_;
    aug_edges_ += (size_t)s.aug_path_edges;
    expand_edges_ += (size_t)s.expand_edges;
    blossom_edges_ += (size_t)s.blossom_scan_edges;
    blossoms_ += (size_t)s.blossoms;
    reused_ += (size_t)s.reused_tree_vertices;
    reset_ += (size_t)s.reset_tree_vertices;
    t_aug_ms_ += s.kernel_ms[XB_K_MT_SCAN] + s.kernel_ms[XB_K_MT_TREE];
    t_expand_ms_ += s.kernel_ms[XB_K_MT_EXPAND];
    t_blossom_ms_ += s.kernel_ms[XB_K_MT_BLOSSOM];
    t_pre_ms_ += s.kernel_ms[XB_K_MT_INIT];
    return mate;
  }
  const xb_match_stats& LastMatchStats() const { return match_stats_; }
  void RegrowForest(bool on) { regrow_ = on; }

  // ---- statistics and timing of xblossom_gpu_engine_10.h:39-95 ----------
  // Times are device time per kernel class (HIP events), in SECONDS like the
  // reference's (which divides nanoseconds by 1e9 despite the _ms names):
  // augmenting path = scan, bridge, pairing and flip passes; expand = claim
  // and commit; blossom = candidate collection, sort and labelling;
  // preprocessing = the phase set-up (forest reset or reuse).
  void GetTimingSummary(double& augmenting_path_ms, double& expand_ms, double& blossom_ms,
                        double& total_ms) const {
    augmenting_path_ms = t_aug_ms_ / 1e3;
    expand_ms = t_expand_ms_ / 1e3;
    blossom_ms = t_blossom_ms_ / 1e3;
    total_ms = augmenting_path_ms + expand_ms + blossom_ms;
  }
  void PrintTimingSummary() const {
    double a = 0, e = 0, b = 0, t = 0;
    GetTimingSummary(a, e, b, t);
    std::printf("%s Timing Summary:\n", EngineName().c_str());
    std::printf("Augmenting Path: %g\n", a);
    std::printf("Expand: %g\n", e);
    std::printf("Blossom: %g\n", b);
    std::printf("Preprocessing Time: %g\n", t_pre_ms_ / 1e3);
    std::printf("Total Time: %g\n", t);
  }
  size_t NumEdgesProcessed() const { return aug_edges_ + expand_edges_ + blossom_edges_; }
  size_t NumAugmentingPathEdgesProcessed() const { return aug_edges_; }
  size_t NumExpandEdgesProcessed() const { return expand_edges_; }
  size_t NumBlossomEdgesProcessed() const { return blossom_edges_; }
  size_t NumBlossoms() const { return blossoms_; }
  size_t NumReusedTreeNodes() const { return reused_; }
  size_t NumResetTreeNodes() const { return reset_; }

  xb_graph* handle() const { return graph_; }

 private:
  xb_graph* graph_ = nullptr;
  size_t n_ = 0;
  size_t path_table_buffer_ratio_ = 0;
  xb_msbfs_opts opts_{};
  xb_run_stats stats_{};
  xb_match_stats match_stats_{};
  bool regrow_ = false;
  size_t aug_edges_ = 0, expand_edges_ = 0, blossom_edges_ = 0, blossoms_ = 0, reused_ = 0, reset_ = 0;
  double t_aug_ms_ = 0, t_expand_ms_ = 0, t_blossom_ms_ = 0, t_pre_ms_ = 0;
};

}  // namespace zblossom

// The reference's global aliases (apps/xb/src/include/utils/mm_types_gpu.h:9,20).
#ifndef XB_NO_GLOBAL_ALIASES
using Matching = zblossom::Matching;
using host_graph_t = zblossom::HGraph;
#endif
