// xb_internal.hpp -- shared declarations between the host engine
// (xb_host.cpp) and the HIP kernels (xb_kernels.hip).  Not installed.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace xb {

// The traversal works on engine ids: vertices relabelled by descending
// out-degree (ties by original id), so the frontier words of the hubs that
// most in-edges point at sit together at the low ids and stay in cache.
// Engine in-edge lists keep ascending ORIGINAL id order, and parents are
// recorded as original ids, so "first provider in list order" is still the
// lowest original id.  All tile state, the level-code planes included, is
// on engine ids.  The emit walks vertices in output (original) order over an
// original-order in-CSR whose entries are engine ids, so its in-list reads
// are sequential streams and its neighbour-plane reads hit the clustered
// hubs; a parent found there is translated back with orig_of.
//
// Bit-parallel tile state.  seen, front and nextw are vertex-major with T
// 64-bit source words per vertex ([v*T + t], engine ids).  Level codes
// (code = level + 1, 0 = unreached) are 8 bit-sliced planes, bit b of plane
// k = bit k of code(source 64t+b, v), in word groups of
// W = min(T, XB_PLANE_GROUP_MAX) words (2 by default) and two arrays:
//   plane k < 4:   planes[((t/W * n + v) * W + t%W) * 4 + k]
//   plane k >= 4:  planes_hi[((t/W * n + v) * W + t%W) * 4 + k - 4]
// so the first 4 planes of a vertex's W words of one group are 32W
// contiguous bytes (codes up to 15, i.e. levels <= 14): one gathered line per
// in-neighbour serves an emit pass over the group's W words, and that pass
// reads a single n*32W-byte region.  planes_hi serves only tiles deeper than
// 14 levels; the host engine allocates it when a tile first needs it.
struct BPState {
  int64_t n;            // vertices (plane addressing)
  uint64_t* seen;
  uint64_t* front[2];
  uint64_t* nextw;
  uint64_t* planes;
  uint64_t* planes_hi;  // planes 4-7 (null until a tile reaches level 15 on the host engine)
  uint32_t* tflag;
  int32_t* tlist;       // vertices whose next words were OR-ed atomically
  uint32_t* tcount;     // device count of tlist
  uint32_t* hpar;       // heavy-vertex parents (original ids) [(h*T + t)*64 + bit], 0xFFFFFFFF = unresolved
  const int32_t* hidx;  // vertex -> heavy index (in-degree above the medium pull cut), else -1
  const int32_t* new_of;    // original id -> engine id (degree order)
  const int32_t* orig_of;   // engine id -> original id
  const int32_t* in_col_o;  // original id of each in-edge's source, parallel to the engine's in_col
};

// Frontier queue: vertex ids plus the exclusive prefix of their out-degrees
// in queue order; ctr packs (edges << 32 | count) so one 64-bit atomic
// reserves both a queue slot range and an edge range.
struct Queue {
  int32_t* v;
  int32_t* eo;
  unsigned long long* ctr;
};

// A hub adjacency piece for the pull: in-edges [jb, je) of vertex v.
struct HubChunk {
  int32_t v, jb, je, pad;
};

// ---- persistent batch kernel ----
// Read-only view of one device graph and its in-degree layout.
struct GraphView {
  const int32_t* off;  // engine-id out-CSR (push)
  const int32_t* col;
  const int32_t* in_off;  // engine-id in-CSR, lists in ascending original id (pull, emit)
  const int32_t* in_col;
  const int32_t* in_col_o;  // original ids parallel to in_col
  const int32_t* new_of;    // original id -> engine id
  const int32_t* orig_of;   // engine id -> original id
  const int32_t* oin_off;   // in-CSR in original vertex and list order, entries as engine ids (emit)
  const int32_t* oin_col;
  const int32_t* oin_colo;  // the same entries as original ids
  const int32_t* perm;
  int32_t bucket_end[3];
  // Per bucket: its id list, or null when the bucket is the id range
  // [begin, end) + bucket_shift (degree-ordered ids of a symmetric graph).
  const int32_t* bucket_perm[3];
  int32_t bucket_shift[3];
  int32_t n_hub_chunks;
  const HubChunk* hubs;
  const int32_t* hidx;
  int32_t n_heavy;
  int32_t n;
  int32_t cut_medium;
  int32_t pad;
  int64_t m;
};

// One tile (up to 64*T sources) of one graph.
// Where the emit writes row r of a tile (element offsets from the level and
// parent base pointers).  Null tables: row r at r * n, the S x n layout of one
// run; a serving tile maps its rows into the tables of several requests.
struct RowMap {
  const int64_t* lvl_row = nullptr;
  const int64_t* par_row = nullptr;
};

struct TileJob {
  int32_t graph;           // index into the GraphView array
  int32_t cnt;             // sources in the tile
  const int32_t* src;      // device, cnt ids
  const int32_t* q0v;      // first frontier: distinct source vertices
  const int32_t* q0eo;     // and the exclusive prefix of their out-degrees
  int32_t qn0, E0;
  int32_t* lvl_out;        // rows of this tile (row stride n); may be null
  int32_t* par_out;
  int32_t status;          // out: 0 done, 1 deeper than the level code (re-run on the plain path)
  int32_t levels;          // out: deepest level reached
  int32_t words;           // 64-bit words per vertex of this tile (1, 2, 4, 8 or 16; 64 * words >= cnt)
  int32_t pad_;
};

static_assert(offsetof(TileJob, levels) == offsetof(TileJob, status) + sizeof(int32_t),
              "a captured emit reads {status, levels} as one pair");

// Team barrier counters live in 256-byte slots of their own: arrivals from
// thousands of workgroups serialize per address, so every XCD gets its own
// arrival counter and only the last arriver of each XCD touches the team's.
constexpr int kMaxXcd = 16;
struct alignas(256) XcdSlot {
  unsigned long long arrivals;  // monotonic: the team's workgroups on this XCD arrive once per barrier
  unsigned long long members;   // the team's workgroups on this XCD (counted at launch)
  unsigned long long pad_[30];
};

// Control block of one team (zeroed by a kernel at the start of every launch).
struct alignas(256) TeamCtl {
  unsigned long long done;       // XCD arrivals: barrier k of the team opens at (k+1) * (XCDs in the team)
  unsigned long long joined;     // workgroups registered at launch
  unsigned long long pad0_[30];
  XcdSlot xcd[kMaxXcd];
  unsigned long long ctr[3];  // packed (edges << 32 | count) queue counters, level d uses ctr[d % 3]
  uint32_t tcount[2];         // touched-list counts, level d uses tcount[d & 1]
  uint32_t emit_chunk;        // the light emit's shared (word group, chunk) counter, reset per tile
  uint32_t pad1_;
  // Phase clocks (wall_clock64 ticks) kept by workgroup 0 of the team: the
  // span between two team barriers is charged to the phase that ended at the
  // second one, indexed by the XB_K_* class of that phase.
  unsigned long long tlast;
  unsigned long long tphase[16];
};
static_assert(sizeof(TeamCtl) % 8 == 0, "TeamCtl is cleared as 64-bit words");

// A team: a contiguous range of workgroups running a contiguous range of jobs
// (the tiles of one graph) with its own workspace.
struct Team {
  int32_t first_block, nblocks;
  int32_t job_begin, job_end;
  // 1: the team runs one tile and leaves its emit to the standalone emit
  // launches that follow the persistent kernel in the captured graph (the
  // tile's planes stay in the team's workspace until then).
  int32_t emit_after;
  int32_t pad_;
  BPState s;
  int32_t* qv[2];
  int32_t* qeo[2];
  TeamCtl* ctl;
};

int persistent_blocks_per_cu();
void launch_bp_persistent(hipStream_t st, int grid_blocks, const Team* teams, int32_t n_teams,
                          const GraphView* graphs, TileJob* jobs, int32_t n_jobs, TeamCtl* ctl, int32_t alpha,
                          unsigned* gerr);

// ---- plain per-source engine ----
void launch_plain_init(hipStream_t st, int32_t* level, int32_t* parent, int32_t n);
void launch_plain_set_source(hipStream_t st, int32_t* level, int32_t* parent, int32_t src, Queue q);
void launch_plain_push(hipStream_t st, const int32_t* off, const int32_t* col, Queue cur, int32_t qn,
                       int32_t E, int32_t* level, int32_t* parent, int32_t d, Queue nxt);

// ---- batched bit-parallel engine (T = words per tile: 1, 2, 4, 8 or 16) ----
void launch_bp_seed(hipStream_t st, int T, BPState s, const int32_t* src, int32_t cnt);
void launch_bp_push(hipStream_t st, int T, BPState s, int cur, const int32_t* off, const int32_t* col,
                    Queue q, int32_t qn, int32_t E);
void launch_bp_commit(hipStream_t st, int T, BPState s, int cur, int32_t code, const int32_t* out_off,
                      Queue nxt, int32_t max_touched);
// cls 0: T lanes per vertex, 1: 4T lanes, 2: one wave per vertex.  perm null:
// vertex = index + shift.
void launch_bp_pull_bucket(hipStream_t st, int T, int cls, BPState s, int cur, int32_t code, int32_t cnt,
                           const int32_t* in_off, const int32_t* in_col, const int32_t* out_off,
                           const int32_t* perm, int32_t shift, int32_t begin, int32_t end, Queue nxt);
// One pull level in one launch: the three in-degree buckets (workgroup ranges
// of their own) and the hub pieces.  Bucket c covers [begin[c], end[c]) of
// perm[c] (null: the id range itself, shifted by shift[c]).
struct PullPlan {
  const int32_t* perm[3];
  int32_t shift[3], begin[3], end[3];
};
void launch_bp_pull_fused(hipStream_t st, int T, BPState s, int cur, int32_t code, int32_t cnt,
                          const int32_t* in_off, const int32_t* in_col, const int32_t* out_off, const PullPlan& plan,
                          Queue nxt, const HubChunk* chunks, int32_t n_chunks);
void launch_bp_pull_hub(hipStream_t st, int T, BPState s, int cur, int32_t cnt, const int32_t* in_col,
                        const HubChunk* chunks, int32_t n_chunks);
void launch_bp_clear(hipStream_t st, int T, BPState s, int cur, Queue q, int32_t qn);
// src: the tile's source ids (level-1 parents are the sources themselves).
// Light-vertex emit of word group `grp` (words grp*W .. grp*W+W-1, W =
// emit_pass_words(T)); symmetric: the graph's in- and out-lists are the same
// (a two-plane level test per in-neighbour).
int emit_pass_words(int T);
// gate (optional): a tile's {status, levels} pair in device memory; the
// launch then writes nothing unless status is 0 and gate_lo <= levels <=
// gate_hi (a captured emit whose tile depth is known only on the device).
void launch_bp_emit(hipStream_t st, int T, BPState s, int32_t grp, int32_t ngrp, int32_t K, int32_t rows,
                    const int32_t* src,
                    const int32_t* in_off, const int32_t* in_col, const int32_t* in_col_orig, int32_t n,
                    int32_t heavy_cut, int32_t* lvl_out, int32_t* par_out, RowMap map,
                    bool symmetric, const int32_t* gate = nullptr, int32_t gate_lo = 0, int32_t gate_hi = 0);

// ---- forest mode (alternating-forest search of maximum matching) ----
// Pool of pending second calls of the path flip's recursion (x, y, next
// entry); ctr counts the entries taken this round.
struct FlipStack {
  int4* e;
  uint32_t* ctr;
  uint32_t cap;
};
void launch_mt_init(hipStream_t st, int32_t n, const int32_t* mate, const int32_t* off, int32_t* root, int32_t* lvl,
                    int32_t* par, unsigned long long* claim, int32_t* best, int32_t* acc, unsigned long long* bridge,
                    uint8_t* used, int32_t* cursor, int32_t* fp, unsigned long long* elab, Queue q);
// Phase set-up with tree reuse: resets the trees of augmented roots, queues
// the outer vertices of the kept trees (q), counts reused / reset vertices
// (counts[0], counts[1]); then the pairing state of every vertex is cleared.
void launch_mt_reuse_init(hipStream_t st, int32_t n, const int32_t* mate, const int32_t* off, int32_t* root,
                          int32_t* lvl, int32_t* par, int32_t* fp, unsigned long long* elab, unsigned long long* claim,
                          const uint8_t* used, Queue q, unsigned long long* counts);
void launch_mt_reuse_clear(hipStream_t st, int32_t n, int32_t* best, int32_t* acc, unsigned long long* bridge,
                           uint8_t* used);
// Level 0: each root's lowest neighbouring untaken root, by a forward-only
// cursor over its sorted adjacency; adds the entries read to `examined`.
void launch_mt_scan0(hipStream_t st, const int32_t* roots, int32_t nroots, const int32_t* off, const int32_t* col,
                     const int32_t* lvl, const uint8_t* used, int32_t* cursor, int32_t* best,
                     unsigned long long* examined);
void launch_mt_scan(hipStream_t st, const int32_t* off, const int32_t* col, Queue q, int32_t qn, int32_t E,
                    const int32_t* root, const int32_t* lvl, const uint8_t* used, int32_t* best,
                    unsigned long long* blossoms);
void launch_mt_accept(hipStream_t st, const int32_t* roots, int32_t nroots, const int32_t* best, int32_t* acc,
                      int32_t* pairs, uint32_t* npairs, unsigned long long* bridge0);
void launch_mt_bridge(hipStream_t st, const int32_t* off, const int32_t* col, Queue q, int32_t qn, int32_t E,
                      const int32_t* root, const int32_t* lvl, const int32_t* acc, unsigned long long* bridge);
void launch_mt_flip(hipStream_t st, const int32_t* pairs, const uint32_t* npairs, int32_t max_pairs,
                    const int32_t* acc, const unsigned long long* bridge, const int32_t* par,
                    const unsigned long long* elab, int32_t* mate, uint8_t* used, FlipStack stk, int32_t max_steps,
                    unsigned* err);
// Blossom step: candidates (same-tree even-even frontier edges), sorted by
// (root, edge position), then marked and applied.
void launch_mt_collect(hipStream_t st, const int32_t* off, const int32_t* col, Queue q, int32_t qn, int32_t E,
                       const int32_t* root, const int32_t* lvl, unsigned long long* cand, uint32_t* ncand,
                       uint32_t cap, unsigned* err);
// Marks each candidate's blossom path (lowest sorted candidate per inner
// vertex), then makes the marked vertices outer and appends them to nxt.
void launch_mt_blossom(hipStream_t st, const int32_t* off, const int32_t* col, int32_t n, const int32_t* par,
                       int32_t* lvl, int32_t* fp, unsigned long long* elab, unsigned long long* lab,
                       const unsigned long long* cand, int32_t C, int32_t* join_of, int32_t* marked,
                       uint32_t* nmarked, uint32_t* formed, Queue nxt, unsigned long long* counts, unsigned* err);
// Ascending sort of 64-bit keys (device radix sort); temp_bytes(n) sizes the scratch.
size_t sort_u64_temp_bytes(int32_t n);
void sort_u64(void* temp, size_t bytes, const unsigned long long* in, unsigned long long* out, int32_t n,
              hipStream_t st);
void launch_mt_reset_best(hipStream_t st, const int32_t* roots, int32_t nroots, int32_t* best);
// One launch zeroing a round's four counters.
void launch_mt_zero_ctrs(hipStream_t st, uint32_t* a, uint32_t* b, uint32_t* c, unsigned long long* d);
// Keeps the entries of `in` whose tree is not taken: into the queue `out`
// (list == nullptr) or into list/ctr32.
void launch_mt_compact(hipStream_t st, const int32_t* in, int32_t cnt, const int32_t* root, const uint8_t* used,
                       const int32_t* off, Queue out, int32_t* list, uint32_t* ctr32);
void launch_mt_claim(hipStream_t st, const int32_t* off, const int32_t* col, Queue q, int32_t qn, int32_t E,
                     const int32_t* root, const int32_t* lvl, unsigned long long* claim, int32_t* tlist,
                     uint32_t* tcount);
void launch_mt_commit(hipStream_t st, const int32_t* tlist, const uint32_t* tcount, int32_t max_t,
                      const int32_t* mate, const int32_t* off, const unsigned long long* claim, int32_t* root,
                      int32_t* lvl, int32_t* par, int32_t* fp, Queue nxt);
void launch_mt_validate(hipStream_t st, int32_t n, const int32_t* mate, const int32_t* off, const int32_t* col,
                        unsigned long long* bad);
void launch_mt_count(hipStream_t st, int32_t n, const int32_t* mate, unsigned long long* out);

// ---- utilities ----
void launch_count_traversed(hipStream_t st, const int32_t* off, const int32_t* levels, int32_t S, int32_t n,
                            unsigned long long* out);

void launch_checksum(hipStream_t st, const int32_t* a, int64_t count, unsigned long long* out);

// ---- test hooks ----
void launch_test_build_queue(hipStream_t st, const int32_t* deg, int32_t k, Queue q);
void launch_test_edge_map(hipStream_t st, Queue q, int32_t qn, int32_t E, int32_t* owner, int32_t* offset);

}  // namespace xb
