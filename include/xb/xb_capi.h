/*
 * xb_capi.h -- C ABI of the MI355X multi-source BFS engine (libxbmsbfs.so).
 *
 * Every entry point returns an int32_t status (XB_OK == 0).  The message of
 * the last failure on the calling thread is available from xb_last_error().
 * Callers own every input and output buffer; the library owns the device copy
 * of the graph and its scratch memory.  A graph handle may be used by one
 * host thread at a time; distinct handles may run concurrently (each has its
 * own HIP stream).
 *
 * Reference interfaces replaced (paths relative to the reference repository):
 *   xb_graph_create_csr / xb_graph_load_text
 *       zblossom::HGraph(row_path, col_path)
 *         apps/xb/src/include/graph/hgraph_gpu.h:37-48
 *       ReadFileIntoHVector  apps/xb/src/include/common/utils_gpu.h:9-25
 *       zblossom::DGraph(const HGraph&)  apps/xb/src/include/graph/dgraph.h:50-63
 *   xb_msbfs_run
 *       the search step of XBlossomGpuEngine10: source seeding
 *         apps/xb/src/xblsm_gpu/xblossom_gpu_engine_10.cu:83-185,
 *       the level loop :72-81 and the expand step :270-359, with the path
 *       table write of apps/xb/src/include/path_table/path_table.h:256-271
 *       turned into per-source level/parent tables.
 *   xb_graph_destroy   ~HGraph / ~DGraph (RAII in the reference)
 *   xb_last_error      replaces CUDA_CHECK's terminate
 *                      (apps/xb/src/include/utils/utils_gpu.h:54-63)
 */
#ifndef XB_CAPI_H
#define XB_CAPI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------- */
#define XB_OK 0
#define XB_ERR_INVALID_ARG (-1)   /* null pointer, negative size, bad option */
#define XB_ERR_BAD_CSR (-2)       /* non-monotone offsets, ids out of range   */
#define XB_ERR_IO (-3)            /* unreadable or malformed text file        */
#define XB_ERR_NOMEM (-4)         /* host or device allocation failed         */
#define XB_ERR_DEVICE (-5)        /* HIP runtime error                        */
#define XB_ERR_SOURCE_RANGE (-6)  /* a source id is outside [0, n)            */
#define XB_ERR_BUDGET (-7)        /* request exceeds the HBM admission budget */
#define XB_ERR_UNSUPPORTED (-8)   /* option combination not supported         */
#define XB_ERR_NO_DEVICE (-9)     /* no HIP device visible                    */
#define XB_ERR_TIMEOUT (-10)      /* xb_server_wait: not done within timeout  */

/* ---- graph flags ------------------------------------------------------ */
#define XB_GRAPH_SYMMETRIC 0u /* undirected: the CSR holds both directions */
#define XB_GRAPH_DIRECTED 1u  /* directed: the transpose is built on load  */

/* ---- run modes -------------------------------------------------------- */
#define XB_MODE_AUTO 0    /* batched bit-parallel engine (falls back to plain) */
#define XB_MODE_PLAIN 1   /* one top-down push BFS launch sequence per source  */
#define XB_MODE_BATCHED 2 /* bit-parallel, 64 sources per 64-bit word          */

/* ---- run flags -------------------------------------------------------- */
#define XB_OUT_DEVICE 1u  /* levels/parents are device pointers (else host)  */
#define XB_PROFILE 2u     /* time every kernel class with HIP events         */
#define XB_NO_LEVELS 4u   /* levels == NULL allowed: only parents written     */
#define XB_FIXED_TEAMS 8u /* xb_batch_*: keep the planned team sizes (no
                             re-split of the grid from each run's clocks)    */
#define XB_MATCH_REGROW 16u /* xb_match_run: regrow the whole forest every
                               phase instead of keeping the trees of roots
                               that are still exposed                       */
#define XB_NO_PIPELINE 32u /* xb_msbfs_run: run the tiles of a multi-tile
                              batched run one after another on one stream
                              (default with XB_OUT_DEVICE: tiles alternate
                              between two workspaces and streams, so a
                              tile's traversal overlaps the previous tile's
                              emit; costs a second tile workspace)          */
#define XB_BATCH_EMIT_INSIDE 64u /* xb_batch_*: every team emits its tiles
                                    inside the persistent kernel (default: a
                                    team of one tile leaves its emit to
                                    standalone launches captured after the
                                    kernel)                                   */

typedef struct xb_graph xb_graph;

/* Tunables of the batched engine.  Zero means "use the default".  Pull
 * degree buckets (by in-degree): deg <= pull_cut_small gets T lanes per
 * vertex (one per source word), deg <= pull_cut_medium 4T lanes, deg <=
 * pull_cut_hub a whole wave, larger ones are split into hub_chunk-edge pieces
 * with one wave each.  Vertices above pull_cut_medium are "heavy": the pull
 * and the push record their parents as they find them (lowest original id),
 * and the emit copies them out instead of rescanning long in-lists. */
typedef struct xb_msbfs_opts {
  int32_t mode;            /* XB_MODE_*                                        */
  uint32_t flags;          /* XB_OUT_DEVICE | XB_PROFILE | ...                 */
  int32_t words_per_tile;  /* 64-bit source words traversed together: 1,2,4,8,16 */
  int32_t alpha;           /* pull when |F| + outdeg(F) > m / alpha            */
  int32_t pull_cut_small;
  int32_t pull_cut_medium;
  int32_t pull_cut_hub;
  int32_t hub_chunk;
  int32_t reserved[8];
} xb_msbfs_opts;

/* Kernel classes timed by XB_PROFILE (index into the xb_run_stats arrays). */
#define XB_K_MEMSET 0     /* tile state clears (hipMemsetAsync)                */
#define XB_K_SEED 1       /* source seeding                                     */
#define XB_K_PUSH 2       /* bit-parallel top-down push                         */
#define XB_K_COMMIT 3     /* commit of pushed / hub-pulled next words            */
#define XB_K_PULL 4       /* bit-parallel bottom-up pull: degree buckets and, in  */
                          /* the one-launch level (default), hub chunks too       */
#define XB_K_PULL_HUB 5   /* bottom-up pull, hub adjacency chunks (when a level  */
                          /* launches buckets and hub chunks separately)         */
#define XB_K_CLEAR 6      /* frontier clear                                     */
#define XB_K_EMIT_HEAVY 7 /* reserved: heavy parents are recorded by pull/push */
#define XB_K_EMIT 8       /* level/parent table write-out                       */
#define XB_K_PLAIN_INIT 9 /* plain engine: row initialisation                   */
#define XB_K_PLAIN_PUSH 10 /* plain engine: top-down push level                 */
#define XB_K_PERSISTENT 11 /* persistent batch kernel (whole batch, one launch) */
#define XB_K_MT_SCAN 12    /* matching: augmenting-edge scan and bridge passes  */
#define XB_K_MT_EXPAND 13  /* matching: claim and commit (forest growth)        */
#define XB_K_MT_TREE 14    /* matching: tree pairing, path flips, compaction    */
#define XB_K_MT_BLOSSOM 15 /* matching: blossom labelling                       */
#define XB_K_MT_INIT 16    /* matching: phase set-up (forest reset or reuse)   */
#define XB_K_CLASSES 17

/* Per-run statistics (filled when the pointer is non-null).  kernel_ms and
 * kernel_launches are filled only with XB_PROFILE; kernel_bytes always holds
 * the algorithmic bytes of each class (the bytes a launch must move at
 * minimum: CSR and state streams plus output, without redundant re-reads).
 * For a batch (xb_batch_run) kernel_ms[XB_K_PERSISTENT] is the launch's
 * HIP-event time; with XB_PROFILE the other classes hold the phase clocks of
 * the persistent kernel (each team's time between barriers, weighted by its
 * share of the grid, so they add up to about the launch time). */
typedef struct xb_run_stats {
  double ms_total;                     /* HIP-event time of the whole run    */
  double kernel_ms[XB_K_CLASSES];      /* summed HIP-event time per class    */
  double kernel_bytes[XB_K_CLASSES];   /* summed algorithmic bytes per class */
  int64_t kernel_launches[XB_K_CLASSES];
  int32_t max_level;        /* deepest level reached over all sources        */
  int32_t tiles;            /* number of source tiles run                    */
  int32_t push_levels;
  int32_t pull_levels;
  int32_t fallback_sources; /* sources re-run on the plain path (depth > 254) */
  int32_t reserved[11];
} xb_run_stats;

/* Fills *o with the defaults. */
int32_t xb_opts_default(xb_msbfs_opts* o);

/* Copies a CSR (row_offsets int32[n+1], col_indices int32[m]) to the device.
 * Validates monotone offsets, row_offsets[0] == 0, row_offsets[n] == m and
 * ids in [0, n).  Adjacency lists are sorted ascending on ingest if needed
 * (the lowest-id parent rule relies on it); duplicate edges are kept. */
int32_t xb_graph_create_csr(const int32_t* row_offsets, int64_t n_plus_1,
                            const int32_t* col_indices, int64_t m,
                            uint32_t flags, xb_graph** out);

/* Reads the two whitespace-separated integer files the reference's
 * ReadFileIntoHVector reads, then behaves as xb_graph_create_csr. */
int32_t xb_graph_load_text(const char* row_offsets_path,
                           const char* col_indices_path, uint32_t flags,
                           xb_graph** out);

/* Binary CSR files: a 40-byte header (magic "XBCSR01", graph flags, n+1, m,
 * a 64-bit FNV-1a checksum over the 32-bit words of both arrays), then the
 * int32 row offsets and column indices, little-endian.  No counterpart in
 * the reference, which only reads the text pair (ReadFileIntoHVector,
 * apps/xb/src/include/common/utils_gpu.h:9-25) and converts datasets with
 * scripts (scripts/2-datasets/...); loading a binary file skips the text
 * parse.  The xb_csr_* calls are host-only (no device needed). */
int32_t xb_csr_save(const char* path, const int32_t* row_offsets, int64_t n_plus_1,
                    const int32_t* col_indices, int64_t m, uint32_t flags);
int32_t xb_csr_load_header(const char* path, int64_t* n_plus_1, int64_t* m, uint32_t* flags);
/* Buffers must hold exactly the header's n_plus_1 and m entries; XB_ERR_IO
 * on a truncated file or checksum mismatch. */
int32_t xb_csr_load(const char* path, int32_t* row_offsets, int64_t n_plus_1,
                    int32_t* col_indices, int64_t m);
/* xb_csr_load + xb_graph_create_csr with the file's flags. */
int32_t xb_graph_load_binary(const char* path, xb_graph** out);
/* Writes the graph's CSR (adjacency lists as stored: sorted on ingest). */
int32_t xb_graph_save_binary(const xb_graph* g, const char* path);

int32_t xb_graph_info(const xb_graph* g, int64_t* n, int64_t* m,
                      uint32_t* flags);

int32_t xb_graph_destroy(xb_graph* g);

/* Runs S single-source BFS traversals.  levels and parents are S x n int32,
 * row s = sources[s]:
 *   levels[s][v]  = hop distance, -1 if unreached;
 *   parents[s][v] = lowest-id u with (u->v) in E and levels[s][u] ==
 *                   levels[s][v]-1; parents[s][sources[s]] = sources[s];
 *                   -1 if unreached.
 * Results are identical for every mode and option setting. */
int32_t xb_msbfs_run(xb_graph* g, const int32_t* sources, int32_t S,
                     int32_t* levels, int32_t* parents,
                     const xb_msbfs_opts* opts, xb_run_stats* stats);

/* ---- allocation-free run -----------------------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory: serving loops with a fixed memory footprint and harnesses that
 * forbid hipMalloc.  Same tables as xb_msbfs_run with XB_OUT_DEVICE (levels
 * and parents are caller-owned device S x n int32 arrays, both required),
 * but the call makes no device or pinned allocation and frees nothing: its
 * queues and tile state live in a static device array of
 * XB_NOALLOC_SCRATCH_BYTES per GPU, which each GPU the library runs on holds
 * until the process ends.  The batched engine takes the widest tile (at most
 * opts->words_per_tile words) whose state fits that array.  Limits:
 *   - XB_ERR_BUDGET when the graph is too large: the plain engine needs
 *     about 16 bytes per vertex, a 64-source bit-parallel tile about
 *     128 bytes per vertex (XB_MODE_AUTO then runs the plain engine);
 *   - XB_ERR_UNSUPPORTED for XB_NO_LEVELS or XB_PROFILE, and when the graph's
 *     pull layout was built for other bucket cut-offs than opts' (a graph is
 *     created with the default cut-offs; xb_msbfs_run with other cut-offs
 *     rebuilds the layout, which allocates).
 * Calls on one GPU share its array and are serialized by a lock inside the
 * library; every call clears the state it reads, so nothing carries over from
 * the previous call. */
#define XB_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
int32_t xb_msbfs_run_noalloc(xb_graph* g, const int32_t* sources, int32_t S,
                             int32_t* levels_dev, int32_t* parents_dev,
                             const xb_msbfs_opts* opts, xb_run_stats* stats);

/* ---- batched multi-graph runs: one hipGraph-captured replay -------------
 *
 * A batch is G graphs, each with its own source set and device output
 * tables.  xb_batch_create plans it once: every graph gets a team of
 * workgroups sized by its estimated cost, its source tiles (64*T sources
 * each) run on that team with the level loop on the device, and the whole
 * batch is one persistent kernel launch captured in a hipGraph.  A graph
 * whose sources fit one tile has its tables written by emit launches
 * captured after the kernel in the same graph (XB_BATCH_EMIT_INSIDE keeps
 * every emit inside the kernel).
 * xb_batch_run replays the graph; tiles deeper than the 8-bit level code are
 * re-run on the plain path afterwards.  After each run the workgroups are
 * re-split between the teams from their measured spans (unless the options
 * carry XB_FIXED_TEAMS), so the first few replays get faster; the split never
 * changes results.  The graphs must stay alive while the batch exists.  A
 * run or autotune on one of them with other bucket cut-offs rebuilds its
 * layout arrays, which the captured launch reads: after any rebuild (even one
 * that restores the original cut-offs) xb_batch_run returns
 * XB_ERR_INVALID_ARG without launching; create a new batch.  Results are
 * those of xb_msbfs_run on each graph. */
typedef struct xb_batch xb_batch;

int32_t xb_batch_create(xb_graph* const* graphs, int32_t G,
                        const int32_t* const* sources, const int32_t* S,
                        int32_t* const* levels_dev, int32_t* const* parents_dev,
                        const xb_msbfs_opts* opts, xb_batch** out);
int32_t xb_batch_run(xb_batch* b, xb_run_stats* stats);
int32_t xb_batch_destroy(xb_batch* b);

/* Convenience: create + run + destroy.  Outputs are host arrays unless
 * opts->flags has XB_OUT_DEVICE. */
int32_t xb_msbfs_run_batched(xb_graph* const* graphs, int32_t G,
                             const int32_t* const* sources, const int32_t* S,
                             int32_t* const* levels, int32_t* const* parents,
                             const xb_msbfs_opts* opts, xb_run_stats* stats);

/* Sum over sources of the out-degrees of the vertices each source reaches
 * (the traversed-edge count of a full top-down BFS), computed on the device
 * from a levels table already in device memory. */
int32_t xb_count_traversed(xb_graph* g, const int32_t* levels_dev, int32_t S,
                           int64_t* edges_out);

/* Test hook (mirrors apps/xb/test/device/scan_bs_test.cu:58-137, which tests
 * the reference's exclusive scan + FindRightmostLTEQ mapping): appends
 * (i, degrees[i]) for i < k through the device frontier compaction, then maps
 * every tasklet t < sum(degrees) to (owner[t], offset[t]) with the device
 * owner search the push kernels use.  Outputs are host arrays of
 * sum(degrees) entries; tasklet order follows the (unordered) device queue,
 * so the pairs, not their positions, are the result. */
int32_t xb_test_edge_map(const int32_t* degrees, int32_t k, int32_t* owner_out,
                         int32_t* offset_out, int64_t* total_out);

/* Order-independent checksum of an int32 device array (sum over i of a 64-bit
 * mix of (i, a[i])): equal tables give equal sums; used to compare S x n
 * tables without copying them to the host. */
int32_t xb_device_checksum(const int32_t* a_dev, int64_t count, uint64_t* out);

/* Device memory helpers for callers without their own allocator. */
int32_t xb_device_alloc(size_t bytes, void** ptr);
int32_t xb_device_free(void* ptr);
int32_t xb_memcpy_d2h(void* dst, const void* src, size_t bytes);
int32_t xb_memcpy_h2d(void* dst, const void* src, size_t bytes);

/* Copies the last error message of the calling thread into buf. */
int32_t xb_last_error(char* buf, size_t len);

/* ---- forest mode: maximum matching -------------------------------------
 * Replaces the phase loop of XBlossomGpuEngine10::FindMaximumMatch
 * (apps/xb/src/include/xbsm_gpu/max_match_engine.h:93-113) with its
 * alternating-forest search (apps/xb/src/xblsm_gpu/xblossom_gpu_engine_10.cu:
 * 83-359: InitAlternatingForest, FindAndFlipAugmentingPathInAlternatingForest,
 * ExpandAlternatingForest, TransformOddNodesInBlossom) and path flip
 * (path_table.h:273-330).  Each phase grows a forest from every exposed
 * vertex; an even-even edge inside a tree forms a blossom whose odd vertices
 * become even; an even-even edge between two trees closes an augmenting path.
 * A phase flips a maximal set of tree-disjoint augmenting paths; phases repeat
 * until one finds none, and the result is a maximum matching on every graph.
 * Between phases the trees whose root is still exposed are kept (the
 * reference's lazy reset, xblossom_gpu_engine_10.cu:119-175) and the next
 * phase grows from all their outer vertices; XB_MATCH_REGROW resets the
 * whole forest every phase instead.
 * All choices are deterministic (lowest root id, then lowest vertex or edge
 * position), so a run's matching is reproducible.  Symmetric graphs only;
 * unmatched vertices are -1. */
typedef struct xb_match_stats {
  int64_t matched_vertices;  /* 2 x the matching's cardinality               */
  int64_t augmentations;     /* paths flipped over all phases                */
  int64_t blossom_edges;     /* same-tree even-even edges met in the final phase (informational) */
  int64_t edges_scanned;     /* frontier edges read by the scan, bridge, claim and blossom passes */
  int64_t blossoms;          /* blossoms formed over the run                  */
  int64_t blossom_vertices;  /* odd vertices made even by them               */
  int32_t phases;
  int32_t levels;            /* forest levels grown over all phases          */
  double ms_total;
  /* Per kernel class (XB_K_MT_*): HIP-event time with XB_PROFILE, algorithmic
   * bytes and launches always. */
  double kernel_ms[XB_K_CLASSES];
  double kernel_bytes[XB_K_CLASSES];
  int64_t kernel_launches[XB_K_CLASSES];
  /* edges_scanned split as the reference engine splits its counters
   * (num_aug_path_edges_processed_, num_expand_edges_processed_,
   * num_blossom_edges_processed_; xblossom_gpu_engine_10.cu:220-221,
   * 303-304, 400-401). */
  int64_t aug_path_edges;     /* level-0 cursor scans, augmenting-edge scans, bridge passes */
  int64_t expand_edges;       /* claim passes                                 */
  int64_t blossom_scan_edges; /* blossom-candidate passes                     */
  /* Per phase set-up, summed over phases (the reference's
   * num_reused_tree_nodes_ / num_reset_tree_nodes_): matched vertices kept
   * in a tree whose root is still exposed, and every other vertex. */
  int64_t reused_tree_vertices;
  int64_t reset_tree_vertices;
  int32_t reserved[8];
} xb_match_stats;

/* init_mate: NULL (empty matching) or n entries (a valid matching, -1 for
 * exposed); mate_out: n entries.  flags: XB_OUT_DEVICE (both are device
 * pointers) | XB_PROFILE (per-class kernel times) | XB_MATCH_REGROW.  An invalid init_mate (out of range, asymmetric, or a pair that
 * is not an edge) is XB_ERR_INVALID_ARG. */
int32_t xb_match_run(xb_graph* g, const int32_t* init_mate, int32_t* mate_out, uint32_t flags,
                     xb_match_stats* stats);

/* ---- autotuner ---------------------------------------------------------
 * No counterpart in the reference, whose thresholds are hand-set constants
 * (the |E|/20 push/pull switch of apps/ligra/ligra/ligra.h:236-277 and the
 * XB engine's fixed launch shapes).  Coordinate descent over words per
 * tile (1, 2, 4, 8, 16), the push/pull switch alpha and the pull bucket
 * cut-offs: each trial runs the caller's sample sources (at most 1,024 of
 * them) on the device as xb_msbfs_run would, narrower tiles as several tiles
 * on two streams, and is scored in milliseconds per source (best of two
 * runs); a value replaces the current one only if it is at least 2% faster.
 * The trials write into two scratch
 * tables of min(S, 1024) x n int32 (fewer rows if those would exceed half
 * the free device memory) that are freed on return; the sample's results
 * are discarded.  *best receives the tuned options (mode
 * XB_MODE_BATCHED). */
typedef struct xb_tune_trial {
  int32_t words_per_tile, alpha, pull_cut_small, pull_cut_medium, pull_cut_hub, hub_chunk;
  double ms_per_source;
} xb_tune_trial;

#define XB_TUNE_MAX_TRIALS 128
typedef struct xb_tune_report {
  int32_t trials;                         /* entries of trial[] filled      */
  int32_t passes;
  double default_ms_per_source;           /* the base options' score        */
  double best_ms_per_source;
  xb_tune_trial trial[XB_TUNE_MAX_TRIALS];
} xb_tune_report;

int32_t xb_autotune(xb_graph* g, const int32_t* sample, int32_t S, const xb_msbfs_opts* base,
                    xb_msbfs_opts* best, xb_tune_report* report);

/* ---- serving: continuous batching over one device-resident graph -------
 * No counterpart in the reference, whose drivers run one dataset and one
 * source list per process (apps/xb/src/run_xb_pp_by_dataset.cu); these
 * entry points serve multi-source BFS queries against a resident graph.
 *
 * Requests (source sets) are queued on submit and admitted in arrival order
 * as their level/parent tables (S x n int32 each, 8*S*n bytes) fit in the
 * server's device budget.  Worker slots (one HIP stream each) pack the rows
 * of admitted requests into bit-parallel tiles of up to 64*words_per_tile
 * sources -- rows of several requests share a tile, a request larger than a
 * tile spans several -- so small queries share traversals and tiles on
 * different slots overlap on the device.  A row's result does not depend on
 * the tile it ran in: every table equals the one xb_msbfs_run writes for
 * the same sources.  The server owns the graph's layout while it lives: do
 * not run other work on the same graph handle meanwhile (a rebuild of the
 * layout by a run with other cut-offs makes every later tile fail with
 * XB_ERR_INVALID_ARG). */
typedef struct xb_server xb_server;

#define XB_SERVE_COUNT_EDGES 1u /* count traversed edges per request (GTEPS) */

typedef struct xb_server_opts {
  int32_t slots;          /* worker slots (streams), default 4                 */
  int32_t max_queued;     /* submitted and unreleased requests, default 65536  */
  int64_t budget_bytes;   /* device bytes for request tables; 0 = 7/8 of free  */
  uint32_t flags;         /* XB_SERVE_COUNT_EDGES (default on)                 */
  int32_t reserved[7];
} xb_server_opts;

typedef struct xb_request_info {
  int32_t status;         /* XB_OK once done, else the failure's status code   */
  int32_t S;              /* sources (rows)                                    */
  int64_t n;              /* vertices (columns)                                */
  int32_t* levels;        /* device S x n, valid until xb_server_release       */
  int32_t* parents;       /* device S x n                                      */
  int64_t edges;          /* traversed edges (XB_SERVE_COUNT_EDGES), else -1   */
  double t_submit_us;     /* times since the server was created                */
  double t_admit_us;      /* tables allocated                                  */
  double t_start_us;      /* first tile carrying it started                    */
  double t_done_us;       /* last tile carrying it finished                    */
  int32_t tiles;          /* tiles that carried its rows                       */
  int32_t reserved[7];
} xb_request_info;

typedef struct xb_server_counters {
  int64_t submitted, completed, failed, released;
  int64_t tiles, tile_rows;             /* tiles run and the rows they carried  */
  int64_t budget_bytes, used_bytes, peak_used_bytes;
  int32_t queued;                       /* waiting for admission or a tile      */
  int32_t running_tiles;
  int32_t reserved[8];
} xb_server_counters;

int32_t xb_server_opts_default(xb_server_opts* o);
/* engine_opts may be null (defaults); its mode must not be XB_MODE_PLAIN.  */
int32_t xb_server_create(xb_graph* g, const xb_server_opts* opts, const xb_msbfs_opts* engine_opts,
                         xb_server** out);
/* Copies the sources.  XB_ERR_BUDGET if 8*S*n exceeds the whole budget (it
 * could never be admitted), XB_ERR_SOURCE_RANGE for an id outside [0, n),
 * XB_ERR_INVALID_ARG for S < 1 or a full queue. */
int32_t xb_server_submit(xb_server* s, const int32_t* sources, int32_t S, uint64_t* id);
/* Waits up to timeout_ms (< 0: no limit) for the request; XB_ERR_TIMEOUT if
 * it is not done by then.  Returns the request's own status once done. */
int32_t xb_server_wait(xb_server* s, uint64_t id, int32_t timeout_ms, xb_request_info* info);
/* Frees the request's tables (waits for it first if it is still running). */
int32_t xb_server_release(xb_server* s, uint64_t id);
int32_t xb_server_stats(xb_server* s, xb_server_counters* st);
/* Finishes the tiles in flight, fails what is still queued, frees everything. */
int32_t xb_server_destroy(xb_server* s);

#ifdef __cplusplus
}
#endif

#endif /* XB_CAPI_H */
