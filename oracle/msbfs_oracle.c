/*
 * msbfs_oracle.c -- CPU oracle for the multi-source BFS path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this file.  Only tests/, __graft_entry__.smoke() and the cpu_baseline
 * leg of bench.py use it, and only as the checker (or as the timed CPU
 * baseline), never as the thing that is measured on the GPU.
 *
 * What it restates
 * ----------------
 * The reference's search step grows BFS trees level by level from a set of
 * source vertices ("exposed vertices become roots",
 * apps/xb/src/xblsm_gpu/xblossom_gpu_engine_10.cu:119-129) and records, for
 * every claimed vertex, an anchor that is adjacent to it and exactly one
 * level closer to the source (apps/xb/src/include/path_table/path_table.h:
 * 256-271: paths_[off..off+1] = [w, v], root_distance[x] = root_distance[v]+2).
 * The CPU engine's expand step does the same with a std::thread pool
 * (apps/xb/src/xblsm_cpu/par_pro_pernode.cu:178-234).  The reference lets the
 * first atomic winner pick the anchor; this restatement fixes the choice to the
 * LOWEST-ID in-neighbour one level closer, so every run is bit-identical.
 *
 *   level[s][v]  = hop distance from sources[s] to v, -1 if unreached
 *   parent[s][v] = min{ u : (u->v) in E, level[s][u] == level[s][v]-1 },
 *                  parent[s][sources[s]] = sources[s], -1 if unreached.
 *
 * Per-source parent convention Parents[start] = start follows
 * apps/ligra/apps/BFS.C:47.
 *
 * The edge-balanced work mapping (exclusive scan of frontier degrees, then a
 * rightmost-<= search per edge tasklet) restates
 * apps/xb/src/include/xbsm_gpu/xblossom_gpu_engine_10.h:237-263 and
 * apps/xb/src/include/utils/cuda_algo.h:7-20.
 *
 * Parity pins (see tests/test_oracle.py): networkx shortest-path lengths and
 * shortest-path predecessor sets (committed fixtures), closed forms on paths,
 * cycles, stars, complete graphs, grids and binary trees, and the reference's
 * own known-answer test for the edge mapping
 * (apps/xb/test/device/scan_bs_test.cu:58-137).
 */
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* Single-source BFS: levels by a FIFO queue, then lowest-id parents.       */
/* ------------------------------------------------------------------------ */

/* Returns the number of reached vertices, or -1 when src is out of range.
 * queue must hold n entries. */
int64_t oracle_bfs_one(int32_t n, const int32_t* off, const int32_t* col,
                       int32_t src, int32_t* level, int32_t* parent,
                       int32_t* queue) {
  for (int32_t v = 0; v < n; ++v) {
    level[v] = -1;
    parent[v] = -1;
  }
  if (src < 0 || src >= n) return -1;
  int64_t head = 0, tail = 0;
  level[src] = 0;
  parent[src] = src;
  queue[tail++] = src;
  while (head < tail) {
    int32_t v = queue[head++];
    int32_t lv = level[v];
    for (int32_t j = off[v]; j < off[v + 1]; ++j) {
      int32_t u = col[j];
      if (level[u] < 0) {
        level[u] = lv + 1;
        queue[tail++] = u;
      }
    }
  }
  /* Parent = lowest-id in-neighbour one level closer.  Scanning every edge
   * out of every reached vertex visits each candidate (u -> v) pair once. */
  for (int64_t i = 0; i < tail; ++i) {
    int32_t u = queue[i];
    int32_t lu = level[u];
    for (int32_t j = off[u]; j < off[u + 1]; ++j) {
      int32_t v = col[j];
      if (level[v] == lu + 1 && (uint32_t)u < (uint32_t)parent[v])
        parent[v] = u;
    }
  }
  return tail;
}

/* ------------------------------------------------------------------------ */
/* Many sources, one thread pool (dynamic scheduling, chunk of one source). */
/* This mirrors the chunked fetch_add scheduling of                          */
/* apps/xb/src/xblsm_cpu/par_pro_pernode.cu:537,602-615.                     */
/* ------------------------------------------------------------------------ */

typedef struct {
  int32_t n;
  const int32_t* off;
  const int32_t* col;
  const int32_t* sources;
  int32_t S;
  int32_t* levels;
  int32_t* parents;
  int64_t next; /* shared work counter */
  int64_t reached_edges; /* sum over sources of out-degrees of reached vertices */
  pthread_mutex_t mu;
} msbfs_job;

static void* msbfs_worker(void* arg) {
  msbfs_job* job = (msbfs_job*)arg;
  int32_t* queue = (int32_t*)malloc(sizeof(int32_t) * (size_t)(job->n > 0 ? job->n : 1));
  int64_t local_edges = 0;
  for (;;) {
    int64_t s = __atomic_fetch_add(&job->next, 1, __ATOMIC_RELAXED);
    if (s >= job->S) break;
    int32_t* lv = job->levels + (size_t)s * (size_t)job->n;
    int32_t* pa = job->parents + (size_t)s * (size_t)job->n;
    int64_t r = oracle_bfs_one(job->n, job->off, job->col, job->sources[s], lv, pa, queue);
    for (int64_t i = 0; i < r; ++i) {
      int32_t v = queue[i];
      local_edges += job->off[v + 1] - job->off[v];
    }
  }
  free(queue);
  pthread_mutex_lock(&job->mu);
  job->reached_edges += local_edges;
  pthread_mutex_unlock(&job->mu);
  return NULL;
}

/* Fills levels/parents (S x n, row-major by source).  Returns the number of
 * directed edges a full top-down BFS scans, summed over sources. */
int64_t oracle_msbfs(int32_t n, const int32_t* off, const int32_t* col,
                     const int32_t* sources, int32_t S, int32_t* levels,
                     int32_t* parents, int32_t nthreads) {
  msbfs_job job;
  memset(&job, 0, sizeof(job));
  job.n = n;
  job.off = off;
  job.col = col;
  job.sources = sources;
  job.S = S;
  job.levels = levels;
  job.parents = parents;
  pthread_mutex_init(&job.mu, NULL);
  if (nthreads < 1) nthreads = 1;
  if (nthreads == 1) {
    msbfs_worker(&job);
  } else {
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)nthreads);
    for (int i = 0; i < nthreads; ++i) pthread_create(&th[i], NULL, msbfs_worker, &job);
    for (int i = 0; i < nthreads; ++i) pthread_join(th[i], NULL);
    free(th);
  }
  pthread_mutex_destroy(&job.mu);
  return job.reached_edges;
}

/* ------------------------------------------------------------------------ */
/* Bit-parallel restatement: 64 sources per uint64 word, top-down only.      */
/* Each level sweeps frontier vertices in ascending id order, so the first  */
/* frontier vertex that offers bit b to an unseen vertex is its lowest-id   */
/* parent for source b.  Used to cross-check the per-source oracle.         */
/* ------------------------------------------------------------------------ */

int64_t oracle_msbfs_bitparallel(int32_t n, const int32_t* off,
                                 const int32_t* col, const int32_t* sources,
                                 int32_t S, int32_t* levels, int32_t* parents) {
  int64_t edges = 0;
  uint64_t* seen = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)(n > 0 ? n : 1));
  uint64_t* front = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)(n > 0 ? n : 1));
  uint64_t* next = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)(n > 0 ? n : 1));
  for (int32_t w0 = 0; w0 < S; w0 += 64) {
    int32_t cnt = S - w0 < 64 ? S - w0 : 64;
    memset(seen, 0, sizeof(uint64_t) * (size_t)n);
    memset(front, 0, sizeof(uint64_t) * (size_t)n);
    for (int32_t b = 0; b < cnt; ++b) {
      int32_t* lv = levels + (size_t)(w0 + b) * (size_t)n;
      int32_t* pa = parents + (size_t)(w0 + b) * (size_t)n;
      for (int32_t v = 0; v < n; ++v) {
        lv[v] = -1;
        pa[v] = -1;
      }
      int32_t s = sources[w0 + b];
      if (s >= 0 && s < n) {
        seen[s] |= 1ull << b;
        front[s] |= 1ull << b;
        lv[s] = 0;
        pa[s] = s;
      }
    }
    int32_t d = 0;
    for (;;) {
      int any = 0;
      memset(next, 0, sizeof(uint64_t) * (size_t)n);
      for (int32_t v = 0; v < n; ++v) {
        uint64_t f = front[v];
        if (!f) continue;
        for (int32_t j = off[v]; j < off[v + 1]; ++j) {
          int32_t u = col[j];
          uint64_t nb = f & ~seen[u];
          if (!nb) continue;
          next[u] |= nb;
          /* v ascends, so the first v to offer bit b is the lowest id. */
          uint64_t fresh = nb;
          while (fresh) {
            int b = __builtin_ctzll(fresh);
            fresh &= fresh - 1;
            int32_t* pa = parents + (size_t)(w0 + b) * (size_t)n;
            if (pa[u] < 0) pa[u] = v;
          }
        }
      }
      for (int32_t v = 0; v < n; ++v) {
        uint64_t nb = next[v] & ~seen[v];
        front[v] = nb;
        if (!nb) continue;
        any = 1;
        seen[v] |= nb;
        while (nb) {
          int b = __builtin_ctzll(nb);
          nb &= nb - 1;
          levels[(size_t)(w0 + b) * (size_t)n + v] = d + 1;
        }
      }
      if (!any) break;
      ++d;
    }
    for (int32_t v = 0; v < n; ++v)
      edges += (int64_t)__builtin_popcountll(seen[v]) * (off[v + 1] - off[v]);
  }
  free(seen);
  free(front);
  free(next);
  return edges;
}

/* ------------------------------------------------------------------------ */
/* Edge-balanced work mapping: exclusive scan + rightmost-<= search.         */
/* ------------------------------------------------------------------------ */

/* presum must hold k+1 entries; returns the total. */
int64_t oracle_exclusive_scan(const int32_t* deg, int64_t k, int64_t* presum) {
  int64_t acc = 0;
  for (int64_t i = 0; i < k; ++i) {
    presum[i] = acc;
    acc += deg[i];
  }
  presum[k] = acc;
  return acc;
}

/* Index i of the rightmost presum[i] <= target among i in [0, k), skipping
 * zero-length segments (apps/xb/src/include/utils/cuda_algo.h:7-20 searches
 * presum[mid+1] <= target, which lands on the last segment that starts at or
 * before the target and is non-empty). */
int64_t oracle_rightmost_lteq(const int64_t* presum, int64_t k, int64_t target) {
  int64_t lo = 0, hi = k - 1;
  while (lo < hi) {
    int64_t mid = lo + (hi - lo) / 2;
    if (presum[mid + 1] <= target)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

/* For every tasklet t in [0, total): owner[t] = segment index, offset[t] =
 * t - presum[owner]. */
int64_t oracle_edge_map(const int32_t* deg, int64_t k, int32_t* owner,
                        int32_t* offset) {
  int64_t* presum = (int64_t*)malloc(sizeof(int64_t) * (size_t)(k + 1));
  int64_t total = oracle_exclusive_scan(deg, k, presum);
  for (int64_t t = 0; t < total; ++t) {
    int64_t i = oracle_rightmost_lteq(presum, k, t);
    owner[t] = (int32_t)i;
    offset[t] = (int32_t)(t - presum[i]);
  }
  free(presum);
  return total;
}
