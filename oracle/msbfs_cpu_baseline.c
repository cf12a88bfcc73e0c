/*
 * msbfs_cpu_baseline.c -- multi-threaded CPU restatements of the multi-source
 * BFS path, timed as the CPU baseline.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this file.  Only tests/ (which pin it against oracle_msbfs) and the
 * cpu_baseline leg of bench.py use it.
 *
 * Both functions produce exactly the tables oracle_msbfs produces (levels =
 * hop distance, parent = lowest-id in-neighbour one level closer,
 * parent[s][s] = s, -1 unreached) and return the same traversed-edge count.
 * They follow the reference CPU engine's schedule and the reference BFS
 * baseline's direction switch:
 *   - threads take work by fetch_add on a shared counter, in chunks of 64
 *     items (apps/xb/src/xblsm_cpu/par_pro_pernode.cu:537,602-615);
 *   - a level runs bottom-up (pull) when |F| + outdeg(F) > m / 20, else
 *     top-down (push) (Ligra's edgeMapData, apps/ligra/ligra/ligra.h:
 *     236-277, threshold |E|/20 at :239,262).
 * A pull scans a vertex's in-neighbours in ascending id order, so the first
 * one found in the frontier is its lowest-id parent; a push keeps the minimum
 * over every frontier in-neighbour it sees.
 *
 * oracle_cpu_msbfs_dir     one source per BFS, threads over sources.
 * oracle_cpu_msbfs_bitpar  64 sources per uint64 word, one word at a time;
 *                          each level is parallel over vertices (chunks of
 *                          64), with a barrier between phases.
 */
#define _POSIX_C_SOURCE 200809L /* pthread barriers under -std=c11 */
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define CHUNK 64

static void* xmalloc(size_t bytes) {
  void* p = malloc(bytes ? bytes : 1);
  if (!p) abort();
  return p;
}

/* ---- per-source, direction-optimizing --------------------------------- */

typedef struct {
  int32_t n;
  int64_t m;
  const int32_t *off, *col, *in_off, *in_col;
  const int32_t* sources;
  int32_t S;
  int32_t *levels, *parents;
  int64_t next;   /* shared source counter */
  int64_t edges;  /* traversed edges, summed */
  pthread_mutex_t mu;
} dir_job;

/* One BFS from src into level/parent (n entries each); returns the reached
 * vertices' out-degree sum.  q holds 2n entries (two frontier lists). */
static int64_t bfs_dir(const dir_job* J, int32_t src, int32_t* level, int32_t* parent, int32_t* q) {
  const int32_t n = J->n;
  for (int32_t v = 0; v < n; ++v) {
    level[v] = -1;
    parent[v] = -1;
  }
  int32_t* cur = q;
  int32_t* nxt = q + n;
  int32_t qn = 1, d = 0;
  int64_t reached_deg = J->off[src + 1] - J->off[src];
  cur[0] = src;
  level[src] = 0;
  parent[src] = src;
  const int64_t thresh = J->m / 20;
  while (qn > 0) {
    int64_t E = 0;
    for (int32_t i = 0; i < qn; ++i) E += J->off[cur[i] + 1] - J->off[cur[i]];
    int32_t nn = 0;
    if ((int64_t)qn + E > thresh) {
      for (int32_t v = 0; v < n; ++v) {
        if (level[v] >= 0) continue;
        for (int32_t j = J->in_off[v]; j < J->in_off[v + 1]; ++j) {
          const int32_t u = J->in_col[j];
          if (level[u] == d) {
            level[v] = d + 1;
            parent[v] = u;
            nxt[nn++] = v;
            break;
          }
        }
      }
    } else {
      for (int32_t i = 0; i < qn; ++i) {
        const int32_t v = cur[i];
        for (int32_t j = J->off[v]; j < J->off[v + 1]; ++j) {
          const int32_t u = J->col[j];
          if (level[u] < 0) {
            level[u] = d + 1;
            parent[u] = v;
            nxt[nn++] = u;
          } else if (level[u] == d + 1 && v < parent[u]) {
            parent[u] = v;
          }
        }
      }
    }
    for (int32_t i = 0; i < nn; ++i) reached_deg += J->off[nxt[i] + 1] - J->off[nxt[i]];
    int32_t* t = cur;
    cur = nxt;
    nxt = t;
    qn = nn;
    ++d;
  }
  return reached_deg;
}

static void* dir_worker(void* arg) {
  dir_job* J = (dir_job*)arg;
  int32_t* q = (int32_t*)xmalloc(sizeof(int32_t) * 2 * (size_t)(J->n > 0 ? J->n : 1));
  int64_t local = 0;
  for (;;) {
    const int64_t s = __atomic_fetch_add(&J->next, 1, __ATOMIC_RELAXED);
    if (s >= J->S) break;
    local += bfs_dir(J, J->sources[s], J->levels + (size_t)s * (size_t)J->n,
                     J->parents + (size_t)s * (size_t)J->n, q);
  }
  free(q);
  pthread_mutex_lock(&J->mu);
  J->edges += local;
  pthread_mutex_unlock(&J->mu);
  return NULL;
}

/* in_off/in_col: the transpose (the same arrays for a symmetric graph). */
int64_t oracle_cpu_msbfs_dir(int32_t n, const int32_t* off, const int32_t* col, const int32_t* in_off,
                             const int32_t* in_col, const int32_t* sources, int32_t S, int32_t* levels,
                             int32_t* parents, int32_t nthreads) {
  dir_job J;
  memset(&J, 0, sizeof(J));
  J.n = n;
  J.m = off[n];
  J.off = off;
  J.col = col;
  J.in_off = in_off;
  J.in_col = in_col;
  J.sources = sources;
  J.S = S;
  J.levels = levels;
  J.parents = parents;
  pthread_mutex_init(&J.mu, NULL);
  if (nthreads < 1) nthreads = 1;
  pthread_t* th = (pthread_t*)xmalloc(sizeof(pthread_t) * (size_t)nthreads);
  for (int i = 0; i < nthreads; ++i) pthread_create(&th[i], NULL, dir_worker, &J);
  for (int i = 0; i < nthreads; ++i) pthread_join(th[i], NULL);
  free(th);
  pthread_mutex_destroy(&J.mu);
  return J.edges;
}

/* ---- bit-parallel, 64 sources per word -------------------------------- */

typedef struct {
  int32_t n;
  int64_t m;
  const int32_t *off, *col, *in_off, *in_col;
  const int32_t* sources;
  int32_t S;
  int32_t *levels, *parents;
  int nthreads;
  pthread_barrier_t bar;  /* all nthreads threads, the caller's included */
  /* word state */
  uint64_t *seen, *front, *next;
  uint8_t* touched;     /* push levels: vertex has OR-ed next bits */
  int32_t *flist, *tlist;
  int64_t fn, tn;       /* frontier / touched list sizes */
  int64_t ctr;          /* chunk counter of the phase */
  int64_t fedges;       /* out-degree sum of the frontier */
  int32_t w0, cnt, d, pull, done;
  int64_t edges;
  pthread_mutex_t mu;
} bp_job;

static int64_t take_chunk(bp_job* J) { return __atomic_fetch_add(&J->ctr, CHUNK, __ATOMIC_RELAXED); }

static void set_entry(bp_job* J, int b, int32_t v, int32_t u) {
  const size_t at = (size_t)(J->w0 + b) * (size_t)J->n + (size_t)v;
  J->parents[at] = u;
  J->levels[at] = J->d + 1;
}

/* Pull: every vertex with unseen bits scans its in-neighbours in ascending
 * order; the first frontier in-neighbour with a bit is that bit's parent. */
static void phase_pull(bp_job* J) {
  const int32_t n = J->n;
  const uint64_t mask = J->cnt >= 64 ? ~0ull : ((1ull << J->cnt) - 1ull);
  for (int64_t c0; (c0 = take_chunk(J)) < n;) {
    const int64_t c1 = c0 + CHUNK < n ? c0 + CHUNK : n;
    for (int64_t v = c0; v < c1; ++v) {
      uint64_t need = ~J->seen[v] & mask, got = 0;
      for (int32_t j = J->in_off[v]; j < J->in_off[v + 1] && need; ++j) {
        const int32_t u = J->in_col[j];
        uint64_t f = J->front[u] & need;
        need &= ~f;
        got |= f;
        while (f) {
          const int b = __builtin_ctzll(f);
          f &= f - 1;
          set_entry(J, b, (int32_t)v, u);
        }
      }
      J->next[v] = got;
    }
  }
}

/* Push: the frontier ORs its new bits into its out-neighbours' next words;
 * every target is listed once (tlist). */
static void phase_push(bp_job* J) {
  int32_t lbuf[CHUNK];
  for (int64_t c0; (c0 = take_chunk(J)) < J->fn;) {
    const int64_t c1 = c0 + CHUNK < J->fn ? c0 + CHUNK : J->fn;
    int k = 0;
    for (int64_t i = c0; i < c1; ++i) {
      const int32_t v = J->flist[i];
      const uint64_t f = J->front[v];
      for (int32_t j = J->off[v]; j < J->off[v + 1]; ++j) {
        const int32_t u = J->col[j];
        const uint64_t nb = f & ~J->seen[u];
        if (!nb) continue;
        if ((__atomic_load_n(&J->next[u], __ATOMIC_RELAXED) & nb) != nb)
          __atomic_fetch_or(&J->next[u], nb, __ATOMIC_RELAXED);
        if (!__atomic_load_n(&J->touched[u], __ATOMIC_RELAXED) &&
            !__atomic_exchange_n(&J->touched[u], 1, __ATOMIC_RELAXED)) {
          lbuf[k++] = u;
          if (k == CHUNK) {
            const int64_t at = __atomic_fetch_add(&J->tn, k, __ATOMIC_RELAXED);
            memcpy(J->tlist + at, lbuf, sizeof(int32_t) * (size_t)k);
            k = 0;
          }
        }
      }
    }
    if (k) {
      const int64_t at = __atomic_fetch_add(&J->tn, k, __ATOMIC_RELAXED);
      memcpy(J->tlist + at, lbuf, sizeof(int32_t) * (size_t)k);
    }
  }
}

/* After a push: the parents of each listed vertex's new bits, from its
 * in-neighbours in ascending order (the first frontier one per bit). */
static void phase_resolve(bp_job* J) {
  for (int64_t c0; (c0 = take_chunk(J)) < J->tn;) {
    const int64_t c1 = c0 + CHUNK < J->tn ? c0 + CHUNK : J->tn;
    for (int64_t i = c0; i < c1; ++i) {
      const int32_t v = J->tlist[i];
      uint64_t nb = J->next[v];
      for (int32_t j = J->in_off[v]; j < J->in_off[v + 1] && nb; ++j) {
        const int32_t u = J->in_col[j];
        uint64_t f = J->front[u] & nb;
        nb &= ~f;
        while (f) {
          const int b = __builtin_ctzll(f);
          f &= f - 1;
          set_entry(J, b, v, u);
        }
      }
      J->touched[v] = 0;
    }
  }
}

/* One level, run by every thread (the caller's thread is thread 0).
 * Barriers: A the level starts, B the push is complete, C the chunk counter
 * is reset, D the next words are complete, E thread 0 has committed. */
static void run_level(bp_job* J, int tid) {
  if (J->pull) {
    phase_pull(J);
  } else {
    phase_push(J);
    pthread_barrier_wait(&J->bar);  /* B */
    if (tid == 0) J->ctr = 0;
    pthread_barrier_wait(&J->bar);  /* C */
    phase_resolve(J);
  }
  pthread_barrier_wait(&J->bar);  /* D */
}

static void* bp_helper(void* arg) {
  bp_job* J = (bp_job*)arg;
  for (;;) {
    pthread_barrier_wait(&J->bar);  /* A */
    if (J->done) break;
    run_level(J, 1);
    pthread_barrier_wait(&J->bar);  /* E */
  }
  return NULL;
}

int64_t oracle_cpu_msbfs_bitpar(int32_t n, const int32_t* off, const int32_t* col, const int32_t* in_off,
                                const int32_t* in_col, const int32_t* sources, int32_t S, int32_t* levels,
                                int32_t* parents, int32_t nthreads) {
  bp_job J;
  memset(&J, 0, sizeof(J));
  J.n = n;
  J.m = off[n];
  J.off = off;
  J.col = col;
  J.in_off = in_off;
  J.in_col = in_col;
  J.sources = sources;
  J.S = S;
  J.levels = levels;
  J.parents = parents;
  if (nthreads < 1) nthreads = 1;
  J.nthreads = nthreads;
  const size_t N = (size_t)(n > 0 ? n : 1);
  J.seen = (uint64_t*)xmalloc(sizeof(uint64_t) * N);
  J.front = (uint64_t*)xmalloc(sizeof(uint64_t) * N);
  J.next = (uint64_t*)xmalloc(sizeof(uint64_t) * N);
  J.touched = (uint8_t*)calloc(N, 1);
  if (!J.touched) abort();
  J.flist = (int32_t*)xmalloc(sizeof(int32_t) * N);
  J.tlist = (int32_t*)xmalloc(sizeof(int32_t) * N);
  pthread_barrier_init(&J.bar, NULL, (unsigned)nthreads);
  pthread_t* th = (pthread_t*)xmalloc(sizeof(pthread_t) * (size_t)nthreads);
  for (int i = 1; i < nthreads; ++i) pthread_create(&th[i], NULL, bp_helper, &J);
  for (size_t i = 0; i < (size_t)S * (size_t)n; ++i) {
    levels[i] = -1;
    parents[i] = -1;
  }
  const int64_t thresh = J.m / 20;
  for (int32_t w0 = 0; w0 < S; w0 += 64) {
    J.w0 = w0;
    J.cnt = S - w0 < 64 ? S - w0 : 64;
    memset(J.seen, 0, sizeof(uint64_t) * N);
    memset(J.front, 0, sizeof(uint64_t) * N);
    memset(J.next, 0, sizeof(uint64_t) * N);
    J.fn = 0;
    for (int32_t b = 0; b < J.cnt; ++b) {
      const int32_t s = sources[w0 + b];
      if (!J.front[s]) J.flist[J.fn++] = s;
      J.seen[s] |= 1ull << b;
      J.front[s] |= 1ull << b;
      levels[(size_t)(w0 + b) * N + (size_t)s] = 0;
      parents[(size_t)(w0 + b) * N + (size_t)s] = s;
    }
    J.d = 0;
    while (J.fn > 0) {
      int64_t E = 0;
      for (int64_t i = 0; i < J.fn; ++i) E += off[J.flist[i] + 1] - off[J.flist[i]];
      J.pull = (J.fn + E > thresh);
      J.ctr = 0;
      J.tn = 0;
      pthread_barrier_wait(&J.bar);  /* A */
      run_level(&J, 0);
      /* Commit: the next words become the frontier. */
      J.fn = 0;
      for (int32_t v = 0; v < n; ++v) {
        const uint64_t nb = J.next[v];
        J.front[v] = nb;
        J.next[v] = 0;
        if (nb) {
          J.seen[v] |= nb;
          J.flist[J.fn++] = v;
        }
      }
      ++J.d;
      pthread_barrier_wait(&J.bar);  /* E */
    }
    for (int32_t v = 0; v < n; ++v) J.edges += (int64_t)__builtin_popcountll(J.seen[v]) * (off[v + 1] - off[v]);
  }
  J.done = 1;
  pthread_barrier_wait(&J.bar);  /* A: the helpers see done and leave */
  for (int i = 1; i < nthreads; ++i) pthread_join(th[i], NULL);
  pthread_barrier_destroy(&J.bar);
  free(th);
  free(J.seen);
  free(J.front);
  free(J.next);
  free(J.touched);
  free(J.flist);
  free(J.tlist);
  return J.edges;
}
