/*
 * match_oracle.c -- CPU oracle for forest mode (maximum matching).
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this file.  Only tests/ and the cpu_baseline leg of bench.py use it,
 * as the checker or as the timed CPU baseline.
 *
 * Functions
 * ---------
 * oracle_match_forest   Sequential restatement of the deterministic
 *                       alternating-forest phase loop that the engine runs on
 *                       the GPU.  It follows the reference's
 *                       FindMaximumMatch loop (apps/xb/src/include/xbsm_gpu/
 *                       max_match_engine.h:93-113) and its search
 *                       (apps/xb/src/xblsm_gpu/xblossom_gpu_engine_10.cu:
 *                       83-359): exposed vertices become even roots
 *                       (:119-129), the even frontier claims its unvisited
 *                       matched neighbours (:270-359), and an edge between
 *                       even vertices of two trees closes an augmenting path,
 *                       which is flipped back to both roots (path_table.h:
 *                       273-330).  The reference lets the first atomic winner
 *                       decide every race; this restatement fixes each
 *                       choice:
 *                         - a vertex is claimed by the even neighbour with the
 *                           lowest key (root << 32 | claimer);
 *                         - of a matched pair claimed from both sides, the
 *                           member with the lower (key, id) becomes odd;
 *                         - per level, rounds take every pair of trees that
 *                           are each other's lowest-rooted neighbour tree
 *                           over even-even edges; the path's bridge is the
 *                           lowest (vertex in the lower tree << 32 | vertex in
 *                           the higher tree); taken trees leave the later
 *                           rounds of the phase.
 *                       Between phases the trees whose root is still exposed
 *                       are kept, as the reference's lazy reset keeps them
 *                       (xblossom_gpu_engine_10.cu:119-175; CPU:
 *                       apps/xb/src/include/xblsm_cpu/xb_cpu_reuse.h:8-71),
 *                       and the next phase grows from all their outer
 *                       vertices; the regrow mode (reuse = 0) resets the
 *                       whole forest every phase.
 *                       The GPU result must equal this one entry for entry.
 * oracle_match_edmonds  Maximum cardinality matching of a general graph by
 *                       Edmonds' blossom algorithm (J. Edmonds, "Paths, trees,
 *                       and flowers", 1965), in the textbook O(n^3) form: a
 *                       BFS from each exposed vertex with blossom bases, LCA
 *                       by marking, and path relabelling on contraction.
 * oracle_match_hk       Maximum matching of a bipartite graph by Hopcroft and
 *                       Karp ("An n^{5/2} algorithm for maximum matchings in
 *                       bipartite graphs", 1973): BFS layering from the free
 *                       left vertices, then augmenting DFS on the layers (with
 *                       an explicit stack).  Returns -1 when the graph is not
 *                       bipartite (2-colouring by BFS).
 * oracle_match_validate Counts the entries that break the matching rules.
 *
 * Parity pins (tests/test_match_oracle.py): closed forms (paths, even and
 * odd cycles, complete bipartite and complete graphs, grids, stars, the
 * Petersen graph), networkx's maximum cardinality matching and scipy's
 * Hopcroft-Karp on random graphs, and the certificate that the forest
 * result is maximum whenever its final phase met no same-tree even-even
 * edge (checked against oracle_match_edmonds).
 *
 * All graphs are symmetric int32 CSR; mate[v] = -1 marks an exposed vertex.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct oracle_match_stats {
  int64_t matched_vertices;
  int64_t augmentations;
  int64_t blossom_edges;
  int64_t edges_scanned; /* frontier edges read by the scan, bridge, claim and blossom passes */
  int64_t blossoms;         /* blossoms formed */
  int64_t blossom_vertices; /* odd vertices made even by them */
  int32_t phases;
  int32_t levels;
  /* edges_scanned split as the reference's counters split it
   * (xblossom_gpu_engine_10.cu:220-221, 303-304, 400-401): */
  int64_t aug_path_edges;     /* level-0 cursor scans, augmenting-edge scans, bridge passes */
  int64_t expand_edges;       /* claim passes */
  int64_t blossom_scan_edges; /* blossom-candidate passes */
  /* Per phase, summed (the reference's lazy reset, xblossom_gpu_engine_10.cu:
   * 119-155): matched vertices kept in a tree whose root is still exposed,
   * and the others (exposed roots, vertices of augmented trees, vertices in
   * no tree); the two add up to n per phase. */
  int64_t reused_tree_vertices;
  int64_t reset_tree_vertices;
} oracle_match_stats;

#define NO_KEY (~0ull)

static void* xcalloc(size_t count, size_t size) {
  void* p = calloc(count ? count : 1, size);
  if (!p) abort();
  return p;
}

int64_t oracle_match_validate(int32_t n, const int32_t* off, const int32_t* col, const int32_t* mate) {
  int64_t bad = 0;
  for (int32_t v = 0; v < n; ++v) {
    const int32_t m = mate[v];
    if (m < 0) {
      bad += (m != -1);
      continue;
    }
    if (m >= n || m == v || mate[m] != v) {
      ++bad;
      continue;
    }
    int edge = 0;
    for (int32_t j = off[v]; j < off[v + 1]; ++j) edge |= (col[j] == m);
    bad += !edge;
  }
  return bad;
}

static int64_t count_matched(int32_t n, const int32_t* mate) {
  int64_t c = 0;
  for (int32_t v = 0; v < n; ++v) c += mate[v] >= 0;
  return c;
}

/* ---- deterministic alternating forest --------------------------------- */

static int64_t frontier_edges(const int32_t* off, const int32_t* front, int32_t fn) {
  int64_t e = 0;
  for (int32_t i = 0; i < fn; ++i) e += off[front[i] + 1] - off[front[i]];
  return e;
}

#define DUMMY (-1) /* the non-outer vertex above a root on every path */

typedef struct {
  const int32_t *off, *col;
  int32_t *lvl, *par, *fp;
  uint64_t *elab, *flag;
} Forest;

/* Outer (even) vertices have an even level; transformed odd vertices get one. */
static int is_outer(const Forest* f, int32_t u) { return f->lvl[u] >= 0 && !(f->lvl[u] & 1); }

/* FIRST(v): the first non-outer vertex on P(v), DUMMY above the root. */
static int32_t first_of(const Forest* f, int32_t v) {
  int32_t u = f->fp[v];
  while (u != DUMMY && is_outer(f, u)) u = f->fp[u];
  return u;
}

static int is_flagged(const Forest* f, int32_t u, uint64_t stamp, int dflag) {
  return u == DUMMY ? dflag : f->flag[u] == stamp;
}

/* Gabow's procedure L for the even-even edge (x, y) of one tree: walks up
 * from FIRST(x) and FIRST(y) alternately to their join, then makes every
 * non-outer vertex below the join outer with the edge label of its side
 * (x side: (x, y); y side: (y, x)), FIRST = join, its (odd) level plus one,
 * appended to next[].  Returns the number of vertices made outer (0: x and y already
 * share a blossom). */
static int64_t blossom_step(Forest* f, int32_t x, int32_t y, uint64_t stamp, int32_t* next, int32_t* nn) {
  int32_t r = first_of(f, x), s = first_of(f, y);
  if (r == s) return 0;
  int dflag = 0;
  if (r == DUMMY) dflag = 1; else f->flag[r] = stamp;
  if (s == DUMMY) dflag = 1; else f->flag[s] = stamp;
  int32_t join;
  for (;;) {
    if (s != DUMMY) {
      const int32_t t = r;
      r = s;
      s = t;
    }
    r = first_of(f, f->par[r]);
    if (is_flagged(f, r, stamp, dflag)) {
      join = r;
      break;
    }
    if (r == DUMMY) dflag = 1; else f->flag[r] = stamp;
  }
  int64_t made = 0;
  for (int side = 0; side < 2; ++side) {
    const int32_t p = side ? y : x, q = side ? x : y;
    const uint64_t label = ((uint64_t)(uint32_t)p << 32) | (uint32_t)q;
    int32_t v = first_of(f, p);
    while (v != join) {
      const int32_t nv = first_of(f, f->par[v]);
      f->elab[v] = label;
      f->fp[v] = join;
      f->lvl[v] += 1;  /* inner (odd) -> outer (even), still above its FIRST chain */
      next[(*nn)++] = v;
      ++made;
      v = nv;
    }
  }
  return made;
}

/* Rematch(v, w): match v to w and flip P(v) behind it (Gabow's procedure R,
 * H. N. Gabow, JACM 23(2), 1976).  Paths: a root's is itself; a grown
 * vertex v with mate t has P(v) = v, t, P(par[t]); a vertex made outer by
 * the edge label (a, b) has P(v) = reverse(P(a) from a to v), P(b).  The
 * walk stops at a vertex whose mate has already changed. */
static void rematch(const Forest* f, int32_t* mate, int32_t v, int32_t w, int32_t* stk, int64_t cap) {
  int64_t top = 0;
  for (;;) {
    const int32_t t = mate[v];
    mate[v] = w;
    if (t >= 0 && mate[t] == v) {
      if (f->elab[v] == NO_KEY) {
        const int32_t c = f->par[t];
        mate[t] = c;
        v = c;
        w = t;
      } else {
        const int32_t a = (int32_t)(f->elab[v] >> 32), b = (int32_t)(f->elab[v] & 0xFFFFFFFFu);
        if (top + 2 > cap) abort();
        stk[top++] = b;
        stk[top++] = a;
        v = a;
        w = b;
      }
      continue;
    }
    if (top == 0) return;
    w = stk[--top];
    v = stk[--top];
  }
}

static int cmp_u64(const void* a, const void* b) {
  const uint64_t x = *(const uint64_t*)a, y = *(const uint64_t*)b;
  return x < y ? -1 : x > y;
}

/* Phase set-up with tree reuse (the reference's lazy reset,
 * apps/xb/src/xblsm_gpu/xblossom_gpu_engine_10.cu:119-175, and the CPU
 * engine's parReuseRemainingTrees, apps/xb/src/include/xblsm_cpu/
 * xb_cpu_reuse.h:8-71): the vertices of the trees that augmented in the last
 * phase (their roots are matched now) are reset; the trees whose root is still
 * exposed keep every label.  The phase's first frontier is every outer vertex
 * of those trees (the reference's ENode2 list), in ascending id order, and
 * its roots are the exposed vertices.  Returns the frontier size. */
static int32_t reuse_forest(int32_t n, const int32_t* mate, int32_t* root, int32_t* lvl, int32_t* par, int32_t* fp,
                            uint64_t* elab, uint64_t* claim, uint8_t* used, int32_t* front, int32_t* roots,
                            int32_t* nroots, oracle_match_stats* s) {
  for (int32_t v = 0; v < n; ++v) {
    const int32_t r = root[v];
    if (r >= 0 && used[r]) {
      root[v] = -1;
      lvl[v] = -1;
      par[v] = -1;
      fp[v] = DUMMY;
      elab[v] = NO_KEY;
      claim[v] = NO_KEY;
      ++s->reset_tree_vertices;
    } else if (r >= 0 && mate[v] >= 0) {
      ++s->reused_tree_vertices;
    } else {
      ++s->reset_tree_vertices;  /* an exposed root, or a vertex in no tree */
    }
  }
  int32_t fn = 0, nr = 0;
  for (int32_t v = 0; v < n; ++v) {
    used[v] = 0;
    if (lvl[v] >= 0 && !(lvl[v] & 1)) front[fn++] = v;
    if (mate[v] < 0) roots[nr++] = v;
  }
  *nroots = nr;
  return fn;
}

/* max_phases > 0 stops after that many phases (a bounded sample for timing).
 * reuse != 0 keeps the trees of still-exposed roots from one phase to the
 * next (reuse_forest); reuse == 0 regrows the whole forest every phase. */
int64_t oracle_match_forest(int32_t n, const int32_t* off, const int32_t* col, int32_t* mate,
                            oracle_match_stats* st, int32_t max_phases, int32_t reuse) {
  oracle_match_stats s;
  memset(&s, 0, sizeof(s));
  int32_t* root = xcalloc((size_t)n, sizeof(int32_t));
  int32_t* lvl = xcalloc((size_t)n, sizeof(int32_t));
  int32_t* par = xcalloc((size_t)n, sizeof(int32_t));
  int32_t* fp = xcalloc((size_t)n, sizeof(int32_t));
  int32_t* best = xcalloc((size_t)n, sizeof(int32_t));
  int32_t* acc = xcalloc((size_t)n, sizeof(int32_t));
  int32_t* roots = xcalloc((size_t)n, sizeof(int32_t));
  int32_t* front = xcalloc((size_t)n, sizeof(int32_t));
  int32_t* next = xcalloc((size_t)n, sizeof(int32_t));
  int32_t* claimed = xcalloc((size_t)n, sizeof(int32_t));
  int32_t* pairs = xcalloc((size_t)n, sizeof(int32_t));
  int32_t* cursor = xcalloc((size_t)n, sizeof(int32_t));
  uint64_t* claim = xcalloc((size_t)n, sizeof(uint64_t));
  uint64_t* bridge = xcalloc((size_t)n, sizeof(uint64_t));
  uint64_t* elab = xcalloc((size_t)n, sizeof(uint64_t));
  uint64_t* flag = xcalloc((size_t)n, sizeof(uint64_t));
  uint8_t* used = xcalloc((size_t)n, 1);
  const int64_t stk_cap = 2 * (int64_t)n + 64;
  int32_t* stk = xcalloc((size_t)stk_cap, sizeof(int32_t));
  int64_t cand_cap = 1024;
  uint64_t* cand = xcalloc((size_t)cand_cap, sizeof(uint64_t));
  uint64_t stamp_base = 0;
  Forest F = {off, col, lvl, par, fp, elab, flag};
  /* Once a phase's level 0 has run to a round with no pair, no two exposed
   * vertices are adjacent, and that stays so: an augmentation matches both of
   * its exposed ends and exposes nothing.  Later phases skip the level-0
   * rounds, which could only scan the roots' lists and find nothing. */
  int roots_independent = 0;
  for (;;) {
    if (max_phases > 0 && s.phases == max_phases) break;
    ++s.phases;
    int64_t same_tree = 0;
    int32_t nroots = 0, fn = 0;
    /* roots_only: the frontier is the phase's roots, every tree one root. */
    int roots_only;
    if (reuse && s.phases > 1) {
      fn = reuse_forest(n, mate, root, lvl, par, fp, elab, claim, used, front, roots, &nroots, &s);
      for (int32_t v = 0; v < n; ++v) {
        best[v] = INT32_MAX;
        acc[v] = -1;
        bridge[v] = NO_KEY;
      }
      roots_only = 0;
    } else {
      for (int32_t v = 0; v < n; ++v) {
        const int ex = mate[v] < 0;
        root[v] = ex ? v : -1;
        lvl[v] = ex ? 0 : -1;
        par[v] = -1;
        fp[v] = DUMMY;
        elab[v] = NO_KEY;
        claim[v] = NO_KEY;
        best[v] = INT32_MAX;
        acc[v] = -1;
        bridge[v] = NO_KEY;
        used[v] = 0;
        cursor[v] = off[v];
        if (ex) roots[nroots++] = v;
      }
      s.reset_tree_vertices += n;
      memcpy(front, roots, sizeof(int32_t) * (size_t)nroots);
      fn = nroots;
      roots_only = 1;
    }
    if (nroots < 2) break;
    int64_t found = 0;
    while (fn > 0) {
      int64_t fe = frontier_edges(off, front, fn);
      int64_t level_same_tree = 0;
      for (;;) {
        if (roots_only && roots_independent) break;
        if (roots_only) {
          /* Every tree is one root: its lowest neighbouring tree is its first
           * adjacency entry that is an untaken root, found by a cursor that
           * only moves forward (taken roots never return). */
          for (int32_t i = 0; i < nroots; ++i) {
            const int32_t t = roots[i], e = off[t + 1];
            int32_t j = cursor[t];
            const int32_t j0 = j;
            while (j < e) {
              const int32_t w = col[j];
              if (w != t && lvl[w] == 0 && !used[w]) break;
              ++j;
            }
            s.aug_path_edges += (j - j0) + (j < e);
            cursor[t] = j;
            best[t] = j < e ? col[j] : INT32_MAX;
          }
        } else {
          s.aug_path_edges += fe;
          for (int32_t i = 0; i < nroots; ++i) best[roots[i]] = INT32_MAX;
          level_same_tree = 0;
          for (int32_t i = 0; i < fn; ++i) {
            const int32_t v = front[i];
            for (int32_t j = off[v]; j < off[v + 1]; ++j) {
              const int32_t w = col[j];
              if (!is_outer(&F, w) || w == v) continue;
              const int32_t rv = root[v], rw = root[w];
              if (rv == rw) {
                ++level_same_tree;
                continue;
              }
              if (used[rv] || used[rw]) continue;
              if (rw < best[rv]) best[rv] = rw;
              if (rv < best[rw]) best[rw] = rv;
            }
          }
          same_tree += level_same_tree;
        }
        int32_t np = 0;
        for (int32_t i = 0; i < nroots; ++i) {
          const int32_t t = roots[i], b = best[t];
          if (b != INT32_MAX && t < b && best[b] == t) {
            acc[t] = b;
            if (roots_only) bridge[t] = ((uint64_t)(uint32_t)t << 32) | (uint32_t)b;
            pairs[np++] = t;
          }
        }
        if (np == 0) break;
        if (!roots_only) {
          s.aug_path_edges += fe;
          for (int32_t i = 0; i < fn; ++i) {
            const int32_t v = front[i];
            for (int32_t j = off[v]; j < off[v + 1]; ++j) {
              const int32_t w = col[j];
              if (!is_outer(&F, w)) continue;
              const int32_t rv = root[v], rw = root[w];
              if (rv == rw) continue;
              const int32_t lo = rv < rw ? rv : rw, hi = rv < rw ? rw : rv;
              if (acc[lo] != hi) continue;
              const uint32_t a = (uint32_t)(rv == lo ? v : w), c = (uint32_t)(rv == lo ? w : v);
              const uint64_t key = ((uint64_t)a << 32) | c;
              if (key < bridge[lo]) bridge[lo] = key;
            }
          }
        }
        for (int32_t i = 0; i < np; ++i) {
          const int32_t t = pairs[i];
          const int32_t a = (int32_t)(bridge[t] >> 32), c = (int32_t)(bridge[t] & 0xFFFFFFFFu);
          rematch(&F, mate, a, c, stk, stk_cap);
          rematch(&F, mate, c, a, stk, stk_cap);
          used[t] = 1;
          used[acc[t]] = 1;
        }
        found += np;
        /* The next round sees only the frontier and roots of trees not taken. */
        int32_t k = 0;
        for (int32_t i = 0; i < fn; ++i)
          if (!used[root[front[i]]]) front[k++] = front[i];
        fn = k;
        k = 0;
        for (int32_t i = 0; i < nroots; ++i)
          if (!used[roots[i]]) roots[k++] = roots[i];
        nroots = k;
        fe = frontier_edges(off, front, fn);
      }
      if (roots_only) roots_independent = 1;
      if (found) break;
      /* Blossom candidates: the same-tree even-even frontier edges, keyed
       * (root << 32 | edge position), taken before the claims. */
      int64_t C = 0;
      if (level_same_tree > 0) {
        if (level_same_tree > cand_cap) {
          free(cand);
          cand_cap = level_same_tree;
          cand = xcalloc((size_t)cand_cap, sizeof(uint64_t));
        }
        s.blossom_scan_edges += fe;
        for (int32_t i = 0; i < fn; ++i) {
          const int32_t v = front[i];
          for (int32_t j = off[v]; j < off[v + 1]; ++j) {
            const int32_t w = col[j];
            if (!is_outer(&F, w) || w == v || root[w] != root[v]) continue;
            cand[C++] = ((uint64_t)(uint32_t)root[v] << 32) | (uint32_t)j;
          }
        }
      }
      s.expand_edges += fe;
      int32_t tn = 0;
      for (int32_t i = 0; i < fn; ++i) {
        const int32_t v = front[i];
        const uint64_t key = ((uint64_t)(uint32_t)root[v] << 32) | (uint32_t)v;
        for (int32_t j = off[v]; j < off[v + 1]; ++j) {
          const int32_t w = col[j];
          if (lvl[w] >= 0) continue;
          if (claim[w] == NO_KEY) claimed[tn++] = w;
          if (key < claim[w]) claim[w] = key;
        }
      }
      int32_t nn = 0;
      for (int32_t i = 0; i < tn; ++i) {
        const int32_t w = claimed[i], x = mate[w];
        const uint64_t kw = claim[w], kx = claim[x];
        if (kw < kx || (kw == kx && w < x)) {
          /* Levels follow the claimer (the frontier of a reused forest mixes
           * levels): odd for the claimed vertex, even for its mate, so levels
           * rise strictly down every tree path. */
          const int32_t rv = (int32_t)(kw >> 32), cl = (int32_t)(kw & 0xFFFFFFFFu);
          root[w] = rv;
          par[w] = cl;
          lvl[w] = lvl[cl] + 1;
          root[x] = rv;
          par[x] = w;
          fp[x] = w;
          lvl[x] = lvl[cl] + 2;
          next[nn++] = x;
        }
      }
      if (C > 0) {
        /* Trees are independent; within a tree, candidates in edge order. */
        qsort(cand, (size_t)C, sizeof(uint64_t), cmp_u64);
        for (int64_t i = 0; i < C; ++i) {
          const int32_t j = (int32_t)(cand[i] & 0xFFFFFFFFu);
          int32_t lo = 0, hi = n;
          while (lo < hi) { /* owner of entry j: the last vertex whose offset is <= j */
            const int32_t mid = (lo + hi + 1) >> 1;
            if (off[mid] <= j) lo = mid; else hi = mid - 1;
          }
          const int64_t made = blossom_step(&F, lo, col[j], stamp_base + (uint64_t)i + 1, next, &nn);
          if (made > 0) {
            ++s.blossoms;
            s.blossom_vertices += made;
          }
        }
        stamp_base += (uint64_t)C;
      }
      int32_t* t = front;
      front = next;
      next = t;
      fn = nn;
      roots_only = 0;
      ++s.levels;
    }
    s.augmentations += found;
    if (!found) {
      s.blossom_edges = same_tree;
      break;
    }
  }
  s.edges_scanned = s.aug_path_edges + s.expand_edges + s.blossom_scan_edges;
  s.matched_vertices = count_matched(n, mate);
  free(root);
  free(lvl);
  free(par);
  free(fp);
  free(best);
  free(acc);
  free(roots);
  free(front);
  free(next);
  free(claimed);
  free(pairs);
  free(cursor);
  free(claim);
  free(bridge);
  free(elab);
  free(flag);
  free(used);
  free(stk);
  free(cand);
  if (st) *st = s;
  return s.matched_vertices;
}

/* ---- Edmonds' blossom algorithm ----------------------------------------- */

typedef struct {
  int32_t n;
  const int32_t *off, *col;
  int32_t *match, *p, *base, *queue;
  uint8_t *used, *blossom, *seen;
  int32_t qh, qt;
} Edm;

static int32_t edm_lca(Edm* e, int32_t a, int32_t b) {
  memset(e->seen, 0, (size_t)e->n);
  for (;;) {
    a = e->base[a];
    e->seen[a] = 1;
    if (e->match[a] < 0) break;
    a = e->p[e->match[a]];
  }
  for (;;) {
    b = e->base[b];
    if (e->seen[b]) return b;
    b = e->p[e->match[b]];
  }
}

static void edm_mark(Edm* e, int32_t v, int32_t b, int32_t child) {
  while (e->base[v] != b) {
    e->blossom[e->base[v]] = 1;
    e->blossom[e->base[e->match[v]]] = 1;
    e->p[v] = child;
    child = e->match[v];
    v = e->p[e->match[v]];
  }
}

/* BFS from root over alternating paths; returns an exposed end or -1. */
static int32_t edm_find(Edm* e, int32_t root) {
  const int32_t n = e->n;
  memset(e->used, 0, (size_t)n);
  for (int32_t i = 0; i < n; ++i) {
    e->p[i] = -1;
    e->base[i] = i;
  }
  e->used[root] = 1;
  e->qh = e->qt = 0;
  e->queue[e->qt++] = root;
  while (e->qh < e->qt) {
    const int32_t v = e->queue[e->qh++];
    for (int32_t j = e->off[v]; j < e->off[v + 1]; ++j) {
      const int32_t to = e->col[j];
      if (to == v || e->base[v] == e->base[to] || e->match[v] == to) continue;
      if (to == root || (e->match[to] >= 0 && e->p[e->match[to]] >= 0)) {
        const int32_t cb = edm_lca(e, v, to);
        memset(e->blossom, 0, (size_t)n);
        edm_mark(e, v, cb, to);
        edm_mark(e, to, cb, v);
        for (int32_t i = 0; i < n; ++i) {
          if (e->blossom[e->base[i]]) {
            e->base[i] = cb;
            if (!e->used[i]) {
              e->used[i] = 1;
              e->queue[e->qt++] = i;
            }
          }
        }
      } else if (e->p[to] < 0) {
        e->p[to] = v;
        if (e->match[to] < 0) return to;
        e->used[e->match[to]] = 1;
        e->queue[e->qt++] = e->match[to];
      }
    }
  }
  return -1;
}

int64_t oracle_match_edmonds(int32_t n, const int32_t* off, const int32_t* col, int32_t* mate) {
  Edm e;
  e.n = n;
  e.off = off;
  e.col = col;
  e.match = mate;
  e.p = xcalloc((size_t)n, sizeof(int32_t));
  e.base = xcalloc((size_t)n, sizeof(int32_t));
  e.queue = xcalloc((size_t)n + 1, sizeof(int32_t));
  e.used = xcalloc((size_t)n, 1);
  e.blossom = xcalloc((size_t)n, 1);
  e.seen = xcalloc((size_t)n, 1);
  for (int32_t v = 0; v < n; ++v) mate[v] = -1;
  for (int32_t v = 0; v < n; ++v) {
    if (mate[v] >= 0) continue;
    int32_t u = edm_find(&e, v);
    while (u >= 0) {
      const int32_t pv = e.p[u], ppv = mate[pv];
      mate[u] = pv;
      mate[pv] = u;
      u = ppv;
    }
  }
  free(e.p);
  free(e.base);
  free(e.queue);
  free(e.used);
  free(e.blossom);
  free(e.seen);
  return count_matched(n, mate);
}

/* ---- Hopcroft-Karp -------------------------------------------------------- */

int64_t oracle_match_hk(int32_t n, const int32_t* off, const int32_t* col, int32_t* mate) {
  int32_t* color = xcalloc((size_t)n, sizeof(int32_t));
  int32_t* queue = xcalloc((size_t)n + 1, sizeof(int32_t));
  int32_t* dist = xcalloc((size_t)n, sizeof(int32_t));
  int32_t* it = xcalloc((size_t)n, sizeof(int32_t));
  int32_t* stk = xcalloc((size_t)n + 1, sizeof(int32_t));
  int32_t* via = xcalloc((size_t)n + 1, sizeof(int32_t));
  int64_t result = 0;
  for (int32_t v = 0; v < n; ++v) color[v] = -1;
  for (int32_t s = 0; s < n && result >= 0; ++s) {
    if (color[s] >= 0) continue;
    color[s] = 0;
    int32_t qh = 0, qt = 0;
    queue[qt++] = s;
    while (qh < qt && result >= 0) {
      const int32_t v = queue[qh++];
      for (int32_t j = off[v]; j < off[v + 1]; ++j) {
        const int32_t w = col[j];
        if (color[w] < 0) {
          color[w] = color[v] ^ 1;
          queue[qt++] = w;
        } else if (color[w] == color[v]) {
          result = -1;
          break;
        }
      }
    }
  }
  if (result < 0) goto done;
  for (int32_t v = 0; v < n; ++v) mate[v] = -1;
  const int32_t INF = INT32_MAX;
  for (;;) {
    int32_t qh = 0, qt = 0, found = 0;
    for (int32_t u = 0; u < n; ++u) {
      if (color[u] != 0) continue;
      if (mate[u] < 0) {
        dist[u] = 0;
        queue[qt++] = u;
      } else {
        dist[u] = INF;
      }
    }
    while (qh < qt) {
      const int32_t u = queue[qh++];
      for (int32_t j = off[u]; j < off[u + 1]; ++j) {
        const int32_t x = mate[col[j]];
        if (x < 0) {
          found = 1;
        } else if (dist[x] == INF) {
          dist[x] = dist[u] + 1;
          queue[qt++] = x;
        }
      }
    }
    if (!found) break;
    for (int32_t u = 0; u < n; ++u) it[u] = off[u];
    for (int32_t u0 = 0; u0 < n; ++u0) {
      if (color[u0] != 0 || mate[u0] >= 0) continue;
      int32_t top = 0;
      stk[0] = u0;
      while (top >= 0) {
        const int32_t x = stk[top];
        if (it[x] == off[x + 1]) {
          dist[x] = INF;
          --top;
          continue;
        }
        const int32_t w = col[it[x]++];
        const int32_t y = mate[w];
        if (y < 0) {
          via[top] = w;
          for (int32_t k = 0; k <= top; ++k) {
            mate[stk[k]] = via[k];
            mate[via[k]] = stk[k];
          }
          break;
        }
        if (dist[y] != INF && dist[y] == dist[x] + 1) {
          via[top] = w;
          stk[++top] = y;
        }
      }
    }
  }
  result = count_matched(n, mate);
done:
  free(color);
  free(queue);
  free(dist);
  free(it);
  free(stk);
  free(via);
  return result;
}
