// xb_kernels.hip -- CDNA4 (gfx950, wave64) kernels of the multi-source BFS
// engine.
//
// Two engines share the frontier-queue machinery:
//
//  * plain per-source engine: one top-down push BFS per source.  Each level is
//    one edge-balanced launch over the frontier queue: the queue carries the
//    exclusive prefix of its vertices' out-degrees, and every lane finds the
//    (frontier vertex, neighbour offset) of its edge by a rightmost-<= search
//    (the mapping of apps/xb/src/include/xbsm_gpu/xblossom_gpu_engine_10.h:
//    237-263 + apps/xb/src/include/utils/cuda_algo.h:7-20, here without a
//    separate scan pass: the prefix is produced by the compaction itself).
//    Claims use atomicCAS on the level, parents atomicMin (lowest id wins),
//    replacing the first-winner atomicExch lock of
//    apps/xb/src/xblsm_gpu/xblossom_gpu_engine_10.cu:325-343.
//
//  * batched bit-parallel engine: 64 sources per 64-bit word, T words per
//    tile.  Levels alternate between a top-down push (edge-balanced, atomicOr
//    into a next-word array, then a commit over the touched vertices) and a
//    bottom-up pull over bitmap frontiers (lane groups per vertex for light
//    vertices in id order, one wave per heavy vertex).  Level codes are kept
//    as 8 bit-sliced planes per (vertex, word).  After the last level an emit
//    pass writes the int32 level and parent tables row by row, resolving each
//    parent as the lowest-id in-neighbour whose code is one less.
//
// Frontier compaction is wave-aggregated: a ballot gives each lane its rank,
// a wave prefix sum gives its edge offset, and one 64-bit atomicAdd on a
// packed (edges << 32 | count) counter reserves both ranges
// (apps/xb/src/include/enode_queue/enode_queue.h:46-64 does one atomic per
// warp for the slot range only).
#include <hip/hip_runtime.h>

#include <cstdint>

#include "xb_internal.hpp"

namespace xb {
namespace {

constexpr int kWave = 64;
constexpr int kBlock = 256;
constexpr int kMaxBlocks = 4096;

__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }
__device__ __forceinline__ uint64_t lanemask_lt() { return (1ull << lane_id()) - 1ull; }

__device__ __forceinline__ int64_t global_wave() {
  return ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
}
__device__ __forceinline__ int64_t num_waves() { return ((int64_t)gridDim.x * blockDim.x) / kWave; }

__device__ __forceinline__ int wave_incl_scan(int x) {
  const int l = lane_id();
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    int y = __shfl_up(x, d, kWave);
    if (l >= d) x += y;
  }
  return x;
}

__device__ __forceinline__ uint64_t wave_excl_or(uint64_t x) {
  const int l = lane_id();
  uint64_t incl = x;
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    uint64_t y = __shfl_up(incl, d, kWave);
    if (l >= d) incl |= y;
  }
  uint64_t ex = __shfl_up(incl, 1, kWave);
  return l == 0 ? 0ull : ex;
}

// Wave-aggregated append of (v, deg) for lanes with pred set.  Must be called
// by every lane of the wave (wave-uniform control flow).
__device__ __forceinline__ void wave_append(bool pred, int32_t v, int32_t deg, Queue q) {
  const uint64_t mask = __ballot(pred);
  if (mask == 0) return;
  const int d = pred ? deg : 0;
  const int incl = wave_incl_scan(d);
  const int total = __shfl(incl, kWave - 1, kWave);
  const int rank = __popcll(mask & lanemask_lt());
  const int leader = __ffsll((unsigned long long)mask) - 1;
  unsigned long long old = 0;
  if (lane_id() == leader)
    old = atomicAdd(q.ctr, ((unsigned long long)(uint32_t)total << 32) | (unsigned long long)__popcll(mask));
  old = __shfl(old, leader, kWave);
  if (pred) {
    const uint32_t pos = (uint32_t)old + (uint32_t)rank;
    q.v[pos] = v;
    q.eo[pos] = (int32_t)(old >> 32) + incl - d;
  }
}

// Wave-aggregated append of v to a plain list with a 32-bit counter.
__device__ __forceinline__ void wave_push_list(bool pred, int32_t v, int32_t* list, uint32_t* ctr) {
  const uint64_t mask = __ballot(pred);
  if (mask == 0) return;
  const int rank = __popcll(mask & lanemask_lt());
  const int leader = __ffsll((unsigned long long)mask) - 1;
  uint32_t old = 0;
  if (lane_id() == leader) old = atomicAdd(ctr, (uint32_t)__popcll(mask));
  old = __shfl(old, leader, kWave);
  if (pred) list[old + rank] = v;
}

// Rightmost i in [0, qn) with eo[i] <= e (eo[0] == 0 <= e).  Zero-degree
// entries share their successor's offset and are skipped by "rightmost".
__device__ __forceinline__ int owner_of(const int32_t* __restrict__ eo, int qn, int e) {
  int lo = 0, hi = qn;
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (eo[mid] <= e)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

inline int grid_for(int64_t items, int per_block) {
  int64_t b = (items + per_block - 1) / per_block;
  if (b < 1) b = 1;
  if (b > kMaxBlocks) b = kMaxBlocks;
  return (int)b;
}

// ============================================================================
// Plain per-source engine
// ============================================================================

__global__ __launch_bounds__(kBlock) void k_plain_init(int32_t* __restrict__ level, int32_t* __restrict__ parent,
                                                       int32_t n, int32_t src) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const bool s = (i == src);
    level[i] = s ? 0 : -1;
    parent[i] = s ? src : -1;
  }
}

// One top-down level: every lane takes one frontier edge (v -> u).
__global__ __launch_bounds__(kBlock) void k_plain_push(const int32_t* __restrict__ off,
                                                       const int32_t* __restrict__ col, Queue cur, int32_t qn,
                                                       int32_t E, int32_t* __restrict__ level,
                                                       int32_t* __restrict__ parent, int32_t d, Queue nxt) {
  const int l = lane_id();
  const int64_t stride = num_waves() * kWave;
  for (int64_t e0 = global_wave() * kWave; e0 < E; e0 += stride) {
    const int e = (int)(e0 + l);
    bool fresh = false;
    int32_t u = 0, du = 0;
    if (e < E) {
      const int i = owner_of(cur.eo, qn, e);
      const int32_t v = cur.v[i];
      u = col[off[v] + (e - cur.eo[i])];
      int32_t lu = level[u];
      if (lu < 0) {
        const int32_t old = atomicCAS(&level[u], -1, d + 1);
        if (old < 0) {
          fresh = true;
          lu = d + 1;
        } else {
          lu = old;
        }
      }
      if (lu == d + 1) {
        // Parents only decrease, so a stale read can only overestimate.
        if ((uint32_t)parent[u] > (uint32_t)v) atomicMin((uint32_t*)&parent[u], (uint32_t)v);
      }
      if (fresh) du = off[u + 1] - off[u];
    }
    wave_append(fresh, u, du, nxt);
  }
}

// ============================================================================
// Batched bit-parallel engine
// ============================================================================

__device__ __forceinline__ uint64_t word_mask(int32_t cnt, int t) {
  const int nb = cnt - 64 * t;
  return nb >= 64 ? ~0ull : (nb <= 0 ? 0ull : ((1ull << nb) - 1ull));
}

// OR `bits` into the planes whose bit is set in `code`.
__device__ __forceinline__ void planes_or(uint64_t* __restrict__ p, uint32_t code, uint64_t bits) {
#pragma unroll
  for (int k = 0; k < 8; ++k)
    if ((code >> k) & 1u) p[k] |= bits;
}

template <int T>
__global__ __launch_bounds__(kBlock) void k_bp_seed(BPState s, const int32_t* __restrict__ src, int32_t cnt) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < cnt; i += gridDim.x * blockDim.x) {
    const int32_t v = src[i];
    const int t = i >> 6;
    const uint64_t bit = 1ull << (i & 63);
    const int64_t w = (int64_t)v * T + t;
    atomicOr((unsigned long long*)&s.seen[w], (unsigned long long)bit);
    atomicOr((unsigned long long*)&s.front[0][w], (unsigned long long)bit);
    atomicOr((unsigned long long*)&s.planes[w * 8 + 0], (unsigned long long)bit);  // code 1 = level 0
  }
}

// Top-down push over the frontier queue, one edge per lane.
template <int T>
__global__ __launch_bounds__(kBlock) void k_bp_push(BPState s, int cur, const int32_t* __restrict__ off,
                                                    const int32_t* __restrict__ col, Queue q, int32_t qn,
                                                    int32_t E) {
  const int l = lane_id();
  const uint64_t* __restrict__ front = s.front[cur];
  const int64_t stride = num_waves() * kWave;
  for (int64_t e0 = global_wave() * kWave; e0 < E; e0 += stride) {
    const int e = (int)(e0 + l);
    bool first = false;
    int32_t u = 0;
    if (e < E) {
      const int i = owner_of(q.eo, qn, e);
      const int32_t v = q.v[i];
      u = col[off[v] + (e - q.eo[i])];
      bool any = false;
#pragma unroll
      for (int t = 0; t < T; ++t) {
        const uint64_t f = front[(int64_t)v * T + t];
        if (!f) continue;
        uint64_t nb = f & ~s.seen[(int64_t)u * T + t];
        if (!nb) continue;
        nb &= ~s.nextw[(int64_t)u * T + t];
        if (!nb) {
          any = true;  // another lane already offered these bits
          continue;
        }
        atomicOr((unsigned long long*)&s.nextw[(int64_t)u * T + t], (unsigned long long)nb);
        any = true;
      }
      if (any && s.tflag[u] == 0u) first = (atomicExch(&s.tflag[u], 1u) == 0u);
    }
    wave_push_list(first, u, s.tlist, s.tcount);
  }
}

// Commit a push level: move next-words into seen/front, record level codes.
template <int T>
__global__ __launch_bounds__(kBlock) void k_bp_commit(BPState s, int cur, uint32_t code,
                                                      const int32_t* __restrict__ out_off, Queue nxt) {
  const int l = lane_id();
  const uint32_t cnt = *s.tcount;
  uint64_t* __restrict__ fnext = s.front[cur ^ 1];
  const int64_t stride = num_waves() * kWave;
  for (int64_t b0 = global_wave() * kWave; b0 < (int64_t)cnt; b0 += stride) {
    const int64_t idx = b0 + l;
    bool any = false;
    int32_t u = 0, du = 0;
    if (idx < (int64_t)cnt) {
      u = s.tlist[idx];
#pragma unroll
      for (int t = 0; t < T; ++t) {
        const int64_t w = (int64_t)u * T + t;
        const uint64_t nw = s.nextw[w];
        if (!nw) continue;
        s.nextw[w] = 0;
        const uint64_t old = s.seen[w];
        const uint64_t nb = nw & ~old;
        if (!nb) continue;
        s.seen[w] = old | nb;
        fnext[w] = nb;
        planes_or(&s.planes[w * 8], code, nb);
        any = true;
      }
      s.tflag[u] = 0u;
      if (any) du = out_off[u + 1] - out_off[u];
    }
    wave_append(any, u, du, nxt);
  }
}

// Bottom-up pull for light vertices (in-degree <= heavy_cut), in id order.
// T lanes per vertex, one source word each; lanes of a vertex walk the same
// in-neighbour list and stop when their word has nothing left to find.
template <int T>
__global__ __launch_bounds__(kBlock) void k_bp_pull_light(BPState s, int cur, uint32_t code, int32_t cnt,
                                                          const int32_t* __restrict__ in_off,
                                                          const int32_t* __restrict__ in_col,
                                                          const int32_t* __restrict__ out_off, int32_t n,
                                                          int32_t heavy_cut, Queue nxt) {
  constexpr int VPW = kWave / T;
  const int l = lane_id();
  const int t = l % T;
  const int slot = l / T;
  const uint64_t vmask = word_mask(cnt, t);
  const uint64_t* __restrict__ front = s.front[cur];
  uint64_t* __restrict__ fnext = s.front[cur ^ 1];
  const int64_t stride = num_waves() * VPW;
  for (int64_t vb = global_wave() * VPW; vb < n; vb += stride) {
    const int32_t v = (int32_t)(vb + slot);
    uint64_t got = 0;
    bool act = v < n;
    int32_t b = 0, e = 0;
    if (act) {
      b = in_off[v];
      e = in_off[v + 1];
      if (e - b > heavy_cut) act = false;
    }
    if (act) {
      const int64_t w = (int64_t)v * T + t;
      const uint64_t old = s.seen[w];
      uint64_t need = ~old & vmask;
      for (int32_t j = b; j < e && need; ++j) {
        const int32_t u = in_col[j];
        const uint64_t g = front[(int64_t)u * T + t] & need;
        got |= g;
        need &= ~g;
      }
      if (got) {
        s.seen[w] = old | got;
        fnext[w] = got;
        planes_or(&s.planes[w * 8], code, got);
      }
    }
    const uint64_t bal = __ballot(got != 0);
    const uint64_t grp = (T == 64) ? bal : ((bal >> (slot * T)) & ((1ull << T) - 1ull));
    const bool lead = act && t == 0 && grp != 0;
    const int32_t dv = lead ? out_off[v + 1] - out_off[v] : 0;
    wave_append(lead, v, dv, nxt);
  }
}

// Bottom-up pull for heavy vertices: one wave per vertex, 64/T in-neighbours
// per step, OR-reduced across the wave.
template <int T>
__global__ __launch_bounds__(kBlock) void k_bp_pull_heavy(BPState s, int cur, uint32_t code, int32_t cnt,
                                                          const int32_t* __restrict__ in_off,
                                                          const int32_t* __restrict__ in_col,
                                                          const int32_t* __restrict__ out_off,
                                                          const int32_t* __restrict__ heavy, int32_t n_heavy,
                                                          Queue nxt) {
  constexpr int NS = kWave / T;
  const int l = lane_id();
  const int t = l % T;
  const int slot = l / T;
  const uint64_t vmask = word_mask(cnt, t);
  const uint64_t* __restrict__ front = s.front[cur];
  uint64_t* __restrict__ fnext = s.front[cur ^ 1];
  for (int64_t h = global_wave(); h < n_heavy; h += num_waves()) {
    const int32_t v = heavy[h];
    const int32_t b = in_off[v], e = in_off[v + 1];
    const int64_t w = (int64_t)v * T + t;
    const uint64_t old = s.seen[w];
    uint64_t need = ~old & vmask;
    uint64_t got = 0;
    for (int32_t j0 = b; j0 < e; j0 += NS) {
      if (!__any(need != 0)) break;
      const int32_t j = j0 + slot;
      uint64_t g = 0;
      if (j < e && need) g = front[(int64_t)in_col[j] * T + t] & need;
#pragma unroll
      for (int x = T; x < kWave; x <<= 1) g |= __shfl_xor(g, x, kWave);
      got |= g;
      need &= ~g;
    }
    if (slot == 0 && got) {
      s.seen[w] = old | got;
      fnext[w] = got;
      planes_or(&s.planes[w * 8], code, got);
    }
    const bool vany = __ballot(slot == 0 && got != 0) != 0;
    const bool lead = (l == 0) && vany;
    const int32_t dv = lead ? out_off[v + 1] - out_off[v] : 0;
    wave_append(lead, v, dv, nxt);
  }
}

// Clear the frontier words of the vertices of a finished level.
template <int T>
__global__ __launch_bounds__(kBlock) void k_bp_clear(BPState s, int cur, const int32_t* __restrict__ qv,
                                                     int32_t qn) {
  uint64_t* __restrict__ f = s.front[cur];
  const int64_t total = (int64_t)qn * T;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
    const int32_t v = qv[i / T];
    f[(int64_t)v * T + (i % T)] = 0ull;
  }
}

// ---- emit: level/parent tables for one source word ----

// Heavy vertices: one wave per vertex; lane i tests in-neighbour j0+i, and the
// lowest lane offering a bit (exclusive prefix-OR) owns that bit's parent.
template <int T>
__global__ __launch_bounds__(kBlock) void k_bp_emit_heavy(BPState s, int32_t word, int32_t K,
                                                          const int32_t* __restrict__ in_off,
                                                          const int32_t* __restrict__ in_col,
                                                          const int32_t* __restrict__ heavy, int32_t n_heavy,
                                                          int32_t* __restrict__ side) {
  const int l = lane_id();
  for (int64_t h = global_wave(); h < n_heavy; h += num_waves()) {
    const int32_t v = heavy[h];
    const uint64_t* pv = &s.planes[((int64_t)v * T + word) * 8];
    uint64_t pk[8], tk[8];
    uint64_t reached = 0, hi = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      pk[k] = (k < K) ? pv[k] : 0ull;
      reached |= pk[k];
      if (k > 0) hi |= pk[k];
    }
    const uint64_t lvl0 = pk[0] & ~hi;
    uint64_t need = reached & ~lvl0;
    uint64_t borrow = need;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      tk[k] = pk[k] ^ borrow;
      borrow &= ~pk[k];
    }
    const int32_t b = in_off[v], e = in_off[v + 1];
    for (int32_t j0 = b; j0 < e && need; j0 += kWave) {
      const int32_t j = j0 + l;
      uint64_t eq = 0;
      int32_t u = 0;
      if (j < e) {
        u = in_col[j];
        const uint64_t* pu = &s.planes[((int64_t)u * T + word) * 8];
        eq = need;
#pragma unroll
        for (int k = 0; k < 8; ++k)
          if (k < K) eq &= ~(pu[k] ^ tk[k]);
      }
      const uint64_t ex = wave_excl_or(eq);
      uint64_t first = eq & ~ex;
      while (first) {
        const int bit = __builtin_ctzll(first);
        first &= first - 1;
        side[h * 64 + bit] = u;
      }
      const uint64_t all = __shfl(ex | eq, kWave - 1, kWave);
      need &= ~all;
    }
  }
}

// Light vertices in id order: 4 lanes per vertex, each owning 16 sources of
// the word, 16 vertices per wave, so each store instruction writes four full
// 64-byte row segments of the level and parent tables.
template <int T>
__global__ __launch_bounds__(kBlock) void k_bp_emit_light(BPState s, int32_t word, int32_t K, int32_t rows,
                                                          const int32_t* __restrict__ in_off,
                                                          const int32_t* __restrict__ in_col, int32_t n,
                                                          int32_t heavy_cut, const int32_t* __restrict__ hidx,
                                                          const int32_t* __restrict__ side,
                                                          int32_t* __restrict__ lvl_out,
                                                          int32_t* __restrict__ par_out) {
  const int l = lane_id();
  const int g = l >> 2;
  const int q = l & 3;
  const int sh = 16 * q;
  const int row0 = 64 * word + sh;
  const int64_t stride = num_waves() * 16;
  for (int64_t vb = global_wave() * 16; vb < n; vb += stride) {
    const int32_t v = (int32_t)(vb + g);
    if (v >= n) continue;
    uint32_t pk[8], tk[8];
    uint32_t reached = 0, hi = 0;
    const uint64_t* pv = &s.planes[((int64_t)v * T + word) * 8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      pk[k] = (k < K) ? (uint32_t)((pv[k] >> sh) & 0xFFFFull) : 0u;
      reached |= pk[k];
      if (k > 0) hi |= pk[k];
    }
    const uint32_t lvl0 = pk[0] & ~hi;
    uint32_t need = reached & ~lvl0;
    uint32_t borrow = need;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      tk[k] = pk[k] ^ borrow;
      borrow &= ~pk[k];
    }
    int32_t par[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) par[i] = -1;
    const int32_t b = in_off[v], e = in_off[v + 1];
    if (e - b > heavy_cut) {
      if (need) {
        const int32_t* sp = side + (int64_t)hidx[v] * 64 + sh;
#pragma unroll
        for (int i = 0; i < 16; ++i) par[i] = sp[i];
      }
    } else {
      for (int32_t j = b; j < e && need; ++j) {
        const int32_t u = in_col[j];
        const uint64_t* pu = &s.planes[((int64_t)u * T + word) * 8];
        uint32_t eq = need;
#pragma unroll
        for (int k = 0; k < 8; ++k)
          if (k < K) eq &= ~((uint32_t)((pu[k] >> sh) & 0xFFFFull) ^ tk[k]);
        if (eq) {
#pragma unroll
          for (int i = 0; i < 16; ++i) par[i] = ((eq >> i) & 1u) ? u : par[i];
          need &= ~eq;
        }
      }
    }
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const int r = row0 + i;
      if (r >= rows) break;
      uint32_t c = 0;
#pragma unroll
      for (int k = 0; k < 8; ++k) c |= ((pk[k] >> i) & 1u) << k;
      const int64_t o = (int64_t)r * n + v;
      if (lvl_out) lvl_out[o] = (int32_t)c - 1;
      par_out[o] = (c == 0) ? -1 : (c == 1 ? v : par[i]);
    }
  }
}

// Sum over (source, vertex) with level >= 0 of the vertex's out-degree.
__global__ __launch_bounds__(kBlock) void k_count_traversed(const int32_t* __restrict__ off,
                                                            const int32_t* __restrict__ levels, int64_t total,
                                                            int32_t n, unsigned long long* out) {
  unsigned long long acc = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
    if (levels[i] >= 0) {
      const int32_t v = (int32_t)(i % n);
      acc += (unsigned long long)(off[v + 1] - off[v]);
    }
  }
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) acc += __shfl_xor(acc, d, kWave);
  if (lane_id() == 0) atomicAdd(out, acc);
}

}  // namespace

// ============================================================================
// Launch wrappers
// ============================================================================

void launch_plain_init(hipStream_t st, int32_t* level, int32_t* parent, int32_t n, int32_t src) {
  hipLaunchKernelGGL(k_plain_init, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, st, level, parent, n, src);
}

void launch_plain_push(hipStream_t st, const int32_t* off, const int32_t* col, Queue cur, int32_t qn, int32_t E,
                       int32_t* level, int32_t* parent, int32_t d, Queue nxt) {
  if (E <= 0) return;
  hipLaunchKernelGGL(k_plain_push, dim3(grid_for(E, kBlock)), dim3(kBlock), 0, st, off, col, cur, qn, E, level,
                     parent, d, nxt);
}

#define XB_DISPATCH_T(T, CALL)      \
  switch (T) {                      \
    case 1: { constexpr int TT = 1; CALL; } break; \
    case 2: { constexpr int TT = 2; CALL; } break; \
    case 4: { constexpr int TT = 4; CALL; } break; \
    case 8: { constexpr int TT = 8; CALL; } break; \
    default: break;                 \
  }

void launch_bp_seed(hipStream_t st, int T, BPState s, const int32_t* src, int32_t cnt) {
  XB_DISPATCH_T(T, hipLaunchKernelGGL(k_bp_seed<TT>, dim3(grid_for(cnt, kBlock)), dim3(kBlock), 0, st, s, src,
                                      cnt));
}

void launch_bp_push(hipStream_t st, int T, BPState s, int cur, const int32_t* off, const int32_t* col, Queue q,
                    int32_t qn, int32_t E) {
  if (E <= 0) return;
  XB_DISPATCH_T(T, hipLaunchKernelGGL(k_bp_push<TT>, dim3(grid_for(E, kBlock)), dim3(kBlock), 0, st, s, cur, off,
                                      col, q, qn, E));
}

void launch_bp_commit(hipStream_t st, int T, BPState s, int cur, int32_t code, const int32_t* out_off, Queue nxt,
                      int32_t max_touched) {
  XB_DISPATCH_T(T, hipLaunchKernelGGL(k_bp_commit<TT>, dim3(grid_for(max_touched, kBlock)), dim3(kBlock), 0, st,
                                      s, cur, (uint32_t)code, out_off, nxt));
}

void launch_bp_pull(hipStream_t st, int T, BPState s, int cur, int32_t code, int32_t cnt, const int32_t* in_off,
                    const int32_t* in_col, const int32_t* out_off, int32_t n, int32_t heavy_cut,
                    const int32_t* heavy, int32_t n_heavy, Queue nxt) {
  const int vpw = kWave / T;
  const int64_t waves = ((int64_t)n + vpw - 1) / vpw;
  XB_DISPATCH_T(T, hipLaunchKernelGGL(k_bp_pull_light<TT>, dim3(grid_for(waves * kWave, kBlock)), dim3(kBlock),
                                      0, st, s, cur, (uint32_t)code, cnt, in_off, in_col, out_off, n, heavy_cut,
                                      nxt));
  if (n_heavy > 0) {
    XB_DISPATCH_T(T, hipLaunchKernelGGL(k_bp_pull_heavy<TT>, dim3(grid_for((int64_t)n_heavy * kWave, kBlock)),
                                        dim3(kBlock), 0, st, s, cur, (uint32_t)code, cnt, in_off, in_col, out_off,
                                        heavy, n_heavy, nxt));
  }
}

void launch_bp_clear(hipStream_t st, int T, BPState s, int cur, Queue q, int32_t qn) {
  if (qn <= 0) return;
  XB_DISPATCH_T(T, hipLaunchKernelGGL(k_bp_clear<TT>, dim3(grid_for((int64_t)qn * TT, kBlock)), dim3(kBlock), 0,
                                      st, s, cur, q.v, qn));
}

void launch_bp_emit(hipStream_t st, int T, BPState s, int32_t word, int32_t K, int32_t rows, const int32_t* in_off,
                    const int32_t* in_col, int32_t n, int32_t heavy_cut, const int32_t* heavy,
                    const int32_t* hidx, int32_t n_heavy, int32_t* side, int32_t* lvl_out, int32_t* par_out) {
  if (n_heavy > 0) {
    XB_DISPATCH_T(T, hipLaunchKernelGGL(k_bp_emit_heavy<TT>, dim3(grid_for((int64_t)n_heavy * kWave, kBlock)),
                                        dim3(kBlock), 0, st, s, word, K, in_off, in_col, heavy, n_heavy, side));
  }
  const int64_t waves = ((int64_t)n + 15) / 16;
  XB_DISPATCH_T(T, hipLaunchKernelGGL(k_bp_emit_light<TT>, dim3(grid_for(waves * kWave, kBlock)), dim3(kBlock), 0,
                                      st, s, word, K, rows, in_off, in_col, n, heavy_cut, hidx, side, lvl_out,
                                      par_out));
}

void launch_count_traversed(hipStream_t st, const int32_t* off, const int32_t* levels, int32_t S, int32_t n,
                            unsigned long long* out) {
  const int64_t total = (int64_t)S * n;
  hipLaunchKernelGGL(k_count_traversed, dim3(grid_for(total, kBlock)), dim3(kBlock), 0, st, off, levels, total, n,
                     out);
}

}  // namespace xb
