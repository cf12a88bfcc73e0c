// xb_kernels.hip -- CDNA4 (gfx950, wave64) kernels of the multi-source BFS
// engine.
//
// Two engines share the frontier-queue machinery:
//
//  * plain per-source engine: one top-down push BFS per source.  Each level is
//    one edge-balanced launch over the frontier queue.  The queue carries the
//    exclusive prefix of its vertices' out-degrees, so a block takes a fixed
//    chunk of edges, finds the owners of the chunk's first and last edge by a
//    rightmost-<= search, stages the owners' offsets in LDS and maps each edge
//    to (frontier vertex, neighbour offset) with an LDS search: the mapping of
//    apps/xb/src/include/xbsm_gpu/xblossom_gpu_engine_10.h:237-263 and
//    apps/xb/src/include/utils/cuda_algo.h:7-20 without a separate scan pass
//    (the prefix comes out of the compaction itself).  Claims use atomicCAS on
//    the level, parents atomicMin (lowest id wins), replacing the first-winner
//    atomicExch lock of apps/xb/src/xblsm_gpu/xblossom_gpu_engine_10.cu:325-343.
//
//  * batched bit-parallel engine: 64 sources per 64-bit word, T words per
//    tile.  Levels are either a top-down push (edge-balanced as above,
//    atomicOr into next words, then a commit over the touched vertices) or a
//    bottom-up pull over the frontier words: vertices are taken in in-degree
//    bucket order (T, 4T or 64 lanes per vertex, each lane issuing U
//    independent neighbour loads per step, OR-reduced across the vertex's
//    lanes), and hub adjacency lists are cut into fixed pieces with one wave
//    each.  Level codes are kept as 8 bit-sliced planes per (vertex, word).
//    After the last level an emit pass writes the int32 level and parent
//    tables in vertex order.
//
// Every phase is a __device__ function over a "team" of workgroups (Grid:
// block index within the team, team size).  The host-driven engine launches
// each phase as its own kernel (the team is the whole grid); the persistent
// batch kernel gives every graph of a batch its own team of workgroups and runs
// the level loop on the device, with team barriers between phases.
//
// Frontier compaction is staged in LDS: a ballot gives each lane its rank, one
// LDS atomic per wave reserves staging slots, and the block flushes with a
// block prefix sum and ONE 64-bit global atomic on a packed (edges << 32 |
// count) counter that reserves both the queue slots and the edge range
// (apps/xb/src/include/enode_queue/enode_queue.h:46-64 does one global atomic
// per warp for the slot range only).
#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>

#include <cstdint>
#include <stdexcept>
#include <string>

#include "xb_internal.hpp"
#include "../../include/xb/xb_capi.h"

namespace xb {
namespace {

constexpr int kWave = 64;
constexpr int kBlock = 256;
constexpr int kMaxBlocks = 8192;
constexpr int kPushR = 4;                // edges per thread per chunk
constexpr int kChunk = kBlock * kPushR;  // edges per block chunk
// Independent neighbour loads per lane per step of the pull and the emit.
#ifndef XB_PULL_U
#define XB_PULL_U 4
#endif
// Write phase of the light emit: 16-byte row stores (4 vertices per lane)
// where the row segment is aligned.
#ifndef XB_EMIT_VEC_STORE
#define XB_EMIT_VEC_STORE 1
#endif

#ifndef XB_EMIT_U
#define XB_EMIT_U 4
#endif
// Source words per pass of the standalone light emit (k_bp_emit_light): its
// workgroups have 256 * W threads and take chunks of 64 vertices in W words,
// so every output row segment stays 64 vertices (256 bytes) wide while one
// gathered plane line serves W words.  The persistent kernel's emit keeps
// one word per pass (256-thread workgroups).
#ifndef XB_EMIT_WORDS
#define XB_EMIT_WORDS 2
#endif
#ifndef XB_EMIT_NT
#define XB_EMIT_NT 1
#endif
// A one-launch emit of several word groups takes them in turn across the
// whole grid (all workgroups on group g, then g + 1), not chunk by chunk, so
// the neighbours' plane lines in the caches are those of one group at a time.
#ifndef XB_EMIT_GROUP_MAJOR
#define XB_EMIT_GROUP_MAJOR 1
#endif
// The persistent kernel's levels leave their frontier words set; each tile
// clears both frontier arrays at its start.
#ifndef XB_PERSIST_LAZY_FRONT_CLEAR
#define XB_PERSIST_LAZY_FRONT_CLEAR 1
#endif
constexpr int kPullU = XB_PULL_U;
constexpr int kEmitU = XB_EMIT_U;
constexpr int kStageCap = 1536;

// ---------------------------------------------------------------------------
// Team geometry
// ---------------------------------------------------------------------------
struct Grid {
  int bid;  // block index within the team
  int nb;   // blocks in the team
};

__device__ __forceinline__ Grid self_grid() { return Grid{(int)blockIdx.x, (int)gridDim.x}; }
__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }
__device__ __forceinline__ uint64_t lanemask_lt() { return (1ull << lane_id()) - 1ull; }
__device__ __forceinline__ int64_t gwave(Grid g) { return ((int64_t)g.bid * kBlock + threadIdx.x) / kWave; }
__device__ __forceinline__ int64_t gnwaves(Grid g) { return (int64_t)g.nb * (kBlock / kWave); }
__device__ __forceinline__ int64_t gthread(Grid g) { return (int64_t)g.bid * kBlock + threadIdx.x; }
__device__ __forceinline__ int64_t gnthreads(Grid g) { return (int64_t)g.nb * kBlock; }

__device__ __forceinline__ unsigned long long wave_sum64(unsigned long long x) {
#pragma unroll
  for (int d = kWave / 2; d > 0; d >>= 1) x += __shfl_xor(x, d, kWave);
  return x;
}

__device__ __forceinline__ int wave_incl_scan(int x) {
  const int l = lane_id();
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    int y = __shfl_up(x, d, kWave);
    if (l >= d) x += y;
  }
  return x;
}

// DPP quad permutation of a 64-bit value: lane 4q+i receives lane 4q+sel(i)
// of its quad (a VALU move, no LDS crossbar round trip).
constexpr int quad_sel(int a, int b, int c, int d) { return a | (b << 2) | (c << 4) | (d << 6); }
// DPP row shifts inside each 16-lane row: row_shr(n): lane i receives lane
// i - n, row_shl(n): lane i + n; a lane whose source falls outside its row
// receives 0 (bound_ctrl, so no register has to hold the 0 beforehand).
constexpr int dpp_row_shr(int n) { return 0x110 + n; }
constexpr int dpp_row_shl(int n) { return 0x100 + n; }
// A 64-bit DPP move (every lane of the wave must be active).
template <int CTRL>
__device__ __forceinline__ uint64_t dpp64(uint64_t x) {
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)x, CTRL, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)(x >> 32), CTRL, 0xF, 0xF, true);
  return ((uint64_t)(uint32_t)hi << 32) | (uint64_t)(uint32_t)lo;
}
template <int CTRL>
__device__ __forceinline__ uint64_t quad_perm64(uint64_t x) {
  return dpp64<CTRL>(x);
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

// Block-aggregated append of v to a plain list with a 32-bit counter: ONE
// global atomic per block and call (a per-wave atomic on one counter
// serializes when most waves append).  Block-uniform call.
__device__ __forceinline__ void block_push_list(bool pred, int32_t v, int32_t* list, uint32_t* ctr) {
  __shared__ uint32_t s_cnt[kBlock / kWave];
  __shared__ uint32_t s_base;
  const uint64_t mask = __ballot(pred);
  const int wv = threadIdx.x / kWave;
  if (lane_id() == 0) s_cnt[wv] = (uint32_t)__popcll(mask);
  __syncthreads();
  uint32_t off = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < kBlock / kWave; ++w) {
    if (w < wv) off += s_cnt[w];
    tot += s_cnt[w];
  }
  if (threadIdx.x == 0 && tot) s_base = atomicAdd(ctr, tot);
  __syncthreads();
  if (pred) list[s_base + off + (uint32_t)__popcll(mask & lanemask_lt())] = v;
  __syncthreads();  // s_cnt and s_base are reused by the next call
}

// Adds a block's sum of x to *dst with ONE atomic.  Block-uniform call.
__device__ __forceinline__ void block_add(unsigned long long x, unsigned long long* dst) {
  __shared__ unsigned long long s_part[kBlock / kWave];
  x = wave_sum64(x);
  if (lane_id() == 0) s_part[threadIdx.x / kWave] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
#pragma unroll
    for (int w = 0; w < kBlock / kWave; ++w) t += s_part[w];
    if (t) atomicAdd(dst, t);
  }
  __syncthreads();
}

// Rightmost i in [lo, hi) with eo[i] <= e, given eo[lo] <= e.  Zero-degree
// entries share their successor's offset and are skipped by "rightmost".
__device__ __forceinline__ int owner_in(const int32_t* __restrict__ eo, int lo, int hi, int e) {
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (eo[mid] <= e)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

// ---------------------------------------------------------------------------
// LDS: one union per block, reused phase after phase.  Both views stay under
// 20 KB so that 8 workgroups of 256 threads fit in a CU's 160 KB.
// ---------------------------------------------------------------------------
struct AppendStage {
  int32_t v[kStageCap];
  int32_t d[kStageCap];
  int cnt;
  int wsum[kBlock / kWave];
  unsigned long long base;
};

struct QueueLds {
  AppendStage sg;
  int32_t s_eo[kChunk + 1];
  int s_lohi[2];
};

// The light emit's workgroup state for W words per pass: the parent tile
// [64W rows][64 vertices + 1 column of padding] and planes 0-3 of every
// (word, vertex) of the chunk.
template <int W>
struct EmitLdsT {
  int32_t tile[64 * W * 65];
  uint64_t pl[W][4][64];  // plane k of (word w, vertex), scan -> write phase
  int64_t next;           // dynamic schedule: the block's next (word group, chunk)
};
using EmitLds = EmitLdsT<1>;

union Lds {
  QueueLds q;
  EmitLds e;
};

// One LDS union per workgroup, shared by every phase a kernel runs.
__shared__ Lds g_lds;
static_assert(sizeof(Lds) <= 20 * 1024 - 256, "LDS union must leave room for 8 workgroups per CU");

// ---- block-level LDS staging of queue appends ----
// Lanes stage (vertex, degree) entries in LDS (one LDS atomic per wave); the
// block flushes the stage to the global queue with ONE global atomic per
// flush, after a block prefix sum over the staged degrees.  Callers run in
// block-uniform loops and call stage_sync_flush() once per iteration, which
// flushes when another iteration's worth of entries might not fit.
__device__ __forceinline__ void stage_init(AppendStage& sg) {
  __syncthreads();
  if (threadIdx.x == 0) sg.cnt = 0;
  __syncthreads();
}

// Wave-uniform call.
__device__ __forceinline__ void stage_put(AppendStage& sg, bool pred, int32_t v, int32_t deg) {
  const uint64_t mask = __ballot(pred);
  if (mask == 0) return;
  const int rank = __popcll(mask & lanemask_lt());
  const int leader = __ffsll((unsigned long long)mask) - 1;
  int base = 0;
  if (lane_id() == leader) base = atomicAdd(&sg.cnt, (int)__popcll(mask));
  base = __shfl(base, leader, kWave);
  if (pred) {
    sg.v[base + rank] = v;
    sg.d[base + rank] = deg;
  }
}

// Block-uniform call.  list == nullptr: frontier queue with the degree prefix;
// otherwise a plain id list with a 32-bit counter ctr32.
__device__ void stage_flush(AppendStage& sg, int cnt, Queue q, int32_t* list, uint32_t* ctr32) {
  const int per = (cnt + kBlock - 1) / kBlock;
  const int i0 = min((int)threadIdx.x * per, cnt);
  const int i1 = min(i0 + per, cnt);
  int sum = 0;
  if (!list)
    for (int i = i0; i < i1; ++i) sum += sg.d[i];
  const int incl = wave_incl_scan(sum);
  const int wv = threadIdx.x / kWave;
  if (lane_id() == kWave - 1) sg.wsum[wv] = incl;
  __syncthreads();
  int woff = 0, total = 0;
#pragma unroll
  for (int w = 0; w < kBlock / kWave; ++w) {
    if (w < wv) woff += sg.wsum[w];
    total += sg.wsum[w];
  }
  if (threadIdx.x == 0) {
    if (list)
      sg.base = atomicAdd(ctr32, (uint32_t)cnt);
    else
      sg.base = atomicAdd(q.ctr, ((unsigned long long)(uint32_t)total << 32) | (unsigned long long)(uint32_t)cnt);
  }
  __syncthreads();
  const uint32_t bc = (uint32_t)(sg.base & 0xFFFFFFFFull);
  if (list) {
    for (int i = threadIdx.x; i < cnt; i += kBlock) list[bc + i] = sg.v[i];
  } else {
    int run = (int)(sg.base >> 32) + woff + incl - sum;
    for (int i = i0; i < i1; ++i) {
      q.v[bc + i] = sg.v[i];
      q.eo[bc + i] = run;
      run += sg.d[i];
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) sg.cnt = 0;
  __syncthreads();
}

// Block-uniform call once per loop iteration (force = last iteration).
__device__ __forceinline__ void stage_sync_flush(AppendStage& sg, bool force, Queue q, int32_t* list = nullptr,
                                                 uint32_t* ctr32 = nullptr) {
  __syncthreads();
  const int cnt = sg.cnt;
  __syncthreads();
  if (cnt == 0 || (!force && cnt <= kStageCap - kBlock)) return;
  stage_flush(sg, cnt, q, list, ctr32);
}

// Block-level owner staging for an edge chunk [c0, c1) of a frontier queue.
// On return s_eo[0..cnt) = eo[lo..lo+cnt) when cnt <= kChunk + 1 (else the
// owners are searched in global memory between lo and lo+cnt).
struct ChunkOwners {
  int lo;
  int cnt;
};

__device__ __forceinline__ ChunkOwners stage_owners(const int32_t* __restrict__ eo, int qn, int c0, int c1,
                                                    QueueLds& L) {
  if (threadIdx.x == 0) L.s_lohi[0] = owner_in(eo, 0, qn, c0);
  if (threadIdx.x == kWave) L.s_lohi[1] = owner_in(eo, 0, qn, c1 - 1);
  __syncthreads();
  ChunkOwners co{L.s_lohi[0], L.s_lohi[1] - L.s_lohi[0] + 1};
  if (co.cnt <= kChunk + 1)
    for (int i = threadIdx.x; i < co.cnt; i += blockDim.x) L.s_eo[i] = eo[co.lo + i];
  __syncthreads();
  return co;
}

__device__ __forceinline__ int chunk_owner(const ChunkOwners& co, const int32_t* __restrict__ eo,
                                           const int32_t* s_eo, int e) {
  if (co.cnt <= kChunk + 1) return co.lo + owner_in(s_eo, 0, co.cnt, e);
  return owner_in(eo, co.lo, co.lo + co.cnt, e);
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
                                                       int32_t n) {
  const int64_t n4 = n / 4;
  int4* l4 = reinterpret_cast<int4*>(level);
  int4* p4 = reinterpret_cast<int4*>(parent);
  const bool aligned = ((reinterpret_cast<uintptr_t>(level) | reinterpret_cast<uintptr_t>(parent)) & 15) == 0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t done = 0;
  if (aligned) {
    const int4 m1 = make_int4(-1, -1, -1, -1);
    for (int64_t i = tid; i < n4; i += stride) {
      l4[i] = m1;
      p4[i] = m1;
    }
    done = n4 * 4;
  }
  for (int64_t i = done + tid; i < n; i += stride) {
    level[i] = -1;
    parent[i] = -1;
  }
}

__global__ void k_plain_set_source(int32_t* __restrict__ level, int32_t* __restrict__ parent, int32_t src,
                                   Queue q) {
  if (threadIdx.x == 0) {
    level[src] = 0;
    parent[src] = src;
    q.v[0] = src;
    q.eo[0] = 0;
  }
}

// One top-down level: every thread takes kPushR frontier edges (v -> u).
__global__ __launch_bounds__(kBlock) void k_plain_push(const int32_t* __restrict__ off,
                                                       const int32_t* __restrict__ col, Queue cur, int32_t qn,
                                                       int32_t E, int32_t* __restrict__ level,
                                                       int32_t* __restrict__ parent, int32_t d, Queue nxt) {
  __shared__ QueueLds L;
  AppendStage& sg = L.sg;
  stage_init(sg);
  for (int64_t c0 = (int64_t)blockIdx.x * kChunk; c0 < E; c0 += (int64_t)gridDim.x * kChunk) {
    const int c1 = (int)((c0 + kChunk < E) ? c0 + kChunk : E);
    const ChunkOwners co = stage_owners(cur.eo, qn, (int)c0, c1, L);
#pragma unroll 1
    for (int r = 0; r < kPushR; ++r) {
      const int e = (int)c0 + r * kBlock + (int)threadIdx.x;
      bool fresh = false;
      int32_t u = 0, du = 0;
      if (e < c1) {
        const int i = chunk_owner(co, cur.eo, L.s_eo, e);
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
      stage_put(sg, fresh, u, du);
      stage_sync_flush(sg, false, nxt);
    }
  }
  stage_sync_flush(sg, true, nxt);
}

// ============================================================================
// Batched bit-parallel engine: phases
// ============================================================================

__device__ __forceinline__ uint64_t word_mask(int32_t cnt, int t) {
  const int nb = cnt - 64 * t;
  return nb >= 64 ? ~0ull : (nb <= 0 ? 0ull : ((1ull << nb) - 1ull));
}

// Words per pass of the standalone emit for tiles of T words.
template <int T>
constexpr int kEmitWords = T < XB_EMIT_WORDS ? T : XB_EMIT_WORDS;

// Level-code planes of (v, word t): groups of W = min(T, XB_PLANE_GROUP_MAX)
// words, vertex-major within a group.  With XB_PLANE_SPLIT (the default, for
// tiles of XB_PLANE_SPLIT_MIN_T words and more) the array holds two halves:
// planes 0-1 of every (group, vertex) first, then planes 2-3, so planes 0-1
// of a vertex's W words of one group are 16W contiguous bytes.  The light emit on a symmetric graph compares planes 0-1
// of every in-neighbour only: its gathers then move 16 bytes per word, not
// the 32 of planes 0-3.  Without the split, planes 0-3 of (v, t) are 32
// contiguous bytes.  Planes 4-7 (planes_hi) use the same layout.  The emit's
// words per pass must divide the group.
#ifndef XB_PLANE_GROUP_MAX
#define XB_PLANE_GROUP_MAX XB_EMIT_WORDS
#endif
#ifndef XB_PLANE_SPLIT
#define XB_PLANE_SPLIT 1
#endif
// Narrowest tile (words) that takes the split layout.  The split puts 4
// vertices' plane pairs of a group in one line where the pull writes
// consecutive vertices, but an emit reading a vertex's own planes 0-3 then
// touches two lines: on the LiveJournal-shaped graph's 16-word tile the pull
// gains more (9.9 against 10.8 ms per step), on the uneven batch's 1-8-word
// tiles the emit loses more (34.5 against 33.6 ms per replay).
#ifndef XB_PLANE_SPLIT_MIN_T
#define XB_PLANE_SPLIT_MIN_T 16
#endif
template <int T>
constexpr int kPlaneGroup = T < XB_PLANE_GROUP_MAX ? T : XB_PLANE_GROUP_MAX;
template <int T>
constexpr bool kPlaneSplit = XB_PLANE_SPLIT && T >= XB_PLANE_SPLIT_MIN_T;

// Index of plane pair `half` (planes 2*half and 2*half + 1; without the split,
// of plane 2*half) of (v, t) in a plane array.
template <int T>
__device__ __forceinline__ int64_t plane_pair_index(const BPState& s, int64_t v, int t, int half) {
  constexpr int W = kPlaneGroup<T>;
  static_assert(kPlaneGroup<T> % kEmitWords<T> == 0, "emit words per pass must divide the plane group");
  if constexpr (kPlaneSplit<T>) {
    constexpr int G = T / W;
    return ((((int64_t)half * G + t / W) * s.n + v) * W + (t % W)) * 2;
  }
  return (((int64_t)(t / W) * s.n + v) * W + (t % W)) * 4 + 2 * half;
}
// Plane k (0-7) of (v, t).
template <int T>
__device__ __forceinline__ uint64_t* plane_ptr(const BPState& s, int64_t v, int t, int k) {
  uint64_t* a = k < 4 ? s.planes : s.planes_hi;
  return a + plane_pair_index<T>(s, v, t, (k >> 1) & 1) + (k & 1);
}

// OR `bits` into the planes of (v, t) whose bit is set in `code`.  One lane
// owns a (v, t) in a level; with XB_PLANES_ATOMIC the OR is a no-return
// atomic done in the L2, so the lane does not wait for the line.
#ifndef XB_PLANES_ATOMIC
#define XB_PLANES_ATOMIC 0
#endif
__device__ __forceinline__ void or_word(uint64_t* p, uint64_t bits) {
  if (XB_PLANES_ATOMIC)
    (void)__hip_atomic_fetch_or(p, bits, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  else
    *p |= bits;
}
template <int T>
__device__ __forceinline__ void planes_or(const BPState& s, int64_t v, int t, uint32_t code, uint64_t bits) {
  if (XB_PLANES_ATOMIC) {
#pragma unroll
    for (int k = 0; k < 8; ++k)
      if ((code >> k) & 1u) or_word(plane_ptr<T>(s, v, t, k), bits);
    return;
  }
  // One vector read-modify-write per 16-byte plane pair with a set code bit
  // (per 32-byte group of planes 0-3 or 4-7 without the split), so the lane
  // waits on one load per pair instead of one per set bit.
  typedef uint64_t u64x2 __attribute__((ext_vector_type(2)));
  typedef uint64_t u64x4 __attribute__((ext_vector_type(4)));
  auto pair_or = [&](uint64_t* a, uint32_t c) {
    if constexpr (kPlaneSplit<T>) {
      if (c & 3u) {
        u64x2* p0 = reinterpret_cast<u64x2*>(a + plane_pair_index<T>(s, v, t, 0));
        const u64x2 m = {(c & 1u) ? bits : 0ull, (c & 2u) ? bits : 0ull};
        *p0 = *p0 | m;
      }
      if (c & 0xCu) {
        u64x2* p1 = reinterpret_cast<u64x2*>(a + plane_pair_index<T>(s, v, t, 1));
        const u64x2 m = {(c & 4u) ? bits : 0ull, (c & 8u) ? bits : 0ull};
        *p1 = *p1 | m;
      }
    } else if (c & 0xFu) {
      u64x4* p4 = reinterpret_cast<u64x4*>(a + plane_pair_index<T>(s, v, t, 0));
      const u64x4 m = {(c & 1u) ? bits : 0ull, (c & 2u) ? bits : 0ull, (c & 4u) ? bits : 0ull,
                       (c & 8u) ? bits : 0ull};
      *p4 = *p4 | m;
    }
  };
  pair_or(s.planes, code & 0xFu);
  if (code >> 4) pair_or(s.planes_hi, code >> 4);
}

// The first K planes of (v, t).
template <int T, int K>
__device__ __forceinline__ void load_planes(const BPState& s, int64_t v, int t, uint64_t (&pk)[K]) {
#pragma unroll
  for (int h = 0; h < (K + 1) / 2; ++h) {
    const uint64_t* a = (h < 2 ? s.planes : s.planes_hi) + plane_pair_index<T>(s, v, t, h & 1);
    pk[2 * h] = a[0];
    if (2 * h + 1 < K) pk[2 * h + 1] = a[1];
  }
}

// Tile-state fills of the persistent kernel (gigabytes per tile, read back
// from HBM by later phases) as streaming, non-temporal stores.
#ifndef XB_FILL_NT
#define XB_FILL_NT 1
#endif

// 16-byte stores over the 16-byte aligned body, 8-byte stores for the
// unaligned first word and an odd last one.
__device__ void fill_u64(Grid g, uint64_t* __restrict__ p, int64_t count, uint64_t val) {
  const int64_t head = ((uintptr_t)p & 15) && count > 0 ? 1 : 0;
  const int64_t pairs = (count - head) / 2;
  typedef uint64_t v2u __attribute__((ext_vector_type(2)));
  v2u* __restrict__ q = reinterpret_cast<v2u*>(p + head);
  const v2u vv = {val, val};
  for (int64_t i = gthread(g); i < pairs; i += gnthreads(g)) {
    if (XB_FILL_NT)
      __builtin_nontemporal_store(vv, &q[i]);
    else
      q[i] = vv;
  }
  if (gthread(g) == 0) {
    if (head) p[0] = val;
    if (head + 2 * pairs < count) p[count - 1] = val;
  }
}

// 16-byte stores over the 16-byte aligned body, 4-byte stores for the up to
// 3 words before it and after it.
__device__ void fill_u32(Grid g, uint32_t* __restrict__ p, int64_t count, uint32_t val) {
  int64_t head = (int64_t)((16u - ((uintptr_t)p & 15u)) & 15u) / 4;
  if (head > count) head = count;
  const int64_t quads = (count - head) / 4;
  typedef uint32_t v4u __attribute__((ext_vector_type(4)));
  v4u* __restrict__ q = reinterpret_cast<v4u*>(p + head);
  const v4u vv = {val, val, val, val};
  for (int64_t i = gthread(g); i < quads; i += gnthreads(g)) {
    if (XB_FILL_NT)
      __builtin_nontemporal_store(vv, &q[i]);
    else
      q[i] = vv;
  }
  const int64_t t = gthread(g);
  if (t < head) p[t] = val;
  const int64_t tail0 = head + 4 * quads;
  if (t < count - tail0) p[tail0 + t] = val;
}

template <int T>
__device__ void seed_phase(Grid g, BPState s, const int32_t* __restrict__ src, int32_t cnt) {
  for (int64_t i = gthread(g); i < cnt; i += gnthreads(g)) {
    const int32_t vo = src[i];
    const int32_t v = s.new_of[vo];
    const int t = (int)(i >> 6);
    const uint64_t bit = 1ull << (i & 63);
    const int64_t w = (int64_t)v * T + t;
    atomicOr((unsigned long long*)&s.seen[w], (unsigned long long)bit);
    atomicOr((unsigned long long*)&s.front[0][w], (unsigned long long)bit);
    atomicOr((unsigned long long*)plane_ptr<T>(s, v, t, 0), (unsigned long long)bit);  // code 1 = level 0
  }
}

// Top-down push over the frontier queue, kPushR edges per thread.
template <int T>
__device__ void push_phase(Grid g, BPState s, int cur, const int32_t* __restrict__ off,
                           const int32_t* __restrict__ col, Queue q, int32_t qn, int32_t E) {
  QueueLds& L = g_lds.q;
  AppendStage& sg = L.sg;
  stage_init(sg);
  const uint64_t* __restrict__ front = s.front[cur];
  for (int64_t c0 = (int64_t)g.bid * kChunk; c0 < E; c0 += (int64_t)g.nb * kChunk) {
    const int c1 = (int)((c0 + kChunk < E) ? c0 + kChunk : E);
    const ChunkOwners co = stage_owners(q.eo, qn, (int)c0, c1, L);
#pragma unroll 1
    for (int r = 0; r < kPushR; ++r) {
      const int e = (int)c0 + r * kBlock + (int)threadIdx.x;
      bool first = false;
      int32_t u = 0;
      if (e < c1) {
        const int i = chunk_owner(co, q.eo, L.s_eo, e);
        const int32_t v = q.v[i];
        u = col[off[v] + (e - q.eo[i])];
        // A heavy vertex takes its parents here as in the pull: every pusher
        // offers itself for each bit it brings, and the lowest original id
        // wins, so the emit need not scan its long in-list for push levels.
        const int32_t h = s.hidx[u];
        const uint32_t vo = h >= 0 ? (uint32_t)s.orig_of[v] : 0u;
        bool any = false;
#pragma unroll
        for (int t = 0; t < T; ++t) {
          const uint64_t f = front[(int64_t)v * T + t];
          if (!f) continue;
          uint64_t nb = f & ~s.seen[(int64_t)u * T + t];
          if (!nb) continue;
          any = true;
          if (h >= 0) {
            uint32_t* hp = s.hpar + ((int64_t)h * T + t) * 64;
            for (uint64_t r = nb; r; r &= r - 1) atomicMin(&hp[__builtin_ctzll(r)], vo);
          }
          nb &= ~s.nextw[(int64_t)u * T + t];
          if (nb) atomicOr((unsigned long long*)&s.nextw[(int64_t)u * T + t], (unsigned long long)nb);
        }
        if (any && s.tflag[u] == 0u) first = (atomicExch(&s.tflag[u], 1u) == 0u);
      }
      stage_put(sg, first, u, 0);
      stage_sync_flush(sg, false, Queue{}, s.tlist, s.tcount);
    }
  }
  stage_sync_flush(sg, true, Queue{}, s.tlist, s.tcount);
}

// Commit atomically gathered next words (push levels and pulled hubs): move
// them into seen/front, record level codes, append to the next queue.
template <int T>
__device__ void commit_phase(Grid g, BPState s, int cur, uint32_t code,
                             const int32_t* __restrict__ out_off, Queue nxt) {
  AppendStage& sg = g_lds.q.sg;
  stage_init(sg);
  const uint32_t cnt = __hip_atomic_load(s.tcount, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  uint64_t* __restrict__ fnext = s.front[cur ^ 1];
  for (int64_t b0 = (int64_t)g.bid * kBlock; b0 < (int64_t)cnt; b0 += (int64_t)g.nb * kBlock) {
    const int64_t idx = b0 + threadIdx.x;
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
        planes_or<T>(s, u, t, code, nb);
        any = true;
      }
      s.tflag[u] = 0u;
      if (any) du = out_off[u + 1] - out_off[u];
    }
    stage_put(sg, any, u, du);
    stage_sync_flush(sg, false, nxt);
  }
  stage_sync_flush(sg, true, nxt);
}

// Bottom-up pull over one in-degree bucket of the vertex order perm[begin,
// end).  G = T*P lanes per vertex: lane (slot, p, t) handles source word t and
// in-neighbours j = b + k + i*P + p (i < kPullU) of step k; the P partitions
// are OR-reduced after every step so every lane of the vertex sees the same
// remaining need and the group stops together.  With REC (heavy vertices,
// one wave each) the neighbours of a step are taken in ascending position and
// the first provider of every new bit -- the lowest-id parent -- is recorded
// in the heavy-parent table s.hpar.
template <int T, int P, bool REC>
__device__ void pull_bucket_phase(Grid g, BPState s, int cur, uint32_t code, int32_t cnt,
                                  const int32_t* __restrict__ in_off, const int32_t* __restrict__ in_col,
                                  const int32_t* __restrict__ out_off, const int32_t* __restrict__ perm,
                                  int32_t shift, int32_t begin, int32_t end, Queue nxt) {
  constexpr int G = T * P;
  constexpr int VPW = kWave / G;
  constexpr int VPB = (kBlock / kWave) * VPW;
  AppendStage& sg = g_lds.q.sg;
  stage_init(sg);
  const int l = lane_id();
  const int t = l % T;
  const int p = (l / T) % P;
  const int slot = l / G;
  const int wv = threadIdx.x / kWave;
  const uint64_t vmask = word_mask(cnt, t);
  const uint64_t* __restrict__ front = s.front[cur];
  uint64_t* __restrict__ fnext = s.front[cur ^ 1];
  for (int64_t bb = begin + (int64_t)g.bid * VPB; bb < end; bb += (int64_t)g.nb * VPB) {
    const int64_t idx = bb + wv * VPW + slot;
    const bool act = idx < end;
    int32_t v = 0, b = 0, e = 0;
    uint64_t old = 0, need = 0;
    uint32_t* hp = nullptr;
    if (act) {
      v = perm ? perm[idx] : (int32_t)(idx + shift);  // no list: the bucket is the id range [begin, end) + shift
      b = in_off[v];
      e = in_off[v + 1];
      old = s.seen[(int64_t)v * T + t];
      need = ~old & vmask;
      if (REC) hp = s.hpar + ((int64_t)s.hidx[v] * T + t) * 64;
    }
    uint64_t got = 0;
    for (int32_t k = 0;; k += P * kPullU) {
      const bool more = need != 0 && (b + k + p < e);
      if (!__any(more)) break;
      int32_t us[kPullU];
      uint64_t fs[kPullU];
#pragma unroll
      for (int i = 0; i < kPullU; ++i) {
        const int32_t j = b + k + i * P + p;
        us[i] = (more && j < e) ? in_col[j] : -1;
      }
#pragma unroll
      for (int i = 0; i < kPullU; ++i) fs[i] = (us[i] >= 0) ? front[(int64_t)us[i] * T + t] : 0ull;
      if (!REC) {
        uint64_t gg = 0;
#pragma unroll
        for (int i = 0; i < kPullU; ++i) gg |= fs[i];
        gg &= need;
#pragma unroll
        for (int x = T; x < G; x <<= 1) gg |= __shfl_xor(gg, x, kWave);
        got |= gg;
        need &= ~gg;
      } else {
#pragma unroll
        for (int i = 0; i < kPullU; ++i) {
          const uint64_t c = fs[i] & need;
          uint64_t incl = c;
#pragma unroll
          for (int d = 1; d < P; d <<= 1) {
            const uint64_t y = __shfl_up(incl, d * T, kWave);
            if (p >= d) incl |= y;
          }
          uint64_t excl = __shfl_up(incl, T, kWave);
          if (p == 0) excl = 0;
          uint64_t first = c & ~excl;
          if (first) {
            const uint32_t uo = (uint32_t)s.in_col_o[b + k + i * P + p];
            while (first) {
              const int bit = __builtin_ctzll(first);
              first &= first - 1;
              hp[bit] = uo;
            }
          }
          const uint64_t all = __shfl(incl, slot * G + (P - 1) * T + t, kWave);
          got |= all;
          need &= ~all;
        }
      }
    }
    if (act && p == 0 && got) {
      const int64_t w = (int64_t)v * T + t;
      s.seen[w] = old | got;
      fnext[w] = got;
      planes_or<T>(s, v, t, code, got);
    }
    const uint64_t bal = __ballot(act && p == 0 && got != 0);
    const uint64_t grp = (bal >> (slot * G)) & ((T == 64) ? ~0ull : ((1ull << T) - 1ull));
    const bool lead = act && (l % G) == 0 && grp != 0;
    // A symmetric graph's out-lists are its in-lists: the out-degree is e - b,
    // already in registers.
    const int32_t dv = lead ? (out_off == in_off ? e - b : out_off[v + 1] - out_off[v]) : 0;
    stage_put(sg, lead, v, dv);
    stage_sync_flush(sg, false, nxt);
  }
  stage_sync_flush(sg, true, nxt);
}

// Bottom-up pull over hub adjacency pieces: one wave per piece, results
// OR-ed atomically into the next words and committed by commit_phase.  Each
// piece records its first provider of every new bit with atomicMin into the
// heavy-parent table, so the lowest id over all pieces wins.
template <int T>
__device__ void pull_hub_phase(Grid g, BPState s, int cur, int32_t cnt, const int32_t* __restrict__ in_col,
                               const HubChunk* __restrict__ chunks, int32_t n_chunks) {
  constexpr int P = kWave / T;
  const int l = lane_id();
  const int t = l % T;
  const int p = l / T;
  const uint64_t vmask = word_mask(cnt, t);
  const uint64_t* __restrict__ front = s.front[cur];
  for (int64_t c = gwave(g); c < n_chunks; c += gnwaves(g)) {
    const HubChunk hc = chunks[c];
    const int64_t w = (int64_t)hc.v * T + t;
    uint32_t* hp = s.hpar + ((int64_t)s.hidx[hc.v] * T + t) * 64;
    uint64_t need = ~s.seen[w] & vmask;
    uint64_t got = 0;
    for (int32_t k = hc.jb; k < hc.je; k += P * kPullU) {
      if (!__any(need != 0)) break;
      int32_t us[kPullU];
      uint64_t fs[kPullU];
#pragma unroll
      for (int i = 0; i < kPullU; ++i) {
        const int32_t j = k + i * P + p;
        us[i] = (need && j < hc.je) ? in_col[j] : -1;
      }
#pragma unroll
      for (int i = 0; i < kPullU; ++i) fs[i] = (us[i] >= 0) ? front[(int64_t)us[i] * T + t] : 0ull;
#pragma unroll
      for (int i = 0; i < kPullU; ++i) {
        const uint64_t cbits = fs[i] & need;
        uint64_t incl = cbits;
#pragma unroll
        for (int d = 1; d < P; d <<= 1) {
          const uint64_t y = __shfl_up(incl, d * T, kWave);
          if (p >= d) incl |= y;
        }
        uint64_t excl = __shfl_up(incl, T, kWave);
        if (p == 0) excl = 0;
        uint64_t first = cbits & ~excl;
        if (first) {
          const uint32_t uo = (uint32_t)s.in_col_o[k + i * P + p];
          while (first) {
            const int bit = __builtin_ctzll(first);
            first &= first - 1;
            atomicMin(&hp[bit], uo);
          }
        }
        const uint64_t all = __shfl(incl, (P - 1) * T + t, kWave);
        got |= all;
        need &= ~all;
      }
    }
    if (p == 0 && got) atomicOr((unsigned long long*)&s.nextw[w], (unsigned long long)got);
    const bool any = __ballot(p == 0 && got != 0) != 0;
    bool first = false;
    if (l == 0 && any && s.tflag[hc.v] == 0u) first = (atomicExch(&s.tflag[hc.v], 1u) == 0u);
    wave_push_list(first, hc.v, s.tlist, s.tcount);
  }
}

// Clear the frontier words of the vertices of a finished level.
template <int T>
__device__ void clear_phase(Grid g, BPState s, int cur, const int32_t* __restrict__ qv, int32_t qn) {
  uint64_t* __restrict__ f = s.front[cur];
  const int64_t total = (int64_t)qn * T;
  for (int64_t i = gthread(g); i < total; i += gnthreads(g)) {
    const int32_t v = qv[i / T];
    f[(int64_t)v * T + (i % T)] = 0ull;
  }
}

// ---- emit: level/parent tables for one source word ----

// Level-code masks of 64 sources from K bit-sliced planes.
template <int K>
struct CodeMasks {
  uint64_t reached, lvl0, lvl1;
  uint64_t tk[K];  // planes of (code - 1), valid for bits with code >= 2
};

template <int K>
__device__ __forceinline__ CodeMasks<K> code_masks(const uint64_t (&pk)[K]) {
  CodeMasks<K> m;
  uint64_t hi = 0, hi2 = 0;
  m.reached = 0;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    m.reached |= pk[k];
    if (k >= 1) hi |= pk[k];
    if (k >= 2) hi2 |= pk[k];
  }
  m.lvl0 = pk[0] & ~hi;
  m.lvl1 = (K >= 2) ? (pk[K >= 2 ? 1 : 0] & ~pk[0] & ~hi2) : 0ull;
  uint64_t borrow = m.reached & ~m.lvl0;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    m.tk[k] = pk[k] ^ borrow;
    borrow &= ~pk[k];
  }
  return m;
}

// The value of partition 0's lane of the same (vertex, word) in the emit's
// lane layout: lanes 4q + p for one word per pass, 8q + 2p + w for two.  DPP
// moves on every lane of the wave (the caller's loop is wave-uniform).
template <int W>
__device__ __forceinline__ uint64_t from_partition0(uint64_t x, int p) {
  if constexpr (W == 1) {
    return quad_perm64<quad_sel(0, 0, 0, 0)>(x);
  } else if constexpr (W == 2) {
    const uint64_t lo = quad_perm64<quad_sel(0, 1, 0, 1)>(x);  // partitions 0-1: lanes 8q + w
    const uint64_t hi = dpp64<dpp_row_shr(4)>(lo);                // partitions 2-3: the same, 4 lanes back
    return p < 2 ? lo : hi;
  } else {
    return __shfl(x, (lane_id() & ~(4 * W - 1)) + lane_id() % W, kWave);
  }
}

// Light-vertex emit, compute phase of one chunk of 64 consecutive vertices
// (output order) and one word group (W words, 64*W rows).  Lane (vertex vi,
// partition p, word w) takes kEmitU consecutive in-neighbours of the vertex
// per step (partition p the p-th run of the step's 4 * kEmitU) and tests the
// 64 sources of its word at once.  The W word-lanes of a (vertex, partition)
// read the same neighbour's plane line.  Inside a lane the positions are
// taken in order; a prefix-OR over the 4 partitions gives the first
// (lowest-id) provider of every bit, staged in the LDS tile [row][vertex]
// (pitch 65); heavy vertices copy their recorded parents into it.  Planes 0-3
// of each (word, vertex) go to pl[w][k][vi] for the write phase.  No block
// barrier inside: lanes write disjoint tile entries.
//
// SYM (symmetric graph): an in-neighbour of a vertex reached at level L is
// reached at level L-1, L or L+1, three codes that differ modulo 4, so
// planes 0-1 alone decide whether it is one level closer: one 16-byte gather
// per in-neighbour instead of the 32 bytes of planes 0-3.
template <int T, int K, int W, bool SYM>
__device__ __forceinline__ void emit_scan_chunk(const BPState& s, int t, int64_t vb, int w, int p, int vi,
                                                const int32_t* __restrict__ in_off,
                                                const int32_t* __restrict__ in_col,
                                                const int32_t* __restrict__ in_col_orig, int32_t n, int32_t heavy_cut,
                                                int32_t* tile, uint64_t (*pl)[4][64]) {
  constexpr int PART = 4;
  constexpr int LD = 65;
  constexpr int KL = K < 4 ? K : 4;
  const int l = lane_id();
  const int rb = 64 * w;  // this lane's rows in the tile
  const int32_t v = (int32_t)(vb + vi);
  const bool act = v < n;
  uint64_t pk[K];
#pragma unroll
  for (int k = 0; k < K; ++k) pk[k] = 0;
  int32_t b = 0, e = 0, vn = 0;
  if (act) {
    vn = s.new_of[v];
    load_planes<T, K>(s, vn, t, pk);
    b = in_off[v];
    e = in_off[v + 1];
  }
  if (p == 0) {
#pragma unroll
    for (int k = 0; k < KL; ++k) pl[w][k][vi] = pk[k];
  }
  const CodeMasks<K> cm = code_masks<K>(pk);
  const bool heavy = act && (e - b > heavy_cut);
  uint64_t need = (act && !heavy) ? (cm.reached & ~cm.lvl0 & ~cm.lvl1) : 0ull;
  if (heavy) {
    const uint32_t* hp = s.hpar + ((int64_t)s.hidx[vn] * T + t) * 64;
#pragma unroll
    for (int i = 0; i < 16; ++i) tile[(rb + 16 * p + i) * LD + vi] = (int32_t)hp[16 * p + i];
  }
  // Step k0 takes list positions b + k0 .. b + k0 + 4 * kEmitU - 1; lane p
  // the kEmitU consecutive ones from b + k0 + p * kEmitU.
  for (int32_t k0 = 0;; k0 += PART * kEmitU) {
    const int32_t j0 = b + k0 + p * kEmitU;  // this lane's first position of the step
    const bool more = need != 0 && j0 < e;
    if (!__any(more)) break;
    int32_t us[kEmitU], uo[kEmitU];
    uint64_t eqs[kEmitU];
    if constexpr (W == 2) {
      // The two word lanes of a (vertex, partition) are the lane pair 2q,
      // 2q+1 and read the same in-list positions: one loads their engine
      // ids, the other their original ids, and a DPP swap hands each lane
      // the other half (one index load per position instead of two per
      // lane).  Both DPP moves run with every lane of the wave active (the
      // loop is wave-uniform), so each lane reads a live partner.
      const int mo = __builtin_amdgcn_update_dpp(0, (int)more, quad_sel(1, 0, 3, 2), 0xF, 0xF, true);
      const bool pm = more || mo != 0;
      int32_t x[kEmitU];
#pragma unroll
      for (int i = 0; i < kEmitU; ++i) x[i] = (pm && j0 + i < e) ? (w ? in_col_orig : in_col)[j0 + i] : -1;
#pragma unroll
      for (int i = 0; i < kEmitU; ++i) {
        const int32_t y = __builtin_amdgcn_update_dpp(0, x[i], quad_sel(1, 0, 3, 2), 0xF, 0xF, true);
        const bool in = more && j0 + i < e;
        us[i] = in ? (w ? y : x[i]) : -1;
        uo[i] = in ? (w ? x[i] : y) : -1;
      }
    } else {
#pragma unroll
      for (int i = 0; i < kEmitU; ++i) {
        const bool in = more && j0 + i < e;
        us[i] = in ? in_col[j0 + i] : -1;
        uo[i] = in ? in_col_orig[j0 + i] : -1;
      }
    }
    constexpr int KC = (SYM && K > 2) ? 2 : K;  // planes compared
    if constexpr (KC <= 2) {
      // The kEmitU neighbours' planes are gathered first and compared after,
      // so the gathers are in flight together instead of each waiting on its
      // own (8 VGPRs of planes here; more planes would cost occupancy).
      uint64_t pu[kEmitU][KC];
#pragma unroll
      for (int i = 0; i < kEmitU; ++i) {
#pragma unroll
        for (int k = 0; k < KC; ++k) pu[i][k] = 0;
        if (us[i] >= 0) load_planes<T, KC>(s, us[i], t, pu[i]);
      }
#pragma unroll
      for (int i = 0; i < kEmitU; ++i) {
        uint64_t eq = us[i] >= 0 ? ~0ull : 0ull;
#pragma unroll
        for (int k = 0; k < KC; ++k) eq &= ~(pu[i][k] ^ cm.tk[k]);
        eqs[i] = eq;
      }
    } else {
#pragma unroll
      for (int i = 0; i < kEmitU; ++i) {
        uint64_t eq = 0;
        if (us[i] >= 0) {
          uint64_t pu[KC];
          load_planes<T, KC>(s, us[i], t, pu);
          eq = ~0ull;
#pragma unroll
          for (int k = 0; k < KC; ++k) eq &= ~(pu[k] ^ cm.tk[k]);
        }
        eqs[i] = eq;
      }
    }
    // Rows each of the lane's positions provides before its earlier ones,
    // in list order inside the lane; then one prefix-OR over the partitions
    // (lower partitions hold lower positions) removes the rows an earlier
    // partition provides.
    uint64_t fl[kEmitU], run = 0;
#pragma unroll
    for (int i = 0; i < kEmitU; ++i) {
      const uint64_t c = eqs[i] & need;
      fl[i] = c & ~run;
      run |= c;
    }
    uint64_t incl = run, excl, all;
    if constexpr (W == 1) {
      // The 4 partitions of a vertex are one lane quad.
      incl |= quad_perm64<quad_sel(0, 0, 1, 2)>(incl);  // lane p <- p-1 (lane 0 <- itself)
      incl |= quad_perm64<quad_sel(0, 1, 0, 1)>(incl);  // lane p <- p-2 (lanes 0, 1 <- themselves)
      excl = quad_perm64<quad_sel(0, 0, 1, 2)>(incl);
      if (p == 0) excl = 0;
      all = quad_perm64<quad_sel(3, 3, 3, 3)>(incl);
    } else if constexpr (W == 2) {
      // The 8 lanes (p, w) of a vertex are an aligned 8-lane group, w the
      // lane parity: DPP row shifts by 2 and 4 lanes move between
      // partitions of the same word, with no LDS crossbar round trip.
      // Every DPP move runs on all lanes; the selects apply the partition
      // bounds after.
      const uint64_t y1 = dpp64<dpp_row_shr(2)>(incl);  // partition p-1
      incl |= (p >= 1) ? y1 : 0ull;
      const uint64_t y2 = dpp64<dpp_row_shr(4)>(incl);  // partition p-2
      incl |= (p >= 2) ? y2 : 0ull;
      excl = dpp64<dpp_row_shr(2)>(incl);
      if (p == 0) excl = 0;
      // Partition 3's total: within each quad take quad lanes 2-3 (the
      // upper quad of the group holds partitions 2-3), then the lower quad
      // reads the upper one 4 lanes on.
      const uint64_t q3 = quad_perm64<quad_sel(2, 3, 2, 3)>(incl);
      const uint64_t q3n = dpp64<dpp_row_shl(4)>(q3);
      all = (p >= 2) ? q3 : q3n;
    } else {
      uint64_t y = __shfl_up(incl, W, kWave);
      if (p >= 1) incl |= y;
      y = __shfl_up(incl, 2 * W, kWave);
      if (p >= 2) incl |= y;
      excl = __shfl_up(incl, W, kWave);
      if (p == 0) excl = 0;
      all = __shfl(incl, (l & ~(PART * W - 1)) | ((PART - 1) * W + w), kWave);
    }
    // Rows a position resolves first: one LDS write each.  Partition 0's
    // positions (the step's lowest, which resolve most rows: its first one
    // is the vertex's lowest-id in-neighbour at step 0) are shared out by
    // row range: their masks and ids go to all 4 partition lanes of the
    // vertex and word, and lane p writes their rows 16p .. 16p + 15.  The
    // other partitions write their own rows, the two 32-bit halves in loops
    // of their own.  A wave's loops run as long as its busiest lane, so this
    // halves the iterations on the LiveJournal-shaped graph.
    int32_t* col = tile + rb * LD + vi;
#pragma unroll
    for (int i = 0; i < kEmitU; ++i) {
      const uint64_t first = fl[i] & ~excl;
      const uint64_t f0 = from_partition0<W>(first, p);
      const int32_t id0 = (int32_t)from_partition0<W>((uint64_t)(uint32_t)uo[i], p);
      for (uint32_t f = (uint32_t)(f0 >> (16 * p)) & 0xFFFFu; f; f &= f - 1)
        col[(16 * p + __builtin_ctz(f)) * LD] = id0;
      if (p != 0) {
        for (uint32_t f = (uint32_t)first; f; f &= f - 1) col[__builtin_ctz(f) * LD] = uo[i];
        for (uint32_t f = (uint32_t)(first >> 32); f; f &= f - 1) col[(32 + __builtin_ctz(f)) * LD] = uo[i];
      }
    }
    need &= ~all;
  }
}

// Light-vertex emit, write phase of one 64-vertex chunk of one word (W = 1,
// K <= 4), by 256 lanes: lane (row group, quad c4) writes 4 consecutive
// vertices of a row as one 16-byte non-temporal store when the row segment
// is aligned, else one by one; a wave writes 4 rows x 256 contiguous bytes
// per store instruction.  The level codes of the lane's 4 vertices in its 4
// rows are packed from the LDS planes once.
// Output-table store of the light emit (streaming, non-temporal by default).
template <class V>
__device__ __forceinline__ void out_store(V v, V* p) {
  if (XB_EMIT_NT)
    __builtin_nontemporal_store(v, p);
  else
    *p = v;
}

template <int K>
__device__ __forceinline__ void emit_write_chunk(int tid, int64_t vb, int32_t grp, int32_t rows,
                                                 const int32_t* __restrict__ src, int32_t n, const int32_t* tile,
                                                 const uint64_t (*pl)[64], int32_t* __restrict__ lvl_out,
                                                 int32_t* __restrict__ par_out, RowMap map) {
  static_assert(K <= 4, "the packed write covers codes below 16");
  constexpr int LD = 65;
  constexpr int RSTEP = 16;  // rows per pass: 256 lanes / 16 quads per row
  const int c4 = 4 * (tid % 16);
  const int row0 = tid / 16;
  // Level codes of vertex c4 + a in the lane's rows row0 + 16 * it: the
  // 4-bit code of row it is bits 16 * it .. 16 * it + 3 of cm[a] (plane k's
  // bits of those rows, masked out of the plane word in one step each).
  constexpr uint64_t kRowBits = 0x0001000100010001ull;
  uint64_t cm[4];
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    uint64_t m = 0;
#pragma unroll
    for (int k = 0; k < K; ++k) m |= ((pl[k][c4 + a] >> row0) & kRowBits) << k;
    cm[a] = m;
  }
#pragma unroll 1
  for (int it = 0; it < 64 / RSTEP; ++it) {
    const int rr = row0 + RSTEP * it;
    const int r = 64 * grp + rr;
    if (r >= rows) break;
    const int32_t sr = src[r];  // the parent of level-1 cells: one load per row, not one per cell
    int32_t lv[4], pv[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      const uint32_t c = (uint32_t)cm[a] & 0xFu;
      cm[a] >>= RSTEP;
      const int32_t va = (int32_t)(vb + c4 + a);
      const int32_t tv = tile[rr * LD + c4 + a];  // read for every cell: selects, no branches
      lv[a] = (int32_t)c - 1;
      pv[a] = c >= 3 ? tv : (c == 2 ? sr : (c == 1 ? va : -1));
    }
    const int64_t v0 = vb + c4;
    const int64_t lo = (map.lvl_row ? map.lvl_row[r] : (int64_t)r * n) + v0;
    const int64_t po = (map.par_row ? map.par_row[r] : (int64_t)r * n) + v0;
    if (v0 + 3 < n && ((lo | po) & 3) == 0) {
      typedef int32_t v4i __attribute__((ext_vector_type(4)));
      if (lvl_out) {
        const v4i lq = {lv[0], lv[1], lv[2], lv[3]};
        out_store(lq, reinterpret_cast<v4i*>(lvl_out + lo));
      }
      const v4i pq = {pv[0], pv[1], pv[2], pv[3]};
      out_store(pq, reinterpret_cast<v4i*>(par_out + po));
    } else {
#pragma unroll
      for (int a = 0; a < 4; ++a) {
        if (v0 + a < n) {
          if (lvl_out) out_store(lv[a], &lvl_out[lo + a]);
          out_store(pv[a], &par_out[po + a]);
        }
      }
    }
  }
}

// Light vertices, one word group (W words, 64*W rows) per pass, by
// workgroups of 256*W threads; a workgroup takes chunks of 64 consecutive
// vertices: emit_scan_chunk, then the write phase -- emit_write_chunk per
// word (256 lanes each) for codes below 16, else 64 consecutive lanes write
// one row's 64 vertices (256 contiguous bytes), 4W rows per store.
//
// Static schedule (DYN false): chunk bid, bid + nb, ... of word group grp.
// Dynamic (DYN): the team takes (word group, chunk) pairs of ngrp groups from
// the shared counter *dyn, so no workgroup idles behind a slower one.
template <int T, int K, int W, bool DYN, bool SYM>
__device__ void emit_light_phase(Grid g, EmitLdsT<W>& L, BPState s, int32_t grp, int32_t rows,
                                 const int32_t* __restrict__ src, const int32_t* __restrict__ in_off,
                                 const int32_t* __restrict__ in_col, const int32_t* __restrict__ in_col_orig,
                                 int32_t n, int32_t heavy_cut, int32_t* __restrict__ lvl_out,
                                 int32_t* __restrict__ par_out, RowMap map, unsigned* dyn, int32_t ngrp) {
  constexpr int NT = kBlock * W;  // threads per workgroup
  constexpr int PART = 4;
  constexpr int LD = 65;  // tile row pitch
  constexpr int KL = K < 4 ? K : 4;
  static_assert(kPlaneGroup<T> % W == 0, "emit words per pass must divide the plane group");
  // Dynamic schedule: a workgroup takes a chunk of a whole plane group and
  // passes its SUB word groups back to back, so the later passes find the
  // in-neighbours' plane lines (shared by the group's words) in the L2.
  constexpr int SUB = DYN ? kPlaneGroup<T> / W : 1;
  // Static schedule: ngrp word groups from grp in one launch, group by group
  // across the grid (XB_EMIT_GROUP_MAJOR) or back to back per chunk.
  const int32_t nsub = DYN ? SUB : (ngrp > 1 ? ngrp : 1);
  const int tid = threadIdx.x;
  const int w = tid % W;
  const int p = (tid / W) % PART;
  const int vi = tid / (W * PART);
  const int wcol = tid % 64;     // write phase: vertex column
  const int wrow = tid / 64;     // write phase: first row
  const int64_t nchunks = ((int64_t)n + 63) / 64;
  const int32_t nsup = (ngrp + SUB - 1) / SUB;  // plane groups (dynamic schedule)
  __syncthreads();
  for (int64_t it = g.bid;; it += g.nb) {
    int64_t vb;
    int32_t grp0 = grp;
    if (DYN) {
      if (tid == 0) L.next = (int64_t)atomicAdd(dyn, 1u);
      __syncthreads();
      const int64_t c = L.next;
      if (c >= (int64_t)nsup * nchunks) break;
      grp0 = (int32_t)(c / nchunks) * SUB;
      vb = (c % nchunks) * 64;
    } else if (XB_EMIT_GROUP_MAJOR && nsub > 1) {
      // Word groups in turn across the whole grid: item it is chunk
      // it % nchunks of group grp + it / nchunks.
      if (it >= (int64_t)nsub * nchunks) break;
      grp0 = grp + (int32_t)(it / nchunks);
      vb = (it % nchunks) * 64;
    } else {
      vb = it * 64;
      if (vb >= n) break;
    }
    const int32_t npass = (!DYN && XB_EMIT_GROUP_MAJOR) ? 1 : nsub;
    for (int sub = 0; sub < npass; ++sub) {
      const int32_t gq = grp0 + sub;  // this pass's word group
      if (DYN && gq >= ngrp) break;
      const int t = gq * W + w;       // this lane's word of the tile
      emit_scan_chunk<T, K, W, SYM>(s, t, vb, w, p, vi, in_off, in_col, in_col_orig, n, heavy_cut, L.tile, L.pl);
      __syncthreads();
#if XB_EMIT_VEC_STORE
      if constexpr (K <= 4) {
        const int w2 = tid / kBlock;
        emit_write_chunk<K>(tid % kBlock, vb, gq * W + w2, rows, src, n, L.tile + w2 * 64 * LD, L.pl[w2], lvl_out,
                            par_out, map);
        __syncthreads();
        continue;
      }
#endif
      const int32_t v2 = (int32_t)(vb + wcol);
      if (v2 < n) {
        int wq = -1;
        uint64_t q[K];
#pragma unroll 4
        for (int rr = wrow; rr < 64 * W; rr += NT / 64) {
          const int w2 = rr >> 6;
          const int bit = rr & 63;
          const int r = 64 * (gq * W + w2) + bit;
          if (r >= rows) break;
          if (w2 != wq) {
            wq = w2;
            if (K <= 4) {
#pragma unroll
              for (int k = 0; k < KL; ++k) q[k] = L.pl[w2][k][wcol];
            } else {
              load_planes<T, K>(s, s.new_of[v2], gq * W + w2, q);
            }
          }
          uint32_t c = 0;
#pragma unroll
          for (int k = 0; k < K; ++k) c |= (uint32_t)((q[k] >> bit) & 1ull) << k;
          const int32_t par = (c == 0) ? -1 : (c == 1 ? v2 : (c == 2 ? src[r] : L.tile[rr * LD + wcol]));
          const int64_t lo = (map.lvl_row ? map.lvl_row[r] : (int64_t)r * n) + v2;
          const int64_t po = (map.par_row ? map.par_row[r] : (int64_t)r * n) + v2;
          if (lvl_out) __builtin_nontemporal_store((int32_t)c - 1, &lvl_out[lo]);
          __builtin_nontemporal_store(par, &par_out[po]);
        }
      }
      __syncthreads();
    }
  }
}

// ============================================================================
// Batched bit-parallel engine: standalone kernels (host-driven level loop)
// ============================================================================

template <int T>
__global__ __launch_bounds__(kBlock) void k_bp_seed(BPState s, const int32_t* __restrict__ src, int32_t cnt) {
  seed_phase<T>(self_grid(), s, src, cnt);
}

template <int T>
__global__ __launch_bounds__(kBlock) void k_bp_push(BPState s, int cur, const int32_t* __restrict__ off,
                                                    const int32_t* __restrict__ col, Queue q, int32_t qn,
                                                    int32_t E) {
  push_phase<T>(self_grid(), s, cur, off, col, q, qn, E);
}

template <int T>
__global__ __launch_bounds__(kBlock) void k_bp_commit(BPState s, int cur, uint32_t code,
                                                      const int32_t* __restrict__ out_off, Queue nxt) {
  commit_phase<T>(self_grid(), s, cur, code, out_off, nxt);
}

// Minimum workgroups per CU the pull's register budget is sized for: 8 (64
// VGPRs, 8 waves per SIMD) hides more of its dependent gathers than the 7 its
// natural 65-69 VGPRs allow (LiveJournal-shaped graph: 13.2 -> 12.7 ms of pull
// per 1,024-source step).
#ifndef XB_PULL_MIN_BLOCKS
#define XB_PULL_MIN_BLOCKS 8
#endif

template <int T, int P, bool REC>
__global__ __launch_bounds__(kBlock, XB_PULL_MIN_BLOCKS) void k_bp_pull_bucket(BPState s, int cur, uint32_t code, int32_t cnt,
                                                           const int32_t* __restrict__ in_off,
                                                           const int32_t* __restrict__ in_col,
                                                           const int32_t* __restrict__ out_off,
                                                           const int32_t* __restrict__ perm, int32_t shift,
                                                           int32_t begin, int32_t end, Queue nxt) {
  pull_bucket_phase<T, P, REC>(self_grid(), s, cur, code, cnt, in_off, in_col, out_off, perm, shift, begin, end,
                               nxt);
}

template <int T>
__global__ __launch_bounds__(kBlock) void k_bp_pull_hub(BPState s, int cur, int32_t cnt,
                                                        const int32_t* __restrict__ in_col,
                                                        const HubChunk* __restrict__ chunks, int32_t n_chunks) {
  pull_hub_phase<T>(self_grid(), s, cur, cnt, in_col, chunks, n_chunks);
}

// One pull level in one launch, heaviest work first: workgroups [0, nb_hub)
// take the hub pieces, the next nb[2] bucket 2 (one wave per vertex), then
// nb[1] bucket 1 and the rest bucket 0.  The four touch disjoint vertices, so
// they run side by side, and the level ends in one grid tail instead of four.
struct PullSplit {
  int32_t nb[3];
  int32_t nb_hub;
};
template <int T>
__global__ __launch_bounds__(kBlock, XB_PULL_MIN_BLOCKS) void k_bp_pull_fused(
    BPState s, int cur, uint32_t code, int32_t cnt, const int32_t* __restrict__ in_off,
    const int32_t* __restrict__ in_col, const int32_t* __restrict__ out_off, PullPlan plan, PullSplit split,
    Queue nxt, const HubChunk* __restrict__ chunks, int32_t n_chunks) {
  constexpr int P1 = T * 4 <= kWave ? 4 : kWave / T;
  constexpr int P2 = kWave / T;
  int bid = (int)blockIdx.x;
  if (bid < split.nb_hub) {
    pull_hub_phase<T>(Grid{bid, split.nb_hub}, s, cur, cnt, in_col, chunks, n_chunks);
    return;
  }
  bid -= split.nb_hub;
  if (bid < split.nb[2]) {
    pull_bucket_phase<T, P2, true>(Grid{bid, split.nb[2]}, s, cur, code, cnt, in_off, in_col, out_off, plan.perm[2],
                                   plan.shift[2], plan.begin[2], plan.end[2], nxt);
    return;
  }
  bid -= split.nb[2];
  if (bid < split.nb[1]) {
    pull_bucket_phase<T, P1, false>(Grid{bid, split.nb[1]}, s, cur, code, cnt, in_off, in_col, out_off,
                                    plan.perm[1], plan.shift[1], plan.begin[1], plan.end[1], nxt);
    return;
  }
  bid -= split.nb[1];
  pull_bucket_phase<T, 1, false>(Grid{bid, split.nb[0]}, s, cur, code, cnt, in_off, in_col, out_off, plan.perm[0],
                                 plan.shift[0], plan.begin[0], plan.end[0], nxt);
}

template <int T>
__global__ __launch_bounds__(kBlock) void k_bp_clear(BPState s, int cur, const int32_t* __restrict__ qv,
                                                     int32_t qn) {
  clear_phase<T>(self_grid(), s, cur, qv, qn);
}

// Waves per SIMD the standalone emit's register budget is sized for: 8 (64
// VGPRs) hides more of its dependent gathers than the 6 waves per SIMD its
// natural 76 VGPRs allow (LiveJournal-shaped graph: 27.5 -> 24.8 ms of emit
// per 1,024-source step).
#ifndef XB_EMIT_WAVES_PER_SIMD
#define XB_EMIT_WAVES_PER_SIMD 8
#endif

// Standalone light emit of word group `grp` (words grp*W .. grp*W + W - 1,
// W = kEmitWords<T>); its LDS is its own, not the shared union.  SYM: the
// graph is symmetric (two-plane neighbour compare).  The bound is in waves
// per SIMD: 8 (64 VGPRs) for codes below 16, which the 2-word kernels meet
// with at most 20 bytes of spill (a 65th VGPR would leave 7 waves per SIMD:
// LiveJournal-shaped graph 18.6 -> 20.3 ms of emit per 1,024-source step);
// the deep-level 2-word kernels get 4, taking 72-98 VGPRs instead of
// spilling.
template <int T, int K, bool SYM>
__global__ __launch_bounds__(kBlock * kEmitWords<T>, XB_EMIT_WAVES_PER_SIMD / (K <= 4 ? 1 : kEmitWords<T>)) void k_bp_emit_light(
    BPState s, int32_t grp, int32_t ngrp, int32_t rows, const int32_t* __restrict__ src,
    const int32_t* __restrict__ in_off, const int32_t* __restrict__ in_col, const int32_t* __restrict__ in_col_orig,
    int32_t n, int32_t heavy_cut, int32_t* __restrict__ lvl_out, int32_t* __restrict__ par_out, RowMap map,
    const int32_t* __restrict__ gate, int32_t gate_lo, int32_t gate_hi) {
  if (gate != nullptr) {
    const int32_t status = gate[0], levels = gate[1];
    if (status != 0 || levels < gate_lo || levels > gate_hi) return;
  }
  __shared__ EmitLdsT<kEmitWords<T>> lds;
  emit_light_phase<T, K, kEmitWords<T>, false, SYM>(self_grid(), lds, s, grp, rows, src, in_off, in_col,
                                                    in_col_orig, n, heavy_cut, lvl_out, par_out, map, nullptr, ngrp);
}

// ============================================================================
// Persistent batch kernel: one team of workgroups per graph, device-side
// level loop, team barriers between phases
// ============================================================================

// Team barrier over the nblocks workgroups of a team.
//
// Arrival is two-level: a workgroup arrives on its XCD's counter (XCD read
// from the hardware XCC_ID register, so nothing assumes how workgroups are
// dispatched); the last arriver of each XCD writes back that XCD's L2 and
// arrives on the team counter; barrier k opens once (k+1) * (XCDs of the
// team) XCD arrivals are in.  One counter per XCD, each in its own 256-byte
// slot, keeps the arrivals of 256 workgroups per XCD from serializing behind
// each other on one address, and one L2 write-back per XCD replaces one per
// workgroup (2048 workgroups: 28 us per barrier against 145 us).
//
// Visibility across XCDs (each XCD has its own L2, not coherent with the
// others for ordinary device memory):
//  * every wave waits for its own stores (s_waitcnt 0), so after the
//    workgroup barrier all of the workgroup's stores are in its XCD's L2
//    before it arrives;
//  * the last arriver of an XCD has seen every arrival of that XCD, so its
//    agent-scope release writes back all of them before it arrives;
//  * waiting polls with relaxed agent-scope loads (an acquire per poll would
//    invalidate the XCD's L2 under every workgroup still working), then one
//    agent-scope acquire per workgroup drops stale lines.
// Waiting is bounded: after kSpinLimit polls (about 4 s) the team sets the
// batch error flag and every workgroup returns at its next barrier, so the
// kernel always terminates.  Workgroup 0 of the team charges the time since
// the previous barrier to `phase` (an XB_K_* class); the barrier opens only
// when the slowest workgroup arrives, so the charge is the phase's critical
// path on the team.
constexpr long kSpinLimit = 1L << 22;

struct TeamSync {
  TeamCtl* ctl;
  int xcd;                    // this workgroup's XCD
  unsigned long long members; // the team's workgroups on that XCD
  unsigned long long xcds;    // XCDs with workgroups of the team
};

__device__ __forceinline__ int xcc_id() {
  unsigned x;
  asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(x));
  return (int)(x & (kMaxXcd - 1));
}

__device__ __forceinline__ bool spin_at_least(unsigned long long* p, unsigned long long target, unsigned* gerr) {
  long spins = 0;
  while (__hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
    __builtin_amdgcn_s_sleep(2);
    if (++spins > kSpinLimit) {
      atomicOr(gerr, 1u);
      return false;
    }
    if ((spins & 1023) == 0 && __hip_atomic_load(gerr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return false;
  }
  return true;
}

// Launch-time registration: every workgroup of the team counts itself on its
// XCD, then all wait for the whole team (a flat barrier, once per launch).
__device__ bool team_join(TeamCtl* ctl, Grid g, unsigned* gerr, TeamSync* ts) {
  __shared__ int s_ok;
  __shared__ unsigned long long s_members, s_xcds;
  const int xc = xcc_id();
  if (threadIdx.x == 0) {
    int ok = 1;
    unsigned long long members = 1, xcds = 1;
    if (g.nb > 1) {
      (void)__hip_atomic_fetch_add(&ctl->xcd[xc].members, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __builtin_amdgcn_s_waitcnt(0);  // the XCD count lands before the join is seen
      (void)__hip_atomic_fetch_add(&ctl->joined, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      ok = spin_at_least(&ctl->joined, (unsigned long long)g.nb, gerr) ? 1 : 0;
      members = __hip_atomic_load(&ctl->xcd[xc].members, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      xcds = 0;
      for (int i = 0; i < kMaxXcd; ++i)
        xcds += __hip_atomic_load(&ctl->xcd[i].members, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ? 1ull : 0ull;
    }
    s_ok = ok;
    s_members = members;
    s_xcds = xcds;
  }
  __syncthreads();
  *ts = TeamSync{ctl, xc, s_members, s_xcds};
  const bool ok = s_ok != 0;
  __syncthreads();
  return ok;
}

__device__ bool team_barrier(const TeamSync& ts, Grid g, unsigned* gerr, int phase) {
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  __shared__ int s_ok;
  TeamCtl* ctl = ts.ctl;
  if (threadIdx.x == 0) {
    int ok = 1;
    if (g.nb > 1) {
      const unsigned long long old =
          __hip_atomic_fetch_add(&ctl->xcd[ts.xcd].arrivals, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const unsigned long long k = old / ts.members;  // index of this barrier
      if ((old + 1ull) % ts.members == 0ull) {
        // Last of this XCD: every workgroup of the XCD has its stores in L2.
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __builtin_amdgcn_s_waitcnt(0);  // the write-back completes before the arrival is seen
        (void)__hip_atomic_fetch_add(&ctl->done, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
      ok = spin_at_least(&ctl->done, (k + 1ull) * ts.xcds, gerr) ? 1 : 0;
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    if (__hip_atomic_load(gerr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) ok = 0;
    s_ok = ok;
    if (g.bid == 0) {
      const unsigned long long now = (unsigned long long)wall_clock64();
      ctl->tphase[phase] += now - ctl->tlast;
      ctl->tlast = now;
    }
  }
  __syncthreads();
  return s_ok != 0;
}

// Entry points of the phases for the persistent kernel.  They are inlined:
// as out-of-line calls every call passed the phase state (BPState, Queue) by
// value through per-lane scratch and spilled the caller's live values per
// lane, a round trip of scratch traffic for all 524,288 lanes of the team at
// every phase (LiveJournal-shaped 16-word tile: 39.4 -> 38.0 ms).  Values the
// tile loop derives from memory may sit in vector registers although every
// lane holds the same value; the wrappers move them to scalar registers
// (readfirstlane) so the phase code addresses memory as it does in a
// standalone kernel, where they are kernel arguments.
__device__ __forceinline__ int32_t uni(int32_t v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ uint32_t uni(uint32_t v) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)v); }
__device__ __forceinline__ int64_t uni(int64_t v) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)((uint64_t)v >> 32));
  return (int64_t)(((uint64_t)hi << 32) | lo);
}
template <class P>
__device__ __forceinline__ P* uni(P* p) {
  return reinterpret_cast<P*>((uintptr_t)uni((int64_t)(uintptr_t)p));
}
__device__ __forceinline__ Grid uni(Grid g) { return Grid{uni(g.bid), uni(g.nb)}; }
__device__ __forceinline__ Queue uni(Queue q) { return Queue{uni(q.v), uni(q.eo), uni(q.ctr)}; }
__device__ __forceinline__ BPState uni(const BPState& s) {
  BPState u;
  u.n = uni(s.n);
  u.seen = uni(s.seen);
  u.front[0] = uni(s.front[0]);
  u.front[1] = uni(s.front[1]);
  u.nextw = uni(s.nextw);
  u.planes = uni(s.planes);
  u.planes_hi = uni(s.planes_hi);
  u.tflag = uni(s.tflag);
  u.tlist = uni(s.tlist);
  u.tcount = uni(s.tcount);
  u.hpar = uni(s.hpar);
  u.hidx = uni(s.hidx);
  u.new_of = uni(s.new_of);
  u.orig_of = uni(s.orig_of);
  u.in_col_o = uni(s.in_col_o);
  return u;
}

template <int T>
__device__ __forceinline__ void nl_seed(Grid g, BPState s, const int32_t* src, int32_t cnt) {
  seed_phase<T>(uni(g), uni(s), uni(src), uni(cnt));
}
template <int T>
__device__ __forceinline__ void nl_push(Grid g, BPState s, int cur, const int32_t* off, const int32_t* col, Queue q,
                                     int32_t qn, int32_t E) {
  push_phase<T>(uni(g), uni(s), uni(cur), uni(off), uni(col), uni(q), uni(qn), uni(E));
}
template <int T>
__device__ __forceinline__ void nl_commit(Grid g, BPState s, int cur, uint32_t code, const int32_t* out_off,
                                       Queue nxt) {
  commit_phase<T>(uni(g), uni(s), uni(cur), uni(code), uni(out_off), uni(nxt));
}
template <int T, int P, bool REC>
__device__ __forceinline__ void nl_pull_bucket(Grid g, BPState s, int cur, uint32_t code, int32_t cnt,
                                            const int32_t* in_off, const int32_t* in_col, const int32_t* out_off,
                                            const int32_t* perm, int32_t shift, int32_t begin, int32_t end,
                                            Queue nxt) {
  pull_bucket_phase<T, P, REC>(uni(g), uni(s), uni(cur), uni(code), uni(cnt), uni(in_off), uni(in_col), uni(out_off),
                               uni(perm), uni(shift), uni(begin), uni(end), uni(nxt));
}
template <int T>
__device__ __forceinline__ void nl_pull_hub(Grid g, BPState s, int cur, int32_t cnt, const int32_t* in_col,
                                         const HubChunk* chunks, int32_t n_chunks) {
  pull_hub_phase<T>(uni(g), uni(s), uni(cur), uni(cnt), uni(in_col), uni(chunks), uni(n_chunks));
}
template <int T>
__device__ __forceinline__ void nl_clear(Grid g, BPState s, int cur, const int32_t* qv, int32_t qn) {
  clear_phase<T>(uni(g), uni(s), uni(cur), uni(qv), uni(qn));
}
__device__ __forceinline__ void nl_fill_u64(Grid g, uint64_t* p, int64_t count, uint64_t val) {
  fill_u64(uni(g), uni(p), uni(count), (uint64_t)uni((int64_t)val));
}
__device__ __forceinline__ void nl_fill_u32(Grid g, uint32_t* p, int64_t count, uint32_t val) {
  fill_u32(uni(g), uni(p), uni(count), uni(val));
}
template <int T, int K>
__device__ __forceinline__ void nl_emit_light(Grid g, BPState s, int32_t word, int32_t rows, const int32_t* src,
                                           const int32_t* in_off, const int32_t* in_col, const int32_t* in_col_orig,
                                           int32_t n, int32_t heavy_cut, int32_t* lvl_out, int32_t* par_out,
                                           unsigned* dyn, int32_t ngrp) {
  emit_light_phase<T, K, 1, true, false>(uni(g), g_lds.e, uni(s), uni(word), uni(rows), uni(src), uni(in_off),
                                         uni(in_col),
                                  uni(in_col_orig), uni(n), uni(heavy_cut), uni(lvl_out), uni(par_out), RowMap{},
                                  uni(dyn), uni(ngrp));
}

template <int T, int K>
__device__ bool emit_all_words(Grid g, BPState s, const GraphView& gv, const TileJob& job, const TeamSync& ts,
                               unsigned* gerr) {
  const int words = (job.cnt + 63) / 64;
  // The light emit of every word group shares one counter, reset before the
  // barrier that precedes it (the reset is visible to the team after it).
  if (g.bid == 0 && threadIdx.x == 0) atomicExch(&ts.ctl->emit_chunk, 0u);
  if (!team_barrier(ts, g, gerr, XB_K_EMIT)) return false;
  const int32_t ngrp = words;  // one word per pass
  nl_emit_light<T, K>(g, s, 0, job.cnt, job.src, gv.oin_off, gv.oin_col, gv.oin_colo, gv.n, gv.cut_medium,
                      job.lvl_out, job.par_out, &ts.ctl->emit_chunk, ngrp);
  return true;
}

// Runs one tile of up to 64*T sources of one graph on a team.  Returns false
// if the batch was aborted.
template <int T>
__device__ bool run_tile(Grid g, const Team& tm, const TeamSync& ts, const GraphView& gv, TileJob& job, int32_t alpha,
                         unsigned* gerr) {
  TeamCtl* ctl = tm.ctl;
  BPState s = tm.s;
  s.hidx = gv.hidx;
  s.new_of = gv.new_of;
  s.orig_of = gv.orig_of;
  s.in_col_o = gv.in_col_o;
  const int32_t n = gv.n;
  const int64_t nw = (int64_t)n * T;
  // Tile set-up: seen, planes and fronts cleared, heavy parents unresolved;
  // next words and touched flags are clean between tiles by construction.
  nl_fill_u64(g, s.seen, nw, 0ull);
  // The team's plane block holds both arrays, planes 4-7 after this tile's
  // 0-3.  Planes 0-3 are cleared here, 4-7 only before the first level whose
  // code needs them (level 15), which most tiles never reach.
  s.planes_hi = s.planes + nw * 4;
  nl_fill_u64(g, s.planes, nw * 4, 0ull);
  if (XB_PERSIST_LAZY_FRONT_CLEAR) {
    // Levels leave their frontier words set (masked by seen in pull and
    // push); both frontier arrays are cleared here instead, once per tile.
    nl_fill_u64(g, s.front[0], nw, 0ull);
    nl_fill_u64(g, s.front[1], nw, 0ull);
  }
  if (gv.n_heavy > 0) nl_fill_u32(g, s.hpar, (int64_t)gv.n_heavy * T * 64, 0xFFFFFFFFu);
  if (g.bid == 0 && threadIdx.x == 0) {
    ctl->ctr[1] = 0ull;
    ctl->tcount[0] = 0u;
  }
  if (!team_barrier(ts, g, gerr, XB_K_MEMSET)) return false;
  nl_seed<T>(g, s, job.src, job.cnt);
  if (!team_barrier(ts, g, gerr, XB_K_SEED)) return false;

  int32_t qn = job.qn0, E = job.E0, d = 0;
  const int64_t thresh = gv.m / (alpha > 0 ? alpha : 1);
  while (qn > 0) {
    const uint32_t code = (uint32_t)d + 2u;  // level d+1
    const int cur = d & 1;
    const Queue qc = (d == 0) ? Queue{const_cast<int32_t*>(job.q0v), const_cast<int32_t*>(job.q0eo), nullptr}
                              : Queue{tm.qv[d & 1], tm.qeo[d & 1], &ctl->ctr[d % 3]};
    if (code > 255u) {
      // Deeper than the 8-bit level code: clear this frontier and hand the
      // tile to the plain path on the host.
      nl_clear<T>(g, s, cur, qc.v, qn);
      if (g.bid == 0 && threadIdx.x == 0) job.status = 1;
      return team_barrier(ts, g, gerr, XB_K_CLEAR);
    }
    if (code == 16u) {
      nl_fill_u64(g, s.planes_hi, nw * 4, 0ull);
      if (!team_barrier(ts, g, gerr, XB_K_MEMSET)) return false;
    }
    const Queue qx{tm.qv[(d + 1) & 1], tm.qeo[(d + 1) & 1], &ctl->ctr[(d + 1) % 3]};
    s.tcount = &ctl->tcount[d & 1];
    if (g.bid == 0 && threadIdx.x == 0) {
      ctl->ctr[(d + 2) % 3] = 0ull;   // the counter of level d+2's queue
      ctl->tcount[(d + 1) & 1] = 0u;  // next level's touched count
    }
    const bool pull = (int64_t)qn + E > thresh;
    if (pull) {
      int32_t begin = 0;
      if (gv.bucket_end[0] > begin)
        nl_pull_bucket<T, 1, false>(g, s, cur, code, job.cnt, gv.in_off, gv.in_col, gv.off, gv.bucket_perm[0],
                                       gv.bucket_shift[0], begin, gv.bucket_end[0], qx);
      begin = gv.bucket_end[0];
      if (gv.bucket_end[1] > begin)
        nl_pull_bucket<T, (T * 4 <= kWave ? 4 : kWave / T), false>(g, s, cur, code, job.cnt, gv.in_off,
                                                                      gv.in_col, gv.off, gv.bucket_perm[1],
                                                                      gv.bucket_shift[1], begin,
                                                                      gv.bucket_end[1], qx);
      begin = gv.bucket_end[1];
      if (gv.bucket_end[2] > begin)
        nl_pull_bucket<T, kWave / T, true>(g, s, cur, code, job.cnt, gv.in_off, gv.in_col, gv.off,
                                              gv.bucket_perm[2], gv.bucket_shift[2], begin, gv.bucket_end[2], qx);
      if (gv.n_hub_chunks > 0) nl_pull_hub<T>(g, s, cur, job.cnt, gv.in_col, gv.hubs, gv.n_hub_chunks);
    } else {
      nl_push<T>(g, s, cur, gv.off, gv.col, qc, qn, E);
    }
    if (!team_barrier(ts, g, gerr, pull ? XB_K_PULL : XB_K_PUSH)) return false;
    nl_commit<T>(g, s, cur, code, gv.off, qx);
    if (!XB_PERSIST_LAZY_FRONT_CLEAR) nl_clear<T>(g, s, cur, qc.v, qn);
    if (!team_barrier(ts, g, gerr, XB_K_COMMIT)) return false;
    const unsigned long long c =
        __hip_atomic_load(&ctl->ctr[(d + 1) % 3], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    qn = (int32_t)(uint32_t)(c & 0xFFFFFFFFull);
    E = (int32_t)(uint32_t)(c >> 32);
    if (qn > 0) ++d;
  }
  int K = 0;
  while ((1 << K) <= d + 1) ++K;
  bool ok = true;
  // A team of one tile leaves the emit to the standalone launches after the
  // kernel (two source words per pass in 512-thread workgroups, which this
  // kernel's 256-thread workgroups and 20 KB LDS union cannot hold).
  if (!tm.emit_after) switch (K) {
    case 1: ok = emit_all_words<T, 1>(g, s, gv, job, ts, gerr); break;
    case 2: ok = emit_all_words<T, 2>(g, s, gv, job, ts, gerr); break;
    case 3: ok = emit_all_words<T, 3>(g, s, gv, job, ts, gerr); break;
    case 4: ok = emit_all_words<T, 4>(g, s, gv, job, ts, gerr); break;
    case 5: ok = emit_all_words<T, 5>(g, s, gv, job, ts, gerr); break;
    case 6: ok = emit_all_words<T, 6>(g, s, gv, job, ts, gerr); break;
    case 7: ok = emit_all_words<T, 7>(g, s, gv, job, ts, gerr); break;
    default: ok = emit_all_words<T, 8>(g, s, gv, job, ts, gerr); break;
  }
  if (!ok) return false;
  if (g.bid == 0 && threadIdx.x == 0) {
    job.status = 0;
    job.levels = d;
  }
  return team_barrier(ts, g, gerr, XB_K_EMIT);
}

// Workgroups of the persistent kernel per CU the register budget is sized for
// (8 x 256 threads = 32 waves per CU, 64 VGPRs per lane; the LDS union is
// sized to allow it).  Measured on the uneven 32-graph batch: 90 ms at 8 per
// CU against 117 ms at 4 per CU (128 VGPRs), whose extra registers do not pay
// for the halved latency hiding of the gathers.
#ifndef XB_PERSISTENT_MIN_BLOCKS
#define XB_PERSISTENT_MIN_BLOCKS 8
#endif

__global__ __launch_bounds__(kBlock, XB_PERSISTENT_MIN_BLOCKS) void k_bp_persistent(
    const Team* __restrict__ teams, int32_t n_teams, const GraphView* __restrict__ graphs,
    TileJob* __restrict__ jobs, int32_t alpha, unsigned* __restrict__ gerr) {
  int ti = 0;
  while (ti + 1 < n_teams && (int)blockIdx.x >= teams[ti + 1].first_block) ++ti;
  const Team tm = teams[ti];
  const Grid g{(int)blockIdx.x - tm.first_block, tm.nblocks};
  if (g.bid < 0 || g.bid >= tm.nblocks) return;  // block outside every team (the grid is the sum of the teams)
  TeamSync ts;
  if (!team_join(tm.ctl, g, gerr, &ts)) return;
  if (g.bid == 0 && threadIdx.x == 0) tm.ctl->tlast = (unsigned long long)wall_clock64();
  for (int32_t j = tm.job_begin; j < tm.job_end; ++j) {
    TileJob& job = jobs[j];
    const GraphView& gv = graphs[job.graph];
    bool ok;
    switch (job.words) {  // uniform over the team: every block reads the same job
      case 1: ok = run_tile<1>(g, tm, ts, gv, job, alpha, gerr); break;
      case 2: ok = run_tile<2>(g, tm, ts, gv, job, alpha, gerr); break;
      case 4: ok = run_tile<4>(g, tm, ts, gv, job, alpha, gerr); break;
      case 8: ok = run_tile<8>(g, tm, ts, gv, job, alpha, gerr); break;
      default: ok = run_tile<16>(g, tm, ts, gv, job, alpha, gerr); break;
    }
    if (!ok) return;
  }
}

// Zeroes the team control blocks (as 64-bit words), the error flag and the
// job statuses inside the captured graph (a kernel rather than memset nodes).
__global__ void k_reset_ctl(TeamCtl* ctl, int32_t n_teams, TileJob* jobs, int32_t n_jobs, unsigned* gerr) {
  const int64_t words = (int64_t)n_teams * (int64_t)(sizeof(TeamCtl) / 8);
  unsigned long long* w = reinterpret_cast<unsigned long long*>(ctl);
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < words; i += (int64_t)gridDim.x * blockDim.x)
    w[i] = 0ull;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n_jobs) {
    jobs[i].status = -1;
    jobs[i].levels = -1;
  }
  if (i == 0) *gerr = 0u;
}

// ============================================================================
// Forest mode: the alternating-forest search of maximum matching
// ============================================================================
//
// One phase grows an alternating forest from every exposed vertex (the even
// roots) and looks for edges between even vertices of different trees (an
// augmenting path each).  Reference semantics: InitAlternatingForest,
// FindAndFlipAugmentingPathInAlternatingForest and ExpandAlternatingForest of
// apps/xb/src/xblsm_gpu/xblossom_gpu_engine_10.cu:83-359, with the path
// flip of apps/xb/src/include/path_table/path_table.h:273-330.  Every choice
// the reference leaves to the first atomic is fixed here:
//  * an unvisited matched vertex is claimed by the even neighbour with the
//    lowest (root, vertex) key, and of a matched pair claimed from both
//    sides the member with the lower (claim, id) becomes odd;
//  * augmenting paths are chosen in rounds: every tree records its lowest
//    neighbouring tree, mutually lowest pairs are taken, their bridge is the
//    lowest (vertex in the lower tree, vertex in the higher tree) edge, and
//    taken trees drop out of later rounds -- so each phase flips a maximal
//    set of tree-disjoint paths, identically on every run.
// Edges closing an odd cycle inside one tree form blossoms, whose odd vertices
// become outer (the blossom step below); between phases the trees whose root
// is still exposed are kept (k_mt_reuse_init).

constexpr unsigned long long kNoKey = ~0ull;
constexpr int32_t kDummy = -1;  // the non-outer vertex above a root on every path (fp of a root)

// Resets the forest and queues every exposed vertex as an even root.
__global__ __launch_bounds__(kBlock) void k_mt_init(int32_t n, const int32_t* __restrict__ mate,
                                                    const int32_t* __restrict__ off, int32_t* __restrict__ root,
                                                    int32_t* __restrict__ lvl, int32_t* __restrict__ par,
                                                    unsigned long long* __restrict__ claim,
                                                    int32_t* __restrict__ best, int32_t* __restrict__ acc,
                                                    unsigned long long* __restrict__ bridge,
                                                    uint8_t* __restrict__ used, int32_t* __restrict__ cursor,
                                                    int32_t* __restrict__ fp, unsigned long long* __restrict__ elab,
                                                    Queue q) {
  AppendStage& sg = g_lds.q.sg;
  stage_init(sg);
  for (int64_t b0 = (int64_t)blockIdx.x * kBlock; b0 < n; b0 += (int64_t)gridDim.x * kBlock) {
    const int64_t v = b0 + threadIdx.x;
    bool ex = false;
    int32_t dv = 0;
    if (v < n) {
      ex = mate[v] < 0;
      cursor[v] = off[v];
      fp[v] = kDummy;
      elab[v] = kNoKey;
      root[v] = ex ? (int32_t)v : -1;
      lvl[v] = ex ? 0 : -1;
      par[v] = -1;
      claim[v] = kNoKey;
      best[v] = INT32_MAX;
      acc[v] = -1;
      bridge[v] = kNoKey;
      used[v] = 0;
      if (ex) dv = off[v + 1] - off[v];
    }
    stage_put(sg, ex, (int32_t)v, dv);
    stage_sync_flush(sg, false, q);
  }
  stage_sync_flush(sg, true, q);
}

// Phase set-up with tree reuse, first pass (the reference's lazy reset,
// apps/xb/src/xblsm_gpu/xblossom_gpu_engine_10.cu:119-175): a vertex of a
// tree whose root augmented last phase (used[root]) is reset; every other
// label is kept.  Outer vertices of the kept trees are queued as the phase's
// first frontier (the reference's ENode2 list); the phase's roots are the
// previous phase's untaken roots, whose list the rounds already keep.
// counts[0] += matched vertices kept in a tree (reused), counts[1] += the
// others (exposed roots, reset vertices, vertices in no tree).  `used` is
// only read here; the second pass clears it.
__global__ __launch_bounds__(kBlock) void k_mt_reuse_init(int32_t n, const int32_t* __restrict__ mate,
                                                          const int32_t* __restrict__ off,
                                                          int32_t* __restrict__ root, int32_t* __restrict__ lvl,
                                                          int32_t* __restrict__ par, int32_t* __restrict__ fp,
                                                          unsigned long long* __restrict__ elab,
                                                          unsigned long long* __restrict__ claim,
                                                          const uint8_t* __restrict__ used, Queue q,
                                                          unsigned long long* __restrict__ counts) {
  AppendStage& sg = g_lds.q.sg;
  stage_init(sg);
  unsigned long long reused = 0, reset = 0;
  for (int64_t b0 = (int64_t)blockIdx.x * kBlock; b0 < n; b0 += (int64_t)gridDim.x * kBlock) {
    const int64_t v = b0 + threadIdx.x;
    bool outer = false;
    int32_t dv = 0;
    if (v < n) {
      const int32_t r = root[v];
      const bool matched = mate[v] >= 0;
      if (r >= 0 && used[r]) {
        root[v] = -1;
        lvl[v] = -1;
        par[v] = -1;
        fp[v] = kDummy;
        elab[v] = kNoKey;
        claim[v] = kNoKey;
        ++reset;
      } else {
        if (r >= 0 && matched)
          ++reused;
        else
          ++reset;
        const int32_t l = lvl[v];
        outer = l >= 0 && !(l & 1);
        if (outer) dv = off[v + 1] - off[v];
      }
    }
    stage_put(sg, outer, (int32_t)v, dv);
    stage_sync_flush(sg, false, q);
  }
  stage_sync_flush(sg, true, q);
  block_add(reused, &counts[0]);
  block_add(reset, &counts[1]);
}

// Second pass: the per-phase pairing state of every vertex.
__global__ __launch_bounds__(kBlock) void k_mt_reuse_clear(int32_t n, int32_t* __restrict__ best,
                                                           int32_t* __restrict__ acc,
                                                           unsigned long long* __restrict__ bridge,
                                                           uint8_t* __restrict__ used) {
  for (int64_t v = (int64_t)blockIdx.x * kBlock + threadIdx.x; v < n; v += (int64_t)gridDim.x * kBlock) {
    best[v] = INT32_MAX;
    acc[v] = -1;
    bridge[v] = kNoKey;
    used[v] = 0;
  }
}

// Calls f(v, w, j) for every edge (v, w) = col[j] out of the frontier queue
// (edge-balanced chunks of kChunk edges per block, as in the push).
template <class F>
__device__ __forceinline__ void for_frontier_edges(const int32_t* __restrict__ off, const int32_t* __restrict__ col,
                                                   Queue q, int32_t qn, int32_t E, F&& f) {
  QueueLds& L = g_lds.q;
  for (int64_t c0 = (int64_t)blockIdx.x * kChunk; c0 < E; c0 += (int64_t)gridDim.x * kChunk) {
    const int c1 = (int)((c0 + kChunk < E) ? c0 + kChunk : E);
    const ChunkOwners co = stage_owners(q.eo, qn, (int)c0, c1, L);
#pragma unroll 1
    for (int r = 0; r < kPushR; ++r) {
      const int e = (int)c0 + r * kBlock + (int)threadIdx.x;
      if (e < c1) {
        const int i = chunk_owner(co, q.eo, L.s_eo, e);
        const int32_t v = q.v[i];
        const int32_t j = off[v] + (e - q.eo[i]);
        f(v, col[j], j);
      }
    }
    __syncthreads();  // the chunk's owner staging is reused by the next chunk
  }
}

// Augmenting-path scan: every tree records its lowest neighbouring tree over
// even-even edges (both trees not taken yet); same-tree even-even edges are
// counted as blossom edges.
__global__ __launch_bounds__(kBlock) void k_mt_scan(const int32_t* __restrict__ off, const int32_t* __restrict__ col,
                                                    Queue q, int32_t qn, int32_t E,
                                                    const int32_t* __restrict__ root,
                                                    const int32_t* __restrict__ lvl, const uint8_t* __restrict__ used,
                                                    int32_t* __restrict__ best,
                                                    unsigned long long* __restrict__ blossoms) {
  unsigned long long same = 0;
  for_frontier_edges(off, col, q, qn, E, [&](int32_t v, int32_t w, int32_t) {
    const int32_t lw = lvl[w];
    if (lw < 0 || (lw & 1) || w == v) return;
    const int32_t rv = root[v], rw = root[w];
    if (rv == rw) {
      ++same;
      return;
    }
    if (used[rv] || used[rw]) return;
    atomicMin(&best[rv], rw);
    atomicMin(&best[rw], rv);
  });
  block_add(same, blossoms);
}

// Level-0 scan.  At level 0 every even vertex is a root, so a root's lowest
// neighbouring tree is its first adjacency entry (lists ascend) that is an
// untaken root other than itself.  A per-root cursor keeps the position of
// that entry across rounds: taken roots never return, so entries passed once
// are never read again.  Each lane walks a short remainder itself; a long one
// is walked by the whole wave, 64 entries per step.  `examined` counts the
// entries read up to and including the one found.
constexpr int32_t kShortWalk = 16;

__global__ __launch_bounds__(kBlock) void k_mt_scan0(const int32_t* __restrict__ roots, int32_t nroots,
                                                     const int32_t* __restrict__ off, const int32_t* __restrict__ col,
                                                     const int32_t* __restrict__ lvl, const uint8_t* __restrict__ used,
                                                     int32_t* __restrict__ cursor, int32_t* __restrict__ best,
                                                     unsigned long long* __restrict__ examined) {
  unsigned long long cnt = 0;
  for (int64_t b0 = (int64_t)blockIdx.x * kBlock; b0 < nroots; b0 += (int64_t)gridDim.x * kBlock) {
    const int64_t i = b0 + threadIdx.x;
    int32_t v = -1, j = 0, e = 0;
    if (i < nroots) {
      v = roots[i];
      j = cursor[v];
      e = off[v + 1];
    }
    const bool lng = v >= 0 && e - j > kShortWalk;
    if (v >= 0 && !lng) {
      const int32_t j0 = j;
      int32_t hit = INT32_MAX;  // the entry found, kept rather than read again
      while (j < e) {
        // Both flags are loaded at once (w < n always): one dependent round
        // trip after col[j] instead of two.
        const int32_t w = col[j];
        const int32_t lw = lvl[w];
        const uint8_t uw = used[w];
        if (w != v && lw == 0 && !uw) {
          hit = w;
          break;
        }
        ++j;
      }
      cnt += (unsigned long long)(j - j0) + (j < e ? 1ull : 0ull);
      cursor[v] = j;
      best[v] = hit;
    }
    uint64_t m = __ballot(lng);
    while (m) {
      const int src = __ffsll((unsigned long long)m) - 1;
      m &= m - 1;
      const int32_t vv = __shfl(v, src, kWave), j0 = __shfl(j, src, kWave), ee = __shfl(e, src, kWave);
      int32_t found = ee, fw = INT32_MAX;
      for (int32_t k = j0; k < ee; k += kWave) {
        const int32_t idx = k + lane_id();
        bool ok = false;
        int32_t w = 0;
        if (idx < ee) {
          w = col[idx];
          const int32_t lw = lvl[w];
          const uint8_t uw = used[w];
          ok = w != vv && lw == 0 && !uw;
        }
        const uint64_t hit = __ballot(ok);
        if (hit) {
          const int first = __ffsll((unsigned long long)hit) - 1;
          found = k + first;
          fw = __shfl(w, first, kWave);
          break;
        }
      }
      if (lane_id() == src) {
        cnt += (unsigned long long)(found - j0) + (found < ee ? 1ull : 0ull);
        cursor[vv] = found;
        best[vv] = fw;
      }
    }
  }
  block_add(cnt, examined);
}

// Mutually lowest tree pairs are taken (recorded at the lower root).  At
// level 0 (bridge0 non-null) both trees are single roots and the pair's
// bridge is the edge between them.
__global__ __launch_bounds__(kBlock) void k_mt_accept(const int32_t* __restrict__ roots, int32_t nroots,
                                                      const int32_t* __restrict__ best, int32_t* __restrict__ acc,
                                                      int32_t* __restrict__ pairs, uint32_t* __restrict__ npairs,
                                                      unsigned long long* __restrict__ bridge0) {
  for (int64_t b0 = (int64_t)blockIdx.x * kBlock; b0 < nroots; b0 += (int64_t)gridDim.x * kBlock) {
    const int64_t i = b0 + threadIdx.x;
    bool take = false;
    int32_t t = 0;
    if (i < nroots) {
      t = roots[i];
      const int32_t b = best[t];
      if (b != INT32_MAX && t < b && best[b] == t) {
        take = true;
        acc[t] = b;
        if (bridge0) bridge0[t] = ((unsigned long long)(uint32_t)t << 32) | (uint32_t)b;
      }
    }
    block_push_list(take, t, pairs, npairs);
  }
}

// Drops the entries whose tree was taken in this round: with a queue
// (list == nullptr) the frontier and its degree prefix, else a plain list of
// roots counted in ctr32.
__global__ __launch_bounds__(kBlock) void k_mt_compact(const int32_t* __restrict__ in, int32_t cnt,
                                                       const int32_t* __restrict__ root,
                                                       const uint8_t* __restrict__ used,
                                                       const int32_t* __restrict__ off, Queue out,
                                                       int32_t* __restrict__ list, uint32_t* __restrict__ ctr32) {
  AppendStage& sg = g_lds.q.sg;
  stage_init(sg);
  for (int64_t b0 = (int64_t)blockIdx.x * kBlock; b0 < cnt; b0 += (int64_t)gridDim.x * kBlock) {
    const int64_t i = b0 + threadIdx.x;
    bool keep = false;
    int32_t v = 0, dv = 0;
    if (i < cnt) {
      v = in[i];
      keep = !used[root[v]];
      if (keep && !list) dv = off[v + 1] - off[v];
    }
    stage_put(sg, keep, v, dv);
    stage_sync_flush(sg, false, out, list, ctr32);
  }
  stage_sync_flush(sg, true, out, list, ctr32);
}

// Clears the trees' lowest-neighbour records before the next round.
// Zeroes a round's counters in one launch (pairs, flip-stack pool, kept
// roots, the next frontier queue's packed counter).
__global__ void k_mt_zero_ctrs(uint32_t* a, uint32_t* b, uint32_t* c, unsigned long long* d) {
  if (threadIdx.x == 0) {
    *a = 0u;
    *b = 0u;
    *c = 0u;
    *d = 0ull;
  }
}

__global__ __launch_bounds__(kBlock) void k_mt_reset_best(const int32_t* __restrict__ roots, int32_t nroots,
                                                          int32_t* __restrict__ best) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < nroots; i += (int64_t)gridDim.x * kBlock)
    best[roots[i]] = INT32_MAX;
}

// Bridge of every taken pair: the lowest (vertex in the lower tree, vertex in
// the higher tree) even-even edge between the two trees.
__global__ __launch_bounds__(kBlock) void k_mt_bridge(const int32_t* __restrict__ off, const int32_t* __restrict__ col,
                                                      Queue q, int32_t qn, int32_t E,
                                                      const int32_t* __restrict__ root,
                                                      const int32_t* __restrict__ lvl, const int32_t* __restrict__ acc,
                                                      unsigned long long* __restrict__ bridge) {
  for_frontier_edges(off, col, q, qn, E, [&](int32_t v, int32_t w, int32_t) {
    const int32_t lw = lvl[w];
    if (lw < 0 || (lw & 1)) return;
    const int32_t rv = root[v], rw = root[w];
    if (rv == rw) return;
    const int32_t lo = rv < rw ? rv : rw, hi = rv < rw ? rw : rv;
    if (acc[lo] != hi) return;
    const uint32_t a = (uint32_t)(rv == lo ? v : w), c = (uint32_t)(rv == lo ? w : v);
    atomicMin(&bridge[lo], ((unsigned long long)a << 32) | c);
  });
}

// Flips each taken path (one thread per path; paths lie in two trees of
// their own).  Every outer vertex v has a path P(v) to its root that starts
// with v's matched edge:
//   root:                     P(v) = v
//   grown (mate t claimed by par[t]):  P(v) = v, t, P(par[t])
//   made outer by blossom edge (a, b) with v on P(a):
//                             P(v) = reverse(P(a) from a to v), P(b)
// Rematch(v, w) matches v to w and flips P(v) behind it; augmenting the
// bridge (a, c) is Rematch(a, c) then Rematch(c, a).  The recursion
// (H. N. Gabow, "An efficient implementation of Edmonds' algorithm for
// maximum matching on graphs", JACM 23(2), 1976, procedure R) stops where
// the walk meets a vertex whose mate was already changed; the second call
// of a blossom label waits on a stack linked through a shared pool.
__global__ __launch_bounds__(kBlock) void k_mt_flip(const int32_t* __restrict__ pairs, const uint32_t* __restrict__ npairs,
                                                    const int32_t* __restrict__ acc,
                                                    const unsigned long long* __restrict__ bridge,
                                                    const int32_t* __restrict__ par,
                                                    const unsigned long long* __restrict__ elab,
                                                    int32_t* __restrict__ mate, uint8_t* __restrict__ used,
                                                    FlipStack stk, int32_t max_steps, unsigned* __restrict__ err) {
  const uint32_t np = *npairs;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < np; i += (int64_t)gridDim.x * kBlock) {
    const int32_t t0 = pairs[i];
    const unsigned long long key = bridge[t0];
    if (key == kNoKey) {  // a taken pair always has a bridge; never walk from an empty one
      atomicOr(err, 1u);
      continue;
    }
    const int32_t a = (int32_t)(key >> 32), c = (int32_t)(key & 0xFFFFFFFFull);
    int32_t top = -1;
    bool ok = true;
    // Pending call Rematch(c, a) under the running Rematch(a, c).
    {
      const uint32_t slot = atomicAdd(stk.ctr, 1u);
      if (slot >= stk.cap) {
        ok = false;
      } else {
        stk.e[slot] = make_int4(c, a, top, 0);
        top = (int32_t)slot;
      }
    }
    int32_t v = a, w = c;
    for (int32_t step = 0; ok; ++step) {
      if (step > max_steps) {
        ok = false;
        break;
      }
      const int32_t t = mate[v];
      mate[v] = w;
      bool next = false;
      if (t >= 0 && mate[t] == v) {
        const unsigned long long el = elab[v];
        if (el == kNoKey) {
          const int32_t cl = par[t];  // the outer vertex that claimed t
          if (cl < 0) {
            ok = false;
            break;
          }
          mate[t] = cl;
          v = cl;
          w = t;
          next = true;
        } else {
          const int32_t la = (int32_t)(el >> 32), lb = (int32_t)(el & 0xFFFFFFFFull);
          const uint32_t slot = atomicAdd(stk.ctr, 1u);
          if (slot >= stk.cap) {
            ok = false;
            break;
          }
          stk.e[slot] = make_int4(lb, la, top, 0);
          top = (int32_t)slot;
          v = la;
          w = lb;
          next = true;
        }
      }
      if (!next) {
        if (top < 0) break;
        const int4 e = stk.e[top];
        v = e.x;
        w = e.y;
        top = e.z;
      }
    }
    if (!ok) atomicOr(err, 2u);
    used[t0] = 1;
    used[acc[t0]] = 1;
  }
}

// Blossom candidates: same-tree even-even edges out of the frontier, keyed
// (root << 32 | edge position) so a sort groups them by tree in adjacency
// order.
__global__ __launch_bounds__(kBlock) void k_mt_collect(const int32_t* __restrict__ off, const int32_t* __restrict__ col,
                                                       Queue q, int32_t qn, int32_t E,
                                                       const int32_t* __restrict__ root,
                                                       const int32_t* __restrict__ lvl,
                                                       unsigned long long* __restrict__ cand,
                                                       uint32_t* __restrict__ ncand, uint32_t cap,
                                                       unsigned* __restrict__ err) {
  for_frontier_edges(off, col, q, qn, E, [&](int32_t v, int32_t w, int32_t j) {
    const int32_t lw = lvl[w];
    if (lw < 0 || (lw & 1) || w == v) return;
    const int32_t rv = root[v];
    if (rv != root[w]) return;
    const uint32_t k = atomicAdd(ncand, 1u);
    if (k < cap)
      cand[k] = ((unsigned long long)(uint32_t)rv << 32) | (uint32_t)j;
    else
      atomicOr(err, 4u);
  });
}

// Forest state the blossom step reads and writes.
struct BlossomState {
  const int32_t* off;
  const int32_t* col;
  int32_t n;
  const int32_t* par;
  int32_t* lvl;
  int32_t* fp;  // outer vertex -> a vertex above it on its path; following fp over outer vertices ends at FIRST
  unsigned long long* elab;
  unsigned long long* lab;  // inner vertex -> lowest (candidate << 1 | side) whose path holds it, ~0 if none
};

__device__ __forceinline__ bool mt_outer(const int32_t* lvl, int32_t u) {
  const int32_t l = lvl[u];
  return l >= 0 && !(l & 1);
}

// FIRST(v): the first non-outer vertex on P(v) (kDummy above the root).
__device__ __forceinline__ int32_t mt_first(const BlossomState& b, int32_t v, int32_t& budget) {
  int32_t u = b.fp[v];
  while (u != kDummy && mt_outer(b.lvl, u) && --budget > 0) u = b.fp[u];
  return u;
}

// The next non-outer vertex above inner vertex u on every path through it:
// FIRST of its claimer.  Levels fall strictly along this chain (the claimer's
// path existed before u was claimed), and kDummy sits at level -1.
__device__ __forceinline__ int32_t mt_up(const BlossomState& b, int32_t u, int32_t& budget) {
  return mt_first(b, b.par[u], budget);
}

__device__ __forceinline__ int32_t mt_depth(const BlossomState& b, int32_t u) { return u == kDummy ? -1 : b.lvl[u]; }

// Blossom step, marking pass.  Processing a tree's candidate edges one by
// one in key order (Gabow's procedure L) makes an inner vertex outer with the
// label of the FIRST candidate whose path, in the forest as it stood before
// the step, runs through it: earlier candidates only merge whole subtrees, so
// a later candidate's path keeps every inner vertex of its original path and
// gains none.  So each candidate (x, y) -- in sorted order i -- finds the
// join of FIRST(x) and FIRST(y) by levels, and writes min(i << 1 | side)
// into every inner vertex strictly below it; a vertex's first mark lists it.
__global__ __launch_bounds__(kBlock) void k_mt_blossom_mark(BlossomState b, const unsigned long long* __restrict__ cand,
                                                            int32_t C, int32_t* __restrict__ join_of,
                                                            int32_t* __restrict__ marked,
                                                            uint32_t* __restrict__ nmarked,
                                                            unsigned* __restrict__ err) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < C; i += (int64_t)gridDim.x * kBlock) {
    const int32_t j = (int32_t)(cand[i] & 0xFFFFFFFFull);
    // Owner of adjacency entry j: the last vertex whose offset is <= j.
    int32_t lo = 0, hi = b.n;
    while (lo < hi) {
      const int32_t mid = (lo + hi + 1) >> 1;
      if (b.off[mid] <= j)
        lo = mid;
      else
        hi = mid - 1;
    }
    const int32_t x = lo, y = b.col[j];
    int32_t budget = 4 * b.n + 16;
    const int32_t r = mt_first(b, x, budget), s = mt_first(b, y, budget);
    join_of[i] = r;
    if (r == s) continue;  // x and y already share a blossom
    int32_t a = r, c = s;
    while (a != c && budget > 0) {
      if (mt_depth(b, a) >= mt_depth(b, c)) {
        if (a == kDummy || b.par[a] < 0) break;
        a = mt_up(b, a, budget);
      } else {
        if (c == kDummy || b.par[c] < 0) break;
        c = mt_up(b, c, budget);
      }
    }
    if (a != c || budget <= 0) {
      atomicOr(err, 8u);
      continue;
    }
    const int32_t join = a;
    join_of[i] = join;
    for (int side = 0; side < 2; ++side) {
      int32_t u = side ? s : r;
      const unsigned long long code = ((unsigned long long)i << 1) | (unsigned long long)side;
      while (u != join) {
        if (u == kDummy || b.par[u] < 0 || budget <= 0) {
          atomicOr(err, 8u);
          break;
        }
        if (atomicMin(&b.lab[u], code) == kNoKey) marked[atomicAdd(nmarked, 1u)] = u;
        u = mt_up(b, u, budget);
      }
    }
  }
}

// Blossom step, applying pass: every marked vertex becomes outer with its
// candidate's edge label (oriented to its side), FIRST pointer = that
// candidate's join, its odd level plus one, and a place in the next frontier.  These
// are the reference's odd nodes in a blossom made even
// (TransformOddNodesInBlossom, xblossom_gpu_engine_10.cu:362-566).
__global__ __launch_bounds__(kBlock) void k_mt_blossom_apply(BlossomState b, const unsigned long long* __restrict__ cand,
                                                             const int32_t* __restrict__ join_of,
                                                             const int32_t* __restrict__ marked,
                                                             const uint32_t* __restrict__ nmarked,
                                                             uint32_t* __restrict__ formed, Queue nxt,
                                                             unsigned long long* __restrict__ counts) {
  const uint32_t cnt = *nmarked;
  unsigned long long made = 0, blossoms = 0;
  for (int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x; k < cnt; k += (int64_t)gridDim.x * kBlock) {
    const int32_t u = marked[k];
    const unsigned long long code = b.lab[u];
    b.lab[u] = kNoKey;
    const int64_t i = (int64_t)(code >> 1);
    const int side = (int)(code & 1ull);
    const int32_t j = (int32_t)(cand[i] & 0xFFFFFFFFull);
    int32_t lo = 0, hi = b.n;
    while (lo < hi) {
      const int32_t mid = (lo + hi + 1) >> 1;
      if (b.off[mid] <= j)
        lo = mid;
      else
        hi = mid - 1;
    }
    const int32_t x = lo, y = b.col[j];
    const int32_t p = side ? y : x, q = side ? x : y;
    b.elab[u] = ((unsigned long long)(uint32_t)p << 32) | (uint32_t)q;
    b.fp[u] = join_of[i];
    b.lvl[u] += 1;  // inner (odd) -> outer (even), still above its FIRST chain
    const int32_t du = b.off[u + 1] - b.off[u];
    const unsigned long long old = atomicAdd(nxt.ctr, ((unsigned long long)(uint32_t)du << 32) | 1ull);
    nxt.v[(uint32_t)(old & 0xFFFFFFFFull)] = u;
    nxt.eo[(uint32_t)(old & 0xFFFFFFFFull)] = (int32_t)(uint32_t)(old >> 32);
    ++made;
    if (atomicExch(&formed[i], 1u) == 0u) ++blossoms;
  }
  made = wave_sum64(made);
  blossoms = wave_sum64(blossoms);
  if (lane_id() == 0) {
    if (blossoms) atomicAdd(&counts[0], blossoms);
    if (made) atomicAdd(&counts[1], made);
  }
}

// Expansion, claim step: unvisited (hence matched) neighbours of the even
// frontier take the lowest (root, claimer) key; first claims are listed.
__global__ __launch_bounds__(kBlock) void k_mt_claim(const int32_t* __restrict__ off, const int32_t* __restrict__ col,
                                                     Queue q, int32_t qn, int32_t E,
                                                     const int32_t* __restrict__ root,
                                                     const int32_t* __restrict__ lvl,
                                                     unsigned long long* __restrict__ claim,
                                                     int32_t* __restrict__ tlist, uint32_t* __restrict__ tcount) {
  QueueLds& L = g_lds.q;
  AppendStage& sg = L.sg;
  stage_init(sg);
  for (int64_t c0 = (int64_t)blockIdx.x * kChunk; c0 < E; c0 += (int64_t)gridDim.x * kChunk) {
    const int c1 = (int)((c0 + kChunk < E) ? c0 + kChunk : E);
    const ChunkOwners co = stage_owners(q.eo, qn, (int)c0, c1, L);
#pragma unroll 1
    for (int r = 0; r < kPushR; ++r) {
      const int e = (int)c0 + r * kBlock + (int)threadIdx.x;
      bool first = false;
      int32_t w = 0;
      if (e < c1) {
        const int i = chunk_owner(co, q.eo, L.s_eo, e);
        const int32_t v = q.v[i];
        w = col[off[v] + (e - q.eo[i])];
        if (lvl[w] < 0) {
          const unsigned long long key = ((unsigned long long)(uint32_t)root[v] << 32) | (uint32_t)v;
          first = atomicMin(&claim[w], key) == kNoKey;
        }
      }
      stage_put(sg, first, w, 0);
      stage_sync_flush(sg, false, Queue{}, tlist, tcount);
    }
  }
  stage_sync_flush(sg, true, Queue{}, tlist, tcount);
}

// Expansion, commit step: of a claimed matched pair (w, x) the member with
// the lower (claim, id) becomes odd (its claimer's level + 1) and its mate
// even (claimer's level + 2, queued for the next step).  Levels follow the
// claimer because the frontier of a reused forest mixes levels; they rise
// strictly down every tree path, which is all the blossom step relies on.
__global__ __launch_bounds__(kBlock) void k_mt_commit(const int32_t* __restrict__ tlist,
                                                      const uint32_t* __restrict__ tcount,
                                                      const int32_t* __restrict__ mate, const int32_t* __restrict__ off,
                                                      const unsigned long long* __restrict__ claim,
                                                      int32_t* __restrict__ root, int32_t* __restrict__ lvl,
                                                      int32_t* __restrict__ par, int32_t* __restrict__ fp,
                                                      Queue nxt) {
  AppendStage& sg = g_lds.q.sg;
  stage_init(sg);
  const uint32_t cnt = *tcount;
  for (int64_t b0 = (int64_t)blockIdx.x * kBlock; b0 < (int64_t)cnt; b0 += (int64_t)gridDim.x * kBlock) {
    const int64_t idx = b0 + threadIdx.x;
    bool odd = false;
    int32_t x = 0, dx = 0;
    if (idx < (int64_t)cnt) {
      const int32_t w = tlist[idx];
      x = mate[w];
      const unsigned long long kw = claim[w], kx = claim[x];
      odd = kw < kx || (kw == kx && w < x);
      if (odd) {
        const int32_t rv = (int32_t)(kw >> 32), cl = (int32_t)(kw & 0xFFFFFFFFull);
        const int32_t lc = lvl[cl];  // an outer vertex: not written by this pass
        root[w] = rv;
        par[w] = cl;
        lvl[w] = lc + 1;
        root[x] = rv;
        par[x] = w;
        fp[x] = w;
        lvl[x] = lc + 2;
        dx = off[x + 1] - off[x];
      }
    }
    stage_put(sg, odd, x, dx);
    stage_sync_flush(sg, false, nxt);
  }
  stage_sync_flush(sg, true, nxt);
}

// Validates a matching: ids in range, symmetric, and every pair an edge
// (binary search in the sorted adjacency list); counts violations.
__global__ __launch_bounds__(kBlock) void k_mt_validate(int32_t n, const int32_t* __restrict__ mate,
                                                        const int32_t* __restrict__ off,
                                                        const int32_t* __restrict__ col,
                                                        unsigned long long* __restrict__ bad) {
  for (int64_t v = (int64_t)blockIdx.x * kBlock + threadIdx.x; v < n; v += (int64_t)gridDim.x * kBlock) {
    const int32_t m = mate[v];
    if (m < 0) {
      if (m != -1) atomicAdd(bad, 1ull);
      continue;
    }
    if (m >= n || m == v || mate[m] != (int32_t)v) {
      atomicAdd(bad, 1ull);
      continue;
    }
    int32_t lo = off[v], hi = off[v + 1];
    while (lo < hi) {
      const int32_t mid = (lo + hi) >> 1;
      if (col[mid] < m)
        lo = mid + 1;
      else
        hi = mid;
    }
    if (lo >= off[v + 1] || col[lo] != m) atomicAdd(bad, 1ull);
  }
}

// Matched vertices: mate[v] >= 0 (for the result count).
__global__ __launch_bounds__(kBlock) void k_mt_count(int32_t n, const int32_t* __restrict__ mate,
                                                     unsigned long long* __restrict__ out) {
  unsigned long long c = 0;
  for (int64_t v = (int64_t)blockIdx.x * kBlock + threadIdx.x; v < n; v += (int64_t)gridDim.x * kBlock)
    c += mate[v] >= 0 ? 1ull : 0ull;
  c = wave_sum64(c);
  if (lane_id() == 0 && c) atomicAdd(out, c);
}

// ============================================================================
// Utilities and test hooks
// ============================================================================

// Sum over (source, vertex) with level >= 0 of the vertex's out-degree.
// blockIdx.y strides over sources, the x dimension over vertices.
__global__ __launch_bounds__(kBlock) void k_count_traversed(const int32_t* __restrict__ off,
                                                            const int32_t* __restrict__ levels, int32_t S,
                                                            int32_t n, unsigned long long* out) {
  unsigned long long acc = 0;
  for (int32_t s = blockIdx.y; s < S; s += gridDim.y) {
    const int32_t* row = levels + (int64_t)s * n;
    for (int32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x)
      if (row[v] >= 0) acc += (unsigned long long)(off[v + 1] - off[v]);
  }
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) acc += __shfl_xor(acc, d, kWave);
  if (lane_id() == 0) atomicAdd(out, acc);
}

// Order-independent checksum of an int32 array: sum over i of
// splitmix64(i * 0x9E3779B97F4A7C15 ^ (uint32)a[i]), modulo 2^64.
__device__ __forceinline__ uint64_t mix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

__global__ __launch_bounds__(kBlock) void k_checksum(const int32_t* __restrict__ a, int64_t count,
                                                     unsigned long long* out) {
  unsigned long long acc = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x)
    acc += mix64((uint64_t)i * 0x9E3779B97F4A7C15ull ^ (uint64_t)(uint32_t)a[i]);
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) acc += __shfl_xor(acc, d, kWave);
  if (lane_id() == 0) atomicAdd(out, acc);
}

// Appends (i, deg[i]) for i < k through the block-staged compaction.
__global__ __launch_bounds__(kBlock) void k_test_build_queue(const int32_t* __restrict__ deg, int32_t k, Queue q) {
  __shared__ QueueLds L;
  AppendStage& sg = L.sg;
  stage_init(sg);
  for (int64_t b0 = (int64_t)blockIdx.x * kBlock; b0 < k; b0 += (int64_t)gridDim.x * kBlock) {
    const int64_t i = b0 + threadIdx.x;
    const bool p = i < k;
    stage_put(sg, p, p ? (int32_t)i : 0, p ? deg[i] : 0);
    stage_sync_flush(sg, false, q);
  }
  stage_sync_flush(sg, true, q);
}

// Maps every edge tasklet e < E to (queue vertex, offset) with the chunked
// owner staging the push kernels use.
__global__ __launch_bounds__(kBlock) void k_test_edge_map(Queue q, int32_t qn, int32_t E, int32_t* __restrict__ owner,
                                                          int32_t* __restrict__ offset) {
  __shared__ QueueLds L;
  for (int64_t c0 = (int64_t)blockIdx.x * kChunk; c0 < E; c0 += (int64_t)gridDim.x * kChunk) {
    const int c1 = (int)((c0 + kChunk < E) ? c0 + kChunk : E);
    const ChunkOwners co = stage_owners(q.eo, qn, (int)c0, c1, L);
    for (int r = 0; r < kPushR; ++r) {
      const int e = (int)c0 + r * kBlock + (int)threadIdx.x;
      if (e < c1) {
        const int i = chunk_owner(co, q.eo, L.s_eo, e);
        owner[e] = q.v[i];
        offset[e] = e - q.eo[i];
      }
    }
    __syncthreads();
  }
}

}  // namespace

// ============================================================================
// Launch wrappers
// ============================================================================

void launch_plain_init(hipStream_t st, int32_t* level, int32_t* parent, int32_t n) {
  hipLaunchKernelGGL(k_plain_init, dim3(grid_for(((int64_t)n + 3) / 4, kBlock)), dim3(kBlock), 0, st, level,
                     parent, n);
}

void launch_plain_set_source(hipStream_t st, int32_t* level, int32_t* parent, int32_t src, Queue q) {
  hipLaunchKernelGGL(k_plain_set_source, dim3(1), dim3(64), 0, st, level, parent, src, q);
}

void launch_plain_push(hipStream_t st, const int32_t* off, const int32_t* col, Queue cur, int32_t qn, int32_t E,
                       int32_t* level, int32_t* parent, int32_t d, Queue nxt) {
  if (E <= 0) return;
  hipLaunchKernelGGL(k_plain_push, dim3(grid_for(E, kChunk)), dim3(kBlock), 0, st, off, col, cur, qn, E, level,
                     parent, d, nxt);
}

#define XB_DISPATCH_T(T, CALL) \
  switch (T) {                 \
    case 1: {                  \
      constexpr int TT = 1;    \
      CALL;                    \
    } break;                   \
    case 2: {                  \
      constexpr int TT = 2;    \
      CALL;                    \
    } break;                   \
    case 4: {                  \
      constexpr int TT = 4;    \
      CALL;                    \
    } break;                   \
    case 8: {                  \
      constexpr int TT = 8;    \
      CALL;                    \
    } break;                   \
    case 16: {                 \
      constexpr int TT = 16;   \
      CALL;                    \
    } break;                   \
    default:                   \
      break;                   \
  }

#define XB_DISPATCH_K(K, CALL) \
  switch (K) {                 \
    case 1: {                  \
      constexpr int KK = 1;    \
      CALL;                    \
    } break;                   \
    case 2: {                  \
      constexpr int KK = 2;    \
      CALL;                    \
    } break;                   \
    case 3: {                  \
      constexpr int KK = 3;    \
      CALL;                    \
    } break;                   \
    case 4: {                  \
      constexpr int KK = 4;    \
      CALL;                    \
    } break;                   \
    case 5: {                  \
      constexpr int KK = 5;    \
      CALL;                    \
    } break;                   \
    case 6: {                  \
      constexpr int KK = 6;    \
      CALL;                    \
    } break;                   \
    case 7: {                  \
      constexpr int KK = 7;    \
      CALL;                    \
    } break;                   \
    default: {                 \
      constexpr int KK = 8;    \
      CALL;                    \
    } break;                   \
  }

void launch_bp_seed(hipStream_t st, int T, BPState s, const int32_t* src, int32_t cnt) {
  XB_DISPATCH_T(T, hipLaunchKernelGGL(k_bp_seed<TT>, dim3(grid_for(cnt, kBlock)), dim3(kBlock), 0, st, s, src,
                                      cnt));
}

void launch_bp_push(hipStream_t st, int T, BPState s, int cur, const int32_t* off, const int32_t* col, Queue q,
                    int32_t qn, int32_t E) {
  if (E <= 0) return;
  XB_DISPATCH_T(T, hipLaunchKernelGGL(k_bp_push<TT>, dim3(grid_for(E, kChunk)), dim3(kBlock), 0, st, s, cur, off,
                                      col, q, qn, E));
}

void launch_bp_commit(hipStream_t st, int T, BPState s, int cur, int32_t code, const int32_t* out_off, Queue nxt,
                      int32_t max_touched) {
  XB_DISPATCH_T(T, hipLaunchKernelGGL(k_bp_commit<TT>, dim3(grid_for(max_touched, kBlock)), dim3(kBlock), 0, st,
                                      s, cur, (uint32_t)code, out_off, nxt));
}

template <int T, int P, bool REC>
static void pull_bucket_launch(hipStream_t st, BPState s, int cur, int32_t code, int32_t cnt, const int32_t* in_off,
                               const int32_t* in_col, const int32_t* out_off, const int32_t* perm, int32_t shift,
                               int32_t begin, int32_t end, Queue nxt) {
  constexpr int VPB = (kBlock / kWave) * (kWave / (T * P));
  hipLaunchKernelGGL((k_bp_pull_bucket<T, P, REC>), dim3(grid_for((int64_t)(end - begin), VPB)), dim3(kBlock), 0,
                     st, s, cur, (uint32_t)code, cnt, in_off, in_col, out_off, perm, shift, begin, end, nxt);
}

void launch_bp_pull_bucket(hipStream_t st, int T, int cls, BPState s, int cur, int32_t code, int32_t cnt,
                           const int32_t* in_off, const int32_t* in_col, const int32_t* out_off,
                           const int32_t* perm, int32_t shift, int32_t begin, int32_t end, Queue nxt) {
  if (end <= begin) return;
  if (cls == 0) {
    XB_DISPATCH_T(T, (pull_bucket_launch<TT, 1, false>(st, s, cur, code, cnt, in_off, in_col, out_off, perm, shift,
                                                       begin, end, nxt)));
  } else if (cls == 1) {
    XB_DISPATCH_T(T, (pull_bucket_launch<TT, (TT * 4 <= kWave ? 4 : kWave / TT), false>(
                         st, s, cur, code, cnt, in_off, in_col, out_off, perm, shift, begin, end, nxt)));
  } else {
    XB_DISPATCH_T(T, (pull_bucket_launch<TT, kWave / TT, true>(st, s, cur, code, cnt, in_off, in_col, out_off, perm,
                                                               shift, begin, end, nxt)));
  }
}

void launch_bp_pull_hub(hipStream_t st, int T, BPState s, int cur, int32_t cnt, const int32_t* in_col,
                        const HubChunk* chunks, int32_t n_chunks) {
  if (n_chunks <= 0) return;
  XB_DISPATCH_T(T, hipLaunchKernelGGL(k_bp_pull_hub<TT>, dim3(grid_for((int64_t)n_chunks * kWave, kBlock)),
                                      dim3(kBlock), 0, st, s, cur, cnt, in_col, chunks, n_chunks));
}

template <int T>
static void pull_fused_launch(hipStream_t st, BPState s, int cur, int32_t code, int32_t cnt, const int32_t* in_off,
                              const int32_t* in_col, const int32_t* out_off, const PullPlan& plan, Queue nxt,
                              const HubChunk* chunks, int32_t n_chunks) {
  constexpr int P1 = T * 4 <= kWave ? 4 : kWave / T;
  constexpr int P2 = kWave / T;
  const int vpb[3] = {(kBlock / kWave) * (kWave / T), (kBlock / kWave) * (kWave / (T * P1)),
                      (kBlock / kWave) * (kWave / (T * P2))};
  PullSplit split;
  int64_t total = 0;
  for (int c = 0; c < 3; ++c) {
    split.nb[c] = plan.end[c] > plan.begin[c] ? grid_for((int64_t)(plan.end[c] - plan.begin[c]), vpb[c]) : 0;
    total += split.nb[c];
  }
  split.nb_hub = n_chunks > 0 ? grid_for((int64_t)n_chunks * kWave, kBlock) : 0;
  total += split.nb_hub;
  if (total == 0) return;
  hipLaunchKernelGGL(k_bp_pull_fused<T>, dim3((unsigned)total), dim3(kBlock), 0, st, s, cur, (uint32_t)code, cnt,
                     in_off, in_col, out_off, plan, split, nxt, chunks, n_chunks);
}

void launch_bp_pull_fused(hipStream_t st, int T, BPState s, int cur, int32_t code, int32_t cnt,
                          const int32_t* in_off, const int32_t* in_col, const int32_t* out_off, const PullPlan& plan,
                          Queue nxt, const HubChunk* chunks, int32_t n_chunks) {
  XB_DISPATCH_T(T, (pull_fused_launch<TT>(st, s, cur, code, cnt, in_off, in_col, out_off, plan, nxt, chunks,
                                          n_chunks)));
}

void launch_bp_clear(hipStream_t st, int T, BPState s, int cur, Queue q, int32_t qn) {
  if (qn <= 0) return;
  XB_DISPATCH_T(T, hipLaunchKernelGGL(k_bp_clear<TT>, dim3(grid_for((int64_t)qn * TT, kBlock)), dim3(kBlock), 0,
                                      st, s, cur, q.v, qn));
}

template <int T, int K>
static void emit_light_launch(hipStream_t st, BPState s, int32_t grp, int32_t ngrp, int32_t rows, const int32_t* src,
                              const int32_t* in_off, const int32_t* in_col, const int32_t* in_col_orig, int32_t n,
                              int32_t heavy_cut, int32_t* lvl_out, int32_t* par_out, RowMap map, bool symmetric,
                              const int32_t* gate, int32_t gate_lo, int32_t gate_hi) {
  if (symmetric)
    hipLaunchKernelGGL((k_bp_emit_light<T, K, true>), dim3(grid_for(n, 64)), dim3(kBlock * kEmitWords<T>), 0, st, s,
                       grp, ngrp, rows, src, in_off, in_col, in_col_orig, n, heavy_cut, lvl_out, par_out, map, gate,
                       gate_lo, gate_hi);
  else
    hipLaunchKernelGGL((k_bp_emit_light<T, K, false>), dim3(grid_for(n, 64)), dim3(kBlock * kEmitWords<T>), 0, st, s,
                       grp, ngrp, rows, src, in_off, in_col, in_col_orig, n, heavy_cut, lvl_out, par_out, map, gate,
                       gate_lo, gate_hi);
}

int emit_pass_words(int T) { return T < XB_EMIT_WORDS ? T : XB_EMIT_WORDS; }

void launch_bp_emit(hipStream_t st, int T, BPState s, int32_t word, int32_t ngrp, int32_t K, int32_t rows,
                    const int32_t* src, const int32_t* in_off, const int32_t* in_col, const int32_t* in_col_orig,
                    int32_t n, int32_t heavy_cut, int32_t* lvl_out, int32_t* par_out, RowMap map, bool symmetric,
                    const int32_t* gate, int32_t gate_lo, int32_t gate_hi) {
  XB_DISPATCH_T(T, XB_DISPATCH_K(K, (emit_light_launch<TT, KK>(st, s, word, ngrp, rows, src, in_off, in_col,
                                                               in_col_orig, n, heavy_cut, lvl_out, par_out, map,
                                                               symmetric, gate, gate_lo, gate_hi))));
}

int persistent_blocks_per_cu() {
  int nb = 0;
  (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_bp_persistent, kBlock, 0);
  return nb;
}

void launch_bp_persistent(hipStream_t st, int grid_blocks, const Team* teams, int32_t n_teams,
                          const GraphView* graphs, TileJob* jobs, int32_t n_jobs, TeamCtl* ctl, int32_t alpha,
                          unsigned* gerr) {
  const int nr = n_jobs > 1024 ? n_jobs : 1024;
  hipLaunchKernelGGL(k_reset_ctl, dim3((nr + 255) / 256), dim3(256), 0, st, ctl, n_teams, jobs, n_jobs, gerr);
  hipLaunchKernelGGL(k_bp_persistent, dim3(grid_blocks), dim3(kBlock), 0, st, teams, n_teams, graphs, jobs, alpha,
                     gerr);
}

// ---- forest mode ----
static int edge_grid(int32_t E) {
  const int64_t g = ((int64_t)E + kChunk - 1) / kChunk;
  return (int)(g < 1 ? 1 : (g > kMaxBlocks ? kMaxBlocks : g));
}
static int item_grid(int64_t n) {
  const int64_t g = (n + kBlock - 1) / kBlock;
  return (int)(g < 1 ? 1 : (g > kMaxBlocks ? kMaxBlocks : g));
}
// Kernels that append through a block stage or end with a per-block atomic:
// 2,048 blocks (8 waves per SIMD on 256 CUs) and a grid-stride loop, so
// that stages flush full (one global atomic per ~1,280 entries) and the
// counters see one atomic per block, not per 256 items.
static int stage_grid(int64_t n) {
  const int g = item_grid(n);
  return g < 2048 ? g : 2048;
}

void launch_mt_init(hipStream_t st, int32_t n, const int32_t* mate, const int32_t* off, int32_t* root, int32_t* lvl,
                    int32_t* par, unsigned long long* claim, int32_t* best, int32_t* acc, unsigned long long* bridge,
                    uint8_t* used, int32_t* cursor, int32_t* fp, unsigned long long* elab, Queue q) {
  hipLaunchKernelGGL(k_mt_init, dim3(stage_grid(n)), dim3(kBlock), 0, st, n, mate, off, root, lvl, par, claim, best,
                     acc, bridge, used, cursor, fp, elab, q);
}
void launch_mt_reuse_init(hipStream_t st, int32_t n, const int32_t* mate, const int32_t* off, int32_t* root,
                          int32_t* lvl, int32_t* par, int32_t* fp, unsigned long long* elab, unsigned long long* claim,
                          const uint8_t* used, Queue q, unsigned long long* counts) {
  hipLaunchKernelGGL(k_mt_reuse_init, dim3(stage_grid(n)), dim3(kBlock), 0, st, n, mate, off, root, lvl, par, fp, elab,
                     claim, used, q, counts);
}
void launch_mt_reuse_clear(hipStream_t st, int32_t n, int32_t* best, int32_t* acc, unsigned long long* bridge,
                           uint8_t* used) {
  hipLaunchKernelGGL(k_mt_reuse_clear, dim3(item_grid(n)), dim3(kBlock), 0, st, n, best, acc, bridge, used);
}
void launch_mt_scan0(hipStream_t st, const int32_t* roots, int32_t nroots, const int32_t* off, const int32_t* col,
                     const int32_t* lvl, const uint8_t* used, int32_t* cursor, int32_t* best,
                     unsigned long long* examined) {
  if (nroots <= 0) return;
  hipLaunchKernelGGL(k_mt_scan0, dim3(stage_grid(nroots)), dim3(kBlock), 0, st, roots, nroots, off, col, lvl, used,
                     cursor, best, examined);
}
void launch_mt_scan(hipStream_t st, const int32_t* off, const int32_t* col, Queue q, int32_t qn, int32_t E,
                    const int32_t* root, const int32_t* lvl, const uint8_t* used, int32_t* best,
                    unsigned long long* blossoms) {
  if (E <= 0) return;
  hipLaunchKernelGGL(k_mt_scan, dim3(edge_grid(E)), dim3(kBlock), 0, st, off, col, q, qn, E, root, lvl, used, best,
                     blossoms);
}
void launch_mt_accept(hipStream_t st, const int32_t* roots, int32_t nroots, const int32_t* best, int32_t* acc,
                      int32_t* pairs, uint32_t* npairs, unsigned long long* bridge0) {
  if (nroots <= 0) return;
  hipLaunchKernelGGL(k_mt_accept, dim3(stage_grid(nroots)), dim3(kBlock), 0, st, roots, nroots, best, acc, pairs,
                     npairs, bridge0);
}
void launch_mt_bridge(hipStream_t st, const int32_t* off, const int32_t* col, Queue q, int32_t qn, int32_t E,
                      const int32_t* root, const int32_t* lvl, const int32_t* acc, unsigned long long* bridge) {
  if (E <= 0) return;
  hipLaunchKernelGGL(k_mt_bridge, dim3(edge_grid(E)), dim3(kBlock), 0, st, off, col, q, qn, E, root, lvl, acc,
                     bridge);
}
void launch_mt_flip(hipStream_t st, const int32_t* pairs, const uint32_t* npairs, int32_t max_pairs,
                    const int32_t* acc, const unsigned long long* bridge, const int32_t* par,
                    const unsigned long long* elab, int32_t* mate, uint8_t* used, FlipStack stk, int32_t max_steps,
                    unsigned* err) {
  if (max_pairs <= 0) return;
  hipLaunchKernelGGL(k_mt_flip, dim3(item_grid(max_pairs)), dim3(kBlock), 0, st, pairs, npairs, acc, bridge, par,
                     elab, mate, used, stk, max_steps, err);
}
void launch_mt_collect(hipStream_t st, const int32_t* off, const int32_t* col, Queue q, int32_t qn, int32_t E,
                       const int32_t* root, const int32_t* lvl, unsigned long long* cand, uint32_t* ncand,
                       uint32_t cap, unsigned* err) {
  if (E <= 0) return;
  hipLaunchKernelGGL(k_mt_collect, dim3(edge_grid(E)), dim3(kBlock), 0, st, off, col, q, qn, E, root, lvl, cand,
                     ncand, cap, err);
}
// The collected blossom candidates (packed 64-bit keys) in ascending order,
// so the mark step can keep the lowest candidate per inner vertex: a rocPRIM
// device radix sort over all 64 key bits.
size_t sort_u64_temp_bytes(int32_t n) {
  size_t bytes = 0;
  const hipError_t e = rocprim::radix_sort_keys(nullptr, bytes, (const unsigned long long*)nullptr,
                                                (unsigned long long*)nullptr, (size_t)(n > 0 ? n : 1), 0, 64);
  if (e != hipSuccess) throw std::runtime_error(std::string("radix sort size query: ") + hipGetErrorString(e));
  return bytes;
}
void sort_u64(void* temp, size_t bytes, const unsigned long long* in, unsigned long long* out, int32_t n,
              hipStream_t st) {
  if (n <= 0) return;
  const hipError_t e = rocprim::radix_sort_keys(temp, bytes, in, out, (size_t)n, 0, 64, st);
  if (e != hipSuccess) throw std::runtime_error(std::string("radix sort: ") + hipGetErrorString(e));
}
void launch_mt_blossom(hipStream_t st, const int32_t* off, const int32_t* col, int32_t n, const int32_t* par,
                       int32_t* lvl, int32_t* fp, unsigned long long* elab, unsigned long long* lab,
                       const unsigned long long* cand, int32_t C, int32_t* join_of, int32_t* marked,
                       uint32_t* nmarked, uint32_t* formed, Queue nxt, unsigned long long* counts, unsigned* err) {
  if (C <= 0) return;
  const BlossomState b{off, col, n, par, lvl, fp, elab, lab};
  hipLaunchKernelGGL(k_mt_blossom_mark, dim3(item_grid(C)), dim3(kBlock), 0, st, b, cand, C, join_of, marked, nmarked,
                     err);
  // Every vertex is marked at most once per step: n bounds the list.
  hipLaunchKernelGGL(k_mt_blossom_apply, dim3(item_grid(n)), dim3(kBlock), 0, st, b, cand, join_of, marked, nmarked,
                     formed, nxt, counts);
}
void launch_mt_compact(hipStream_t st, const int32_t* in, int32_t cnt, const int32_t* root, const uint8_t* used,
                       const int32_t* off, Queue out, int32_t* list, uint32_t* ctr32) {
  if (cnt <= 0) return;
  hipLaunchKernelGGL(k_mt_compact, dim3(stage_grid(cnt)), dim3(kBlock), 0, st, in, cnt, root, used, off, out, list,
                     ctr32);
}
void launch_mt_zero_ctrs(hipStream_t st, uint32_t* a, uint32_t* b, uint32_t* c, unsigned long long* d) {
  hipLaunchKernelGGL(k_mt_zero_ctrs, dim3(1), dim3(64), 0, st, a, b, c, d);
}
void launch_mt_reset_best(hipStream_t st, const int32_t* roots, int32_t nroots, int32_t* best) {
  if (nroots <= 0) return;
  hipLaunchKernelGGL(k_mt_reset_best, dim3(item_grid(nroots)), dim3(kBlock), 0, st, roots, nroots, best);
}
void launch_mt_claim(hipStream_t st, const int32_t* off, const int32_t* col, Queue q, int32_t qn, int32_t E,
                     const int32_t* root, const int32_t* lvl, unsigned long long* claim, int32_t* tlist,
                     uint32_t* tcount) {
  if (E <= 0) return;
  hipLaunchKernelGGL(k_mt_claim, dim3(edge_grid(E)), dim3(kBlock), 0, st, off, col, q, qn, E, root, lvl, claim,
                     tlist, tcount);
}
void launch_mt_commit(hipStream_t st, const int32_t* tlist, const uint32_t* tcount, int32_t max_t,
                      const int32_t* mate, const int32_t* off, const unsigned long long* claim, int32_t* root,
                      int32_t* lvl, int32_t* par, int32_t* fp, Queue nxt) {
  if (max_t <= 0) return;
  hipLaunchKernelGGL(k_mt_commit, dim3(stage_grid(max_t)), dim3(kBlock), 0, st, tlist, tcount, mate, off, claim, root,
                     lvl, par, fp, nxt);
}
void launch_mt_validate(hipStream_t st, int32_t n, const int32_t* mate, const int32_t* off, const int32_t* col,
                        unsigned long long* bad) {
  hipLaunchKernelGGL(k_mt_validate, dim3(item_grid(n)), dim3(kBlock), 0, st, n, mate, off, col, bad);
}
void launch_mt_count(hipStream_t st, int32_t n, const int32_t* mate, unsigned long long* out) {
  hipLaunchKernelGGL(k_mt_count, dim3(item_grid(n) < 1024 ? item_grid(n) : 1024), dim3(kBlock), 0, st, n, mate, out);
}

void launch_count_traversed(hipStream_t st, const int32_t* off, const int32_t* levels, int32_t S, int32_t n,
                            unsigned long long* out) {
  const int gx = grid_for(n, kBlock) < 64 ? grid_for(n, kBlock) : 64;
  const int gy = S < 256 ? S : 256;
  hipLaunchKernelGGL(k_count_traversed, dim3(gx, gy), dim3(kBlock), 0, st, off, levels, S, n, out);
}

void launch_checksum(hipStream_t st, const int32_t* a, int64_t count, unsigned long long* out) {
  hipLaunchKernelGGL(k_checksum, dim3(grid_for(count, kBlock * 8)), dim3(kBlock), 0, st, a, count, out);
}

void launch_test_build_queue(hipStream_t st, const int32_t* deg, int32_t k, Queue q) {
  hipLaunchKernelGGL(k_test_build_queue, dim3(grid_for(k, kBlock)), dim3(kBlock), 0, st, deg, k, q);
}

void launch_test_edge_map(hipStream_t st, Queue q, int32_t qn, int32_t E, int32_t* owner, int32_t* offset) {
  if (E <= 0) return;
  hipLaunchKernelGGL(k_test_edge_map, dim3(grid_for(E, kChunk)), dim3(kBlock), 0, st, q, qn, E, owner, offset);
}

}  // namespace xb
