// xb_internal.hpp -- shared declarations between the host engine
// (xb_host.cpp) and the HIP kernels (xb_kernels.hip).  Not installed.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace xb {

// Bit-parallel tile state.  All arrays are vertex-major with T 64-bit source
// words per vertex ([v*T + t]); planes hold 8 bit-sliced planes of the level
// code (code = level + 1, 0 = unreached) per (vertex, word):
// planes[(v*T + t)*8 + k] bit b = bit k of code(source 64t+b, v).
struct BPState {
  uint64_t* seen;
  uint64_t* front[2];
  uint64_t* nextw;
  uint64_t* planes;
  uint32_t* tflag;
  int32_t* tlist;          // touched vertices of a push level
  uint32_t* tcount;        // device count of tlist
};

// Frontier queue: vertex ids plus the exclusive prefix of their out-degrees
// in queue order; ctr packs (edges << 32 | count) so one 64-bit atomic
// reserves both a queue slot range and an edge range.
struct Queue {
  int32_t* v;
  int32_t* eo;
  unsigned long long* ctr;
};

// ---- plain per-source engine ----
void launch_plain_init(hipStream_t st, int32_t* level, int32_t* parent, int32_t n, int32_t src);
void launch_plain_push(hipStream_t st, const int32_t* off, const int32_t* col, Queue cur, int32_t qn,
                       int32_t E, int32_t* level, int32_t* parent, int32_t d, Queue nxt);

// ---- batched bit-parallel engine (T = words per tile: 1, 2, 4 or 8) ----
void launch_bp_seed(hipStream_t st, int T, BPState s, const int32_t* src, int32_t cnt);
void launch_bp_push(hipStream_t st, int T, BPState s, int cur, const int32_t* off, const int32_t* col,
                    Queue q, int32_t qn, int32_t E);
void launch_bp_commit(hipStream_t st, int T, BPState s, int cur, int32_t code, const int32_t* out_off,
                      Queue nxt, int32_t max_touched);
void launch_bp_pull(hipStream_t st, int T, BPState s, int cur, int32_t code, int32_t cnt,
                    const int32_t* in_off, const int32_t* in_col, const int32_t* out_off, int32_t n,
                    int32_t heavy_cut, const int32_t* heavy, int32_t n_heavy, Queue nxt);
void launch_bp_clear(hipStream_t st, int T, BPState s, int cur, Queue q, int32_t qn);
void launch_bp_emit(hipStream_t st, int T, BPState s, int32_t word, int32_t K, int32_t rows,
                    const int32_t* in_off, const int32_t* in_col, int32_t n, int32_t heavy_cut,
                    const int32_t* heavy, const int32_t* hidx, int32_t n_heavy, int32_t* side,
                    int32_t* lvl_out, int32_t* par_out);

// ---- utilities ----
void launch_count_traversed(hipStream_t st, const int32_t* off, const int32_t* levels, int32_t S, int32_t n,
                            unsigned long long* out);

}  // namespace xb
