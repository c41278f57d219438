// fp_round.h -- device pieces shared by the round schedules whose apply
// tracks a "max set" (the parts holding the maximum of a per-part value) and
// stops a round when it gains a part: MSV, MSRV (send / send + receive
// volume, MGV/partition.cpp:977-1003) and MSM (messages,
// MSM/partition.cpp:1335-1370). Included by one translation unit each; the
// kernels live in an anonymous namespace.
#pragma once
#include "fp_internal.h"

namespace fpe {
namespace {

// Accepted moves of the round in visiting order: slot_mi[slot] = move index
// (or -1), mv[i] = (vertex, own part, new part, first event slot), and after
// the last move a sentinel whose event slot is the round's event total (mv
// holds one more entry than the round has slots).
constexpr int COMPACT_RUN = 16;  // consecutive slots per thread per step of k_round_compact

__global__ void __launch_bounds__(1024) k_round_compact(const int* __restrict__ slot_dec, const int* __restrict__ slot_own,
                                                      const int* __restrict__ order,
                                                      const unsigned* __restrict__ vtx_ptr, int r0, int nslots,
                                                      int* __restrict__ slot_mi, int4* __restrict__ mv,
                                                      int* __restrict__ mstat) {
  // each thread owns a run of COMPACT_RUN consecutive slots, read and written
  // with 16-byte accesses; accepted slots are few, so their vertices and
  // degrees are loaded per set bit of the run's acceptance mask
  static_assert(COMPACT_RUN == 16, "the run is four int4 words");
  __shared__ int s_w[16];
  int base = 0, ebase = 0;
  const int step = (int)blockDim.x * COMPACT_RUN;
  for (int s0 = 0; s0 < nslots; s0 += step) {
    const int first = s0 + (int)threadIdx.x * COMPACT_RUN;
    const bool whole = first + COMPACT_RUN <= nslots;
    unsigned acc = 0;
    if (whole) {
      const int4* p = reinterpret_cast<const int4*>(slot_dec + first);
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int4 x = p[q];
        acc |= (x.x >= 0 ? 1u : 0u) << (4 * q);
        acc |= (x.y >= 0 ? 2u : 0u) << (4 * q);
        acc |= (x.z >= 0 ? 4u : 0u) << (4 * q);
        acc |= (x.w >= 0 ? 8u : 0u) << (4 * q);
      }
    } else {
      for (int j = 0; j < COMPACT_RUN && first + j < nslots; j++)
        if (slot_dec[first + j] >= 0) acc |= 1u << j;
    }
    int deg = 0;
    for (unsigned m = acc; m; m &= m - 1) {
      const int s = first + __builtin_ctz(m);
      const int v = order ? order[r0 + s] : r0 + s;
      deg += (int)(vtx_ptr[v + 1] - vtx_ptr[v]);
    }
    int tot, totd;
    const int ex = block_excl_scan((int)__popc(acc), s_w, &tot);
    int exd = block_excl_scan(deg, s_w, &totd);
    if (whole) {
      int4* q4 = reinterpret_cast<int4*>(slot_mi + first);
#pragma unroll
      for (int q = 0; q < 4; q++) {
        int w[4];
#pragma unroll
        for (int t = 0; t < 4; t++) {
          const int j = 4 * q + t;
          w[t] = (acc >> j) & 1u ? base + ex + __popc(acc & ((1u << j) - 1u)) : -1;
        }
        q4[q] = make_int4(w[0], w[1], w[2], w[3]);
      }
    } else {
      for (int j = 0; j < COMPACT_RUN && first + j < nslots; j++)
        slot_mi[first + j] = (acc >> j) & 1u ? base + ex + __popc(acc & ((1u << j) - 1u)) : -1;
    }
    int i = base + ex;
    for (unsigned m = acc; m; m &= m - 1) {
      const int s = first + __builtin_ctz(m);
      const int v = order ? order[r0 + s] : r0 + s;
      mv[i++] = make_int4(v, slot_own[s], slot_dec[s], ebase + exd);
      exd += (int)(vtx_ptr[v + 1] - vtx_ptr[v]);
    }
    base += tot;
    ebase += totd;
  }
  if (threadIdx.x == 0) {
    mstat[2] = base;
    mstat[3] = base;
    mv[base] = make_int4(-1, -1, -1, ebase);  // end sentinel: move i has mv[i + 1].w - mv[i].w events
  }
}

// Max-set tracking by one wavefront over val[0..K) in LDS (K <= 64 * 64).
// Lane l holds the parts l, l + 64, l + 128, ...: their maximum (lmax) and
// the bitmask of those holding it (lbits, bit j = part 64 j + l). The max
// set is the union of lbits over the lanes whose lmax equals the wave max;
// inS is this lane's share of it.
struct LaneMax {
  int lmax;
  unsigned long long lbits, inS;
};

__device__ __forceinline__ void lane_rescan(const int* val, int K, LaneMax& s) {
  int m = INT_MIN;
  unsigned long long bits = 0;
  int j = 0;
  for (int p = lane_id(); p < K; p += WAVE, j++) {
    const int x = val[p];
    const unsigned long long bj = 1ull << j;
    bits = x > m ? bj : (x == m ? bits | bj : bits);
    m = max(m, x);
  }
  s.lmax = m;
  s.lbits = bits;
}

// Initial state from val; returns the max.
__device__ __forceinline__ int maxset_init(const int* val, int K, LaneMax& s) {
  lane_rescan(val, K, s);
  const int M = wave_max(s.lmax);
  s.inS = s.lmax == M ? s.lbits : 0ull;
  return M;
}

// One move's step: val holds the move's changes and dirty[l] != 0 marks the
// lanes l whose parts changed (cleared here). Updates the max M and returns
// true iff the new max set holds a part the previous one did not.
__device__ __forceinline__ bool maxset_step(const int* val, int* dirty, int K, LaneMax& s, int& M) {
  const int lane = lane_id();
  if (dirty[lane]) {
    dirty[lane] = 0;
    lane_rescan(val, K, s);
  }
  M = wave_max(s.lmax);
  const unsigned long long nS = s.lmax == M ? s.lbits : 0ull;
  const bool enter = __ballot((nS & ~s.inS) != 0ull) != 0ull;
  s.inS = nS;
  return enter;
}

// number of parts in the max set
__device__ __forceinline__ int maxset_count(const LaneMax& s) { return wave_sum(__popcll(s.inS)); }

}  // namespace
}  // namespace fpe
