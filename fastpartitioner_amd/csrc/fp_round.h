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
// holds one more entry than the round has slots). WEIGHTS rounds (TM, whose
// apply reads no event slots and most of whose slots hold a move) store the
// vertex weight in place of the event slot, and 0 in the sentinel. One block
// per tile of COMPACT_TILE slots; with several tiles, k_compact_count first
// writes each tile's move and event totals and every block adds up the ones
// before its own.
constexpr int COMPACT_RUN = 16;  // consecutive slots per thread
constexpr int COMPACT_BLOCK = 1024;
constexpr int COMPACT_TILE = COMPACT_BLOCK * COMPACT_RUN;

// One thread's run of COMPACT_RUN slots: the acceptance mask (bit j: slot
// first + j holds a move) and the decisions. WEIGHTS: the accepted slots'
// vertices and weights are loaded here, all loads of the run issued
// together, the indexes of the slots without a move clamped to the run's
// first slot so that every load stays in bounds. Otherwise (few moves per
// round) vertices and degrees are loaded per set bit where they are used.
struct CompactRun {
  unsigned acc;
  int dec[COMPACT_RUN], v[COMPACT_RUN], w[COMPACT_RUN];
};

template <bool WEIGHTS>
__device__ __forceinline__ void compact_load(const int* __restrict__ slot_dec, const int* __restrict__ order,
                                             const int* __restrict__ vwgt, int r0, int nslots, int first,
                                             CompactRun& r) {
  static_assert(COMPACT_RUN == 16, "the run is four int4 words");
  r.acc = 0;
  if (first >= nslots) return;
  if (first + COMPACT_RUN <= nslots) {
    const int4* p = reinterpret_cast<const int4*>(slot_dec + first);
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int4 x = p[q];
      r.dec[4 * q] = x.x;
      r.dec[4 * q + 1] = x.y;
      r.dec[4 * q + 2] = x.z;
      r.dec[4 * q + 3] = x.w;
    }
  } else {
#pragma unroll
    for (int j = 0; j < COMPACT_RUN; j++) {
      const bool in = first + j < nslots;
      const int x = slot_dec[in ? first + j : first];
      r.dec[j] = in ? x : -1;
    }
  }
#pragma unroll
  for (int j = 0; j < COMPACT_RUN; j++) r.acc |= (r.dec[j] >= 0 ? 1u : 0u) << j;
  if (!WEIGHTS || !r.acc) return;
#pragma unroll
  for (int j = 0; j < COMPACT_RUN; j++) {
    const int s = (r.acc >> j) & 1u ? first + j : first;
    r.v[j] = order ? order[r0 + s] : r0 + s;
  }
#pragma unroll
  for (int j = 0; j < COMPACT_RUN; j++) r.w[j] = vwgt ? vwgt[r.v[j]] : 1;
}

// events of the run's moves (their vertices' degrees); none for WEIGHTS
template <bool WEIGHTS>
__device__ __forceinline__ int compact_events(const CompactRun& r, const int* __restrict__ order,
                                              const unsigned* __restrict__ vtx_ptr, int r0, int first) {
  int d = 0;
  if (!WEIGHTS) {
    for (unsigned m = r.acc; m; m &= m - 1) {
      const int s = first + __builtin_ctz(m);
      const int v = order ? order[r0 + s] : r0 + s;
      d += (int)(vtx_ptr[v + 1] - vtx_ptr[v]);
    }
  }
  return d;
}

template <bool WEIGHTS>
__global__ void __launch_bounds__(COMPACT_BLOCK) k_compact_count(const int* __restrict__ slot_dec,
                                                                 const int* __restrict__ order,
                                                                 const unsigned* __restrict__ vtx_ptr, int r0,
                                                                 int nslots, int2* __restrict__ tiles) {
  __shared__ int s_w[COMPACT_BLOCK / WAVE];
  const int first = (int)blockIdx.x * COMPACT_TILE + (int)threadIdx.x * COMPACT_RUN;
  CompactRun r;
  compact_load<false>(slot_dec, order, nullptr, r0, nslots, first, r);
  int tot, totd;
  block_excl_scan((int)__popc(r.acc), s_w, &tot);
  block_excl_scan(compact_events<WEIGHTS>(r, order, vtx_ptr, r0, first), s_w, &totd);
  if (threadIdx.x == 0) tiles[blockIdx.x] = make_int2(tot, totd);
}

// tiles: the per-tile totals of k_compact_count, or null for a single tile
template <bool WEIGHTS>
__global__ void __launch_bounds__(COMPACT_BLOCK) k_round_compact(const int* __restrict__ slot_dec,
                                                                 const int* __restrict__ slot_own,
                                                                 const int* __restrict__ order,
                                                                 const unsigned* __restrict__ vtx_ptr,
                                                                 const int* __restrict__ vwgt, int r0, int nslots,
                                                                 const int2* __restrict__ tiles,
                                                                 int* __restrict__ slot_mi, int4* __restrict__ mv,
                                                                 int* __restrict__ mstat) {
  __shared__ int s_w[COMPACT_BLOCK / WAVE];
  int base = 0, ebase = 0;
  if (tiles) {
    int a = 0, e = 0;
    for (int t = (int)threadIdx.x; t < (int)blockIdx.x; t += COMPACT_BLOCK) {
      const int2 x = tiles[t];
      a += x.x;
      e += x.y;
    }
    block_excl_scan(a, s_w, &base);
    block_excl_scan(e, s_w, &ebase);
  }
  const int first = (int)blockIdx.x * COMPACT_TILE + (int)threadIdx.x * COMPACT_RUN;
  CompactRun r;
  compact_load<WEIGHTS>(slot_dec, order, vwgt, r0, nslots, first, r);
  int tot, totd;
  const int ex = block_excl_scan((int)__popc(r.acc), s_w, &tot);
  const int exd = block_excl_scan(compact_events<WEIGHTS>(r, order, vtx_ptr, r0, first), s_w, &totd);
  const unsigned acc = r.acc;
  if (first + COMPACT_RUN <= nslots) {
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
  if (WEIGHTS && acc) {
    int own[COMPACT_RUN];
#pragma unroll
    for (int j = 0; j < COMPACT_RUN; j++) own[j] = slot_own[(acc >> j) & 1u ? first + j : first];
    int i = base + ex;
#pragma unroll
    for (int j = 0; j < COMPACT_RUN; j++) {
      if ((acc >> j) & 1u) mv[i++] = make_int4(r.v[j], own[j], r.dec[j], r.w[j]);
    }
  } else if (!WEIGHTS) {
    int i = base + ex, e = ebase + exd;
    for (unsigned m = acc; m; m &= m - 1) {
      const int s = first + __builtin_ctz(m);
      const int v = order ? order[r0 + s] : r0 + s;
      mv[i++] = make_int4(v, slot_own[s], slot_dec[s], e);
      e += (int)(vtx_ptr[v + 1] - vtx_ptr[v]);
    }
  }
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) {
    mstat[2] = base + tot;
    mstat[3] = base + tot;
    mv[base + tot] = make_int4(-1, -1, -1, ebase + totd);  // end sentinel: move i has mv[i + 1].w - mv[i].w events
  }
}

// Tile totals for rounds of more than one compaction tile (allocated at plan
// time, so that recording a pass allocates nothing).
inline fp_status ensure_compact_tiles(fp_ctx* ctx, int R) {
  const int nb = (R + COMPACT_TILE - 1) / COMPACT_TILE;
  if (nb > ctx->compact_cap) {
    FP_TRY(dalloc(ctx, &ctx->compact_tiles, (size_t)nb));
    ctx->compact_cap = nb;
  }
  return FP_OK;
}

template <bool WEIGHTS>
inline void launch_round_compact(fp_ctx* ctx, const int* order, int r0, int nslots, int* slot_mi, int4* mv,
                                 int* mstat) {
  const int nb = std::max(1, (nslots + COMPACT_TILE - 1) / COMPACT_TILE);
  if (nb > 1)
    k_compact_count<WEIGHTS><<<nb, COMPACT_BLOCK, 0, ctx->stream>>>(ctx->slot_dec, order, ctx->vtx_ptr, r0, nslots,
                                                                    ctx->compact_tiles);
  k_round_compact<WEIGHTS><<<nb, COMPACT_BLOCK, 0, ctx->stream>>>(
      ctx->slot_dec, ctx->slot_own, order, ctx->vtx_ptr, ctx->vwgt, r0, nslots, nb > 1 ? ctx->compact_tiles : nullptr,
      slot_mi, mv, mstat);
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
