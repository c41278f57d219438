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

__device__ __forceinline__ void wave_max_count(int& m, int& c) {
  for (int off = WAVE / 2; off > 0; off >>= 1) {
    const int om = __shfl_xor(m, off), oc = __shfl_xor(c, off);
    if (om > m) {
      m = om;
      c = oc;
    } else if (om == m) {
      c += oc;
    }
  }
}

// mstat[0] = max of val, mstat[1] = number of parts holding it (MGV/partition.cpp:712-729)
__global__ void __launch_bounds__(1024) k_round_maxset(const int* __restrict__ val, int K, int* __restrict__ mstat) {
  __shared__ int s_m[16], s_c[16];
  int m = INT_MIN, c = 0;
  for (int q = (int)threadIdx.x; q < K; q += (int)blockDim.x) {
    const int x = val[q];
    if (x > m) {
      m = x;
      c = 1;
    } else if (x == m) {
      c++;
    }
  }
  wave_max_count(m, c);
  const int lane = lane_id(), w = (int)(threadIdx.x / WAVE);
  if (lane == 0) {
    s_m[w] = m;
    s_c[w] = c;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int bm = INT_MIN, bc = 0;
    for (int i = 0; i < (int)(blockDim.x / WAVE); i++) {
      if (s_m[i] > bm) {
        bm = s_m[i];
        bc = s_c[i];
      } else if (s_m[i] == bm) {
        bc += s_c[i];
      }
    }
    mstat[0] = bm;
    mstat[1] = bc;
  }
}

// Accepted moves of the round in visiting order: slot_mi[slot] = move index
// (or -1), mv[i] = (vertex, own part, new part, first event slot).
constexpr int COMPACT_RUN = 16;  // consecutive slots per thread per step of k_round_compact

__global__ void __launch_bounds__(1024) k_round_compact(const int* __restrict__ slot_dec, const int* __restrict__ slot_own,
                                                      const int* __restrict__ order,
                                                      const unsigned* __restrict__ vtx_ptr, int r0, int nslots,
                                                      int* __restrict__ slot_mi, int4* __restrict__ mv,
                                                      int* __restrict__ mstat) {
  // each thread owns a run of COMPACT_RUN consecutive slots, so one step of
  // the block covers 16K slots and the block-wide scans are few
  __shared__ int s_w[16];
  int base = 0, ebase = 0;
  const int step = (int)blockDim.x * COMPACT_RUN;
  for (int s0 = 0; s0 < nslots; s0 += step) {
    const int first = s0 + (int)threadIdx.x * COMPACT_RUN;
    int f = 0, deg = 0;
    for (int j = 0; j < COMPACT_RUN; j++) {
      const int s = first + j;
      if (s < nslots && slot_dec[s] >= 0) {
        const int v = order ? order[r0 + s] : r0 + s;
        f++;
        deg += (int)(vtx_ptr[v + 1] - vtx_ptr[v]);
      }
    }
    int tot, totd;
    int ex = block_excl_scan(f, s_w, &tot);
    int exd = block_excl_scan(deg, s_w, &totd);
    for (int j = 0; j < COMPACT_RUN; j++) {
      const int s = first + j;
      if (s >= nslots) break;
      const int d = slot_dec[s];
      if (d >= 0) {
        const int v = order ? order[r0 + s] : r0 + s;
        slot_mi[s] = base + ex;
        mv[base + ex] = make_int4(v, slot_own[s], d, ebase + exd);
        ex++;
        exd += (int)(vtx_ptr[v + 1] - vtx_ptr[v]);
      } else {
        slot_mi[s] = -1;
      }
    }
    base += tot;
    ebase += totd;
  }
  if (threadIdx.x == 0) {
    mstat[2] = base;
    mstat[3] = base;
  }
}

// One move's step of the max-set bookkeeping, run by a single wavefront.
// val[] (LDS) already holds the move's changes; delta[p] holds the change of
// val[p] for every part p that item(it), it in [0, nit), names (duplicates
// allowed, -1 = no part); claim[] is all INT_MAX on entry. Updates the max M
// and its multiplicity cntM, clears delta/claim of the items, and returns true
// iff the new max set holds a part the previous one did not. A move that
// leaves every affected part off the max and below it changes nothing (the
// common case, decided with one read per item). Otherwise: the max rose (the
// new set is the affected parts at the new value); it stayed (the untouched
// parts at M plus affected parts landing on it); or it fell (full reduction).
template <class Item>
__device__ bool maxset_step(int* val, int* delta, int* claim, int K, int nit, Item item, int& M, int& cntM) {
  const int lane = lane_id();
  bool touch = false;
  for (int b = 0; b < nit; b += WAVE) {
    const int p = item(b + lane);
    if (p >= 0) {
      const int nw = val[p], old = nw - delta[p];
      touch = touch || old == M || nw >= M;
    }
  }
  if (__ballot(touch) == 0) {
    wave_lds_sync();
    for (int b = 0; b < nit; b += WAVE) {
      const int p = item(b + lane);
      if (p >= 0) delta[p] = 0;
    }
    wave_lds_sync();
    return false;
  }
  for (int b = 0; b < nit; b += WAVE) {
    const int p = item(b + lane);
    if (p >= 0) atomicMin(&claim[p], b + lane);
  }
  wave_lds_sync();
  // one representative item per affected part: old and new values
  int up = INT_MIN, inS = 0, newM = 0;
  bool enterM = false;
  for (int b = 0; b < nit; b += WAVE) {
    const int it = b + lane;
    const int p = item(it);
    if (p >= 0 && claim[p] == it) {
      const int nw = val[p], old = nw - delta[p];
      up = max(up, nw);
      inS += old == M;
      newM += nw == M;
      enterM = enterM || (nw == M && old != M);
    }
  }
  {
    int c0 = 0;
    wave_max_count(up, c0);
  }
  inS = wave_sum(inS);
  newM = wave_sum(newM);
  enterM = __ballot(enterM) != 0;
  bool stop;
  int nM, nC;
  if (up > M) {
    int cu = 0;
    bool enter = false;
    for (int b = 0; b < nit; b += WAVE) {
      const int it = b + lane;
      const int p = item(it);
      if (p >= 0 && claim[p] == it && val[p] == up) {
        cu++;
        enter = enter || (val[p] - delta[p] != M);
      }
    }
    stop = __ballot(enter) != 0;
    nM = up;
    nC = wave_sum(cu);
  } else if (cntM - inS > 0 || newM > 0) {
    stop = enterM;
    nM = M;
    nC = cntM - inS + newM;
  } else {
    int mm = INT_MIN, cc = 0;
    for (int q = lane; q < K; q += WAVE) {
      const int x = val[q];
      if (x > mm) {
        mm = x;
        cc = 1;
      } else if (x == mm) {
        cc++;
      }
    }
    wave_max_count(mm, cc);
    int stay = 0;
    for (int b = 0; b < nit; b += WAVE) {
      const int it = b + lane;
      const int p = item(it);
      if (p >= 0 && claim[p] == it && val[p] == mm && val[p] - delta[p] == M) stay++;
    }
    stay = wave_sum(stay);
    stop = cc > stay;
    nM = mm;
    nC = cc;
  }
  M = nM;
  cntM = nC;
  wave_lds_sync();
  for (int b = 0; b < nit; b += WAVE) {
    const int p = item(b + lane);
    if (p >= 0) {
      delta[p] = 0;
      claim[p] = INT_MAX;
    }
  }
  wave_lds_sync();
  return stop;
}

}  // namespace
}  // namespace fpe
