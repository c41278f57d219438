// fp_internal.h -- shared internals of the HIP engine (not part of the C-ABI).
//
// Device layout of one context (all in HBM, allocated once per upload):
//   net_ptr[noe+1], net_pins[pins]   net->pins CSR (reader order)
//   vtx_ptr[nov+1], vtx_nets[pins]   vertex->nets CSR (ascending nets)
//   apply_pins[pins]                 per net, the ordering ranks of its pins, ascending
//   src[noe], vwgt[nov], order[nov]  net sources, vertex weights (nullable), visiting order (nullable)
//   part[nov]                        part of each vertex
//   ent[pins + noe + 64]             net n's segment starts at seg(n) = net_ptr[n] + n:
//                                    ent[seg] = lambda(n), then its connectivity list,
//                                    one u32 per entry: (pin count << pb) | part,
//                                    pb = ceil(log2 K); 64 words of tail padding let a
//                                    wavefront read a whole 64-word window of any segment
//   ev[pins + noe + 64] (u16)        evaluation view of the same segments: ev[seg] =
//                                    lambda(n), then per entry part | (count == 1) << 15;
//                                    a 64-lane window of it is one 128-byte line
//   vtx_seg[pins]                    per (vertex, net) incidence: seg(net), so one random
//                                    read brings a net's lambda and first 63 entries
//   sizes[K] (int64)                 part weights
#pragma once
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <vector>

#include "../../include/fastpart.h"

namespace fpe {

constexpr int WAVE = 64;
constexpr int LDS_BUDGET = 64 * 1024;
constexpr int LDS_PER_CU = 160 * 1024;  // gfx950: LDS of a compute unit, shared by its workgroups
constexpr int MAX_WAVES_PER_CU = 32;    // 8 wavefronts per SIMD
constexpr unsigned NOFIRST = 0xFFFFFFFFu;
constexpr int ENT_PAD = 64;  // tail padding of the segment arrays
constexpr unsigned EV_SINGLE = 0x8000u;  // evaluation view: the part has one pin in the net

// evaluation-view word of a packed (count << pb | part) list entry
__host__ __device__ __forceinline__ unsigned short ev_word(unsigned x, int pb, unsigned pmask) {
  return (unsigned short)((x & pmask) | ((x >> pb) == 1u ? EV_SINGLE : 0u));
}

// A part size as an unsigned key that orders like the signed size: the
// reference compares int part sizes (TV/partition.cpp:108), and its round
// schedule's source-only size update (TV/partition.cpp:827-829) can drive a
// size below zero. Sizes outside the int range, which the reference's own
// counters cannot hold, saturate.
__host__ __device__ __forceinline__ unsigned size_key(long long sz) {
  const long long c = sz < (long long)INT_MIN ? (long long)INT_MIN : (sz > (long long)INT_MAX ? (long long)INT_MAX : sz);
  return (unsigned)(int)c ^ 0x80000000u;
}

// ---- wave helpers ---------------------------------------------------------
__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & (WAVE - 1)); }
// index of the wavefront in its block, as a wave-uniform (scalar) value
__device__ __forceinline__ int wave_id() { return __builtin_amdgcn_readfirstlane((int)(threadIdx.x / WAVE)); }

// orders the LDS traffic of one wavefront: everything before is visible to
// every lane of the wave after. The LDS performs one wavefront's operations
// in issue order, so wavefront-scope fences need no wait: this only keeps the
// compiler from moving LDS accesses across it (a workgroup-scope fence would
// wait for every outstanding LDS access here). Not for data shared between
// the waves of a workgroup: that needs __syncthreads().
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// A load at a wave-uniform address, through the scalar cache: for arrays the
// kernel only reads (the scalar cache starts every launch empty).
template <typename T>
__device__ __forceinline__ T uniform_load(const T* p) {
  return *(const __attribute__((address_space(4))) T*)p;
}

template <typename T>
__device__ __forceinline__ T wave_sum(T x) {
  for (int off = WAVE / 2; off > 0; off >>= 1) x += __shfl_xor(x, off);
  return x;
}

// 32-bit lane exchange by a DPP pattern (no LDS traffic); lanes whose row is
// not in ROW_MASK, or whose source lane lies outside the row, receive `old`.
template <int CTRL, int ROW_MASK = 0xF>
__device__ __forceinline__ int dpp_i32(int old, int v) {
  return __builtin_amdgcn_update_dpp(old, v, CTRL, ROW_MASK, 0xF, false);
}

// Wave-wide sum and max of an int through DPP: quad_perm [1,0,3,2] and
// [2,3,0,1], row_ror 4 and 8 (row-wide), row_bcast15 into rows 1 and 3,
// row_bcast31 into rows 2 and 3; lane 63 then holds the result, which is
// returned to every lane. The value a step moves into lanes it does not
// reach is the operation's identity (0, INT_MIN), which lets the compiler
// fold each step into one DPP-modified add / max.
__device__ __forceinline__ int wave_sum(int x) {
  x += dpp_i32<0xB1>(0, x);
  x += dpp_i32<0x4E>(0, x);
  x += dpp_i32<0x124>(0, x);
  x += dpp_i32<0x128>(0, x);
  x += dpp_i32<0x142, 0xA>(0, x);
  x += dpp_i32<0x143, 0xC>(0, x);
  return __builtin_amdgcn_readlane(x, WAVE - 1);
}

__device__ __forceinline__ int wave_max(int x) {
  x = max(x, dpp_i32<0xB1>(INT_MIN, x));
  x = max(x, dpp_i32<0x4E>(INT_MIN, x));
  x = max(x, dpp_i32<0x124>(INT_MIN, x));
  x = max(x, dpp_i32<0x128>(INT_MIN, x));
  x = max(x, dpp_i32<0x142, 0xA>(INT_MIN, x));
  x = max(x, dpp_i32<0x143, 0xC>(INT_MIN, x));
  return __builtin_amdgcn_readlane(x, WAVE - 1);
}

// The same for unsigned max (identity 0) and min (identity ~0).
__device__ __forceinline__ unsigned wave_max_u32(unsigned x) {
  x = max(x, (unsigned)dpp_i32<0xB1>(0, (int)x));
  x = max(x, (unsigned)dpp_i32<0x4E>(0, (int)x));
  x = max(x, (unsigned)dpp_i32<0x124>(0, (int)x));
  x = max(x, (unsigned)dpp_i32<0x128>(0, (int)x));
  x = max(x, (unsigned)dpp_i32<0x142, 0xA>(0, (int)x));
  x = max(x, (unsigned)dpp_i32<0x143, 0xC>(0, (int)x));
  return (unsigned)__builtin_amdgcn_readlane((int)x, WAVE - 1);
}

__device__ __forceinline__ unsigned wave_min_u32(unsigned x) {
  x = min(x, (unsigned)dpp_i32<0xB1>(-1, (int)x));
  x = min(x, (unsigned)dpp_i32<0x4E>(-1, (int)x));
  x = min(x, (unsigned)dpp_i32<0x124>(-1, (int)x));
  x = min(x, (unsigned)dpp_i32<0x128>(-1, (int)x));
  x = min(x, (unsigned)dpp_i32<0x142, 0xA>(-1, (int)x));
  x = min(x, (unsigned)dpp_i32<0x143, 0xC>(-1, (int)x));
  return (unsigned)__builtin_amdgcn_readlane((int)x, WAVE - 1);
}

// Wave-wide AND of an unsigned (identity ~0), and of a 64-bit word as two
// 32-bit reductions.
__device__ __forceinline__ unsigned wave_and_u32(unsigned x) {
  x &= (unsigned)dpp_i32<0xB1>(-1, (int)x);
  x &= (unsigned)dpp_i32<0x4E>(-1, (int)x);
  x &= (unsigned)dpp_i32<0x124>(-1, (int)x);
  x &= (unsigned)dpp_i32<0x128>(-1, (int)x);
  x &= (unsigned)dpp_i32<0x142, 0xA>(-1, (int)x);
  x &= (unsigned)dpp_i32<0x143, 0xC>(-1, (int)x);
  return (unsigned)__builtin_amdgcn_readlane((int)x, WAVE - 1);
}

__device__ __forceinline__ unsigned long long wave_and_u64(unsigned long long x) {
  return ((unsigned long long)wave_and_u32((unsigned)(x >> 32)) << 32) | wave_and_u32((unsigned)x);
}

// lane l's 64-bit value, to every lane
__device__ __forceinline__ unsigned long long readlane_u64(unsigned long long v, int l) {
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)v, l);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(v >> 32), l);
  return ((unsigned long long)hi << 32) | lo;
}

// Inclusive prefix sum over the wave through DPP: row_shr 1, 2, 4, 8 (lanes
// whose source is outside the row add 0), then row_bcast15 into rows 1 and 3
// and row_bcast31 into rows 2 and 3 carry the row totals.
__device__ __forceinline__ int wave_incl_scan(int x) {
  x += dpp_i32<0x111>(0, x);
  x += dpp_i32<0x112>(0, x);
  x += dpp_i32<0x114>(0, x);
  x += dpp_i32<0x118>(0, x);
  x += dpp_i32<0x142, 0xA>(0, x);
  x += dpp_i32<0x143, 0xC>(0, x);
  return x;
}

template <typename T>
__device__ __forceinline__ T wave_incl_scan(T x) {
  const int lane = lane_id();
  for (int off = 1; off < WAVE; off <<= 1) {
    const T y = __shfl_up(x, off);
    if (lane >= off) x += y;
  }
  return x;
}

// Exclusive scan over a whole block (blockDim.x multiple of 64, <= 1024).
// `scratch` holds blockDim.x/64 values. Returns the exclusive prefix; *total
// receives the block total. Must be called by every thread of the block.
template <typename T>
__device__ __forceinline__ T block_excl_scan(T x, T* scratch, T* total) {
  const int lane = lane_id(), wave = (int)(threadIdx.x / WAVE), nw = (int)(blockDim.x / WAVE);
  const T incl = wave_incl_scan(x);
  if (lane == WAVE - 1) scratch[wave] = incl;
  __syncthreads();
  T base = 0, tot = 0;
  for (int w = 0; w < nw; w++) {
    const T s = scratch[w];
    if (w < wave) base += s;
    tot += s;
  }
  __syncthreads();
  *total = tot;
  return base + incl - x;
}

// Wave-wide max / min of a 64-bit key: the high words first, then the low
// words of the lanes holding the extreme high word (two 32-bit reductions).
__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long x) {
  const unsigned hi = (unsigned)(x >> 32), lo = (unsigned)x;
  const unsigned mh = wave_max_u32(hi);
  const unsigned ml = wave_max_u32(hi == mh ? lo : 0u);
  return ((unsigned long long)mh << 32) | ml;
}

__device__ __forceinline__ unsigned long long wave_min_u64(unsigned long long x) {
  const unsigned hi = (unsigned)(x >> 32), lo = (unsigned)x;
  const unsigned mh = wave_min_u32(hi);
  const unsigned ml = wave_min_u32(hi == mh ? lo : ~0u);
  return ((unsigned long long)mh << 32) | ml;
}

// first index k in [lo, hi) with arr[k] >= key (arr ascending); wave-uniform result
__device__ __forceinline__ unsigned wave_lower_bound(const int* __restrict__ arr, unsigned lo, unsigned hi, int key) {
  const int lane = lane_id();
  while (hi - lo > WAVE) {
    const unsigned step = (hi - lo + WAVE - 1) / WAVE;
    const unsigned k = lo + (unsigned)lane * step;
    const bool lt = k < hi && arr[k] < key;
    const int c = __popcll(__ballot(lt));
    if (c == 0) return lo;
    const unsigned nlo = lo + (unsigned)(c - 1) * step + 1;
    const unsigned nhi = min(hi, lo + (unsigned)c * step);
    lo = nlo;
    hi = nhi;
  }
  const unsigned k = lo + (unsigned)lane;
  const bool lt = k < hi && arr[k] < key;
  return lo + (unsigned)__popcll(__ballot(lt));
}

// index of the entry of part p in list[0..L) or -1; wave-uniform
__device__ __forceinline__ int wave_find(const unsigned* list, int L, int p, unsigned pmask) {
  const int lane = lane_id();
  for (int t0 = 0; t0 < L; t0 += WAVE) {
    const int t = t0 + lane;
    const bool hit = t < L && (int)(list[t] & pmask) == p;
    const unsigned long long m = __ballot(hit);
    if (m) return t0 + __ffsll((long long)m) - 1;
  }
  return -1;
}

// One move's edit of one net's connectivity list held in LDS (one wavefront):
// remove one pin of part p (decrement, or swap the last entry into its slot
// when its count reaches 0) then add one pin of part q (increment, or append);
// TV/partition.cpp:388-435. L is the list length (wave-uniform).
__device__ __forceinline__ void wave_list_move(unsigned* list, int& L, int p, int q, int pb, unsigned pmask) {
  const int lane = lane_id();
  const unsigned inc = 1u << pb;
  int j = wave_find(list, L, p, pmask);
  if (j >= 0) {
    const unsigned x = list[j];
    if ((x >> pb) > 1u) {
      if (lane == 0) list[j] = x - inc;
    } else {
      if (lane == 0) list[j] = list[L - 1];
      L--;
    }
    wave_lds_sync();
  }
  j = wave_find(list, L, q, pmask);
  if (j >= 0) {
    if (lane == 0) list[j] += inc;
  } else {
    if (lane == 0) list[L] = inc | (unsigned)q;
    L++;
  }
  wave_lds_sync();
}

// A net's list in LDS with a part -> position index (one wavefront): at[p]
// is the position of part p's entry, or LIST_NONE. Between nets every slot of
// at holds LIST_NONE: wave_index_set fills the slots of a loaded list and
// wave_index_clear empties those of the final one.
constexpr unsigned short LIST_NONE = 0xFFFF;

__device__ __forceinline__ void wave_index_init(unsigned short* at, int K) {
  for (int q = lane_id(); q < K; q += WAVE) at[q] = LIST_NONE;
  wave_lds_sync();
}

__device__ __forceinline__ void wave_index_set(const unsigned* list, int L, unsigned short* at, unsigned pmask) {
  for (int t = lane_id(); t < L; t += WAVE) at[list[t] & pmask] = (unsigned short)t;
  wave_lds_sync();
}

__device__ __forceinline__ void wave_index_clear(const unsigned* list, int L, unsigned short* at, unsigned pmask) {
  for (int t = lane_id(); t < L; t += WAVE) at[list[t] & pmask] = LIST_NONE;
  wave_lds_sync();
}

// wave_list_move through the index (same edits, same list order): two
// dependent LDS reads per move instead of scans of the list. Returns in co
// the pin count of p's entry before the move (0 without one) and in found
// whether q had an entry.
__device__ __forceinline__ void wave_ilist_move(unsigned* list, unsigned short* at, int& L, int p, int q, int pb,
                                                unsigned pmask, unsigned& co, bool& found) {
  const int lane = lane_id();
  const unsigned inc = 1u << pb;
  const int jp = p < 0 ? (int)LIST_NONE : __builtin_amdgcn_readfirstlane((int)at[p]);  // p -1: no part yet
  int jq = __builtin_amdgcn_readfirstlane((int)at[q]);
  const bool hasp = jp != (int)LIST_NONE;
  const unsigned x = (unsigned)__builtin_amdgcn_readfirstlane((int)list[hasp ? jp : 0]);
  const unsigned y = (unsigned)__builtin_amdgcn_readfirstlane((int)list[L > 0 ? L - 1 : 0]);
  co = hasp ? x >> pb : 0u;
  found = jq != (int)LIST_NONE;
  bool qin = found;  // q has an entry after the removal
  if (co > 1u) {
    if (lane == 0) list[jp] = x - inc;
  } else if (hasp) {
    // the last entry takes p's slot
    if (lane == 0) {
      list[jp] = y;
      at[y & pmask] = (unsigned short)jp;
      at[p] = LIST_NONE;
    }
    L--;
    if (q == p) qin = false;       // q's entry was the one removed
    else if (jq == L) jq = jp;     // q's entry was the last one: it moved into p's slot
  }
  if (qin) {
    if (lane == 0) atomicAdd(&list[jq], inc);
  } else {
    if (lane == 0) {
      list[L] = inc | (unsigned)q;
      at[q] = (unsigned short)L;
    }
    L++;
  }
  wave_lds_sync();
}

// First 64 words of net n's segment (seg = net_ptr[n] + n): the segment holds
// the length and then np >= length entries, so one load covers the length and
// the first 63 entries.
__device__ __forceinline__ unsigned wave_list_head(const unsigned* ent, unsigned seg, unsigned np) {
  const int lane = lane_id();
  return (unsigned)lane <= np ? ent[seg + (unsigned)lane] : 0u;
}

// Stages a list into LDS from its head x (wave_list_head) and the rest of its
// segment; returns lambda.
__device__ __forceinline__ int wave_list_stage(const unsigned* ent, unsigned seg, unsigned x, unsigned* list) {
  const int lane = lane_id();
  const int L = __builtin_amdgcn_readfirstlane((int)x);
  if (lane >= 1 && lane <= L) list[lane - 1] = x;
  for (int t = WAVE - 1 + lane; t < L; t += WAVE) list[t] = ent[seg + 1 + (unsigned)t];
  wave_lds_sync();
  return L;
}

// Stages net n's list (segment at seg = net_ptr[n] + n) into LDS; returns lambda.
__device__ __forceinline__ int wave_list_load(const unsigned* ent, unsigned seg, unsigned np, unsigned* list) {
  return wave_list_stage(ent, seg, wave_list_head(ent, seg, np), list);
}

// Per-lane lower bound: first k in [lo, hi) with arr[k] >= key (arr ascending).
__device__ __forceinline__ unsigned lane_lower_bound(const unsigned* __restrict__ arr, unsigned lo, unsigned hi,
                                                     unsigned key) {
  while (lo < hi) {
    const unsigned mid = (lo + hi) >> 1;
    if (arr[mid] < key) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// Sweeps the 16-bit lists (evaluation view) of nn <= 64 nets whose segment
// offsets lane j of `sg` holds: NIF segment heads are loaded together (a
// 64-word window brings lambda in lane 0 and the first 63 entries). The
// further windows of the lists longer than 63 entries follow as one task list
// over all those nets (numbered net by net by a scan of their window counts),
// LT windows loaded together, with the lambdas of the first windows. For every
// window f(j, L, t, ok, x) runs on all lanes: j = net index (wave-uniform), L =
// its lambda, t = this lane's list position, ok = t < L, x = the entry. A
// window may read up to 63 words past a segment: the view has ENT_PAD words of
// tail padding.
#ifndef FP_SWEEP_LONG_TASKS
#define FP_SWEEP_LONG_TASKS 2
#endif
constexpr int SWEEP_LONG_TASKS = FP_SWEEP_LONG_TASKS;  // long-list windows in flight in the sweeps

template <int NIF, int LT, class F>
__device__ __forceinline__ void wave_sweep_lists(const unsigned short* __restrict__ ev, unsigned sg, int nn, F f) {
  const int lane = lane_id();
  int lmax = 0;  // longest list
  int lj = 0;    // lane j: lambda of net j
  for (int j0 = 0; j0 < nn; j0 += NIF) {
    unsigned xv[NIF];
#pragma unroll
    for (int u = 0; u < NIF; u++) {
      const unsigned sj = (unsigned)__builtin_amdgcn_readlane((int)sg, min(j0 + u, nn - 1));
      xv[u] = ev[sj + (unsigned)lane];
    }
#pragma unroll
    for (int u = 0; u < NIF; u++) {
      const int j = j0 + u;
      if (j < nn) {
        const int Lj = __builtin_amdgcn_readlane((int)xv[u], 0);
        lmax = max(lmax, Lj);
        lj = lane == j ? Lj : lj;
        f(j, Lj, lane - 1, lane >= 1 && lane - 1 < Lj, xv[u]);
      }
    }
  }
  if (lmax > WAVE - 1) {
    const int nwin = lj >> 6;  // ceil((lambda - 63) / 64) windows past the first
    const int incl = wave_incl_scan(nwin);
    const int ntask = __builtin_amdgcn_readlane(incl, WAVE - 1);
    for (int k0 = 0; k0 < ntask; k0 += LT) {
      unsigned xw[LT];
      int tj[LT], tw[LT], tl[LT];
#pragma unroll
      for (int u = 0; u < LT; u++) {
        const int k = min(k0 + u, ntask - 1);
        const int j = (int)__popcll(__ballot(incl <= k));  // its net: incl is non-decreasing
        const int w = k - __builtin_amdgcn_readlane(incl - nwin, j);
        const unsigned sj = (unsigned)__builtin_amdgcn_readlane((int)sg, j);
        const int Lj = __builtin_amdgcn_readlane(lj, j);
        tj[u] = j;
        tw[u] = WAVE - 1 + w * WAVE;
        tl[u] = Lj;
        const int t = tw[u] + lane;
        xw[u] = ev[sj + 1u + (unsigned)(t < Lj ? t : 0)];
      }
#pragma unroll
      for (int u = 0; u < LT; u++) {
        const int t = tw[u] + lane;
        if (k0 + u < ntask) f(tj[u], tl[u], t, t < tl[u], xw[u]);
      }
    }
  }
}

// Adds a vertex's nets vn[0..deg) to the round's touched-net list, each net
// once per round (net_stamp holds the round tag of its last addition), with
// one counter atomic per 64 nets instead of one per net. Called by the whole
// wavefront.
__device__ __forceinline__ void wave_touch_nets(const unsigned* __restrict__ vn, int deg, unsigned tag,
                                                unsigned* __restrict__ net_stamp, int* __restrict__ touched,
                                                int* __restrict__ tcount) {
  const int lane = lane_id();
  for (int i0 = 0; i0 < deg; i0 += WAVE) {
    const int i = i0 + lane;
    unsigned n = 0;
    bool isnew = false;
    if (i < deg) {
      n = vn[i];
      isnew = atomicExch(&net_stamp[n], tag) != tag;
    }
    const unsigned long long b = __ballot(isnew);
    if (b) {
      int base = 0;
      if (lane == 0) base = atomicAdd(tcount, (int)__popcll(b));
      base = __builtin_amdgcn_readfirstlane(base);
      if (isnew) touched[base + (int)__popcll(b & ((1ull << lane) - 1ull))] = (int)n;
    }
  }
}

// A wave's newly touched nets, staged in its own LDS buffer of TOUCH_BUF
// entries and appended to the round's touched list a buffer at a time: one
// atomic on the list's counter per flush instead of one per vertex (atomics
// from every wave on that one address queue at one L2 slice). The wave must
// flush before it exits.
constexpr int TOUCH_BUF = 256;

struct WaveTouch {
  unsigned* buf;  // this wave's LDS buffer
  int n;          // staged nets (wave-uniform)
};

__device__ __forceinline__ void wave_touch_flush(WaveTouch& t, int* __restrict__ touched, int* __restrict__ tcount) {
  if (t.n == 0) return;
  const int lane = lane_id();
  int base = 0;
  if (lane == 0) base = atomicAdd(tcount, t.n);
  base = __builtin_amdgcn_readfirstlane(base);
  wave_lds_sync();
  for (int i = lane; i < t.n; i += WAVE) touched[base + i] = (int)t.buf[i];
  wave_lds_sync();
  t.n = 0;
}

// Appends the lanes' items (where take) to a list through the wave's LDS
// buffer (a WaveTouch of TOUCH_BUF entries): one atomic on the list's counter
// per full buffer instead of one per call. The order of the list is not kept.
__device__ __forceinline__ void wave_stage_append(bool take, unsigned item, WaveTouch& t, int* __restrict__ list,
                                                  int* __restrict__ count) {
  const unsigned long long b = __ballot(take);
  const int c = (int)__popcll(b);
  if (c == 0) return;
  if (t.n + c > TOUCH_BUF) wave_touch_flush(t, list, count);
  if (take) t.buf[t.n + (int)__popcll(b & ((1ull << lane_id()) - 1ull))] = item;
  t.n += c;
}

// The last flush of a kernel, for every wave of the workgroup together: one
// atomic on the list's counter per workgroup. Every thread of the workgroup
// must call it (it synchronizes the workgroup); s_tn holds waves + 1 ints.
__device__ __forceinline__ void block_touch_flush(WaveTouch& t, int* __restrict__ touched, int* __restrict__ tcount,
                                                  int* s_tn) {
  const int lane = lane_id(), wave = wave_id(), nw = (int)(blockDim.x / WAVE);
  if (lane == 0) s_tn[wave] = t.n;
  __syncthreads();
  if (threadIdx.x == 0) {
    int tot = 0;
    for (int w = 0; w < nw; w++) {
      const int c = s_tn[w];
      s_tn[w] = tot;
      tot += c;
    }
    s_tn[nw] = tot ? atomicAdd(tcount, tot) : 0;
  }
  __syncthreads();
  const int base = s_tn[nw] + s_tn[wave];
  for (int i = lane; i < t.n; i += WAVE) touched[base + i] = (int)t.buf[i];
  t.n = 0;
}

// wave_touch_nets through the wave's buffer
__device__ __forceinline__ void wave_touch_nets_staged(const unsigned* __restrict__ vn, int deg, unsigned tag,
                                                       unsigned* __restrict__ net_stamp, WaveTouch& t,
                                                       int* __restrict__ touched, int* __restrict__ tcount) {
  const int lane = lane_id();
  for (int i0 = 0; i0 < deg; i0 += WAVE) {
    const int i = i0 + lane;
    unsigned n = 0;
    bool isnew = false;
    if (i < deg) {
      n = vn[i];
      isnew = atomicExch(&net_stamp[n], tag) != tag;
    }
    const unsigned long long b = __ballot(isnew);
    const int c = (int)__popcll(b);
    if (c == 0) continue;
    if (t.n + c > TOUCH_BUF) wave_touch_flush(t, touched, tcount);
    if (isnew) t.buf[t.n + (int)__popcll(b & ((1ull << lane) - 1ull))] = n;
    t.n += c;
  }
}

// Range [lo, hi) of apply_pins[s..e) (a net's pin ranks, ascending) that
// may hold ranks in [r0, r1): the exact range for nets longer than a wave,
// the whole net otherwise (the caller tests each rank against [r0, r1)).
__device__ __forceinline__ void wave_rank_window(const int* apply_pins, unsigned s, unsigned e, int r0, int r1,
                                                 unsigned& lo, unsigned& hi) {
  if (e - s <= (unsigned)WAVE) {
    lo = s;
    hi = e;
    return;
  }
  lo = wave_lower_bound(apply_pins, s, e, r0);
  hi = wave_lower_bound(apply_pins, lo, e, r1);
}

// Walks a wave's share of the touched nets (entries first, first + stride,
// ... of touched): their ids and pin ranges are loaded 64 nets at a time,
// one net per lane, then f(n, s, e) runs on every lane for each net in
// order, with [s, e) the net's range in net_ptr.
template <class F>
__device__ __forceinline__ void wave_for_touched(const int* __restrict__ touched, int ntouched,
                                                 const unsigned* __restrict__ net_ptr, int first, int stride, F f) {
  const int lane = lane_id();
  for (long long b0 = first; b0 < ntouched; b0 += (long long)WAVE * stride) {
    const long long ij = b0 + (long long)lane * stride;
    int nj = 0;
    unsigned sj = 0, ej = 0;
    if (ij < ntouched) {
      nj = touched[ij];
      sj = net_ptr[nj];
      ej = net_ptr[nj + 1];
    }
    const int cnt = (int)min((long long)WAVE, ((long long)ntouched - b0 + stride - 1) / stride);
    for (int j = 0; j < cnt; j++)
      f(__builtin_amdgcn_readlane(nj, j), (unsigned)__builtin_amdgcn_readlane((int)sj, j),
        (unsigned)__builtin_amdgcn_readlane((int)ej, j));
  }
}

// Writes a list back to both the packed segment and its evaluation view.
__device__ __forceinline__ void wave_list_store(unsigned* ent, unsigned short* ev, unsigned seg, const unsigned* list,
                                                int L, int pb, unsigned pmask) {
  const int lane = lane_id();
  for (int t = lane; t < L; t += WAVE) {
    const unsigned x = list[t];
    ent[seg + 1 + (unsigned)t] = x;
    ev[seg + 1 + (unsigned)t] = ev_word(x, pb, pmask);
  }
  if (lane == 0) {
    ent[seg] = (unsigned)L;
    ev[seg] = (unsigned short)L;
  }
}

// Adds the number of the block's threads with `flag` set to *counter, with
// one global atomic per block (same-address atomics from every wave queue at
// one L2 slice). Every thread of the block calls it.
__device__ __forceinline__ void block_count_add(unsigned long long* counter, bool flag) {
  __shared__ unsigned long long s_cnt;
  if (threadIdx.x == 0) s_cnt = 0;
  __syncthreads();
  const unsigned long long b = __ballot(flag);
  if (lane_id() == 0 && b) atomicAdd(&s_cnt, (unsigned long long)__popcll(b));
  __syncthreads();
  if (threadIdx.x == 0 && s_cnt) atomicAdd(counter, s_cnt);
}

__device__ __forceinline__ void atomic_add_i64(long long* p, long long v) {
  atomicAdd(reinterpret_cast<unsigned long long*>(p), (unsigned long long)v);
}

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

// The start of every refinement pass in one launch: clears the net stamps
// (round tags restart at 1 each pass) and the touched counters; with val
// given, block 0 also writes max(val[0..K)) and the number of parts holding
// it to mstat[0..1] (the max set of MGV/partition.cpp:712-729 and
// MSM/partition.cpp:1030-1045). 1024 threads per block.
__global__ void __launch_bounds__(1024) k_pass_begin(unsigned* __restrict__ stamp, size_t n, int* __restrict__ counters,
                                                     const int* __restrict__ val, int K, int* __restrict__ mstat);

template <typename T>
__global__ void k_fill(T* p, size_t n, T v) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

}  // namespace fpe

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
struct fp_ctx {
  int device = -1;
  hipStream_t stream = nullptr;
  char err[512] = {0};
  long long dev_bytes = 0;
  std::vector<void*> owned;
  // bumped by every change that can invalidate recorded kernel arguments
  // (device pointers, K, the list layout): allocation, freeing, a new K
  unsigned long long gen = 0;

  // hypergraph
  int nov = 0, noe = 0;
  long long pins = 0;
  unsigned *net_ptr = nullptr, *net_pins = nullptr, *vtx_ptr = nullptr, *vtx_nets = nullptr;
  int *src = nullptr, *vwgt = nullptr, *order = nullptr, *apply_pins = nullptr;
  bool has_graph = false, dup_pins = false;
  bool src_on_pins = false;  // every non-empty net's source is one of its pins
  long long total_weight = 0;
  unsigned max_net = 0, max_vdeg = 0;
  unsigned long long max_src_pins = 0;  // max over vertices of the pins of the nets whose source it is
  unsigned long long sum_deg2 = 0;  // sum over vertices of degree^2 (bounds CN's per-pass edit list)
  std::vector<unsigned> h_net_ptr;
  std::vector<int> h_rank_pins;  // per net, ordering ranks of its pins ascending
  std::vector<int> h_order;      // rank -> vertex
  std::vector<int> h_vwgt;
  // pinned host staging of large uploads from pageable memory: two buffers
  // of STAGE_BYTES (allocated on the first such upload, freed with the
  // context), each with the event of the last copy out of it
  void* stage[2] = {nullptr, nullptr};
  hipEvent_t stage_ev[2] = {nullptr, nullptr};

  // state
  int K = 0, pb = 0;
  int kcap = 0;  // capacity of the K-sized arrays
  unsigned pmask = 0;
  int* part = nullptr;
  unsigned* ent = nullptr;
  unsigned short* ev = nullptr;
  unsigned* vtx_seg = nullptr;
  long long* sizes = nullptr;
  bool has_state = false;

  // round scratch
  int rmax = 0;
  int *slot_own = nullptr, *slot_dec = nullptr, *touched = nullptr, *counters = nullptr;
  int* flags = nullptr;  // [0] CN edit list overflow: the size it needed (0 = none), [1] a repeated
                         // pin's net over CN_MAXOPS ops, [2] CN passes applied (counted by the commit)
  unsigned* net_stamp = nullptr;  // round tag of the last round that touched the net (cleared every pass)
  long long *cand_tot = nullptr, *sizes_frozen = nullptr, *pref = nullptr;
  int* hist = nullptr;
  size_t hist_cap = 0;
  int *over_idx = nullptr, *over_count = nullptr;
  unsigned long long* moves = nullptr;
  int2* compact_tiles = nullptr;  // per-tile move / event totals of a round's compaction
  int compact_cap = 0;

  // CN refinement scratch (allocated on the first CN refinement of an upload)
  struct {
    bool ready = false;
    int *old_part = nullptr, *new_part = nullptr, *npush = nullptr;
    long long *pbase = nullptr, *sums = nullptr, *opoff = nullptr;
    int *mv_v = nullptr, *mv_net = nullptr, *mv_from = nullptr, *mv_last = nullptr;
    int *opcnt = nullptr, *opfill = nullptr, *ops = nullptr;
    long long ops_cap = 0;
    unsigned long long* totals = nullptr;  // [0] ops needed [1] ops placed [2] max per net [3] pushes [4] big nets
    unsigned* bigbits = nullptr;           // bitmap over a pass's pushes (k_cn_bigsort)
    int* big = nullptr;                    // nets with more ops than an LDS sort takes
    long long big_cap = 0;
  } cn;

  // MSV / MSRV refinement scratch (allocated on first use)
  struct {
    int* vol = nullptr;      // K: send (+ receive) volume per part
    int* mstat = nullptr;    // [0] max, [1] parts at the max, [2] round moves, [3] moves applied
    int* slot_mi = nullptr;  // rmax: move index of a slot or -1
    int4* mv = nullptr;      // rmax moves: vertex, own part, new part, event offset
    int4* evbuf = nullptr;   // per (move, incident net): source part, volume changes
    int rmax = 0;
    int kcap = 0;
    long long evcap = 0;
  } vol;

  // TM / MSM refinement scratch (allocated on first use)
  struct {
    int* pc = nullptr;                 // K*K: nets with source in part a and pins in part b
    int* msg = nullptr;                // K: parts each part sends to
    unsigned long long* B = nullptr;   // K rows of ceil(K/64) words: bit (a, b) = pc[a][b] > 0
    unsigned long long* BT = nullptr;  // transpose of B
    int* mstat = nullptr;              // as vol.mstat, over msg
    int* slot_mi = nullptr;
    int4* mv = nullptr;
    int2* evbuf = nullptr;             // MSM: per incidence, pc changes (cell, +-1)
    unsigned* incoff = nullptr;        // MSM: first event slot of each incidence (pins + 1)
    int kcap = 0, rmax = 0;
    long long evcap = 0;
    bool incoff_ready = false;
  } conn;

  // metric scratch (grown on demand by fp_eval)
  unsigned* met_scratch = nullptr;  // work counter, message bitmap, per-block partial sums
  size_t met_words = 0;
  long long* met_out = nullptr;     // 6 scalars + 5 K-vectors
  size_t met_out_cap = 0;
  int* lam = nullptr;               // noe: per-net lambda of the last fp_eval that asked for it

  // timing
  bool kernel_timing = false;
  long long timed_launches = 0;
  double timed_ms = 0;
  std::vector<hipEvent_t> ev_pool;
  size_t timed_next = 0;  // next pooled event pair for per-launch timing (events 0, 1 bracket a refine)
};

namespace fpe {

fp_status fail(fp_ctx* c, fp_status st, const char* fmt, ...);
// A prepared refinement: what every pass of one objective on one context
// needs, checked and allocated once. The launch shape (waves per evaluation
// block, evaluation grid cap) does not change results; 0 = the default.
struct RefinePlan {
  fp_metric metric = FP_METRIC_TV;
  fp_schedule sched{};
  int R = 0;
  int eval_waves = 0;
  int eval_grid = 0;
};

// plan_*: checks, scratch, and the per-refinement initialisation kernels
// (volumes, connections). pass_*: enqueues one pass on ctx->stream with no
// host synchronisation and no host-visible state change, so a pass can be
// captured into a HIP graph and replayed (every pass resets its own round
// counters and net stamps on the device).
fp_status plan_tv(fp_ctx* ctx, RefinePlan& p);
fp_status pass_tv(fp_ctx* ctx, const RefinePlan& p);
fp_status plan_cn(fp_ctx* ctx, RefinePlan& p);
fp_status pass_cn(fp_ctx* ctx, const RefinePlan& p);
fp_status plan_vol(fp_ctx* ctx, RefinePlan& p);
fp_status pass_vol(fp_ctx* ctx, const RefinePlan& p);
fp_status plan_conn(fp_ctx* ctx, RefinePlan& p);
fp_status pass_conn(fp_ctx* ctx, const RefinePlan& p);
// dispatch on p.metric
fp_status plan_refine(fp_ctx* ctx, fp_metric metric, const fp_schedule* sched, RefinePlan& p);
fp_status record_pass(fp_ctx* ctx, const RefinePlan& p);
// CN only: after a pass, grows the edit list if the pass overflowed it (the
// pass then left the state unchanged); *redo = 1 when it must run again
fp_status cn_after_pass(fp_ctx* ctx, int* redo);
// CN only: `passes` passes enqueued back to back with one host check at the
// end. A pass that overflows the edit list leaves the state unchanged, and so
// does every later pass of the burst (the same state needs the same list);
// the commit counts the applied passes, so the rest run again after the list
// has grown.
fp_status cn_passes(fp_ctx* ctx, const RefinePlan& p, int passes);
fp_status dalloc_bytes(fp_ctx* ctx, void** p, size_t bytes);
void dfree_all(fp_ctx* ctx);
hipEvent_t pooled_event(fp_ctx* ctx, size_t i);

template <typename T>
fp_status dalloc(fp_ctx* ctx, T** p, size_t count) {
  void* q = nullptr;
  fp_status s = dalloc_bytes(ctx, &q, (count ? count : 1) * sizeof(T));
  *p = static_cast<T*>(q);
  return s;
}

#define FP_HIP_OK(call)                                                                              \
  do {                                                                                               \
    hipError_t e_ = (call);                                                                          \
    if (e_ != hipSuccess) return ::fpe::fail(ctx, FP_ERR_HIP, "%s: %s", #call, hipGetErrorString(e_)); \
  } while (0)

#define FP_TRY(x)               \
  do {                          \
    fp_status s_ = (x);         \
    if (s_ != FP_OK) return s_; \
  } while (0)

// Host to device on ctx->stream, returning when the host array may be
// reused. Pinned host arrays are copied directly; large pageable ones go
// through the context's pinned staging buffers, filled by all host threads
// while the previous piece is in flight.
fp_status h2d_bulk(fp_ctx* ctx, void* d, const void* h, size_t bytes);

template <typename T>
fp_status h2d(fp_ctx* ctx, T* d, const T* h, size_t count) {
  return h2d_bulk(ctx, d, h, count * sizeof(T));
}
template <typename T>
fp_status d2h(fp_ctx* ctx, T* h, const T* d, size_t count) {
  if (count == 0) return FP_OK;
  FP_HIP_OK(hipMemcpyAsync(h, d, count * sizeof(T), hipMemcpyDeviceToHost, ctx->stream));
  return FP_OK;
}
template <typename T>
void fill_async(fp_ctx* ctx, T* p, size_t n, T v) {
  if (n == 0) return;
  const int blocks = (int)std::min<size_t>((n + 255) / 256, 4096);
  k_fill<T><<<blocks, 256, 0, ctx->stream>>>(p, n, v);
}

// Launches a pass's dominant kernel, kern<<<grid, block, lds>>>(arg) on the
// context's stream. With ctx->kernel_timing the launch's own start and end
// go into the next two pooled events (fp_last_kernel_timing adds them up after
// fp_refine): the runtime stamps the dispatch itself, so no event packets sit
// between the pass's kernels.
template <class Arg>
fp_status launch_timed(fp_ctx* ctx, void (*kern)(Arg), int grid, int block, size_t lds, Arg a) {
  if (!ctx->kernel_timing) {
    kern<<<grid, block, lds, ctx->stream>>>(a);
    return FP_OK;
  }
  hipEvent_t ta = pooled_event(ctx, ctx->timed_next), tb = pooled_event(ctx, ctx->timed_next + 1);
  if (!ta || !tb) return fail(ctx, FP_ERR_HIP, "hipEventCreate failed");
  ctx->timed_next += 2;
  void* args[] = {&a};
  FP_HIP_OK(hipExtLaunchKernel(reinterpret_cast<const void*>(kern), dim3((unsigned)grid), dim3((unsigned)block), args,
                               lds, ctx->stream, ta, tb, 0));
  return FP_OK;
}

// K-sized arrays and the packed-entry layout for K parts
fp_status ensure_k(fp_ctx* ctx, int K);
// k_pass_begin on the context's stream (val may be null)
void pass_begin(fp_ctx* ctx, const int* val, int K, int* mstat);

// blocks of a per-touched-net kernel with `waves` wavefronts per block: a
// round touches at most noe nets
inline int net_grid(const fp_ctx* ctx, int cap, int waves) {
  return std::max(1, std::min(cap, (ctx->noe + waves - 1) / waves));
}

}  // namespace fpe
