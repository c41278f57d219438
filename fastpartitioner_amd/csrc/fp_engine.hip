// fp_engine.hip -- FastPartitioner refinement + metric engine for MI355X (gfx950).
//
// One context holds one hypergraph resident in HBM:
//   net_ptr[noe+1], net_pins[pins]   net->pins CSR (reader order)
//   vtx_ptr[nov+1], vtx_nets[pins]   vertex->nets CSR (ascending nets)
//   apply_pins[pins]                 per net, the ordering ranks of its pins, ascending
//   src[noe], vwgt[nov], order[nov]  net sources, vertex weights, visiting order
// and its connectivity state:
//   part[nov]                        part of each vertex
//   lam[noe]                         lambda(n) = length of net n's list
//   ent[pins]                        list of net n in ent[net_ptr[n] .. +lam[n]), one u32
//                                    per entry: (pin count << pb) | part
//   sizes[K] (int64)                 part weights
//
// A TV refinement pass runs in rounds of R consecutive ordering positions:
//   k_tv_eval    one wavefront per vertex: sweep of the lists of its nets into
//                LDS histograms (count + first-appearance rank per part), LDG2
//                decision by a lexicographic wave reduction, TV gain test
//   [BAL only]   k_bal_hist / k_bal_scan / k_bal_admit: per-part prefix-sum
//                admission under the cap, in ordering order
//   k_ref_commit part vector + part sizes of the accepted moves (REF)
//   k_apply      one wavefront per touched net: replays the round's moves on
//                the net's list in ordering order (remove = swap with last,
//                add = increment or append), so list order matches the
//                reference's sequential edits exactly
// Everything is integer except the LDG2 score, which uses the same IEEE
// double operations as the reference (no contraction, correctly rounded
// division).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <climits>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <numeric>
#include <vector>

#include "../../include/fastpart.h"

namespace {

constexpr int WAVE = 64;
constexpr int EVAL_BLOCK = 256;   // 4 wavefronts
constexpr int APPLY_BLOCK = 256;
constexpr int METRIC_BLOCK = 256;
constexpr int LDS_BUDGET = 64 * 1024;
constexpr unsigned NOFIRST = 0xFFFFFFFFu;

// ---------------------------------------------------------------------------
// wave helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & (WAVE - 1)); }

// orders LDS traffic of one wavefront: everything before is visible to every
// lane of the wave after
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

template <typename T>
__device__ __forceinline__ T wave_sum(T x) {
  for (int off = WAVE / 2; off > 0; off >>= 1) x += __shfl_xor(x, off);
  return x;
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

// ---------------------------------------------------------------------------
// kernels: TV round
// ---------------------------------------------------------------------------
struct EvalArgs {
  const unsigned* vtx_ptr;
  const unsigned* vtx_nets;
  const unsigned* net_ptr;
  const int* lam;
  const unsigned* ent;
  const int* part;
  const int* order;  // nullable: identity
  const int* vwgt;   // nullable: unit
  const long long* sizes;
  int K, pb;
  unsigned pmask;
  int r0, nslots;
  double C;
  long long max_pw, min_pw;
  int* slot_own;
  int* slot_dec;
  int* dec_out;   // nullable (fp_eval_moves)
  int* gain_out;  // nullable
  // REF: stamp the nets of accepted movers into the touched list
  int stamp_nets;
  unsigned tag;
  unsigned* net_stamp;
  int* touched;
  int* tcount;
  int* tcount_next;
  // BAL: per-part candidate weight totals
  long long* cand_tot;
};

// Per vertex: sweep, LDG2 decision, TV gain test (TV/partition.cpp:732-805).
__global__ void __launch_bounds__(EVAL_BLOCK) k_tv_eval(EvalArgs a) {
  extern __shared__ __align__(16) unsigned char lds[];
  const int K = a.K;
  double* s_pen = reinterpret_cast<double*>(lds);
  long long* s_size = reinterpret_cast<long long*>(s_pen + K);
  unsigned* wbase = reinterpret_cast<unsigned*>(s_size + K);
  __shared__ long long red_sz[EVAL_BLOCK / WAVE];
  __shared__ int red_ix[EVAL_BLOCK / WAVE];
  const int lane = lane_id(), wave = (int)(threadIdx.x / WAVE), nw = (int)(blockDim.x / WAVE);
  unsigned* cnt = wbase + (size_t)wave * 2 * (size_t)K;
  unsigned* first = cnt + K;

  // penalties of the frozen part sizes: penalty = 1 - (1.0 * size / C)  (TV/partition.cpp:103)
  long long bs = LLONG_MAX;
  int bi = INT_MAX;
  for (int q = (int)threadIdx.x; q < K; q += (int)blockDim.x) {
    const long long sz = a.sizes[q];
    s_size[q] = sz;
    s_pen[q] = 1.0 - ((double)sz / a.C);
    if (sz < bs) { bs = sz; bi = q; }
  }
  for (int q = lane; q < K; q += WAVE) { cnt[q] = 0; first[q] = NOFIRST; }
  // fallback: lowest-index part of minimum size (TV/partition.cpp:118-132)
  for (int off = WAVE / 2; off > 0; off >>= 1) {
    const long long os = __shfl_xor(bs, off);
    const int oi = __shfl_xor(bi, off);
    if (os < bs || (os == bs && oi < bi)) { bs = os; bi = oi; }
  }
  if (lane == 0) { red_sz[wave] = bs; red_ix[wave] = bi; }
  if (blockIdx.x == 0 && threadIdx.x == 0 && a.tcount_next) *a.tcount_next = 0;
  __syncthreads();
  int fb = red_ix[0];
  {
    long long b = red_sz[0];
    for (int w = 1; w < nw; w++)
      if (red_sz[w] < b || (red_sz[w] == b && red_ix[w] < fb)) { b = red_sz[w]; fb = red_ix[w]; }
  }

  for (int slot = (int)blockIdx.x * nw + wave; slot < a.nslots; slot += (int)gridDim.x * nw) {
    const int z = a.r0 + slot;
    const int v = a.order ? a.order[z] : z;
    const int own = a.part[v];
    const unsigned vb = a.vtx_ptr[v];
    const int deg = (int)(a.vtx_ptr[v + 1] - vb);
    int lg = 0;
    unsigned ebase = 0;
    // sweep: histogram of parts over the lists of v's nets, first-appearance
    // rank = position in the concatenation of those lists (TV/partition.cpp:318-338)
    for (int c0 = 0; c0 < deg; c0 += WAVE) {
      const int i = c0 + lane;
      unsigned s = 0;
      int L = 0;
      if (i < deg) {
        const unsigned n = a.vtx_nets[vb + (unsigned)i];
        s = a.net_ptr[n];
        L = a.lam[n];
      }
      const int nn = min(WAVE, deg - c0);
      for (int j = 0; j < nn; j++) {
        const unsigned sj = (unsigned)__shfl((int)s, j);
        const int Lj = __shfl(L, j);
        for (int t = lane; t < Lj; t += WAVE) {
          const unsigned x = a.ent[sj + (unsigned)t];
          const int q = (int)(x & a.pmask);
          atomicAdd(&cnt[q], 1u);
          atomicMin(&first[q], ebase + (unsigned)t);
          lg += (q == own && (x >> a.pb) == 1u) ? 1 : 0;   // leave gain (TV/partition.cpp:365-368)
        }
        ebase += (unsigned)Lj;
      }
    }
    wave_lds_sync();
    // candidates: lexicographic max of (score, -size, -first rank) over score > 0
    // equals the sequential LDG2 scan over active parts (TV/partition.cpp:98-116)
    double bsc = 0;
    long long bsz = 0;
    unsigned bfi = NOFIRST, bc = 0, cfb = 0;
    int bq = -1;
    for (int q = lane; q < K; q += WAVE) {
      const unsigned c = cnt[q];
      if (c) {
        const unsigned f = first[q];
        cnt[q] = 0;
        first[q] = NOFIRST;
        if (q == fb) cfb = c;
        const int adj = (int)c - (q == own ? 1 : 0);   // scores[part_id]-- (TV/partition.cpp:345)
        const double sc = (double)adj * s_pen[q];
        if (sc > 0) {
          const long long sz = s_size[q];
          if (bq < 0 || sc > bsc || (sc == bsc && (sz < bsz || (sz == bsz && f < bfi)))) {
            bsc = sc; bsz = sz; bfi = f; bq = q; bc = c;
          }
        }
      }
    }
    for (int off = WAVE / 2; off > 0; off >>= 1) {
      const double osc = __shfl_xor(bsc, off);
      const long long osz = __shfl_xor(bsz, off);
      const unsigned ofi = (unsigned)__shfl_xor((int)bfi, off);
      const int oq = __shfl_xor(bq, off);
      const unsigned oc = (unsigned)__shfl_xor((int)bc, off);
      if (oq >= 0 && (bq < 0 || osc > bsc || (osc == bsc && (osz < bsz || (osz == bsz && ofi < bfi))))) {
        bsc = osc; bsz = osz; bfi = ofi; bq = oq; bc = oc;
      }
    }
    cfb = wave_sum(cfb);
    lg = wave_sum(lg);
    wave_lds_sync();
    int d = bq >= 0 ? bq : fb;
    const unsigned craw = bq >= 0 ? bc : cfb;
    if (s_size[d] > a.max_pw) d = own;   // TV/partition.cpp:135-138
    if (s_size[d] < a.min_pw) d = own;   // TV/partition.cpp:140-143
    const int al = deg - (int)craw;      // arrival loss: nets of v lacking d (TV/partition.cpp:369-377)
    const bool accept = d != own && lg >= al;   // TV/partition.cpp:380
    if (lane == 0) {
      a.slot_own[slot] = own;
      a.slot_dec[slot] = accept ? d : -1;
      if (a.dec_out) a.dec_out[v] = d;
      if (a.gain_out) a.gain_out[v] = d == own ? INT_MIN : lg - al;
      if (accept && a.cand_tot)
        atomicAdd(reinterpret_cast<unsigned long long*>(&a.cand_tot[d]),
                  (unsigned long long)(a.vwgt ? a.vwgt[v] : 1));
    }
    if (accept && a.stamp_nets) {
      for (int i = lane; i < deg; i += WAVE) {
        const unsigned n = a.vtx_nets[vb + (unsigned)i];
        if (atomicExch(&a.net_stamp[n], a.tag) != a.tag) a.touched[atomicAdd(a.tcount, 1)] = (int)n;
      }
    }
  }
}

// REF apply bookkeeping: part vector and part sizes (TV/partition.cpp:818-830):
// the source part shrinks, the destination's size is left unchanged.
__global__ void k_ref_commit(const int* __restrict__ slot_own, const int* __restrict__ slot_dec,
                             const int* __restrict__ order, int r0, int nslots, int* __restrict__ part,
                             long long* __restrict__ sizes, unsigned long long* __restrict__ moves) {
  const int slot = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  bool mv = false;
  if (slot < nslots) {
    const int d = slot_dec[slot];
    if (d >= 0) {
      const int z = r0 + slot;
      const int v = order ? order[z] : z;
      part[v] = d;
      atomicAdd(reinterpret_cast<unsigned long long*>(&sizes[slot_own[slot]]), (unsigned long long)(-1LL));
      mv = true;
    }
  }
  const unsigned long long m = __ballot(mv);
  if (lane_id() == 0 && m) atomicAdd(moves, (unsigned long long)__popcll(m));
}

// BAL step 1, one wavefront per tile of 64 slots: over-subscribed parts
// (candidate weight > cap - size) get a dense index (same in every block);
// the tile's candidate weight per over-subscribed part goes to tile_sum[o][tile].
__global__ void __launch_bounds__(256) k_bal_hist(const int* __restrict__ slot_dec, const int* __restrict__ order,
                                                  const int* __restrict__ vwgt, int r0, int nslots, int K,
                                                  const long long* __restrict__ sizes,
                                                  const long long* __restrict__ cand_tot, long long cap,
                                                  int* __restrict__ over_idx, int* __restrict__ over_count,
                                                  long long* __restrict__ tile_sum, int ntiles) {
  extern __shared__ __align__(16) unsigned char lds[];
  int* s_over = reinterpret_cast<int*>(lds);                 // K
  long long* s_acc = reinterpret_cast<long long*>(s_over + K + (K & 1));  // per wave: K
  __shared__ int s_nover;
  const int lane = lane_id(), wave = (int)(threadIdx.x / WAVE), nw = (int)(blockDim.x / WAVE);
  // dense index of over-subscribed parts: exclusive prefix count in part order
  if (threadIdx.x == 0) {
    int o = 0;
    for (int d = 0; d < K; d++) {
      const bool ov = cand_tot[d] > cap - sizes[d];
      s_over[d] = ov ? o++ : -1;
    }
    s_nover = o;
  }
  __syncthreads();
  const int nover = s_nover;
  if (blockIdx.x == 0) {
    for (int d = (int)threadIdx.x; d < K; d += (int)blockDim.x) over_idx[d] = s_over[d];
    if (threadIdx.x == 0) *over_count = nover;
  }
  if (nover == 0) return;
  long long* acc = s_acc + (size_t)wave * (size_t)K;
  for (int tile = (int)blockIdx.x * nw + wave; tile < ntiles; tile += (int)gridDim.x * nw) {
    for (int o = lane; o < nover; o += WAVE) acc[o] = 0;
    wave_lds_sync();
    const int slot = tile * WAVE + lane;
    if (slot < nslots) {
      const int d = slot_dec[slot];
      if (d >= 0 && s_over[d] >= 0) {
        const int z = r0 + slot;
        const long long w = vwgt ? vwgt[order ? order[z] : z] : 1;
        atomicAdd(reinterpret_cast<unsigned long long*>(&acc[s_over[d]]), (unsigned long long)w);
      }
    }
    wave_lds_sync();
    for (int o = lane; o < nover; o += WAVE) tile_sum[(size_t)o * (size_t)ntiles + (size_t)tile] = acc[o];
    wave_lds_sync();
  }
}

// BAL step 2: exclusive scan of tile_sum over tiles, one block per
// over-subscribed part; also clears the candidate totals for the next round.
__global__ void __launch_bounds__(256) k_bal_scan(long long* __restrict__ tile_sum, int ntiles,
                                                  const int* __restrict__ over_count,
                                                  long long* __restrict__ cand_tot, int K) {
  __shared__ long long s[256];
  __shared__ long long carry;
  if (blockIdx.x == 0)
    for (int d = (int)threadIdx.x; d < K; d += (int)blockDim.x) cand_tot[d] = 0;
  const int nover = *over_count;
  for (int o = (int)blockIdx.x; o < nover; o += (int)gridDim.x) {
    long long* row = tile_sum + (size_t)o * (size_t)ntiles;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int t0 = 0; t0 < ntiles; t0 += (int)blockDim.x) {
      const int t = t0 + (int)threadIdx.x;
      const long long x = t < ntiles ? row[t] : 0;
      s[threadIdx.x] = x;
      __syncthreads();
      for (int off = 1; off < (int)blockDim.x; off <<= 1) {
        const long long y = (int)threadIdx.x >= off ? s[threadIdx.x - off] : 0;
        __syncthreads();
        s[threadIdx.x] += y;
        __syncthreads();
      }
      const long long incl = s[threadIdx.x];
      if (t < ntiles) row[t] = carry + incl - x;
      __syncthreads();
      if (threadIdx.x == blockDim.x - 1) carry += incl;
      __syncthreads();
    }
  }
}

// BAL step 3, one wavefront per tile: a candidate into an over-subscribed part
// d is admitted iff sizes[d] + (weight of earlier candidates into d) + w <= cap;
// admitted moves commit part/sizes and stamp their nets for the apply.
__global__ void __launch_bounds__(256) k_bal_admit(int* __restrict__ slot_dec, const int* __restrict__ slot_own,
                                                   const int* __restrict__ order, const int* __restrict__ vwgt,
                                                   int r0, int nslots, const long long* __restrict__ sizes_frozen,
                                                   long long cap, const int* __restrict__ over_idx,
                                                   const int* __restrict__ over_count,
                                                   const long long* __restrict__ tile_sum, int ntiles,
                                                   int* __restrict__ part, long long* __restrict__ sizes_next,
                                                   const unsigned* __restrict__ vtx_ptr,
                                                   const unsigned* __restrict__ vtx_nets, unsigned tag,
                                                   unsigned* __restrict__ net_stamp, int* __restrict__ touched,
                                                   int* __restrict__ tcount, unsigned long long* __restrict__ moves) {
  const int lane = lane_id(), wave = (int)(threadIdx.x / WAVE), nw = (int)(blockDim.x / WAVE);
  const int nover = *over_count;
  for (int tile = (int)blockIdx.x * nw + wave; tile < ntiles; tile += (int)gridDim.x * nw) {
    const int slot = tile * WAVE + lane;
    int d = -1, v = -1;
    long long w = 0;
    if (slot < nslots) {
      d = slot_dec[slot];
      const int z = r0 + slot;
      v = order ? order[z] : z;
      w = vwgt ? vwgt[v] : 1;
    }
    const int o = (d >= 0 && nover > 0) ? over_idx[d] : -1;
    // weight of earlier candidates into the same part within the tile
    long long pre = 0;
    for (int j = 0; j < WAVE; j++) {
      const int dj = __shfl(d, j);
      const long long wj = __shfl(w, j);
      if (j < lane && dj == d) pre += wj;
    }
    bool adm = d >= 0;
    if (o >= 0) {
      const long long before = tile_sum[(size_t)o * (size_t)ntiles + (size_t)tile] + pre;
      adm = sizes_frozen[d] + before + w <= cap;
    }
    if (d >= 0 && !adm) slot_dec[slot] = -1;
    if (adm) {
      const int own = slot_own[slot];
      part[v] = d;
      atomicAdd(reinterpret_cast<unsigned long long*>(&sizes_next[own]), (unsigned long long)(-w));
      atomicAdd(reinterpret_cast<unsigned long long*>(&sizes_next[d]), (unsigned long long)w);
    }
    const unsigned long long am = __ballot(adm);
    if (lane == 0 && am) atomicAdd(moves, (unsigned long long)__popcll(am));
    // stamp the nets of the admitted movers of this tile, one mover at a time
    unsigned long long m = am;
    while (m) {
      const int l = __ffsll((long long)m) - 1;
      m &= m - 1;
      const int vv = __shfl(v, l);
      const unsigned vb = vtx_ptr[vv], ve = vtx_ptr[vv + 1];
      for (unsigned i = vb + (unsigned)lane; i < ve; i += WAVE) {
        const unsigned n = vtx_nets[i];
        if (atomicExch(&net_stamp[n], tag) != tag) touched[atomicAdd(tcount, 1)] = (int)n;
      }
    }
  }
}

struct ApplyArgs {
  const unsigned* net_ptr;
  const int* apply_pins;
  const int* slot_own;
  const int* slot_dec;
  const int* touched;
  const int* tcount;
  int* lam;
  unsigned* ent;
  int K, pb;
  unsigned pmask;
  int r0, r1;
};

// Replays the round's moves on every touched net, in ordering order
// (TV/partition.cpp:846-890: remove = decrement or swap-with-last, add =
// increment or append). One wavefront per net; the list lives in LDS.
__global__ void __launch_bounds__(APPLY_BLOCK) k_apply(ApplyArgs a) {
  extern __shared__ __align__(16) unsigned char lds[];
  const int lane = lane_id(), wave = (int)(threadIdx.x / WAVE), nw = (int)(blockDim.x / WAVE);
  unsigned* list = reinterpret_cast<unsigned*>(lds) + (size_t)wave * (size_t)a.K;
  const unsigned inc = 1u << a.pb;
  const int ntouched = *a.tcount;
  for (int i = (int)blockIdx.x * nw + wave; i < ntouched; i += (int)gridDim.x * nw) {
    const int n = a.touched[i];
    const unsigned s = a.net_ptr[n], e = a.net_ptr[n + 1];
    int L = a.lam[n];
    for (int t = lane; t < L; t += WAVE) list[t] = a.ent[s + (unsigned)t];
    const unsigned lo = wave_lower_bound(a.apply_pins, s, e, a.r0);
    const unsigned hi = wave_lower_bound(a.apply_pins, lo, e, a.r1);
    wave_lds_sync();
    for (unsigned k0 = lo; k0 < hi; k0 += WAVE) {
      const unsigned k = k0 + (unsigned)lane;
      int dd = -1, oo = -1;
      if (k < hi) {
        const int slot = a.apply_pins[k] - a.r0;
        dd = a.slot_dec[slot];
        if (dd >= 0) oo = a.slot_own[slot];
      }
      unsigned long long m = __ballot(dd >= 0);
      while (m) {
        const int l = __ffsll((long long)m) - 1;
        m &= m - 1;
        const int p = __shfl(oo, l), q = __shfl(dd, l);
        // remove one pin of p
        int j = wave_find(list, L, p, a.pmask);
        if (j >= 0) {
          const unsigned x = list[j];
          if ((x >> a.pb) > 1u) {
            if (lane == 0) list[j] = x - inc;
          } else {
            if (lane == 0) list[j] = list[L - 1];
            L--;
          }
          wave_lds_sync();
        }
        // add one pin of q
        j = wave_find(list, L, q, a.pmask);
        if (j >= 0) {
          if (lane == 0) list[j] += inc;
        } else {
          if (lane == 0) list[L] = inc | (unsigned)q;
          L++;
        }
        wave_lds_sync();
      }
    }
    for (int t = lane; t < L; t += WAVE) a.ent[s + (unsigned)t] = list[t];
    if (lane == 0) a.lam[n] = L;
    wave_lds_sync();
  }
}

// ---------------------------------------------------------------------------
// kernels: metrics from scratch (shared/metric.cpp), one wavefront per net
// ---------------------------------------------------------------------------
struct MetricArgs {
  const unsigned* net_ptr;
  const unsigned* net_pins;
  const int* part;
  const int* src;
  int noe, K, nov;
  int* lam;
  unsigned* p2p;               // K*K bits, zeroed
  unsigned long long* send;    // MGAV send, K
  unsigned long long* recv;    // K
  unsigned long long* sm;      // K
  unsigned long long* rm;      // K
  unsigned long long* send_msv;  // MGV send, K
  unsigned long long* tvcn;    // [0] TV, [1] CN
};

__global__ void __launch_bounds__(METRIC_BLOCK) k_metrics(MetricArgs a) {
  extern __shared__ __align__(16) unsigned char lds[];
  const int K = a.K;
  unsigned long long* s_send = reinterpret_cast<unsigned long long*>(lds);
  unsigned long long* s_recv = s_send + K;
  unsigned long long* s_msv = s_recv + K;
  int* marks = reinterpret_cast<int*>(s_msv + K);
  const int lane = lane_id(), wave = (int)(threadIdx.x / WAVE), nw = (int)(blockDim.x / WAVE);
  int* mark = marks + (size_t)wave * (size_t)K;
  for (int q = (int)threadIdx.x; q < K; q += (int)blockDim.x) { s_send[q] = 0; s_recv[q] = 0; s_msv[q] = 0; }
  for (int q = lane; q < K; q += WAVE) mark[q] = -1;
  __syncthreads();
  unsigned long long tv = 0, cn = 0;
  for (int n = (int)blockIdx.x * nw + wave; n < a.noe; n += (int)gridDim.x * nw) {
    const unsigned s = a.net_ptr[n], e = a.net_ptr[n + 1];
    const int sv = a.src[n];
    const int sp = (sv >= 0 && sv < a.nov) ? a.part[sv] : -1;
    int l = 0, occ = 0;
    for (unsigned k0 = s; k0 < e; k0 += WAVE) {
      const unsigned k = k0 + (unsigned)lane;
      bool fresh = false;
      int q = -1;
      if (k < e) {
        const int v = (int)a.net_pins[k];
        q = a.part[v];
        occ += (v == sv) ? 1 : 0;
        fresh = atomicExch(&mark[q], n) != n;   // first pin of part q in this net
      }
      l += __popcll(__ballot(fresh));
      if (fresh && sp >= 0 && q != sp) {
        atomicAdd(&s_recv[q], 1ull);
        const unsigned bit = (unsigned)sp * (unsigned)K + (unsigned)q;
        const unsigned old = atomicOr(&a.p2p[bit >> 5], 1u << (bit & 31));
        if (!(old & (1u << (bit & 31)))) {
          atomicAdd(&a.sm[sp], 1ull);
          atomicAdd(&a.rm[q], 1ull);
        }
      }
    }
    occ = wave_sum(occ);
    if (lane == 0) {
      a.lam[n] = l;
      if (l != 0) tv += (unsigned long long)(l - 1);
      if (l > 1) cn++;
      if (sp >= 0) {
        atomicAdd(&s_send[sp], (unsigned long long)(long long)(l - 1));
        if (occ) atomicAdd(&s_msv[sp], (unsigned long long)((long long)(l - 1) * occ));
      }
    }
  }
  __syncthreads();
  for (int q = (int)threadIdx.x; q < K; q += (int)blockDim.x) {
    if (s_send[q]) atomicAdd(&a.send[q], s_send[q]);
    if (s_recv[q]) atomicAdd(&a.recv[q], s_recv[q]);
    if (s_msv[q]) atomicAdd(&a.send_msv[q], s_msv[q]);
  }
  if (lane == 0) {
    if (tv) atomicAdd(&a.tvcn[0], tv);
    if (cn) atomicAdd(&a.tvcn[1], cn);
  }
}

template <typename T>
__global__ void k_fill(T* p, size_t n, T v) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

}  // namespace

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
struct fp_ctx {
  int device = -1;
  hipStream_t stream = nullptr;
  char err[512] = {0};
  long long dev_bytes = 0;
  std::vector<void*> owned;

  // hypergraph
  int nov = 0, noe = 0;
  long long pins = 0;
  unsigned *net_ptr = nullptr, *net_pins = nullptr, *vtx_ptr = nullptr, *vtx_nets = nullptr;
  int *src = nullptr, *vwgt = nullptr, *order = nullptr, *apply_pins = nullptr;
  bool has_graph = false, dup_pins = false;
  long long total_weight = 0;
  std::vector<unsigned> h_net_ptr, h_vtx_ptr, h_vtx_nets;
  std::vector<int> h_rank_pins;  // per net, ordering ranks of its pins ascending

  // state
  int K = 0, pb = 0;
  unsigned pmask = 0;
  int *part = nullptr, *lam = nullptr;
  unsigned* ent = nullptr;
  long long* sizes = nullptr;
  bool has_state = false;

  // round scratch
  int rmax = 0;
  int *slot_own = nullptr, *slot_dec = nullptr, *touched = nullptr, *counters = nullptr;
  unsigned* net_stamp = nullptr;
  unsigned tag = 0;
  long long* cand_tot = nullptr;
  long long* tile_sum = nullptr;
  long long* sizes_frozen = nullptr;
  size_t tile_cap = 0;
  int *over_idx = nullptr, *over_count = nullptr;
  unsigned long long* moves = nullptr;

  // metric scratch
  unsigned* p2p = nullptr;
  unsigned long long* mvec = nullptr;  // 5*K + 2

  // timing
  bool kernel_timing = false;
  long long timed_launches = 0;
  double timed_ms = 0;
  std::vector<hipEvent_t> ev_pool;
};

namespace {

fp_status fail(fp_ctx* c, fp_status st, const char* fmt, ...) {
  if (c) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(c->err, sizeof c->err, fmt, ap);
    va_end(ap);
  }
  return st;
}

#define HIP_OK(call)                                                                        \
  do {                                                                                      \
    hipError_t e_ = (call);                                                                 \
    if (e_ != hipSuccess) return fail(ctx, FP_ERR_HIP, "%s: %s", #call, hipGetErrorString(e_)); \
  } while (0)

template <typename T>
fp_status dalloc(fp_ctx* ctx, T** p, size_t count) {
  *p = nullptr;
  if (count == 0) count = 1;
  void* q = nullptr;
  hipError_t e = hipMalloc(&q, count * sizeof(T));
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return fail(ctx, FP_ERR_OOM, "hipMalloc(%zu bytes) failed: %s", count * sizeof(T), hipGetErrorString(e));
  }
  ctx->owned.push_back(q);
  ctx->dev_bytes += (long long)(count * sizeof(T));
  *p = static_cast<T*>(q);
  return FP_OK;
}

void dfree_all(fp_ctx* ctx) {
  for (void* p : ctx->owned) (void)hipFree(p);
  ctx->owned.clear();
  ctx->dev_bytes = 0;
}

#define TRY(x)                        \
  do {                                \
    fp_status s_ = (x);               \
    if (s_ != FP_OK) return s_;       \
  } while (0)

template <typename T>
fp_status h2d(fp_ctx* ctx, T* d, const T* h, size_t count) {
  if (count == 0) return FP_OK;
  HIP_OK(hipMemcpyAsync(d, h, count * sizeof(T), hipMemcpyHostToDevice, ctx->stream));
  return FP_OK;
}
template <typename T>
fp_status d2h(fp_ctx* ctx, T* h, const T* d, size_t count) {
  if (count == 0) return FP_OK;
  HIP_OK(hipMemcpyAsync(h, d, count * sizeof(T), hipMemcpyDeviceToHost, ctx->stream));
  return FP_OK;
}

template <typename T>
void fill_async(fp_ctx* ctx, T* p, size_t n, T v) {
  if (n == 0) return;
  const int blocks = (int)std::min<size_t>((n + 255) / 256, 4096);
  k_fill<T><<<blocks, 256, 0, ctx->stream>>>(p, n, v);
}

int bits_for(int K) {
  int b = 1;
  while ((1 << b) < K) b++;
  return b;
}

int eval_waves(int K) {
  // LDS: 16 bytes per part shared + 8 bytes per part per wavefront
  for (int w = EVAL_BLOCK / WAVE; w >= 1; w--)
    if ((size_t)K * 16 + (size_t)w * 8 * (size_t)K <= (size_t)LDS_BUDGET) return w;
  return 0;
}


}  // namespace

extern "C" {

fp_status fp_device_count(int* count) {
  if (!count) return FP_ERR_ARG;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) {
    (void)hipGetLastError();
    *count = 0;
    return FP_ERR_NODEV;
  }
  *count = n;
  return n > 0 ? FP_OK : FP_ERR_NODEV;
}

fp_status fp_ctx_create(fp_ctx** out, int device) {
  if (!out) return FP_ERR_ARG;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
    (void)hipGetLastError();
    return FP_ERR_NODEV;
  }
  if (device < 0 || device >= n) return FP_ERR_ARG;
  fp_ctx* ctx = new fp_ctx();
  ctx->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
    (void)hipGetLastError();
    delete ctx;
    return FP_ERR_HIP;
  }
  *out = ctx;
  return FP_OK;
}

void fp_ctx_destroy(fp_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  dfree_all(ctx);
  for (hipEvent_t e : ctx->ev_pool) (void)hipEventDestroy(e);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

const char* fp_ctx_last_error(const fp_ctx* ctx) { return ctx ? ctx->err : "null context"; }

long long fp_ctx_device_bytes(const fp_ctx* ctx) { return ctx ? ctx->dev_bytes : 0; }

fp_status fp_upload(fp_ctx* ctx, int nov, int noe, const unsigned* net_ptr, const unsigned* net_pins,
                    const unsigned* vtx_ptr, const unsigned* vtx_nets, const int* src, const int* vwgt,
                    const int* ordering) {
  if (!ctx) return FP_ERR_ARG;
  if (nov < 1 || noe < 0 || !net_ptr || (!net_pins && net_ptr[noe] > 0))
    return fail(ctx, FP_ERR_ARG, "fp_upload: bad sizes or null arrays");
  HIP_OK(hipSetDevice(ctx->device));
  HIP_OK(hipStreamSynchronize(ctx->stream));
  dfree_all(ctx);
  ctx->has_graph = ctx->has_state = false;
  const long long pins = net_ptr[noe];
  if (net_ptr[0] != 0) return fail(ctx, FP_ERR_ARG, "fp_upload: net_ptr[0] != 0");
  for (int n = 0; n < noe; n++)
    if (net_ptr[n + 1] < net_ptr[n]) return fail(ctx, FP_ERR_ARG, "fp_upload: net_ptr not monotone at %d", n);
  for (long long k = 0; k < pins; k++)
    if (net_pins[k] >= (unsigned)nov) return fail(ctx, FP_ERR_ARG, "fp_upload: pin %lld out of range", k);
  if (src)
    for (int n = 0; n < noe; n++)
      if (src[n] < -1 || src[n] >= nov) return fail(ctx, FP_ERR_ARG, "fp_upload: src[%d] out of range", n);
  long long tw = 0;
  if (vwgt) {
    for (int v = 0; v < nov; v++) {
      if (vwgt[v] < 0) return fail(ctx, FP_ERR_ARG, "fp_upload: negative weight at %d", v);
      tw += vwgt[v];
    }
  } else {
    tw = nov;
  }
  std::vector<int> rank((size_t)nov);
  if (ordering) {
    std::vector<char> seen((size_t)nov, 0);
    for (int z = 0; z < nov; z++) {
      const int v = ordering[z];
      if (v < 0 || v >= nov || seen[(size_t)v]) return fail(ctx, FP_ERR_ARG, "fp_upload: ordering is not a permutation");
      seen[(size_t)v] = 1;
      rank[(size_t)v] = z;
    }
  } else {
    std::iota(rank.begin(), rank.end(), 0);
  }
  ctx->nov = nov;
  ctx->noe = noe;
  ctx->pins = pins;
  ctx->total_weight = tw;
  ctx->h_net_ptr.assign(net_ptr, net_ptr + noe + 1);
  // vertex->nets view
  if (vtx_ptr && vtx_nets) {
    if (vtx_ptr[0] != 0 || (long long)vtx_ptr[nov] != pins) return fail(ctx, FP_ERR_ARG, "fp_upload: vtx_ptr inconsistent");
    ctx->h_vtx_ptr.assign(vtx_ptr, vtx_ptr + nov + 1);
    ctx->h_vtx_nets.assign(vtx_nets, vtx_nets + pins);
    for (long long k = 0; k < pins; k++)
      if (vtx_nets[k] >= (unsigned)noe) return fail(ctx, FP_ERR_ARG, "fp_upload: vtx_nets out of range");
  } else {
    ctx->h_vtx_ptr.assign((size_t)nov + 1, 0);
    ctx->h_vtx_nets.assign((size_t)pins, 0);
    for (long long k = 0; k < pins; k++) ctx->h_vtx_ptr[net_pins[k] + 1]++;
    for (int v = 0; v < nov; v++) ctx->h_vtx_ptr[(size_t)v + 1] += ctx->h_vtx_ptr[(size_t)v];
    std::vector<unsigned> fill(ctx->h_vtx_ptr.begin(), ctx->h_vtx_ptr.end() - 1);
    for (int n = 0; n < noe; n++)
      for (unsigned k = net_ptr[n]; k < net_ptr[n + 1]; k++) ctx->h_vtx_nets[fill[net_pins[k]]++] = (unsigned)n;
  }
  // per net: ordering ranks of the pins, ascending (the apply replays in this order)
  ctx->h_rank_pins.resize((size_t)pins);
  bool dup = false;
#pragma omp parallel for schedule(dynamic, 1024) reduction(|| : dup)
  for (int n = 0; n < noe; n++) {
    int* r = ctx->h_rank_pins.data() + net_ptr[n];
    const int len = (int)(net_ptr[n + 1] - net_ptr[n]);
    bool sorted = true;
    for (int k = 0; k < len; k++) {
      r[k] = rank[net_pins[net_ptr[n] + (unsigned)k]];
      if (k && r[k] <= r[k - 1]) sorted = false;
    }
    if (!sorted) {
      std::sort(r, r + len);
      for (int k = 1; k < len; k++)
        if (r[k] == r[k - 1]) dup = true;
    }
  }
  ctx->dup_pins = dup;

  TRY(dalloc(ctx, &ctx->net_ptr, (size_t)noe + 1));
  TRY(dalloc(ctx, &ctx->net_pins, (size_t)pins));
  TRY(dalloc(ctx, &ctx->vtx_ptr, (size_t)nov + 1));
  TRY(dalloc(ctx, &ctx->vtx_nets, (size_t)pins));
  TRY(dalloc(ctx, &ctx->apply_pins, (size_t)pins));
  TRY(dalloc(ctx, &ctx->src, (size_t)std::max(noe, 1)));
  TRY(h2d(ctx, ctx->net_ptr, net_ptr, (size_t)noe + 1));
  TRY(h2d(ctx, ctx->net_pins, net_pins, (size_t)pins));
  TRY(h2d(ctx, ctx->vtx_ptr, ctx->h_vtx_ptr.data(), (size_t)nov + 1));
  TRY(h2d(ctx, ctx->vtx_nets, ctx->h_vtx_nets.data(), (size_t)pins));
  TRY(h2d(ctx, ctx->apply_pins, ctx->h_rank_pins.data(), (size_t)pins));
  if (src) {
    TRY(h2d(ctx, ctx->src, src, (size_t)noe));
  } else {
    fill_async<int>(ctx, ctx->src, (size_t)std::max(noe, 1), -1);
  }
  ctx->vwgt = nullptr;
  if (vwgt) {
    TRY(dalloc(ctx, &ctx->vwgt, (size_t)nov));
    TRY(h2d(ctx, ctx->vwgt, vwgt, (size_t)nov));
  }
  ctx->order = nullptr;
  if (ordering) {
    bool ident = true;
    for (int z = 0; z < nov && ident; z++) ident = ordering[z] == z;
    if (!ident) {
      TRY(dalloc(ctx, &ctx->order, (size_t)nov));
      TRY(h2d(ctx, ctx->order, ordering, (size_t)nov));
    }
  }
  // state and scratch that do not depend on K
  TRY(dalloc(ctx, &ctx->part, (size_t)nov));
  TRY(dalloc(ctx, &ctx->lam, (size_t)std::max(noe, 1)));
  TRY(dalloc(ctx, &ctx->ent, (size_t)pins));
  TRY(dalloc(ctx, &ctx->net_stamp, (size_t)std::max(noe, 1)));
  TRY(dalloc(ctx, &ctx->touched, (size_t)std::max(noe, 1)));
  TRY(dalloc(ctx, &ctx->counters, 4));
  TRY(dalloc(ctx, &ctx->moves, 1));
  TRY(dalloc(ctx, &ctx->over_count, 1));
  fill_async<unsigned>(ctx, ctx->net_stamp, (size_t)std::max(noe, 1), 0u);
  ctx->tag = 0;
  ctx->rmax = 0;
  ctx->K = 0;
  ctx->tile_cap = 0;
  HIP_OK(hipGetLastError());
  HIP_OK(hipStreamSynchronize(ctx->stream));
  ctx->has_graph = true;
  return FP_OK;
}

}  // extern "C"

namespace {

fp_status ensure_k(fp_ctx* ctx, int K) {
  if (K < 1) return fail(ctx, FP_ERR_ARG, "K must be >= 1");
  if (K > 65536) return fail(ctx, FP_ERR_UNSUPPORTED, "K = %d above 65536", K);
  if (eval_waves(K) == 0) return fail(ctx, FP_ERR_UNSUPPORTED, "K = %d needs more LDS than a workgroup has", K);
  const int pb = bits_for(K);
  unsigned maxdeg = 0;
  for (int n = 0; n < ctx->noe; n++) maxdeg = std::max(maxdeg, ctx->h_net_ptr[(size_t)n + 1] - ctx->h_net_ptr[(size_t)n]);
  if (pb < 32 && (unsigned long long)maxdeg >= (1ull << (32 - pb)))
    return fail(ctx, FP_ERR_UNSUPPORTED, "net of %u pins too large for K = %d in the packed list layout", maxdeg, K);
  if (ctx->K != K) {
    // (re)allocate K-sized arrays; old ones stay owned until the next upload
    TRY(dalloc(ctx, &ctx->sizes, (size_t)K));
    TRY(dalloc(ctx, &ctx->sizes_frozen, (size_t)K));
    TRY(dalloc(ctx, &ctx->cand_tot, (size_t)K));
    TRY(dalloc(ctx, &ctx->over_idx, (size_t)K));
    TRY(dalloc(ctx, &ctx->p2p, ((size_t)K * (size_t)K + 31) / 32));
    TRY(dalloc(ctx, &ctx->mvec, (size_t)5 * (size_t)K + 2));
    fill_async<long long>(ctx, ctx->cand_tot, (size_t)K, 0);
    ctx->K = K;
  }
  ctx->pb = pb;
  ctx->pmask = (pb >= 32) ? 0xFFFFFFFFu : ((1u << pb) - 1u);
  return FP_OK;
}

fp_status upload_state(fp_ctx* ctx, const int* part, const std::vector<int>& lam, const std::vector<unsigned>& ent,
                       const long long* sizes_in) {
  std::vector<long long> sz((size_t)ctx->K, 0);
  if (sizes_in) {
    std::copy(sizes_in, sizes_in + ctx->K, sz.begin());
  } else {
    // weights from host copy of vwgt are not kept; recompute on host from part
    // using unit weights unless a weight array was uploaded
    std::vector<int> w;
    if (ctx->vwgt) {
      w.resize((size_t)ctx->nov);
      HIP_OK(hipMemcpy(w.data(), ctx->vwgt, sizeof(int) * (size_t)ctx->nov, hipMemcpyDeviceToHost));
    }
    for (int v = 0; v < ctx->nov; v++) sz[(size_t)part[v]] += ctx->vwgt ? w[(size_t)v] : 1;
  }
  TRY(h2d(ctx, ctx->part, part, (size_t)ctx->nov));
  TRY(h2d(ctx, ctx->lam, lam.data(), (size_t)ctx->noe));
  TRY(h2d(ctx, ctx->ent, ent.data(), (size_t)ctx->pins));
  TRY(h2d(ctx, ctx->sizes, sz.data(), (size_t)ctx->K));
  HIP_OK(hipStreamSynchronize(ctx->stream));
  ctx->has_state = true;
  return FP_OK;
}

}  // namespace

extern "C" {

fp_status fp_set_partition(fp_ctx* ctx, int K, const int* part) {
  if (!ctx) return FP_ERR_ARG;
  if (!ctx->has_graph) return fail(ctx, FP_ERR_STATE, "fp_set_partition before fp_upload");
  if (!part) return fail(ctx, FP_ERR_ARG, "null part");
  HIP_OK(hipSetDevice(ctx->device));
  for (int v = 0; v < ctx->nov; v++)
    if (part[v] < 0 || part[v] >= K) return fail(ctx, FP_ERR_ARG, "part[%d] = %d out of [0, K)", v, part[v]);
  TRY(ensure_k(ctx, K));
  const int noe = ctx->noe;
  std::vector<int> lam((size_t)noe, 0);
  std::vector<unsigned> ent((size_t)ctx->pins, 0);
  const unsigned inc = 1u << ctx->pb;
  // list order = first appearance of each part along the ordering: walk the
  // pins of each net in rank order (h_rank_pins) and append unseen parts
  std::vector<int> order_of_rank((size_t)ctx->nov);
  {
    std::vector<int> ord((size_t)ctx->nov);
    if (ctx->order) {
      HIP_OK(hipMemcpy(ord.data(), ctx->order, sizeof(int) * (size_t)ctx->nov, hipMemcpyDeviceToHost));
    } else {
      std::iota(ord.begin(), ord.end(), 0);
    }
    order_of_rank.swap(ord);
  }
#pragma omp parallel
  {
    std::vector<int> slot((size_t)K, -1), owner((size_t)K, -1);
#pragma omp for schedule(dynamic, 1024)
    for (int n = 0; n < noe; n++) {
      const unsigned s = ctx->h_net_ptr[(size_t)n], e = ctx->h_net_ptr[(size_t)n + 1];
      int L = 0;
      for (unsigned k = s; k < e; k++) {
        const int q = part[order_of_rank[(size_t)ctx->h_rank_pins[k]]];
        if (owner[(size_t)q] != n) {
          owner[(size_t)q] = n;
          slot[(size_t)q] = L;
          ent[s + (unsigned)L] = inc | (unsigned)q;
          L++;
        } else {
          ent[s + (unsigned)slot[(size_t)q]] += inc;
        }
      }
      lam[(size_t)n] = L;
    }
  }
  return upload_state(ctx, part, lam, ent, nullptr);
}

fp_status fp_set_state(fp_ctx* ctx, int K, const int* part, const int* ep_ends, const int* ep_parts,
                       const int* ep_cnt, const long long* sizes) {
  if (!ctx) return FP_ERR_ARG;
  if (!ctx->has_graph) return fail(ctx, FP_ERR_STATE, "fp_set_state before fp_upload");
  if (!part || !ep_ends || !ep_parts || !ep_cnt) return fail(ctx, FP_ERR_ARG, "null state array");
  HIP_OK(hipSetDevice(ctx->device));
  for (int v = 0; v < ctx->nov; v++)
    if (part[v] < 0 || part[v] >= K) return fail(ctx, FP_ERR_ARG, "part[%d] = %d out of [0, K)", v, part[v]);
  TRY(ensure_k(ctx, K));
  std::vector<int> lam(ep_ends, ep_ends + ctx->noe);
  std::vector<unsigned> ent((size_t)ctx->pins, 0);
  for (int n = 0; n < ctx->noe; n++) {
    const unsigned s = ctx->h_net_ptr[(size_t)n], len = ctx->h_net_ptr[(size_t)n + 1] - s;
    if (lam[(size_t)n] < 0 || (unsigned)lam[(size_t)n] > len) return fail(ctx, FP_ERR_ARG, "ep_ends[%d] out of range", n);
    for (int j = 0; j < lam[(size_t)n]; j++) {
      const int q = ep_parts[s + (unsigned)j], c = ep_cnt[s + (unsigned)j];
      if (q < 0 || q >= K || c < 1 || (unsigned)c > len)
        return fail(ctx, FP_ERR_ARG, "list entry %u of net %d invalid", j, n);
      ent[s + (unsigned)j] = ((unsigned)c << ctx->pb) | (unsigned)q;
    }
  }
  return upload_state(ctx, part, lam, ent, sizes);
}

fp_status fp_get_state(fp_ctx* ctx, int* part, int* ep_ends, int* ep_parts, int* ep_cnt, long long* sizes) {
  if (!ctx) return FP_ERR_ARG;
  if (!ctx->has_state) return fail(ctx, FP_ERR_STATE, "no state");
  HIP_OK(hipSetDevice(ctx->device));
  std::vector<int> lam((size_t)ctx->noe);
  std::vector<unsigned> ent;
  TRY(d2h(ctx, lam.data(), ctx->lam, (size_t)ctx->noe));
  if (ep_parts || ep_cnt) {
    ent.resize((size_t)ctx->pins);
    TRY(d2h(ctx, ent.data(), ctx->ent, (size_t)ctx->pins));
  }
  if (part) TRY(d2h(ctx, part, ctx->part, (size_t)ctx->nov));
  if (sizes) TRY(d2h(ctx, sizes, ctx->sizes, (size_t)ctx->K));
  HIP_OK(hipStreamSynchronize(ctx->stream));
  if (ep_ends) std::copy(lam.begin(), lam.end(), ep_ends);
  if (ep_parts || ep_cnt) {
    for (int n = 0; n < ctx->noe; n++) {
      const unsigned s = ctx->h_net_ptr[(size_t)n], e = ctx->h_net_ptr[(size_t)n + 1];
      for (unsigned k = s; k < e; k++) {
        const bool live = (int)(k - s) < lam[(size_t)n];
        if (ep_parts) ep_parts[k] = live ? (int)(ent[k] & ctx->pmask) : -1;
        if (ep_cnt) ep_cnt[k] = live ? (int)(ent[k] >> ctx->pb) : 0;
      }
    }
  }
  return FP_OK;
}

}  // extern "C"

namespace {

fp_status ensure_round_scratch(fp_ctx* ctx, int R) {
  if (R > ctx->rmax) {
    TRY(dalloc(ctx, &ctx->slot_own, (size_t)R));
    TRY(dalloc(ctx, &ctx->slot_dec, (size_t)R));
    ctx->rmax = R;
  }
  return FP_OK;
}

fp_status check_sched(fp_ctx* ctx, const fp_schedule* s) {
  if (!s) return fail(ctx, FP_ERR_ARG, "null schedule");
  if (s->kind != FP_SCHED_REF && s->kind != FP_SCHED_BAL) return fail(ctx, FP_ERR_ARG, "unknown schedule kind %d", s->kind);
  if (s->round < 1) return fail(ctx, FP_ERR_ARG, "round size must be >= 1");
  if (!(s->ldg_C > 0)) return fail(ctx, FP_ERR_ARG, "ldg_C must be > 0");
  return FP_OK;
}

EvalArgs make_eval_args(fp_ctx* ctx, const fp_schedule* s) {
  EvalArgs a{};
  a.vtx_ptr = ctx->vtx_ptr;
  a.vtx_nets = ctx->vtx_nets;
  a.net_ptr = ctx->net_ptr;
  a.lam = ctx->lam;
  a.ent = ctx->ent;
  a.part = ctx->part;
  a.order = ctx->order;
  a.vwgt = ctx->vwgt;
  a.sizes = ctx->sizes;
  a.K = ctx->K;
  a.pb = ctx->pb;
  a.pmask = ctx->pmask;
  a.C = s->ldg_C;
  a.max_pw = s->max_pw;
  a.min_pw = s->min_pw;
  a.slot_own = ctx->slot_own;
  a.slot_dec = ctx->slot_dec;
  a.net_stamp = ctx->net_stamp;
  a.touched = ctx->touched;
  return a;
}

size_t eval_lds(int K, int waves) { return (size_t)K * 16 + (size_t)waves * 8 * (size_t)K; }

int eval_grid(int nslots, int waves) {
  const int blocks = (nslots + waves - 1) / waves;
  return std::max(1, std::min(blocks, 8192));
}

hipEvent_t pooled_event(fp_ctx* ctx, size_t i) {
  while (ctx->ev_pool.size() <= i) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    ctx->ev_pool.push_back(e);
  }
  return ctx->ev_pool[i];
}

}  // namespace

extern "C" {

fp_status fp_set_kernel_timing(fp_ctx* ctx, int enable) {
  if (!ctx) return FP_ERR_ARG;
  ctx->kernel_timing = enable != 0;
  return FP_OK;
}

fp_status fp_last_kernel_timing(fp_ctx* ctx, long long* launches, double* total_ms) {
  if (!ctx) return FP_ERR_ARG;
  if (launches) *launches = ctx->timed_launches;
  if (total_ms) *total_ms = ctx->timed_ms;
  return FP_OK;
}

fp_status fp_refine(fp_ctx* ctx, fp_metric metric, int passes, const fp_schedule* sched, fp_refine_stats* stats) {
  if (!ctx) return FP_ERR_ARG;
  if (!ctx->has_state) return fail(ctx, FP_ERR_STATE, "fp_refine before a state was set");
  TRY(check_sched(ctx, sched));
  if (passes < 0) return fail(ctx, FP_ERR_ARG, "passes < 0");
  if (metric != FP_METRIC_TV) return fail(ctx, FP_ERR_UNSUPPORTED, "refinement for metric %d not available", (int)metric);
  if (ctx->dup_pins) return fail(ctx, FP_ERR_UNSUPPORTED, "a net lists the same vertex twice; refinement needs distinct pins");
  HIP_OK(hipSetDevice(ctx->device));
  const int nov = ctx->nov, K = ctx->K;
  const int R = std::min(sched->round, nov);
  TRY(ensure_round_scratch(ctx, R));
  const bool bal = sched->kind == FP_SCHED_BAL;
  const int ntiles_max = (R + WAVE - 1) / WAVE;
  if (bal && (size_t)K * (size_t)ntiles_max > ctx->tile_cap) {
    TRY(dalloc(ctx, &ctx->tile_sum, (size_t)K * (size_t)ntiles_max));
    ctx->tile_cap = (size_t)K * (size_t)ntiles_max;
  }
  const int waves = eval_waves(K);
  const size_t lds_e = eval_lds(K, waves);
  const int apply_waves = APPLY_BLOCK / WAVE;
  const size_t lds_a = (size_t)apply_waves * (size_t)K * sizeof(unsigned);
  if (lds_a > (size_t)LDS_BUDGET) return fail(ctx, FP_ERR_UNSUPPORTED, "K too large for the apply kernel");
  const int apply_grid = 2048;

  fill_async<int>(ctx, ctx->counters, 4, 0);
  fill_async<unsigned long long>(ctx, ctx->moves, 1, 0ull);
  hipEvent_t e0 = pooled_event(ctx, 0), e1 = pooled_event(ctx, 1);
  if (!e0 || !e1) return fail(ctx, FP_ERR_HIP, "hipEventCreate failed");
  HIP_OK(hipEventRecord(e0, ctx->stream));
  ctx->timed_launches = 0;
  ctx->timed_ms = 0;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> timed;
  long long rounds = 0;
  for (int it = 0; it < passes; it++) {
    for (int r0 = 0; r0 < nov; r0 += R) {
      const int nslots = std::min(R, nov - r0);
      ctx->tag++;
      const int par = (int)(ctx->tag & 1u);
      EvalArgs a = make_eval_args(ctx, sched);
      a.r0 = r0;
      a.nslots = nslots;
      a.stamp_nets = bal ? 0 : 1;
      a.tag = ctx->tag;
      a.tcount = ctx->counters + par;
      a.tcount_next = ctx->counters + (par ^ 1);
      a.cand_tot = bal ? ctx->cand_tot : nullptr;
      const int grid_e = eval_grid(nslots, waves);
      hipEvent_t ta = nullptr, tb = nullptr;
      if (ctx->kernel_timing) {
        ta = pooled_event(ctx, 2 + 2 * timed.size());
        tb = pooled_event(ctx, 3 + 2 * timed.size());
        if (!ta || !tb) return fail(ctx, FP_ERR_HIP, "hipEventCreate failed");
        HIP_OK(hipEventRecord(ta, ctx->stream));
      }
      k_tv_eval<<<grid_e, waves * WAVE, lds_e, ctx->stream>>>(a);
      if (ctx->kernel_timing) {
        HIP_OK(hipEventRecord(tb, ctx->stream));
        timed.emplace_back(ta, tb);
      }
      if (bal) {
        const int ntiles = (nslots + WAVE - 1) / WAVE;
        const int g_tiles = std::max(1, std::min((ntiles + 3) / 4, 4096));
        const size_t lds_h = (size_t)(K + 2) * sizeof(int) + 4 * (size_t)K * sizeof(long long);
        HIP_OK(hipMemcpyAsync(ctx->sizes_frozen, ctx->sizes, sizeof(long long) * (size_t)K,
                              hipMemcpyDeviceToDevice, ctx->stream));
        k_bal_hist<<<g_tiles, 256, lds_h, ctx->stream>>>(ctx->slot_dec, ctx->order, ctx->vwgt, r0, nslots, K,
                                                         ctx->sizes_frozen, ctx->cand_tot, sched->cap,
                                                         ctx->over_idx, ctx->over_count, ctx->tile_sum, ntiles);
        k_bal_scan<<<std::max(1, std::min(K, 1024)), 256, 0, ctx->stream>>>(ctx->tile_sum, ntiles,
                                                                            ctx->over_count, ctx->cand_tot, K);
        k_bal_admit<<<g_tiles, 256, 0, ctx->stream>>>(ctx->slot_dec, ctx->slot_own, ctx->order, ctx->vwgt, r0,
                                                      nslots, ctx->sizes_frozen, sched->cap, ctx->over_idx,
                                                      ctx->over_count, ctx->tile_sum, ntiles, ctx->part,
                                                      ctx->sizes, ctx->vtx_ptr, ctx->vtx_nets, ctx->tag,
                                                      ctx->net_stamp, ctx->touched, ctx->counters + par,
                                                      ctx->moves);
      } else {
        k_ref_commit<<<(nslots + 255) / 256, 256, 0, ctx->stream>>>(ctx->slot_own, ctx->slot_dec, ctx->order, r0,
                                                                    nslots, ctx->part, ctx->sizes, ctx->moves);
      }
      ApplyArgs ap{};
      ap.net_ptr = ctx->net_ptr;
      ap.apply_pins = ctx->apply_pins;
      ap.slot_own = ctx->slot_own;
      ap.slot_dec = ctx->slot_dec;
      ap.touched = ctx->touched;
      ap.tcount = ctx->counters + par;
      ap.lam = ctx->lam;
      ap.ent = ctx->ent;
      ap.K = K;
      ap.pb = ctx->pb;
      ap.pmask = ctx->pmask;
      ap.r0 = r0;
      ap.r1 = r0 + nslots;
      k_apply<<<apply_grid, APPLY_BLOCK, lds_a, ctx->stream>>>(ap);
      rounds++;
    }
  }
  HIP_OK(hipGetLastError());
  HIP_OK(hipEventRecord(e1, ctx->stream));
  unsigned long long mv = 0;
  TRY(d2h(ctx, &mv, ctx->moves, 1));
  HIP_OK(hipStreamSynchronize(ctx->stream));
  float ms = 0;
  HIP_OK(hipEventElapsedTime(&ms, e0, e1));
  for (auto& pr : timed) {
    float t = 0;
    HIP_OK(hipEventElapsedTime(&t, pr.first, pr.second));
    ctx->timed_ms += t;
  }
  ctx->timed_launches = (long long)timed.size();
  if (stats) {
    stats->moves = (long long)mv;
    stats->rounds = rounds;
    stats->pin_visits = (long long)passes * ctx->pins;
    stats->device_ms = ms;
  }
  return FP_OK;
}

fp_status fp_eval_moves(fp_ctx* ctx, fp_metric metric, const fp_schedule* sched, int* dec, int* gain) {
  if (!ctx) return FP_ERR_ARG;
  if (!ctx->has_state) return fail(ctx, FP_ERR_STATE, "fp_eval_moves before a state was set");
  TRY(check_sched(ctx, sched));
  if (metric != FP_METRIC_TV) return fail(ctx, FP_ERR_UNSUPPORTED, "move evaluation for metric %d not available", (int)metric);
  HIP_OK(hipSetDevice(ctx->device));
  const int nov = ctx->nov, K = ctx->K;
  TRY(ensure_round_scratch(ctx, nov));
  int *d_dec = nullptr, *d_gain = nullptr;
  HIP_OK(hipMalloc(&d_dec, sizeof(int) * (size_t)nov));
  if (hipMalloc(&d_gain, sizeof(int) * (size_t)nov) != hipSuccess) {
    (void)hipFree(d_dec);
    return fail(ctx, FP_ERR_OOM, "hipMalloc failed");
  }
  const int waves = eval_waves(K);
  EvalArgs a = make_eval_args(ctx, sched);
  a.r0 = 0;
  a.nslots = nov;
  a.dec_out = d_dec;
  a.gain_out = d_gain;
  k_tv_eval<<<eval_grid(nov, waves), waves * WAVE, eval_lds(K, waves), ctx->stream>>>(a);
  fp_status st = FP_OK;
  if (hipGetLastError() != hipSuccess) st = fail(ctx, FP_ERR_HIP, "k_tv_eval launch failed");
  if (st == FP_OK && dec) st = d2h(ctx, dec, d_dec, (size_t)nov);
  if (st == FP_OK && gain) st = d2h(ctx, gain, d_gain, (size_t)nov);
  if (hipStreamSynchronize(ctx->stream) != hipSuccess && st == FP_OK) st = fail(ctx, FP_ERR_HIP, "sync failed");
  (void)hipFree(d_dec);
  (void)hipFree(d_gain);
  return st;
}

fp_status fp_eval(fp_ctx* ctx, fp_metrics* out, int* lam, long long* sv, long long* rv, long long* sm, long long* rm,
                  long long* sv_msv) {
  if (!ctx) return FP_ERR_ARG;
  if (!ctx->has_state) return fail(ctx, FP_ERR_STATE, "fp_eval before a state was set");
  HIP_OK(hipSetDevice(ctx->device));
  const int K = ctx->K;
  int* d_lam = nullptr;
  HIP_OK(hipMalloc(&d_lam, sizeof(int) * (size_t)std::max(ctx->noe, 1)));
  fill_async<unsigned>(ctx, ctx->p2p, ((size_t)K * (size_t)K + 31) / 32, 0u);
  fill_async<unsigned long long>(ctx, ctx->mvec, (size_t)5 * (size_t)K + 2, 0ull);
  MetricArgs m{};
  m.net_ptr = ctx->net_ptr;
  m.net_pins = ctx->net_pins;
  m.part = ctx->part;
  m.src = ctx->src;
  m.noe = ctx->noe;
  m.K = K;
  m.nov = ctx->nov;
  m.lam = d_lam;
  m.p2p = ctx->p2p;
  m.send = ctx->mvec;
  m.recv = ctx->mvec + K;
  m.sm = ctx->mvec + 2 * (size_t)K;
  m.rm = ctx->mvec + 3 * (size_t)K;
  m.send_msv = ctx->mvec + 4 * (size_t)K;
  m.tvcn = ctx->mvec + 5 * (size_t)K;
  const int waves = METRIC_BLOCK / WAVE;
  const size_t lds = (size_t)3 * K * sizeof(unsigned long long) + (size_t)waves * K * sizeof(int);
  if (lds > (size_t)LDS_BUDGET) {
    (void)hipFree(d_lam);
    return fail(ctx, FP_ERR_UNSUPPORTED, "K too large for the metric kernel");
  }
  const int grid = std::max(1, std::min((ctx->noe + waves - 1) / waves, 4096));
  if (ctx->noe > 0) k_metrics<<<grid, METRIC_BLOCK, lds, ctx->stream>>>(m);
  std::vector<unsigned long long> h((size_t)5 * K + 2);
  fp_status st = FP_OK;
  if (hipGetLastError() != hipSuccess) st = fail(ctx, FP_ERR_HIP, "k_metrics launch failed");
  if (st == FP_OK) st = d2h(ctx, h.data(), ctx->mvec, h.size());
  if (st == FP_OK && lam) st = d2h(ctx, lam, d_lam, (size_t)ctx->noe);
  if (hipStreamSynchronize(ctx->stream) != hipSuccess && st == FP_OK) st = fail(ctx, FP_ERR_HIP, "sync failed");
  (void)hipFree(d_lam);
  if (st != FP_OK) return st;
  const long long* send = reinterpret_cast<const long long*>(h.data());
  const long long* recv = send + K;
  const long long* smv = send + 2 * (size_t)K;
  const long long* rmv = send + 3 * (size_t)K;
  const long long* msv = send + 4 * (size_t)K;
  if (out) {
    out->tv = (long long)h[5 * (size_t)K];
    out->cn = (long long)h[5 * (size_t)K + 1];
    long long tm = 0, m_sm = -1, m_sv = -1, m_srv = -1;
    for (int k = 0; k < K; k++) {
      tm += smv[k];
      m_sm = std::max(m_sm, smv[k]);
      m_sv = std::max(m_sv, msv[k]);
      m_srv = std::max(m_srv, send[k] + recv[k]);
    }
    out->tm = tm;
    out->msm = m_sm;
    out->msv = m_sv;
    out->msrv = m_srv;
  }
  if (sv) std::copy(send, send + K, sv);
  if (rv) std::copy(recv, recv + K, rv);
  if (sm) std::copy(smv, smv + K, sm);
  if (rm) std::copy(rmv, rmv + K, rm);
  if (sv_msv) std::copy(msv, msv + K, sv_msv);
  return FP_OK;
}

}  // extern "C"
