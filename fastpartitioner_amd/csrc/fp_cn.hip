// fp_cn.hip -- CN (cut-net) refinement on the device: the lambda = 2 push of
// CN/partition.cpp:618-779 with one pass per round, evaluated on the state
// frozen at the pass start and applied in (net, pin) order, i.e. the
// reference's partition_par with one thread (CN/partition.cpp:436-836).
//
// A pass:
//   k_cn_eval     one wavefront filters 64 nets at a time; every lambda = 2
//                 net picks old = the part with the smaller pin count (ties:
//                 the second entry) and is accepted iff none of its old-side
//                 pins lies on an uncut net (CN/partition.cpp:650-690)
//   scan          push offsets: exclusive scan of the accepted nets' old-side
//                 pin counts, in net order
//   k_cn_pushes   push list in (net, pin) order
//   k_cn_prep     per push: its source part (the new part of the vertex's
//                 previous pushing net, else its part at pass start -- a
//                 vertex can be pushed by several nets) and the ops it puts
//                 on each of the vertex's nets
//   k_cn_alloc / k_cn_scatter   per touched net, the op list (push indices)
//   k_cn_bigsort  nets with more than CN_MAXOPS ops: their push indices in
//                 order through a bitmap over the pass's pushes (one block)
//   k_cn_apply_small / k_cn_apply   one wavefront per touched net: sorts
//                 its ops by push index (up to 64 in registers, up to
//                 CN_MAXOPS in LDS; a big net's ordered ops are read) and
//                 replays remove/add on the list
//   k_cn_commit   part vector and part sizes (CN/partition.cpp:786-789)
#include "fp_internal.h"

namespace fpe {
namespace {

constexpr int CN_MAXOPS = 4096;  // ops per net and pass sorted in LDS; more go through k_cn_bigsort
constexpr int SCAN_BLOCK = 1024;
constexpr int CN_EVAL_GROUP = 4;  // a pin's nets whose lambdas k_cn_eval loads together

struct CnEvalArgs {
  const unsigned* net_ptr;
  const unsigned* net_pins;
  const unsigned* vtx_ptr;
  const unsigned* vtx_seg;
  const unsigned* ent;
  const unsigned short* ev;
  const int* part;
  int noe, pb;
  unsigned pmask;
  int* cn_old;  // accepted: old part; else -1
  int* cn_new;
  int* npush;
};

__global__ void __launch_bounds__(256) k_cn_eval(CnEvalArgs a) {
  const int lane = lane_id(), wave = wave_id(), nw = (int)(blockDim.x / WAVE);
  for (int n0 = ((int)blockIdx.x * nw + wave) * WAVE; n0 < a.noe; n0 += (int)gridDim.x * nw * WAVE) {
    const int n = n0 + lane;
    bool two = false;
    if (n < a.noe) {
      two = a.ev[a.net_ptr[n] + (unsigned)n] == 2;
      if (!two) { a.cn_old[n] = -1; a.npush[n] = 0; }
    }
    unsigned long long m = __ballot(two);
    while (m) {
      const int nn = n0 + __ffsll((long long)m) - 1;
      m &= m - 1;
      const unsigned s = a.net_ptr[nn], e = a.net_ptr[nn + 1];
      const unsigned seg = s + (unsigned)nn;
      const unsigned x1 = a.ent[seg + 1], x2 = a.ent[seg + 2];
      const int p1 = (int)(x1 & a.pmask), c1 = (int)(x1 >> a.pb);
      const int p2 = (int)(x2 & a.pmask), c2 = (int)(x2 >> a.pb);
      const int oldp = c1 < c2 ? p1 : p2, newp = c1 < c2 ? p2 : p1;
      bool bad = false;
      int cnt = 0;
      for (unsigned k0 = s; k0 < e && !bad; k0 += WAVE) {
        const unsigned k = k0 + (unsigned)lane;
        bool isold = false, b = false;
        if (k < e) {
          const int v = (int)a.net_pins[k];
          isold = a.part[v] == oldp;
          if (isold) {
            // v's nets CN_EVAL_GROUP at a time: their segment offsets load
            // together, then their lambdas, instead of one dependent pair of
            // loads per net
            const unsigned jb = a.vtx_ptr[v], je = a.vtx_ptr[v + 1];
            for (unsigned j = jb; j < je && !b; j += CN_EVAL_GROUP) {
              unsigned sg[CN_EVAL_GROUP];
#pragma unroll
              for (int u = 0; u < CN_EVAL_GROUP; u++) sg[u] = a.vtx_seg[min(j + (unsigned)u, je - 1)];
#pragma unroll
              for (int u = 0; u < CN_EVAL_GROUP; u++) b |= a.ev[sg[u]] == 1;
            }
          }
        }
        cnt += __popcll(__ballot(isold));
        bad = __ballot(b) != 0;
      }
      if (lane == 0) {
        a.cn_old[nn] = bad ? -1 : oldp;
        a.cn_new[nn] = newp;
        a.npush[nn] = bad ? 0 : cnt;
      }
    }
  }
}

// exclusive scan of an int array, three launches
__global__ void __launch_bounds__(SCAN_BLOCK) k_scan_sums(const int* __restrict__ x, int n, long long* __restrict__ sums) {
  __shared__ long long scratch[SCAN_BLOCK / WAVE];
  const int i = (int)(blockIdx.x * SCAN_BLOCK + threadIdx.x);
  long long tot;
  (void)block_excl_scan<long long>(i < n ? x[i] : 0, scratch, &tot);
  if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(SCAN_BLOCK) k_scan_top(long long* __restrict__ sums, int nblocks,
                                                         long long* __restrict__ total) {
  __shared__ long long scratch[SCAN_BLOCK / WAVE];
  __shared__ long long carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int b0 = 0; b0 < nblocks; b0 += SCAN_BLOCK) {
    const int b = b0 + (int)threadIdx.x;
    const long long v = b < nblocks ? sums[b] : 0;
    long long tot;
    const long long ex = block_excl_scan(v, scratch, &tot);
    if (b < nblocks) sums[b] = carry + ex;
    __syncthreads();
    if (threadIdx.x == 0) carry += tot;
    __syncthreads();
  }
  if (threadIdx.x == 0) *total = carry;
}

__global__ void __launch_bounds__(SCAN_BLOCK) k_scan_apply(const int* __restrict__ x, int n,
                                                           const long long* __restrict__ sums,
                                                           long long* __restrict__ out) {
  __shared__ long long scratch[SCAN_BLOCK / WAVE];
  const int i = (int)(blockIdx.x * SCAN_BLOCK + threadIdx.x);
  long long tot;
  const long long ex = block_excl_scan<long long>(i < n ? x[i] : 0, scratch, &tot);
  if (i < n) out[i] = sums[blockIdx.x] + ex;
}

// the pushes of every accepted net, at its offset, in pin order
__global__ void __launch_bounds__(256) k_cn_pushes(const unsigned* __restrict__ net_ptr,
                                                   const unsigned* __restrict__ net_pins, const int* __restrict__ part,
                                                   const int* __restrict__ cn_old, const long long* __restrict__ pbase,
                                                   int noe, int* __restrict__ mv_v, int* __restrict__ mv_net) {
  const int lane = lane_id(), wave = wave_id(), nw = (int)(blockDim.x / WAVE);
  for (int n0 = ((int)blockIdx.x * nw + wave) * WAVE; n0 < noe; n0 += (int)gridDim.x * nw * WAVE) {
    const int n = n0 + lane;
    unsigned long long m = __ballot(n < noe && cn_old[n] >= 0);
    while (m) {
      const int nn = n0 + __ffsll((long long)m) - 1;
      m &= m - 1;
      const int oldp = cn_old[nn];
      long long pos = pbase[nn];
      for (unsigned k0 = net_ptr[nn]; k0 < net_ptr[nn + 1]; k0 += WAVE) {
        const unsigned k = k0 + (unsigned)lane;
        int v = -1;
        bool isold = false;
        if (k < net_ptr[nn + 1]) {
          v = (int)net_pins[k];
          isold = part[v] == oldp;
        }
        const unsigned long long b = __ballot(isold);
        if (isold) {
          const long long at = pos + __popcll(b & ((1ull << lane) - 1ull));
          mv_v[at] = v;
          mv_net[at] = nn;
        }
        pos += __popcll(b);
      }
    }
  }
}

struct CnPrepArgs {
  const int* mv_v;
  const int* mv_net;
  const unsigned long long* nmoves;  // device: the pass's push count
  const unsigned* vtx_ptr;
  const unsigned* vtx_nets;
  const int* part;
  const int* cn_old;
  const int* cn_new;
  const long long* pbase;  // first push of each accepted net
  const int* npush;        // pushes of each accepted net
  int* mv_from;
  int* mv_last;
  int* touched;
  int* tcount;
  int* opcnt;
  unsigned long long* totals;  // [0] ops
};

// Source part of each push: a vertex v can be pushed by several nets of the
// pass (every accepted net whose old part is v's part at pass start); pushes
// apply in net order, so the source of the push by net n is the new part of
// the largest such net below n, else v's part at pass start.
// The touched-net appends go through a per-wave LDS buffer (one atomic on
// the list's counter per full buffer, one per workgroup at the end) and the
// op total is summed per block: same-address atomics whose returns the waves
// wait for queue at one L2 slice. A wave takes 64 consecutive pushes and
// walks their net lists in step, so that its appends stay wave-uniform.
__global__ void __launch_bounds__(256) k_cn_prep(CnPrepArgs a) {
  __shared__ unsigned long long s_ops;
  __shared__ unsigned s_buf[256 / WAVE][TOUCH_BUF];
  __shared__ int s_tn[256 / WAVE + 1];
  if (threadIdx.x == 0) s_ops = 0;
  __syncthreads();
  const int lane = lane_id(), wave = wave_id();
  WaveTouch st{s_buf[wave], 0};
  unsigned long long ops = 0;
  const long long nm = (long long)*a.nmoves;
  for (long long m0 = (long long)blockIdx.x * blockDim.x + (long long)wave * WAVE; m0 < nm;
       m0 += (long long)gridDim.x * blockDim.x) {
    const long long m = m0 + lane;
    const bool act = m < nm;
    int v = 0, n = -1;
    unsigned vb = 0, deg = 0;
    if (act) {
      v = a.mv_v[m];
      n = a.mv_net[m];
      const int p0 = a.part[v];
      int prev = -1;
      bool last = true;
      vb = a.vtx_ptr[v];
      deg = a.vtx_ptr[v + 1] - vb;
      // v's nets CN_EVAL_GROUP at a time: ids, then their accepted old parts
      for (unsigned j = vb; j < vb + deg; j += CN_EVAL_GROUP) {
        int kk[CN_EVAL_GROUP], co[CN_EVAL_GROUP];
#pragma unroll
        for (int u = 0; u < CN_EVAL_GROUP; u++) kk[u] = (int)a.vtx_nets[min(j + (unsigned)u, vb + deg - 1)];
#pragma unroll
        for (int u = 0; u < CN_EVAL_GROUP; u++) co[u] = a.cn_old[kk[u]];
#pragma unroll
        for (int u = 0; u < CN_EVAL_GROUP; u++) {
          const int k = kk[u];
          if (j + (unsigned)u < vb + deg && k != n && co[u] == p0) {
            if (k < n) prev = max(prev, k);
            else last = false;
          }
        }
      }
      a.mv_from[m] = prev >= 0 ? a.cn_new[prev] : p0;
      a.mv_last[m] = last ? 1 : 0;
      ops += deg;
    }
    const int dmax = wave_max((int)deg);
    for (int jj = 0; jj < dmax; jj++) {
      // the net's first op of the pass puts it on the touched list (op counts
      // are 0 between passes: the apply resets them). The pushing net's own
      // ops are its pushes (its pins are distinct: refinement rejects repeated
      // pins), added once by its first push instead of one atomic per push on
      // one word
      bool isnew = false;
      unsigned t = 0;
      if (act && (unsigned)jj < deg) {
        t = a.vtx_nets[vb + (unsigned)jj];
        if ((int)t != n) isnew = atomicAdd(&a.opcnt[t], 1) == 0;
        else if (m == a.pbase[n]) isnew = atomicAdd(&a.opcnt[t], a.npush[n]) == 0;
      }
      wave_stage_append(isnew, t, st, a.touched, a.tcount);
    }
  }
  block_touch_flush(st, a.touched, a.tcount, s_tn);
  if (ops) atomicAdd(&s_ops, ops);
  __syncthreads();
  if (threadIdx.x == 0 && s_ops) atomicAdd(&a.totals[0], s_ops);
}

// The pass's edit lists must fit: total ops <= cap. A net's ops are distinct
// push indices unless a pin repeats in a net (dup_pins): then a net with more
// than CN_MAXOPS ops cannot be ordered by k_cn_bigsort's bitmap. Otherwise
// flags[0] = the total needed, flags[1] = 1, and the rest of the pass leaves
// the state unchanged.
__global__ void k_cn_check(const unsigned long long* __restrict__ totals, long long cap, int dup_pins,
                           int* __restrict__ flags) {
  const unsigned long long need = totals[0];
  const bool ops_over = need > (unsigned long long)cap;
  const bool net_over = dup_pins && totals[2] > (unsigned long long)CN_MAXOPS;
  flags[0] = ops_over ? (int)min(need, (unsigned long long)INT_MAX) : 0;
  flags[1] = net_over ? 1 : 0;
}

// Each touched net's op range: the block's nets take consecutive ranges
// (a block scan), the block's start comes from one atomic on the running total.
__global__ void __launch_bounds__(256) k_cn_alloc(const int* __restrict__ touched, const int* __restrict__ tcount,
                                                  const int* __restrict__ opcnt, const int* __restrict__ cn_old,
                                                  const int* __restrict__ npush, long long* __restrict__ opoff,
                                                  int* __restrict__ opfill, unsigned long long* __restrict__ totals,
                                                  int* __restrict__ big, long long big_cap) {
  __shared__ long long s_w[256 / WAVE];
  __shared__ long long s_base;
  __shared__ int s_max;
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  const bool valid = i < *tcount;
  const int t = valid ? touched[i] : 0;
  const int c = valid ? opcnt[t] : 0;
  if (threadIdx.x == 0) s_max = 0;
  long long tot;
  const long long ex = block_excl_scan((long long)c, s_w, &tot);
  if (c) atomicMax(&s_max, c);
  if (threadIdx.x == 0) s_base = tot ? (long long)atomicAdd(&totals[1], (unsigned long long)tot) : 0;
  __syncthreads();
  if (valid) {
    opoff[t] = s_base + ex;
    // an accepted net's own pushes take the first slots of its range, by push index
    opfill[t] = cn_old[t] >= 0 ? npush[t] : 0;
    // too many ops to sort in LDS: the net goes on the big list (totals[4])
    if (c > CN_MAXOPS) {
      const unsigned long long k = atomicAdd(&totals[4], 1ull);
      if ((long long)k < big_cap) big[k] = t;
    }
  }
  if (threadIdx.x == 0 && s_max) atomicMax(reinterpret_cast<unsigned long long*>(&totals[2]), (unsigned long long)s_max);
}

__global__ void k_cn_scatter(const int* __restrict__ mv_v, const int* __restrict__ mv_net,
                             const long long* __restrict__ pbase, const unsigned long long* __restrict__ nmoves,
                             const int* __restrict__ flags, const unsigned* __restrict__ vtx_ptr,
                             const unsigned* __restrict__ vtx_nets, const long long* __restrict__ opoff,
                             int* __restrict__ opfill, int* __restrict__ ops) {
  if (flags[0] | flags[1]) return;
  const long long nm = (long long)*nmoves;
  for (long long m = (long long)blockIdx.x * blockDim.x + threadIdx.x; m < nm; m += (long long)gridDim.x * blockDim.x) {
    const int v = mv_v[m], n = mv_net[m];
    const unsigned jb = vtx_ptr[v], je = vtx_ptr[v + 1];
    // CN_EVAL_GROUP nets at a time: their ids, then their list offsets and
    // slots (independent loads and atomics in flight together)
    for (unsigned j = jb; j < je; j += CN_EVAL_GROUP) {
      unsigned t[CN_EVAL_GROUP];
      long long off[CN_EVAL_GROUP], slot[CN_EVAL_GROUP];
#pragma unroll
      for (int u = 0; u < CN_EVAL_GROUP; u++) t[u] = vtx_nets[min(j + (unsigned)u, je - 1)];
#pragma unroll
      for (int u = 0; u < CN_EVAL_GROUP; u++) {
        off[u] = opoff[t[u]];
        slot[u] = 0;
        if (j + (unsigned)u < je) slot[u] = (int)t[u] == n ? m - pbase[n] : (long long)atomicAdd(&opfill[t[u]], 1);
      }
#pragma unroll
      for (int u = 0; u < CN_EVAL_GROUP; u++)
        if (j + (unsigned)u < je) ops[off[u] + slot[u]] = (int)m;
    }
  }
}

// The big nets' op lists in push order: for each net on the big list, its
// push indices set bits of a bitmap over the pass's pushes, then the block
// writes the set bits back in order (each thread a run of words, a block scan
// of the runs' counts). One block; the big nets one after another.
constexpr int BIGSORT_BLOCK = 1024;

__global__ void __launch_bounds__(BIGSORT_BLOCK) k_cn_bigsort(const int* __restrict__ big, long long big_cap,
                                                              const unsigned long long* __restrict__ totals,
                                                              const unsigned long long* __restrict__ nmoves,
                                                              const int* __restrict__ flags, const int* __restrict__ opcnt,
                                                              const long long* __restrict__ opoff,
                                                              unsigned* __restrict__ bits, int* __restrict__ ops) {
  __shared__ long long s_w[BIGSORT_BLOCK / WAVE];
  if (flags[0] | flags[1]) return;
  const int nbig = (int)min(totals[4], (unsigned long long)big_cap);
  const long long W = ((long long)*nmoves + 31) / 32;
  const int tid = (int)threadIdx.x;
  const long long per = (W + BIGSORT_BLOCK - 1) / BIGSORT_BLOCK;
  const long long w0 = min(W, (long long)tid * per), w1 = min(W, w0 + per);
  for (int b = 0; b < nbig; b++) {
    const int t = big[b];
    const int c = opcnt[t];
    const long long off = opoff[t];
    for (long long w = tid; w < W; w += BIGSORT_BLOCK) bits[w] = 0u;
    __syncthreads();
    for (int j = tid; j < c; j += BIGSORT_BLOCK) {
      const int m = ops[off + j];
      atomicOr(&bits[m >> 5], 1u << (m & 31));
    }
    __syncthreads();
    long long cnt = 0;
    for (long long w = w0; w < w1; w++) cnt += __popc(bits[w]);
    long long tot;
    long long at = block_excl_scan(cnt, s_w, &tot);
    for (long long w = w0; w < w1; w++) {
      unsigned x = bits[w];
      while (x) {
        const int k = __ffs((int)x) - 1;
        x &= x - 1u;
        ops[off + at++] = (int)(w * 32 + k);
      }
    }
    __syncthreads();
  }
}

struct CnApplyArgs {
  const unsigned* net_ptr;
  const int* touched;
  const int* tcount;
  const int* opcnt;
  const long long* opoff;
  const int* ops;
  const int* mv_from;
  const int* mv_net;
  const int* cn_new;
  unsigned* ent;
  unsigned short* ev;
  int K, pb;
  unsigned pmask;
  int* opcnt_reset;
  const int* flags;  // overflow: only reset the op counts
};

// ascending sort of keys[0..P) in LDS by one wavefront (P a power of two)
__device__ void wave_bitonic_sort(unsigned* keys, int P) {
  for (int k = 2; k <= P; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = lane_id(); i < P; i += WAVE) {
        const int l = i ^ j;
        if (l > i) {
          const unsigned x = keys[i], y = keys[l];
          const bool up = (i & k) == 0;
          if ((x > y) == up) { keys[i] = y; keys[l] = x; }
        }
      }
      wave_lds_sync();
    }
  }
}

// ascending sort of one key per lane across the wavefront (bitonic network
// over lane exchanges)
__device__ __forceinline__ unsigned wave_sort_lanes(unsigned x) {
  const int lane = lane_id();
  for (int k = 2; k <= WAVE; k <<= 1)
    for (int j = k >> 1; j > 0; j >>= 1) {
      const unsigned y = (unsigned)__shfl_xor((int)x, j);
      // the lower lane of a pair keeps the minimum in an ascending block
      x = (((lane & j) == 0) == ((lane & k) == 0)) ? min(x, y) : max(x, y);
    }
  return x;
}

// The touched nets with at most 64 ops (most of them), one wavefront per net,
// four per block: the ops are sorted in registers, so the LDS holds only the
// list (and its index), and more waves fit a CU than in k_cn_apply, whose
// sort buffer they do not need. In an overflowed pass (flags) every touched
// net's op count is reset and nothing else is done.
template <bool INDEXED>
__global__ void __launch_bounds__(256) k_cn_apply_small(CnApplyArgs a) {
  extern __shared__ __align__(16) unsigned char lds[];
  const int lane = lane_id(), wave = wave_id(), nw = (int)(blockDim.x / WAVE);
  unsigned* list = reinterpret_cast<unsigned*>(lds) + (size_t)wave * (size_t)a.K;
  unsigned short* at = reinterpret_cast<unsigned short*>(reinterpret_cast<unsigned*>(lds) + (size_t)nw * (size_t)a.K) +
                       (size_t)wave * (size_t)a.K;
  if (INDEXED) wave_index_init(at, a.K);
  const int nt = *a.tcount;
  const bool skip = (a.flags[0] | a.flags[1]) != 0;
  for (int i = (int)blockIdx.x * nw + wave; i < nt; i += (int)gridDim.x * nw) {
    const int t = a.touched[i];
    if (skip) {
      if (lane == 0) a.opcnt_reset[t] = 0;
      continue;
    }
    const int c = a.opcnt[t];
    if (c > WAVE) continue;  // k_cn_apply's
    const long long off = a.opoff[t];
    const unsigned m = wave_sort_lanes(lane < c ? (unsigned)a.ops[off + lane] : 0xFFFFFFFFu);
    int from_l = 0, to_l = 0;
    if (lane < c) {
      from_l = a.mv_from[m];
      to_l = a.cn_new[a.mv_net[m]];
    }
    const unsigned s0 = a.net_ptr[t], seg = s0 + (unsigned)t;
    int L = wave_list_load(a.ent, seg, a.net_ptr[t + 1] - s0, list);
    if (INDEXED) wave_index_set(list, L, at, a.pmask);
    for (int l = 0; l < c; l++) {
      const int p = __builtin_amdgcn_readlane(from_l, l), q = __builtin_amdgcn_readlane(to_l, l);
      if (INDEXED) {
        unsigned co;
        bool found;
        wave_ilist_move(list, at, L, p, q, a.pb, a.pmask, co, found);
      } else {
        wave_list_move(list, L, p, q, a.pb, a.pmask);
      }
    }
    wave_list_store(a.ent, a.ev, seg, list, L, a.pb, a.pmask);
    if (INDEXED) wave_index_clear(list, L, at, a.pmask);
    if (lane == 0) a.opcnt_reset[t] = 0;
    wave_lds_sync();
  }
}

// INDEXED: a part -> position index of the list in LDS (2 bytes per part
// after the keys), for K where it fits.
template <bool INDEXED>
__global__ void __launch_bounds__(WAVE) k_cn_apply(CnApplyArgs a) {
  extern __shared__ __align__(16) unsigned char lds[];
  unsigned* list = reinterpret_cast<unsigned*>(lds);
  unsigned* keys = list + a.K;
  unsigned short* at = reinterpret_cast<unsigned short*>(keys + CN_MAXOPS);
  const int lane = lane_id();
  if (INDEXED) wave_index_init(at, a.K);
  const int nt = *a.tcount;
  if ((a.flags[0] | a.flags[1]) != 0) return;  // k_cn_apply_small reset the op counts
  for (int i = (int)blockIdx.x; i < nt; i += (int)gridDim.x) {
    const int t = a.touched[i];
    const int c = a.opcnt[t];
    if (c <= WAVE) continue;  // k_cn_apply_small's (reset to 0 there)
    const long long off = a.opoff[t];
    // up to CN_MAXOPS ops are sorted here; a big net's ops come in order
    // from k_cn_bigsort and are read straight from the op list
    const bool small = c <= CN_MAXOPS;
    if (small) {
      int P = 1;
      while (P < c) P <<= 1;
      for (int j = lane; j < P; j += WAVE) keys[j] = j < c ? (unsigned)a.ops[off + j] : 0xFFFFFFFFu;
      wave_lds_sync();
      if (c > 1) wave_bitonic_sort(keys, P);
    }
    const unsigned s0 = a.net_ptr[t], seg = s0 + (unsigned)t;
    int L = wave_list_load(a.ent, seg, a.net_ptr[t + 1] - s0, list);
    if (INDEXED) wave_index_set(list, L, at, a.pmask);
    for (int j0 = 0; j0 < c; j0 += WAVE) {
      // 64 moves at a time: each lane loads one move's old and new part
      int from_l = 0, to_l = 0;
      if (j0 + lane < c) {
        const int m = small ? (int)keys[j0 + lane] : a.ops[off + j0 + lane];
        from_l = a.mv_from[m];
        to_l = a.cn_new[a.mv_net[m]];
      }
      const int nj = min(WAVE, c - j0);
      for (int l = 0; l < nj; l++) {
        const int p = __builtin_amdgcn_readlane(from_l, l), q = __builtin_amdgcn_readlane(to_l, l);
        if (INDEXED) {
          unsigned co;
          bool found;
          wave_ilist_move(list, at, L, p, q, a.pb, a.pmask, co, found);
        } else {
          wave_list_move(list, L, p, q, a.pb, a.pmask);
        }
      }
    }
    wave_list_store(a.ent, a.ev, seg, list, L, a.pb, a.pmask);
    if (INDEXED) wave_index_clear(list, L, at, a.pmask);
    if (lane == 0) a.opcnt_reset[t] = 0;
    wave_lds_sync();
  }
}

// LDS of k_cn_apply_small: per wave a K-entry list and, where four waves'
// fit, the part -> position index; fewer waves per block for large K
bool cn_small_indexed(int K) {
  return (size_t)4 * (size_t)K * (sizeof(unsigned) + sizeof(unsigned short)) <= (size_t)LDS_BUDGET;
}
int cn_small_waves(int K) {
  if (cn_small_indexed(K)) return 4;
  return (int)std::max<size_t>(1, std::min<size_t>(4, (size_t)LDS_BUDGET / ((size_t)K * sizeof(unsigned))));
}
size_t cn_small_lds(int K) {
  return (size_t)cn_small_waves(K) * (size_t)K *
         (sizeof(unsigned) + (cn_small_indexed(K) ? sizeof(unsigned short) : 0));
}

// the CN apply's LDS: list, sort keys, and the part -> position index where
// it fits the budget
bool cn_apply_indexed(int K) {
  return (size_t)K * (sizeof(unsigned) + sizeof(unsigned short)) + (size_t)CN_MAXOPS * sizeof(unsigned) <=
         (size_t)LDS_BUDGET;
}

size_t cn_apply_lds(int K) {
  return (size_t)K * sizeof(unsigned) + (size_t)CN_MAXOPS * sizeof(unsigned) +
         (cn_apply_indexed(K) ? (size_t)K * sizeof(unsigned short) : 0);
}

// AGG: the block sums its size changes per part in LDS and adds them to the
// sizes with one atomic per part (a pass's pushes otherwise queue on the K
// size words at a few L2 slices); launched with few blocks, each covering
// many pushes.
template <bool AGG>
__global__ void __launch_bounds__(256) k_cn_commit(const int* __restrict__ mv_v, const int* __restrict__ mv_net,
                                                   const int* __restrict__ mv_from, const int* __restrict__ mv_last,
                                                   const unsigned long long* __restrict__ nmoves,
                                                   int* __restrict__ flags, const int* __restrict__ cn_new,
                                                   const int* __restrict__ vwgt, int K, int* __restrict__ part,
                                                   long long* __restrict__ sizes, unsigned long long* __restrict__ moves) {
  extern __shared__ __align__(16) unsigned char lds[];
  unsigned long long* s_dsz = reinterpret_cast<unsigned long long*>(lds);
  if (flags[0] | flags[1]) return;
  if (AGG) {
    for (int q = (int)threadIdx.x; q < K; q += (int)blockDim.x) s_dsz[q] = 0;
    __syncthreads();
  }
  const long long nm = (long long)*nmoves;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    atomicAdd(moves, (unsigned long long)nm);
    atomicAdd(&flags[2], 1);  // passes applied since the host last cleared it
  }
  for (long long m = (long long)blockIdx.x * blockDim.x + threadIdx.x; m < nm; m += (long long)gridDim.x * blockDim.x) {
    const int v = mv_v[m];
    const int to = cn_new[mv_net[m]];
    const long long w = vwgt ? vwgt[v] : 1;
    if (AGG) {
      atomicAdd(&s_dsz[mv_from[m]], (unsigned long long)-w);
      atomicAdd(&s_dsz[to], (unsigned long long)w);
    } else {
      atomic_add_i64(&sizes[mv_from[m]], -w);
      atomic_add_i64(&sizes[to], w);
    }
    if (mv_last[m]) part[v] = to;
  }
  if (AGG) {
    __syncthreads();
    for (int q = (int)threadIdx.x; q < K; q += (int)blockDim.x) {
      const long long x = (long long)s_dsz[q];
      if (x) atomic_add_i64(&sizes[q], x);
    }
  }
}

}  // namespace

fp_status plan_cn(fp_ctx* ctx, RefinePlan& p) {
  const int noe = ctx->noe;
  const long long pins = ctx->pins;
  const int nblk = (noe + SCAN_BLOCK - 1) / SCAN_BLOCK;
  auto& s = ctx->cn;
  if (cn_apply_lds(ctx->K) > (size_t)LDS_BUDGET)
    return fail(ctx, FP_ERR_UNSUPPORTED, "K = %d too large for the CN apply", ctx->K);
  if (!s.ready) {
    const size_t ne = (size_t)std::max(noe, 1), np = (size_t)std::max(pins, 1ll);
    FP_TRY(dalloc(ctx, &s.old_part, ne));
    FP_TRY(dalloc(ctx, &s.new_part, ne));
    FP_TRY(dalloc(ctx, &s.npush, ne));
    FP_TRY(dalloc(ctx, &s.pbase, ne));
    FP_TRY(dalloc(ctx, &s.opoff, ne));
    FP_TRY(dalloc(ctx, &s.opcnt, ne));
    FP_TRY(dalloc(ctx, &s.opfill, ne));
    FP_TRY(dalloc(ctx, &s.sums, (size_t)std::max(nblk, 1)));
    FP_TRY(dalloc(ctx, &s.mv_v, np));
    FP_TRY(dalloc(ctx, &s.mv_net, np));
    FP_TRY(dalloc(ctx, &s.mv_from, np));
    FP_TRY(dalloc(ctx, &s.mv_last, np));
    FP_TRY(dalloc(ctx, &s.totals, 8));
    // a pass's pushes are at most its pins (one per (net, pin)): the big
    // nets' bitmap; its ops at most sum of degree^2 (below): the big list, a
    // big net holding more than CN_MAXOPS of them
    FP_TRY(dalloc(ctx, &s.bigbits, (size_t)(np / 32 + 1)));
    s.big_cap = (long long)(std::max<unsigned long long>(ctx->sum_deg2, 1ull) / (unsigned long long)CN_MAXOPS + 1);
    FP_TRY(dalloc(ctx, &s.big, (size_t)s.big_cap));
    fill_async<int>(ctx, s.opcnt, ne, 0);
    // a pass edits the lists of a pushed vertex's nets, and a vertex is pushed
    // by at most its degree many nets: sum of degree^2 bounds the edit list
    const long long cap = (long long)std::max<unsigned long long>(
        1ull, std::min<unsigned long long>(ctx->sum_deg2, 4ull * (unsigned long long)np));
    FP_TRY(dalloc(ctx, &s.ops, (size_t)cap));
    s.ops_cap = cap;
    s.ready = true;
  }
  fill_async<int>(ctx, ctx->flags, 4, 0);
  p.R = ctx->nov;
  return FP_OK;
}

fp_status pass_cn(fp_ctx* ctx, const RefinePlan&) {
  const int noe = ctx->noe;
  const int nblk = (noe + SCAN_BLOCK - 1) / SCAN_BLOCK;
  auto& s = ctx->cn;
  const size_t lds_apply = cn_apply_lds(ctx->K);
  auto cn_apply = cn_apply_indexed(ctx->K) ? k_cn_apply<true> : k_cn_apply<false>;
  const unsigned long long* nm = s.totals + 3;
  // one round per pass: the touched-net counters start at 0
  pass_begin(ctx, nullptr, 0, nullptr);
  fill_async<unsigned long long>(ctx, s.totals, 5, 0ull);
  CnEvalArgs ea{ctx->net_ptr, ctx->net_pins, ctx->vtx_ptr, ctx->vtx_seg, ctx->ent, ctx->ev, ctx->part,
                noe, ctx->pb, ctx->pmask, s.old_part, s.new_part, s.npush};
  const int waves_needed = (noe + WAVE - 1) / WAVE;
  k_cn_eval<<<std::max(1, std::min((waves_needed + 3) / 4, 8192)), 256, 0, ctx->stream>>>(ea);
  k_scan_sums<<<std::max(nblk, 1), SCAN_BLOCK, 0, ctx->stream>>>(s.npush, noe, s.sums);
  k_scan_top<<<1, SCAN_BLOCK, 0, ctx->stream>>>(s.sums, nblk, reinterpret_cast<long long*>(s.totals + 3));
  k_scan_apply<<<std::max(nblk, 1), SCAN_BLOCK, 0, ctx->stream>>>(s.npush, noe, s.sums, s.pbase);
  k_cn_pushes<<<std::max(1, std::min((waves_needed + 3) / 4, 8192)), 256, 0, ctx->stream>>>(
      ctx->net_ptr, ctx->net_pins, ctx->part, s.old_part, s.pbase, noe, s.mv_v, s.mv_net);
  const int par = 1;
  CnPrepArgs pa{s.mv_v, s.mv_net, nm, ctx->vtx_ptr, ctx->vtx_nets, ctx->part, s.old_part, s.new_part, s.pbase,
                s.npush, s.mv_from, s.mv_last, ctx->touched, ctx->counters + par, s.opcnt, s.totals};
  const int gm = (int)std::min<long long>((std::max(ctx->pins, 1ll) + 255) / 256, 8192);
  k_cn_prep<<<gm, 256, 0, ctx->stream>>>(pa);
  k_cn_alloc<<<(noe + 255) / 256, 256, 0, ctx->stream>>>(ctx->touched, ctx->counters + par, s.opcnt, s.old_part,
                                                         s.npush, s.opoff, s.opfill, s.totals, s.big, s.big_cap);
  k_cn_check<<<1, 1, 0, ctx->stream>>>(s.totals, s.ops_cap, ctx->dup_pins ? 1 : 0, ctx->flags);
  k_cn_scatter<<<gm, 256, 0, ctx->stream>>>(s.mv_v, s.mv_net, s.pbase, nm, ctx->flags, ctx->vtx_ptr, ctx->vtx_nets,
                                            s.opoff, s.opfill, s.ops);
  k_cn_bigsort<<<1, BIGSORT_BLOCK, 0, ctx->stream>>>(s.big, s.big_cap, s.totals, nm, ctx->flags, s.opcnt, s.opoff,
                                                     s.bigbits, s.ops);
  CnApplyArgs aa{ctx->net_ptr, ctx->touched, ctx->counters + par, s.opcnt, s.opoff, s.ops, s.mv_from, s.mv_net,
                 s.new_part, ctx->ent, ctx->ev, ctx->K, ctx->pb, ctx->pmask, s.opcnt, ctx->flags};
  auto cn_small = cn_small_indexed(ctx->K) ? k_cn_apply_small<true> : k_cn_apply_small<false>;
  cn_small<<<4096, cn_small_waves(ctx->K) * WAVE, cn_small_lds(ctx->K), ctx->stream>>>(aa);
  cn_apply<<<8192, WAVE, lds_apply, ctx->stream>>>(aa);
  if ((size_t)ctx->K * sizeof(long long) <= (size_t)LDS_BUDGET)
    k_cn_commit<true><<<std::min(gm, 256), 256, (size_t)ctx->K * sizeof(long long), ctx->stream>>>(
        s.mv_v, s.mv_net, s.mv_from, s.mv_last, nm, ctx->flags, s.new_part, ctx->vwgt, ctx->K, ctx->part, ctx->sizes,
        ctx->moves);
  else
    k_cn_commit<false><<<gm, 256, 0, ctx->stream>>>(s.mv_v, s.mv_net, s.mv_from, s.mv_last, nm, ctx->flags,
                                                    s.new_part, ctx->vwgt, ctx->K, ctx->part, ctx->sizes, ctx->moves);
  return FP_OK;
}

fp_status cn_passes(fp_ctx* ctx, const RefinePlan& p, int passes) {
  int done = 0;
  while (done < passes) {
    fill_async<int>(ctx, ctx->flags + 2, 1, 0);
    const int run = passes - done;
    for (int it = 0; it < run; it++) FP_TRY(pass_cn(ctx, p));
    int f[3] = {0, 0, 0};
    FP_TRY(d2h(ctx, f, ctx->flags, 3));
    FP_HIP_OK(hipStreamSynchronize(ctx->stream));
    if (f[1])
      return fail(ctx, FP_ERR_UNSUPPORTED,
                  "a net with a repeated pin receives more than %d moves in one CN pass; that pass was not applied",
                  CN_MAXOPS);
    if (f[2] < 0 || f[2] > run) return fail(ctx, FP_ERR_HIP, "CN pass count %d out of [0, %d]", f[2], run);
    done += f[2];
    if (f[0]) {
      auto& s = ctx->cn;
      s.ops_cap = std::max<long long>((long long)f[0], 2 * s.ops_cap);
      FP_TRY(dalloc(ctx, &s.ops, (size_t)s.ops_cap));
      fill_async<int>(ctx, ctx->flags, 4, 0);
    } else if (f[2] != run) {
      return fail(ctx, FP_ERR_HIP, "CN passes: %d of %d applied without an overflow", f[2], run);
    }
  }
  return FP_OK;
}

fp_status cn_after_pass(fp_ctx* ctx, int* redo) {
  *redo = 0;
  int f[2] = {0, 0};
  FP_TRY(d2h(ctx, f, ctx->flags, 2));
  FP_HIP_OK(hipStreamSynchronize(ctx->stream));
  if (f[1])
    return fail(ctx, FP_ERR_UNSUPPORTED,
                "a net with a repeated pin receives more than %d moves in one CN pass; that pass was not applied",
                CN_MAXOPS);
  if (f[0]) {
    auto& s = ctx->cn;
    s.ops_cap = std::max<long long>((long long)f[0], 2 * s.ops_cap);
    FP_TRY(dalloc(ctx, &s.ops, (size_t)s.ops_cap));
    fill_async<int>(ctx, ctx->flags, 4, 0);
    *redo = 1;
  }
  return FP_OK;
}

}  // namespace fpe
