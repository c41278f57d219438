// fp_tv.hip -- TV refinement on the device (TV/partition.cpp:307-440 decisions,
// TV/partition.cpp:669-896 round schedule).
//
// One pass = rounds of R consecutive ordering positions. A round is:
//   k_tv_eval    one wavefront per vertex (grid-stride over the round,
//                software-pipelined over the wave's vertices) sweeps the lists
//                of the vertex's nets in their 16-bit evaluation view -- one
//                128-byte read per net brings its lambda and first 63
//                entries, 3 nets in flight; the further windows of long
//                lists follow as one task list over all of the vertex's long
//                nets, 2 windows in flight -- into LDS histograms of pin
//                counts per part, then picks the LDG2 decision by a
//                lexicographic wave reduction (a vertex whose best score and
//                size two parts share is swept again with the
//                first-appearance ranks that break the tie) and applies the
//                TV gain test. Candidates stamp their nets into the
//                round's touched-net list; BAL also accumulates candidate
//                weights per part and per (part, 64-slot tile).
//   k_bal_scan   BAL: parts whose candidate weight exceeds cap - size get a
//                dense index; for each, an exclusive scan over tiles
//   k_bal_admit  BAL: prefix admission per part in ordering order, commit of
//                part vector and sizes
//   k_apply      one wavefront per touched net (one lane per net of at most 8
//                pins) replays the round's moves on its list in ordering
//                order (plus, for REF, the commit of the part vector and the
//                source-part sizes in extra blocks)
#include "fp_internal.h"

namespace fpe {
namespace {

constexpr int EVAL_BLOCK = 256;  // 4 wavefronts
constexpr int TILE = 64;         // slots per admission tile
#ifndef FP_NETS_IN_FLIGHT
#define FP_NETS_IN_FLIGHT 3
#endif
constexpr int NETS_IN_FLIGHT = FP_NETS_IN_FLIGHT;  // segment head windows loaded together per wave
// nets per chunk of a vertex's net list: a multiple of NETS_IN_FLIGHT, so that
// no group of a chunk reaches past the wave's last lane (lanes from nn on hold
// the zero padding: lists of length 0)
constexpr int CHUNK = (WAVE / NETS_IN_FLIGHT) * NETS_IN_FLIGHT;
#ifndef FP_LONG_TASKS
#define FP_LONG_TASKS 2
#endif
constexpr int LONG_TASKS = FP_LONG_TASKS;  // windows of long lists (any of the chunk's nets) loaded together
constexpr int SCAN_RUN = 16;     // consecutive tiles per thread in the BAL scan
constexpr int APPLY_BLOCK = 256;
constexpr int SMALL_NET = 8;     // nets of at most this many pins are replayed one per lane
constexpr int APPLY_NET_BLOCKS = 8192;
constexpr int APPLY_GROUP = 4;   // 64-rank windows of a net whose loads the apply issues together

struct EvalArgs {
  const unsigned* vtx_ptr;
  const unsigned* vtx_nets;
  const unsigned* vtx_seg;
  const unsigned short* ev;
  unsigned pad_seg;  // the 64 zero words after the last segment: a list of length 0
  const int* part;
  const int* order;  // nullable: identity
  const int* vwgt;   // nullable: unit
  const long long* sizes;
  int K, pb;
  unsigned pmask;
  int r0, nslots, ntiles;
  double C;
  long long max_pw, min_pw;
  int* slot_own;
  int* slot_dec;
  int* dec_out;   // nullable (fp_eval_moves)
  int* gain_out;  // nullable
  // touched-net list of the round (nullable: no stamping)
  unsigned tag;
  unsigned* net_stamp;
  int* touched;
  int* tcount;
  int* tcount_next;
  // BAL (nullable): candidate weight per part, per (part, tile); frozen sizes copy
  long long* cand_tot;
  int* hist;
  long long* sizes_frozen;
  // initial partition (nullable): the lowest-index part of minimum frozen size
  int* fb_out;
};

// slot decision of an unplaced vertex whose part the fallback assigns
constexpr int SLOT_FALLBACK = -2;

__device__ __forceinline__ bool better(double sc, long long sz, unsigned fi, double bsc, long long bsz, unsigned bfi) {
  return sc > bsc || (sc == bsc && (sz < bsz || (sz == bsz && fi < bfi)));
}

// The LDG2 decision of one vertex with first-appearance ranks, for a vertex
// whose best (score, size) two parts share: its lists are swept again into
// pin counts and ranks (i << pb | t, i the net's index in the vertex's list,
// t the part's position in the net's list: the order in which the reference's
// scan meets the parts, TV/partition.cpp:318-338), and the lexicographic max
// of (score, -size, -rank) over score > 0 picks the part. cnt and first are
// left clear (0, NOFIRST).
template <int MODE>
__device__ __forceinline__ void tie_decision(const EvalArgs& a, unsigned vb, int deg, int own, unsigned* cnt,
                                          unsigned* first, const double* s_pen, const unsigned* s_skey, int& d,
                                          unsigned& craw) {
  const int lane = lane_id(), K = a.K;
  for (int c0 = 0; c0 < deg; c0 += WAVE) {
    const int nn = min(WAVE, deg - c0);
    const unsigned sg = lane < nn ? a.vtx_seg[vb + (unsigned)(c0 + lane)] : a.pad_seg;
    wave_sweep_lists<1, 1>(a.ev, sg, nn, [&](int j, int L, int t, bool ok, unsigned x) {
      if (ok && (MODE != 2 || L == 1)) {  // mode 2, CN's stream: uncut nets only
        const int q = (int)(x & ~EV_SINGLE);
        atomicAdd(&cnt[q], 1u);
        atomicMin(&first[q], ((unsigned)(c0 + j) << a.pb) | (unsigned)t);
      }
    });
  }
  wave_lds_sync();
  unsigned long long b1 = 0, b2 = ~0ull;
  int bq = -1;
  unsigned bc = 0;
  for (int q = lane; q < K; q += WAVE) {
    const unsigned c = cnt[q], f = first[q];
    cnt[q] = 0;
    first[q] = NOFIRST;
    const int adj = (int)c - (q == own ? 1 : 0);
    const double sc = (double)adj * s_pen[q];
    const bool ok = c != 0 && sc > 0;
    const unsigned long long k1 = ok ? (unsigned long long)__double_as_longlong(sc) : 0ull;
    const unsigned long long k2 = ((unsigned long long)s_skey[q] << 32) | f;
    const bool better = ok && (k1 > b1 || (k1 == b1 && k2 < b2));
    b1 = better ? k1 : b1;
    b2 = better ? k2 : b2;
    bq = better ? q : bq;
    bc = better ? c : bc;
  }
  const unsigned long long m1 = wave_max_u64(b1);
  const unsigned long long m2 = wave_min_u64(b1 == m1 ? b2 : ~0ull);
  const int wl = __ffsll((long long)__ballot(b1 == m1 && b2 == m2)) - 1;
  d = __builtin_amdgcn_readlane(bq, wl);
  craw = (unsigned)__builtin_amdgcn_readlane((int)bc, wl);
  wave_lds_sync();
}

// The evaluation's modes: a refinement round (TV/partition.cpp:732-805), or a
// sub-round of the streaming initial partition, whose unplaced vertices
// (part -1) are scored over every list of their nets (LDG2,
// TV/partition.cpp:214-233; plain LDG, MGV/partition.cpp:28-65) or only over
// the lists of their uncut nets (CN/partition.cpp:138-156).
enum EvalMode { EVAL_REFINE = 0, EVAL_INIT = 1, EVAL_INIT_CN = 2 };

// Per vertex: sweep, LDG2 decision, TV gain test (TV/partition.cpp:732-805).
// Eight waves per SIMD (the hardware's maximum) hide more of the dependent
// list reads than the registers a freer allocation would save: the compiler
// keeps within 64 VGPRs by parking cold scalar values in VGPR lanes.
// EVAL_INIT*: a vertex already placed is skipped (no candidate); an unplaced
// one is a candidate for its LDG choice, the lowest-index part of minimum
// size standing in where the choice is above max_pw (where the reference
// would index part_sizes[-1]); the min_pw band does not apply (own == -1).
template <int MODE>
__global__ void __launch_bounds__(EVAL_BLOCK) __attribute__((amdgpu_waves_per_eu(8))) k_tv_eval(EvalArgs a) {
  extern __shared__ __align__(16) unsigned char lds[];
  const int K = a.K;
  double* s_pen = reinterpret_cast<double*>(lds);
  unsigned* s_skey = reinterpret_cast<unsigned*>(s_pen + K);  // size_key of each frozen part size
  unsigned* wbase = s_skey + K;
  __shared__ long long red_sz[EVAL_BLOCK / WAVE];
  __shared__ int red_ix[EVAL_BLOCK / WAVE];
  __shared__ int s_tn[EVAL_BLOCK / WAVE + 1];
  const int lane = lane_id(), wave = wave_id(), nw = (int)(blockDim.x / WAVE);
  // per wavefront: cnt[K] and first[K], each followed by 64 spare slots that
  // take the atomics of lanes past a list's end (no branch around them)
  unsigned* cnt = wbase + (size_t)wave * 2 * (size_t)(K + WAVE);
  unsigned* first = cnt + K + WAVE;
  const int spare = K + lane_id();
  // per wavefront: the nets its accepted vertices touch, before they go to the touched list
  WaveTouch tb{wbase + (size_t)nw * 2 * (size_t)(K + WAVE) + (size_t)wave * TOUCH_BUF, 0};

  // penalties of the frozen part sizes: penalty = 1 - (1.0 * size / C)  (TV/partition.cpp:103)
  long long bs = LLONG_MAX;
  int bi = INT_MAX;
  for (int q = (int)threadIdx.x; q < K; q += (int)blockDim.x) {
    const long long sz = a.sizes[q];
    s_skey[q] = size_key(sz);
    s_pen[q] = 1.0 - ((double)sz / a.C);
    if (sz < bs) { bs = sz; bi = q; }
    if (blockIdx.x == 0 && a.sizes_frozen) a.sizes_frozen[q] = sz;
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
  if (MODE != EVAL_REFINE && blockIdx.x == 0 && threadIdx.x == 0) *a.fb_out = fb;

  // One wavefront per vertex, grid-stride over the round's slots, software
  // pipelined over the wave's vertices: while vertex X is swept, the segment
  // addresses of X+1's nets and the vertex header of X+2 are in flight.
  const int stride = (int)gridDim.x * nw;
  const int slot_end = a.nslots;
  int slot = (int)blockIdx.x * nw + wave;
  int hv[3], hown[3], hdeg[3];
  unsigned hvb[3];
  auto load_vertex = [&](int k, int sl) {
    if (sl < slot_end) {
      const int vv = a.order ? uniform_load(a.order + (a.r0 + sl)) : a.r0 + sl;
      hv[k] = vv;
      hown[k] = uniform_load(a.part + vv);
      hvb[k] = uniform_load(a.vtx_ptr + vv);
      hdeg[k] = (int)(uniform_load(a.vtx_ptr + vv + 1) - hvb[k]);
    } else {
      hv[k] = 0; hown[k] = 0; hvb[k] = 0; hdeg[k] = 0;
    }
  };
  auto load_segs = [&](int k) -> unsigned {
    return lane < min(hdeg[k], CHUNK) ? a.vtx_seg[hvb[k] + (unsigned)lane] : a.pad_seg;
  };
  // the first window group of the vertex about to be swept, loaded while the
  // previous vertex picks its candidate
  unsigned xpre[NETS_IN_FLIGHT];
  auto load_first = [&](unsigned sgv) {
#pragma unroll
    for (int u = 0; u < NETS_IN_FLIGHT; u++) {
      const unsigned sj = (unsigned)__builtin_amdgcn_readlane((int)sgv, u);
      xpre[u] = a.ev[sj + (unsigned)lane];
    }
  };
  load_vertex(0, slot);
  load_vertex(1, slot + stride);
  unsigned sg0 = load_segs(0);
  load_first(sg0);
  while (slot < slot_end) {
    load_vertex(2, slot + 2 * stride);
    const unsigned sg1 = load_segs(1);
    const int v = hv[0], own = hown[0], deg = hdeg[0];
    const unsigned vb = hvb[0];

    // sweep: histogram of parts over the lists of v's nets. The first
    // appearance of a part is ranked by (net index i in v's net list, position
    // t in that net's list), packed as i << pb | t (t < lambda <= K <= 2^pb):
    // the order in which the reference's scan meets the parts
    // (TV/partition.cpp:318-338). One 64-word read of a segment brings lambda
    // (lane 0) and the first 63 entries (lanes 1..63); NETS_IN_FLIGHT such
    // reads are issued together, longer lists are finished after. Lanes past
    // the vertex's nets hold the zero padding's offset, so the windows past
    // the last net see lambda = 0. The leave-gain count is kept per lane.
    int lgl = 0;
    for (int c0 = 0; c0 < deg; c0 += CHUNK) {
      const int nn = min(CHUNK, deg - c0);
      const unsigned sg = c0 == 0 ? sg0 : (lane < nn ? a.vtx_seg[vb + (unsigned)(c0 + lane)] : a.pad_seg);
      int lmax = 0;  // longest list of the chunk
      int lj = 0;    // lane j: lambda of the chunk's net j
      // the next group's windows are in flight while a group is counted
      unsigned xv[NETS_IN_FLIGHT];
      auto load_group = [&](int j0, unsigned (&dst)[NETS_IN_FLIGHT]) {
#pragma unroll
        for (int u = 0; u < NETS_IN_FLIGHT; u++) {
          const unsigned sj = (unsigned)__builtin_amdgcn_readlane((int)sg, j0 + u);
          dst[u] = a.ev[sj + (unsigned)lane];
        }
      };
      if (c0 == 0) {
#pragma unroll
        for (int u = 0; u < NETS_IN_FLIGHT; u++) xv[u] = xpre[u];
      } else {
        load_group(0, xv);
      }
      for (int j0 = 0; j0 < nn; j0 += NETS_IN_FLIGHT) {
        unsigned xn[NETS_IN_FLIGHT];
        const bool more = j0 + NETS_IN_FLIGHT < nn;
        if (more) load_group(j0 + NETS_IN_FLIGHT, xn);
#pragma unroll
        for (int u = 0; u < NETS_IN_FLIGHT; u++) {
          const int j = j0 + u;
          const int Lj = __builtin_amdgcn_readlane((int)xv[u], 0);
          lmax = max(lmax, Lj);
          lj = lane == j ? Lj : lj;
          const int t = lane - 1;
          // EVAL_INIT_CN: only nets of one part vote (lambda == 1)
          const bool ok = lane >= 1 && t < Lj && (MODE != EVAL_INIT_CN || Lj == 1);
          const unsigned x = xv[u];
          const int q = (int)(x & ~EV_SINGLE);
          const int qs = ok ? q : spare;
          atomicAdd(&cnt[qs], 1u);
          // leave gain: nets where v is the only pin of its part (TV/partition.cpp:365-368)
          lgl += ok && q == own && (x & EV_SINGLE) ? 1 : 0;
        }
        if (more) {
#pragma unroll
          for (int u = 0; u < NETS_IN_FLIGHT; u++) xv[u] = xn[u];
        }
      }
      // the rest of the long lists (more than 63 entries): entries 63.. in
      // windows of 64. Each (net, window) pair of the chunk is a task,
      // numbered net by net by a scan of the per-net window counts, and
      // LONG_TASKS tasks' windows are loaded together, whichever nets they
      // belong to; the lambdas come from the first windows.
      if (MODE != EVAL_INIT_CN && lmax > WAVE - 1) {
        const int nwin = lj >> 6;  // ceil((lambda - 63) / 64) windows past the first
        const int incl = wave_incl_scan(nwin);
        const int ntask = __builtin_amdgcn_readlane(incl, WAVE - 1);
        for (int k0 = 0; k0 < ntask; k0 += LONG_TASKS) {
          unsigned xw[LONG_TASKS];
          int tj[LONG_TASKS], tw[LONG_TASKS], tl[LONG_TASKS];
#pragma unroll
          for (int u = 0; u < LONG_TASKS; u++) {
            // past the last task: a repeat of it, with no entries counted
            const int k = min(k0 + u, ntask - 1);
            const int j = (int)__popcll(__ballot(incl <= k));  // its net: incl is non-decreasing
            const int w = k - __builtin_amdgcn_readlane(incl - nwin, j);
            const unsigned sj = (unsigned)__builtin_amdgcn_readlane((int)sg, j);
            const int Lj = __builtin_amdgcn_readlane(lj, j);
            tj[u] = j;
            tw[u] = WAVE - 1 + w * WAVE;
            tl[u] = k0 + u < ntask ? Lj : 0;
            const int t = tw[u] + lane;
            xw[u] = a.ev[sj + 1u + (unsigned)(t < Lj ? t : 0)];
          }
#pragma unroll
          for (int u = 0; u < LONG_TASKS; u++) {
            const int t = tw[u] + lane;
            const bool ok = t < tl[u];
            const unsigned x = xw[u];
            const int q = (int)(x & ~EV_SINGLE);
            const int qs = ok ? q : spare;
            atomicAdd(&cnt[qs], 1u);
            lgl += ok && q == own && (x & EV_SINGLE) ? 1 : 0;
          }
        }
      }
    }
    const int lg = wave_sum(lgl);
    load_first(sg1);
    wave_lds_sync();
    // candidates: lexicographic max of (score, -size, -first rank) over score >
    // 0 equals the sequential LDG2 scan over the active parts
    // (TV/partition.cpp:98-116). Scores are positive doubles, so their bit
    // patterns order like the values: key1 = score bits (max), key2 =
    // size_key(size) << 32 | first rank (min among key1 ties).
    const unsigned cfb = cnt[fb];
    // candidates: lexicographic max of (score, -size, -first rank) over score >
    // 0 (TV/partition.cpp:98-116). The first rank decides only between parts
    // of equal score and size: the sweep counts pins only, and a vertex whose
    // best (score, size) is shared by two parts is decided again with ranks
    // (tie_decision). Scores are positive doubles, so their bit patterns order
    // like the values: key1 = score bits (max), key2 = size_key(size) (min
    // among key1 ties); ltie: this lane met its best keys twice.
    unsigned long long b1 = 0;
    unsigned b2 = ~0u;
    int bq = -1;
    unsigned bc = 0;
    bool ltie = false;
    // four 64-part chunks per step, without branches: their LDS reads issue
    // together; every slot is cleared for the next vertex
    for (int q0 = 0; q0 < K; q0 += 4 * WAVE) {
      unsigned cv[4];
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const int q = min(q0 + u * WAVE + lane, K - 1);
        cv[u] = q0 + u * WAVE + lane < K ? cnt[q] : 0u;
      }
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const int q = q0 + u * WAVE + lane;
        const int qc = min(q, K - 1);
        if (q < K) cnt[q] = 0;
        const int adj = (int)cv[u] - (q == own ? 1 : 0);   // scores[part_id]-- (TV/partition.cpp:345)
        const double sc = (double)adj * s_pen[qc];
        const bool ok = cv[u] != 0 && sc > 0;
        const unsigned long long k1 = ok ? (unsigned long long)__double_as_longlong(sc) : 0ull;
        const unsigned k2 = s_skey[qc];
        const bool better = ok && (k1 > b1 || (k1 == b1 && k2 < b2));
        const bool same = ok && k1 == b1 && k2 == b2;
        ltie = better ? false : (ltie || same);
        b1 = better ? k1 : b1;
        b2 = better ? k2 : b2;
        bq = better ? q : bq;
        bc = better ? cv[u] : bc;
      }
    }
    const unsigned long long m1 = wave_max_u64(b1);
    wave_lds_sync();
    int d = fb;
    unsigned craw = cfb;
    if (m1 != 0) {
      // the second key is needed only when several lanes hold the best score
      unsigned long long cand = __ballot(b1 == m1);
      if (__popcll(cand) > 1) {
        const unsigned m2 = wave_min_u32(b1 == m1 ? b2 : ~0u);
        cand = __ballot(b1 == m1 && b2 == m2);
      }
      const int wl = __ffsll((long long)cand) - 1;
      if (__popcll(cand) > 1 || __builtin_amdgcn_readlane((int)ltie, wl)) {
        tie_decision<MODE>(a, vb, deg, own, cnt, first, s_pen, s_skey, d, craw);
      } else {
        d = __builtin_amdgcn_readlane(bq, wl);
        craw = (unsigned)__builtin_amdgcn_readlane((int)bc, wl);
      }
    }
    // the bands, compared as size keys (the same order)
    bool accept;
    int dd;  // the slot's decision: a part, -1 (none) or SLOT_FALLBACK
    if (MODE == EVAL_REFINE) {
      if (s_skey[d] > size_key(a.max_pw)) d = own;   // TV/partition.cpp:135-138
      if (s_skey[d] < size_key(a.min_pw)) d = own;   // TV/partition.cpp:140-143
      const int al = deg - (int)craw;      // arrival loss: nets of v lacking d (TV/partition.cpp:369-377)
      accept = d != own && lg >= al;       // TV/partition.cpp:380
      dd = accept ? d : -1;
    } else {
      // an unplaced vertex with no part above score 0, or whose choice is
      // above the band, takes its turn in the fallback
      accept = own < 0;
      dd = !accept ? -1 : (m1 == 0 || s_skey[d] > size_key(a.max_pw) ? SLOT_FALLBACK : d);
    }
    if (lane == 0) {
      a.slot_own[slot] = own;
      a.slot_dec[slot] = dd;
      if (MODE == EVAL_REFINE && a.dec_out) a.dec_out[v] = d;
      if (MODE == EVAL_REFINE && a.gain_out) a.gain_out[v] = d == own ? INT_MIN : lg - (deg - (int)craw);
      if (dd >= 0 && a.cand_tot) {
        const int w = a.vwgt ? a.vwgt[v] : 1;
        atomic_add_i64(&a.cand_tot[d], w);
        atomicAdd(&a.hist[(size_t)d * (size_t)a.ntiles + (size_t)(slot / TILE)], w);
      }
    }
    if (accept && a.touched)
      wave_touch_nets_staged(a.vtx_nets + vb, deg, a.tag, a.net_stamp, tb, a.touched, a.tcount);
    slot += stride;
    hv[0] = hv[1]; hown[0] = hown[1]; hvb[0] = hvb[1]; hdeg[0] = hdeg[1];
    hv[1] = hv[2]; hown[1] = hown[2]; hvb[1] = hvb[2]; hdeg[1] = hdeg[2];
    sg0 = sg1;
  }
  if (a.touched) block_touch_flush(tb, a.touched, a.tcount, s_tn);
}

// BAL step 1: over-subscribed parts (candidate weight > cap - frozen size) get
// a dense index; one block per such part scans its tile histogram (exclusive)
// into pref and clears the histogram row.
__global__ void __launch_bounds__(256) k_bal_scan(int* __restrict__ hist, long long* __restrict__ pref, int ntiles,
                                                  int K, const long long* __restrict__ sizes_frozen,
                                                  const long long* __restrict__ cand_tot, long long cap,
                                                  int* __restrict__ over_idx, int* __restrict__ over_count) {
  extern __shared__ __align__(16) unsigned char lds[];
  int* s_list = reinterpret_cast<int*>(lds);  // K
  __shared__ int s_wi[16];
  __shared__ long long s_wl[16];
  int base = 0;
  for (int d0 = 0; d0 < K; d0 += (int)blockDim.x) {
    const int d = d0 + (int)threadIdx.x;
    const int f = (d < K && cand_tot[d] > cap - sizes_frozen[d]) ? 1 : 0;
    int tot;
    const int ex = block_excl_scan(f, s_wi, &tot);
    if (d < K) {
      if (f) s_list[base + ex] = d;
      if (blockIdx.x == 0) over_idx[d] = f ? base + ex : -1;
    }
    base += tot;
  }
  const int nover = base;
  if (blockIdx.x == 0 && threadIdx.x == 0) *over_count = nover;
  __syncthreads();
  // each thread scans SCAN_RUN consecutive tiles, loaded together, so a
  // 4096-tile row takes one round trip to memory and one block scan
  for (int o = (int)blockIdx.x; o < nover; o += (int)gridDim.x) {
    const int d = s_list[o];
    int* row = hist + (size_t)d * (size_t)ntiles;
    long long* out = pref + (size_t)o * (size_t)ntiles;
    long long carry = 0;
    for (int t0 = 0; t0 < ntiles; t0 += (int)blockDim.x * SCAN_RUN) {
      const int tb = t0 + (int)threadIdx.x * SCAN_RUN;
      int x[SCAN_RUN];
#pragma unroll
      for (int p = 0; p < SCAN_RUN; p++) {
        const int v = row[min(tb + p, ntiles - 1)];
        x[p] = tb + p < ntiles ? v : 0;
      }
      long long run = 0;
#pragma unroll
      for (int p = 0; p < SCAN_RUN; p++) run += x[p];
      long long tot;
      long long acc = carry + block_excl_scan(run, s_wl, &tot);
#pragma unroll
      for (int p = 0; p < SCAN_RUN; p++) {
        if (tb + p < ntiles) {
          out[tb + p] = acc;
          row[tb + p] = 0;
        }
        acc += x[p];
      }
      carry += tot;
    }
  }
}

// Initial partition, between the evaluation and the admission (one block):
// the j-th fallback slot of the sub-round, in slot order, takes part
// (fb + j) mod K, fb the lowest-index part of minimum frozen size, and its
// weight joins the candidate totals and tile histogram.
constexpr int FALLBACK_RUN = 16;  // consecutive slots per thread in k_init_fallback

__global__ void __launch_bounds__(1024) k_init_fallback(int* __restrict__ slot_dec, const int* __restrict__ order,
                                                        const int* __restrict__ vwgt, int r0, int nslots, int ntiles,
                                                        int K, const int* __restrict__ fb_in,
                                                        long long* __restrict__ cand_tot, int* __restrict__ hist) {
  __shared__ int s_wi[16];
  const int fb = *fb_in;
  int carry = 0;
  // each thread takes FALLBACK_RUN consecutive slots, their decisions loaded
  // together: a sub-round of 16,384 slots is one load round trip and one
  // block scan
  for (int b0 = 0; b0 < nslots; b0 += (int)blockDim.x * FALLBACK_RUN) {
    const int base = b0 + (int)threadIdx.x * FALLBACK_RUN;
    int dec[FALLBACK_RUN];
#pragma unroll
    for (int q = 0; q < FALLBACK_RUN; q++) dec[q] = base + q < nslots ? slot_dec[base + q] : 0;
    int cnt = 0;
#pragma unroll
    for (int q = 0; q < FALLBACK_RUN; q++) cnt += dec[q] == SLOT_FALLBACK ? 1 : 0;
    int tot;
    int j = carry + block_excl_scan(cnt, s_wi, &tot);
    if (cnt)
      for (int q = 0; q < FALLBACK_RUN; q++) {
        if (dec[q] != SLOT_FALLBACK) continue;
        const int slot = base + q;
        const int d = (int)(((long long)fb + j) % K);
        j++;
        slot_dec[slot] = d;
        const int z = r0 + slot;
        const int w = vwgt ? vwgt[order ? order[z] : z] : 1;
        atomic_add_i64(&cand_tot[d], w);
        atomicAdd(&hist[(size_t)d * (size_t)ntiles + (size_t)(slot / TILE)], w);
      }
    carry += tot;
  }
}

// BAL step 2, one wavefront per tile: a candidate into an over-subscribed part
// d is admitted iff frozen_size[d] + (weight of earlier candidates into d) + w
// <= cap (first_free: or it is the first candidate into d); every other
// candidate is admitted. Commits part vector and sizes of admitted moves (a
// vertex without a part, own -1, only adds to its new part's size); clears
// the tile's histogram cells of non-over parts. AGG:
// the block sums its size changes per part in LDS first, so that sizes takes
// one atomic per part and block (same-address atomics from every wave would
// queue at one L2 slice).
template <bool AGG>
__global__ void __launch_bounds__(256) k_bal_admit(int* __restrict__ slot_dec, const int* __restrict__ slot_own,
                                                   const int* __restrict__ order, const int* __restrict__ vwgt,
                                                   int r0, int nslots, int ntiles, int K,
                                                   const long long* __restrict__ sizes_frozen, long long cap,
                                                   bool first_free, const int* __restrict__ over_idx,
                                                   const int* __restrict__ over_count,
                                                   const long long* __restrict__ pref, int* __restrict__ hist,
                                                   int* __restrict__ part, long long* __restrict__ sizes,
                                                   long long* __restrict__ cand_tot,
                                                   unsigned long long* __restrict__ moves) {
  const int lane = lane_id(), wave = wave_id(), nw = (int)(blockDim.x / WAVE);
  const int nover = *over_count;
  // the block's admitted moves, added to the global count once per block
  // (same-address atomics from every wave would queue at one L2 slice)
  __shared__ unsigned long long s_moves;
  // AGG: the block's size changes per part, added to sizes once per block
  extern __shared__ __align__(16) unsigned char lds[];
  unsigned long long* s_dsz = reinterpret_cast<unsigned long long*>(lds);
  if (threadIdx.x == 0) s_moves = 0;
  if (AGG)
    for (int d = (int)threadIdx.x; d < K; d += (int)blockDim.x) s_dsz[d] = 0;
  __syncthreads();
  if (blockIdx.x == 0)
    for (int d = (int)threadIdx.x; d < K; d += (int)blockDim.x) cand_tot[d] = 0;
  for (int tile = (int)blockIdx.x * nw + wave; tile < ntiles; tile += (int)gridDim.x * nw) {
    const int slot = tile * TILE + lane;
    int d = -1, v = -1;
    long long w = 0;
    if (slot < nslots) {
      d = slot_dec[slot];
      const int z = r0 + slot;
      v = order ? order[z] : z;
      w = vwgt ? vwgt[v] : 1;
    }
    const int o = (d >= 0 && nover > 0) ? over_idx[d] : -1;
    bool adm = d >= 0;
    if (__ballot(o >= 0)) {
      // weight of earlier candidates into the same part within the tile
      long long pre = 0;
      const int wi = (int)w;  // vertex weights are int
      for (int j = 0; j < WAVE; j++) {
        const int dj = __builtin_amdgcn_readlane(d, j), wj = __builtin_amdgcn_readlane(wi, j);
        if (j < lane && dj == d) pre += wj;
      }
      if (o >= 0) {
        const long long before = pref[(size_t)o * (size_t)ntiles + (size_t)tile] + pre;
        adm = sizes_frozen[d] + before + w <= cap || (first_free && before == 0);
      }
    }
    if (d >= 0) {
      if (o < 0) hist[(size_t)d * (size_t)ntiles + (size_t)tile] = 0;
      if (!adm) slot_dec[slot] = -1;
    }
    if (adm) {
      part[v] = d;
      const int own = slot_own[slot];
      if (AGG) {
        if (own >= 0) atomicAdd(&s_dsz[own], (unsigned long long)-w);
        atomicAdd(&s_dsz[d], (unsigned long long)w);
      } else {
        if (own >= 0) atomic_add_i64(&sizes[own], -w);
        atomic_add_i64(&sizes[d], w);
      }
    }
    const unsigned long long am = __ballot(adm);
    if (lane == 0 && am) atomicAdd(&s_moves, (unsigned long long)__popcll(am));
  }
  __syncthreads();
  if (threadIdx.x == 0 && s_moves) atomicAdd(moves, s_moves);
  if (AGG)
    for (int d = (int)threadIdx.x; d < K; d += (int)blockDim.x) {
      const long long x = (long long)s_dsz[d];
      if (x) atomic_add_i64(&sizes[d], x);
    }
}

struct ApplyArgs {
  const unsigned* net_ptr;
  const int* apply_pins;
  const int* slot_own;
  const int* slot_dec;
  const int* touched;
  const int* tcount;
  unsigned* ent;
  unsigned short* ev;
  int K, pb;
  unsigned pmask;
  int r0, r1;
  // REF commit in blocks >= net_blocks (nullable part: no commit)
  int net_blocks;
  const int* order;
  int* part;
  long long* sizes;
  unsigned long long* moves;
};

// Replays the round's moves on every touched net in ordering order
// (TV/partition.cpp:846-890: remove = decrement or swap-with-last, add =
// increment or append); one wavefront per net, the list staged in LDS.
// Blocks past net_blocks commit the REF part vector and source-part sizes
// (TV/partition.cpp:818-830: the destination's size is left unchanged).
// INDEXED: each wave keeps a part -> position index of its list in LDS (2
// bytes per part after the lists), for K where it fits.
// bytes of the per-wave lists (and indexes) at the start of the apply's LDS
template <bool INDEXED>
__host__ __device__ __forceinline__ size_t apply_list_bytes(int K, int nw) {
  return (size_t)nw * (size_t)K * (sizeof(unsigned) + (INDEXED ? sizeof(unsigned short) : 0));
}

// One lane replays the round's moves on a net of at most SMALL_NET pins, in
// the rank order of its pins, on a copy of the list in the lane's own LDS
// slots: the same edits, in the same order, as the wavefront replay (remove
// = decrement or swap-with-last, add = increment or append). The list and the
// ranks load together, then the moves of the ranked pins.
__device__ __forceinline__ void lane_replay_small(const ApplyArgs& a, int n, unsigned s, unsigned e, unsigned* my) {
  const unsigned np = e - s;
  const unsigned seg = s + (unsigned)n;
  int r[SMALL_NET];
  unsigned x[SMALL_NET + 1];
#pragma unroll
  for (int i = 0; i < SMALL_NET; i++) r[i] = (unsigned)i < np ? a.apply_pins[s + (unsigned)i] : INT_MAX;
#pragma unroll
  for (int t = 0; t <= SMALL_NET; t++) x[t] = (unsigned)t <= np ? a.ent[seg + (unsigned)t] : 0u;
  int dd[SMALL_NET], oo[SMALL_NET];
#pragma unroll
  for (int i = 0; i < SMALL_NET; i++) {
    const bool in = r[i] >= a.r0 && r[i] < a.r1;
    dd[i] = in ? a.slot_dec[r[i] - a.r0] : -1;
    oo[i] = in ? a.slot_own[r[i] - a.r0] : 0;
  }
  int L = (int)x[0];
#pragma unroll
  for (int t = 0; t < SMALL_NET; t++)
    if (t < L) my[t] = x[t + 1];
  bool any = false;
#pragma unroll
  for (int i = 0; i < SMALL_NET; i++) any |= dd[i] >= 0;
  if (!any) return;
  const unsigned inc = 1u << a.pb;
#pragma unroll
  for (int i = 0; i < SMALL_NET; i++) {
    if (dd[i] < 0) continue;
    const int p = oo[i], q = dd[i];
    int jp = -1;
    for (int t = 0; t < L; t++)
      if ((int)(my[t] & a.pmask) == p) jp = t;
    if (jp >= 0) {
      const unsigned y = my[jp];
      if ((y >> a.pb) > 1u) {
        my[jp] = y - inc;
      } else {
        my[jp] = my[L - 1];
        L--;
      }
    }
    int jq = -1;
    for (int t = 0; t < L; t++)
      if ((int)(my[t] & a.pmask) == q) jq = t;
    if (jq >= 0) {
      my[jq] += inc;
    } else {
      my[L] = inc | (unsigned)q;
      L++;
    }
  }
  for (int t = 0; t < L; t++) {
    const unsigned y = my[t];
    a.ent[seg + 1 + (unsigned)t] = y;
    a.ev[seg + 1 + (unsigned)t] = ev_word(y, a.pb, a.pmask);
  }
  a.ent[seg] = (unsigned)L;
  a.ev[seg] = (unsigned short)L;
}

template <bool INDEXED>
__global__ void __launch_bounds__(APPLY_BLOCK) k_apply(ApplyArgs a) {
  extern __shared__ __align__(16) unsigned char lds[];
  const int lane = lane_id(), wave = wave_id(), nw = (int)(blockDim.x / WAVE);
  if ((int)blockIdx.x >= a.net_blocks) {
    const int slot = ((int)blockIdx.x - a.net_blocks) * (int)blockDim.x + (int)threadIdx.x;
    bool mv = false;
    if (slot < a.r1 - a.r0) {
      const int d = a.slot_dec[slot];
      if (d >= 0) {
        const int z = a.r0 + slot;
        a.part[a.order ? a.order[z] : z] = d;
        atomic_add_i64(&a.sizes[a.slot_own[slot]], -1);
        mv = true;
      }
    }
    block_count_add(a.moves, mv);
    return;
  }
  unsigned* list = reinterpret_cast<unsigned*>(lds) + (size_t)wave * (size_t)a.K;
  unsigned short* at = reinterpret_cast<unsigned short*>(reinterpret_cast<unsigned*>(lds) + (size_t)nw * (size_t)a.K) +
                       (size_t)wave * (size_t)a.K;
  // per lane: SMALL_NET + 1 words (the odd stride keeps the lanes on distinct banks)
  unsigned* my = reinterpret_cast<unsigned*>(lds + apply_list_bytes<INDEXED>(a.K, nw)) +
                 (size_t)(wave * WAVE + lane) * (size_t)(SMALL_NET + 1);
  if (INDEXED) wave_index_init(at, a.K);
  const int ntouched = *a.tcount;
  // a net's head: the first 64 words of its segment and, for a net of at
  // most 64 pins, its ranks; loaded while the wave's previous net is replayed
  struct NetHead {
    unsigned x;
    int r;
  };
  auto head = [&](int n, unsigned s, unsigned e) -> NetHead {
    NetHead h;
    h.x = wave_list_head(a.ent, s + (unsigned)n, e - s);
    h.r = e - s <= (unsigned)WAVE && (unsigned)lane < e - s ? a.apply_pins[s + (unsigned)lane] : INT_MAX;
    return h;
  };
  auto replay_net = [&](int n, unsigned s, unsigned e, const NetHead& h) {
    const unsigned seg = s + (unsigned)n;
    const bool small = e - s <= (unsigned)WAVE;
    const int r_small = h.r;
    unsigned lo, hi;
    wave_rank_window(a.apply_pins, s, e, a.r0, a.r1, lo, hi);
    int L = wave_list_stage(a.ent, seg, h.x, list);
    if (INDEXED) wave_index_set(list, L, at, a.pmask);
    bool dirty = false;
    // the round's pins of the net, APPLY_GROUP windows of 64 ranks at a time:
    // their ranks load together, then their slots' decisions, so that a big
    // net (a Zipf net of thousands of pins, a quarter of them in the round)
    // takes a few dependent load rounds instead of one per window
    for (unsigned k0 = lo; k0 < hi; k0 += APPLY_GROUP * WAVE) {
      int rr[APPLY_GROUP], dd[APPLY_GROUP], oo[APPLY_GROUP];
#pragma unroll
      for (int u = 0; u < APPLY_GROUP; u++) {
        const unsigned k = k0 + (unsigned)(u * WAVE + lane);
        rr[u] = small ? (u == 0 ? r_small : INT_MAX) : (k < hi ? a.apply_pins[k] : INT_MAX);
      }
#pragma unroll
      for (int u = 0; u < APPLY_GROUP; u++) {
        dd[u] = -1;
        oo[u] = -1;
        if (rr[u] >= a.r0 && rr[u] < a.r1) {
          dd[u] = a.slot_dec[rr[u] - a.r0];
          oo[u] = a.slot_own[rr[u] - a.r0];
        }
      }
#pragma unroll
      for (int u = 0; u < APPLY_GROUP; u++) {
        unsigned long long m = __ballot(dd[u] >= 0);
        while (m) {
          dirty = true;
          const int l = __builtin_amdgcn_readfirstlane(__ffsll((long long)m) - 1);
          m &= m - 1;
          const int o = __builtin_amdgcn_readlane(oo[u], l), d = __builtin_amdgcn_readlane(dd[u], l);
          if (INDEXED) {
            unsigned co;
            bool found;
            wave_ilist_move(list, at, L, o, d, a.pb, a.pmask, co, found);
          } else {
            wave_list_move(list, L, o, d, a.pb, a.pmask);
          }
        }
      }
    }
    if (dirty) wave_list_store(a.ent, a.ev, seg, list, L, a.pb, a.pmask);
    if (INDEXED) wave_index_clear(list, L, at, a.pmask);
    wave_lds_sync();
  };
  // the wave's share of the touched nets, 64 at a time, one per lane: nets of
  // at most SMALL_NET pins are replayed by their lanes, the others by the
  // whole wavefront one after another, the next one's head loading meanwhile
  const int stride = a.net_blocks * nw;
  for (long long b0 = (int)blockIdx.x * nw + wave; b0 < ntouched; b0 += (long long)WAVE * stride) {
    const long long ij = b0 + (long long)lane * stride;
    int nj = 0;
    unsigned sj = 0, ej = 0;
    if (ij < ntouched) {
      nj = a.touched[ij];
      sj = a.net_ptr[nj];
      ej = a.net_ptr[nj + 1];
    }
    const bool sm = ij < ntouched && ej - sj <= (unsigned)SMALL_NET;
    unsigned long long big = __ballot(ij < ntouched && !sm);
    auto net_at = [&](int j, int& n, unsigned& s, unsigned& e) {
      n = __builtin_amdgcn_readlane(nj, j);
      s = (unsigned)__builtin_amdgcn_readlane((int)sj, j);
      e = (unsigned)__builtin_amdgcn_readlane((int)ej, j);
    };
    int j = -1, n = 0;
    unsigned s = 0, e = 0;
    NetHead cur{0u, INT_MAX};
    if (big) {
      j = __builtin_amdgcn_readfirstlane(__ffsll((long long)big) - 1);
      big &= big - 1;
      net_at(j, n, s, e);
      cur = head(n, s, e);
    }
    if (sm) lane_replay_small(a, nj, sj, ej, my);
    while (j >= 0) {
      int jn = -1, nn = 0;
      unsigned sn = 0, en = 0;
      NetHead nxt = cur;
      if (big) {
        jn = __builtin_amdgcn_readfirstlane(__ffsll((long long)big) - 1);
        big &= big - 1;
        net_at(jn, nn, sn, en);
        nxt = head(nn, sn, en);
      }
      replay_net(n, s, e, cur);
      j = jn;
      n = nn;
      s = sn;
      e = en;
      cur = nxt;
    }
  }
}

// the apply's LDS: a K-entry list per wave, and the part -> position index
// where it fits the budget
// LDS of the apply: per wave a K-entry list and, where it fits the budget,
// the part -> position index; then SMALL_NET + 1 words per lane
constexpr size_t APPLY_SMALL_BYTES = (size_t)APPLY_BLOCK * (SMALL_NET + 1) * sizeof(unsigned);

bool apply_indexed(int K) {
  return apply_list_bytes<true>(K, APPLY_BLOCK / WAVE) + APPLY_SMALL_BYTES <= (size_t)LDS_BUDGET;
}

size_t apply_lds(int K) {
  return (apply_indexed(K) ? apply_list_bytes<true>(K, APPLY_BLOCK / WAVE)
                           : apply_list_bytes<false>(K, APPLY_BLOCK / WAVE)) +
         APPLY_SMALL_BYTES;
}

size_t eval_lds(int K, int waves) {
  return (size_t)K * 12 + (size_t)waves * (8 * (size_t)(K + WAVE) + 4 * (size_t)TOUCH_BUF);
}

int eval_waves(int K) {
  // LDS: 12 bytes per part shared by the block + 8 bytes per part (and per
  // spare slot) and the touched-net buffer per wavefront
  for (int w = EVAL_BLOCK / WAVE; w >= 1; w--)
    if (eval_lds(K, w) <= (size_t)LDS_BUDGET) return w;
  return 0;
}

int eval_grid(int nslots, int waves) { return std::max(1, std::min((nslots + waves - 1) / waves, 8192)); }

fp_status ensure_round_scratch(fp_ctx* ctx, int R) {
  if (R > ctx->rmax) {
    FP_TRY(dalloc(ctx, &ctx->slot_own, (size_t)R));
    FP_TRY(dalloc(ctx, &ctx->slot_dec, (size_t)R));
    ctx->rmax = R;
  }
  return FP_OK;
}

fp_status ensure_bal_scratch(fp_ctx* ctx, int R) {
  const size_t ntiles = (size_t)(R + TILE - 1) / TILE;
  const size_t need = ntiles * (size_t)ctx->K;
  if (need > ctx->hist_cap) {
    FP_TRY(dalloc(ctx, &ctx->hist, need));
    FP_TRY(dalloc(ctx, &ctx->pref, need));
    fill_async<int>(ctx, ctx->hist, need, 0);
    ctx->hist_cap = need;
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

// what the evaluation kernel's packed keys assume
fp_status check_envelope(fp_ctx* ctx) {
  if (eval_waves(ctx->K) == 0) return fail(ctx, FP_ERR_UNSUPPORTED, "K = %d needs more LDS than a workgroup has", ctx->K);
  if ((unsigned long long)ctx->max_vdeg + 1 >= (1ull << (32 - ctx->pb)))
    return fail(ctx, FP_ERR_UNSUPPORTED, "a vertex of degree %u is too large for K = %d", ctx->max_vdeg, ctx->K);
  if (ctx->total_weight >= (1ll << 31)) return fail(ctx, FP_ERR_UNSUPPORTED, "total vertex weight >= 2^31");
  return FP_OK;
}

EvalArgs make_eval_args(fp_ctx* ctx, const fp_schedule* s) {
  EvalArgs a{};
  a.pad_seg = (unsigned)((size_t)ctx->pins + (size_t)ctx->noe);
  a.vtx_ptr = ctx->vtx_ptr;
  a.vtx_nets = ctx->vtx_nets;
  a.vtx_seg = ctx->vtx_seg;
  a.ev = ctx->ev;
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
  return a;
}

}  // namespace

fp_status plan_tv(fp_ctx* ctx, RefinePlan& p) {
  const int K = ctx->K;
  const int waves = p.eval_waves > 0 ? std::min(p.eval_waves, eval_waves(K)) : eval_waves(K);
  if (waves == 0) return fail(ctx, FP_ERR_UNSUPPORTED, "K = %d needs more LDS than a workgroup has", K);
  p.eval_waves = waves;
  if (apply_lds(K) > (size_t)LDS_BUDGET) return fail(ctx, FP_ERR_UNSUPPORTED, "K = %d too large for the apply kernel", K);
  p.R = std::min(p.sched.round, ctx->nov);
  FP_TRY(ensure_round_scratch(ctx, p.R));
  if (p.sched.kind == FP_SCHED_BAL) FP_TRY(ensure_bal_scratch(ctx, p.R));
  return FP_OK;
}

fp_status pass_tv(fp_ctx* ctx, const RefinePlan& p) {
  const int nov = ctx->nov, K = ctx->K, R = p.R, waves = p.eval_waves;
  const fp_schedule* sched = &p.sched;
  const bool bal = sched->kind == FP_SCHED_BAL;
  const size_t lds_e = eval_lds(K, waves);
  const size_t lds_a = apply_lds(K);
  auto apply = apply_indexed(K) ? k_apply<true> : k_apply<false>;
  const size_t lds_s = (size_t)K * sizeof(int);
  // every pass restarts its round tags: clear the stamps and the touched counters
  pass_begin(ctx, nullptr, 0, nullptr);
  unsigned tag = 0;
  for (int r0 = 0; r0 < nov; r0 += R) {
    const int nslots = std::min(R, nov - r0);
    const int ntiles = (nslots + TILE - 1) / TILE;
    tag++;
    const int par = (int)(tag & 1u);
    EvalArgs a = make_eval_args(ctx, sched);
    a.r0 = r0;
    a.nslots = nslots;
    a.ntiles = ntiles;
    a.tag = tag;
    a.net_stamp = ctx->net_stamp;
    a.touched = ctx->touched;
    a.tcount = ctx->counters + par;
    a.tcount_next = ctx->counters + (par ^ 1);
    if (bal) {
      a.cand_tot = ctx->cand_tot;
      a.hist = ctx->hist;
      a.sizes_frozen = ctx->sizes_frozen;
    }
    int grid = eval_grid(nslots, waves);
    if (p.eval_grid > 0) grid = std::min(grid, p.eval_grid);
    FP_TRY(launch_timed(ctx, k_tv_eval<EVAL_REFINE>, grid, waves * WAVE, lds_e, a));
    if (bal) {
      k_bal_scan<<<std::max(1, std::min(K, 512)), 256, lds_s, ctx->stream>>>(
          ctx->hist, ctx->pref, ntiles, K, ctx->sizes_frozen, ctx->cand_tot, sched->cap, ctx->over_idx,
          ctx->over_count);
      // the size changes are summed per block in LDS where K parts fit
      const int agrid = std::max(1, std::min((ntiles + 3) / 4, 4096));
      if ((size_t)K * sizeof(long long) <= (size_t)LDS_BUDGET)
        k_bal_admit<true><<<agrid, 256, (size_t)K * sizeof(long long), ctx->stream>>>(
            ctx->slot_dec, ctx->slot_own, ctx->order, ctx->vwgt, r0, nslots, ntiles, K, ctx->sizes_frozen, sched->cap,
            false, ctx->over_idx, ctx->over_count, ctx->pref, ctx->hist, ctx->part, ctx->sizes, ctx->cand_tot, ctx->moves);
      else
        k_bal_admit<false><<<agrid, 256, 0, ctx->stream>>>(
            ctx->slot_dec, ctx->slot_own, ctx->order, ctx->vwgt, r0, nslots, ntiles, K, ctx->sizes_frozen, sched->cap,
            false, ctx->over_idx, ctx->over_count, ctx->pref, ctx->hist, ctx->part, ctx->sizes, ctx->cand_tot, ctx->moves);
    }
    ApplyArgs ap{};
    ap.net_ptr = ctx->net_ptr;
    ap.apply_pins = ctx->apply_pins;
    ap.slot_own = ctx->slot_own;
    ap.slot_dec = ctx->slot_dec;
    ap.touched = ctx->touched;
    ap.tcount = ctx->counters + par;
    ap.ent = ctx->ent;
    ap.ev = ctx->ev;
    ap.K = K;
    ap.pb = ctx->pb;
    ap.pmask = ctx->pmask;
    ap.r0 = r0;
    ap.r1 = r0 + nslots;
    ap.net_blocks = net_grid(ctx, APPLY_NET_BLOCKS, APPLY_BLOCK / WAVE);
    int grid_a = ap.net_blocks;
    if (!bal) {
      ap.order = ctx->order;
      ap.part = ctx->part;
      ap.sizes = ctx->sizes;
      ap.moves = ctx->moves;
      grid_a += (nslots + APPLY_BLOCK - 1) / APPLY_BLOCK;
    }
    apply<<<grid_a, APPLY_BLOCK, lds_a, ctx->stream>>>(ap);
  }
  return FP_OK;
}

fp_status plan_refine(fp_ctx* ctx, fp_metric metric, const fp_schedule* sched, RefinePlan& p) {
  if (!ctx->has_state) return fail(ctx, FP_ERR_STATE, "refinement before a state was set");
  if (ctx->dup_pins) return fail(ctx, FP_ERR_UNSUPPORTED, "a net lists the same vertex twice; refinement needs distinct pins");
  const int waves = p.eval_waves, grid = p.eval_grid;
  p = RefinePlan{};
  p.metric = metric;
  p.eval_waves = waves;
  p.eval_grid = grid;
  FP_HIP_OK(hipSetDevice(ctx->device));
  if (metric == FP_METRIC_CN) return plan_cn(ctx, p);  // the lambda = 2 push takes no schedule
  FP_TRY(check_sched(ctx, sched));
  p.sched = *sched;
  FP_TRY(check_envelope(ctx));
  switch (metric) {
    case FP_METRIC_TV: return plan_tv(ctx, p);
    case FP_METRIC_MSV:
    case FP_METRIC_MSRV: return plan_vol(ctx, p);
    case FP_METRIC_TM:
    case FP_METRIC_MSM: return plan_conn(ctx, p);
    default: return fail(ctx, FP_ERR_UNSUPPORTED, "refinement for metric %d not available", (int)metric);
  }
}

fp_status record_pass(fp_ctx* ctx, const RefinePlan& p) {
  switch (p.metric) {
    case FP_METRIC_TV: return pass_tv(ctx, p);
    case FP_METRIC_CN: return pass_cn(ctx, p);
    case FP_METRIC_MSV:
    case FP_METRIC_MSRV: return pass_vol(ctx, p);
    default: return pass_conn(ctx, p);
  }
}

}  // namespace fpe

using namespace fpe;

extern "C" {

fp_status fp_refine(fp_ctx* ctx, fp_metric metric, int passes, const fp_schedule* sched, fp_refine_stats* stats) {
  if (!ctx) return FP_ERR_ARG;
  if (passes < 0) return fail(ctx, FP_ERR_ARG, "passes < 0");
  RefinePlan p;
  FP_TRY(plan_refine(ctx, metric, sched, p));
  fill_async<unsigned long long>(ctx, ctx->moves, 1, 0ull);
  ctx->timed_next = 2;
  ctx->timed_launches = 0;
  ctx->timed_ms = 0;
  hipEvent_t e0 = pooled_event(ctx, 0), e1 = pooled_event(ctx, 1);
  if (!e0 || !e1) return fail(ctx, FP_ERR_HIP, "hipEventCreate failed");
  FP_HIP_OK(hipEventRecord(e0, ctx->stream));
  if (metric == FP_METRIC_CN) {
    FP_TRY(cn_passes(ctx, p, passes));
  } else {
    for (int it = 0; it < passes; it++) FP_TRY(record_pass(ctx, p));
  }
  FP_HIP_OK(hipGetLastError());
  FP_HIP_OK(hipEventRecord(e1, ctx->stream));
  unsigned long long mv = 0;
  FP_TRY(d2h(ctx, &mv, ctx->moves, 1));
  FP_HIP_OK(hipStreamSynchronize(ctx->stream));
  float ms = 0;
  FP_HIP_OK(hipEventElapsedTime(&ms, e0, e1));
  for (size_t i = 2; i + 1 < ctx->timed_next; i += 2) {
    float t = 0;
    FP_HIP_OK(hipEventElapsedTime(&t, ctx->ev_pool[i], ctx->ev_pool[i + 1]));
    ctx->timed_ms += t;
    ctx->timed_launches++;
  }
  if (stats) {
    stats->moves = (long long)mv;
    stats->rounds = metric == FP_METRIC_CN ? passes : (long long)passes * ((ctx->nov + p.R - 1) / p.R);
    stats->pin_visits = (long long)passes * ctx->pins;
    stats->device_ms = ms;
  }
  return FP_OK;
}

fp_status fp_initial_partition_device(fp_ctx* ctx, int mode, int K, int round, double C, long long max_pw,
                                      long long cap, fp_init_stats* stats) {
  if (!ctx) return FP_ERR_ARG;
  if (!ctx->has_graph) return fail(ctx, FP_ERR_STATE, "fp_initial_partition_device before fp_upload");
  if (mode < 0 || mode > 2) return fail(ctx, FP_ERR_ARG, "unknown initial partition mode %d", mode);
  if (round < 1) return fail(ctx, FP_ERR_ARG, "round size must be >= 1");
  if (!(C > 0)) return fail(ctx, FP_ERR_ARG, "C must be > 0");
  if (ctx->dup_pins) return fail(ctx, FP_ERR_UNSUPPORTED, "a net lists the same vertex twice");
  FP_HIP_OK(hipSetDevice(ctx->device));
  FP_TRY(ensure_k(ctx, K));
  FP_TRY(check_envelope(ctx));
  if (apply_lds(K) > (size_t)LDS_BUDGET) return fail(ctx, FP_ERR_UNSUPPORTED, "K = %d too large for the apply kernel", K);
  const int nov = ctx->nov;
  const int R = std::max(1, std::min(round, nov));
  const int waves = eval_waves(K);
  FP_TRY(ensure_round_scratch(ctx, R));
  FP_TRY(ensure_bal_scratch(ctx, R));
  // the empty state: no part, empty lists, zero sizes
  const size_t segs = (size_t)ctx->pins + (size_t)ctx->noe + ENT_PAD;
  fill_async<int>(ctx, ctx->part, (size_t)std::max(nov, 1), -1);
  fill_async<unsigned>(ctx, ctx->ent, segs, 0u);
  fill_async<unsigned short>(ctx, ctx->ev, segs, (unsigned short)0);
  fill_async<long long>(ctx, ctx->sizes, (size_t)K, 0);
  fill_async<long long>(ctx, ctx->cand_tot, (size_t)K, 0);
  fill_async<unsigned long long>(ctx, ctx->moves, 1, 0ull);
  pass_begin(ctx, nullptr, 0, nullptr);
  hipEvent_t e0 = pooled_event(ctx, 0), e1 = pooled_event(ctx, 1);
  if (!e0 || !e1) return fail(ctx, FP_ERR_HIP, "hipEventCreate failed");
  FP_HIP_OK(hipEventRecord(e0, ctx->stream));
  const int eval_mode = mode == 2 ? EVAL_INIT_CN : EVAL_INIT;
  const size_t lds_e = eval_lds(K, waves);
  const size_t lds_a = apply_lds(K);
  auto apply = apply_indexed(K) ? k_apply<true> : k_apply<false>;
  const bool agg = (size_t)K * sizeof(long long) <= (size_t)LDS_BUDGET;
  int* fb = ctx->counters + 2;
  unsigned tag = 0;
  long long subrounds = 0, rounds = 0;
  int burst = 1;  // sub-rounds enqueued before the host looks at the progress
  auto subround = [&](int r0, int nslots) -> fp_status {
    const int ntiles = (nslots + TILE - 1) / TILE;
    tag++;
    const int par = (int)(tag & 1u);
    EvalArgs a{};
    a.pad_seg = (unsigned)((size_t)ctx->pins + (size_t)ctx->noe);
    a.vtx_ptr = ctx->vtx_ptr;
    a.vtx_nets = ctx->vtx_nets;
    a.vtx_seg = ctx->vtx_seg;
    a.ev = ctx->ev;
    a.part = ctx->part;
    a.order = ctx->order;
    a.vwgt = ctx->vwgt;
    a.sizes = ctx->sizes;
    a.K = K;
    a.pb = ctx->pb;
    a.pmask = ctx->pmask;
    a.C = C;
    a.max_pw = mode == 0 ? max_pw : LLONG_MAX;
    a.min_pw = LLONG_MIN;
    a.slot_own = ctx->slot_own;
    a.slot_dec = ctx->slot_dec;
    a.r0 = r0;
    a.nslots = nslots;
    a.ntiles = ntiles;
    a.tag = tag;
    a.net_stamp = ctx->net_stamp;
    a.touched = ctx->touched;
    a.tcount = ctx->counters + par;
    a.tcount_next = ctx->counters + (par ^ 1);
    a.cand_tot = ctx->cand_tot;
    a.hist = ctx->hist;
    a.sizes_frozen = ctx->sizes_frozen;
    a.fb_out = fb;
    const int grid = eval_grid(nslots, waves);
    if (eval_mode == EVAL_INIT_CN)
      k_tv_eval<EVAL_INIT_CN><<<grid, waves * WAVE, lds_e, ctx->stream>>>(a);
    else
      k_tv_eval<EVAL_INIT><<<grid, waves * WAVE, lds_e, ctx->stream>>>(a);
    k_init_fallback<<<1, 1024, 0, ctx->stream>>>(ctx->slot_dec, ctx->order, ctx->vwgt, r0, nslots, ntiles, K, fb,
                                                 ctx->cand_tot, ctx->hist);
    k_bal_scan<<<std::max(1, std::min(K, 512)), 256, (size_t)K * sizeof(int), ctx->stream>>>(
        ctx->hist, ctx->pref, ntiles, K, ctx->sizes_frozen, ctx->cand_tot, cap, ctx->over_idx, ctx->over_count);
    const int agrid = std::max(1, std::min((ntiles + 3) / 4, 4096));
    if (agg)
      k_bal_admit<true><<<agrid, 256, (size_t)K * sizeof(long long), ctx->stream>>>(
          ctx->slot_dec, ctx->slot_own, ctx->order, ctx->vwgt, r0, nslots, ntiles, K, ctx->sizes_frozen, cap, true,
          ctx->over_idx, ctx->over_count, ctx->pref, ctx->hist, ctx->part, ctx->sizes, ctx->cand_tot, ctx->moves);
    else
      k_bal_admit<false><<<agrid, 256, 0, ctx->stream>>>(
          ctx->slot_dec, ctx->slot_own, ctx->order, ctx->vwgt, r0, nslots, ntiles, K, ctx->sizes_frozen, cap, true,
          ctx->over_idx, ctx->over_count, ctx->pref, ctx->hist, ctx->part, ctx->sizes, ctx->cand_tot, ctx->moves);
    ApplyArgs ap{};
    ap.net_ptr = ctx->net_ptr;
    ap.apply_pins = ctx->apply_pins;
    ap.slot_own = ctx->slot_own;
    ap.slot_dec = ctx->slot_dec;
    ap.touched = ctx->touched;
    ap.tcount = ctx->counters + par;
    ap.ent = ctx->ent;
    ap.ev = ctx->ev;
    ap.K = K;
    ap.pb = ctx->pb;
    ap.pmask = ctx->pmask;
    ap.r0 = r0;
    ap.r1 = r0 + nslots;
    ap.net_blocks = net_grid(ctx, APPLY_NET_BLOCKS, APPLY_BLOCK / WAVE);
    apply<<<ap.net_blocks, APPLY_BLOCK, lds_a, ctx->stream>>>(ap);
    FP_HIP_OK(hipGetLastError());
    subrounds++;
    return FP_OK;
  };
  for (int r0 = 0; r0 < nov; r0 += R) {
    const int nslots = std::min(R, nov - r0);
    rounds++;
    // every sub-round places at least one vertex of the round: a round takes
    // at most nslots of them
    int launched = 0;
    for (;;) {
      const int n = std::min(burst, nslots - launched);
      for (int i = 0; i < n; i++) FP_TRY(subround(r0, nslots));
      launched += n;
      unsigned long long placed = 0;
      FP_TRY(d2h(ctx, &placed, ctx->moves, 1));
      FP_HIP_OK(hipStreamSynchronize(ctx->stream));
      if (placed == (unsigned long long)r0 + (unsigned long long)nslots) break;
      if (launched >= nslots)
        return fail(ctx, FP_ERR_STATE, "initial partition round at %d did not place its vertices", r0);
    }
    // the next round starts with as many sub-rounds as this one took (one
    // fewer when this one was done at its first look, to probe downwards)
    burst = launched == burst ? std::max(1, burst - 1) : launched;
  }
  FP_HIP_OK(hipEventRecord(e1, ctx->stream));
  FP_HIP_OK(hipStreamSynchronize(ctx->stream));
  ctx->has_state = true;
  if (stats) {
    float ms = 0;
    FP_HIP_OK(hipEventElapsedTime(&ms, e0, e1));
    stats->rounds = rounds;
    stats->subrounds = subrounds;
    stats->device_ms = ms;
  }
  return FP_OK;
}

fp_status fp_eval_moves(fp_ctx* ctx, fp_metric metric, const fp_schedule* sched, int* dec, int* gain) {
  if (!ctx) return FP_ERR_ARG;
  if (!ctx->has_state) return fail(ctx, FP_ERR_STATE, "fp_eval_moves before a state was set");
  FP_TRY(check_sched(ctx, sched));
  if (metric != FP_METRIC_TV) return fail(ctx, FP_ERR_UNSUPPORTED, "move evaluation for metric %d not available", (int)metric);
  const int nov = ctx->nov, K = ctx->K;
  FP_TRY(check_envelope(ctx));
  const int waves = eval_waves(K);
  FP_HIP_OK(hipSetDevice(ctx->device));
  FP_TRY(ensure_round_scratch(ctx, nov));
  int *d_dec = nullptr, *d_gain = nullptr;
  FP_HIP_OK(hipMalloc(&d_dec, sizeof(int) * (size_t)nov));
  if (hipMalloc(&d_gain, sizeof(int) * (size_t)nov) != hipSuccess) {
    (void)hipFree(d_dec);
    return fail(ctx, FP_ERR_OOM, "hipMalloc failed");
  }
  EvalArgs a = make_eval_args(ctx, sched);
  a.r0 = 0;
  a.nslots = nov;
  a.ntiles = (nov + TILE - 1) / TILE;
  a.dec_out = d_dec;
  a.gain_out = d_gain;
  k_tv_eval<EVAL_REFINE><<<eval_grid(nov, waves), waves * WAVE, eval_lds(K, waves), ctx->stream>>>(a);
  fp_status st = FP_OK;
  if (hipGetLastError() != hipSuccess) st = fail(ctx, FP_ERR_HIP, "k_tv_eval launch failed");
  if (st == FP_OK && dec) st = d2h(ctx, dec, d_dec, (size_t)nov);
  if (st == FP_OK && gain) st = d2h(ctx, gain, d_gain, (size_t)nov);
  if (hipStreamSynchronize(ctx->stream) != hipSuccess && st == FP_OK) st = fail(ctx, FP_ERR_HIP, "sync failed");
  (void)hipFree(d_dec);
  (void)hipFree(d_gain);
  return st;
}

}  // extern "C"
