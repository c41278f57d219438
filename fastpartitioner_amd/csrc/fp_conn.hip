// fp_conn.hip -- TM and MSM refinement on the device: the round schedules of
// partition_par in TM/partition.cpp:947-1265 (total messages) and
// MSM/partition.cpp:1010-1400 (max send messages).
//
// Both keep the part-connection matrix pc[a][b] = number of nets whose source
// is in part a and that have pins in part b != a; msg[a] = number of b with
// pc[a][b] > 0. Besides pc (int, K x K) the engine keeps its nonzero pattern
// as two bitmaps, B (rows a) and BT (rows b), so that the loss terms of a
// candidate part k -- "receiver r with pc[k][r] == 0" and "sender s with
// pc[s][k] == 0" -- are one bit of a 64-bit word per (r or s, 64 parts k).
//
// A TM round (the reference runs a whole pass as one round):
//   k_conn_eval   one wavefront per vertex: receivers, senders and gain on the
//                 frozen state, then the best part over all k (lanes = k)
//   k_round_compact the candidates in vertex order, with their weights
//   k_tm_admit    one wavefront: candidates in vertex order, skipping a
//                 destination whose size reached C * 1.02 (TM/partition.cpp:
//                 1115-1122), decided 64 at a time by size bounds where they
//                 allow; commits part vector and sizes
//   k_conn_apply  (mode 0) one wavefront per touched net replays the admitted
//                 moves on its list and adds each move's pc changes (atomics:
//                 pc's final value does not depend on their order)
//   k_conn_bits   B, BT and msg from pc
// An MSM round:
//   k_conn_eval, k_round_compact
//   k_conn_apply  (mode 1, dry) per touched net and move, the pc changes the
//                 reference's apply makes, into per-incidence event slots
//   k_msm_scan    one wavefront applies the moves' pc changes in order, counts
//                 each part's messages from the 0 <-> 1 crossings, and stops
//                 after the first move whose max-message set gains a part
//   k_conn_apply  (mode 2, commit) list edits, part vector and sizes up to the stop
//   k_conn_bits
#include "fp_round.h"

namespace fpe {
namespace {

constexpr int CONN_BLOCK = 256;
constexpr int CONN_KMAX = 4096;
constexpr int REPLAY_BLOCK = 256;
constexpr int REPLAY_NET_BLOCKS = 2048;

// one wavefront per net: pc[source part][listed part] += 1 for listed parts
// other than the source's
__global__ void __launch_bounds__(256) k_conn_init(const unsigned* __restrict__ net_ptr,
                                                   const unsigned* __restrict__ ent, const int* __restrict__ src,
                                                   const int* __restrict__ part, int noe, unsigned pmask, int K,
                                                   int* __restrict__ pc) {
  const int lane = lane_id();
  const int nw = (int)(gridDim.x * blockDim.x / WAVE);
  for (int n = (int)((blockIdx.x * blockDim.x + threadIdx.x) / WAVE); n < noe; n += nw) {
    const unsigned seg = net_ptr[n] + (unsigned)n;
    const int L = (int)ent[seg];
    if (L < 2) continue;
    const int sp = part[src[n]];
    for (int t = lane; t < L; t += WAVE) {
      const int q = (int)(ent[seg + 1 + (unsigned)t] & pmask);
      if (q != sp) atomicAdd(&pc[(size_t)sp * (size_t)K + (size_t)q], 1);
    }
  }
}

// One block (four wavefronts) per 64 x 64 tile of pc: the tile's words of B
// (wavefront 0) and BT (wavefront 1), and the popcounts of its B words into
// msg (zeroed before; msg null: left as it is).
__global__ void __launch_bounds__(256) k_conn_bits(const int* __restrict__ pc, int K, int W,
                                                    unsigned long long* __restrict__ B,
                                                    unsigned long long* __restrict__ BT, int* __restrict__ msg) {
  __shared__ int tile[WAVE][WAVE + 1];
  const int lane = lane_id(), wave = (int)(threadIdx.x / WAVE);
  const int ta = (int)blockIdx.y, tb = (int)blockIdx.x;
  for (int r = wave; r < WAVE; r += (int)(blockDim.x / WAVE)) {
    const int a = ta * WAVE + r, b = tb * WAVE + lane;
    tile[r][lane] = (a < K && b < K) ? pc[(size_t)a * (size_t)K + (size_t)b] : 0;
  }
  __syncthreads();
  if (wave == 0) {
    // row a = ta*64 + lane, columns b = tb*64 + c
    unsigned long long row = 0;
    for (int c = 0; c < WAVE; c++)
      if (tile[lane][c] > 0 && ta * WAVE + lane != tb * WAVE + c) row |= 1ull << c;
    const int a = ta * WAVE + lane;
    if (a < K) {
      B[(size_t)a * (size_t)W + (size_t)tb] = row;
      if (row && msg) atomicAdd(&msg[a], __popcll(row));
    }
  } else if (wave == 1) {
    // row b = tb*64 + lane of the transpose, columns a = ta*64 + c
    unsigned long long col = 0;
    for (int c = 0; c < WAVE; c++)
      if (tile[c][lane] > 0 && ta * WAVE + c != tb * WAVE + lane) col |= 1ull << c;
    const int b = tb * WAVE + lane;
    if (b < K) BT[(size_t)b * (size_t)W + (size_t)ta] = col;
  }
}

struct ConnEvalArgs {
  const unsigned* vtx_ptr;
  const unsigned* vtx_nets;
  const unsigned* vtx_seg;
  const unsigned short* ev;
  const int* src;
  const int* part;
  const long long* sizes;
  const int* pc;
  const unsigned long long* B;
  const unsigned long long* BT;
  const int* msg;
  const int* mstat;
  int K, W, msm;
  int rcap, scap;  // LDS entries per wavefront for the receiver and sender lists
  int r0, nslots;
  double C;
  int* slot_own;
  int* slot_dec;
  unsigned tag;
  unsigned* net_stamp;
  int* touched;
  int* tcount;
  int* tcount_next;
};

__device__ __forceinline__ bool bit_of(const unsigned* m, int p) { return (m[p >> 5] >> (p & 31)) & 1u; }

// Per vertex (TM/partition.cpp:978-1084; MSM/partition.cpp:1058-1190):
// receivers = parts v's source nets reach after v leaves (own part if it
// keeps pins there, every other listed part); senders = source parts of v's
// other nets; gain = source-net targets t with pc[own][t] == 1 plus, for the
// FIRST net of each sender part s, 1 if v is own's only pin there and
// pc[s][own] == 1. TM: best k by gain - loss; MSM: the max-set rule.
__global__ void __launch_bounds__(CONN_BLOCK) __attribute__((amdgpu_waves_per_eu(8))) k_conn_eval(ConnEvalArgs a) {
  extern __shared__ __align__(16) unsigned char lds[];
  __shared__ int s_tn[CONN_BLOCK / WAVE + 1];
  const int K = a.K, KW = (K + 31) / 32;
  const int lane = lane_id(), wave = wave_id(), nw = (int)(blockDim.x / WAVE);
  // block-shared: bit k of vmask = part k is lighter than C (TM's candidates)
  unsigned long long* vmask = reinterpret_cast<unsigned long long*>(lds);
  const size_t per = (size_t)(a.rcap + a.scap + 3 * KW);
  int* recv = reinterpret_cast<int*>(vmask + a.W) + (size_t)wave * per;
  int* send = recv + a.rcap;
  unsigned* rmark = reinterpret_cast<unsigned*>(send + a.scap);
  unsigned* smark = rmark + KW;
  unsigned* sgain = smark + KW;
  // per wavefront: the nets its moving vertices touch, before they go to the touched list
  WaveTouch tb{reinterpret_cast<unsigned*>(vmask + a.W) + (size_t)nw * per +
                   (size_t)wave * TOUCH_BUF,
               0};
  for (int q = lane; q < KW; q += WAVE) {
    rmark[q] = 0;
    smark[q] = 0;
    sgain[q] = 0;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0 && a.tcount_next) *a.tcount_next = 0;
  if (!a.msm)
    for (int w = wave; w < a.W; w += nw) {
      const int k = w * WAVE + lane;
      const unsigned long long b = __ballot(k < K && (double)a.sizes[k] < a.C);
      if (lane == 0) vmask[w] = b;
    }
  __syncthreads();
  const int M = a.msm ? a.mstat[0] : 0, cntM = a.msm ? a.mstat[1] : 0;
  const unsigned long long lt = (1ull << lane) - 1ull;

  for (int slot = (int)blockIdx.x * nw + wave; slot < a.nslots; slot += (int)gridDim.x * nw) {
    const int v = a.r0 + slot;
    const int own = a.part[v];
    const unsigned vb = a.vtx_ptr[v];
    const int deg = (int)(a.vtx_ptr[v + 1] - vb);
    int nrecv = 0, nsend = 0, gain_own = 0, gain_s = 0;
    // lane j takes net j of each chunk of v's nets: its segment, source and
    // source part; the lists are swept with several nets in flight
    for (int c0 = 0; c0 < deg; c0 += WAVE) {
      const int nn = min(WAVE, deg - c0);
      unsigned sg = 0;
      int ps = -1;
      bool isrc = false;
      if (lane < nn) {
        const unsigned k = vb + (unsigned)(c0 + lane);
        sg = a.vtx_seg[k];
        const int s = a.src[a.vtx_nets[k]];
        isrc = s == v;
        if (!isrc) ps = a.part[s];
      }
      const unsigned long long srcmask = __ballot(isrc);
      unsigned long long smask = 0;  // nets where v is own's only pin
      wave_sweep_lists<8, SWEEP_LONG_TASKS>(a.ev, sg, nn, [&](int j, int, int, bool ok, unsigned x) {
        const int q = (int)(x & ~EV_SINGLE);
        if ((srcmask >> j) & 1ull) {
          // v's own source net: receivers are the listed parts other than own,
          // and own if it keeps other pins there
          const bool rec = ok && (q != own || !(x & EV_SINGLE));
          bool isnew = false;
          if (rec) {
            const unsigned bit = 1u << (q & 31);
            isnew = (atomicOr(&rmark[q >> 5], bit) & bit) == 0u;
          }
          const unsigned long long bm = __ballot(isnew);
          const int at = nrecv + __popcll(bm & lt);
          if (isnew && at < a.rcap) recv[at] = q;
          nrecv += __popcll(bm);
          if (ok && q != own && a.pc[(size_t)own * (size_t)K + (size_t)q] == 1) gain_own++;
        } else if (__ballot(ok && q == own && (x & EV_SINGLE))) {
          smask |= 1ull << j;
        }
      });
      // senders: the first net of each source part in v's net order
      bool first = lane < nn && !isrc && !bit_of(smark, ps);
      for (unsigned long long m = ~srcmask & (nn == WAVE ? ~0ull : ((1ull << nn) - 1ull)); m; m &= m - 1) {
        const int jj = __builtin_amdgcn_readfirstlane(__ffsll((long long)m) - 1);
        if (lane > jj && ps == __builtin_amdgcn_readlane(ps, jj)) first = false;
      }
      wave_lds_sync();
      const unsigned long long fm = __ballot(first);
      if (first) {
        const int at = nsend + __popcll(fm & lt);
        if (at < a.scap) send[at] = ps;
        atomicOr(&smark[ps >> 5], 1u << (ps & 31));
        if (((smask >> lane) & 1ull) && a.pc[(size_t)ps * (size_t)K + (size_t)own] == 1) {
          gain_s++;
          atomicOr(&sgain[ps >> 5], 1u << (ps & 31));
        }
      }
      nsend += __popcll(fm);
      wave_lds_sync();
    }
    gain_s = wave_sum(gain_s);
    gain_own = wave_sum(gain_own);
    const int gain = gain_own + gain_s;
    wave_lds_sync();
    int dec = -1;
    if (gain > 0) {
      // The items' bitmap rows are read 64 items at a time, one item per lane,
      // so that the loads are in flight together; a lane (a part k of the 64
      // being scored) then reads the words from the lanes that hold them.
      // Losses of k: the items i != k whose row word has bit k clear.
      auto count_losses = [&](const int* items, int nitems, const unsigned long long* rows, int w, int k) {
        int loss = 0;
        for (int i0 = 0; i0 < nitems; i0 += WAVE) {
          const int n = min(WAVE, nitems - i0);
          int it = -1;
          unsigned long long x = ~0ull;
          if (lane < n) {
            it = items[i0 + lane];
            x = rows[(size_t)it * (size_t)a.W + (size_t)w];
          }
          for (int i = 0; i < n; i++) {
            const int r = __builtin_amdgcn_readlane(it, i);
            const unsigned long long xi = readlane_u64(x, i);
            loss += (r != k && !((xi >> lane) & 1ull)) ? 1 : 0;
          }
        }
        return loss;
      };
      if (!a.msm) {
        // TM: the largest gain - loss > 0 over parts lighter than C, the first
        // k on ties. A part connected to every receiver and every sender has
        // loss 0, so improvement = gain, the largest possible: the first such
        // candidate is found by ANDing their bitmap words, 64 parts at a time
        // (an item's own bit counts as connected: its loss term skips k ==
        // the item)
        auto and_rows = [&](const int* items, int nitems, const unsigned long long* rows, int w) {
          unsigned long long all = ~0ull;
          for (int i0 = 0; i0 < nitems; i0 += WAVE) {
            unsigned long long x = ~0ull;
            if (i0 + lane < nitems) {
              const int it = items[i0 + lane];
              x = rows[(size_t)it * (size_t)a.W + (size_t)w];
              if ((it >> 6) == w) x |= 1ull << (it & 63);
            }
            all &= wave_and_u64(x);
          }
          return all;
        };
        int kfast = -1;
        for (int w = 0; w < a.W && kfast < 0; w++) {
          unsigned long long allc = vmask[w];
          if ((own >> 6) == w) allc &= ~(1ull << (own & 63));
          if (allc) allc &= and_rows(recv, nrecv, a.BT, w);
          if (allc) allc &= and_rows(send, nsend, a.B, w);
          if (allc) kfast = w * WAVE + __builtin_ctzll(allc);
        }
        int bimp = 0, bk = INT_MAX;
        if (kfast >= 0) {
          bimp = gain;
          bk = kfast;
        }
        for (int k0 = 0; kfast < 0 && k0 < K; k0 += WAVE) {
          const int k = k0 + lane, w = k0 / WAVE;
          const int loss = count_losses(recv, nrecv, a.BT, w, k) + count_losses(send, nsend, a.B, w, k);
          const bool valid = k < K && k != own && (double)a.sizes[k] < a.C;
          if (valid && gain - loss > bimp) {
            bimp = gain - loss;
            bk = k;
          }
        }
        int mimp = bimp, dummy = 0;
        wave_max_count(mimp, dummy);
        if (mimp > 0) {
          const unsigned long long cand = __ballot(bimp == mimp);
          int kk = INT_MAX;
          if ((cand >> lane) & 1ull) kk = bk;
          for (int off = WAVE / 2; off > 0; off >>= 1) kk = min(kk, __shfl_xor(kk, off));
          dec = kk;
        }
      } else {
        // MSM: every max part must gain strictly more than it loses, no other
        // part may reach the max; improvement = sum over the max parts.
        // The max parts among the items number X (own and the senders) plus
        // one if k is a max part outside them, and must be all cntM of them:
        // with X + 1 < cntM no k passes
        const bool own_s = bit_of(smark, own);
        const int g_own = gain_own + ((own_s && bit_of(sgain, own)) ? 1 : 0);
        int X = 0;
        for (int i = lane; i < nsend; i += WAVE) {
          const int sp = send[i];
          X += (sp != own && a.msg[sp] == M) ? 1 : 0;
        }
        X = wave_sum(X) + (a.msg[own] == M ? 1 : 0);
        const int m_own = a.msg[own];
        int bimp = -1, bk = INT_MAX;
        for (int k0 = 0; X + 1 >= cntM && k0 < K; k0 += WAVE) {
          const int k = k0 + lane, w = k0 / WAVE;
          const bool valid = k < K && k != own;
          const int kc = valid ? k : own;
          // receivers r with pc[k][r] == 0 (the diagonal never counts)
          const int lk = count_losses(recv, nrecv, a.BT, w, kc);
          bool bad = false;
          int imp = 0, nmax = 0;
          auto item = [&](int mp, int g, int l) {
            if (mp == M) {
              if (g <= l) bad = true;
              imp += g - l;
              nmax++;
            } else if (mp + l - g >= M) {
              bad = true;
            }
          };
          item(m_own, g_own, own_s ? (int)!((a.B[(size_t)own * (size_t)a.W + (size_t)w] >> lane) & 1ull) : 0);
          const bool k_s = valid && bit_of(smark, kc);
          item(a.msg[kc], (k_s && bit_of(sgain, kc)) ? 1 : 0, lk);
          // the senders other than own and k, 64 at a time: part, its
          // messages, its gain bit and its B row word
          for (int i0 = 0; i0 < nsend; i0 += WAVE) {
            const int n = min(WAVE, nsend - i0);
            int sl = -1, ml = 0, gl = 0;
            unsigned long long xl = ~0ull;
            if (lane < n) {
              sl = send[i0 + lane];
              ml = a.msg[sl];
              gl = bit_of(sgain, sl) ? 1 : 0;
              xl = a.B[(size_t)sl * (size_t)a.W + (size_t)w];
            }
            for (int i = 0; i < n; i++) {
              const int sp = __builtin_amdgcn_readlane(sl, i);
              if (sp == own) continue;
              const unsigned long long xi = readlane_u64(xl, i);
              if (sp != kc)
                item(__builtin_amdgcn_readlane(ml, i), __builtin_amdgcn_readlane(gl, i), (int)!((xi >> lane) & 1ull));
            }
          }
          if (nmax != cntM) bad = true;
          if (valid && !bad && imp > bimp) {
            bimp = imp;
            bk = k;
          }
        }
        int mimp = bimp, dummy = 0;
        wave_max_count(mimp, dummy);
        if (mimp > -1) {
          const unsigned long long cand = __ballot(bimp == mimp);
          int kk = INT_MAX;
          if ((cand >> lane) & 1ull) kk = bk;
          for (int off = WAVE / 2; off > 0; off >>= 1) kk = min(kk, __shfl_xor(kk, off));
          dec = kk;
        }
      }
    }
    if (lane == 0) {
      a.slot_own[slot] = own;
      a.slot_dec[slot] = dec;
    }
    if (dec >= 0) wave_touch_nets_staged(a.vtx_nets + vb, deg, a.tag, a.net_stamp, tb, a.touched, a.tcount);
    // clear this vertex's marks
    wave_lds_sync();
    for (int i = lane; i < nrecv; i += WAVE) rmark[recv[i] >> 5] = 0;
    for (int i = lane; i < nsend; i += WAVE) {
      smark[send[i] >> 5] = 0;
      sgain[send[i] >> 5] = 0;
    }
    wave_lds_sync();
  }
  block_touch_flush(tb, a.touched, a.tcount, s_tn);
}

// TM apply order (TM/partition.cpp:1111-1125): the round's candidates (the
// compacted move list) in vertex order; one whose destination size is at
// least thr = C * 1.02 at that point is dropped. Commits part vector and both
// sizes of the others. Sizes are integers, so size < thr is size < T with
// T = ceil(thr). Chunks of 64 candidates: with in[p] / out[p] the chunk's
// candidate weight into / out of part p, a candidate into d is dropped
// whatever the chunk's other decisions are when sz[d] - out[d] >= T, and
// admitted when sz[d] + in[d] < T. The destinations of the undecided
// candidates are "open"; each open part gets an owner lane that holds its
// size in a register. The undecided candidates and the admitted ones that
// touch an open part are walked in order, reading and updating only those
// registers; then every admitted move of the chunk is added to the sizes.
// The moves (vertex, own part, new part, weight) are read TM_PREFETCH chunks
// at a time, the next group's loads in flight while a group is decided.
constexpr int TM_PREFETCH = 4;

__global__ void __launch_bounds__(WAVE) k_tm_admit(int* __restrict__ slot_dec, const int4* __restrict__ mv,
                                                   const int* __restrict__ mstat, int r0, int K, double thr,
                                                   long long* __restrict__ sizes, int* __restrict__ part,
                                                   unsigned long long* __restrict__ moves) {
  extern __shared__ __align__(16) unsigned char lds[];
  long long* sz = reinterpret_cast<long long*>(lds);
  int* inw = reinterpret_cast<int*>(sz + K);
  int* outw = inw + K;
  const int lane = lane_id();
  const long long T = thr < 9.0e18 ? (long long)ceil(thr) : LLONG_MAX;
  for (int q = lane; q < K; q += WAVE) {
    sz[q] = sizes[q];
    inw[q] = 0;
    outw[q] = 0;
  }
  wave_lds_sync();
  const int nmv = mstat[2];
  unsigned long long nm = 0;
  // decides one chunk in LDS and registers only; returns this lane's decision
  auto chunk = [&](int m0, int4 x) -> bool {
    const bool valid = m0 + lane < nmv;
    const int v = x.x, o = x.y, d = x.z, w = valid ? x.w : 0;
    // chunk sums stay exact in int for weights up to 2^24; otherwise every
    // candidate of the chunk stays undecided
    const bool bounded = __ballot(valid && (unsigned)w > (1u << 24)) == 0ull;
    int status = 0;  // 1 admitted, -1 dropped, 0 undecided
    if (bounded) {
      if (valid) {
        atomicAdd(&inw[d], w);
        atomicAdd(&outw[o], w);
      }
      wave_lds_sync();
      if (valid) {
        const long long sd = sz[d];
        if (sd - (long long)outw[d] >= T) status = -1;
        else if (sd + (long long)inw[d] < T) status = 1;
      }
      wave_lds_sync();
      if (valid) {
        inw[d] = 0;
        outw[o] = 0;
      }
      wave_lds_sync();
    }
    // owners of the open parts: inw[p] = owner lane + 1 (0: p is not open)
    const bool und = valid && status == 0;
    if (und) inw[d] = lane + 1;
    wave_lds_sync();
    const int down = valid ? inw[d] : 0, oown = valid ? inw[o] : 0;
    long long mysz = und && down == lane + 1 ? sz[d] : 0;
    const bool walk = und || (valid && status > 0 && (down | oown) != 0);
    bool adm = valid && status > 0;
    const int pk = down | (oown << 8) | (und ? 1 << 16 : 0);
    for (unsigned long long b = __ballot(walk); b; b &= b - 1) {
      const int l = __builtin_amdgcn_readfirstlane((int)__builtin_ctzll(b));
      const int q = __builtin_amdgcn_readlane(pk, l), ww = __builtin_amdgcn_readlane(w, l);
      const int dl = (q & 0xFF) - 1, ol = ((q >> 8) & 0xFF) - 1;
      bool a = true;
      if (q >> 16) {  // undecided: its destination's size at this point decides
        const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(unsigned long long)mysz, dl);
        const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)((unsigned long long)mysz >> 32), dl);
        a = (long long)(((unsigned long long)hi << 32) | lo) < T;
      }
      if (a) {
        if (lane == dl) mysz += ww;
        if (lane == ol) mysz -= ww;
      }
      if (lane == l) adm = a;
    }
    wave_lds_sync();
    if (und) inw[d] = 0;
    if (valid && adm) {
      atomicAdd(reinterpret_cast<unsigned long long*>(&sz[o]), (unsigned long long)(-(long long)w));
      atomicAdd(reinterpret_cast<unsigned long long*>(&sz[d]), (unsigned long long)(long long)w);
    }
    nm += (unsigned long long)__popcll(__ballot(valid && adm));
    wave_lds_sync();
    return valid && adm;
  };
  // Two register buffers in turn, and a group's decisions written after the
  // group (a store per candidate into part and slot_dec, the unchanged value
  // where nothing changes), so that deciding a group issues no global memory
  // operation: the wait for the next group's moves then covers only
  // operations issued before them. Loads past the end read the sentinel.
  int4 ga[TM_PREFETCH], gb[TM_PREFETCH];
  auto load_group = [&](int g0, int4 (&dst)[TM_PREFETCH]) {
#pragma unroll
    for (int j = 0; j < TM_PREFETCH; j++) dst[j] = mv[min(g0 + j * WAVE + lane, nmv)];
  };
  auto run_group = [&](int g0, const int4 (&src)[TM_PREFETCH]) {
    bool adm[TM_PREFETCH];
#pragma unroll
    for (int j = 0; j < TM_PREFETCH; j++) adm[j] = g0 + j * WAVE < nmv ? chunk(g0 + j * WAVE, src[j]) : false;
#pragma unroll
    for (int j = 0; j < TM_PREFETCH; j++) {
      if (g0 + j * WAVE + lane < nmv) {
        const int4 x = src[j];
        part[x.x] = adm[j] ? x.z : x.y;
        slot_dec[x.x - r0] = adm[j] ? x.z : -1;
      }
    }
  };
  constexpr int G = TM_PREFETCH * WAVE;
  load_group(0, ga);
  for (int b0 = 0; b0 < nmv; b0 += 2 * G) {
    load_group(b0 + G, gb);
    run_group(b0, ga);
    load_group(b0 + 2 * G, ga);
    run_group(b0 + G, gb);
  }
  for (int q = lane; q < K; q += WAVE) sizes[q] = sz[q];
  if (lane == 0 && nm) atomicAdd(moves, nm);
}

// part of net n's source at the round start: slot_own for a vertex of the
// round (part[] may already hold its new part), part[] otherwise
__device__ __forceinline__ int frozen_part(int s, const int* part, const int* slot_own, int r0, int r1) {
  return (s >= r0 && s < r1) ? slot_own[s - r0] : part[s];
}

struct ConnApplyArgs {
  const unsigned* net_ptr;
  const int* apply_pins;
  const int* slot_dec;
  const int* slot_own;
  const int* slot_mi;  // MSM: move index of a slot
  const int4* mv;
  const int* touched;
  const int* tcount;
  const int* src;
  const unsigned* vtx_ptr;
  const unsigned* vtx_nets;
  const unsigned* incoff;
  const int* mstat;
  const int* vwgt;
  unsigned* ent;
  unsigned short* ev;
  int* part;
  long long* sizes;
  int* pc;
  int2* evbuf;
  unsigned long long* moves;
  int K, pb;
  unsigned pmask;
  int r0, r1;
  int mode;  // 0: TM apply, 1: MSM dry, 2: MSM commit
  int net_blocks;
};

// The pc changes of one move v: own -> d on one net, from its list after the
// move's edit (TM/partition.cpp:1178-1240, MSM/partition.cpp:1257-1330):
//   v is the source:  +1 (d, t) for listed t != d; -1 (own, t) for listed
//                     t not in {own, d}; -1 (own, d) if d has other pins
//   else (source part sp):  +1 (sp, d) if d is new in the net; -1 (sp, own)
//                     if own left the net
// emit(cell, delta) is called by lanes; returns the number of calls.
template <class Emit>
__device__ void conn_changes(const unsigned* list, int L, int own, int d, bool is_src, int sp, int K, int pb,
                             unsigned pmask, Emit emit) {
  const int lane = lane_id();
  if (is_src) {
    for (int t0 = 0; t0 < L; t0 += WAVE) {
      const int t = t0 + lane;
      if (t < L) {
        const unsigned x = list[t];
        const int q = (int)(x & pmask);
        const unsigned c = x >> pb;
        emit(2 * t, q != d ? d * K + q : -1, 1);
        emit(2 * t + 1, (q != own && q != d) ? own * K + q : ((q == d && c > 1u) ? own * K + d : -1), -1);
      }
    }
  } else {
    const int jd = wave_find(list, L, d, pmask);
    const bool dnew = jd >= 0 && (list[jd] >> pb) == 1u;
    const bool ogone = wave_find(list, L, own, pmask) < 0;
    if (lane == 0) {
      emit(0, dnew ? sp * K + d : -1, 1);
      emit(1, ogone ? sp * K + own : -1, -1);
    }
  }
}

__global__ void __launch_bounds__(REPLAY_BLOCK) k_conn_apply(ConnApplyArgs a) {
  extern __shared__ __align__(16) unsigned char lds[];
  const int lane = lane_id(), wave = wave_id(), nw = (int)(blockDim.x / WAVE);
  const int nstop = a.mode == 2 ? a.mstat[3] : 0;
  if ((int)blockIdx.x >= a.net_blocks) {
    // MSM commit: part vector and both sizes of the moves before the stop
    const int mi = ((int)blockIdx.x - a.net_blocks) * (int)blockDim.x + (int)threadIdx.x;
    const bool mvd = mi < nstop;
    if (mvd) {
      const int4 m = a.mv[mi];
      const long long w = a.vwgt ? a.vwgt[m.x] : 1;
      a.part[m.x] = m.z;
      atomic_add_i64(&a.sizes[m.y], -w);
      atomic_add_i64(&a.sizes[m.z], w);
    }
    block_count_add(a.moves, mvd);
    return;
  }
  unsigned* list = reinterpret_cast<unsigned*>(lds) + (size_t)wave * (size_t)a.K;
  const int ntouched = *a.tcount;
  for (int i = (int)blockIdx.x * nw + wave; i < ntouched; i += a.net_blocks * nw) {
    const int n = a.touched[i];
    const unsigned s = a.net_ptr[n], e = a.net_ptr[n + 1];
    const unsigned seg = s + (unsigned)n;
    unsigned lo, hi;
    // a net of at most 64 pins: its ranks (and its source) are read together with its list
    const bool small = e - s <= (unsigned)WAVE;
    const int r_small = small && (unsigned)lane < e - s ? a.apply_pins[s + (unsigned)lane] : INT_MAX;
    const int srcv = a.src[n];
    wave_rank_window(a.apply_pins, s, e, a.r0, a.r1, lo, hi);
    int L = wave_list_load(a.ent, seg, e - s, list);
    int sp = frozen_part(srcv, a.part, a.slot_own, a.r0, a.r1);
    bool dirty = false;
    for (unsigned k0 = lo; k0 < hi; k0 += WAVE) {
      const unsigned k = k0 + (unsigned)lane;
      int slot = -1, dd = -1;
      const int r = small ? r_small : (k < hi ? a.apply_pins[k] : INT_MAX);
      if (r >= a.r0 && r < a.r1) {
        slot = r - a.r0;
        if (a.mode == 0) {
          dd = a.slot_dec[slot];
        } else {
          const int mi = a.slot_mi[slot];
          dd = (mi >= 0 && (a.mode == 1 || mi < nstop)) ? mi : -1;
        }
      }
      // each lane fetches its own move's parts and, for the dry replay, the
      // event slots of (move, n); the ordered replay below reads them from the lanes
      int own_l = 0, d_l = 0;
      unsigned e0_l = 0, e1_l = 0;
      if (dd >= 0) {
        if (a.mode == 0) {
          own_l = a.slot_own[slot];
          d_l = dd;
        } else {
          const int4 m = a.mv[dd];
          own_l = m.y;
          d_l = m.z;
          if (a.mode == 1) {
            // event slots of incidence (v, n): [incoff[j], incoff[j + 1]), j = position of n in v's list
            const int v = a.r0 + slot;
            const unsigned j = lane_lower_bound(a.vtx_nets, a.vtx_ptr[v], a.vtx_ptr[v + 1], (unsigned)n);
            e0_l = a.incoff[j];
            e1_l = a.incoff[j + 1];
          }
        }
      }
      unsigned long long bm = __ballot(dd >= 0);
      while (bm) {
        const int l = __builtin_amdgcn_readfirstlane(__ffsll((long long)bm) - 1);
        bm &= bm - 1;
        const int v = a.r0 + __builtin_amdgcn_readlane(slot, l);
        const int own = __builtin_amdgcn_readlane(own_l, l), d = __builtin_amdgcn_readlane(d_l, l);
        wave_list_move(list, L, own, d, a.pb, a.pmask);
        dirty = true;
        const bool is_src = srcv == v;
        if (a.mode == 0) {
          conn_changes(list, L, own, d, is_src, sp, a.K, a.pb, a.pmask, [&](int, int cell, int dl) {
            if (cell >= 0) atomicAdd(&a.pc[cell], dl);
          });
        } else if (a.mode == 1) {
          const unsigned e0 = (unsigned)__builtin_amdgcn_readlane((int)e0_l, l),
                         e1 = (unsigned)__builtin_amdgcn_readlane((int)e1_l, l);
          for (unsigned t = e0 + (unsigned)lane; t < e1; t += WAVE) a.evbuf[t] = make_int2(0, 0);
          wave_lds_sync();
          conn_changes(list, L, own, d, is_src, sp, a.K, a.pb, a.pmask, [&](int slot_i, int cell, int dl) {
            if (cell >= 0) a.evbuf[e0 + (unsigned)slot_i] = make_int2(cell, dl);
          });
        }
        if (is_src) sp = d;
      }
    }
    if (a.mode != 1 && dirty) wave_list_store(a.ent, a.ev, seg, list, L, a.pb, a.pmask);
    wave_lds_sync();
  }
}

// MSM: the round's moves in order (MSM/partition.cpp:1230-1370). Each move's
// pc changes go to pc with atomics; a cell crossing 0 <-> 1 changes its row's
// message count; then the max-set step decides whether the round stops.
__global__ void __launch_bounds__(WAVE) k_msm_scan(int* __restrict__ pc, int* __restrict__ msg_g, int K,
                                                   const int4* __restrict__ mv, const int2* __restrict__ evbuf,
                                                   const unsigned* __restrict__ incoff,
                                                   const unsigned* __restrict__ vtx_ptr, int* __restrict__ mstat) {
  extern __shared__ __align__(16) unsigned char lds[];
  int* msg = reinterpret_cast<int*>(lds);
  int* dirty = msg + K;
  const int lane = lane_id();
  for (int q = lane; q < K; q += WAVE) msg[q] = msg_g[q];
  dirty[lane] = 0;
  wave_lds_sync();
  LaneMax ms;
  int M = maxset_init(msg, K, ms);
  const int nmoves = mstat[2];
  int nstop = nmoves;
  for (int mi = 0; mi < nmoves; mi++) {
    const int v = mv[mi].x;
    const unsigned e0 = incoff[vtx_ptr[v]], e1 = incoff[vtx_ptr[v + 1]];
    const int nit = (int)(e1 - e0);
    for (int b = 0; b < nit; b += WAVE) {
      const int it = b + lane;
      if (it < nit) {
        const int2 e = evbuf[e0 + (unsigned)it];
        if (e.y != 0) {
          const int old = atomicAdd(&pc[e.x], e.y);
          const int cross = (e.y > 0 && old == 0) ? 1 : ((e.y < 0 && old == 1) ? -1 : 0);
          if (cross) {
            const int r = e.x / K;
            atomicAdd(&msg[r], cross);
            dirty[r & (WAVE - 1)] = 1;
          }
        }
      }
    }
    wave_lds_sync();
    if (maxset_step(msg, dirty, K, ms, M)) {
      nstop = mi + 1;
      break;
    }
  }
  const int cntM = maxset_count(ms);
  // the messages after the moves up to the stop: what a recount from pc gives
  for (int q = lane; q < K; q += WAVE) msg_g[q] = msg[q];
  if (lane == 0) {
    mstat[0] = M;
    mstat[1] = cntM;
    mstat[3] = nstop;
  }
}

// LDS of k_conn_eval: the block's candidate mask, then per wavefront the
// receiver and sender lists and three K-bit marks, then per wavefront the
// touched-net buffer
// Receiver and sender list capacities: a vertex's receivers are distinct
// parts listed in the nets whose source it is (at most their pins), its
// senders distinct source parts of its other nets (at most its degree).
// Sized by the instance, not by K, the lists leave the LDS to more waves.
int conn_rcap(const fp_ctx* ctx) {
  return (int)std::max<unsigned long long>(1, std::min<unsigned long long>(ctx->max_src_pins, (unsigned long long)ctx->K));
}
int conn_scap(const fp_ctx* ctx) { return std::max(1, std::min((int)ctx->max_vdeg, ctx->K)); }

size_t conn_eval_lds(int K, int waves, int rcap, int scap) {
  const size_t per = ((size_t)rcap + (size_t)scap + 3 * (size_t)((K + 31) / 32) + TOUCH_BUF) * sizeof(int);
  return (size_t)((K + 63) / 64) * sizeof(unsigned long long) + (size_t)waves * per;
}

int conn_eval_waves(int K, int rcap, int scap) {
  for (int w = CONN_BLOCK / WAVE; w >= 1; w--)
    if (conn_eval_lds(K, w, rcap, scap) <= (size_t)LDS_BUDGET) return w;
  return 0;
}

// MSM event slots per incidence (v, n): 2 * min(|n|, K) when v is n's source,
// else 2 (the bounds of conn_changes). Built once per upload on the host.
fp_status ensure_incoff(fp_ctx* ctx) {
  auto& c = ctx->conn;
  if (c.incoff_ready) return FP_OK;
  const int nov = ctx->nov, noe = ctx->noe;
  std::vector<unsigned> vptr((size_t)nov + 1), vnets((size_t)std::max(ctx->pins, 1ll));
  std::vector<int> srcs((size_t)std::max(noe, 1));
  FP_TRY(d2h(ctx, vptr.data(), ctx->vtx_ptr, (size_t)nov + 1));
  if (ctx->pins) FP_TRY(d2h(ctx, vnets.data(), ctx->vtx_nets, (size_t)ctx->pins));
  if (noe) FP_TRY(d2h(ctx, srcs.data(), ctx->src, (size_t)noe));
  FP_HIP_OK(hipStreamSynchronize(ctx->stream));
  std::vector<unsigned> off((size_t)ctx->pins + 1);
  unsigned long long acc = 0;
  for (int v = 0; v < nov; v++)
    for (unsigned j = vptr[(size_t)v]; j < vptr[(size_t)v + 1]; j++) {
      off[j] = (unsigned)acc;
      const int n = (int)vnets[j];
      const unsigned sz = ctx->h_net_ptr[(size_t)n + 1] - ctx->h_net_ptr[(size_t)n];
      acc += srcs[(size_t)n] == v ? 2ull * std::min<unsigned long long>(sz, (unsigned long long)ctx->K) : 2ull;
    }
  if (acc >= (1ull << 32)) return fail(ctx, FP_ERR_UNSUPPORTED, "MSM event slots exceed 2^32");
  off[(size_t)ctx->pins] = (unsigned)acc;
  FP_TRY(dalloc(ctx, &c.incoff, (size_t)ctx->pins + 1));
  FP_TRY(h2d(ctx, c.incoff, off.data(), (size_t)ctx->pins + 1));
  FP_TRY(dalloc(ctx, &c.evbuf, (size_t)std::max(acc, 1ull)));
  c.evcap = (long long)std::max(acc, 1ull);
  c.incoff_ready = true;
  return FP_OK;
}

fp_status ensure_conn_scratch(fp_ctx* ctx, int R) {
  auto& c = ctx->conn;
  const int K = ctx->K, W = (K + 63) / 64;
  if (K > c.kcap) {
    FP_TRY(dalloc(ctx, &c.pc, (size_t)K * (size_t)K));
    FP_TRY(dalloc(ctx, &c.msg, (size_t)K));
    FP_TRY(dalloc(ctx, &c.B, (size_t)K * (size_t)W));
    FP_TRY(dalloc(ctx, &c.BT, (size_t)K * (size_t)W));
    FP_TRY(dalloc(ctx, &c.mstat, 8));
    c.kcap = K;
  }
  if (R > c.rmax) {
    FP_TRY(dalloc(ctx, &c.slot_mi, (size_t)R));
    FP_TRY(dalloc(ctx, &c.mv, (size_t)R + 1));  // + the end sentinel k_round_compact writes
    c.rmax = R;
  }
  FP_TRY(ensure_compact_tiles(ctx, R));
  if (R > ctx->rmax) {
    FP_TRY(dalloc(ctx, &ctx->slot_own, (size_t)R));
    FP_TRY(dalloc(ctx, &ctx->slot_dec, (size_t)R));
    ctx->rmax = R;
  }
  return FP_OK;
}

// B and BT from pc, and msg recounted unless the caller has it (an MSM round's
// scan leaves the exact counts)
void launch_bits(fp_ctx* ctx, bool recount_msg = true) {
  auto& c = ctx->conn;
  const int K = ctx->K, W = (K + 63) / 64;
  if (recount_msg) fill_async<int>(ctx, c.msg, (size_t)K, 0);
  k_conn_bits<<<dim3((unsigned)W, (unsigned)W), 256, 0, ctx->stream>>>(c.pc, K, W, c.B, c.BT,
                                                                       recount_msg ? c.msg : nullptr);
}

}  // namespace

fp_status plan_conn(fp_ctx* ctx, RefinePlan& p) {
  const bool msm = p.metric == FP_METRIC_MSM;
  const int nov = ctx->nov, K = ctx->K;
  if (K > CONN_KMAX) return fail(ctx, FP_ERR_UNSUPPORTED, "message refinement supports K <= %d", CONN_KMAX);
  if (!ctx->src_on_pins)
    return fail(ctx, FP_ERR_UNSUPPORTED, "message refinement needs net sources that are pins of their nets");
  if (ctx->order)
    return fail(ctx, FP_ERR_UNSUPPORTED,
                "TM/MSM refinement visits vertex ids in order (TM/partition.cpp:976); upload with the identity ordering");
  if (p.sched.kind != FP_SCHED_REF)
    return fail(ctx, FP_ERR_UNSUPPORTED, "message refinement runs the reference round schedule (FP_SCHED_REF) only");
  const int cw = conn_eval_waves(K, conn_rcap(ctx), conn_scap(ctx));
  const int waves = p.eval_waves > 0 ? std::min(p.eval_waves, cw) : cw;
  if (waves == 0) return fail(ctx, FP_ERR_UNSUPPORTED, "K = %d needs more LDS than a workgroup has", K);
  p.eval_waves = waves;
  p.R = std::min(p.sched.round, nov);
  FP_TRY(ensure_conn_scratch(ctx, p.R));
  if (msm) FP_TRY(ensure_incoff(ctx));
  // the connections of the resident state, kept incrementally by the passes
  auto& c = ctx->conn;
  fill_async<int>(ctx, c.pc, (size_t)K * (size_t)K, 0);
  const int ngrid = std::max(1, std::min((ctx->noe + 255) / 256, 4096));
  k_conn_init<<<std::min(ngrid * 8, 4096), 256, 0, ctx->stream>>>(ctx->net_ptr, ctx->ent, ctx->src, ctx->part, ctx->noe, ctx->pmask, K,
                                               c.pc);
  launch_bits(ctx);
  FP_HIP_OK(hipGetLastError());
  return FP_OK;
}

fp_status pass_conn(fp_ctx* ctx, const RefinePlan& p) {
  const bool msm = p.metric == FP_METRIC_MSM;
  const int nov = ctx->nov, K = ctx->K, W = (K + 63) / 64, R = p.R, waves = p.eval_waves;
  const int rcap = conn_rcap(ctx), scap = conn_scap(ctx);
  const size_t lds_e = conn_eval_lds(K, waves, rcap, scap);
  const size_t lds_r = (size_t)(REPLAY_BLOCK / WAVE) * (size_t)K * sizeof(unsigned);
  auto& c = ctx->conn;
  const double thr = p.sched.ldg_C * (1.02);
  pass_begin(ctx, msm ? c.msg : nullptr, K, c.mstat);
  unsigned tag = 0;
  for (int r0 = 0; r0 < nov; r0 += R) {
    const int nslots = std::min(R, nov - r0);
    tag++;
    const int par = (int)(tag & 1u);
    ConnEvalArgs a{};
    a.vtx_ptr = ctx->vtx_ptr;
    a.vtx_nets = ctx->vtx_nets;
    a.vtx_seg = ctx->vtx_seg;
    a.ev = ctx->ev;
    a.src = ctx->src;
    a.part = ctx->part;
    a.sizes = ctx->sizes;
    a.pc = c.pc;
    a.B = c.B;
    a.BT = c.BT;
    a.msg = c.msg;
    a.mstat = c.mstat;
    a.K = K;
    a.W = W;
    a.msm = msm ? 1 : 0;
    a.rcap = rcap;
    a.scap = scap;
    a.r0 = r0;
    a.nslots = nslots;
    a.C = p.sched.ldg_C;
    a.slot_own = ctx->slot_own;
    a.slot_dec = ctx->slot_dec;
    a.tag = tag;
    a.net_stamp = ctx->net_stamp;
    a.touched = ctx->touched;
    a.tcount = ctx->counters + par;
    a.tcount_next = ctx->counters + (par ^ 1);
    int egrid = std::max(1, std::min((nslots + waves - 1) / waves, 8192));
    if (p.eval_grid > 0) egrid = std::min(egrid, p.eval_grid);
    FP_TRY(launch_timed(ctx, k_conn_eval, egrid, waves * WAVE, lds_e, a));
    ConnApplyArgs ap{};
    ap.net_ptr = ctx->net_ptr;
    ap.apply_pins = ctx->apply_pins;
    ap.slot_dec = ctx->slot_dec;
    ap.slot_own = ctx->slot_own;
    ap.slot_mi = c.slot_mi;
    ap.mv = c.mv;
    ap.touched = ctx->touched;
    ap.tcount = ctx->counters + par;
    ap.src = ctx->src;
    ap.vtx_ptr = ctx->vtx_ptr;
    ap.vtx_nets = ctx->vtx_nets;
    ap.incoff = c.incoff;
    ap.mstat = c.mstat;
    ap.vwgt = ctx->vwgt;
    ap.ent = ctx->ent;
    ap.ev = ctx->ev;
    ap.part = ctx->part;
    ap.sizes = ctx->sizes;
    ap.pc = c.pc;
    ap.evbuf = c.evbuf;
    ap.moves = ctx->moves;
    ap.K = K;
    ap.pb = ctx->pb;
    ap.pmask = ctx->pmask;
    ap.r0 = r0;
    ap.r1 = r0 + nslots;
    ap.net_blocks = net_grid(ctx, REPLAY_NET_BLOCKS, REPLAY_BLOCK / WAVE);
    if (!msm) {
      launch_round_compact<true>(ctx, nullptr, r0, nslots, c.slot_mi, c.mv, c.mstat);
      k_tm_admit<<<1, WAVE, (size_t)K * (sizeof(long long) + 2 * sizeof(int)), ctx->stream>>>(
          ctx->slot_dec, c.mv, c.mstat, r0, K, thr, ctx->sizes, ctx->part, ctx->moves);
      ap.mode = 0;
      k_conn_apply<<<ap.net_blocks, REPLAY_BLOCK, lds_r, ctx->stream>>>(ap);
    } else {
      launch_round_compact<false>(ctx, nullptr, r0, nslots, c.slot_mi, c.mv, c.mstat);
      ap.mode = 1;
      k_conn_apply<<<ap.net_blocks, REPLAY_BLOCK, lds_r, ctx->stream>>>(ap);
      k_msm_scan<<<1, WAVE, ((size_t)K + WAVE) * sizeof(int), ctx->stream>>>(c.pc, c.msg, K, c.mv, c.evbuf,
                                                                                 c.incoff, ctx->vtx_ptr, c.mstat);
      ap.mode = 2;
      k_conn_apply<<<ap.net_blocks + (nslots + REPLAY_BLOCK - 1) / REPLAY_BLOCK, REPLAY_BLOCK, lds_r,
                     ctx->stream>>>(ap);
    }
    launch_bits(ctx, !msm);
  }
  return FP_OK;
}

}  // namespace fpe

using namespace fpe;

extern "C" fp_status fp_get_connections(fp_ctx* ctx, int* pc, int* msg) {
  if (!ctx) return FP_ERR_ARG;
  if (!ctx->has_state) return fail(ctx, FP_ERR_STATE, "fp_get_connections before a state was set");
  if (ctx->K > CONN_KMAX) return fail(ctx, FP_ERR_UNSUPPORTED, "part connections support K <= %d", CONN_KMAX);
  if (!ctx->src_on_pins)
    return fail(ctx, FP_ERR_UNSUPPORTED, "part connections need net sources that are pins of their nets");
  FP_HIP_OK(hipSetDevice(ctx->device));
  const int K = ctx->K;
  FP_TRY(ensure_conn_scratch(ctx, 1));
  auto& c = ctx->conn;
  fill_async<int>(ctx, c.pc, (size_t)K * (size_t)K, 0);
  const int ngrid = std::max(1, std::min((ctx->noe + 255) / 256, 4096));
  k_conn_init<<<std::min(ngrid * 8, 4096), 256, 0, ctx->stream>>>(ctx->net_ptr, ctx->ent, ctx->src, ctx->part, ctx->noe, ctx->pmask, K,
                                               c.pc);
  launch_bits(ctx);
  FP_HIP_OK(hipGetLastError());
  if (pc) FP_TRY(d2h(ctx, pc, c.pc, (size_t)K * (size_t)K));
  if (msg) FP_TRY(d2h(ctx, msg, c.msg, (size_t)K));
  FP_HIP_OK(hipStreamSynchronize(ctx->stream));
  return FP_OK;
}
