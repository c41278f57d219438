// fp_vol.hip -- MSV and MSRV refinement on the device: the round schedule of
// partition_par in MGV/partition.cpp:706-1040 (max send volume) and
// MSRV/partition.cpp:769-1110 (max send + receive volume).
//
// vol[p] is part p's send volume (MSV) or send + receive volume (MSRV); the
// objective is max_p vol[p], and the "max set" is the set of parts holding it.
// A round covers R consecutive visiting positions (MSV: vertex ids, the
// reference's refinement ignores the ordering, MGV/partition.cpp:755-757):
//   k_vol_eval     one wavefront per vertex: LDG decision on the frozen state
//                  (the band-free LDG of MGV/partition.cpp:28-65), then the
//                  reference's per-part gain/loss terms, accepted iff every
//                  max part gains at least (MSV) / more than (MSRV) it loses
//                  and no other part reaches the max
//   k_round_compact accepted moves in visiting order, with event offsets
//   k_vol_replay   (dry) one wavefront per touched net replays the round's
//                  moves on its list in order and records, per (move, net),
//                  the volume changes the reference's apply makes
//   k_vol_scan     one wavefront applies the moves' volume changes in order,
//                  tracks the max set, and stops after the first move whose
//                  max set holds a part the previous one did not
//                  (MGV/partition.cpp:977-1003); batches of 64 moves are
//                  decided together where bounds allow
//   k_vol_replay   (commit) list edits, part vector and both part sizes of the
//                  moves up to the stop
#include "fp_round.h"

namespace fpe {
namespace {

constexpr int VOL_BLOCK = 512;  // at most 8 wavefronts; vol_eval_waves picks the count
constexpr int VOL_KMAX = 4096;  // k_vol_scan: one lane of the max-set tracking holds at most 64 parts
constexpr int REPLAY_BLOCK = 256;
constexpr int REPLAY_NET_BLOCKS = 2048;

// per part: volume of every net whose source is in the part, lambda - 1 (plus,
// for MSRV, one per other part of the net: what it receives); one wavefront
// per net, block-private sums in LDS
__global__ void __launch_bounds__(256) k_vol_init(const unsigned* __restrict__ net_ptr,
                                                  const unsigned* __restrict__ ent, const int* __restrict__ src,
                                                  const int* __restrict__ part, int noe, unsigned pmask, int K,
                                                  int msrv, int* __restrict__ vol) {
  extern __shared__ int s_vol[];
  const int lane = lane_id();
  for (int q = (int)threadIdx.x; q < K; q += (int)blockDim.x) s_vol[q] = 0;
  __syncthreads();
  const int nw = (int)(gridDim.x * blockDim.x / WAVE);
  for (int n = (int)((blockIdx.x * blockDim.x + threadIdx.x) / WAVE); n < noe; n += nw) {
    const unsigned seg = net_ptr[n] + (unsigned)n;
    const int L = (int)ent[seg];
    if (L == 0) continue;
    const int sp = part[src[n]];
    if (L > 1 && lane == 0) atomicAdd(&s_vol[sp], L - 1);
    if (msrv)
      for (int t = lane; t < L; t += WAVE) {
        const int q = (int)(ent[seg + 1 + (unsigned)t] & pmask);
        if (q != sp) atomicAdd(&s_vol[q], 1);
      }
  }
  __syncthreads();
  for (int q = (int)threadIdx.x; q < K; q += (int)blockDim.x)
    if (s_vol[q]) atomicAdd(&vol[q], s_vol[q]);
}

struct VolEvalArgs {
  const unsigned* vtx_ptr;
  const unsigned* vtx_nets;
  const unsigned* vtx_seg;
  const unsigned short* ev;
  const int* src;
  const int* part;
  const int* order;  // nullable: identity
  const long long* sizes;
  const int* vol;
  const int* mstat;
  int K, pb, msrv;
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

// The first-appearance rank of the LDG tie-break (MGV/partition.cpp:765-790)
// for the parts tied at the best (score, size): the count array is rebuilt
// by a second sweep, the tied parts are marked in it, and a third sweep
// keeps each marked part's lowest rank + 1 (ranks (net index, list position)
// are distinct, so the minimum picks one part). Leaves cnt all zero.
__device__ __forceinline__ int vol_tie_decision(const VolEvalArgs& a, unsigned vb, int deg, int own, unsigned* cnt,
                                                const double* s_pen, const unsigned* s_skey,
                                                unsigned long long m1, unsigned m2) {
  const int lane = lane_id(), K = a.K;
  for (int c0 = 0; c0 < deg; c0 += WAVE) {
    const int nn = min(WAVE, deg - c0);
    const unsigned sg = lane < nn ? a.vtx_seg[vb + (unsigned)(c0 + lane)] : 0u;
    wave_sweep_lists<1, 1>(a.ev, sg, nn, [&](int, int, int, bool ok, unsigned x) {
      if (ok) atomicAdd(&cnt[(int)(x & ~EV_SINGLE)], 1u);
    });
  }
  wave_lds_sync();
  for (int q = lane; q < K; q += WAVE) {
    const unsigned c = cnt[q];
    const int adj = (int)c - (q == own ? 1 : 0);
    const double sc = (double)adj * s_pen[q];
    const bool ok = c != 0 && sc > 0;
    const unsigned long long k1 = ok ? (unsigned long long)__double_as_longlong(sc) : 0ull;
    cnt[q] = ok && k1 == m1 && s_skey[q] == m2 ? NOFIRST : 0u;
  }
  wave_lds_sync();
  for (int c0 = 0; c0 < deg; c0 += WAVE) {
    const int nn = min(WAVE, deg - c0);
    const unsigned sg = lane < nn ? a.vtx_seg[vb + (unsigned)(c0 + lane)] : 0u;
    wave_sweep_lists<1, 1>(a.ev, sg, nn, [&](int j, int, int t, bool ok, unsigned x) {
      const int q = (int)(x & ~EV_SINGLE);
      if (ok && cnt[q] != 0u) atomicMin(&cnt[q], (((unsigned)(c0 + j) << a.pb) | (unsigned)t) + 1u);
    });
  }
  wave_lds_sync();
  unsigned br = NOFIRST;
  int bq = -1;
  for (int q = lane; q < K; q += WAVE) {
    const unsigned r = cnt[q];
    cnt[q] = 0u;
    if (r != 0u && r < br) {
      br = r;
      bq = q;
    }
  }
  const unsigned mr = wave_min_u32(br);
  const int wl = __ffsll((long long)__ballot(br == mr && bq >= 0)) - 1;
  wave_lds_sync();
  return __builtin_amdgcn_readlane(bq, wl);
}

__global__ void __launch_bounds__(VOL_BLOCK) k_vol_eval(VolEvalArgs a) {
  extern __shared__ __align__(16) unsigned char lds[];
  const int K = a.K;
  double* s_pen = reinterpret_cast<double*>(lds);
  unsigned* s_skey = reinterpret_cast<unsigned*>(s_pen + K);  // size_key of each frozen part size
  unsigned* wbase = s_skey + K;
  __shared__ long long red_sz[VOL_BLOCK / WAVE];
  __shared__ int red_ix[VOL_BLOCK / WAVE];
  const int lane = lane_id(), wave = wave_id(), nw = (int)(blockDim.x / WAVE);
  const int KW = (K + 31) / 32;
  unsigned* cnt = wbase + (size_t)wave * (size_t)(K + KW);
  unsigned* hitb = cnt + K;                  // max parts the move affects (MSRV), one bit per part
  int* delta = reinterpret_cast<int*>(cnt);  // reused once the decision is made (cnt is all 0 again)

  // penalty = 1 - (1.0 * size / C) (MGV/partition.cpp:36); fallback = lowest-index lightest part (:48-60)
  long long bs = LLONG_MAX;
  int bi = INT_MAX;
  for (int q = (int)threadIdx.x; q < K; q += (int)blockDim.x) {
    const long long sz = a.sizes[q];
    s_skey[q] = size_key(sz);
    s_pen[q] = 1.0 - ((double)sz / a.C);
    if (sz < bs) {
      bs = sz;
      bi = q;
    }
  }
  for (int q = lane; q < K; q += WAVE) cnt[q] = 0;
  for (int q = lane; q < KW; q += WAVE) hitb[q] = 0;
  for (int off = WAVE / 2; off > 0; off >>= 1) {
    const long long os = __shfl_xor(bs, off);
    const int oi = __shfl_xor(bi, off);
    if (os < bs || (os == bs && oi < bi)) {
      bs = os;
      bi = oi;
    }
  }
  if (lane == 0) {
    red_sz[wave] = bs;
    red_ix[wave] = bi;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0 && a.tcount_next) *a.tcount_next = 0;
  __syncthreads();
  int fb = red_ix[0];
  {
    long long b = red_sz[0];
    for (int w = 1; w < nw; w++)
      if (red_sz[w] < b || (red_sz[w] == b && red_ix[w] < fb)) {
        b = red_sz[w];
        fb = red_ix[w];
      }
  }
  const int M = a.mstat[0], cntM = a.mstat[1];

  for (int slot = (int)blockIdx.x * nw + wave; slot < a.nslots; slot += (int)gridDim.x * nw) {
    const int z = a.r0 + slot;
    const int v = a.order ? a.order[z] : z;
    const int own = a.part[v];
    const unsigned vb = a.vtx_ptr[v];
    const int deg = (int)(a.vtx_ptr[v + 1] - vb);
    // this lane's net of the first chunk: segment offset and source part
    // (own for v's own source nets), kept for the later passes
    const unsigned sg0 = lane < deg ? a.vtx_seg[vb + (unsigned)lane] : 0u;
    int ip0 = -1;
    bool isrc0 = false;
    if (lane < deg) {
      const int s = a.src[a.vtx_nets[vb + (unsigned)lane]];
      isrc0 = s == v;
      ip0 = isrc0 ? own : a.part[s];
    }

    // scores: pin counts per part over v's nets (MGV/partition.cpp:765-790)
    for (int c0 = 0; c0 < deg; c0 += WAVE) {
      const int nn = min(WAVE, deg - c0);
      const unsigned sg = c0 == 0 ? sg0 : (lane < nn ? a.vtx_seg[vb + (unsigned)(c0 + lane)] : 0u);
      wave_sweep_lists<8, SWEEP_LONG_TASKS>(a.ev, sg, nn, [&](int, int, int, bool ok, unsigned x) {
        if (ok) atomicAdd(&cnt[(int)(x & ~EV_SINGLE)], 1u);
      });
    }
    wave_lds_sync();
    // LDG: lexicographic max of (score, -size, -first rank) over score > 0.
    // Scores are positive doubles, so their bit patterns order like the
    // values: key1 = score bits (max), key2 = size_key(size) (min among key1
    // ties); ltie: this lane met its best keys twice. The first-appearance
    // rank only matters between parts tied on both keys (vol_tie_decision).
    // Four 64-part chunks per step, without branches: their LDS reads issue
    // together; every slot is cleared for the next vertex
    unsigned long long b1 = 0;
    unsigned b2 = ~0u;
    int bq = -1;
    bool ltie = false;
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
        const int adj = (int)cv[u] - (q == own ? 1 : 0);  // scores[part_id]-- (MGV/partition.cpp:792-796)
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
      }
    }
    const unsigned long long m1 = wave_max_u64(b1);
    const unsigned m2 = wave_min_u32(b1 == m1 ? b2 : ~0u);
    wave_lds_sync();
    int d = fb;
    if (m1 != 0) {
      const unsigned long long cand = __ballot(b1 == m1 && b2 == m2);
      if (__popcll(cand) > 1 || __ballot(b1 == m1 && b2 == m2 && ltie) != 0ull) {
        d = vol_tie_decision(a, vb, deg, own, cnt, s_pen, s_skey, m1, m2);
      } else {
        d = __builtin_amdgcn_readlane(bq, __ffsll((long long)cand) - 1);
      }
    }

    int dec = -1;
    if (d != own) {
      // gains / losses (MGV/partition.cpp:800-840; MSRV/partition.cpp:852-897),
      // accumulated as delta[p] = losses[p] - gains[p]; lane j takes net j of
      // each chunk: whether v is own's only pin there (single), whether d is
      // listed (found), and the net's source part
      int d_own = 0, d_dec = 0;
      for (int c0 = 0; c0 < deg; c0 += WAVE) {
        const int nn = min(WAVE, deg - c0);
        unsigned sg = sg0;
        int ip = ip0;
        bool isrc = isrc0;
        if (c0 > 0) {
          sg = lane < nn ? a.vtx_seg[vb + (unsigned)(c0 + lane)] : 0u;
          ip = -1;
          if (lane < nn) {
            const int s = a.src[a.vtx_nets[vb + (unsigned)(c0 + lane)]];
            isrc = s == v;
            ip = isrc ? own : a.part[s];
          }
        }
        unsigned long long smask = 0, fmask = 0;
        int myL = 0;
        wave_sweep_lists<8, SWEEP_LONG_TASKS>(a.ev, sg, nn, [&](int j, int L, int t, bool ok, unsigned x) {
          const int q = (int)(x & ~EV_SINGLE);
          if (__ballot(ok && q == own && (x & EV_SINGLE))) smask |= 1ull << j;
          if (__ballot(ok && q == d)) fmask |= 1ull << j;
          if (lane == j) myL = L;
          (void)t;
        });
        if (lane < nn) {
          const bool single = (smask >> lane) & 1ull, found = (fmask >> lane) & 1ull;
          if (isrc) {
            d_own -= myL - 1;  // gains[part_id] += end_cnt - 1
            d_dec += myL - 1;  // losses[decision] += end_cnt - 1
            if (single) d_dec -= 1;
            if (!found) d_dec += 1;
          } else {
            const int dl = (found ? 0 : 1) - (single ? 1 : 0);
            if (dl) atomicAdd(&delta[ip], dl);
            if (a.msrv) {
              if (single) d_own -= 1;
              if (!found) d_dec += 1;
            }
          }
        }
      }
      d_own = wave_sum(d_own);
      d_dec = wave_sum(d_dec);
      if (lane == 0) {
        atomicAdd(&delta[own], d_own);
        atomicAdd(&delta[d], d_dec);
      }
      wave_lds_sync();
      // the affected parts: item 0 = own, 1 = d, 2 + i = source part of net i
      // (own for v's own source nets); every other part has gains = losses = 0.
      // g(item, part) runs for every item, net items on lane i of their chunk
      auto for_items = [&](auto g) {
        if (lane < 2) g(lane, lane == 0 ? own : d);
        for (int c0 = 0; c0 < deg; c0 += WAVE) {
          const int i = c0 + lane;
          if (i < deg) {
            int p = ip0;
            if (c0 > 0) {
              const int s = a.src[a.vtx_nets[vb + (unsigned)i]];
              p = s == v ? own : a.part[s];
            }
            g(2 + i, p);
          }
        }
      };
      // MSRV: max parts untouched by the move have gains = losses = 0 and fail
      // the strict test, so every max part must be affected: maxhit counts
      // the distinct max parts among the items
      bool bad = false;
      int maxhit = 0;
      for_items([&](int, int p) {
        const int dl = delta[p];
        const int vv = a.vol[p];
        if (vv == M) {
          if (a.msrv ? dl >= 0 : dl > 0) bad = true;  // a max part must gain (MSRV: strictly)
          const unsigned bit = 1u << (p & 31);
          if (!(atomicOr(&hitb[p >> 5], bit) & bit)) maxhit++;
        } else if (vv + dl >= M) {
          bad = true;  // no other part may reach the max
        }
      });
      bad = __ballot(bad) != 0;
      maxhit = wave_sum(maxhit);
      wave_lds_sync();
      for_items([&](int, int p) {
        delta[p] = 0;
        hitb[p >> 5] = 0u;
      });
      wave_lds_sync();
      if (!bad && (!a.msrv || maxhit == cntM)) dec = d;
    }
    if (lane == 0) {
      a.slot_own[slot] = own;
      a.slot_dec[slot] = dec;
    }
    if (dec >= 0) wave_touch_nets(a.vtx_nets + vb, deg, a.tag, a.net_stamp, a.touched, a.tcount);
  }
}

// First word of a volume event: the net's source part (or -1) and its change
// (-1, 0 or +1) in one int.
__device__ __forceinline__ int vol_event_word(int sp, int dsp) { return ((sp + 1) << 2) | (dsp + 1); }
__device__ __forceinline__ int vol_event_part(int x) { return (x >> 2) - 1; }
__device__ __forceinline__ int vol_event_change(int x) { return (x & 3) - 1; }

struct VolReplayArgs {
  const unsigned* net_ptr;
  const int* apply_pins;
  const int* slot_mi;
  const int4* mv;
  const int* touched;
  const int* tcount;
  const int* src;
  const unsigned* vtx_ptr;
  const unsigned* vtx_nets;
  const int* mstat;
  const int* vwgt;
  unsigned* ent;
  unsigned short* ev;
  int* part;
  long long* sizes;
  unsigned long long* moves;
  int4* evbuf;
  int K, pb;
  unsigned pmask;
  int r0, r1, msrv, commit, net_blocks;
};

// dry (commit = 0): per (move, net) volume changes of the reference's apply
//   (MGV/partition.cpp:897-975; MSRV/partition.cpp:952-1052), recorded as
//   (source part and its change, change of own part, change of new part,
//   move index) at event slot mv.evoff + (position of the net in the move's
//   vertex->nets list); the list edits stay in LDS
// commit (commit = 1): list edits of the moves before the stop; blocks past
//   net_blocks commit part vector and both part sizes
// INDEXED: each wave keeps a part -> position index of its list in LDS
//   (2 bytes per part after the lists), for K where it fits
template <bool INDEXED>
__global__ void __launch_bounds__(REPLAY_BLOCK) k_vol_replay(VolReplayArgs a) {
  extern __shared__ __align__(16) unsigned char lds[];
  const int lane = lane_id(), wave = wave_id(), nw = (int)(blockDim.x / WAVE);
  const int nstop = a.mstat[3];
  if ((int)blockIdx.x >= a.net_blocks) {
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
  unsigned short* at = reinterpret_cast<unsigned short*>(reinterpret_cast<unsigned*>(lds) + (size_t)nw * (size_t)a.K) +
                       (size_t)wave * (size_t)a.K;
  if (INDEXED) wave_index_init(at, a.K);
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
    if (INDEXED) wave_index_set(list, L, at, a.pmask);
    int sp_cur = a.commit ? 0 : a.part[srcv];  // the source's part as the replay goes
    bool dirty = false;
    for (unsigned k0 = lo; k0 < hi; k0 += WAVE) {
      const unsigned k = k0 + (unsigned)lane;
      int mi = -1;
      const int r = small ? r_small : (k < hi ? a.apply_pins[k] : INT_MAX);
      if (r >= a.r0 && r < a.r1) mi = a.slot_mi[r - a.r0];
      if (a.commit && mi >= nstop) mi = -1;
      // each lane fetches its own move (v, own, d, evoff) and, for the dry
      // replay, the event slot of (move, n); the ordered replay below reads
      // them from the lanes
      int4 mm = make_int4(0, 0, 0, 0);
      unsigned evslot = 0;
      if (mi >= 0) {
        mm = a.mv[mi];
        if (!a.commit) {
          const unsigned vb = a.vtx_ptr[mm.x], ve = a.vtx_ptr[mm.x + 1];
          evslot = (unsigned)mm.w + (lane_lower_bound(a.vtx_nets, vb, ve, (unsigned)n) - vb);
        }
      }
      unsigned long long bm = __ballot(mi >= 0);
      while (bm) {
        const int l = __builtin_amdgcn_readfirstlane(__ffsll((long long)bm) - 1);
        bm &= bm - 1;
        const int v = __builtin_amdgcn_readlane(mm.x, l), own = __builtin_amdgcn_readlane(mm.y, l),
                  d = __builtin_amdgcn_readlane(mm.z, l);
        if (!a.commit) {
          const bool isrc = srcv == v;
          const int lam_b = L;
          unsigned co;
          bool found;
          if (INDEXED) {
            wave_ilist_move(list, at, L, own, d, a.pb, a.pmask, co, found);
          } else {
            const int jo = wave_find(list, L, own, a.pmask);
            co = jo >= 0 ? (list[jo] >> a.pb) : 0u;
            found = wave_find(list, L, d, a.pmask) >= 0;
            wave_list_move(list, L, own, d, a.pb, a.pmask);
          }
          const int lam_mid = co == 1u ? lam_b - 1 : lam_b;
          int e_own = 0, e_dec = 0, e_sp = 0, spx = -1;
          if (isrc) e_own -= lam_b - 1;                       // send_volumes[part_id] -= end_cnt - 1
          if (co == 1u) {
            if (!isrc) {
              spx = sp_cur;
              e_sp -= 1;                                      // send_volumes[net_src_part]--
              if (a.msrv) e_own -= 1;                         // received_volumes[part_id]--
            }
          } else if (isrc && a.msrv) {
            e_own += 1;                                       // received_volumes[part_id]++
          }
          if (isrc) e_dec += lam_mid - 1;                     // send_volumes[net_src_part] += end_cnt - 1
          if (found) {
            if (isrc && a.msrv) e_dec -= 1;                   // received_volumes[decision]--
          } else if (isrc) {
            e_dec += 1;                                       // send_volumes[net_src_part]++ (source part = d)
          } else {
            spx = sp_cur;
            e_sp += 1;                                        // send_volumes[net_src_part]++
            if (a.msrv) e_dec += 1;                           // received_volumes[decision]++
          }
          if (lane == 0)
            a.evbuf[(unsigned)__builtin_amdgcn_readlane((int)evslot, l)] =
                make_int4(vol_event_word(spx, e_sp), e_own, e_dec, __builtin_amdgcn_readlane(mi, l));
          if (isrc) sp_cur = d;
        } else if (INDEXED) {
          unsigned co;
          bool found;
          wave_ilist_move(list, at, L, own, d, a.pb, a.pmask, co, found);
        } else {
          wave_list_move(list, L, own, d, a.pb, a.pmask);
        }
        dirty = true;
      }
    }
    if (a.commit && dirty) wave_list_store(a.ent, a.ev, seg, list, L, a.pb, a.pmask);
    if (INDEXED) wave_index_clear(list, L, at, a.pmask);
    wave_lds_sync();
  }
}

// LDS of k_vol_scan's batched step, after vol[K] and dirty[64]: per-part
// bounds of a batch's changes, candidate map, and the candidates x moves
// table of values.
struct VolBatchLds {
  int* pos;    // [K] sum of the batch's positive changes of a part
  int* neg;    // [K] sum of its negative changes
  int* cidx;   // [K] candidate index of a part or -1
  int* cpart;  // [WAVE] part of a candidate
  int* cflag;  // [WAVE] the candidate is in the max set at the batch start
  int2* mvod;  // [WAVE] (own part, new part) of the batch's moves
  int* D;      // [WAVE][WAVE + 1] per candidate, change and then value after each move
};
constexpr int VOL_STAGE_LOADS = 8;      // event loads per lane in flight while staging
constexpr int VOL_BATCH_KMAX = 1024;  // the batched step's lane holds at most 16 parts
constexpr int DROW = WAVE + 1;        // padded row of D: a lane per row is conflict-free

__host__ __device__ constexpr size_t vol_batch_ints(int K) { return (size_t)3 * (size_t)K + 4 * WAVE + WAVE * DROW; }

// Applies f(part, change) to the three part changes of a staged event of
// move j: own part, new part and the net's source part.
template <class F>
__device__ __forceinline__ void vol_event_changes(const int4 x, const int2 od, F f) {
  if (x.y) f(od.x, x.y);
  if (x.z) f(od.y, x.z);
  const int sp = vol_event_part(x.x), dsp = vol_event_change(x.x);
  if (sp >= 0 && dsp) f(sp, dsp);
}

// The moves mi0 .. mi0 + nb - 1 of the round, their tot events staged (in
// order) in stage, the max set tracked by ms. Exact without going move by
// move: with A_low = max over the max set S of (value + negative changes),
// a part outside S whose value + positive changes stays below A_low never
// reaches the max within the batch (the max is at least A_low until the
// round stops). The others (the candidates, S included) are followed move by
// move in a candidates x moves table of changes, summed along the moves by
// wave scans: before the round stops, the max after
// move t is the largest value of S's candidates, the max set is the S
// candidates holding it, and move t stops the round iff a candidate outside
// S reaches it or an S candidate not in the max set after move t - 1 is in
// it after t. Applies the moves up to the stop (all if none) and returns the
// batch index of the stop move, -1 if none, or -2 (nothing applied) when
// the candidates do not fit a wave.
// Runs f(event, (own, new) of its move, its move index - mi0) over the staged
// events [0, n), VOL_EV_UNROLL per lane at a time: their reads are issued
// before f touches LDS (the compiler keeps LDS reads and atomics in program
// order, so one event at a time would wait out each read).
constexpr int VOL_EV_UNROLL = 4;

template <class F>
__device__ __forceinline__ void for_staged_events(const int4* stage, int n, const int2* mvod, int mi0, F f) {
  const int lane = lane_id();
  for (int e0 = 0; e0 < n; e0 += VOL_EV_UNROLL * WAVE) {
    int4 x[VOL_EV_UNROLL];
    int2 od[VOL_EV_UNROLL];
#pragma unroll
    for (int u = 0; u < VOL_EV_UNROLL; u++) x[u] = stage[min(e0 + u * WAVE + lane, n - 1)];
#pragma unroll
    for (int u = 0; u < VOL_EV_UNROLL; u++) od[u] = mvod[x[u].w - mi0];
#pragma unroll
    for (int u = 0; u < VOL_EV_UNROLL; u++)
      if (e0 + u * WAVE + lane < n) f(x[u], od[u], x[u].w - mi0);
  }
}

__device__ int vol_scan_batch(int* vol, int K, const int4* stage, int tot, int nb, int mi0, int own_l, int d_l,
                              int cum_l, const VolBatchLds& B, LaneMax& ms, int& M) {
  const int lane = lane_id();
  if (lane < nb) B.mvod[lane] = make_int2(own_l, d_l);
  wave_lds_sync();
  for_staged_events(stage, tot, B.mvod, mi0, [&](int4 x, int2 od, int) {
    vol_event_changes(x, od, [&](int p, int w) {
      if (w > 0) atomicAdd(&B.pos[p], w);
      else atomicAdd(&B.neg[p], w);
    });
  });
  wave_lds_sync();
  int al = INT_MIN;
  for (unsigned long long b = ms.inS; b; b &= b - 1) {
    const int p = lane + WAVE * __builtin_ctzll(b);
    al = max(al, vol[p] + B.neg[p]);
  }
  const int alow = wave_max(al);
  unsigned long long cb = 0;
  {
    int j = 0;
    for (int p = lane; p < K; p += WAVE, j++) {
      const bool inS = (ms.inS >> j) & 1ull;
      if (inS || vol[p] + B.pos[p] >= alow) cb |= 1ull << j;
      B.pos[p] = 0;
      B.neg[p] = 0;
    }
  }
  const int nl = __popcll(cb);
  const int incl = wave_incl_scan(nl);
  const int nc = __builtin_amdgcn_readlane(incl, WAVE - 1);
  if (nc > WAVE) return -2;
  {
    int c = incl - nl;
    for (unsigned long long b = cb; b; b &= b - 1, c++) {
      const int j = __builtin_ctzll(b);
      const int p = lane + WAVE * j;
      B.cidx[p] = c;
      B.cpart[c] = p;
      B.cflag[c] = (int)((ms.inS >> j) & 1ull);
    }
  }
  if (lane < nb)
    for (int r = 0; r < nc; r++) B.D[r * DROW + lane] = 0;
  wave_lds_sync();
  for_staged_events(stage, tot, B.mvod, mi0, [&](int4 x, int2 od, int j) {
    vol_event_changes(x, od, [&](int p, int w) {
      const int ci = B.cidx[p];
      if (ci >= 0) atomicAdd(&B.D[ci * DROW + j], w);
    });
  });
  wave_lds_sync();
  // lane t: the max and max set after move t, and whether t stops the
  // round; candidate r's value after move t is its value before the batch
  // plus the inclusive prefix of its row up to column t
  int A = INT_MIN, Bm = INT_MIN;
  unsigned long long mask = 0, m0 = 0;
  for (int r = 0; r < nc; r++) {
    const bool s0 = B.cflag[r] != 0;
    const int v = vol[B.cpart[r]] + wave_incl_scan(lane < nb ? B.D[r * DROW + lane] : 0);
    if (s0) {
      m0 |= 1ull << r;
      mask = v > A ? 1ull << r : (v == A ? mask | 1ull << r : mask);
      A = max(A, v);
    } else {
      Bm = max(Bm, v);
    }
  }
  unsigned long long prev = ((unsigned long long)(unsigned)__shfl_up((int)(unsigned)(mask >> 32), 1) << 32) |
                            (unsigned)__shfl_up((int)(unsigned)(mask & 0xFFFFFFFFull), 1);
  if (lane == 0) prev = m0;
  const unsigned long long sb = __ballot(lane < nb && (Bm >= A || (mask & ~prev) != 0ull));
  const int tstop = sb ? (int)__builtin_ctzll(sb) : -1;
  const int eend = __builtin_amdgcn_readlane(cum_l, sb ? tstop : nb - 1);
  for_staged_events(stage, eend, B.mvod, mi0,
                    [&](int4 x, int2 od, int) { vol_event_changes(x, od, [&](int p, int w) { atomicAdd(&vol[p], w); }); });
  if (lane < nc) B.cidx[B.cpart[lane]] = -1;
  wave_lds_sync();
  M = maxset_init(vol, K, ms);
  return tstop;
}

// One wavefront: the round's moves in order, each followed by the max-set
// update of MGV/partition.cpp:977-1003; stops after the first move whose max
// set holds a part the previous max set did not. The events of consecutive
// moves are contiguous in evbuf, so they are staged into LDS a batch of up to
// 64 moves at a time (stage_cap events). With batched != 0 (K <= 1024) a
// staged batch goes through vol_scan_batch; otherwise, or when its
// candidates do not fit, move by move (a move with more events than the
// stage holds is read from HBM). Writes back vol, the max and its
// multiplicity, and the number of moves to commit (mstat[3]).
__global__ void __launch_bounds__(WAVE) k_vol_scan(int* __restrict__ vol_g, int K, const int4* __restrict__ mv,
                                                   const int4* __restrict__ evbuf, int* __restrict__ mstat,
                                                   int stage_cap, int batched) {
  extern __shared__ __align__(16) unsigned char lds[];
  int4* stage = reinterpret_cast<int4*>(lds);
  int* vol = reinterpret_cast<int*>(stage + stage_cap);
  int* dirty = vol + K;
  VolBatchLds B{};
  B.pos = dirty + WAVE;
  B.neg = B.pos + K;
  B.cidx = B.neg + K;
  B.cpart = B.cidx + K;
  B.cflag = B.cpart + WAVE;
  B.mvod = reinterpret_cast<int2*>(B.cflag + WAVE);
  B.D = reinterpret_cast<int*>(B.mvod + WAVE);
  const int lane = lane_id();
  for (int q = lane; q < K; q += WAVE) {
    vol[q] = vol_g[q];
    if (batched) {
      B.pos[q] = 0;
      B.neg[q] = 0;
      B.cidx[q] = -1;
    }
  }
  dirty[lane] = 0;
  wave_lds_sync();
  LaneMax ms;
  int M = maxset_init(vol, K, ms);
  const int nmoves = mstat[2];
  int nstop = nmoves;
  bool stopped = false;
  // headers of 64 moves from m0: (vertex, own, new, event offset) and the
  // event count, the next offset minus this one (mv ends with a sentinel)
  auto load_head = [&](int m0, int4& h, int& dg) {
    const bool in = m0 + lane < nmoves;
    h = in ? mv[m0 + lane] : make_int4(0, 0, 0, 0);
    dg = in ? mv[m0 + lane + 1].w - h.w : 0;
  };
  int4 hm;
  int hdeg;
  load_head(0, hm, hdeg);
  for (int mi = 0; mi < nmoves && !stopped;) {
    // this batch: the leading moves of the 64 headers at hand whose events fit the stage
    const int cum = wave_incl_scan(hdeg);
    int nb = __popcll(__ballot(mi + lane < nmoves && cum <= stage_cap));
    const bool staged = nb > 0;
    if (!staged) nb = 1;
    const unsigned eb0 = (unsigned)__builtin_amdgcn_readfirstlane(hm.w);
    // the next batch's headers, in flight while this batch runs
    int4 hm_n;
    int hdeg_n;
    load_head(mi + nb, hm_n, hdeg_n);
    int tot = 0;
    if (staged) {
      tot = __builtin_amdgcn_readlane(cum, nb - 1);
      // VOL_STAGE_LOADS loads per lane issued together, then stored
      for (int t0 = 0; t0 < tot; t0 += VOL_STAGE_LOADS * WAVE) {
        int4 e[VOL_STAGE_LOADS];
        // past the end a lane reads and writes the batch's last event again
        // (no branch, so the loads are not moved next to their stores)
#pragma unroll
        for (int u = 0; u < VOL_STAGE_LOADS; u++) e[u] = evbuf[eb0 + (unsigned)min(t0 + u * WAVE + lane, tot - 1)];
        // an empty asm that takes every loaded value: all loads issue before
        // the first store instead of each waiting next to its own store
#pragma unroll
        for (int u = 0; u < VOL_STAGE_LOADS; u++) asm volatile("" ::"v"(e[u].x), "v"(e[u].y), "v"(e[u].z), "v"(e[u].w));
#pragma unroll
        for (int u = 0; u < VOL_STAGE_LOADS; u++) stage[min(t0 + u * WAVE + lane, tot - 1)] = e[u];
      }
      wave_lds_sync();
    }
    int bstop = -2;
    if (batched && staged) bstop = vol_scan_batch(vol, K, stage, tot, nb, mi, hm.y, hm.z, cum, B, ms, M);
    if (bstop >= 0) {
      nstop = mi + bstop + 1;
      stopped = true;
    }
    for (int j = 0; bstop == -2 && j < nb; j++) {
      const int own = __builtin_amdgcn_readlane(hm.y, j), d = __builtin_amdgcn_readlane(hm.z, j);
      const unsigned eb = (unsigned)__builtin_amdgcn_readlane(hm.w, j);
      const int deg = __builtin_amdgcn_readlane(hdeg, j);
      const int4* evs = staged ? stage + (eb - eb0) : evbuf + eb;
      int s_own = 0, s_dec = 0;
      for (int b = 0; b < deg; b += WAVE) {
        const int i = b + lane;
        if (i < deg) {
          const int4 e = evs[i];
          s_own += e.y;
          s_dec += e.z;
          const int sp = vol_event_part(e.x), dsp = vol_event_change(e.x);
          if (sp >= 0 && dsp != 0) {
            atomicAdd(&vol[sp], dsp);
            dirty[sp & (WAVE - 1)] = 1;
          }
        }
      }
      s_own = wave_sum(s_own);
      s_dec = wave_sum(s_dec);
      if (lane == 0) {
        if (s_own) {
          atomicAdd(&vol[own], s_own);
          dirty[own & (WAVE - 1)] = 1;
        }
        if (s_dec) {
          atomicAdd(&vol[d], s_dec);
          dirty[d & (WAVE - 1)] = 1;
        }
      }
      wave_lds_sync();
      if (maxset_step(vol, dirty, K, ms, M)) {
        nstop = mi + j + 1;
        stopped = true;
        break;
      }
    }
    mi += nb;
    hm = hm_n;
    hdeg = hdeg_n;
    wave_lds_sync();
  }
  const int cntM = maxset_count(ms);
  for (int q = lane; q < K; q += WAVE) vol_g[q] = vol[q];
  if (lane == 0) {
    mstat[0] = M;
    mstat[1] = cntM;
    mstat[3] = nstop;
  }
}

// LDS of an evaluation workgroup: 12 bytes per part shared by the workgroup,
// per wavefront a count per part and a bit per part
size_t vol_eval_lds(int K, int waves) {
  return (size_t)K * 12 + (size_t)waves * 4 * ((size_t)K + (size_t)((K + 31) / 32));
}

// Wavefronts per evaluation workgroup: the count that keeps the most
// wavefronts resident on a CU (K = 1024: 8 per workgroup, 3 workgroups, 24
// wavefronts), the larger on ties.
int vol_eval_waves(int K) {
  int best = 0, bw = 0;
  for (int w = 1; w <= VOL_BLOCK / WAVE; w++) {
    const size_t lds = vol_eval_lds(K, w);
    if (lds > (size_t)LDS_BUDGET) break;
    const int per_cu = w * std::min((int)((size_t)LDS_PER_CU / lds), MAX_WAVES_PER_CU / w);
    if (per_cu >= best) {
      best = per_cu;
      bw = w;
    }
  }
  return bw;
}

fp_status ensure_vol_scratch(fp_ctx* ctx, int R) {
  // capacities are reset by every upload, which frees the old buffers
  auto& s = ctx->vol;
  if (ctx->K > s.kcap) {
    FP_TRY(dalloc(ctx, &s.vol, (size_t)ctx->K));
    FP_TRY(dalloc(ctx, &s.mstat, 8));
    s.kcap = ctx->K;
  }
  if (R > s.rmax) {
    FP_TRY(dalloc(ctx, &s.slot_mi, (size_t)R));
    FP_TRY(dalloc(ctx, &s.mv, (size_t)R + 1));  // + the end sentinel k_round_compact writes
    s.rmax = R;
  }
  FP_TRY(ensure_compact_tiles(ctx, R));
  const long long need = std::max(ctx->pins, 1ll);
  if (need > s.evcap) {
    FP_TRY(dalloc(ctx, &s.evbuf, (size_t)need));
    s.evcap = need;
  }
  return FP_OK;
}

}  // namespace

fp_status plan_vol(fp_ctx* ctx, RefinePlan& p) {
  const bool msrv = p.metric == FP_METRIC_MSRV;
  const int nov = ctx->nov, K = ctx->K;
  if (K > VOL_KMAX) return fail(ctx, FP_ERR_UNSUPPORTED, "volume refinement supports K <= %d", VOL_KMAX);
  if (!ctx->src_on_pins)
    return fail(ctx, FP_ERR_UNSUPPORTED, "volume refinement needs net sources that are pins of their nets");
  if (!msrv && ctx->order)
    return fail(ctx, FP_ERR_UNSUPPORTED,
                "MSV refinement visits vertex ids in order (MGV/partition.cpp:755-757); upload with the identity ordering");
  if (ctx->pins >= (1ll << 31)) return fail(ctx, FP_ERR_UNSUPPORTED, "volume refinement needs pins < 2^31");
  if (p.sched.kind != FP_SCHED_REF)
    return fail(ctx, FP_ERR_UNSUPPORTED, "volume refinement runs the reference round schedule (FP_SCHED_REF) only");
  const int waves = p.eval_waves > 0 ? std::min(p.eval_waves, vol_eval_waves(K)) : vol_eval_waves(K);
  if (waves == 0) return fail(ctx, FP_ERR_UNSUPPORTED, "K = %d needs more LDS than a workgroup has", K);
  p.eval_waves = waves;
  p.R = std::min(p.sched.round, nov);
  FP_TRY(ensure_vol_scratch(ctx, p.R));
  if (p.R > ctx->rmax) {
    FP_TRY(dalloc(ctx, &ctx->slot_own, (size_t)p.R));
    FP_TRY(dalloc(ctx, &ctx->slot_dec, (size_t)p.R));
    ctx->rmax = p.R;
  }
  // the volumes of the resident state, kept incrementally by the passes
  auto& vs = ctx->vol;
  fill_async<int>(ctx, vs.vol, (size_t)K, 0);
  const int ngrid = std::max(1, std::min((ctx->noe + 255) / 256, 4096));
  k_vol_init<<<std::min(ngrid * 8, 1024), 256, (size_t)K * sizeof(int), ctx->stream>>>(
      ctx->net_ptr, ctx->ent, ctx->src, ctx->part, ctx->noe, ctx->pmask, K, msrv ? 1 : 0, vs.vol);
  FP_HIP_OK(hipGetLastError());
  return FP_OK;
}

fp_status pass_vol(fp_ctx* ctx, const RefinePlan& p) {
  const bool msrv = p.metric == FP_METRIC_MSRV;
  const int nov = ctx->nov, K = ctx->K, R = p.R, waves = p.eval_waves;
  const size_t lds_e = vol_eval_lds(K, waves);
  // the replay's part -> position index where it fits the LDS budget
  const bool indexed = (size_t)(REPLAY_BLOCK / WAVE) * (size_t)K * (sizeof(unsigned) + sizeof(unsigned short)) <=
                       (size_t)LDS_BUDGET;
  const size_t lds_r = (size_t)(REPLAY_BLOCK / WAVE) * (size_t)K *
                       (sizeof(unsigned) + (indexed ? sizeof(unsigned short) : 0));
  auto replay = indexed ? k_vol_replay<true> : k_vol_replay<false>;
  const int batched = K <= VOL_BATCH_KMAX ? 1 : 0;
  const size_t lds_vol = ((size_t)K + WAVE + (batched ? vol_batch_ints(K) : 0)) * sizeof(int);
  const int stage_cap = (int)std::min<size_t>(4096, ((size_t)LDS_BUDGET - lds_vol) / sizeof(int4));
  const size_t lds_s = lds_vol + (size_t)stage_cap * sizeof(int4);
  auto& vs = ctx->vol;
  pass_begin(ctx, vs.vol, K, vs.mstat);
  unsigned tag = 0;
  for (int r0 = 0; r0 < nov; r0 += R) {
    const int nslots = std::min(R, nov - r0);
    tag++;
    const int par = (int)(tag & 1u);
    VolEvalArgs a{};
    a.vtx_ptr = ctx->vtx_ptr;
    a.vtx_nets = ctx->vtx_nets;
    a.vtx_seg = ctx->vtx_seg;
    a.ev = ctx->ev;
    a.src = ctx->src;
    a.part = ctx->part;
    a.order = ctx->order;
    a.sizes = ctx->sizes;
    a.vol = vs.vol;
    a.mstat = vs.mstat;
    a.K = K;
    a.pb = ctx->pb;
    a.msrv = msrv ? 1 : 0;
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
    FP_TRY(launch_timed(ctx, k_vol_eval, egrid, waves * WAVE, lds_e, a));
    launch_round_compact<false>(ctx, ctx->order, r0, nslots, vs.slot_mi, vs.mv, vs.mstat);
    VolReplayArgs ra{};
    ra.net_ptr = ctx->net_ptr;
    ra.apply_pins = ctx->apply_pins;
    ra.slot_mi = vs.slot_mi;
    ra.mv = vs.mv;
    ra.touched = ctx->touched;
    ra.tcount = ctx->counters + par;
    ra.src = ctx->src;
    ra.vtx_ptr = ctx->vtx_ptr;
    ra.vtx_nets = ctx->vtx_nets;
    ra.mstat = vs.mstat;
    ra.vwgt = ctx->vwgt;
    ra.ent = ctx->ent;
    ra.ev = ctx->ev;
    ra.part = ctx->part;
    ra.sizes = ctx->sizes;
    ra.moves = ctx->moves;
    ra.evbuf = vs.evbuf;
    ra.K = K;
    ra.pb = ctx->pb;
    ra.pmask = ctx->pmask;
    ra.r0 = r0;
    ra.r1 = r0 + nslots;
    ra.msrv = msrv ? 1 : 0;
    ra.net_blocks = net_grid(ctx, REPLAY_NET_BLOCKS, REPLAY_BLOCK / WAVE);
    ra.commit = 0;
    replay<<<ra.net_blocks, REPLAY_BLOCK, lds_r, ctx->stream>>>(ra);
    k_vol_scan<<<1, WAVE, lds_s, ctx->stream>>>(vs.vol, K, vs.mv, vs.evbuf, vs.mstat, stage_cap, batched);
    ra.commit = 1;
    replay<<<ra.net_blocks + (nslots + REPLAY_BLOCK - 1) / REPLAY_BLOCK, REPLAY_BLOCK, lds_r, ctx->stream>>>(ra);
  }
  return FP_OK;
}

}  // namespace fpe
