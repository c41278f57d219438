// fp_metrics.hip -- the six metrics of shared/metric.cpp from scratch on the
// device, in three launches:
//
//  k_metrics   wavefronts take chunks of 64 consecutive nets round robin.
//              Nets of at most 64 pins are packed whole into 64-lane
//              windows, one pin per lane (coalesced pin loads, one part gather
//              per pin); the lanes' (net, part) keys are sorted across the
//              wavefront (bitonic, in registers), so a key that differs from
//              its left neighbour is the first pin of its part in its net.
//              Longer nets go 64 pins at a time through a per-wavefront LDS
//              bitmap of the parts seen. Per-part send / receive / MGV send
//              volumes are summed in block-private LDS counters and stored
//              as one row of partial sums per block (above K = 4096, in one
//              row in HBM by atomics); each (source part,
//              receiving part) pair sets a bit of the message bitmap (a plain
//              read first, so known pairs cost no atomic).
//  k_metrics_colsum   column sums of the partial rows;
//  k_metrics_msgcount the messages sent and received per part, popcounts of
//              the bitmap's rows and columns;
//  k_metrics_final   TM, MSM, MSV and MSRV from the per-part vectors.
//
// Algorithmic bytes per evaluation: 8 per pin (pin id + its part) + 12 per
// net (two offsets read per chunk lane, source id and its part; lambda when
// requested).
#include "fp_internal.h"

#include <mutex>

namespace fpe {
namespace {

constexpr int MET_WAVES = 8;             // wavefronts per block
constexpr int MET_BLOCK = MET_WAVES * WAVE;
constexpr int MET_CHUNK = 64;            // nets per work item
constexpr int MET_GRID_MAX = 1024;       // blocks at most (rows of partial sums)
constexpr int MET_LDS_KMAX = 4096;      // per-part counters in LDS up to here, in HBM above
constexpr int MET_KMAX = 16384;
constexpr unsigned KEY_NONE = 0xFFFFFFFFu;

constexpr int MET_LDS_MSG_KMAX = 512;   // per-block message bitmaps in LDS up to here

// Offsets (in words) of the per-evaluation scratch: the message bitmap (K
// rows of W = ceil(K/32) words) and the partial rows of PR words: PS = 3K + 2
// counters (send as int, receive, MGV send, TV, CN), then, with per-block
// message bitmaps, the block's K x W bitmap.
struct MetLayout {
  int K, W, PS, PR, grid;
  bool lds_acc, lds_msg;
  size_t bitmap_off, partial_off, words;
};

MetLayout met_layout(int K, int noe, int grid_max) {
  MetLayout L{};
  L.K = K;
  L.W = (K + 31) / 32;
  L.PS = 3 * K + 2;
  L.lds_acc = K <= MET_LDS_KMAX;
  L.lds_msg = K <= MET_LDS_MSG_KMAX;
  L.PR = L.PS + (L.lds_msg ? K * L.W : 0);
  const int chunks = std::max(1, (noe + MET_CHUNK - 1) / MET_CHUNK);
  L.grid = std::max(1, std::min(grid_max, (chunks + MET_WAVES - 1) / MET_WAVES));
  L.bitmap_off = 0;
  L.partial_off = L.bitmap_off + (size_t)K * (size_t)L.W;
  L.words = L.partial_off + (size_t)(L.lds_acc ? L.grid : 1) * (size_t)L.PR;
  return L;
}

struct MetricArgs {
  const unsigned* net_ptr;
  const unsigned* net_pins;
  const int* part;
  const int* src;
  int noe, K, nov, pb, W, PS, PR;
  int* lam;             // noe, nullable
  unsigned* bitmap;     // K x W words, zeroed
  unsigned* partial;    // gridDim.x rows of PR words
};

// Inclusive max-scan over the wave (values >= -1); the DPP steps move -1 into
// lanes they do not reach.
__device__ __forceinline__ int wave_incl_max(int x) {
  x = max(x, dpp_i32<0x111>(-1, x));
  x = max(x, dpp_i32<0x112>(-1, x));
  x = max(x, dpp_i32<0x114>(-1, x));
  x = max(x, dpp_i32<0x118>(-1, x));
  x = max(x, dpp_i32<0x142, 0xA>(-1, x));
  x = max(x, dpp_i32<0x143, 0xC>(-1, x));
  return x;
}

// Ascending bitonic sort of one key per lane across the wavefront.
__device__ __forceinline__ unsigned wave_sort_u32(unsigned key) {
  const int lane = lane_id();
#pragma unroll
  for (int k = 2; k <= WAVE; k <<= 1) {
#pragma unroll
    for (int j = k >> 1; j > 0; j >>= 1) {
      const unsigned other = (unsigned)__shfl_xor((int)key, j);
      const bool up = (lane & k) == 0;
      const bool lower = (lane & j) == 0;
      key = (lower == up) ? min(key, other) : max(key, other);
    }
  }
  return key;
}

// Each wavefront queues the messages (sp -> q) it finds in LDS as bit
// indices of the message bitmap and sets them MSG_Q at a time: the bitmap
// words load together, then an atomic OR, whose result nobody waits for,
// goes to each bit a word does not hold yet. A plain read can be stale
// (another L2 took the atomic); that costs a redundant atomic, never a bit.
constexpr int MSG_Q = 128;

__device__ __forceinline__ void msg_flush(unsigned* __restrict__ bitmap, const unsigned* q, int& n) {
  const int lane = lane_id();
  unsigned idx[MSG_Q / WAVE], w[MSG_Q / WAVE];
#pragma unroll
  for (int u = 0; u < MSG_Q / WAVE; u++) {
    const int t = u * WAVE + lane;
    idx[u] = t < n ? q[t] : 0xFFFFFFFFu;
    w[u] = idx[u] != 0xFFFFFFFFu ? bitmap[idx[u] >> 5] : ~0u;
  }
#pragma unroll
  for (int u = 0; u < MSG_Q / WAVE; u++) {
    const unsigned bit = 1u << (idx[u] & 31);
    if (idx[u] != 0xFFFFFFFFu && !(w[u] & bit)) atomicOr(&bitmap[idx[u] >> 5], bit);
  }
  wave_lds_sync();
  n = 0;
}

// Appends the message (sp -> q) of every lane with `want` to the wave's
// queue (flushed first when it could overflow).
__device__ __forceinline__ void msg_push(unsigned* __restrict__ bitmap, unsigned* q, int& n, bool want, int sp,
                                         int qp, int W) {
  const unsigned long long m = __ballot(want);
  if (!m) return;
  const int c = (int)__popcll(m);
  if (n + c > MSG_Q) msg_flush(bitmap, q, n);
  if (want) q[n + (int)__popcll(m & ((1ull << lane_id()) - 1ull))] = (unsigned)sp * (unsigned)W * 32u + (unsigned)qp;
  n += c;
  wave_lds_sync();
}

constexpr int MET_GROUP = 4;  // windows (or long-net steps of 64 pins) loaded together

// LDS_ACC: per-part counters block-private in LDS, stored as row blockIdx.x
// of the partial sums; otherwise (K above MET_LDS_KMAX) atomics go straight
// to partial row 0, zeroed beforehand. LDS_MSG: the block's messages are bits
// of its own bitmap in LDS, stored after its counters (K up to
// MET_LDS_MSG_KMAX); otherwise each wavefront queues them for the global
// bitmap.
template <bool LDS_ACC, bool LDS_MSG>
__global__ void __launch_bounds__(MET_BLOCK) __attribute__((amdgpu_waves_per_eu(8))) k_metrics(MetricArgs a) {
  extern __shared__ __align__(16) unsigned char lds[];
  const int K = a.K, W = a.W, pb = a.pb;
  const unsigned pmask = (1u << pb) - 1u;
  unsigned* acc = LDS_ACC ? reinterpret_cast<unsigned*>(lds) : a.partial;
  int* s_send = reinterpret_cast<int*>(acc);
  unsigned* s_recv = acc + K;
  unsigned* s_msv = s_recv + K;
  unsigned* s_tvcn = LDS_ACC ? s_msv + K : reinterpret_cast<unsigned*>(lds);  // [0] TV, [1] CN
  unsigned* s_msg = s_tvcn + 2;                                                // LDS_MSG: K x W bits
  const int lane = lane_id(), wave = wave_id();
  const int per_wave = W + WAVE + (LDS_MSG ? 0 : MSG_Q);
  unsigned* seen = s_msg + (LDS_MSG ? (size_t)K * (size_t)W : 0) + (size_t)wave * (size_t)per_wave;  // W words
  int* winbuf = reinterpret_cast<int*>(seen + W);  // 64: net starting at each window lane
  unsigned* mq = seen + W + WAVE;                  // !LDS_MSG: MSG_Q queued message bits
  if (LDS_ACC)
    for (int t = (int)threadIdx.x; t < 3 * K; t += (int)blockDim.x) acc[t] = 0;
  if (LDS_MSG)
    for (int t = (int)threadIdx.x; t < K * W; t += (int)blockDim.x) s_msg[t] = 0;
  if (threadIdx.x < 2) s_tvcn[threadIdx.x] = 0;
  for (int t = lane; t < W; t += WAVE) seen[t] = 0;
  __syncthreads();
  unsigned tv = 0, cn = 0;  // this wave's running sums (every lane holds them)
  int mqn = 0;              // queued messages
  // the message (sp -> qp) of every lane with `want`
  auto message = [&](bool want, int sp, int qp) {
    if (LDS_MSG) {
      if (want) atomicOr(&s_msg[(size_t)sp * (size_t)W + (unsigned)(qp >> 5)], 1u << (qp & 31));
    } else {
      msg_push(a.bitmap, mq, mqn, want, sp, qp, W);
    }
  };

  // chunks go to the wavefronts round robin: one counter that every
  // wavefront took its next chunk from serialised those atomics at one
  // address and cost the kernel half its time
  const int waves_total = (int)gridDim.x * MET_WAVES;
  for (int chunk = (int)blockIdx.x * MET_WAVES + wave; (long long)chunk * MET_CHUNK < a.noe; chunk += waves_total) {
    const int c = chunk * MET_CHUNK;
    const int nc = min(MET_CHUNK, a.noe - c);
    // lane j: net c + j
    unsigned start = 0, end = 0;
    int sj = -1, spj = -1;
    if (lane < nc) {
      start = a.net_ptr[c + lane];
      end = a.net_ptr[c + lane + 1];
      sj = a.src[c + lane];
    }
    if (sj >= 0 && sj < a.nov) spj = a.part[sj];
    const int len = (int)(end - start);
    int lamj = 0, occj = 0;  // lane j: lambda and source occurrences of net c + j
    int i = 0;
    while (i < nc) {
      // up to MET_GROUP windows of whole nets of at most 64 pins, from net i on
      int wi[MET_GROUP], wc[MET_GROUP];
      unsigned wbase[MET_GROUP], wp[MET_GROUP];
      int ng = 0;
#pragma unroll
      for (int u = 0; u < MET_GROUP; u++) {
        wi[u] = i;
        wc[u] = 0;
        wbase[u] = 0;
        wp[u] = 0;
        if (ng == u && i < nc) {
          const unsigned base = (unsigned)__builtin_amdgcn_readlane((int)start, i);
          const unsigned long long fit = __ballot(lane >= i && lane < nc && end - base <= (unsigned)WAVE);
          const int cnt = (int)__popcll(fit);
          if (cnt > 0) {
            wbase[u] = base;
            wc[u] = cnt;
            wp[u] = (unsigned)__builtin_amdgcn_readlane((int)end, i + cnt - 1) - base;
            i += cnt;
            ng++;
          }
        }
      }
      int vv[MET_GROUP], qq[MET_GROUP];
#pragma unroll
      for (int u = 0; u < MET_GROUP; u++) vv[u] = (u < ng && (unsigned)lane < wp[u]) ? (int)a.net_pins[wbase[u] + lane] : -1;
#pragma unroll
      for (int u = 0; u < MET_GROUP; u++) qq[u] = vv[u] >= 0 ? a.part[vv[u]] : 0;
#pragma unroll
      for (int u = 0; u < MET_GROUP; u++) {
        if (u >= ng) break;
        const int i0 = wi[u], cnt = wc[u];
        const unsigned base = wbase[u];
        const bool in_win = lane >= i0 && lane < i0 + cnt;
        winbuf[lane] = -1;
        wave_lds_sync();
        if (in_win && len > 0) winbuf[start - base] = lane;
        wave_lds_sync();
        const int owner = max(wave_incl_max(winbuf[lane]), 0);  // net (chunk lane) of this lane's pin
        const int srcp = __shfl(sj, owner);
        unsigned key = KEY_NONE;
        bool is_src = false;
        if (vv[u] >= 0) {
          key = ((unsigned)(owner - i0) << pb) | (unsigned)qq[u];
          is_src = vv[u] == srcp;
        }
        const unsigned long long occm = __ballot(is_src);
        // sorting keeps every net on its own lanes: net j's keys stay in [start_j, end_j)
        key = wave_sort_u32(key);
        const unsigned prev = (unsigned)__shfl_up((int)key, 1);
        const bool fresh = key != KEY_NONE && (lane == 0 || prev != key);
        const unsigned long long fm = __ballot(fresh);
        const int kn = i0 + (int)(key == KEY_NONE ? 0u : key >> pb);
        const int spk = __shfl(spj, min(kn, WAVE - 1));
        const int q = (int)(key & pmask);
        const bool recv = fresh && spk >= 0 && q != spk;
        if (recv) atomicAdd(&s_recv[q], 1u);
        message(recv, spk, q);
        if (in_win) {
          const unsigned o = start - base;
          const unsigned long long seg = len >= WAVE ? ~0ull : (((1ull << len) - 1ull) << o);
          lamj = (int)__popcll(fm & seg);
          occj = (int)__popcll(occm & seg);
        }
      }
      if (ng < MET_GROUP && i < nc) {
        // net i is longer than a wave: MET_GROUP x 64 pins at a time through
        // the LDS bitmap of the parts seen
        const unsigned s0 = (unsigned)__builtin_amdgcn_readlane((int)start, i);
        const unsigned e = (unsigned)__builtin_amdgcn_readlane((int)end, i);
        const int srcL = __builtin_amdgcn_readlane(sj, i);
        const int spL = __builtin_amdgcn_readlane(spj, i);
        int l = 0, occ = 0;
        for (unsigned k0 = s0; k0 < e; k0 += MET_GROUP * WAVE) {
          int lv[MET_GROUP], lq[MET_GROUP];
#pragma unroll
          for (int u = 0; u < MET_GROUP; u++) {
            const unsigned k = k0 + (unsigned)(u * WAVE + lane);
            lv[u] = k < e ? (int)a.net_pins[k] : -1;
          }
#pragma unroll
          for (int u = 0; u < MET_GROUP; u++) lq[u] = lv[u] >= 0 ? a.part[lv[u]] : 0;
#pragma unroll
          for (int u = 0; u < MET_GROUP; u++) {
            bool fresh = false;
            if (lv[u] >= 0) {
              occ += lv[u] == srcL ? 1 : 0;
              const unsigned bit = 1u << (lq[u] & 31);
              fresh = !(atomicOr(&seen[lq[u] >> 5], bit) & bit);
            }
            l += (int)__popcll(__ballot(fresh));
            const bool recv = fresh && spL >= 0 && lq[u] != spL;
            if (recv) atomicAdd(&s_recv[lq[u]], 1u);
            message(recv, spL, lq[u]);
          }
        }
        occ = wave_sum(occ);
        wave_lds_sync();
        for (int t = lane; t < W; t += WAVE) seen[t] = 0;
        wave_lds_sync();
        if (lane == i) {
          lamj = l;
          occj = occ;
        }
        i++;
      }
    }
    // per net: lambda, TV, CN, send volumes
    if (lane < nc) {
      if (a.lam) a.lam[c + lane] = lamj;
      if (spj >= 0) {
        atomicAdd(&s_send[spj], lamj - 1);
        if (occj) atomicAdd(&s_msv[spj], (unsigned)((lamj - 1) * occj));
      }
    }
    tv += (unsigned)wave_sum(lane < nc && lamj != 0 ? lamj - 1 : 0);
    cn += (unsigned)wave_sum(lane < nc && lamj > 1 ? 1 : 0);
  }
  if (!LDS_MSG && mqn) msg_flush(a.bitmap, mq, mqn);
  if (lane == 0) {
    atomicAdd(&s_tvcn[0], tv);
    atomicAdd(&s_tvcn[1], cn);
  }
  __syncthreads();
  if (LDS_ACC) {
    // counters, TV, CN and (LDS_MSG) the block's bitmap are contiguous in LDS
    unsigned* row = a.partial + (size_t)blockIdx.x * (size_t)a.PR;
    for (int t = (int)threadIdx.x; t < a.PR; t += (int)blockDim.x) row[t] = acc[t];
  } else if (threadIdx.x < 2) {
    atomicAdd(&a.partial[3 * (size_t)K + threadIdx.x], s_tvcn[threadIdx.x]);
  }
}

// Column sums of the partial rows into out (zeroed before): block (c, r)
// sums the 64 columns c*64.. over its share of the rows, 16 row groups of 64
// threads with RED_UNROLL loads in flight each, then one atomic per column.
// out = [6 scalars][send K][recv K][sm K][rm K][msv K].
constexpr int RED_ROW_BLOCKS = 16;
constexpr int RED_UNROLL = 4;

__global__ void __launch_bounds__(1024) k_metrics_colsum(const unsigned* __restrict__ partial, int rows, int PR,
                                                         int K, long long* __restrict__ out) {
  const int PS = 3 * K + 2;
  __shared__ long long acc[16][64];
  const int cl = (int)threadIdx.x & 63, g = (int)threadIdx.x >> 6;
  const int col = (int)blockIdx.x * 64 + cl;
  const int per = (rows + (int)gridDim.y - 1) / (int)gridDim.y;
  const int r_lo = (int)blockIdx.y * per, r_hi = min(rows, r_lo + per);
  long long s = 0;
  if (col < PS) {
    const bool is_signed = col < K;  // send volumes: int
    for (int r0 = r_lo + g; r0 < r_hi; r0 += 16 * RED_UNROLL) {
      unsigned x[RED_UNROLL];
#pragma unroll
      for (int u = 0; u < RED_UNROLL; u++) {
        const int r = r0 + 16 * u;
        x[u] = r < r_hi ? partial[(size_t)r * (size_t)PR + (size_t)col] : 0u;
      }
#pragma unroll
      for (int u = 0; u < RED_UNROLL; u++) s += is_signed ? (long long)(int)x[u] : (long long)x[u];
    }
  }
  acc[g][cl] = s;
  __syncthreads();
  if (g == 0 && col < PS) {
    for (int h = 1; h < 16; h++) s += acc[h][cl];
    long long* dst;
    if (col < K) dst = out + 6 + col;                                   // send
    else if (col < 2 * K) dst = out + 6 + K + (col - K);                // recv
    else if (col < 3 * K) dst = out + 6 + 4 * (size_t)K + (col - 2 * K);  // MGV send
    else dst = out + (col - 3 * K);                                     // TV, CN
    if (s) atomicAdd(reinterpret_cast<unsigned long long*>(dst), (unsigned long long)s);
  }
}

// The message bitmap (zeroed before) as the OR of the blocks' bitmaps, which
// follow their counters in the partial rows: block (c, r) ORs words c*1024..
// over its share of the rows, RED_UNROLL loads in flight, then one atomic OR
// per word with bits.
__global__ void __launch_bounds__(1024) k_metrics_bitor(const unsigned* __restrict__ partial, int rows, int PR,
                                                        int PS, int words, unsigned* __restrict__ bitmap) {
  const int w = (int)blockIdx.x * 1024 + (int)threadIdx.x;
  if (w >= words) return;
  const int per = (rows + (int)gridDim.y - 1) / (int)gridDim.y;
  const int r_lo = (int)blockIdx.y * per, r_hi = min(rows, r_lo + per);
  unsigned m = 0;
  for (int r0 = r_lo; r0 < r_hi; r0 += RED_UNROLL) {
    unsigned x[RED_UNROLL];
#pragma unroll
    for (int u = 0; u < RED_UNROLL; u++)
      x[u] = r0 + u < r_hi ? partial[(size_t)(r0 + u) * (size_t)PR + (size_t)PS + (size_t)w] : 0u;
#pragma unroll
    for (int u = 0; u < RED_UNROLL; u++) m |= x[u];
  }
  if (m) atomicOr(&bitmap[w], m);
}

// Messages sent and received per part from the message bitmap (K rows of W
// words, zeroed out[6 + 2K ..] before): block b takes rows b*16 .. b*16+15;
// sent: one wavefront per row sums its words' popcounts; received: thread t
// counts column t (and t + 1024, ...) over the block's rows, one atomic per
// column with a count.
constexpr int MSG_ROWS = 16;

__global__ void __launch_bounds__(1024) k_metrics_msgcount(const unsigned* __restrict__ bitmap, int K, int W,
                                                           long long* __restrict__ out) {
  const int p0 = (int)blockIdx.x * MSG_ROWS, p1 = min(K, p0 + MSG_ROWS);
  const int lane = lane_id(), wave = (int)threadIdx.x / WAVE;
  for (int p = p0 + wave; p < p1; p += (int)blockDim.x / WAVE) {
    int c = 0;
    for (int w = lane; w < W; w += WAVE) c += __popc(bitmap[(size_t)p * (size_t)W + w]);
    c = wave_sum(c);
    if (lane == 0) out[6 + 2 * (size_t)K + p] = c;
  }
  for (int q = (int)threadIdx.x; q < K; q += (int)blockDim.x) {
    const unsigned bit = 1u << (q & 31);
    unsigned x[MSG_ROWS];
#pragma unroll
    for (int u = 0; u < MSG_ROWS; u++) x[u] = p0 + u < p1 ? bitmap[(size_t)(p0 + u) * (size_t)W + (q >> 5)] : 0u;
    int c = 0;
#pragma unroll
    for (int u = 0; u < MSG_ROWS; u++) c += (x[u] & bit) ? 1 : 0;
    if (c) atomicAdd(reinterpret_cast<unsigned long long*>(out + 6 + 3 * (size_t)K + q), (unsigned long long)c);
  }
}

// TM = sum of messages sent, MSM = their max, MSV = max MGV send volume,
// MSRV = max of send + receive volume (shared/metric.cpp:122-391).
__global__ void __launch_bounds__(1024) k_metrics_final(int K, long long* __restrict__ out) {
  __shared__ long long s_r[4][16];
  const long long* send = out + 6;
  const long long* recv = send + K;
  const long long* sm = send + 2 * (size_t)K;
  const long long* msv = send + 4 * (size_t)K;
  long long tm = 0, msm = -1, mv = -1, msrv = -1;
  for (int k = (int)threadIdx.x; k < K; k += (int)blockDim.x) {
    tm += sm[k];
    msm = max(msm, sm[k]);
    mv = max(mv, msv[k]);
    msrv = max(msrv, send[k] + recv[k]);
  }
  for (int off = WAVE / 2; off > 0; off >>= 1) {
    tm += __shfl_xor(tm, off);
    msm = max(msm, __shfl_xor(msm, off));
    mv = max(mv, __shfl_xor(mv, off));
    msrv = max(msrv, __shfl_xor(msrv, off));
  }
  const int w = (int)threadIdx.x / WAVE;
  if (lane_id() == 0) {
    s_r[0][w] = tm;
    s_r[1][w] = msm;
    s_r[2][w] = mv;
    s_r[3][w] = msrv;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const int nw = (int)blockDim.x / WAVE;
    for (int h = 1; h < nw; h++) {
      tm += s_r[0][h];
      msm = max(msm, s_r[1][h]);
      mv = max(mv, s_r[2][h]);
      msrv = max(msrv, s_r[3][h]);
    }
    out[2] = tm;
    out[3] = msm;
    out[4] = mv;
    out[5] = msrv;
  }
}

size_t met_lds_bytes(const MetLayout& L) {
  const size_t acc = L.lds_acc ? (size_t)3 * L.K : 0;
  const size_t msg = L.lds_msg ? (size_t)L.K * (size_t)L.W : 0;
  const size_t per_wave = (size_t)L.W + WAVE + (L.lds_msg ? 0 : MSG_Q);
  return (acc + 2 + msg + (size_t)MET_WAVES * per_wave) * sizeof(unsigned);
}

fp_status met_check(fp_ctx* ctx, int K) {
  if (K > MET_KMAX) return fail(ctx, FP_ERR_UNSUPPORTED, "K = %d above %d for the metric kernel", K, MET_KMAX);
  // per-block partial sums are 32-bit: a block's sums stay below the pin
  // count, except the MGV send volume (lambda - 1) * occurrences of the
  // source among a net's pins, which only repeated pins can push past it
  if (ctx->pins >= (1ll << 31) || (ctx->dup_pins && (unsigned long long)ctx->max_net * ctx->max_net >= (1ull << 31)))
    return fail(ctx, FP_ERR_UNSUPPORTED, "hypergraph too large for the metric kernel's 32-bit partial sums");
  return FP_OK;
}

// Enqueues the three launches on ctx->stream. scratch: L.words words; out: 6 + 5K.
fp_status met_launch(fp_ctx* ctx, const MetLayout& L, unsigned* scratch, long long* out, int* lam) {
  // the bitmap; without LDS counters also partial row 0
  fill_async<unsigned>(ctx, scratch, L.partial_off + (L.lds_acc ? 0 : (size_t)L.PR), 0u);
  MetricArgs m{};
  m.net_ptr = ctx->net_ptr;
  m.net_pins = ctx->net_pins;
  m.part = ctx->part;
  m.src = ctx->src;
  m.noe = ctx->noe;
  m.K = L.K;
  m.nov = ctx->nov;
  m.pb = ctx->pb;
  m.W = L.W;
  m.PS = L.PS;
  m.PR = L.PR;
  m.lam = lam;
  m.bitmap = scratch + L.bitmap_off;
  m.partial = scratch + L.partial_off;
  const size_t lds = met_lds_bytes(L);
  if (L.lds_msg) k_metrics<true, true><<<L.grid, MET_BLOCK, lds, ctx->stream>>>(m);
  else if (L.lds_acc) k_metrics<true, false><<<L.grid, MET_BLOCK, lds, ctx->stream>>>(m);
  else k_metrics<false, false><<<L.grid, MET_BLOCK, lds, ctx->stream>>>(m);
  // the output vectors are sums of atomics: zero first (a fill kernel, so
  // that the evaluation can be captured and replayed)
  fill_async<long long>(ctx, out, (size_t)6 + 5 * (size_t)L.K, 0ll);
  const int rows = L.lds_acc ? L.grid : 1;
  const unsigned ry = (unsigned)std::min(RED_ROW_BLOCKS, rows);
  k_metrics_colsum<<<dim3((unsigned)((L.PS + 63) / 64), ry), 1024, 0, ctx->stream>>>(m.partial, rows, L.PR, L.K, out);
  if (L.lds_msg) {
    const int words = L.K * L.W;
    k_metrics_bitor<<<dim3((unsigned)((words + 1023) / 1024), ry), 1024, 0, ctx->stream>>>(m.partial, rows, L.PR,
                                                                                           L.PS, words, m.bitmap);
  }
  k_metrics_msgcount<<<(L.K + MSG_ROWS - 1) / MSG_ROWS, 1024, 0, ctx->stream>>>(m.bitmap, L.K, L.W, out);
  k_metrics_final<<<1, 1024, 0, ctx->stream>>>(L.K, out);
  if (hipGetLastError() != hipSuccess) return fail(ctx, FP_ERR_HIP, "metric kernel launch failed");
  return FP_OK;
}

// ---- the allocation-free evaluation's scratch -------------------------------
// fp_eval_noalloc keeps the message bitmap and the partial rows in this
// file-scope array instead of context memory: K <= NOALLOC_KMAX, at most
// NOALLOC_GRID blocks. Each GPU the library runs on holds its own copy from
// the library's load until the process ends (5 MB).
constexpr int NOALLOC_KMAX = 1024;
constexpr int NOALLOC_GRID = 128;

constexpr size_t noalloc_words(int K) {
  return (size_t)K * (size_t)((K + 31) / 32) +
         (size_t)NOALLOC_GRID * ((size_t)(3 * K + 2) + (K <= MET_LDS_MSG_KMAX ? (size_t)K * (size_t)((K + 31) / 32) : 0));
}
// the largest layout over K <= NOALLOC_KMAX: at the last K with block
// bitmaps or at the maximum K
constexpr size_t NOALLOC_WORDS =
    noalloc_words(MET_LDS_MSG_KMAX) > noalloc_words(NOALLOC_KMAX) ? noalloc_words(MET_LDS_MSG_KMAX)
                                                                   : noalloc_words(NOALLOC_KMAX);
__device__ __align__(256) unsigned scratch[NOALLOC_WORDS];

constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
unsigned* g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

// The device address of `scratch` on ctx's device, looked up once per device
// with that device current (ctx's calls set it) and kept; host code uses
// only this address, never the array's name.
fp_status noalloc_scratch(fp_ctx* ctx, unsigned** out) {
  if (ctx->device < 0 || ctx->device >= NOALLOC_MAX_DEVICES)
    return fail(ctx, FP_ERR_UNSUPPORTED, "device %d beyond the scratch table", ctx->device);
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (!g_scratch_ptr[ctx->device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      (void)hipGetLastError();
      return fail(ctx, FP_ERR_HIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    }
    g_scratch_ptr[ctx->device] = static_cast<unsigned*>(p);
  }
  *out = g_scratch_ptr[ctx->device];
  return FP_OK;
}

}  // namespace
}  // namespace fpe

using namespace fpe;

extern "C" {

fp_status fp_eval(fp_ctx* ctx, fp_metrics* out, int* lam, long long* sv, long long* rv, long long* sm, long long* rm,
                  long long* sv_msv) {
  if (!ctx) return FP_ERR_ARG;
  if (!ctx->has_state) return fail(ctx, FP_ERR_STATE, "fp_eval before a state was set");
  FP_HIP_OK(hipSetDevice(ctx->device));
  const int K = ctx->K;
  FP_TRY(met_check(ctx, K));
  const MetLayout L = met_layout(K, ctx->noe, MET_GRID_MAX);
  if (L.words > ctx->met_words) {
    FP_TRY(dalloc(ctx, &ctx->met_scratch, L.words));
    ctx->met_words = L.words;
  }
  if ((size_t)(6 + 5 * K) > ctx->met_out_cap) {
    FP_TRY(dalloc(ctx, &ctx->met_out, (size_t)(6 + 5 * K)));
    ctx->met_out_cap = (size_t)(6 + 5 * K);
  }
  if (lam && !ctx->lam) FP_TRY(dalloc(ctx, &ctx->lam, (size_t)std::max(ctx->noe, 1)));
  FP_TRY(met_launch(ctx, L, ctx->met_scratch, ctx->met_out, lam ? ctx->lam : nullptr));
  std::vector<long long> h((size_t)6 + 5 * (size_t)K);
  fp_status st = d2h(ctx, h.data(), ctx->met_out, h.size());
  if (st == FP_OK && lam) st = d2h(ctx, lam, ctx->lam, (size_t)ctx->noe);
  if (hipStreamSynchronize(ctx->stream) != hipSuccess && st == FP_OK) st = fail(ctx, FP_ERR_HIP, "sync failed");
  if (st != FP_OK) return st;
  const long long* v = h.data() + 6;
  if (out) {
    out->tv = h[0];
    out->cn = h[1];
    out->tm = h[2];
    out->msm = h[3];
    out->msv = h[4];
    out->msrv = h[5];
  }
  if (sv) std::copy(v, v + K, sv);
  if (rv) std::copy(v + K, v + 2 * (size_t)K, rv);
  if (sm) std::copy(v + 2 * (size_t)K, v + 3 * (size_t)K, sm);
  if (rm) std::copy(v + 3 * (size_t)K, v + 4 * (size_t)K, rm);
  if (sv_msv) std::copy(v + 4 * (size_t)K, v + 5 * (size_t)K, sv_msv);
  return FP_OK;
}

fp_status fp_eval_noalloc(fp_ctx* ctx, long long* d_out, int* d_lam) {
  if (!ctx || !d_out) return FP_ERR_ARG;
  if (!ctx->has_state) return fail(ctx, FP_ERR_STATE, "fp_eval_noalloc before a state was set");
  FP_HIP_OK(hipSetDevice(ctx->device));
  const int K = ctx->K;
  if (K > NOALLOC_KMAX)
    return fail(ctx, FP_ERR_UNSUPPORTED, "K = %d above %d, the allocation-free evaluation's maximum", K, NOALLOC_KMAX);
  FP_TRY(met_check(ctx, K));
  const MetLayout L = met_layout(K, ctx->noe, NOALLOC_GRID);
  if (L.words > NOALLOC_WORDS) return fail(ctx, FP_ERR_UNSUPPORTED, "metric scratch beyond the static maximum");
  unsigned* sp = nullptr;
  FP_TRY(noalloc_scratch(ctx, &sp));
  return met_launch(ctx, L, sp, d_out, d_lam);
}

}  // extern "C"
