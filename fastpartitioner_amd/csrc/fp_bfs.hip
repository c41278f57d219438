// fp_bfs.hip -- BFS vertex ordering on the device (shared/utils.cpp:433-590,
// driven as TV/tv.cpp:110-128: a BFS from the start vertex, one from the
// farthest vertex it found, then the visit order of a BFS from the farthest
// vertex of that one).
//
// The reference's BFS on one thread visits, level by level, each vertex of the
// level in queue order, its nets in vertex -> nets order and each net's pins
// in the net view's order; an unvisited pin joins the next level. Each net
// needs expanding once only (the first time: afterwards all its pins are
// visited), by the first frontier vertex that holds it. Level-synchronous on
// the device, with the same result:
//   k_bfs_claim   one wavefront per frontier vertex (position p): every net j
//                 of it not expanded yet keeps the smallest (p, j) claim; the
//                 first claim of a net puts it on the level's net list (for a
//                 large frontier, k_bfs_claim_nets finds the same claims from
//                 the nets' side, without atomics)
//   k_bfs_expand  each claimed net (a lane for a small net, the wavefront
//                 for a large one): every unvisited pin k keeps
//                 the smallest key (p, j, k) -- its first discovery in the
//                 sequential order; the first key of a pin puts it on the
//                 level's candidate list
//   k_bfs_mark    candidates get their level, and their keys go to a sort
//   radix sort    by key (rocPRIM through hipCUB): the next level, in the
//                 sequential visit order, written straight into the ordering
#include <hipcub/hipcub.hpp>

#include "fp_internal.h"

namespace fpe {
namespace {

constexpr int BFS_BLOCK = 256;
constexpr unsigned long long NO_KEY = ~0ull;
// nets of at most this many pins are taken by one lane each (four pins per
// load round trip) in k_bfs_claim_nets and k_bfs_expand; larger ones by the
// whole wavefront
constexpr unsigned BFS_CLAIM_LANE_PINS = 32;
constexpr unsigned BFS_EXPAND_LANE_PINS = 8;

__global__ void __launch_bounds__(BFS_BLOCK) k_bfs_claim(const int* __restrict__ frontier, int nf,
                                                         const unsigned* __restrict__ vtx_ptr,
                                                         const unsigned* __restrict__ vtx_nets,
                                                         const unsigned char* __restrict__ net_done,
                                                         unsigned long long* __restrict__ net_owner,
                                                         int* __restrict__ level_nets, int* __restrict__ nlevel) {
  const int lane = lane_id(), nw = (int)(blockDim.x / WAVE);
  __shared__ unsigned s_buf[BFS_BLOCK / WAVE][TOUCH_BUF];
  __shared__ int s_tn[BFS_BLOCK / WAVE + 1];
  WaveTouch st{s_buf[wave_id()], 0};
  for (int p = (int)blockIdx.x * nw + wave_id(); p < nf; p += (int)gridDim.x * nw) {
    const int v = frontier[p];
    const unsigned vb = vtx_ptr[v], deg = vtx_ptr[v + 1] - vb;
    for (unsigned j0 = 0; j0 < deg; j0 += WAVE) {
      const unsigned j = j0 + (unsigned)lane;
      bool first = false;
      int e = 0;
      if (j < deg) {
        e = (int)vtx_nets[vb + j];
        // a claim at or above the one already read cannot win: no atomic
        // (a stale read only costs an atomic, as claims only decrease)
        const unsigned long long key = ((unsigned long long)p << 32) | j;
        if (!net_done[e] && key < net_owner[e]) first = atomicMin(&net_owner[e], key) == NO_KEY;
      }
      wave_stage_append(first, (unsigned)e, st, level_nets, nlevel);
    }
  }
  block_touch_flush(st, level_nets, nlevel, s_tn);
}

// (position of v in the frontier, index of net e in v's ascending net list)
__device__ __forceinline__ unsigned long long bfs_claim_key(int v, int e, const unsigned* __restrict__ vtx_ptr,
                                                            const unsigned* __restrict__ vtx_nets,
                                                            const int* __restrict__ pos) {
  unsigned lo = vtx_ptr[v], hi = vtx_ptr[v + 1];
  const unsigned base = lo;
  while (lo < hi) {
    const unsigned mid = (lo + hi) >> 1;
    if (vtx_nets[mid] < (unsigned)e) lo = mid + 1;
    else hi = mid;
  }
  return ((unsigned long long)pos[v] << 32) | (lo - base);
}

// The same claims from the nets' side, for a large frontier (claims from the
// frontier would put thousands of atomics on each big net's owner word):
// one wavefront per net not expanded yet reads its pins, and the pins in the
// frontier (plev[v] == level, position pos[v]) give the net's smallest
// (position, index of the net in the pin's net list) claim.
__global__ void __launch_bounds__(BFS_BLOCK) k_bfs_claim_nets(const unsigned* __restrict__ net_ptr,
                                                              const unsigned* __restrict__ pins, int noe,
                                                              const unsigned* __restrict__ vtx_ptr,
                                                              const unsigned* __restrict__ vtx_nets,
                                                              const unsigned char* __restrict__ net_done,
                                                              const int* __restrict__ plev,
                                                              const int* __restrict__ pos, int level,
                                                              unsigned long long* __restrict__ net_owner,
                                                              int* __restrict__ level_nets, int* __restrict__ nlevel) {
  const int lane = lane_id(), nw = (int)(blockDim.x / WAVE);
  __shared__ unsigned s_buf[BFS_BLOCK / WAVE][TOUCH_BUF];
  __shared__ int s_tn[BFS_BLOCK / WAVE + 1];
  WaveTouch st{s_buf[wave_id()], 0};
  for (int e0 = ((int)blockIdx.x * nw + wave_id()) * WAVE; e0 < noe; e0 += (int)gridDim.x * nw * WAVE) {
    // 64 nets per wave, one per lane: a small open net is scanned by its
    // lane, four pins per load round trip; the large ones by the wave in turn
    const int ee = e0 + lane;
    const bool open = ee < noe && !net_done[ee];
    unsigned ns = 0, nlen = 0;
    if (open) {
      ns = net_ptr[ee];
      nlen = net_ptr[ee + 1] - ns;
    }
    const bool mine = open && nlen <= BFS_CLAIM_LANE_PINS;
    bool any = false;
    if (mine) {
      unsigned long long best = NO_KEY;
      for (unsigned k0 = 0; k0 < nlen; k0 += 4) {
        int v[4], lv[4];
#pragma unroll
        for (int q = 0; q < 4; q++) v[q] = k0 + q < nlen ? (int)pins[ns + k0 + q] : -1;
#pragma unroll
        for (int q = 0; q < 4; q++) lv[q] = v[q] >= 0 ? plev[v[q]] : -2;
#pragma unroll
        for (int q = 0; q < 4; q++)
          if (lv[q] == level) best = min(best, bfs_claim_key(v[q], ee, vtx_ptr, vtx_nets, pos));
      }
      if (best != NO_KEY) {
        net_owner[ee] = best;
        any = true;
      }
    }
    unsigned long long big = __ballot(open && !mine);
    while (big) {
      const int e = e0 + __ffsll((long long)big) - 1;
      big &= big - 1;
      const unsigned s = net_ptr[e], len = net_ptr[e + 1] - s;
      unsigned long long best = NO_KEY;
      for (unsigned k0 = 0; k0 < len; k0 += WAVE) {
        const unsigned k = k0 + (unsigned)lane;
        if (k < len) {
          const int v = (int)pins[s + k];
          if (plev[v] == level) best = min(best, bfs_claim_key(v, e, vtx_ptr, vtx_nets, pos));
        }
      }
      best = wave_min_u64(best);
      if (best != NO_KEY && lane == e - e0) {
        net_owner[e] = best;
        any = true;
      }
    }
    wave_stage_append(any, (unsigned)ee, st, level_nets, nlevel);
  }
  block_touch_flush(st, level_nets, nlevel, s_tn);
}

// positions of the frontier's vertices for k_bfs_claim_nets
__global__ void __launch_bounds__(BFS_BLOCK) k_bfs_positions(const int* __restrict__ frontier, int nf, int level,
                                                             int* __restrict__ plev, int* __restrict__ pos) {
  for (int i = (int)(blockIdx.x * blockDim.x + threadIdx.x); i < nf; i += (int)(gridDim.x * blockDim.x)) {
    const int v = frontier[i];
    plev[v] = level;
    pos[v] = i;
  }
}

__global__ void __launch_bounds__(BFS_BLOCK) k_bfs_expand(const int* __restrict__ level_nets,
                                                          const int* __restrict__ nlevel,
                                                          const unsigned* __restrict__ net_ptr,
                                                          const unsigned* __restrict__ pins,
                                                          const unsigned long long* __restrict__ net_owner,
                                                          unsigned char* __restrict__ net_done,
                                                          const int* __restrict__ dist, int kb,
                                                          unsigned long long* __restrict__ vkey,
                                                          int* __restrict__ cand, int* __restrict__ ncand) {
  const int lane = lane_id(), nw = (int)(blockDim.x / WAVE);
  __shared__ unsigned s_buf[BFS_BLOCK / WAVE][TOUCH_BUF];
  __shared__ int s_tn[BFS_BLOCK / WAVE + 1];
  WaveTouch st{s_buf[wave_id()], 0};
  const int n = *nlevel;
  // c of the level's nets per wave (c = the level's share per wave of the
  // grid, 1 to 64), one per lane: a small net is expanded by its lane, four
  // pins per load round trip; the large ones by the wave in turn (so a level
  // of few nets still spreads them over the grid)
  const int tw = (int)gridDim.x * nw;
  const int c = (int)min((long long)WAVE, max(1LL, ((long long)n + tw - 1) / tw));
  for (int i0 = ((int)blockIdx.x * nw + wave_id()) * c; i0 < n; i0 += tw * c) {
    const int i = i0 + lane;
    int e = -1;
    unsigned s = 0, len = 0;
    unsigned long long hi = 0, jk = 0;
    if (lane < c && i < n) {
      e = level_nets[i];
      const unsigned long long own = net_owner[e];
      // (p, j) of the owner, then the pin's index in the net
      hi = own & 0xFFFFFFFF00000000ull;
      jk = (own & 0xFFFFFFFFull) << kb;
      net_done[e] = 1;
      s = net_ptr[e];
      len = net_ptr[e + 1] - s;
    }
    const bool mine = e >= 0 && len <= BFS_EXPAND_LANE_PINS;
    const unsigned mlen = (unsigned)wave_max(mine ? (int)len : 0);
    for (unsigned k0 = 0; k0 < mlen; k0 += 4) {
      int u[4];
      bool first[4];
#pragma unroll
      for (int q = 0; q < 4; q++) u[q] = mine && k0 + q < len ? (int)pins[s + k0 + q] : -1;
#pragma unroll
      for (int q = 0; q < 4; q++) first[q] = u[q] >= 0 && dist[u[q]] < 0;
#pragma unroll
      for (int q = 0; q < 4; q++)
        if (first[q]) first[q] = atomicMin(&vkey[u[q]], hi | jk | (k0 + q)) == NO_KEY;
#pragma unroll
      for (int q = 0; q < 4; q++) wave_stage_append(first[q], (unsigned)u[q], st, cand, ncand);
    }
    unsigned long long big = __ballot(e >= 0 && !mine);
    while (big) {
      const int b = __ffsll((long long)big) - 1;
      big &= big - 1;
      const unsigned bs = (unsigned)__shfl((int)s, b), blen = (unsigned)__shfl((int)len, b);
      const unsigned long long bhi = (unsigned long long)(unsigned)__shfl((int)(hi >> 32), b) << 32;
      const unsigned long long bjk = ((unsigned long long)(unsigned)__shfl((int)(jk >> 32), b) << 32) |
                                     (unsigned)__shfl((int)(unsigned)jk, b);
      for (unsigned k0 = 0; k0 < blen; k0 += WAVE) {
        const unsigned k = k0 + (unsigned)lane;
        bool first = false;
        int u = 0;
        if (k < blen) {
          u = (int)pins[bs + k];
          if (dist[u] < 0) first = atomicMin(&vkey[u], bhi | bjk | k) == NO_KEY;
        }
        wave_stage_append(first, (unsigned)u, st, cand, ncand);
      }
    }
  }
  block_touch_flush(st, cand, ncand, s_tn);
}

__global__ void __launch_bounds__(BFS_BLOCK) k_bfs_mark(const int* __restrict__ cand, int n, int level,
                                                        int* __restrict__ dist,
                                                        const unsigned long long* __restrict__ vkey,
                                                        unsigned long long* __restrict__ keys, int* __restrict__ vals) {
  for (int i = (int)(blockIdx.x * blockDim.x + threadIdx.x); i < n; i += (int)(gridDim.x * blockDim.x)) {
    const int u = cand[i];
    dist[u] = level;
    keys[i] = vkey[u];
    vals[i] = u;
  }
}

// the lowest vertex id of a level (the farthest vertex: shared/utils.cpp:466-473)
__global__ void __launch_bounds__(1024) k_bfs_min(const int* __restrict__ level, int n, int* __restrict__ out) {
  __shared__ int s_m[16];
  int m = INT_MAX;
  for (int i = (int)threadIdx.x; i < n; i += (int)blockDim.x) m = min(m, level[i]);
  for (int off = WAVE / 2; off > 0; off >>= 1) m = min(m, __shfl_xor(m, off));
  if (lane_id() == 0) s_m[threadIdx.x / WAVE] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < (int)(blockDim.x / WAVE); w++) m = min(m, s_m[w]);
    *out = min(m, s_m[0]);
  }
}

int bits_of(unsigned long long x) {
  int b = 0;
  while (b < 64 && (x >> b) != 0) b++;
  return b;
}

struct BfsScratch {
  int* dist = nullptr;
  int* plev = nullptr;
  int* pos = nullptr;
  unsigned long long* vkey = nullptr;
  unsigned long long* net_owner = nullptr;
  unsigned char* net_done = nullptr;
  int* level_nets = nullptr;
  int* cand = nullptr;
  unsigned long long* keys = nullptr;
  unsigned long long* keys_out = nullptr;
  int* vals = nullptr;
  int* order = nullptr;
  int* counts = nullptr;  // [0] level nets, [1] candidates, [2] min
  void* temp = nullptr;
  size_t temp_bytes = 0;
  std::vector<void*> all;
  ~BfsScratch() {
    for (void* p : all) (void)hipFree(p);
  }
  template <typename T>
  bool get(T** p, size_t n) {
    void* q = nullptr;
    if (hipMalloc(&q, sizeof(T) * std::max<size_t>(n, 1)) != hipSuccess) return false;
    all.push_back(q);
    *p = static_cast<T*>(q);
    return true;
  }
};

// One BFS from start: the visit order into s.order[0 .. *reached), the last
// level's lowest vertex into *farthest.
// ordered: each level in the sequential visit order (radix sort by key);
// otherwise in any order, which gives the same levels and farthest vertex
fp_status bfs_run(fp_ctx* ctx, BfsScratch& s, const unsigned* pins, int kb, int start, int* reached, int* farthest,
                  int* levels, bool ordered) {
  const int nov = ctx->nov, noe = ctx->noe;
  fill_async<int>(ctx, s.dist, (size_t)nov, -1);
  fill_async<int>(ctx, s.plev, (size_t)nov, -1);
  fill_async<unsigned long long>(ctx, s.vkey, (size_t)nov, NO_KEY);
  fill_async<unsigned long long>(ctx, s.net_owner, (size_t)std::max(noe, 1), NO_KEY);
  fill_async<unsigned char>(ctx, s.net_done, (size_t)std::max(noe, 1), (unsigned char)0);
  fill_async<int>(ctx, s.dist + start, 1, 0);
  fill_async<int>(ctx, s.order, 1, start);
  int count = 1, nf = 1, level = 0;
  int* frontier = s.order;
  for (;;) {
    fill_async<int>(ctx, s.counts, 2, 0);
    // a frontier whose incidences reach an eighth of the nets claims from
    // the nets' side
    if ((double)nf * (double)ctx->pins / (double)std::max(nov, 1) * 8.0 >= (double)noe) {
      k_bfs_positions<<<std::max(1, std::min((nf + BFS_BLOCK - 1) / BFS_BLOCK, 4096)), BFS_BLOCK, 0, ctx->stream>>>(
          frontier, nf, level, s.plev, s.pos);
      k_bfs_claim_nets<<<std::max(1, std::min((noe + 255) / 256, 8192)), BFS_BLOCK, 0, ctx->stream>>>(
          ctx->net_ptr, pins, noe, ctx->vtx_ptr, ctx->vtx_nets, s.net_done, s.plev, s.pos, level, s.net_owner,
          s.level_nets, s.counts);
    } else {
      const int cgrid = std::max(1, std::min((nf + 3) / 4, 8192));
      k_bfs_claim<<<cgrid, BFS_BLOCK, 0, ctx->stream>>>(frontier, nf, ctx->vtx_ptr, ctx->vtx_nets, s.net_done,
                                                        s.net_owner, s.level_nets, s.counts);
    }
    k_bfs_expand<<<8192, BFS_BLOCK, 0, ctx->stream>>>(s.level_nets, s.counts, ctx->net_ptr, pins, s.net_owner,
                                                      s.net_done, s.dist, kb, s.vkey, s.cand, s.counts + 1);
    FP_HIP_OK(hipGetLastError());
    int nc = 0;
    FP_TRY(d2h(ctx, &nc, s.counts + 1, 1));
    FP_HIP_OK(hipStreamSynchronize(ctx->stream));
    if (nc == 0) break;
    level++;
    k_bfs_mark<<<std::max(1, std::min((nc + BFS_BLOCK - 1) / BFS_BLOCK, 4096)), BFS_BLOCK, 0, ctx->stream>>>(
        s.cand, nc, level, s.dist, s.vkey, s.keys, ordered ? s.vals : s.order + count);
    if (ordered) {
      // keys: position of the parent in its level (< 2^32) above (net index, pin index)
      const int end_bit = std::min(64, 32 + bits_of((unsigned long long)nf));
      size_t tb = s.temp_bytes;
      if (hipcub::DeviceRadixSort::SortPairs(s.temp, tb, s.keys, s.keys_out, s.vals, s.order + count, nc, 0, end_bit,
                                             ctx->stream) != hipSuccess)
        return fail(ctx, FP_ERR_HIP, "BFS level sort failed");
    }
    frontier = s.order + count;
    nf = nc;
    count += nc;
  }
  k_bfs_min<<<1, 1024, 0, ctx->stream>>>(frontier, nf, s.counts + 2);
  FP_HIP_OK(hipGetLastError());
  FP_TRY(d2h(ctx, farthest, s.counts + 2, 1));
  FP_HIP_OK(hipStreamSynchronize(ctx->stream));
  *reached = count;
  if (levels) *levels = level;
  return FP_OK;
}

}  // namespace
}  // namespace fpe

using namespace fpe;

extern "C" {

fp_status fp_bfs_ordering(fp_ctx* ctx, int start, int view, int* ordering, int* farthest_first, int* farthest_all,
                          fp_bfs_stats* stats) {
  if (!ctx) return FP_ERR_ARG;
  if (!ctx->has_graph) return fail(ctx, FP_ERR_STATE, "fp_bfs_ordering before fp_upload");
  if (!ordering) return fail(ctx, FP_ERR_ARG, "null ordering");
  const int nov = ctx->nov, noe = ctx->noe;
  if (start < 0 || start >= nov) return fail(ctx, FP_ERR_ARG, "start vertex %d out of [0, %d)", start, nov);
  if (view != FP_BFS_PINS_ASCENDING && view != FP_BFS_PINS_UPLOADED)
    return fail(ctx, FP_ERR_ARG, "unknown net view %d", view);
  if (view == FP_BFS_PINS_ASCENDING && ctx->order)
    return fail(ctx, FP_ERR_STATE, "the ascending net view needs a context uploaded with the identity ordering");
  // the sort key packs (net index in the vertex's list, pin index in the net)
  // into 32 bits below the parent's position
  const int kb = bits_of((unsigned long long)ctx->max_net);
  if (kb >= 32 || (unsigned long long)ctx->max_vdeg > (1ull << (32 - kb)))
    return fail(ctx, FP_ERR_UNSUPPORTED, "vertex degree %u with nets of %u pins exceeds the BFS key", ctx->max_vdeg,
                ctx->max_net);
  FP_HIP_OK(hipSetDevice(ctx->device));
  BfsScratch s;
  const size_t nv = (size_t)nov, ne = (size_t)std::max(noe, 1);
  if (!s.get(&s.dist, nv) || !s.get(&s.plev, nv) || !s.get(&s.pos, nv) || !s.get(&s.vkey, nv) || !s.get(&s.net_owner, ne) || !s.get(&s.net_done, ne) ||
      !s.get(&s.level_nets, ne) || !s.get(&s.cand, nv) || !s.get(&s.keys, nv) || !s.get(&s.keys_out, nv) ||
      !s.get(&s.vals, nv) || !s.get(&s.order, nv) || !s.get(&s.counts, 4))
    return fail(ctx, FP_ERR_OOM, "BFS scratch allocation failed");
  if (hipcub::DeviceRadixSort::SortPairs(nullptr, s.temp_bytes, s.keys, s.keys_out, s.vals, s.order, nov, 0, 64,
                                         ctx->stream) != hipSuccess)
    return fail(ctx, FP_ERR_HIP, "BFS sort sizing failed");
  if (!s.get(reinterpret_cast<unsigned char**>(&s.temp), s.temp_bytes))
    return fail(ctx, FP_ERR_OOM, "BFS sort scratch allocation failed");
  const unsigned* pins =
      view == FP_BFS_PINS_UPLOADED ? ctx->net_pins : reinterpret_cast<const unsigned*>(ctx->apply_pins);
  hipEvent_t e0 = pooled_event(ctx, 0), e1 = pooled_event(ctx, 1);
  if (!e0 || !e1) return fail(ctx, FP_ERR_HIP, "hipEventCreate failed");
  FP_HIP_OK(hipEventRecord(e0, ctx->stream));
  int reached = 0, f1 = 0, f2 = 0, far = 0, levels = 0;
  // the first two BFS only find the farthest vertex: their levels need no order
  FP_TRY(bfs_run(ctx, s, pins, kb, start, &reached, &f1, nullptr, false));
  FP_TRY(bfs_run(ctx, s, pins, kb, f1, &reached, &f2, nullptr, false));
  FP_TRY(bfs_run(ctx, s, pins, kb, f2, &reached, &far, &levels, true));
  FP_HIP_OK(hipEventRecord(e1, ctx->stream));
  FP_TRY(d2h(ctx, ordering, s.order, (size_t)reached));
  FP_HIP_OK(hipStreamSynchronize(ctx->stream));
  // vertices the BFS does not reach (the reference leaves them -1) follow in
  // ascending order
  if (reached < nov) {
    std::vector<char> seen(nv, 0);
    for (int i = 0; i < reached; i++) seen[(size_t)ordering[i]] = 1;
    int n = reached;
    for (int v = 0; v < nov; v++)
      if (!seen[(size_t)v]) ordering[n++] = v;
  }
  if (farthest_first) *farthest_first = f1;
  if (farthest_all) *farthest_all = f2;
  if (stats) {
    float ms = 0;
    FP_HIP_OK(hipEventElapsedTime(&ms, e0, e1));
    stats->reached = reached;
    stats->levels = levels;
    stats->device_ms = ms;
  }
  return FP_OK;
}

}  // extern "C"
