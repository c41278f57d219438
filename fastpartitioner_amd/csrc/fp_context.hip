// fp_context.hip -- context lifetime, hypergraph upload and connectivity state
// transfer of the HIP engine (include/fastpart.h).
#include "fp_internal.h"

namespace fpe {

__global__ void __launch_bounds__(1024) k_pass_begin(unsigned* __restrict__ stamp, size_t n, int* __restrict__ counters,
                                                     const int* __restrict__ val, int K, int* __restrict__ mstat) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    stamp[i] = 0u;
  if (blockIdx.x != 0) return;
  if (threadIdx.x < 4) counters[threadIdx.x] = 0;
  if (!val) return;
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
  const int w = (int)(threadIdx.x / WAVE);
  if (lane_id() == 0) {
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

void pass_begin(fp_ctx* ctx, const int* val, int K, int* mstat) {
  const size_t n = (size_t)std::max(ctx->noe, 1);
  const int blocks = (int)std::min<size_t>((n + 1023) / 1024, 1024);
  k_pass_begin<<<blocks, 1024, 0, ctx->stream>>>(ctx->net_stamp, n, ctx->counters, val, K, mstat);
}


fp_status fail(fp_ctx* c, fp_status st, const char* fmt, ...) {
  if (c) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(c->err, sizeof c->err, fmt, ap);
    va_end(ap);
  }
  return st;
}

fp_status dalloc_bytes(fp_ctx* ctx, void** p, size_t bytes) {
  *p = nullptr;
  void* q = nullptr;
  hipError_t e = hipMalloc(&q, bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return fail(ctx, FP_ERR_OOM, "hipMalloc(%zu bytes) failed: %s", bytes, hipGetErrorString(e));
  }
  ctx->owned.push_back(q);
  ctx->dev_bytes += (long long)bytes;
  ctx->gen++;
  *p = q;
  return FP_OK;
}

void dfree_all(fp_ctx* ctx) {
  for (void* p : ctx->owned) (void)hipFree(p);
  ctx->owned.clear();
  ctx->dev_bytes = 0;
  ctx->gen++;
}

hipEvent_t pooled_event(fp_ctx* ctx, size_t i) {
  while (ctx->ev_pool.size() <= i) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    ctx->ev_pool.push_back(e);
  }
  return ctx->ev_pool[i];
}

namespace {

int bits_for(int K) {
  int b = 1;
  while ((1 << b) < K) b++;
  return b;
}

}  // namespace

namespace {

constexpr size_t STAGE_BYTES = size_t{16} << 20;  // per staging buffer
constexpr size_t STAGE_MIN = size_t{16} << 20;    // smaller copies go straight from pageable memory
constexpr size_t STAGE_PIECE = size_t{1} << 20;   // host threads fill a staging buffer in pieces of this

bool host_pinned(const void* h) {
  hipPointerAttribute_t a;
  if (hipPointerGetAttributes(&a, h) != hipSuccess) {
    (void)hipGetLastError();  // pageable memory the runtime does not know
    return false;
  }
  return a.type == hipMemoryTypeHost;
}

void free_stage(fp_ctx* ctx) {
  for (int b = 0; b < 2; b++) {
    if (ctx->stage_ev[b]) (void)hipEventDestroy(ctx->stage_ev[b]);
    if (ctx->stage[b]) (void)hipHostFree(ctx->stage[b]);
    ctx->stage_ev[b] = nullptr;
    ctx->stage[b] = nullptr;
  }
}

// the staging buffers, or false (then copies go straight from pageable memory)
bool ensure_stage(fp_ctx* ctx) {
  if (ctx->stage_ev[1]) return true;
  for (int b = 0; b < 2; b++) {
    void* p = nullptr;
    hipEvent_t e = nullptr;
    const bool ok = hipHostMalloc(&p, STAGE_BYTES, hipHostMallocDefault) == hipSuccess;
    if (ok) ctx->stage[b] = p;
    if (!ok || hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
      (void)hipGetLastError();
      free_stage(ctx);
      return false;
    }
    ctx->stage_ev[b] = e;
    if (hipEventRecord(e, ctx->stream) != hipSuccess) {
      (void)hipGetLastError();
      free_stage(ctx);
      return false;
    }
  }
  return true;
}

}  // namespace

fp_status h2d_bulk(fp_ctx* ctx, void* d, const void* h, size_t bytes) {
  if (bytes == 0) return FP_OK;
  if (bytes < STAGE_MIN || host_pinned(h) || !ensure_stage(ctx)) {
    FP_HIP_OK(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, ctx->stream));
    return FP_OK;
  }
  for (size_t off = 0, i = 0; off < bytes; off += STAGE_BYTES, i++) {
    const size_t n = std::min(STAGE_BYTES, bytes - off);
    const int b = (int)(i & 1);
    // the copy out of this buffer two pieces ago has finished
    FP_HIP_OK(hipEventSynchronize(ctx->stage_ev[b]));
    char* dst = static_cast<char*>(ctx->stage[b]);
    const char* src = static_cast<const char*>(h) + off;
    const long long pieces = (long long)((n + STAGE_PIECE - 1) / STAGE_PIECE);
#pragma omp parallel for schedule(static)
    for (long long q = 0; q < pieces; q++) {
      const size_t o = (size_t)q * STAGE_PIECE;
      std::memcpy(dst + o, src + o, std::min(STAGE_PIECE, n - o));
    }
    FP_HIP_OK(hipMemcpyAsync(static_cast<char*>(d) + off, dst, n, hipMemcpyHostToDevice, ctx->stream));
    FP_HIP_OK(hipEventRecord(ctx->stage_ev[b], ctx->stream));
  }
  return FP_OK;
}

// K-sized arrays and the packed-entry layout for K parts
fp_status ensure_k(fp_ctx* ctx, int K) {
  if (K < 1) return fail(ctx, FP_ERR_ARG, "K must be >= 1");
  if (K > 16384) return fail(ctx, FP_ERR_UNSUPPORTED, "K = %d above 16384", K);  // parts fit the 15-bit view
  const int pb = bits_for(K);
  if ((unsigned long long)ctx->max_net >= (1ull << (32 - pb)))
    return fail(ctx, FP_ERR_UNSUPPORTED, "a net of %u pins is too large for K = %d in the packed list layout",
                ctx->max_net, K);
  if (K > ctx->kcap) {
    // K-sized arrays grow only; a smaller K reuses them
    FP_TRY(dalloc(ctx, &ctx->sizes, (size_t)K));
    FP_TRY(dalloc(ctx, &ctx->sizes_frozen, (size_t)K));
    FP_TRY(dalloc(ctx, &ctx->cand_tot, (size_t)K));
    FP_TRY(dalloc(ctx, &ctx->over_idx, (size_t)K));
    ctx->kcap = K;
  }
  if (ctx->K != K) {
    // the BAL histogram is all zero between rounds whatever its shape; the
    // candidate totals restart from zero
    fill_async<long long>(ctx, ctx->cand_tot, (size_t)K, 0);
    ctx->K = K;
    ctx->gen++;
  }
  if (ctx->pb != pb) ctx->gen++;
  ctx->pb = pb;
  ctx->pmask = (1u << pb) - 1u;
  return FP_OK;
}

namespace {

constexpr int BUILD_KMAX = 4096;  // k_build_lists keeps four K-int arrays per wavefront in LDS

// Per incidence of the vertex view: the segment of its net in the packed
// connectivity layout (the net's list starts at net_ptr[n] + n).
__global__ void __launch_bounds__(256) k_vtx_seg(const unsigned* __restrict__ net_ptr,
                                                 const unsigned* __restrict__ vtx_nets, size_t pins,
                                                 unsigned* __restrict__ seg) {
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < pins; k += (size_t)gridDim.x * blockDim.x) {
    const unsigned n = vtx_nets[k];
    seg[k] = net_ptr[n] + n;
  }
}

// One wavefront per net: the net's connectivity list from the part vector,
// in first-appearance order along the ordering (the state the reference's
// streaming initial partition leaves, TV/partition.cpp:240-263). The pins are
// visited in rank order (apply_pins holds each net's pin ranks ascending), 64
// at a time: the first lane of each part not yet listed appends it, then every
// lane counts its pin. Writes the packed segment and its evaluation view.
__global__ void __launch_bounds__(256) k_build_lists(const unsigned* __restrict__ net_ptr,
                                                     const int* __restrict__ apply_pins, const int* __restrict__ order,
                                                     const int* __restrict__ part, int noe, int K, int pb,
                                                     unsigned pmask, unsigned* __restrict__ ent,
                                                     unsigned short* __restrict__ ev) {
  extern __shared__ __align__(16) unsigned char lds[];
  const int lane = lane_id(), wave = wave_id(), nw = (int)(blockDim.x / WAVE);
  int* slot = reinterpret_cast<int*>(lds) + (size_t)wave * 4 * (size_t)K;  // list position of a part, -1
  int* claim = slot + K;                                                   // first lane of a new part, 64
  int* lcnt = claim + K;                                                   // pin count per list position
  int* lq = lcnt + K;                                                      // part per list position
  for (int q = lane; q < K; q += WAVE) {
    slot[q] = -1;
    claim[q] = WAVE;
  }
  wave_lds_sync();
  const unsigned long long lt = (1ull << lane) - 1ull;
  for (int n = (int)blockIdx.x * nw + wave; n < noe; n += (int)gridDim.x * nw) {
    const unsigned s = net_ptr[n], e = net_ptr[n + 1];
    int L = 0;
    for (unsigned k0 = s; k0 < e; k0 += WAVE) {
      const unsigned k = k0 + (unsigned)lane;
      const bool ok = k < e;
      int q = 0;
      if (ok) {
        const int r = apply_pins[k];
        q = part[order ? order[r] : r];
      }
      const bool fresh = ok && slot[q] < 0;
      if (fresh) atomicMin(&claim[q], lane);
      wave_lds_sync();
      const bool first = fresh && claim[q] == lane;
      const unsigned long long bm = __ballot(first);
      if (first) {
        const int pos = L + __popcll(bm & lt);
        slot[q] = pos;
        lcnt[pos] = 0;
        lq[pos] = q;
        claim[q] = WAVE;
      }
      L += __popcll(bm);
      wave_lds_sync();
      if (ok) atomicAdd(&lcnt[slot[q]], 1);
      wave_lds_sync();
    }
    const unsigned seg = s + (unsigned)n;
    for (int t = lane; t < L; t += WAVE) {
      const unsigned x = ((unsigned)lcnt[t] << pb) | (unsigned)lq[t];
      ent[seg + 1 + (unsigned)t] = x;
      ev[seg + 1 + (unsigned)t] = ev_word(x, pb, pmask);
      slot[lq[t]] = -1;
    }
    if (lane == 0) {
      ent[seg] = (unsigned)L;
      ev[seg] = (unsigned short)L;
    }
    wave_lds_sync();
  }
}

// ent: the segment array (lambda + list per net), host copy of the device layout
fp_status upload_state(fp_ctx* ctx, const int* part, const std::vector<unsigned>& ent, const long long* sizes_in) {
  std::vector<long long> sz((size_t)ctx->K, 0);
  if (sizes_in) {
    std::copy(sizes_in, sizes_in + ctx->K, sz.begin());
  } else {
    for (int v = 0; v < ctx->nov; v++) sz[(size_t)part[v]] += ctx->h_vwgt.empty() ? 1 : ctx->h_vwgt[(size_t)v];
  }
  FP_TRY(h2d(ctx, ctx->part, part, (size_t)ctx->nov));
  FP_TRY(h2d(ctx, ctx->ent, ent.data(), ent.size()));
  // evaluation view of every segment
  std::vector<unsigned short> ev(ent.size(), 0);
  for (int n = 0; n < ctx->noe; n++) {
    const size_t seg = (size_t)ctx->h_net_ptr[(size_t)n] + (size_t)n;
    const unsigned L = ent[seg];
    ev[seg] = (unsigned short)L;
    for (unsigned t = 0; t < L; t++) ev[seg + 1 + t] = ev_word(ent[seg + 1 + t], ctx->pb, ctx->pmask);
  }
  FP_TRY(h2d(ctx, ctx->ev, ev.data(), ev.size()));
  FP_TRY(h2d(ctx, ctx->sizes, sz.data(), (size_t)ctx->K));
  FP_HIP_OK(hipStreamSynchronize(ctx->stream));
  ctx->has_state = true;
  return FP_OK;
}

}  // namespace
}  // namespace fpe

using namespace fpe;

extern "C" {

const char* fp_status_string(fp_status st) {
  switch (st) {
    case FP_OK: return "ok";
    case FP_ERR_ARG: return "invalid argument";
    case FP_ERR_HIP: return "HIP runtime error";
    case FP_ERR_OOM: return "device allocation failed";
    case FP_ERR_STATE: return "call out of order";
    case FP_ERR_UNSUPPORTED: return "input outside the supported envelope";
    case FP_ERR_BUDGET: return "rejected by the HBM footprint budget";
    case FP_ERR_IO: return "file I/O error";
    case FP_ERR_NODEV: return "no usable GPU";
  }
  return "unknown status";
}

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
  free_stage(ctx);
  for (hipEvent_t e : ctx->ev_pool) (void)hipEventDestroy(e);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

const char* fp_ctx_last_error(const fp_ctx* ctx) { return ctx ? ctx->err : "null context"; }

void* fp_host_alloc_pinned(size_t bytes) {
  void* p = nullptr;
  if (hipHostMalloc(&p, bytes ? bytes : 1, hipHostMallocDefault) != hipSuccess) {
    (void)hipGetLastError();
    return nullptr;
  }
  return p;
}

void fp_host_free_pinned(void* p) {
  if (p) (void)hipHostFree(p);
}

long long fp_ctx_device_bytes(const fp_ctx* ctx) { return ctx ? ctx->dev_bytes : 0; }

fp_status fp_upload(fp_ctx* ctx, int nov, int noe, const unsigned* net_ptr, const unsigned* net_pins,
                    const unsigned* vtx_ptr, const unsigned* vtx_nets, const int* src, const int* vwgt,
                    const int* ordering) {
  if (!ctx) return FP_ERR_ARG;
  if (nov < 1 || noe < 0 || !net_ptr) return fail(ctx, FP_ERR_ARG, "fp_upload: bad sizes or null arrays");
  const long long pins = net_ptr[noe];
  if (!net_pins && pins > 0) return fail(ctx, FP_ERR_ARG, "fp_upload: null net_pins");
  FP_HIP_OK(hipSetDevice(ctx->device));
  FP_HIP_OK(hipStreamSynchronize(ctx->stream));
  dfree_all(ctx);
  ctx->has_graph = ctx->has_state = false;
  ctx->K = 0;
  ctx->kcap = 0;
  ctx->rmax = 0;
  ctx->hist_cap = 0;
  ctx->compact_cap = 0;
  ctx->cn.ready = false;
  ctx->cn.ops_cap = 0;
  ctx->vol.rmax = 0;
  ctx->vol.kcap = 0;
  ctx->vol.evcap = 0;
  ctx->conn.kcap = 0;
  ctx->conn.rmax = 0;
  ctx->conn.evcap = 0;
  ctx->conn.incoff_ready = false;
  ctx->lam = nullptr;
  ctx->met_scratch = nullptr;
  ctx->met_words = 0;
  ctx->met_out = nullptr;
  ctx->met_out_cap = 0;
  if (net_ptr[0] != 0) return fail(ctx, FP_ERR_ARG, "fp_upload: net_ptr[0] != 0");
  unsigned max_net = 0;
  for (int n = 0; n < noe; n++) {
    if (net_ptr[n + 1] < net_ptr[n]) return fail(ctx, FP_ERR_ARG, "fp_upload: net_ptr not monotone at %d", n);
    max_net = std::max(max_net, net_ptr[n + 1] - net_ptr[n]);
  }
  {
    // the first pin out of range, on every host thread
    long long bad = LLONG_MAX;
#pragma omp parallel for schedule(static) reduction(min : bad)
    for (long long k = 0; k < pins; k++)
      if (net_pins[k] >= (unsigned)nov) bad = std::min(bad, k);
    if (bad != LLONG_MAX) return fail(ctx, FP_ERR_ARG, "fp_upload: pin %lld out of range", bad);
  }
  if (src)
    for (int n = 0; n < noe; n++)
      if (src[n] < -1 || src[n] >= nov) return fail(ctx, FP_ERR_ARG, "fp_upload: src[%d] out of range", n);
  long long tw = 0;
  ctx->h_vwgt.clear();
  if (vwgt) {
    for (int v = 0; v < nov; v++) {
      if (vwgt[v] < 0) return fail(ctx, FP_ERR_ARG, "fp_upload: negative weight at %d", v);
      tw += vwgt[v];
    }
    ctx->h_vwgt.assign(vwgt, vwgt + nov);
  } else {
    tw = nov;
  }
  std::vector<int> rank((size_t)nov);
  ctx->h_order.resize((size_t)nov);
  bool ident = true;
  if (ordering) {
    std::vector<char> seen((size_t)nov, 0);
    for (int z = 0; z < nov; z++) {
      const int v = ordering[z];
      if (v < 0 || v >= nov || seen[(size_t)v]) return fail(ctx, FP_ERR_ARG, "fp_upload: ordering is not a permutation");
      seen[(size_t)v] = 1;
      rank[(size_t)v] = z;
      ctx->h_order[(size_t)z] = v;
      ident = ident && v == z;
    }
  } else {
    std::iota(rank.begin(), rank.end(), 0);
    std::iota(ctx->h_order.begin(), ctx->h_order.end(), 0);
  }
  ctx->nov = nov;
  ctx->noe = noe;
  ctx->pins = pins;
  ctx->total_weight = tw;
  ctx->max_net = max_net;
  ctx->h_net_ptr.assign(net_ptr, net_ptr + noe + 1);
  // vertex->nets view: the caller's, or a transposition on the host
  std::vector<unsigned> t_vptr, t_vnets;
  const unsigned* h_vptr = vtx_ptr;
  const unsigned* h_vnets = vtx_nets;
  if (vtx_ptr && vtx_nets) {
    if (vtx_ptr[0] != 0 || (long long)vtx_ptr[nov] != pins) return fail(ctx, FP_ERR_ARG, "fp_upload: vtx_ptr inconsistent");
    for (int v = 0; v < nov; v++)
      if (vtx_ptr[v + 1] < vtx_ptr[v]) return fail(ctx, FP_ERR_ARG, "fp_upload: vtx_ptr not monotone");
    bool bad = false;
#pragma omp parallel for schedule(static) reduction(|| : bad)
    for (long long k = 0; k < pins; k++) bad = bad || vtx_nets[k] >= (unsigned)noe;
    if (bad) return fail(ctx, FP_ERR_ARG, "fp_upload: vtx_nets out of range");
  } else {
    // counting-sort transpose: ascending nets per vertex (TV/tv.cpp:79-100)
    t_vptr.assign((size_t)nov + 1, 0);
    t_vnets.assign((size_t)pins, 0);
    for (long long k = 0; k < pins; k++) t_vptr[net_pins[k] + 1]++;
    for (int v = 0; v < nov; v++) t_vptr[(size_t)v + 1] += t_vptr[(size_t)v];
    std::vector<unsigned> fill(t_vptr.begin(), t_vptr.end() - 1);
    for (int n = 0; n < noe; n++)
      for (unsigned k = net_ptr[n]; k < net_ptr[n + 1]; k++) t_vnets[fill[net_pins[k]]++] = (unsigned)n;
    h_vptr = t_vptr.data();
    h_vnets = t_vnets.data();
  }
  unsigned max_vdeg = 0;
  unsigned long long sum_deg2 = 0;
#pragma omp parallel for schedule(static) reduction(max : max_vdeg) reduction(+ : sum_deg2)
  for (int v = 0; v < nov; v++) {
    const unsigned d = h_vptr[v + 1] - h_vptr[v];
    max_vdeg = std::max(max_vdeg, d);
    sum_deg2 += (unsigned long long)d * d;
  }
  ctx->max_vdeg = max_vdeg;
  ctx->sum_deg2 = sum_deg2;
  // per net: ordering ranks of its pins, ascending (the apply replays moves in this order)
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
  // the volume objectives read a net's source part through its pins
  // (MGV/partition.cpp:158-186): every non-empty net's source must be a pin
  bool src_ok = src != nullptr;
  if (src_ok) {
#pragma omp parallel for schedule(dynamic, 1024) reduction(&& : src_ok)
    for (int n = 0; n < noe; n++) {
      if (net_ptr[n + 1] == net_ptr[n]) continue;
      bool hit = false;
      for (unsigned k = net_ptr[n]; k < net_ptr[n + 1] && !hit; k++) hit = (int)net_pins[k] == src[n];
      src_ok = src_ok && hit;
    }
  }
  ctx->src_on_pins = src_ok;
  ctx->max_src_pins = 0;
  if (src) {
    std::vector<unsigned long long> sp((size_t)nov, 0);
#pragma omp parallel for schedule(static)
    for (int n = 0; n < noe; n++)
      if (src[n] >= 0) {
#pragma omp atomic
        sp[(size_t)src[n]] += net_ptr[n + 1] - net_ptr[n];
      }
    unsigned long long m = 0;
#pragma omp parallel for schedule(static) reduction(max : m)
    for (int v = 0; v < nov; v++) m = std::max(m, sp[(size_t)v]);
    ctx->max_src_pins = m;
  }

  FP_TRY(dalloc(ctx, &ctx->net_ptr, (size_t)noe + 1));
  FP_TRY(dalloc(ctx, &ctx->net_pins, (size_t)pins));
  FP_TRY(dalloc(ctx, &ctx->vtx_ptr, (size_t)nov + 1));
  FP_TRY(dalloc(ctx, &ctx->vtx_nets, (size_t)pins));
  FP_TRY(dalloc(ctx, &ctx->apply_pins, (size_t)pins));
  FP_TRY(dalloc(ctx, &ctx->src, (size_t)std::max(noe, 1)));
  FP_TRY(h2d(ctx, ctx->net_ptr, net_ptr, (size_t)noe + 1));
  FP_TRY(h2d(ctx, ctx->net_pins, net_pins, (size_t)pins));
  FP_TRY(h2d(ctx, ctx->vtx_ptr, h_vptr, (size_t)nov + 1));
  FP_TRY(h2d(ctx, ctx->vtx_nets, h_vnets, (size_t)pins));
  FP_TRY(h2d(ctx, ctx->apply_pins, ctx->h_rank_pins.data(), (size_t)pins));
  if (src) {
    FP_TRY(h2d(ctx, ctx->src, src, (size_t)noe));
  } else {
    fill_async<int>(ctx, ctx->src, (size_t)std::max(noe, 1), -1);
  }
  ctx->vwgt = nullptr;
  if (vwgt) {
    FP_TRY(dalloc(ctx, &ctx->vwgt, (size_t)nov));
    FP_TRY(h2d(ctx, ctx->vwgt, vwgt, (size_t)nov));
  }
  ctx->order = nullptr;
  if (!ident) {
    FP_TRY(dalloc(ctx, &ctx->order, (size_t)nov));
    FP_TRY(h2d(ctx, ctx->order, ordering, (size_t)nov));
  }
  // state and scratch that do not depend on K
  FP_TRY(dalloc(ctx, &ctx->part, (size_t)nov));
  FP_TRY(dalloc(ctx, &ctx->ent, (size_t)pins + (size_t)noe + ENT_PAD));
  FP_TRY(dalloc(ctx, &ctx->ev, (size_t)pins + (size_t)noe + ENT_PAD));
  fill_async<unsigned>(ctx, ctx->ent, (size_t)pins + (size_t)noe + ENT_PAD, 0u);
  fill_async<unsigned short>(ctx, ctx->ev, (size_t)pins + (size_t)noe + ENT_PAD, (unsigned short)0);
  FP_TRY(dalloc(ctx, &ctx->vtx_seg, (size_t)pins));
  if (pins > 0)
    k_vtx_seg<<<(int)std::min<long long>((pins + 255) / 256, 8192), 256, 0, ctx->stream>>>(
        ctx->net_ptr, ctx->vtx_nets, (size_t)pins, ctx->vtx_seg);
  FP_TRY(dalloc(ctx, &ctx->net_stamp, (size_t)std::max(noe, 1)));
  FP_TRY(dalloc(ctx, &ctx->touched, (size_t)std::max(noe, 1)));
  FP_TRY(dalloc(ctx, &ctx->counters, 4));
  FP_TRY(dalloc(ctx, &ctx->flags, 4));
  fill_async<int>(ctx, ctx->flags, 4, 0);
  FP_TRY(dalloc(ctx, &ctx->moves, 1));
  FP_TRY(dalloc(ctx, &ctx->over_count, 1));
  fill_async<unsigned>(ctx, ctx->net_stamp, (size_t)std::max(noe, 1), 0u);
  FP_HIP_OK(hipGetLastError());
  FP_HIP_OK(hipStreamSynchronize(ctx->stream));
  ctx->has_graph = true;
  return FP_OK;
}

fp_status fp_set_partition(fp_ctx* ctx, int K, const int* part) {
  if (!ctx) return FP_ERR_ARG;
  if (!ctx->has_graph) return fail(ctx, FP_ERR_STATE, "fp_set_partition before fp_upload");
  if (!part) return fail(ctx, FP_ERR_ARG, "null part");
  FP_HIP_OK(hipSetDevice(ctx->device));
  for (int v = 0; v < ctx->nov; v++)
    if (part[v] < 0 || part[v] >= K) return fail(ctx, FP_ERR_ARG, "part[%d] = %d out of [0, K)", v, part[v]);
  FP_TRY(ensure_k(ctx, K));
  const int noe = ctx->noe;
  if (K <= BUILD_KMAX) {
    // built on the device: only the part vector and the part sizes travel
    std::vector<long long> sz((size_t)K, 0);
    for (int v = 0; v < ctx->nov; v++) sz[(size_t)part[v]] += ctx->h_vwgt.empty() ? 1 : ctx->h_vwgt[(size_t)v];
    FP_TRY(h2d(ctx, ctx->part, part, (size_t)ctx->nov));
    FP_TRY(h2d(ctx, ctx->sizes, sz.data(), (size_t)K));
    int waves = 4;
    while (waves > 1 && (size_t)waves * 16 * (size_t)K > (size_t)LDS_BUDGET) waves--;
    const int grid = std::max(1, std::min((noe + waves - 1) / waves, 8192));
    k_build_lists<<<grid, waves * WAVE, (size_t)waves * 16 * (size_t)K, ctx->stream>>>(
        ctx->net_ptr, ctx->apply_pins, ctx->order, ctx->part, noe, K, ctx->pb, ctx->pmask, ctx->ent, ctx->ev);
    FP_HIP_OK(hipGetLastError());
    FP_HIP_OK(hipStreamSynchronize(ctx->stream));
    ctx->has_state = true;
    return FP_OK;
  }
  std::vector<unsigned> ent((size_t)ctx->pins + (size_t)noe + ENT_PAD, 0);
  const unsigned inc = 1u << ctx->pb;
  // list order = first appearance of each part along the ordering (the state
  // the streaming initial partition leaves, TV/partition.cpp:240-263): walk
  // each net's pins in rank order and append unseen parts
#pragma omp parallel
  {
    std::vector<int> slot((size_t)K, -1), owner((size_t)K, -1);
#pragma omp for schedule(dynamic, 1024)
    for (int n = 0; n < noe; n++) {
      const unsigned s = ctx->h_net_ptr[(size_t)n], e = ctx->h_net_ptr[(size_t)n + 1];
      unsigned* list = ent.data() + s + (unsigned)n + 1;
      int L = 0;
      for (unsigned k = s; k < e; k++) {
        const int q = part[ctx->h_order[(size_t)ctx->h_rank_pins[k]]];
        if (owner[(size_t)q] != n) {
          owner[(size_t)q] = n;
          slot[(size_t)q] = L;
          list[L] = inc | (unsigned)q;
          L++;
        } else {
          list[slot[(size_t)q]] += inc;
        }
      }
      list[-1] = (unsigned)L;
    }
  }
  return upload_state(ctx, part, ent, nullptr);
}

fp_status fp_set_state(fp_ctx* ctx, int K, const int* part, const int* ep_ends, const int* ep_parts,
                       const int* ep_cnt, const long long* sizes) {
  if (!ctx) return FP_ERR_ARG;
  if (!ctx->has_graph) return fail(ctx, FP_ERR_STATE, "fp_set_state before fp_upload");
  if (!part || !ep_ends || !ep_parts || !ep_cnt) return fail(ctx, FP_ERR_ARG, "null state array");
  FP_HIP_OK(hipSetDevice(ctx->device));
  for (int v = 0; v < ctx->nov; v++)
    if (part[v] < 0 || part[v] >= K) return fail(ctx, FP_ERR_ARG, "part[%d] = %d out of [0, K)", v, part[v]);
  FP_TRY(ensure_k(ctx, K));
  std::vector<unsigned> ent((size_t)ctx->pins + (size_t)ctx->noe + ENT_PAD, 0);
  for (int n = 0; n < ctx->noe; n++) {
    const unsigned s = ctx->h_net_ptr[(size_t)n], len = ctx->h_net_ptr[(size_t)n + 1] - s;
    const int L = ep_ends[n];
    if (L < 0 || (unsigned)L > len) return fail(ctx, FP_ERR_ARG, "ep_ends[%d] out of range", n);
    unsigned* seg = ent.data() + s + (unsigned)n;
    seg[0] = (unsigned)L;
    for (int j = 0; j < L; j++) {
      const int q = ep_parts[s + (unsigned)j], c = ep_cnt[s + (unsigned)j];
      if (q < 0 || q >= K || c < 1 || (unsigned)c > len)
        return fail(ctx, FP_ERR_ARG, "list entry %d of net %d invalid", j, n);
      seg[1 + j] = ((unsigned)c << ctx->pb) | (unsigned)q;
    }
  }
  return upload_state(ctx, part, ent, sizes);
}

fp_status fp_get_state(fp_ctx* ctx, int* part, int* ep_ends, int* ep_parts, int* ep_cnt, long long* sizes) {
  if (!ctx) return FP_ERR_ARG;
  if (!ctx->has_state) return fail(ctx, FP_ERR_STATE, "no state");
  FP_HIP_OK(hipSetDevice(ctx->device));
  // the lists travel only when asked for (a caller after the part vector
  // alone, as the server's requests, skips their download and unpacking)
  const bool lists = ep_ends || ep_parts || ep_cnt;
  std::vector<unsigned> ent(lists ? (size_t)ctx->pins + (size_t)ctx->noe : 0);
  if (lists) FP_TRY(d2h(ctx, ent.data(), ctx->ent, ent.size()));
  if (part) FP_TRY(d2h(ctx, part, ctx->part, (size_t)ctx->nov));
  if (sizes) FP_TRY(d2h(ctx, sizes, ctx->sizes, (size_t)ctx->K));
  FP_HIP_OK(hipStreamSynchronize(ctx->stream));
  if (!lists) return FP_OK;
#pragma omp parallel for schedule(dynamic, 4096)
  for (int n = 0; n < ctx->noe; n++) {
    const unsigned s = ctx->h_net_ptr[(size_t)n], e = ctx->h_net_ptr[(size_t)n + 1];
    const unsigned* seg = ent.data() + s + (unsigned)n;
    const int L = (int)seg[0];
    if (ep_ends) ep_ends[n] = L;
    for (unsigned k = s; k < e; k++) {
      const int j = (int)(k - s);
      const bool live = j < L;
      if (ep_parts) ep_parts[k] = live ? (int)(seg[1 + j] & ctx->pmask) : -1;
      if (ep_cnt) ep_cnt[k] = live ? (int)(seg[1 + j] >> ctx->pb) : 0;
    }
  }
  return FP_OK;
}

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

}  // extern "C"
