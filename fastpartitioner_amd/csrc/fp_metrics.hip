// fp_metrics.hip -- the six metrics of shared/metric.cpp from scratch on the
// device: one wavefront per net sweeps its pins (coalesced), deduplicates the
// pins' parts through an LDS mark array, and accumulates lambda, TV, CN and
// the per-part int64 send / receive volume and message vectors (block-private
// in LDS up to K = 1024, straight to HBM above).
#include "fp_internal.h"

namespace fpe {
namespace {

constexpr int METRIC_BLOCK = 256;

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

constexpr int METRIC_LDS_KMAX = 1024;
#ifndef FP_METRIC_GRID
#define FP_METRIC_GRID 4096
#endif
constexpr int METRIC_GRID = FP_METRIC_GRID;  // blocks at most: each flushes its per-part sums once

template <bool LDS_ACC>
__global__ void __launch_bounds__(METRIC_BLOCK) k_metrics(MetricArgs a) {
  extern __shared__ __align__(16) unsigned char lds[];
  const int K = a.K;
  unsigned long long* s_send = LDS_ACC ? reinterpret_cast<unsigned long long*>(lds) : a.send;
  unsigned long long* s_recv = LDS_ACC ? s_send + K : a.recv;
  unsigned long long* s_msv = LDS_ACC ? s_recv + K : a.send_msv;
  int* marks = LDS_ACC ? reinterpret_cast<int*>(s_msv + K) : reinterpret_cast<int*>(lds);
  const int lane = lane_id(), wave = wave_id(), nw = (int)(blockDim.x / WAVE);
  int* mark = marks + (size_t)wave * (size_t)K;
  if (LDS_ACC)
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
        // the pair is usually known already: a plain read skips the atomic
        // (a stale 0 only costs the atomic, whose result decides)
        const unsigned m = 1u << (bit & 31);
        if (!(__builtin_nontemporal_load(&a.p2p[bit >> 5]) & m) && !(atomicOr(&a.p2p[bit >> 5], m) & m)) {
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
  if (LDS_ACC) {
    __syncthreads();
    for (int q = (int)threadIdx.x; q < K; q += (int)blockDim.x) {
      if (s_send[q]) atomicAdd(&a.send[q], s_send[q]);
      if (s_recv[q]) atomicAdd(&a.recv[q], s_recv[q]);
      if (s_msv[q]) atomicAdd(&a.send_msv[q], s_msv[q]);
    }
  }
  // TV and CN: summed over the block, then one atomic each
  __shared__ unsigned long long s_tv, s_cn;
  if (threadIdx.x == 0) {
    s_tv = 0;
    s_cn = 0;
  }
  __syncthreads();
  if (lane == 0) {
    if (tv) atomicAdd(&s_tv, tv);
    if (cn) atomicAdd(&s_cn, cn);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    if (s_tv) atomicAdd(&a.tvcn[0], s_tv);
    if (s_cn) atomicAdd(&a.tvcn[1], s_cn);
  }
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
  const bool lds_acc = K <= METRIC_LDS_KMAX;
  // mark arrays (K ints per wavefront) plus, up to METRIC_LDS_KMAX, three
  // block-private int64 vectors
  const size_t acc_bytes = lds_acc ? (size_t)3 * K * sizeof(unsigned long long) : 0;
  int waves = METRIC_BLOCK / WAVE;
  while (waves > 0 && acc_bytes + (size_t)waves * K * sizeof(int) > (size_t)LDS_BUDGET) waves--;
  if (waves == 0) return fail(ctx, FP_ERR_UNSUPPORTED, "K = %d too large for the metric kernel", K);
  if (!ctx->lam) FP_TRY(dalloc(ctx, &ctx->lam, (size_t)std::max(ctx->noe, 1)));
  int* d_lam = ctx->lam;
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
  const size_t lds = acc_bytes + (size_t)waves * K * sizeof(int);
  const int grid = std::max(1, std::min((ctx->noe + waves - 1) / waves, METRIC_GRID));
  if (ctx->noe > 0) {
    if (lds_acc) k_metrics<true><<<grid, waves * WAVE, lds, ctx->stream>>>(m);
    else k_metrics<false><<<grid, waves * WAVE, lds, ctx->stream>>>(m);
  }
  std::vector<unsigned long long> h((size_t)5 * K + 2);
  fp_status st = FP_OK;
  if (hipGetLastError() != hipSuccess) st = fail(ctx, FP_ERR_HIP, "k_metrics launch failed");
  if (st == FP_OK) st = d2h(ctx, h.data(), ctx->mvec, h.size());
  if (st == FP_OK && lam) st = d2h(ctx, lam, d_lam, (size_t)ctx->noe);
  if (hipStreamSynchronize(ctx->stream) != hipSuccess && st == FP_OK) st = fail(ctx, FP_ERR_HIP, "sync failed");
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
