// fp_legacy.cpp -- the reference's own device entry points (partition_gpu_wrapper
// and TM's two variants) implemented over the C-ABI of libfastpart.so, so that
// the reference's partition_gpu() links against this library unchanged.
//
// The reference declares one partition_gpu_wrapper per objective, with C++
// linkage and a signature of its own:
//   TV    TV/partition.cpp:6-26    (defined at TV/gpu_partition.cu:332)
//   CN    CN/partition.cpp:7-27    (defined at CN/gpu_partition.cu:263-283)
//   MSV   MGV/partition.cpp:4-25   (its .cu definition disagrees; the declaration is what links)
//   MSRV  MSRV/partition.cpp:4-24  (the TV signature)
//   TM    TM/partition.cpp:3-7     (partition_gpu_wrapper, partition_gpu_optimized_wrapper,
//                                   partition_gpu_threads_wrapper)
//   MSM   no device entry in the reference: MSM/msm.cpp:67 calls a
//         partition_gpu(...) that MSM/partition.h does not declare and no
//         file defines. This library defines it (the signature msm.cpp calls,
//         that of TM/partition.h:16), covering what TM's partition_gpu does
//         around its wrapper: the message greedy stream, the device
//         refinement, the six metrics printed in the reference's format.
// MSRV declares the TV signature for another objective, so each objective gets
// a library of its own: this file is compiled once per objective with
// -DFP_LEGACY_<TV|CN|MSV|MSRV|TM|MSM>.
//
// Contract kept from the reference:
//   * all arrays are the caller's host arrays, borrowed for the call; refined
//     results come back in the same arrays: part_vec, ep_ends, ep_parts,
//     ep_cnt (slots past a net's lambda read part -1, count 0), part_sizes and
//     the objective's bookkeeping (send_volumes; part_connections, part_messages)
//   * errors print a message and exit with the status code, as the
//     reference's gpuAssert does (TV/gpu_partition.cu:9-17)
// Choices the reference leaves open (its GPU kernels race, so no schedule of
// theirs can be reproduced): each objective runs the round schedule of its
// partition_par without the data race, with rounds of FASTPART_ROUND vertices.
// The default is the round partition_par itself uses at the process's OpenMP
// thread count (TV 8, MSV/MSRV 256 per thread; TM a whole pass), so `-g` gives
// the partition partition_par gives at that thread count. FASTPART_DEVICE
// selects the GPU (default 0).
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include <omp.h>

#include "../../include/fastpart.h"
#include "../../include/fastpart_host.h"

namespace {

void die(fp_ctx* ctx, int st, const char* what) {
  std::fprintf(stderr, "fastpart: %s failed (status %d, %s): %s\n", what, st, fp_status_string((fp_status)st),
               ctx ? fp_ctx_last_error(ctx) : "");
  std::fflush(stderr);
  if (ctx) fp_ctx_destroy(ctx);
  std::exit(st);
}

long env_long(const char* name, long dflt) {
  const char* s = std::getenv(name);
  if (!s || !*s) return dflt;
  char* end = nullptr;
  long v = std::strtol(s, &end, 10);
  return (end && *end == 0) ? v : dflt;
}

struct Legacy {
  fp_metric metric;
  const unsigned *vtx_ptr, *vtx_nets;
  const unsigned *net_ptr, *net_pins;  // null: transpose of the vertex->nets view
  int nov, noe;
  long long pins;
  const int* src;
  int K;
  const int* ordering;  // null: identity
  int *part_vec, *ep_ends, *ep_parts, *ep_cnt, *part_sizes;
  int passes;
  double C;
  long default_round;
};

// Uploads the hypergraph and the caller's state, runs the passes and copies
// the state back into the caller's arrays. Returns the live context (for the
// objective's extra outputs); the caller destroys it.
fp_ctx* refine_in_place(const Legacy& g) {
  std::vector<unsigned> net_ptr, net_pins;
  const unsigned *np = g.net_ptr, *pp = g.net_pins;
  if (!np || !pp) {
    net_ptr.resize((size_t)g.noe + 1);
    net_pins.resize((size_t)(g.pins > 0 ? g.pins : 1));
    int st = fp_csr_transpose(g.noe, g.nov, g.vtx_ptr, g.vtx_nets, net_ptr.data(), net_pins.data());
    if (st) die(nullptr, st, "transpose of the vertex->nets view");
    np = net_ptr.data();
    pp = net_pins.data();
  }
  fp_ctx* ctx = nullptr;
  int st = fp_ctx_create(&ctx, (int)env_long("FASTPART_DEVICE", 0));
  if (st) die(ctx, st, "fp_ctx_create");
  st = fp_upload(ctx, g.nov, g.noe, np, pp, g.vtx_ptr, g.vtx_nets, g.src, nullptr, g.ordering);
  if (st) die(ctx, st, "fp_upload");
  std::vector<long long> sizes((size_t)g.K);
  for (int k = 0; k < g.K; k++) sizes[(size_t)k] = g.part_sizes[k];
  st = fp_set_state(ctx, g.K, g.part_vec, g.ep_ends, g.ep_parts, g.ep_cnt, sizes.data());
  if (st) die(ctx, st, "fp_set_state");

  fp_schedule sched{};  // read by every objective but CN
  sched.kind = FP_SCHED_REF;
  sched.round = (int)env_long("FASTPART_ROUND", g.default_round);
  sched.ldg_C = g.C;
  sched.max_pw = fp_ref_max_pw(g.nov, g.K);
  sched.min_pw = fp_ref_min_pw(g.nov, g.K);
  sched.cap = 0;
  fp_refine_stats stats{};
  st = fp_refine(ctx, g.metric, g.passes, &sched, &stats);
  if (st) die(ctx, st, "fp_refine");
  st = fp_get_state(ctx, g.part_vec, g.ep_ends, g.ep_parts, g.ep_cnt, sizes.data());
  if (st) die(ctx, st, "fp_get_state");
  for (int k = 0; k < g.K; k++) g.part_sizes[k] = (int)sizes[(size_t)k];
  std::printf("fastpart: %d passes on the GPU, %lld moves, %.3f ms on the device\n", g.passes, stats.moves,
              stats.device_ms);
  return ctx;
}

// (nov / K) * 1.05: the capacity every reference directory hard-codes (e.g.
// MGV/partition.cpp:544-545), for the entry points that are not handed C
double ref_C(int nov, int K) { return fp_ref_capacity(nov, K, 0.05); }

long threads() { return (long)omp_get_max_threads(); }

}  // namespace

#if defined(FP_LEGACY_TV) || defined(FP_LEGACY_MSRV)
// TV/partition.cpp:6-26 and MSRV/partition.cpp:4-24. ep_starts (noe+1) are the
// net offsets; row_ptr/col_ind the vertex->nets view; marker, active_parts,
// scores and edge_wghts are the reference's scratch and net weights, which
// the engine does not need.
void partition_gpu_wrapper(unsigned int* row_ptr, unsigned int* col_ind, unsigned int* /*net_appearence*/, int nov,
                           int noe, long long pins, int* /*edge_wghts*/, int* src_nodes, int K, int* ordering,
                           int* /*marker*/, int* /*active_parts*/, int* part_sizes, int* /*ep_starts*/, int* ep_ends,
                           int* ep_parts, int* part_vec, int* ep_cnt, double* /*scores*/, int REFI_ITER, double C) {
#if defined(FP_LEGACY_TV)
  Legacy g{FP_METRIC_TV, row_ptr, col_ind, nullptr, nullptr, nov, noe, pins, src_nodes, K, ordering, part_vec,
           ep_ends, ep_parts, ep_cnt, part_sizes, REFI_ITER, C, 8 * threads()};
#else
  Legacy g{FP_METRIC_MSRV, row_ptr, col_ind, nullptr, nullptr, nov, noe, pins, src_nodes, K, ordering, part_vec,
           ep_ends, ep_parts, ep_cnt, part_sizes, REFI_ITER, C, 256 * threads()};
#endif
  fp_ctx_destroy(refine_in_place(g));
}
#elif defined(FP_LEGACY_CN)
// CN/partition.cpp:7-27. row_ptr_inv/col_ind_inv are the net->pins view in
// file order, which the lambda = 2 push scans.
void partition_gpu_wrapper(unsigned int* row_ptr, unsigned int* col_ind, unsigned int* /*net_appearence*/, int nov,
                           int noe, long long pins, int* /*edge_wgths*/, int* src_nodes, int K, int* ordering,
                           int /*REFI_OPT*/, int REFI_ITER, unsigned int* row_ptr_inv, unsigned int* col_ind_inv,
                           int* part_sizes, int* /*ep_starts*/, int* ep_parts, int* ep_ends, int* ep_cnt,
                           int* part_vec) {
  Legacy g{FP_METRIC_CN, row_ptr, col_ind, row_ptr_inv, col_ind_inv, nov, noe, pins, src_nodes, K, ordering, part_vec,
           ep_ends, ep_parts, ep_cnt, part_sizes, REFI_ITER, ref_C(nov, K), 1};
  fp_ctx_destroy(refine_in_place(g));
}
#elif defined(FP_LEGACY_MSV)
// MGV/partition.cpp:4-25. The refinement visits vertex ids in order
// (MGV/partition.cpp:755-757), so the ordering is not uploaded.
// send_volumes[K] comes back as the refined send volumes.
void partition_gpu_wrapper(unsigned int* row_ptr, unsigned int* col_ind, unsigned int* /*net_appearence*/, int nov,
                           int noe, long long pins, int* /*edge_wgths*/, int* src_nodes, int K, int* /*ordering*/,
                           int /*REFI_OPT*/, int REFI_ITER, int* send_volumes, double* /*scores*/, int* /*marker*/,
                           int* /*active_parts*/, int* /*ep_starts*/, int* ep_parts, int* ep_ends, int* ep_cnt,
                           int* part_sizes, int* part_vec) {
  Legacy g{FP_METRIC_MSV, row_ptr, col_ind, nullptr, nullptr, nov, noe, pins, src_nodes, K, nullptr, part_vec,
           ep_ends, ep_parts, ep_cnt, part_sizes, REFI_ITER, ref_C(nov, K), 256 * threads()};
  fp_ctx* ctx = refine_in_place(g);
  std::vector<long long> sv((size_t)K);
  fp_metrics m{};
  int st = fp_eval(ctx, &m, nullptr, nullptr, nullptr, nullptr, nullptr, sv.data());
  if (st) die(ctx, st, "fp_eval");
  for (int k = 0; k < K; k++) send_volumes[k] = (int)sv[(size_t)k];
  fp_ctx_destroy(ctx);
}
#elif defined(FP_LEGACY_TM)
namespace {
// TM/partition.cpp:3-7. The refinement visits vertex ids in order
// (TM/partition.cpp:976); part_connections[K][K] and part_messages[K] come
// back as the refined connections.
void tm_wrapper(unsigned int* row_ptr, unsigned int* col_ind, int nov, int noe, long long pins, int* src_nodes, int K,
                int REFI_ITER, int* part_vec, int* ep_parts, int* ep_ends, int* ep_cnt, int** part_connections,
                int* part_sizes, int* part_messages) {
  Legacy g{FP_METRIC_TM, row_ptr, col_ind, nullptr, nullptr, nov, noe, pins, src_nodes, K, nullptr, part_vec,
           ep_ends, ep_parts, ep_cnt, part_sizes, REFI_ITER, ref_C(nov, K), (long)nov};
  fp_ctx* ctx = refine_in_place(g);
  std::vector<int> pc((size_t)K * (size_t)K);
  int st = fp_get_connections(ctx, pc.data(), part_messages);
  if (st) die(ctx, st, "fp_get_connections");
  for (int a = 0; a < K; a++)
    for (int b = 0; b < K; b++) part_connections[a][b] = pc[(size_t)a * (size_t)K + (size_t)b];
  fp_ctx_destroy(ctx);
}
}  // namespace

#define FP_TM_ENTRY(name)                                                                                          \
  void name(unsigned int* row_ptr, unsigned int* col_ind, unsigned int* /*net_appearence*/, int nov, int noe,      \
            long long pins, int* /*edge_wgths*/, int* src_nodes, int K, int* /*ordering*/, int /*REFI_OPT*/,       \
            int REFI_ITER, int* part_vec, int* /*ep_starts*/, int* ep_parts, int* ep_ends, int* ep_cnt,            \
            int** part_connections, int* part_sizes, int* part_messages) {                                         \
    tm_wrapper(row_ptr, col_ind, nov, noe, pins, src_nodes, K, REFI_ITER, part_vec, ep_parts, ep_ends, ep_cnt,     \
               part_connections, part_sizes, part_messages);                                                      \
  }
FP_TM_ENTRY(partition_gpu_wrapper)
FP_TM_ENTRY(partition_gpu_optimized_wrapper)
FP_TM_ENTRY(partition_gpu_threads_wrapper)
#elif defined(FP_LEGACY_MSM)
// The MSM driver's device path (see the top of this file). row_ptr/col_ind:
// the vertex->nets view; src_nodes: net sources; ordering: the stream's
// visiting order. Rounds of 128 vertices per OpenMP thread, as MSM's
// partition_par (MSM/partition.cpp:1048-1056), unless FASTPART_ROUND is set.
// FASTPART_PARTVEC, when set, names a file that receives the refined part
// vector (one id per line); the reference's MSM driver writes none.
void partition_gpu(unsigned int* row_ptr, unsigned int* col_ind, unsigned int* /*net_appearence*/, int nov, int noe,
                   long long pins, int* /*edge_wgths*/, int* src_nodes, int K, int* ordering, std::string /*f_name*/,
                   int REFI_OPT, int REFI_ITER) {
  std::vector<unsigned> net_ptr((size_t)noe + 1), net_pins((size_t)(pins > 0 ? pins : 1));
  int st = fp_csr_transpose(noe, nov, row_ptr, col_ind, net_ptr.data(), net_pins.data());
  if (st) die(nullptr, st, "transpose of the vertex->nets view");
  const double C = ref_C(nov, K);
  std::vector<int> part((size_t)nov);
  auto t0 = std::chrono::steady_clock::now();
  st = fp_initial_partition_msg(nov, noe, net_ptr.data(), row_ptr, col_ind, src_nodes, ordering, nullptr, K, C,
                                part.data());
  if (st) die(nullptr, st, "fp_initial_partition_msg");
  auto t1 = std::chrono::steady_clock::now();
  std::printf("Paritioning took : %g seconds.\n", std::chrono::duration<double>(t1 - t0).count());
  fp_ctx* ctx = nullptr;
  st = fp_ctx_create(&ctx, (int)env_long("FASTPART_DEVICE", 0));
  if (st) die(ctx, st, "fp_ctx_create");
  st = fp_upload(ctx, nov, noe, net_ptr.data(), net_pins.data(), row_ptr, col_ind, src_nodes, nullptr, ordering);
  if (st) die(ctx, st, "fp_upload");
  st = fp_set_partition(ctx, K, part.data());
  if (st) die(ctx, st, "fp_set_partition");
  fp_metrics m{};
  st = fp_eval(ctx, &m, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
  if (st) die(ctx, st, "fp_eval");
  // the reference's report (MSM/partition.cpp:318-330 and :664-700)
  std::printf("TV : %lld\nCN : %lld\nTM : %lld\nMGM : %lld\nMGV : %lld\nMGAV : %lld\n", m.tv, m.cn, m.tm, m.msm,
              m.msv, m.msrv);
  if (REFI_OPT) {
    fp_schedule sched{};
    sched.kind = FP_SCHED_REF;
    sched.round = (int)env_long("FASTPART_ROUND", 128 * threads());
    sched.ldg_C = C;
    fp_refine_stats stats{};
    st = fp_refine(ctx, FP_METRIC_MSM, REFI_ITER, &sched, &stats);
    if (st) die(ctx, st, "fp_refine");
    std::printf("Refinement took : %g seconds.\n", stats.device_ms / 1000.0);
    fp_metrics r{};
    st = fp_eval(ctx, &r, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
    if (st) die(ctx, st, "fp_eval");
    std::printf("After Refinement TV : %lld\nAfter Refinement CN : %lld\nAfter Refinement TM : %lld\n"
                "After Refinement MGM : %lld\nAfter Refinement MGV : %lld\nAfter Refinement MGAV : %lld\n",
                r.tv, r.cn, r.tm, r.msm, r.msv, r.msrv);
    st = fp_get_state(ctx, part.data(), nullptr, nullptr, nullptr, nullptr);
    if (st) die(ctx, st, "fp_get_state");
  } else {
    std::printf("After Refinement TV : -1\nAfter Refinement CN : -1\nAfter Refinement TM : -1\n"
                "After Refinement MGM : -1\nAfter Refinement MGV : -1\nAfter Refinement MGAV : -1\n");
  }
  std::printf("Iterated for : %d times\n", REFI_ITER);
  std::fflush(stdout);
  const char* pv = std::getenv("FASTPART_PARTVEC");
  if (pv && *pv) {
    st = fp_write_part_vector(pv, nov, part.data());
    if (st) die(ctx, st, "fp_write_part_vector");
  }
  fp_ctx_destroy(ctx);
}
#else
#error "compile with -DFP_LEGACY_TV, -DFP_LEGACY_CN, -DFP_LEGACY_MSV, -DFP_LEGACY_MSRV, -DFP_LEGACY_TM or -DFP_LEGACY_MSM"
#endif
