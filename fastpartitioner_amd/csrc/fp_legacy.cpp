// fp_legacy.cpp -- the reference's own device entry point, partition_gpu_wrapper,
// implemented over the C-ABI of libfastpart.so, so that the reference's
// partition_gpu() links against this library unchanged.
//
// The reference declares one partition_gpu_wrapper per objective, with C++
// linkage and a signature of its own:
//   TV    TV/partition.cpp:6-26    (defined at TV/gpu_partition.cu:332)
//   CN    CN/partition.cpp:7-27    (defined at CN/gpu_partition.cu:263-283)
// MSRV declares the TV signature for another objective (MSRV/partition.cpp:4-24),
// so each objective gets a library of its own: this file is compiled once per
// objective with -DFP_LEGACY_TV or -DFP_LEGACY_CN.
//
// Contract kept from the reference:
//   * all arrays are the caller's host arrays, borrowed for the call; refined
//     results come back in the same arrays: part_vec, ep_ends, ep_parts,
//     ep_cnt (slots past a net's lambda read part -1, count 0) and part_sizes
//   * errors print a message and exit with the status code, as the
//     reference's gpuAssert does (TV/gpu_partition.cu:9-17)
// Choices the reference leaves open (its GPU kernels race, so no schedule of
// theirs can be reproduced):
//   * TV runs the round schedule of partition_par (TV/partition.cpp:669-896)
//     with rounds of FASTPART_ROUND vertices; the default 8 x OpenMP threads
//     is the round partition_par itself uses, so `-g` gives the partition that
//     partition_par gives at that thread count without its data race
//   * CN runs the lambda = 2 push of partition_par (CN/partition.cpp:436-836)
//   * FASTPART_DEVICE selects the GPU (default 0)
#include <cstdio>
#include <cstdlib>
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

// Uploads the hypergraph and the caller's state, runs `passes` passes of
// `metric`, and copies the state back into the caller's arrays.
// net_ptr/net_pins may be null: the net->pins view is then the transpose of
// the vertex->nets view (pins ascending per net).
void refine_in_place(fp_metric metric, const unsigned* vtx_ptr, const unsigned* vtx_nets, const unsigned* net_ptr_in,
                     const unsigned* net_pins_in, int nov, int noe, long long pins, const int* src, int K,
                     const int* ordering, int* part_vec, int* ep_ends, int* ep_parts, int* ep_cnt, int* part_sizes,
                     int passes, double C) {
  std::vector<unsigned> net_ptr, net_pins;
  if (!net_ptr_in || !net_pins_in) {
    net_ptr.resize((size_t)noe + 1);
    net_pins.resize((size_t)(pins > 0 ? pins : 1));
    int st = fp_csr_transpose(noe, nov, vtx_ptr, vtx_nets, net_ptr.data(), net_pins.data());
    if (st) die(nullptr, st, "transpose of the vertex->nets view");
    net_ptr_in = net_ptr.data();
    net_pins_in = net_pins.data();
  }
  fp_ctx* ctx = nullptr;
  int st = fp_ctx_create(&ctx, (int)env_long("FASTPART_DEVICE", 0));
  if (st) die(ctx, st, "fp_ctx_create");
  st = fp_upload(ctx, nov, noe, net_ptr_in, net_pins_in, vtx_ptr, vtx_nets, src, nullptr, ordering);
  if (st) die(ctx, st, "fp_upload");
  std::vector<long long> sizes((size_t)K);
  for (int k = 0; k < K; k++) sizes[(size_t)k] = part_sizes[k];
  st = fp_set_state(ctx, K, part_vec, ep_ends, ep_parts, ep_cnt, sizes.data());
  if (st) die(ctx, st, "fp_set_state");

  fp_schedule sched{};  // read by TV only
  sched.kind = FP_SCHED_REF;
  sched.round = (int)env_long("FASTPART_ROUND", 8L * omp_get_max_threads());
  sched.ldg_C = C;
  sched.max_pw = fp_ref_max_pw(nov, K);
  sched.min_pw = fp_ref_min_pw(nov, K);
  sched.cap = 0;
  fp_refine_stats stats{};
  st = fp_refine(ctx, metric, passes, &sched, &stats);
  if (st) die(ctx, st, "fp_refine");
  st = fp_get_state(ctx, part_vec, ep_ends, ep_parts, ep_cnt, sizes.data());
  if (st) die(ctx, st, "fp_get_state");
  for (int k = 0; k < K; k++) part_sizes[k] = (int)sizes[(size_t)k];
  std::printf("fastpart: %d passes on the GPU, %lld moves, %.3f ms on the device\n", passes, stats.moves,
              stats.device_ms);
  fp_ctx_destroy(ctx);
}

}  // namespace

#if defined(FP_LEGACY_TV)
// TV/partition.cpp:6-26. ep_starts (noe+1) are the net offsets; row_ptr/col_ind
// the vertex->nets view; marker, active_parts, scores and edge_wghts are the
// reference's scratch and net weights, which the engine does not need.
void partition_gpu_wrapper(unsigned int* row_ptr, unsigned int* col_ind, unsigned int* /*net_appearence*/, int nov,
                           int noe, long long pins, int* /*edge_wghts*/, int* src_nodes, int K, int* ordering,
                           int* /*marker*/, int* /*active_parts*/, int* part_sizes, int* /*ep_starts*/, int* ep_ends,
                           int* ep_parts, int* part_vec, int* ep_cnt, double* /*scores*/, int REFI_ITER, double C) {
  refine_in_place(FP_METRIC_TV, row_ptr, col_ind, nullptr, nullptr, nov, noe, pins, src_nodes, K, ordering, part_vec,
                  ep_ends, ep_parts, ep_cnt, part_sizes, REFI_ITER, C);
}
#elif defined(FP_LEGACY_CN)
// CN/partition.cpp:7-27. row_ptr_inv/col_ind_inv are the net->pins view in
// file order, which the lambda = 2 push scans.
void partition_gpu_wrapper(unsigned int* row_ptr, unsigned int* col_ind, unsigned int* /*net_appearence*/, int nov,
                           int noe, long long pins, int* /*edge_wgths*/, int* src_nodes, int K, int* ordering,
                           int /*REFI_OPT*/, int REFI_ITER, unsigned int* row_ptr_inv, unsigned int* col_ind_inv,
                           int* part_sizes, int* /*ep_starts*/, int* ep_parts, int* ep_ends, int* ep_cnt,
                           int* part_vec) {
  refine_in_place(FP_METRIC_CN, row_ptr, col_ind, row_ptr_inv, col_ind_inv, nov, noe, pins, src_nodes, K, ordering,
                  part_vec, ep_ends, ep_parts, ep_cnt, part_sizes, REFI_ITER, 0.0);
}
#else
#error "compile with -DFP_LEGACY_TV or -DFP_LEGACY_CN"
#endif
