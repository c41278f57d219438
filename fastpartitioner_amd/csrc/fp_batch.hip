// fp_batch.hip -- batched refinement: many resident hypergraphs (contexts) of
// uneven size, K and objective refined together.
//
// Members are spread over L lanes by pass cost (the autotuner's measured pass
// time, else pins; longest first onto the least loaded lane). One pass of the
// whole batch (the kernel sequences fp_refine enqueues for every member, with
// no host synchronisation in them) is captured once into a single HIP graph:
// the batch stream forks into the lanes, each lane holds its members' passes
// back to back, and the lanes join again. A pass of the batch is one launch
// of that graph; its lanes are independent branches that the runtime runs
// side by side. Results are those of fp_refine on each member alone: lanes
// and launch shapes only change how the work is spread over the GPU.
//
// Autotuning: the first passes of a batch are run eagerly, each with another
// candidate launch shape (wavefronts per evaluation block, evaluation grid
// cap), timed per member with events on the member's stream; every member
// then keeps its fastest shape for the captured passes.
#include "fp_internal.h"

struct fp_batch {
  int device = -1;
  hipStream_t stream = nullptr;
  char err[512] = {0};
  struct Member {
    fp_ctx* ctx;
    fp_metric metric;
    fp_schedule sched;
    bool has_sched;
    fpe::RefinePlan plan;
    long long moves = 0;
    float tuned_ms = 0;
  };
  std::vector<Member> members;
  int lanes = 8;
  std::vector<hipStream_t> lane_streams;
  hipGraph_t graph = nullptr;          // one pass of every member, lanes as parallel branches
  hipGraphExec_t exec = nullptr;
  std::vector<int> lane_of;  // member -> lane
  std::vector<hipEvent_t> events;
  std::vector<unsigned long long> gen_at_capture;  // members' generations the captured graphs were recorded at
};

namespace {

fp_status bfail(fp_batch* b, fp_status st, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(b->err, sizeof(b->err), fmt, ap);
  va_end(ap);
  return st;
}

#define B_HIP_OK(call)                                                                  \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess) return bfail(b, FP_ERR_HIP, "%s: %s", #call, hipGetErrorString(e_)); \
  } while (0)

void drop_graphs(fp_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
}

hipEvent_t batch_event(fp_batch* b, size_t i) {
  while (b->events.size() <= i) {
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    b->events.push_back(e);
  }
  return b->events[i];
}

// a member's failure becomes the batch's, with the member named
fp_status member_fail(fp_batch* b, size_t i, fp_status st) {
  return bfail(b, st, "member %zu: %s", i, b->members[i].ctx->err);
}

// candidate launch shapes: (waves per evaluation block, evaluation grid cap)
constexpr int TUNE_SHAPES[][2] = {{0, 0}, {1, 0}, {2, 0}, {4, 2048}, {2, 1024}, {1, 4096}};
constexpr int NUM_SHAPES = (int)(sizeof(TUNE_SHAPES) / sizeof(TUNE_SHAPES[0]));

}  // namespace

using namespace fpe;

extern "C" {

fp_status fp_batch_create(fp_batch** out, int device) {
  if (!out) return FP_ERR_ARG;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
    (void)hipGetLastError();
    return FP_ERR_NODEV;
  }
  if (device < 0 || device >= n) return FP_ERR_ARG;
  fp_batch* b = new fp_batch();
  b->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
    (void)hipGetLastError();
    delete b;
    return FP_ERR_HIP;
  }
  *out = b;
  return FP_OK;
}

void fp_batch_destroy(fp_batch* b) {
  if (!b) return;
  (void)hipSetDevice(b->device);
  (void)hipStreamSynchronize(b->stream);
  for (hipStream_t s : b->lane_streams) (void)hipStreamSynchronize(s);
  drop_graphs(b);
  for (hipEvent_t e : b->events) (void)hipEventDestroy(e);
  for (hipStream_t s : b->lane_streams) (void)hipStreamDestroy(s);
  (void)hipStreamDestroy(b->stream);
  delete b;
}

fp_status fp_batch_set_lanes(fp_batch* b, int lanes) {
  if (!b || lanes < 1 || lanes > 64) return FP_ERR_ARG;
  if (lanes != b->lanes) drop_graphs(b);
  b->lanes = lanes;
  return FP_OK;
}

const char* fp_batch_last_error(const fp_batch* b) { return b ? b->err : "null batch"; }

fp_status fp_batch_add(fp_batch* b, fp_ctx* ctx, fp_metric metric, const fp_schedule* sched) {
  if (!b || !ctx) return FP_ERR_ARG;
  if (ctx->device != b->device) return bfail(b, FP_ERR_ARG, "context on device %d, batch on %d", ctx->device, b->device);
  if (metric != FP_METRIC_CN && !sched) return bfail(b, FP_ERR_ARG, "a schedule is needed for metric %d", (int)metric);
  for (auto& m : b->members)
    if (m.ctx == ctx) return bfail(b, FP_ERR_ARG, "context added twice");
  fp_batch::Member m{};
  m.ctx = ctx;
  m.metric = metric;
  m.has_sched = sched != nullptr;
  if (sched) m.sched = *sched;
  b->members.push_back(m);
  drop_graphs(b);
  return FP_OK;
}

int fp_batch_size(const fp_batch* b) { return b ? (int)b->members.size() : 0; }

fp_status fp_batch_refine(fp_batch* b, int passes, int tune_passes, double* device_ms) {
  if (!b) return FP_ERR_ARG;
  if (passes < 0 || tune_passes < 0) return bfail(b, FP_ERR_ARG, "negative pass count");
  B_HIP_OK(hipSetDevice(b->device));
  const size_t nm = b->members.size();
  // a captured batch graph is reused when no tuning is asked for and no
  // member changed anything its recorded kernel arguments rely on (device
  // buffers, K, the list layout) since the capture
  if (tune_passes > 0) drop_graphs(b);
  // plan every member (initialisation kernels run on the member streams)
  for (size_t i = 0; i < nm; i++) {
    auto& m = b->members[i];
    RefinePlan p;  // keeps the launch shape a previous call tuned
    p.eval_waves = m.plan.eval_waves;
    p.eval_grid = m.plan.eval_grid;
    m.plan = p;
    fp_status st = plan_refine(m.ctx, m.metric, m.has_sched ? &m.sched : nullptr, m.plan);
    if (st) return member_fail(b, i, st);
    fill_async<unsigned long long>(m.ctx, m.ctx->moves, 1, 0ull);
    m.moves = 0;
    if (tune_passes > 0) m.tuned_ms = 0;  // a call without trials keeps the last measured times
  }
  hipEvent_t t0 = batch_event(b, 0), t1 = batch_event(b, 1);
  if (!t0 || !t1) return bfail(b, FP_ERR_HIP, "hipEventCreate failed");
  for (size_t i = 0; i < nm; i++) B_HIP_OK(hipStreamSynchronize(b->members[i].ctx->stream));
  B_HIP_OK(hipEventRecord(t0, b->stream));

  // autotuning passes: real passes, each member with the next candidate shape
  const int ntune = std::min(tune_passes, passes);
  std::vector<std::vector<float>> shape_ms(nm, std::vector<float>(NUM_SHAPES, -1.f));
  for (int t = 0; t < ntune; t++) {
    const int sh = t % NUM_SHAPES;
    for (size_t i = 0; i < nm; i++) {
      auto& m = b->members[i];
      RefinePlan p = m.plan;
      p.eval_waves = TUNE_SHAPES[sh][0];
      p.eval_grid = TUNE_SHAPES[sh][1];
      fp_status st = plan_refine(m.ctx, m.metric, m.has_sched ? &m.sched : nullptr, p);
      if (st) return member_fail(b, i, st);
      hipEvent_t e0 = batch_event(b, 2 + 2 * i), e1 = batch_event(b, 3 + 2 * i);
      if (!e0 || !e1) return bfail(b, FP_ERR_HIP, "hipEventCreate failed");
      B_HIP_OK(hipEventRecord(e0, m.ctx->stream));
      st = record_pass(m.ctx, p);
      if (st) return member_fail(b, i, st);
      B_HIP_OK(hipEventRecord(e1, m.ctx->stream));
    }
    for (size_t i = 0; i < nm; i++) {
      auto& m = b->members[i];
      B_HIP_OK(hipEventSynchronize(batch_event(b, 3 + 2 * i)));
      float ms = 0;
      B_HIP_OK(hipEventElapsedTime(&ms, batch_event(b, 2 + 2 * i), batch_event(b, 3 + 2 * i)));
      float& slot = shape_ms[i][sh];
      slot = slot < 0 ? ms : std::min(slot, ms);
      if (m.metric == FP_METRIC_CN) {
        // an overflowing pass left the state unchanged: grow and run it again
        int redo = 0;
        fp_status st = cn_after_pass(m.ctx, &redo);
        while (st == FP_OK && redo) {
          RefinePlan p = m.plan;
          st = record_pass(m.ctx, p);
          if (st == FP_OK) st = cn_after_pass(m.ctx, &redo);
        }
        if (st) return member_fail(b, i, st);
      }
    }
  }
  // each member keeps its fastest measured shape (its plan is re-made for it:
  // the initialisation kernels of a plan recompute the same values)
  for (size_t i = 0; i < nm && ntune > 0; i++) {
    auto& m = b->members[i];
    int best = 0;
    for (int sh = 1; sh < NUM_SHAPES; sh++)
      if (shape_ms[i][sh] >= 0 && (shape_ms[i][best] < 0 || shape_ms[i][sh] < shape_ms[i][best])) best = sh;
    m.tuned_ms = shape_ms[i][best];
    RefinePlan p;
    p.eval_waves = TUNE_SHAPES[best][0];
    p.eval_grid = TUNE_SHAPES[best][1];
    fp_status st = plan_refine(m.ctx, m.metric, m.has_sched ? &m.sched : nullptr, p);
    if (st) return member_fail(b, i, st);
    m.plan = p;
  }

  const int rest = passes - ntune;
  if (b->exec) {
    bool same = b->gen_at_capture.size() == nm;
    for (size_t i = 0; same && i < nm; i++) same = b->gen_at_capture[i] == b->members[i].ctx->gen;
    if (!same) drop_graphs(b);
  }
  const int L = std::max(1, std::min(b->lanes, (int)nm));
  if (rest > 0 && nm > 0 && !b->exec) {
    while ((int)b->lane_streams.size() < L) {
      hipStream_t s = nullptr;
      B_HIP_OK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
      b->lane_streams.push_back(s);
    }
    // longest first onto the least loaded lane; the pass cost is the time a
    // member's tuned shape took in the autotuner trials when every member has
    // one, its pins + vertices otherwise
    bool timed = true;
    for (size_t i = 0; i < nm; i++) timed = timed && b->members[i].tuned_ms > 0;
    std::vector<double> cost(nm);
    for (size_t i = 0; i < nm; i++)
      cost[i] = timed ? (double)b->members[i].tuned_ms : (double)(b->members[i].ctx->pins + b->members[i].ctx->nov);
    std::vector<size_t> order(nm);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) { return cost[x] > cost[y]; });
    std::vector<double> load((size_t)L, 0.0);
    b->lane_of.assign(nm, 0);
    for (size_t i : order) {
      const int l = (int)(std::min_element(load.begin(), load.end()) - load.begin());
      b->lane_of[i] = l;
      load[(size_t)l] += cost[i];
    }
    for (size_t i = 0; i < nm; i++) B_HIP_OK(hipStreamSynchronize(b->members[i].ctx->stream));
    {
      // one graph: the batch stream forks into the lanes, each lane records
      // its members' passes back to back, and the lanes join back
      hipStream_t os = b->stream;
      hipEvent_t fork = batch_event(b, 2 + 4 * nm);
      if (!fork) return bfail(b, FP_ERR_HIP, "hipEventCreate failed");
      for (int l = 0; l < L; l++)
        if (!batch_event(b, 2 + 4 * nm + 1 + (size_t)l)) return bfail(b, FP_ERR_HIP, "hipEventCreate failed");
      B_HIP_OK(hipStreamBeginCapture(os, hipStreamCaptureModeThreadLocal));
      fp_status st = FP_OK;
      size_t bad = 0;
      hipError_t he = hipEventRecord(fork, os);
      for (int l = 0; l < L && he == hipSuccess; l++) he = hipStreamWaitEvent(b->lane_streams[(size_t)l], fork, 0);
      for (size_t i = 0; i < nm && st == FP_OK && he == hipSuccess; i++) {
        auto& m = b->members[i];
        const hipStream_t own = m.ctx->stream;
        const bool timing = m.ctx->kernel_timing;
        m.ctx->stream = b->lane_streams[(size_t)b->lane_of[i]];
        m.ctx->kernel_timing = false;
        st = record_pass(m.ctx, m.plan);
        m.ctx->stream = own;
        m.ctx->kernel_timing = timing;
        bad = i;
      }
      for (int l = 0; l < L && he == hipSuccess; l++) {
        hipEvent_t j = b->events[2 + 4 * nm + 1 + (size_t)l];
        he = hipEventRecord(j, b->lane_streams[(size_t)l]);
        if (he == hipSuccess) he = hipStreamWaitEvent(os, j, 0);
      }
      hipGraph_t g = nullptr;
      const hipError_t ce = hipStreamEndCapture(os, &g);
      if (st) {
        if (g) (void)hipGraphDestroy(g);
        return member_fail(b, bad, st);
      }
      if (he != hipSuccess || ce != hipSuccess || !g) {
        if (g) (void)hipGraphDestroy(g);
        return bfail(b, FP_ERR_HIP, "pass capture failed: %s", hipGetErrorString(he != hipSuccess ? he : ce));
      }
      b->graph = g;
      B_HIP_OK(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
    }
    b->gen_at_capture.resize(nm);
    for (size_t i = 0; i < nm; i++) b->gen_at_capture[i] = b->members[i].ctx->gen;
  }
  if (rest > 0 && nm > 0) {
    // the pass graph starts after the members' planning work
    for (size_t i = 0; i < nm; i++) {
      hipEvent_t e = batch_event(b, 2 + 3 * nm + i);
      if (!e) return bfail(b, FP_ERR_HIP, "hipEventCreate failed");
      B_HIP_OK(hipEventRecord(e, b->members[i].ctx->stream));
      B_HIP_OK(hipStreamWaitEvent(b->stream, e, 0));
    }
    for (int it = 0; it < rest; it++) B_HIP_OK(hipGraphLaunch(b->exec, b->stream));
  }
  // the batch stream's end event waits for work left on member streams
  for (size_t i = 0; i < nm; i++) {
    hipEvent_t e = batch_event(b, 2 + 2 * nm + i);
    if (!e) return bfail(b, FP_ERR_HIP, "hipEventCreate failed");
    B_HIP_OK(hipEventRecord(e, b->members[i].ctx->stream));
    B_HIP_OK(hipStreamWaitEvent(b->stream, e, 0));
  }
  B_HIP_OK(hipEventRecord(t1, b->stream));
  B_HIP_OK(hipStreamSynchronize(b->stream));
  for (size_t i = 0; i < nm; i++) B_HIP_OK(hipStreamSynchronize(b->members[i].ctx->stream));
  float ms = 0;
  B_HIP_OK(hipEventElapsedTime(&ms, t0, t1));
  if (device_ms) *device_ms = ms;
  // per-member results; a CN member whose edit list overflowed in a captured
  // pass skipped that pass, which is reported as an error
  for (size_t i = 0; i < nm; i++) {
    auto& m = b->members[i];
    unsigned long long mv = 0;
    if (d2h(m.ctx, &mv, m.ctx->moves, 1) != FP_OK) return member_fail(b, i, FP_ERR_HIP);
    if (m.metric == FP_METRIC_CN) {
      int f[2] = {0, 0};
      if (d2h(m.ctx, f, m.ctx->flags, 2) != FP_OK) return member_fail(b, i, FP_ERR_HIP);
      B_HIP_OK(hipStreamSynchronize(m.ctx->stream));
      if (f[0] || f[1]) return bfail(b, FP_ERR_UNSUPPORTED, "member %zu: a CN pass overflowed its edit list and was skipped", i);
    }
    B_HIP_OK(hipStreamSynchronize(m.ctx->stream));
    m.moves = (long long)mv;
  }
  return FP_OK;
}

fp_status fp_batch_member(const fp_batch* b, int i, long long* moves, int* eval_waves, int* eval_grid,
                          double* tuned_pass_ms) {
  if (!b || i < 0 || i >= (int)b->members.size()) return FP_ERR_ARG;
  const auto& m = b->members[(size_t)i];
  if (moves) *moves = m.moves;
  if (eval_waves) *eval_waves = m.plan.eval_waves;
  if (eval_grid) *eval_grid = m.plan.eval_grid;
  if (tuned_pass_ms) *tuned_pass_ms = m.tuned_ms;
  return FP_OK;
}

}  // extern "C"
