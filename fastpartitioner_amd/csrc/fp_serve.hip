// fp_serve.hip -- the serving loop: resident hypergraphs, a request queue and
// worker threads that refine requests concurrently on one GPU.
//
// Every registered hypergraph is uploaded into a context of its own (both CSR
// views plus the connectivity state arrays stay in HBM); a hypergraph whose
// contexts are all busy gets another one (a replica, uploaded from the host
// copy the server keeps) while it has fewer than the server's replica limit.
// A request names a hypergraph, K, an objective, a schedule, an initial part
// vector and a pass count. Workers take requests in arrival order, skipping a
// request whose hypergraph has no free context and may not get another (that
// one waits for it), and admit a request only while the HBM footprint of the
// requests in flight plus its own stays within the server's budget. A running request sets its state
// from the part vector, refines, and evaluates the six metrics on its
// hypergraph's stream; requests on different hypergraphs overlap on the device,
// and new requests are admitted as slots and memory free up (continuous
// batching). A request's result depends only on the request: it is the same
// alone and in any mix.
#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>

#include "fp_internal.h"

struct fp_server {
  int device = 0;
  long long budget = 0;
  char err[512] = {0};
  std::chrono::steady_clock::time_point t0;

  struct Graph {
    std::vector<fp_ctx*> ctxs;  // the hypergraph's contexts (replicas), all with the same upload
    std::vector<char> busy;     // per context
    int creating = 0;           // replicas being uploaded by workers
    int nov = 0, noe = 0;
    long long pins = 0;
    long long bytes = 0;  // device bytes of one resident copy with its state arrays (after upload)
    std::vector<unsigned> net_ptr, net_pins;  // host copy for replicas
    std::vector<int> src, vwgt;
  };
  struct Request {
    long long id = 0;
    int graph = 0;
    fp_metric metric = FP_METRIC_TV;
    int K = 0;
    int passes = 0;
    bool has_sched = false;
    fp_schedule sched{};
    std::vector<int> part;
    long long footprint = 0;
    // results
    bool done = false;
    fp_status status = FP_OK;
    char err[256] = {0};
    fp_metrics metrics{};
    long long moves = 0;
    double submit_ms = 0, start_ms = 0, end_ms = 0;
  };

  int max_replicas = 1;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::vector<std::unique_ptr<Graph>> graphs;  // stable addresses: workers read host copies unlocked
  std::deque<long long> queue;
  std::unordered_map<long long, std::unique_ptr<Request>> reqs;
  long long next_id = 1;
  long long inflight_bytes = 0, peak_bytes = 0;
  int inflight = 0, peak_inflight = 0;
  long long completed = 0;
  bool stop = false;
  std::vector<std::thread> workers;

  double now_ms() const {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
};

namespace {

fp_status sfail(fp_server* s, fp_status st, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(s->err, sizeof(s->err), fmt, ap);
  va_end(ap);
  return st;
}

// HBM a request needs beyond its resident hypergraph: the per-pass scratch of
// its objective, an upper estimate that depends on the request alone (the
// context keeps its scratch between requests, so this over-counts a
// hypergraph that already served a request).
long long scratch_bytes(const fp_server::Graph& g, fp_metric metric, int K, int R) {
  const long long pins = g.pins, nov = g.nov;
  const long long r = std::min<long long>(R > 0 ? R : nov, nov);
  long long b = 8 * (long long)K * 8;  // sizes, frozen sizes, candidate totals, metric vectors
  b += (long long)K * (long long)K / 8;  // message bitmap of the metric evaluation
  switch (metric) {
    case FP_METRIC_TV: b += 8 * r + 8 * (long long)K * ((r + 63) / 64); break;
    case FP_METRIC_CN: b += 16 * pins + 40 * (long long)g.nov + 16 * pins; break;
    case FP_METRIC_MSV:
    case FP_METRIC_MSRV: b += 16 * pins + 24 * r; break;
    default: b += 4 * (long long)K * K + 16 * (long long)K * ((K + 63) / 64) + 8 * 4 * pins + 24 * r; break;
  }
  return b;
}

// a new context holding graph g's upload (the caller holds no lock)
fp_status upload_replica(fp_server* s, const fp_server::Graph& g, fp_ctx** out) {
  fp_ctx* ctx = nullptr;
  fp_status st = fp_ctx_create(&ctx, s->device);
  if (st) return st;
  st = fp_upload(ctx, g.nov, g.noe, g.net_ptr.data(), g.net_pins.data(), nullptr, nullptr,
                 g.src.empty() ? nullptr : g.src.data(), g.vwgt.empty() ? nullptr : g.vwgt.data(), nullptr);
  if (st) {
    fp_ctx_destroy(ctx);
    return st;
  }
  *out = ctx;
  return FP_OK;
}

void run_request(fp_server* s, fp_server::Request& q, fp_ctx* ctx) {
  fp_status st = fp_set_partition(ctx, q.K, q.part.data());
  fp_refine_stats stats{};
  if (st == FP_OK) st = fp_refine(ctx, q.metric, q.passes, q.has_sched ? &q.sched : nullptr, &stats);
  if (st == FP_OK) st = fp_eval(ctx, &q.metrics, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
  if (st == FP_OK) st = fp_get_state(ctx, q.part.data(), nullptr, nullptr, nullptr, nullptr);
  q.status = st;
  q.moves = stats.moves;
  if (st != FP_OK) std::snprintf(q.err, sizeof(q.err), "%s", fp_ctx_last_error(ctx));
  (void)s;
}

void worker_loop(fp_server* s) {
  (void)hipSetDevice(s->device);
  for (;;) {
    fp_server::Request* q = nullptr;
    int gi = -1, ci = -1;
    fp_ctx* ctx = nullptr;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      for (;;) {
        if (s->stop) return;
        // first request in arrival order whose hypergraph has a free context,
        // or may get another, and that fits the budget (a request larger
        // than the budget runs alone)
        for (auto it = s->queue.begin(); it != s->queue.end(); ++it) {
          fp_server::Request* c = s->reqs[*it].get();
          fp_server::Graph& cg = *s->graphs[(size_t)c->graph];
          int free_ci = -1;
          for (size_t k = 0; k < cg.ctxs.size() && free_ci < 0; k++)
            if (!cg.busy[k]) free_ci = (int)k;
          const bool grow = free_ci < 0 && (int)cg.ctxs.size() + cg.creating < s->max_replicas;
          if (free_ci < 0 && !grow) continue;
          if (s->inflight > 0 && s->inflight_bytes + c->footprint > s->budget) break;
          q = c;
          gi = c->graph;
          ci = free_ci;
          s->queue.erase(it);
          break;
        }
        if (q) break;
        s->cv_work.wait(lk);
      }
      fp_server::Graph& g = *s->graphs[(size_t)gi];
      if (ci >= 0) {
        g.busy[(size_t)ci] = 1;
        ctx = g.ctxs[(size_t)ci];
      } else {
        g.creating++;
      }
      s->inflight++;
      s->inflight_bytes += q->footprint;
      s->peak_inflight = std::max(s->peak_inflight, s->inflight);
      s->peak_bytes = std::max(s->peak_bytes, s->inflight_bytes);
      q->start_ms = s->now_ms();
    }
    if (ci < 0) {
      // a new replica for this request, uploaded outside the lock from the
      // host copy (which the server never changes after registration)
      const fp_server::Graph* gp = nullptr;
      {
        std::lock_guard<std::mutex> lk(s->mu);
        gp = s->graphs[(size_t)gi].get();
      }
      fp_ctx* nc = nullptr;
      const fp_status st = upload_replica(s, *gp, &nc);
      std::lock_guard<std::mutex> lk(s->mu);
      fp_server::Graph& g = *s->graphs[(size_t)gi];
      g.creating--;
      if (st == FP_OK) {
        g.ctxs.push_back(nc);
        g.busy.push_back(1);
        ci = (int)g.ctxs.size() - 1;
        ctx = nc;
      } else {
        q->status = st;
        std::snprintf(q->err, sizeof(q->err), "replica upload failed");
      }
    }
    if (ctx) run_request(s, *q, ctx);
    {
      std::lock_guard<std::mutex> lk(s->mu);
      q->end_ms = s->now_ms();
      q->done = true;
      if (ci >= 0) s->graphs[(size_t)gi]->busy[(size_t)ci] = 0;
      s->inflight--;
      s->inflight_bytes -= q->footprint;
      s->completed++;
    }
    s->cv_work.notify_all();
    s->cv_done.notify_all();
  }
}

}  // namespace

using namespace fpe;

extern "C" {

fp_status fp_server_create(fp_server** out, int device, long long hbm_budget, int workers) {
  if (!out || hbm_budget <= 0 || workers < 1 || workers > 256) return FP_ERR_ARG;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
    (void)hipGetLastError();
    return FP_ERR_NODEV;
  }
  if (device < 0 || device >= n) return FP_ERR_ARG;
  fp_server* s = new fp_server();
  s->device = device;
  s->budget = hbm_budget;
  s->t0 = std::chrono::steady_clock::now();
  for (int i = 0; i < workers; i++) s->workers.emplace_back(worker_loop, s);
  *out = s;
  return FP_OK;
}

void fp_server_destroy(fp_server* s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  for (auto& t : s->workers) t.join();
  for (auto& g : s->graphs)
    for (fp_ctx* c : g->ctxs) fp_ctx_destroy(c);
  delete s;
}

fp_status fp_server_set_replicas(fp_server* s, int max_replicas) {
  if (!s || max_replicas < 1 || max_replicas > 64) return FP_ERR_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  s->max_replicas = max_replicas;
  return FP_OK;
}

const char* fp_server_last_error(const fp_server* s) { return s ? s->err : "null server"; }

fp_status fp_server_add_graph(fp_server* s, int nov, int noe, const unsigned* net_ptr, const unsigned* net_pins,
                              const int* src, const int* vwgt, int* graph_id) {
  if (!s || !graph_id) return FP_ERR_ARG;
  fp_ctx* ctx = nullptr;
  fp_status st = fp_ctx_create(&ctx, s->device);
  if (st) return sfail(s, st, "fp_ctx_create failed");
  st = fp_upload(ctx, nov, noe, net_ptr, net_pins, nullptr, nullptr, src, vwgt, nullptr);
  if (st) {
    sfail(s, st, "fp_upload: %s", fp_ctx_last_error(ctx));
    fp_ctx_destroy(ctx);
    return st;
  }
  auto g = std::make_unique<fp_server::Graph>();
  g->ctxs.push_back(ctx);
  g->busy.push_back(0);
  g->nov = nov;
  g->noe = noe;
  g->pins = net_ptr[noe];
  g->bytes = fp_ctx_device_bytes(ctx);
  // host copy for replicas
  g->net_ptr.assign(net_ptr, net_ptr + noe + 1);
  g->net_pins.assign(net_pins, net_pins + g->pins);
  if (src) g->src.assign(src, src + noe);
  if (vwgt) g->vwgt.assign(vwgt, vwgt + nov);
  std::lock_guard<std::mutex> lk(s->mu);
  s->graphs.push_back(std::move(g));
  *graph_id = (int)s->graphs.size() - 1;
  return FP_OK;
}

fp_status fp_server_submit(fp_server* s, int graph_id, fp_metric metric, int K, const int* part,
                           const fp_schedule* sched, int passes, long long* request_id) {
  if (!s || !part || !request_id || passes < 0 || K < 1) return FP_ERR_ARG;
  if (metric != FP_METRIC_CN && !sched) return sfail(s, FP_ERR_ARG, "a schedule is needed for metric %d", (int)metric);
  auto q = std::make_unique<fp_server::Request>();
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (graph_id < 0 || graph_id >= (int)s->graphs.size()) return sfail(s, FP_ERR_ARG, "no hypergraph %d", graph_id);
    const fp_server::Graph& g = *s->graphs[(size_t)graph_id];
    q->part.assign(part, part + g.nov);
    q->footprint = g.bytes + scratch_bytes(g, metric, K, sched ? sched->round : 0);
    if (q->footprint > s->budget)
      return sfail(s, FP_ERR_BUDGET, "request needs %lld bytes, the budget is %lld", q->footprint, s->budget);
  }
  q->graph = graph_id;
  q->metric = metric;
  q->K = K;
  q->passes = passes;
  q->has_sched = sched != nullptr;
  if (sched) q->sched = *sched;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    q->id = s->next_id++;
    q->submit_ms = s->now_ms();
    *request_id = q->id;
    s->queue.push_back(q->id);
    s->reqs[q->id] = std::move(q);
  }
  s->cv_work.notify_all();
  return FP_OK;
}

fp_status fp_server_wait(fp_server* s, long long request_id, fp_request_info* info, int* part_out) {
  if (!s) return FP_ERR_ARG;
  std::unique_ptr<fp_server::Request> q;
  {
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(request_id);
    if (it == s->reqs.end()) return sfail(s, FP_ERR_ARG, "no request %lld", request_id);
    fp_server::Request* r = it->second.get();
    s->cv_done.wait(lk, [&] { return r->done; });
    q = std::move(it->second);
    s->reqs.erase(it);
  }
  if (info) {
    info->id = q->id;
    info->graph = q->graph;
    info->metric = (int)q->metric;
    info->K = q->K;
    info->passes = q->passes;
    info->status = q->status;
    info->moves = q->moves;
    info->metrics = q->metrics;
    info->submit_ms = q->submit_ms;
    info->start_ms = q->start_ms;
    info->end_ms = q->end_ms;
    info->footprint = q->footprint;
  }
  if (part_out) std::memcpy(part_out, q->part.data(), sizeof(int) * q->part.size());
  if (q->status != FP_OK) sfail(s, q->status, "request %lld: %s", request_id, q->err);
  return q->status;
}

fp_status fp_server_stats(fp_server* s, long long* completed, int* peak_inflight, long long* peak_bytes) {
  if (!s) return FP_ERR_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  if (completed) *completed = s->completed;
  if (peak_inflight) *peak_inflight = s->peak_inflight;
  if (peak_bytes) *peak_bytes = s->peak_bytes;
  return FP_OK;
}

}  // extern "C"
