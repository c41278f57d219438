/* fastpart.h -- C-ABI of the FastPartitioner refinement and metric engine for
 * MI355X (gfx950), libfastpart.so.
 *
 * The engine keeps one hypergraph per context resident in HBM (both CSR
 * views, net sources, vertex weights) together with its connectivity state
 * (part vector, per-net lists of (part, pin count) and their lengths lambda,
 * part weights) and runs refinement passes and metric evaluation on it.
 *
 * It replaces the device side of the reference's per-metric
 * partition_gpu_wrapper(...) entry points (TV/partition.cpp:6-26 declared,
 * TV/gpu_partition.cu:332-553 defined; CN/partition.cpp:7-27; TM/partition.cpp:3-7;
 * MGV/partition.cpp:4-25; MSRV/partition.cpp:4-24). A caller that already holds
 * the reference's host arrays maps them as:
 *   row_ptr/col_ind (vertex->nets)      -> fp_upload vtx_ptr/vtx_nets
 *   the reader's net->pins CSR          -> fp_upload net_ptr/net_pins
 *   src_nodes                           -> fp_upload src
 *   part_vec, ep_ends, ep_parts, ep_cnt -> fp_set_state / fp_get_state
 *   ep_starts                           == net_ptr (the list of net n starts at net_ptr[n])
 *   REFI_ITER, C                        -> fp_refine passes, fp_schedule.ldg_C
 *
 * Conventions: every entry point returns fp_status (0 = FP_OK) and never
 * exits the process; fp_ctx_last_error() explains the last failure. Arrays are
 * host pointers to plain int/unsigned/long long data; the engine copies them.
 * One context = one HIP stream; calls on one context must not overlap, calls
 * on different contexts may run from different threads.
 */
#ifndef FASTPART_H
#define FASTPART_H
#include <stddef.h>
#ifdef __cplusplus
extern "C" {
#endif

#ifndef FP_STATUS_DEFINED
#define FP_STATUS_DEFINED
typedef enum {
  FP_OK = 0,
  FP_ERR_ARG = 1,         /* invalid argument or malformed input */
  FP_ERR_HIP = 2,         /* HIP runtime error */
  FP_ERR_OOM = 3,         /* device allocation failed */
  FP_ERR_STATE = 4,       /* call out of order (e.g. refine before upload) */
  FP_ERR_UNSUPPORTED = 5, /* input outside the supported envelope */
  FP_ERR_BUDGET = 6,      /* rejected by an HBM footprint budget */
  FP_ERR_IO = 7,          /* file I/O error */
  FP_ERR_NODEV = 8        /* no usable GPU */
} fp_status;
#endif

typedef struct fp_ctx fp_ctx;

/* The six objectives (shared/metric.h:7-21). */
typedef enum {
  FP_METRIC_CN = 0,   /* cut nets */
  FP_METRIC_TV = 1,   /* total volume, sum of (lambda - 1) */
  FP_METRIC_TM = 2,   /* total messages */
  FP_METRIC_MSM = 3,  /* max send messages */
  FP_METRIC_MSV = 4,  /* max send volume */
  FP_METRIC_MSRV = 5  /* max send + receive volume */
} fp_metric;

/* Refinement schedules. Both are deterministic: results do not depend on
 * thread timing or atomic ordering.
 *
 * FP_SCHED_REF: the reference's round schedule partition_par
 *   (TV/partition.cpp:669-896) without its data race (TV/partition.cpp:806-814):
 *   rounds of `round` consecutive positions of the ordering; every vertex of a
 *   round decides on the state frozen at the round start; accepted moves are
 *   applied in ordering order; as in the reference, applying a move lowers the
 *   source part's size and leaves the destination's unchanged
 *   (TV/partition.cpp:827-829). round = 8 * threads reproduces partition_par.
 *
 * FP_SCHED_BAL: the balanced round schedule. Same per-round decisions on
 *   weighted part sizes; then for each destination part d the candidates of
 *   the round, in ordering order, are admitted while
 *   size[d] + (weights of earlier candidates into d) + w <= cap
 *   (a prefix sum per part); admitted moves update both part sizes.
 */
typedef enum { FP_SCHED_REF = 0, FP_SCHED_BAL = 1 } fp_sched_kind;

typedef struct {
  int kind;          /* fp_sched_kind */
  int round;         /* vertices per round, >= 1 */
  double ldg_C;      /* LDG capacity constant C of TV/partition.cpp:103 */
  long long max_pw;  /* a destination heavier than this is refused (TV/partition.cpp:135-138) */
  long long min_pw;  /* a destination lighter than this is refused (TV/partition.cpp:140-143) */
  long long cap;     /* FP_SCHED_BAL: per-part admission cap */
} fp_schedule;

typedef struct {
  long long moves;        /* applied moves */
  long long rounds;       /* rounds executed */
  long long pin_visits;   /* passes * pins: (vertex, net) incidences scored */
  double device_ms;       /* device time of the passes (HIP events on the context stream) */
} fp_refine_stats;

typedef struct {
  long long tv, cn, tm, msm, msv, msrv;
} fp_metrics;

/* ---- context ------------------------------------------------------------ */
/* A fixed English name of a status code. */
const char* fp_status_string(fp_status st);
fp_status fp_device_count(int* count);
fp_status fp_ctx_create(fp_ctx** ctx, int device);
void fp_ctx_destroy(fp_ctx* ctx);
const char* fp_ctx_last_error(const fp_ctx* ctx);
/* bytes of device memory the context holds now */
long long fp_ctx_device_bytes(const fp_ctx* ctx);

/* ---- hypergraph --------------------------------------------------------- */
/* Uploads both CSR views once (shared/hypergraph_rw.cpp:97-242 and the
 * transpose of TV/tv.cpp:76-100). vtx_ptr/vtx_nets may be NULL: the engine
 * then builds the vertex->nets view itself (ascending nets per vertex, as the
 * reference transpose). src (net source vertex, -1 = none) and vwgt (vertex
 * weights, NULL = unit) are optional; ordering is the vertex visiting order
 * of the refinement (NULL = identity; TV/tv.cpp:122-128). Replaces any
 * previous hypergraph and state of the context. */
fp_status fp_upload(fp_ctx* ctx, int nov, int noe, const unsigned* net_ptr, const unsigned* net_pins,
                    const unsigned* vtx_ptr, const unsigned* vtx_nets, const int* src, const int* vwgt,
                    const int* ordering);
/* Host arrays travel by DMA: arrays in pinned memory (fp_host_alloc_pinned,
 * or any other page-locked allocation) are copied directly; pageable arrays
 * of 16 MiB and more pass through two 16 MiB pinned staging buffers of the
 * context (allocated on the first such copy, freed by fp_ctx_destroy), filled
 * by all host threads while the previous piece is in flight. A reader that
 * fills both CSR views in pinned memory uploads at the PCIe rate. */
void* fp_host_alloc_pinned(size_t bytes); /* NULL on failure */
void fp_host_free_pinned(void* p);

/* ---- connectivity state ------------------------------------------------- */
/* State from a part vector alone. Each net's list is ordered by first
 * appearance along the ordering, which is the state the reference's
 * streaming initial partition leaves (TV/partition.cpp:240-263). */
fp_status fp_set_partition(fp_ctx* ctx, int K, const int* part);

/* Streaming initial partition on the device, leaving the part vector, the
 * connectivity lists (in placement order, as the reference's stream leaves
 * them: TV/partition.cpp:240-263) and the part sizes resident, ready for
 * fp_refine. Replaces the host stream of TV/partition.cpp:207-264 (mode
 * FP_INIT_LDG_BANDS = 0, LDG2 with the max_pw band), MGV/partition.cpp:122-189
 * and MSRV/partition.cpp:131-208 (FP_INIT_LDG = 1, band-free LDG) and
 * CN/partition.cpp:132-221 (FP_INIT_CN_LDG = 2, votes of uncut nets only);
 * the mode values are those of fastpart_host.h's fp_initial_partition.
 * Buffered rounds: the ordering is cut into rounds of `round` positions; each
 * sub-round decides every unplaced vertex of the round on the frozen state;
 * candidates with no positive score fall back in turn, the j-th (ordering
 * order) taking part (f + j) mod K, f the lowest-index part of minimum size;
 * per part, the longest ordering-order prefix of its candidates within cap is
 * admitted (the first candidate into a part always); the admitted are placed
 * in ordering order and the rest decided again. round = 1 is the reference's
 * sequential stream exactly (where the reference is defined: a choice above
 * max_pw, where it would index part_sizes[-1], falls back here). C is the
 * LDG capacity (TV/partition.cpp:159-160), max_pw the band (mode 0 only),
 * cap the admission cap (fp_balance_cap). stats (nullable) receives the
 * rounds, the sub-rounds and the device time. */
typedef struct {
  long long rounds;
  long long subrounds;
  double device_ms;
} fp_init_stats;
fp_status fp_initial_partition_device(fp_ctx* ctx, int mode, int K, int round, double C, long long max_pw,
                                      long long cap, fp_init_stats* stats);

/* BFS vertex ordering on the device: the reference's -o 1 ordering
 * (shared/utils.cpp:433-590 as TV/tv.cpp:110-128 drives it: a BFS from
 * start, one from the farthest vertex it found -- the lowest-index vertex of
 * the largest distance -- and the visit order of a BFS from the farthest
 * vertex of that one), equal to the reference's with one thread (its
 * several-thread order depends on thread timing). The BFS goes vertex -> its
 * nets -> each net's pins in the order of `view`:
 *   FP_BFS_PINS_ASCENDING: ascending vertex order, the createInverse view
 *     TV/MGV/TM expand (shared/utils.cpp:32-56); needs a context uploaded
 *     with the identity ordering
 *   FP_BFS_PINS_UPLOADED: the uploaded (reader) order, CN/cn.cpp:345-354's view
 * Vertices the BFS does not reach, which the reference leaves at -1, follow
 * in ascending order, so ordering[nov] is always a permutation to hand to
 * fp_upload. The reference starts from rand() % nov (glibc: 1804289383 %
 * nov). farthest_first / farthest_all / stats (nullable) receive the two
 * farthest vertices, and the reached count, levels of the last BFS and
 * device time. Needs max vertex degree * max net size within 32 bits. */
enum { FP_BFS_PINS_ASCENDING = 0, FP_BFS_PINS_UPLOADED = 1 };
typedef struct {
  long long reached;
  long long levels;
  double device_ms;
} fp_bfs_stats;
fp_status fp_bfs_ordering(fp_ctx* ctx, int start, int view, int* ordering, int* farthest_first, int* farthest_all,
                          fp_bfs_stats* stats);

/* State in the reference's layout (TV/partition.cpp:150-158): ep_ends[noe]
 * (lambda), and per net n the slots net_ptr[n] .. net_ptr[n]+ep_ends[n]-1 of
 * ep_parts[pins]/ep_cnt[pins]. sizes (K, nullable) are the part sizes to
 * refine from; NULL = recomputed from part and the vertex weights. */
fp_status fp_set_state(fp_ctx* ctx, int K, const int* part, const int* ep_ends, const int* ep_parts,
                       const int* ep_cnt, const long long* sizes);

/* Downloads the state in the reference's layout; slots past a net's lambda
 * read part -1 / count 0, as the reference leaves them. Any pointer may be
 * NULL; with ep_ends, ep_parts and ep_cnt all NULL the lists are not read. */
fp_status fp_get_state(fp_ctx* ctx, int* part, int* ep_ends, int* ep_parts, int* ep_cnt, long long* sizes);

/* ---- refinement --------------------------------------------------------- */
/* `passes` refinement passes of `metric` under `sched` on the resident state.
 * TV: vertex-centric LDG refinement of TV/partition.cpp:307-440 (decision and
 * gain test), rounds per sched (FP_SCHED_REF or FP_SCHED_BAL).
 * CN: the lambda = 2 push of CN/partition.cpp:436-836; sched is not used and
 * may be NULL.
 * MSV / MSRV: partition_par of MGV/partition.cpp:706-1040 / MSRV/partition.cpp:
 * 769-1110, FP_SCHED_REF only: rounds of sched->round positions (MSV: vertex
 * ids, so the context's ordering must be the identity; MSRV: the ordering),
 * band-free LDG with capacity sched->ldg_C, the max-volume guard, and the
 * round cut after the first move that lets a new part into the max set.
 * round = 256 * threads reproduces the reference.
 * TM / MSM: partition_par of TM/partition.cpp:947-1265 / MSM/partition.cpp:
 * 1010-1400, FP_SCHED_REF only, vertex ids in order (identity ordering).
 * TM: capacity sched->ldg_C, destinations at or above 1.02 * ldg_C skipped at
 * apply; round >= nov reproduces the reference (one frozen round per pass).
 * MSM: the max-message guard and round cut; round = 128 * threads reproduces
 * the reference. The volume and message objectives need K <= 4096 and net
 * sources that are pins of their nets.
 * stats may be NULL. */
fp_status fp_refine(fp_ctx* ctx, fp_metric metric, int passes, const fp_schedule* sched, fp_refine_stats* stats);

/* One frozen-state evaluation of every vertex without applying anything:
 * dec[v] = LDG decision (TV/partition.cpp:349), gain[v] = leave_gain -
 * arrival_loss (TV/partition.cpp:355-380) or INT_MIN when dec[v] is v's own
 * part. Either pointer may be NULL. */
fp_status fp_eval_moves(fp_ctx* ctx, fp_metric metric, const fp_schedule* sched, int* dec, int* gain);

/* ---- metrics ------------------------------------------------------------ */
/* All six metrics of the resident part vector, from scratch on the device
 * (shared/metric.cpp: TV :57, CN :3, TM_scratch :122, MGM_scratch :216,
 * MGV_Scratch :301, MGAV_Scratch :353). Optional outputs: lam[noe] (lambda per
 * net); K-long int64 vectors sv/rv (MGAV_Scratch send/receive volumes),
 * sm/rm (distinct messages sent/received per part, CN/cn.cpp:54-58) and
 * sv_msv (MGV_Scratch send volumes). */
fp_status fp_eval(fp_ctx* ctx, fp_metrics* out, int* lam, long long* sv, long long* rv, long long* sm,
                  long long* rm, long long* sv_msv);

/* fp_eval without allocating: no counterpart in the reference. For callers
 * that cannot allocate device memory (code under HIP graph capture, serving
 * loops that must not call hipMalloc). The same six metrics and per-part
 * vectors, written to DEVICE memory on the context's stream, with no host
 * synchronisation: d_out[6 + 5K] = tv, cn, tm, msm, msv, msrv, then sv[K],
 * rv[K], sm[K], rm[K], sv_msv[K] (as fp_eval's outputs); d_lam (nullable,
 * device, noe ints) = lambda per net. The kernels' scratch is a static device
 * array, one per GPU, held until the process ends: K <= 1024 (larger K:
 * FP_ERR_UNSUPPORTED). Every call on a GPU shares that GPU's array, so calls
 * that overlap on one GPU (contexts on different streams or threads) race;
 * calls in sequence on one stream do not. The context must already hold a
 * state (fp_set_partition / fp_set_state, which allocate). */
fp_status fp_eval_noalloc(fp_ctx* ctx, long long* d_out, int* d_lam);

/* The part-connection matrix of the resident state (TM/partition.cpp:190-235,
 * the bookkeeping of TM and MSM): pc[a*K + b] = number of nets whose source is
 * in part a and that have pins in part b != a (pc[a*K + a] = 0); msg[a] = the
 * number of b with pc[a*K + b] > 0, the messages part a sends. Either pointer
 * may be NULL. Needs K <= 4096 and net sources that are pins of their nets. */
fp_status fp_get_connections(fp_ctx* ctx, int* pc, int* msg);

/* ---- batched refinement ------------------------------------------------- */
/* A batch refines many resident hypergraphs of uneven size, K and objective
 * together. Members are spread over lanes by pass cost; one pass of every
 * member is captured once into a single HIP graph whose lanes are parallel
 * branches, and a batch pass is one launch of that graph. Every member
 * ends exactly where fp_refine(ctx, metric, passes, sched) alone would leave
 * it. The reference has no counterpart (it refines one hypergraph per
 * process). */
typedef struct fp_batch fp_batch;
fp_status fp_batch_create(fp_batch** batch, int device);
void fp_batch_destroy(fp_batch* batch);
const char* fp_batch_last_error(const fp_batch* batch);
/* Adds a context that holds a hypergraph and a state, with the objective and
 * schedule it is refined under (sched may be NULL for CN). The batch borrows
 * the context; it must outlive the batch's use of it. */
fp_status fp_batch_add(fp_batch* batch, fp_ctx* ctx, fp_metric metric, const fp_schedule* sched);
int fp_batch_size(const fp_batch* batch);
/* Number of lanes, the parallel branches of the pass graph (1..64, default
 * 8). */
fp_status fp_batch_set_lanes(fp_batch* batch, int lanes);
/* `passes` passes of every member. The first tune_passes of them run eagerly
 * with a different candidate launch shape each (wavefronts per evaluation
 * block, evaluation grid cap), timed per member; each member then keeps its
 * fastest shape for the remaining passes, which run as one batch-graph launch
 * per pass. device_ms (nullable) receives the device time of the whole call. */
fp_status fp_batch_refine(fp_batch* batch, int passes, int tune_passes, double* device_ms);
/* Member i after fp_batch_refine: applied moves, the launch shape it kept and
 * its fastest tuning pass (ms). Any pointer may be NULL. */
fp_status fp_batch_member(const fp_batch* batch, int i, long long* moves, int* eval_waves, int* eval_grid,
                          double* tuned_pass_ms);

/* ---- serving ------------------------------------------------------------ */
/* A request loop over resident hypergraphs. Each registered hypergraph is
 * uploaded once into a context of its own (CSR views and connectivity state
 * stay in HBM). Worker threads take requests in arrival order, skipping one
 * whose hypergraph is busy, and admit a request only while the HBM footprint
 * of the requests in flight plus its own stays within the budget; a request
 * sets its state from its part vector, refines and evaluates on its
 * hypergraph's stream, so requests on different hypergraphs overlap on the
 * device and new ones start as slots and memory free up. A request's result
 * depends only on the request. The reference has no counterpart. */
typedef struct fp_server fp_server;
typedef struct {
  long long id;
  int graph, metric, K, passes;
  fp_status status;
  long long moves;
  fp_metrics metrics;          /* the six metrics after refinement */
  double submit_ms, start_ms, end_ms;  /* since fp_server_create */
  long long footprint;         /* HBM bytes the admission charged */
} fp_request_info;
/* hbm_budget: bytes the requests in flight may hold together; workers: 1..256 */
fp_status fp_server_create(fp_server** server, int device, long long hbm_budget, int workers);
/* Stops the workers after their current requests; queued requests are
 * dropped, so wait for every submitted request first. */
void fp_server_destroy(fp_server* server);
const char* fp_server_last_error(const fp_server* server);
/* Up to max_replicas contexts (1..64, default 1) per hypergraph: a request
 * for a hypergraph whose contexts are all busy gets a new one, uploaded from
 * the server's host copy, instead of waiting. Replicas stay resident until
 * fp_server_destroy, outside the budget of the requests in flight. No
 * counterpart in the reference. */
fp_status fp_server_set_replicas(fp_server* server, int max_replicas);
/* Registers (uploads) a hypergraph: net->pins CSR, net sources (NULL: none),
 * vertex weights (NULL: unit); the server keeps a host copy for replicas.
 * graph_id receives its index. */
fp_status fp_server_add_graph(fp_server* server, int nov, int noe, const unsigned* net_ptr, const unsigned* net_pins,
                              const int* src, const int* vwgt, int* graph_id);
/* Queues a request: `passes` passes of `metric` under `sched` (NULL for CN) on
 * hypergraph graph_id from the part vector part[nov] (copied) with K parts.
 * FP_ERR_BUDGET if the request alone exceeds the budget. Does not block. */
fp_status fp_server_submit(fp_server* server, int graph_id, fp_metric metric, int K, const int* part,
                           const fp_schedule* sched, int passes, long long* request_id);
/* Blocks until the request is done and releases it: its info, and its final
 * part vector into part_out[nov] (nullable). Returns the request's status. */
fp_status fp_server_wait(fp_server* server, long long request_id, fp_request_info* info, int* part_out);
fp_status fp_server_stats(fp_server* server, long long* completed, int* peak_inflight, long long* peak_bytes);

/* Event timing of the dominant refinement kernel in the last fp_refine (the
 * evaluation: TV k_tv_eval, MSV/MSRV k_vol_eval, TM/MSM k_conn_eval; CN has
 * none): launches, total device ms of those launches (HIP events around each
 * launch, on the context stream). */
fp_status fp_last_kernel_timing(fp_ctx* ctx, long long* launches, double* total_ms);
/* Enables (1) or disables (0) per-launch event timing of fp_refine's
 * dominant kernel (adds events around every launch). */
fp_status fp_set_kernel_timing(fp_ctx* ctx, int enable);

#ifdef __cplusplus
}
#endif
#endif
