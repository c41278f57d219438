/* fastpart_host.h -- host-side C-ABI of the FastPartitioner MI355X engine
 * (libfastpart_host.so; pure C++, no GPU). Everything on either side of the
 * device refinement path that a caller needs to get a hypergraph and an
 * initial partition to it:
 *   - seeded synthetic column-net generators (the benchmark workloads),
 *   - the net->pins -> vertex->nets CSR transpose,
 *   - PaToH-like text reading/writing, the .plbin cache, .ws weight and
 *     source files, part-vector files,
 *   - the streaming initial partition, and the reference's balance constants.
 * All functions return an fp_status code (0 = FP_OK); none calls exit().
 */
#ifndef FASTPART_HOST_H
#define FASTPART_HOST_H
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

/* degree laws for the generated rows (nets) */
enum { FP_DEG_FIXED = 0, FP_DEG_UNIFORM = 1, FP_DEG_ZIPF = 2 };

/* Column-net hypergraph of an n x n sparse matrix: net i = row i = the
 * diagonal column i plus distinct uniform off-diagonal columns, pins sorted
 * ascending; vertex j = column j; the net source is the diagonal vertex.
 *   FP_DEG_FIXED:   deg_lo off-diagonals per row
 *   FP_DEG_UNIFORM: U{deg_lo..deg_hi} off-diagonals per row
 *   FP_DEG_ZIPF:    total row degree d in [deg_lo, deg_hi] with P(d) ~ d^-alpha
 * Every row draws from its own splitmix64 stream keyed by (seed, row), so the
 * result does not depend on the thread count. */
typedef struct {
  int n;
  int degree_mode;
  int deg_lo;
  int deg_hi;
  double zipf_alpha;
  unsigned long long seed;
} fp_gen_params;

/* pass 1: net_ptr[n+1] and the pin count */
int fp_gen_colnet_sizes(const fp_gen_params* p, unsigned* net_ptr, long long* pins);
/* pass 2: net_pins[pins] and src[n] (src may be NULL) */
int fp_gen_colnet_pins(const fp_gen_params* p, const unsigned* net_ptr, unsigned* net_pins, int* src);
/* vertex weights U{lo..hi} */
int fp_gen_weights(int n, unsigned long long seed, int lo, int hi, int* vwgt);

/* vertex->nets CSR (ascending nets per vertex) from the net->pins CSR */
int fp_csr_transpose(int nov, int noe, const unsigned* net_ptr, const unsigned* net_pins, unsigned* vtx_ptr,
                     unsigned* vtx_nets);

/* Streaming initial partition over `ordering` (NULL = identity), the state
 * each objective's refinement starts from in the reference.
 *   FP_INIT_LDG_BANDS: LDG2 of TV/partition.cpp:89-145 (stream at :207-264)
 *   FP_INIT_LDG:       band-free LDG of MGV/partition.cpp:28-65 (MSV, MSRV:
 *                      MGV/partition.cpp:122-189, MSRV/partition.cpp:123-196)
 *   FP_INIT_CN_LDG:    the CN stream, CN/partition.cpp:132-221 (votes from
 *                      nets with one part so far only)
 * vwgt NULL = unit weights; sizes grow by the vertex weight. Writes part[nov].
 * FP_INIT_MSG_GREEDY is fp_initial_partition_msg's (it needs the sources). */
enum { FP_INIT_LDG_BANDS = 0, FP_INIT_LDG = 1, FP_INIT_CN_LDG = 2, FP_INIT_MSG_GREEDY = 3 };
int fp_initial_partition(int mode, int nov, int noe, const unsigned* net_ptr, const unsigned* vtx_ptr,
                         const unsigned* vtx_nets, const int* ordering, const int* vwgt, int K, double C,
                         long long max_pw, long long min_pw, int* part);

/* The TM / MSM greedy stream (TM/partition.cpp:87-250, the MSM stream is the
 * same): each vertex goes to the part lighter than C that saves the most
 * messages given the parts placed so far (ties: smaller size, then lower
 * index; none lighter than C: lowest-index part of minimum size). src[noe]
 * = net sources. K <= 4096. Writes part[nov]. */
int fp_initial_partition_msg(int nov, int noe, const unsigned* net_ptr, const unsigned* vtx_ptr,
                             const unsigned* vtx_nets, const int* src, const int* ordering, const int* vwgt, int K,
                             double C, int* part);

/* Balance constants. With eps = 0.05 and unit weights these are exactly the
 * reference's hard-coded C (TV/partition.cpp:159-160), max_pw and min_pw
 * (TV/partition.cpp:95-96). fp_balance_cap = floor((1+eps) * ceil(W/K)). */
double fp_ref_capacity(long long total_weight, int K, double eps);
long long fp_ref_max_pw(long long total_weight, int K);
long long fp_ref_min_pw(long long total_weight, int K);
long long fp_balance_cap(long long total_weight, int K, double eps);

/* PaToH-like text (header "index_base nov noe pins", one net per line),
 * read with the rules of shared/hypergraph_rw.cpp:97-183 (the first noe
 * non-blank lines are the nets; noe is corrected to the nets read; a pin
 * count other than the header's fails), parsed on all threads. The header
 * gives the capacities to allocate. */
int fp_read_patoh_header(const char* path, int* index_base, int* nov, int* noe, long long* pins);
int fp_read_patoh(const char* path, int nov_cap, int noe_cap, long long pins_cap, int* nov, int* noe,
                  long long* pins, unsigned* net_ptr, unsigned* net_pins);
int fp_write_patoh(const char* path, int index_base, int nov, int noe, const unsigned* net_ptr,
                   const unsigned* net_pins);

/* The reference's binary cache of a text hypergraph, <file>.plbin
 * (shared/hypergraph_rw.cpp:185-240): u32 noe, u32 nov, i64 pins, then per
 * net its u32 end offset followed by its u32 pins. Files written here are
 * byte-identical to the reference's. */
int fp_write_plbin(const char* path, int nov, int noe, const unsigned* net_ptr, const unsigned* net_pins);
int fp_read_plbin_header(const char* path, int* nov, int* noe, long long* pins);
int fp_read_plbin(const char* path, int nov_cap, int noe_cap, long long pins_cap, int* nov, int* noe, long long* pins,
                  unsigned* net_ptr, unsigned* net_pins);
/* read_patohlike's behaviour (shared/hypergraph_rw.cpp:97-242): <path>.plbin
 * when it exists, else the text, then (write_cache != 0) the cache is
 * written. *from_cache (nullable) tells which was read. Capacities as for
 * fp_read_patoh (the cache's header, or the text's, gives them). */
int fp_read_hypergraph(const char* path, int write_cache, int nov_cap, int noe_cap, long long pins_cap, int* nov,
                       int* noe, long long* pins, unsigned* net_ptr, unsigned* net_pins, int* from_cache);

/* Net weights and sources, <base>.ws with one "weight source" line per net
 * (shared/hypergraph_rw.cpp:434-494; base = the hypergraph path up to its
 * last '.'). fp_read_ws reads noe lines (edge_w / src nullable);
 * fp_write_ws writes them (edge_w NULL = weight 1). */
int fp_ws_path(const char* hg_path, char* out, int cap);
int fp_read_ws(const char* path, int noe, int* edge_w, int* src);
int fp_write_ws(const char* path, int noe, const int* edge_w, const int* src);

/* Part vectors as the reference writes them: one part id per line, vertex
 * order (CN/partition.cpp:425-433 "<file>.fastpart.<K>",
 * shared/utils.cpp:592-608). */
int fp_write_part_vector(const char* path, int nov, const int* part);
int fp_read_part_vector(const char* path, int nov, int* part);

#ifdef __cplusplus
}
#endif
#endif
