// fp_host.cpp -- host-side helpers of the FastPartitioner MI355X engine:
// seeded synthetic column-net hypergraph generators, the net->pins /
// vertex->nets CSR transpose, PaToH-like text I/O and the streaming initial
// partition that precedes refinement. Pure C++ (no HIP); built into
// libfastpart_host.so. Declarations: include/fastpart_host.h.
#include "../../include/fastpart_host.h"

#include <algorithm>
#include <cfloat>
#include <cstdint>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include <omp.h>

namespace {

// splitmix64: one independent stream per (seed, row)
inline uint64_t mix(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
inline uint64_t stream_seed(uint64_t seed, uint64_t row, uint64_t salt) {
  uint64_t s = seed * 0xD1B54A32D192ED03ULL ^ (row + 1) * 0x8CB92BA72F3D8DD7ULL ^ salt;
  mix(s);
  return s;
}
inline uint64_t below(uint64_t& s, uint64_t m) {  // uniform in [0, m)
  return (uint64_t)(((unsigned __int128)mix(s) * m) >> 64);
}
inline double unit(uint64_t& s) { return (double)(mix(s) >> 11) * (1.0 / 9007199254740992.0); }

struct DegreeLaw {
  const fp_gen_params* p;
  std::vector<double> cdf;  // zipf CDF over [lo, hi]
  explicit DegreeLaw(const fp_gen_params* prm) : p(prm) {
    if (p->degree_mode == FP_DEG_ZIPF) {
      int lo = p->deg_lo, hi = p->deg_hi;
      cdf.resize((size_t)(hi - lo + 1));
      double acc = 0;
      for (int d = lo; d <= hi; d++) {
        acc += std::pow((double)d, -p->zipf_alpha);
        cdf[(size_t)(d - lo)] = acc;
      }
      for (double& c : cdf) c /= acc;
    }
  }
  // number of off-diagonal pins of row i (stream s positioned at row start)
  int offdiag(uint64_t& s) const {
    switch (p->degree_mode) {
      case FP_DEG_FIXED: return p->deg_lo;
      case FP_DEG_UNIFORM: return p->deg_lo + (int)below(s, (uint64_t)(p->deg_hi - p->deg_lo + 1));
      default: {
        double u = unit(s);
        size_t k = (size_t)(std::lower_bound(cdf.begin(), cdf.end(), u) - cdf.begin());
        if (k >= cdf.size()) k = cdf.size() - 1;
        return p->deg_lo + (int)k - 1;  // total degree includes the diagonal
      }
    }
  }
};

int check_params(const fp_gen_params* p) {
  if (!p || p->n < 1) return FP_ERR_ARG;
  if (p->degree_mode == FP_DEG_ZIPF && (p->deg_lo < 1 || p->deg_hi < p->deg_lo || !(p->zipf_alpha > 0)))
    return FP_ERR_ARG;
  if (p->degree_mode != FP_DEG_ZIPF && (p->deg_lo < 0 || (p->degree_mode == FP_DEG_UNIFORM && p->deg_hi < p->deg_lo)))
    return FP_ERR_ARG;
  return FP_OK;
}

}  // namespace

extern "C" {

int fp_gen_colnet_sizes(const fp_gen_params* p, unsigned* net_ptr, long long* pins_out) {
  int rc = check_params(p);
  if (rc) return rc;
  DegreeLaw law(p);
  const int n = p->n;
  std::vector<unsigned> deg((size_t)n);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; i++) {
    uint64_t s = stream_seed(p->seed, (uint64_t)i, 0);
    int od = law.offdiag(s);
    if (od > n - 1) od = n - 1;
    if (od < 0) od = 0;
    deg[(size_t)i] = (unsigned)(od + 1);
  }
  unsigned long long acc = 0;
  net_ptr[0] = 0;
  for (int i = 0; i < n; i++) {
    acc += deg[(size_t)i];
    if (acc > 0xFFFFFFFFULL) return FP_ERR_UNSUPPORTED;
    net_ptr[i + 1] = (unsigned)acc;
  }
  if (pins_out) *pins_out = (long long)acc;
  return FP_OK;
}

int fp_gen_colnet_pins(const fp_gen_params* p, const unsigned* net_ptr, unsigned* net_pins, int* src) {
  int rc = check_params(p);
  if (rc) return rc;
  DegreeLaw law(p);
  const int n = p->n;
  int bad = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(| : bad)
  for (int i = 0; i < n; i++) {
    uint64_t s = stream_seed(p->seed, (uint64_t)i, 0);
    int od = law.offdiag(s);
    if (od > n - 1) od = n - 1;
    if (od < 0) od = 0;
    if ((unsigned)(od + 1) != net_ptr[i + 1] - net_ptr[i]) { bad = 1; continue; }
    unsigned* row = net_pins + net_ptr[i];
    // distinct uniform off-diagonal columns: draw from [0, n-1) and skip i
    int have = 0;
    if (od == n - 1) {
      for (int c = 0; c < n; c++)
        if (c != i) row[have++] = (unsigned)c;
    } else if (od <= 64) {
      while (have < od) {
        unsigned c = (unsigned)below(s, (uint64_t)(n - 1));
        if (c >= (unsigned)i) c++;
        bool dup = false;
        for (int t = 0; t < have; t++)
          if (row[t] == c) { dup = true; break; }
        if (!dup) row[have++] = c;
      }
    } else {
      std::vector<unsigned> buf;
      buf.reserve((size_t)od);
      while ((int)buf.size() < od) {
        size_t need = (size_t)od - buf.size();
        for (size_t t = 0; t < need; t++) {
          unsigned c = (unsigned)below(s, (uint64_t)(n - 1));
          if (c >= (unsigned)i) c++;
          buf.push_back(c);
        }
        std::sort(buf.begin(), buf.end());
        buf.erase(std::unique(buf.begin(), buf.end()), buf.end());
      }
      for (int t = 0; t < od; t++) row[t] = buf[(size_t)t];
      have = od;
    }
    row[have++] = (unsigned)i;  // the diagonal: always a pin, and the net's source
    std::sort(row, row + have);
    if (src) src[i] = i;
  }
  return bad ? FP_ERR_ARG : FP_OK;
}

int fp_gen_weights(int n, unsigned long long seed, int lo, int hi, int* vwgt) {
  if (n < 0 || lo < 1 || hi < lo || !vwgt) return FP_ERR_ARG;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; i++) {
    uint64_t s = stream_seed(seed, (uint64_t)i, 0x5EEDULL);
    vwgt[i] = lo + (int)below(s, (uint64_t)(hi - lo + 1));
  }
  return FP_OK;
}

int fp_csr_transpose(int nov, int noe, const unsigned* net_ptr, const unsigned* net_pins, unsigned* vtx_ptr,
                     unsigned* vtx_nets) {
  if (nov < 0 || noe < 0 || !net_ptr || !vtx_ptr) return FP_ERR_ARG;
  // counting sort over pins: vertex->nets lists come out in ascending net
  // order, as the reference's transpose produces them (TV/tv.cpp:79-100)
  std::memset(vtx_ptr, 0, sizeof(unsigned) * (size_t)(nov + 1));
  for (int i = 0; i < noe; i++)
    for (unsigned j = net_ptr[i]; j < net_ptr[i + 1]; j++) {
      if (net_pins[j] >= (unsigned)nov) return FP_ERR_ARG;
      vtx_ptr[net_pins[j] + 1]++;
    }
  for (int v = 0; v < nov; v++) vtx_ptr[v + 1] += vtx_ptr[v];
  std::vector<unsigned> fill(vtx_ptr, vtx_ptr + nov);
  for (int i = 0; i < noe; i++)
    for (unsigned j = net_ptr[i]; j < net_ptr[i + 1]; j++) vtx_nets[fill[net_pins[j]]++] = (unsigned)i;
  return FP_OK;
}

// Streaming initial partitions (only the part vector is produced; the
// connectivity lists a stream leaves are rebuilt on the device by
// fp_set_partition in first-appearance order).
//
//   FP_INIT_LDG_BANDS  LDG2 of TV/partition.cpp:89-145 as the TV stream uses it
//                      (TV/partition.cpp:207-264)
//   FP_INIT_LDG        band-free LDG of MGV/partition.cpp:28-65 (the MSV and
//                      MSRV streams, MGV/partition.cpp:122-189)
//   FP_INIT_CN_LDG     the CN stream (CN/partition.cpp:132-221): only nets that
//                      are not cut (one part so far) vote, no bands
//   FP_INIT_MSG_GREEDY the TM/MSM greedy (TM/partition.cpp:87-250; the MSM
//                      stream is the same, MSM/partition.cpp:146-308): the part
//                      among those lighter than C that saves the most messages
int fp_initial_partition(int mode, int nov, int noe, const unsigned* net_ptr, const unsigned* vtx_ptr,
                         const unsigned* vtx_nets, const int* ordering, const int* vwgt, int K, double C,
                         long long max_pw, long long min_pw, int* part) {
  if (nov < 1 || noe < 0 || K < 1 || !net_ptr || !vtx_ptr || !vtx_nets || !part) return FP_ERR_ARG;
  if (mode == FP_INIT_MSG_GREEDY) return FP_ERR_ARG;  // needs the net sources: fp_initial_partition_msg
  if (mode != FP_INIT_LDG_BANDS && mode != FP_INIT_LDG && mode != FP_INIT_CN_LDG) return FP_ERR_ARG;
  const bool cut_free = mode == FP_INIT_CN_LDG;
  const long long pins = net_ptr[noe];
  std::vector<int> lam((size_t)noe, 0), lp((size_t)pins, -1);
  std::vector<long long> sizes((size_t)K, 0);
  std::vector<int> seen((size_t)K, -1), cnt((size_t)K, 0), active;
  active.reserve((size_t)K);
  for (int v = 0; v < nov; v++) part[v] = -1;
  for (int z = 0; z < nov; z++) {
    int v = ordering ? ordering[z] : z;
    if (v < 0 || v >= nov || part[v] != -1) return FP_ERR_ARG;
    active.clear();
    for (unsigned i = vtx_ptr[v]; i < vtx_ptr[v + 1]; i++) {
      int n = (int)vtx_nets[i];
      const int* L = &lp[net_ptr[n]];
      const int l = lam[(size_t)n];
      if (cut_free && l != 1) continue;  // CN/partition.cpp:140-143: a cut net has no say
      for (int j = 0; j < l; j++) {
        int q = L[j];
        if (seen[(size_t)q] != v) {
          seen[(size_t)q] = v;
          cnt[(size_t)q] = 1;
          active.push_back(q);
        } else {
          cnt[(size_t)q]++;
        }
      }
    }
    double best = 0;
    int d = -1;
    for (int q : active) {
      double penalty = 1 - (1.0 * (double)sizes[(size_t)q] / C);
      double score = (double)cnt[(size_t)q] * penalty;
      if (score > 0 && (d == -1 || score > best || (score == best && sizes[(size_t)q] < sizes[(size_t)d]))) {
        best = score;
        d = q;
      }
    }
    if (d == -1) {
      long long m = LLONG_MAX;
      for (int q = 0; q < K; q++)
        if (sizes[(size_t)q] < m) { m = sizes[(size_t)q]; d = q; }
    }
    if (mode == FP_INIT_LDG_BANDS && sizes[(size_t)d] > max_pw) return FP_ERR_UNSUPPORTED;  // reference: undefined
    (void)min_pw;  // the lower band never applies to an unassigned vertex
    part[v] = d;
    sizes[(size_t)d] += vwgt ? vwgt[v] : 1;
    for (unsigned i = vtx_ptr[v]; i < vtx_ptr[v + 1]; i++) {
      int n = (int)vtx_nets[i];
      int* L = &lp[net_ptr[n]];
      int l = lam[(size_t)n], j = 0;
      while (j < l && L[j] != d) j++;
      if (j == l) { L[l] = d; lam[(size_t)n] = l + 1; }
    }
  }
  return FP_OK;
}

namespace {

// K counters of up to 2^levels - 1 each, bit-sliced over K-bit rows: adding a
// row of K bits costs W = ceil(K/64) words per level the carry reaches.
struct SlicedCounts {
  int W = 0, levels = 0;
  std::vector<uint64_t> c;  // levels x W
  void reset(int w, int lv) {
    W = w;
    levels = lv;
    c.assign((size_t)lv * (size_t)w, 0ull);
  }
  void add(const uint64_t* row) {
    for (int w = 0; w < W; w++) {
      uint64_t carry = row[w];
      for (int b = 0; b < levels && carry; b++) {
        uint64_t& x = c[(size_t)b * (size_t)W + (size_t)w];
        const uint64_t t = x & carry;
        x ^= carry;
        carry = t;
      }
    }
  }
  int get(int k) const {
    int v = 0;
    const int w = k >> 6, sh = k & 63;
    for (int b = 0; b < levels; b++) v |= (int)((c[(size_t)b * (size_t)W + (size_t)w] >> sh) & 1ull) << b;
    return v;
  }
};

}  // namespace

// The TM/MSM greedy stream (TM/partition.cpp:87-250). The message bookkeeping
// of the reference -- part_connections[a][b], the number of nets with their
// source in a and a pin in b != a, and part_messages[a] -- is kept as counts
// plus the rows and columns of its nonzero pattern as bitmaps, so that a
// vertex's savings for all K parts are sums of bitmap rows:
//   saved[k] = #{source parts s of v's nets placed so far : pc[s][k] > 0}
//            + sum over v's own nets (v is their source) of
//              #{parts q in the net's list : pc[k][q] > 0}
// The vertex goes to the part lighter than C with the most savings, ties to
// the smaller size, then the lower index; with none lighter than C, to the
// lowest-index part of minimum size.
int fp_initial_partition_msg(int nov, int noe, const unsigned* net_ptr, const unsigned* vtx_ptr,
                             const unsigned* vtx_nets, const int* src, const int* ordering, const int* vwgt, int K,
                             double C, int* part) {
  if (nov < 1 || noe < 0 || K < 1 || !net_ptr || !vtx_ptr || !vtx_nets || !src || !part) return FP_ERR_ARG;
  if (K > 4096) return FP_ERR_UNSUPPORTED;  // a K x K count matrix
  const long long pins = net_ptr[noe];
  const int W = (K + 63) / 64;
  std::vector<int> pc((size_t)K * (size_t)K, 0);
  std::vector<uint64_t> rows((size_t)K * (size_t)W, 0ull), cols((size_t)K * (size_t)W, 0ull);
  std::vector<int> lam((size_t)noe, 0), lp((size_t)pins, -1), lc((size_t)pins, 0), src_part((size_t)noe, -1);
  std::vector<long long> sizes((size_t)K, 0);
  std::vector<int> smark((size_t)K, -1);
  std::vector<int> sparts, own_nets;
  SlicedCounts cnt;
  for (int v = 0; v < nov; v++) part[v] = -1;
  auto set_pc = [&](int a, int b) {  // pc[a][b]++ (a != b)
    int& x = pc[(size_t)a * (size_t)K + (size_t)b];
    if (x++ == 0) {
      rows[(size_t)a * (size_t)W + (size_t)(b >> 6)] |= 1ull << (b & 63);
      cols[(size_t)b * (size_t)W + (size_t)(a >> 6)] |= 1ull << (a & 63);
    }
  };
  for (int z = 0; z < nov; z++) {
    const int v = ordering ? ordering[z] : z;
    if (v < 0 || v >= nov || part[v] != -1) return FP_ERR_ARG;
    sparts.clear();
    own_nets.clear();
    int terms = 0;
    for (unsigned i = vtx_ptr[v]; i < vtx_ptr[v + 1]; i++) {
      const int n = (int)vtx_nets[i];
      if (src[n] == v) {
        own_nets.push_back(n);
        terms += lam[(size_t)n];
      }
      const int sp = src_part[(size_t)n];
      if (sp >= 0 && smark[(size_t)sp] != v) {
        smark[(size_t)sp] = v;
        sparts.push_back(sp);
        terms++;
      }
    }
    int levels = 1;
    while ((1 << levels) <= terms) levels++;
    cnt.reset(W, levels);
    for (int sp : sparts) cnt.add(&rows[(size_t)sp * (size_t)W]);
    for (int n : own_nets) {
      const int* L = &lp[net_ptr[n]];
      for (int j = 0; j < lam[(size_t)n]; j++) cnt.add(&cols[(size_t)L[j] * (size_t)W]);
    }
    int d = -1, best = -1;
    for (int k = 0; k < K; k++) {
      if (!((double)sizes[(size_t)k] < C)) continue;
      const int saved = cnt.get(k);
      if (saved > best || (saved == best && sizes[(size_t)k] < sizes[(size_t)d])) {
        best = saved;
        d = k;
      }
    }
    if (d == -1) {
      long long m = LLONG_MAX;
      for (int k = 0; k < K; k++)
        if (sizes[(size_t)k] < m) { m = sizes[(size_t)k]; d = k; }
    }
    part[v] = d;
    sizes[(size_t)d] += vwgt ? vwgt[v] : 1;
    // the vertex joins its nets' lists, then the messages follow
    // (TM/partition.cpp:175-235): a net whose source is placed gains a
    // message to d when d is new to it; v's own nets send from d to every
    // other part in them
    for (unsigned i = vtx_ptr[v]; i < vtx_ptr[v + 1]; i++) {
      const int n = (int)vtx_nets[i];
      int* L = &lp[net_ptr[n]];
      int* Lc = &lc[net_ptr[n]];
      const int l = lam[(size_t)n];
      int j = 0;
      while (j < l && L[j] != d) j++;
      const bool fresh = j == l;
      if (fresh) {
        L[l] = d;
        Lc[l] = 1;
        lam[(size_t)n] = l + 1;
      } else {
        Lc[j]++;
      }
      const int sp = src_part[(size_t)n];
      if (sp >= 0 && sp != d && (fresh || pc[(size_t)sp * (size_t)K + (size_t)d] == 0)) set_pc(sp, d);
    }
    for (int n : own_nets) {
      src_part[(size_t)n] = d;
      const int* L = &lp[net_ptr[n]];
      for (int j = 0; j < lam[(size_t)n]; j++)
        if (L[j] != d) set_pc(d, L[j]);
    }
  }
  return FP_OK;
}

double fp_ref_capacity(long long total_weight, int K, double eps) {
  // TV/partition.cpp:159-160: C = (nov / K) * 1.05 with integer division first
  return (double)(total_weight / K) * (1.0 + eps);
}
long long fp_ref_max_pw(long long total_weight, int K) {
  // TV/partition.cpp:95: int max_pw = (nov/K) + ((nov/K)*0.1)
  long long a = total_weight / K;
  return (long long)((double)a + (double)a * 0.1);
}
long long fp_ref_min_pw(long long total_weight, int K) {
  // TV/partition.cpp:96: int min_pw = (nov/K) - ((nov/K)*0.1)
  long long a = total_weight / K;
  return (long long)((double)a - (double)a * 0.1);
}
long long fp_balance_cap(long long total_weight, int K, double eps) {
  // (1 + eps) * ceil(W / K), rounded down
  long long avg = (total_weight + K - 1) / K;
  return (long long)std::floor((1.0 + eps) * (double)avg);
}

// ---- PaToH-like text I/O (format of shared/hypergraph_rw.cpp:97-183) -----

int fp_read_patoh_header(const char* path, int* index_base, int* nov, int* noe, long long* pins) {
  FILE* f = std::fopen(path, "r");
  if (!f) return FP_ERR_IO;
  int ok = std::fscanf(f, "%d %d %d %lld", index_base, nov, noe, pins) == 4;
  std::fclose(f);
  return ok ? FP_OK : FP_ERR_IO;
}

namespace {

bool slurp(const char* path, std::vector<char>& buf) {
  FILE* f = std::fopen(path, "rb");
  if (!f) return false;
  std::fseek(f, 0, SEEK_END);
  const long sz = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  if (sz < 0) { std::fclose(f); return false; }
  buf.resize((size_t)sz + 1);
  const size_t got = std::fread(buf.data(), 1, (size_t)sz, f);
  std::fclose(f);
  buf.resize(got + 1);
  buf[got] = '\0';
  return true;
}

inline bool is_space(char c) { return c == ' ' || c == '\t' || c == '\r' || c == '\n' || c == '\v' || c == '\f'; }

}  // namespace

// PaToH-like text, parsed in parallel with the rules of read_patohlike
// (shared/hypergraph_rw.cpp:97-183): header "index_base nov noe pins" (the
// rest of its line ignored); then the first noe non-blank lines are the nets
// (blank lines skipped), ids minus 1 when the base is not 0; noe becomes the
// number of nets read; a pin count other than the header's fails. Lines are
// found, then their tokens counted, then parsed into place, each step over
// all threads.
int fp_read_patoh(const char* path, int nov_cap, int noe_cap, long long pins_cap, int* nov, int* noe,
                  long long* pins, unsigned* net_ptr, unsigned* net_pins) {
  std::vector<char> buf;
  if (!slurp(path, buf)) return FP_ERR_IO;
  const char* text = buf.data();
  const size_t len = buf.size() - 1;
  int base, hv, he;
  long long hp;
  int hdr_end = 0;
  if (std::sscanf(text, "%d %d %d %lld%n", &base, &hv, &he, &hp, &hdr_end) != 4) return FP_ERR_IO;
  size_t pos = (size_t)hdr_end;
  while (pos < len && text[pos] != '\n') pos++;  // rest of the header line (:138)
  if (pos < len) pos++;
  // line starts: per-thread newline counts over equal chunks, then positions
  const int T = std::max(1, omp_get_max_threads());
  const size_t body = len - pos;
  std::vector<size_t> nl((size_t)T + 1, 0);
#pragma omp parallel num_threads(T)
  {
    const int t = omp_get_thread_num();
    const size_t a = pos + body * (size_t)t / (size_t)T, b = pos + body * (size_t)(t + 1) / (size_t)T;
    size_t c = 0;
    for (size_t i = a; i < b; i++) c += text[i] == '\n';
    nl[(size_t)t + 1] = c;
  }
  for (int t = 0; t < T; t++) nl[(size_t)t + 1] += nl[(size_t)t];
  const size_t nlines = nl[(size_t)T] + 1;  // the last line may lack a newline
  std::vector<size_t> start(nlines + 1);
  start[0] = pos;
#pragma omp parallel num_threads(T)
  {
    const int t = omp_get_thread_num();
    const size_t a = pos + body * (size_t)t / (size_t)T, b = pos + body * (size_t)(t + 1) / (size_t)T;
    size_t k = nl[(size_t)t];
    for (size_t i = a; i < b; i++)
      if (text[i] == '\n') start[++k] = i + 1;
  }
  start[nlines] = len + 1;
  // tokens per line
  std::vector<int> tok(nlines, 0);
#pragma omp parallel for num_threads(T) schedule(dynamic, 4096)
  for (long long l = 0; l < (long long)nlines; l++) {
    const size_t e = std::min(start[(size_t)l + 1], len);
    int c = 0;
    bool in = false;
    for (size_t i = start[(size_t)l]; i < e; i++) {
      const bool sp = is_space(text[i]);
      if (!sp && !in) c++;
      in = !sp;
    }
    tok[(size_t)l] = c;
  }
  // the first `he` non-blank lines are the nets
  std::vector<size_t> net_line;
  net_line.reserve((size_t)std::max(he, 0));
  long long total = 0;
  for (size_t l = 0; l < nlines && (int)net_line.size() < he; l++) {
    if (tok[l] == 0) continue;
    if ((int)net_line.size() >= noe_cap) return FP_ERR_ARG;
    net_ptr[net_line.size()] = (unsigned)total;
    total += tok[l];
    if (total > pins_cap || total >= (1ll << 32)) return FP_ERR_ARG;
    net_line.push_back(l);
  }
  const int n_read = (int)net_line.size();
  net_ptr[n_read] = (unsigned)total;
  if (total != hp) return FP_ERR_ARG;  // pin-count mismatch fails the read (:167-173)
  if (hv > nov_cap) return FP_ERR_ARG;
  const long salt = base == 0 ? 0 : -1;
  int bad = 0;
#pragma omp parallel for num_threads(T) schedule(dynamic, 1024) reduction(| : bad)
  for (int n = 0; n < n_read; n++) {
    const size_t l = net_line[(size_t)n];
    const char* q = text + start[l];
    const char* e = text + std::min(start[l + 1], len);
    unsigned k = net_ptr[n];
    while (q < e) {
      while (q < e && is_space(*q)) q++;
      if (q >= e) break;
      char* endp;
      const long x = std::strtol(q, &endp, 10);
      if (endp == q) { bad = 1; break; }
      q = endp;
      const long v = x + salt;
      if (v < 0 || v >= hv) { bad = 1; break; }
      net_pins[k++] = (unsigned)v;
    }
  }
  if (bad) return FP_ERR_ARG;
  *nov = hv;
  *noe = n_read;  // noe is corrected to the nets actually read (:159-165)
  *pins = total;
  return FP_OK;
}

// ---- the .plbin cache (shared/hypergraph_rw.cpp:185-240) ------------------
// u32 noe, u32 nov, i64 pins, then per net: u32 end offset, u32 pins[...].

int fp_write_plbin(const char* path, int nov, int noe, const unsigned* net_ptr, const unsigned* net_pins) {
  if (!path || nov < 0 || noe < 0 || !net_ptr) return FP_ERR_ARG;
  FILE* f = std::fopen(path, "wb");
  if (!f) return FP_ERR_IO;
  const unsigned un = (unsigned)noe, uv = (unsigned)nov;
  const long long pins = net_ptr[noe];
  bool ok = std::fwrite(&un, 4, 1, f) == 1 && std::fwrite(&uv, 4, 1, f) == 1 && std::fwrite(&pins, 8, 1, f) == 1;
  std::vector<unsigned> rec;
  for (int n = 0; n < noe && ok; n++) {
    rec.assign(1, net_ptr[n + 1]);
    rec.insert(rec.end(), net_pins + net_ptr[n], net_pins + net_ptr[n + 1]);
    ok = std::fwrite(rec.data(), 4, rec.size(), f) == rec.size();
  }
  ok = std::fclose(f) == 0 && ok;
  return ok ? FP_OK : FP_ERR_IO;
}

int fp_read_plbin_header(const char* path, int* nov, int* noe, long long* pins) {
  FILE* f = std::fopen(path, "rb");
  if (!f) return FP_ERR_IO;
  unsigned un = 0, uv = 0;
  long long p = 0;
  const bool ok = std::fread(&un, 4, 1, f) == 1 && std::fread(&uv, 4, 1, f) == 1 && std::fread(&p, 8, 1, f) == 1;
  std::fclose(f);
  if (!ok) return FP_ERR_IO;
  *noe = (int)un;
  *nov = (int)uv;
  *pins = p;
  return FP_OK;
}

int fp_read_plbin(const char* path, int nov_cap, int noe_cap, long long pins_cap, int* nov, int* noe, long long* pins,
                  unsigned* net_ptr, unsigned* net_pins) {
  std::vector<char> buf;
  if (!slurp(path, buf) || buf.size() < 17) return FP_ERR_IO;
  const char* b = buf.data();
  unsigned un, uv;
  long long p;
  std::memcpy(&un, b, 4);
  std::memcpy(&uv, b + 4, 4);
  std::memcpy(&p, b + 8, 8);
  if ((long long)un > noe_cap || (long long)uv > nov_cap || p > pins_cap || p < 0) return FP_ERR_ARG;
  const size_t words = (buf.size() - 1 - 16) / 4;
  const unsigned* w = reinterpret_cast<const unsigned*>(b + 16);
  std::vector<unsigned> tmp;
  if (reinterpret_cast<uintptr_t>(w) % alignof(unsigned)) {
    tmp.resize(words);
    std::memcpy(tmp.data(), b + 16, words * 4);
    w = tmp.data();
  }
  size_t k = 0;
  net_ptr[0] = 0;
  for (unsigned n = 0; n < un; n++) {
    if (k >= words) return FP_ERR_IO;
    const unsigned end = w[k++];
    if (end < net_ptr[n] || (long long)end > p || k + (end - net_ptr[n]) > words) return FP_ERR_IO;
    for (unsigned j = net_ptr[n]; j < end; j++) {
      const unsigned v = w[k++];
      if (v >= uv) return FP_ERR_ARG;
      net_pins[j] = v;
    }
    net_ptr[n + 1] = end;
  }
  if ((long long)net_ptr[un] != p) return FP_ERR_IO;
  *nov = (int)uv;
  *noe = (int)un;
  *pins = p;
  return FP_OK;
}

// ---- .ws weights and sources (shared/hypergraph_rw.cpp:434-494) ------------

int fp_ws_path(const char* hg_path, char* out, int cap) {
  if (!hg_path || !out || cap < 1) return FP_ERR_ARG;
  std::string s(hg_path);
  const size_t dot = s.rfind('.');
  const std::string w = (dot == std::string::npos ? s : s.substr(0, dot)) + ".ws";
  if ((int)w.size() + 1 > cap) return FP_ERR_ARG;
  std::memcpy(out, w.c_str(), w.size() + 1);
  return FP_OK;
}

int fp_read_ws(const char* path, int noe, int* edge_w, int* src) {
  FILE* f = std::fopen(path, "r");
  if (!f) return FP_ERR_IO;
  int n = 0, w, v;
  char line[256];
  while (n < noe && std::fgets(line, sizeof line, f)) {
    if (std::sscanf(line, "%d %d", &w, &v) != 2) continue;
    if (edge_w) edge_w[n] = w;
    if (src) src[n] = v;
    n++;
  }
  std::fclose(f);
  return n == noe ? FP_OK : FP_ERR_IO;
}

int fp_write_ws(const char* path, int noe, const int* edge_w, const int* src) {
  if (!path || noe < 0 || !src) return FP_ERR_ARG;
  FILE* f = std::fopen(path, "w");
  if (!f) return FP_ERR_IO;
  for (int n = 0; n < noe; n++) std::fprintf(f, "%d %d\n", edge_w ? edge_w[n] : 1, src[n]);
  return std::fclose(f) == 0 ? FP_OK : FP_ERR_IO;
}

// ---- part vectors (CN/partition.cpp:425-433, shared/utils.cpp:592-608) -----

int fp_write_part_vector(const char* path, int nov, const int* part) {
  if (!path || nov < 0 || !part) return FP_ERR_ARG;
  FILE* f = std::fopen(path, "w");
  if (!f) return FP_ERR_IO;
  std::vector<char> out;
  out.reserve((size_t)nov * 4);
  char tmp[16];
  for (int v = 0; v < nov; v++) {
    const int k = std::snprintf(tmp, sizeof tmp, "%d\n", part[v]);
    out.insert(out.end(), tmp, tmp + k);
  }
  const bool ok = std::fwrite(out.data(), 1, out.size(), f) == out.size();
  return (std::fclose(f) == 0 && ok) ? FP_OK : FP_ERR_IO;
}

int fp_read_part_vector(const char* path, int nov, int* part) {
  std::vector<char> buf;
  if (!slurp(path, buf)) return FP_ERR_IO;
  const char* q = buf.data();
  for (int v = 0; v < nov; v++) {
    char* e;
    const long x = std::strtol(q, &e, 10);
    if (e == q) return FP_ERR_IO;
    part[v] = (int)x;
    q = e;
  }
  return FP_OK;
}

// The reference's reader with its cache: <path>.plbin when it exists, else
// the text, after which the cache is written when write_cache is set
// (read_patohlike, shared/hypergraph_rw.cpp:97-242).
int fp_read_hypergraph(const char* path, int write_cache, int nov_cap, int noe_cap, long long pins_cap, int* nov,
                       int* noe, long long* pins, unsigned* net_ptr, unsigned* net_pins, int* from_cache) {
  if (!path) return FP_ERR_ARG;
  const std::string bin = std::string(path) + ".plbin";
  FILE* probe = std::fopen(bin.c_str(), "rb");
  if (probe) {
    std::fclose(probe);
    if (from_cache) *from_cache = 1;
    return fp_read_plbin(bin.c_str(), nov_cap, noe_cap, pins_cap, nov, noe, pins, net_ptr, net_pins);
  }
  if (from_cache) *from_cache = 0;
  int st = fp_read_patoh(path, nov_cap, noe_cap, pins_cap, nov, noe, pins, net_ptr, net_pins);
  if (st == FP_OK && write_cache) st = fp_write_plbin(bin.c_str(), *nov, *noe, net_ptr, net_pins);
  return st;
}

int fp_write_patoh(const char* path, int index_base, int nov, int noe, const unsigned* net_ptr,
                   const unsigned* net_pins) {
  FILE* f = std::fopen(path, "w");
  if (!f) return FP_ERR_IO;
  std::fprintf(f, "%d %d %d %u\n", index_base, nov, noe, net_ptr[noe]);
  for (int n = 0; n < noe; n++) {
    for (unsigned j = net_ptr[n]; j < net_ptr[n + 1]; j++)
      std::fprintf(f, j + 1 < net_ptr[n + 1] ? "%u " : "%u", net_pins[j] + (unsigned)index_base);
    std::fputc('\n', f);
  }
  return std::fclose(f) == 0 ? FP_OK : FP_ERR_IO;
}

}  // extern "C"
