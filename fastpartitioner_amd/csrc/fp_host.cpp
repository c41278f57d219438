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

int fp_read_patoh(const char* path, int nov_cap, int noe_cap, long long pins_cap, int* nov, int* noe,
                  long long* pins, unsigned* net_ptr, unsigned* net_pins) {
  FILE* f = std::fopen(path, "r");
  if (!f) return FP_ERR_IO;
  int base, hv, he;
  long long hp;
  if (std::fscanf(f, "%d %d %d %lld", &base, &hv, &he, &hp) != 4) { std::fclose(f); return FP_ERR_IO; }
  // rest of the header line is ignored (shared/hypergraph_rw.cpp:138)
  int ch;
  while ((ch = std::fgetc(f)) != EOF && ch != '\n') {}
  const int salt = base == 0 ? 0 : -1;
  long long v_count = 0;
  int n_count = 0;
  net_ptr[0] = 0;
  std::string line;
  char buf[1 << 16];
  while (n_count < he) {
    line.clear();
    bool got = false;
    while (std::fgets(buf, sizeof buf, f)) {
      got = true;
      line += buf;
      if (!line.empty() && line.back() == '\n') break;
    }
    if (!got) break;
    const char* s = line.c_str();
    char* e;
    int local = 0;
    for (;;) {
      while (*s == ' ' || *s == '\t' || *s == '\r' || *s == '\n') s++;
      if (!*s) break;
      long x = std::strtol(s, &e, 10);
      if (e == s) break;
      s = e;
      long v = x + salt;
      if (v < 0 || v >= hv || v_count >= pins_cap) { std::fclose(f); return FP_ERR_ARG; }
      net_pins[v_count++] = (unsigned)v;
      local++;
    }
    if (local == 0) continue;  // blank lines are skipped (shared/hypergraph_rw.cpp:152-156)
    if (n_count >= noe_cap) { std::fclose(f); return FP_ERR_ARG; }
    net_ptr[n_count + 1] = net_ptr[n_count] + (unsigned)local;
    n_count++;
  }
  std::fclose(f);
  if (v_count != hp) return FP_ERR_ARG;  // pin-count mismatch fails the read (:167-173)
  if (hv > nov_cap) return FP_ERR_ARG;
  *nov = hv;
  *noe = n_count;  // noe is corrected to the nets actually read (:159-165)
  *pins = v_count;
  return FP_OK;
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
