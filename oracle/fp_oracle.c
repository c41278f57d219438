/* fp_oracle.c -- CPU restatement of FastPartitioner's refinement and metric
 * path. TEST INFRASTRUCTURE ONLY (the checker): used by tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg. The product
 * library never links, calls or falls back to this file.
 *
 * Plain C, scalar, single threaded. Each function cites the reference lines
 * it restates; the header documents the contract of each entry point.
 */
#include "fp_oracle.h"

#include <limits.h>
#include <stdlib.h>
#include <string.h>

/* Vertex weights of the refinements and initial partitions below (NULL =
 * unit, the reference's only case): a move changes the part sizes by the
 * vertex's weight, as the engine does. Set by or_set_vertex_weights. */
static const int* g_vwgt = NULL;
/* test hook: a planted mistake (0 = none), see or_plant */
static int g_plant = 0;
void or_plant(int code) { g_plant = code; }
#define VW(v) (g_vwgt ? g_vwgt[(v)] : 1)
void or_set_vertex_weights(const int* vwgt) { g_vwgt = vwgt; }

/* ---------------------------------------------------------------------- */
/* connectivity list edits (identical in every reference directory)        */
/* ---------------------------------------------------------------------- */

/* remove one pin of part p from net n: TV/partition.cpp:388-411 */
static void list_remove(int n, int p, const unsigned* net_ptr, int* ep_ends, int* ep_parts, int* ep_cnt) {
  int s = (int)net_ptr[n];
  int end = ep_ends[n];
  for (int j = 0; j < end; j++) {
    if (ep_parts[s + j] == p) {
      ep_cnt[s + j]--;
      if (ep_cnt[s + j] == 0) {
        ep_parts[s + j] = ep_parts[s + end - 1];
        ep_cnt[s + j] = ep_cnt[s + end - 1];
        ep_parts[s + end - 1] = -1;
        ep_cnt[s + end - 1] = 0;
        ep_ends[n]--;
      }
      return;
    }
  }
}

/* add one pin of part q to net n: TV/partition.cpp:413-435 */
static void list_add(int n, int q, const unsigned* net_ptr, int* ep_ends, int* ep_parts, int* ep_cnt) {
  int s = (int)net_ptr[n];
  int end = ep_ends[n];
  for (int j = 0; j < end; j++) {
    if (ep_parts[s + j] == q) {
      ep_cnt[s + j]++;
      return;
    }
  }
  ep_parts[s + end] = q;
  ep_cnt[s + end] = 1;
  ep_ends[n]++;
}

/* move vertex v from p to q: all removals over v's nets, then all adds
 * (TV/partition.cpp:387-436). */
static void move_vertex(int v, int p, int q, const unsigned* net_ptr, const unsigned* vtx_ptr,
                        const unsigned* vtx_nets, int* ep_ends, int* ep_parts, int* ep_cnt) {
  for (unsigned i = vtx_ptr[v]; i < vtx_ptr[v + 1]; i++) list_remove((int)vtx_nets[i], p, net_ptr, ep_ends, ep_parts, ep_cnt);
  for (unsigned i = vtx_ptr[v]; i < vtx_ptr[v + 1]; i++) list_add((int)vtx_nets[i], q, net_ptr, ep_ends, ep_parts, ep_cnt);
}

/* ---------------------------------------------------------------------- */

void or_transpose(int nov, int noe, const unsigned* net_ptr, const unsigned* net_pins, unsigned* vtx_ptr,
                  unsigned* vtx_nets) {
  /* TV/tv.cpp:79-100 */
  memset(vtx_ptr, 0, sizeof(unsigned) * (size_t)(nov + 1));
  for (int i = 0; i < noe; i++)
    for (unsigned j = net_ptr[i]; j < net_ptr[i + 1]; j++) vtx_ptr[net_pins[j] + 1]++;
  for (int i = 1; i <= nov; i++) vtx_ptr[i] += vtx_ptr[i - 1];
  for (int i = 0; i < noe; i++)
    for (unsigned j = net_ptr[i]; j < net_ptr[i + 1]; j++) vtx_nets[vtx_ptr[net_pins[j]]++] = (unsigned)i;
  for (int i = nov; i > 0; i--) vtx_ptr[i] = vtx_ptr[i - 1];
  vtx_ptr[0] = 0;
}

int or_build_state(int nov, int noe, const unsigned* net_ptr, const unsigned* vtx_ptr, const unsigned* vtx_nets,
                   const int* ordering, int K, const int* part, const int* vwgt, int* ep_ends, int* ep_parts,
                   int* ep_cnt, long long* sizes) {
  long long pins = net_ptr[noe];
  for (long long i = 0; i < pins; i++) { ep_parts[i] = -1; ep_cnt[i] = 0; }
  for (int n = 0; n < noe; n++) ep_ends[n] = 0;
  for (int k = 0; k < K; k++) sizes[k] = 0;
  for (int z = 0; z < nov; z++) {
    int v = ordering ? ordering[z] : z;
    int q = part[v];
    if (q < 0 || q >= K) return 1;
    sizes[q] += vwgt ? vwgt[v] : 1;
    for (unsigned i = vtx_ptr[v]; i < vtx_ptr[v + 1]; i++) list_add((int)vtx_nets[i], q, net_ptr, ep_ends, ep_parts, ep_cnt);
  }
  return 0;
}

/* ---------------------------------------------------------------------- */
/* metrics                                                                 */
/* ---------------------------------------------------------------------- */

int or_metrics(int nov, int noe, const unsigned* net_ptr, const unsigned* net_pins, const int* src, int K,
               const int* part, long long* out, int* lam, long long* sv, long long* rv, long long* sm,
               long long* rm, long long* sv_msv) {
  int* mark = (int*)malloc(sizeof(int) * (size_t)K);
  unsigned char* p2p = (unsigned char*)calloc((size_t)K * (size_t)K, 1);
  long long* s_send = (long long*)calloc((size_t)K, sizeof(long long));
  long long* s_recv = (long long*)calloc((size_t)K, sizeof(long long));
  long long* s_sm = (long long*)calloc((size_t)K, sizeof(long long));
  long long* s_rm = (long long*)calloc((size_t)K, sizeof(long long));
  long long* s_msv = (long long*)calloc((size_t)K, sizeof(long long));
  if (!mark || !p2p || !s_send || !s_recv || !s_sm || !s_rm || !s_msv) return 2;
  for (int k = 0; k < K; k++) mark[k] = -1;
  long long tv = 0, cn = 0;
  for (int n = 0; n < noe; n++) {
    /* lambda(n): distinct parts among the pins (CN/cn.cpp:42-50) */
    int l = 0;
    for (unsigned j = net_ptr[n]; j < net_ptr[n + 1]; j++) {
      int q = part[net_pins[j]];
      if (mark[q] != n) { mark[q] = n; l++; }
    }
    lam[n] = l;
    if (l != 0) tv += l - 1;    /* metric.cpp:62-65 */
    if (l > 1) cn++;            /* metric.cpp:8-11 */
    int s = src[n];
    if (s < 0 || s >= nov) { free(mark); free(p2p); free(s_send); free(s_recv); free(s_sm); free(s_rm); free(s_msv); return 1; }
    int sp = part[s];
    /* MGAV_Scratch send: (metric.cpp:365-366) */
    s_send[sp] += l - 1;
    /* MGV_Scratch: adds lambda-1 once per occurrence of src among the pins (metric.cpp:311-318) */
    for (unsigned j = net_ptr[n]; j < net_ptr[n + 1]; j++)
      if ((int)net_pins[j] == s) s_msv[sp] += l - 1;
    /* receive volumes and messages: each distinct part != sp (metric.cpp:369-375, TM_scratch :169-176) */
    for (unsigned j = net_ptr[n]; j < net_ptr[n + 1]; j++) {
      int q = part[net_pins[j]];
      if (mark[q] == -2 - n) continue;  /* already counted for this net */
      mark[q] = -2 - n;
      if (q == sp) continue;
      s_recv[q]++;
      unsigned char* e = &p2p[(size_t)sp * (size_t)K + (size_t)q];
      if (!*e) { *e = 1; s_sm[sp]++; s_rm[q]++; }
    }
  }
  long long tm = 0, msm = -1, msv = -1, msrv = -1;
  for (int k = 0; k < K; k++) {
    tm += s_sm[k];
    if (s_sm[k] > msm) msm = s_sm[k];             /* MGM_scratch max, metric.cpp:273-283 */
    if (s_msv[k] > msv) msv = s_msv[k];           /* MGV_Scratch max, metric.cpp:321-333 */
    if (s_send[k] + s_recv[k] > msrv) msrv = s_send[k] + s_recv[k]; /* metric.cpp:379-389 */
  }
  out[0] = tv; out[1] = cn; out[2] = tm; out[3] = msm; out[4] = msv; out[5] = msrv;
  for (int k = 0; k < K; k++) {
    if (sv) sv[k] = s_send[k];
    if (rv) rv[k] = s_recv[k];
    if (sm) sm[k] = s_sm[k];
    if (rm) rm[k] = s_rm[k];
    if (sv_msv) sv_msv[k] = s_msv[k];
  }
  free(mark); free(p2p); free(s_send); free(s_recv); free(s_sm); free(s_rm); free(s_msv);
  return 0;
}

/* ---------------------------------------------------------------------- */
/* TV                                                                      */
/* ---------------------------------------------------------------------- */

/* LDG2 decision (TV/partition.cpp:89-145). Returns the decision, or -2 when
 * the reference would index part_sizes[-1] (max band hit while the vertex is
 * still unassigned: undefined in the reference). One body serves int part
 * sizes (the reference's own) and int64 weighted sizes; the score is the same
 * double expression either way (1.0 * size converts exactly below 2^53). */
#define DEFINE_LDG2(NAME, T, T_MAX)                                                                   \
  static int NAME(int K, int apc, const int* active, const double* scores, const T* sizes, double C, \
                  T max_pw, T min_pw, int own) {                                                      \
    double max_val = 0;                                                                               \
    int max_index = -1;                                                                               \
    for (int i = 0; i < apc; i++) {                                                                   \
      int q = active[i];                                                                              \
      double penalty = 1 - (1.0 * sizes[q] / C);                                                      \
      double score = scores[q] * penalty;                                                             \
      if (score > 0) {                                                                                \
        if (max_index == -1 || score > max_val || (score == max_val && sizes[q] < sizes[max_index])) { \
          max_val = score;                                                                            \
          max_index = q;                                                                              \
        }                                                                                             \
      }                                                                                               \
    }                                                                                                 \
    if (max_index == -1) {                                                                            \
      T min_load = T_MAX;                                                                             \
      for (int i = 0; i < K; i++)                                                                     \
        if (sizes[i] < min_load) { min_load = sizes[i]; max_index = i; }                              \
    }                                                                                                 \
    int d = max_index;                                                                                \
    if (sizes[d] > max_pw) {                                                                          \
      d = own;                                                                                        \
      if (d < 0) return -2;                                                                           \
    }                                                                                                 \
    if (sizes[d] < min_pw && own != -1) d = own;                                                      \
    return d;                                                                                         \
  }
DEFINE_LDG2(ldg2, int, INT_MAX)
DEFINE_LDG2(ldg2_w, long long, LLONG_MAX)

int or_tv_ldg(int nov, int noe, const unsigned* net_ptr, const unsigned* vtx_ptr, const unsigned* vtx_nets,
              const int* ordering, int K, double C, int max_pw, int min_pw, int par_skip, int* part, int* ep_ends,
              int* ep_parts, int* ep_cnt, int* sizes, int* marker) {
  long long pins = net_ptr[noe];
  int* active = (int*)malloc(sizeof(int) * (size_t)K);
  double* scores = (double*)malloc(sizeof(double) * (size_t)K);
  if (!active || !scores) return 2;
  /* TV/partition.cpp:166-174 */
  for (long long i = 0; i < pins; i++) { ep_cnt[i] = 0; ep_parts[i] = -1; }
  for (int k = 0; k < K; k++) { marker[k] = -1; sizes[k] = 0; scores[k] = 0.0; active[k] = -1; }
  for (int v = 0; v < nov; v++) part[v] = -1;
  for (int n = 0; n < noe; n++) ep_ends[n] = 0;
  int rc = 0;
  for (int z = 0; z < nov; z++) {
    int v = ordering[z];
    int apc = 0;
    /* TV/partition.cpp:214-233 */
    for (unsigned i = vtx_ptr[v]; i < vtx_ptr[v + 1]; i++) {
      int n = (int)vtx_nets[i];
      int s = (int)net_ptr[n];
      for (int j = 0; j < ep_ends[n]; j++) {
        int q = ep_parts[s + j];
        if (marker[q] != v) { scores[q] = 1; marker[q] = v; active[apc++] = q; }
        else scores[q]++;
      }
    }
    int d = ldg2(K, apc, active, scores, sizes, C, max_pw, min_pw, part[v]);
    if (d == -2) { rc = 3; break; }
    if (par_skip && sizes[d] > max_pw) continue;   /* TV/partition.cpp:598-602 */
    sizes[d] += VW(v);
    part[v] = d;
    for (unsigned i = vtx_ptr[v]; i < vtx_ptr[v + 1]; i++) list_add((int)vtx_nets[i], d, net_ptr, ep_ends, ep_parts, ep_cnt);
  }
  free(active); free(scores);
  return rc;
}

/* Buffered round streaming initial partition: the engine's device stream.
 * The ordering is cut into rounds of R positions. Each sub-round of a round
 * decides every still unplaced vertex of the round against the state frozen
 * at the sub-round's start (lists and part sizes), with the reference's
 * stream scoring for the mode:
 *   mode 0: LDG2 over every list of the vertex's nets (TV/partition.cpp:89-145,
 *           214-233), a choice above max_pw counting as no choice (where the
 *           reference would index part_sizes[-1])
 *   mode 1: band-free LDG over every list (MGV/partition.cpp:28-65)
 *   mode 2: LDG over the lists of uncut nets only (CN/partition.cpp:138-191)
 * Candidates without a choice (no part scores above 0) fall back in turn:
 * the j-th of them in ordering order takes part (f + j) mod K, f the
 * lowest-index part of minimum size (the reference's fallback when j = 0).
 * Then per part, in ordering order, the longest prefix of its candidates
 * whose weight fits is admitted: candidate i into part d is admitted iff no
 * earlier candidate of the sub-round chose d, or the candidates into d all
 * fit (their total <= cap - size[d]), or size[d] + (weight of the earlier
 * candidates into d) + w(i) <= cap. The admitted are placed in ordering
 * order (part, size, list appends: TV/partition.cpp:240-263); the rest are
 * decided again in the next sub-round. Each sub-round places at least the
 * first candidate into each chosen part, so a round ends; R = 1 is the
 * reference's sequential stream exactly. Sizes are weights (int64). */
int or_round_ldg(int mode, int nov, int noe, const unsigned* net_ptr, const unsigned* vtx_ptr,
                 const unsigned* vtx_nets, const int* ordering, int K, double C, long long max_pw, int R,
                 long long cap, int* part, int* ep_ends, int* ep_parts, int* ep_cnt, long long* sizes,
                 long long* subrounds) {
  long long pins = net_ptr[noe];
  if (R < 1 || K < 1 || mode < 0 || mode > 2) return 1;
  const int RR = R < nov ? R : (nov > 0 ? nov : 1);
  long long* marker = (long long*)malloc(sizeof(long long) * (size_t)K);
  int* active = (int*)malloc(sizeof(int) * (size_t)K);
  double* scores = (double*)calloc((size_t)K, sizeof(double));
  long long* tot = (long long*)malloc(sizeof(long long) * (size_t)K);
  long long did = 0;
  long long* run = (long long*)malloc(sizeof(long long) * (size_t)K);
  int* cz = (int*)malloc(sizeof(int) * (size_t)RR);
  int* cd = (int*)malloc(sizeof(int) * (size_t)RR);
  unsigned char* adm = (unsigned char*)malloc((size_t)RR);
  if (!marker || !active || !scores || !tot || !run || !cz || !cd || !adm) return 2;
  for (long long i = 0; i < pins; i++) { ep_cnt[i] = 0; ep_parts[i] = -1; }
  for (int k = 0; k < K; k++) { marker[k] = -1; sizes[k] = 0; }
  for (int v = 0; v < nov; v++) part[v] = -1;
  for (int n = 0; n < noe; n++) ep_ends[n] = 0;
  *subrounds = 0;
  int rc = 0;
  const long long band = mode == 0 ? max_pw : LLONG_MAX;
  for (int r0 = 0; r0 < nov; r0 += R) {
    const int ns = R < nov - r0 ? R : nov - r0;
    for (;;) {
      int nc = 0;
      for (int z = r0; z < r0 + ns; z++) {
        const int v = ordering[z];
        if (part[v] >= 0) continue;
        int apc = 0;
        did++;
        for (unsigned i = vtx_ptr[v]; i < vtx_ptr[v + 1]; i++) {
          const int n = (int)vtx_nets[i];
          const int s = (int)net_ptr[n];
          if (mode == 2 && ep_ends[n] != 1) continue;
          for (int j = 0; j < ep_ends[n]; j++) {
            const int q = ep_parts[s + j];
            /* the marker is the decision's number, unique as v is per
             * vertex in the sequential stream */
            if (marker[q] != did) { scores[q] = 1; marker[q] = did; active[apc++] = q; }
            else scores[q]++;
          }
        }
        /* LDG2's own scan: a positive score, or -1 (no part scores) */
        double best = 0;
        int d = -1;
        for (int i = 0; i < apc; i++) {
          const int q = active[i];
          const double sc = scores[q] * (1 - (1.0 * sizes[q] / C));
          if (sc > 0 && (d == -1 || sc > best || (sc == best && sizes[q] < sizes[d]))) { best = sc; d = q; }
        }
        if (d >= 0 && sizes[d] > band) d = -1;
        cz[nc] = z;
        cd[nc] = d;
        nc++;
      }
      if (nc == 0) break;
      /* fallback candidates (no positive score, or a choice above the band):
       * the j-th of the sub-round, in ordering order, takes part (f + j) mod
       * K, f the lowest-index part of minimum size -- with one candidate the
       * reference's fallback (TV/partition.cpp:118-132) */
      {
        int nf = 0;
        for (int i = 0; i < nc; i++) nf += cd[i] < 0;
        if (nf > 0) {
          int f0 = 0;
          for (int k = 1; k < K; k++) if (sizes[k] < sizes[f0]) f0 = k;
          int j = 0;
          for (int i = 0; i < nc; i++)
            if (cd[i] < 0) cd[i] = (f0 + (g_plant == 6 ? 0 : j++)) % K;
        }
      }
      (*subrounds)++;
      /* admission on the frozen sizes */
      for (int k = 0; k < K; k++) { tot[k] = 0; run[k] = 0; }
      for (int i = 0; i < nc; i++) tot[cd[i]] += VW(ordering[cz[i]]);
      for (int i = 0; i < nc; i++) {
        const int d = cd[i];
        const long long w = VW(ordering[cz[i]]);
        const long long before = run[d];
        run[d] += w;
        adm[i] = (unsigned char)((g_plant != 4 && before == 0) || tot[d] <= cap - sizes[d] ||
                                 (g_plant == 5 ? sizes[d] + before + w < cap : sizes[d] + before + w <= cap));
      }
      int placed = 0;
      for (int ii = 0; ii < nc; ii++) {
        const int i = g_plant == 7 ? nc - 1 - ii : ii;
        if (!adm[i]) continue;
        placed++;
        const int v = ordering[cz[i]], d = cd[i];
        part[v] = d;
        sizes[d] += VW(v);
        for (unsigned j = vtx_ptr[v]; j < vtx_ptr[v + 1]; j++)
          list_add((int)vtx_nets[j], d, net_ptr, ep_ends, ep_parts, ep_cnt);
      }
      if (placed == 0) { rc = 4; goto done; }   /* only a planted mistake stalls a round */
    }
  }
done:
  free(marker); free(active); free(scores); free(tot); free(run); free(cz); free(cd); free(adm);
  return rc;
}

int or_tv_refine_serial(int nov, int noe, const unsigned* net_ptr, const unsigned* vtx_ptr,
                        const unsigned* vtx_nets, const int* ordering, int K, double C, int max_pw, int min_pw,
                        int passes, int* part, int* ep_ends, int* ep_parts, int* ep_cnt, int* sizes, int* marker) {
  (void)noe;
  int* active = (int*)malloc(sizeof(int) * (size_t)K);
  double* scores = (double*)calloc((size_t)K, sizeof(double));
  if (!active || !scores) return 2;
  for (int it = 0; it < passes; it++) {
    for (int z = 0; z < nov; z++) {
      int v = ordering[z];
      int own = part[v];
      int apc = 0;
      /* TV/partition.cpp:318-338: marker is NOT reset between vertices or passes */
      for (unsigned i = vtx_ptr[v]; i < vtx_ptr[v + 1]; i++) {
        int n = (int)vtx_nets[i];
        int s = (int)net_ptr[n];
        int end = ep_ends[n];
        for (int j = 0; j < end; j++) {
          int q = ep_parts[s + j];
          if (marker[q] != v) { scores[q] = 1; marker[q] = v; active[apc++] = q; }
          else scores[q]++;
        }
      }
      scores[own]--;   /* TV/partition.cpp:343-346 */
      int d = ldg2(K, apc, active, scores, sizes, C, max_pw, min_pw, own);
      if (d == -2) { free(active); free(scores); return 3; }
      if (d == own) continue;
      /* TV/partition.cpp:355-378 */
      int leave_gain = 0, arrival_loss = 0;
      for (unsigned i = vtx_ptr[v]; i < vtx_ptr[v + 1]; i++) {
        int n = (int)vtx_nets[i];
        int s = (int)net_ptr[n];
        int end = ep_ends[n];
        int found = 0;
        for (int j = 0; j < end; j++) {
          if (ep_parts[s + j] == own && ep_cnt[s + j] == 1) leave_gain++;
          else if (ep_parts[s + j] == d) found = 1;
        }
        if (!found) arrival_loss++;
      }
      if (leave_gain >= arrival_loss) {   /* TV/partition.cpp:380-386 */
        part[v] = d;
        sizes[own] -= VW(v);
        sizes[d] += VW(v);
        move_vertex(v, own, d, net_ptr, vtx_ptr, vtx_nets, ep_ends, ep_parts, ep_cnt);
      }
    }
  }
  free(active); free(scores);
  return 0;
}

/* One frozen-state TV evaluation of vertex v (TV/partition.cpp:732-805).
 * Scores start fresh for every vertex (the per-thread marker of partition_par
 * is reset each pass and each vertex is visited once per pass,
 * TV/partition.cpp:714). Returns the LDG2 decision and writes the gain test
 * terms. Part sizes are int64 so that the weighted balanced schedule
 * evaluates through this same function. */
static int tv_eval(int v, const unsigned* net_ptr, const unsigned* vtx_ptr, const unsigned* vtx_nets, int K,
                   double C, long long max_pw, long long min_pw, const int* part, const int* ep_ends,
                   const int* ep_parts, const int* ep_cnt, const long long* sizes, int* marker, int* active,
                   double* scores, int* lg, int* al) {
  int own = part[v];
  int apc = 0;
  for (unsigned i = vtx_ptr[v]; i < vtx_ptr[v + 1]; i++) {
    int n = (int)vtx_nets[i];
    int s = (int)net_ptr[n];
    for (int j = 0; j < ep_ends[n]; j++) {
      int q = ep_parts[s + j];
      if (marker[q] != v) { scores[q] = 1; marker[q] = v; active[apc++] = q; }
      else scores[q]++;
    }
  }
  scores[own]--;
  int d = ldg2_w(K, apc, active, scores, sizes, C, max_pw, min_pw, own);
  /* reset this vertex's marks so the next vertex starts fresh */
  for (int i = 0; i < apc; i++) marker[active[i]] = -1;
  *lg = 0; *al = 0;
  if (d < 0 || d == own) return d;
  for (unsigned i = vtx_ptr[v]; i < vtx_ptr[v + 1]; i++) {
    int n = (int)vtx_nets[i];
    int s = (int)net_ptr[n];
    int found = 0;
    for (int j = 0; j < ep_ends[n]; j++) {
      int t = ep_parts[s + j];
      if (t == own && ep_cnt[s + j] == 1) (*lg)++;
      else if (t == d) found = 1;
    }
    if (!found) (*al)++;
  }
  return d;
}

int or_tv_refine_par(int nov, int noe, const unsigned* net_ptr, const unsigned* vtx_ptr, const unsigned* vtx_nets,
                     const int* ordering, int K, double C, int max_pw, int min_pw, int passes, int R, int chunk,
                     int* part, int* ep_ends, int* ep_parts, int* ep_cnt, int* sizes, int* gains) {
  (void)noe;
  if (R < 1 || chunk < 1) return 1;
  int* marker = (int*)malloc(sizeof(int) * (size_t)K);
  int* active = (int*)malloc(sizeof(int) * (size_t)K);
  double* scores = (double*)calloc((size_t)K, sizeof(double));
  int* dec = (int*)malloc(sizeof(int) * (size_t)(R < nov ? R : (nov > 0 ? nov : 1)));  /* a round holds at most nov */
  long long* frozen = (long long*)malloc(sizeof(long long) * (size_t)K);
  if (!marker || !active || !scores || !dec || !frozen) return 2;
  for (int k = 0; k < K; k++) marker[k] = -1;
  for (int it = 0; it < passes; it++) {
    for (int r0 = 0; r0 < nov; r0 += R) {
      int r1 = r0 + R < nov ? r0 + R : nov;
      /* compute phase on the frozen state (TV/partition.cpp:725-805) */
      for (int k = 0; k < K; k++) frozen[k] = sizes[k];
      for (int z = r0; z < r1; z++) {
        int v = ordering[z];
        int lg, al;
        int d = tv_eval(v, net_ptr, vtx_ptr, vtx_nets, K, C, max_pw, min_pw, part, ep_ends, ep_parts, ep_cnt,
                        frozen, marker, active, scores, &lg, &al);
        if (d == -2) { free(marker); free(active); free(scores); free(dec); free(frozen); return 3; }
        if (gains) gains[v] = (d == part[v]) ? INT_MIN : lg - al;
        dec[z - r0] = (d != part[v] && lg >= al) ? d : -1;
      }
      /* apply in ordering order (TV/partition.cpp:810-892) */
      for (int c0 = r0; c0 < r1; c0 += chunk) {
        int c1 = c0 + chunk < r1 ? c0 + chunk : r1;
        for (int z = c0; z < c1; z++) {
          int d = dec[z - r0];
          if (d < 0) continue;
          int v = ordering[z];
          int own = part[v];
          if (sizes[d] > max_pw) break;   /* "Weird": ends this thread's list */
          part[v] = d;
          sizes[own] -= VW(v);
          move_vertex(v, own, d, net_ptr, vtx_ptr, vtx_nets, ep_ends, ep_parts, ep_cnt);
        }
      }
    }
  }
  free(marker); free(active); free(scores); free(dec); free(frozen);
  return 0;
}

/* Test hook: a nonzero code makes the balanced admission below deliberately
 * wrong, so that tests can show their checks catch such a mistake. 0 = the
 * rule as specified. */

void or_apply_move(int v, int p, int q, const unsigned* net_ptr, const unsigned* vtx_ptr, const unsigned* vtx_nets,
                   int* ep_ends, int* ep_parts, int* ep_cnt) {
  move_vertex(v, p, q, net_ptr, vtx_ptr, vtx_nets, ep_ends, ep_parts, ep_cnt);
}

int or_tv_refine_bal(int nov, int noe, const unsigned* net_ptr, const unsigned* vtx_ptr, const unsigned* vtx_nets,
                     const int* ordering, const int* vwgt, int K, double C, long long max_pw, long long min_pw,
                     long long cap, int passes, int R, int* part, int* ep_ends, int* ep_parts, int* ep_cnt,
                     long long* sizes, int* gains) {
  (void)noe;
  if (R < 1) return 1;
  int* marker = (int*)malloc(sizeof(int) * (size_t)K);
  int* active = (int*)malloc(sizeof(int) * (size_t)K);
  double* scores = (double*)calloc((size_t)K, sizeof(double));
  int* dec = (int*)malloc(sizeof(int) * (size_t)(R < nov ? R : (nov > 0 ? nov : 1)));  /* a round holds at most nov */
  long long* inflow = (long long*)calloc((size_t)K, sizeof(long long));
  if (!marker || !active || !scores || !dec || !inflow) return 2;
  for (int k = 0; k < K; k++) marker[k] = -1;
  for (int it = 0; it < passes; it++) {
    for (int r0 = 0; r0 < nov; r0 += R) {
      int r1 = r0 + R < nov ? r0 + R : nov;
      /* decisions and gain test of the reference round schedule on the frozen state */
      for (int z = r0; z < r1; z++) {
        int v = ordering[z];
        int lg, al;
        int d = tv_eval(v, net_ptr, vtx_ptr, vtx_nets, K, C, max_pw, min_pw, part, ep_ends, ep_parts, ep_cnt,
                        sizes, marker, active, scores, &lg, &al);
        if (d == -2) { free(marker); free(active); free(scores); free(dec); free(inflow); return 3; }
        if (gains) gains[v] = (d == part[v]) ? INT_MIN : lg - al;
        dec[z - r0] = (d != part[v] && lg >= al) ? d : -1;
      }
      /* per-destination prefix admission in ordering order: a candidate into d
       * is admitted iff size[d] + (weight of every earlier candidate into d,
       * admitted or not) + its own weight <= cap */
      for (int k = 0; k < K; k++) inflow[k] = 0;
      for (int z = r0; z < r1; z++) {
        int d = dec[z - r0];
        if (d < 0) continue;
        long long w = vwgt ? vwgt[ordering[z]] : 1;
        long long before = inflow[d];
        int admit = sizes[d] + before + w <= cap;
        if (g_plant == 1) admit = 1;
        if (g_plant == 3) admit = sizes[d] + before + w < cap;
        if (g_plant != 2 || admit) inflow[d] += w;
        if (!admit) dec[z - r0] = -1;
      }
      for (int z = r0; z < r1; z++) {
        int d = dec[z - r0];
        if (d < 0) continue;
        int v = ordering[z];
        int own = part[v];
        long long w = vwgt ? vwgt[v] : 1;
        part[v] = d;
        sizes[own] -= w;
        sizes[d] += w;
        move_vertex(v, own, d, net_ptr, vtx_ptr, vtx_nets, ep_ends, ep_parts, ep_cnt);
      }
    }
  }
  free(marker); free(active); free(scores); free(dec); free(inflow);
  return 0;
}

/* ---------------------------------------------------------------------- */
/* CN                                                                      */
/* ---------------------------------------------------------------------- */

int or_cn_ldg(int nov, int noe, const unsigned* net_ptr, const unsigned* vtx_ptr, const unsigned* vtx_nets,
              const int* ordering, int K, double C, int* part, int* ep_ends, int* ep_parts, int* ep_cnt,
              int* sizes) {
  long long pins = net_ptr[noe];
  int* marker = (int*)malloc(sizeof(int) * (size_t)K);
  int* active = (int*)malloc(sizeof(int) * (size_t)K);
  double* scores = (double*)calloc((size_t)K, sizeof(double));
  if (!marker || !active || !scores) return 2;
  for (long long i = 0; i < pins; i++) { ep_cnt[i] = 0; ep_parts[i] = -1; }
  for (int k = 0; k < K; k++) { marker[k] = -1; sizes[k] = 0; }
  for (int v = 0; v < nov; v++) part[v] = -1;
  for (int n = 0; n < noe; n++) ep_ends[n] = 0;
  for (int z = 0; z < nov; z++) {
    int v = ordering[z];
    int apc = 0;
    /* CN/partition.cpp:138-156: only nets that are not cut vote */
    for (unsigned i = vtx_ptr[v]; i < vtx_ptr[v + 1]; i++) {
      int n = (int)vtx_nets[i];
      if (ep_ends[n] == 1) {
        int q = ep_parts[net_ptr[n]];
        if (marker[q] != v) { scores[q] = 1; marker[q] = v; active[apc++] = q; }
        else scores[q]++;
      }
    }
    /* CN/partition.cpp:159-191 */
    double max_val = 0;
    int max_index = -1;
    for (int i = 0; i < apc; i++) {
      int q = active[i];
      double penalty = 1 - (1.0 * sizes[q] / C);
      double score = scores[q] * penalty;
      if (score > 0 && (max_index == -1 || score > max_val || (score == max_val && sizes[q] < sizes[max_index]))) {
        max_val = score;
        max_index = q;
      }
    }
    if (max_index == -1) {
      int min_load = INT_MAX;
      for (int i = 0; i < K; i++)
        if (sizes[i] < min_load) { min_load = sizes[i]; max_index = i; }
    }
    int d = max_index;
    sizes[d] += VW(v);
    part[v] = d;
    for (unsigned i = vtx_ptr[v]; i < vtx_ptr[v + 1]; i++) list_add((int)vtx_nets[i], d, net_ptr, ep_ends, ep_parts, ep_cnt);
  }
  free(marker); free(active); free(scores);
  return 0;
}

/* lambda=2 push decision for net n on the current state (CN/partition.cpp:254-310).
 * Returns 1 and fills push[] (count *np) with the old-side pins in net->pins
 * order if the push is allowed, else 0. */
static int cn_net_decision(int n, const unsigned* net_ptr, const unsigned* net_pins, const unsigned* vtx_ptr,
                           const unsigned* vtx_nets, const int* part, const int* ep_ends, const int* ep_parts,
                           const int* ep_cnt, int* push, int* np, int* newd) {
  int s = (int)net_ptr[n];
  int p1 = ep_parts[s], c1 = ep_cnt[s];
  int p2 = ep_parts[s + 1], c2 = ep_cnt[s + 1];
  int oldd, nd;
  if (c1 < c2) { oldd = p1; nd = p2; }
  else { oldd = p2; nd = p1; }
  *np = 0;
  for (unsigned k = net_ptr[n]; k < net_ptr[n + 1]; k++) {
    int v = (int)net_pins[k];
    if (part[v] != oldd) continue;
    push[(*np)++] = v;
    for (unsigned j = vtx_ptr[v]; j < vtx_ptr[v + 1]; j++)
      if (ep_ends[vtx_nets[j]] == 1) return 0;
  }
  *newd = nd;
  return 1;
}

int or_cn_refine_serial(int nov, int noe, const unsigned* net_ptr, const unsigned* net_pins,
                        const unsigned* vtx_ptr, const unsigned* vtx_nets, int K, int passes, int* part,
                        int* ep_ends, int* ep_parts, int* ep_cnt, int* sizes) {
  (void)nov; (void)K;
  int maxdeg = 0;
  for (int n = 0; n < noe; n++)
    if ((int)(net_ptr[n + 1] - net_ptr[n]) > maxdeg) maxdeg = (int)(net_ptr[n + 1] - net_ptr[n]);
  int* push = (int*)malloc(sizeof(int) * (size_t)(maxdeg + 1));
  if (!push) return 2;
  for (int it = 0; it < passes; it++) {
    for (int n = 0; n < noe; n++) {
      if (ep_ends[n] != 2) continue;
      int np, nd;
      if (!cn_net_decision(n, net_ptr, net_pins, vtx_ptr, vtx_nets, part, ep_ends, ep_parts, ep_cnt, push, &np, &nd))
        continue;
      /* CN/partition.cpp:313-371: move immediately, no balance check */
      for (int i = 0; i < np; i++) {
        int v = push[i];
        int own = part[v];
        part[v] = nd;
        sizes[own] -= VW(v);
        sizes[nd] += VW(v);
        move_vertex(v, own, nd, net_ptr, vtx_ptr, vtx_nets, ep_ends, ep_parts, ep_cnt);
      }
    }
  }
  free(push);
  return 0;
}

int or_cn_refine_par(int nov, int noe, const unsigned* net_ptr, const unsigned* net_pins, const unsigned* vtx_ptr,
                     const unsigned* vtx_nets, int K, int passes, int* part, int* ep_ends, int* ep_parts,
                     int* ep_cnt, int* sizes) {
  (void)nov; (void)K;
  long long pins = net_ptr[noe];
  int maxdeg = 0;
  for (int n = 0; n < noe; n++)
    if ((int)(net_ptr[n + 1] - net_ptr[n]) > maxdeg) maxdeg = (int)(net_ptr[n + 1] - net_ptr[n]);
  int* push = (int*)malloc(sizeof(int) * (size_t)(maxdeg + 1));
  int* mv_v = (int*)malloc(sizeof(int) * (size_t)(pins + 1));
  int* mv_d = (int*)malloc(sizeof(int) * (size_t)(pins + 1));
  if (!push || !mv_v || !mv_d) return 2;
  for (int it = 0; it < passes; it++) {
    /* evaluation on the frozen state (CN/partition.cpp:629-772) */
    long long m = 0;
    for (int n = 0; n < noe; n++) {
      if (ep_ends[n] != 2) continue;
      int np, nd;
      if (!cn_net_decision(n, net_ptr, net_pins, vtx_ptr, vtx_nets, part, ep_ends, ep_parts, ep_cnt, push, &np, &nd))
        continue;
      for (int i = 0; i < np; i++) { mv_v[m] = push[i]; mv_d[m] = nd; m++; }
    }
    /* apply in push order; the source part is read at apply time (CN/partition.cpp:778-840) */
    for (long long i = 0; i < m; i++) {
      int v = mv_v[i], nd = mv_d[i];
      int own = part[v];
      part[v] = nd;
      sizes[own] -= VW(v);
      sizes[nd] += VW(v);
      move_vertex(v, own, nd, net_ptr, vtx_ptr, vtx_nets, ep_ends, ep_parts, ep_cnt);
    }
  }
  free(push); free(mv_v); free(mv_d);
  return 0;
}

/* ---------------------------------------------------------------------- */
/* MSV (MGV directory)                                                     */
/* ---------------------------------------------------------------------- */

/* band-free LDG (MGV/partition.cpp:28-65) */
static int ldg_plain(int K, int apc, const int* active, const double* scores, const int* sizes, double C) {
  double max_val = 0;
  int max_index = -1;
  for (int i = 0; i < apc; i++) {
    int q = active[i];
    double penalty = 1 - (1.0 * sizes[q] / C);
    double score = scores[q] * penalty;
    if (score > 0) {
      if (max_index == -1 || score > max_val || (score == max_val && sizes[q] < sizes[max_index])) {
        max_val = score;
        max_index = q;
      }
    }
  }
  if (max_index == -1) {
    int min_load = INT_MAX;
    for (int i = 0; i < K; i++)
      if (sizes[i] < min_load) { min_load = sizes[i]; max_index = i; }
  }
  return max_index;
}

/* Initial partition of MGV/partition.cpp:122-189 and MSRV/partition.cpp:131-208:
 * band-free LDG stream; the volumes are accumulated as the stream goes (the
 * source's part is read after the decision). rv NULL: send volumes only. */
static int vol_ldg(int nov, int noe, const unsigned* net_ptr, const unsigned* vtx_ptr, const unsigned* vtx_nets,
                   const int* src, const int* ordering, int K, double C, int* part, int* ep_ends, int* ep_parts,
                   int* ep_cnt, int* sizes, int* sv, int* rv) {
  long long pins = net_ptr[noe];
  int* marker = (int*)malloc(sizeof(int) * (size_t)K);
  int* active = (int*)malloc(sizeof(int) * (size_t)K);
  double* scores = (double*)calloc((size_t)K, sizeof(double));
  if (!marker || !active || !scores) return 2;
  for (long long i = 0; i < pins; i++) { ep_cnt[i] = 0; ep_parts[i] = -1; }
  for (int k = 0; k < K; k++) { marker[k] = -1; sizes[k] = 0; sv[k] = 0; if (rv) rv[k] = 0; }
  for (int v = 0; v < nov; v++) part[v] = -1;
  for (int n = 0; n < noe; n++) ep_ends[n] = 0;
  for (int z = 0; z < nov; z++) {
    int v = ordering[z];
    int apc = 0;
    for (unsigned i = vtx_ptr[v]; i < vtx_ptr[v + 1]; i++) {
      int n = (int)vtx_nets[i];
      int s = (int)net_ptr[n];
      for (int j = 0; j < ep_ends[n]; j++) {
        int q = ep_parts[s + j];
        if (marker[q] != v) { scores[q] = 1; marker[q] = v; active[apc++] = q; }
        else scores[q]++;
      }
    }
    int d = ldg_plain(K, apc, active, scores, sizes, C);
    sizes[d] += VW(v);
    part[v] = d;
    for (unsigned i = vtx_ptr[v]; i < vtx_ptr[v + 1]; i++) {
      int n = (int)vtx_nets[i];
      int s = (int)net_ptr[n];
      int end = ep_ends[n];
      int sn = src[n];
      int sp = part[sn];
      int found = 0;
      for (int j = 0; j < end; j++) {
        int q = ep_parts[s + j];
        if (sn == v && q != sp) { sv[sp]++; if (rv) rv[q]++; }
        if (q == d) { found = 1; ep_cnt[s + j]++; }
      }
      if (!found) {
        ep_parts[s + end] = d;
        ep_cnt[s + end] = 1;
        ep_ends[n]++;
        if (sn != v && sp != -1) { sv[sp]++; if (rv) rv[d]++; }
      }
    }
  }
  free(marker); free(active); free(scores);
  return 0;
}

int or_msv_ldg(int nov, int noe, const unsigned* net_ptr, const unsigned* vtx_ptr, const unsigned* vtx_nets,
               const int* src, const int* ordering, int K, double C, int* part, int* ep_ends, int* ep_parts,
               int* ep_cnt, int* sizes, int* sv) {
  return vol_ldg(nov, noe, net_ptr, vtx_ptr, vtx_nets, src, ordering, K, C, part, ep_ends, ep_parts, ep_cnt, sizes,
                 sv, NULL);
}

int or_msrv_ldg(int nov, int noe, const unsigned* net_ptr, const unsigned* vtx_ptr, const unsigned* vtx_nets,
                const int* src, const int* ordering, int K, double C, int* part, int* ep_ends, int* ep_parts,
                int* ep_cnt, int* sizes, int* sv, int* rv) {
  return vol_ldg(nov, noe, net_ptr, vtx_ptr, vtx_nets, src, ordering, K, C, part, ep_ends, ep_parts, ep_cnt, sizes,
                 sv, rv);
}

static int vol_of(const int* sv, const int* rv, int k) { return sv[k] + (rv ? rv[k] : 0); }

/* The max set (MGV/partition.cpp:712-729; MSRV/partition.cpp:772-791 on
 * send + receive): returns the max and sets inmax[k] = 1 for its parts. */
static int vol_max_set(int K, const int* sv, const int* rv, unsigned char* inmax) {
  int m = -1;
  for (int k = 0; k < K; k++) if (vol_of(sv, rv, k) > m) m = vol_of(sv, rv, k);
  for (int k = 0; k < K; k++) inmax[k] = (unsigned char)(vol_of(sv, rv, k) == m);
  return m;
}

/* The reference's gains and losses of moving v from its part to d
 * (MGV/partition.cpp:800-840; with recv, MSRV/partition.cpp:852-897), added
 * into gains[K] / losses[K]. */
static void vol_gains(int v, int d, const unsigned* net_ptr, const unsigned* vtx_ptr, const unsigned* vtx_nets,
                      const int* src, const int* part, const int* ep_ends, const int* ep_parts, const int* ep_cnt,
                      int recv, int* gains, int* losses) {
  int own = part[v];
  for (unsigned i = vtx_ptr[v]; i < vtx_ptr[v + 1]; i++) {
    int n = (int)vtx_nets[i];
    int s = (int)net_ptr[n];
    int end = ep_ends[n];
    int sn = src[n];
    int sp = part[sn];
    if (sn == v) { gains[own] += end - 1; losses[d] += end - 1; }
    int found = 0;
    for (int j = 0; j < end; j++) {
      if (ep_parts[s + j] == own && ep_cnt[s + j] == 1) {
        if (sn != v) { gains[sp]++; if (recv) gains[own]++; }
        else losses[d]--;
      }
      if (ep_parts[s + j] == d) found = 1;
    }
    if (!found) {
      if (sn != v) { losses[sp]++; if (recv) losses[d]++; }
      else losses[d]++;
    }
  }
}

int or_vol_move_delta(int nov, int noe, const unsigned* net_ptr, const unsigned* vtx_ptr, const unsigned* vtx_nets,
                      const int* src, int K, const int* part, const int* ep_ends, const int* ep_parts,
                      const int* ep_cnt, int recv, int v, int d, int* delta) {
  (void)noe;
  if (v < 0 || v >= nov || d < 0 || d >= K || d == part[v]) return 1;
  int* gains = (int*)calloc((size_t)K, sizeof(int));
  int* losses = (int*)calloc((size_t)K, sizeof(int));
  if (!gains || !losses) return 2;
  vol_gains(v, d, net_ptr, vtx_ptr, vtx_nets, src, part, ep_ends, ep_parts, ep_cnt, recv, gains, losses);
  for (int k = 0; k < K; k++) delta[k] = losses[k] - gains[k];
  free(gains); free(losses);
  return 0;
}

/* LDG decision and max-set test of vertex v (MGV/partition.cpp:758-856;
 * with rv, MSRV/partition.cpp:818-920: receive terms in the gains and losses
 * and a strict gain for the max parts). Returns the decision if the move is
 * accepted, else -1. marker: all -1 on entry and exit; gains/losses: all 0 on
 * entry and exit. The max test reads inmax[]/max_v; with sticky != NULL it
 * reads the max-part list sticky[0..nsticky) instead (MSRV's serial
 * bookkeeping, see or_msrv_refine_serial). */
static int vol_eval(int v, const unsigned* net_ptr, const unsigned* vtx_ptr, const unsigned* vtx_nets,
                    const int* src, int K, double C, const int* part, const int* ep_ends, const int* ep_parts,
                    const int* ep_cnt, const int* sizes, const int* sv, const int* rv, int max_v,
                    const unsigned char* inmax, const int* sticky, int nsticky, int* marker, int* active,
                    double* scores, int* gains, int* losses) {
  int own = part[v];
  int apc = 0;
  for (unsigned i = vtx_ptr[v]; i < vtx_ptr[v + 1]; i++) {
    int n = (int)vtx_nets[i];
    int s = (int)net_ptr[n];
    for (int j = 0; j < ep_ends[n]; j++) {
      int q = ep_parts[s + j];
      if (marker[q] != v) { scores[q] = 1; marker[q] = v; active[apc++] = q; }
      else scores[q]++;
    }
  }
  scores[own]--;
  int d = ldg_plain(K, apc, active, scores, sizes, C);
  for (int i = 0; i < apc; i++) marker[active[i]] = -1;
  if (d == own) return -1;
  vol_gains(v, d, net_ptr, vtx_ptr, vtx_nets, src, part, ep_ends, ep_parts, ep_cnt, rv != NULL, gains, losses);
  int good = 1;
  if (sticky) {
    for (int i = 0; i < nsticky && good; i++)
      if (gains[sticky[i]] <= losses[sticky[i]]) good = 0;
  } else {
    for (int k = 0; k < K && good; k++)
      if (inmax[k] && (rv ? gains[k] <= losses[k] : gains[k] < losses[k])) good = 0;
  }
  for (int k = 0; k < K && good; k++)
    if (!inmax[k] && vol_of(sv, rv, k) + (losses[k] - gains[k]) >= max_v) good = 0;
  for (int k = 0; k < K; k++) { gains[k] = 0; losses[k] = 0; }
  return good ? d : -1;
}

/* Applies v: own -> d with the volume bookkeeping of MGV/partition.cpp:897-975
 * (with rv, MSRV/partition.cpp:952-1052); the source's part is read after
 * part[v] = d. */
static void vol_apply(int v, int d, const unsigned* net_ptr, const unsigned* vtx_ptr, const unsigned* vtx_nets,
                      const int* src, int* part, int* ep_ends, int* ep_parts, int* ep_cnt, int* sizes, int* sv,
                      int* rv) {
  int own = part[v];
  part[v] = d;
  sizes[own] -= VW(v);
  sizes[d] += VW(v);
  for (unsigned i = vtx_ptr[v]; i < vtx_ptr[v + 1]; i++) {
    int n = (int)vtx_nets[i];
    int s = (int)net_ptr[n];
    int end = ep_ends[n];
    int sn = src[n];
    int sp = part[sn];
    if (sn == v) sv[own] -= end - 1;
    for (int j = 0; j < end; j++) {
      if (ep_parts[s + j] == own) {
        ep_cnt[s + j]--;
        if (ep_cnt[s + j] == 0) {
          ep_parts[s + j] = ep_parts[s + end - 1];
          ep_cnt[s + j] = ep_cnt[s + end - 1];
          ep_parts[s + end - 1] = -1;
          ep_cnt[s + end - 1] = 0;
          ep_ends[n]--;
          end--;
          if (sn != v) { sv[sp]--; if (rv) rv[own]--; }
        } else if (rv && sn == v) {
          rv[own]++;
        }
      }
    }
  }
  for (unsigned i = vtx_ptr[v]; i < vtx_ptr[v + 1]; i++) {
    int n = (int)vtx_nets[i];
    int s = (int)net_ptr[n];
    int end = ep_ends[n];
    int sn = src[n];
    int sp = part[sn];
    if (sn == v) sv[sp] += end - 1;
    int found = 0;
    for (int j = 0; j < end; j++)
      if (ep_parts[s + j] == d) {
        found = 1;
        ep_cnt[s + j]++;
        if (rv && sn == v) rv[d]--;
        break;
      }
    if (!found) {
      ep_parts[s + end] = d;
      ep_cnt[s + end] = 1;
      ep_ends[n]++;
      sv[sp]++;
      if (rv && sn != v) rv[d]++;
    }
  }
}

typedef struct {
  int *marker, *active, *gains, *losses;
  double* scores;
  unsigned char *inmax, *inmax2;
} vol_scratch;

static int vol_scratch_init(vol_scratch* w, int K) {
  w->marker = (int*)malloc(sizeof(int) * (size_t)K);
  w->active = (int*)malloc(sizeof(int) * (size_t)K);
  w->gains = (int*)calloc((size_t)K, sizeof(int));
  w->losses = (int*)calloc((size_t)K, sizeof(int));
  w->scores = (double*)calloc((size_t)K, sizeof(double));
  w->inmax = (unsigned char*)malloc((size_t)K);
  w->inmax2 = (unsigned char*)malloc((size_t)K);
  if (!w->marker || !w->active || !w->gains || !w->losses || !w->scores || !w->inmax || !w->inmax2) return 2;
  for (int k = 0; k < K; k++) w->marker[k] = -1;
  return 0;
}

static void vol_scratch_free(vol_scratch* w) {
  free(w->marker); free(w->active); free(w->gains); free(w->losses); free(w->scores); free(w->inmax);
  free(w->inmax2);
}

int or_msv_refine_serial(int nov, int noe, const unsigned* net_ptr, const unsigned* vtx_ptr,
                         const unsigned* vtx_nets, const int* src, const int* ordering, int K, double C,
                         int passes, int* part, int* ep_ends, int* ep_parts, int* ep_cnt, int* sizes, int* sv) {
  (void)noe;
  vol_scratch w;
  if (vol_scratch_init(&w, K)) return 2;
  for (int it = 0; it < passes; it++) {
    int max_v = vol_max_set(K, sv, NULL, w.inmax);
    for (int z = 0; z < nov; z++) {
      int v = ordering[z];
      int d = vol_eval(v, net_ptr, vtx_ptr, vtx_nets, src, K, C, part, ep_ends, ep_parts, ep_cnt, sizes, sv, NULL,
                       max_v, w.inmax, NULL, 0, w.marker, w.active, w.scores, w.gains, w.losses);
      if (d < 0) continue;
      vol_apply(v, d, net_ptr, vtx_ptr, vtx_nets, src, part, ep_ends, ep_parts, ep_cnt, sizes, sv, NULL);
      max_v = vol_max_set(K, sv, NULL, w.inmax);
    }
  }
  vol_scratch_free(&w);
  return 0;
}

/* shared round schedule of both partition_par's (vertex = ordering ? ordering[z] : z) */
static int vol_refine_par(int nov, const unsigned* net_ptr, const unsigned* vtx_ptr, const unsigned* vtx_nets,
                          const int* src, const int* ordering, int K, double C, int passes, int R, int* part,
                          int* ep_ends, int* ep_parts, int* ep_cnt, int* sizes, int* sv, int* rv, long long* moves) {
  if (R < 1) return 1;
  vol_scratch w;
  if (vol_scratch_init(&w, K)) return 2;
  int* dec = (int*)malloc(sizeof(int) * (size_t)(R < nov ? R : (nov > 0 ? nov : 1)));  /* a round holds at most nov */
  if (!dec) return 2;
  long long applied = 0;
  for (int it = 0; it < passes; it++) {
    int max_v = vol_max_set(K, sv, rv, w.inmax);
    for (int r0 = 0; r0 < nov; r0 += R) {
      int r1 = r0 + R < nov ? r0 + R : nov;
      for (int z = r0; z < r1; z++) {
        int v = ordering ? ordering[z] : z;
        dec[z - r0] = vol_eval(v, net_ptr, vtx_ptr, vtx_nets, src, K, C, part, ep_ends, ep_parts, ep_cnt, sizes,
                               sv, rv, max_v, w.inmax, NULL, 0, w.marker, w.active, w.scores, w.gains, w.losses);
      }
      for (int z = r0; z < r1; z++) {
        int d = dec[z - r0];
        if (d < 0) continue;
        int v = ordering ? ordering[z] : z;
        vol_apply(v, d, net_ptr, vtx_ptr, vtx_nets, src, part, ep_ends, ep_parts, ep_cnt, sizes, sv, rv);
        applied++;
        /* MGV/partition.cpp:977-1003: a new part in the max set ends the round */
        memcpy(w.inmax2, w.inmax, (size_t)K);
        max_v = vol_max_set(K, sv, rv, w.inmax);
        int entered = 0;
        for (int k = 0; k < K; k++) if (w.inmax[k] && !w.inmax2[k]) entered = 1;
        if (entered) break;
      }
    }
  }
  if (moves) *moves = applied;
  free(dec);
  vol_scratch_free(&w);
  return 0;
}

int or_msv_refine_par(int nov, int noe, const unsigned* net_ptr, const unsigned* vtx_ptr, const unsigned* vtx_nets,
                      const int* src, int K, double C, int passes, int R, int* part, int* ep_ends, int* ep_parts,
                      int* ep_cnt, int* sizes, int* sv, long long* moves) {
  (void)noe;
  return vol_refine_par(nov, net_ptr, vtx_ptr, vtx_nets, src, NULL, K, C, passes, R, part, ep_ends, ep_parts, ep_cnt,
                        sizes, sv, NULL, moves);
}

int or_msrv_refine_par(int nov, int noe, const unsigned* net_ptr, const unsigned* vtx_ptr, const unsigned* vtx_nets,
                       const int* src, const int* ordering, int K, double C, int passes, int R, int* part,
                       int* ep_ends, int* ep_parts, int* ep_cnt, int* sizes, int* sv, int* rv, long long* moves) {
  (void)noe;
  return vol_refine_par(nov, net_ptr, vtx_ptr, vtx_nets, src, ordering, K, C, passes, R, part, ep_ends, ep_parts,
                        ep_cnt, sizes, sv, rv, moves);
}

int or_msrv_refine_serial(int nov, int noe, const unsigned* net_ptr, const unsigned* vtx_ptr,
                          const unsigned* vtx_nets, const int* src, const int* ordering, int K, double C,
                          int passes, int* part, int* ep_ends, int* ep_parts, int* ep_cnt, int* sizes, int* sv,
                          int* rv) {
  (void)noe;
  vol_scratch w;
  if (vol_scratch_init(&w, K)) return 2;
  /* the max-part list can grow past K (parts are appended again at every
   * recomputation, MSRV/partition.cpp:496-512); capacity grows as needed */
  int cap = K > 0 ? 2 * K : 2;
  int* maxp = (int*)malloc(sizeof(int) * (size_t)cap);
  int* mark = (int*)malloc(sizeof(int) * (size_t)K);
  if (!maxp || !mark) return 2;
  for (int it = 0; it < passes; it++) {
    /* MSRV/partition.cpp:240-262: fresh max set at the pass start */
    int max_v = -1, nmax = 0, changed = 0;
    for (int k = 0; k < K; k++) mark[k] = -1;
    for (int k = 0; k < K; k++) {
      int a = vol_of(sv, rv, k);
      if (a > max_v) { nmax = 1; maxp[0] = k; max_v = a; changed++; mark[k] = changed; }
      else if (a == max_v) { maxp[nmax++] = k; mark[k] = changed; }
    }
    for (int z = 0; z < nov; z++) {
      int v = ordering[z];
      for (int k = 0; k < K; k++) w.inmax[k] = (unsigned char)(mark[k] == changed);
      int d = vol_eval(v, net_ptr, vtx_ptr, vtx_nets, src, K, C, part, ep_ends, ep_parts, ep_cnt, sizes, sv, rv,
                       max_v, w.inmax, maxp, nmax, w.marker, w.active, w.scores, w.gains, w.losses);
      if (d < 0) continue;
      vol_apply(v, d, net_ptr, vtx_ptr, vtx_nets, src, part, ep_ends, ep_parts, ep_cnt, sizes, sv, rv);
      /* MSRV/partition.cpp:486-512: recomputed only when a listed max part
       * changed its volume, and without resetting the list, the max or the
       * marks: the max never falls within a pass and the list keeps growing */
      int redo = 0;
      for (int i = 0; i < nmax; i++) if (max_v != vol_of(sv, rv, maxp[i])) redo = 1;
      for (int k = 0; redo && k < K; k++) {
        int a = vol_of(sv, rv, k);
        if (a > max_v) { nmax = 1; maxp[0] = k; max_v = a; changed++; mark[k] = changed; }
        else if (a == max_v) {
          if (nmax == cap) {
            cap *= 2;
            int* g = (int*)realloc(maxp, sizeof(int) * (size_t)cap);
            if (!g) return 2;
            maxp = g;
          }
          maxp[nmax++] = k;
          mark[k] = changed;
        }
      }
    }
  }
  free(maxp); free(mark);
  vol_scratch_free(&w);
  return 0;
}

/* ---------------------------------------------------------------------- */
/* TM and MSM: part connections                                            */
/* ---------------------------------------------------------------------- */
/* pc[a*K + b] = number of nets whose source is in part a and that have pins
 * in part b (b != a); msg[a] = number of b with pc[a][b] > 0. */

/* Greedy initial partition (TM/partition.cpp:87-235; MSM/partition.cpp:
 * identical stream): the part saving most messages under the capacity C,
 * ties to the lighter part, else the lightest part. */
static int conn_init(int nov, int noe, const unsigned* net_ptr, const unsigned* vtx_ptr, const unsigned* vtx_nets,
                     const int* src, const int* ordering, int K, double C, int diag, int* part, int* ep_ends,
                     int* ep_parts, int* ep_cnt, int* sizes, int* pc, int* msg) {
  long long pins = net_ptr[noe];
  int maxdeg = 0;
  for (int v = 0; v < nov; v++)
    if ((int)(vtx_ptr[v + 1] - vtx_ptr[v]) > maxdeg) maxdeg = (int)(vtx_ptr[v + 1] - vtx_ptr[v]);
  int* src_marker = (int*)malloc(sizeof(int) * (size_t)K);
  int* src_part_of = (int*)malloc(sizeof(int) * (size_t)(noe > 0 ? noe : 1));
  int* active_src = (int*)malloc(sizeof(int) * (size_t)(maxdeg + 1));
  int* active_net = (int*)malloc(sizeof(int) * (size_t)(maxdeg + 1));
  int* src_update = (int*)malloc(sizeof(int) * (size_t)(maxdeg + 1));
  if (!src_marker || !src_part_of || !active_src || !active_net || !src_update) return 2;
  for (long long i = 0; i < pins; i++) { ep_cnt[i] = 0; ep_parts[i] = -1; }
  for (int k = 0; k < K; k++) { src_marker[k] = -1; sizes[k] = 0; msg[k] = 0; }
  for (long long i = 0; i < (long long)K * K; i++) pc[i] = 0;
  for (int k = 0; k < K; k++) pc[(size_t)k * K + k] = diag;   /* MSM/partition.cpp:99: -1 */
  for (int v = 0; v < nov; v++) part[v] = -1;
  for (int n = 0; n < noe; n++) { ep_ends[n] = 0; src_part_of[n] = -1; }
  for (int z = 0; z < nov; z++) {
    int v = ordering[z];
    int nsrc = 0, nupd = 0, nact = 0;
    for (unsigned i = vtx_ptr[v]; i < vtx_ptr[v + 1]; i++) {
      int n = (int)vtx_nets[i];
      if (src[n] == v) src_update[nupd++] = n;
      if (src_part_of[n] != -1) {
        int sp = src_part_of[n];
        if (src_marker[sp] != v) { src_marker[sp] = v; active_src[nsrc++] = sp; }
        active_net[nact++] = n;
      }
    }
    int best = -1, best_saved = -1;
    for (int k = 0; k < K; k++) {
      if (!(sizes[k] < C)) continue;
      int saved = 0;
      for (int i = 0; i < nsrc; i++) if (pc[(size_t)active_src[i] * K + k] > 0) saved++;
      for (int i = 0; i < nupd; i++) {
        int n = src_update[i];
        int s = (int)net_ptr[n];
        for (int j = 0; j < ep_ends[n]; j++) if (pc[(size_t)k * K + ep_parts[s + j]] > 0) saved++;
      }
      if (saved > best_saved || (saved == best_saved && sizes[k] < sizes[best])) { best = k; best_saved = saved; }
    }
    if (best == -1) {
      int min_load = INT_MAX;
      for (int k = 0; k < K; k++) if (sizes[k] < min_load) { min_load = sizes[k]; best = k; }
    }
    int d = best;
    part[v] = d;
    sizes[d] += VW(v);
    for (unsigned i = vtx_ptr[v]; i < vtx_ptr[v + 1]; i++) list_add((int)vtx_nets[i], d, net_ptr, ep_ends, ep_parts, ep_cnt);
    /* TM/partition.cpp:190-215 */
    for (int i = 0; i < nact; i++) {
      int n = active_net[i];
      int sp = src_part_of[n];
      if (sp == d) continue;
      size_t c = (size_t)sp * K + d;
      if (pc[c] == 0) { msg[sp]++; pc[c] = 1; }
      else {
        int s = (int)net_ptr[n];
        for (int j = 0; j < ep_ends[n]; j++)
          if (ep_parts[s + j] == d) { if (ep_cnt[s + j] == 1) pc[c]++; break; }
      }
    }
    /* TM/partition.cpp:216-235 */
    for (int i = 0; i < nupd; i++) {
      int n = src_update[i];
      src_part_of[n] = d;
      int s = (int)net_ptr[n];
      for (int j = 0; j < ep_ends[n]; j++) {
        int q = ep_parts[s + j];
        if (q == d) continue;
        size_t c = (size_t)d * K + q;
        if (pc[c] == 0) { msg[d]++; pc[c] = 1; }
        else pc[c]++;
      }
    }
  }
  free(src_marker); free(src_part_of); free(active_src); free(active_net); free(src_update);
  return 0;
}

int or_tm_init(int nov, int noe, const unsigned* net_ptr, const unsigned* vtx_ptr, const unsigned* vtx_nets,
               const int* src, const int* ordering, int K, double C, int* part, int* ep_ends, int* ep_parts,
               int* ep_cnt, int* sizes, int* pc, int* msg) {
  return conn_init(nov, noe, net_ptr, vtx_ptr, vtx_nets, src, ordering, K, C, 0, part, ep_ends, ep_parts, ep_cnt,
                   sizes, pc, msg);
}

int or_msm_init(int nov, int noe, const unsigned* net_ptr, const unsigned* vtx_ptr, const unsigned* vtx_nets,
                const int* src, const int* ordering, int K, double C, int* part, int* ep_ends, int* ep_parts,
                int* ep_cnt, int* sizes, int* pc, int* msg) {
  return conn_init(nov, noe, net_ptr, vtx_ptr, vtx_nets, src, ordering, K, C, -1, part, ep_ends, ep_parts, ep_cnt,
                   sizes, pc, msg);
}

/* Receivers, senders and the TM gain of moving v out of its part
 * (TM/partition.cpp:318-391). gains (nullable, K): MSM's per-part split of
 * the same terms (MSM/partition.cpp:1075-1130: the source-net terms go to
 * the own part, the sender terms to the sender part). Returns the total gain.
 * mrecv/msend: K scratch marks, -1 or another vertex on entry. */
static int conn_gain(int v, const unsigned* net_ptr, const unsigned* vtx_ptr, const unsigned* vtx_nets,
                     const int* src, int K, const int* part, const int* ep_ends, const int* ep_parts,
                     const int* ep_cnt, const int* pc, int* mrecv, int* msend, int* recv, int* nrecv, int* send,
                     int* nsend, int* gains) {
  int own = part[v];
  int gain = 0;
  *nrecv = 0;
  *nsend = 0;
  for (unsigned i = vtx_ptr[v]; i < vtx_ptr[v + 1]; i++) {
    int n = (int)vtx_nets[i];
    int s = (int)net_ptr[n];
    int end = ep_ends[n];
    if (src[n] == v) {
      for (int j = 0; j < end; j++) {
        int t = ep_parts[s + j];
        if (t == own && ep_cnt[s + j] > 1 && mrecv[own] != v) { recv[(*nrecv)++] = own; mrecv[own] = v; }
        if (t != own) {
          if (mrecv[t] != v) { recv[(*nrecv)++] = t; mrecv[t] = v; }
          /* the reference's gain_loss_2d counter is 1 here (it is raised and
           * lowered within one net), so the test is pc[own][t] == 1 */
          if (pc[(size_t)own * K + t] == 1) { gain++; if (gains) gains[own]++; }
        }
      }
    } else {
      int ps = part[src[n]];
      if (msend[ps] != v) {
        send[(*nsend)++] = ps;
        msend[ps] = v;
        for (int j = 0; j < end; j++)
          if (ep_parts[s + j] == own && ep_cnt[s + j] == 1 && pc[(size_t)ps * K + own] == 1) {
            gain++;
            if (gains) gains[ps]++;
          }
      }
    }
  }
  return gain;
}

/* TM decision of v on the given state (TM/partition.cpp:318-440 and
 * :978-1084): -1 or the part with the largest improvement gain - loss > 0
 * (the first such part on ties), among parts lighter than C. */
static int tm_decide(int v, const unsigned* net_ptr, const unsigned* vtx_ptr, const unsigned* vtx_nets,
                     const int* src, int K, double C, const int* part, const int* ep_ends, const int* ep_parts,
                     const int* ep_cnt, const int* sizes, const int* pc, int* mrecv, int* msend, int* recv,
                     int* send) {
  int nrecv, nsend;
  int own = part[v];
  int gain = conn_gain(v, net_ptr, vtx_ptr, vtx_nets, src, K, part, ep_ends, ep_parts, ep_cnt, pc, mrecv, msend, recv,
                       &nrecv, send, &nsend, NULL);
  if (gain <= 0) return -1;
  int best = -1, best_imp = 0;
  for (int k = 0; k < K; k++) {
    if (k == own || !(sizes[k] < C)) continue;
    int loss = 0;
    for (int i = 0; i < nrecv; i++) if (recv[i] != k && pc[(size_t)k * K + recv[i]] == 0) loss++;
    for (int i = 0; i < nsend; i++) if (send[i] != k && pc[(size_t)send[i] * K + k] == 0) loss++;
    if (gain > loss && gain - loss > best_imp) { best = k; best_imp = gain - loss; }
  }
  return best;
}

/* Moves v to d: list edits then the connection update of
 * TM/partition.cpp:1178-1240 (identical in MSM/partition.cpp:1257-1330). */
static void conn_apply(int v, int d, const unsigned* net_ptr, const unsigned* vtx_ptr, const unsigned* vtx_nets,
                       const int* src, int K, int* part, int* ep_ends, int* ep_parts, int* ep_cnt, int* sizes,
                       int* pc, int* msg) {
  int own = part[v];
  part[v] = d;
  sizes[own] -= VW(v);
  sizes[d] += VW(v);
  move_vertex(v, own, d, net_ptr, vtx_ptr, vtx_nets, ep_ends, ep_parts, ep_cnt);
  for (unsigned i = vtx_ptr[v]; i < vtx_ptr[v + 1]; i++) {
    int n = (int)vtx_nets[i];
    int s = (int)net_ptr[n];
    int end = ep_ends[n];
    if (src[n] == v) {
      for (int j = 0; j < end; j++) {
        int t = ep_parts[s + j];
        if (t != d) { if (++pc[(size_t)d * K + t] == 1) msg[d]++; }
        if (t != own && t != d) { if (--pc[(size_t)own * K + t] == 0) msg[own]--; }
        if (t == d && ep_cnt[s + j] > 1) { if (--pc[(size_t)own * K + t] == 0) msg[own]--; }
      }
    } else {
      int sp = part[src[n]];
      int found = 0;
      for (int j = 0; j < end; j++) {
        if (ep_parts[s + j] == d && ep_cnt[s + j] == 1) { if (++pc[(size_t)sp * K + d] == 1) msg[sp]++; }
        if (ep_parts[s + j] == own) found = 1;
      }
      if (!found) { if (--pc[(size_t)sp * K + own] == 0) msg[sp]--; }
    }
  }
}

typedef struct {
  int *mrecv, *msend, *recv, *send;
} conn_scratch;

static int conn_scratch_init(conn_scratch* w, int K) {
  w->mrecv = (int*)malloc(sizeof(int) * (size_t)K);
  w->msend = (int*)malloc(sizeof(int) * (size_t)K);
  w->recv = (int*)malloc(sizeof(int) * (size_t)K);
  w->send = (int*)malloc(sizeof(int) * (size_t)K);
  if (!w->mrecv || !w->msend || !w->recv || !w->send) return 2;
  for (int k = 0; k < K; k++) { w->mrecv[k] = -1; w->msend[k] = -1; }
  return 0;
}

static void conn_scratch_free(conn_scratch* w) { free(w->mrecv); free(w->msend); free(w->recv); free(w->send); }

int or_tm_refine_serial(int nov, int noe, const unsigned* net_ptr, const unsigned* vtx_ptr, const unsigned* vtx_nets,
                        const int* src, int K, double C, int passes, int* part, int* ep_ends, int* ep_parts,
                        int* ep_cnt, int* sizes, int* pc, int* msg) {
  (void)noe;
  conn_scratch w;
  if (conn_scratch_init(&w, K)) return 2;
  for (int it = 0; it < passes; it++) {
    for (int k = 0; k < K; k++) { w.mrecv[k] = -1; w.msend[k] = -1; }
    for (int v = 0; v < nov; v++) {
      int d = tm_decide(v, net_ptr, vtx_ptr, vtx_nets, src, K, C, part, ep_ends, ep_parts, ep_cnt, sizes, pc,
                        w.mrecv, w.msend, w.recv, w.send);
      if (d < 0) continue;
      conn_apply(v, d, net_ptr, vtx_ptr, vtx_nets, src, K, part, ep_ends, ep_parts, ep_cnt, sizes, pc, msg);
    }
  }
  conn_scratch_free(&w);
  return 0;
}

int or_tm_refine_par(int nov, int noe, const unsigned* net_ptr, const unsigned* vtx_ptr, const unsigned* vtx_nets,
                     const int* src, int K, double C, int passes, int R, int* part, int* ep_ends, int* ep_parts,
                     int* ep_cnt, int* sizes, int* pc, int* msg, long long* moves) {
  (void)noe;
  if (R < 1) return 1;
  conn_scratch w;
  if (conn_scratch_init(&w, K)) return 2;
  int* dec = (int*)malloc(sizeof(int) * (size_t)(R < nov ? R : (nov > 0 ? nov : 1)));  /* a round holds at most nov */
  if (!dec) return 2;
  const double thr = C * (1.02);
  long long applied = 0;
  for (int it = 0; it < passes; it++) {
    for (int k = 0; k < K; k++) { w.mrecv[k] = -1; w.msend[k] = -1; }
    for (int r0 = 0; r0 < nov; r0 += R) {
      int r1 = r0 + R < nov ? r0 + R : nov;
      for (int v = r0; v < r1; v++)
        dec[v - r0] = tm_decide(v, net_ptr, vtx_ptr, vtx_nets, src, K, C, part, ep_ends, ep_parts, ep_cnt, sizes, pc,
                                w.mrecv, w.msend, w.recv, w.send);
      /* TM/partition.cpp:1115-1122: a destination at or above C * 1.02 is skipped */
      for (int v = r0; v < r1; v++) {
        int d = dec[v - r0];
        if (d < 0 || sizes[d] >= thr) continue;
        conn_apply(v, d, net_ptr, vtx_ptr, vtx_nets, src, K, part, ep_ends, ep_parts, ep_cnt, sizes, pc, msg);
        applied++;
      }
    }
  }
  if (moves) *moves = applied;
  free(dec);
  conn_scratch_free(&w);
  return 0;
}

/* MSM decision of v (MSM/partition.cpp:1058-1190; serial :364-500): among
 * k != own, the part with the largest improvement = sum over the max parts of
 * gains - losses, provided every max part gains strictly more than it loses
 * and no other part reaches the max; the first such k on ties; -1 if none.
 * losses[k] counts receivers r with pc[k][r] == 0, losses[s] senders s with
 * pc[s][k] == 0 (the diagonal of pc is -1, so it never counts). */
static int msm_decide(int v, const unsigned* net_ptr, const unsigned* vtx_ptr, const unsigned* vtx_nets,
                      const int* src, int K, const int* part, const int* ep_ends, const int* ep_parts,
                      const int* ep_cnt, const int* pc, const int* msg, int max_m, const unsigned char* inmax,
                      int* mrecv, int* msend, int* recv, int* send, int* gains, int* losses) {
  int nrecv, nsend;
  int own = part[v];
  for (int k = 0; k < K; k++) gains[k] = 0;
  int gain = conn_gain(v, net_ptr, vtx_ptr, vtx_nets, src, K, part, ep_ends, ep_parts, ep_cnt, pc, mrecv, msend, recv,
                       &nrecv, send, &nsend, gains);
  if (gain <= 0) return -1;   /* has_gain */
  int best = -1, best_imp = -1;
  for (int k = 0; k < K; k++) {
    if (k == own) continue;
    for (int i = 0; i < K; i++) losses[i] = 0;
    for (int i = 0; i < nrecv; i++) if (pc[(size_t)k * K + recv[i]] == 0) losses[k]++;
    for (int i = 0; i < nsend; i++) if (pc[(size_t)send[i] * K + k] == 0) losses[send[i]]++;
    int good = 1, imp = 0;
    for (int m = 0; m < K && good; m++) {
      if (!inmax[m]) continue;
      if (gains[m] <= losses[m]) good = 0;
      else imp += gains[m] - losses[m];
    }
    for (int i = 0; i < K && good; i++)
      if (!inmax[i] && msg[i] + (losses[i] - gains[i]) >= max_m) good = 0;
    if (good && imp > best_imp) { best = k; best_imp = imp; }
  }
  return best;
}

static int msm_max_set(int K, const int* msg, unsigned char* inmax) {
  int m = -1;
  for (int k = 0; k < K; k++) if (msg[k] > m) m = msg[k];
  for (int k = 0; k < K; k++) inmax[k] = (unsigned char)(msg[k] == m);
  return m;
}

int or_msm_refine_serial(int nov, int noe, const unsigned* net_ptr, const unsigned* vtx_ptr, const unsigned* vtx_nets,
                         const int* src, int K, int passes, int* part, int* ep_ends, int* ep_parts, int* ep_cnt,
                         int* sizes, int* pc, int* msg) {
  (void)noe;
  conn_scratch w;
  if (conn_scratch_init(&w, K)) return 2;
  int* gains = (int*)calloc((size_t)K, sizeof(int));
  int* losses = (int*)calloc((size_t)K, sizeof(int));
  unsigned char* inmax = (unsigned char*)malloc((size_t)K);
  if (!gains || !losses || !inmax) return 2;
  for (int it = 0; it < passes; it++) {
    for (int k = 0; k < K; k++) { w.mrecv[k] = -1; w.msend[k] = -1; }
    for (int v = 0; v < nov; v++) {
      /* MSM/partition.cpp:364-380: the max set is recomputed before every vertex */
      int max_m = msm_max_set(K, msg, inmax);
      int d = msm_decide(v, net_ptr, vtx_ptr, vtx_nets, src, K, part, ep_ends, ep_parts, ep_cnt, pc, msg, max_m,
                         inmax, w.mrecv, w.msend, w.recv, w.send, gains, losses);
      if (d < 0) continue;
      conn_apply(v, d, net_ptr, vtx_ptr, vtx_nets, src, K, part, ep_ends, ep_parts, ep_cnt, sizes, pc, msg);
    }
  }
  free(gains); free(losses); free(inmax);
  conn_scratch_free(&w);
  return 0;
}

int or_msm_refine_par(int nov, int noe, const unsigned* net_ptr, const unsigned* vtx_ptr, const unsigned* vtx_nets,
                      const int* src, int K, int passes, int R, int* part, int* ep_ends, int* ep_parts, int* ep_cnt,
                      int* sizes, int* pc, int* msg, long long* moves) {
  (void)noe;
  if (R < 1) return 1;
  conn_scratch w;
  if (conn_scratch_init(&w, K)) return 2;
  int* gains = (int*)calloc((size_t)K, sizeof(int));
  int* losses = (int*)calloc((size_t)K, sizeof(int));
  unsigned char* inmax = (unsigned char*)malloc((size_t)K);
  unsigned char* inmax2 = (unsigned char*)malloc((size_t)K);
  int* dec = (int*)malloc(sizeof(int) * (size_t)(R < nov ? R : (nov > 0 ? nov : 1)));  /* a round holds at most nov */
  if (!gains || !losses || !inmax || !inmax2 || !dec) return 2;
  long long applied = 0;
  for (int it = 0; it < passes; it++) {
    for (int k = 0; k < K; k++) { w.mrecv[k] = -1; w.msend[k] = -1; }
    int max_m = msm_max_set(K, msg, inmax);
    for (int r0 = 0; r0 < nov; r0 += R) {
      int r1 = r0 + R < nov ? r0 + R : nov;
      for (int v = r0; v < r1; v++)
        dec[v - r0] = msm_decide(v, net_ptr, vtx_ptr, vtx_nets, src, K, part, ep_ends, ep_parts, ep_cnt, pc, msg,
                                 max_m, inmax, w.mrecv, w.msend, w.recv, w.send, gains, losses);
      for (int v = r0; v < r1; v++) {
        int d = dec[v - r0];
        if (d < 0) continue;
        conn_apply(v, d, net_ptr, vtx_ptr, vtx_nets, src, K, part, ep_ends, ep_parts, ep_cnt, sizes, pc, msg);
        applied++;
        /* MSM/partition.cpp:1335-1370: a new part in the max set ends the round */
        memcpy(inmax2, inmax, (size_t)K);
        max_m = msm_max_set(K, msg, inmax);
        int entered = 0;
        for (int k = 0; k < K; k++) if (inmax[k] && !inmax2[k]) entered = 1;
        if (entered) break;
      }
    }
  }
  if (moves) *moves = applied;
  free(gains); free(losses); free(inmax); free(inmax2); free(dec);
  conn_scratch_free(&w);
  return 0;
}

/* ---------------------------------------------------------------------- */
/* BFS vertex ordering                                                     */
/* ---------------------------------------------------------------------- */

/* One BFS over the vertex -> net -> pin adjacency as topDown
 * (shared/utils.cpp:543-590) runs it on one thread: level by level, each
 * vertex of the level in queue order, its nets in vertex -> nets order, each
 * net's pins in the order of the net view given, an unvisited pin joining the
 * next level. queue[] receives the visit order; returns the vertices reached.
 * A net is expanded the first time only: after that every pin of it is
 * visited, so the reference's later expansions of it find nothing. */
static int bfs_run(int start, int nov, int noe, const unsigned* vtx_ptr, const unsigned* vtx_nets,
                   const unsigned* net_ptr, const unsigned* pins, int* dist, int* queue, unsigned char* done) {
  for (int v = 0; v < nov; v++) dist[v] = -1;
  for (int n = 0; n < noe; n++) done[n] = 0;
  int head = 0, tail = 0;
  dist[start] = 0;
  queue[tail++] = start;
  while (head < tail) {
    const int v = queue[head++];
    for (unsigned j = vtx_ptr[v]; j < vtx_ptr[v + 1]; j++) {
      const unsigned e = vtx_nets[j];
      if (done[e]) continue;
      done[e] = 1;
      for (unsigned k = net_ptr[e]; k < net_ptr[e + 1]; k++) {
        const int u = (int)pins[k];
        if (dist[u] < 0) { dist[u] = dist[v] + 1; queue[tail++] = u; }
      }
    }
  }
  return tail;
}

/* the lowest-index vertex of the largest distance (shared/utils.cpp:466-473) */
static int bfs_farthest(int nov, const int* dist) {
  int far = -1, fd = -1;
  for (int i = 0; i < nov; i++)
    if (dist[i] > fd) { far = i; fd = dist[i]; }
  return far;
}

int or_bfs_ordering(int nov, int noe, const unsigned* vtx_ptr, const unsigned* vtx_nets, const unsigned* net_ptr,
                    const unsigned* pins, int start, int* ordering, int* farthest_first, int* farthest_all) {
  if (nov <= 0 || start < 0 || start >= nov) return 1;
  int* dist = (int*)malloc(sizeof(int) * (size_t)nov);
  unsigned char* seen = (unsigned char*)calloc((size_t)nov, 1);
  unsigned char* done = (unsigned char*)malloc((size_t)(noe > 0 ? noe : 1));
  if (!dist || !seen || !done) return 2;
  bfs_run(start, nov, noe, vtx_ptr, vtx_nets, net_ptr, pins, dist, ordering, done);
  *farthest_first = bfs_farthest(nov, dist);
  bfs_run(*farthest_first, nov, noe, vtx_ptr, vtx_nets, net_ptr, pins, dist, ordering, done);
  *farthest_all = bfs_farthest(nov, dist);
  int n = bfs_run(*farthest_all, nov, noe, vtx_ptr, vtx_nets, net_ptr, pins, dist, ordering, done);
  /* vertices the BFS does not reach (the reference leaves them -1) follow
   * in ascending order */
  for (int i = 0; i < n; i++) seen[ordering[i]] = 1;
  for (int v = 0; v < nov; v++)
    if (!seen[v]) ordering[n++] = v;
  free(dist);
  free(seen);
  free(done);
  return 0;
}
