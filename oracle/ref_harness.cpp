// Test infrastructure only (oracle side). Drives the REFERENCE's own CPU
// partitioner, compiled unchanged from /root/reference/<M>/partition.cpp and
// shared/*.cpp, on an instance file, and records the connectivity state the
// reference reaches after its initial partition and after refinement.
//
// Nothing here is product code; the product never links or runs it.
//
// How the state is observed without touching the reference sources: every
// reference partition()/partition_par() calls the metric helpers
//   TM_scratch(K, part_vec, ep_parts, ep_starts, ep_ends, ...)   (shared/metric.cpp:122)
// right after the initial partition and again after refinement, and
//   TV_Refined(ep_starts, ep_cnt, noe, K)                         (shared/metric.cpp:70)
// after refinement (e.g. TV/partition.cpp:273,449,451). The recipe links with
// -Wl,--wrap=<mangled name> so these calls land here first; we copy the arrays
// and forward to the real function.
//
// The vertex->nets CSR is built exactly as TV/tv.cpp:76-100 does (counting
// sort over the net->pins CSR); ordering is the identity (TV/tv.cpp:122-128)
// unless an ordering file is given.
// net_appearence is given NET degrees (consistent arrays), not the vertex
// degrees tv.cpp passes (TV/tv.cpp:71), so MGV_Scratch's inverse is valid.
//
// Usage: ref_<m> <instance.fphg> <K> <passes> <threads> <out.fprs> [ordering.i32]
//   threads == 0 -> partition()            (serial schedule)
//   threads >= 1 -> partition_par(threads)  (round schedule, OpenMP)
//   ordering.i32: nov raw int32 vertex ids, the visiting order handed to the
//   reference in place of the identity (TV/tv.cpp:122-128 builds the identity
//   or, with -o 1, a BFS order).
//
// Built with -DREF_DROPIN (binaries dropin_<m>): threads == -1 runs the
// reference's partition_gpu() (TV/partition.cpp:984, CN/partition.cpp:1599),
// whose partition_gpu_wrapper then resolves to this repository's
// libfastpart_legacy_<m>.so -- the reference driving the engine through its
// own device entry point. dropin_msm calls the partition_gpu that
// libfastpart_legacy_msm.so adds for MSM (MSM/msm.cpp:67 calls one the
// reference never defines).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>
#include <chrono>

using namespace std;

// The reference's partition.h declares `extern chrono::duration<double, milli> total;`
// (TV/partition.h:13); its driver (TV/tv.cpp:35) defines it. This harness is the driver.
chrono::duration<double, milli> total(0);

#if defined(REF_CN)
void partition(unsigned int*, unsigned int*, unsigned int*, int, int, long long, int*, int*, int, int*,
               string&, int&, int&, unsigned int*, unsigned int*);
void partition_par(unsigned int*, unsigned int*, unsigned int*, int, int, long long, int*, int*, int, int*,
                   string, int, int, int, unsigned int*, unsigned int*);
void partition_gpu(unsigned int*, unsigned int*, unsigned int*, int, int, long long, int*, int*, int, int*,
                   string, int, int, unsigned int*, unsigned int*);
#else
void partition(unsigned int*, unsigned int*, unsigned int*, int, int, long long, int*, int*, int, int*,
               string, int, int);
void partition_par(unsigned int*, unsigned int*, unsigned int*, int, int, long long, int*, int*, int, int*,
                   string, int, int, int);
void partition_gpu(unsigned int*, unsigned int*, unsigned int*, int, int, long long, int*, int*, int, int*,
                   string, int, int);
#endif

extern "C" int __real__Z10TM_scratchiPiS_S_S_iiS_(int, int*, int*, int*, int*, int, int, int*);
extern "C" int __real__Z10TV_RefinedPiS_ii(int*, int*, int, int);

static int g_nov, g_noe;
static long long g_pins;
static int g_K;
static int g_tm_calls = 0;
static vector<int> st_part[2], st_ends[2], st_parts[2], st_cnt;
static bool have_cnt = false;

extern "C" int __wrap__Z10TM_scratchiPiS_S_S_iiS_(int K, int* labels, int* ep_parts, int* ep_starts, int* ep_ends,
                                                    int nov, int noe, int* src) {
  int slot = g_tm_calls == 0 ? 0 : 1;
  st_part[slot].assign(labels, labels + g_nov);
  st_ends[slot].assign(ep_ends, ep_ends + g_noe);
  st_parts[slot].assign(ep_parts, ep_parts + g_pins);
  g_tm_calls++;
  return __real__Z10TM_scratchiPiS_S_S_iiS_(K, labels, ep_parts, ep_starts, ep_ends, nov, noe, src);
}

extern "C" int __wrap__Z10TV_RefinedPiS_ii(int* ep_starts, int* ep_cnt, int noe, int K) {
  st_cnt.assign(ep_cnt, ep_cnt + g_pins);
  have_cnt = true;
  return __real__Z10TV_RefinedPiS_ii(ep_starts, ep_cnt, noe, K);
}

static void die(const char* m) { fprintf(stderr, "ref_harness: %s\n", m); exit(2); }

int main(int argc, char** argv) {
  if (argc != 6 && argc != 7) die("usage: ref_<m> <instance.fphg> <K> <passes> <threads> <out.fprs> [ordering.i32]");
  FILE* f = fopen(argv[1], "rb");
  if (!f) die("cannot open instance");
  char magic[4];
  int ver, nov, noe;
  long long pins;
  if (fread(magic, 1, 4, f) != 4 || memcmp(magic, "FPHG", 4)) die("bad magic");
  if (fread(&ver, 4, 1, f) != 1 || fread(&nov, 4, 1, f) != 1 || fread(&noe, 4, 1, f) != 1 ||
      fread(&pins, 8, 1, f) != 1)
    die("bad header");
  unsigned int* col_ptr = new unsigned int[noe + 1];
  unsigned int* row_ind = new unsigned int[pins];
  int* src_nodes = new int[noe];
  if (fread(col_ptr, 4, noe + 1, f) != (size_t)(noe + 1)) die("short net_ptr");
  if (fread(row_ind, 4, pins, f) != (size_t)pins) die("short net_pins");
  if (fread(src_nodes, 4, noe, f) != (size_t)noe) die("short src");
  fclose(f);
  int K = atoi(argv[2]), passes = atoi(argv[3]), nt = atoi(argv[4]);
  g_nov = nov; g_noe = noe; g_pins = pins; g_K = K;

  // vertex->nets CSR: counting sort as in TV/tv.cpp:76-100
  unsigned int* row_ptr = new unsigned int[nov + 1];
  unsigned int* col_ind = new unsigned int[pins];
  memset(row_ptr, 0, sizeof(unsigned int) * (nov + 1));
  for (int i = 0; i < noe; i++)
    for (unsigned j = col_ptr[i]; j < col_ptr[i + 1]; j++) row_ptr[row_ind[j] + 1]++;
  for (int i = 1; i <= nov; i++) row_ptr[i] += row_ptr[i - 1];
  for (int i = 0; i < noe; i++)
    for (unsigned j = col_ptr[i]; j < col_ptr[i + 1]; j++) col_ind[row_ptr[row_ind[j]]++] = i;
  for (int i = nov; i > 0; i--) row_ptr[i] = row_ptr[i - 1];
  row_ptr[0] = 0;

  unsigned int* net_app = new unsigned int[noe];
  for (int i = 0; i < noe; i++) net_app[i] = col_ptr[i + 1] - col_ptr[i];
  int* edge_w = new int[noe];
  for (int i = 0; i < noe; i++) edge_w[i] = 1;
  int* ordering = new int[nov];
  for (int i = 0; i < nov; i++) ordering[i] = i;
  if (argc == 7) {
    FILE* fo = fopen(argv[6], "rb");
    if (!fo || fread(ordering, 4, nov, fo) != (size_t)nov) die("short ordering file");
    fclose(fo);
    vector<char> seen(nov, 0);
    for (int i = 0; i < nov; i++) {
      if (ordering[i] < 0 || ordering[i] >= nov || seen[ordering[i]]) die("ordering is not a permutation");
      seen[ordering[i]] = 1;
    }
  }
  string fname = string(argv[5]) + ".ref";
  int refi_opt = passes > 0 ? 1 : 0, refi_iter = passes;

  auto t0 = chrono::steady_clock::now();
#if defined(REF_DROPIN) && defined(REF_CN)
  if (nt == -1)
    partition_gpu(row_ptr, col_ind, net_app, nov, noe, pins, edge_w, src_nodes, K, ordering, fname, refi_opt,
                  refi_iter, col_ptr, row_ind);
  else
#elif defined(REF_DROPIN)
  if (nt == -1)
    partition_gpu(row_ptr, col_ind, net_app, nov, noe, pins, edge_w, src_nodes, K, ordering, fname, refi_opt,
                  refi_iter);
  else
#endif
#if defined(REF_CN)
  if (nt == 0)
    partition(row_ptr, col_ind, net_app, nov, noe, pins, edge_w, src_nodes, K, ordering, fname, refi_opt,
              refi_iter, col_ptr, row_ind);
  else
    partition_par(row_ptr, col_ind, net_app, nov, noe, pins, edge_w, src_nodes, K, ordering, fname, refi_opt,
                  refi_iter, nt, col_ptr, row_ind);
#else
  if (nt == 0)
    partition(row_ptr, col_ind, net_app, nov, noe, pins, edge_w, src_nodes, K, ordering, fname, refi_opt,
              refi_iter);
  else
    partition_par(row_ptr, col_ind, net_app, nov, noe, pins, edge_w, src_nodes, K, ordering, fname, refi_opt,
                  refi_iter, nt);
#endif
  auto t1 = chrono::steady_clock::now();
  fflush(stdout);
#if defined(REF_DROPIN) && defined(REF_MSM)
  // MSM's partition_gpu is the library's own (the reference has none): it
  // reports its part vector through FASTPART_PARTVEC, not through the
  // reference's metric helpers
  if (nt == -1) return 0;
#endif
  if (g_tm_calls < 2 || !have_cnt) die("reference did not report both states");

  FILE* o = fopen(argv[5], "wb");
  if (!o) die("cannot open output");
  fwrite("FPRS", 1, 4, o);
  int hdr[3] = {nov, noe, K};
  fwrite(hdr, 4, 3, o);
  fwrite(&pins, 8, 1, o);
  double wall = chrono::duration<double>(t1 - t0).count();
  fwrite(&wall, 8, 1, o);
  for (int s = 0; s < 2; s++) {
    fwrite(st_part[s].data(), 4, nov, o);
    fwrite(st_ends[s].data(), 4, noe, o);
    fwrite(st_parts[s].data(), 4, pins, o);
  }
  fwrite(st_cnt.data(), 4, pins, o);
  fclose(o);
  return 0;
}
