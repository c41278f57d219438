// Test infrastructure only (oracle side). Runs the REFERENCE's BFS ordering
// (shared/utils.cpp:433-587, compiled unchanged) the way TV/tv.cpp:110-128
// drives it: bfsFarthest_queue_par from a start vertex, again from the
// farthest vertex found, then bfsOrdering_queue_par from that one.
//
// Input (binary, little endian): u32 nov, u32 noe, i64 pins,
//   u32 vtx_ptr[nov + 1], u32 vtx_nets[pins]   (vertex -> nets)
//   u32 net_ptr[noe + 1], u32 net_pins[pins]   (net -> pins, the view expanded)
//   i32 start, i32 threads
// Output: i32 farthest_first, i32 farthest_all, i32 ordering[nov]
// (the reference leaves -1 for vertices the BFS does not reach).
//
// Usage: ref_bfs <in.bin> <out.bin>
#include <cstdio>
#include <vector>

#include "utils.h"

template <typename T>
static bool rd(FILE* f, T* p, size_t n) {
  return std::fread(p, sizeof(T), n, f) == n;
}

int main(int argc, char** argv) {
  if (argc != 3) {
    std::fprintf(stderr, "usage: ref_bfs <in.bin> <out.bin>\n");
    return 2;
  }
  FILE* f = std::fopen(argv[1], "rb");
  if (!f) return 2;
  unsigned nov = 0, noe = 0;
  long long pins = 0;
  if (!rd(f, &nov, 1) || !rd(f, &noe, 1) || !rd(f, &pins, 1)) return 2;
  std::vector<unsigned> vptr(nov + 1), vnets((size_t)pins), nptr(noe + 1), npins((size_t)pins);
  int start = 0, threads = 1;
  if (!rd(f, vptr.data(), vptr.size()) || !rd(f, vnets.data(), vnets.size()) || !rd(f, nptr.data(), nptr.size()) ||
      !rd(f, npins.data(), npins.size()) || !rd(f, &start, 1) || !rd(f, &threads, 1))
    return 2;
  std::fclose(f);
  std::vector<int> distance(nov), queue(nov), ordering(nov, -1);
  const int f1 = bfsFarthest_queue_par(start, vptr.data(), vnets.data(), nptr.data(), npins.data(), (int)nov,
                                       distance.data(), queue.data(), threads);
  const int f2 = bfsFarthest_queue_par(f1, vptr.data(), vnets.data(), nptr.data(), npins.data(), (int)nov,
                                       distance.data(), queue.data(), threads);
  bfsOrdering_queue_par(f2, vptr.data(), vnets.data(), nptr.data(), npins.data(), (int)nov, distance.data(),
                        queue.data(), ordering.data(), threads);
  FILE* o = std::fopen(argv[2], "wb");
  if (!o) return 2;
  std::fwrite(&f1, 4, 1, o);
  std::fwrite(&f2, 4, 1, o);
  std::fwrite(ordering.data(), 4, nov, o);
  std::fclose(o);
  return 0;
}
