// Test infrastructure only (oracle side). Reads a PaToH-like hypergraph file
// with the REFERENCE's own reader, read_patohlike (shared/hypergraph_rw.cpp:
// 97-242, compiled unchanged), and writes what it returned to a binary file:
//   i32 status (1 = the reader succeeded), u32 nov, u32 noe, i64 pins,
//   u32 col_ptr[noe + 1], u32 row_ind[pins], u32 vertex_appearence[nov]
// The reader itself writes <file>.plbin after a text read and reads that
// cache instead of the text when it exists.
//
// Usage: ref_reader <hypergraph> <out.bin>
#include <cstdio>
#include <string>

#include "hypergraph_rw.h"

int main(int argc, char** argv) {
  if (argc != 3) {
    std::fprintf(stderr, "usage: ref_reader <hypergraph> <out.bin>\n");
    return 2;
  }
  unsigned int *col_ptr = nullptr, *row_ind = nullptr, *app = nullptr;
  unsigned int nov = 0, noe = 0;
  long long pins = 0;
  const bool ok = read_patohlike(std::string(argv[1]), &col_ptr, &row_ind, &app, nov, noe, pins);
  FILE* f = std::fopen(argv[2], "wb");
  if (!f) return 2;
  const int st = ok ? 1 : 0;
  std::fwrite(&st, 4, 1, f);
  std::fwrite(&nov, 4, 1, f);
  std::fwrite(&noe, 4, 1, f);
  std::fwrite(&pins, 8, 1, f);
  if (ok) {
    std::fwrite(col_ptr, 4, (size_t)noe + 1, f);
    std::fwrite(row_ind, 4, (size_t)pins, f);
    std::fwrite(app, 4, nov, f);
  }
  std::fclose(f);
  return 0;
}
