// Test infrastructure only (oracle side). Linked into oracle/_ref/cli_tv, the
// reference's own TV driver (TV/tv.cpp, unchanged) whose partition_gpu_wrapper
// resolves to this repository's libfastpart_legacy_tv.so.
//
// The reference writes its final part vector with writePartVector
// (shared/utils.cpp:592-608) into a fixed folder of its author's machine
// (/home/fatih/new_tensors...), which does not exist here. The recipe links
// with -Wl,--wrap=<writePartVector's mangled name>, so the call lands here
// first: with FASTPART_PARTVEC set, the vector is written to that path in the
// reference's format (one part id per line); otherwise the call goes on to
// the real function.
#include <cstdio>
#include <cstdlib>
#include <string>

// The wrapper and the real function are named by the mangled symbol, so they
// take C linkage (no second mangling); the parameters keep the C++ ABI of
// writePartVector(std::string, int, int, int*, std::string).
extern "C" void __real__Z15writePartVectorNSt7__cxx1112basic_stringIcSt11char_traitsIcESaIcEEEiiPiS4_(
    std::string f_name, int nov, int K, int* part_vec, std::string algo_name);

extern "C" void __wrap__Z15writePartVectorNSt7__cxx1112basic_stringIcSt11char_traitsIcESaIcEEEiiPiS4_(
    std::string f_name, int nov, int K, int* part_vec, std::string algo_name) {
  const char* path = std::getenv("FASTPART_PARTVEC");
  if (!path || !*path) {
    __real__Z15writePartVectorNSt7__cxx1112basic_stringIcSt11char_traitsIcESaIcEEEiiPiS4_(f_name, nov, K, part_vec,
                                                                                         algo_name);
    return;
  }
  FILE* f = std::fopen(path, "w");
  if (!f) {
    std::fprintf(stderr, "cli_capture: cannot open %s\n", path);
    std::exit(3);
  }
  for (int i = 0; i < nov; i++) std::fprintf(f, "%d\n", part_vec[i]);
  std::fclose(f);
}
