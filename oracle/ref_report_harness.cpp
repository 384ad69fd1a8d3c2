// ref_report_harness.cpp -- test infrastructure only.  Linked against the
// reference host driver (src/common/AP26.cpp, compiled with its main renamed)
// so that the reference's own ReportSolution (AP26.cpp:549-602) can be called
// on synthetic (k, K, first) reports, e.g. APs holding base-2 pseudoprimes
// (the Non-Solution recursion no shipped fixture reaches).
//
//   ref_report_harness k K first [k K first ...]
// prints result_hash after each call; SOL lines go to ./SOL-AP26.txt.
#include <cstdint>
#include <cstdio>
#include <cstdlib>

extern int result_hash;
void ReportSolution(int AP_Length, int difference, int64_t First_Term);

int main(int argc, char **argv) {
  result_hash = 0;
  for (int i = 1; i + 2 < argc; i += 3) {
    ReportSolution(atoi(argv[i]), atoi(argv[i + 1]), strtoll(argv[i + 2], nullptr, 10));
    printf("hash %d\n", result_hash);
    fflush(stdout);
  }
  return 0;
}
