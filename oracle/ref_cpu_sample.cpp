// ref_cpu_sample.cpp -- test and baseline infrastructure only.  Times the
// reference CPU build's own AVX2 search (Search_avx2, src/cpu/cpuavx2.cpp:
// 72-870, compiled from the reference's sources by oracle/Makefile_ref) on a
// bounded sample of one K.  The reference reports progress once per (i31,
// i37) block of the n43 enumeration (src/cpu/cpuavx2.cpp:379-388): 85 blocks
// per group of 4 shifts, each 4/850 of a K while the first group (shifts
// 0..3) runs.  This harness provides that progress callback (the reference
// host driver is compiled with its own renamed), notes the time of every
// block boundary, and stops after the requested seconds:
//
//   ref_cpu_sample K SHIFT SECONDS [ITER]
// prints "blocks B seconds T" (B whole blocks in T seconds).  ITER (default
// 0) starts the search at shift group ITER/4 like a resumed checkpoint: 8
// times the last group, whose vectors carry shifts 8 and 9 and two shifts
// past the K's ten, which the reference masks out.
// SOL lines the search finds go to ./SOL-AP26.txt, as in the reference.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <unistd.h>

void Search_avx2(int K, int startSHIFT, int ITER, int K_COUNT, int K_DONE);

static double g_limit = 20.0;
static double g_t0 = -1.0, g_tlast = 0.0;
static int g_blocks = -1;

static double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

/* called by Search_avx2 at the start of every (i31, i37) block */
void Progress(double) {
  const double t = now_s();
  if (g_blocks < 0) {
    g_t0 = t;
    g_blocks = 0;
    return;
  }
  g_blocks++;
  g_tlast = t;
  if (t - g_t0 >= g_limit || g_blocks >= 84) { /* stay inside the group the search started in */
    printf("blocks %d seconds %.6f\n", g_blocks, g_tlast - g_t0);
    fflush(stdout);
    _exit(0);
  }
}

int main(int argc, char **argv) {
  if (argc < 4) {
    fprintf(stderr, "usage: %s K SHIFT SECONDS [ITER]\n", argv[0]);
    return 2;
  }
  const int K = atoi(argv[1]), SHIFT = atoi(argv[2]);
  const int ITER = argc > 4 ? atoi(argv[4]) : 0;
  if (ITER != 0 && ITER != 4 && ITER != 8) {
    fprintf(stderr, "ITER must be 0, 4 or 8 (a shift group boundary)\n");
    return 2;
  }
  g_limit = atof(argv[3]);
  Search_avx2(K, SHIFT, ITER, 1, 0);
  printf("blocks %d seconds %.6f\n", g_blocks, g_tlast - g_t0);
  return 0;
}
