/* ap26_oracle_main.c -- command-line driver for the oracle (test
 * infrastructure only): search one K over a span of shifts and n43 starts and
 * print the raw reports ("k K first") and the survivor count.
 *
 *   ap26_oracle K SHIFT [it0 it1 j0 j1 threads]
 */
#include <stdio.h>
#include <stdlib.h>
#include <time.h>
#include "ap26_oracle.h"

int main(int argc, char **argv) {
  if (argc < 3) {
    fprintf(stderr, "usage: %s K SHIFT [it0 it1 j0 j1 threads]\n", argv[0]);
    return 1;
  }
  int K = atoi(argv[1]), SHIFT = atoi(argv[2]);
  int it0 = argc > 3 ? atoi(argv[3]) : 0, it1 = argc > 4 ? atoi(argv[4]) : 10;
  int j0 = argc > 5 ? atoi(argv[5]) : 0, j1 = argc > 6 ? atoi(argv[6]) : AP26O_NN43;
  int th = argc > 7 ? atoi(argv[7]) : 1;
  static ap26o_sol sols[1 << 16];
  int64_t nsurv = 0;
  struct timespec a, b;
  clock_gettime(CLOCK_MONOTONIC, &a);
  int64_t n = ap26o_search_span(K, SHIFT, it0, it1, j0, j1, sols, 1 << 16, &nsurv, th);
  clock_gettime(CLOCK_MONOTONIC, &b);
  for (int64_t i = 0; i < n && i < (1 << 16); i++)
    printf("%d %d %lld\n", sols[i].k, sols[i].K, (long long)sols[i].first);
  fprintf(stderr, "reports %lld survivors %lld seconds %.3f\n", (long long)n, (long long)nsurv,
          (b.tv_sec - a.tv_sec) + 1e-9 * (b.tv_nsec - a.tv_nsec));
  return 0;
}
