/* ap26_main.c -- the AP26 command: `AP26 KMIN KMAX SHIFT [options]`, a drop-in
 * for the reference application's operator surface (src/common/AP26.cpp:660-1041):
 * same arguments, same AP26-state.{a,b}.txt checkpoints, same SOL-AP26.txt
 * records and result hash, exit status 1 on failure.  The search itself runs
 * on the GPU through libap26gpu.so; there is no CPU search path.
 *
 * Options (the reference's -compute and CPU-ISA flags are accepted):
 *   -device N      HIP device (default 0)
 *   -report N      SOL threshold, the reference's MINIMUM_AP_LENGTH_TO_REPORT (20)
 *   -shiftckpt     checkpoint after every shift (ITER = shifts done), as the
 *                  reference CPU build does; default is one checkpoint per K
 *   -batch         search the whole unit with the batched persistent engine
 *   -q             no Solution/Checkpoint lines on stdout
 */
#define _POSIX_C_SOURCE 200809L
#include <inttypes.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "ap26_gpu.h"
#include "ap26_host.h"

#define SOL_CAP 65536

static int cmp_sol(const void *a, const void *b) {
  const ap26_sol *x = (const ap26_sol *)a, *y = (const ap26_sol *)b;
  if (x->first != y->first) return x->first < y->first ? -1 : 1;
  return x->k - y->k;
}

static int report_all(ap26_host *h, ap26_sol *sols, int32_t n) {
  /* record order is not part of the contract (AP26.cpp:293); sort for
     deterministic files */
  qsort(sols, (size_t)n, sizeof(ap26_sol), cmp_sol);
  for (int32_t i = 0; i < n; i++) {
    /* the device verified every term (9 bases): ReportSolution on its bits */
    int rc = ap26h_report_masks(h, sols[i].k, sols[i].K, sols[i].first, sols[i].prp2_mask, sols[i].valid_mask);
    if (rc) return rc;
  }
  return 0;
}

int main(int argc, char *argv[]) {
  int device = 0, min_report = AP26H_DEFAULT_MIN_REPORT, shiftckpt = 0, batch = 0, quiet = 0;
  printf("AP26 MI355X (HIP gfx950) 10-shift search, %s\n", ap26_version());
  fprintf(stderr, "Command line: ");
  for (int i = 0; i < argc; i++) fprintf(stderr, "%s ", argv[i]);
  fprintf(stderr, "\n");

  if (argc < 4) {
    printf("Usage: %s KMIN KMAX SHIFT [-device N] [-report N] [-shiftckpt] [-batch] [-q]\n", argv[0]);
    return EXIT_FAILURE;
  }
  int32_t KMIN = 0, KMAX = 0, SHIFT = 0;
  sscanf(argv[1], "%d", &KMIN);
  sscanf(argv[2], "%d", &KMAX);
  sscanf(argv[3], "%d", &SHIFT);
  for (int i = 4; i < argc; i++) {
    if (!strcmp(argv[i], "-compute")) {
      fprintf(stderr, "Compute mode enabled.\n");
    } else if (!strcmp(argv[i], "-device") && i + 1 < argc) {
      device = atoi(argv[++i]);
    } else if (!strcmp(argv[i], "-report") && i + 1 < argc) {
      min_report = atoi(argv[++i]);
    } else if (!strcmp(argv[i], "-shiftckpt")) {
      shiftckpt = 1;
    } else if (!strcmp(argv[i], "-batch")) {
      batch = 1;
    } else if (!strcmp(argv[i], "-q")) {
      quiet = 1;
    } else if (!strcmp(argv[i], "-sse2") || !strcmp(argv[i], "-sse41") || !strcmp(argv[i], "-avx") ||
               !strcmp(argv[i], "-avx2")) {
      fprintf(stderr, "ignoring CPU option %s (GPU build)\n", argv[i]);
    }
  }

  ap26_host h;
  ap26h_init(&h, KMIN, KMAX, SHIFT, NULL, min_report);
  h.quiet = quiet;
  int32_t K, ITER;
  if (ap26h_read_state(&h, &K, &ITER)) {
    printf("Resuming search from checkpoint.\n");
    fprintf(stderr, "Resuming from checkpoint. K: %d ITER: %d\n", K, ITER);
    /* any shift iteration 0..9 can be resumed (a superset of the reference
       CPU builds' ITER rules, AP26.cpp:802-822) */
    if (ITER < 0 || ITER > 9) {
      printf("Checkpoint was from a different version. Restarting workunit.\n");
      fprintf(stderr, "Checkpoint was from a different version. Restarting workunit.\n");
      K = KMIN;
      ITER = 0;
      h.result_hash = 0;
      h.last_trickle = (double)time(NULL);
    }
  } else {
    printf("Beginning a new search with parameters from the command line\n");
    K = KMIN;
    ITER = 0;
    h.result_hash = 0;
    h.last_trickle = (double)time(NULL);
    h.write_state_a_next = 1;
  }

  ap26_ctx *ctx = NULL;
  int rc = ap26_open(device, NULL, &ctx);
  if (rc) {
    fprintf(stderr, "Error: cannot open GPU %d: %s\n", device, ap26_strerror(rc));
    return EXIT_FAILURE;
  }
  char name[64];
  int32_t cus = 0;
  ap26_device_info(ctx, &cus, NULL, NULL, name, sizeof(name));
  fprintf(stderr, "GPU Info:\n  Arch: \t%s\n  CUs: \t%d\n", name, cus);

  ap26h_count(&h, K);
  printf("%d K in this range remaining to be searched (%d skipped, %d done).\n", h.K_COUNT - h.K_DONE,
         KMAX + 1 - KMIN - h.K_COUNT, h.K_DONE);

  ap26_sol *sols = (ap26_sol *)malloc(sizeof(ap26_sol) * SOL_CAP);
  int status = EXIT_SUCCESS;
  if (!sols) return EXIT_FAILURE;

  if (batch && K <= KMAX) {
    /* one unit: the remaining range, resumed shift for its first K */
    ap26_range r = {K, KMAX, SHIFT, ITER};
    int32_t n = 0;
    int64_t tot = 0;
    ap26h_checkpoint(&h, K, 0, ITER);
    rc = ap26_search_batch(ctx, &r, 1, sols, SOL_CAP, &n, &tot, NULL);
    if (rc) {
      fprintf(stderr, "Error: GPU search failed: %s\n", ap26_strerror(rc));
      status = EXIT_FAILURE;
    } else if ((rc = report_all(&h, sols, n)) != 0) {
      status = EXIT_FAILURE;
    } else {
      K = KMAX + 1;
      ITER = 0;
      h.K_DONE = h.K_COUNT;
    }
  } else {
    for (; K <= KMAX && status == EXIT_SUCCESS; ++K) {
      if (!ap26h_will_search(K)) continue;
      ap26h_checkpoint(&h, K, 0, ITER);
      time_t t0 = time(NULL);
      if (shiftckpt) {
        for (int it = ITER; it < 10; it++) {
          int32_t n = 0;
          rc = ap26_search_span(ctx, K, SHIFT, it, it + 1, 0, 10840, sols, SOL_CAP, &n, NULL);
          if (rc) break;
          if ((rc = report_all(&h, sols, n)) != 0) break;
          if (it + 1 < 10) ap26h_checkpoint(&h, K, 0, it + 1);
        }
      } else {
        int32_t n = 0;
        rc = ap26_search_k(ctx, K, SHIFT, ITER, sols, SOL_CAP, &n, NULL);
        if (!rc) rc = report_all(&h, sols, n);
      }
      if (rc) {
        fprintf(stderr, "Error: search of K %d failed: %s\n", K,
                rc <= -10 ? "host reporting error" : ap26_strerror(rc));
        status = EXIT_FAILURE;
        break;
      }
      if (!quiet) printf("Computation of K: %d complete in %d seconds\n", K, (int)(time(NULL) - t0));
      ITER = 0;
      h.K_DONE++;
    }
  }

  if (status == EXIT_SUCCESS) {
    ap26h_checkpoint(&h, K, 1, ITER);
    if (ap26h_write_hash(&h)) status = EXIT_FAILURE;
    else fprintf(stderr, "Workunit complete.\n");
  }
  free(sols);
  ap26h_close(&h);
  ap26_close(ctx);
  return status;
}
