/* ap26_main.c -- the AP26 command: `AP26 KMIN KMAX SHIFT [options]`, a drop-in
 * for the reference application's operator surface (src/common/AP26.cpp:660-1041):
 * same arguments, same AP26-state.{a,b}.txt checkpoints, same SOL-AP26.txt
 * records and result hash, exit status 1 on failure.  The search itself runs
 * on the GPU through libap26gpu.so; there is no CPU search path.
 *
 * Options (the reference's -compute and CPU-ISA flags are accepted):
 *   -device N      HIP device (default 0)
 *   -report N      SOL threshold, the reference's MINIMUM_AP_LENGTH_TO_REPORT (20)
 *   -shiftckpt     checkpoint inside each K at ITER 4 and 8, the reference
 *                  AVX2 build's cadence (src/cpu/cpuavx2.cpp:858-861); the
 *                  spans [0,4), [4,8), [8,10) keep the shift pairs of the
 *                  sieve whole; default is one checkpoint per K
 *   -batch         search the unit with the batched persistent engine, 256
 *                  searched K per call (AP26_BATCH_K_PER_CALL overrides); the
 *                  state files, records and hash line are those of the per-K
 *                  run, one checkpoint per K
 *   -q             no Solution/Checkpoint lines on stdout
 *   -metrics FILE  append a progress record (progress, cputime, runtime, K,
 *                  ITER, counts: the trickle-up fields as metrics) to FILE at
 *                  every checkpoint; with -manifest, FILE inside each unit's
 *                  directory
 *
 * Batch manifest: `AP26 -manifest FILE [options]` reads work units, one
 * "KMIN KMAX SHIFT" per line ('#' starts a comment), and runs them together on
 * the batched engine.  Unit u keeps its files in ./unit_<KMIN>_<KMAX>_<SHIFT>/:
 * the AP26-state.{a,b}.txt and SOL-AP26.txt that `AP26 KMIN KMAX SHIFT` run in
 * that directory writes (same checkpoint sequence, records and hash line), so
 * each unit resumes from its own checkpoint and compares with a single run.
 */
#define _POSIX_C_SOURCE 200809L
#include <inttypes.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <errno.h>
#include <sys/stat.h>
#include <time.h>

#include "ap26_gpu.h"
#include "ap26_host.h"

#define SOL_CAP 65536
#define BATCH_K_PER_CALL 256     /* -batch: searched K per batched call */
#define BATCH_SOL_CAP (1 << 20)  /* -batch: report buffer per call */

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

static int cmp_sol_unit(const void *a, const void *b) {
  const ap26_sol *x = (const ap26_sol *)a, *y = (const ap26_sol *)b;
  if (x->range_id != y->range_id) return x->range_id < y->range_id ? -1 : 1;
  if (x->K != y->K) return x->K < y->K ? -1 : 1;
  return cmp_sol(a, b);
}

/* The reference CPU build's PrimeQ exits on a walked term >= 2^62
   (src/common/PrimeQ_x86.h:160-166); the library reports that case as
   AP26_E_RANGE and the command exits with the reference's message. */
static void walk_limit_message(int32_t K, int32_t SHIFT) {
  fprintf(stderr, "PrimeQ: modulus out of range.\n");
  fprintf(stderr, "Error: K %d at SHIFT %d walks terms at or past 2^62 (the reference's PrimeQ limit)\n", K,
          SHIFT);
}

/* One K of a unit from shift iteration ITER: checkpoint, search, report.
   Returns 0, AP26_E_RANGE (the walk reached 2^62: the reference would exit
   here), another AP26_E_* code, or a host reporting error. */
static int search_one_k(ap26_ctx *ctx, ap26_host *h, int32_t K, int32_t ITER, int shiftckpt, ap26_sol *sols) {
  ap26h_checkpoint(h, K, 0, ITER);
  int rc;
  if (shiftckpt) {
    for (int it = ITER; it < 10;) {
      const int it1 = it < 8 ? (it / 4 + 1) * 4 : 10;
      int32_t n = 0;
      rc = ap26_search_span(ctx, K, h->SHIFT, it, it1, 0, 10840, sols, SOL_CAP, &n, NULL);
      if (!rc) rc = report_all(h, sols, n);
      if (rc) return rc;
      if (it1 < 10) ap26h_checkpoint(h, K, 0, it1);
      it = it1;
    }
    return 0;
  }
  int32_t n = 0;
  rc = ap26_search_k(ctx, K, h->SHIFT, ITER, sols, SOL_CAP, &n, NULL);
  if (!rc) rc = report_all(h, sols, n);
  return rc;
}

/* Open a unit's state as AP26.cpp:789-832 does: resume (K, ITER) from the
   newest state file, or start at KMIN. */
static void unit_begin(ap26_host *h, int32_t *K, int32_t *ITER) {
  if (ap26h_read_state(h, K, ITER)) {
    printf("Resuming search from checkpoint.\n");
    fprintf(stderr, "Resuming from checkpoint. K: %d ITER: %d\n", *K, *ITER);
    /* any shift iteration 0..9 can be resumed (a superset of the reference
       CPU builds' ITER rules, AP26.cpp:802-822) */
    if (*ITER < 0 || *ITER > 9) {
      printf("Checkpoint was from a different version. Restarting workunit.\n");
      fprintf(stderr, "Checkpoint was from a different version. Restarting workunit.\n");
      *K = h->KMIN;
      *ITER = 0;
      h->result_hash = 0;
      h->last_trickle = (double)time(NULL);
    }
  } else {
    printf("Beginning a new search with parameters from the command line\n");
    *K = h->KMIN;
    *ITER = 0;
    h->result_hash = 0;
    h->last_trickle = (double)time(NULL);
    h->write_state_a_next = 1;
  }
}

typedef struct {
  ap26_host h;
  int32_t K, ITER, Kend;
  int done;
} unit_t;

/* last checkpoint and hash line of a finished manifest unit */
static int finish_unit(unit_t *x) {
  int status = EXIT_SUCCESS;
  ap26h_checkpoint(&x->h, x->K, 1, 0);
  if (ap26h_write_hash(&x->h)) status = EXIT_FAILURE;
  else fprintf(stderr, "Workunit %d %d %d complete.\n", x->h.KMIN, x->h.KMAX, x->h.SHIFT);
  ap26h_close(&x->h);
  x->done = 1;
  return status;
}

#define MANIFEST_SOL_CAP (1 << 20)
#define MANIFEST_K_PER_UNIT 16  /* searched K per unit per batched call */
#define MANIFEST_K_PER_CALL 256 /* searched K per batched call */

/* The batch manifest: every unit advances MANIFEST_K_PER_UNIT searched K per
   batched call; the host then replays the single-unit loop for those K
   (checkpoint before each K, its reports in order, K_DONE) from the call's
   results, and closes a unit (final checkpoint, hash line) when its range
   is done. */
static int run_manifest(ap26_ctx *ctx, const char *path, int32_t min_report, int quiet, const char *metrics) {
  FILE *f = fopen(path, "r");
  if (!f) {
    fprintf(stderr, "Cannot open manifest %s\n", path);
    return EXIT_FAILURE;
  }
  unit_t *u = NULL;
  int nu = 0, cap_u = 0, status = EXIT_SUCCESS;
  char line[512];
  while (fgets(line, sizeof(line), f)) {
    char *c = line;
    while (*c == ' ' || *c == '\t') c++;
    if (*c == '#' || *c == '\n' || *c == '\r' || *c == 0) continue;
    int32_t a, b, sh;
    if (sscanf(c, "%d %d %d", &a, &b, &sh) != 3 || a < 1 || b < a || sh < 0) {
      fprintf(stderr, "Bad manifest line: %s", line);
      status = EXIT_FAILURE;
      break;
    }
    if (nu == cap_u) {
      cap_u = cap_u ? 2 * cap_u : 16;
      unit_t *nuv = (unit_t *)realloc(u, sizeof(unit_t) * (size_t)cap_u);
      if (!nuv) {
        status = EXIT_FAILURE;
        break;
      }
      u = nuv;
    }
    char dir[128];
    snprintf(dir, sizeof(dir), "unit_%d_%d_%d", a, b, sh);
    if (mkdir(dir, 0777) != 0 && errno != EEXIST) {
      fprintf(stderr, "Cannot create %s\n", dir);
      status = EXIT_FAILURE;
      break;
    }
    unit_t *x = &u[nu++];
    memset(x, 0, sizeof(*x));
    ap26h_init(&x->h, a, b, sh, dir, min_report);
    x->h.quiet = quiet;
    unit_begin(&x->h, &x->K, &x->ITER);
    ap26h_count(&x->h, x->K);
  }
  fclose(f);
  ap26_sol *sols = status == EXIT_SUCCESS ? (ap26_sol *)malloc(sizeof(ap26_sol) * MANIFEST_SOL_CAP) : NULL;
  ap26_range *r = status == EXIT_SUCCESS ? (ap26_range *)malloc(sizeof(ap26_range) * (size_t)(nu + 1)) : NULL;
  int *owner = status == EXIT_SUCCESS ? (int *)malloc(sizeof(int) * (size_t)(nu + 1)) : NULL;
  if (status == EXIT_SUCCESS && (!sols || !r || !owner)) status = EXIT_FAILURE;
  printf("%d work units in %s\n", nu, path);
  int failed = 0; /* a unit stopped at the reference's 2^62 limit */
  while (status == EXIT_SUCCESS) {
    int nr = 0, kcall = 0;
    for (int i = 0; i < nu && kcall < MANIFEST_K_PER_CALL; i++) {
      unit_t *x = &u[i];
      if (x->done) continue;
      int cnt = 0;
      int32_t k2 = x->K, last = x->K - 1;
      for (; k2 <= x->h.KMAX && cnt < MANIFEST_K_PER_UNIT; k2++)
        if (ap26h_will_search(k2)) {
          cnt++;
          last = k2;
        }
      x->Kend = cnt ? last : x->h.KMAX;
      if (metrics && !x->h.metrics_file) {
        /* opened when the unit starts and closed when it ends, so a long
           manifest holds files only for the units in flight */
        char mp[1024];
        snprintf(mp, sizeof(mp), "unit_%d_%d_%d/%s", x->h.KMIN, x->h.KMAX, x->h.SHIFT, metrics);
        if (ap26h_open_metrics(&x->h, mp)) {
          status = EXIT_FAILURE;
          break;
        }
      }
      if (cnt) {
        r[nr] = (ap26_range){x->K, x->Kend, x->h.SHIFT, x->ITER};
        owner[nr++] = i;
        kcall += cnt;
      } else {
        owner[nr] = i; /* nothing left to search: close below */
        r[nr++] = (ap26_range){x->K, x->K - 1, x->h.SHIFT, 0};
      }
    }
    if (status != EXIT_SUCCESS || nr == 0) break;
    int32_t n = 0;
    int rc = ap26_search_batch(ctx, r, nr, sols, MANIFEST_SOL_CAP, &n, NULL, NULL);
    if (rc == AP26_E_RANGE) {
      /* a walk of some unit reached 2^62: search this call's units K by K,
         so each stops where its own `AP26 KMIN KMAX SHIFT` run would */
      for (int ri = 0; ri < nr && status == EXIT_SUCCESS; ri++) {
        unit_t *x = &u[owner[ri]];
        int urc = 0;
        for (int32_t K = x->K; K <= x->Kend && r[ri].kmax >= r[ri].kmin && !urc; K++) {
          if (!ap26h_will_search(K)) continue;
          urc = search_one_k(ctx, &x->h, K, x->ITER, 0, sols);
          if (urc == AP26_E_RANGE) walk_limit_message(K, x->h.SHIFT);
          if (!urc) {
            x->ITER = 0;
            x->h.K_DONE++;
          }
        }
        if (urc) {
          fprintf(stderr, "Workunit %d %d %d stopped.\n", x->h.KMIN, x->h.KMAX, x->h.SHIFT);
          ap26h_close(&x->h);
          x->done = 1;
          failed = 1;
          if (urc != AP26_E_RANGE) status = EXIT_FAILURE;
          continue;
        }
        x->K = x->Kend + 1;
        if (x->K > x->h.KMAX) status = finish_unit(x);
      }
      continue;
    }
    if (rc) {
      fprintf(stderr, "Error: GPU search failed: %s\n", ap26_strerror(rc));
      status = EXIT_FAILURE;
      break;
    }
    qsort(sols, (size_t)n, sizeof(ap26_sol), cmp_sol_unit);
    int32_t s = 0;
    for (int ri = 0; ri < nr && status == EXIT_SUCCESS; ri++) {
      unit_t *x = &u[owner[ri]];
      for (int32_t K = x->K; K <= x->Kend && r[ri].kmax >= r[ri].kmin; K++) {
        if (!ap26h_will_search(K)) continue;
        ap26h_checkpoint(&x->h, K, 0, x->ITER);
        for (; s < n && sols[s].range_id == ri && sols[s].K == K; s++) {
          rc = ap26h_report_masks(&x->h, sols[s].k, sols[s].K, sols[s].first, sols[s].prp2_mask,
                                  sols[s].valid_mask);
          if (rc) break;
        }
        if (rc) {
          fprintf(stderr, "Error: reporting for unit %d..%d failed\n", x->h.KMIN, x->h.KMAX);
          status = EXIT_FAILURE;
          break;
        }
        x->ITER = 0;
        x->h.K_DONE++;
      }
      x->K = x->Kend + 1;
      if (status == EXIT_SUCCESS && x->K > x->h.KMAX) status = finish_unit(x);
    }
  }
  for (int i = 0; i < nu; i++)
    if (!u[i].done) ap26h_close(&u[i].h);
  free(sols);
  free(r);
  free(owner);
  free(u);
  return failed ? EXIT_FAILURE : status;
}

int main(int argc, char *argv[]) {
  int device = 0, min_report = AP26H_DEFAULT_MIN_REPORT, shiftckpt = 0, batch = 0, quiet = 0;
  printf("AP26 MI355X (HIP gfx950) 10-shift search, %s\n", ap26_version());
  fprintf(stderr, "Command line: ");
  for (int i = 0; i < argc; i++) fprintf(stderr, "%s ", argv[i]);
  fprintf(stderr, "\n");

  const char *manifest = NULL, *metrics = NULL;
  for (int i = 1; i + 1 < argc; i++) {
    if (!strcmp(argv[i], "-manifest")) manifest = argv[i + 1];
    if (!strcmp(argv[i], "-metrics")) metrics = argv[i + 1];
  }
  if (manifest) {
    for (int i = 1; i < argc; i++) {
      if (!strcmp(argv[i], "-device") && i + 1 < argc) device = atoi(argv[++i]);
      else if (!strcmp(argv[i], "-report") && i + 1 < argc) min_report = atoi(argv[++i]);
      else if (!strcmp(argv[i], "-q")) quiet = 1;
    }
    ap26_ctx *mctx = NULL;
    int mrc = ap26_open(device, NULL, &mctx);
    if (mrc) {
      fprintf(stderr, "Error: cannot open GPU %d: %s\n", device, ap26_strerror(mrc));
      return EXIT_FAILURE;
    }
    int st = run_manifest(mctx, manifest, min_report, quiet, metrics);
    ap26_close(mctx);
    return st;
  }
  if (argc < 4) {
    printf("Usage: %s KMIN KMAX SHIFT [-device N] [-report N] [-shiftckpt] [-batch] [-q] [-metrics FILE]\n"
           "       %s -manifest FILE [-device N] [-report N] [-q] [-metrics FILE]\n", argv[0], argv[0]);
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
    } else if (!strcmp(argv[i], "-metrics") && i + 1 < argc) {
      i++; /* taken above */
    } else if (!strcmp(argv[i], "-sse2") || !strcmp(argv[i], "-sse41") || !strcmp(argv[i], "-avx") ||
               !strcmp(argv[i], "-avx2")) {
      fprintf(stderr, "ignoring CPU option %s (GPU build)\n", argv[i]);
    }
  }

  ap26_host h;
  ap26h_init(&h, KMIN, KMAX, SHIFT, NULL, min_report);
  h.quiet = quiet;
  int32_t K, ITER;
  unit_begin(&h, &K, &ITER);
  if (metrics && ap26h_open_metrics(&h, metrics)) return EXIT_FAILURE;

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

  ap26_sol *sols = (ap26_sol *)malloc(sizeof(ap26_sol) * (batch ? BATCH_SOL_CAP : SOL_CAP));
  int status = EXIT_SUCCESS;
  if (!sols) return EXIT_FAILURE;

  if (batch) {
    /* The batched engine on chunks of up to BATCH_K_PER_CALL searched K (the
       first resumed at ITER).  Each chunk's reports are replayed K by K as
       the per-K loop below writes them (checkpoint before the K, its reports
       sorted by first term, K_DONE), so the state files advance per K and an
       interrupted run resumes at the K after the last replayed one.  A chunk
       whose reports exceed the buffer is searched again in halves. */
    int32_t per_call = BATCH_K_PER_CALL;
    const char *env = getenv("AP26_BATCH_K_PER_CALL");
    if (env && atoi(env) > 0) per_call = atoi(env);
    while (K <= KMAX && status == EXIT_SUCCESS) {
      int cnt = 0;
      int32_t Kend = K - 1;
      for (int32_t k2 = K; k2 <= KMAX && cnt < per_call; k2++)
        if (ap26h_will_search(k2)) {
          cnt++;
          Kend = k2;
        }
      if (cnt == 0) {
        K = KMAX + 1;
        ITER = 0;
        break;
      }
      ap26_range r = {K, Kend, SHIFT, ITER};
      int32_t n = 0;
      rc = ap26_search_batch(ctx, &r, 1, sols, BATCH_SOL_CAP, &n, NULL, NULL);
      if (rc == AP26_E_OVERFLOW && per_call > 1) {
        per_call = (per_call + 1) / 2;
        continue;
      }
      if (rc == AP26_E_RANGE) {
        /* a walk in this chunk reached 2^62: go on K by K (below), which
           stops at the K where the reference would */
        batch = 0;
        break;
      }
      if (rc) {
        fprintf(stderr, "Error: GPU search failed: %s\n", ap26_strerror(rc));
        status = EXIT_FAILURE;
        break;
      }
      qsort(sols, (size_t)n, sizeof(ap26_sol), cmp_sol_unit);
      int32_t s = 0;
      for (int32_t k = K; k <= Kend && status == EXIT_SUCCESS; k++) {
        if (!ap26h_will_search(k)) continue;
        ap26h_checkpoint(&h, k, 0, ITER);
        for (; s < n && sols[s].K == k; s++) {
          rc = ap26h_report_masks(&h, sols[s].k, sols[s].K, sols[s].first, sols[s].prp2_mask,
                                  sols[s].valid_mask);
          if (rc) {
            status = EXIT_FAILURE;
            break;
          }
        }
        ITER = 0;
        if (status == EXIT_SUCCESS) h.K_DONE++;
      }
      if (status != EXIT_SUCCESS) {
        fprintf(stderr, "Error: reporting for K %d..%d failed\n", K, Kend);
        break;
      }
      K = Kend + 1;
    }
  }
  if (!batch) {
    for (; K <= KMAX && status == EXIT_SUCCESS; ++K) {
      if (!ap26h_will_search(K)) continue;
      time_t t0 = time(NULL);
      rc = search_one_k(ctx, &h, K, ITER, shiftckpt, sols);
      if (rc == AP26_E_RANGE) {
        walk_limit_message(K, SHIFT);
        status = EXIT_FAILURE;
        break;
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
