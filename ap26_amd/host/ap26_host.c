/* ap26_host.c -- host side of the AP26 operator surface (see
 * include/ap26_host.h): state files, ReportSolution, result hash.
 * Behaviour follows src/common/AP26.cpp of the reference, cited per function.
 */
#define _POSIX_C_SOURCE 200809L
#include "ap26_host.h"

#include <inttypes.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <unistd.h>

#include "ap26_consts.h"

static void join_path(char *dst, size_t n, const char *dir, const char *name) {
  if (dir && dir[0])
    snprintf(dst, n, "%s/%s", dir, name);
  else
    snprintf(dst, n, "%s", name);
}

void ap26h_init(ap26_host *h, int32_t KMIN, int32_t KMAX, int32_t SHIFT, const char *dir,
                int32_t min_report) {
  memset(h, 0, sizeof(*h));
  h->KMIN = KMIN;
  h->KMAX = KMAX;
  h->SHIFT = SHIFT;
  h->min_report = min_report > 0 ? min_report : AP26H_DEFAULT_MIN_REPORT;
  h->write_state_a_next = 1;
  join_path(h->state_a, sizeof(h->state_a), dir, "AP26-state.a.txt");
  join_path(h->state_b, sizeof(h->state_b), dir, "AP26-state.b.txt");
  join_path(h->results_name, sizeof(h->results_name), dir, "SOL-AP26.txt");
}

void ap26h_close(ap26_host *h) {
  if (h->results_file) {
    fclose(h->results_file);
    h->results_file = NULL;
  }
  if (h->metrics_file) {
    fclose(h->metrics_file);
    h->metrics_file = NULL;
  }
}

static double wall_seconds(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

static double cpu_seconds(void) {
  struct timespec ts;
  clock_gettime(CLOCK_PROCESS_CPUTIME_ID, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

double ap26h_progress(const ap26_host *h, int32_t ITER) {
  if (h->K_COUNT <= 0) return 1.0;
  return (double)(h->K_DONE * 10 + ITER) / (double)(h->K_COUNT * 10);
}

int ap26h_open_metrics(ap26_host *h, const char *path) {
  if (h->metrics_file) fclose(h->metrics_file);
  h->metrics_file = fopen(path, "a");
  if (!h->metrics_file) {
    fprintf(stderr, "Cannot open %s !!!\n", path);
    return AP26H_E_IO;
  }
  h->t0_wall = wall_seconds();
  h->t0_cpu = cpu_seconds();
  return AP26H_OK;
}

int ap26h_will_search(int32_t K) {
  return (K % 29) && (K % 31) && (K % 37) && (K % 41) && (K % 43) && (K % 47) && (K % 53) &&
         (K % 59);
}

void ap26h_count(ap26_host *h, int32_t K_resume) {
  h->K_COUNT = 0;
  h->K_DONE = 0;
  for (int64_t i = h->KMIN; i <= h->KMAX; i++) {
    if (ap26h_will_search((int32_t)i)) {
      h->K_COUNT++;
      if (K_resume > i) h->K_DONE++;
    }
  }
}

/* ---- state files (AP26.cpp:360-476) ------------------------------------ */

typedef struct {
  int good;
  int K, ITER, hash;
  double trickle;
} state_rec;

static state_rec read_one(const ap26_host *h, const char *name) {
  state_rec s;
  memset(&s, 0, sizeof(s));
  FILE *in = fopen(name, "r");
  if (!in) return s;
  int a, b, c;
  if (fscanf(in, "%d %d %d %d %d %d %lf\n", &a, &b, &c, &s.K, &s.ITER, &s.hash, &s.trickle) != 7) {
    fprintf(stderr, "Cannot parse %s !!!\n", name);
    fclose(in);
    return s;
  }
  fclose(in);
  s.good = (a == h->KMIN && b == h->KMAX && c == h->SHIFT);
  return s;
}

int ap26h_read_state(ap26_host *h, int32_t *K, int32_t *ITER) {
  state_rec A = read_one(h, h->state_a);
  state_rec B = read_one(h, h->state_b);
  if (A.good && B.good) {
    if (A.K > B.K || (A.K == B.K && A.ITER > B.ITER))
      B.good = 0;
    else
      A.good = 0;
  }
  const state_rec *use = A.good ? &A : (B.good ? &B : NULL);
  if (!use) return 0;
  *K = use->K;
  *ITER = use->ITER;
  h->result_hash = use->hash;
  h->last_trickle = (double)(time_t)use->trickle;
  /* write to the other file next (A read -> B next) */
  h->write_state_a_next = A.good ? 0 : 1;
  return 1;
}

void ap26h_write_state(ap26_host *h, int32_t K, int32_t ITER) {
  const char *name = h->write_state_a_next ? h->state_a : h->state_b;
  FILE *out = fopen(name, "w");
  if (!out) {
    fprintf(stderr, "Cannot open %s !!!\n", name);
    return;
  }
  if (fprintf(out, "%d %d %d %d %d %d %lf\n", h->KMIN, h->KMAX, h->SHIFT, K, ITER, h->result_hash,
              h->last_trickle) < 0) {
    fprintf(stderr, "Cannot write to %s !!! Continuing...\n", name);
    fclose(out);
    return;
  }
  if (fclose(out) == 0) h->write_state_a_next = !h->write_state_a_next;
}

void ap26h_checkpoint(ap26_host *h, int32_t K, int32_t force, int32_t ITER) {
  (void)force; /* without BOINC every checkpoint is taken (AP26.cpp:611-633) */
  if (h->results_file) {
    fflush(h->results_file);
    fsync(fileno(h->results_file));
  }
  ap26h_write_state(h, K, ITER);
  if (!h->quiet)
    printf("Checkpoint: KMIN:%d KMAX:%d SHIFT:%d K:%d ITER:%d\n", h->KMIN, h->KMAX, h->SHIFT, K,
           ITER);
  if (h->metrics_file) {
    const double prog = ap26h_progress(h, ITER);
    if (!h->quiet) printf("progress: %.2f%%\n", prog * 100.0);
    fprintf(h->metrics_file,
            "{\"KMIN\": %d, \"KMAX\": %d, \"SHIFT\": %d, \"K\": %d, \"ITER\": %d, \"progress\": %f, "
            "\"cputime\": %f, \"runtime\": %f, \"k_done\": %d, \"k_count\": %d, \"result_hash\": %d, "
            "\"reports\": %" PRId64 ", \"records\": %" PRId64 ", \"non_solutions\": %" PRId64 "}\n",
            h->KMIN, h->KMAX, h->SHIFT, K, ITER, prog, cpu_seconds() - h->t0_cpu, wall_seconds() - h->t0_wall,
            h->K_DONE, h->K_COUNT, h->result_hash, h->n_reports, h->n_records, h->n_nonsol);
    fflush(h->metrics_file);
  }
}

/* ---- verification (AP26.cpp:484-543, PrimeQ_x86.h:114-153) -------------- */

static uint64_t mulmod_u128(uint64_t a, uint64_t b, uint64_t n) {
  return (uint64_t)(((unsigned __int128)a * b) % n);
}

int ap26h_sprp(uint64_t a, uint64_t N) {
  if (N < 3 || !(N & 1)) return 0;
  uint64_t d = N - 1;
  int t = 0;
  while (!(d & 1)) {
    d >>= 1;
    t++;
  }
  uint64_t base = a % N;
  if (base == 0) return 1;
  /* right-to-left: r = a^d */
  uint64_t r = 1, sq = base;
  for (uint64_t e = d; e; e >>= 1) {
    if (e & 1) r = mulmod_u128(r, sq, N);
    sq = mulmod_u128(sq, sq, N);
  }
  if (r == 1 || r == N - 1) return 1;
  for (int s = 1; s < t; s++) {
    r = mulmod_u128(r, r, N);
    if (r == N - 1) return 1;
  }
  return 0;
}

int32_t ap26h_validate(int32_t k, int32_t K, int64_t first, int32_t nbases, int32_t *err) {
  static const uint64_t bases[9] = {2, 3, 5, 7, 11, 13, 17, 19, 23};
  if (err) *err = 0;
  if (first % 2 == 0) return 0;
  int64_t N = first;
  int32_t j;
  for (j = 0; j < k; j++) {
    if (N < 0 || N >= (INT64_C(1) << 62)) {
      fprintf(stderr, "ap26_host.c: modulus out of range.\n");
      if (err) *err = AP26H_E_RANGE;
      return j;
    }
    for (int b = 0; b < nbases; b++)
      if (!ap26h_sprp(bases[b], (uint64_t)N)) return j;
    N += (int64_t)K * AP26_PRIM23;
  }
  return j;
}

/* ---- ReportSolution (AP26.cpp:549-602) ---------------------------------- */

static int open_results(ap26_host *h) {
  if (!h->results_file) h->results_file = fopen(h->results_name, "a");
  if (!h->results_file) {
    fprintf(stderr, "Cannot open %s !!!\n", h->results_name);
    return AP26H_E_IO;
  }
  return AP26H_OK;
}

int ap26h_report(ap26_host *h, int32_t k, int32_t K, int64_t first) {
  h->n_reports++;
  /* every report enters the hash, whatever its length */
  h->result_hash += (int32_t)(first % 1000);
  h->result_hash += k;
  if (h->result_hash > AP26_MAXINTV) h->result_hash -= AP26_MAXINTV;

  int32_t err = 0;
  int32_t i = ap26h_validate(k, K, first, 9, &err);
  if (err) return err;
  if (i < k) {
    h->n_nonsol++;
    if (!h->quiet) printf("Non-Solution: %d %d %" PRId64 "\n", k, K, first);
    int32_t i2 = ap26h_validate(k, K, first, 1, &err);
    if (err) return err;
    if (i2 < k) {
      printf("Error: Computation error, found invalid AP, exiting...\n");
      fprintf(stderr, "Error: Computation error, found invalid AP\n");
      return AP26H_E_INVALID;
    }
    /* the leading i terms and the trailing k-(i+1) terms may still hold APs */
    int rc = ap26h_report(h, i, K, first);
    if (rc) return rc;
    return ap26h_report(h, k - (i + 1), K, first + (int64_t)(i + 1) * K * AP26_PRIM23);
  }
  if (k >= h->min_report) {
    int rc = open_results(h);
    if (rc) return rc;
    if (!h->quiet) printf("Solution: %d %d %" PRId64 "\n", k, K, first);
    if (fprintf(h->results_file, "%d %d %" PRId64 "\n", k, K, first) < 0) {
      fprintf(stderr, "Cannot write to %s !!!\n", h->results_name);
      return AP26H_E_IO;
    }
    h->n_records++;
  }
  return AP26H_OK;
}

/* validate_ap26 / val_base2_ap26 read from per-term bits computed on the
   device: bit (off + j) of mask covers term j of this (sub-)AP.  Same order
   as ap26h_validate: parity, then per term the range check and the test.
   Terms past bit 31 are tested here. */
static int32_t validate_bits(int32_t k, int32_t K, int64_t first, uint32_t mask, int off, int32_t nbases,
                             int32_t *err) {
  static const uint64_t bases[9] = {2, 3, 5, 7, 11, 13, 17, 19, 23};
  if (err) *err = 0;
  if (first % 2 == 0) return 0;
  int64_t N = first;
  int32_t j;
  for (j = 0; j < k; j++) {
    if (N < 0 || N >= (INT64_C(1) << 62)) {
      fprintf(stderr, "ap26_host.c: modulus out of range.\n");
      if (err) *err = AP26H_E_RANGE;
      return j;
    }
    if (off + j < 32) {
      if (!((mask >> (off + j)) & 1u)) return j;
    } else {
      for (int b = 0; b < nbases; b++)
        if (!ap26h_sprp(bases[b], (uint64_t)N)) return j;
    }
    N += (int64_t)K * AP26_PRIM23;
  }
  return j;
}

static int report_bits(ap26_host *h, int32_t k, int32_t K, int64_t first, int off, uint32_t prp2_mask,
                       uint32_t valid_mask) {
  h->n_reports++;
  h->result_hash += (int32_t)(first % 1000);
  h->result_hash += k;
  if (h->result_hash > AP26_MAXINTV) h->result_hash -= AP26_MAXINTV;

  int32_t err = 0;
  int32_t i = validate_bits(k, K, first, valid_mask, off, 9, &err);
  if (err) return err;
  if (i < k) {
    h->n_nonsol++;
    if (!h->quiet) printf("Non-Solution: %d %d %" PRId64 "\n", k, K, first);
    int32_t i2 = validate_bits(k, K, first, prp2_mask, off, 1, &err);
    if (err) return err;
    if (i2 < k) {
      printf("Error: Computation error, found invalid AP, exiting...\n");
      fprintf(stderr, "Error: Computation error, found invalid AP\n");
      return AP26H_E_INVALID;
    }
    int rc = report_bits(h, i, K, first, off, prp2_mask, valid_mask);
    if (rc) return rc;
    return report_bits(h, k - (i + 1), K, first + (int64_t)(i + 1) * K * AP26_PRIM23, off + i + 1, prp2_mask,
                       valid_mask);
  }
  if (k >= h->min_report) {
    int rc = open_results(h);
    if (rc) return rc;
    if (!h->quiet) printf("Solution: %d %d %" PRId64 "\n", k, K, first);
    if (fprintf(h->results_file, "%d %d %" PRId64 "\n", k, K, first) < 0) {
      fprintf(stderr, "Cannot write to %s !!!\n", h->results_name);
      return AP26H_E_IO;
    }
    h->n_records++;
  }
  return AP26H_OK;
}

int ap26h_report_masks(ap26_host *h, int32_t k, int32_t K, int64_t first, uint32_t prp2_mask,
                       uint32_t valid_mask) {
  return report_bits(h, k, K, first, 0, prp2_mask, valid_mask);
}

/* ---- result hash (AP26.cpp:295-357) ------------------------------------- */

uint64_t ap26h_hash_word(int32_t KMIN, int32_t KMAX, int32_t hash) {
  uint64_t minmax = (uint64_t)((int64_t)KMIN + (int64_t)KMAX);
  while (minmax > (uint64_t)AP26_MAXINTV) minmax -= AP26_MAXINTV;
  return (minmax << 32) | (uint64_t)(uint32_t)hash;
}

int32_t ap26h_cap_total(int64_t T) {
  if (T <= 0) return 0;
  return (int32_t)((T - 1) % AP26_MAXINTV) + 1;
}

int ap26h_write_hash(ap26_host *h) {
  int rc = open_results(h);
  if (rc) return rc;
  uint64_t w = ap26h_hash_word(h->KMIN, h->KMAX, h->result_hash);
  if (fprintf(h->results_file, "%016" PRIX64 "\n", w) < 0) {
    fprintf(stderr, "Cannot write to %s !!!\n", h->results_name);
    return AP26H_E_IO;
  }
  fclose(h->results_file);
  h->results_file = NULL;
  return AP26H_OK;
}
