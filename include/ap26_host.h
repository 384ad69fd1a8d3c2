/* ap26_host.h -- the host side of the AP26 operator surface (libap26host.so,
 * linked into the AP26 command): checkpoint state files, result reporting
 * (hash, 9-base verification, Non-Solution splits, SOL records) and the final
 * result hash.  CPU-only C; it never touches the GPU.
 *
 * Each function keeps the behaviour of the reference host driver
 * (src/common/AP26.cpp) that it replaces, cited per declaration.  Unlike the
 * reference, state is held in an ap26_host object instead of globals, and
 * failures return a negative code instead of calling exit() -- the AP26
 * command exits with status 1 on them, as the reference does.
 */
#ifndef AP26_HOST_H
#define AP26_HOST_H

#include <stdint.h>
#include <stdio.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AP26H_OK 0
#define AP26H_E_IO (-10)        /* cannot open/write SOL file */
#define AP26H_E_INVALID (-11)   /* reported AP is not base-2 PRP ("Computation error") */
#define AP26H_E_RANGE (-12)     /* modulus out of range (term >= 2^62) */

#define AP26H_DEFAULT_MIN_REPORT 20 /* MINIMUM_AP_LENGTH_TO_REPORT, AP26.cpp:103-105 */

typedef struct {
  int32_t KMIN, KMAX, SHIFT;
  int32_t K_COUNT, K_DONE;
  int32_t result_hash;      /* capped running hash (AP26.cpp:110, :557-561) */
  double last_trickle;
  int32_t write_state_a_next;
  int32_t min_report;       /* SOL threshold; the hash takes every report */
  int32_t quiet;            /* suppress stdout progress lines */
  int64_t n_reports;        /* ReportSolution calls, recursion included */
  int64_t n_records;        /* SOL lines written */
  int64_t n_nonsol;         /* Non-Solution splits */
  FILE *results_file;
  char state_a[1024], state_b[1024], results_name[1024];
  FILE *metrics_file;       /* progress records, one JSON line per checkpoint */
  double t0_wall, t0_cpu;   /* run start (seconds) for runtime / cputime */
} ap26_host;

/* Initialise for a work unit; dir (may be NULL or "") prefixes the file names
   AP26-state.a.txt, AP26-state.b.txt, SOL-AP26.txt (AP26.cpp:92-98). */
void ap26h_init(ap26_host *h, int32_t KMIN, int32_t KMAX, int32_t SHIFT, const char *dir,
                int32_t min_report);
void ap26h_close(ap26_host *h);

/* K admission (AP26.cpp:649-653). */
int ap26h_will_search(int32_t K);

/* Count searched / already-done K for the unit (AP26.cpp:945-951). */
void ap26h_count(ap26_host *h, int32_t K_resume);

/* read_state (AP26.cpp:396-476): newest of the two state files whose
   KMIN KMAX SHIFT match; 1 and *K, *ITER, hash, trickle set if found. */
int ap26h_read_state(ap26_host *h, int32_t *K, int32_t *ITER);

/* write_state (AP26.cpp:360-390): alternate A/B; the next file flips only
   when the write and close succeed. */
void ap26h_write_state(ap26_host *h, int32_t K, int32_t ITER);

/* checkpoint without BOINC (AP26.cpp:607-645): flush+fsync SOL file, write
   state, print the Checkpoint line. */
void ap26h_checkpoint(ap26_host *h, int32_t K, int32_t force, int32_t ITER);

/* ReportSolution (AP26.cpp:549-602). */
int ap26h_report(ap26_host *h, int32_t k, int32_t K, int64_t first);

/* ReportSolution with the per-term verification done on the device: bit j of
   prp2_mask / valid_mask says term first + j*K*23# is a strong probable prime
   to base 2 / to all 9 bases (ap26_sol's masks).  Same files, hash, messages
   and return codes as ap26h_report, without host SPRP work for terms 0..31. */
int ap26h_report_masks(ap26_host *h, int32_t k, int32_t K, int64_t first, uint32_t prp2_mask,
                       uint32_t valid_mask);

/* Fraction of the unit done at (K_DONE, ITER): (K_DONE*10 + ITER) /
   (K_COUNT*10), 1 for an empty unit (AP26.cpp:636-641, boinc_fraction_done). */
double ap26h_progress(const ap26_host *h, int32_t ITER);

/* Progress records, the trickle-up message's fields (AP26.cpp:251-287:
   progress, cputime, runtime) as metrics: from this call on, every
   checkpoint appends one JSON line to path (append mode) and prints the
   non-BOINC "progress: x.xx%" line (AP26.cpp:240-248) unless quiet. */
int ap26h_open_metrics(ap26_host *h, const char *path);

/* write_hash (AP26.cpp:295-357): append the 16-hex-digit hash line. */
int ap26h_write_hash(ap26_host *h);

/* The 64-bit hash word before hex formatting (AP26.cpp:308-319). */
uint64_t ap26h_hash_word(int32_t KMIN, int32_t KMAX, int32_t hash);

/* Capped-sum rule of the hash (AP26.cpp:557-561) applied to an exact total:
   the value result_hash holds after adding contributions summing to T. */
int32_t ap26h_cap_total(int64_t T);

/* strong_prp(a, N) of PrimeQ_x86.h:114-153 (textbook, exact 128-bit products),
   and validate_ap26 / val_base2_ap26 (AP26.cpp:484-543); nbases = 9 or 1.
   Return AP26H_E_RANGE via *err when a term leaves [0, 2^62). */
int ap26h_sprp(uint64_t a, uint64_t N);
int32_t ap26h_validate(int32_t k, int32_t K, int64_t first, int32_t nbases, int32_t *err);

#ifdef __cplusplus
}
#endif
#endif
