/* ap26_oracle.h -- CPU restatement of the AP26 search, TEST INFRASTRUCTURE ONLY.
 *
 * This is the parity oracle.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it, and only as the checker (or as the
 * timed CPU comparator); the product path (ap26_amd/) never links or calls it.
 *
 * Every function restates the reference algorithm from its definition and
 * cites the reference file:line it follows.  How each part is pinned:
 *   ap26o_will_search, ap26o_kconsts, ap26o_n43_table  closed forms, recomputed
 *       independently in tests (Python big ints) and via the reference fixtures;
 *   ap26o_ok_table, ap26o_okok_table   closed-form definition checked by brute
 *       force (gcd of the 24 AP terms) in tests;
 *   ap26o_sieve_span     brute-force completeness on small spans + the
 *       reference's own survivor count comment (K=366384: 38,838,420) and the
 *       golden SOL files;
 *   ap26o_sprp           textbook Miller-Rabin base b; pinned by trial division
 *       for N < 2^20, known base-2 strong pseudoprimes and primes near 2^62/2^64;
 *   ap26o_walk, ap26o_report, ap26o_hash_word   golden SOL/hash files written by
 *       the reference CPU build compiled from source (tests/golden/).
 */
#ifndef AP26_ORACLE_H
#define AP26_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AP26O_NPRIMES 83      /* primes 61..541 */
#define AP26O_NOK 23693       /* sum of those primes = packed table length */
#define AP26O_NN43 10840      /* n43 starts per K */

typedef struct {
  int32_t K;
  int64_t STEP, n0, S31, S37, S41, S43, S47, S53, S59;
} ap26o_kconst;

typedef struct {
  int32_t k;        /* AP length */
  int32_t K;        /* difference multiplier */
  int64_t first;    /* first term */
} ap26o_sol;

/* K admission (src/common/AP26.cpp:649-653). */
int ap26o_will_search(int32_t K);
/* Per-K constants (src/cpu/cpusse2.cpp:233-243, src/opencl/AP26.h:112-120). */
void ap26o_kconsts(int32_t K, ap26o_kconst *kc);
/* n43 starts in the reference enumeration order (src/opencl/AP26.h:125-136). */
void ap26o_n43_table(const ap26o_kconst *kc, int64_t out[AP26O_NN43]);
/* The 83 sieve primes and their packed-table offsets (src/opencl/kernels/offset.cl). */
const int32_t *ap26o_primes(void);
const int32_t *ap26o_offsets(void);
/* OK_p tables, packed (src/cpu/cpusse2.cpp:31-35 MAKE_OK). */
void ap26o_ok_table(int32_t K, uint8_t out[AP26O_NOK]);
/* OKOK_p tables for one SHIFT, packed (src/cpu/cpusse2.cpp:37-43 MAKE_OKOK). */
void ap26o_okok_table(int32_t K, int32_t SHIFT, uint64_t out[AP26O_NOK]);

/* Strong probable-prime test to base b, textbook (src/common/PrimeQ_x86.h:114-153).
   Defined for odd N >= 3 (any size < 2^64); returns 0 for N < 3 or even N. */
int ap26o_sprp(uint64_t b, uint64_t N);
/* Base-2 test the search uses (src/common/PrimeQ_x86.h:160-169, without the
   2^62 exit). */
int ap26o_prp2(uint64_t N);
/* ap26o_prp2 of count values on `threads` threads */
void ap26o_prp2_array(const uint64_t *n, uint8_t *out, size_t count, int32_t threads);

/* AP walk from sieve survivor n (src/cpu/cpusse2.cpp:570-590): returns k and
   writes the first term; k < 10 means nothing is reported. */
int32_t ap26o_walk(int64_t n, int64_t STEP, int64_t *first);

/* Sieve survivors of shifts [SHIFT0 + 64*it0, SHIFT0 + 64*it1) over n43 indices
   [j0, j1) (src/cpu/cpusse2.cpp:389-621): every n = n59 + (b+SHIFT)*MOD with all 83
   table bits set and no factor 7..23.  Writes up to cap survivors, returns the
   total count (which may exceed cap). */
int64_t ap26o_sieve_span(int32_t K, int32_t SHIFT0, int32_t it0, int32_t it1,
                         int32_t j0, int32_t j1, int64_t *out, int64_t cap);

/* Full search of that span: sieve + walk.  Writes up to cap reports (k >= 10)
   and returns the count; *n_surv gets the number of sieve survivors.
   nthreads > 1 splits the n43 span across threads. */
int64_t ap26o_search_span(int32_t K, int32_t SHIFT0, int32_t it0, int32_t it1,
                          int32_t j0, int32_t j1, ap26o_sol *out, int64_t cap,
                          int64_t *n_surv, int32_t nthreads);

/* Host-side reporting (src/common/AP26.cpp:484-602).  Adds this report's
   contribution(s) to *hash (capped sum), appends the SOL records with
   k >= min_report to recs (up to cap, count in *nrec), counts Non-Solution
   splits in *nonsol.  Returns 0, or -1 where the reference exits ("Computation
   error": the AP is not base-2 PRP). */
int ap26o_report(int32_t k, int32_t K, int64_t first, int32_t *hash,
                 int32_t min_report, ap26o_sol *recs, int64_t cap, int64_t *nrec,
                 int64_t *nonsol);
/* validate_ap26 / val_base2_ap26 (src/common/AP26.cpp:484-543). */
int32_t ap26o_validate(int32_t k, int32_t K, int64_t first, int nbases);

/* Table lookups of the span under the reference GPU kernel's early-exit
   grouping (src/opencl/kernels/sieve.cl:32-114), each shift's mask on its own:
   the algorithmic work count the sieve's roofline is priced on. */
int64_t ap26o_lookup_count(int32_t K, int32_t SHIFT0, int32_t it0, int32_t it1, int32_t j0,
                           int32_t j1);

/* 64-bit result hash word (src/common/AP26.cpp:295-357). */
uint64_t ap26o_hash_word(int32_t KMIN, int32_t KMAX, int32_t hash);

#ifdef __cplusplus
}
#endif
#endif
