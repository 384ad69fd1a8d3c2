/* ap26_oracle.c -- CPU restatement of the AP26 search.  TEST INFRASTRUCTURE
 * ONLY: loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg as the checker / CPU comparator; never by the product (ap26_amd/).
 *
 * Written from the reference's definitions, not from its vector code.
 * Reference files cited below are relative to the ibethune/ap26 tree.
 */
#include "ap26_oracle.h"

#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* Wroblewski's constants (src/common/CONST.H:7-27).  They define the search
   space, so they are reproduced as data. */
#define O_PRIM23 INT64_C(223092870)
#define O_MOD INT64_C(258559632607830)
#define O_N0 INT64_C(106990415896110)
#define O_N30 INT64_C(94805198622871)
#define O_S3 INT64_C(172373088405220)
#define O_S5 INT64_C(51711926521566)
static const int64_t O_PRES[7] = {
    INT64_C(16681266619860),  /* 31 */
    INT64_C(181690552643340), /* 37 */
    INT64_C(132432982555230), /* 41 */
    INT64_C(126273308948010), /* 43 */
    INT64_C(115526644356690), /* 47 */
    INT64_C(48784836341100),  /* 53 */
    INT64_C(100794433050510)  /* 59 */
};

static int32_t g_primes[AP26O_NPRIMES];
static int32_t g_offsets[AP26O_NPRIMES];
static uint64_t g_magic[AP26O_NPRIMES]; /* floor(2^64/p)+1: exact n mod p for n < 2^48 */
static pthread_once_t g_once = PTHREAD_ONCE_INIT;

static void init_primes(void) {
  int n = 0, off = 0;
  for (int p = 61; p <= 541; p += 2) {
    int isp = 1;
    for (int d = 3; d * d <= p; d += 2)
      if (p % d == 0) { isp = 0; break; }
    if (!isp) continue;
    g_primes[n] = p;
    g_offsets[n] = off;
    g_magic[n] = UINT64_MAX / (uint64_t)p + 1;
    off += p;
    n++;
  }
}

const int32_t *ap26o_primes(void) { pthread_once(&g_once, init_primes); return g_primes; }
const int32_t *ap26o_offsets(void) { pthread_once(&g_once, init_primes); return g_offsets; }

static inline uint32_t mod_small(uint64_t n, int i) {
  /* n < 2^48: the quotient estimate is exact (error < 2^-16 < 1/p). */
  uint64_t q = (uint64_t)(((unsigned __int128)n * g_magic[i]) >> 64);
  return (uint32_t)(n - q * (uint64_t)g_primes[i]);
}

/* src/common/AP26.cpp:649-653 */
int ap26o_will_search(int32_t K) {
  static const int f[8] = {29, 31, 37, 41, 43, 47, 53, 59};
  for (int i = 0; i < 8; i++)
    if (K % f[i] == 0) return 0;
  return 1;
}

static int64_t mulmod_mod(int64_t a, int64_t K) {
  return (int64_t)(((__int128)a * K) % O_MOD);
}

/* src/cpu/cpusse2.cpp:233-243 (= src/opencl/AP26.h:112-120).  The reference
   splits K = 17835q + r to stay inside int64; exact 128-bit products give the
   same values for every K where the split does not overflow (K < ~6.3e8). */
void ap26o_kconsts(int32_t K, ap26o_kconst *kc) {
  kc->K = K;
  kc->STEP = (int64_t)K * O_PRIM23;
  kc->n0 = (mulmod_mod(O_N0, K) + O_N30) % O_MOD;
  kc->S31 = mulmod_mod(O_PRES[0], K);
  kc->S37 = mulmod_mod(O_PRES[1], K);
  kc->S41 = mulmod_mod(O_PRES[2], K);
  kc->S43 = mulmod_mod(O_PRES[3], K);
  kc->S47 = mulmod_mod(O_PRES[4], K);
  kc->S53 = mulmod_mod(O_PRES[5], K);
  kc->S59 = mulmod_mod(O_PRES[6], K);
}

/* src/opencl/AP26.h:125-136 (CPU inline form src/cpu/cpusse2.cpp:389-407):
   residue classes mod 31, 37, 41 under the three guards, times 2 (S3) x 4 (S5). */
void ap26o_n43_table(const ap26o_kconst *kc, int64_t out[AP26O_NN43]) {
  int c = 0;
  for (int i31 = 0; i31 < 7; i31++)
    for (int i37 = 0; i37 < 13; i37++) {
      if (!(i37 - i31 <= 10 && i31 - i37 <= 4)) continue;
      for (int i41 = 0; i41 < 17; i41++) {
        if (!(i41 - i31 <= 14 && i41 - i37 <= 14 && i31 - i41 <= 4 && i37 - i41 <= 10))
          continue;
        for (int i3 = 0; i3 < 2; i3++)
          for (int i5 = 0; i5 < 4; i5++) {
            int64_t v = kc->n0 + i3 * O_S3 + i5 * O_S5 + i31 * kc->S31 + i37 * kc->S37 +
                        i41 * kc->S41;
            if (c < AP26O_NN43) out[c] = v % O_MOD;
            c++;
          }
      }
    }
}

/* OK_p[r] == 0 iff some term n + i*STEP, i in [0,23], is divisible by p when
   n == r (mod p); the reference marks r = j*STEP mod p for j = p-23..p
   (src/cpu/cpusse2.cpp:31-35, src/opencl/kernels/setok.cl:20-22). */
void ap26o_ok_table(int32_t K, uint8_t out[AP26O_NOK]) {
  pthread_once(&g_once, init_primes);
  int64_t STEP = (int64_t)K * O_PRIM23;
  for (int i = 0; i < AP26O_NPRIMES; i++) {
    int p = g_primes[i];
    uint8_t *ok = out + g_offsets[i];
    memset(ok, 1, (size_t)p);
    int sp = (int)(STEP % p);
    for (int j = p - 23; j <= p; j++) ok[(j * sp) % p] = 0;
  }
}

/* OKOK_p[r] bit jj = OK_p[(r + (jj+SHIFT)*MOD) mod p]
   (src/cpu/cpusse2.cpp:37-43, src/opencl/kernels/setupokok.cl:16-20). */
void ap26o_okok_table(int32_t K, int32_t SHIFT, uint64_t out[AP26O_NOK]) {
  uint8_t ok[AP26O_NOK];
  ap26o_ok_table(K, ok);
  for (int i = 0; i < AP26O_NPRIMES; i++) {
    int p = g_primes[i];
    for (int r = 0; r < p; r++) {
      uint64_t w = 0;
      for (int jj = 0; jj < 64; jj++) {
        int64_t idx = ((int64_t)r + (int64_t)(jj + SHIFT) * O_MOD) % p;
        if (ok[g_offsets[i] + idx]) w |= UINT64_C(1) << jj;
      }
      out[g_offsets[i] + r] = w;
    }
  }
}

/* ---------------------------------------------------------------------- */
/* Strong probable-prime test, textbook (src/common/PrimeQ_x86.h:114-153). */

static inline uint64_t mulmod128(uint64_t a, uint64_t b, uint64_t n) {
  return (uint64_t)(((unsigned __int128)a * b) % n);
}

int ap26o_sprp(uint64_t b, uint64_t N) {
  if (N < 3 || !(N & 1)) return 0;
  uint64_t d = N - 1;
  int t = 0;
  while (!(d & 1)) { d >>= 1; t++; }
  uint64_t a = b % N;
  if (a == 0) return 1; /* N divides the base: not a witness (N prime <= b) */
  uint64_t r = 1;
  /* r = a^d mod N, left to right */
  for (int bit = 63 - __builtin_clzll(d); bit >= 0; bit--) {
    r = mulmod128(r, r, N);
    if ((d >> bit) & 1) r = mulmod128(r, a, N);
  }
  if (r == 1 || r == N - 1) return 1;
  for (int s = 1; s < t; s++) {
    r = mulmod128(r, r, N);
    if (r == N - 1) return 1;
  }
  return 0;
}

int ap26o_prp2(uint64_t N) { return ap26o_sprp(2, N); }

/* ap26o_prp2 over an array, split over threads (contiguous slices): the
   checker of whole SPRP batches such as the 2^26-value micro-benchmark. */
typedef struct {
  const uint64_t *n;
  uint8_t *out;
  size_t b, e;
} prp_slice;

static void *prp_worker(void *arg) {
  prp_slice *sl = (prp_slice *)arg;
  for (size_t i = sl->b; i < sl->e; i++) sl->out[i] = (uint8_t)ap26o_prp2(sl->n[i]);
  return NULL;
}

void ap26o_prp2_array(const uint64_t *n, uint8_t *out, size_t count, int32_t threads) {
  if (threads < 1) threads = 1;
  if ((size_t)threads > count) threads = count ? (int32_t)count : 1;
  pthread_t *th = (pthread_t *)calloc((size_t)threads, sizeof(pthread_t));
  prp_slice *sl = (prp_slice *)calloc((size_t)threads, sizeof(prp_slice));
  for (int32_t t = 0; t < threads; t++) {
    sl[t].n = n;
    sl[t].out = out;
    sl[t].b = count * (size_t)t / (size_t)threads;
    sl[t].e = count * (size_t)(t + 1) / (size_t)threads;
    pthread_create(&th[t], NULL, prp_worker, &sl[t]);
  }
  for (int32_t t = 0; t < threads; t++) pthread_join(th[t], NULL);
  free(th);
  free(sl);
}

/* src/cpu/cpusse2.cpp:570-590 (GPU state machine src/opencl/kernels/checkn.cl:160-202):
   count base-2 SPRPs forward from n+5*STEP; only if that count reaches 10 extend
   backward from n+4*STEP while the term is positive. */
int32_t ap26o_walk(int64_t n, int64_t STEP, int64_t *first) {
  int32_t k = 0;
  int64_t m = n + 5 * STEP;
  while (ap26o_prp2((uint64_t)m)) { k++; m += STEP; }
  if (k >= 10) {
    m = n + 4 * STEP;
    while (m > 0 && ap26o_prp2((uint64_t)m)) { k++; m -= STEP; }
  }
  *first = m + STEP;
  return k;
}

/* ---------------------------------------------------------------------- */
/* Sieve (src/cpu/cpusse2.cpp:389-621, src/opencl/kernels/sieve.cl:28-152). */

typedef struct {
  int32_t K, SHIFT0, it0, it1, j0, j1;
  int walk;               /* 0: collect survivors; 1: walk them */
  int64_t *surv; int64_t surv_cap; int64_t nsurv;
  ap26o_sol *sols; int64_t sol_cap; int64_t nsol;
} span_job;

static int pass7to23(int64_t n) {
  return (n % 7) && (n % 11) && (n % 13) && (n % 17) && (n % 19) && (n % 23);
}

static void run_span(span_job *job) {
  pthread_once(&g_once, init_primes);
  ap26o_kconst kc;
  ap26o_kconsts(job->K, &kc);
  int64_t *n43 = (int64_t *)malloc(sizeof(int64_t) * AP26O_NN43);
  ap26o_n43_table(&kc, n43);
  int ns = job->it1 - job->it0;
  uint64_t *okok = (uint64_t *)malloc(sizeof(uint64_t) * AP26O_NOK * (size_t)(ns > 0 ? ns : 1));
  for (int s = 0; s < ns; s++)
    ap26o_okok_table(job->K, job->SHIFT0 + 64 * (job->it0 + s), okok + (size_t)s * AP26O_NOK);

  uint64_t mask[10];
  for (int j = job->j0; j < job->j1; j++) {
    int64_t a43 = n43[j];
    for (int i43 = 0; i43 < 19; i43++) {
      int64_t a47 = a43;
      for (int i47 = 0; i47 < 23; i47++) {
        int64_t a53 = a47;
        for (int i53 = 0; i53 < 29; i53++) {
          int64_t n59 = a53;
          for (int i59 = 0; i59 < 35; i59++) {
            uint64_t any = 0;
            for (int s = 0; s < ns; s++) mask[s] = ~UINT64_C(0);
            for (int i = 0; i < AP26O_NPRIMES; i++) {
              uint32_t r = mod_small((uint64_t)n59, i);
              const uint64_t *t = okok + g_offsets[i] + r;
              any = 0;
              for (int s = 0; s < ns; s++) {
                mask[s] &= t[(size_t)s * AP26O_NOK];
                any |= mask[s];
              }
              if (!any) break;
            }
            if (any) {
              for (int s = 0; s < ns; s++) {
                uint64_t w = mask[s];
                int SHIFT = job->SHIFT0 + 64 * (job->it0 + s);
                while (w) {
                  int b = __builtin_ctzll(w);
                  w &= w - 1;
                  int64_t n = n59 + (int64_t)(b + SHIFT) * O_MOD;
                  if (!pass7to23(n)) continue;
                  if (!job->walk) {
                    if (job->nsurv < job->surv_cap) job->surv[job->nsurv] = n;
                    job->nsurv++;
                  } else {
                    job->nsurv++;
                    int64_t first;
                    int32_t k = ap26o_walk(n, kc.STEP, &first);
                    if (k >= 10) {
                      if (job->nsol < job->sol_cap) {
                        job->sols[job->nsol].k = k;
                        job->sols[job->nsol].K = job->K;
                        job->sols[job->nsol].first = first;
                      }
                      job->nsol++;
                    }
                  }
                }
              }
            }
            n59 += kc.S59;
            if (n59 >= O_MOD) n59 -= O_MOD;
          }
          a53 += kc.S53;
          if (a53 >= O_MOD) a53 -= O_MOD;
        }
        a47 += kc.S47;
        if (a47 >= O_MOD) a47 -= O_MOD;
      }
      a43 += kc.S43;
      if (a43 >= O_MOD) a43 -= O_MOD;
    }
  }
  free(okok);
  free(n43);
}

int64_t ap26o_sieve_span(int32_t K, int32_t SHIFT0, int32_t it0, int32_t it1, int32_t j0,
                         int32_t j1, int64_t *out, int64_t cap) {
  span_job job;
  memset(&job, 0, sizeof(job));
  job.K = K; job.SHIFT0 = SHIFT0; job.it0 = it0; job.it1 = it1; job.j0 = j0; job.j1 = j1;
  job.walk = 0; job.surv = out; job.surv_cap = cap;
  run_span(&job);
  return job.nsurv;
}

static void *span_thread(void *arg) { run_span((span_job *)arg); return NULL; }

int64_t ap26o_search_span(int32_t K, int32_t SHIFT0, int32_t it0, int32_t it1, int32_t j0,
                          int32_t j1, ap26o_sol *out, int64_t cap, int64_t *n_surv,
                          int32_t nthreads) {
  if (nthreads < 1) nthreads = 1;
  if (nthreads > j1 - j0) nthreads = (j1 - j0) > 0 ? (j1 - j0) : 1;
  span_job *jobs = (span_job *)calloc((size_t)nthreads, sizeof(span_job));
  pthread_t *th = (pthread_t *)calloc((size_t)nthreads, sizeof(pthread_t));
  int span = j1 - j0;
  for (int t = 0; t < nthreads; t++) {
    span_job *jb = &jobs[t];
    jb->K = K; jb->SHIFT0 = SHIFT0; jb->it0 = it0; jb->it1 = it1;
    jb->j0 = j0 + (int)((int64_t)span * t / nthreads);
    jb->j1 = j0 + (int)((int64_t)span * (t + 1) / nthreads);
    jb->walk = 1;
    jb->sol_cap = 4096;
    jb->sols = (ap26o_sol *)malloc(sizeof(ap26o_sol) * (size_t)jb->sol_cap);
  }
  if (nthreads == 1) {
    run_span(&jobs[0]);
  } else {
    for (int t = 0; t < nthreads; t++) pthread_create(&th[t], NULL, span_thread, &jobs[t]);
    for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
  }
  int64_t total = 0, surv = 0;
  for (int t = 0; t < nthreads; t++) {
    int64_t m = jobs[t].nsol < jobs[t].sol_cap ? jobs[t].nsol : jobs[t].sol_cap;
    for (int64_t e = 0; e < m; e++) {
      if (total < cap) out[total] = jobs[t].sols[e];
      total++;
    }
    surv += jobs[t].nsurv;
    free(jobs[t].sols);
  }
  free(jobs);
  free(th);
  if (n_surv) *n_surv = surv;
  return total;
}

/* ---------------------------------------------------------------------- */
/* Host-side reporting (src/common/AP26.cpp:484-602). */

int32_t ap26o_validate(int32_t k, int32_t K, int64_t first, int nbases) {
  static const uint64_t bases[9] = {2, 3, 5, 7, 11, 13, 17, 19, 23};
  if (first % 2 == 0) return 0;
  int64_t N = first;
  int32_t j;
  for (j = 0; j < k; j++) {
    for (int b = 0; b < nbases; b++)
      if (!ap26o_sprp(bases[b], (uint64_t)N)) return j;
    N += (int64_t)K * O_PRIM23;
  }
  return j;
}

#define O_MAXINTV 2000000000

int ap26o_report(int32_t k, int32_t K, int64_t first, int32_t *hash, int32_t min_report,
                 ap26o_sol *recs, int64_t cap, int64_t *nrec, int64_t *nonsol) {
  /* every report enters the hash, whatever its length (AP26.cpp:557-561) */
  *hash += (int32_t)(first % 1000);
  *hash += k;
  if (*hash > O_MAXINTV) *hash -= O_MAXINTV;
  int32_t i = ap26o_validate(k, K, first, 9);
  if (i < k) {
    if (nonsol) (*nonsol)++;
    if (ap26o_validate(k, K, first, 1) < k) return -1; /* AP26.cpp:569-574 */
    if (ap26o_report(i, K, first, hash, min_report, recs, cap, nrec, nonsol)) return -1;
    return ap26o_report(k - (i + 1), K, first + (int64_t)(i + 1) * K * O_PRIM23, hash,
                        min_report, recs, cap, nrec, nonsol);
  }
  if (k >= min_report) {
    if (*nrec < cap) {
      recs[*nrec].k = k;
      recs[*nrec].K = K;
      recs[*nrec].first = first;
    }
    (*nrec)++;
  }
  return 0;
}

/* src/common/AP26.cpp:308-319: top 32 bits = KMIN+KMAX folded below 2e9,
   bottom 32 bits = the capped hash. */
uint64_t ap26o_hash_word(int32_t KMIN, int32_t KMAX, int32_t hash) {
  uint64_t minmax = (uint64_t)(int64_t)(KMIN + KMAX);
  while (minmax > O_MAXINTV) minmax -= O_MAXINTV;
  return (minmax << 32) | (uint32_t)hash;
}

/* ---------------------------------------------------------------------- */
/* Algorithmic work count of the sieve: the number of table lookups a mask
   evaluation performs when every shift's 64-bit mask is evaluated on its own
   with the reference GPU kernel's early-exit grouping (groups of 5,5,5,4,4
   primes, then pairs: src/opencl/kernels/sieve.cl:32-114).  Summed over the
   positions of the span; used to price the sieve against the INT32 roofline. */
int64_t ap26o_lookup_count(int32_t K, int32_t SHIFT0, int32_t it0, int32_t it1, int32_t j0,
                           int32_t j1) {
  pthread_once(&g_once, init_primes);
  static const int group_end[] = {5, 10, 15, 19, 23, 25, 27, 29, 31, 33, 35, 37, 39, 41, 43, 45, 47,
                                  49, 51, 53, 55, 57, 59, 61, 63, 65, 67, 69, 71, 73, 75, 77, 79, 81,
                                  83};
  const int ngroups = (int)(sizeof(group_end) / sizeof(group_end[0]));
  ap26o_kconst kc;
  ap26o_kconsts(K, &kc);
  int64_t *n43 = (int64_t *)malloc(sizeof(int64_t) * AP26O_NN43);
  ap26o_n43_table(&kc, n43);
  int ns = it1 - it0;
  uint64_t *okok = (uint64_t *)malloc(sizeof(uint64_t) * AP26O_NOK * (size_t)(ns > 0 ? ns : 1));
  for (int s = 0; s < ns; s++)
    ap26o_okok_table(K, SHIFT0 + 64 * (it0 + s), okok + (size_t)s * AP26O_NOK);
  int64_t total = 0;
  for (int j = j0; j < j1; j++) {
    int64_t a43 = n43[j];
    for (int i43 = 0; i43 < 19; i43++) {
      int64_t a47 = a43;
      for (int i47 = 0; i47 < 23; i47++) {
        int64_t a53 = a47;
        for (int i53 = 0; i53 < 29; i53++) {
          int64_t n59 = a53;
          for (int i59 = 0; i59 < 35; i59++) {
            uint32_t r[AP26O_NPRIMES];
            for (int i = 0; i < AP26O_NPRIMES; i++) r[i] = mod_small((uint64_t)n59, i);
            for (int s = 0; s < ns; s++) {
              const uint64_t *t = okok + (size_t)s * AP26O_NOK;
              uint64_t m = ~UINT64_C(0);
              int i = 0;
              for (int g = 0; g < ngroups; g++) {
                for (; i < group_end[g]; i++) m &= t[g_offsets[i] + r[i]];
                total += group_end[g] - (g ? group_end[g - 1] : 0);
                if (!m) break;
              }
            }
            n59 += kc.S59;
            if (n59 >= O_MOD) n59 -= O_MOD;
          }
          a53 += kc.S53;
          if (a53 >= O_MOD) a53 -= O_MOD;
        }
        a47 += kc.S47;
        if (a47 >= O_MOD) a47 -= O_MOD;
      }
      a43 += kc.S43;
      if (a43 >= O_MOD) a43 -= O_MOD;
    }
  }
  free(okok);
  free(n43);
  return total;
}
