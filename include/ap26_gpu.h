/* ap26_gpu.h -- C ABI of the MI355X (gfx950) AP26 search library
 * (libap26gpu.so).  Plain C types only: the host (AP26 CLI, Python ctypes
 * mirror) never sees HIP or torch types.
 *
 * The library owns device memory and streams inside an ap26_ctx; callers own
 * every host buffer they pass.  Calls on one ctx must not overlap (no internal
 * lock); several ctx may coexist on one device.  Every entry point returns 0 or
 * a negative AP26_E* code (ap26_strerror); none exits the process -- the host
 * decides, as the reference's exit(EXIT_FAILURE) paths did.
 */
#ifndef AP26_GPU_H
#define AP26_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AP26_OK 0
#define AP26_E_DEVICE (-1)    /* HIP runtime / device error */
#define AP26_E_NOMEM (-2)     /* device allocation failed */
#define AP26_E_OVERFLOW (-3)  /* caller's output capacity too small (count still set) */
#define AP26_E_RANGE (-4)     /* argument out of the supported range */
#define AP26_E_INTERNAL (-5)  /* internal buffer overflow (should not happen) */
#define AP26_E_BUSY (-6)      /* serving queue full */

/* One reported progression (k terms, difference K*23#, first term).
   range_id names the request/range it came from (batch/serving), else -1. */
typedef struct {
  int32_t k;
  int32_t K;
  int64_t first;
  int32_t range_id;
  int32_t shift_iter;  /* which of the 10 shifts found it (0..9) */
  uint32_t prp2_mask;  /* bit j: first + j*K*23# is a base-2 strong PRP */
  uint32_t valid_mask; /* bit j: ... and a strong PRP to 3, 5, ..., 23 too
                          (validate_ap26, src/common/AP26.cpp:513-543);
                          terms 0..31, computed on the device */
} ap26_sol;

/* One work unit: the reference's `AP26 KMIN KMAX SHIFT`, restricted to K in
   [kmin, kmax] that pass will_search, starting at shift iteration iter_start
   for K == kmin (a resumed unit), 0 for the others. */
typedef struct {
  int32_t kmin;
  int32_t kmax;
  int32_t shift;
  int32_t iter_start;
} ap26_range;

/* Tuning knobs (0 = library default / autotuner cache). */
typedef struct {
  int32_t sieve_variant;    /* 0 default (53: shift-pair queued-compaction
                               sieve with the bank-rotated level-0 group in
                               512-byte rows, seven levels), 1 simple
                               per-shift sieve, 2 simple sieve with 2 shifts
                               per pass, 10..53 tuning candidates */
  int32_t block;            /* (reported) threads per workgroup of the sieve */
  int32_t blocks_per_cu;    /* resident workgroups per CU (grid = CUs x this) */
  int32_t verbose;
} ap26_opts;

/* Per-call measurements (device time from HIP events on the ctx stream). */
typedef struct {
  int64_t k_searched;       /* K that passed will_search */
  int64_t k_skipped;
  int64_t positions;        /* n59 positions x shifts evaluated */
  int64_t survivors;        /* sieve survivors (candidates walked) */
  int64_t prp_tests;        /* base-2 SPRP tests run by the walk */
  int64_t reports;          /* (k >= 10) reports */
  double ms_tables;         /* table-build kernel time */
  double ms_sieve;          /* sieve kernel time; batch: the sum over the
                               timed rounds (the first 1024 of each segment)
                               of the round's sieve graph (n43 starts, tables,
                               persistent sieve), sieve_launches = the number
                               of rounds timed */
  double ms_walk;           /* walk kernel time (batch: device time of all
                               rounds) */
  double ms_total;          /* wall time of the call */
  int32_t sieve_launches;
  int32_t walk_launches;
  int64_t walk_past_limit;  /* walked terms >= 2^62: the reference CPU build's
                               PrimeQ exits on the first
                               (src/common/PrimeQ_x86.h:160-166); a call that
                               meets one returns AP26_E_RANGE and no reports */
} ap26_stats;

typedef struct ap26_ctx ap26_ctx;

const char *ap26_strerror(int code);
const char *ap26_version(void);

/* Open a context on HIP device `device` (opts may be NULL). */
int ap26_open(int device, const ap26_opts *opts, ap26_ctx **out);
int ap26_close(ap26_ctx *ctx);
/* Device properties the tuning uses. */
int ap26_device_info(ap26_ctx *ctx, int32_t *cu_count, int64_t *lds_per_block,
                     int64_t *global_mem, char *name, int32_t name_len);

/* Search one K over its 10 shifts SHIFT+64*iter, iter in [iter_start, 10):
   the per-K seam of the reference, replacing
     SearchAP26(int K, int startSHIFT, int s_local, int compute)  src/opencl/AP26.h:78
     Search_avx2(int K, int startSHIFT, int ITER, int K_COUNT, int K_DONE)  src/cpu/cpuavx2.h:15
   Raw reports are written to out (cap entries, *n_out = total found); the
   caller applies ReportSolution (hash, 9-base check, SOL records).  Blocking. */
int ap26_search_k(ap26_ctx *ctx, int32_t K, int32_t SHIFT, int32_t iter_start, ap26_sol *out,
                  int32_t cap, int32_t *n_out, ap26_stats *stats);

/* As ap26_search_k over shift iterations [it0, it1) and n43 starts [j0, j1)
   (0 <= j0 < j1 <= 10840) -- a tile of one K's search, used by tests and the
   autotuner. */
int ap26_search_span(ap26_ctx *ctx, int32_t K, int32_t SHIFT, int32_t it0, int32_t it1,
                     int32_t j0, int32_t j1, ap26_sol *out, int32_t cap, int32_t *n_out,
                     ap26_stats *stats);

/* ap26_search_span for callers that cannot allocate (code under HIP graph
   capture, tight serving loops, harnesses that forbid hipMalloc): no
   counterpart in the reference.  It calls no hipMalloc, hipFree or other
   allocating routine, host or device: its device working memory is one
   file-scope array of about 34 MB that every device the library runs on
   holds until the process ends (tables of 10 shifts, n43 starts, 2^22
   survivor slots, counters, 4096 reports), cleared by kernels as each call
   needs.  The caller passes inputs and outputs as for ap26_search_span;
   results, return codes and stats are the same.  Limit: at most 64 n43
   starts per call (j1 - j0 <= 64), enough survivor slots for any K <= 6e8;
   a larger span returns AP26_E_RANGE.  All calls on one GPU share that
   GPU's array: calls that overlap on one GPU (from several threads or
   contexts) race and must be serialised by the caller. */
int ap26_search_span_noalloc(ap26_ctx *ctx, int32_t K, int32_t SHIFT, int32_t it0, int32_t it1,
                             int32_t j0, int32_t j1, ap26_sol *out, int32_t cap, int32_t *n_out,
                             ap26_stats *stats);

/* Sieve survivors only (the candidates the walk would test) of that tile,
   in no particular order.  *n_out = total; AP26_E_OVERFLOW if > cap. */
int ap26_sieve_span(ap26_ctx *ctx, int32_t K, int32_t SHIFT, int32_t it0, int32_t it1,
                    int32_t j0, int32_t j1, int64_t *out, int64_t cap, int64_t *n_out);

/* Packed OKOK tables of one (K, SHIFT) as the sieve uses them, re-centred to
   the reference's layout: out[offset(p) + r] = OKOK_p[r] (23693 words). */
int ap26_okok_table(ap26_ctx *ctx, int32_t K, int32_t SHIFT, uint64_t *out);

/* Base-2 strong probable-prime test of count odd 64-bit values (host buffers):
   out[i] = 1 iff n[i] is a base-2 SPRP (textbook; 0 for n < 3 or even n).
   Replaces PrimeQ(int64_t) (src/common/PrimeQ_x86.h:160), without its 2^62 cap. */
int ap26_prp_batch(ap26_ctx *ctx, const uint64_t *n, uint8_t *out, size_t count);

/* Same on device-resident buffers (from ap26_dev_alloc), on the ctx stream;
   *ms = kernel time from HIP events around the launch (NULL to skip). */
int ap26_prp_batch_dev(ap26_ctx *ctx, const void *d_n, void *d_out, size_t count, double *ms);

/* Many work units in one persistent launch sequence (work-stealing K queue,
   captured hipGraph per round of 16 K).  out gets every report with range_id =
   index into r; hash_totals[i] = exact sum over range i of (first % 1000 + k)
   for the raw reports (the host applies the reference's cap and Non-Solution
   rules).  The searched K run in segments of 2048 (device verification and
   one report copy per segment), so a call may hold any number of K; *n_out =
   total reports, AP26_E_OVERFLOW if > cap (the search is complete; call
   again with a larger buffer). */
int ap26_search_batch(ap26_ctx *ctx, const ap26_range *r, int32_t nr, ap26_sol *out,
                      int32_t cap, int32_t *n_out, int64_t *hash_totals, ap26_stats *stats);

/* The batched engine's sieve alone for one K (shift iterations iter_start..9):
   its survivors, in no particular order, for parity against ap26_sieve_span.
   *n_out = total; AP26_E_OVERFLOW if > cap.  Tile of the same search as
   ap26_search_batch (the reference's sieve, src/opencl/kernels/sieve.cl:12-156). */
int ap26_batch_sieve_k(ap26_ctx *ctx, int32_t K, int32_t SHIFT, int32_t iter_start, int64_t *out,
                       int64_t cap, int64_t *n_out);

/* 9-base verification on the device of n caller-supplied reports (k, K,
   first): fills prp2_mask and valid_mask of each (validate_ap26 and
   val_base2_ap26, src/common/AP26.cpp:484-543).  The search entry points
   fill them for their own reports. */
int ap26_verify_sols(ap26_ctx *ctx, ap26_sol *sols, int32_t n);

/* ---- serving: continuous batching of work units ------------------------
   Units are admitted while unit slots and device memory allow; every round
   takes the next K of each active unit in turn (round-robin), so units that
   arrive while others run join the next round and share the device with
   them.  Reports and per-unit hash totals are kept per unit (totals on the
   device) until collected.  Single host thread per server. */
typedef struct ap26_server ap26_server;

/* max_units: admitted units (active + waiting); mem_budget_bytes: device
   memory the server may use for in-flight K (0 = 1/8 of the GPU); it sets how
   many K a round holds. */
int ap26_server_open(ap26_ctx *ctx, int32_t max_units, int64_t mem_budget_bytes, ap26_server **out);
int ap26_server_close(ap26_server *s);
/* Queue a unit; AP26_E_BUSY if max_units are already admitted. */
int ap26_server_submit(ap26_server *s, const ap26_range *unit, int64_t *ticket);
/* Launch the next round if work is waiting and harvest the finished one;
   *completed = units finished by this call.  Blocks for one round. */
int ap26_server_pump(ap26_server *s, int32_t *completed);
/* Result of a finished unit (done = 0 while it runs): its raw reports (range_id
   = -1), the exact hash total, and its latency from submit to completion. */
int ap26_server_result(ap26_server *s, int64_t ticket, ap26_sol *out, int32_t cap, int32_t *n_out,
                       int64_t *hash_total, double *latency_ms, int32_t *done);
/* Counters: units admitted/finished, K searched, rounds, K per round. */
/* Progress of one admitted unit: searched K finished / searched K in the
   unit (the reference's fraction done, AP26.cpp:636-641, at K granularity). */
int ap26_server_progress(ap26_server *s, int64_t ticket, double *fraction, int64_t *k_done, int64_t *k_total);

int ap26_server_stats(ap26_server *s, int64_t *units_done, int64_t *k_done, int64_t *rounds,
                      int32_t *k_per_round, int32_t *active);

/* ---- autotuner --------------------------------------------------------
   Times the batched engine on one K at SHIFT: first its sieve variants
   (LDS-staged primes, wavefronts per CU, compaction levels, one shift or a
   shift pair per work item) and work-item sizes, keeping the fastest by
   device time; then, with that sieve, the walk's compaction depth (terms a
   lane tests before a run is queued again: 1, 2 or 4), keeping the fastest
   by the walk's time.  The choice holds for this ctx and is written to
   cache_path (if not NULL); ap26_open adopts a cache named by the
   AP26_TUNE_CACHE environment variable.  quick = 1 tries 10 + 3 candidates,
   results[] lists every candidate timed in that order. */
typedef struct {
  int32_t variant;
  int32_t chunk;      /* n43 starts per work item */
  int32_t walk_depth; /* walk compaction depth */
  int32_t reserved;
  double ms_per_k;    /* device time of the sample K */
  double ms_walk;     /* of which after the sieve (walk and verification) */
} ap26_tune_entry;

int ap26_autotune(ap26_ctx *ctx, int32_t K, int32_t SHIFT, int32_t quick, const char *cache_path,
                  ap26_tune_entry *results, int32_t cap, int32_t *n_out, ap26_tune_entry *best);

/* Device buffer helpers (so callers need no HIP/torch to stage inputs). */
int ap26_dev_alloc(ap26_ctx *ctx, size_t bytes, void **dptr);
int ap26_dev_free(ap26_ctx *ctx, void *dptr);
int ap26_copy_h2d(ap26_ctx *ctx, void *dst, const void *src, size_t bytes);
int ap26_copy_d2h(ap26_ctx *ctx, void *dst, const void *src, size_t bytes);
int ap26_sync(ap26_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif
