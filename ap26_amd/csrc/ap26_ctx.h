/* ap26_ctx.h -- library-private: the ap26_ctx object and helpers shared by
 * the per-K path (ap26_gpu.hip) and the batched engine (ap26_batch.hip).
 */
#ifndef AP26_CTX_H
#define AP26_CTX_H

#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

#include "../../include/ap26_gpu.h"
#include "ap26_kernels.h"

using namespace ap26;

#define HIPCHK(x)                                                                   \
  do {                                                                              \
    hipError_t e_ = (x);                                                            \
    if (e_ != hipSuccess) {                                                         \
      if (ctx && ctx->opts.verbose)                                                 \
        fprintf(stderr, "ap26: %s failed: %s\n", #x, hipGetErrorString(e_));       \
      return AP26_E_DEVICE;                                                         \
    }                                                                               \
  } while (0)


constexpr int64_t SURV_CAP = INT64_C(1) << 26; /* initial survivor buffer: 67M (~40M per K) */
constexpr int SOL_CAP = 1 << 16;
constexpr int NCOUNT = 16; /* device counters */
/* C_SLAB + p: blocks claimed by the sieve launch of shift pair p */
enum { C_NSURV = 0, C_NTESTS = 1, C_NSOL = 2, C_LIVE = 3, C_BIG = 4, C_ERR = 7, C_SLAB = 8 };

using SieveFn = void (*)(SieveArgs);

struct SieveVariant {
  int S, NL, block;
  SieveFn fn;
  int gq_entries = 0; /* global queue entries per wave (SieveArgs::gq) */
};

template <int S, int NL>
SieveVariant make_variant() {
  return SieveVariant{S, NL, 256, &sieve_kernel<S, NL>};
}

template <int NL, int WAVES, class LV, bool VC, unsigned G = 0>
SieveVariant make_qvariant() {
  return SieveVariant{1, NL, 64 * WAVES, &sieve_queue_kernel<NL, WAVES, LV, VC, G>};
}

template <int NL, int WAVES, class LV, int V = 0>
SieveVariant make_q2variant() {
  return SieveVariant{2, NL, 64 * WAVES, &sieve_queue2_kernel<NL, WAVES, LV, V>};
}

/* shift pairs with the bank-rotated level-0 group (primes 61..137) */
template <int NL, int WAVES, class LV, class TU = Tune<>>
SieveVariant make_q2variantr() {
  return SieveVariant{2, NL, 64 * WAVES, &sieve_queue2_kernel<NL, WAVES, LV, 0, 0, ROT_N, TU>};
}

/* the same with queues GQ.. in global memory, so that more primes fit in LDS */
template <int NL, int WAVES, class LV, class TU, int GQ>
SieveVariant make_q2variantrg() {
  return SieveVariant{2, NL, 64 * WAVES, &sieve_queue2_kernel<NL, WAVES, LV, 0, 0, ROT_N, TU, GQ>,
                      qsieve2_global_entries<LV, GQ>()};
}

/* the same with 512-byte rotation rows (no per-lane column registers) */
template <int NL, int WAVES, class LV, class TU, int GQ>
SieveVariant make_q2variantrw() {
  return SieveVariant{2, NL, 64 * WAVES, &sieve_queue2_kernel<NL, WAVES, LV, 0, 0, ROT_WIDE, TU, GQ>,
                      qsieve2_global_entries<LV, GQ>()};
}


struct BatchState;
struct ap26_ctx {
  int device = 0;
  ap26_opts opts{};
  hipStream_t stream = nullptr;
  hipEvent_t ev[8] = {};
  int cu_count = 0;
  size_t lds_per_block = 0;
  size_t global_mem = 0;
  char name[256] = {0};
  /* device buffers */
  uint64_t *d_tab = nullptr;     /* 10 shifts of tables */
  int64_t *d_n43 = nullptr;      /* [10840] */
  int64_t *d_surv = nullptr;     /* surv_cap entries, grown for K with more survivors */
  int64_t surv_cap = 0;
  unsigned long long *d_cnt = nullptr; /* NCOUNT counters (sol count stored as u32 at C_NSOL) */
  DevSol *d_sols = nullptr;      /* SOL_CAP */
  void *d_gq = nullptr;          /* the sieve variant's global queues: grid x waves regions */
  SieveVariant sieve{};
  SieveVariant sieve_lds{};      /* all queues in LDS: the allocation-free entry point's sieve */
  BatchState *batch = nullptr;
  int tune_variant = -1; /* batched sieve variant / work-item chunk chosen by */
  int tune_chunk = -1;   /* ap26_autotune (or its cache); -1 = default       */
  int tune_walk_depth = -1; /* walk compaction depth chosen the same way */
  void *scratch = nullptr; /* this device's address of the allocation-free entry
                              point's scratch array (hipGetSymbolAddress) */
};


SieveVariant pick_variant(int id);

/* Walk compaction depth: AP26_WALK_DEPTH, else the autotuned choice, else 1;
   one of WALK_DEPTHS. */
static inline int walk_depth_of(const ap26_ctx *ctx) {
  const char *e = getenv("AP26_WALK_DEPTH");
  const int d = e ? atoi(e) : ctx->tune_walk_depth;
  for (int w : WALK_DEPTHS)
    if (d == w) return d;
  return 1;
}

template <class Args>
using WalkFn = void (*)(Args);

/* the walk kernel of a depth */
static inline WalkFn<WalkArgs> pick_walk(int depth) {
  return depth == 4 ? &walk_kernel<4> : depth == 2 ? &walk_kernel<2> : &walk_kernel<1>;
}
static inline WalkFn<BatchWalkArgs> pick_batch_walk(int depth) {
  return depth == 4 ? &batch_walk_kernel<4> : depth == 2 ? &batch_walk_kernel<2> : &batch_walk_kernel<1>;
}

/* batched engine state (ap26_batch.hip) */
struct BatchState;
void batch_free(ap26_ctx *ctx);
void tune_load_cache(ap26_ctx *ctx);

#endif
