/* ap26_gpu.hip -- C ABI of libap26gpu.so (declared in include/ap26_gpu.h):
 * context, per-K search (the plain path), tiles, SPRP batches.
 */
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstddef>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "../../include/ap26_gpu.h"
#include "ap26_host_math.h"
#include "ap26_kernels.h"

using namespace ap26;

#include "ap26_ctx.h"

/* Sieve variants: 0 = default (53: shift-pair queued sieve with the
   bank-rotated level-0 group in 512-byte rows, level 0 on 20 primes, seven
   compaction levels, 12 waves per CU, queues 1.. in global memory), 1 =
   simple S=1, 2 =
   simple S=2, 10..30 = one-shift queued-sieve tuning candidates (LDS
   primes, waves per CU, levels, VGPR constants), 31..33 = shift-pair queued
   sieves without the rotation group, 34..36 = shift-pair with the first
   1..3 level-0 primes from register bit windows instead of LDS, 40..53 =
   rotation-group forms (45: every queue in LDS, the allocation-free entry
   point's sieve; 47: four levels; 48, 49: 47 with packed level-0 primes
   through L1; 50..52: live-lane packed gathers, at 12 and 16 waves). */
SieveVariant pick_variant(int id) {
  switch (id) {
    case 1: return make_variant<1, 45>();
    case 2: return make_variant<2, 45>();
    case 10: return make_qvariant<57, 8, Levels4, false>();
    case 11: return make_qvariant<57, 8, Levels4, true>();
    case 12: return make_qvariant<61, 8, Levels4, true>();
    case 13: return make_qvariant<45, 12, Levels4, false>();
    case 14: return make_qvariant<57, 16, Levels3, false>();
    case 15: return make_qvariant<45, 16, Levels3, false>();
    case 16: return make_qvariant<45, 8, Levels5, true>();
    case 17: return make_qvariant<65, 8, Levels3, true>();
    case 18: return make_qvariant<45, 12, Levels3, false>();
    case 19: return make_qvariant<31, 16, Levels4, false>();
    case 20: return make_qvariant<29, 16, Levels4, false>();
    case 21: return make_qvariant<35, 16, Levels4, false>();
    case 22: return make_qvariant<20, 16, Levels5, false>();
    case 23: return make_qvariant<38, 16, Levels4, false>();
    case 24: return make_qvariant<25, 16, Levels5, false>();
    case 25: return make_qvariant<31, 12, Levels5, false>();
    case 26: return make_qvariant<38, 16, Levels4, false, 0x5555>();
    case 27: return make_qvariant<38, 16, Levels4, false, 0x1111>();
    case 28: return make_qvariant<38, 16, Levels4, false, 0x7fff>();
    case 29: return make_qvariant<38, 16, Levels4, false, 0x0fff>();
    case 30: return make_qvariant<38, 16, Levels4, false, 0x00ff>();
    case 31: return make_q2variant<29, 16, Levels<0, 18, 29, 51, 83>>();
    case 32: return make_q2variant<29, 16, Levels<0, 20, 31, 51, 83>>();
    case 33: return make_q2variant<29, 16, Levels4>();
    case 34: return make_q2variant<29, 16, Levels<0, 18, 29, 51, 83>, 1>();
    case 35: return make_q2variant<29, 16, Levels<0, 18, 29, 51, 83>, 2>();
    case 36: return make_q2variant<29, 16, Levels<0, 18, 29, 51, 83>, 3>();
    case 40: return make_q2variantr<29, 12, Levels<0, 16, 29, 51, 83>>();
    case 41: return make_q2variantr<29, 12, Levels<0, 18, 29, 51, 83>>();
    case 42: return make_q2variantr<30, 12, Levels<0, 18, 30, 51, 83>>();
    case 43: return make_q2variantr<36, 8, Levels<0, 18, 36, 55, 83>>();
    case 44: return make_q2variantr<21, 16, Levels<0, 16, 21, 51, 83>>();
    case 45: return make_q2variantr<30, 12, Levels<0, 20, 30, 51, 83>, Tune<2, 3, 1, 1>>();
    /* 46: the last queue in global memory, 37 primes in LDS */
    case 46: return make_q2variantrg<37, 12, Levels<0, 20, 33, 51, 83>, Tune<2, 3, 1, 1>, 2>();
    /* 47: 512-byte rotation rows, queues 1.. in global memory, 34 primes in LDS */
    case 47: return make_q2variantrw<34, 12, Levels<0, 20, 33, 51, 83>, Tune<2, 3, 1, 1>, 1>();
    /* 48, 49: 47 with the last 4 or 2 level-0 primes read through L1 */
    case 48: return make_q2variantrw<34, 12, Levels<0, 20, 33, 51, 83>, TuneL1<2, 3, 1, 1, 4>, 1>();
    case 49: return make_q2variantrw<34, 12, Levels<0, 20, 33, 51, 83>, TuneL1<2, 3, 1, 1, 2>, 1>();
    /* 50: 47 with the packed level-0 gathers on live lanes only */
    case 50: return make_q2variantrw<34, 12, Levels<0, 20, 33, 51, 83>, TuneSkip<2, 3, 1, 1>, 1>();
    /* 51, 52: the same at 16 waves per CU */
    case 51: return make_q2variantrw<30, 16, Levels<0, 20, 30, 51, 83>, TuneSkip<2, 3, 1, 1>, 1>();
    case 52: return make_q2variantrw<31, 16, Levels<0, 20, 31, 51, 83>, TuneSkip<2, 3, 1, 2>, 1>();
    /* 53: 47 with seven compaction levels (the batched engine's 117) */
    case 53: return make_q2variantrw<34, 12, Levels<0, 20, 28, 34, 42, 51, 66, 83>, Tune<2, 3, 1, 1>, 1>();
    default: return make_q2variantrw<34, 12, Levels<0, 20, 28, 34, 42, 51, 66, 83>, Tune<2, 3, 1, 1>, 1>(); /* = 53 */
  }
}

/* ---- scratch of the allocation-free entry point -------------------------- */
/* ap26_search_span_noalloc works in one file-scope device array: the tables
   of 10 shifts, the n43 starts, NOALLOC_SURV survivor slots, the counters and
   NOALLOC_SOLS reports (about 34 MB, held on every device the library runs
   on until the process ends).  Its survivor slots hold every survivor of
   NOALLOC_MAX_N43 n43 starts of any K <= 6e8: at most 40.6 M per K x 5.45
   (K = 61*67*71*73) x 64/10840, plus the slots the sieve's waves reserve and
   leave unused (up to 256 CUs x 12 waves x 128 per shift-pair launch, 5
   launches): 3.4 M < 2^22. */
constexpr int NOALLOC_MAX_N43 = 64;
constexpr int64_t NOALLOC_SURV = INT64_C(1) << 22;
constexpr int NOALLOC_SOLS = 4096;

constexpr size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }
constexpr size_t NA_TAB = 0;
constexpr size_t NA_N43 = align256(NA_TAB + sizeof(uint64_t) * NTAB * AP26_NSHIFT);
constexpr size_t NA_SURV = align256(NA_N43 + sizeof(int64_t) * AP26_NN43);
constexpr size_t NA_CNT = align256(NA_SURV + sizeof(int64_t) * NOALLOC_SURV);
constexpr size_t NA_SOLS = align256(NA_CNT + sizeof(unsigned long long) * NCOUNT);
constexpr size_t NA_BYTES = align256(NA_SOLS + sizeof(DevSol) * NOALLOC_SOLS);

static __device__ __align__(256) unsigned char scratch[NA_BYTES];

/* Each device holds its own copy of `scratch` at its own address; the host
   name is only a stand-in.  The address is looked up once per device (with
   that device current) and kept here; host code uses it, never the name. */
constexpr int MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[MAX_DEVICES];

static int scratch_address(int device, void **out) {
  *out = nullptr;
  if (device < 0 || device >= MAX_DEVICES) return AP26_E_RANGE;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return AP26_E_DEVICE;
    g_scratch_ptr[device] = p;
  }
  *out = g_scratch_ptr[device];
  return AP26_OK;
}

extern "C" {

const char *ap26_strerror(int code) {
  switch (code) {
    case AP26_OK: return "ok";
    case AP26_E_DEVICE: return "HIP device/runtime error";
    case AP26_E_NOMEM: return "device memory allocation failed";
    case AP26_E_OVERFLOW: return "output capacity too small";
    case AP26_E_RANGE: return "argument out of range";
    case AP26_E_INTERNAL: return "internal buffer overflow";
    case AP26_E_BUSY: return "queue full";
    default: return "unknown error";
  }
}

const char *ap26_version(void) { return "ap26-mi355x 0.1 (gfx950)"; }

int ap26_open(int device, const ap26_opts *opts, ap26_ctx **out) {
  if (!out) return AP26_E_RANGE;
  *out = nullptr;
  ap26_ctx *ctx = new (std::nothrow) ap26_ctx();
  if (!ctx) return AP26_E_NOMEM;
  if (opts) ctx->opts = *opts;
  ctx->device = device;
  int rc = AP26_OK;
  do {
    hipError_t e;
    if ((e = hipSetDevice(device)) != hipSuccess) { rc = AP26_E_DEVICE; break; }
    hipDeviceProp_t prop;
    if ((e = hipGetDeviceProperties(&prop, device)) != hipSuccess) { rc = AP26_E_DEVICE; break; }
    ctx->cu_count = prop.multiProcessorCount;
    ctx->lds_per_block = prop.sharedMemPerBlock;
    ctx->global_mem = prop.totalGlobalMem;
    snprintf(ctx->name, sizeof(ctx->name), "%s", prop.gcnArchName);
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) { rc = AP26_E_DEVICE; break; }
    for (auto &ev : ctx->ev)
      if (hipEventCreate(&ev) != hipSuccess) { rc = AP26_E_DEVICE; break; }
    if (rc) break;
    if (hipMalloc(&ctx->d_tab, sizeof(uint64_t) * NTAB * AP26_NSHIFT) != hipSuccess ||
        hipMalloc(&ctx->d_n43, sizeof(int64_t) * AP26_NN43) != hipSuccess ||
        hipMalloc(&ctx->d_surv, sizeof(int64_t) * SURV_CAP) != hipSuccess ||
        hipMalloc(&ctx->d_cnt, sizeof(unsigned long long) * NCOUNT) != hipSuccess ||
        hipMalloc(&ctx->d_sols, sizeof(DevSol) * SOL_CAP) != hipSuccess) {
      rc = AP26_E_NOMEM;
      break;
    }
    ctx->surv_cap = SURV_CAP;
    if ((rc = scratch_address(device, &ctx->scratch)) != AP26_OK) break;
    ctx->sieve = pick_variant(ctx->opts.sieve_variant);
    ctx->sieve_lds = ctx->sieve.gq_entries > 0 ? pick_variant(45) : ctx->sieve;
    tune_load_cache(ctx);
    ctx->opts.block = ctx->sieve.block;
    if (ctx->opts.blocks_per_cu <= 0) ctx->opts.blocks_per_cu = ctx->sieve.block > 256 ? 1 : 2;
    /* one global queue region per wave of enqueue_sieve's grid */
    if (ctx->sieve.gq_entries > 0 &&
        hipMalloc(&ctx->d_gq, sizeof(QEntry) * (size_t)ctx->cu_count * ctx->opts.blocks_per_cu *
                                  (ctx->sieve.block / 64) * ctx->sieve.gq_entries) != hipSuccess) {
      rc = AP26_E_NOMEM;
      break;
    }
  } while (0);
  if (rc) {
    ap26_close(ctx);
    return rc;
  }
  *out = ctx;
  return AP26_OK;
}

int ap26_close(ap26_ctx *ctx) {
  if (!ctx) return AP26_OK;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  (void)hipFree(ctx->d_tab);
  (void)hipFree(ctx->d_n43);
  (void)hipFree(ctx->d_surv);
  (void)hipFree(ctx->d_cnt);
  (void)hipFree(ctx->d_sols);
  (void)hipFree(ctx->d_gq);
  batch_free(ctx);
  for (auto &ev : ctx->ev)
    if (ev) (void)hipEventDestroy(ev);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return AP26_OK;
}

int ap26_device_info(ap26_ctx *ctx, int32_t *cu_count, int64_t *lds_per_block, int64_t *global_mem,
                     char *name, int32_t name_len) {
  if (!ctx) return AP26_E_RANGE;
  if (cu_count) *cu_count = ctx->cu_count;
  if (lds_per_block) *lds_per_block = (int64_t)ctx->lds_per_block;
  if (global_mem) *global_mem = (int64_t)ctx->global_mem;
  if (name && name_len > 0) snprintf(name, (size_t)name_len, "%s", ctx->name);
  return AP26_OK;
}

}  // extern "C"

/* ---- one tile of one K --------------------------------------------------- */

static double elapsed_ms(hipEvent_t a, hipEvent_t b) {
  float ms = 0.f;
  if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0.0;
  return (double)ms;
}

/* Device buffers of one tile search: the context's own, or the
   allocation-free entry point's carve of `scratch`. */
struct SpanBufs {
  uint64_t *tab;             /* NTAB x 10 words */
  int64_t *n43;              /* [10840] */
  int64_t *surv;
  int64_t surv_cap;
  unsigned long long *cnt;   /* NCOUNT counters */
  DevSol *sols;
  int32_t sol_cap;
  const SieveVariant *sv;    /* the sieve to launch */
  void *gq;                  /* its global queues (sv->gq_entries > 0) */
};

static SpanBufs ctx_bufs(ap26_ctx *ctx) {
  return SpanBufs{ctx->d_tab, ctx->d_n43, ctx->d_surv, ctx->surv_cap, ctx->d_cnt, ctx->d_sols, SOL_CAP,
                  &ctx->sieve, ctx->d_gq};
}

/* Enqueue sieve of shifts [it0,it1) x n43 [j0,j1) into bufs.surv (counter
   C_NSURV).  Counters must have been cleared. */
static int enqueue_sieve(ap26_ctx *ctx, const SpanBufs &bf, const KParams &kp, int it0, int it1, int j0, int j1,
                         ap26_stats *st) {
  const SieveVariant &sv = *bf.sv;
  const int S = sv.S;
  const int nslot = it1 - it0;
  const int npass = (nslot + S - 1) / S;
  /* tables for every shift slot of every pass */
  dim3 tgrid((541 + 63) / 64, NPT, npass * S);
  hipLaunchKernelGGL(tables_kernel, tgrid, dim3(64), 0, ctx->stream, kp, it0, it1, S, bf.tab);
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(ctx->ev[1], ctx->stream));
  const int grid = ctx->cu_count * ctx->opts.blocks_per_cu;
  for (int p = 0; p < npass; p++) {
    SieveArgs a;
    a.kp = kp;
    a.n43 = bf.n43;
    a.tab = bf.tab + (size_t)p * NTAB * S;
    a.shift0 = kp.SHIFT + 64 * (it0 + p * S);
    a.nslot = (it1 - (it0 + p * S)) < S ? (it1 - (it0 + p * S)) : S;
    a.blk0 = (int64_t)j0 * AP26_N53_PER_N43;
    a.blk1 = (int64_t)j1 * AP26_N53_PER_N43;
    a.surv = bf.surv;
    a.nsurv = bf.cnt + C_NSURV;
    a.cap = bf.surv_cap;
    a.surv_job = nullptr;
    a.job = 0;
    a.err = (unsigned int *)(bf.cnt + C_ERR);
    for (int v = 0; v < NVW; v++) a.vc[v] = vw_offset(a.shift0, VW_P[v]);
    a.gq = bf.gq; /* sized for `grid` workgroups of sv.block threads (ap26_open) */
    /* the shift-pair sieves claim their blocks through a counter per launch,
       cleared with the other counters (at most 5 passes of 2 shifts) */
    a.slab = nullptr;
    if (S == 2) {
      if (p >= NCOUNT - C_SLAB) return AP26_E_INTERNAL;
      a.slab = (unsigned int *)(bf.cnt + C_SLAB + p);
    }
    hipLaunchKernelGGL(sv.fn, dim3(grid), dim3(sv.block), 0, ctx->stream, a);
    HIPCHK(hipGetLastError());
    if (st) st->sieve_launches++;
  }
  HIPCHK(hipEventRecord(ctx->ev[2], ctx->stream));
  if (st) st->positions += (int64_t)(j1 - j0) * AP26_N53_PER_N43 * AP26_N59_STEPS * nslot;
  return AP26_OK;
}

static int prepare_k(ap26_ctx *ctx, int32_t K, int32_t SHIFT, KParams *kp) {
  HostK hk;
  host_kconsts(K, &hk);
  kp->STEP = hk.STEP;
  kp->n0 = hk.n0;
  kp->S31 = hk.S[0];
  kp->S37 = hk.S[1];
  kp->S41 = hk.S[2];
  kp->S43 = hk.S[3];
  kp->S47 = hk.S[4];
  kp->S53 = hk.S[5];
  kp->S59 = hk.S[6];
  kp->K = K;
  kp->SHIFT = SHIFT;
  for (int v = 0; v < NVW; v++) host_vwin(K, VW_P[v], kp->vw[v]);
  std::vector<int64_t> n43(AP26_NN43);
  host_n43(&hk, n43.data());
  HIPCHK(hipMemcpyAsync(ctx->d_n43, n43.data(), sizeof(int64_t) * AP26_NN43, hipMemcpyHostToDevice,
                        ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream)); /* n43 is a stack-owned host buffer */
  hipLaunchKernelGGL(zero_kernel, dim3(1), dim3(64), 0, ctx->stream, ctx->d_cnt, NCOUNT);
  HIPCHK(hipGetLastError());
  return AP26_OK;
}

/* Grow the survivor buffer to at least `need` entries (contents dropped). */
static int ensure_surv(ap26_ctx *ctx, int64_t need) {
  if (need <= ctx->surv_cap) return AP26_OK;
  HIPCHK(hipStreamSynchronize(ctx->stream));
  (void)hipFree(ctx->d_surv);
  ctx->d_surv = nullptr;
  ctx->surv_cap = 0;
  if (hipMalloc(&ctx->d_surv, sizeof(int64_t) * need) != hipSuccess) {
    ctx->d_surv = nullptr;
    if (hipMalloc(&ctx->d_surv, sizeof(int64_t) * SURV_CAP) != hipSuccess) return AP26_E_NOMEM;
    ctx->surv_cap = SURV_CAP;
    return AP26_E_NOMEM;
  }
  ctx->surv_cap = need;
  return AP26_OK;
}

/* walk: the span is searched (walk and reports), not only sieved */
static bool span_ok(int32_t K, int32_t SHIFT, int32_t it0, int32_t it1, int32_t j0, int32_t j1, bool walk) {
  if (K < 1 || K > 600000000) return false;       /* per-K constants exact below ~6.3e8 */
  if (!host_shift_ok(K, SHIFT)) return false;   /* every term fits in int64 */
  if (walk && !host_walk_ok(K, SHIFT)) return false; /* the reference's 2^62 PrimeQ limit */
  if (it0 < 0 || it1 > AP26_NSHIFT || it0 > it1) return false;
  if (j0 < 0 || j1 > AP26_NN43 || j0 > j1) return false;
  return true;
}

extern "C" {

int ap26_search_span(ap26_ctx *ctx, int32_t K, int32_t SHIFT, int32_t it0, int32_t it1, int32_t j0,
                     int32_t j1, ap26_sol *out, int32_t cap, int32_t *n_out, ap26_stats *stats) {
  if (!ctx || !n_out || (cap > 0 && !out)) return AP26_E_RANGE;
  if (!span_ok(K, SHIFT, it0, it1, j0, j1, true)) return AP26_E_RANGE;
  auto t0 = std::chrono::steady_clock::now();
  HIPCHK(hipSetDevice(ctx->device));
  ap26_stats st{};
  *n_out = 0;
  if (it0 == it1 || j0 == j1) {
    if (stats) *stats = st;
    return AP26_OK;
  }
  KParams kp;
  int rc = ensure_surv(ctx, host_surv_budget(K));
  if (rc) return rc;
  unsigned long long cnt[NCOUNT];
  for (int attempt = 0;; attempt++) {
    st = ap26_stats{};
    rc = prepare_k(ctx, K, SHIFT, &kp);
    if (rc) return rc;
    HIPCHK(hipEventRecord(ctx->ev[0], ctx->stream));
    rc = enqueue_sieve(ctx, ctx_bufs(ctx), kp, it0, it1, j0, j1, &st);
    if (rc) return rc;
    WalkArgs w;
    w.surv = ctx->d_surv;
    w.nsurv = ctx->d_cnt + C_NSURV;
    w.cap = ctx->surv_cap;
    w.STEP = kp.STEP;
    w.K = K;
    w.SHIFT = SHIFT;
    w.range_id = -1;
    w.sols = ctx->d_sols;
    w.nsol = (unsigned int *)(ctx->d_cnt + C_NSOL);
    w.sol_cap = SOL_CAP;
    w.ntests = ctx->d_cnt + C_NTESTS;
    w.err = (unsigned int *)(ctx->d_cnt + C_ERR);
    w.nlive = ctx->d_cnt + C_LIVE;
    w.big = ctx->d_cnt + C_BIG;
    const WalkFn<WalkArgs> wfn = pick_walk(walk_depth_of(ctx));
    hipLaunchKernelGGL(wfn, dim3(ctx->cu_count * 8), dim3(256), 0, ctx->stream, w);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(verify_kernel, dim3(256), dim3(256), 0, ctx->stream, ctx->d_sols, w.nsol, SOL_CAP);
    hipLaunchKernelGGL(verify_finish_kernel, dim3(16), dim3(256), 0, ctx->stream, ctx->d_sols, w.nsol, SOL_CAP);
    HIPCHK(hipGetLastError());
    st.walk_launches++;
    HIPCHK(hipEventRecord(ctx->ev[3], ctx->stream));
    HIPCHK(hipMemcpyAsync(cnt, ctx->d_cnt, sizeof(cnt), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    /* more survivors than the buffer held (not seen for any K <= 6e8 within
       the budget): grow it to the count seen and search the span again */
    if (cnt[C_ERR]) return AP26_E_INTERNAL; /* a queue step ran on part of a wave */
    if (cnt[C_NSURV] <= (unsigned long long)ctx->surv_cap) break;
    if (attempt) return AP26_E_INTERNAL;
    rc = ensure_surv(ctx, (int64_t)(cnt[C_NSURV] + (cnt[C_NSURV] >> 3)));
    if (rc) return rc;
  }
  const unsigned int nsol = (unsigned int)(cnt[C_NSOL] & 0xffffffffu);
  if (nsol > (unsigned int)SOL_CAP) return AP26_E_INTERNAL;
  if (cnt[C_BIG]) { /* the walk reached 2^62, where the reference's PrimeQ exits */
    st.walk_past_limit = (int64_t)cnt[C_BIG];
    if (stats) *stats = st;
    return AP26_E_RANGE;
  }
  std::vector<DevSol> hs(nsol);
  if (nsol) {
    HIPCHK(hipMemcpy(hs.data(), ctx->d_sols, sizeof(DevSol) * nsol, hipMemcpyDeviceToHost));
  }
  int ncopy = (int)nsol < cap ? (int)nsol : cap;
  for (int i = 0; i < ncopy; i++) {
    out[i].k = hs[i].k;
    out[i].K = hs[i].K;
    out[i].first = hs[i].first;
    out[i].range_id = hs[i].range_id;
    out[i].shift_iter = hs[i].shift_iter;
    out[i].prp2_mask = hs[i].prp2;
    out[i].valid_mask = hs[i].valid;
  }
  *n_out = (int32_t)nsol;
  st.k_searched = 1;
  st.survivors = (int64_t)cnt[C_LIVE];
  st.prp_tests = (int64_t)cnt[C_NTESTS];
  st.reports = nsol;
  st.ms_tables = elapsed_ms(ctx->ev[0], ctx->ev[1]);
  st.ms_sieve = elapsed_ms(ctx->ev[1], ctx->ev[2]);
  st.ms_walk = elapsed_ms(ctx->ev[2], ctx->ev[3]);
  st.ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (stats) *stats = st;
  return (int)nsol > cap ? AP26_E_OVERFLOW : AP26_OK;
}

static_assert(sizeof(DevSol) == sizeof(ap26_sol) && offsetof(DevSol, first) == offsetof(ap26_sol, first) &&
                  offsetof(DevSol, range_id) == offsetof(ap26_sol, range_id) &&
                  offsetof(DevSol, prp2) == offsetof(ap26_sol, prp2_mask) &&
                  offsetof(DevSol, valid) == offsetof(ap26_sol, valid_mask),
              "device reports are copied out as ap26_sol");

int ap26_search_span_noalloc(ap26_ctx *ctx, int32_t K, int32_t SHIFT, int32_t it0, int32_t it1, int32_t j0,
                             int32_t j1, ap26_sol *out, int32_t cap, int32_t *n_out, ap26_stats *stats) {
  if (!ctx || !n_out || (cap > 0 && !out)) return AP26_E_RANGE;
  if (!span_ok(K, SHIFT, it0, it1, j0, j1, true)) return AP26_E_RANGE;
  if (j1 - j0 > NOALLOC_MAX_N43) { /* more than the scratch's survivor slots are sized for */
    if (ctx->opts.verbose)
      fprintf(stderr, "ap26_search_span_noalloc: %d n43 starts, at most %d\n", j1 - j0, NOALLOC_MAX_N43);
    return AP26_E_RANGE;
  }
  if (!ctx->scratch) return AP26_E_DEVICE;
  auto t0 = std::chrono::steady_clock::now();
  HIPCHK(hipSetDevice(ctx->device));
  ap26_stats st{};
  *n_out = 0;
  if (it0 == it1 || j0 == j1) {
    if (stats) *stats = st;
    return AP26_OK;
  }
  unsigned char *base = (unsigned char *)ctx->scratch;
  const SpanBufs bf{(uint64_t *)(base + NA_TAB), (int64_t *)(base + NA_N43), (int64_t *)(base + NA_SURV),
                    NOALLOC_SURV, (unsigned long long *)(base + NA_CNT), (DevSol *)(base + NA_SOLS),
                    NOALLOC_SOLS, &ctx->sieve_lds, nullptr}; /* every sieve queue in LDS */
  KParams kp;
  HostK hk;
  host_kconsts(K, &hk);
  kp.STEP = hk.STEP;
  kp.n0 = hk.n0;
  kp.S31 = hk.S[0];
  kp.S37 = hk.S[1];
  kp.S41 = hk.S[2];
  kp.S43 = hk.S[3];
  kp.S47 = hk.S[4];
  kp.S53 = hk.S[5];
  kp.S59 = hk.S[6];
  kp.K = K;
  kp.SHIFT = SHIFT;
  for (int v = 0; v < NVW; v++) host_vwin(K, VW_P[v], kp.vw[v]);
  /* the array keeps the last call's contents: clear the counters (a kernel,
     no memset node), compute this K's n43 starts of the span on the device */
  hipLaunchKernelGGL(zero_kernel, dim3(1), dim3(64), 0, ctx->stream, bf.cnt, NCOUNT);
  hipLaunchKernelGGL(n43_span_kernel, dim3((j1 - j0 + 63) / 64), dim3(64), 0, ctx->stream, kp, j0, j1, bf.n43);
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(ctx->ev[0], ctx->stream));
  int rc = enqueue_sieve(ctx, bf, kp, it0, it1, j0, j1, &st);
  if (rc) return rc;
  WalkArgs w;
  w.surv = bf.surv;
  w.nsurv = bf.cnt + C_NSURV;
  w.cap = bf.surv_cap;
  w.STEP = kp.STEP;
  w.K = K;
  w.SHIFT = SHIFT;
  w.range_id = -1;
  w.sols = bf.sols;
  w.nsol = (unsigned int *)(bf.cnt + C_NSOL);
  w.sol_cap = bf.sol_cap;
  w.ntests = bf.cnt + C_NTESTS;
  w.err = (unsigned int *)(bf.cnt + C_ERR);
  w.nlive = bf.cnt + C_LIVE;
  w.big = bf.cnt + C_BIG;
  const WalkFn<WalkArgs> wfn = pick_walk(walk_depth_of(ctx));
  hipLaunchKernelGGL(wfn, dim3(ctx->cu_count * 8), dim3(256), 0, ctx->stream, w);
  hipLaunchKernelGGL(verify_kernel, dim3(256), dim3(256), 0, ctx->stream, bf.sols, w.nsol, bf.sol_cap);
  hipLaunchKernelGGL(verify_finish_kernel, dim3(16), dim3(256), 0, ctx->stream, bf.sols, w.nsol, bf.sol_cap);
  HIPCHK(hipGetLastError());
  st.walk_launches++;
  HIPCHK(hipEventRecord(ctx->ev[3], ctx->stream));
  unsigned long long cnt[NCOUNT];
  HIPCHK(hipMemcpyAsync(cnt, bf.cnt, sizeof(cnt), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  if (cnt[C_ERR]) return AP26_E_INTERNAL;                                 /* a queue step on part of a wave */
  if (cnt[C_NSURV] > (unsigned long long)bf.surv_cap) return AP26_E_INTERNAL; /* exceeds the sizing bound */
  const unsigned int nsol = (unsigned int)(cnt[C_NSOL] & 0xffffffffu);
  if (nsol > (unsigned int)bf.sol_cap) return AP26_E_INTERNAL;
  st.k_searched = 1;
  st.survivors = (int64_t)cnt[C_LIVE];
  st.prp_tests = (int64_t)cnt[C_NTESTS];
  st.reports = nsol;
  st.ms_tables = elapsed_ms(ctx->ev[0], ctx->ev[1]);
  st.ms_sieve = elapsed_ms(ctx->ev[1], ctx->ev[2]);
  st.ms_walk = elapsed_ms(ctx->ev[2], ctx->ev[3]);
  if (cnt[C_BIG]) { /* the walk reached 2^62, where the reference's PrimeQ exits */
    st.walk_past_limit = (int64_t)cnt[C_BIG];
    if (stats) *stats = st;
    return AP26_E_RANGE;
  }
  const int ncopy = (int)nsol < cap ? (int)nsol : cap;
  if (ncopy > 0)
    HIPCHK(hipMemcpy(out, bf.sols, sizeof(ap26_sol) * (size_t)ncopy, hipMemcpyDeviceToHost));
  *n_out = (int32_t)nsol;
  st.ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (stats) *stats = st;
  return (int)nsol > cap ? AP26_E_OVERFLOW : AP26_OK;
}

int ap26_search_k(ap26_ctx *ctx, int32_t K, int32_t SHIFT, int32_t iter_start, ap26_sol *out, int32_t cap,
                  int32_t *n_out, ap26_stats *stats) {
  return ap26_search_span(ctx, K, SHIFT, iter_start, AP26_NSHIFT, 0, AP26_NN43, out, cap, n_out, stats);
}

int ap26_sieve_span(ap26_ctx *ctx, int32_t K, int32_t SHIFT, int32_t it0, int32_t it1, int32_t j0, int32_t j1,
                    int64_t *out, int64_t cap, int64_t *n_out) {
  if (!ctx || !n_out || (cap > 0 && !out)) return AP26_E_RANGE;
  if (!span_ok(K, SHIFT, it0, it1, j0, j1, false)) return AP26_E_RANGE;
  HIPCHK(hipSetDevice(ctx->device));
  *n_out = 0;
  if (it0 == it1 || j0 == j1) return AP26_OK;
  KParams kp;
  int rc = ensure_surv(ctx, host_surv_budget(K));
  if (rc) return rc;
  unsigned long long nsurv = 0;
  for (int attempt = 0;; attempt++) {
    rc = prepare_k(ctx, K, SHIFT, &kp);
    if (rc) return rc;
    HIPCHK(hipEventRecord(ctx->ev[0], ctx->stream));
    rc = enqueue_sieve(ctx, ctx_bufs(ctx), kp, it0, it1, j0, j1, nullptr);
    if (rc) return rc;
    unsigned long long err = 0;
    HIPCHK(hipMemcpyAsync(&nsurv, ctx->d_cnt + C_NSURV, sizeof(nsurv), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipMemcpyAsync(&err, ctx->d_cnt + C_ERR, sizeof(err), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (err) return AP26_E_INTERNAL;
    if (nsurv <= (unsigned long long)ctx->surv_cap) break;
    if (attempt) return AP26_E_INTERNAL;
    rc = ensure_surv(ctx, (int64_t)(nsurv + (nsurv >> 3)));
    if (rc) return rc;
  }
  /* the shift-pair sieve leaves unused slots (-1) at the end of each wave's
     last chunk: keep the survivors only */
  std::vector<int64_t> all((size_t)nsurv);
  if (nsurv) HIPCHK(hipMemcpy(all.data(), ctx->d_surv, sizeof(int64_t) * (size_t)nsurv, hipMemcpyDeviceToHost));
  int64_t n = 0;
  for (int64_t v : all)
    if (v >= 0) {
      if (n < cap) out[n] = v;
      n++;
    }
  *n_out = n;
  return n > cap ? AP26_E_OVERFLOW : AP26_OK;
}

int ap26_verify_sols(ap26_ctx *ctx, ap26_sol *sols, int32_t n) {
  if (!ctx || n < 0 || (n > 0 && !sols)) return AP26_E_RANGE;
  if (n > SOL_CAP) return AP26_E_RANGE;
  if (n == 0) return AP26_OK;
  for (int32_t i = 0; i < n; i++)
    if (sols[i].k < 0 || sols[i].K < 1 || sols[i].K > 600000000 || sols[i].first < 0) return AP26_E_RANGE;
  HIPCHK(hipSetDevice(ctx->device));
  std::vector<DevSol> hs(n);
  for (int32_t i = 0; i < n; i++) hs[i] = DevSol{sols[i].k, sols[i].K, sols[i].first, sols[i].range_id,
                                                sols[i].shift_iter, 0u, 0u};
  HIPCHK(hipMemcpyAsync(ctx->d_sols, hs.data(), sizeof(DevSol) * n, hipMemcpyHostToDevice, ctx->stream));
  const unsigned long long cnt = (unsigned long long)(uint32_t)n;
  HIPCHK(hipMemcpyAsync(ctx->d_cnt + C_NSOL, &cnt, sizeof(cnt), hipMemcpyHostToDevice, ctx->stream));
  hipLaunchKernelGGL(verify_kernel, dim3(256), dim3(256), 0, ctx->stream,
                     ctx->d_sols, (const unsigned int *)(ctx->d_cnt + C_NSOL), SOL_CAP);
  hipLaunchKernelGGL(verify_finish_kernel, dim3(16), dim3(256), 0, ctx->stream,
                     ctx->d_sols, (const unsigned int *)(ctx->d_cnt + C_NSOL), SOL_CAP);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(hs.data(), ctx->d_sols, sizeof(DevSol) * n, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  for (int32_t i = 0; i < n; i++) {
    sols[i].prp2_mask = hs[i].prp2;
    sols[i].valid_mask = hs[i].valid;
  }
  return AP26_OK;
}

int ap26_okok_table(ap26_ctx *ctx, int32_t K, int32_t SHIFT, uint64_t *out) {
  if (!ctx || !out) return AP26_E_RANGE;
  if (!span_ok(K, SHIFT, 0, 1, 0, 1, false)) return AP26_E_RANGE;
  HIPCHK(hipSetDevice(ctx->device));
  KParams kp;
  HostK hk;
  host_kconsts(K, &hk);
  kp.STEP = hk.STEP;
  kp.K = K;
  kp.SHIFT = SHIFT;
  dim3 tgrid((541 + 63) / 64, NPT, 1);
  hipLaunchKernelGGL(tables_kernel, tgrid, dim3(64), 0, ctx->stream, kp, 0, 1, 1, ctx->d_tab);
  HIPCHK(hipGetLastError());
  std::vector<uint64_t> t(NTAB);
  HIPCHK(hipMemcpyAsync(t.data(), ctx->d_tab, sizeof(uint64_t) * NTAB, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  for (int i = 0; i < NSP; i++) {
    int p = PT.p[i];
    for (int r = 0; r < p; r++) {
      int c = (r + (p - 1) / 2) % p; /* centred index of residue r */
      out[PT.off[i] + r] = t[PT.off[i] + c];
    }
  }
  return AP26_OK;
}

int ap26_prp_batch_dev(ap26_ctx *ctx, const void *d_n, void *d_out, size_t count, double *ms) {
  if (!ctx || (count && (!d_n || !d_out))) return AP26_E_RANGE;
  HIPCHK(hipSetDevice(ctx->device));
  if (count == 0) {
    if (ms) *ms = 0.0;
    return AP26_OK;
  }
  size_t blocks = (count + 255) / 256;
  size_t maxb = (size_t)ctx->cu_count * 64;
  if (blocks > maxb) blocks = maxb;
  HIPCHK(hipEventRecord(ctx->ev[4], ctx->stream));
  hipLaunchKernelGGL(prp_kernel, dim3((unsigned)blocks), dim3(256), 0, ctx->stream, (const uint64_t *)d_n,
                     (uint8_t *)d_out, (uint64_t)count);
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(ctx->ev[5], ctx->stream));
  if (ms) {
    HIPCHK(hipEventSynchronize(ctx->ev[5]));
    *ms = elapsed_ms(ctx->ev[4], ctx->ev[5]);
  }
  return AP26_OK;
}

int ap26_prp_batch(ap26_ctx *ctx, const uint64_t *n, uint8_t *out, size_t count) {
  if (!ctx || (count && (!n || !out))) return AP26_E_RANGE;
  if (count == 0) return AP26_OK;
  HIPCHK(hipSetDevice(ctx->device));
  void *dn = nullptr, *dout = nullptr;
  if (hipMalloc(&dn, count * 8) != hipSuccess) return AP26_E_NOMEM;
  if (hipMalloc(&dout, count) != hipSuccess) {
    (void)hipFree(dn);
    return AP26_E_NOMEM;
  }
  int rc = AP26_OK;
  if (hipMemcpy(dn, n, count * 8, hipMemcpyHostToDevice) != hipSuccess) rc = AP26_E_DEVICE;
  if (!rc) rc = ap26_prp_batch_dev(ctx, dn, dout, count, nullptr);
  if (!rc && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = AP26_E_DEVICE;
  if (!rc && hipMemcpy(out, dout, count, hipMemcpyDeviceToHost) != hipSuccess) rc = AP26_E_DEVICE;
  (void)hipFree(dn);
  (void)hipFree(dout);
  return rc;
}

int ap26_dev_alloc(ap26_ctx *ctx, size_t bytes, void **dptr) {
  if (!ctx || !dptr) return AP26_E_RANGE;
  HIPCHK(hipSetDevice(ctx->device));
  if (hipMalloc(dptr, bytes ? bytes : 1) != hipSuccess) return AP26_E_NOMEM;
  return AP26_OK;
}

int ap26_dev_free(ap26_ctx *ctx, void *dptr) {
  if (!ctx) return AP26_E_RANGE;
  HIPCHK(hipSetDevice(ctx->device));
  HIPCHK(hipFree(dptr));
  return AP26_OK;
}

int ap26_copy_h2d(ap26_ctx *ctx, void *dst, const void *src, size_t bytes) {
  if (!ctx) return AP26_E_RANGE;
  HIPCHK(hipSetDevice(ctx->device));
  HIPCHK(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return AP26_OK;
}

int ap26_copy_d2h(ap26_ctx *ctx, void *dst, const void *src, size_t bytes) {
  if (!ctx) return AP26_E_RANGE;
  HIPCHK(hipSetDevice(ctx->device));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  HIPCHK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return AP26_OK;
}

int ap26_sync(ap26_ctx *ctx) {
  if (!ctx) return AP26_E_RANGE;
  HIPCHK(hipSetDevice(ctx->device));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return AP26_OK;
}

}  // extern "C"
