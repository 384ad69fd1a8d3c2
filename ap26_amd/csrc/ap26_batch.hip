/* ap26_batch.hip -- the batched engine behind ap26_search_batch: many work
 * units (KMIN..KMAX, SHIFT, resumed shift) expanded into K jobs and searched
 * in rounds of up to `per_round` K.  One round is a captured hipGraph
 *   n43 starts -> tables -> persistent work-stealing sieve -> walk -> advance
 * whose kernels read the round number from device memory, so the same graph
 * is replayed for every round with no host synchronisation in between.
 * Reports carry their range id; per-range hash totals are accumulated on the
 * device as exact 64-bit sums.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <map>
#include <vector>

#include "ap26_ctx.h"
#include "ap26_host_math.h"
#include "ap26_pvariants.h"

using namespace ap26;

namespace {

constexpr int BATCH_SOL_CAP = 1 << 20;
/* n43 starts per work item: small items keep the end of a round short (the
   CUs finish their last items within a few ms of each other), and restaging
   the LDS tables when a CU moves to another (K, shift pair) costs little */
constexpr int DEFAULT_CHUNK = 11;
constexpr int BATCH_SEG_K = 2048; /* searched K per segment of one batched call */
constexpr int MAX_RANGES = 1 << 20;

/* The variant `id` of the persistent sieve (ap26_pvariants.h); 0 or an
   unknown id: 117, the default. */
PersistVariant pick_pvariant(int id) {
  bool (*const parts[])(int, PersistVariant &) = {pick_pvariant_1, pick_pvariant_2, pick_pvariant_3,
                                                   pick_pvariant_4, pick_pvariant_5};
  PersistVariant v{0, 0, 0, nullptr};
  for (const int want : {id, 117})
    for (auto part : parts)
      if (part(want, v)) return v;
  return v; /* not reached: 117 is in the table */
}

}  // namespace

/* Per-round inputs and outputs.  The batched call uses io[0]; the server
   alternates io[0] and io[1], so that one round runs while the host harvests
   the one before.  The large buffers (tables, n43 starts, survivors) are
   shared: rounds run one after another on the context's stream. */
struct RoundIO {
  BatchJobDev *d_jobs = nullptr;
  RoundState *d_rs = nullptr;
  unsigned long long *d_cnt = nullptr; /* [0] nsurv [1] tests [2] nsol(u32) [3] work(u32) [4] overflow [5] survivors [6] err */
  DevSol *d_sols = nullptr;
  unsigned long long *d_hash = nullptr; /* per range: exact sum of first%1000 + k */
  unsigned long long *d_big = nullptr;  /* per range: walked terms >= 2^62 */
  /* one round = graph A (n43 starts, tables, persistent sieve) then graph B
     (walk, advance); events recorded between them time the sieve side */
  hipGraphExec_t graph = nullptr;
  hipGraphExec_t graph_b = nullptr;
};

struct BatchState {
  int per_round = 0;
  int chunk = 0, nchunk = 0;
  int pvariant = -1;
  PersistVariant pv{};
  int walk_depth = 1;        /* compaction depth of the batched walk */
  int64_t surv_cap = 0;      /* survivor slots per round */
  int64_t surv_slack = 0;    /* slots the sieve's waves may reserve and leave unused */
  int max_jobs = 0;          /* capacity of each io's d_jobs */
  int max_ranges = 0;        /* capacity of each io's d_hash and d_big */
  int nio = 1;
  RoundIO io[2];
  uint32_t *d_triples = nullptr;
  int64_t *d_n43 = nullptr;
  uint64_t *d_tab = nullptr;
  int64_t *d_surv = nullptr;
  int32_t *d_surv_job = nullptr;
  QEntry *d_gq = nullptr;     /* global queue regions: cu_count x waves x pv.gq_entries */
  std::vector<hipEvent_t> ev; /* 2 per timed round */
};

static void drop_graphs(BatchState *b) {
  for (RoundIO &o : b->io) {
    if (o.graph) (void)hipGraphExecDestroy(o.graph);
    if (o.graph_b) (void)hipGraphExecDestroy(o.graph_b);
    o.graph = o.graph_b = nullptr;
  }
}

static void batch_state_free(BatchState *b) {
  if (!b) return;
  drop_graphs(b);
  for (auto &e : b->ev) (void)hipEventDestroy(e);
  for (RoundIO &o : b->io) {
    (void)hipFree(o.d_jobs);
    (void)hipFree(o.d_rs);
    (void)hipFree(o.d_cnt);
    (void)hipFree(o.d_sols);
    (void)hipFree(o.d_hash);
    (void)hipFree(o.d_big);
  }
  (void)hipFree(b->d_triples);
  (void)hipFree(b->d_n43);
  (void)hipFree(b->d_tab);
  (void)hipFree(b->d_surv);
  (void)hipFree(b->d_surv_job);
  (void)hipFree(b->d_gq);
  delete b;
}

/* per_round = K held by one round (device memory ~0.61 GB per K) */
void batch_free(ap26_ctx *ctx) {
  batch_state_free(ctx->batch);
  ctx->batch = nullptr;
}

static int batch_state_new(ap26_ctx *ctx, int per_round, int nio, BatchState **out) {
  *out = nullptr;
  BatchState *b = new (std::nothrow) BatchState();
  if (!b) return AP26_E_NOMEM;
  int rc = AP26_OK;
  b->per_round = per_round;
  b->nio = nio;
  /* environment overrides > autotuned choice > defaults */
  const char *e = getenv("AP26_BATCH_CHUNK");
  b->chunk = e ? atoi(e) : (ctx->tune_chunk > 0 ? ctx->tune_chunk : DEFAULT_CHUNK);
  if (b->chunk < 1 || b->chunk > AP26_NN43) b->chunk = DEFAULT_CHUNK;
  b->nchunk = (AP26_NN43 + b->chunk - 1) / b->chunk;
  e = getenv("AP26_BATCH_VARIANT");
  b->pvariant = e ? atoi(e) : (ctx->tune_variant >= 0 ? ctx->tune_variant : 0);
  b->pv = pick_pvariant(b->pvariant);
  b->walk_depth = walk_depth_of(ctx);
  b->surv_slack = (int64_t)ctx->cu_count * b->pv.waves * SURV_CHUNK;
  b->surv_cap = (int64_t)b->per_round * (INT64_C(48) << 20) + b->surv_slack; /* ~40M survivors per K */
  /* guarded (i31, i37, i41) triples in the reference order */
  std::vector<uint32_t> trip;
  for (int i31 = 0; i31 < 7; i31++)
    for (int i37 = 0; i37 < 13; i37++) {
      if (i37 - i31 > 10 || i31 - i37 > 4) continue;
      for (int i41 = 0; i41 < 17; i41++) {
        if (i41 - i31 > 14 || i41 - i37 > 14 || i31 - i41 > 4 || i37 - i41 > 10) continue;
        trip.push_back((uint32_t)i31 | ((uint32_t)i37 << 8) | ((uint32_t)i41 << 16));
      }
    }
  if (trip.size() * 8 != AP26_NN43) rc = AP26_E_INTERNAL;
  for (int i = 0; i < nio && !rc; i++)
    if (hipMalloc(&b->io[i].d_rs, sizeof(RoundState)) != hipSuccess ||
        hipMalloc(&b->io[i].d_cnt, sizeof(unsigned long long) * 8) != hipSuccess ||
        hipMalloc(&b->io[i].d_sols, sizeof(DevSol) * BATCH_SOL_CAP) != hipSuccess)
      rc = AP26_E_NOMEM;
  if (!rc && (hipMalloc(&b->d_triples, trip.size() * 4) != hipSuccess ||
              hipMalloc(&b->d_n43, sizeof(int64_t) * AP26_NN43 * b->per_round) != hipSuccess ||
              hipMalloc(&b->d_tab, sizeof(uint64_t) * NTAB * AP26_NSHIFT * b->per_round) != hipSuccess ||
              hipMalloc(&b->d_surv, sizeof(int64_t) * b->surv_cap) != hipSuccess ||
              hipMalloc(&b->d_surv_job, sizeof(int32_t) * b->surv_cap) != hipSuccess))
    rc = AP26_E_NOMEM;
  /* one region per wave of the one-workgroup-per-CU launch (enqueue_round) */
  if (!rc && b->pv.gq_entries > 0 &&
      hipMalloc(&b->d_gq, sizeof(QEntry) * (size_t)ctx->cu_count * b->pv.waves * b->pv.gq_entries) != hipSuccess)
    rc = AP26_E_NOMEM;
  if (!rc && hipMemcpy(b->d_triples, trip.data(), trip.size() * 4, hipMemcpyHostToDevice) != hipSuccess)
    rc = AP26_E_DEVICE;
  if (rc) {
    batch_state_free(b);
    return rc;
  }
  *out = b;
  return AP26_OK;
}

static int batch_init(ap26_ctx *ctx) {
  if (ctx->batch) return AP26_OK;
  const char *e = getenv("AP26_BATCH_ROUND");
  int per_round = e ? atoi(e) : 16;
  if (per_round < 1 || per_round > 64) per_round = 16;
  return batch_state_new(ctx, per_round, 1, &ctx->batch);
}

static int ensure_capacity(BatchState *b, int njobs, int nranges) {
  if (njobs > b->max_jobs) {
    drop_graphs(b);
    b->max_jobs = 0;
    const int cap = njobs < 1024 ? 1024 : njobs;
    for (int i = 0; i < b->nio; i++) {
      (void)hipFree(b->io[i].d_jobs);
      b->io[i].d_jobs = nullptr;
      if (hipMalloc(&b->io[i].d_jobs, sizeof(BatchJobDev) * cap) != hipSuccess) return AP26_E_NOMEM;
    }
    b->max_jobs = cap;
  }
  if (nranges > b->max_ranges) {
    drop_graphs(b);
    b->max_ranges = 0;
    const int cap = nranges < 1024 ? 1024 : nranges;
    for (int i = 0; i < b->nio; i++) {
      RoundIO &o = b->io[i];
      (void)hipFree(o.d_hash);
      (void)hipFree(o.d_big);
      o.d_hash = o.d_big = nullptr;
      if (hipMalloc(&o.d_hash, sizeof(unsigned long long) * cap) != hipSuccess ||
          hipMalloc(&o.d_big, sizeof(unsigned long long) * cap) != hipSuccess)
        return AP26_E_NOMEM;
    }
    b->max_ranges = cap;
  }
  return AP26_OK;
}

/* Grow the per-round survivor buffers to `need` slots; the graphs hold the
   old pointers and are captured again.  Waits for the rounds in flight. */
static int ensure_surv_cap(ap26_ctx *ctx, BatchState *b, int64_t need) {
  if (need <= b->surv_cap) return AP26_OK;
  HIPCHK(hipStreamSynchronize(ctx->stream));
  (void)hipFree(b->d_surv);
  (void)hipFree(b->d_surv_job);
  b->d_surv = nullptr;
  b->d_surv_job = nullptr;
  b->surv_cap = 0;
  drop_graphs(b);
  if (hipMalloc(&b->d_surv, sizeof(int64_t) * need) != hipSuccess ||
      hipMalloc(&b->d_surv_job, sizeof(int32_t) * need) != hipSuccess) {
    (void)hipFree(b->d_surv);
    (void)hipFree(b->d_surv_job);
    b->d_surv = nullptr;
    b->d_surv_job = nullptr;
    return AP26_E_NOMEM;
  }
  b->surv_cap = need;
  return AP26_OK;
}

/* enqueue one round on ctx->stream (captured into the graphs): part 0 =
   n43 starts, tables, sieve; part 1 = walk, advance */
static int enqueue_round(ap26_ctx *ctx, BatchState *b, int part, const RoundIO &o) {
  const int R = b->per_round;
  if (part == 0) {
    hipLaunchKernelGGL(n43_kernel, dim3((R * AP26_NN43 + 255) / 256), dim3(256), 0, ctx->stream, o.d_jobs, o.d_rs,
                       b->d_triples, b->d_n43);
    hipLaunchKernelGGL(batch_tables_kernel, dim3((541 + 63) / 64, NPT, R * AP26_NSHIFT), dim3(64), 0,
                       ctx->stream, o.d_jobs, o.d_rs, b->d_tab, b->pv.S);
    PersistArgs p;
    p.jobs = o.d_jobs;
    p.rs = o.d_rs;
    p.n43 = b->d_n43;
    p.tab = b->d_tab;
    p.surv = b->d_surv;
    p.surv_job = b->d_surv_job;
    p.nsurv = o.d_cnt + 0;
    p.cap = b->surv_cap;
    p.work = (unsigned int *)(o.d_cnt + 3);
    p.err = (unsigned int *)(o.d_cnt + 6);
    p.chunk = b->chunk;
    p.nchunk = b->nchunk;
    p.gq = b->d_gq; /* sized for dim3(ctx->cu_count) workgroups of pv.waves waves */
    hipLaunchKernelGGL(b->pv.fn, dim3(ctx->cu_count), dim3(64 * b->pv.waves), 0, ctx->stream, p);
    return hipGetLastError() == hipSuccess ? AP26_OK : AP26_E_DEVICE;
  }
  BatchWalkArgs w;
  w.jobs = o.d_jobs;
  w.rs = o.d_rs;
  w.surv = b->d_surv;
  w.surv_job = b->d_surv_job;
  w.nsurv = o.d_cnt + 0;
  w.cap = b->surv_cap;
  w.sols = o.d_sols;
  w.nsol = (unsigned int *)(o.d_cnt + 2);
  w.sol_cap = BATCH_SOL_CAP;
  w.ntests = o.d_cnt + 1;
  w.hash_tot = o.d_hash;
  w.overflow = o.d_cnt + 4;
  w.nlive = o.d_cnt + 5;
  w.err = (unsigned int *)(o.d_cnt + 6);
  w.big = o.d_big;
  const WalkFn<BatchWalkArgs> wfn = pick_batch_walk(b->walk_depth);
  hipLaunchKernelGGL(wfn, dim3(ctx->cu_count * 8), dim3(256), 0, ctx->stream, w);
  hipLaunchKernelGGL(round_advance_kernel, dim3(1), dim3(64), 0, ctx->stream, o.d_rs, o.d_cnt + 0,
                     (unsigned int *)(o.d_cnt + 3));
  return hipGetLastError() == hipSuccess ? AP26_OK : AP26_E_DEVICE;
}

static int capture(ap26_ctx *ctx, BatchState *b, int part, const RoundIO &o, hipGraphExec_t *out) {
  hipGraph_t g;
  HIPCHK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
  int rc = enqueue_round(ctx, b, part, o);
  hipError_t ce = hipStreamEndCapture(ctx->stream, &g);
  if (rc || ce != hipSuccess) return AP26_E_DEVICE;
  hipError_t ie = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  return ie == hipSuccess ? AP26_OK : AP26_E_DEVICE;
}

static int ensure_graph(ap26_ctx *ctx, BatchState *b, int i) {
  RoundIO &o = b->io[i];
  if (o.graph && o.graph_b) return AP26_OK;
  int rc = o.graph ? AP26_OK : capture(ctx, b, 0, o, &o.graph);
  if (!rc && !o.graph_b) rc = capture(ctx, b, 1, o, &o.graph_b);
  return rc;
}

/* launch one round of io i; with ev0/ev1 recorded around the sieve graph */
static int launch_round(ap26_ctx *ctx, BatchState *b, int i, hipEvent_t ev0, hipEvent_t ev1) {
  if (ev0) HIPCHK(hipEventRecord(ev0, ctx->stream));
  HIPCHK(hipGraphLaunch(b->io[i].graph, ctx->stream));
  if (ev1) HIPCHK(hipEventRecord(ev1, ctx->stream));
  HIPCHK(hipGraphLaunch(b->io[i].graph_b, ctx->stream));
  return AP26_OK;
}

static BatchJobDev make_job(int32_t K, int32_t shift, int32_t it0, int32_t range_id) {
  BatchJobDev j;
  HostK hk;
  host_kconsts(K, &hk);
  j.kp.STEP = hk.STEP;
  j.kp.n0 = hk.n0;
  j.kp.S31 = hk.S[0];
  j.kp.S37 = hk.S[1];
  j.kp.S41 = hk.S[2];
  j.kp.S43 = hk.S[3];
  j.kp.S47 = hk.S[4];
  j.kp.S53 = hk.S[5];
  j.kp.S59 = hk.S[6];
  j.kp.K = K;
  j.kp.SHIFT = shift;
  for (int v = 0; v < NVW; v++) host_vwin(K, VW_P[v], j.kp.vw[v]);
  j.it0 = it0;
  j.range_id = range_id;
  return j;
}

/* walk: the unit is searched (walked and reported), not only sieved; STEP,
   and with it every bound, is largest at kmax */
static bool unit_ok(const ap26_range &u, bool walk = true) {
  return !(u.kmin < 1 || u.kmax > 600000000 || u.kmin > u.kmax + 1 || !host_shift_ok(u.kmax, u.shift) ||
           (walk && !host_walk_ok(u.kmax, u.shift)) || u.iter_start < 0 || u.iter_start > AP26_NSHIFT);
}

__global__ void batch_reset_kernel(RoundState *rs, int njobs, int per_round, unsigned long long *cnt,
                                   unsigned long long *hash, unsigned long long *big, int nranges) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) {
    rs->base = 0;
    rs->njobs = njobs;
    rs->per_round = per_round;
    rs->pad = 0;
  }
  if (i < 8) cnt[i] = 0;
  if (i < nranges) {
    hash[i] = 0;
    big[i] = 0;
  }
}

extern "C" int ap26_search_batch(ap26_ctx *ctx, const ap26_range *r, int32_t nr, ap26_sol *out, int32_t cap,
                                 int32_t *n_out, int64_t *hash_totals, ap26_stats *stats) {
  if (!ctx || !n_out || nr < 0 || (nr > 0 && !r) || (cap > 0 && !out)) return AP26_E_RANGE;
  if (nr > MAX_RANGES) return AP26_E_RANGE;
  auto t0 = std::chrono::steady_clock::now();
  HIPCHK(hipSetDevice(ctx->device));
  *n_out = 0;
  ap26_stats st{};
  /* expand units into searched K */
  std::vector<BatchJobDev> jobs;
  for (int i = 0; i < nr; i++) {
    const ap26_range &u = r[i];
    if (!unit_ok(u)) return AP26_E_RANGE;
    for (int64_t K = u.kmin; K <= u.kmax; K++) {
      if (!host_will_search((int32_t)K)) {
        st.k_skipped++;
        continue;
      }
      jobs.push_back(make_job((int32_t)K, u.shift, (K == u.kmin) ? u.iter_start : 0, i));
    }
  }
  st.k_searched = (int64_t)jobs.size();
  if (hash_totals)
    for (int i = 0; i < nr; i++) hash_totals[i] = 0;
  if (jobs.empty()) {
    if (stats) *stats = st;
    return AP26_OK;
  }
  int rc = batch_init(ctx);
  if (rc) return rc;
  BatchState *b = ctx->batch;
  /* Segments of at most seg_k searched K run back to back; each ends with
     the device verification and one copy of its reports, so the device
     report buffer (BATCH_SOL_CAP, ~180 reports per K) bounds a segment, not
     the call.  A segment whose reports overflow it runs again in halves. */
  int seg_k = BATCH_SEG_K;
  unsigned int sol_limit = BATCH_SOL_CAP;
  {
    /* test hooks: a smaller segment, a smaller report budget per segment */
    const char *e = getenv("AP26_BATCH_SEG_K");
    if (e && atoi(e) > 0) seg_k = atoi(e);
    e = getenv("AP26_BATCH_SOL_LIMIT");
    if (e && atoi(e) > 0 && atoi(e) < BATCH_SOL_CAP) sol_limit = (unsigned int)atoi(e);
  }
  rc = ensure_capacity(b, (int)std::min<size_t>(jobs.size(), (size_t)seg_k), nr);
  if (rc) return rc;
  std::vector<DevSol> hs;
  std::vector<unsigned long long> tot_all(nr, 0), tot(nr), big(nr);
  unsigned long long cnt[8];
  int timed_total = 0;
  for (size_t seg0 = 0; seg0 < jobs.size();) {
    const int segn = (int)std::min<size_t>(jobs.size() - seg0, (size_t)seg_k);
    const BatchJobDev *sj = jobs.data() + seg0;
    /* survivor slots: the heaviest round's budget (K divisible by sieve
       primes keep more survivors) */
    int64_t need = 0;
    for (int r0 = 0; r0 < segn; r0 += b->per_round) {
      int64_t sum = 0;
      for (int j = r0; j < segn && j < r0 + b->per_round; j++) sum += host_surv_budget(sj[j].kp.K);
      need = sum > need ? sum : need;
    }
    need += b->surv_slack;
    const int rounds = (segn + b->per_round - 1) / b->per_round;
    const int timed = rounds < 1024 ? rounds : 1024;
    while ((int)b->ev.size() < 2 * timed) {
      hipEvent_t e;
      HIPCHK(hipEventCreate(&e));
      b->ev.push_back(e);
    }
    for (int attempt = 0;; attempt++) {
      rc = ensure_surv_cap(ctx, b, need);
      if (rc) return rc;
      HIPCHK(hipMemcpyAsync(b->io[0].d_jobs, sj, sizeof(BatchJobDev) * segn, hipMemcpyHostToDevice, ctx->stream));
      const int nres = nr > 8 ? nr : 8;
      hipLaunchKernelGGL(batch_reset_kernel, dim3((nres + 255) / 256), dim3(256), 0, ctx->stream, b->io[0].d_rs, segn,
                         b->per_round, b->io[0].d_cnt, b->io[0].d_hash, b->io[0].d_big, nr);
      HIPCHK(hipGetLastError());
      rc = ensure_graph(ctx, b, 0);
      if (rc) return rc;
      HIPCHK(hipEventRecord(ctx->ev[6], ctx->stream));
      for (int i = 0; i < rounds; i++) {
        rc = i < timed ? launch_round(ctx, b, 0, b->ev[2 * i], b->ev[2 * i + 1])
                       : launch_round(ctx, b, 0, nullptr, nullptr);
        if (rc) return rc;
      }
      hipLaunchKernelGGL(verify_kernel, dim3(256), dim3(256), 0, ctx->stream, b->io[0].d_sols,
                         (const unsigned int *)(b->io[0].d_cnt + 2), BATCH_SOL_CAP);
      hipLaunchKernelGGL(verify_finish_kernel, dim3(16), dim3(256), 0, ctx->stream, b->io[0].d_sols,
                         (const unsigned int *)(b->io[0].d_cnt + 2), BATCH_SOL_CAP);
      HIPCHK(hipGetLastError());
      HIPCHK(hipEventRecord(ctx->ev[7], ctx->stream));
      HIPCHK(hipMemcpyAsync(cnt, b->io[0].d_cnt, sizeof(cnt), hipMemcpyDeviceToHost, ctx->stream));
      HIPCHK(hipMemcpyAsync(tot.data(), b->io[0].d_hash, sizeof(unsigned long long) * nr, hipMemcpyDeviceToHost,
                            ctx->stream));
      HIPCHK(hipMemcpyAsync(big.data(), b->io[0].d_big, sizeof(unsigned long long) * nr, hipMemcpyDeviceToHost,
                            ctx->stream));
      HIPCHK(hipStreamSynchronize(ctx->stream));
      if (cnt[6]) return AP26_E_INTERNAL; /* a queue step ran on part of a wave */
      {
        /* a walk reached 2^62, where the reference's PrimeQ exits: no reports */
        int64_t nbig = 0;
        for (int i = 0; i < nr; i++) nbig += (int64_t)big[i];
        if (nbig) {
          st.walk_past_limit = nbig;
          if (stats) *stats = st;
          return AP26_E_RANGE;
        }
      }
      /* a round kept more survivors than its budget: grow the buffers to the
         largest count seen and run the segment again */
      if (cnt[4] == 0) break;
      if (attempt) return AP26_E_INTERNAL;
      need = (int64_t)(cnt[4] + (cnt[4] >> 3)) + b->surv_slack;
    }
    const unsigned int nsol = (unsigned int)(cnt[2] & 0xffffffffu);
    if (nsol > sol_limit && segn > 1) {
      seg_k = (segn + 1) / 2;
      continue;
    }
    if (nsol > (unsigned int)BATCH_SOL_CAP) return AP26_E_INTERNAL; /* one K, > 2^20 reports */
    const size_t h0 = hs.size();
    hs.resize(h0 + nsol);
    if (nsol) HIPCHK(hipMemcpy(hs.data() + h0, b->io[0].d_sols, sizeof(DevSol) * nsol, hipMemcpyDeviceToHost));
    for (int i = 0; i < nr; i++) tot_all[i] += tot[i];
    st.prp_tests += (int64_t)cnt[1];
    st.survivors += (int64_t)cnt[5];
    {
      /* ms_walk: device time of all rounds; ms_sieve: sum over the timed
         rounds of the sieve graph (n43 starts + tables + persistent sieve) */
      float ms = 0;
      if (hipEventElapsedTime(&ms, ctx->ev[6], ctx->ev[7]) == hipSuccess) st.ms_walk += ms;
      for (int i = 0; i < timed; i++)
        if (hipEventElapsedTime(&ms, b->ev[2 * i], b->ev[2 * i + 1]) == hipSuccess) st.ms_sieve += ms;
    }
    timed_total += timed;
    seg0 += (size_t)segn;
  }
  const int64_t nsol = (int64_t)hs.size();
  if (nsol > INT32_MAX) return AP26_E_INTERNAL;
  const int ncopy = nsol < cap ? (int)nsol : cap;
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
  if (hash_totals)
    for (int i = 0; i < nr; i++) hash_totals[i] = (int64_t)tot_all[i];
  st.reports = nsol;
  for (const BatchJobDev &j : jobs)
    st.positions += (int64_t)AP26_NN43 * AP26_N53_PER_N43 * AP26_N59_STEPS * (AP26_NSHIFT - j.it0);
  st.sieve_launches = timed_total;
  st.ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (stats) *stats = st;
  return nsol > cap ? AP26_E_OVERFLOW : AP26_OK;
}

extern "C" int ap26_batch_sieve_k(ap26_ctx *ctx, int32_t K, int32_t SHIFT, int32_t iter_start, int64_t *out,
                                  int64_t cap, int64_t *n_out) {
  if (!ctx || !n_out || cap < 0 || (cap > 0 && !out)) return AP26_E_RANGE;
  *n_out = 0;
  const ap26_range u{K, K, SHIFT, iter_start};
  if (!unit_ok(u, false) || K < 1 || !host_will_search(K)) return AP26_E_RANGE;
  HIPCHK(hipSetDevice(ctx->device));
  int rc = batch_init(ctx);
  if (rc) return rc;
  BatchState *b = ctx->batch;
  rc = ensure_capacity(b, 1, 1);
  if (!rc) rc = ensure_surv_cap(ctx, b, host_surv_budget(K) + b->surv_slack);
  if (rc) return rc;
  const BatchJobDev job = make_job(K, SHIFT, iter_start, 0);
  HIPCHK(hipMemcpy(b->io[0].d_jobs, &job, sizeof(job), hipMemcpyHostToDevice));
  hipLaunchKernelGGL(batch_reset_kernel, dim3(1), dim3(256), 0, ctx->stream, b->io[0].d_rs, 1, b->per_round, b->io[0].d_cnt,
                     b->io[0].d_hash, b->io[0].d_big, 1);
  HIPCHK(hipGetLastError());
  rc = ensure_graph(ctx, b, 0);
  if (rc) return rc;
  HIPCHK(hipGraphLaunch(b->io[0].graph, ctx->stream)); /* n43 starts, tables, sieve: no walk */
  unsigned long long cnt[8];
  HIPCHK(hipMemcpyAsync(cnt, b->io[0].d_cnt, sizeof(cnt), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  const unsigned long long nres = cnt[0];
  if (cnt[6] || nres > (unsigned long long)b->surv_cap) return AP26_E_INTERNAL;
  std::vector<int64_t> all(nres);
  if (nres) HIPCHK(hipMemcpy(all.data(), b->d_surv, sizeof(int64_t) * nres, hipMemcpyDeviceToHost));
  int64_t n = 0;
  for (int64_t v : all)
    if (v >= 0) {
      if (n < cap) out[n] = v;
      n++;
    }
  *n_out = n;
  return n > cap ? AP26_E_OVERFLOW : AP26_OK;
}

/* ---- serving ------------------------------------------------------------ */
/* ---- serving ------------------------------------------------------------ */
/* Continuous batching.  Units are admitted by slots and by a device-memory
   budget (R K per round); each round takes the next K of every active unit
   in turn.  Up to two rounds are in flight: when the units waiting fill a
   whole second round, pump() queues it on the stream behind the running one
   before it waits for the older round, so a saturated device never idles
   while the host harvests a round, admits new units and builds the next;
   below saturation it queues nothing behind the running round, so a unit
   that arrives meanwhile joins the very next round.  A round's outputs (counters, reports,
   per-slot hash sums and 2^62 counts) live in its own RoundIO and are copied
   to pinned host buffers on the stream; per-unit hash totals are summed on
   the host from the per-round sums, so a round can be run again (a survivor
   buffer that overflowed is grown first) without undoing anything on the
   device. */

constexpr int SERVER_SOL_COPY = 1 << 14; /* reports per round copied back with the round */

struct SUnit {
  ap26_range r{};
  std::vector<int32_t> ks; /* searched K */
  size_t next = 0;         /* next K to schedule */
  int outstanding = 0;     /* K in rounds in flight */
  int slot = -1;
  double t_submit = 0, t_done = 0;
  std::vector<ap26_sol> sols;
  int64_t hash = 0;        /* exact total over the harvested rounds */
  int status = AP26_OK;    /* AP26_E_RANGE: a walk reached 2^62 */
  bool done = false;
};

struct SRound {
  int io = 0;
  std::vector<int64_t> owner;     /* ticket of each job */
  std::vector<BatchJobDev> jobs;
  int64_t slots = 0;              /* survivor slots it needs */
  bool rerun = false;
};

struct SPinned {                  /* pinned host side of one RoundIO */
  BatchJobDev *jobs = nullptr;
  RoundState *rs = nullptr;
  unsigned long long *cnt = nullptr;
  unsigned long long *hash = nullptr;
  unsigned long long *big = nullptr;
  DevSol *sols = nullptr;
  hipEvent_t done = nullptr;
};

struct ap26_server {
  ap26_ctx *ctx = nullptr;
  BatchState *b = nullptr;
  int max_units = 0;
  int R = 0;
  std::map<int64_t, SUnit> units; /* admitted, not yet collected */
  std::deque<int64_t> active;     /* admission order, units with K left */
  std::vector<int> free_slots;
  int64_t next_ticket = 1;
  int64_t units_done = 0, k_done = 0, rounds = 0;
  size_t rr = 0;                  /* round-robin cursor */
  SPinned pin[2];
  std::deque<SRound> inflight;    /* oldest first, at most 2 */
  int next_io = 0;
};

static double now_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

/* clear one RoundIO's counters and per-slot sums before its round */
__global__ void server_zero_kernel(unsigned long long *cnt, unsigned long long *hash, unsigned long long *big,
                                   int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < 8) cnt[i] = 0;
  if (i < n) {
    hash[i] = 0;
    big[i] = 0;
  }
}

static void server_free(ap26_server *s) {
  if (!s) return;
  batch_state_free(s->b);
  for (SPinned &p : s->pin) {
    if (p.jobs) (void)hipHostFree(p.jobs);
    if (p.rs) (void)hipHostFree(p.rs);
    if (p.cnt) (void)hipHostFree(p.cnt);
    if (p.hash) (void)hipHostFree(p.hash);
    if (p.big) (void)hipHostFree(p.big);
    if (p.sols) (void)hipHostFree(p.sols);
    if (p.done) (void)hipEventDestroy(p.done);
  }
  delete s;
}

extern "C" int ap26_server_open(ap26_ctx *ctx, int32_t max_units, int64_t mem_budget_bytes, ap26_server **out) {
  if (!ctx || !out || max_units < 1 || max_units > (1 << 16)) return AP26_E_RANGE;
  *out = nullptr;
  HIPCHK(hipSetDevice(ctx->device));
  ap26_server *s = new (std::nothrow) ap26_server();
  if (!s) return AP26_E_NOMEM;
  s->ctx = ctx;
  s->max_units = max_units;
  int64_t budget = mem_budget_bytes > 0 ? mem_budget_bytes : (int64_t)(ctx->global_mem / 8);
  const int64_t per_k = (int64_t)NTAB * AP26_NSHIFT * 8 + (int64_t)AP26_NN43 * 8 + (INT64_C(48) << 20) * 12;
  int R = (int)(budget / per_k);
  if (R < 1) {
    delete s;
    return AP26_E_NOMEM; /* not even one K fits the budget */
  }
  if (R > 64) R = 64;
  s->R = R;
  int rc = batch_state_new(ctx, R, 2, &s->b);
  if (!rc) rc = ensure_capacity(s->b, R, max_units);
  for (int i = 0; i < 2 && !rc; i++) {
    SPinned &p = s->pin[i];
    if (hipHostMalloc((void **)&p.jobs, sizeof(BatchJobDev) * R, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void **)&p.rs, sizeof(RoundState), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void **)&p.cnt, sizeof(unsigned long long) * 8, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void **)&p.hash, sizeof(unsigned long long) * max_units, hipHostMallocDefault) !=
            hipSuccess ||
        hipHostMalloc((void **)&p.big, sizeof(unsigned long long) * max_units, hipHostMallocDefault) !=
            hipSuccess ||
        hipHostMalloc((void **)&p.sols, sizeof(DevSol) * SERVER_SOL_COPY, hipHostMallocDefault) != hipSuccess)
      rc = AP26_E_NOMEM;
    else if (hipEventCreateWithFlags(&p.done, hipEventDisableTiming) != hipSuccess)
      rc = AP26_E_DEVICE;
  }
  if (!rc) rc = ensure_graph(ctx, s->b, 0);
  if (!rc) rc = ensure_graph(ctx, s->b, 1);
  if (rc) {
    server_free(s);
    return rc;
  }
  for (int i = max_units - 1; i >= 0; i--) s->free_slots.push_back(i);
  *out = s;
  return AP26_OK;
}

extern "C" int ap26_server_close(ap26_server *s) {
  if (!s) return AP26_OK;
  ap26_ctx *ctx = s->ctx;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  server_free(s);
  return AP26_OK;
}

extern "C" int ap26_server_submit(ap26_server *s, const ap26_range *unit, int64_t *ticket) {
  if (!s || !unit || !ticket) return AP26_E_RANGE;
  if (!unit_ok(*unit)) return AP26_E_RANGE;
  if (s->free_slots.empty()) return AP26_E_BUSY;
  SUnit u;
  u.r = *unit;
  for (int64_t K = unit->kmin; K <= unit->kmax; K++)
    if (host_will_search((int32_t)K)) u.ks.push_back((int32_t)K);
  u.t_submit = now_ms();
  const int64_t t = s->next_ticket++;
  if (u.ks.empty() || (u.ks.size() == 1 && unit->iter_start >= AP26_NSHIFT)) { /* nothing to search */
    u.done = true;
    u.t_done = u.t_submit;
    s->units_done++;
  } else {
    u.slot = s->free_slots.back();
    s->free_slots.pop_back();
    s->active.push_back(t);
  }
  s->units.emplace(t, std::move(u));
  *ticket = t;
  return AP26_OK;
}

/* Enqueue round r on its io: clear the io's outputs, upload the jobs, run
   the round's graphs and the report verification, copy the outputs to the
   io's pinned buffers and record the io's event.  Nothing waits. */
static int server_launch(ap26_server *s, const SRound &r) {
  ap26_ctx *ctx = s->ctx;
  BatchState *b = s->b;
  const RoundIO &o = b->io[r.io];
  SPinned &p = s->pin[r.io];
  const int n = (int)r.jobs.size();
  int rc = ensure_surv_cap(ctx, b, r.slots);
  if (!rc) rc = ensure_graph(ctx, b, r.io);
  if (rc) return rc;
  std::memcpy(p.jobs, r.jobs.data(), sizeof(BatchJobDev) * n);
  *p.rs = RoundState{0, n, s->R, 0};
  const int nz = s->max_units > 8 ? s->max_units : 8;
  hipLaunchKernelGGL(server_zero_kernel, dim3((nz + 255) / 256), dim3(256), 0, ctx->stream, o.d_cnt, o.d_hash, o.d_big,
                     s->max_units);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(o.d_jobs, p.jobs, sizeof(BatchJobDev) * n, hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(hipMemcpyAsync(o.d_rs, p.rs, sizeof(RoundState), hipMemcpyHostToDevice, ctx->stream));
  rc = launch_round(ctx, b, r.io, nullptr, nullptr);
  if (rc) return rc;
  hipLaunchKernelGGL(verify_kernel, dim3(256), dim3(256), 0, ctx->stream, o.d_sols,
                     (const unsigned int *)(o.d_cnt + 2), BATCH_SOL_CAP);
  hipLaunchKernelGGL(verify_finish_kernel, dim3(16), dim3(256), 0, ctx->stream, o.d_sols,
                     (const unsigned int *)(o.d_cnt + 2), BATCH_SOL_CAP);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(p.cnt, o.d_cnt, sizeof(unsigned long long) * 8, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipMemcpyAsync(p.hash, o.d_hash, sizeof(unsigned long long) * s->max_units, hipMemcpyDeviceToHost,
                        ctx->stream));
  HIPCHK(hipMemcpyAsync(p.big, o.d_big, sizeof(unsigned long long) * s->max_units, hipMemcpyDeviceToHost,
                        ctx->stream));
  HIPCHK(hipMemcpyAsync(p.sols, o.d_sols, sizeof(DevSol) * SERVER_SOL_COPY, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipEventRecord(p.done, ctx->stream));
  return AP26_OK;
}

/* Build the next round (one K from each active unit in turn until R jobs or
   the survivor slots are taken; a K divisible by sieve primes takes more
   slots, and one that needs more than the buffers hold alone grows them) and
   launch it.  With require_full, only a round that is full (R jobs, or the
   survivor slots taken) is launched.  Returns 1 if a round was launched, 0
   if not, or an error, in which case the units' scheduling is left as it
   was. */
static int server_enqueue(ap26_server *s, bool require_full) {
  BatchState *b = s->b;
  SRound r;
  r.io = s->next_io;
  r.slots = b->surv_slack;
  std::map<int64_t, size_t> taken; /* K taken per unit, in order from its next */
  bool progress = true, full = false;
  while ((int)r.jobs.size() < s->R && progress && !full) {
    progress = false;
    for (size_t c = 0; c < s->active.size() && (int)r.jobs.size() < s->R; c++) {
      const int64_t t = s->active[(s->rr + c) % s->active.size()];
      SUnit &u = s->units[t];
      const size_t at = u.next + taken[t];
      if (at >= u.ks.size()) continue;
      const int32_t K = u.ks[at];
      const int64_t need = host_surv_budget(K);
      if (!r.jobs.empty() && r.slots + need > b->surv_cap) {
        full = true;
        break;
      }
      r.slots += need;
      const int it0 = (at == 0 && K == u.r.kmin) ? u.r.iter_start : 0;
      r.jobs.push_back(make_job(K, u.r.shift, it0, u.slot));
      r.owner.push_back(t);
      taken[t]++;
      progress = true;
    }
  }
  if (r.jobs.empty() || (require_full && !full && (int)r.jobs.size() < s->R)) return 0;
  s->rr = s->active.empty() ? 0 : (s->rr + 1) % s->active.size();
  const int rc = server_launch(s, r);
  if (rc) return rc;
  for (const auto &kv : taken) {
    SUnit &u = s->units[kv.first];
    u.next += kv.second;
    u.outstanding += (int)kv.second;
  }
  s->inflight.push_back(std::move(r));
  s->next_io ^= 1;
  return 1;
}

static void server_finish_unit(ap26_server *s, int64_t t, SUnit &u, double when) {
  u.done = true;
  u.t_done = when;
  s->free_slots.push_back(u.slot);
  u.slot = -1;
  s->units_done++;
  for (size_t i = 0; i < s->active.size(); i++)
    if (s->active[i] == t) {
      s->active.erase(s->active.begin() + (long)i);
      break;
    }
  if (s->rr >= s->active.size()) s->rr = 0;
}

/* Wait for the oldest round in flight and hand its results to its units. */
static int server_harvest(ap26_server *s, int32_t *completed) {
  ap26_ctx *ctx = s->ctx;
  BatchState *b = s->b;
  SRound &r = s->inflight.front();
  SPinned &p = s->pin[r.io];
  HIPCHK(hipEventSynchronize(p.done));
  if (p.cnt[6] != 0) return AP26_E_INTERNAL; /* a queue step ran on part of a wave */
  if (p.cnt[4] != 0) {
    /* survivors beyond the buffer: grow it to the count seen and run the
       round again, after the other round in flight (if any) */
    if (r.rerun) return AP26_E_INTERNAL;
    r.rerun = true;
    const int64_t need = (int64_t)(p.cnt[4] + (p.cnt[4] >> 3)) + b->surv_slack;
    r.slots = need > r.slots ? need : r.slots;
    int rc = server_launch(s, r);
    if (rc) return rc;
    HIPCHK(hipEventSynchronize(p.done));
    if (p.cnt[6] != 0 || p.cnt[4] != 0) return AP26_E_INTERNAL;
  }
  const unsigned int nsol = (unsigned int)(p.cnt[2] & 0xffffffffu);
  if (nsol > (unsigned int)BATCH_SOL_CAP) return AP26_E_INTERNAL;
  std::vector<DevSol> extra;
  const DevSol *sols = p.sols;
  if (nsol > (unsigned int)SERVER_SOL_COPY) { /* rare: copy the rest, after the round behind it */
    extra.resize(nsol);
    HIPCHK(hipMemcpyAsync(extra.data(), b->io[r.io].d_sols, sizeof(DevSol) * nsol, hipMemcpyDeviceToHost,
                          ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    sols = extra.data();
  }
  std::map<int, int64_t> slot_owner;
  std::map<int64_t, int> per_unit;
  for (int64_t t : r.owner) {
    slot_owner[s->units[t].slot] = t;
    per_unit[t]++;
  }
  for (unsigned int i = 0; i < nsol; i++) {
    const DevSol &d = sols[i];
    auto it = slot_owner.find(d.range_id);
    if (it == slot_owner.end()) return AP26_E_INTERNAL;
    SUnit &u = s->units[it->second];
    if (u.status == AP26_OK) u.sols.push_back(ap26_sol{d.k, d.K, d.first, -1, d.shift_iter, d.prp2, d.valid});
  }
  const double when = now_ms();
  for (const auto &kv : per_unit) {
    SUnit &u = s->units[kv.first];
    u.outstanding -= kv.second;
    u.hash += (int64_t)p.hash[u.slot];
    if (p.big[u.slot] && u.status == AP26_OK) {
      /* the walk reached 2^62, where the reference's PrimeQ exits: the
         unit fails with AP26_E_RANGE and is not scheduled further */
      u.status = AP26_E_RANGE;
      u.next = u.ks.size();
      u.sols.clear();
    }
    if (u.outstanding == 0 && u.next >= u.ks.size() && !u.done) {
      server_finish_unit(s, kv.first, u, when);
      if (completed) (*completed)++;
    }
  }
  s->rounds++;
  s->k_done += (int64_t)r.jobs.size();
  s->inflight.pop_front();
  return AP26_OK;
}

extern "C" int ap26_server_pump(ap26_server *s, int32_t *completed) {
  if (!s) return AP26_E_RANGE;
  ap26_ctx *ctx = s->ctx;
  if (completed) *completed = 0;
  HIPCHK(hipSetDevice(ctx->device));
  while (s->inflight.size() < 2) {
    const int rc = server_enqueue(s, !s->inflight.empty());
    if (rc < 0) return rc;
    if (rc == 0) break;
  }
  if (s->inflight.empty()) return AP26_OK;
  return server_harvest(s, completed);
}

extern "C" int ap26_server_result(ap26_server *s, int64_t ticket, ap26_sol *out, int32_t cap, int32_t *n_out,
                                  int64_t *hash_total, double *latency_ms, int32_t *done) {
  if (!s || !done) return AP26_E_RANGE;
  auto it = s->units.find(ticket);
  if (it == s->units.end()) return AP26_E_RANGE;
  SUnit &u = it->second;
  *done = u.done ? 1 : 0;
  if (!u.done) return AP26_OK;
  if (latency_ms) *latency_ms = u.t_done - u.t_submit;
  if (u.status != AP26_OK) { /* collected: the unit failed */
    const int st = u.status;
    if (n_out) *n_out = 0;
    if (hash_total) *hash_total = 0;
    s->units.erase(it);
    return st;
  }
  const int n = (int)u.sols.size();
  if (n_out) *n_out = n;
  if (hash_total) *hash_total = u.hash;
  if (n > cap) return AP26_E_OVERFLOW; /* kept: call again with a larger buffer */
  for (int i = 0; i < n; i++) out[i] = u.sols[i];
  s->units.erase(it);
  return AP26_OK;
}

extern "C" int ap26_server_progress(ap26_server *s, int64_t ticket, double *fraction, int64_t *k_done,
                                    int64_t *k_total) {
  if (!s) return AP26_E_RANGE;
  auto it = s->units.find(ticket);
  if (it == s->units.end()) return AP26_E_RANGE;
  const SUnit &u = it->second;
  const int64_t total = (int64_t)u.ks.size();
  const int64_t done = u.done ? total : (int64_t)u.next - u.outstanding;
  if (k_done) *k_done = done;
  if (k_total) *k_total = total;
  if (fraction) *fraction = total > 0 ? (double)done / (double)total : 1.0;
  return AP26_OK;
}

extern "C" int ap26_server_stats(ap26_server *s, int64_t *units_done, int64_t *k_done, int64_t *rounds,
                                 int32_t *k_per_round, int32_t *active) {
  if (!s) return AP26_E_RANGE;
  if (units_done) *units_done = s->units_done;
  if (k_done) *k_done = s->k_done;
  if (rounds) *rounds = s->rounds;
  if (k_per_round) *k_per_round = s->R;
  if (active) *active = (int32_t)s->active.size();
  return AP26_OK;
}

/* ---- autotuner ------------------------------------------------------------ */
/* Times the batched engine's sieve variants (LDS primes, waves per CU,
   compaction levels, one shift or a shift pair per item) and work-item sizes
   (n43 starts per item) on one K, keeps the fastest for this context and, if
   a path is given, writes it to a cache file that ap26_open loads again when
   AP26_TUNE_CACHE names it. */

static const int TUNE_VARIANTS[] = {117, 126, 119, 95, 121, 98, 101, 85, 100, 70, 75, 28};
static const int TUNE_CHUNKS[] = {11, 21, 43, 85};

extern "C" int ap26_autotune(ap26_ctx *ctx, int32_t K, int32_t SHIFT, int32_t quick, const char *cache_path,
                             ap26_tune_entry *results, int32_t cap, int32_t *n_out, ap26_tune_entry *best) {
  if (!ctx || K < 1 || K > 600000000 || !host_will_search(K) || !host_walk_ok(K, SHIFT)) return AP26_E_RANGE;
  HIPCHK(hipSetDevice(ctx->device));
  const int nv = quick ? 5 : (int)(sizeof(TUNE_VARIANTS) / sizeof(int));
  const int nc = quick ? 2 : (int)(sizeof(TUNE_CHUNKS) / sizeof(int));
  const int nd = (int)(sizeof(WALK_DEPTHS) / sizeof(int));
  const int save_v = ctx->tune_variant, save_c = ctx->tune_chunk, save_d = ctx->tune_walk_depth;
  BatchState *save_b = ctx->batch;
  ap26_tune_entry top{-1, -1, 1, 0, 1e30, 1e30};
  int n = 0, rc = AP26_OK;
  ap26_range r{K, K, SHIFT, 0};
  ap26_sol *sols = (ap26_sol *)malloc(sizeof(ap26_sol) * 4096);
  if (!sols) return AP26_E_NOMEM;
  /* one candidate: a fresh batched state built with the context's choice,
     one warm call (graph capture), one timed call */
  auto time_one = [&](ap26_tune_entry &t) {
    BatchState *tmp = nullptr;
    int e = batch_state_new(ctx, 1, 1, &tmp);
    if (e) return e;
    ctx->batch = tmp;
    int32_t ns = 0;
    int64_t tot = 0;
    ap26_stats st{};
    e = ap26_search_batch(ctx, &r, 1, sols, 4096, &ns, &tot, &st);
    if (!e) e = ap26_search_batch(ctx, &r, 1, sols, 4096, &ns, &tot, &st);
    ctx->batch = save_b;
    batch_state_free(tmp);
    t = ap26_tune_entry{ctx->tune_variant, ctx->tune_chunk, walk_depth_of(ctx), 0, st.ms_walk,
                        st.ms_walk - st.ms_sieve};
    return e;
  };
  /* stage 1: sieve variant x work-item size, at walk depth 1 */
  ctx->tune_walk_depth = 1;
  for (int vi = 0; vi < nv && !rc; vi++)
    for (int ci = 0; ci < nc && !rc; ci++) {
      ctx->tune_variant = TUNE_VARIANTS[vi];
      ctx->tune_chunk = TUNE_CHUNKS[ci];
      ap26_tune_entry t{};
      if ((rc = time_one(t))) break;
      if (results && n < cap) results[n] = t;
      n++;
      if (t.ms_per_k < top.ms_per_k) top = t;
    }
  /* stage 2: the walk's compaction depth with that sieve, by the walk's own
     time (the device time after the sieve graph) */
  if (!rc) {
    ctx->tune_variant = top.variant;
    ctx->tune_chunk = top.chunk;
    double best_walk = 1e30;
    for (int di = 0; di < nd && !rc; di++) {
      ctx->tune_walk_depth = WALK_DEPTHS[di];
      ap26_tune_entry t{};
      if ((rc = time_one(t))) break;
      if (results && n < cap) results[n] = t;
      n++;
      if (t.ms_walk < best_walk) {
        best_walk = t.ms_walk;
        top = t;
      }
    }
  }
  free(sols);
  ctx->batch = save_b;
  if (rc) {
    ctx->tune_variant = save_v;
    ctx->tune_chunk = save_c;
    ctx->tune_walk_depth = save_d;
    return rc;
  }
  ctx->tune_variant = top.variant;
  ctx->tune_chunk = top.chunk;
  ctx->tune_walk_depth = top.walk_depth;
  batch_free(ctx); /* the next batch is built with the tuned choice */
  if (n_out) *n_out = n;
  if (best) *best = top;
  if (cache_path && cache_path[0]) {
    FILE *f = fopen(cache_path, "w");
    if (!f) return AP26_E_RANGE;
    fprintf(f, "ap26-tune 2 %s %d %d %d %.3f\n", ctx->name, top.variant, top.chunk, top.walk_depth, top.ms_per_k);
    fclose(f);
  }
  return AP26_OK;
}

/* called by ap26_open: adopt a cached choice for this device */
void tune_load_cache(ap26_ctx *ctx) {
  const char *path = getenv("AP26_TUNE_CACHE");
  if (!path || !path[0]) return;
  FILE *f = fopen(path, "r");
  if (!f) return;
  char arch[256];
  int ver = 0, v = -1, c = -1, d = 1;
  double ms = 0;
  /* version 1: variant, chunk; version 2 adds the walk depth */
  if (fscanf(f, "ap26-tune %d %255s %d %d", &ver, arch, &v, &c) == 4 && (ver == 1 || ver == 2) &&
      (ver == 1 || fscanf(f, "%d", &d) == 1) && fscanf(f, "%lf", &ms) == 1 && strcmp(arch, ctx->name) == 0 &&
      c > 0 && c <= AP26_NN43) {
    ctx->tune_variant = v;
    ctx->tune_chunk = c;
    ctx->tune_walk_depth = d;
  }
  fclose(f);
}
