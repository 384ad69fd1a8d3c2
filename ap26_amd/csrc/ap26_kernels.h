/* ap26_kernels.h -- the AP26 kernels for gfx950: sieve tables, LDS-staged
 * sieve, survivor walk, batched SPRP.
 *
 * Data layout in HBM (per K being searched):
 *   tab   : for each pass of S shifts, NTAB*S uint64 words laid out
 *           [prime table][centred residue][shift in pass]; the first NOK*S
 *           words are the 83 sieve primes, the last 90*S the 7..23 filters.
 *   n43   : int64[10840] starts of the K.
 *   surv  : int64 survivor list (n values), appended per wave.
 *   sols  : reports (k >= 10).
 */
#ifndef AP26_KERNELS_H
#define AP26_KERNELS_H

#include <utility>

#include "ap26_device.h"

namespace ap26 {

/* level-0 primes whose masks can come from registers (host_vwin) */
constexpr int NVW = 3;
constexpr int VW_P[NVW] = {61, 67, 71};

struct KParams {
  int64_t STEP, n0, S31, S37, S41, S43, S47, S53, S59;
  int32_t K, SHIFT;
  uint64_t vw[NVW][4]; /* window strings of 61, 67, 71 (host_vwin) */
};

struct DevSol {
  int32_t k, K;
  int64_t first;
  int32_t range_id, shift_iter;
  uint32_t prp2, valid; /* bit j: term j is a base-2 SPRP / passes the 9 bases */
};

template <int B, int E, class F>
__device__ __forceinline__ void static_for(F &&f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    static_for<B + 1, E>(f);
  }
}

/* Number of set bits of the wave mask bal in the lanes below this one
   (v_mbcnt_lo / v_mbcnt_hi): a lane's slot among the lanes that append. */
__device__ __forceinline__ int lanes_below(uint64_t bal) {
  return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(bal >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0u));
}

/* ---- table build ------------------------------------------------------- */
/* OK_p[v] is 0 iff some term n + i*STEP (i in 0..23) is divisible by p when
 * n == v (mod p) (src/cpu/cpusse2.cpp:31-35): v == -i*STEP.  With sp = STEP mod p
 * != 0 that is ((p - v) * sp^-1 mod p) <= 23; with sp == 0 it is v == 0.
 * OKOK_p[r] bit jj = OK_p[(r + (jj+SHIFT)*MOD) mod p] (src/cpu/cpusse2.cpp:37-43).
 * For the filter primes q = 7..23 bit jj = ((r + (jj+SHIFT)*MOD) mod q != 0),
 * the reference's per-candidate n%7..n%23 test (src/cpu/cpusse2.cpp:572-577).
 * Grid: x = centred residue chunk (64), y = table index, z = shift slot. */
__device__ __forceinline__ uint64_t table_word(const KParams &kp, int ti, int c, int shift) {
  const int p = PT.p[ti];
  const int r = (c + (p + 1) / 2) % p; /* centred entry c holds residue c-(p-1)/2 */
  const int mp = (int)(AP26_MOD % p);
  const bool sieve = ti < NSP;
  int sp = 0, inv = 0;
  if (sieve) {
    sp = (int)(((int64_t)(kp.K % p) * (AP26_PRIM23 % p)) % p);
    if (sp) { /* sp^(p-2) mod p */
      int e = p - 2, b = sp;
      inv = 1;
      while (e) {
        if (e & 1) inv = inv * b % p;
        b = b * b % p;
        e >>= 1;
      }
    }
  }
  uint64_t w = 0;
  int js = shift % p; /* (jj + shift) mod p, advanced per bit */
  for (int jj = 0; jj < 64; jj++) {
    int v = (r + js * mp) % p;
    bool ok;
    if (!sieve || sp == 0) ok = v != 0;
    else ok = ((p - v) * inv) % p > 23;
    w |= (uint64_t)ok << jj;
    if (++js == p) js = 0;
  }
  return w;
}

static __global__ void tables_kernel(KParams kp, int32_t it0, int32_t it1, int32_t S,
                              uint64_t *__restrict__ tab) {
  const int ti = blockIdx.y;
  const int slot = blockIdx.z;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  const int p = PT.p[ti];
  if (c >= p) return;
  const int pass = slot / S, s = slot % S;
  const int it = it0 + slot;
  uint64_t *dst = tab + (size_t)pass * NTAB * S + (size_t)(PT.off[ti] + c) * S + s;
  *dst = it >= it1 ? 0 : table_word(kp, ti, c, kp.SHIFT + 64 * it);
}

/* ---- sieve ------------------------------------------------------------- */

struct SieveArgs {
  KParams kp;
  const int64_t *__restrict__ n43;  /* [10840] */
  const uint64_t *__restrict__ tab; /* this pass: NTAB*S words */
  int32_t shift0;                   /* SHIFT of the pass's first slot */
  int32_t nslot;                    /* live slots in this pass (<= S) */
  int64_t blk0, blk1;               /* n53 blocks [blk0, blk1) */
  int64_t *__restrict__ surv;
  unsigned long long *__restrict__ nsurv;
  int64_t cap;
  int32_t *__restrict__ surv_job; /* batched engine: job index of each survivor */
  int32_t job;
  unsigned int *__restrict__ err;   /* set if a queue step ran on part of a wave */
  int32_t vc[NVW];                  /* (shift0 mod p) + p*((p+1)/2), p = 61, 67, 71 */
  void *__restrict__ gq;            /* QEntry regions, one per wave of the grid (variants with GQ) */
  unsigned int *__restrict__ slab;  /* plain path: blocks of [blk0, blk1) claimed so far (zero at launch) */
};

__host__ __device__ constexpr int32_t vw_offset(int32_t shift0, int p) { return shift0 % p + p * ((p + 1) / 2); }

/* n53 = (n43[j] + i43*S43 + i47*S47 + i53*S53) mod MOD for block
   b = j*12673 + i43*667 + i47*29 + i53, i.e. the reference's nested
   "+= S mod MOD" loops (src/cpu/cpusse2.cpp:397-407, setupn.cl:396-418). */
__device__ __forceinline__ int64_t n53_of_block(const KParams &kp, const int64_t *n43, int64_t b) {
  int j = (int)(b / AP26_N53_PER_N43);
  int rem = (int)(b - (int64_t)j * AP26_N53_PER_N43);
  int i43 = rem / 667;
  rem -= i43 * 667;
  int i47 = rem / 29;
  int i53 = rem - i47 * 29;
  int64_t v = n43[j] + i43 * kp.S43 + i47 * kp.S47 + i53 * kp.S53; /* < 70*MOD < 2^56 */
  int64_t q = (int64_t)((double)v * (1.0 / (double)AP26_MOD));
  v -= q * AP26_MOD;
  if (v < 0) v += AP26_MOD;
  if (v >= AP26_MOD) v -= AP26_MOD;
  return v;
}

/* Early-exit points: after the first 15 primes (61..131), then every 4. */
__host__ __device__ constexpr bool check_after(int i) { return i == 14 || (i > 14 && ((i - 14) % 4 == 0)); }

/* One thread walks the 35 positions n59 = n53 + i59*S59 (mod MOD) of an n53
 * block and, for each, ANDs the 64-bit masks of S shifts over the 83 sieve
 * primes (LDS for the first NL, the rest from L2), with early exit; survivors
 * of the 7..23 filter are appended per wave with one atomic.  The reference's
 * sieve (src/opencl/kernels/sieve.cl:28-152, src/cpu/cpusse2.cpp:410-614). */
template <int S, int NL>
static __global__ void __launch_bounds__(256) sieve_kernel(SieveArgs a) {
  constexpr int LMAIN = PT.off[NL] * S;
  constexpr int LFILT = 90 * S;
  __shared__ uint64_t lds[LMAIN + LFILT];
  for (int i = threadIdx.x; i < LMAIN; i += blockDim.x) lds[i] = a.tab[i];
  for (int i = threadIdx.x; i < LFILT; i += blockDim.x) lds[LMAIN + i] = a.tab[NOK * S + i];
  __syncthreads();

  const uint64_t *__restrict__ gtab = a.tab;
  const int lane = __lane_id();
  const double S59d = (double)a.kp.S59;
  const double MODd = (double)AP26_MOD;
  const double WRAP = MODd + TWO52;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;

  constexpr int SCALE = 8 * S; /* bytes per centred entry */
  const char *ldsb = (const char *)lds;
  const char *gtb = (const char *)gtab;

  for (int64_t b = a.blk0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; b < a.blk1; b += stride) {
    double xd = (double)n53_of_block(a.kp, a.n43, b) + TWO52;
    for (int i59 = 0; i59 < AP26_N59_STEPS; i59++) {
      const double x = xd - TWO52;
      const uint32_t xs = lo32(xd) * (uint32_t)SCALE;
      uint64_t m[S];
#pragma unroll
      for (int s = 0; s < S; s++) m[s] = ~(uint64_t)0;
      bool alive = true;
      static_for<0, NSP>([&](auto IC) {
        constexpr int I = decltype(IC)::value;
        constexpr int P = PT.p[I];
        if (alive) {
          const int bo = centred_byteoff<P, SCALE>(x, xs);
          constexpr int C0 = SCALE * (PT.off[I] + (P - 1) / 2); /* entry of residue 0 */
          if constexpr (I < NL) {
#pragma unroll
            for (int s = 0; s < S; s++) m[s] &= *(const uint64_t *)(ldsb + C0 + bo + 8 * s);
          } else {
#pragma unroll
            for (int s = 0; s < S; s++) m[s] &= *(const uint64_t *)(gtb + (uint32_t)(C0 + bo + 8 * s));
          }
          if constexpr (check_after(I) || I == NSP - 1) {
            uint64_t any = 0;
#pragma unroll
            for (int s = 0; s < S; s++) any |= m[s];
            alive = any != 0;
          }
        }
      });
      if (alive) {
        static_for<NSP, NPT>([&](auto IC) {
          constexpr int I = decltype(IC)::value;
          constexpr int P = PT.p[I];
          const int bo = centred_byteoff<P, SCALE>(x, xs);
          constexpr int C0 = 8 * LMAIN + SCALE * (PT.off[I] - NOK + (P - 1) / 2);
#pragma unroll
          for (int s = 0; s < S; s++) m[s] &= *(const uint64_t *)(ldsb + C0 + bo + 8 * s);
        });
        const int64_t n59 = (int64_t)(__double_as_longlong(xd) & ((INT64_C(1) << 52) - 1));
#pragma unroll
        for (int s = 0; s < S; s++) {
          uint64_t w = m[s];
          const int64_t nbase = n59 + (int64_t)(a.shift0 + 64 * s) * AP26_MOD;
          while (true) {
            const bool has = w != 0;
            const uint64_t bal = __ballot(has);
            if (bal == 0) break;
            const int cnt = __popcll(bal);
            const int leader = __ffsll((unsigned long long)bal) - 1;
            const int pre = __popcll(bal & ((UINT64_C(1) << lane) - 1));
            unsigned long long base = 0;
            if (lane == leader) base = atomicAdd(a.nsurv, (unsigned long long)cnt);
            base = __shfl(base, leader);
            if (has) {
              const int bit = __ffsll((unsigned long long)w) - 1;
              w &= w - 1;
              const unsigned long long idx = base + (unsigned long long)pre;
              if (idx < (unsigned long long)a.cap) a.surv[idx] = nbase + (int64_t)bit * AP26_MOD;
            }
          }
        }
      }
      xd += S59d;
      if (xd >= WRAP) xd -= MODd;
    }
  }
}

/* ---- compacted sieve --------------------------------------------------- */
/* The mask of a position dies after ~18 primes on average but the tail is
 * long, so lanes that evaluate positions side by side diverge badly.  Here the
 * 83 primes are cut into levels; every lane runs level 0 for its own position,
 * and the positions whose mask is still non-zero are appended (ballot + mbcnt)
 * to a per-wave LDS queue.  Whenever a queue holds 64 entries the wave pops a
 * full wave of them and runs the next level on dense lanes, appending its
 * survivors to the next queue; the last level also applies the 7..23 filter
 * and writes the survivors out.  Levels have no internal exits.  Tables of the
 * first NL primes live in LDS, the rest are read through L1/L2. */

/* Level sets: prime-index boundaries of the levels. */
template <int... Bs>
struct Levels {
  static constexpr int N = sizeof...(Bs) - 1; /* levels; N-1 queues */
  static constexpr int B[sizeof...(Bs)] = {Bs...};
};
using Levels4 = Levels<0, 15, 29, 45, 83>;     /* 61..131 | 137..199 | 211..307 | 311..541 */
using Levels3 = Levels<0, 15, 29, 83>;
using Levels5 = Levels<0, 12, 20, 30, 45, 83>;

struct QEntry {
  double xd;  /* n59 + 2^52 */
  uint64_t m; /* mask so far */
};

/* Per-lane copies of the level-0 constants held in VGPRs: with ~45 scalar
   constants live the compiler otherwise recycles one SGPR pair for all of them
   and serialises the lookups. */
template <int B, int E>
struct VConsts {
  double inv[E - B];
  int m8p[E - B];
  __device__ __forceinline__ void init() {
    static_for<B, E>([&](auto IC) {
      constexpr int I = decltype(IC)::value;
      constexpr double c = 1.0 / (double)PT.p[I];
      asm("v_mov_b64 %0, %1" : "=v"(inv[I - B]) : "s"(c));
      asm("v_mov_b32 %0, %1" : "=v"(m8p[I - B]) : "s"(-8 * PT.p[I]));
    });
  }
};

template <int NL, int B, int E, unsigned G = 0>
__device__ __forceinline__ uint64_t apply_primes(uint64_t m, double xd, const char *ldsb, const char *gtb) {
  const double x = xd - TWO52;
  const uint32_t xs = lo32(xd) * 8u;
  int bo[E - B];
  uint64_t w[E - B];
  static_for<B, E>([&](auto IC) {
    constexpr int I = decltype(IC)::value;
    bo[I - B] = centred_byteoff<PT.p[I], 8>(x, xs);
  });
  static_for<B, E>([&](auto IC) {
    constexpr int I = decltype(IC)::value;
    constexpr int P = PT.p[I];
    constexpr int C0 = 8 * (PT.off[I] + (P - 1) / 2);
    if constexpr (I < NL && !(I < 32 && ((G >> I) & 1)))
      w[I - B] = *(const uint64_t *)(ldsb + C0 + bo[I - B]);
    else /* C0 + bo >= 0: an unsigned offset from the wave-uniform base */
      w[I - B] = *(const uint64_t *)(gtb + (uint32_t)(C0 + bo[I - B]));
  });
  static_for<B, E>([&](auto IC) { m &= w[decltype(IC)::value - B]; });
  return m;
}

/* same, with the constants in VGPRs (all primes in LDS); offsets, loads and
   ANDs are issued as three batches so the LDS reads overlap */
template <int B, int E>
__device__ __forceinline__ uint64_t apply_primes_vc(uint64_t m, double xd, const char *ldsb,
                                                    const VConsts<B, E> &vc) {
  const double x = xd - TWO52;
  const uint32_t xs = lo32(xd) * 8u;
  int bo[E - B];
  uint64_t w[E - B];
  static_for<B, E>([&](auto IC) {
    constexpr int I = decltype(IC)::value;
    const double qd = __fma_rn(x, vc.inv[I - B], SHIFTER);
    int r;
    asm("v_mad_i32_i24 %0, %1, %2, %3" : "=v"(r) : "v"((int)lo32(qd)), "v"(vc.m8p[I - B]), "v"(xs));
    bo[I - B] = (r << 5) >> 5;
  });
  static_for<B, E>([&](auto IC) {
    constexpr int I = decltype(IC)::value;
    constexpr int P = PT.p[I];
    constexpr int C0 = 8 * (PT.off[I] + (P - 1) / 2);
    w[I - B] = *(const uint64_t *)(ldsb + C0 + bo[I - B]);
  });
  static_for<B, E>([&](auto IC) { m &= w[decltype(IC)::value - B]; });
  return m;
}

template <int NL>
__device__ __forceinline__ uint64_t apply_filters(uint64_t m, double xd, const char *ldsb) {
  const double x = xd - TWO52;
  const uint32_t xs = lo32(xd) * 8u;
  constexpr int LMAIN = PT.off[NL];
  static_for<NSP, NPT>([&](auto IC) {
    constexpr int I = decltype(IC)::value;
    constexpr int P = PT.p[I];
    const int bo = centred_byteoff<P, 8>(x, xs);
    constexpr int C0 = 8 * (LMAIN + PT.off[I] - NOK + (P - 1) / 2);
    m &= *(const uint64_t *)(ldsb + C0 + bo);
  });
  return m;
}

/* Append the survivors n = n59 + (bit + shift0)*MOD of this lane's mask; one
   atomic per wave per round. */
__device__ __forceinline__ void emit_survivors(uint64_t w, double xd, const SieveArgs &a, int lane) {
  const int64_t n59 = (int64_t)(__double_as_longlong(xd) & ((INT64_C(1) << 52) - 1));
  const int64_t nbase = n59 + (int64_t)a.shift0 * AP26_MOD;
  while (true) {
    const bool has = w != 0;
    const uint64_t bal = __ballot(has);
    if (bal == 0) break;
    const int cnt = __popcll(bal);
    const int leader = __ffsll((unsigned long long)bal) - 1;
    const int pre = __popcll(bal & ((UINT64_C(1) << lane) - 1));
    unsigned long long base = 0;
    if (lane == leader) base = atomicAdd(a.nsurv, (unsigned long long)cnt);
    base = __shfl(base, leader);
    if (has) {
      const int bit = __ffsll((unsigned long long)w) - 1;
      w &= w - 1;
      const unsigned long long idx = base + (unsigned long long)pre;
      if (idx < (unsigned long long)a.cap) {
        a.surv[idx] = nbase + (int64_t)bit * AP26_MOD;
        if (a.surv_job) a.surv_job[idx] = a.job;
      }
    }
  }
}

template <int NL, class LV>
struct QSieve {
  static constexpr int QCAP = 128;
  static constexpr int NQ = LV::N - 1;
  QEntry (*q)[QCAP];  /* this wave's NQ queues */
  int qn[NQ];
  bool partial; /* a push ran with lanes switched off (must never happen) */
  const char *ldsb, *gtb;
  int lane;
  const SieveArgs *a;

  __device__ __forceinline__ void push(int qi, double xd, uint64_t m, bool pred) {
    const uint64_t bal = __ballot(pred);
    const int pre = __popcll(bal & ((UINT64_C(1) << lane) - 1));
    if (pred) q[qi][qn[qi] + pre] = QEntry{xd, m};
    /* every lane keeps the same count only if every lane runs every push:
       the callers' loops are uniform, and a push on part of the wave is
       reported (the queue would lose entries) */
    qn[qi] += __popcll(bal);
    partial |= __builtin_amdgcn_read_exec() != ~UINT64_C(0);
    __builtin_amdgcn_wave_barrier();
  }

  /* run level L+1 on entries popped from queue L (lanes with pred) */
  template <int L>
  __device__ __forceinline__ void run_level(const QEntry &e, bool pred) {
    uint64_t m = e.m;
    if (pred) m = apply_primes<NL, LV::B[L + 1], LV::B[L + 2]>(m, e.xd, ldsb, gtb);
    if constexpr (L + 2 == LV::N) {
      if (pred && m) m = apply_filters<NL>(m, e.xd, ldsb);
      emit_survivors(pred ? m : 0, e.xd, *a, lane);
    } else {
      push(L + 1, e.xd, m, pred && m != 0);
      drain<L + 1>();
    }
  }

  /* pop full waves from queue L while it holds >= 64 entries */
  template <int L>
  __device__ __forceinline__ void drain() {
    while (qn[L] >= 64) {
      qn[L] -= 64;
      const QEntry e = q[L][qn[L] + lane];
      __builtin_amdgcn_wave_barrier();
      run_level<L>(e, true);
    }
  }

  /* flush the partial queues, lowest level first */
  template <int L>
  __device__ __forceinline__ void flush() {
    if constexpr (L < NQ) {
      drain<L>();
      const int n = qn[L];
      if (n > 0) {
        const bool pred = lane < n;
        QEntry e{TWO52, 0};
        if (pred) e = q[L][lane];
        __builtin_amdgcn_wave_barrier();
        qn[L] = 0;
        run_level<L>(e, pred);
      }
      flush<L + 1>();
    } else {
      if (partial && lane == 0) atomicOr(a->err, 1u);
    }
  }
};

template <int NL, int WAVES, class LV>
constexpr int qsieve_lds_bytes() {
  return (PT.off[NL] + 90) * 8 + WAVES * (LV::N - 1) * 128 * (int)sizeof(QEntry);
}

/* The positions of n53 blocks [blk0, blk1) taken by this thread at tid,
   tid+stride, ...: level 0 on every lane, queues for the rest, then flush. */
template <int NL, class LV, bool VC, unsigned G = 0>
__device__ __forceinline__ void qsieve_blocks(QSieve<NL, LV> &qs, const VConsts<0, LV::B[1]> &vc,
                                              const SieveArgs &a, int64_t tid, int64_t stride) {
  const double S59d = (double)a.kp.S59;
  const double MODd = (double)AP26_MOD;
  const double WRAP = MODd + TWO52;
  /* Every lane runs every step of every stride that holds a block for any
     lane of the wave: the queues need the whole wave, so the loops are
     uniform and lanes past the end work on a valid block and drop it. */
  for (int64_t base = a.blk0; base < a.blk1; base += stride) {
    const int64_t b = base + tid;
    const bool has = b < a.blk1;
    if (__ballot(has) == 0) break;
    double xd = (double)n53_of_block(a.kp, a.n43, has ? b : a.blk0) + TWO52;
    for (int i59 = 0; i59 < AP26_N59_STEPS; i59++) {
      uint64_t m;
      if constexpr (VC)
        m = apply_primes_vc<0, LV::B[1]>(~(uint64_t)0, xd, qs.ldsb, vc);
      else
        m = apply_primes<NL, 0, LV::B[1], G>(~(uint64_t)0, xd, qs.ldsb, qs.gtb);
      m = has ? m : 0;
      qs.push(0, xd, m, m != 0);
      xd += S59d;
      xd = xd >= WRAP ? xd - MODd : xd;
      qs.template drain<0>();
    }
  }
  qs.template flush<0>();
}

template <int NL, int WAVES, class LV>
struct QSieveLds {
  uint64_t tab[PT.off[NL] + 90];
  QEntry queues[WAVES][LV::N - 1][128];
};

/* One launch per (K, shift): the plain path.  VC: level-0 constants in VGPRs;
   G: level-0 primes (bit i = prime index i) read through the vector L1 cache
   instead of LDS, so that the two gather paths work in parallel. */
template <int NL, int WAVES, class LV, bool VC, unsigned G = 0>
static __global__ void __launch_bounds__(64 * WAVES) __attribute__((amdgpu_waves_per_eu(WAVES / 4, WAVES / 4)))
sieve_queue_kernel(SieveArgs a) {
  static_assert(qsieve_lds_bytes<NL, WAVES, LV>() <= 160 * 1024, "LDS budget");
  static_assert(LV::B[1] <= NL, "level 0 must be in LDS");
  constexpr int LMAIN = PT.off[NL];
  __shared__ QSieveLds<NL, WAVES, LV> L;
  for (int i = threadIdx.x; i < LMAIN; i += blockDim.x) L.tab[i] = a.tab[i];
  for (int i = threadIdx.x; i < 90; i += blockDim.x) L.tab[LMAIN + i] = a.tab[NOK + i];
  __syncthreads();

  QSieve<NL, LV> qs;
  qs.lane = threadIdx.x & 63;
  qs.q = L.queues[threadIdx.x >> 6];
  for (int i = 0; i < QSieve<NL, LV>::NQ; i++) qs.qn[i] = 0;
  qs.partial = false;
  qs.ldsb = (const char *)L.tab;
  qs.gtb = (const char *)a.tab;
  qs.a = &a;
  VConsts<0, LV::B[1]> vc;
  if constexpr (VC) vc.init();
  qsieve_blocks<NL, LV, VC, G>(qs, vc, a, (int64_t)blockIdx.x * blockDim.x + threadIdx.x,
                               (int64_t)gridDim.x * blockDim.x);
}

/* ---- batched persistent sieve ------------------------------------------ */

struct BatchJobDev {
  KParams kp;        /* K, SHIFT (the unit's), STEP, n0, S31..S59 */
  int32_t it0;       /* first shift iteration searched */
  int32_t range_id;
};

struct RoundState {
  int32_t base;      /* first job of the current round */
  int32_t njobs;     /* jobs in the whole batch */
  int32_t per_round; /* jobs per round */
  int32_t pad;
};

struct PersistArgs {
  const BatchJobDev *__restrict__ jobs;
  const RoundState *__restrict__ rs;
  const int64_t *__restrict__ n43;   /* per_round x 10840 */
  const uint64_t *__restrict__ tab;  /* per_round x 10 x NTAB */
  int64_t *__restrict__ surv;
  int32_t *__restrict__ surv_job;
  unsigned long long *__restrict__ nsurv;
  int64_t cap;
  unsigned int *__restrict__ work;   /* work-stealing item counter */
  unsigned int *__restrict__ err;    /* queue steps on part of a wave */
  int32_t chunk;                     /* n43 starts per item */
  int32_t nchunk;
  QEntry *__restrict__ gq;           /* per-wave global queue regions (variants with GQ) */
};

/* Work items (job, shift, n43 chunk) of one round, taken from a global atomic
   counter by whichever workgroup is free: long and short ranges, mixed SHIFTs
   and resumed units share one launch.  A workgroup restages its LDS tables
   only when the (job, shift) of its next item differs. */
template <int NL, int WAVES, class LV>
static __global__ void __launch_bounds__(64 * WAVES) __attribute__((amdgpu_waves_per_eu(WAVES / 4, WAVES / 4)))
sieve_persistent_kernel(PersistArgs p) {
  static_assert(qsieve_lds_bytes<NL, WAVES, LV>() <= 160 * 1024, "LDS budget");
  constexpr int LMAIN = PT.off[NL];
  __shared__ QSieveLds<NL, WAVES, LV> L;
  __shared__ unsigned int s_item;
  const RoundState rs = *p.rs;
  const int njobs = min(rs.per_round, rs.njobs - rs.base);
  const unsigned int nitems = (unsigned int)(njobs > 0 ? njobs : 0) * AP26_NSHIFT * p.nchunk;
  int cur = -1;
  VConsts<0, LV::B[1]> vc;
  while (true) {
    __syncthreads();
    if (threadIdx.x == 0) s_item = atomicAdd(p.work, 1u);
    __syncthreads();
    const unsigned int item = __builtin_amdgcn_readfirstlane(s_item); /* uniform: keeps job data scalar */
    if (item >= nitems) break;
    const int jl = (int)(item / (AP26_NSHIFT * p.nchunk));
    const int rem = (int)(item - (unsigned int)jl * AP26_NSHIFT * p.nchunk);
    const int it = rem / p.nchunk;
    const int chunk = rem - it * p.nchunk;
    const BatchJobDev &job = p.jobs[rs.base + jl];
    if (it < job.it0) continue;
    const uint64_t *tab = p.tab + ((size_t)jl * AP26_NSHIFT + it) * NTAB;
    const int key = jl * AP26_NSHIFT + it;
    if (key != cur) {
      for (int i = threadIdx.x; i < LMAIN; i += blockDim.x) L.tab[i] = tab[i];
      for (int i = threadIdx.x; i < 90; i += blockDim.x) L.tab[LMAIN + i] = tab[NOK + i];
      __syncthreads();
      cur = key;
    }
    SieveArgs a;
    a.kp = job.kp;
    a.n43 = p.n43 + (size_t)jl * AP26_NN43;
    a.tab = tab;
    a.shift0 = job.kp.SHIFT + 64 * it;
    a.nslot = 1;
    const int j0 = chunk * p.chunk;
    const int j1 = min(AP26_NN43, j0 + p.chunk);
    a.blk0 = (int64_t)j0 * AP26_N53_PER_N43;
    a.blk1 = (int64_t)j1 * AP26_N53_PER_N43;
    a.surv = p.surv;
    a.nsurv = p.nsurv;
    a.cap = p.cap;
    a.surv_job = p.surv_job;
    a.err = p.err;
    a.job = jl;
    QSieve<NL, LV> qs;
    qs.lane = threadIdx.x & 63;
    qs.q = L.queues[threadIdx.x >> 6];
    for (int i = 0; i < QSieve<NL, LV>::NQ; i++) qs.qn[i] = 0;
    qs.partial = false;
    qs.ldsb = (const char *)L.tab;
    qs.gtb = (const char *)tab;
    qs.a = &a;
    qsieve_blocks<NL, LV, false>(qs, vc, a, threadIdx.x, blockDim.x);
  }
}

/* n43 starts of every job in the round, from the constant list of guarded
   (i31, i37, i41) triples (src/opencl/AP26.h:125-136). */
static __global__ void n43_kernel(const BatchJobDev *__restrict__ jobs, const RoundState *__restrict__ rsp,
                           const uint32_t *__restrict__ triples, int64_t *__restrict__ n43) {
  const RoundState rs = *rsp;
  const int njobs = min(rs.per_round, rs.njobs - rs.base);
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  const int jl = idx / AP26_NN43;
  if (jl >= njobs) return;
  const int j = idx - jl * AP26_NN43;
  const uint32_t t = triples[j >> 3];
  const int i31 = t & 0xff, i37 = (t >> 8) & 0xff, i41 = (t >> 16) & 0xff;
  const int i3 = (j >> 2) & 1, i5 = j & 3;
  const KParams &kp = jobs[rs.base + jl].kp;
  int64_t v = kp.n0 + i3 * AP26_S3 + i5 * AP26_S5 + i31 * kp.S31 + i37 * kp.S37 + i41 * kp.S41;
  int64_t q = (int64_t)((double)v * (1.0 / (double)AP26_MOD));
  v -= q * AP26_MOD;
  if (v < 0) v += AP26_MOD;
  if (v >= AP26_MOD) v -= AP26_MOD;
  n43[(size_t)jl * AP26_NN43 + j] = v;
}

/* The guarded (i31, i37, i41) triples of the n43 enumeration in the
   reference's order (src/opencl/AP26.h:125-136): n43 index j = 8*t + 4*i3
   + i5 for triple t.  Built at compile time for kernels that compute the n43
   starts of one K without a host table. */
struct TripleTable {
  uint32_t t[AP26_NN43 / 8];
  int n;
};

constexpr TripleTable make_triples() {
  TripleTable T{};
  for (int i31 = 0; i31 < 7; i31++)
    for (int i37 = 0; i37 < 13; i37++) {
      if (i37 - i31 > 10 || i31 - i37 > 4) continue;
      for (int i41 = 0; i41 < 17; i41++) {
        if (i41 - i31 > 14 || i41 - i37 > 14 || i31 - i41 > 4 || i37 - i41 > 10) continue;
        if (T.n < AP26_NN43 / 8) T.t[T.n] = (uint32_t)i31 | ((uint32_t)i37 << 8) | ((uint32_t)i41 << 16);
        T.n++;
      }
    }
  return T;
}

constexpr TripleTable TRIPLES = make_triples();
static_assert(TRIPLES.n * 8 == AP26_NN43, "n43 starts per K");

/* n43 starts [j0, j1) of one K into n43[j] (as n43_kernel does per job) */
static __global__ void n43_span_kernel(KParams kp, int32_t j0, int32_t j1, int64_t *__restrict__ n43) {
  const int j = j0 + (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (j >= j1) return;
  const uint32_t t = TRIPLES.t[j >> 3];
  const int i31 = t & 0xff, i37 = (t >> 8) & 0xff, i41 = (t >> 16) & 0xff;
  const int i3 = (j >> 2) & 1, i5 = j & 3;
  int64_t v = kp.n0 + i3 * AP26_S3 + i5 * AP26_S5 + i31 * kp.S31 + i37 * kp.S37 + i41 * kp.S41;
  int64_t q = (int64_t)((double)v * (1.0 / (double)AP26_MOD));
  v -= q * AP26_MOD;
  if (v < 0) v += AP26_MOD;
  if (v >= AP26_MOD) v -= AP26_MOD;
  n43[j] = v;
}

/* Tables of every (job, shift) of the round; grid z = job*10 + shift.  S = 1:
   [job][shift][entry]; S = 2: [job][shift pair][entry][shift in pair].  Shifts
   before the job's resume point get all-zero masks. */
static __global__ void batch_tables_kernel(const BatchJobDev *__restrict__ jobs, const RoundState *__restrict__ rsp,
                                           uint64_t *__restrict__ tab, int S) {
  const RoundState rs = *rsp;
  const int njobs = min(rs.per_round, rs.njobs - rs.base);
  const int jl = blockIdx.z / AP26_NSHIFT;
  const int it = blockIdx.z - jl * AP26_NSHIFT;
  if (jl >= njobs) return;
  const int ti = blockIdx.y;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= PT.p[ti]) return;
  const BatchJobDev &job = jobs[rs.base + jl];
  const uint64_t w = it < job.it0 ? 0 : table_word(job.kp, ti, c, job.kp.SHIFT + 64 * it);
  if (S == 1)
    tab[((size_t)jl * AP26_NSHIFT + it) * NTAB + PT.off[ti] + c] = w;
  else
    tab[((size_t)jl * (AP26_NSHIFT / 2) + it / 2) * NTAB * 2 + (size_t)(PT.off[ti] + c) * 2 + (it & 1)] = w;
}

/* ---- shift-pair persistent sieve (S = 2) -------------------------------- */
/* Level 0 reads both shifts of a pair with one 16-byte LDS load per prime and
   one residue; the surviving (position, shift) pairs are queued separately,
   the shift bit stored in bit 48 of the position's double (n59 < 2^48). */

constexpr double TWO48 = 281474976710656.0;

__device__ __forceinline__ void decode2(double xd, double &x, uint32_t &xs, int &e) {
  e = (int)((__double_as_longlong(xd) >> 48) & 1);
  x = xd - (e ? TWO52 + TWO48 : TWO52);
  xs = lo32(xd) * 16u + 8u * (uint32_t)e;
}

/* Scheduling hint: interleave the residue arithmetic with the LDS gathers
   (9 VALU instructions, then 3 DS reads, repeated), so that a wave's first
   gathers are in flight while it computes the later residues (sieve 177 ->
   174 ms/K against all residues first, then all reads). */
template <int NREADS>
__device__ __forceinline__ void interleave_residues_and_reads() {
  static_for<0, (NREADS + 2) / 3>([&](auto) {
    __builtin_amdgcn_sched_group_barrier(0x2, 9, 0);   /* VALU */
    __builtin_amdgcn_sched_group_barrier(0x100, 3, 0); /* DS read */
  });
}

/* ---- bank-rotated level 0 (shift-pair sieve) ----------------------------- */
/* A random 16-byte LDS gather serves the 16 lanes of a pass in about three
   times the cycles of a conflict-free one: lanes land on the same 4-bank group
   with different addresses.  The rotated layout removes that for the first
   ROT_N = 16 sieve primes (61..137): LDS row rho (256 bytes, one per bank
   sweep) holds in its 16-byte column q the shift-pair entry of prime q for
   the residue rho - ROT_HALF, so prime q lives in bank group q only.  At
   step k of level 0, lane j looks up prime (j + k) mod 16 for its own
   position; the 16 lanes of each pass then read 16 different bank groups,
   and after 16 steps every lane has ANDed every prime of the group.  The
   per-lane prime constants (reciprocal, row multiplier, column) live in
   VGPRs (RotConsts).  With 256-byte rows the i24 multiply-add gives the row
   offset 256*r exactly modulo 2^32 (256 * 2^24 = 2^32), so no sign
   extension is needed: |256 r| < 2^31. */
constexpr int ROT_N = 16;
constexpr int ROT_HALF = (PT.p[ROT_N - 1] - 1) / 2; /* 68 */
constexpr int ROT_ROWS = 2 * ROT_HALF + 1;          /* 137 */
constexpr int ROT_BYTES = 256 * ROT_ROWS;           /* 35,072 */
static_assert(ROT_N == 16 && PT.p[ROT_N - 1] == 137, "rotation group is 61..137");

/* The rotation layouts, chosen by R0: 0 (none), ROT_N (rows of 256 bytes,
   one column per prime) or ROT_WIDE (rows of 512 bytes holding every column
   twice: at step k lane j reads column (j mod 16) + k, whose bank group is
   (j + k) mod 16 as before, but whose offset 16 (j mod 16) + 16 k is a
   per-lane constant plus an immediate, so the position's multiply-add
   produces the address and no per-lane column registers are needed). */
constexpr int ROT_WIDE = 2 * ROT_N;
constexpr int rot_primes(int R0) { return R0 ? ROT_N : 0; }            /* first packed prime */
constexpr int rot_row_bytes(int R0) { return R0 == ROT_WIDE ? 512 : 256; }

/* LDS bytes of the rotation rows. */
template <int R0>
constexpr int rot_lds_bytes() {
  return R0 ? rot_row_bytes(R0) * ROT_ROWS : 0;
}

/* LDS bytes of the packed shift-pair tables of primes rot_primes(R0)..NL-1
   and the 7..23 filters.  They come first in the table area, so that every
   packed entry's offset fits the 16-bit immediate of the LDS read (no add
   per lookup); the rotation rows follow, and their base is folded into the
   per-position multiply-add addend of level 0. */
template <int R0, int NL>
constexpr int packed_lds_bytes() {
  return 16 * (PT.off[NL] - PT.off[rot_primes(R0)] + 90);
}

template <int W>
struct RotConstsT {
  double inv[ROT_N]; /* RN(1/p), the constant centred_byteoff uses */
  int m[ROT_N];      /* -W p */
  int col[ROT_N];    /* 16 q: byte offset of prime q's column in a 256-byte row */
  __device__ __forceinline__ void init(int lane) {
    static_for<0, ROT_N>([&](auto KC) {
      constexpr int k = decltype(KC)::value;
      const int q = (lane + k) & (ROT_N - 1);
      int p = PT.p[0];
      static_for<1, ROT_N>([&](auto QC) {
        constexpr int Q = decltype(QC)::value;
        p = q == Q ? PT.p[Q] : p;
      });
      inv[k] = 1.0 / (double)p; /* correctly rounded, = the compile-time 1.0/P */
      m[k] = -W * p;
      col[k] = 16 * q;
    });
  }
};
using RotConsts = RotConstsT<256>;

/* Stage one shift pair's tables into LDS: the packed entries of primes
   rot_primes(R0)..NL-1, the 7..23 filters, then the rotation rows (R0 != 0).
   tab is the pair's global table, [entry][shift in pair], 16 bytes per
   entry. */
template <int R0, int NL>
__device__ __forceinline__ void stage_pair_tables(char *lds, const uint64_t *__restrict__ tab) {
  constexpr int RB = packed_lds_bytes<R0, NL>();
  if constexpr (R0) {
    constexpr int NC = rot_row_bytes(R0) / 16; /* columns per row: 16 or 32 */
    const uint4 *t4 = (const uint4 *)tab;
    uint4 *rows = (uint4 *)(lds + RB);
    for (int i = threadIdx.x; i < ROT_ROWS * NC; i += blockDim.x) {
      const int q = i & (ROT_N - 1);
      const int p = PT.p[q];
      int c = i / NC - ROT_HALF + (p - 1) / 2; /* centred entry of residue row - ROT_HALF */
      c = (c % p + p) % p;                      /* rows past (p-1)/2 are never read; keep them valid */
      rows[i] = t4[PT.off[q] + c];
    }
  }
  uint64_t *packed = (uint64_t *)lds;
  constexpr int P0 = rot_primes(R0);
  constexpr int NPK = 2 * (PT.off[NL] - PT.off[P0]);
  for (int i = threadIdx.x; i < NPK; i += blockDim.x) packed[i] = tab[2 * PT.off[P0] + i];
  for (int i = threadIdx.x; i < 180; i += blockDim.x) packed[NPK + i] = tab[2 * NOK + i];
}

/* a & b & c in one instruction (v_bitop3_b32 with truth table 0x80). */
__device__ __forceinline__ uint32_t and3(uint32_t a, uint32_t b, uint32_t c) {
  uint32_t r;
  asm("v_bitop3_b32 %0, %1, %2, %3 bitop3:0x80" : "=v"(r) : "v"(a), "v"(b), "v"(c));
  return r;
}

/* AND of a 64-bit mask with N words, two words per three-input instruction
   and half.  (Written with 64-bit ANDs, the chain compiles to one v_and_b32
   per word and half.) */
template <int N>
__device__ __forceinline__ uint64_t and_words(uint64_t m, const uint64_t (&w)[N]) {
  uint32_t lo = (uint32_t)m, hi = (uint32_t)(m >> 32);
  static_for<0, N / 2>([&](auto KC) {
    constexpr int k = 2 * decltype(KC)::value;
    lo = and3(lo, (uint32_t)w[k], (uint32_t)w[k + 1]);
    hi = and3(hi, (uint32_t)(w[k] >> 32), (uint32_t)(w[k + 1] >> 32));
  });
  if constexpr (N % 2) {
    lo &= (uint32_t)w[N - 1];
    hi &= (uint32_t)(w[N - 1] >> 32);
  }
  return ((uint64_t)hi << 32) | lo;
}

/* The primes [B, E) of a queued (position, shift) entry, in chunks of CH
   primes (residues, loads, ANDs per chunk): with the rotation group's
   constants resident, the deep levels must not hold all their loads at once. */
template <int NL, int B, int E, int R0 = 0, int CH = 64>
__device__ __forceinline__ uint64_t apply_primes2(uint64_t m, double xd, const char *ldsb, const char *gtb) {
  static_assert(B >= rot_primes(R0), "rotated primes are level 0 only");
  double x;
  uint32_t xs;
  int e;
  decode2(xd, x, xs, e);
  if constexpr (E - B > CH) {
    m = apply_primes2<NL, B, B + CH, R0, CH>(m, xd, ldsb, gtb);
    return apply_primes2<NL, B + CH, E, R0, CH>(m, xd, ldsb, gtb);
  } else {
    int bo[E - B];
    uint64_t w[E - B];
    static_for<B, E>([&](auto IC) {
      constexpr int I = decltype(IC)::value;
      bo[I - B] = centred_byteoff<PT.p[I], 16>(x, xs);
    });
    static_for<B, E>([&](auto IC) {
      constexpr int I = decltype(IC)::value;
      if constexpr (I < NL) {
        constexpr int C0 = 16 * (PT.off[I] - PT.off[rot_primes(R0)] + (PT.p[I] - 1) / 2);
        w[I - B] = *(const uint64_t *)(ldsb + C0 + bo[I - B]);
      } else {
        /* C0 + bo is a byte offset within the prime's p entries, never
           negative: as an unsigned 32-bit offset from the wave-uniform table
           base it takes one add (the scalar-base form of the load) instead
           of a 64-bit sign extension and add */
        constexpr int C0 = 16 * (PT.off[I] + (PT.p[I] - 1) / 2);
        w[I - B] = *(const uint64_t *)(gtb + (uint32_t)(C0 + bo[I - B]));
      }
    });
    if constexpr (E <= NL) interleave_residues_and_reads<E - B>();
    return and_words(m, w);
  }
}

template <int NL, int R0 = 0>
__device__ __forceinline__ uint64_t apply_filters2(uint64_t m, double xd, const char *ldsb) {
  double x;
  uint32_t xs;
  int e;
  decode2(xd, x, xs, e);
  constexpr int LMAIN = PT.off[NL] - PT.off[rot_primes(R0)];
  static_for<NSP, NPT>([&](auto IC) {
    constexpr int I = decltype(IC)::value;
    const int bo = centred_byteoff<PT.p[I], 16>(x, xs);
    constexpr int C0 = 16 * (LMAIN + PT.off[I] - NOK + (PT.p[I] - 1) / 2);
    m &= *(const uint64_t *)(ldsb + C0 + bo);
  });
  return m;
}

/* Level 0 of a fresh position with the rotation group: both masks of the
   pair over primes 0..E-1, the first ROT_N through the rotated rows, the rest
   (ROT_N..E-1) from the packed LDS tables.  init (all ones, or zero for a
   lane whose position is only a stand-in) starts the AND chains, so such a
   lane's masks come out zero at no extra cost.  RBASE: byte offset of the
   rotation rows in the table area (after the packed tables). */
template <int E, int SPLIT, int W, int RBASE, int GL = 0, bool SKIP = false>
__device__ __forceinline__ void level0_rot(double xd, const char *ldsb, const char *gtb, const RotConstsT<W> &rc,
                                           uint32_t init, uint64_t &m0, uint64_t &m1) {
  static_assert(ROT_N % SPLIT == 0, "equal parts of the rotation group");
  static_assert(W == 256 || W == 512, "row width");
  constexpr int PART = ROT_N / SPLIT;
  static_assert(E >= ROT_N, "level 0 holds the rotation group");
  static_assert(GL >= 0 && GL <= E - ROT_N, "vector-memory primes are packed level-0 primes");
  const double x = xd - TWO52;
  const uint32_t xs16 = lo32(xd) * 16u;
  /* the last GL primes from the pair's global table, issued first so that
     their longer latency runs under the rotated gathers; the byte offset is
     non-negative (a centred entry within its prime's p entries) */
  uint4 g[GL > 0 ? GL : 1];
  if constexpr (GL > 0) {
    static_for<E - GL, E>([&](auto IC) {
      constexpr int I = decltype(IC)::value;
      const int bo = centred_byteoff<PT.p[I], 16>(x, xs16);
      constexpr int C0 = 16 * (PT.off[I] + (PT.p[I] - 1) / 2);
      g[I - (E - GL)] = *(const uint4 *)(gtb + (uint32_t)(C0 + bo));
    });
  }
  /* W * x mod 2^32 (exact: W * 2^24 = 0 mod 2^32 for the i24 multiply-add)
     plus the offset of the rows' residue 0 (the rows' base and ROT_HALF
     rows); wide rows: plus this lane's column */
  const uint32_t xsw = lo32(xd) * (uint32_t)W + (uint32_t)(RBASE + W * ROT_HALF) +
                       (W == 512 ? 16u * (threadIdx.x & 15) : 0u);
  uint32_t a0 = init, a1 = init, a2 = init, a3 = init;
  /* SPLIT parts of the group, so that at most 16/SPLIT gathers are in flight */
  static_for<0, SPLIT>([&](auto HC) {
    constexpr int H = decltype(HC)::value;
    uint4 w[PART];
    static_for<0, PART>([&](auto KC) {
      constexpr int k = H * PART + decltype(KC)::value;
      const double qd = __fma_rn(x, rc.inv[k], SHIFTER);
      int r; /* W * (centred residue of prime (lane + k) mod 16) [+ the wide row's offsets] */
      asm("v_mad_i32_i24 %0, %1, %2, %3" : "=v"(r) : "v"((int)lo32(qd)), "v"(rc.m[k]), "v"(xsw));
      if constexpr (W == 512)
        w[decltype(KC)::value] = *(const uint4 *)(ldsb + r + 16 * k); /* column (lane mod 16) + k */
      else
        w[decltype(KC)::value] = *(const uint4 *)(ldsb + r + rc.col[k]);
    });
    interleave_residues_and_reads<PART>();
    static_for<0, PART>([&](auto KC) {
      constexpr int k = decltype(KC)::value;
      a0 &= w[k].x;
      a1 &= w[k].y;
      a2 &= w[k].z;
      a3 &= w[k].w;
    });
  });
  if constexpr (E - GL > ROT_N) {
    auto packed = [&]() {
      uint4 w[E - GL - ROT_N];
      static_for<ROT_N, E - GL>([&](auto IC) {
        constexpr int I = decltype(IC)::value;
        const int bo = centred_byteoff<PT.p[I], 16>(x, xs16);
        constexpr int C0 = 16 * (PT.off[I] - PT.off[ROT_N] + (PT.p[I] - 1) / 2);
        w[I - ROT_N] = *(const uint4 *)(ldsb + C0 + bo);
      });
      static_for<0, E - GL - ROT_N>([&](auto IC) {
        constexpr int k = decltype(IC)::value;
        a0 &= w[k].x;
        a1 &= w[k].y;
        a2 &= w[k].z;
        a3 &= w[k].w;
      });
    };
    if constexpr (SKIP) {
      if ((a0 | a1 | a2 | a3) != 0u) packed(); /* a zero mask stays zero */
    } else {
      packed();
    }
  }
  if constexpr (GL > 0) {
    static_for<0, GL>([&](auto IC) {
      constexpr int k = decltype(IC)::value;
      a0 &= g[k].x;
      a1 &= g[k].y;
      a2 &= g[k].z;
      a3 &= g[k].w;
    });
  }
  m0 = ((uint64_t)a1 << 32) | a0;
  m1 = ((uint64_t)a3 << 32) | a2;
}

constexpr int pow_mod_const(int a, int e, int p) {
  long long r = 1, b = a % p;
  while (e) {
    if (e & 1) r = r * b % p;
    b = b * b % p;
    e >>= 1;
  }
  return (int)r;
}

/* Both 64-bit masks of a shift pair for the prime P from its window string
   W (host_vwin) in registers: r = n59 mod P (centred), rho = (r*M^-1 +
   shift0) mod P by one multiply-add and a reciprocal multiply, and the
   masks are bits [rho, rho+128) of W, by 64-bit funnel shifts. */
template <int P>
__device__ __forceinline__ void vwin_pair(double x, uint32_t xs1, const uint64_t *W, int C, uint32_t &a0,
                                          uint32_t &a1, uint32_t &a2, uint32_t &a3) {
  constexpr int U = pow_mod_const((int)(AP26_MOD % P), P - 2, P);         /* M^-1 mod P */
  constexpr uint32_t M32 = (uint32_t)((((uint64_t)1 << 32) + P - 1) / P); /* ceil(2^32/P) */
  const int r = centred_byteoff<P, 8>(x, xs1 * 8u) >> 3;
  const int xv = r * U + C;                                 /* in [0, 2^13) */
  const int q = (int)__umulhi((uint32_t)xv, M32);           /* exact: xv*P < 2^32/P */
  const int rho = xv - q * P;                               /* [0, P) */
  const uint32_t sh = (uint32_t)rho & 63u, ns = 63u - sh;
  const uint64_t lo_a = (W[0] >> sh) | ((W[1] << 1) << ns); /* bits [rho, rho+64) for rho < 64 */
  const uint64_t hi_a = (W[1] >> sh) | ((W[2] << 1) << ns);
  uint64_t lo = lo_a, hi = hi_a;
  if constexpr (P > 64) {
    const uint64_t hi_b = (W[2] >> sh) | ((W[3] << 1) << ns);
    const bool up = rho >= 64;
    lo = up ? hi_a : lo_a;
    hi = up ? hi_b : hi_a;
  }
  a0 &= (uint32_t)lo;
  a1 &= (uint32_t)(lo >> 32);
  a2 &= (uint32_t)hi;
  a3 &= (uint32_t)(hi >> 32);
}

/* level 0 of a fresh position (shift bit 0): both masks of the pair; the
   first V primes (61, 67, 71) from their window strings, the last G through
   the vector-memory path (L1; same table in global memory), the rest from
   LDS: the three paths run in parallel, and LDS is the busiest. */
template <int B, int E, int V = 0, int G = 0>
__device__ __forceinline__ void level0_pair(double xd, const char *ldsb, const char *gtb, uint64_t &m0,
                                            uint64_t &m1, const KParams &kp, const int32_t *vc) {
  static_assert(B == 0 && V <= NVW && V <= E, "register primes lead level 0");
  const double x = xd - TWO52;
  const uint32_t xs = lo32(xd) * 16u;
  int bo[E - B];
  uint4 w[E - B];
  static_for<B + V, E>([&](auto IC) {
    constexpr int I = decltype(IC)::value;
    bo[I - B] = centred_byteoff<PT.p[I], 16>(x, xs);
  });
  static_for<B + V, E>([&](auto IC) {
    constexpr int I = decltype(IC)::value;
    constexpr int C0 = 16 * (PT.off[I] + (PT.p[I] - 1) / 2);
    if constexpr (I >= E - G)
      w[I - B] = *(const uint4 *)(gtb + (uint32_t)(C0 + bo[I - B]));
    else
      w[I - B] = *(const uint4 *)(ldsb + C0 + bo[I - B]);
  });
  interleave_residues_and_reads<E - B - V>();
  uint32_t a0 = ~0u, a1 = ~0u, a2 = ~0u, a3 = ~0u;
  static_for<0, V>([&](auto IC) {
    constexpr int I = decltype(IC)::value;
    static_assert(PT.p[I] == VW_P[I], "register primes are 61, 67, 71");
    vwin_pair<PT.p[I]>(x, lo32(xd), kp.vw[I], vc[I], a0, a1, a2, a3);
  });
  static_for<B + V, E>([&](auto IC) {
    constexpr int I = decltype(IC)::value;
    a0 &= w[I - B].x;
    a1 &= w[I - B].y;
    a2 &= w[I - B].z;
    a3 &= w[I - B].w;
  });
  m0 = ((uint64_t)a1 << 32) | a0;
  m1 = ((uint64_t)a3 << 32) | a2;
}

/* Survivor slots of one wave in the batched sieve: the wave reserves
   SURV_CHUNK consecutive slots with one atomic and fills them before taking
   more, so the survivor list is written in whole cache lines and the global
   counter sees one atomic per SURV_CHUNK survivors.  Slots a wave leaves
   unused when it finishes hold -1, which the walk skips. */
constexpr int SURV_CHUNK = 128;
struct SurvCursor {
  unsigned long long base; /* next free slot */
  int left;                /* free slots left at base */
};

__device__ __forceinline__ void emit_survivors2(uint64_t w, double xd, const SieveArgs &a, int lane,
                                                SurvCursor &cur) {
  const uint64_t bits = (uint64_t)__double_as_longlong(xd);
  const int e = (int)((bits >> 48) & 1);
  const int64_t n59 = (int64_t)(bits & ((INT64_C(1) << 48) - 1));
  const int64_t nbase = n59 + (int64_t)(a.shift0 + 64 * e) * AP26_MOD;
  while (true) {
    const bool has = w != 0;
    const uint64_t bal = __ballot(has);
    if (bal == 0) break;
    const int cnt = __popcll(bal);
    const int leader = __ffsll((unsigned long long)bal) - 1;
    const int pre = lanes_below(bal);
    unsigned long long fresh = 0;
    if (cnt > cur.left) { /* wave-uniform */
      if (lane == leader) fresh = atomicAdd(a.nsurv, (unsigned long long)SURV_CHUNK);
      fresh = __shfl(fresh, leader);
    }
    if (has) {
      const int bit = __ffsll((unsigned long long)w) - 1;
      w &= w - 1;
      const unsigned long long idx =
          pre < cur.left ? cur.base + (unsigned long long)pre : fresh + (unsigned long long)(pre - cur.left);
      if (idx < (unsigned long long)a.cap) {
        a.surv[idx] = nbase + (int64_t)bit * AP26_MOD;
        if (a.surv_job) a.surv_job[idx] = a.job;
      }
    }
    if (cnt > cur.left) {
      cur.base = fresh + (unsigned long long)(cnt - cur.left);
      cur.left = SURV_CHUNK - (cnt - cur.left);
    } else {
      cur.base += (unsigned long long)cnt;
      cur.left -= cnt;
    }
  }
}

/* Mark the wave's unused reserved slots as empty (-1). */
__device__ __forceinline__ void close_survivors2(const SurvCursor &cur, int64_t *surv, int64_t cap, int lane) {
  for (int i = lane; i < cur.left; i += 64) {
    const unsigned long long idx = cur.base + (unsigned long long)i;
    if (idx < (unsigned long long)cap) surv[idx] = -1;
  }
}

/* Wave issue priorities by phase of the pair sieve (s_setprio): level-0
   gathers 2, level-1 drains 3, deeper drains 1, queue bookkeeping 0.  When
   several waves of a SIMD are ready, the arbiter then interleaves phases with
   different unit mixes (LDS gathers, residue VALU work, L2 loads) instead of
   taking waves round-robin: 185 -> 177 ms/K of sieve on the 16-K round, with
   equal priorities for the two phases 180, level 0 alone raised 181. */
constexpr int PRIO_LEVEL0 = 2, PRIO_LEVEL1 = 3, PRIO_DEEP = 1;

/* Tuning of the rotated shift-pair sieve: wave priorities of the level-0
   step, the level-1 drains and the deeper drains, and the number of parts
   the rotated level-0 gathers are issued in. */
template <int P0 = PRIO_LEVEL0, int P1 = PRIO_LEVEL1, int PD = PRIO_DEEP, int SPLIT = 2>
struct Tune {
  static constexpr int prio0 = P0, prio1 = P1, prio_deep = PD, split = SPLIT;
  static constexpr int l1 = 0;          /* level-0 primes read through vector memory */
  static constexpr bool lane_skip = false; /* packed level-0 gathers only on live lanes */
  static constexpr int nt_queue = 99;    /* global queues from this one on: non-temporal */
};

/* As Tune, and the last L1 primes of level 0 (the packed ones after the
   rotated group) are read from the pair's global table through the vector
   memory path (L1) instead of LDS: their random 16-byte gathers are the
   level-0 reads that meet bank conflicts, and the L1 is otherwise lightly
   used. */
template <int P0, int P1, int PD, int SPLIT, int L1>
struct TuneL1 : Tune<P0, P1, PD, SPLIT> {
  static constexpr int l1 = L1;
};

/* As Tune, and the global queues NTQ.. (the deep ones, which fill slowly)
   are written and read with non-temporal hints, so that their lines are the
   first the L2 evicts and the busier queues' lines stay. */
template <int P0, int P1, int PD, int SPLIT, int NTQ>
struct TuneNT : Tune<P0, P1, PD, SPLIT> {
  static constexpr int nt_queue = NTQ;
};

/* As Tune, and the packed level-0 gathers (after the rotated group) run only
   on the lanes whose pair mask is still non-zero: the LDS resolves bank
   conflicts among the active lanes only, at the price of waiting for the
   rotated group's result before those gathers are issued. */
template <int P0, int P1, int PD, int SPLIT>
struct TuneSkip : Tune<P0, P1, PD, SPLIT> {
  static constexpr bool lane_skip = true;
};

/* Queue capacity per wave: queue 0 holds Q0 entries, the deeper ones 128.
   With Q0 = 192 both pushes of a position (one per shift of the pair) go in
   before one drain: a single call site of the drain chain per position
   (its code and the registers it needs appear once).  Queues GQ and deeper
   live in a per-wave region of global memory instead of LDS (GQ >= the queue
   count: all in LDS); the LDS holds the others. */
constexpr int NO_GLOBAL_QUEUES = 99;

template <int Q0, class LV, int GQ = NO_GLOBAL_QUEUES>
constexpr int qsieve2_wave_entries() {
  constexpr int NQ = LV::N - 1;
  constexpr int NL_Q = NQ < GQ ? NQ : GQ; /* queues in LDS */
  return Q0 + 128 * (NL_Q - 1);
}

/* Global-memory queue entries per wave (queues GQ .. NQ-1, 128 each). */
template <class LV, int GQ>
constexpr int qsieve2_global_entries() {
  constexpr int NQ = LV::N - 1;
  return GQ < NQ ? 128 * (NQ - GQ) : 0;
}

template <int NL, class LV, int R0 = 0, int Q0 = 128, class TU = Tune<>, int GQ = NO_GLOBAL_QUEUES>
struct QSieve2 {
  static_assert(Q0 == 128 || Q0 == 192, "queue 0 holds one or two pushes past a full wave");
  static_assert(GQ >= 1, "queue 0 stays in LDS");
  static constexpr int NQ = LV::N - 1;
  QEntry *qb;  /* this wave's LDS queues: queue 0 (Q0 entries), then 128 per level below GQ */
  QEntry *gqb; /* this wave's global queues GQ.. (128 entries each) */
  int qn[NQ];
  uint64_t seen_exec; /* AND of the exec masks of every queue step: not all ones if
                         a step ran with lanes switched off (must never happen);
                         wave-uniform, so it stays in scalar registers */
  const char *ldsb, *gtb;
  int lane;
  const SieveArgs *a;
  SurvCursor *cur;

  /* Every lane keeps the same counts only if every lane runs every push.
     Pushes happen in straight-line code of the position steps, drain
     iterations and flushes, whose loops are uniform; each of those checks
     once that the whole wave runs it, and a failure is reported (the queue
     would lose entries). */
  __device__ __forceinline__ void check_full_wave() {
    seen_exec &= __builtin_amdgcn_read_exec();
  }

  /* A wave reads back its own global queue entries, written by other lanes
     of the same wave: its vector memory operations are processed in order,
     and the CU's L1 is write-through, so no fence is needed between a push
     and the drain that pops it. */
  template <int QI>
  __device__ __forceinline__ QEntry *q() const {
    if constexpr (QI >= GQ)
      return gqb + 128 * (QI - GQ);
    else
      return qb + (QI == 0 ? 0 : Q0 + 128 * (QI - 1));
  }

  /* Entry idx (0 <= idx < the queue's capacity) of queue QI; in global
     memory as an unsigned 32-bit byte offset from the wave-uniform base, so
     that the access takes the scalar-base form (no 64-bit address add). */
  template <int QI>
  __device__ __forceinline__ QEntry &qe(int idx) const {
    if constexpr (QI >= GQ)
      return *(QEntry *)((char *)gqb + (uint32_t)((128 * (QI - GQ) + idx) * (int)sizeof(QEntry)));
    else
      return q<QI>()[idx];
  }

  template <int QI>
  __device__ __forceinline__ void put(int idx, double xd, uint64_t m) const {
    if constexpr (QI >= GQ && QI >= TU::nt_queue) {
      QEntry &e = qe<QI>(idx);
      __builtin_nontemporal_store(xd, &e.xd);
      __builtin_nontemporal_store(m, &e.m);
    } else {
      qe<QI>(idx) = QEntry{xd, m};
    }
  }

  template <int QI>
  __device__ __forceinline__ QEntry get(int idx) const {
    if constexpr (QI >= GQ && QI >= TU::nt_queue) {
      const QEntry &e = qe<QI>(idx);
      return QEntry{__builtin_nontemporal_load(&e.xd), __builtin_nontemporal_load(&e.m)};
    } else {
      return qe<QI>(idx);
    }
  }

  template <int QI>
  __device__ __forceinline__ void push(double xd, uint64_t m, bool pred) {
    const uint64_t bal = __ballot(pred);
    if (pred) put<QI>(qn[QI] + lanes_below(bal), xd, m);
    qn[QI] += __popcll(bal);
    __builtin_amdgcn_wave_barrier();
  }

  template <int L>
  __device__ __forceinline__ void run_level(const QEntry &e, bool pred) {
    uint64_t m = e.m;
    if (pred) m = apply_primes2<NL, LV::B[L + 1], LV::B[L + 2], R0, (R0 ? 8 : 64)>(m, e.xd, ldsb, gtb);
    if constexpr (L + 2 == LV::N) {
      if (pred && m) m = apply_filters2<NL, R0>(m, e.xd, ldsb);
      emit_survivors2(pred ? m : 0, e.xd, *a, lane, *cur);
    } else {
      push<L + 1>(e.xd, m, pred && m != 0);
      drain<L + 1>();
    }
  }

  template <int L>
  __device__ __forceinline__ void drain() {
    while (qn[L] >= 64) {
      check_full_wave();
      qn[L] -= 64;
      const QEntry e = get<L>(qn[L] + lane);
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_s_setprio(L == 0 ? TU::prio1 : TU::prio_deep);
      run_level<L>(e, true);
      __builtin_amdgcn_s_setprio(0);
    }
  }

  template <int L>
  __device__ __forceinline__ void flush() {
    if constexpr (L < NQ) {
      drain<L>();
      const int n = qn[L];
      if (n > 0) {
        check_full_wave();
        const bool pred = lane < n;
        QEntry e{TWO52, 0};
        if (pred) e = get<L>(lane);
        __builtin_amdgcn_wave_barrier();
        qn[L] = 0;
        run_level<L>(e, pred);
      }
      flush<L + 1>();
    } else {
      if (seen_exec != ~UINT64_C(0) && lane == 0) atomicOr(a->err, 1u);
    }
  }
};

template <int NL, class LV, int V = 0, int G = 0, int R0 = 0, int Q0 = 128, class TU = Tune<>,
          int GQ = NO_GLOBAL_QUEUES, int DYN = 0>
__device__ __forceinline__ void qsieve2_blocks(QSieve2<NL, LV, R0, Q0, TU, GQ> &qs, const SieveArgs &a, int64_t tid,
                                               int64_t stride,
                                               const RotConstsT<rot_row_bytes(R0)> *rc = nullptr,
                                               unsigned int *slab = nullptr) {
  const double S59d = (double)a.kp.S59;
  const double MODd = (double)AP26_MOD;
  const double WRAP = MODd + TWO52;
  /* the 35 positions of block b (lanes with b >= blk1 work on a stand-in and
     keep masks of zero); as in qsieve_blocks, uniform loops: the whole wave
     in every step */
  auto sieve_block = [&](int64_t b) {
    const bool has = b < a.blk1;
    double xd = (double)n53_of_block(a.kp, a.n43, has ? b : a.blk0) + TWO52;
    uint32_t init = has ? ~0u : 0u; /* lanes past the last block: masks of zero */
    /* opaque to the optimiser, which would otherwise turn the AND with init
       back into a select per mask word after the chains */
    asm volatile("" : "+v"(init));
    for (int i59 = 0; i59 < AP26_N59_STEPS; i59++) {
      qs.check_full_wave();
      uint64_t m0, m1;
      __builtin_amdgcn_s_setprio(TU::prio0);
      if constexpr (R0) {
        level0_rot<LV::B[1], TU::split, rot_row_bytes(R0), packed_lds_bytes<R0, NL>(), TU::l1, TU::lane_skip>(
            xd, qs.ldsb, qs.gtb, *rc, init, m0, m1);
      } else {
        level0_pair<0, LV::B[1], V, G>(xd, qs.ldsb, qs.gtb, m0, m1, a.kp, a.vc);
        m0 = has ? m0 : 0;
        m1 = has ? m1 : 0;
      }
      __builtin_amdgcn_s_setprio(0);
      qs.template push<0>(xd, m0, m0 != 0);
      if constexpr (Q0 == 128) qs.template drain<0>();
      qs.template push<0>(xd + TWO48, m1, m1 != 0);
      qs.template drain<0>();
      xd += S59d;
      xd = xd >= WRAP ? xd - MODd : xd;
    }
  };
  if constexpr (DYN > 0) {
    /* `slab`: a counter of claimed blocks, zero at the start -- in LDS for
       the workgroup's item (DYN = 1) or in global memory for a launch's
       whole grid (DYN = 4, fewer atomics).  Each wave claims the next DYN
       slabs of 64 blocks when it is ready for them instead of taking every
       stride-th slab, so the waves reach the end of the item (or launch)
       together however their deeper levels fared. */
    while (true) {
      unsigned int s0 = 0;
      if (qs.lane == 0) s0 = atomicAdd(slab, 64u * DYN);
      s0 = __builtin_amdgcn_readfirstlane(s0); /* lane 0's claim, wave-uniform */
      const int64_t base = a.blk0 + (int64_t)s0;
      if (base >= a.blk1) break;
      for (int k = 0; k < DYN; k++) {
        if (base + 64 * k >= a.blk1) break;
        sieve_block(base + 64 * k + qs.lane);
      }
    }
  } else {
    for (int64_t base = a.blk0; base < a.blk1; base += stride) {
      if (__ballot(base + tid < a.blk1) == 0) break;
      sieve_block(base + tid);
    }
  }
  qs.template flush<0>();
}

template <int NL, int WAVES, class LV, int R0 = 0, int Q0 = 128, int GQ = NO_GLOBAL_QUEUES>
struct QSieve2Lds {
  /* packed pair entries of primes rot_primes(R0)..NL-1, filters, rotation rows */
  alignas(16) char tab[packed_lds_bytes<R0, NL>() + rot_lds_bytes<R0>()];
  QEntry queues[WAVES][qsieve2_wave_entries<Q0, LV, GQ>()];
};

template <int NL, int WAVES, class LV, int R0 = 0, int Q0 = 128, int GQ = NO_GLOBAL_QUEUES>
constexpr int qsieve2_lds_bytes() {
  return (int)sizeof(QSieve2Lds<NL, WAVES, LV, R0, Q0, GQ>);
}

/* Items (job, shift pair, n43 chunk); tables [entry][shift in pair].  V: the
   first V level-0 primes from registers (vwin_pair) instead of LDS.  R0 =
   ROT_N: level 0 starts with the bank-rotated group (level0_rot). */
template <int NL, int WAVES, class LV, int V = 0, int G = 0, int R0 = 0, class TU = Tune<>,
          int GQ = NO_GLOBAL_QUEUES>
static __global__ void __launch_bounds__(64 * WAVES) __attribute__((amdgpu_waves_per_eu(WAVES / 4, (WAVES + 3) / 4)))
sieve_persistent2_kernel(PersistArgs p) {
  constexpr int Q0 = R0 ? 192 : 128;
  static_assert(qsieve2_lds_bytes<NL, WAVES, LV, R0, Q0, GQ>() + 64 + 16 * WAVES <= 160 * 1024, "LDS budget");
  static_assert(LV::B[1] <= NL, "level 0 must be in LDS");
  static_assert(R0 == 0 || ((R0 == ROT_N || R0 == ROT_WIDE) && V == 0 && G == 0),
                "rotation group replaces V/G level-0 forms");
  constexpr int NPASS = AP26_NSHIFT / 2;
  __shared__ QSieve2Lds<NL, WAVES, LV, R0, Q0, GQ> L;
  __shared__ unsigned int s_item;
  __shared__ unsigned int s_slab; /* blocks of the current item claimed by its waves */
  const RoundState rs = *p.rs;
  const int njobs = min(rs.per_round, rs.njobs - rs.base);
  const unsigned int nitems = (unsigned int)(njobs > 0 ? njobs : 0) * NPASS * p.nchunk;
  int cur = -1;
  __shared__ SurvCursor s_cur[WAVES]; /* per wave, in LDS to spare registers */
  SurvCursor &sc = s_cur[threadIdx.x >> 6];
  sc = SurvCursor{0, 0};
  RotConstsT<rot_row_bytes(R0)> rc;
  if constexpr (R0) rc.init(threadIdx.x & 63);
  while (true) {
    __syncthreads();
    if (threadIdx.x == 0) {
      s_item = atomicAdd(p.work, 1u);
      s_slab = 0;
    }
    __syncthreads();
    const unsigned int item = __builtin_amdgcn_readfirstlane(s_item); /* uniform: keeps job data scalar */
    if (item >= nitems) {
      close_survivors2(sc, p.surv, p.cap, threadIdx.x & 63);
      break;
    }
    const int jl = (int)(item / (NPASS * p.nchunk));
    const int rem = (int)(item - (unsigned int)jl * NPASS * p.nchunk);
    const int t = rem / p.nchunk;
    const int chunk = rem - t * p.nchunk;
    const BatchJobDev &job = p.jobs[rs.base + jl];
    if (2 * t + 1 < job.it0) continue;
    const uint64_t *tab = p.tab + ((size_t)jl * NPASS + t) * NTAB * 2;
    const int key = jl * NPASS + t;
    if (key != cur) {
      stage_pair_tables<R0, NL>(L.tab, tab);
      __syncthreads();
      cur = key;
    }
    SieveArgs a;
    a.kp = job.kp;
    a.n43 = p.n43 + (size_t)jl * AP26_NN43;
    a.tab = tab;
    a.shift0 = job.kp.SHIFT + 128 * t;
    a.nslot = 2;
    for (int v = 0; v < NVW; v++) a.vc[v] = vw_offset(a.shift0, VW_P[v]);
    const int j0 = chunk * p.chunk;
    const int j1 = min(AP26_NN43, j0 + p.chunk);
    a.blk0 = (int64_t)j0 * AP26_N53_PER_N43;
    a.blk1 = (int64_t)j1 * AP26_N53_PER_N43;
    a.surv = p.surv;
    a.nsurv = p.nsurv;
    a.cap = p.cap;
    a.surv_job = p.surv_job;
    a.err = p.err;
    a.job = jl;
    QSieve2<NL, LV, R0, Q0, TU, GQ> qs;
    qs.lane = threadIdx.x & 63;
    qs.qb = L.queues[threadIdx.x >> 6];
    /* p.gq holds gridDim.x * WAVES regions of qsieve2_global_entries entries;
       the wave index is uniform, so the global base stays in scalar registers */
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    qs.gqb = p.gq + ((size_t)blockIdx.x * WAVES + wave) * qsieve2_global_entries<LV, GQ>();
    for (int i = 0; i < QSieve2<NL, LV, R0, Q0, TU, GQ>::NQ; i++) qs.qn[i] = 0;
    qs.seen_exec = ~UINT64_C(0);
    qs.ldsb = (const char *)L.tab;
    qs.gtb = (const char *)tab;
    qs.a = &a;
    qs.cur = &sc;
    qsieve2_blocks<NL, LV, V, G, R0, Q0, TU, GQ, 1>(qs, a, threadIdx.x, blockDim.x, &rc, &s_slab);
  }
}

/* One launch per (K, shift pair): the plain path with the shift-pair sieve
   (tables [entry][shift in pair], level 0 with one 16-byte read per prime for
   both shifts), survivors in per-wave slot chunks as in the batched engine. */
template <int NL, int WAVES, class LV, int V = 0, int G = 0, int R0 = 0, class TU = Tune<>,
          int GQ = NO_GLOBAL_QUEUES>
static __global__ void __launch_bounds__(64 * WAVES) __attribute__((amdgpu_waves_per_eu(WAVES / 4, (WAVES + 3) / 4)))
sieve_queue2_kernel(SieveArgs a) {
  constexpr int Q0 = R0 ? 192 : 128;
  static_assert(qsieve2_lds_bytes<NL, WAVES, LV, R0, Q0, GQ>() + 16 * WAVES <= 160 * 1024, "LDS budget");
  static_assert(LV::B[1] <= NL, "level 0 must be in LDS");
  static_assert(R0 == 0 || ((R0 == ROT_N || R0 == ROT_WIDE) && V == 0 && G == 0),
                "rotation group replaces V/G level-0 forms");
  __shared__ QSieve2Lds<NL, WAVES, LV, R0, Q0, GQ> L;
  __shared__ SurvCursor s_cur[WAVES];
  stage_pair_tables<R0, NL>(L.tab, a.tab);
  SurvCursor &sc = s_cur[threadIdx.x >> 6];
  sc = SurvCursor{0, 0};
  RotConstsT<rot_row_bytes(R0)> rc;
  if constexpr (R0) rc.init(threadIdx.x & 63);
  __syncthreads();
  QSieve2<NL, LV, R0, Q0, TU, GQ> qs;
  qs.lane = threadIdx.x & 63;
  qs.qb = L.queues[threadIdx.x >> 6];
  /* a.gq holds gridDim.x * WAVES regions of qsieve2_global_entries entries */
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  qs.gqb = (QEntry *)a.gq + ((size_t)blockIdx.x * WAVES + wave) * qsieve2_global_entries<LV, GQ>();
  for (int i = 0; i < QSieve2<NL, LV, R0, Q0, TU, GQ>::NQ; i++) qs.qn[i] = 0;
  qs.seen_exec = ~UINT64_C(0);
  qs.ldsb = (const char *)L.tab;
  qs.gtb = (const char *)a.tab;
  qs.a = &a;
  qs.cur = &sc;
  qsieve2_blocks<NL, LV, V, G, R0, Q0, TU, GQ, 4>(qs, a, (int64_t)blockIdx.x * blockDim.x + threadIdx.x,
                                              (int64_t)gridDim.x * blockDim.x, &rc, a.slab);
  close_survivors2(sc, a.surv, a.cap, threadIdx.x & 63);
}

/* ---- compacted walk ------------------------------------------------------ */
/* The walk of a survivor c counts base-2 SPRPs forward from c+5*STEP; only a
 * run of >= 10 extends backward from c+4*STEP while the term is positive, and
 * is reported as (k, m+STEP) (src/cpu/cpusse2.cpp:570-590,
 * src/opencl/kernels/checkn.cl:160-202).  About 72 % of runs stop at the first
 * term, so one thread per survivor leaves most lanes idle while a few runs go
 * on.  Here runs wait in a per-wave LDS queue and every pass tests one term
 * on 64 dense lanes: runs that go on are queued again, finished ones are
 * closed (backward extension and report, rare, inline). */
struct WEntry {
  int64_t c;   /* survivor */
  int64_t m;   /* next forward term to test */
  int32_t k;   /* SPRPs found forward so far */
  int32_t job; /* batched engine: job index (0 on the plain path) */
};

constexpr int WALK_QCAP = 128;

/* Walked terms at or past 2^62: the reference CPU build's PrimeQ exits on
   the first (src/common/PrimeQ_x86.h:160-166).  Only forward terms matter:
   the backward extension tests terms below the run's first forward term. */
constexpr uint64_t WALK_LIMIT = UINT64_C(1) << 62;

/* DEPTH: terms a lane tests in a row before a run that goes on is queued
   again (the compaction depth).  1 keeps every pass dense; larger depths
   trade idle lanes (runs that stop early in the pass) for fewer queue
   round trips.  The terms tested, and so every result, are the same. */
constexpr int WALK_DEPTHS[] = {1, 2, 4};

template <class StepOf, class Close, class Big, int DEPTH = 1>
struct WalkQueue {
  WEntry *q; /* this wave's WALK_QCAP entries */
  int qn = 0;
  int lane;
  uint64_t seen_exec = ~UINT64_C(0); /* AND of the exec masks of every push (all ones
                                        unless a push ran on part of the wave) */
  unsigned long long tests = 0;
  StepOf step_of;
  Close close;
  Big big; /* called with the job of a walked term >= WALK_LIMIT */

  __device__ __forceinline__ void push(const WEntry &e, bool pred) {
    const uint64_t bal = __ballot(pred);
    const int pre = __popcll(bal & ((UINT64_C(1) << lane) - 1));
    if (pred) q[qn + pre] = e;
    qn += __popcll(bal);
    seen_exec &= __builtin_amdgcn_read_exec();
    __builtin_amdgcn_wave_barrier();
  }

  /* test up to DEPTH terms of each lane's run (pred lanes), requeue runs
     that go on */
  __device__ __forceinline__ void step(WEntry e, bool pred) {
    bool go = pred;
    for (int d = 0; d < DEPTH; d++) {
      if (go) {
        tests++;
        if ((uint64_t)e.m >= WALK_LIMIT) big(e.job);
        go = sprp2((uint64_t)e.m);
        if (go) {
          e.k++;
          e.m += step_of(e.job);
        }
      }
    }
    push(e, go);
    if (pred && !go && e.k >= 10) close(e, tests);
  }

  __device__ __forceinline__ void drain() {
    while (qn >= 64) {
      qn -= 64;
      const WEntry e = q[qn + lane];
      __builtin_amdgcn_wave_barrier();
      step(e, true);
    }
  }

  __device__ __forceinline__ void flush() {
    while (qn > 0) {
      const int cnt = qn < 64 ? qn : 64;
      qn -= cnt;
      const bool pred = lane < cnt;
      WEntry e{0, 0, 0, 0};
      if (pred) e = q[qn + lane];
      __builtin_amdgcn_wave_barrier();
      step(e, pred);
    }
  }
};

/* backward extension of a forward run of k >= 10 ending before e.m; returns
   k and the first term */
__device__ __forceinline__ int walk_back(const WEntry &e, int64_t STEP, int64_t *first, unsigned long long &tests) {
  int k = e.k;
  int64_t m = e.c + 4 * STEP;
  while (m > 0) {
    tests++;
    if (!sprp2((uint64_t)m)) break;
    k++;
    m -= STEP;
  }
  *first = m + STEP;
  return k;
}

struct BatchWalkArgs {
  const BatchJobDev *__restrict__ jobs;
  const RoundState *__restrict__ rs;
  const int64_t *__restrict__ surv;
  const int32_t *__restrict__ surv_job;
  const unsigned long long *__restrict__ nsurv;
  int64_t cap;
  DevSol *__restrict__ sols;
  unsigned int *__restrict__ nsol;
  int32_t sol_cap;
  unsigned long long *__restrict__ ntests;
  unsigned long long *__restrict__ hash_tot; /* per range: exact sum of first%1000 + k */
  unsigned long long *__restrict__ overflow; /* largest survivor count beyond cap */
  unsigned long long *__restrict__ nlive;    /* survivors walked, over all rounds */
  unsigned int *__restrict__ err;            /* a queue step ran on part of a wave */
  unsigned long long *__restrict__ big;      /* per range: walked terms >= 2^62 */
};

template <int DEPTH>
static __global__ void __launch_bounds__(256) batch_walk_kernel(BatchWalkArgs a) {
  __shared__ WEntry s_q[4][WALK_QCAP];
  const RoundState rs = *a.rs;
  unsigned long long n = *a.nsurv;
  if (n > (unsigned long long)a.cap) {
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicMax(a.overflow, n);
    n = (unsigned long long)a.cap;
  }
  const BatchJobDev *jobs = a.jobs + rs.base;
  auto step_of = [&](int32_t job) { return jobs[job].kp.STEP; };
  auto close = [&](const WEntry &e, unsigned long long &tests) {
    const BatchJobDev &job = jobs[e.job];
    int64_t first;
    const int k = walk_back(e, job.kp.STEP, &first, tests);
    if (k >= 10) {
      atomicAdd(&a.hash_tot[job.range_id], (unsigned long long)(first % 1000 + k));
      const unsigned int idx = atomicAdd(a.nsol, 1u);
      if (idx < (unsigned int)a.sol_cap) {
        DevSol s;
        s.k = k;
        s.K = job.kp.K;
        s.first = first;
        s.range_id = job.range_id;
        s.shift_iter = (int32_t)((e.c / AP26_MOD - job.kp.SHIFT) / 64);
        s.prp2 = s.valid = 0;
        a.sols[idx] = s;
      }
    }
  };
  auto big = [&](int32_t job) { atomicAdd(&a.big[jobs[job].range_id], 1ull); };
  WalkQueue<decltype(step_of), decltype(close), decltype(big), DEPTH> wq{
      s_q[threadIdx.x >> 6], 0, (int)(threadIdx.x & 63), ~UINT64_C(0), 0, step_of, close, big};
  unsigned long long live = 0;
  const unsigned long long wave = ((unsigned long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const unsigned long long nwaves = ((unsigned long long)gridDim.x * blockDim.x) >> 6;
  /* uniform loops: the queue needs the whole wave in every step */
  for (unsigned long long base = wave * 64; base < n; base += nwaves * 64) {
    const unsigned long long i = base + (unsigned long long)wq.lane;
    const int64_t c = i < n ? a.surv[i] : -1;
    const bool has = c >= 0; /* -1: slot reserved but unused */
    WEntry e{0, 0, 0, 0};
    if (has) {
      e.job = a.surv_job[i];
      e.c = c;
      e.m = c + 5 * jobs[e.job].kp.STEP;
      live++;
    }
    wq.step(e, has);
    wq.drain();
  }
  wq.flush();
  if (wq.seen_exec != ~UINT64_C(0) && wq.lane == 0) atomicOr(a.err, 1u);
  if (wq.tests) atomicAdd(a.ntests, wq.tests);
  if (live) atomicAdd(a.nlive, live);
}

/* End of round: clear the per-round counters, advance to the next round. */
static __global__ void round_advance_kernel(RoundState *rs, unsigned long long *nsurv, unsigned int *work) {
  if (threadIdx.x == 0) {
    rs->base += rs->per_round;
    *nsurv = 0;
    *work = 0;
  }
}

/* ---- walk -------------------------------------------------------------- */
/* For each survivor n: count base-2 SPRPs forward from n+5*STEP; only if the
 * count reaches 10, extend backward from n+4*STEP while the term is positive;
 * report (k, m+STEP) if k >= 10 (src/cpu/cpusse2.cpp:570-590,
 * src/opencl/kernels/checkn.cl:160-202). */
struct WalkArgs {
  const int64_t *__restrict__ surv;
  const unsigned long long *__restrict__ nsurv;
  int64_t cap;
  int64_t STEP;
  int32_t K, SHIFT, range_id;
  DevSol *__restrict__ sols;
  unsigned int *__restrict__ nsol;
  int32_t sol_cap;
  unsigned long long *__restrict__ ntests;
  unsigned int *__restrict__ err; /* a queue step ran on part of a wave */
  unsigned long long *__restrict__ nlive; /* survivors walked */
  unsigned long long *__restrict__ big;   /* walked terms >= 2^62 */
};

template <int DEPTH>
static __global__ void __launch_bounds__(256) walk_kernel(WalkArgs a) {
  __shared__ WEntry s_q[4][WALK_QCAP];
  unsigned long long n = *a.nsurv;
  if (n > (unsigned long long)a.cap) n = (unsigned long long)a.cap;
  const int64_t STEP = a.STEP;
  auto step_of = [&](int32_t) { return STEP; };
  auto close = [&](const WEntry &e, unsigned long long &tests) {
    int64_t first;
    const int k = walk_back(e, STEP, &first, tests);
    if (k >= 10) {
      const unsigned int idx = atomicAdd(a.nsol, 1u);
      if (idx < (unsigned int)a.sol_cap) {
        DevSol s;
        s.k = k;
        s.K = a.K;
        s.first = first;
        s.range_id = a.range_id;
        s.shift_iter = (int32_t)((e.c / AP26_MOD - a.SHIFT) / 64);
        s.prp2 = s.valid = 0;
        a.sols[idx] = s;
      }
    }
  };
  unsigned long long nbig = 0;
  auto big = [&](int32_t) { nbig++; };
  WalkQueue<decltype(step_of), decltype(close), decltype(big), DEPTH> wq{
      s_q[threadIdx.x >> 6], 0, (int)(threadIdx.x & 63), ~UINT64_C(0), 0, step_of, close, big};
  const unsigned long long wave = ((unsigned long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const unsigned long long nwaves = ((unsigned long long)gridDim.x * blockDim.x) >> 6;
  unsigned long long live = 0;
  for (unsigned long long base = wave * 64; base < n; base += nwaves * 64) {
    const unsigned long long i = base + (unsigned long long)wq.lane;
    const int64_t c = i < n ? a.surv[i] : -1;
    const bool has = c >= 0; /* -1: slot reserved but unused */
    WEntry e{0, 0, 0, 0};
    if (has) {
      e.c = c;
      e.m = c + 5 * STEP;
      live++;
    }
    wq.step(e, has);
    wq.drain();
  }
  wq.flush();
  if (wq.seen_exec != ~UINT64_C(0) && wq.lane == 0) atomicOr(a.err, 1u);
  if (wq.tests) atomicAdd(a.ntests, wq.tests);
  if (live) atomicAdd(a.nlive, live);
  if (nbig) atomicAdd(a.big, nbig);
}

/* ---- batched SPRP ------------------------------------------------------ */
static __global__ void __launch_bounds__(256) prp_kernel(const uint64_t *__restrict__ n, uint8_t *__restrict__ out,
                                                   uint64_t count) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
       i += (uint64_t)gridDim.x * blockDim.x)
    out[i] = sprp2(n[i]) ? 1 : 0;
}

/* Clear counters with a kernel (not a memset node), so the sequence can be
   captured into a graph. */
/* 9-base verification of reports (validate_ap26 / val_base2_ap26,
   src/common/AP26.cpp:484-543): per report, bit j of prp2 says term
   first + j*STEP is a base-2 strong probable prime, bit j of valid that it is
   one to all of 2, 3, 5, 7, 11, 13, 17, 19, 23.  The host runs
   ReportSolution's Non-Solution recursion on these bits (terms 0..31).
   One thread per (report, term, base): verify_kernel ORs the base-2 passes
   into prp2 and the other bases' failures into valid (both start at 0), and
   verify_finish_kernel turns valid into prp2 & ~failures. */
static __global__ void __launch_bounds__(256) verify_kernel(DevSol *__restrict__ sols,
                                                            const unsigned int *__restrict__ nsol, int32_t cap) {
  unsigned int n = *nsol;
  if (n > (unsigned int)cap) n = (unsigned int)cap;
  const uint64_t total = (uint64_t)n * 32 * 9;
  for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (uint64_t)gridDim.x * blockDim.x) {
    const unsigned int i = (unsigned int)(t / (32 * 9));
    const int r = (int)(t - (uint64_t)i * (32 * 9));
    const int j = r / 9, bi = r - 9 * (r / 9);
    const DevSol &s = sols[i];
    if (j >= s.k || j >= 32) continue;
    const int64_t N = s.first + (int64_t)j * ((int64_t)s.K * AP26_PRIM23);
    if (N < 3) continue;
    constexpr uint32_t B[9] = {2, 3, 5, 7, 11, 13, 17, 19, 23};
    const bool pass = sprp_base((uint64_t)N, B[bi]);
    if (bi == 0) {
      if (pass) atomicOr(&sols[i].prp2, 1u << j);
    } else if (!pass) {
      atomicOr(&sols[i].valid, 1u << j);
    }
  }
}

static __global__ void __launch_bounds__(256) verify_finish_kernel(DevSol *__restrict__ sols,
                                                                   const unsigned int *__restrict__ nsol,
                                                                   int32_t cap) {
  unsigned int n = *nsol;
  if (n > (unsigned int)cap) n = (unsigned int)cap;
  for (unsigned int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    sols[i].valid = sols[i].prp2 & ~sols[i].valid;
}

static __global__ void zero_kernel(unsigned long long *p, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

}  // namespace ap26

#endif
