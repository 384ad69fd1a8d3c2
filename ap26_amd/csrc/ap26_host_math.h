/* ap26_host_math.h -- per-K constants and n43 starts, computed on the host
 * (C and C++), and the K admission rule.
 */
#ifndef AP26_HOST_MATH_H
#define AP26_HOST_MATH_H

#include <stdint.h>

#include "ap26_consts.h"

typedef struct {
  int32_t K;
  int64_t STEP;
  int64_t n0;
  int64_t S[7]; /* S31, S37, S41, S43, S47, S53, S59 */
} HostK;

/* K is searched iff no prime 29..59 divides it (src/common/AP26.cpp:649-653). */
static inline int host_will_search(int32_t K) {
  return (K % 29) && (K % 31) && (K % 37) && (K % 41) && (K % 43) && (K % 47) && (K % 53) &&
         (K % 59);
}

static inline int64_t host_mulmod_mod(int64_t a, int64_t k) {
  return (int64_t)(((__int128)a * k) % AP26_MOD);
}

/* STEP = K*23#, n0 = (N0*K + N30) mod MOD, S_q = PRES_q*K mod MOD
   (src/opencl/AP26.h:112-120).  Exact 128-bit products; the reference's
   K = 17835q + r split yields the same values for K below ~6.3e8. */
static inline void host_kconsts(int32_t K, HostK *hk) {
  static const int64_t pres[7] = {AP26_PRES31, AP26_PRES37, AP26_PRES41, AP26_PRES43,
                                  AP26_PRES47, AP26_PRES53, AP26_PRES59};
  hk->K = K;
  hk->STEP = (int64_t)K * AP26_PRIM23;
  hk->n0 = (host_mulmod_mod(AP26_N0, K) + AP26_N30) % AP26_MOD;
  for (int i = 0; i < 7; i++) hk->S[i] = host_mulmod_mod(pres[i], K);
}

/* The 10840 n43 starts in the reference's order (src/opencl/AP26.h:125-136):
   i31 < 7, i37 < 13, i41 < 17 under the three guards, then i3 < 2, i5 < 4. */
static inline void host_n43(const HostK *hk, int64_t *out) {
  int c = 0;
  for (int i31 = 0; i31 < 7; i31++)
    for (int i37 = 0; i37 < 13; i37++) {
      if (i37 - i31 > 10 || i31 - i37 > 4) continue;
      for (int i41 = 0; i41 < 17; i41++) {
        if (i41 - i31 > 14 || i41 - i37 > 14 || i31 - i41 > 4 || i37 - i41 > 10) continue;
        const int64_t base = hk->n0 + i31 * hk->S[0] + i37 * hk->S[1] + i41 * hk->S[2];
        for (int i3 = 0; i3 < 2; i3++)
          for (int i5 = 0; i5 < 4; i5++) out[c++] = (base + i3 * AP26_S3 + i5 * AP26_S5) % AP26_MOD;
      }
    }
}

/* Window string of a level-0 prime p computed in registers instead of read
   from a table: with M = MOD mod p, E_p[t] = OK_p[(t*M) mod p], and bit b
   (0..255) of out is E_p[b mod p].  For a position with residue r, bit b of
   the masks of shifts A, A+64 (b < 128) is OK_p[(r + (A+b)*M) mod p] =
   E_p[(r*M^-1 + A + b) mod p], a 128-bit window of out.  OK_p[v] = 0 iff
   (p - v)*STEP^-1 mod p <= 23, or v == 0 when p divides STEP (the tables'
   definition, src/cpu/cpusse2.cpp:31-43). */
static inline void host_vwin(int32_t K, int p, uint64_t out[4]) {
  const int sp = (int)(((int64_t)(K % p) * (AP26_PRIM23 % p)) % p);
  int inv = 0;
  if (sp) {
    int e = p - 2, b = sp;
    inv = 1;
    while (e) {
      if (e & 1) inv = inv * b % p;
      b = b * b % p;
      e >>= 1;
    }
  }
  const int M = (int)(AP26_MOD % p);
  for (int i = 0; i < 4; i++) out[i] = 0;
  for (int b = 0; b < 256; b++) {
    const int v = (b % p) * M % p;
    const int ok = sp ? ((p - v) * inv % p > 23) : (v != 0);
    out[b >> 6] |= (uint64_t)ok << (b & 63);
  }
}

/* Every term the search forms for K and SHIFT fits in int64: survivors are
   below (SHIFT + 640) * MOD, and an AP with difference STEP holds at most 29
   primes in a row (29 does not divide a searched K, so one term in 29 is a
   multiple of 29), so the walk never passes survivor + 34 * STEP. */
static inline int host_shift_ok(int32_t K, int32_t SHIFT) {
  if (K < 1 || SHIFT < 0) return 0;
  const __int128 top = ((__int128)SHIFT + 640) * AP26_MOD + (__int128)35 * K * AP26_PRIM23;
  return top < ((__int128)1 << 63);
}

/* The reference CPU build tests every walked term with PrimeQ, which exits
   on any N >= 2^62 (src/common/PrimeQ_x86.h:160-166, its x87 multiply
   limit).  A survivor's walk starts at survivor + 5*STEP; when that is past
   2^62 for the largest candidates, (SHIFT + 640)*MOD - 1 + 5*STEP >= 2^62,
   the reference stops at the first of them, so the search refuses such a K
   before any GPU work.  Below that bound, whether a walk reaches 2^62 depends
   on the data: the device counts such terms and the call fails with
   AP26_E_RANGE (the reference's exit) instead of reporting. */
#define AP26_WALK_LIMIT_BITS 62
static inline int host_walk_ok(int32_t K, int32_t SHIFT) {
  if (!host_shift_ok(K, SHIFT)) return 0;
  const __int128 first = ((__int128)SHIFT + 640) * AP26_MOD - 1 + (__int128)5 * K * AP26_PRIM23;
  return first < ((__int128)1 << AP26_WALK_LIMIT_BITS);
}

/* Survivor-buffer entries to hold for one K.  A sieve prime p (61..541)
   rejects 24 of its p residues, or only residue 0 when p divides K (then
   STEP = 0 mod p), so each such prime scales the expected survivor count by
   (p - 1) / (p - 24): 1.62 for 61, and 5.45 for K = 61*67*71*73, the most a
   K <= 6e8 can hold.  An ordinary K keeps 38-41 M survivors over its 10
   shifts; the budget is 48 Mi entries times that factor. */
static inline int64_t host_surv_budget(int32_t K) {
  static const int sp[] = {61,  67,  71,  73,  79,  83,  89,  97,  101, 103, 107, 109, 113, 127,
                           131, 137, 139, 149, 151, 157, 163, 167, 173, 179, 181, 191, 193, 197,
                           199, 211, 223, 227, 229, 233, 239, 241, 251, 257, 263, 269, 271, 277,
                           281, 283, 293, 307, 311, 313, 317, 331, 337, 347, 349, 353, 359, 367,
                           373, 379, 383, 389, 397, 401, 409, 419, 421, 431, 433, 439, 443, 449,
                           457, 461, 463, 467, 479, 487, 491, 499, 503, 509, 521, 523, 541};
  double w = 48.0 * (1 << 20);
  for (unsigned i = 0; i < sizeof(sp) / sizeof(sp[0]); i++)
    if (K % sp[i] == 0) w *= (double)(sp[i] - 1) / (double)(sp[i] - 24);
  return (int64_t)w;
}

#endif
