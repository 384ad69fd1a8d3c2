/* ap26_device.h -- device-side building blocks of the AP26 search on gfx950:
 * compile-time prime tables, the fp64 residue step, 64-bit Montgomery
 * arithmetic and the base-2 strong probable-prime test.
 */
#ifndef AP26_DEVICE_H
#define AP26_DEVICE_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ap26_consts.h"

namespace ap26 {

/* ---- prime tables ------------------------------------------------------ */

constexpr int NSP = 83;           /* sieve primes 61..541 */
constexpr int NFP = 6;            /* filter primes 7, 11, 13, 17, 19, 23 */
constexpr int NPT = NSP + NFP;    /* tables per shift */

struct PrimeTable {
  int p[NPT];
  int off[NPT + 1]; /* packed offsets; off[NPT] = total entries */
};

constexpr bool is_prime_c(int n) {
  if (n < 2) return false;
  for (int d = 2; d * d <= n; d++)
    if (n % d == 0) return false;
  return true;
}

constexpr PrimeTable make_prime_table() {
  PrimeTable t{};
  int n = 0, off = 0;
  for (int p = 61; p <= 541; p++)
    if (is_prime_c(p)) { t.p[n] = p; t.off[n] = off; off += p; n++; }
  const int fp[NFP] = {7, 11, 13, 17, 19, 23};
  for (int i = 0; i < NFP; i++) { t.p[n] = fp[i]; t.off[n] = off; off += fp[i]; n++; }
  t.off[n] = off;
  return t;
}

constexpr PrimeTable PT = make_prime_table();
constexpr int NOK = 23693;        /* sieve-table entries (sum of 61..541) */
constexpr int NTAB = NOK + 90;    /* plus the 7..23 filter tables */
static_assert(PT.off[NSP] == NOK, "sieve table size");
static_assert(PT.off[NPT] == NTAB, "table size");
static_assert(PT.p[NSP - 1] == 541, "last sieve prime");

/* ---- residues ---------------------------------------------------------- */
/* A position n59 < MOD < 2^48 is carried as the double xd = n59 + 2^52, whose
 * low 32 mantissa bits are the low 32 bits of n59.  For an odd modulus p,
 *   q = rint(n59 / p)   from one fma against RN(1/p) and the 1.5*2^52 shifter,
 *   r = n59 - q*p       from the low 24 bits (one i24 multiply-add),
 * and |r| <= (p-1)/2: the quotient error is at most n59*2^-53 <= 2^-5, so
 * |n59/p - q| <= 1/2 + 2^-5/p and, r being an integer with p odd, r lies in
 * [-(p-1)/2, (p-1)/2].  Tables are stored "centred": entry c holds residue
 * c - (p-1)/2 (mod p), so the index c = r + (p-1)/2 needs no correction. */

constexpr double TWO52 = 4503599627370496.0;
constexpr double SHIFTER = 6755399441055744.0; /* 1.5 * 2^52 */

__device__ __forceinline__ uint32_t lo32(double d) {
  return (uint32_t)__double_as_longlong(d);
}

/* Byte offset SCALE*r of the centred entry of n59 relative to the entry of
 * residue 0 (SCALE = bytes per centred entry), given x = n59 as a double and
 * xs = SCALE * (low 32 bits of n59).  The multiply-add runs on the full-rate
 * 24-bit multiplier: it is exact modulo SCALE*2^24 (a power of two >= 2^27
 * for SCALE >= 8), and |SCALE*r| < 2^26 (SCALE <= 80, |r| <= 270), so
 * sign-extending 27 bits recovers SCALE*r. */
template <int P, int SCALE>
__device__ __forceinline__ int centred_byteoff(double x, uint32_t xs) {
  static_assert(SCALE * P < (1 << 23), "i24 operand");
  static_assert(SCALE >= 8 && (SCALE & (SCALE - 1)) == 0, "exact modulo 2^27 needs SCALE >= 8");
  constexpr double inv = 1.0 / (double)P;
  const double qd = __fma_rn(x, inv, SHIFTER);
  const int q = (int)lo32(qd);
  int r;
  asm("v_mad_i32_i24 %0, %1, %2, %3" : "=v"(r) : "v"(q), "s"(-(SCALE * P)), "v"(xs));
  return (r << 5) >> 5;
}

/* ---- 64-bit Montgomery arithmetic (R = 2^64), n odd -------------------- */

struct Mont {
  uint64_t n, ninv; /* ninv = -n^-1 mod 2^64 */
};

__device__ __forceinline__ uint64_t neg_inverse(uint64_t n) {
  uint64_t x = n; /* correct to 3 bits: n*n == 1 mod 8 */
  x *= 2 - n * x;
  x *= 2 - n * x;
  x *= 2 - n * x;
  x *= 2 - n * x;
  x *= 2 - n * x; /* 96 >= 64 bits */
  return (uint64_t)0 - x;
}

/* (a*b + m*n) / 2^64 with m = a*b*ninv mod 2^64; a, b < n; result < n. */
__device__ __forceinline__ uint64_t mont_mul(uint64_t a, uint64_t b, const Mont &M) {
  uint64_t tlo = a * b;
  uint64_t thi = __umul64hi(a, b);
  uint64_t m = tlo * M.ninv;
  uint64_t mhi = __umul64hi(m, M.n);
  /* tlo + lo(m*n) == 0 mod 2^64: it carries iff tlo != 0 */
  uint64_t s = thi + mhi;
  uint64_t c1 = s < thi;               /* carry out of bit 63 (only if n > 2^63) */
  uint64_t u = s + (tlo != 0);
  c1 |= (u < s);
  if (c1 || u >= M.n) u -= M.n;
  return u;
}

/* 2x mod n for x < n */
__device__ __forceinline__ uint64_t dbl_mod(uint64_t x, uint64_t n) {
  uint64_t t = n - x;
  return x >= t ? x - t : x + x;
}

/* Montgomery squaring for odd n < 2^63, times 2^s (s = 0 or 1):
   returns 2^s * a^2 / 2^64 mod n for a < n, with ninv = n^-1 mod 2^64.
   Since a < 2^63, a1 = a >> 32 < 2^31 and T = a^2 is three 32x32->64
   multiply-adds with no carries: t = a0*(2*a1) + hi32(a0^2) < 2^64 and
   hi(T) = a1^2 + hi32(t).  T * 2^s < 2 n^2, so hi(T * 2^s) < n when
   n < 2^63.  Reduction: m = lo(T) * ninv, so m*n and T agree in the low
   64 bits and (T - m*n) / 2^64 = hi(T) - hi(m*n) exactly, a value in
   (-n, n) that one conditional add of n brings into [0, n). */
__device__ __forceinline__ uint64_t mont_sqr63(uint64_t a, uint32_t s, uint64_t n,
                                               uint64_t ninv) {
  const uint32_t a0 = (uint32_t)a, a1 = (uint32_t)(a >> 32);
  const uint64_t p0 = (uint64_t)a0 * a0;
  const uint64_t t = (uint64_t)a0 * (a1 << 1) + (p0 >> 32);
  uint64_t hi = (uint64_t)a1 * a1 + (t >> 32);
  uint64_t lo = (t << 32) | (uint32_t)p0;
  hi = (hi << s) | ((lo >> 63) & s);
  lo <<= s;
  const uint64_t m = lo * ninv;
  const uint64_t h = __umul64hi(m, n);
  const uint64_t u = hi - h;
  return hi < h ? u + n : u;
}

/* Base-2 strong probable-prime test (textbook Miller-Rabin, base 2), the
 * predicate of PrimeQ (src/common/PrimeQ_x86.h:114-169): n odd >= 3 with
 * n-1 = d*2^t, d odd, passes iff 2^d == 1 or 2^(d*2^s) == -1 (mod n) for some
 * 0 <= s < t.  Valid for every odd n < 2^64.  Left-to-right exponentiation in
 * Montgomery form: one squaring per bit of d after the first, a modular
 * doubling for each set bit (base 2).  n < 2^63 (every search term and the
 * micro-benchmark's range) takes the squaring above; larger n the general
 * multiply with its carry handling. */
__device__ __forceinline__ bool sprp2(uint64_t n) {
  if (n < 3 || !(n & 1)) return false;
  const uint64_t ninv = neg_inverse(n);
  uint64_t d = n - 1;
  const int t = __builtin_ctzll(d);
  d >>= t;
  const int bl = 64 - __builtin_clzll(n);
  /* x = 2^65 mod n = Montgomery form of 2, by doubling 2^(bl-1) < n */
  uint64_t x = (uint64_t)1 << (bl - 1);
  for (int i = 0; i < 66 - bl; i++) x = dbl_mod(x, n);
  /* one = R mod n = x/2 mod n */
  const uint64_t one = (x & 1) ? (x >> 1) + (n >> 1) + 1 : (x >> 1);
  const uint64_t mone = n - one;
  const int L = 64 - __builtin_clzll(d);
  if (n < (UINT64_C(1) << 63)) {
    const uint64_t ninvp = (uint64_t)0 - ninv;
    /* the bits of d below its leading one, from the top, at bit 63 */
    uint64_t dd = L > 1 ? d << (65 - L) : 0;
    for (int b = L - 2; b >= 0; b--) {
      x = mont_sqr63(x, (uint32_t)(dd >> 63), n, ninvp);
      dd <<= 1;
    }
    if (x == one || x == mone) return true;
    for (int s = 1; s < t; s++) {
      x = mont_sqr63(x, 0, n, ninvp);
      if (x == mone) return true;
      if (x == one) return false;
    }
    return false;
  }
  const Mont M{n, ninv};
  for (int b = L - 2; b >= 0; b--) {
    x = mont_mul(x, x, M);
    if ((d >> b) & 1) x = dbl_mod(x, n);
  }
  if (x == one || x == mone) return true;
  for (int s = 1; s < t; s++) {
    x = mont_mul(x, x, M);
    if (x == mone) return true;
    if (x == one) return false;
  }
  return false;
}

/* Strong probable-prime test to a small base a (2 <= a < 2^32), textbook
 * Miller-Rabin as strong_prp(a, N) (src/common/PrimeQ_x86.h:114-153): n odd
 * >= 3, n-1 = d*2^t; true iff a^d == 1 or a^(d*2^s) == -1 (mod n) for some
 * 0 <= s < t; a == 0 (mod n) counts as passing, as in the host restatement.
 * Montgomery form throughout (general multiply; verification is rare). */
__device__ __noinline__ bool sprp_base(uint64_t n, uint32_t a) {
  if (n < 3 || !(n & 1)) return false;
  const uint64_t ar = (uint64_t)a % n;
  if (ar == 0) return true;
  const Mont M{n, neg_inverse(n)};
  uint64_t d = n - 1;
  const int t = __builtin_ctzll(d);
  d >>= t;
  const int bl = 64 - __builtin_clzll(n);
  uint64_t x2 = (uint64_t)1 << (bl - 1); /* 2^65 mod n by doubling */
  for (int i = 0; i < 66 - bl; i++) x2 = dbl_mod(x2, n);
  const uint64_t one = (x2 & 1) ? (x2 >> 1) + (n >> 1) + 1 : (x2 >> 1);
  const uint64_t mone = n - one;
  /* a*R mod n by double-and-add over the bits of a */
  uint64_t am = 0;
  for (int i = 31 - __builtin_clz((uint32_t)ar); i >= 0; i--) {
    am = dbl_mod(am, n);
    if ((ar >> i) & 1) am = am >= n - one ? am - (n - one) : am + one;
  }
  uint64_t x = am;
  for (int i = 62 - __builtin_clzll(d); i >= 0; i--) {
    x = mont_mul(x, x, M);
    if ((d >> i) & 1) x = mont_mul(x, am, M);
  }
  if (x == one || x == mone) return true;
  for (int s = 1; s < t; s++) {
    x = mont_mul(x, x, M);
    if (x == mone) return true;
    if (x == one) return false;
  }
  return false;
}

}  // namespace ap26

#endif
