/* ap26_pvariants.h -- the batched engine's persistent sieve variants.
 * The table of variants is split over ap26_pvariants_1..5.hip, so that their
 * kernel instantiations compile in parallel; pick_pvariant (ap26_batch.hip)
 * asks each part in turn.
 */
#pragma once

#include "ap26_kernels.h"

namespace ap26 {

using PersistFn = void (*)(PersistArgs);

struct PersistVariant {
  int NL, waves, S;
  PersistFn fn;
  int gq_entries = 0; /* global queue entries per wave (PersistArgs::gq) */
};

/* 1..3, 9: one shift per item; 10..40, 50..56: shift pairs.  Level 0
   (every position, both shifts of a pair per 16-byte read) runs on the first
   16..20 primes: a pair's 128-bit mask stays non-zero longer than one shift's,
   so level 0 pays off further than on one shift (61..131).  35..38: the first
   1..3 level-0 primes from register bit windows; 39..40: the last 2 or 4
   level-0 primes read through L1 instead of LDS.  50..91: level 0 opens
   with the bank-rotated group of 61..137 (level0_rot), whose 64 VGPRs of
   per-lane constants leave room for 12 waves per CU (16 waves spill in the
   hot loop, 8 lose the latency hiding); 57..74 vary the level boundaries,
   the wave priorities and how many rotated gathers are in flight at once
   (Tune).  75..91 move the deeper queues to global memory and spend the
   LDS on more table primes: a level that reads its tables through L1/L2
   costs more than the queue traffic (85: the last queue in global memory,
   primes 61..251 in LDS).  92..102 use the 512-byte rotation rows
   (ROT_WIDE: no per-lane column registers or adds, twice the LDS; 95:
   queues 1 and 2 in global memory, primes 61..233 in LDS); 16 waves still
   spill the rotation constants in the hot loop.  103, 104: 95 with the
   last 4 or 2 level-0 primes read through L1 (slower: the random
   vector-memory gathers cost more than their LDS bank conflicts).
   110..127: 95 with more compaction levels (queues 1.. in global memory):
   an entry whose mask has died leaves before the next run of lookups, so
   each extra level saves lookups for the price of one queue step; 0 = 117,
   the default: seven levels, 61..157 | 163..197 | 199..233 | 239..277 |
   281..337 | 347..431 | 433..541. */
bool pick_pvariant_1(int id, PersistVariant &v);
bool pick_pvariant_2(int id, PersistVariant &v);
bool pick_pvariant_3(int id, PersistVariant &v);
bool pick_pvariant_4(int id, PersistVariant &v);
bool pick_pvariant_5(int id, PersistVariant &v);

/* The kernels are file-local (static), so each part builds its own variants
   with these helpers, kept file-local as well. */
namespace {

template <int NL, int WAVES, class LV>
PersistVariant make_pvariant() {
  return PersistVariant{NL, WAVES, 1, &sieve_persistent_kernel<NL, WAVES, LV>};
}

template <int NL, int WAVES, class LV, int V = 0, int G = 0>
PersistVariant make_pvariant2() {
  return PersistVariant{NL, WAVES, 2, &sieve_persistent2_kernel<NL, WAVES, LV, V, G>};
}

/* shift pairs with the bank-rotated level-0 group (primes 61..137) */
template <int NL, int WAVES, class LV, class TU = Tune<>>
PersistVariant make_pvariant2r() {
  return PersistVariant{NL, WAVES, 2, &sieve_persistent2_kernel<NL, WAVES, LV, 0, 0, ROT_N, TU>};
}

/* the rotated form with 512-byte rows (no per-lane column registers) */
template <int NL, int WAVES, class LV, class TU, int GQ>
PersistVariant make_pvariant2rw() {
  return PersistVariant{NL, WAVES, 2, &sieve_persistent2_kernel<NL, WAVES, LV, 0, 0, ROT_WIDE, TU, GQ>,
                        qsieve2_global_entries<LV, GQ>()};
}

/* the rotated form with queues GQ.. in global memory, so that more primes fit in LDS */
template <int NL, int WAVES, class LV, class TU, int GQ>
PersistVariant make_pvariant2rg() {
  return PersistVariant{NL, WAVES, 2, &sieve_persistent2_kernel<NL, WAVES, LV, 0, 0, ROT_N, TU, GQ>,
                        qsieve2_global_entries<LV, GQ>()};
}

}  // namespace

}  // namespace ap26
