/* ap26_pvariants_5.hip -- part 5 of the batched engine's sieve variant
 * table (ap26_pvariants.h).
 */
#include <hip/hip_runtime.h>

#include "ap26_pvariants.h"

namespace ap26 {

static PersistVariant part(int id) {
  switch (id) {
    case 116: return make_pvariant2rw<34, 12, Levels<0, 20, 28, 34, 42, 51, 83>, Tune<2, 3, 1, 1>, 1>();
    case 117: return make_pvariant2rw<34, 12, Levels<0, 20, 28, 34, 42, 51, 66, 83>, Tune<2, 3, 1, 1>, 1>();
    case 118: return make_pvariant2rw<34, 12, Levels<0, 20, 30, 34, 51, 66, 83>, Tune<2, 3, 1, 1>, 1>();
    /* 119..123: eight to ten levels */
    case 119: return make_pvariant2rw<34, 12, Levels<0, 20, 28, 34, 42, 51, 60, 70, 83>, Tune<2, 3, 1, 1>, 1>();
    case 120: return make_pvariant2rw<34, 12, Levels<0, 20, 28, 34, 40, 46, 51, 60, 70, 83>, Tune<2, 3, 1, 1>, 1>();
    case 121: return make_pvariant2rw<34, 12, Levels<0, 20, 28, 34, 42, 51, 58, 66, 74, 83>, Tune<2, 3, 1, 1>, 1>();
    case 122: return make_pvariant2rw<34, 12, Levels<0, 20, 27, 31, 34, 42, 51, 66, 83>, Tune<2, 3, 1, 1>, 1>();
    case 123: return make_pvariant2rw<34, 12, Levels<0, 20, 28, 34, 40, 46, 53, 61, 71, 83>, Tune<2, 3, 1, 1>, 1>();
    /* 124..127: split points near 117's */
    case 124: return make_pvariant2rw<34, 12, Levels<0, 20, 29, 34, 42, 51, 66, 83>, Tune<2, 3, 1, 1>, 1>();
    case 125: return make_pvariant2rw<34, 12, Levels<0, 20, 28, 34, 43, 53, 68, 83>, Tune<2, 3, 1, 1>, 1>();
    case 126: return make_pvariant2rw<34, 12, Levels<0, 20, 28, 34, 41, 49, 63, 83>, Tune<2, 3, 1, 1>, 1>();
    case 127: return make_pvariant2rw<34, 12, Levels<0, 21, 28, 34, 42, 51, 66, 83>, Tune<2, 3, 1, 1>, 1>();
    /* 128..131: 117 with a level 0 of 19 primes, or other wave priorities */
    case 128: return make_pvariant2rw<34, 12, Levels<0, 19, 28, 34, 42, 51, 66, 83>, Tune<2, 3, 1, 1>, 1>();
    case 129: return make_pvariant2rw<34, 12, Levels<0, 20, 28, 34, 42, 51, 66, 83>, Tune<2, 3, 2, 1>, 1>();
    case 130: return make_pvariant2rw<34, 12, Levels<0, 20, 28, 34, 42, 51, 66, 83>, Tune<2, 3, 0, 1>, 1>();
    case 131: return make_pvariant2rw<34, 12, Levels<0, 20, 28, 34, 42, 51, 66, 83>, Tune<1, 3, 1, 1>, 1>();
    /* 132..134: 117 with the global queues from 3, 2 or 4 on non-temporal */
    case 132: return make_pvariant2rw<34, 12, Levels<0, 20, 28, 34, 42, 51, 66, 83>, TuneNT<2, 3, 1, 1, 3>, 1>();
    case 133: return make_pvariant2rw<34, 12, Levels<0, 20, 28, 34, 42, 51, 66, 83>, TuneNT<2, 3, 1, 1, 2>, 1>();
    case 134: return make_pvariant2rw<34, 12, Levels<0, 20, 28, 34, 42, 51, 66, 83>, TuneNT<2, 3, 1, 1, 4>, 1>();
    /* 135..137: level 0 on the 16 rotated primes only, the packed 139..157
       a level of their own (slower: twice the queue-0 traffic) */
    case 135: return make_pvariant2rw<34, 12, Levels<0, 16, 20, 28, 34, 42, 51, 66, 83>, Tune<2, 3, 1, 1>, 1>();
    case 136: return make_pvariant2rw<34, 12, Levels<0, 16, 20, 28, 34, 42, 51, 66, 83>, Tune<2, 3, 3, 1>, 1>();
    case 137: return make_pvariant2rw<34, 12, Levels<0, 16, 22, 28, 34, 42, 51, 66, 83>, Tune<2, 3, 1, 1>, 1>();
    /* 138, 139: seven levels with 27 primes in LDS, the level-1 queue in LDS
       (138: a third of 117's L2 write-backs, 2.6 % slower) or in global
       memory (139) */
    case 138: return make_pvariant2rw<27, 12, Levels<0, 20, 27, 34, 42, 51, 66, 83>, Tune<2, 3, 1, 1>, 2>();
    case 139: return make_pvariant2rw<27, 12, Levels<0, 20, 27, 34, 42, 51, 66, 83>, Tune<2, 3, 1, 1>, 1>();
    case 28: return make_pvariant2<29, 16, Levels<0, 18, 29, 51, 83>>();
    default: return PersistVariant{0, 0, 0, nullptr};
  }
}

bool pick_pvariant_5(int id, PersistVariant &v) {
  v = part(id);
  return v.fn != nullptr;
}

}  // namespace ap26
