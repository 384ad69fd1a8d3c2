/* ap26_pvariants_4.hip -- part 4 of the batched engine's sieve variant
 * table (ap26_pvariants.h).
 */
#include <hip/hip_runtime.h>

#include "ap26_pvariants.h"

namespace ap26 {

static PersistVariant part(int id) {
  switch (id) {
    case 91: return make_pvariant2rg<37, 12, Levels<0, 21, 33, 51, 83>, Tune<2, 3, 1, 1>, 2>();
    /* 92..97: 512-byte rotation rows; 16 waves where the registers allow */
    case 92: return make_pvariant2rw<30, 12, Levels<0, 20, 30, 51, 83>, Tune<2, 3, 1, 1>, 1>();
    case 93: return make_pvariant2rw<30, 16, Levels<0, 20, 30, 51, 83>, Tune<2, 3, 1, 2>, 1>();
    case 94: return make_pvariant2rw<30, 16, Levels<0, 20, 30, 51, 83>, Tune<2, 3, 1, 1>, 1>();
    case 95: return make_pvariant2rw<34, 12, Levels<0, 20, 33, 51, 83>, Tune<2, 3, 1, 1>, 1>();
    case 96: return make_pvariant2rw<31, 16, Levels<0, 20, 31, 51, 83>, Tune<2, 3, 1, 4>, 1>();
    case 97: return make_pvariant2rw<27, 12, Levels<0, 20, 27, 51, 83>, Tune<2, 3, 1, 1>, 2>();
    case 98: return make_pvariant2rw<34, 12, Levels<0, 20, 34, 51, 83>, Tune<2, 3, 1, 1>, 1>();
    case 99: return make_pvariant2rw<34, 12, Levels<0, 20, 32, 51, 83>, Tune<2, 3, 1, 1>, 1>();
    case 100: return make_pvariant2rw<34, 12, Levels<0, 20, 34, 55, 83>, Tune<2, 3, 1, 1>, 1>();
    case 101: return make_pvariant2rw<34, 12, Levels<0, 20, 33, 47, 83>, Tune<2, 3, 1, 1>, 1>();
    case 102: return make_pvariant2rw<34, 12, Levels<0, 20, 30, 42, 83>, Tune<2, 3, 1, 1>, 1>();
    /* 103, 104: 95 with the last 4 or 2 level-0 primes read through L1 */
    case 103: return make_pvariant2rw<34, 12, Levels<0, 20, 33, 51, 83>, TuneL1<2, 3, 1, 1, 4>, 1>();
    case 104: return make_pvariant2rw<34, 12, Levels<0, 20, 33, 51, 83>, TuneL1<2, 3, 1, 1, 2>, 1>();
    /* 105: 95 with the packed level-0 gathers on live lanes only */
    case 105: return make_pvariant2rw<34, 12, Levels<0, 20, 33, 51, 83>, TuneSkip<2, 3, 1, 1>, 1>();
    /* 106, 107: the same at 16 waves per CU (the registers now fit without
       spills in the position loop), fewer primes in LDS for the queues */
    case 106: return make_pvariant2rw<30, 16, Levels<0, 20, 30, 51, 83>, TuneSkip<2, 3, 1, 1>, 1>();
    case 107: return make_pvariant2rw<31, 16, Levels<0, 20, 31, 51, 83>, TuneSkip<2, 3, 1, 2>, 1>();
    case 108: return make_pvariant2rw<31, 16, Levels<0, 20, 31, 51, 83>, TuneSkip<2, 3, 1, 4>, 1>();
    case 109: return make_pvariant2rw<31, 16, Levels<0, 20, 31, 51, 83>, TuneSkip<1, 2, 0, 2>, 1>();
    /* 110..113: 95 with five levels (one more compaction, its queue in
       global memory) */
    case 110: return make_pvariant2rw<34, 12, Levels<0, 20, 26, 33, 51, 83>, Tune<2, 3, 1, 1>, 1>();
    case 111: return make_pvariant2rw<34, 12, Levels<0, 20, 28, 34, 51, 83>, Tune<2, 3, 1, 1>, 1>();
    case 112: return make_pvariant2rw<34, 12, Levels<0, 20, 33, 42, 51, 83>, Tune<2, 3, 1, 1>, 1>();
    case 113: return make_pvariant2rw<34, 12, Levels<0, 20, 33, 51, 66, 83>, Tune<2, 3, 1, 1>, 1>();
    /* 114..118: six and seven levels */
    case 114: return make_pvariant2rw<34, 12, Levels<0, 20, 28, 34, 51, 66, 83>, Tune<2, 3, 1, 1>, 1>();
    case 115: return make_pvariant2rw<34, 12, Levels<0, 20, 33, 42, 51, 66, 83>, Tune<2, 3, 1, 1>, 1>();
    default: return PersistVariant{0, 0, 0, nullptr};
  }
}

bool pick_pvariant_4(int id, PersistVariant &v) {
  v = part(id);
  return v.fn != nullptr;
}

}  // namespace ap26
