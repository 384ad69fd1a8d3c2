/* ap26_pvariants_3.hip -- part 3 of the batched engine's sieve variant
 * table (ap26_pvariants.h).
 */
#include <hip/hip_runtime.h>

#include "ap26_pvariants.h"

namespace ap26 {

static PersistVariant part(int id) {
  switch (id) {
    case 66: return make_pvariant2r<30, 12, Levels<0, 19, 30, 51, 83>>();
    case 67: return make_pvariant2r<30, 12, Levels<0, 19, 30, 51, 83>, Tune<2, 3, 1, 1>>();
    case 68: return make_pvariant2r<30, 12, Levels<0, 19, 30, 51, 83>, Tune<3, 2, 1, 1>>();
    case 69: return make_pvariant2r<30, 12, Levels<0, 18, 30, 51, 83>, Tune<3, 2, 1, 1>>();
    case 70: return make_pvariant2r<30, 12, Levels<0, 20, 30, 51, 83>, Tune<2, 3, 1, 1>>();
    case 71: return make_pvariant2r<30, 12, Levels<0, 19, 30, 53, 83>, Tune<2, 3, 1, 1>>();
    case 72: return make_pvariant2r<30, 12, Levels<0, 19, 30, 49, 83>, Tune<2, 3, 1, 1>>();
    case 73: return make_pvariant2r<29, 12, Levels<0, 19, 29, 51, 83>, Tune<2, 3, 1, 1>>();
    case 74: return make_pvariant2r<30, 12, Levels<0, 19, 30, 51, 83>, Tune<3, 3, 1, 1>>();
    /* 75..80: queues 1.. in global memory; 42 or 38 primes in LDS */
    case 75: return make_pvariant2rg<42, 12, Levels<0, 20, 30, 42, 83>, Tune<2, 3, 1, 1>, 1>();
    case 76: return make_pvariant2rg<42, 12, Levels<0, 20, 30, 42, 60, 83>, Tune<2, 3, 1, 1>, 1>();
    case 77: return make_pvariant2rg<42, 12, Levels<0, 20, 32, 42, 83>, Tune<2, 3, 1, 1>, 1>();
    case 78: return make_pvariant2rg<42, 12, Levels<0, 20, 42, 83>, Tune<2, 3, 1, 1>, 1>();
    case 79: return make_pvariant2rg<30, 12, Levels<0, 20, 30, 51, 83>, Tune<2, 3, 1, 1>, 1>();
    case 80: return make_pvariant2rg<38, 12, Levels<0, 20, 30, 38, 55, 83>, Tune<2, 3, 1, 1>, 1>();
    /* 81..85: only the last queue in global memory; 37 primes in LDS */
    case 81: return make_pvariant2rg<37, 12, Levels<0, 20, 30, 51, 83>, Tune<2, 3, 1, 1>, 2>();
    case 82: return make_pvariant2rg<37, 12, Levels<0, 20, 30, 37, 83>, Tune<2, 3, 1, 1>, 2>();
    case 83: return make_pvariant2rg<37, 12, Levels<0, 20, 37, 51, 83>, Tune<2, 3, 1, 1>, 2>();
    case 84: return make_pvariant2rg<37, 12, Levels<0, 20, 30, 44, 83>, Tune<2, 3, 1, 1>, 2>();
    case 85: return make_pvariant2rg<37, 12, Levels<0, 20, 33, 51, 83>, Tune<2, 3, 1, 1>, 2>();
    case 86: return make_pvariant2rg<37, 12, Levels<0, 20, 34, 51, 83>, Tune<2, 3, 1, 1>, 2>();
    case 87: return make_pvariant2rg<37, 12, Levels<0, 20, 33, 55, 83>, Tune<2, 3, 1, 1>, 2>();
    case 88: return make_pvariant2rg<37, 12, Levels<0, 20, 33, 47, 83>, Tune<2, 3, 1, 1>, 2>();
    case 89: return make_pvariant2rg<37, 12, Levels<0, 19, 33, 51, 83>, Tune<2, 3, 1, 1>, 2>();
    case 90: return make_pvariant2rg<37, 12, Levels<0, 20, 32, 51, 83>, Tune<2, 3, 1, 1>, 2>();
    default: return PersistVariant{0, 0, 0, nullptr};
  }
}

bool pick_pvariant_3(int id, PersistVariant &v) {
  v = part(id);
  return v.fn != nullptr;
}

}  // namespace ap26
