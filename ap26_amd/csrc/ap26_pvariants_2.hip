/* ap26_pvariants_2.hip -- part 2 of the batched engine's sieve variant
 * table (ap26_pvariants.h).
 */
#include <hip/hip_runtime.h>

#include "ap26_pvariants.h"

namespace ap26 {

static PersistVariant part(int id) {
  switch (id) {
    case 32: return make_pvariant2<29, 16, Levels<0, 20, 31, 51, 83>>();
    case 33: return make_pvariant2<27, 16, Levels<0, 20, 29, 51, 83>>();
    case 34: return make_pvariant2<29, 16, Levels<0, 19, 29, 51, 83>>();
    case 35: return make_pvariant2<29, 16, Levels<0, 18, 29, 51, 83>, 1>();
    case 36: return make_pvariant2<29, 16, Levels<0, 18, 29, 51, 83>, 2>();
    case 37: return make_pvariant2<29, 16, Levels<0, 18, 29, 51, 83>, 3>();
    case 38: return make_pvariant2<29, 16, Levels<0, 20, 31, 51, 83>, 3>();
    case 39: return make_pvariant2<29, 16, Levels<0, 18, 29, 51, 83>, 0, 2>();
    case 40: return make_pvariant2<29, 16, Levels<0, 18, 29, 51, 83>, 0, 4>();
    case 50: return make_pvariant2r<21, 16, Levels<0, 16, 21, 51, 83>>();
    case 51: return make_pvariant2r<21, 16, Levels<0, 18, 21, 51, 83>>();
    case 52: return make_pvariant2r<29, 12, Levels<0, 16, 29, 51, 83>>();
    case 53: return make_pvariant2r<29, 12, Levels<0, 18, 29, 51, 83>>();
    case 54: return make_pvariant2r<30, 12, Levels<0, 18, 30, 51, 83>>();
    case 55: return make_pvariant2r<29, 12, Levels<0, 20, 29, 51, 83>>();
    case 56: return make_pvariant2r<36, 8, Levels<0, 18, 36, 55, 83>>();
    /* 57..66: variant 54 with other wave priorities (level 0, level-1
       drains, deeper drains) and level-0 splits */
    case 57: return make_pvariant2r<30, 12, Levels<0, 18, 30, 51, 83>, Tune<0, 0, 0, 2>>();
    case 58: return make_pvariant2r<30, 12, Levels<0, 18, 30, 51, 83>, Tune<2, 3, 1, 1>>();
    case 59: return make_pvariant2r<30, 12, Levels<0, 18, 30, 51, 83>, Tune<2, 3, 1, 4>>();
    case 60: return make_pvariant2r<30, 12, Levels<0, 18, 30, 51, 83>, Tune<1, 2, 0, 2>>();
    case 61: return make_pvariant2r<30, 12, Levels<0, 18, 30, 51, 83>, Tune<3, 2, 1, 2>>();
    case 62: return make_pvariant2r<30, 12, Levels<0, 18, 30, 51, 83>, Tune<2, 2, 2, 2>>();
    case 63: return make_pvariant2r<30, 12, Levels<0, 18, 30, 55, 83>>();
    case 64: return make_pvariant2r<30, 12, Levels<0, 18, 30, 45, 83>>();
    case 65: return make_pvariant2r<30, 12, Levels<0, 17, 30, 51, 83>>();
    default: return PersistVariant{0, 0, 0, nullptr};
  }
}

bool pick_pvariant_2(int id, PersistVariant &v) {
  v = part(id);
  return v.fn != nullptr;
}

}  // namespace ap26
