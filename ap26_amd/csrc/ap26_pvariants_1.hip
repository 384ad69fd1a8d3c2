/* ap26_pvariants_1.hip -- part 1 of the batched engine's sieve variant
 * table (ap26_pvariants.h).
 */
#include <hip/hip_runtime.h>

#include "ap26_pvariants.h"

namespace ap26 {

static PersistVariant part(int id) {
  switch (id) {
    case 1: return make_pvariant<31, 16, Levels4>();
    case 2: return make_pvariant<45, 12, Levels4>();
    case 3: return make_pvariant<25, 16, Levels5>();
    case 10: return make_pvariant2<29, 16, Levels4>();
    case 11: return make_pvariant2<31, 14, Levels4>();
    case 12: return make_pvariant2<35, 12, Levels4>();
    case 13: return make_pvariant2<20, 12, Levels5>();
    case 14: return make_pvariant2<31, 12, Levels4>();
    case 9: return make_pvariant<38, 16, Levels4>();
    case 15: return make_pvariant2<20, 16, Levels4>();
    case 16: return make_pvariant2<24, 16, Levels4>();
    case 17: return make_pvariant2<18, 16, Levels4>();
    case 18: return make_pvariant2<22, 16, Levels4>();
    case 19: return make_pvariant2<15, 16, Levels4>();
    case 20: return make_pvariant2<18, 16, Levels<0, 15, 22, 29, 45, 83>>();
    case 21: return make_pvariant2<15, 16, Levels<0, 15, 22, 29, 45, 83>>();
    case 22: return make_pvariant2<29, 16, Levels<0, 18, 29, 45, 83>>();
    case 23: return make_pvariant2<29, 16, Levels<0, 20, 29, 45, 83>>();
    case 24: return make_pvariant2<29, 16, Levels<0, 22, 29, 45, 83>>();
    case 25: return make_pvariant2<29, 16, Levels<0, 20, 29, 41, 83>>();
    case 26: return make_pvariant2<29, 16, Levels<0, 20, 29, 51, 83>>();
    case 27: return make_pvariant2<29, 16, Levels<0, 24, 29, 45, 83>>();
    case 29: return make_pvariant2<29, 16, Levels<0, 19, 29, 45, 83>>();
    case 30: return make_pvariant2<29, 16, Levels<0, 20, 29, 55, 83>>();
    case 31: return make_pvariant2<29, 16, Levels<0, 18, 27, 45, 83>>();
    default: return PersistVariant{0, 0, 0, nullptr};
  }
}

bool pick_pvariant_1(int id, PersistVariant &v) {
  v = part(id);
  return v.fn != nullptr;
}

}  // namespace ap26
