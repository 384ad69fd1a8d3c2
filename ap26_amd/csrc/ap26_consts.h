/* ap26_consts.h -- constants of the AP26 search space, shared by the HIP
 * library and the C host.
 *
 * The numeric constants are Wroblewski's (src/common/CONST.H:7-27): they define
 * which progressions are searched, so they are reproduced as data.  The prime
 * tables below are generated at compile time.
 */
#ifndef AP26_CONSTS_H
#define AP26_CONSTS_H

#include <stdint.h>

#define AP26_PRIM23 INT64_C(223092870)          /* 2*3*5*7*11*13*17*19*23 */
#define AP26_MOD INT64_C(258559632607830)       /* 2*3*5*29*31*...*59 */
#define AP26_N0 INT64_C(106990415896110)
#define AP26_N30 INT64_C(94805198622871)
#define AP26_S3 INT64_C(172373088405220)
#define AP26_S5 INT64_C(51711926521566)
#define AP26_PRES31 INT64_C(16681266619860)
#define AP26_PRES37 INT64_C(181690552643340)
#define AP26_PRES41 INT64_C(132432982555230)
#define AP26_PRES43 INT64_C(126273308948010)
#define AP26_PRES47 INT64_C(115526644356690)
#define AP26_PRES53 INT64_C(48784836341100)
#define AP26_PRES59 INT64_C(100794433050510)

#define AP26_NN43 10840       /* n43 starts per K */
#define AP26_N53_PER_N43 12673 /* 19 * 23 * 29 */
#define AP26_N59_STEPS 35     /* i59 steps per n53 */
#define AP26_NSHIFT 10        /* shifts per K: SHIFT + 64*iter */
#define AP26_MAXINTV 2000000000

#endif
