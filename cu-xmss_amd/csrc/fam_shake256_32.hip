// fam_shake256_32.hip -- kernel instantiations for one hash family (ShakeFam<136,32>).
#include "launch.h"

namespace xm {
FamOps* ops_shake256_32() {
    static FamOpsT<ShakeFam<136,32>> ops;
    return &ops;
}
}  // namespace xm
