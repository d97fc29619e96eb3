// fam_shake256_64.hip -- kernel instantiations for one hash family (ShakeFam<136,64>).
#include "launch.h"

namespace xm {
FamOps* ops_shake256_64() {
    static FamOpsT<ShakeFam<136,64>> ops;
    return &ops;
}
}  // namespace xm
