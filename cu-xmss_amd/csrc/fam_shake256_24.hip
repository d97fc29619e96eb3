// fam_shake256_24.hip -- kernel instantiations for one hash family (ShakeFam<136,24>).
#include "launch.h"

namespace xm {
FamOps* ops_shake256_24() {
    static FamOpsT<ShakeFam<136,24>> ops;
    return &ops;
}
}  // namespace xm
