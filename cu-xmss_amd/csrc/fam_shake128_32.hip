// fam_shake128_32.hip -- kernel instantiations for one hash family (ShakeFam<168,32>).
#include "launch.h"

namespace xm {
FamOps* ops_shake128_32() {
    static FamOpsT<ShakeFam<168,32>> ops;
    return &ops;
}
}  // namespace xm
