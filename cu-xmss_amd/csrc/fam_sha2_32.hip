// fam_sha2_32.hip -- kernel instantiations for one hash family (Sha256Fam<32>).
#include "launch.h"

namespace xm {
FamOps* ops_sha2_32() {
    static FamOpsT<Sha256Fam<32>> ops;
    return &ops;
}
}  // namespace xm
