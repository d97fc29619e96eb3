// fam_sha2_24.hip -- kernel instantiations for one hash family (Sha256Fam<24>).
#include "launch.h"

namespace xm {
FamOps* ops_sha2_24() {
    static FamOpsT<Sha256Fam<24>> ops;
    return &ops;
}
}  // namespace xm
