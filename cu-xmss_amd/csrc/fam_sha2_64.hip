// fam_sha2_64.hip -- kernel instantiations for one hash family (Sha512Fam).
#include "launch.h"

namespace xm {
FamOps* ops_sha2_64() {
    static FamOpsT<Sha512Fam> ops;
    return &ops;
}
}  // namespace xm
