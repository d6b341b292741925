// spread_w7_8.hip -- spread launches of kernel widths 7..8 (see spread_impl.h)
#include "spread_impl.h"

namespace mpsq {

MPSQ_SPREAD_W_INST(7)
MPSQ_SPREAD_W_INST(8)

}  // namespace mpsq
