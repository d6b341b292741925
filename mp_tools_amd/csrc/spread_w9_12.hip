// spread_w9_12.hip -- spread launches of kernel widths 9..12 (see spread_impl.h)
#include "spread_impl.h"

namespace mpsq {

MPSQ_SPREAD_W_INST(9)
MPSQ_SPREAD_W_INST(10)
MPSQ_SPREAD_W_INST(11)
MPSQ_SPREAD_W_INST(12)

}  // namespace mpsq
