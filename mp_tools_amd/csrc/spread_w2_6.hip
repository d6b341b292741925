// spread_w2_6.hip -- spread launches of kernel widths 2..6 (see spread_impl.h)
#include "spread_impl.h"

namespace mpsq {

MPSQ_SPREAD_W_INST(2)
MPSQ_SPREAD_W_INST(3)
MPSQ_SPREAD_W_INST(4)
MPSQ_SPREAD_W_INST(5)
MPSQ_SPREAD_W_INST(6)

}  // namespace mpsq
