// spread_w13_16.hip -- spread launches of kernel widths 13..16 (see spread_impl.h)
#include "spread_impl.h"

namespace mpsq {

MPSQ_SPREAD_W_INST(13)
MPSQ_SPREAD_W_INST(14)
MPSQ_SPREAD_W_INST(15)
MPSQ_SPREAD_W_INST(16)

}  // namespace mpsq
