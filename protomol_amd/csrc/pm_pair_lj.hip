// pm_pair_lj.hip -- the pair kernels of LennardJonesForce alone.
#include "pm_pair.h"

namespace pm {

void launch_pair_lj(const Launch &L, const ForceDev &P, int kind) { launch_pair_sw<PM_TERM_LJ, PM_TERM_NONE>(L, P, kind); }

}  // namespace pm
