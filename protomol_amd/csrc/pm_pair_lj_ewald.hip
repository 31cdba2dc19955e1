// pm_pair_lj_ewald.hip -- the pair kernels of LennardJonesForce + the Ewald real-space term.
#include "pm_pair.h"

namespace pm {

void launch_pair_lj_ewald(const Launch &L, const ForceDev &P, int kind) { launch_pair_sw<PM_TERM_LJ, PM_TERM_EWALD_REAL>(L, P, kind); }

}  // namespace pm
