// pm_pair_ewald.hip -- the pair kernels of the Ewald real-space term alone.
#include "pm_pair.h"

namespace pm {

void launch_pair_ewald(const Launch &L, const ForceDev &P, int kind) { launch_pair_t<PM_TERM_NONE, PM_TERM_EWALD_REAL, PM_SWITCH_NONE, PM_SWITCH_CUTOFF>(L, P, kind); }

}  // namespace pm
