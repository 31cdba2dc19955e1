// pm_pair_coulomb.hip -- the pair kernels of CoulombForce alone.
#include "pm_pair.h"

namespace pm {

void launch_pair_coulomb(const Launch &L, const ForceDev &P, int kind) { launch_pair_sw<PM_TERM_COULOMB, PM_TERM_NONE>(L, P, kind); }

}  // namespace pm
