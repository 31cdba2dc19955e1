// pm_pair_lj_coulomb.hip -- the pair kernels of LennardJonesForce + CoulombForce (OneAtomPairTwo).
#include "pm_pair.h"

namespace pm {

void launch_pair_lj_coulomb(const Launch &L, const ForceDev &P, int kind) { launch_pair_sw<PM_TERM_LJ, PM_TERM_COULOMB>(L, P, kind); }

}  // namespace pm
