// pm_gb.h -- Generalized Born (OBC) all-pairs passes, gfx950.
#pragma once
#include "../../include/protomol_amd.h"
#include "pm_kernels.h"

namespace pm {

constexpr int kGBI = 64;  // atoms per tile (one per lane)

inline int gb_tile_count(int n) { return (n + kGBI - 1) / kGBI; }
// the force pass visits each unordered pair once: tile pairs (it, jt), it <= jt
inline int gb_pair_count(int nt) { return nt * (nt + 1) / 2; }
constexpr int kGBPairDoubles = 2 * 3 * kGBI;  // partial forces per tile pair

void gb_setup_force(ForceDev &P, const pm_force_desc &d);
void gb_setup_system(SysDev &S, const double *rho, const double *scaling, double *born, double *drfac,
                     double *partial, double *spare);
void launch_gb(const Launch &L, const ForceDev &P, bool energies);

}  // namespace pm
