// pm_gb.h -- Generalized Born (OBC) all-pairs passes, gfx950.
#pragma once
#include "../../include/protomol_amd.h"
#include "pm_kernels.h"

namespace pm {

constexpr int kGBBlock = 256;  // i atoms per workgroup (one per lane), j staged in LDS

inline int gb_tile_count(int n) { return (n + kGBBlock - 1) / kGBBlock; }

void gb_setup_force(ForceDev &P, const pm_force_desc &d);
void gb_setup_system(SysDev &S, const double *rho, const double *scaling, double *born, double *drfac,
                     double *partial, double *spare);
void launch_gb(const Launch &L, const ForceDev &P, bool energies);

}  // namespace pm
