// pm_kernels.h -- launch interface of the gfx950 kernels (host side).
#pragma once
#include "pm_core.h"

namespace pm {

// Launch geometry of one context: the device SysDev table and prefix arrays
// (nsys + 1 ints each, device memory) over the systems' atoms, cluster
// capacities, columns, slots, Ewald correction pairs and bonded terms.
struct Launch {
  hipStream_t stream = nullptr;
  SysDev *sys = nullptr;
  int nsys = 0;
  const int *atom_prefix = nullptr, *cl_prefix = nullptr, *col_prefix = nullptr, *slot_prefix = nullptr,
            *corr_prefix = nullptr, *bonded_prefix = nullptr, *gb_prefix = nullptr;
  int n_atoms_total = 0, n_cl_total = 0, n_cols_total = 0, n_slots_total = 0, n_corr_total = 0,
      n_bonded_total = 0, n_gb_tiles_total = 0;
};

void launch_import_x(const Launch &L, const double *xaos);
void launch_import_v(const Launch &L, const double *vaos);
void launch_rebuild(const Launch &L, float rlist2);
bool launch_pair(const Launch &L, const ForceDev &P, bool energies, bool census);
void launch_ewald_corr(const Launch &L, const ForceDev &P, bool energies);
void launch_bonded(const Launch &L, bool energies);
void launch_kick_drift(const Launch &L, double h, int mode);
void launch_half_kick(const Launch &L, double hh);
void launch_zero_forces(const Launch &L);
void launch_step_begin(const Launch &L, int mode, float lim2, int kmax);
void launch_export_f(const Launch &L, double *out);
void launch_export_xv(const Launch &L, double *xo, double *vo);

}  // namespace pm
