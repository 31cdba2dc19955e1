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
            *corr_prefix = nullptr, *bonded_prefix = nullptr, *gb_prefix = nullptr, *sc_prefix = nullptr,
            *gbp_prefix = nullptr;  // GB tile pairs
  int n_atoms_total = 0, n_cl_total = 0, n_cols_total = 0, n_slots_total = 0, n_corr_total = 0,
      n_bonded_total = 0, n_gb_tiles_total = 0, n_local_total = 0, n_sc_total = 0, n_gb_pairs_total = 0;
  const struct EwLaunch *ew = nullptr;  // per system with a reciprocal-space term (host array)
  int n_ew = 0;
  int max_mol = 0;                      // most molecules of a system (molecular virial; 0 = off)
  // one device int: some system of the launch rebuilds its list this force
  // pass (raised by the list check, cleared by the kick); null: the rebuild
  // kernels check each system's own decision only
  int *rb_any = nullptr;
  // the list build takes a super-cluster's excluded pairs from LDS up to this
  // many (kHitCap); past it, each entry walks its j atoms' partners.  A test
  // knob (PM_DEBUG_HIT_CAP) lowers it to drive that path
  int hit_cap = 256;
};

// Host-side launch shape of one system's reciprocal-space passes
struct EwLaunch {
  const SysDev *sys = nullptr;  // the system's record in the device table
  int n = 0, nk = 0, ntiles = 0, nchunks = 0, nfb = 0, lds_bytes = 0;
};
// kEwVirial: the energies pass also adds the k-sum virial (to the atomic
// slots, and to the molecular ones with kEwMolVirial); kEwMolVirial: add each
// atom's reciprocal force (x) (its molecule's centre - its minimal position)
// to the molecular virial
enum { kEwForces = 1, kEwEnergies = 2, kEwVirial = 4, kEwMolVirial = 8 };

void launch_import(const Launch &L, int what);
void launch_export(const Launch &L, int what);
void launch_gather_scalars(const Launch &L, long long *out);
void launch_clear_err(const Launch &L);
// the list check of a force pass (the bookkeeping of launch_step_begin with
// the same mode, lim2, kmax) and, where it decides so, the pair-list rebuild.
// local (forces-only passes): its first kernel also runs the local terms'
// forces, as launch_local(L, *local, kLocalForces) would; returns whether it
// did
bool launch_rebuild(const Launch &L, float rlist2, int mode, float lim2, int kmax, const ForceDev *local = nullptr);
// k_bin with the local terms' forces in the same launch (pm_local.hip)
void launch_bin_local(const Launch &L, const ForceDev &P, int mode, float lim2, int kmax);
bool launch_pair(const Launch &L, const ForceDev &P, int kind);
// the pair kernels of one term combination each (pm_pair_*.hip)
void launch_pair_lj(const Launch &L, const ForceDev &P, int kind);
void launch_pair_coulomb(const Launch &L, const ForceDev &P, int kind);
void launch_pair_lj_coulomb(const Launch &L, const ForceDev &P, int kind);
void launch_pair_lj_ewald(const Launch &L, const ForceDev &P, int kind);
void launch_pair_ewald(const Launch &L, const ForceDev &P, int kind);
void launch_local(const Launch &L, const ForceDev &P, int mode);
// LangevinLeapfrogIntegrator::doHalfKick constants for one time step
// (integrator/base/LangevinLeapfrogIntegrator.cpp:46-70); on = 0: the plain
// leapfrog kicks
struct LangevinDev {
  int on = 0;
  double vdt = 1.0, fdt = 0.0, ndt = 0.0, fc = 0.0;  // e^{-g dt/2}, (1 - e^{-g dt/2})/g, noise scale, 2 kB T g
};
// zero_forces: the kick clears the pair forces it consumed (the GB path;
// the cutoff path's list check clears them)
void launch_kick_drift(const Launch &L, double h, const LangevinDev &lv, bool zero_forces);
void launch_md_arm(const Launch &L);
void launch_half_kick(const Launch &L, double hh, const LangevinDev &lv);
void launch_zero_forces(const Launch &L);
// zero `bytes` (a multiple of 8) of device memory with a fill kernel
void launch_fill_zero(void *p, size_t bytes, hipStream_t stream);
// the molecules' centres of mass at the current positions (molecular virial)
void launch_mol_com(const Launch &L);
void launch_step_begin(const Launch &L, int mode, float lim2, int kmax);
// reciprocal-space Ewald of the systems in L.ew: kEwForces (tables, structure
// factors, forces) and/or kEwEnergies (energy and virial from the last
// structure factors)
void launch_ewald(const Launch &L, int mode);

}  // namespace pm
