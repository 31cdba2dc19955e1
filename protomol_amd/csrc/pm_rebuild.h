// pm_rebuild.h -- the list check and the first rebuild step (binning),
// shared by k_bin (pm_kernels.hip) and k_bin_local (pm_local.hip), and the
// per-step bookkeeping that the pair kernel finishes.
#pragma once
#include "pm_device.h"

namespace pm {

// The list check of a force pass (the Verlet criterion: rebuild when an atom
// moved more than lim from where the list was built, or every kmax steps, or
// always with kStepForce), from values no kernel of the rebuild changes
// the step's largest squared displacement (float bits), over the slots
__device__ __forceinline__ unsigned max_disp2(const SysDev &S) {
  unsigned d = 0u;
  for (int k = 0; k < kDispSlots; ++k) d = max(d, S.maxdisp2[k]);
  return d;
}
__device__ __forceinline__ void clear_disp2(const SysDev &S) {
  for (int k = 0; k < kDispSlots; ++k) S.maxdisp2[k] = 0u;
}
__device__ __forceinline__ bool rebuild_due(const SysDev &S, int mode, float lim2, int kmax) {
  if (mode & kStepForce) return true;
  if (!(mode & kStepAuto)) return false;
  return __uint_as_float(max_disp2(S)) > lim2 || (kmax > 0 && S.ctl[kCtlSince] + 1 >= kmax);
}

// the rest of a step's bookkeeping once k_bin has published the decision in
// ctl[kCtlRebuild] (one thread per system, in the pair kernel: after the
// rebuild kernels, before the next kick)
__device__ __forceinline__ void step_finish_ctl(const SysDev &S) {
  int *ctl = S.ctl;
  ctl[kCtlFirst] = 0;  // the step's kick has run
  ctl[kCtlNoise] += 1;
  const unsigned d2 = max_disp2(S);
  *S.maxdisp_run = max(*S.maxdisp_run, d2);
  clear_disp2(S);
  if (ctl[kCtlRebuild]) {
    ctl[kCtlSince] = 0;
    ctl[kCtlCount] += 1;
  } else {
    ctl[kCtlSince] += 1;
  }
}

// Rebuild, step 1 of 3, one thread per atom.  It also starts the force pass:
// it clears the pair forces; every system's thread of atom 0 publishes the
// list check's decision (and
// raises the launch's rb_any flag, cleared by the step's kick, when the
// system rebuilds) and clears the census, and the first threads clear the
// energy accumulators.
// When the list is rebuilt, the atom's (z, index) key goes to its x-y
// column's bucket.  Lanes of a wave that share a (system, column) key are
// served by one atomic: the wave peels off the group of the first remaining
// lane until every lane is done.  Atoms in original order are spatially
// coherent, so a wave usually holds only a few keys.
// one atom of k_bin: the force pass's start and, when the list is rebuilt,
// the atom's column and sort key; returns whether it has a key to place
__device__ __forceinline__ bool bin_atom(const SysDev &S, int a, int mode, float lim2, int kmax, int &col,
                                         unsigned long long &key, int *rb_any) {
  const int n = S.n_atoms;
  const bool rb = rebuild_due(S, mode, lim2, kmax);
  for (int k = a; k < kAccN * kAccCopies; k += n) S.acc[k] = 0;
  // the pass's pair forces start from zero (every slot, before any rebuild
  // reassigns them; slot order, so a wave's stores are contiguous)
  for (int k = a; k < S.slot_cap; k += n) {
    S.fx[k] = 0;
    S.fy[k] = 0;
    S.fz[k] = 0;
  }
  if (a == 0) {
    S.ctl[kCtlRebuild] = rb ? 1 : 0;
    S.census[0] = 0;
    S.census[1] = 0;
    if (rb) {
      S.ctl[kCtlEntries] = 0;  // the list build counts its cluster pairs afresh
      if (rb_any) *rb_any = 1;
    }
  }
  if (!rb) return false;
  double x = S.x[a], y = S.x[n + a], z = S.x[2 * n + a];
  if (!(isfinite(x) && isfinite(y) && isfinite(z))) {
    atomicOr(S.err, (unsigned)kErrNonFinite);
    x = y = z = 0.0;
  }
  if (S.vac && (x < S.vac_lo[0] || x >= S.vac_hi[0] || y < S.vac_lo[1] || y >= S.vac_hi[1] ||
                z < S.vac_lo[2] || z >= S.vac_hi[2]))
    atomicOr(S.err, (unsigned)kErrVacuumRegion);
  const double rx = wrap_rel(x, S.omin[0], S.L[0]);
  const double ry = wrap_rel(y, S.omin[1], S.L[1]);
  const double rz = wrap_rel(z, S.omin[2], S.L[2]);
  int cx = (int)(rx / S.ax), cy = (int)(ry / S.ay);
  cx = cx < 0 ? 0 : (cx >= S.ncx ? S.ncx - 1 : cx);
  cy = cy < 0 ? 0 : (cy >= S.ncy ? S.ncy - 1 : cy);
  col = cy * S.ncx + cx;
  const unsigned zb = __float_as_uint((float)rz);  // z >= 0: bit order == value order
  key = ((unsigned long long)zb << 32) | (unsigned)a;
  return true;
}

// k_bin's work for the thread's atom g over the launch's systems (every lane
// of the wave takes part: the bucket appends are grouped per wave)
__device__ __forceinline__ void bin_thread(const SysDev *__restrict__ sys, const int *__restrict__ ap, int nsys,
                                           int mode, float lim2, int kmax, int g, int *rb_any) {
  const int lane = threadIdx.x & 63;
  const int total = ap[nsys];
  bool todo = false;
  int s = wave_sys(ap, nsys, g, total), col = 0;
  unsigned long long key = 0;
  if (s >= 0) {
    const SysDev *__restrict__ S = sys + s;  // scalar loads of the record
    if (g < total) todo = bin_atom(*S, g - ap[s], mode, lim2, kmax, col, key, rb_any);
  } else {
    s = g < total ? find_sys(ap, nsys, g) : 0;
    if (g < total) todo = bin_atom(sys[s], g - ap[s], mode, lim2, kmax, col, key, rb_any);
  }
  // one atomic per (system, column) group of the wave
  while (true) {
    const unsigned long long rem = wave_ballot(todo);
    if (rem == 0) break;
    const int leader = __ffsll((long long)rem) - 1;
    const int ls = __shfl(s, leader), lc = __shfl(col, leader);
    const bool mine = todo && s == ls && col == lc;
    const unsigned long long grp = wave_ballot(mine);
    int base = 0;
    if (lane == leader) base = atomicAdd(&sys[ls].col_count[lc], __popcll(grp));
    base = __shfl(base, leader);
    if (mine) {
      const int pos = base + __popcll(grp & ((1ull << lane) - 1ull));
      if (pos < kSortCap) sys[s].bucket[(size_t)col * kSortCap + pos] = key;
      else atomicOr(sys[s].err, (unsigned)kErrColumnCap);
    }
    todo = todo && !mine;
  }
}

}  // namespace pm
