// pm_local.hip -- the local terms of the force pass (bonds, angles, Ewald
// excluded-pair corrections), FP64, owner computes, no atomics on forces.
// Built with machine-level LICM off: a loop around these terms would
// otherwise keep the FP64 library constants of erf / exp / atan2 in
// registers for its whole length (178 against 92 VGPRs per lane).
#include <hip/hip_runtime.h>
#include <math.h>

#include "pm_device.h"
#include "pm_rebuild.h"

namespace pm {

// one wave per share (loc_share): an atom waits for the longest of its terms
// instead of their sum
constexpr int kLocWaves = kLocShares;

// An atom's items of share w (kind `kind`), FP64, evaluated from that atom's
// side ("owner computes": each participant evaluates the term with the same
// expressions, so the forces on its atoms are exactly opposite; energies and
// virial are counted by the term's first atom).  The items come as records
// (LocRec): one coalesced load, then the positions of the term's atoms.
//   BondSystemForce::calcBond (force/bonded/BondSystemForce.h:77-108)
//   AngleSystemForce::calcAngle (force/bonded/AngleSystemForce.h:66-140)
//   NonbondedFullEwaldSystemForce::correctionTerm (NonbondedFullEwaldSystemForce.h:1097-1131)
template <bool EN>
__device__ __forceinline__ void local_terms(const SysDev &S, int a, const ForceDev &P, int w, int kind, double f[3],
                                            double &eb, double &ea, double &ec, double v[6]) {
  const int n = S.n_atoms;
  const PM_GLOBAL double *X = gp(S.x);
  const PM_GLOBAL LocRec *R = gp(S.loc_rec) + (size_t)S.loc_base[w] * n + a;
  const int rows = S.loc_rows[w];
  for (int k = 0; k < rows; ++k) {
    const LocRec rec = R[(size_t)k * n];
    const int role = rec.role;
    if (role < 0) break;  // the atom's items of the share come first
    if (kind == kLocBond) {
      const int a1 = rec.a0, a2 = rec.a1;
      const double kb = rec.p0, r0 = rec.p1;
      double r12[3] = {X[a1] - X[a2], X[n + a1] - X[n + a2], X[2 * n + a1] - X[2 * n + a2]};
      min_image_vec(S, r12);
      const double r = sqrt(r12[0] * r12[0] + r12[1] * r12[1] + r12[2] * r12[2]);
      const double dp = 2.0 * kb * (r - r0);
      const double c = -dp / r;
      const double sg = role == 0 ? 1.0 : -1.0;
      for (int d = 0; d < 3; ++d) f[d] += sg * (r12[d] * c);
      if (EN && role == 0) {
        eb += kb * (r - r0) * (r - r0);
        v[0] += r12[0] * c * r12[0];
        v[1] += r12[0] * c * r12[1];
        v[2] += r12[0] * c * r12[2];
        v[3] += r12[1] * c * r12[1];
        v[4] += r12[1] * c * r12[2];
        v[5] += r12[2] * c * r12[2];
      }
    } else if (kind == kLocAngle) {
      const int a1 = rec.a0, a2 = rec.a1, a3 = rec.a2;
      const double ka = rec.p0, t0 = rec.p1;
      double r12[3] = {X[a1] - X[a2], X[n + a1] - X[n + a2], X[2 * n + a1] - X[2 * n + a2]};
      double r32[3] = {X[a3] - X[a2], X[n + a3] - X[n + a2], X[2 * n + a3] - X[2 * n + a2]};
      min_image_vec(S, r12);
      min_image_vec(S, r32);
      const double d12 = sqrt(r12[0] * r12[0] + r12[1] * r12[1] + r12[2] * r12[2]);
      const double d32 = sqrt(r32[0] * r32[0] + r32[1] * r32[1] + r32[2] * r32[2]);
      const double cr0 = r12[1] * r32[2] - r12[2] * r32[1], cr1 = r12[2] * r32[0] - r12[0] * r32[2],
                   cr2 = r12[0] * r32[1] - r12[1] * r32[0];
      const double cn = sqrt(cr0 * cr0 + cr1 * cr1 + cr2 * cr2);
      const double dt = r12[0] * r32[0] + r12[1] * r32[1] + r12[2] * r32[2];
      const double theta = atan2(cn, dt);
      // sin and cos of theta = atan2(|r12 x r32|, r12 . r32), from the same vectors
      const double inv = 1.0 / (d12 * d32);
      const double st = cn * inv, ct = dt * inv;
      const double dp = 2.0 * ka * (theta - t0);
      double f1[3], f3[3];
      for (int d = 0; d < 3; ++d) {
        const double u12 = r12[d] / d12, u32 = r32[d] / d32;
        f1[d] = -((u12 * ct - u32) / (st * d12)) * dp;
        f3[d] = -((u32 * ct - u12) / (st * d32)) * dp;
      }
      for (int d = 0; d < 3; ++d) f[d] += role == 0 ? f1[d] : (role == 2 ? f3[d] : -f1[d] - f3[d]);
      if (EN && role == 0) {
        ea += ka * (theta - t0) * (theta - t0);
        v[0] += f1[0] * r12[0] + f3[0] * r32[0];
        v[1] += f1[0] * r12[1] + f3[0] * r32[1];
        v[2] += f1[0] * r12[2] + f3[0] * r32[2];
        v[3] += f1[1] * r12[1] + f3[1] * r32[1];
        v[4] += f1[1] * r12[2] + f3[1] * r32[2];
        v[5] += f1[2] * r12[2] + f3[2] * r32[2];
      }
    } else {
      const int a1 = rec.a0, a2 = rec.a1;
      double d[3] = {X[a2] - X[a1], X[n + a2] - X[n + a1], X[2 * n + a2] - X[2 * n + a1]};
      min_image_vec(S, d);
      const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
      const double qq = rec.p0;  // q_a q_b, times 1 - csf for a modified pair
      const double r = sqrt(r2), rr = 1.0 / r;
      const double e = erf(P.dalpha * r) * rr;  // C-library erf, as the reference on Linux
      const double fs = qq * (P.dtwo_alpha_pi * exp(-(P.dalpha * P.dalpha) * r2) - e) * rr * rr;
      const double sg = role == 0 ? -1.0 : 1.0;
      for (int q = 0; q < 3; ++q) f[q] += sg * (d[q] * fs);
      if (EN && role == 0) {
        ec += -qq * e;
        v[0] += d[0] * fs * d[0];
        v[1] += d[0] * fs * d[1];
        v[2] += d[0] * fs * d[2];
        v[3] += d[1] * fs * d[1];
        v[4] += d[1] * fs * d[2];
        v[5] += d[2] * fs * d[2];
      }
    }
  }
}

// One block of k_local: the 64 atoms from blockIdx.x * 64 over the launch's
// systems, wave w taking share w (every thread of the block calls it)
template <bool EN>
__device__ __forceinline__ void local_block(const SysDev *__restrict__ sys, const int *__restrict__ ap, int nsys,
                                            const ForceDev &P, int mode, double (*part)[3][64]) {
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int g = blockIdx.x * 64 + lane;  // the atom, over the launch's systems
  const int total = ap[nsys];
  const bool ok = g < total;
  int kind, slot, nslots;
  loc_share(wv, kind, slot, nslots);
  double f[3] = {0.0, 0.0, 0.0};
  double eb = 0.0, ea = 0.0, ec = 0.0, v[6] = {0, 0, 0, 0, 0, 0};
  int s = wave_sys(ap, nsys, g, total);
  if (s >= 0) {
    const SysDev *__restrict__ S = sys + s;  // scalar loads of the record
    if (ok) local_terms<EN>(*S, g - ap[s], P, wv, kind, f, eb, ea, ec, v);
  } else {
    s = ok ? find_sys(ap, nsys, g) : 0;
    if (ok) local_terms<EN>(sys[s], g - ap[s], P, wv, kind, f, eb, ea, ec, v);
  }
  if (mode & kLocalForces) {
    part[wv][0][lane] = f[0];
    part[wv][1][lane] = f[1];
    part[wv][2][lane] = f[2];
    __syncthreads();
    // every atom, zero for one without local terms; the slots in a fixed order
    if (wv < 3 && ok) {
      double t = 0.0;
      for (int w = 0; w < kLocWaves; ++w) t += part[w][wv][lane];
      const SysDev &S = sys[s];
      gp(S.fl)[(size_t)wv * S.n_atoms + (g - ap[s])] = to_fixed(t, kFScale);
    }
  }
  if (!EN) return;
  // bonded and correction terms add to the atomic virial only when it is on
  // (BondSystemForce.h:121, AngleSystemForce.h:155, NonbondedFullEwaldSystemForce.h:1125)
  if (!P.virial)
    for (int k = 0; k < 6; ++k) v[k] = 0.0;
  // each lane's terms go to fixed point first; integer sums are exact, so the
  // totals do not depend on how lanes are grouped
  long long q[9] = {to_fixed(eb, kEScale), to_fixed(ea, kEScale), to_fixed(ec, kEScale),
                    to_fixed(v[0], kEScale), to_fixed(v[1], kEScale), to_fixed(v[2], kEScale),
                    to_fixed(v[3], kEScale), to_fixed(v[4], kEScale), to_fixed(v[5], kEScale)};
  const int slots[9] = {kAccBond, kAccAngle, kAccCoul, kAccVxx, kAccVxy, kAccVxz, kAccVyy, kAccVyz, kAccVzz};
  // the block's 64 atoms are all of one system or not, alike in every wave
  const int s0 = __shfl(s, 0);
  const bool mixed = wave_ballot(ok && s != s0) != 0;
  if (mixed) {
    if (ok)
      for (int k = 0; k < 9; ++k)
        if (q[k]) atomic_add_ll(&acc_copy(sys[s].acc, blockIdx.x)[slots[k]], q[k]);
    return;
  }
  // wave sums, then the block's through LDS: nine atomics per block
  for (int k = 0; k < 9; ++k) q[k] = wave_sum_ll(q[k]);
  __syncthreads();  // every read of part above is done
  long long *red = reinterpret_cast<long long *>(&part[0][0][0]);
  if (lane == 0)
    for (int k = 0; k < 9; ++k) red[wv * 9 + k] = q[k];
  __syncthreads();
  if (wv == 0 && lane == 0) {
    long long *acc = acc_copy(sys[s0].acc, blockIdx.x);
    for (int k = 0; k < 9; ++k) {
      long long t = 0;
      for (int w = 0; w < kLocWaves; ++w) t += red[w * 9 + k];
      if (t) atomic_add_ll(&acc[slots[k]], t);
    }
  }
}

template <bool EN>
__global__ __launch_bounds__(64 * kLocWaves) void k_local(const SysDev *__restrict__ sys, const int *__restrict__ ap,
                                                          int nsys, ForceDev P, int mode) {
  __shared__ double part[kLocWaves][3][64];
  local_block<EN>(sys, ap, nsys, P, mode, part);
}

// The local terms with their energies and virial: one wave per 64 atoms
// walks the shares in order.  Each share's force is added to the atom's sum
// in the order k_local's LDS sum uses, and each share's energies go to fixed
// point on their own, so forces and totals are the same bits as k_local's.
// (With a wave per share, the energy form took 25 us at the 96k box; here
// fewer, longer waves.)
__global__ __launch_bounds__(64) void k_local_seq(const SysDev *__restrict__ sys, const int *__restrict__ ap,
                                                  int nsys, ForceDev P) {
  const int lane = threadIdx.x & 63;
  const int g = blockIdx.x * 64 + lane;
  const int total = ap[nsys];
  const bool ok = g < total;
  int s = wave_sys(ap, nsys, g, total);
  const bool uniform = s >= 0;
  if (!uniform) s = ok ? find_sys(ap, nsys, g) : 0;
  double t[3] = {0.0, 0.0, 0.0};
  long long q[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int sh = 0; sh < kLocWaves; ++sh) {
    int kind, slot, nslots;
    loc_share(sh, kind, slot, nslots);
    double f[3] = {0.0, 0.0, 0.0};
    double eb = 0.0, ea = 0.0, ec = 0.0, v[6] = {0, 0, 0, 0, 0, 0};
    if (ok) {
      if (uniform) {
        const SysDev *__restrict__ S = sys + s;
        local_terms<true>(*S, g - ap[s], P, sh, kind, f, eb, ea, ec, v);
      } else {
        local_terms<true>(sys[s], g - ap[s], P, sh, kind, f, eb, ea, ec, v);
      }
    }
    t[0] += f[0];
    t[1] += f[1];
    t[2] += f[2];
    // bonded and correction terms add to the atomic virial only when it is on
    // (BondSystemForce.h:121, AngleSystemForce.h:155, NonbondedFullEwaldSystemForce.h:1125)
    if (!P.virial)
      for (int k = 0; k < 6; ++k) v[k] = 0.0;
    q[0] += to_fixed(eb, kEScale);
    q[1] += to_fixed(ea, kEScale);
    q[2] += to_fixed(ec, kEScale);
    for (int k = 0; k < 6; ++k) q[3 + k] += to_fixed(v[k], kEScale);
  }
  if (ok) {
    const SysDev &S = sys[s];
    for (int c = 0; c < 3; ++c) gp(S.fl)[(size_t)c * S.n_atoms + (g - ap[s])] = to_fixed(t[c], kFScale);
  }
  const int slots[9] = {kAccBond, kAccAngle, kAccCoul, kAccVxx, kAccVxy, kAccVxz, kAccVyy, kAccVyz, kAccVzz};
  const int s0 = __shfl(s, 0);
  const bool mixed = wave_ballot(ok && s != s0) != 0;
  if (mixed) {
    if (ok)
      for (int k = 0; k < 9; ++k)
        if (q[k]) atomic_add_ll(&acc_copy(sys[s].acc, blockIdx.x)[slots[k]], q[k]);
    return;
  }
  for (int k = 0; k < 9; ++k) q[k] = wave_sum_ll(q[k]);
  if (lane == 0) {
    long long *acc = acc_copy(sys[s0].acc, blockIdx.x);
    for (int k = 0; k < 9; ++k)
      if (q[k]) atomic_add_ll(&acc[slots[k]], q[k]);
  }
}

// k_bin and k_local (forces) in one launch, for a forces-only pass: the list
// check and binning of a block's 64 atoms ride on its last wave, before that
// wave's share of the local terms (both read the positions of the pass, and
// neither writes what the other reads)
__global__ __launch_bounds__(64 * kLocWaves) void k_bin_local(const SysDev *__restrict__ sys,
                                                              const int *__restrict__ ap, int nsys, ForceDev P,
                                                              int mode, float lim2, int kmax, int *rb_any) {
  __shared__ double part[kLocWaves][3][64];
  if ((threadIdx.x >> 6) == kLocWaves - 1)
    bin_thread(sys, ap, nsys, mode, lim2, kmax, blockIdx.x * 64 + (threadIdx.x & 63), rb_any);
  local_block<false>(sys, ap, nsys, P, kLocalForces, part);
}

void launch_bin_local(const Launch &L, const ForceDev &P, int mode, float lim2, int kmax) {
  hipLaunchKernelGGL(k_bin_local, dim3(nblk(L.n_atoms_total, 64)), dim3(64 * kLocWaves), 0, L.stream, L.sys,
                     L.atom_prefix, L.nsys, P, mode, lim2, kmax, L.rb_any);
}

void launch_local(const Launch &L, const ForceDev &P, int mode) {
  if (!L.n_atoms_total || !L.n_local_total) return;
  const dim3 g(nblk(L.n_atoms_total, 64)), b(64 * kLocWaves);
  if (mode & kLocalEnergies)
    hipLaunchKernelGGL(k_local_seq, g, dim3(64), 0, L.stream, L.sys, L.atom_prefix, L.nsys, P);
  else
    hipLaunchKernelGGL(k_local<false>, g, b, 0, L.stream, L.sys, L.atom_prefix, L.nsys, P, mode);
}

}  // namespace pm
