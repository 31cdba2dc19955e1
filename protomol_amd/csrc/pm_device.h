// pm_device.h -- device helpers shared by the gfx950 kernel sources: pointer
// address spaces, system lookup, fixed-point atomics, wave reductions, the
// FP64 minimal image and the switching functions.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

#include "../../include/protomol_amd.h"
#include "pm_kernels.h"

namespace pm {

// workgroups of b threads covering n items (host side)
static inline int nblk(long long n, int b) { return (int)((n + b - 1) / b); }

// ------------------------------------------------------------- helpers

// Pointers read out of the SysDev table are generic; casting them to the
// global (1) or constant (4) address space lets the compiler use global_* and,
// for wave-uniform reads of data no kernel writes meanwhile, s_load_*.
// (the host pass of hipcc sees kernel bodies too, without address spaces)
// PM_LDS marks pointers into the workgroup's LDS, so that a load through one
// can never be merged with a global load into a FLAT load of a generic pointer.
#if defined(__HIP_DEVICE_COMPILE__)
#define PM_GLOBAL __attribute__((address_space(1)))
#define PM_CONST __attribute__((address_space(4)))
#define PM_LDS __attribute__((address_space(3)))
#else
#define PM_GLOBAL
#define PM_CONST
#define PM_LDS
#endif
template <class T>
__device__ __forceinline__ PM_GLOBAL T *gp(T *p) {
  return (PM_GLOBAL T *)p;
}
template <class T>
__device__ __forceinline__ const PM_CONST T *cp(const T *p) {
  return (const PM_CONST T *)p;
}

__device__ __forceinline__ int find_sys(const int *__restrict__ prefix, int nsys, int w) {
  int lo = 0, hi = nsys;  // prefix[lo] <= w < prefix[hi]
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (prefix[mid] <= w) lo = mid;
    else hi = mid;
  }
  return lo;
}

// The system of a per-atom launch's wave (prefix ap over nsys systems, total
// atoms): wave-uniform, in scalar registers, when every live lane of the wave
// belongs to that system; -1 when the wave straddles two systems.  Every lane
// of the wave must be active.
__device__ __forceinline__ int wave_sys(const int *__restrict__ ap, int nsys, int g, int total) {
  const int g0 = __builtin_amdgcn_readfirstlane(g);  // lane 0 holds the wave's lowest index
  const int s = __builtin_amdgcn_readfirstlane(find_sys(ap, nsys, min(g0, max(total - 1, 0))));
  const int end = ap[s + 1];
  return __ballot(g < total && g >= end) == 0 ? s : -1;
}

__device__ __forceinline__ long long to_fixed(double v, double scale) {
  return __double2ll_rn(v * scale);
}

__device__ __forceinline__ void atomic_add_ll(long long *p, long long v) {
  atomicAdd(reinterpret_cast<unsigned long long *>(p), static_cast<unsigned long long>(v));
}
#if defined(__HIP_DEVICE_COMPILE__)
__device__ __forceinline__ void atomic_add_ll(PM_GLOBAL long long *p, long long v) {
  __hip_atomic_fetch_add((PM_GLOBAL unsigned long long *)p, static_cast<unsigned long long>(v), __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
}
#endif

__device__ __forceinline__ unsigned long long pair_hash(unsigned lo, unsigned hi) {
  unsigned long long z = ((unsigned long long)hi << 32) | lo;
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__device__ __forceinline__ unsigned long long wave_ballot(bool p) { return __ballot(p); }
// lane mask of a predicate over the wave (all 64 lanes active), and a lane's
// bit of such a mask; both stay in scalar registers where the compiler can
__device__ __forceinline__ unsigned long long lane_mask(bool p) { return __builtin_amdgcn_ballot_w64(p); }
__device__ __forceinline__ bool lane_bit(unsigned long long m) { return __builtin_amdgcn_inverse_ballot_w64(m); }

__device__ __forceinline__ float wave_max_f(float v) {
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
  return v;
}
// Sum of int64 values over the lanes of a wave (exact, order-free).
__device__ __forceinline__ long long wave_sum_ll(long long v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}
__device__ __forceinline__ double wave_sum_d(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}
__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// wrapped coordinate in [0, L) relative to the box's lower corner
__device__ __forceinline__ double wrap_rel(double x, double omin, double L) {
  double r = x - omin;
  return r - L * floor(r / L);
}

// PeriodicBoundaryConditions::minimalDifference(c1, c2, r2), orthogonal branch
// (PeriodicBoundaryConditions.h:72-110), evaluated without contraction so the
// squared distance rounds exactly as the reference's FP64 arithmetic.
__device__ __forceinline__ double min_image_r2(const SysDev &S, int a, int b) {
#pragma clang fp contract(off)
  const int n = S.n_atoms;
  double d0 = S.x[b] - S.x[a];
  double d1 = S.x[n + b] - S.x[n + a];
  double d2 = S.x[2 * n + b] - S.x[2 * n + a];
  double r2 = d0 * d0 + d1 * d1 + d2 * d2;
  if (S.periodic && r2 > S.D) {
    if (d0 < -S.H2[0]) d0 += S.H[0];
    else if (d0 > S.H2[0]) d0 -= S.H[0];
    if (d1 < -S.H2[1]) d1 += S.H[1];
    else if (d1 > S.H2[1]) d1 -= S.H[1];
    if (d2 < -S.H2[2]) d2 += S.H[2];
    else if (d2 > S.H2[2]) d2 -= S.H[2];
    r2 = d0 * d0 + d1 * d1 + d2 * d2;
    if (r2 > S.D) {
      d0 -= S.L[0] * rint(S.Lr[0] * d0);
      d1 -= S.L[1] * rint(S.Lr[1] * d1);
      d2 -= S.L[2] * rint(S.Lr[2] * d2);
      r2 = d0 * d0 + d1 * d1 + d2 * d2;
    }
  }
  return r2;
}

// the same procedure on a difference vector in place (bonded terms, the
// molecular centres of the molecular virial)
__device__ __forceinline__ void min_image_vec(const SysDev &S, double *d) {
  double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
  if (!S.periodic || r2 <= S.D) return;
  for (int k = 0; k < 3; ++k) {
    if (d[k] < -S.H2[k]) d[k] += S.H[k];
    else if (d[k] > S.H2[k]) d[k] -= S.H[k];
  }
  r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
  if (r2 > S.D)
    for (int k = 0; k < 3; ++k) d[k] -= S.L[k] * rint(S.Lr[k] * d[k]);
}

// Cn coefficient tables (CnSwitchingFunction.cpp:9-31), rows = order 2,3,4,(5),6
static __constant__ double kSwCoef[5][7] = {{10., -15., 6., 0, 0, 0, 0},
                                     {35., -84., 70., -20., 0, 0, 0},
                                     {126., -420., 540., -315., 70., 0, 0},
                                     {0, 0, 0, 0, 0, 0, 0},
                                     {1716., -9009., 20020., -24024., 16380., -6006., 924.0}};
static __constant__ double kDSwCoef[5][7] = {{30., -60., 30., 0, 0, 0, 0},
                                      {140., -420., 420., -140., 0, 0, 0},
                                      {630., -2520., 3780., -2520., 630., 0, 0},
                                      {0, 0, 0, 0, 0, 0, 0},
                                      {12012., -72072., 180180., -240240., 180180., -72072., 12012.0}};

// sum_i coef[i] t^(ord+1+i) r[i] and sum_i dcoef[i] t^(ord+i) r[i], running powers
__device__ __forceinline__ void cn_poly(const SwitchDev &sw, double t, const double *r, double &val, double &der) {
  double p = 1.0;
  for (int k = 0; k < sw.ord_int; ++k) p *= t;
  val = 0.0;
  der = 0.0;
  for (int i = 0; i < sw.ord_int + 1; ++i) {
    der += kDSwCoef[sw.ord_idx][i] * p * r[i];
    p *= t;
    val += kSwCoef[sw.ord_idx][i] * p * r[i];
  }
}

__device__ __forceinline__ void switch_cn_d(const SwitchDev &sw, double r2, double &val, double &der) {
  // CnSwitchingFunction::operator() (CnSwitchingFunction.cpp:149-175) and the
  // two branches of CmpCnCnSwitchingFunction::operator() (CmpCnCnSwitchingFunction.h:52-95)
  val = 0.0;
  der = 0.0;
  if (r2 > sw.dcut2) return;
  const double sd = sqrt(r2);
  if (sw.kind == PM_SWITCH_CN) {
    if (r2 < sw.don2) {
      val = 1.0;
      return;
    }
    cn_poly(sw, sd - sw.dcut, sw.irange, val, der);
    der /= sd;
    return;
  }
  if (r2 > sw.doff2) {
    cn_poly(sw, sd - sw.dcut, sw.itrange, val, der);
    der /= sd;
  } else if (r2 >= sw.don2) {
    double v, d;
    cn_poly(sw, sd - sw.doff, sw.irange, v, d);
    val = 1.0 - v;
    der = -d / sd;
  }
}

// switched force/energy: F = F*S - E*dS, E = E*S (OneAtomPair.h:76-84)
__device__ __forceinline__ void apply_switch(const SwitchDev &sw, float r2, float &E, float &F) {
  switch (sw.kind) {
  case PM_SWITCH_C2: {  // C2SwitchingFunction.h:42-53
    float c2 = sw.cut2 - r2;
    float c4 = c2 * (sw.s2 + 2.0f * r2);
    float S = sw.s1 * (c2 * c4);
    float dS = sw.s3 * (c2 * c2 - c4);
    bool inner = r2 < sw.on2;
    bool outer = r2 > sw.cut2;
    S = inner ? 1.0f : (outer ? 0.0f : S);
    dS = (inner || outer) ? 0.0f : dS;
    F = F * S - E * dS;
    E = E * S;
    break;
  }
  case PM_SWITCH_C1: {  // C1SwitchingFunction.h:41-50
    float d = r2 * rsqrtf(r2);
    float S = 1.0f - d * (sw.c15rc1 - r2 * sw.c05rc3);
    float dS = -sw.c15rc1 / d + d * sw.c15rc3;
    bool outer = r2 > sw.cut2;
    S = outer ? 0.0f : S;
    dS = outer ? 0.0f : dS;
    F = F * S - E * dS;
    E = E * S;
    break;
  }
  case PM_SWITCH_CUTOFF:
    // CutoffSwitchingFunction (CutoffSwitchingFunction.h:41-44) is 1 up to its
    // cutoff, which is this term's inclusion cutoff: the FP64-checked lane
    // masks have already decided it, so nothing is left to scale
    break;
  case PM_SWITCH_CN:
  case PM_SWITCH_CMPCNCN: {
    double S, dS;
    switch_cn_d(sw, (double)r2, S, dS);
    F = (float)((double)F * S - (double)E * dS);
    E = (float)((double)E * S);
    break;
  }
  default:
    break;
  }
}

}  // namespace pm
