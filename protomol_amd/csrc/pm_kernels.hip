// pm_kernels.hip -- gfx950 kernels of the nonbonded force path.
//
//   rebuild:  k_bin -> k_col_scan -> k_col_fill -> k_col_sort -> k_cluster_geom -> k_pairlist
//   force:    k_pair<TermA, TermB, SwitchA, SwitchB, Kind>, k_local (bonded + Ewald correction)
//   step:     k_kick_drift, k_half_kick, k_export_f, k_kinetic
//
// All kernels take the device table of SysDev records plus a prefix array
// over the systems, so one launch covers every system of a batch.
#include "pm_kernels.h"

#include <hip/hip_runtime.h>
#include <math.h>

#include <algorithm>

#include "../../include/protomol_amd.h"

namespace pm {

// ------------------------------------------------------------- helpers

// Pointers read out of the SysDev table are generic; casting them to the
// global (1) or constant (4) address space lets the compiler use global_* and,
// for wave-uniform reads of data no kernel writes meanwhile, s_load_*.
// (the host pass of hipcc sees kernel bodies too, without address spaces)
#if defined(__HIP_DEVICE_COMPILE__)
#define PM_GLOBAL __attribute__((address_space(1)))
#define PM_CONST __attribute__((address_space(4)))
#else
#define PM_GLOBAL
#define PM_CONST
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
__constant__ double kSwCoef[5][7] = {{10., -15., 6., 0, 0, 0, 0},
                                     {35., -84., 70., -20., 0, 0, 0},
                                     {126., -420., 540., -315., 70., 0, 0},
                                     {0, 0, 0, 0, 0, 0, 0},
                                     {1716., -9009., 20020., -24024., 16380., -6006., 924.0}};
__constant__ double kDSwCoef[5][7] = {{30., -60., 30., 0, 0, 0, 0},
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

// ------------------------------------------------------------- rebuild

// host staging (AoS FP64, per system) -> device state; what: kXfer* bits
__global__ void k_import(SysDev *sys, const int *ap, int nsys, int what) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ap[nsys]) return;
  const int s = find_sys(ap, nsys, g);
  SysDev &S = sys[s];
  const int a = g - ap[s], n = S.n_atoms;
  for (int d = 0; d < 3; ++d) {
    if (what & kXferX) S.x[d * n + a] = S.xaos[3 * (size_t)a + d];
    if (what & kXferV) S.v[d * n + a] = S.vaos[3 * (size_t)a + d];
  }
}

// Lanes of a wave that share a (system, column) key are served by one atomic:
// the wave peels off the group of the first remaining lane until every lane is
// done.  Atoms in original order are spatially coherent, so a wave usually
// holds only a few keys.
__global__ void k_bin(SysDev *sys, const int *ap, int nsys) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  bool todo = g < ap[nsys];
  int s = 0, col = 0;
  if (todo) {
    s = find_sys(ap, nsys, g);
    todo = sys[s].ctl[kCtlRebuild] != 0;
  }
  if (todo) {
    SysDev &S = sys[s];
    const int a = g - ap[s], n = S.n_atoms;
    double x = S.x[a], y = S.x[n + a], z = S.x[2 * n + a];
    if (!(isfinite(x) && isfinite(y) && isfinite(z))) {
      atomicOr(S.err, (unsigned)kErrNonFinite);
      x = y = z = 0.0;
    }
    const double rx = wrap_rel(x, S.omin[0], S.L[0]);
    const double ry = wrap_rel(y, S.omin[1], S.L[1]);
    const double rz = wrap_rel(z, S.omin[2], S.L[2]);
    int cx = (int)(rx / S.ax), cy = (int)(ry / S.ay);
    cx = cx < 0 ? 0 : (cx >= S.ncx ? S.ncx - 1 : cx);
    cy = cy < 0 ? 0 : (cy >= S.ncy ? S.ncy - 1 : cy);
    col = cy * S.ncx + cx;
    S.bin_col[a] = col;
    S.bin_z[a] = (float)rz;
  }
  // one atomic per (system, column) group of the wave
  while (true) {
    const unsigned long long rem = wave_ballot(todo);
    if (rem == 0) break;
    const int leader = __ffsll((long long)rem) - 1;
    const int ls = __shfl(s, leader), lc = __shfl(col, leader);
    const bool mine = todo && s == ls && col == lc;
    const unsigned long long grp = wave_ballot(mine);
    if (lane == leader) atomicAdd(&sys[ls].col_count[lc], __popcll(grp));
    todo = todo && !mine;
  }
}

// one block per system: exclusive scans of atoms and clusters over columns
__global__ __launch_bounds__(1024) void k_col_scan(SysDev *sys) {
  SysDev &S = sys[blockIdx.x];
  if (!S.ctl[kCtlRebuild]) return;
  __shared__ int sa[1024], sc[1024], ss[1024];
  const int t = threadIdx.x, nt = blockDim.x;
  const int ncol = S.ncol;
  const int chunk = (ncol + nt - 1) / nt;
  const int c0 = t * chunk, c1 = min(ncol, c0 + chunk);
  int la = 0, lc = 0, ls = 0;
  for (int c = c0; c < c1; ++c) {
    const int n = S.col_count[c];
    const int k = (n + kCluster - 1) / kCluster;
    la += n;
    lc += k;
    ls += (k + kScCl - 1) / kScCl;
  }
  sa[t] = la;
  sc[t] = lc;
  ss[t] = ls;
  __syncthreads();
  for (int o = 1; o < nt; o <<= 1) {
    const int va = t >= o ? sa[t - o] : 0, vc = t >= o ? sc[t - o] : 0, vs = t >= o ? ss[t - o] : 0;
    __syncthreads();
    sa[t] += va;
    sc[t] += vc;
    ss[t] += vs;
    __syncthreads();
  }
  int ba = sa[t] - la, bc = sc[t] - lc, bs = ss[t] - ls;
  for (int c = c0; c < c1; ++c) {
    const int n = S.col_count[c];
    const int k = (n + kCluster - 1) / kCluster;
    const int q = (k + kScCl - 1) / kScCl;
    S.col_atom_start[c] = ba;
    S.col_cl_start[c] = bc;
    S.col_ncl[c] = k;
    S.col_sc_start[c] = bs;
    S.col_nsc[c] = q;
    S.col_cursor[c] = 0;
    for (int j = 0; j < q; ++j)
      if (bs + j < S.sc_cap) S.sc_col[bs + j] = c;
    ba += n;
    bc += k;
    bs += q;
  }
  if (t == nt - 1) {
    int tot = sc[t], tsc = ss[t];
    if (tot > S.cl_cap || tsc > S.sc_cap) {
      atomicOr(S.err, (unsigned)kErrClusterCap);
      tot = min(tot, S.cl_cap);
      tsc = min(tsc, S.sc_cap);
    }
    *S.n_clusters = tot;
    *S.n_sc = tsc;
  }
}

__global__ void k_col_fill(SysDev *sys, const int *ap, int nsys) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  bool todo = g < ap[nsys];
  int s = 0, col = 0, a = 0;
  if (todo) {
    s = find_sys(ap, nsys, g);
    todo = sys[s].ctl[kCtlRebuild] != 0;
  }
  if (todo) {
    a = g - ap[s];
    col = sys[s].bin_col[a];
  }
  while (true) {
    const unsigned long long rem = wave_ballot(todo);
    if (rem == 0) break;
    const int leader = __ffsll((long long)rem) - 1;
    const int ls = __shfl(s, leader), lc = __shfl(col, leader);
    const bool mine = todo && s == ls && col == lc;
    const unsigned long long grp = wave_ballot(mine);
    int base = 0;
    if (lane == leader) base = atomicAdd(&sys[ls].col_cursor[lc], __popcll(grp));
    base = __shfl(base, leader);
    if (mine) {
      SysDev &S = sys[s];
      const int pos = base + __popcll(grp & ((1ull << lane) - 1ull));
      const unsigned zb = __float_as_uint(S.bin_z[a]);  // z >= 0: bit order == value order
      S.colbuf[S.col_atom_start[col] + pos] = ((unsigned long long)zb << 32) | (unsigned)a;
    }
    todo = todo && !mine;
  }
}

// one block per column: bitonic sort by (z, atom) and slot assignment
__global__ __launch_bounds__(1024) void k_col_sort(SysDev *sys, const int *cp, int nsys) {
  __shared__ unsigned long long keys[kSortCap];
  int gcol = blockIdx.x;
  if (gcol >= cp[nsys]) return;
  int s = find_sys(cp, nsys, gcol);
  SysDev &S = sys[s];
  if (!S.ctl[kCtlRebuild]) return;
  int col = gcol - cp[s];
  const int n = S.col_count[col];
  const int t = threadIdx.x, nt = blockDim.x;
  if (n > kSortCap) {
    if (t == 0) atomicOr(S.err, (unsigned)kErrColumnCap);
    return;
  }
  const int start = S.col_atom_start[col];
  int P = 1;
  while (P < n) P <<= 1;
  for (int i = t; i < P; i += nt) keys[i] = i < n ? S.colbuf[start + i] : ~0ull;
  __syncthreads();
  for (int k = 2; k <= P; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = t; i < P; i += nt) {
        int l = i ^ j;
        if (l > i) {
          unsigned long long a = keys[i], b = keys[l];
          bool up = (i & k) == 0;
          if ((a > b) == up) {
            keys[i] = b;
            keys[l] = a;
          }
        }
      }
      __syncthreads();
    }
  }
  const int cl0 = S.col_cl_start[col], ncl = S.col_ncl[col];
  for (int k = t; k < ncl * kCluster; k += nt) {
    int slot = cl0 * kCluster + k;
    if (k < n) {
      int a = (int)(keys[k] & 0xffffffffull);
      S.slot[a] = slot;
      S.orig[slot] = a;
      S.smol[slot] = S.molecule[a];
      if (S.lj_mode == PM_LJ_PERATOM) S.ljp[slot] = make_float2(S.sig_half[a], S.eps_sqrt[a]);
      else S.ljt[slot] = S.type[a];
    } else {
      S.orig[slot] = -1;
      S.smol[slot] = -1;
      if (S.lj_mode == PM_LJ_PERATOM) S.ljp[slot] = make_float2(0.f, 0.f);
      else S.ljt[slot] = 0;
      S.xq[slot] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
  }
  for (int c = t; c < ncl; c += nt) {
    S.cl_col[cl0 + c] = col;
    S.cl_nvalid[cl0 + c] = min(kCluster, n - c * kCluster);
  }
  __syncthreads();
  if (t == 0) S.col_count[col] = 0;  // ready for the next rebuild
}

// one thread per cluster: bounding box, origin, local coordinates
__global__ void k_cluster_geom(SysDev *sys, const int *clp, int nsys) {
  int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= clp[nsys]) return;
  int s = find_sys(clp, nsys, g);
  SysDev &S = sys[s];
  int c = g - clp[s];
  if (!S.ctl[kCtlRebuild] || c >= *S.n_clusters) return;
  const int n = S.n_atoms, nv = S.cl_nvalid[c];
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  double rel[kCluster][3];
  int at[kCluster];
  for (int k = 0; k < nv; ++k) {
    int a = S.orig[c * kCluster + k];
    at[k] = a;
    for (int d = 0; d < 3; ++d) {
      double r = wrap_rel(S.x[d * n + a], S.omin[d], S.L[d]);
      rel[k][d] = r;
      lo[d] = fmin(lo[d], r);
      hi[d] = fmax(hi[d], r);
    }
  }
  double org[3];
  for (int d = 0; d < 3; ++d) {
    org[d] = 0.5 * (lo[d] + hi[d]);
    S.cl_origin[d * S.cl_cap + c] = org[d];
  }
  S.cl_lo[c] = make_float4((float)lo[0] - 1e-3f, (float)lo[1] - 1e-3f, (float)lo[2] - 1e-3f, 0.f);
  S.cl_hi[c] = make_float4((float)hi[0] + 1e-3f, (float)hi[1] + 1e-3f, (float)hi[2] + 1e-3f, 0.f);
  for (int k = 0; k < nv; ++k) {
    int a = at[k];
    float l[3];
    for (int d = 0; d < 3; ++d) {
      double xa = S.x[d * n + a];
      double ref = S.omin[d] + org[d] + (xa - S.omin[d] - rel[k][d]);
      S.xref[d * n + a] = ref;
      l[d] = (float)(xa - ref);
    }
    S.xq[c * kCluster + k] = make_float4(l[0], l[1], l[2], (float)S.charge[a]);
    S.lbuild[a] = make_float4(l[0], l[1], l[2], 0.f);
  }
}

__device__ __forceinline__ int floordiv(int a, int b) {
  int q = a / b;
  return (a % b != 0 && ((a < 0) != (b < 0))) ? q - 1 : q;
}

__device__ __forceinline__ int excl_lookup(const SysDev &S, int a, int b) {
  int e0 = S.excl_start[a], e1 = S.excl_start[a + 1];
  for (int k = e0; k < e1; ++k)
    if (S.excl_partner[k] == b) return S.excl_class[k];
  return 0;
}

// Super-cluster pair list within rlist, half shell over columns.  A
// super-cluster is kScCl z-consecutive clusters of one column; its list holds
// the j clusters within rlist of any of them, with an 8-bit mask of which.
// One workgroup per column: the boxes of the j clusters in the column's window
// of neighbour columns (periodic x/y image applied) are staged in LDS once;
// each wave then tests its super-clusters' clusters against them, one z image
// at a time.  Windows larger than the LDS stage take several passes.  A second
// pass per super-cluster fills in the entries' FP32 image offsets and the
// FULL/MODIFIED lane masks of cluster pairs that share a molecule.
constexpr int kListThreads = 1024;
constexpr int kListWaves = kListThreads / kWave;
constexpr int kCandCap = 1280;   // staged j clusters per pass
constexpr int kWinCap = 128;     // neighbour columns per pass
constexpr int kColCap = kSortCap / kCluster;  // clusters of one column
constexpr int kColSc = kColCap / kScCl;       // super-clusters of one column
static_assert(kColCap < kCandCap, "a column's clusters must fit one LDS stage");

struct ListScratch {  // per wave, in the candidate stage's LDS once pass A is over
  int mol[kWave], at[kWave];
  unsigned char rowF[kScCl][kScCl], rowM[kScCl][kScCl];
};

__global__ __launch_bounds__(kListThreads) void k_pairlist(SysDev *sys, const int *colp, int nsys, float rlist2) {
  __shared__ float4 c_lo[kCandCap], c_hi[kCandCap];
  __shared__ int c_id[kCandCap];             // cj | (nvalid - 1) << 28
  __shared__ int w_col[kWinCap], w_start[kWinCap + 1], w_sxy[kWinCap];
  __shared__ int n_win, next_q, sc_count[kColSc];
  __shared__ float4 ib_lo[kListWaves][kScCl], ib_hi[kListWaves][kScCl];
  static_assert(sizeof(ListScratch) * kListWaves <= sizeof(float4) * kCandCap, "scratch fits the stage");
  const int gcol = blockIdx.x;
  if (gcol >= colp[nsys]) return;
  const int s = find_sys(colp, nsys, gcol);
  SysDev &S = sys[s];
  if (!S.ctl[kCtlRebuild]) return;
  const int coli = gcol - colp[s];
  const int ci0 = S.col_cl_start[coli], nci = S.col_ncl[coli];
  if (nci == 0) return;
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  if (nci > kColCap) {  // the column sort has already flagged it
    if (t == 0) atomicOr(S.err, (unsigned)kErrColumnCap);
    return;
  }
  const int nsc = S.col_nsc[coli], sc0 = S.col_sc_start[coli];
  const int cxi = coli % S.ncx, cyi = coli / S.ncx;
  const float Lx = (float)S.L[0], Ly = (float)S.L[1], Lz = (float)S.L[2];
  const float ax = (float)S.ax, ay = (float)S.ay;
  const float ix0 = cxi * ax, ix1 = (cxi + 1) * ax, iy0 = cyi * ay, iy1 = (cyi + 1) * ay;
  const int cap = S.list_cap;
  const int wx = 2 * S.rx + 1, nwin_all = wx * (2 * S.ry + 1);
  for (int k = t; k < nsc; k += kListThreads) sc_count[k] = 0;
  if (t == 0) next_q = 0;
  __syncthreads();
  // ---- pass A: entries (j cluster, image, i-cluster mask)
  while (true) {
    if (t == 0) {  // this pass's neighbour columns, in window order
      int q = next_q, nw = 0, tot = 0;
      w_start[0] = 0;
      for (; q < nwin_all && nw < kWinCap; ++q) {
        const int dy = q / wx - S.ry, dx = q % wx - S.rx;
        if (dy < 0 || (dy == 0 && dx < 0)) continue;  // half shell over columns
        const int cyu = cyi + dy, cxu = cxi + dx;
        const float gy = fmaxf(0.f, fmaxf(cyu * ay - iy1, iy0 - (cyu + 1) * ay) - 0.01f);
        const float gx = fmaxf(0.f, fmaxf(cxu * ax - ix1, ix0 - (cxu + 1) * ax) - 0.01f);
        if (gx * gx + gy * gy >= rlist2) continue;
        const int sy = floordiv(cyu, S.ncy), sx = floordiv(cxu, S.ncx);
        const int colj = (cyu - sy * S.ncy) * S.ncx + (cxu - sx * S.ncx);
        const int nc = S.col_ncl[colj];
        if (tot + nc > kCandCap) break;  // next pass (nc <= kColCap < kCandCap: a pass never ends empty)
        w_col[nw] = colj;
        w_sxy[nw] = (sx + 1) | ((sy + 1) << 2) | ((dx == 0 && dy == 0) ? 16 : 0);
        tot += nc;
        w_start[++nw] = tot;
      }
      n_win = nw;
      next_q = q;
    }
    __syncthreads();
    const int nw = n_win, ncand = w_start[nw];
    for (int k = t; k < ncand; k += kListThreads) {
      int w = 0;
      while (w_start[w + 1] <= k) ++w;
      const int cj = S.col_cl_start[w_col[w]] + (k - w_start[w]);
      const int sxy = w_sxy[w];
      const float sxL = ((sxy & 3) - 1) * Lx, syL = (((sxy >> 2) & 3) - 1) * Ly;
      const int nvj = S.cl_nvalid[cj];
      float4 lo = S.cl_lo[cj], hi = S.cl_hi[cj];
      lo.x += sxL; hi.x += sxL;
      lo.y += syL; hi.y += syL;
      lo.w = __int_as_float(sxy);
      if (nvj == 0) lo.x = 1e30f;  // never within range
      c_lo[k] = lo;
      c_hi[k] = hi;
      c_id[k] = cj | (max(nvj - 1, 0) << 28);
    }
    __syncthreads();
    for (int scl = wv; scl < nsc; scl += kListWaves) {
      const int cl0 = ci0 + scl * kScCl, nic = min(kScCl, nci - scl * kScCl);
      if (lane < kScCl) {
        const bool ok = lane < nic && S.cl_nvalid[cl0 + lane] > 0;
        ib_lo[wv][lane] = ok ? S.cl_lo[cl0 + lane] : make_float4(1e30f, 1e30f, 1e30f, 0.f);
        ib_hi[wv][lane] = ok ? S.cl_hi[cl0 + lane] : make_float4(1e30f, 1e30f, 1e30f, 0.f);
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      float zlo = 1e30f, zhi = -1e30f;
      for (int c = 0; c < nic; ++c) {
        zlo = fminf(zlo, ib_lo[wv][c].z);
        zhi = fmaxf(zhi, ib_hi[wv][c].z);
      }
      uint4 *out = S.list + 2 * (size_t)(sc0 + scl) * cap;
      int count = sc_count[scl];
      for (int sz = -1; sz <= 1; ++sz) {
        // a z image is a candidate only if the super-cluster's z range comes
        // within rlist of that image of the wrapped frame [0, Lz)
        if (sz == 1 && fmaxf(0.f, Lz - zhi - 0.01f) * fmaxf(0.f, Lz - zhi - 0.01f) >= rlist2) continue;
        if (sz == -1 && fmaxf(0.f, zlo - 0.01f) * fmaxf(0.f, zlo - 0.01f) >= rlist2) continue;
        const float szL = sz * Lz;
        for (int base = 0; base < ncand; base += kWave) {
          const int k = base + lane;
          unsigned im = 0;
          int cj = 0, sxy = 0, idw = 0;
          if (k < ncand) {
            const float4 loj = c_lo[k], hij = c_hi[k];
            idw = c_id[k];
            cj = idw & 0x7FFFFF;
            sxy = __float_as_int(loj.w);
            const bool samecol = (sxy & 16) != 0;
            for (int c = 0; c < nic; ++c) {
              // each cluster pair once: within a column, j at or above i (z image 0)
              if (samecol && !(sz > 0 || (sz == 0 && cj >= cl0 + c))) continue;
              const float4 loi = ib_lo[wv][c], hii = ib_hi[wv][c];
              const float ex = fmaxf(0.f, fmaxf(loj.x - hii.x, loi.x - hij.x));
              const float ey = fmaxf(0.f, fmaxf(loj.y - hii.y, loi.y - hij.y));
              const float ez = fmaxf(0.f, fmaxf(loj.z + szL - hii.z, loi.z - (hij.z + szL)));
              if (ex * ex + ey * ey + ez * ez < rlist2) im |= 1u << c;
            }
          }
          const bool hit = im != 0;
          const unsigned long long b = wave_ballot(hit);
          if (b == 0) continue;
          const int pos = count + __popcll(b & ((1ull << lane) - 1ull));
          if (hit && pos < cap) {
            // the image offset and masks are filled in by pass B
            const int sh = (sxy & 3) * 9 + ((sxy >> 2) & 3) * 3 + (sz + 1);
            out[2 * pos].x = (unsigned)cj | ((unsigned)sh << 23) | ((unsigned)idw & 0x70000000u);
            out[2 * pos + 1] = make_uint4(im, 0u, 0u, 0u);
          }
          count += __popcll(b);
        }
      }
      if (lane == 0) sc_count[scl] = count;
    }
    __syncthreads();
    if (next_q >= nwin_all) break;
    __syncthreads();
  }
  // ---- pass B, per super-cluster: offsets, exclusion masks, lengths
  ListScratch &W = reinterpret_cast<ListScratch *>(c_lo)[wv];
  long long pairs8 = 0;  // 8x8 cluster pairs, for the statistics
  for (int scl = wv; scl < nsc; scl += kListWaves) {
    const int sc = sc0 + scl;
    const int cl0 = ci0 + scl * kScCl, nic = min(kScCl, nci - scl * kScCl);
    int count = sc_count[scl];
    if (count > cap) {
      if (lane == 0) atomicOr(S.err, (unsigned)kErrListCap);
      count = cap;
    }
    uint4 *out = S.list + 2 * (size_t)sc * cap;
    const double ox = S.cl_origin[cl0], oy = S.cl_origin[S.cl_cap + cl0], oz = S.cl_origin[2 * S.cl_cap + cl0];
    {
      const int c = lane >> 3, ii = lane & 7;
      int mol = -1, at = -1;
      if (c < nic) {
        const int si = (cl0 + c) * kCluster + ii;
        mol = S.smol[si];
        at = S.orig[si];
      }
      W.mol[lane] = mol;
      W.at[lane] = at;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    int mslot = 0;
    unsigned long long *mbase = S.list_mask + (size_t)sc * S.mask_cap * (2 * kScCl);
    for (int e0 = 0; e0 < count; e0 += kWave) {
      const int e = e0 + lane;
      unsigned share = 0;
      int cj = 0;
      if (e < count) {
        const unsigned x = out[2 * e].x;
        const unsigned im = out[2 * e + 1].x & 0xFFu;
        cj = (int)(x & 0x7FFFFFu);
        const int sh = (int)((x >> 23) & 31u);
        const int sx = sh / 9 - 1, sy = (sh / 3) % 3 - 1, sz = sh % 3 - 1;
        const double ojx = S.cl_origin[cj] + (double)sx * S.L[0];
        const double ojy = S.cl_origin[S.cl_cap + cj] + (double)sy * S.L[1];
        const double ojz = S.cl_origin[2 * S.cl_cap + cj] + (double)sz * S.L[2];
        out[2 * e].y = __float_as_uint((float)(ojx - ox));
        out[2 * e].z = __float_as_uint((float)(ojy - oy));
        out[2 * e].w = __float_as_uint((float)(ojz - oz));
        pairs8 += __popc(im);
        const int4 m0 = *reinterpret_cast<const int4 *>(S.smol + cj * kCluster);
        const int4 m1 = *reinterpret_cast<const int4 *>(S.smol + cj * kCluster + 4);
        const int mj[kCluster] = {m0.x, m0.y, m0.z, m0.w, m1.x, m1.y, m1.z, m1.w};
        for (int c = 0; c < kScCl; ++c) {
          bool sh_c = false;
          for (int b = 0; b < kCluster; ++b) {
            const int mi = W.mol[c * kCluster + b];
            for (int a = 0; a < kCluster; ++a) sh_c = sh_c || (mi >= 0 && mj[a] == mi);
          }
          if (sh_c && ((im >> c) & 1u)) share |= 1u << c;
        }
      }
      unsigned long long flagged = wave_ballot(share != 0);
      while (flagged) {
        const int g = __ffsll((long long)flagged) - 1;
        flagged &= flagged - 1;
        const unsigned sh_e = __shfl(share, g);
        const int cj_e = __shfl(cj, g);
        const int ee = e0 + g;
        if (mslot >= S.mask_cap) {
          if (lane == 0) atomicOr(S.err, (unsigned)kErrListCap);
          continue;
        }
        // lanes (k, jj): the k-th i cluster sharing a molecule with j cluster cj_e
        const int k = lane >> 3, jj = lane & 7;
        const int ncomb = __popc(sh_e);
        int ck = -1;
        {
          unsigned m = sh_e;
          for (int r = 0; r < k && m; ++r) m &= m - 1;
          if (m && k < ncomb) ck = __ffs(m) - 1;
        }
        unsigned rf = 0, rm = 0;
        if (ck >= 0) {
          const int sj = cj_e * kCluster + jj;
          const int mj = S.smol[sj], bj = S.orig[sj];
          // j atom's exclusions as index-window bits (one load pair per lane);
          // an atom with a partner outside the window looks up the CSR
          const unsigned long long wf = bj >= 0 ? S.excl_win[2 * (size_t)bj] : 0ull;
          const unsigned long long wm = bj >= 0 ? S.excl_win[2 * (size_t)bj + 1] : 0ull;
          const bool far = (wf >> 63) != 0;
          for (int ii = 0; ii < kCluster; ++ii) {
            const int mi = W.mol[ck * kCluster + ii], ai = W.at[ck * kCluster + ii];
            if (mi >= 0 && mi == mj && ai != bj) {
              const int dd = ai - bj + 31;
              const bool inw = dd >= 0 && dd < 63;
              const int cls = far ? excl_lookup(S, ai, bj)
                                  : (inw && ((wf >> dd) & 1ull) ? PM_EXCL_FULL
                                     : (inw && ((wm >> dd) & 1ull) ? PM_EXCL_MODIFIED : 0));
              if (cls == PM_EXCL_FULL) rf |= 1u << ii;
              if (cls == PM_EXCL_MODIFIED) rm |= 1u << ii;
            }
          }
          W.rowF[k][jj] = (unsigned char)rf;
          W.rowM[k][jj] = (unsigned char)rm;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        // 16 words of the slot: FULL, MODIFIED lane masks (bit jj * 8 + ii) per i cluster
        if (lane < 2 * kScCl) {
          const int c = lane >> 1;
          unsigned long long v = 0;
          if ((sh_e >> c) & 1u) {
            const int kc = __popc(sh_e & ((1u << c) - 1u));
            v = (lane & 1) ? *reinterpret_cast<const unsigned long long *>(W.rowM[kc])
                           : *reinterpret_cast<const unsigned long long *>(W.rowF[kc]);
          }
          mbase[(size_t)mslot * (2 * kScCl) + lane] = v;
        }
        if (lane == 0) {
          out[2 * ee].x |= 0x80000000u;
          out[2 * ee + 1].x |= (unsigned)mslot << 8;
        }
        ++mslot;
        __builtin_amdgcn_wave_barrier();
      }
    }
    if (lane == 0) S.list_len[sc] = count;
  }
  pairs8 = wave_sum_ll(pairs8);
  if (lane == 0 && pairs8) atomicAdd(&S.ctl[kCtlEntries], (int)pairs8);
}

// ------------------------------------------------------------ pair force

// compile-time switch kinds for the hot path; kSwGeneric dispatches at run time
constexpr int kSwGeneric = -1;

// OUTER: the switch's cutoff may be shorter than the pair's inclusion cutoff
// (OneAtomPairTwo with two different cutoffs), so S = 0 beyond it
template <int SW, bool OUTER = true>
__device__ __forceinline__ void apply_switch_t(const SwitchDev &sw, float r2, float &E, float &F) {
  if (SW == PM_SWITCH_C2) {  // C2SwitchingFunction.h:42-53
    // inside r_on the switch is 1 with zero derivative: the unswitched values
    // are kept as they are (one select on each result)
    const float c2 = sw.cut2 - r2;
    const float c4 = c2 * (sw.s2 + 2.0f * r2);
    const float S = sw.s1 * (c2 * c4);
    const float dS = sw.s3 * (c2 * c2 - c4);
    const bool inner = r2 < sw.on2;
    float Fs = F * S - E * dS, Es = E * S;
    if (OUTER) {
      const bool outer = r2 > sw.cut2;
      Fs = outer ? 0.0f : Fs;
      Es = outer ? 0.0f : Es;
    }
    F = inner ? F : Fs;
    E = inner ? E : Es;
  } else if (SW == PM_SWITCH_CUTOFF || SW == PM_SWITCH_NONE) {
    // inside the inclusion cutoff the value is 1 (CutoffSwitchingFunction.h:41-44)
  } else {
    apply_switch(sw, r2, E, F);
  }
}

// sum over 8 consecutive lanes with DPP (quad swaps, then half-row mirror);
// every lane of the group ends with the same bits
__device__ __forceinline__ float dpp_sum8(float v) {
  v += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0xB1, 0xF, 0xF, false));
  v += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x4E, 0xF, 0xF, false));
  v += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x141, 0xF, 0xF, false));
  return v;
}

// Sums of (x, y, z) over each group of 8 consecutive lanes, left in lane 0
// (x), lane 1 (z) and lane 2 (y) of the group: halving steps over lane pairs
// (quad swap), quad halves, then the group's two quads.  Fixed order, so the
// result bits do not depend on anything but the inputs.
__device__ __forceinline__ float dpp_sum8_scatter(float x, float y, float z, int g) {
  const bool odd = g & 1;
  // step 1 (lane ^ 1): even lanes keep x and y, odd lanes keep z
  const float send1 = odd ? x : z;
  const float r1 = __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(send1), 0xB1, 0xF, 0xF, false));
  const float r2 = __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(y), 0xB1, 0xF, 0xF, false));
  const float a = (odd ? z : x) + r1;  // even: x pair sum, odd: z pair sum
  const float b = y + r2;              // even: y pair sum
  // step 2 (lane ^ 2): lane 0 keeps x, lane 2 keeps y, lane 1 keeps z
  const bool q0 = (g & 3) == 0;
  const float send2 = q0 ? b : a;
  const float r3 = __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(send2), 0x4E, 0xF, 0xF, false));
  const float c = ((g & 3) == 2 ? b : a) + r3;
  // step 3: lanes 0..2 of the group add lanes 4..6 (row shift left by 4)
  const float r4 = __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(c), 0x104, 0xF, 0xF, false));
  return c + r4;
}

constexpr int kLJLds = 256;  // LJ table entries kept in LDS per wave (T <= 16; larger tables are read from L1/L2)
// LDS per wave of the pair kernel: the LJ table as (A, B, 12 A, 6 B), then the
// super-cluster's 64 i atoms (position + charge, byte offset of the atom's
// LJ table row, atom index, per-atom LJ)
constexpr int kPairWaveLds = kLJLds * 16 + kWave * (16 + 4 + 4 + 8);

// One pair term set (OneAtomPair[Two]::doOneAtomPair, OneAtomPair.h:47-107):
// the scalar force factor F/r for the pair, energies into ea / eb.  qq is the
// charge product with the modified-pair scale already applied; A, B are the
// pair's LJ coefficients and A12 = 12 A, B6 = 6 B.
template <int TA, int TB, int SWA, int SWB, bool SAME>
__device__ __forceinline__ float pair_terms(const ForceDev &P, float r2, float qq, float A, float B, float A12,
                                            float B6, bool inA, bool inB, float &ea, float &eb) {
  const float rinv = __builtin_amdgcn_rsqf(r2);
  const float r2inv = rinv * rinv;
  float fs = 0.f;
  ea = 0.f;
  eb = 0.f;
  if (TA == PM_TERM_LJ) {
    // LennardJonesForce::operator() (LennardJonesForce.h:45-58):
    // E = A r^-12 - B r^-6 = r^-6 (A r^-6 - B), F r = r^-6 (12 A r^-6 - 6 B)
    const float r6 = r2inv * r2inv * r2inv;
    float E = r6 * fmaf(A, r6, -B);
    float F = r6 * fmaf(A12, r6, -B6) * r2inv;
    apply_switch_t<SWA, !SAME>(P.sw_a, r2, E, F);
    fs = inA ? F : 0.f;
    ea = inA ? E : 0.f;
  } else if (TA == PM_TERM_COULOMB) {
    float E = qq * rinv;
    float F = E * r2inv;
    apply_switch_t<SWA, !SAME>(P.sw_a, r2, E, F);
    fs = inA ? F : 0.f;
    ea = inA ? E : 0.f;
  }
  if (TB == PM_TERM_COULOMB) {
    // CoulombForce::operator() (CoulombForce.h:29-40)
    float E = qq * rinv;
    float F = E * r2inv;
    apply_switch_t<SWB, !SAME>(P.sw_b, r2, E, F);
    fs += inB ? F : 0.f;
    eb = inB ? E : 0.f;
  } else if (TB == PM_TERM_EWALD_REAL) {
    // NonbondedFullEwaldSystemForce::realTerm (NonbondedFullEwaldSystemForce.h:818-834):
    // erfc(ar) by Abramowitz & Stegun 7.1.26, the approximation the reference uses,
    // with t = 1 / (1 + p a r) and exp(-(a r)^2) = 2^(r^2 * (-a^2 log2 e)),
    // the constants p a and -a^2 log2 e taken from the host
    const float r = r2 * rinv;
    const float t = __builtin_amdgcn_rcpf(fmaf(P.alpha_p, r, 1.0f));
    const float poly =
        t * (0.254829592f + t * (-0.284496736f + t * (1.421413741f + t * (-1.453152027f + t * 1.061405429f))));
    // E = q^2 erfc(a r) / r, F r = (E + 2 a / sqrt(pi) q^2 exp(-(a r)^2)) / r^2
    const float ee = qq * __builtin_amdgcn_exp2f(r2 * P.alpha2_log2e);
    const float F = ee * fmaf(poly, rinv, P.two_alpha_pi) * r2inv;
    fs += inB ? F : 0.f;
    eb = inB ? poly * ee * rinv : 0.f;
  }
  return fs;
}

// P.pair_split waves share one super-cluster's list, each taking an equal
// share of its entries.  An entry is one j cluster against the super-cluster's i
// clusters set in its 8-bit mask; lane = (jj, ii) = (lane >> 3, lane & 7)
// holds i atom ii of the current i cluster against j atom jj, so one 8x8
// cluster pair per step.  The i atoms (positions relative to the
// super-cluster origin) sit in LDS, the j cluster in registers across its i
// clusters; FP32 arithmetic; pairs within `guard` of a cutoff are decided on
// the reference's FP64 minimal image instead, so the within-cutoff pair set is
// exactly the FP64 one.  i forces accumulate in registers per i cluster;
// j forces are summed over the entry's i clusters and the 8 i lanes (DPP);
// both land in 64-bit fixed point, so the total does not depend on the order
// in which waves finish.
// SAME: both terms share one inclusion cutoff (one test per pair, and no
// "outer" branch in the switch)
// waves per SIMD the register allocation must allow: 6 for the forces-only
// kernel (measured best on gfx950: a little spilling, more waves in flight);
// the kinds that also carry FP64 energy and virial sums (or the census) get
// the registers of 4 waves, so those sums never spill inside the entry loop
#ifndef PM_PAIR_MIN_WAVES
#define PM_PAIR_MIN_WAVES 6
#endif
#ifndef PM_PAIR_MIN_WAVES_EN
#define PM_PAIR_MIN_WAVES_EN 4
#endif
template <int TA, int TB, int SWA, int SWB, int KIND, bool SAME>
__global__ __launch_bounds__(512, (KIND == kPairForces ? PM_PAIR_MIN_WAVES : PM_PAIR_MIN_WAVES_EN)) void k_pair(const SysDev *__restrict__ sys, const int *__restrict__ scp, int nsys,
                                              ForceDev P) {
  constexpr bool MV = KIND == kPairForcesEnergiesMV || KIND == kPairEnergiesMV;
  constexpr bool FORCES = KIND != kPairEnergies && KIND != kPairEnergiesMV;
  constexpr bool EN = KIND == kPairForcesEnergies || KIND == kPairEnergies || MV;
  constexpr bool CEN = KIND == kPairCensus;
  extern __shared__ float4 pair_lds[];
  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;
  char *wb = reinterpret_cast<char *>(pair_lds) + (size_t)wv * kPairWaveLds;
  float4 *ljs = reinterpret_cast<float4 *>(wb);
  float4 *lx = reinterpret_cast<float4 *>(wb + kLJLds * 16);
  int *linfo = reinterpret_cast<int *>(wb + kLJLds * 16 + kWave * 16);  // byte offset of row ti of the LJ table
  int *lorig = linfo + kWave;
  float2 *lpp = reinterpret_cast<float2 *>(lorig + kWave);
  const int w = __builtin_amdgcn_readfirstlane((int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6));
  const int split = P.pair_split;
  const int part = w % split, scg = w / split;
  if (scg >= scp[nsys]) return;
  const int s = __builtin_amdgcn_readfirstlane(find_sys(scp, nsys, scg));
  const SysDev *__restrict__ S = sys + s;
  const int sc = scg - scp[s];
  if (sc >= *cp(S->n_sc)) return;
  const int col = *cp(S->sc_col + sc);
  const int kz = sc - *cp(S->col_sc_start + col);
  const int cl0 = *cp(S->col_cl_start + col) + kz * kScCl;
  const int nic = min(kScCl, *cp(S->col_ncl + col) - kz * kScCl);
  const PM_GLOBAL float4 *xq = gp(S->xq);
  const PM_GLOBAL int *ljt = gp(S->ljt);
  const PM_GLOBAL float2 *ljp = gp(S->ljp);
  const PM_GLOBAL int *orig = gp(S->orig);
  PM_GLOBAL long long *gfx = gp(S->fx);
  PM_GLOBAL long long *gfy = gp(S->fy);
  PM_GLOBAL long long *gfz = gp(S->fz);
  const int ii = lane & 7, jj = lane >> 3;
  // j forces: lanes ii = 0, 1, 2 of each j group hold the x, z, y sums
  PM_GLOBAL long long *gfc = gfx + (size_t)(ii == 1 ? 2 : (ii == 2 ? 1 : 0)) * S->f_stride;
  const bool peratom = S->lj_mode == PM_LJ_PERATOM;
  const int T = S->n_types;
  // more types than the LDS copy holds: rows from the global (A, B, 12 A, 6 B)
  // table through the caches (wave-uniform)
  const bool big_table = T * T > kLJLds;
  const char *gtab12 = reinterpret_cast<const char *>(S->lj_table12);
  // stage the super-cluster's i atoms: lane = c * 8 + atom
  unsigned long long valid_ca;  // bit c * 8 + atom: a real atom
  {
    const int c = lane >> 3, a = lane & 7;
    float4 q = make_float4(0.f, 0.f, 0.f, 0.f);
    int info = 0, og = -1;
    bool valid = false;
    float2 pp = make_float2(0.f, 0.f);
    if (c < nic) {
      const int cl = cl0 + c, si = cl * kCluster + a;
      const float4 x = xq[si];
      const double *O = S->cl_origin;
      const int cc = S->cl_cap;
      q = make_float4(x.x + (float)(O[cl] - O[cl0]), x.y + (float)(O[cc + cl] - O[cc + cl0]),
                      x.z + (float)(O[2 * cc + cl] - O[2 * cc + cl0]), x.w);
      valid = a < S->cl_nvalid[cl];
      if (TA == PM_TERM_LJ && !peratom) info = ljt[si] * T * 16;
      if (TA == PM_TERM_LJ && peratom) pp = ljp[si];
      og = orig[si];
    }
    lx[lane] = q;
    linfo[lane] = info;
    lorig[lane] = og;
    lpp[lane] = pp;
    valid_ca = lane_mask(valid);
  }
  // LJ table (T <= 16; larger ones stay in global memory) as (A, B, 12 A, 6 B); the
  // modified-pair coefficients stay in the global table
  const PM_GLOBAL float4 *gtab = gp(S->lj_table);
  if (TA == PM_TERM_LJ && !peratom && !big_table) {
    for (int k = lane; k < T * T; k += kWave) {
      const float4 g = gtab[k];
      ljs[k] = make_float4(g.x, g.y, 12.0f * g.x, 6.0f * g.y);
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  // wave-uniform lane masks (lane = jj * 8 + ii): real i atoms per i cluster,
  // and the i < j triangle of a cluster's pairs with itself
  unsigned long long vim[kScCl];
#pragma unroll
  for (int c = 0; c < kScCl; ++c) vim[c] = ((valid_ca >> (8 * c)) & 0xFFull) * 0x0101010101010101ull;
  const unsigned long long trim = wave_ballot(ii < jj);
  const float csf = S->csf, ljsc = S->lj_scaling;
  float fi[kScCl][3];
#pragma unroll
  for (int c = 0; c < kScCl; ++c) fi[c][0] = fi[c][1] = fi[c][2] = 0.f;
  double eA = 0.0, eB = 0.0;
  double vxx = 0.0, vxy = 0.0, vxz = 0.0, vyy = 0.0, vyz = 0.0, vzz = 0.0;
  double mxx = 0.0, mxy = 0.0, mxz = 0.0, myy = 0.0, myz = 0.0, mzz = 0.0;
  unsigned long long ccount = 0, chash = 0;
  const int len = *cp(S->list_len + sc);
  const int e_begin = part * len / split, e_end = (part + 1) * len / split;
  const PM_CONST uint4 *Lst = cp(S->list + 2 * (size_t)sc * S->list_cap);
  const PM_CONST unsigned long long *Mb = cp(S->list_mask + (size_t)sc * S->mask_cap * (2 * kScCl));
  for (int e = e_begin; e < e_end; ++e) {
    const uint4 ent = Lst[2 * e];
    const unsigned e1 = Lst[2 * e + 1].x;
    const unsigned w0 = ent.x;
    const int cj = (int)(w0 & 0x7FFFFFu);
    const int sh = (int)((w0 >> 23) & 31u);
    const int nvj = (int)((w0 >> 28) & 7u) + 1;
    const unsigned imask = e1 & 0xFFu;
    const int sj = cj * kCluster + jj;
    const float4 xj = xq[sj];
    const float xjx = xj.x + __uint_as_float(ent.y), xjy = xj.y + __uint_as_float(ent.z),
                xjz = xj.z + __uint_as_float(ent.w);
    int tj = 0;
    float2 pj = make_float2(0.f, 0.f);
    if (TA == PM_TERM_LJ) {
      if (peratom) pj = ljp[sj];
      else tj = ljt[sj];
    }
    const char *ljrow = reinterpret_cast<const char *>(ljs) + tj * 16;  // + row offset = (ti, tj) entry
    const unsigned long long jm = nvj == 8 ? ~0ull : ((1ull << (8 * nvj)) - 1ull);
    const int selfc = sh == kShiftCenter ? cj - cl0 : -1;  // i cluster that is cj itself
    const bool flag = (w0 & 0x80000000u) != 0;
    const PM_CONST unsigned long long *M = Mb + (size_t)(e1 >> 8) * (2 * kScCl);
    float fjx = 0.f, fjy = 0.f, fjz = 0.f;
    bool any = false;  // wave-uniform: some i cluster of the entry had a pair inside the cutoff
#pragma unroll
    for (int c = 0; c < kScCl; ++c) {
      if (!((imask >> c) & 1u)) continue;
      const float4 xi = lx[c * kCluster + ii];
      const int info = linfo[c * kCluster + ii];
      const float dx = xjx - xi.x, dy = xjy - xi.y, dz = xjz - xi.z;
      const float r2 = dx * dx + dy * dy + dz * dz;
      // the pair tests as wave lane masks (scalar ANDs, no per-lane shifts)
      unsigned long long am = vim[c] & jm;
      if (c == selfc) am &= trim;
      unsigned long long mm = 0;  // modified (1-4) pairs
      if (flag) {
        am &= ~M[2 * c];
        mm = M[2 * c + 1];
      }
      unsigned long long mA, mB;
      if (SAME) {
        mA = lane_mask(r2 <= P.cutA2) & am;
        const unsigned long long near = lane_mask(fabsf(r2 - P.cutA2) <= P.guard) & am;
        if (__builtin_expect(near != 0, 0)) {
          bool in = lane_bit(mA);
          if (lane_bit(near)) in = min_image_r2(*S, lorig[c * kCluster + ii], orig[sj]) <= P.cutA2d;
          mA = lane_mask(in);
        }
        mB = mA;
      } else {
        mA = TA != PM_TERM_NONE ? lane_mask(r2 <= P.cutA2) & am : 0ull;
        mB = TB != PM_TERM_NONE ? lane_mask(r2 <= P.cutB2) & am : 0ull;
        const bool nearA = (TA != PM_TERM_NONE) && fabsf(r2 - P.cutA2) <= P.guard;
        const bool nearB = (TB != PM_TERM_NONE) && fabsf(r2 - P.cutB2) <= P.guard;
        const unsigned long long near = lane_mask(nearA || nearB) & am;
        if (__builtin_expect(near != 0, 0)) {
          bool a = lane_bit(mA), b = lane_bit(mB);
          if (lane_bit(near)) {
            const double r2d = min_image_r2(*S, lorig[c * kCluster + ii], orig[sj]);
            a = (TA != PM_TERM_NONE) && r2d <= P.cutA2d;
            b = (TB != PM_TERM_NONE) && r2d <= P.cutB2d;
          }
          mA = lane_mask(a);
          mB = lane_mask(b);
        }
      }
      if ((mA | mB) == 0) continue;
      any = true;
      const bool inA = lane_bit(mA), inB = lane_bit(mB);
      float qq = xi.w * xj.w;
      float A = 0.f, B = 0.f, A12 = 0.f, B6 = 0.f;
      if (TA == PM_TERM_LJ) {
        if (peratom) {
          // GROMACS mixing (BuildTopology.cpp:650-659): sigma = (si+sj)/2, eps = sqrt(ei ej)
          const float2 pi = lpp[c * kCluster + ii];
          const float sg = pi.x + pj.x, ep = pi.y * pj.y;
          const float sg2 = sg * sg, sg6 = sg2 * sg2 * sg2;
          B = 4.0f * ep * sg6;
          A = B * sg6;
          A12 = 12.0f * A;
          B6 = 6.0f * B;
        } else {
          float4 p;
          if (__builtin_expect(big_table, 0))
            p = *gp(reinterpret_cast<const float4 *>(gtab12 + info + tj * 16));
          else
            p = *reinterpret_cast<const float4 *>(ljrow + info);
          A = p.x;
          B = p.y;
          A12 = p.z;
          B6 = p.w;
        }
      }
      // modified pairs (rare: a wave-uniform test first): scaled charges
      // and the 1-4 LJ coefficients
      if (__builtin_expect(mm != 0, 0)) {
        if (lane_bit(mm)) {
          qq *= csf;
          if (TA == PM_TERM_LJ) {
            if (peratom) {
              A *= ljsc;
              B *= ljsc;
            } else {
              const float4 g = gtab[(info >> 4) + tj];
              A = g.z;
              B = g.w;
            }
            A12 = 12.0f * A;
            B6 = 6.0f * B;
          }
        }
      }
      float ea, eb;
      const float fs = pair_terms<TA, TB, SWA, SWB, SAME>(P, r2, qq, A, B, A12, B6, inA, inB, ea, eb);
      const float fx = fs * dx, fy = fs * dy, fz = fs * dz;
      fi[c][0] -= fx;
      fi[c][1] -= fy;
      fi[c][2] -= fz;
      fjx += fx;
      fjy += fy;
      fjz += fz;
      if (EN) {
        eA += (double)ea;
        eB += (double)eb;
        bool atomic = true;
        if (MV) {
          // OneAtomPair.h:95-103: a pair of two molecules adds f (x) (R_J - R_I)
          // to the molecular virial and f (x) d to the atomic one; a pair
          // inside a molecule adds to the atomic virial when that is on
          const bool live = inA || inB;
          const int mi = live ? S->smol[(cl0 + c) * kCluster + ii] : 0;
          const int mj = live ? S->smol[sj] : 0;
          const bool other = live && mi != mj;
          atomic = other || P.virial;
          if (other) {
            const int nm = S->n_mol;
            double cd[3] = {S->com[mj] - S->com[mi], S->com[nm + mj] - S->com[nm + mi],
                            S->com[2 * nm + mj] - S->com[2 * nm + mi]};
            min_image_vec(*S, cd);
            mxx += (double)fx * cd[0];
            mxy += (double)fx * cd[1];
            mxz += (double)fx * cd[2];
            myy += (double)fy * cd[1];
            myz += (double)fy * cd[2];
            mzz += (double)fz * cd[2];
          }
        }
        if (atomic) {
          vxx += (double)(fx * dx);
          vxy += (double)(fx * dy);
          vxz += (double)(fx * dz);
          vyy += (double)(fy * dy);
          vyz += (double)(fy * dz);
          vzz += (double)(fz * dz);
        }
      }
      if (CEN) {
        if (inA || inB) {
          const int ai = lorig[c * kCluster + ii], aj = orig[sj];
          const unsigned lo = (unsigned)min(ai, aj), hi = (unsigned)max(ai, aj);
          ccount += 1;
          chash += pair_hash(lo, hi);
        }
      }
    }
    // Newton's third law: j force = sum over the i clusters and the 8 i lanes
    // (nothing to add when no pair of the entry was inside the cutoff)
    if (FORCES && any) {
      const float g = dpp_sum8_scatter(fjx, fjy, fjz, ii);
      if (ii < 3 && jj < nvj) atomic_add_ll(&gfc[sj], to_fixed((double)g, kFScale));
    }
  }
  // i force = sum over the 8 j lanes that share ii
  if (FORCES) {
#pragma unroll
    for (int c = 0; c < kScCl; ++c) {
      if (c >= nic) break;
      float ax = fi[c][0], ay = fi[c][1], az = fi[c][2];
      ax += __shfl_xor(ax, 8);
      ay += __shfl_xor(ay, 8);
      az += __shfl_xor(az, 8);
      ax += __shfl_xor(ax, 16);
      ay += __shfl_xor(ay, 16);
      az += __shfl_xor(az, 16);
      ax += __shfl_xor(ax, 32);
      ay += __shfl_xor(ay, 32);
      az += __shfl_xor(az, 32);
      if (jj == 0 && lane_bit(vim[c])) {
        const int si = (cl0 + c) * kCluster + ii;
        atomic_add_ll(&gfx[si], to_fixed((double)ax, kFScale));
        atomic_add_ll(&gfy[si], to_fixed((double)ay, kFScale));
        atomic_add_ll(&gfz[si], to_fixed((double)az, kFScale));
      }
    }
  }
  if (EN) {
    eA = wave_sum_d(eA);
    eB = wave_sum_d(eB);
    vxx = wave_sum_d(vxx);
    vxy = wave_sum_d(vxy);
    vxz = wave_sum_d(vxz);
    vyy = wave_sum_d(vyy);
    vyz = wave_sum_d(vyz);
    vzz = wave_sum_d(vzz);
    if (lane == 0) {
      PM_GLOBAL long long *acc = gp(acc_copy(S->acc, w));
      const int slotA = TA == PM_TERM_LJ ? kAccLJ : kAccCoul;
      atomic_add_ll(&acc[slotA], to_fixed(eA, kEScale));
      atomic_add_ll(&acc[kAccCoul], to_fixed(eB, kEScale));
      atomic_add_ll(&acc[kAccVxx], to_fixed(vxx, kEScale));
      atomic_add_ll(&acc[kAccVxy], to_fixed(vxy, kEScale));
      atomic_add_ll(&acc[kAccVxz], to_fixed(vxz, kEScale));
      atomic_add_ll(&acc[kAccVyy], to_fixed(vyy, kEScale));
      atomic_add_ll(&acc[kAccVyz], to_fixed(vyz, kEScale));
      atomic_add_ll(&acc[kAccVzz], to_fixed(vzz, kEScale));
    }
    if (MV) {
      mxx = wave_sum_d(mxx);
      mxy = wave_sum_d(mxy);
      mxz = wave_sum_d(mxz);
      myy = wave_sum_d(myy);
      myz = wave_sum_d(myz);
      mzz = wave_sum_d(mzz);
      if (lane == 0) {
        PM_GLOBAL long long *acc = gp(acc_copy(S->acc, w));
        atomic_add_ll(&acc[kAccMVxx], to_fixed(mxx, kEScale));
        atomic_add_ll(&acc[kAccMVxy], to_fixed(mxy, kEScale));
        atomic_add_ll(&acc[kAccMVxz], to_fixed(mxz, kEScale));
        atomic_add_ll(&acc[kAccMVyy], to_fixed(myy, kEScale));
        atomic_add_ll(&acc[kAccMVyz], to_fixed(myz, kEScale));
        atomic_add_ll(&acc[kAccMVzz], to_fixed(mzz, kEScale));
      }
    }
  }
  if (CEN) {
    ccount = wave_sum_u64(ccount);
    chash = wave_sum_u64(chash);
    if (lane == 0) {
      atomicAdd((unsigned long long *)&S->census[0], ccount);
      atomicAdd((unsigned long long *)&S->census[1], chash);
    }
  }
}

// ------------------------------------------------------- FP64 side terms



// Bonded terms and Ewald excluded-pair corrections, one thread per atom
// ("owner computes"): the thread walks the terms its atom takes part in and
// evaluates each in FP64 from that atom's side.  A term is evaluated once per
// participant with the same expressions, so the forces on its atoms are
// exactly opposite, and each atom's slot is updated by one thread only -- no
// atomics, and the result does not depend on which other systems share the
// launch.  Energies and virial (mode & kLocalEnergies) are counted by the
// term's first atom.
//   BondSystemForce::calcBond (force/bonded/BondSystemForce.h:77-108)
//   AngleSystemForce::calcAngle (force/bonded/AngleSystemForce.h:66-140)
//   NonbondedFullEwaldSystemForce::correctionTerm (NonbondedFullEwaldSystemForce.h:1097-1131)
__global__ __launch_bounds__(256) void k_local(SysDev *sys, const int *ap, int nsys, ForceDev P, int mode) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  const bool ok = g < ap[nsys];
  int s = 0;
  double f[3] = {0.0, 0.0, 0.0};
  double eb = 0.0, ea = 0.0, ec = 0.0, v[6] = {0, 0, 0, 0, 0, 0};
  if (ok) {
    s = find_sys(ap, nsys, g);
    const SysDev &S = sys[s];
    const int a = g - ap[s], n = S.n_atoms;
    const double *X = S.x;
    const int *ls = S.loc_start + 3 * (size_t)a;
    // one kind at a time, so lanes of a wave run the same term code together
    for (int it = ls[kLocBond]; it < ls[kLocBond + 1]; ++it) {
      const int item = S.loc_item[it];
      const int role = item & 3, k = item >> 2;
      const int a1 = S.bonds[2 * k], a2 = S.bonds[2 * k + 1];
      double r12[3] = {X[a1] - X[a2], X[n + a1] - X[n + a2], X[2 * n + a1] - X[2 * n + a2]};
      min_image_vec(S, r12);
      const double r = sqrt(r12[0] * r12[0] + r12[1] * r12[1] + r12[2] * r12[2]);
      const double dp = 2.0 * S.bond_k[k] * (r - S.bond_r0[k]);
      const double c = -dp / r;
      const double sg = role == 0 ? 1.0 : -1.0;
      for (int d = 0; d < 3; ++d) f[d] += sg * (r12[d] * c);
      if (role == 0) {
        eb += S.bond_k[k] * (r - S.bond_r0[k]) * (r - S.bond_r0[k]);
        v[0] += r12[0] * c * r12[0];
        v[1] += r12[0] * c * r12[1];
        v[2] += r12[0] * c * r12[2];
        v[3] += r12[1] * c * r12[1];
        v[4] += r12[1] * c * r12[2];
        v[5] += r12[2] * c * r12[2];
      }
    }
    for (int it = ls[kLocAngle]; it < ls[kLocAngle + 1]; ++it) {
      const int item = S.loc_item[it];
      const int role = item & 3, k = item >> 2;
      const int a1 = S.angles[3 * k], a2 = S.angles[3 * k + 1], a3 = S.angles[3 * k + 2];
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
      const double dp = 2.0 * S.angle_k[k] * (theta - S.angle_t0[k]);
      double f1[3], f3[3];
      for (int d = 0; d < 3; ++d) {
        const double u12 = r12[d] / d12, u32 = r32[d] / d32;
        f1[d] = -((u12 * ct - u32) / (st * d12)) * dp;
        f3[d] = -((u32 * ct - u12) / (st * d32)) * dp;
      }
      for (int d = 0; d < 3; ++d) f[d] += role == 0 ? f1[d] : (role == 2 ? f3[d] : -f1[d] - f3[d]);
      if (role == 0) {
        ea += S.angle_k[k] * (theta - S.angle_t0[k]) * (theta - S.angle_t0[k]);
        v[0] += f1[0] * r12[0] + f3[0] * r32[0];
        v[1] += f1[0] * r12[1] + f3[0] * r32[1];
        v[2] += f1[0] * r12[2] + f3[0] * r32[2];
        v[3] += f1[1] * r12[1] + f3[1] * r32[1];
        v[4] += f1[1] * r12[2] + f3[1] * r32[2];
        v[5] += f1[2] * r12[2] + f3[2] * r32[2];
      }
    }
    for (int it = ls[kLocCorr]; it < ls[kLocCorr + 1]; ++it) {
      const int item = S.loc_item[it];
      const int role = item & 3, k = item >> 2;
      const int a1 = S.corr[3 * k], a2 = S.corr[3 * k + 1], cls = S.corr[3 * k + 2];
      double d[3] = {X[a2] - X[a1], X[n + a2] - X[n + a1], X[2 * n + a2] - X[2 * n + a1]};
      min_image_vec(S, d);
      const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
      double qq = S.charge[a1] * S.charge[a2];
      if (cls == PM_EXCL_MODIFIED) qq *= 1.0 - (double)S.csf;
      const double r = sqrt(r2), rr = 1.0 / r;
      const double e = erf(P.dalpha * r) * rr;  // C-library erf, as the reference on Linux
      const double fs = qq * (P.dtwo_alpha_pi * exp(-(P.dalpha * P.dalpha) * r2) - e) * rr * rr;
      const double sg = role == 0 ? -1.0 : 1.0;
      for (int q = 0; q < 3; ++q) f[q] += sg * (d[q] * fs);
      if (role == 0) {
        ec += -qq * e;
        v[0] += d[0] * fs * d[0];
        v[1] += d[0] * fs * d[1];
        v[2] += d[0] * fs * d[2];
        v[3] += d[1] * fs * d[1];
        v[4] += d[1] * fs * d[2];
        v[5] += d[2] * fs * d[2];
      }
    }
    if ((mode & kLocalForces) && ls[3] > ls[0]) {
      const int sl = S.slot[a];
      S.fx[sl] += to_fixed(f[0], kFScale);
      S.fy[sl] += to_fixed(f[1], kFScale);
      S.fz[sl] += to_fixed(f[2], kFScale);
    }
  }
  if (!(mode & kLocalEnergies)) return;
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
  const int s0 = __shfl(s, 0);
  const bool mixed = wave_ballot(ok && s != s0) != 0;
  const int wid = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  if (!mixed) {
    for (int k = 0; k < 9; ++k) q[k] = wave_sum_ll(q[k]);
    if ((threadIdx.x & 63) == 0)
      for (int k = 0; k < 9; ++k)
        if (q[k]) atomic_add_ll(&acc_copy(sys[s0].acc, wid)[slots[k]], q[k]);
  } else if (ok) {
    for (int k = 0; k < 9; ++k)
      if (q[k]) atomic_add_ll(&acc_copy(sys[s].acc, wid)[slots[k]], q[k]);
  }
}

// ------------------------------------------------------------ integrator

// First step of a run (the system's kCtlFirst word is set):
//   doHalfKickdoDrift (LeapfrogIntegrator.cpp:33-56)  v += F*h*0.5/m; x += v*h
// later steps:
//   doKickdoDrift     (LeapfrogIntegrator.cpp:58-87)  v += F*h/m;     x += v*h
// The force buffer is consumed (zeroed) and the FP32 cluster-local
// coordinates for the next force evaluation are written.
// Philox-4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
// easy as 1, 2, 3", SC 2011): a counter-based generator, so atom a's noise
// at kick k of its system depends on (seed, a, k) only
__device__ __forceinline__ uint4 philox4x32_10(uint4 c, uint2 k) {
  for (int r = 0; r < 10; ++r) {
    const unsigned hi0 = __umulhi(0xD2511F53u, c.x), lo0 = 0xD2511F53u * c.x;
    const unsigned hi1 = __umulhi(0xCD9E8D57u, c.z), lo1 = 0xCD9E8D57u * c.z;
    c = make_uint4(hi1 ^ c.y ^ k.x, lo1, hi0 ^ c.w ^ k.y, lo0);
    k.x += 0x9E3779B9u;
    k.y += 0xBB67AE85u;
  }
  return c;
}

// three N(0, 1) deviates for atom a at noise counter ctr (Box-Muller on the
// four 32-bit outputs, u = (x + 1/2) 2^-32 in (0, 1))
__device__ __forceinline__ void lv_gauss(unsigned seed, int a, unsigned ctr, double g[3]) {
  const uint4 r = philox4x32_10(make_uint4((unsigned)a, ctr, 0x4C414E47u, 0u), make_uint2(seed, 0x50524F54u));
  const double s = 2.3283064365386963e-10;  // 2^-32
  const double u0 = ((double)r.x + 0.5) * s, u1 = ((double)r.y + 0.5) * s;
  const double u2 = ((double)r.z + 0.5) * s, u3 = ((double)r.w + 0.5) * s;
  const double m0 = sqrt(-2.0 * log(u0)), m1 = sqrt(-2.0 * log(u2));
  double s0, c0, s1, c1;
  sincospi(2.0 * u1, &s0, &c0);
  sincospi(2.0 * u3, &s1, &c1);
  g[0] = m0 * c0;
  g[1] = m0 * s0;
  g[2] = m1 * c1;
}

// one Langevin half kick (LangevinLeapfrogIntegrator.cpp:46-70):
// v = v e^{-g dt/2} + F (1 - e^{-g dt/2}) / (g m) + N(0,1) sqrt(2 kB T g / m) ndt
__device__ __forceinline__ void lv_half(double v[3], const double F[3], double m, const LangevinDev &lv,
                                        unsigned seed, int a, unsigned ctr) {
  double g[3];
  lv_gauss(seed, a, ctr, g);
  const double sf = sqrt(lv.fc / m);
  for (int d = 0; d < 3; ++d) v[d] = v[d] * lv.vdt + F[d] * lv.fdt / m + g[d] * sf * lv.ndt;
}

__global__ __launch_bounds__(256) void k_kick_drift(SysDev *sys, const int *ap, int nsys, double h, LangevinDev lv) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  float d2 = 0.f;
  int s = 0;
  bool ok = g < ap[nsys];
  if (ok) {
    s = find_sys(ap, nsys, g);
    SysDev &S = sys[s];
    const int a = g - ap[s], n = S.n_atoms;
    const int sl = S.slot[a];
    const int mode = S.ctl[kCtlFirst] ? 0 : 1;
    const double m = 1.0 / S.inv_mass[a];
    const double F[3] = {(double)S.fx[sl] * kFScaleInv, (double)S.fy[sl] * kFScaleInv,
                         (double)S.fz[sl] * kFScaleInv};
    S.fx[sl] = 0;
    S.fy[sl] = 0;
    S.fz[sl] = 0;
    float l[3];
    double vl[3];
    if (lv.on) {
      // the closing half kick of the previous step (noise 2c) unless this is
      // the first step of a run, then the opening half kick of this one (2c + 1)
      const unsigned c = (unsigned)S.ctl[kCtlNoise];
      for (int d = 0; d < 3; ++d) vl[d] = S.v[d * n + a];
      if (mode == 1) lv_half(vl, F, m, lv, S.lv_seed, a, 2u * c);
      lv_half(vl, F, m, lv, S.lv_seed, a, 2u * c + 1u);
    }
    for (int d = 0; d < 3; ++d) {
      double vv = S.v[d * n + a];
      if (lv.on) vv = vl[d];
      else if (mode == 0) vv += F[d] * h * 0.5 / m;
      else vv += F[d] * h / m;
      S.v[d * n + a] = vv;
      const double xx = S.x[d * n + a] + vv * h;
      S.x[d * n + a] = xx;
      l[d] = (float)(xx - S.xref[d * n + a]);
    }
    const float4 q = S.xq[sl];
    S.xq[sl] = make_float4(l[0], l[1], l[2], q.w);
    const float4 b = S.lbuild[a];
    const float ex = l[0] - b.x, ey = l[1] - b.y, ez = l[2] - b.z;
    d2 = ex * ex + ey * ey + ez * ez;
    if (!isfinite(d2)) d2 = 3.0e38f;
  }
  // max per system: a block whose atoms all belong to one system reduces
  // through LDS to one atomic; a block straddling systems uses one per lane
  __shared__ float red[4];
  __shared__ int bs[2];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int glast = min((int)(blockIdx.x * blockDim.x + blockDim.x), ap[nsys]) - 1;
  if (threadIdx.x == 0) bs[0] = s;
  if (g == glast) bs[1] = s;
  const float mx = wave_max_f(d2);
  if (lane == 0) red[wv] = mx;
  __syncthreads();
  const int nw = blockDim.x >> 6;
  if (bs[0] == bs[1]) {
    if (threadIdx.x == 0) {
      float m = red[0];
      for (int k = 1; k < nw; ++k) m = fmaxf(m, red[k]);
      atomicMax(sys[s].maxdisp2, __float_as_uint(m));
    }
  } else if (ok) {
    atomicMax(sys[s].maxdisp2, __float_as_uint(d2));
  }
}

// StandardIntegrator::doHalfKick (StandardIntegrator.cpp:80-92): h = 0.5*dt*ITF, v += F*h/m
__global__ void k_half_kick(SysDev *sys, const int *ap, int nsys, double hh, LangevinDev lv) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ap[nsys]) return;
  const int s = find_sys(ap, nsys, g);
  SysDev &S = sys[s];
  const int a = g - ap[s], n = S.n_atoms;
  const int sl = S.slot[a];
  const double m = 1.0 / S.inv_mass[a];
  const double F[3] = {(double)S.fx[sl] * kFScaleInv, (double)S.fy[sl] * kFScaleInv,
                       (double)S.fz[sl] * kFScaleInv};
  if (lv.on) {
    // the closing half kick of the last step: noise counter 2c
    double v[3] = {S.v[a], S.v[n + a], S.v[2 * n + a]};
    lv_half(v, F, m, lv, S.lv_seed, a, 2u * (unsigned)S.ctl[kCtlNoise]);
    for (int d = 0; d < 3; ++d) S.v[d * n + a] = v[d];
    return;
  }
  for (int d = 0; d < 3; ++d) S.v[d * n + a] += F[d] * hh / m;
}

// zero the fixed-point forces of every slot (a fill kernel, not a memset)
__global__ void k_zero_forces(SysDev *sys, const int *slp, int nsys) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= slp[nsys]) return;
  const int s = find_sys(slp, nsys, g);
  SysDev &S = sys[s];
  const int k = g - slp[s];
  S.fx[k] = 0;
  S.fy[k] = 0;
  S.fz[k] = 0;
}

// Start of a force evaluation, one block per system: clear the energy/virial
// accumulators and the census, and decide whether the pair list is rebuilt
// (kStepForce: always; kStepAuto: when an atom moved more than lim from where
// the list was built -- the Verlet criterion, with lim = skin/2 -- or kmax
// steps have passed).  The decision is read by every rebuild kernel.
__global__ void k_step_begin(SysDev *sys, int nsys, int mode, float lim2, int kmax) {
  const int s = blockIdx.x;
  if (s >= nsys) return;
  SysDev &S = sys[s];
  for (int k = threadIdx.x; k < kAccN * kAccCopies; k += blockDim.x) S.acc[k] = 0;
  if (threadIdx.x == 0) {
    S.census[0] = 0;
    S.census[1] = 0;
    int *ctl = S.ctl;
    ctl[kCtlFirst] = 0;  // the step's kick has run
    ctl[kCtlNoise] += 1;
    const unsigned d2 = *S.maxdisp2;
    bool rb = (mode & kStepForce) != 0;
    if (!rb && (mode & kStepAuto))
      rb = __uint_as_float(d2) > lim2 || (kmax > 0 && ctl[kCtlSince] + 1 >= kmax);
    *S.maxdisp_run = max(*S.maxdisp_run, d2);
    *S.maxdisp2 = 0u;
    ctl[kCtlRebuild] = rb ? 1 : 0;
    if (rb) {
      ctl[kCtlSince] = 0;
      ctl[kCtlCount] += 1;
      ctl[kCtlEntries] = 0;
    } else {
      ctl[kCtlSince] += 1;
    }
  }
}

// device state -> host staging (AoS FP64, per system); what: kXfer* bits
__global__ void k_export(SysDev *sys, const int *ap, int nsys, int what) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ap[nsys]) return;
  const int s = find_sys(ap, nsys, g);
  const SysDev &S = sys[s];
  const int a = g - ap[s], n = S.n_atoms;
  const int sl = S.slot[a];
  for (int d = 0; d < 3; ++d) {
    if (what & kXferX) S.xaos[3 * (size_t)a + d] = S.x[d * n + a];
    if (what & kXferV) S.vaos[3 * (size_t)a + d] = S.v[d * n + a];
  }
  if (what & kXferF) {
    S.faos[3 * (size_t)a] = (double)S.fx[sl] * kFScaleInv;
    S.faos[3 * (size_t)a + 1] = (double)S.fy[sl] * kFScaleInv;
    S.faos[3 * (size_t)a + 2] = (double)S.fz[sl] * kFScaleInv;
  }
}

// per-system scalars of a launch, packed for one device-to-host copy:
// kScalN int64 per system (see kScal*)
__global__ void k_gather_scalars(const SysDev *sys, int nsys, long long *out) {
  const int s = blockIdx.x;
  if (s >= nsys) return;
  const SysDev &S = sys[s];
  long long *o = out + (size_t)s * kScalN;
  for (int k = threadIdx.x; k < kScalN; k += blockDim.x) {
    long long v = 0;
    if (k < kAccN)
      for (int c = 0; c < kAccCopies; ++c) v += S.acc[(size_t)c * kAccN + k];
    else if (k < kScalCensus + 2) v = (long long)S.census[k - kScalCensus];
    else if (k == kScalErr) v = *S.err;
    else if (k >= kScalCtl && k < kScalCtl + kCtlN) v = S.ctl[k - kScalCtl];
    else if (k == kScalMaxdispRun) v = *S.maxdisp_run;
    o[k] = v;
  }
}

// clear the error bits of the systems of a launch
__global__ void k_clear_err(SysDev *sys, int nsys) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < nsys) *sys[s].err = 0u;
}

// ------------------------------------------------------------- launchers

static inline int nblk(long long n, int b) { return (int)((n + b - 1) / b); }

void launch_import(const Launch &L, int what) {
  if (L.n_atoms_total) hipLaunchKernelGGL(k_import, dim3(nblk(L.n_atoms_total, 256)), dim3(256), 0, L.stream, L.sys, L.atom_prefix, L.nsys, what);
}
void launch_export(const Launch &L, int what) {
  if (L.n_atoms_total) hipLaunchKernelGGL(k_export, dim3(nblk(L.n_atoms_total, 256)), dim3(256), 0, L.stream, L.sys, L.atom_prefix, L.nsys, what);
}
void launch_gather_scalars(const Launch &L, long long *out) {
  hipLaunchKernelGGL(k_gather_scalars, dim3(L.nsys), dim3(64), 0, L.stream, L.sys, L.nsys, out);
}
void launch_clear_err(const Launch &L) {
  hipLaunchKernelGGL(k_clear_err, dim3(nblk(L.nsys, 64)), dim3(64), 0, L.stream, L.sys, L.nsys);
}

void launch_rebuild(const Launch &L, float rlist2) {
  if (!L.n_atoms_total) return;
  hipLaunchKernelGGL(k_bin, dim3(nblk(L.n_atoms_total, 256)), dim3(256), 0, L.stream, L.sys, L.atom_prefix, L.nsys);
  hipLaunchKernelGGL(k_col_scan, dim3(L.nsys), dim3(1024), 0, L.stream, L.sys);
  hipLaunchKernelGGL(k_col_fill, dim3(nblk(L.n_atoms_total, 256)), dim3(256), 0, L.stream, L.sys, L.atom_prefix, L.nsys);
  hipLaunchKernelGGL(k_col_sort, dim3(L.n_cols_total), dim3(1024), 0, L.stream, L.sys, L.col_prefix, L.nsys);
  hipLaunchKernelGGL(k_cluster_geom, dim3(nblk(L.n_cl_total, 128)), dim3(128), 0, L.stream, L.sys, L.cl_prefix, L.nsys);
  hipLaunchKernelGGL(k_pairlist, dim3(L.n_cols_total), dim3(kListThreads), 0, L.stream, L.sys, L.col_prefix, L.nsys, rlist2);
}

template <int TA, int TB, int SWA, int SWB, bool SAME = false>
static void launch_pair_t(const Launch &L, const ForceDev &P, int kind) {
  // P.pair_split waves per super-cluster; P.pair_waves waves per workgroup,
  // each with its own LDS region
  const int waves = P.pair_waves;
  const int threads = kWave * waves;
  dim3 g(nblk((long long)L.n_sc_total * P.pair_split, waves)), b(threads);
  const size_t lds = (size_t)waves * kPairWaveLds;
  switch (kind) {
  case kPairForces:
    hipLaunchKernelGGL((k_pair<TA, TB, SWA, SWB, kPairForces, SAME>), g, b, lds, L.stream, L.sys, L.sc_prefix, L.nsys, P);
    break;
  case kPairForcesEnergies:
    hipLaunchKernelGGL((k_pair<TA, TB, SWA, SWB, kPairForcesEnergies, SAME>), g, b, lds, L.stream, L.sys, L.sc_prefix, L.nsys, P);
    break;
  case kPairEnergies:
    hipLaunchKernelGGL((k_pair<TA, TB, SWA, SWB, kPairEnergies, SAME>), g, b, lds, L.stream, L.sys, L.sc_prefix, L.nsys, P);
    break;
  case kPairForcesEnergiesMV:
    hipLaunchKernelGGL((k_pair<TA, TB, SWA, SWB, kPairForcesEnergiesMV, SAME>), g, b, lds, L.stream, L.sys, L.sc_prefix,
                       L.nsys, P);
    break;
  case kPairEnergiesMV:
    hipLaunchKernelGGL((k_pair<TA, TB, SWA, SWB, kPairEnergiesMV, SAME>), g, b, lds, L.stream, L.sys, L.sc_prefix,
                       L.nsys, P);
    break;
  default:
    hipLaunchKernelGGL((k_pair<TA, TB, SWA, SWB, kPairCensus, SAME>), g, b, lds, L.stream, L.sys, L.sc_prefix, L.nsys, P);
    break;
  }
}

// C2 switches get a compile-time fast path; every other kind runs the generic one
template <int TA, int TB>
static void launch_pair_sw(const Launch &L, const ForceDev &P, int kind) {
  const bool a2 = TA == PM_TERM_NONE || P.sw_a.kind == PM_SWITCH_C2;
  const bool b2 = TB != PM_TERM_COULOMB || P.sw_b.kind == PM_SWITCH_C2;
  constexpr int SB = TB == PM_TERM_COULOMB ? PM_SWITCH_C2 : PM_SWITCH_CUTOFF;
  constexpr int SBG = TB == PM_TERM_COULOMB ? kSwGeneric : PM_SWITCH_CUTOFF;
  // one inclusion cutoff for both terms, and no switch that ends before it
  // (OneAtomPairTwo includes a pair within the larger switch cutoff; the
  // shorter switch must then be zeroed beyond its own cutoff, the OUTER form)
  const bool same = TA != PM_TERM_NONE && TB != PM_TERM_NONE && P.cutA2 == P.cutB2 && P.cutA2d == P.cutB2d &&
                    P.sw_a.cut2 >= P.cutA2 && (TB != PM_TERM_COULOMB || P.sw_b.cut2 >= P.cutB2);
  if (a2 && b2 && same) launch_pair_t<TA, TB, PM_SWITCH_C2, SB, true>(L, P, kind);
  else if (a2 && b2) launch_pair_t<TA, TB, PM_SWITCH_C2, SB>(L, P, kind);
  else launch_pair_t<TA, TB, kSwGeneric, SBG>(L, P, kind);
}

bool launch_pair(const Launch &L, const ForceDev &P, int kind) {
  if (!L.n_sc_total) return true;
  const int a = P.term_a, b = P.term_b;
  if (a == PM_TERM_LJ && b == PM_TERM_NONE) launch_pair_sw<PM_TERM_LJ, PM_TERM_NONE>(L, P, kind);
  else if (a == PM_TERM_COULOMB && b == PM_TERM_NONE) launch_pair_sw<PM_TERM_COULOMB, PM_TERM_NONE>(L, P, kind);
  else if (a == PM_TERM_LJ && b == PM_TERM_COULOMB) launch_pair_sw<PM_TERM_LJ, PM_TERM_COULOMB>(L, P, kind);
  else if (a == PM_TERM_LJ && b == PM_TERM_EWALD_REAL) launch_pair_sw<PM_TERM_LJ, PM_TERM_EWALD_REAL>(L, P, kind);
  else if (a == PM_TERM_NONE && b == PM_TERM_EWALD_REAL) launch_pair_t<PM_TERM_NONE, PM_TERM_EWALD_REAL, PM_SWITCH_NONE, PM_SWITCH_CUTOFF>(L, P, kind);
  else return false;
  return true;
}

void launch_local(const Launch &L, const ForceDev &P, int mode) {
  if (L.n_atoms_total && L.n_local_total)
    hipLaunchKernelGGL(k_local, dim3(nblk(L.n_atoms_total, 256)), dim3(256), 0, L.stream, L.sys, L.atom_prefix, L.nsys, P, mode);
}
void launch_kick_drift(const Launch &L, double h, const LangevinDev &lv) {
  if (L.n_atoms_total) hipLaunchKernelGGL(k_kick_drift, dim3(nblk(L.n_atoms_total, 256)), dim3(256), 0, L.stream, L.sys, L.atom_prefix, L.nsys, h, lv);
}
void launch_half_kick(const Launch &L, double hh, const LangevinDev &lv) {
  if (L.n_atoms_total) hipLaunchKernelGGL(k_half_kick, dim3(nblk(L.n_atoms_total, 256)), dim3(256), 0, L.stream, L.sys, L.atom_prefix, L.nsys, hh, lv);
}
// buildMolecularCenterOfMass (TopologyUtilities.cpp:677-684) with
// molecularCenterOfMass (:645-674): one thread per molecule, the atoms in
// index order, Kahan-compensated mass-weighted sums, FP64
__global__ __launch_bounds__(256) void k_mol_com(SysDev *sys) {
#pragma clang fp contract(off)
  const SysDev &S = sys[blockIdx.y];
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= S.n_mol) return;
  const int n = S.n_atoms, a0 = S.mol_start[m], a1 = S.mol_start[m + 1];
  double c[3] = {0.0, 0.0, 0.0}, t[3] = {0.0, 0.0, 0.0}, sm = 0.0;
  for (int k = a0; k < a1; ++k) {
    const int a = S.mol_atoms[k];
    const double mass = S.mass[a];
    if (k == a0) {
      sm = mass;
      for (int q = 0; q < 3; ++q) c[q] = S.x[q * n + a] * mass;
      continue;
    }
    sm += mass;
    for (int q = 0; q < 3; ++q) {
      const double tx = S.x[q * n + a] * mass;
      const double ty = tx - t[q];
      const double tt = c[q] + ty;
      t[q] = (tt - c[q]) - ty;
      c[q] = tt;
    }
  }
  for (int q = 0; q < 3; ++q) S.com[q * S.n_mol + m] = a1 > a0 ? c[q] / sm : 0.0;
}

void launch_mol_com(const Launch &L) {
  if (L.max_mol > 0)
    hipLaunchKernelGGL(k_mol_com, dim3(nblk(L.max_mol, 256), L.nsys), dim3(256), 0, L.stream, L.sys);
}

// zero `bytes` (a multiple of 8) at p: a fill kernel, not a memset
__global__ void k_fill_zero(unsigned long long *p, size_t n) {
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x) p[k] = 0ull;
}

void launch_fill_zero(void *p, size_t bytes, hipStream_t stream) {
  const size_t n = bytes / 8;
  if (!n) return;
  const int blocks = (int)std::min<size_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_fill_zero, dim3(blocks), dim3(256), 0, stream, static_cast<unsigned long long *>(p), n);
}

void launch_zero_forces(const Launch &L) {
  if (L.n_slots_total) hipLaunchKernelGGL(k_zero_forces, dim3(nblk(L.n_slots_total, 256)), dim3(256), 0, L.stream, L.sys, L.slot_prefix, L.nsys);
}
// start of an MD run for the systems of the launch: the first kick is a half
// kick, and the run's rebuild count and displacement maximum restart
__global__ void k_md_arm(SysDev *sys, int nsys) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nsys) return;
  SysDev &S = sys[s];
  S.ctl[kCtlFirst] = 1;
  S.ctl[kCtlCount] = 0;
  *S.maxdisp_run = 0u;
}

void launch_md_arm(const Launch &L) {
  hipLaunchKernelGGL(k_md_arm, dim3(nblk(L.nsys, 64)), dim3(64), 0, L.stream, L.sys, L.nsys);
}

void launch_step_begin(const Launch &L, int mode, float lim2, int kmax) {
  hipLaunchKernelGGL(k_step_begin, dim3(L.nsys), dim3(64), 0, L.stream, L.sys, L.nsys, mode, lim2, kmax);
}
}  // namespace pm
