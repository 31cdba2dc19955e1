// pm_ewald.hip -- Ewald reciprocal-space sum on gfx950.
//
// NonbondedFullEwaldSystemForce::reciprocalTerm
// (force/nonbonded/NonbondedFullEwaldSystemForce.h:873-1083) sums, over the
// k-vectors k = 2 pi (h/Lx, k/Ly, l/Lz) of the h >= 0 half space with
// |k| < kc,
//   S(k) = sum_j q_j e^{i k.r_j},   E = 4 pi / V sum_k b_k |S(k)|^2,
//   F_j  = sum_k k (8 pi / V) b_k (sin_j(k) Re S - cos_j(k) Im S),
// with b_k = exp(-k^2 / (4 alpha^2)) / k^2.  e^{i k.r_j} factors into
// per-dimension phases X_j(h) Y_j(k) Z_j(l), which the reference tabulates
// per atom by the addition theorem (:900-945); so does k_ew_tables here,
// into one record per atom: X(0..xrows) | Y(0..yrows) | Z(0..zrows).
//
// All in FP64: on this part FP64 vector FMA issues at the rate of FP32, and
// without transcendentals inside the sums nothing is gained in FP32.
//   k_ew_tables  one thread per atom: minimalPosition, sin/cos, recurrences
//   k_ew_sf      S(k) partial sums as a complex matrix product on the FP64
//                matrix cores: S(h, p) = sum_j X_j(h) W_j(p) over atoms j,
//                with p = (k, l) points and W_j(p) = Y_j(k) Z_j(l) formed in
//                registers; rows = (h, re/im) of an h block of 8, columns =
//                16 points, 4 atoms per step, operands read straight from
//                the atoms' records.  One partial per atom chunk, summed in
//                chunk order by k_ew_ssum (deterministic)
//   k_ew_ssum    S(k), the per-k force weights and b |S|^2
//   k_ew_force   one thread per atom over a block of k rows: the phases of a
//                row follow from the previous row's by recurrence (no table
//                reads inside a sweep) and run along the row by the
//                three-term cosine recurrence, the per-k weights are
//                wave-uniform (scalar loads); partial forces per row block
//   k_ew_fsum    forces summed over row blocks in order, into the 64-bit
//                fixed-point force slots
//   k_ew_energy  energy and virial, one workgroup, fixed summation order
#include <hip/hip_runtime.h>
#include <math.h>

#include "../../include/protomol_amd.h"
#include "pm_kernels.h"

namespace pm {

#if defined(__HIP_DEVICE_COMPILE__)
#define EW_CONST __attribute__((address_space(4)))
#else
#define EW_CONST
#endif

template <class T>
__device__ __forceinline__ const EW_CONST T *ew_cp(const T *p) {
  return (const EW_CONST T *)p;
}

constexpr int kEwHB = 8;        // h values per structure-factor block
constexpr int kEwFThreads = 256;

typedef double ew_d4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cmulc(double2 a, double2 b) {  // a * conj(b)
  return make_double2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
// record entry for a signed index: e^{-i n t} = conj(e^{i n t})
__device__ __forceinline__ double2 ew_phase(const double2 *rec, int idx) {
  const double2 v = rec[idx < 0 ? -idx : idx];
  return idx < 0 ? make_double2(v.x, -v.y) : v;
}

// Per-atom phase tables, as NonbondedFullEwaldSystemForce.h:900-945:
// r = minimalPosition(x) (PeriodicBoundaryConditions.h:113-147), t = r 2 pi / L,
// (cos, sin)(n t) by the addition theorem, the x table scaled by the charge.
// PeriodicBoundaryConditions: the wrap shared by minimalPosition (:113-147,
// after the origin is taken off) and minimalDifference (:72-110)
__device__ __forceinline__ void ew_wrap(const SysDev &S, double *d) {
#pragma clang fp contract(off)
  if (d[0] * d[0] + d[1] * d[1] + d[2] * d[2] > S.D) {
    for (int k = 0; k < 3; ++k) {
      if (d[k] < -S.H2[k]) d[k] += S.H[k];
      else if (d[k] > S.H2[k]) d[k] -= S.H[k];
    }
    if (d[0] * d[0] + d[1] * d[1] + d[2] * d[2] > S.D)
      for (int k = 0; k < 3; ++k) d[k] = d[k] - S.L[k] * rint(S.Lr[k] * d[k]);
  }
}

__device__ __forceinline__ void ew_min_position(const SysDev &S, int j, double *d) {
  const int n = S.n_atoms;
  d[0] = S.x[j] - S.ew_org[0];
  d[1] = S.x[n + j] - S.ew_org[1];
  d[2] = S.x[2 * n + j] - S.ew_org[2];
  ew_wrap(S, d);
}

__global__ __launch_bounds__(256) void k_ew_tables(const SysDev *__restrict__ sys) {
#pragma clang fp contract(off)
  const SysDev &S = *sys;
  const int n = S.n_atoms;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  double d[3];
  ew_min_position(S, j, d);
  const double q = S.charge[j];
  double2 *rec = S.ew_tab + (size_t)j * S.ew_rl;
  double2 *sec[3] = {rec, rec + S.ew_xrows, rec + S.ew_xrows + S.ew_yrows};
  const int rows[3] = {S.ew_xrows, S.ew_yrows, S.ew_zrows};
  for (int k = 0; k < 3; ++k) {
    const double t = d[k] * 2.0 * M_PI * S.Lr[k];
    double s1, c1;
    sincos(t, &s1, &c1);
    const double w = k == 0 ? q : 1.0;
    double2 *T = sec[k];
    T[0] = make_double2(w * 1.0, 0.0);  // n = 0
    if (rows[k] < 2) continue;
    double cs = w * s1, cc = w * c1;    // n = 1
    T[1] = make_double2(cc, cs);
    for (int m = 2; m < rows[k]; ++m) {
      const double ns = s1 * cc + c1 * cs;
      const double nc = c1 * cc - s1 * cs;
      cs = ns;
      cc = nc;
      T[m] = make_double2(cc, cs);
    }
  }
}

// Structure-factor partial sums.  Block = (tile, chunk): the tile's
// kEwSFPoints (k, l) points of one h block, the chunk a contiguous range of
// atoms.
// v_mfma_f64_16x16x4f64 computes D += A B with A 16 x 4 (lane l: row
// l % 16, column l / 16), B 4 x 16 (lane l: row l / 16, column l % 16) and
// D 16 x 16 (lane l, register r: row 4 r + l / 16, column l % 16).  Here
// A = x phases of 4 atoms, B = Re W or Im W of those atoms at 16 points, and
// A row i carries (h, part) = ((i & 3) + 4 (i >> 3), (i >> 2) & 1), so lane
// l (q = l / 16) ends with Dr, Di at (q, re), (q, im), (q + 4, re),
// (q + 4, im) -- both parts of h = q and q + 4 -- and
//   Re S(h, p) = Dr(h, re) - Di(h, im),  Im S(h, p) = Di(h, re) + Dr(h, im).
// Atoms past the chunk's end enter with zero weight.
// Workgroups are dealt to the 8 XCDs round robin; every tile of one atom
// chunk is placed on the same XCD (block = ((chunk / 8) tiles + tile) 8 +
// chunk % 8), so that chunk's phase records are fetched into one L2.
__global__ __launch_bounds__(kEwSFWaves * 64) void k_ew_sf(const SysDev *__restrict__ sys) {
  const SysDev &S = *sys;
  const int n = S.n_atoms;
  const int nt = S.ew_ntiles;
  const int b = blockIdx.x;
  const int chunk = (b / (8 * nt)) * 8 + (b & 7), tile = (b >> 3) % nt;
  if (chunk >= S.ew_nchunks) return;
  const int4 T = ew_cp(S.ew_tiles)[tile];   // scalar loads: the tile is the same for the workgroup
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int col = lane & 15, quad = lane >> 4;
  const int RL = S.ew_rl, YO = S.ew_xrows, ZO = S.ew_xrows + S.ew_yrows;
  int ky[kEwSFTiles], kz[kEwSFTiles], pidx[kEwSFTiles];
  double sy[kEwSFTiles], sz[kEwSFTiles];
#pragma unroll
  for (int t = 0; t < kEwSFTiles; ++t) {
    const int p = T.y + (wv * kEwSFTiles + t) * 16 + col;
    pidx[t] = p < T.z ? p : -1;
    const short2 kl = S.ew_pts[p < T.z ? p : T.y];
    ky[t] = YO + (kl.x < 0 ? -kl.x : kl.x);
    kz[t] = ZO + (kl.y < 0 ? -kl.y : kl.y);
    sy[t] = kl.x < 0 ? -1.0 : 1.0;
    sz[t] = kl.y < 0 ? -1.0 : 1.0;
  }
  const int ai = lane & 15;
  const int xrow = T.x + (ai & 3) + 4 * (ai >> 3), xpart = (ai >> 2) & 1;
  ew_d4 dr[kEwSFTiles], di[kEwSFTiles];
#pragma unroll
  for (int t = 0; t < kEwSFTiles; ++t) dr[t] = di[t] = (ew_d4){0.0, 0.0, 0.0, 0.0};
  const int a0 = chunk * S.ew_chunk;
  const int a1 = min(n, a0 + S.ew_chunk);
  // operands of the next atom quad are loaded while the current one's
  // products run (register double buffer)
  double2 cx, cy[kEwSFTiles], cz[kEwSFTiles];
  {
    const double2 *rec = S.ew_tab + (size_t)min(a0 + quad, a1 - 1) * RL;
    cx = rec[xrow];
#pragma unroll
    for (int t = 0; t < kEwSFTiles; ++t) {
      cy[t] = rec[ky[t]];
      cz[t] = rec[kz[t]];
    }
  }
  for (int a = a0; a < a1; a += 4) {
    double2 nx, ny[kEwSFTiles], nz[kEwSFTiles];
    {
      const double2 *rec = S.ew_tab + (size_t)min(a + 4 + quad, a1 - 1) * RL;
      nx = rec[xrow];
#pragma unroll
      for (int t = 0; t < kEwSFTiles; ++t) {
        ny[t] = rec[ky[t]];
        nz[t] = rec[kz[t]];
      }
    }
    const double av = a + quad < a1 ? (xpart ? cx.y : cx.x) : 0.0;
#pragma unroll
    for (int t = 0; t < kEwSFTiles; ++t) {
      const double yi = sy[t] * cy[t].y, zi = sz[t] * cz[t].y;
      const double wr = cy[t].x * cz[t].x - yi * zi, wi = cy[t].x * zi + yi * cz[t].x;
      dr[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, wr, dr[t], 0, 0, 0);
      di[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, wi, di[t], 0, 0, 0);
    }
    cx = nx;
#pragma unroll
    for (int t = 0; t < kEwSFTiles; ++t) {
      cy[t] = ny[t];
      cz[t] = nz[t];
    }
  }
  double2 *out = S.ew_psum + (size_t)chunk * S.ew_nk;
#pragma unroll
  for (int t = 0; t < kEwSFTiles; ++t) {
    if (pidx[t] < 0) continue;
#pragma unroll
    for (int hh = 0; hh < 2; ++hh) {
      const int h = quad + 4 * hh;
      if (h >= T.w) continue;
      const int k = S.ew_kmap[(size_t)pidx[t] * kEwHB + h];
      if (k >= 0) out[k] = make_double2(dr[t][2 * hh] - di[t][2 * hh + 1], di[t][2 * hh] + dr[t][2 * hh + 1]);
    }
  }
}

// S(k) = sum over chunks in chunk order; force weights and b |S|^2
__global__ __launch_bounds__(256) void k_ew_ssum(const SysDev *__restrict__ sys) {
  const SysDev &S = *sys;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= S.ew_nk) return;
  double c = 0.0, s = 0.0;
  for (int ch = 0; ch < S.ew_nchunks; ++ch) {
    const double2 v = S.ew_psum[(size_t)ch * S.ew_nk + k];
    c += v.x;
    s += v.y;
  }
  const double4 kd = S.ew_kd[k];
  const double e = kd.w * (s * s + c * c);
  const double a = S.ew_a8 * kd.w;
  S.ew_sk[k] = make_double4(a * c, a * s, kd.z, e);
}

// Forces of two atoms per thread (j and j + kEwFThreads of the workgroup's
// 2 kEwFThreads) over the k rows of block blockIdx.y: two independent chains
// per wave-uniform weight load.  Within a block, rows of one h run over
// consecutive k and their first l moves by a few steps: Y and Z(first l)
// follow by multiplying with the atom's unit phases; anything else reloads
// from the atom's record.
constexpr int kEwFAtoms = 2;
__global__ __launch_bounds__(kEwFThreads) void k_ew_force(const SysDev *__restrict__ sys) {
  const SysDev &S = *sys;
  const int n = S.n_atoms;
  const int b = blockIdx.y;
  const EW_CONST int *fbr = ew_cp(S.ew_fb_rows);
  const EW_CONST int4 *rows = ew_cp(S.ew_rows);
  const EW_CONST double4 *sk = ew_cp(S.ew_sk);
  const EW_CONST double4 *kd = ew_cp(S.ew_kd);
  const int r0 = fbr[b], r1 = fbr[b + 1];
  int jj[kEwFAtoms];
  const double2 *recx[kEwFAtoms], *recy[kEwFAtoms], *recz[kEwFAtoms];
  double2 ey1[kEwFAtoms], ez1[kEwFAtoms], X[kEwFAtoms], Y[kEwFAtoms], Zl[kEwFAtoms];
  double c2[kEwFAtoms], fx[kEwFAtoms], fy[kEwFAtoms], fz[kEwFAtoms];
#pragma unroll
  for (int u = 0; u < kEwFAtoms; ++u) {
    jj[u] = blockIdx.x * (kEwFAtoms * kEwFThreads) + u * kEwFThreads + (int)threadIdx.x;
    const int j = min(jj[u], n - 1);
    recx[u] = S.ew_tab + (size_t)j * S.ew_rl;
    recy[u] = recx[u] + S.ew_xrows;
    recz[u] = recy[u] + S.ew_yrows;
    ey1[u] = recy[u][1];
    ez1[u] = recz[u][1];
    c2[u] = 2.0 * ez1[u].x;
    fx[u] = fy[u] = fz[u] = 0.0;
    X[u] = Y[u] = Zl[u] = make_double2(0.0, 0.0);
  }
  int ph = -1, pk = 0, plo = 0;
  for (int r = r0; r < r1; ++r) {
    const int4 R = rows[r];
    const int kend = rows[r + 1].w;
    const bool newh = R.x != ph, stepk = R.y == pk + 1;
    const int dl = R.z - plo;
#pragma unroll
    for (int u = 0; u < kEwFAtoms; ++u) {
      if (newh) {
        X[u] = recx[u][R.x];
        Y[u] = ew_phase(recy[u], R.y);
        Zl[u] = ew_phase(recz[u], R.z);
      } else {
        Y[u] = stepk ? cmul(Y[u], ey1[u]) : ew_phase(recy[u], R.y);
        if (dl >= -4 && dl <= 4) {
          for (int m = 0; m < dl; ++m) Zl[u] = cmul(Zl[u], ez1[u]);
          for (int m = 0; m < -dl; ++m) Zl[u] = cmulc(Zl[u], ez1[u]);
        } else {
          Zl[u] = ew_phase(recz[u], R.z);
        }
      }
    }
    ph = R.x;
    pk = R.y;
    plo = R.z;
    // e(l) along the row by the three-term recurrence
    // e(l + 1) = 2 cos(t_z) e(l) - e(l - 1), two FMAs per step
    double2 e[kEwFAtoms], ep[kEwFAtoms];
    double rs[kEwFAtoms];
#pragma unroll
    for (int u = 0; u < kEwFAtoms; ++u) {
      e[u] = cmul(cmul(X[u], Y[u]), Zl[u]);
      ep[u] = cmulc(e[u], ez1[u]);  // e(first l - 1)
      rs[u] = 0.0;
    }
    int q = R.w;
    for (; q + 3 < kend; q += 4) {  // four k per pass: one wait for their weights
      double4 w[4];
#pragma unroll
      for (int v = 0; v < 4; ++v) w[v] = sk[q + v];
#pragma unroll
      for (int v = 0; v < 4; ++v)
#pragma unroll
        for (int u = 0; u < kEwFAtoms; ++u) {
          const double t = e[u].y * w[v].x - e[u].x * w[v].y;
          const double2 en = make_double2(c2[u] * e[u].x - ep[u].x, c2[u] * e[u].y - ep[u].y);
          ep[u] = e[u];
          e[u] = en;
          fz[u] += w[v].z * t;
          rs[u] += t;
        }
    }
    for (; q < kend; ++q) {
      const double4 w = sk[q];
#pragma unroll
      for (int u = 0; u < kEwFAtoms; ++u) {
        const double t = e[u].y * w.x - e[u].x * w.y;
        const double2 en = make_double2(c2[u] * e[u].x - ep[u].x, c2[u] * e[u].y - ep[u].y);
        ep[u] = e[u];
        e[u] = en;
        fz[u] += w.z * t;
        rs[u] += t;
      }
    }
    const double4 k0 = kd[R.w];
#pragma unroll
    for (int u = 0; u < kEwFAtoms; ++u) {
      fx[u] += k0.x * rs[u];
      fy[u] += k0.y * rs[u];
    }
  }
  double *pf = S.ew_pf + (size_t)b * 3 * n;
#pragma unroll
  for (int u = 0; u < kEwFAtoms; ++u) {
    if (jj[u] >= n) continue;
    pf[jj[u]] = fx[u];
    pf[(size_t)n + jj[u]] = fy[u];
    pf[2 * (size_t)n + jj[u]] = fz[u];
  }
}

// sum of the row blocks' partial forces, in block order, into the fixed-point
// force slots (one thread per atom; kernels of a stream run in order)
__global__ __launch_bounds__(256) void k_ew_fsum(const SysDev *__restrict__ sys) {
  const SysDev &S = *sys;
  const int n = S.n_atoms;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  double f[3] = {0.0, 0.0, 0.0};
  for (int b = 0; b < S.ew_nfb; ++b)
    for (int k = 0; k < 3; ++k) f[k] += S.ew_pf[((size_t)b * 3 + k) * n + j];
  const int sl = S.slot[j];
  S.fx[sl] += __double2ll_rn(f[0] * kFScale);
  S.fy[sl] += __double2ll_rn(f[1] * kFScale);
  S.fz[sl] += __double2ll_rn(f[2] * kFScale);
}

// The reciprocal term's share of the molecular virial
// (NonbondedFullEwaldSystemForce.h:1039-1049): each atom's reciprocal force
// (summed over k here, per k there) times minimalDifference(minimalPosition(x),
// R) to its molecule's centre R.  One thread per atom; the per-atom products
// go to fixed point before they are summed, so the total does not depend on
// the grouping.
__global__ __launch_bounds__(256) void k_ew_molvir(const SysDev *__restrict__ sys) {
  const SysDev &S = *sys;
  const int n = S.n_atoms;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  long long q[6] = {0, 0, 0, 0, 0, 0};
  if (j < n) {
    double f[3] = {0.0, 0.0, 0.0};
    for (int b = 0; b < S.ew_nfb; ++b)
      for (int k = 0; k < 3; ++k) f[k] += S.ew_pf[((size_t)b * 3 + k) * n + j];
    double d[3];
    ew_min_position(S, j, d);
    const int m = S.molecule[j], nm = S.n_mol;
    double r[3] = {S.com[m] - d[0], S.com[nm + m] - d[1], S.com[2 * nm + m] - d[2]};
    ew_wrap(S, r);
    const double v[6] = {f[0] * r[0], f[0] * r[1], f[0] * r[2], f[1] * r[1], f[1] * r[2], f[2] * r[2]};
    for (int k = 0; k < 6; ++k) q[k] = __double2ll_rn(v[k] * kEScale);
  }
  for (int k = 0; k < 6; ++k)
    for (int o = 32; o > 0; o >>= 1) q[k] += __shfl_xor(q[k], o);
  if ((threadIdx.x & 63) == 0) {
    long long *acc = acc_copy(S.acc, (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6));
    const int slot[6] = {kAccMVxx, kAccMVxy, kAccMVxz, kAccMVyy, kAccMVyz, kAccMVzz};
    for (int k = 0; k < 6; ++k)
      if (q[k]) atomicAdd(reinterpret_cast<unsigned long long *>(&acc[slot[k]]), static_cast<unsigned long long>(q[k]));
  }
}

// c sum_k b |S|^2 and the virial (NonbondedFullEwaldSystemForce.h:1028-1038,
// 1053-1078): one workgroup, strided partial sums then a fixed tree.  The
// virial goes to the atomic slots with kEwVirial, and also to the molecular
// ones with kEwMolVirial (the reference sums it only with the atomic virial on)
__global__ __launch_bounds__(1024) void k_ew_energy(const SysDev *__restrict__ sys, int mode) {
  __shared__ double red[7][1024];
  const SysDev &S = *sys;
  const int t = threadIdx.x;
  double v[7] = {0, 0, 0, 0, 0, 0, 0};
  for (int k = t; k < S.ew_nk; k += 1024) {
    const double e = S.ew_sk[k].w;
    const double4 kd = S.ew_kd[k];
    const double ksq = kd.x * kd.x + kd.y * kd.y + kd.z * kd.z;
    const double c = 2.0 * (1.0 / ksq + S.ew_fac);
    v[0] += e;
    v[1] += e * (1.0 - c * kd.x * kd.x);
    v[2] -= e * c * kd.x * kd.y;
    v[3] -= e * c * kd.x * kd.z;
    v[4] += e * (1.0 - c * kd.y * kd.y);
    v[5] -= e * c * kd.y * kd.z;
    v[6] += e * (1.0 - c * kd.z * kd.z);
  }
  for (int m = 0; m < 7; ++m) red[m][t] = v[m];
  __syncthreads();
  for (int w = 512; w > 0; w >>= 1) {
    if (t < w)
      for (int m = 0; m < 7; ++m) red[m][t] += red[m][t + w];
    __syncthreads();
  }
  if (t == 0) {
    long long *acc = S.acc;
    const double c = S.ew_c;
    const int slot[7] = {kAccCoul, kAccVxx, kAccVxy, kAccVxz, kAccVyy, kAccVyz, kAccVzz};
    const int mslot[7] = {-1, kAccMVxx, kAccMVxy, kAccMVxz, kAccMVyy, kAccMVyz, kAccMVzz};
    for (int m = 0; m < 7; ++m) {
      const unsigned long long q = static_cast<unsigned long long>(__double2ll_rn(c * red[m][0] * kEScale));
      if (m == 0 || (mode & kEwVirial)) atomicAdd(reinterpret_cast<unsigned long long *>(&acc[slot[m]]), q);
      if (m > 0 && (mode & kEwVirial) && (mode & kEwMolVirial))
        atomicAdd(reinterpret_cast<unsigned long long *>(&acc[mslot[m]]), q);
    }
  }
}

void launch_ewald(const Launch &L, int mode) {
  for (int s = 0; s < L.n_ew; ++s) {
    const EwLaunch &E = L.ew[s];
    if (!E.n || !E.nk) continue;
    if (mode & kEwForces) {
      hipLaunchKernelGGL(k_ew_tables, dim3((E.n + 255) / 256), dim3(256), 0, L.stream, E.sys);
      hipLaunchKernelGGL(k_ew_sf, dim3(E.ntiles * ((E.nchunks + 7) / 8) * 8), dim3(kEwSFWaves * 64), 0, L.stream,
                         E.sys);
      hipLaunchKernelGGL(k_ew_ssum, dim3((E.nk + 255) / 256), dim3(256), 0, L.stream, E.sys);
      hipLaunchKernelGGL(k_ew_force, dim3((E.n + kEwFAtoms * kEwFThreads - 1) / (kEwFAtoms * kEwFThreads), E.nfb),
                         dim3(kEwFThreads), 0, L.stream, E.sys);
      hipLaunchKernelGGL(k_ew_fsum, dim3((E.n + 255) / 256), dim3(256), 0, L.stream, E.sys);
    }
    if (mode & kEwMolVirial)
      hipLaunchKernelGGL(k_ew_molvir, dim3((E.n + 255) / 256), dim3(256), 0, L.stream, E.sys);
    if (mode & kEwEnergies) hipLaunchKernelGGL(k_ew_energy, dim3(1), dim3(1024), 0, L.stream, E.sys, mode);
  }
}

}  // namespace pm
