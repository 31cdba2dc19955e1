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
// per atom by the addition theorem (:900-945); so does k_ew_tables here.
//
// All in FP64: on this part FP64 vector FMA issues at the rate of FP32, and
// without transcendentals inside the sums nothing is gained in FP32.
//   k_ew_tables  one thread per atom: minimalPosition, sin/cos, recurrences
//   k_ew_sf      S(k) partial sums: a lane owns one (k, l) point and 8 h
//                values (one h block), atoms stream through LDS in tiles of
//                16, so S needs no cross-lane reduction; one partial per atom
//                chunk, summed in chunk order by k_ew_ssum (deterministic)
//   k_ew_ssum    S(k), the per-k force weights and b |S|^2
//   k_ew_force   one thread per atom over a block of k rows: the phases of a
//                row follow from the previous row's by recurrence (no table
//                reads inside a sweep), the per-k weights are wave-uniform
//                (scalar loads); partial forces per row block
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
constexpr int kEwTA = 16;       // atoms per LDS tile of the structure-factor pass
constexpr int kEwSFThreads = 256;
constexpr int kEwFThreads = 256;

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cmulc(double2 a, double2 b) {  // a * conj(b)
  return make_double2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
// table entry for a signed index: e^{-i n t} = conj(e^{i n t})
__device__ __forceinline__ double2 ew_phase(const double2 *tab, int n_atoms, int idx, int j) {
  const double2 v = tab[(size_t)(idx < 0 ? -idx : idx) * n_atoms + j];
  return idx < 0 ? make_double2(v.x, -v.y) : v;
}

// Per-atom phase tables, as NonbondedFullEwaldSystemForce.h:900-945:
// r = minimalPosition(x) (PeriodicBoundaryConditions.h:113-147), t = r 2 pi / L,
// (cos, sin)(n t) by the addition theorem, the x table scaled by the charge.
__global__ __launch_bounds__(256) void k_ew_tables(const SysDev *__restrict__ sys) {
#pragma clang fp contract(off)
  const SysDev &S = *sys;
  const int n = S.n_atoms;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  double d[3] = {S.x[j] - S.ew_org[0], S.x[n + j] - S.ew_org[1], S.x[2 * n + j] - S.ew_org[2]};
  if (d[0] * d[0] + d[1] * d[1] + d[2] * d[2] > S.D) {
    for (int k = 0; k < 3; ++k) {
      if (d[k] < -S.H2[k]) d[k] += S.H[k];
      else if (d[k] > S.H2[k]) d[k] -= S.H[k];
    }
    if (d[0] * d[0] + d[1] * d[1] + d[2] * d[2] > S.D)
      for (int k = 0; k < 3; ++k) d[k] = d[k] - S.L[k] * rint(S.Lr[k] * d[k]);
  }
  const double q = S.charge[j];
  double2 *tab[3] = {S.ew_tx, S.ew_ty, S.ew_tz};
  const int rows[3] = {S.ew_xrows, S.ew_yrows, S.ew_zrows};
  for (int k = 0; k < 3; ++k) {
    const double t = d[k] * 2.0 * M_PI * S.Lr[k];
    double s1, c1;
    sincos(t, &s1, &c1);
    const double w = k == 0 ? q : 1.0;
    double2 *T = tab[k] + j;
    double ps = 0.0 * w, pc = w * 1.0;  // n = 0
    T[0] = make_double2(pc, ps);
    if (rows[k] < 2) continue;
    double cs = w * s1, cc = w * c1;    // n = 1
    T[(size_t)n] = make_double2(cc, cs);
    for (int m = 2; m < rows[k]; ++m) {
      const double ns = s1 * cc + c1 * cs;
      const double nc = c1 * cc - s1 * cs;
      cs = ns;
      cc = nc;
      T[(size_t)m * n] = make_double2(cc, cs);
    }
    (void)ps;
    (void)pc;
  }
}

// Structure-factor partial sums.  Block = (tile, chunk): the tile's 256
// lanes own (k, l) points of one h block (up to 8 h values), the chunk is a
// contiguous range of atoms.  Atom tiles of 16 are staged in LDS as
// [atom][row] with rows Y(0..), Z(0..), X(h0..h0+7).
__global__ __launch_bounds__(kEwSFThreads) void k_ew_sf(const SysDev *__restrict__ sys) {
  extern __shared__ double2 ew_lds[];
  const SysDev &S = *sys;
  const int n = S.n_atoms;
  const int nt = S.ew_ntiles;
  const int tile = blockIdx.x % nt, chunk = blockIdx.x / nt;
  const int4 T = S.ew_tiles[tile];
  const int K1 = S.ew_yrows, L1 = S.ew_zrows, RS = S.ew_lds_stride;
  const int pt = T.y + (int)threadIdx.x;
  const bool live = pt < T.z;
  const short2 kl = S.ew_pts[live ? pt : T.y];
  const int ka = kl.x < 0 ? -kl.x : kl.x, la = kl.y < 0 ? -kl.y : kl.y;
  const double sy = kl.x < 0 ? -1.0 : 1.0, sz = kl.y < 0 ? -1.0 : 1.0;
  double ar[kEwHB], ai[kEwHB];
#pragma unroll
  for (int h = 0; h < kEwHB; ++h) ar[h] = ai[h] = 0.0;
  const int a0 = chunk * S.ew_chunk;
  const int a1 = min(n, a0 + S.ew_chunk);
  const int nrow = K1 + L1 + kEwHB;
  for (int t0 = a0; t0 < a1; t0 += kEwTA) {
    const int na = min(kEwTA, a1 - t0);
    __syncthreads();
    for (int idx = threadIdx.x; idx < nrow * kEwTA; idx += kEwSFThreads) {
      const int r = idx / kEwTA, a = idx - r * kEwTA;
      double2 v = make_double2(0.0, 0.0);
      if (a < na) {
        if (r < K1) v = S.ew_ty[(size_t)r * n + t0 + a];
        else if (r < K1 + L1) v = S.ew_tz[(size_t)(r - K1) * n + t0 + a];
        else if (r - K1 - L1 < T.w) v = S.ew_tx[(size_t)(T.x + r - K1 - L1) * n + t0 + a];
      }
      ew_lds[a * RS + r] = v;
    }
    __syncthreads();
    for (int a = 0; a < na; ++a) {
      const double2 *A = ew_lds + a * RS;
      const double2 Y = A[ka], Z = A[K1 + la];
      const double yi = sy * Y.y, zi = sz * Z.y;
      const double wr = Y.x * Z.x - yi * zi, wi = Y.x * zi + yi * Z.x;
#pragma unroll
      for (int h = 0; h < kEwHB; ++h) {
        const double2 X = A[K1 + L1 + h];
        ar[h] += X.x * wr - X.y * wi;
        ai[h] += X.x * wi + X.y * wr;
      }
    }
  }
  if (!live) return;
  double2 *out = S.ew_psum + (size_t)chunk * S.ew_nk;
#pragma unroll
  for (int h = 0; h < kEwHB; ++h) {
    if (h >= T.w) break;
    const int k = S.ew_kmap[(size_t)pt * kEwHB + h];
    if (k >= 0) out[k] = make_double2(ar[h], ai[h]);
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

// Forces of atom j (one per thread) over the k rows of block blockIdx.y.
// Within a block, rows of one h run over consecutive k and their first l
// moves by a few steps: Y and Z(first l) follow by multiplying with the
// atom's unit phases; anything else reloads from the tables.
__global__ __launch_bounds__(kEwFThreads) void k_ew_force(const SysDev *__restrict__ sys) {
  const SysDev &S = *sys;
  const int n = S.n_atoms;
  const int j0 = blockIdx.x * kEwFThreads + (int)threadIdx.x;
  const bool live = j0 < n;
  const int j = live ? j0 : n - 1;
  const int b = blockIdx.y;
  const EW_CONST int *fbr = ew_cp(S.ew_fb_rows);
  const EW_CONST int4 *rows = ew_cp(S.ew_rows);
  const EW_CONST double4 *sk = ew_cp(S.ew_sk);
  const EW_CONST double4 *kd = ew_cp(S.ew_kd);
  const int r0 = fbr[b], r1 = fbr[b + 1];
  const double2 ey1 = S.ew_ty[(size_t)n + j], ez1 = S.ew_tz[(size_t)n + j];
  double fx = 0.0, fy = 0.0, fz = 0.0;
  int ph = -1, pk = 0, plo = 0;
  double2 X = make_double2(0.0, 0.0), Y = X, Zl = X;
  for (int r = r0; r < r1; ++r) {
    const int4 R = rows[r];
    const int kend = rows[r + 1].w;
    if (R.x != ph) {
      X = S.ew_tx[(size_t)R.x * n + j];
      Y = ew_phase(S.ew_ty, n, R.y, j);
      Zl = ew_phase(S.ew_tz, n, R.z, j);
    } else {
      if (R.y == pk + 1) Y = cmul(Y, ey1);
      else Y = ew_phase(S.ew_ty, n, R.y, j);
      const int dl = R.z - plo;
      if (dl >= -4 && dl <= 4) {
        for (int m = 0; m < dl; ++m) Zl = cmul(Zl, ez1);
        for (int m = 0; m < -dl; ++m) Zl = cmulc(Zl, ez1);
      } else {
        Zl = ew_phase(S.ew_tz, n, R.z, j);
      }
    }
    ph = R.x;
    pk = R.y;
    plo = R.z;
    double2 e = cmul(cmul(X, Y), Zl);
    double rs = 0.0;
    for (int q = R.w; q < kend; ++q) {
      const double4 w = sk[q];
      const double t = e.y * w.x - e.x * w.y;
      fz += w.z * t;
      rs += t;
      e = cmul(e, ez1);
    }
    const double4 k0 = kd[R.w];
    fx += k0.x * rs;
    fy += k0.y * rs;
  }
  if (!live) return;
  double *pf = S.ew_pf + (size_t)b * 3 * n;
  pf[j] = fx;
  pf[(size_t)n + j] = fy;
  pf[2 * (size_t)n + j] = fz;
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

// c sum_k b |S|^2 and the virial (NonbondedFullEwaldSystemForce.h:1028-1038,
// 1053-1067): one workgroup, strided partial sums then a fixed tree
__global__ __launch_bounds__(1024) void k_ew_energy(const SysDev *__restrict__ sys) {
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
    for (int m = 0; m < 7; ++m)
      atomicAdd(reinterpret_cast<unsigned long long *>(&acc[slot[m]]),
                static_cast<unsigned long long>(__double2ll_rn(c * red[m][0] * kEScale)));
  }
}

void launch_ewald(const Launch &L, int mode) {
  for (int s = 0; s < L.n_ew; ++s) {
    const EwLaunch &E = L.ew[s];
    if (!E.n || !E.nk) continue;
    if (mode & kEwForces) {
      hipLaunchKernelGGL(k_ew_tables, dim3((E.n + 255) / 256), dim3(256), 0, L.stream, E.sys);
      hipLaunchKernelGGL(k_ew_sf, dim3(E.ntiles * E.nchunks), dim3(kEwSFThreads), E.lds_bytes, L.stream, E.sys);
      hipLaunchKernelGGL(k_ew_ssum, dim3((E.nk + 255) / 256), dim3(256), 0, L.stream, E.sys);
      hipLaunchKernelGGL(k_ew_force, dim3((E.n + kEwFThreads - 1) / kEwFThreads, E.nfb), dim3(kEwFThreads), 0,
                         L.stream, E.sys);
      hipLaunchKernelGGL(k_ew_fsum, dim3((E.n + 255) / 256), dim3(256), 0, L.stream, E.sys);
    }
    if (mode & kEwEnergies) hipLaunchKernelGGL(k_ew_energy, dim3(1), dim3(1024), 0, L.stream, E.sys);
  }
}

}  // namespace pm
