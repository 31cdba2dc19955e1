// pm_gb.hip -- Generalized Born (OBC) implicit solvent on gfx950.
//
// ProtoMol evaluates GB as three dependent all-pairs passes over
// NonbondedSimpleFullSystemForce (GBBornRadii -> GBPartialSum -> GBForce) and
// stores ten N x N intermediate arrays per system (topology/Atom.h:62-83).
// Here each pass recomputes what it needs instead, visiting each unordered
// pair once (64 x 64 tile pairs, one wave each) and keeping per tile pair
// only the 64 + 64 per-atom partial sums, which a per-atom kernel adds in a
// fixed order.  No N x N pair storage, no atomics, a fixed summation order
// per atom (bitwise reproducible), FP32 pair arithmetic on FP64 coordinate
// differences, FP64 accumulation and FP64 per-atom Born-radius algebra.
#include <hip/hip_runtime.h>
#include <math.h>

#include "pm_gb.h"

namespace pm {

// the value GBACEForce.cpp:7 defines as PI (sic), used by its energy and force
constexpr double kAcePi = 3.14169;

void gb_setup_force(ForceDev &P, const pm_force_desc &d) {
  P.gb_offset = d.gb_offset > 0 ? d.gb_offset : 0.09;
  P.gb_a = d.gb_alpha;
  P.gb_b = d.gb_beta;
  P.gb_g = d.gb_gamma;
  const double ein = d.gb_solute_dielec > 0 ? d.gb_solute_dielec : 1.0;
  const double eout = d.gb_solvent_dielec > 0 ? d.gb_solvent_dielec : 80.0;
  P.gb_fac = (1.0 / ein) - (1.0 / eout);
  P.gb_ace = (d.flags & PM_FLAG_GB_ACE) ? 1 : 0;
  P.gb_ace_sigma = d.gb_ace_sigma > 0 ? d.gb_ace_sigma : 2.26 / 418.4;  // GBACEForce.cpp:218-222 defaults
  P.gb_ace_rho_s = d.gb_ace_rho_s > 0 ? d.gb_ace_rho_s : 1.4;
  P.gb_ace_c1 = 24 * kAcePi * P.gb_ace_sigma;
  // PM_FLAG_GB alone: GB + LJ + Coulomb; with PM_FLAG_ALL_PAIRS the pair
  // terms are term_a (LJ or none) and term_b (Coulomb or none), and the GB
  // passes run only with PM_FLAG_GB
  const bool sel = (d.flags & PM_FLAG_ALL_PAIRS) != 0;
  P.gb_on = (d.flags & PM_FLAG_GB) ? 1 : 0;
  P.lj_on = !sel || d.term_a == PM_TERM_LJ ? 1 : 0;
  P.coul_on = !sel || d.term_b == PM_TERM_COULOMB ? 1 : 0;
  P.term_a = P.lj_on ? PM_TERM_LJ : PM_TERM_NONE;
  P.term_b = P.coul_on ? PM_TERM_COULOMB : PM_TERM_NONE;
}

void gb_setup_system(SysDev &S, const double *rho, const double *scaling, double *born, double *drfac,
                     double *partial, double *spare) {
  S.gb_rho = rho;
  S.gb_S = scaling;
  S.gb_born = born;
  S.gb_drfac = drfac;
  S.gb_partial = partial;
  S.gb_spare = spare;
}

namespace {

__device__ __forceinline__ int find_sys_gb(const int *__restrict__ prefix, int nsys, int w) {
  int lo = 0, hi = nsys;
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (prefix[mid] <= w) lo = mid;
    else hi = mid;
  }
  return lo;
}

__device__ __forceinline__ unsigned long long wave_ballot_gb(bool p) { return __ballot(p); }

// reciprocal by the hardware instruction (1 ulp)
__device__ __forceinline__ float rcp(float v) { return __builtin_amdgcn_rcpf(v); }

// contribution of atom j to the burial term of atom i (GBBornRadii.cpp:59-91);
// iori = 1/ori, idist = 1/dist
__device__ __forceinline__ float burial_term(float dist, float idist, float ori, float iori, float orj, float sj) {
  const float sorj = sj * orj;
  float L, U;
  if (ori >= dist + sorj) {
    L = 1.0f;
    U = 1.0f;
  } else {
    const float t = fabsf(dist - sorj);
    L = ori > t ? ori : t;
    U = dist + sorj;
  }
  const float iL = rcp(L), iU = rcp(U), iL2 = iL * iL, iU2 = iU * iU;
  const float C = ori < sorj - dist ? 2.0f * (iori - iL) : 0.0f;
  return (iL - iU) + (0.25f * dist) * (iU2 - iL2) + (0.5f * idist) * __logf(L * iU) +
         (0.25f * sorj * sorj * idist) * (iL2 - iU2) + C;
}

// d(burial_i)/d(r_ij) (GBForce.cpp:55-126).  The reference's sum of ten
// terms is grouped by 1/L and 1/U (the same terms, fewer operations):
//   [dL] g(1/L) - [dU] g(1/U) + (1 + s^2/d^2)/8 (1/U^2 - 1/L^2)
//     - log(L/U) / (4 d^2) + [rho~_i <= s - d] [dL] 2/L^2
// with g(x) = x (1/(4d) - x/2 + (d^2 - s^2) x^2 / (4d)), s = S_j rho~_j,
// dL = [rho~_i <= d - s], dU = [rho~_i < d + s]; the reference's term
// (U/(4 d L))(dL/U - L dU/U^2) is (dL/L - dU/U)/(4 d), the x/(4d) part of g.
// r2 = d^2, id4 = 1/(4d), q4 = 1/(4d^2)
__device__ __forceinline__ float dburial(float dist, float r2, float id4, float q4, float ori, float orj, float sj) {
  const float sorj = sj * orj;
  float L, U;
  if (ori >= dist + sorj) {
    L = 1.0f;
    U = 1.0f;
  } else {
    const float t = fabsf(dist - sorj);
    L = ori > t ? ori : t;
    U = dist + sorj;
  }
  const float iL = rcp(L), iU = rcp(U), iL2 = iL * iL;
  const bool dL = ori <= dist - sorj, dU = ori < dist + sorj;
  const float s2 = sorj * sorj;
  const float k4 = (r2 - s2) * id4;
  const float gL = iL * fmaf(iL, fmaf(k4, iL, -0.5f), id4);
  const float gU = iU * fmaf(iU, fmaf(k4, iU, -0.5f), id4);
  float T = (dL ? gL : 0.0f) - (dU ? gU : 0.0f);
  T = fmaf(fmaf(0.5f * s2, q4, 0.125f), fmaf(iU, iU, -iL2), T);
  T = fmaf(-q4, __logf(L * iU), T);
  return (dL && ori <= sorj - dist) ? T + 2.0f * iL2 : T;
}

}  // namespace

// The three passes visit each unordered pair once, in 64 x 64 tile pairs
// (it <= jt), one wave per tile pair: lane l holds i atom it*64+l in
// registers; at step r it meets j atom jt*64+((l+r)&63), read from the wave's
// copy of the j tile in LDS, and adds the pair's terms to its own i sum and
// to a j sum that travels one lane down after each step, so after the sweep
// every j sum sits in a known lane -- no atomics.  A diagonal tile takes
// steps 1..32 (step 32 on lanes 0..31 only), which visits each of its pairs
// once.  Both sides go to the tile pair's slot of gb_fpart, and a per-atom
// kernel adds an atom's partials over the tile pairs in a fixed order.
constexpr int kGBPWaves = 4;

// tile pair (it, jt) of the wave's index p within its system (row-major over it <= jt)
__device__ __forceinline__ void gb_tile_pair(int p, int nt, int &it, int &jt) {
  it = 0;
  while (p >= nt - it) {
    p -= nt - it;
    ++it;
  }
  jt = it + p;
}

// slot of tile pair (a, b), a <= b
__device__ __forceinline__ double *gb_slot(const SysDev &S, int a, int b) {
  return S.gb_fpart + (size_t)(a * S.gb_nt - a * (a - 1) / 2 + (b - a)) * kGBPairDoubles;
}

// an atom's partials over the tile pairs, in tile order: the j side of (u, t)
// for u < t, both sides of (t, t), the i side of (t, u) for u > t; side
// offsets: i side at `io`, j side at `jo` within a slot
__device__ __forceinline__ double gb_atom_sum(const SysDev &S, int t, int lane, int io, int jo) {
  double v = 0.0;
  for (int u = 0; u < S.gb_nt; ++u) {
    const double *slot = gb_slot(S, min(u, t), max(u, t));
    if (u >= t) v += slot[io + lane];
    if (u <= t) v += slot[jo + lane];
  }
  return v;
}

// passes 1 and 2: PASS 1 the burial terms of both atoms of a pair
// (GBBornRadii.cpp:15-128), PASS 2 their partial-sum terms
// (GBPartialSum.cpp:11-42: q_i q_l f^-3 / 2 e^{-r^2/4RiRl} (R_l + r^2 / 4 R_i))
template <int PASS>
__global__ __launch_bounds__(kGBPWaves * 64) void k_gb_sym(SysDev *sys, const int *pp, int nsys, ForceDev P) {
  __shared__ double4 jpos[kGBPWaves][kGBI];
  __shared__ float4 ja[kGBPWaves][kGBI];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int w = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  if (w >= pp[nsys]) return;
  const int s = find_sys_gb(pp, nsys, w);
  const SysDev &S = sys[s];
  const int n = S.n_atoms;
  int it, jt;
  gb_tile_pair(w - pp[s], S.gb_nt, it, jt);
  const bool diag = it == jt;
  {
    const int j = jt * kGBI + lane;
    double4 pos = make_double4(0.0, 0.0, 0.0, 0.0);
    float4 a = make_float4(1.f, 1.f, 1.f, 1.f);
    if (j < n) {
      pos = make_double4(S.x[j], S.x[n + j], S.x[2 * n + j], 0.0);
      if (PASS == 1) {
        const float o = (float)(S.gb_rho[j] - P.gb_offset);
        a = make_float4(o, (float)S.gb_S[j], rcp(o), 0.f);
      } else {
        const float B = (float)S.gb_born[j];
        a = make_float4((float)S.charge[j], B, rcp(B), 0.f);
      }
    }
    jpos[wv][lane] = pos;
    ja[wv][lane] = a;
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  const int i = it * kGBI + lane;
  const bool vi = i < n;
  double xi = 0, yi = 0, zi = 0;
  float u0 = 1.f, u1 = 1.f, u2 = 1.f;  // PASS 1: rho~, S, 1/rho~; PASS 2: q, B, 1/B
  if (vi) {
    xi = S.x[i];
    yi = S.x[n + i];
    zi = S.x[2 * n + i];
    if (PASS == 1) {
      u0 = (float)(S.gb_rho[i] - P.gb_offset);
      u1 = (float)S.gb_S[i];
      u2 = rcp(u0);
    } else {
      u0 = (float)S.charge[i];
      u1 = (float)S.gb_born[i];
      u2 = rcp(u1);
    }
  }
  double si = 0.0, sj = 0.0;
  const int r0 = diag ? 1 : 0, r1 = diag ? 32 : 63;
  for (int r = r0; r <= r1; ++r) {
    const int jl = (lane + r) & 63;
    const int j = jt * kGBI + jl;
    if (vi && j < n && !(diag && r == 32 && lane >= 32)) {
      const double4 pj = jpos[wv][jl];
      const float4 aj = ja[wv][jl];
      const float dx = (float)(pj.x - xi), dy = (float)(pj.y - yi), dz = (float)(pj.z - zi);
      const float r2 = dx * dx + dy * dy + dz * dz;
      if (PASS == 1) {
        const float idist = rsqrtf(r2), dist = r2 * idist;
        si += (double)burial_term(dist, idist, u0, u2, aj.x, aj.y);
        sj += (double)burial_term(dist, idist, aj.x, aj.z, u0, u1);
      } else {
        const float Bj = aj.y;
        const float ex = __expf(-r2 * (0.25f * u2 * aj.z));  // exp(-r^2 / (4 Bi Bj))
        const float inv = rsqrtf(r2 + u1 * Bj * ex);          // 1 / f_GB
        const float part = u0 * aj.x * inv * inv * 0.5f * inv * ex;
        si += (double)(part * (Bj + r2 * (0.25f * u2)));
        sj += (double)(part * (u1 + r2 * (0.25f * aj.z)));
      }
    }
    sj = __shfl(sj, (lane + 1) & 63);  // lane l now holds the sum of j (l + r + 1) & 63
  }
  double *slot = S.gb_fpart + (size_t)(w - pp[s]) * kGBPairDoubles;
  slot[lane] = si;
  slot[kGBI + ((lane + r1 + 1) & 63)] = sj;
}

// per atom after pass 1: the burial sum, then the OBC Born radius
// (GBBornRadii::postProcess, GBBornRadii.cpp:145-173) and the per-atom factor
// of dR_i/dr_ij (GBForce.cpp:126: R_i^2 rho~_i (1 - tanh^2) dtanh/dpsi / rho_i);
// after pass 2: the partial sum
template <int PASS>
__global__ __launch_bounds__(kGBI) void k_gb_sums(SysDev *sys, const int *gp, int nsys, ForceDev P) {
  const int tile = blockIdx.x;
  if (tile >= gp[nsys]) return;
  const int s = find_sys_gb(gp, nsys, tile);
  SysDev &S = sys[s];
  const int t = tile - gp[s], lane = threadIdx.x;
  const int i = t * kGBI + lane;
  if (i >= S.n_atoms) return;
  const double v = gb_atom_sum(S, t, lane, 0, kGBI);
  if (PASS == 2) {
    S.gb_partial[i] = v;
    return;
  }
  const double rho = S.gb_rho[i], orid = rho - P.gb_offset;
  const double psi = 0.5 * v * orid;
  const double tp = P.gb_a * psi - P.gb_b * psi * psi + P.gb_g * psi * psi * psi;
  const double th = tanh(tp);
  const double B = 1.0 / ((1.0 / orid) - (1.0 / rho) * th);
  const double tderv = P.gb_a - 2.0 * P.gb_b * psi + 3.0 * P.gb_g * psi * psi;
  S.gb_born[i] = B;
  S.gb_drfac[i] = B * B * orid * (1.0 - th * th) * tderv * (1.0 / rho);
  // ACE: c_i / R_i^7 with c_i = (rho_i + rho_s)^2 rho_i^6 (GBACEForce.cpp:52-54, 155)
  if (P.gb_ace) {
    const double rs = rho + P.gb_ace_rho_s, r3 = rho * rho * rho, iB = 1.0 / B, iB3 = iB * iB * iB;
    S.gb_spare[i] = rs * rs * (r3 * r3) * (iB3 * iB3 * iB);
  }
}

// pass 3: GB pair force and energy (GBForce.cpp:14-175) plus the all-pairs
// Lennard-Jones and Coulomb terms (OneAtomPair<Vacuum, Universal, ...> with
// the exclusion table), each unordered pair once, as the reference's pair
// loop visits it.  One wave per 64 x 64 tile pair (it <= jt): lane l holds i
// atom it*64+l in registers; at step r it meets j atom jt*64+((l+r)&63), read
// from the wave's copy of the j tile in LDS, and adds the pair force to its
// own i sum and to a j sum that travels one lane down after each step, so
// after the sweep every j sum sits in a known lane -- no atomics.  A diagonal
// tile takes steps 1..32 (step 32 on lanes 0..31 only), which visits each of
// its pairs once.  Both sides go to the tile pair's slot of gb_fpart;
// k_gb_fsum adds an atom's partials in a fixed order.
// EN: energies; VIR: with the energies, the virials of each pair (OneAtomPairNoExclusion.h:
// 73-81, OneAtomPair.h:95-103): f (x) d to the atomic virial, and f (x)
// (R_J - R_I) to the molecular one for pairs of two molecules.
template <bool VIR, bool EN>
__global__ __launch_bounds__(kGBPWaves * 64) void k_gb_pairs(SysDev *sys, const int *pp, int nsys, ForceDev P) {
  constexpr bool energies = EN;
  __shared__ double4 jpos[kGBPWaves][kGBI];
  __shared__ float4 ja[kGBPWaves][kGBI], jb[kGBPWaves][kGBI], jc[kGBPWaves][kGBI];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int w = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  if (w >= pp[nsys]) return;
  const int s = find_sys_gb(pp, nsys, w);
  SysDev &S = sys[s];
  const int n = S.n_atoms, nt = S.gb_nt;
  int it, jt;
  gb_tile_pair(w - pp[s], nt, it, jt);
  const bool diag = it == jt;
  const bool peratom = S.lj_mode == PM_LJ_PERATOM;
  // the j tile into this wave's LDS records
  {
    const int j = jt * kGBI + lane;
    double4 pos = make_double4(0.0, 0.0, 0.0, 0.0);
    float4 a = make_float4(1.f, 1.f, 1.f, 1.f), b = make_float4(0.f, 0.f, 0.f, 0.f), c = b;
    if (j < n) {
      pos = make_double4(S.x[j], S.x[n + j], S.x[2 * n + j], 0.0);
      const float q = (float)S.charge[j], B = (float)S.gb_born[j], iB = rcp(B);
      a = make_float4((float)(S.gb_rho[j] - P.gb_offset), (float)S.gb_S[j], B, iB);
      b = make_float4(q, 0.5f * q * q * iB * iB, (float)S.gb_drfac[j], (float)S.gb_partial[j]);
      c.x = peratom ? S.sig_half[j] : __int_as_float(S.type[j]);
      c.y = peratom ? S.eps_sqrt[j] : 0.f;
      c.z = P.gb_ace ? (float)S.gb_spare[j] : 0.f;
      c.w = __int_as_float(S.molecule[j]);
    }
    jpos[wv][lane] = pos;
    ja[wv][lane] = a;
    jb[wv][lane] = b;
    jc[wv][lane] = c;
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  const int i = it * kGBI + lane;
  const bool vi = i < n;
  double xi = 0, yi = 0, zi = 0;
  float qi = 0.f, Bi = 1.f, ori = 1.f, Si = 1.f, dri = 0.f, Pi = 0.f, sgi = 0.f, epi = 0.f, aki = 0.f;
  int ti = 0, moli = 0, ep0 = 0, ee = 0, pmin = 0x7fffffff, pmax = -1;
  if (vi) {
    xi = S.x[i];
    yi = S.x[n + i];
    zi = S.x[2 * n + i];
    qi = (float)S.charge[i];
    Bi = (float)S.gb_born[i];
    ori = (float)(S.gb_rho[i] - P.gb_offset);
    Si = (float)S.gb_S[i];
    dri = (float)S.gb_drfac[i];
    Pi = (float)S.gb_partial[i];
    if (peratom) {
      sgi = S.sig_half[i];
      epi = S.eps_sqrt[i];
    } else {
      ti = S.type[i];
    }
    aki = P.gb_ace ? (float)S.gb_spare[i] : 0.f;
    moli = S.molecule[i];
    ep0 = S.excl_start[i];
    ee = S.excl_start[i + 1];
    if (ep0 < ee) {  // partners are sorted
      pmin = S.excl_partner[ep0];
      pmax = S.excl_partner[ee - 1];
    }
  }
  // does any pair of this tile pair carry an exclusion? (wave-uniform)
  const bool any_ex = wave_ballot_gb(pmax >= jt * kGBI && pmin < (jt + 1) * kGBI) != 0;
  const float fac = (float)P.gb_fac;
  const float iBi = rcp(Bi), i4Bi = 0.25f * iBi;
  const float qiB2 = 0.5f * qi * qi * iBi * iBi;  // q_i^2 / (2 B_i^2)
  const float csf = S.csf;
  const float ace_c1 = (float)P.gb_ace_c1;
  const float lj_on = P.lj_on ? 1.0f : 0.0f, coul_on = P.coul_on ? 1.0f : 0.0f;
  double fx = 0.0, fy = 0.0, fz = 0.0, gx = 0.0, gy = 0.0, gz = 0.0, egb = 0.0, elj = 0.0, eco = 0.0;
  double vir[6] = {0, 0, 0, 0, 0, 0}, mvir[6] = {0, 0, 0, 0, 0, 0};
  const int r0 = diag ? 1 : 0, r1 = diag ? 32 : 63;
  for (int r = r0; r <= r1; ++r) {
    const int jl = (lane + r) & 63;
    const int j = jt * kGBI + jl;
    // every lane evaluates (padding records sit far away, so the math stays
    // finite); a lane without a pair adds zeros, and the accumulations need
    // no per-lane predication
    const bool ok = vi && j < n && !(diag && r == 32 && lane >= 32);
    {
      int cls = 0;
      if (any_ex)
        for (int q = ep0; q < ee; ++q)
          if (S.excl_partner[q] == j) cls = S.excl_class[q];
      const double4 pj = jpos[wv][jl];
      const float4 aj = ja[wv][jl], bj4 = jb[wv][jl], cj4 = jc[wv][jl];
      const float dx = (float)(pj.x - xi), dy = (float)(pj.y - yi), dz = (float)(pj.z - zi);
      const float r2 = ok ? dx * dx + dy * dy + dz * dz : 1.0e8f;
      const float rinv = rsqrtf(r2);
      const float qq = qi * bj4.x;
      float fs = 0.0f;
      if (P.gb_on) {  // wave-uniform
        const float dist = r2 * rinv;
        const float Bj = aj.z, iBj = aj.w;
        const float BB = Bi * Bj;
        const float ex = __expf(-r2 * (i4Bi * iBj));  // exp(-r^2 / (4 Bi Bj))
        const float inv = rsqrtf(r2 + BB * ex);       // 1 / f_GB
        const float id4 = 0.25f * rinv, q4 = id4 * rinv;
        const float dRi = dri * dburial(dist, r2, id4, q4, ori, aj.x, aj.y);
        const float dRj = bj4.z * dburial(dist, r2, id4, q4, aj.x, ori, Si);
        const float c0 = qq * inv * inv * 0.5f * inv;
        const float ai = Bj + r2 * i4Bi, bj = Bi + r2 * (0.25f * iBj);
        float force = -c0 * ((2.0f * dist - 0.5f * ex * dist) + ex * dRi * ai + ex * dRj * bj) * rinv;
        force -= qiB2 * dRi * rinv;
        force -= bj4.y * dRj * rinv;
        const float part = c0 * ex;
        force -= (Pi - part * ai) * (dRi * rinv);
        force -= (bj4.w - part * bj) * (dRj * rinv);
        force *= fac;
        // GBACEForce.cpp:155: c1 (c_i R_i^-7 dR_i/dr + c_j R_j^-7 dR_j/dr) / r
        if (P.gb_ace) force += ace_c1 * (aki * dRi + cj4.z * dRj) * rinv;
        fs = force;
        if (energies) egb += ok ? (double)(-qq * inv * fac) : 0.0;
      }
      if (ok && cls != PM_EXCL_FULL) {
        float A, Bc;
        if (peratom) {
          const float sg = sgi + cj4.x, e2 = epi * cj4.y;
          const float sg2 = sg * sg, sg6 = sg2 * sg2 * sg2;
          Bc = 4.0f * e2 * sg6;
          A = Bc * sg6;
          if (cls == PM_EXCL_MODIFIED) {
            A *= S.lj_scaling;
            Bc *= S.lj_scaling;
          }
        } else {
          const float4 q4 = S.lj_table[ti * S.n_types + __float_as_int(cj4.x)];
          A = cls == PM_EXCL_MODIFIED ? q4.z : q4.x;
          Bc = cls == PM_EXCL_MODIFIED ? q4.w : q4.y;
        }
        const float r2inv = rinv * rinv;
        const float r6 = r2inv * r2inv * r2inv;
        const float r6B = lj_on * Bc * r6, r12A = lj_on * A * r6 * r6;
        float ec = coul_on * qq * rinv;
        if (cls == PM_EXCL_MODIFIED) ec *= csf;
        fs += (12.0f * r12A - 6.0f * r6B) * r2inv + ec * r2inv;
        if (energies) {
          elj += (double)(r12A - r6B);
          eco += (double)ec;
        }
      }
      fs = ok ? fs : 0.0f;
      const float px = fs * dx, py = fs * dy, pz = fs * dz;
      fx -= (double)px;
      fy -= (double)py;
      fz -= (double)pz;
      gx += (double)px;
      gy += (double)py;
      gz += (double)pz;
      if (VIR && energies) {
        const bool other = ok && __float_as_int(cj4.w) != moli;
        if (other || P.virial) {
          vir[0] += (double)(px * dx);
          vir[1] += (double)(px * dy);
          vir[2] += (double)(px * dz);
          vir[3] += (double)(py * dy);
          vir[4] += (double)(py * dz);
          vir[5] += (double)(pz * dz);
        }
        if (other && P.molvirial) {
          const int mj = __float_as_int(cj4.w), nm = S.n_mol;
          const double cx = S.com[mj] - S.com[moli], cy = S.com[nm + mj] - S.com[nm + moli],
                       cz = S.com[2 * nm + mj] - S.com[2 * nm + moli];
          mvir[0] += (double)px * cx;
          mvir[1] += (double)px * cy;
          mvir[2] += (double)px * cz;
          mvir[3] += (double)py * cy;
          mvir[4] += (double)py * cz;
          mvir[5] += (double)pz * cz;
        }
      }
    }
    // the j sums move one lane down: lane l then holds the sum of j (l + r + 1) & 63
    gx = __shfl(gx, (lane + 1) & 63);
    gy = __shfl(gy, (lane + 1) & 63);
    gz = __shfl(gz, (lane + 1) & 63);
  }
  double *slot = S.gb_fpart + (size_t)(w - pp[s]) * kGBPairDoubles;
  slot[lane] = fx;
  slot[kGBI + lane] = fy;
  slot[2 * kGBI + lane] = fz;
  const int jf = (lane + r1 + 1) & 63;
  slot[3 * kGBI + jf] = gx;
  slot[4 * kGBI + jf] = gy;
  slot[5 * kGBI + jf] = gz;
  if (!energies) return;
  // wave sums in a fixed order, then fixed-point totals
  constexpr int NV = VIR ? 15 : 3;
  double v[NV];
  v[0] = egb;
  v[1] = elj;
  v[2] = eco;
  if (VIR)
    for (int c = 0; c < 6; ++c) {
      v[3 + c] = vir[c];
      v[9 + c] = mvir[c];
    }
  for (int c = 0; c < NV; ++c)
    for (int o = 32; o > 0; o >>= 1) v[c] += __shfl_xor(v[c], o);
  if (lane == 0) {
    long long *acc = acc_copy(S.acc, w);
    atomicAdd(reinterpret_cast<unsigned long long *>(&acc[kAccCoul]), (unsigned long long)__double2ll_rn((v[0] + v[2]) * kEScale));
    atomicAdd(reinterpret_cast<unsigned long long *>(&acc[kAccLJ]), (unsigned long long)__double2ll_rn(v[1] * kEScale));
    if (VIR) {
      const int vs[12] = {kAccVxx, kAccVxy, kAccVxz, kAccVyy, kAccVyz, kAccVzz,
                          kAccMVxx, kAccMVxy, kAccMVxz, kAccMVyy, kAccMVyz, kAccMVzz};
      for (int c = 0; c < 12; ++c)
        atomicAdd(reinterpret_cast<unsigned long long *>(&acc[vs[c]]), (unsigned long long)__double2ll_rn(v[3 + c] * kEScale));
    }
  }
}

// Per atom: its force partials of every tile pair, in tile order (the
// j side of (u, t) for u < t, both sides of (t, t), the i side of (t, u) for
// u > t), into the fixed-point slots; with the energies, the GB self energy
// (GBForce.cpp:50-60, 230-242) and the ACE energy (GBACEForce.cpp:40-50,
// 160-170) of the atom.
__global__ __launch_bounds__(kGBI) void k_gb_fsum(SysDev *sys, const int *gp, int nsys, ForceDev P, int energies) {
  const int tile = blockIdx.x;
  if (tile >= gp[nsys]) return;
  const int s = find_sys_gb(gp, nsys, tile);
  SysDev &S = sys[s];
  const int n = S.n_atoms, nt = S.gb_nt;
  const int t = tile - gp[s], lane = threadIdx.x;
  const int i = t * kGBI + lane;
  double e = 0.0;
  if (i < n) {
    double f[3];
    for (int c = 0; c < 3; ++c) f[c] = gb_atom_sum(S, t, lane, c * kGBI, (3 + c) * kGBI);
    S.fx[i] += __double2ll_rn(f[0] * kFScale);
    S.fy[i] += __double2ll_rn(f[1] * kFScale);
    S.fz[i] += __double2ll_rn(f[2] * kFScale);
    if (energies && P.gb_on) {
      e = -0.5 * (double)S.charge[i] * (double)S.charge[i] * (1.0 / S.gb_born[i]) * P.gb_fac;
      if (P.gb_ace) {
        // 4 PI sigma (rho_i + rho_s)^2 (rho_i / R_i)^6
        const double rho = S.gb_rho[i], rs = rho + P.gb_ace_rho_s, p = rho / S.gb_born[i], p3 = p * p * p;
        e += 4 * kAcePi * P.gb_ace_sigma * rs * rs * (p3 * p3);
      }
    }
  }
  if (!energies) return;
  for (int o = 32; o > 0; o >>= 1) e += __shfl_xor(e, o);
  if (lane == 0)
    atomicAdd(reinterpret_cast<unsigned long long *>(&acc_copy(S.acc, tile)[kAccCoul]),
              (unsigned long long)__double2ll_rn(e * kEScale));
}

void launch_gb(const Launch &L, const ForceDev &P, bool energies) {
  if (!L.n_gb_tiles_total) return;
  dim3 gt(L.n_gb_tiles_total), bt(kGBI);
  dim3 gp2((L.n_gb_pairs_total + kGBPWaves - 1) / kGBPWaves), bp(kGBPWaves * 64);
  if (P.gb_on) {  // Born radii, then partial sums
    hipLaunchKernelGGL(k_gb_sym<1>, gp2, bp, 0, L.stream, L.sys, L.gbp_prefix, L.nsys, P);
    hipLaunchKernelGGL(k_gb_sums<1>, gt, bt, 0, L.stream, L.sys, L.gb_prefix, L.nsys, P);
    hipLaunchKernelGGL(k_gb_sym<2>, gp2, bp, 0, L.stream, L.sys, L.gbp_prefix, L.nsys, P);
    hipLaunchKernelGGL(k_gb_sums<2>, gt, bt, 0, L.stream, L.sys, L.gb_prefix, L.nsys, P);
  }
  if (energies && (P.virial || P.molvirial))
    hipLaunchKernelGGL((k_gb_pairs<true, true>), gp2, bp, 0, L.stream, L.sys, L.gbp_prefix, L.nsys, P);
  else if (energies)
    hipLaunchKernelGGL((k_gb_pairs<false, true>), gp2, bp, 0, L.stream, L.sys, L.gbp_prefix, L.nsys, P);
  else
    hipLaunchKernelGGL((k_gb_pairs<false, false>), gp2, bp, 0, L.stream, L.sys, L.gbp_prefix, L.nsys, P);
  hipLaunchKernelGGL(k_gb_fsum, dim3(L.n_gb_tiles_total), dim3(kGBI), 0, L.stream, L.sys, L.gb_prefix, L.nsys, P,
                     energies ? 1 : 0);
}

}  // namespace pm
