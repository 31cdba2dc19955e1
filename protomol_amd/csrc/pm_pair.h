// pm_pair.h -- the pair-force kernel k_pair and its launch, included by the
// pm_pair_*.hip sources; each instantiates the kernels of one term
// combination, so the instantiations compile in parallel.
#pragma once
#include "pm_device.h"
#include "pm_rebuild.h"

namespace pm {

// ------------------------------------------------------------ pair force

// compile-time switch kinds for the hot path; kSwGeneric dispatches at run time
constexpr int kSwGeneric = -1;

// OUTER: the switch's cutoff may be shorter than the pair's inclusion cutoff
// (OneAtomPairTwo with two different cutoffs), so S = 0 beyond it.
// C2SwitchingFunction.h:42-53: S = s1 c2 c4 and dS = s3 (c2^2 - c4),
// c2 = cut^2 - r^2, c4 = c2 (s2 + 2 r^2), s2 = cut^2 - 3 on^2, s3 = 4 s1 =
// 4 / D^3, D = cut^2 - on^2.  Factored: with w = r^2 - on^2,
// S = a1 c2^2 (w + D/2) and dS = -6 a1 c2 w (a1 = 2 s1), so
// F S - E dS = a1 c2 (c2 (w + D/2) F + 6 w E).  Inside r_on, w is clamped to
// 0 (an integer max of the bits: no canonicalizing float max), where S = 1
// and dS = 0 to FP32 rounding: the unswitched values.
template <int SW, bool OUTER = true>
__device__ __forceinline__ void apply_switch_t(const SwitchDev &sw, float r2, float &E, float &F) {
  if (SW == PM_SWITCH_C2) {
    const float w = __int_as_float(max(__float_as_int(r2 - sw.on2), 0));
    const float c2 = sw.c2d - w;
    const float cp = c2 * (w + sw.c2hd);
    const float ac = sw.c2a1 * c2;
    float Fs = ac * fmaf(w * E, 6.0f, cp * F);
    float Es = ac * (cp * E);
    if (OUTER) {
      const bool outer = r2 > sw.cut2;
      Fs = outer ? 0.0f : Fs;
      Es = outer ? 0.0f : Es;
    }
    F = Fs;
    E = Es;
  } else if (SW == PM_SWITCH_CUTOFF || SW == PM_SWITCH_NONE) {
    // inside the inclusion cutoff the value is 1 (CutoffSwitchingFunction.h:41-44)
  } else {
    apply_switch(sw, r2, E, F);
  }
}

// The LJ term through the factored C2 switch, with coefficients that come
// pre-scaled (A', B', A12', B6') = (12 s1 A, 12 s1 B, 24 s1 A, 12 s1 B) (the
// LJ rows of the table in LDS are staged that way), so the LJ energy enters the
// switch as 6 a1 E and its force as a1 F.  With r^-6 taken out of both,
// F S - E dS = c2 r^-6 (c2 (w + D/2) r^-2 (A12' r^-6 - B6') + w (A' r^-6 - B')).
// E returns the switched energy E S.
template <bool OUTER>
__device__ __forceinline__ void lj_c2_pre(const SwitchDev &sw, float r2, float r2inv, float r6, float A, float B,
                                          float A12, float B6, float &E, float &F) {
  const float pe = fmaf(A, r6, -B), pg = fmaf(A12, r6, -B6);
  const float w = __int_as_float(max(__float_as_int(r2 - sw.on2), 0));
  const float c2 = sw.c2d - w;
  const float cp = c2 * (w + sw.c2hd);
  float Fs = (c2 * r6) * fmaf(cp * r2inv, pg, w * pe);
  float Es = ((r6 * pe) * c2) * cp * (1.0f / 6.0f);
  if (OUTER) {
    const bool outer = r2 > sw.cut2;
    Fs = outer ? 0.0f : Fs;
    Es = outer ? 0.0f : Es;
  }
  F = Fs;
  E = Es;
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

// gfx950 cross-row swaps: swap32 exchanges the upper 32 lanes of a with the
// lower 32 lanes of b; swap16 exchanges the odd 16-lane rows of a with the
// even rows of b.  Afterwards a + b holds, in each half (row), the sum of one
// register's two halves (rows): the first register's in the lower half (even
// rows), the second's in the upper half (odd rows).
__device__ __forceinline__ void swap32(float &a, float &b) {
  const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(a), __float_as_uint(b), false, false);
  a = __uint_as_float(r[0]);
  b = __uint_as_float(r[1]);
}
__device__ __forceinline__ void swap16(float &a, float &b) {
  const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(a), __float_as_uint(b), false, false);
  a = __uint_as_float(r[0]);
  b = __uint_as_float(r[1]);
}

// The i forces of a super-cluster: v[3 c + k] holds component k of i cluster
// c for the lane's (jj, ii) = (lane >> 3, lane & 7).  Sums each of the 24
// over the 8 lanes of equal ii (the jj lanes) and scatters the results so
// lane (jj, ii) ends with cluster jj's three components for atom ii in v[0..2]:
// halving steps over lane bit 5 (swap32), bit 4 (swap16) and bit 3 (DPP row
// rotate by 8), each keeping the half of the values that the lane's bit
// selects.  Fixed order: the result bits depend on the inputs only.
__device__ __forceinline__ void reduce_scatter_i(float (&v)[3 * kScCl], int lane) {
#pragma unroll
  for (int m = 0; m < 12; ++m) {
    swap32(v[m], v[m + 12]);
    v[m] += v[m + 12];
  }
#pragma unroll
  for (int m = 0; m < 6; ++m) {
    swap16(v[m], v[m + 6]);
    v[m] += v[m + 6];
  }
  const bool b0 = (lane >> 3) & 1;
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    const float send = b0 ? v[m] : v[m + 3];
    const float keep = b0 ? v[m + 3] : v[m];
    v[m] = keep + __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(send), 0x128, 0xF, 0xF, false));
  }
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
// PRE: the LJ coefficients come pre-scaled for the C2 switch's factored form
// (A, B, A12, B6 times 12 s1, 12 s1, 2 s1, 2 s1)
template <int TA, int TB, int SWA, int SWB, bool SAME, bool PRE = false>
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
    float E, F;
    if (PRE) {
      lj_c2_pre<!SAME>(P.sw_a, r2, r2inv, r6, A, B, A12, B6, E, F);
    } else {
      E = r6 * fmaf(A, r6, -B);
      F = r6 * fmaf(A12, r6, -B6) * r2inv;
      apply_switch_t<SWA, !SAME>(P.sw_a, r2, E, F);
    }
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
// where a wave takes its LJ coefficients (wave-uniform, one copy of the entry
// loop each): the (A, B, 12 A, 6 B) table staged in LDS (at most kLJLds
// entries), the same table read through the caches, or per-atom mixing
constexpr int kLjLds = 0, kLjGlobal = 1, kLjPerAtom = 2;

template <int TA, int TB, int SWA, int SWB, int KIND, bool SAME, int LJM>
__device__ __forceinline__ void pair_wave(const SysDev *__restrict__ S, const int sc, const int part, const int split,
                                          const int w, const ForceDev &P) {
  constexpr bool MV = KIND == kPairForcesEnergiesMV;
  constexpr bool EN = KIND == kPairForcesEnergies || MV;
  constexpr bool CEN = KIND == kPairCensus;
  constexpr bool peratom = LJM == kLjPerAtom;
  constexpr bool big_table = LJM == kLjGlobal;
  // the LDS table's rows pre-scaled for the C2 switch's factored form
  constexpr bool PRE = TA == PM_TERM_LJ && SWA == PM_SWITCH_C2 && LJM == kLjLds;
  extern __shared__ float4 pair_lds[];
  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;
  char *wb = reinterpret_cast<char *>(pair_lds) + (size_t)wv * kPairWaveLds;
  float4 *ljs = reinterpret_cast<float4 *>(wb);
  float4 *lx = reinterpret_cast<float4 *>(wb + kLJLds * 16);
  int *linfo = reinterpret_cast<int *>(wb + kLJLds * 16 + kWave * 16);  // byte offset of row ti of the LJ table
  int *lorig = linfo + kWave;
  float2 *lpp = reinterpret_cast<float2 *>(lorig + kWave);
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
  const int T = S->n_types;
  const PM_GLOBAL char *gtab12 = gp(reinterpret_cast<const char *>(S->lj_table12));
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
  // LJ table (T <= 16; larger ones stay in global memory) as (A, B, 12 A, 6 B),
  // or with a C2 switch as (12 s1) (A, B, 2 A, B): the energy 6 a1 E and force
  // a1 F the factored switch takes (a1 = 2 s1); the modified-pair
  // coefficients stay in the global table
  const PM_GLOBAL float4 *gtab = gp(S->lj_table);
  if (TA == PM_TERM_LJ && !peratom && !big_table) {
    const float sc = P.sw_a.c2sc;
    for (int k = lane; k < T * T; k += kWave) {
      const float4 g = gtab[k];
      ljs[k] = PRE ? make_float4(sc * g.x, sc * g.y, 2.0f * (sc * g.x), sc * g.y)
                   : make_float4(g.x, g.y, 12.0f * g.x, 6.0f * g.y);
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
    const PM_LDS char *ljrow = (const PM_LDS char *)(ljs) + tj * 16;  // + row offset = (ti, tj) entry
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
      // two compares per cutoff: inside the band's lower edge is in, above
      // its upper edge out, between them the FP64 recheck decides
      unsigned long long mA, mB;
      if (SAME) {
        mA = lane_mask(r2 <= P.cutA2lo) & am;
        const unsigned long long near = lane_mask(r2 <= P.cutA2hi) & am & ~mA;
        if (__builtin_expect(near != 0, 0)) {
          bool in = lane_bit(mA);
          if (lane_bit(near)) in = min_image_r2(*S, lorig[c * kCluster + ii], orig[sj]) <= P.cutA2d;
          mA = lane_mask(in);
        }
        mB = mA;
      } else {
        mA = TA != PM_TERM_NONE ? lane_mask(r2 <= P.cutA2lo) & am : 0ull;
        mB = TB != PM_TERM_NONE ? lane_mask(r2 <= P.cutB2lo) & am : 0ull;
        const unsigned long long nearA = TA != PM_TERM_NONE ? lane_mask(r2 <= P.cutA2hi) & am & ~mA : 0ull;
        const unsigned long long nearB = TB != PM_TERM_NONE ? lane_mask(r2 <= P.cutB2hi) & am & ~mB : 0ull;
        const unsigned long long near = nearA | nearB;
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
          if (big_table)
            p = *reinterpret_cast<const PM_GLOBAL float4 *>(gtab12 + info + tj * 16);
          else
            p = *reinterpret_cast<const PM_LDS float4 *>(ljrow + info);
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
            if (PRE) {
              A *= P.sw_a.c2sc;
              B *= P.sw_a.c2sc;
              A12 = 2.0f * A;
              B6 = B;
            } else {
              A12 = 12.0f * A;
              B6 = 6.0f * B;
            }
          }
        }
      }
      float ea, eb;
      const float fs = pair_terms<TA, TB, SWA, SWB, SAME, PRE>(P, r2, qq, A, B, A12, B6, inA, inB, ea, eb);
      // fused multiply-adds, written out: every kind of the kernel rounds the
      // forces alike (the energy kinds' virial products do not change them),
      // so a step that also sums energies moves the atoms as one that does not
      fi[c][0] = fmaf(-fs, dx, fi[c][0]);
      fi[c][1] = fmaf(-fs, dy, fi[c][1]);
      fi[c][2] = fmaf(-fs, dz, fi[c][2]);
      fjx = fmaf(fs, dx, fjx);
      fjy = fmaf(fs, dy, fjy);
      fjz = fmaf(fs, dz, fjz);
      const float fx = fs * dx, fy = fs * dy, fz = fs * dz;
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
    if (any) {
      const float g = dpp_sum8_scatter(fjx, fjy, fjz, ii);
      if (ii < 3 && jj < nvj) atomic_add_ll(&gfc[sj], to_fixed((double)g, kFScale));
    }
  }
  // i force = sum over the 8 j lanes that share ii; lane (jj, ii) then holds
  // atom ii of i cluster jj, which is real where valid_ca has its bit
  {
    float v[3 * kScCl];
#pragma unroll
    for (int c = 0; c < kScCl; ++c) {
      v[3 * c] = fi[c][0];
      v[3 * c + 1] = fi[c][1];
      v[3 * c + 2] = fi[c][2];
    }
    reduce_scatter_i(v, lane);
    if (lane_bit(valid_ca)) {
      const int si = (cl0 + jj) * kCluster + ii;
      atomic_add_ll(&gfx[si], to_fixed((double)v[0], kFScale));
      atomic_add_ll(&gfy[si], to_fixed((double)v[1], kFScale));
      atomic_add_ll(&gfz[si], to_fixed((double)v[2], kFScale));
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

template <int TA, int TB, int SWA, int SWB, int KIND, bool SAME>
__global__ __launch_bounds__(512, (KIND == kPairForces ? PM_PAIR_MIN_WAVES : PM_PAIR_MIN_WAVES_EN)) void k_pair(
    const SysDev *__restrict__ sys, const int *__restrict__ scp, int nsys, ForceDev P) {
  const int w = __builtin_amdgcn_readfirstlane((int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6));
  const int split = P.pair_split;
  const int part = w % split, scg = w / split;
  if (scg >= scp[nsys]) return;
  const int s = __builtin_amdgcn_readfirstlane(find_sys(scp, nsys, scg));
  const SysDev *__restrict__ S = sys + s;
  const int sc = scg - scp[s];
  // the step's bookkeeping, after the rebuild kernels have read it
  if (P.finish_step && sc == 0 && part == 0 && (threadIdx.x & 63) == 0) step_finish_ctl(*S);
  if (sc >= *cp(S->n_sc)) return;
  if (TA != PM_TERM_LJ) {
    pair_wave<TA, TB, SWA, SWB, KIND, SAME, kLjLds>(S, sc, part, split, w, P);
  } else if (S->lj_mode == PM_LJ_PERATOM) {
    pair_wave<TA, TB, SWA, SWB, KIND, SAME, kLjPerAtom>(S, sc, part, split, w, P);
  } else if (S->n_types * S->n_types > kLJLds) {
    pair_wave<TA, TB, SWA, SWB, KIND, SAME, kLjGlobal>(S, sc, part, split, w, P);
  } else {
    pair_wave<TA, TB, SWA, SWB, KIND, SAME, kLjLds>(S, sc, part, split, w, P);
  }
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
  case kPairForcesEnergiesMV:
    hipLaunchKernelGGL((k_pair<TA, TB, SWA, SWB, kPairForcesEnergiesMV, SAME>), g, b, lds, L.stream, L.sys, L.sc_prefix,
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

}  // namespace pm
