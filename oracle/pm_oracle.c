/*
 * pm_oracle.c -- FP64 CPU restatement of ProtoMol's nonbonded force path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path and the CPU baseline
 * timed beside it.  Compiled with -ffp-contract=off so every FP64 expression
 * rounds the way ProtoMol's own x86-64 build does (no fused multiply-add).
 *
 * Reference file:line citations are relative to the ProtoMol source tree.
 */
#define _GNU_SOURCE
#include "pm_oracle.h"

#include <math.h>
#include <pthread.h>
#include <limits.h>
#include <stdlib.h>
#include <string.h>

#define ORC_EPSILON 1.0e-14            /* PMConstants.cpp:50 */
#define ORC_MAX_INT_2 (2147483647 / 2) /* PMConstants.cpp:44 */

/* ------------------------------------------------------------------ hash */

uint64_t orc_pair_hash(uint32_t lo, uint32_t hi) {
  /* splitmix64 finaliser of the packed (lo, hi) pair; order independent sum
   * over a pair set gives a census that is independent of traversal order. */
  uint64_t z = ((uint64_t)hi << 32) | (uint64_t)lo;
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

/* Per-thread accumulators of the threaded walks: one 128-byte-aligned block
 * per thread, so no two threads ever write into one cache line (the energy
 * slots and the census are updated once per pair). */
typedef struct orc_acc {
  double e[ORC_NSLOTS];
  orc_census census;
  double real_energy;
} __attribute__((aligned(128))) orc_acc;

static orc_acc *acc_alloc(int nthreads) {
  orc_acc *a = (orc_acc *)aligned_alloc(128, sizeof(orc_acc) * (size_t)nthreads);
  if (a) memset(a, 0, sizeof(orc_acc) * (size_t)nthreads);
  return a;
}

static inline void census_add(orc_census *c, int a, int b) {
  if (!c) return;
  uint32_t lo = (uint32_t)(a < b ? a : b), hi = (uint32_t)(a < b ? b : a);
  c->count += 1;
  c->hash += orc_pair_hash(lo, hi);
}

/* -------------------------------------------------------------- switches */

/* Cn coefficient tables (CnSwitchingFunction.cpp:9-31); rows = order 2,3,4,5(unused),6 */
static const double orc_swcoef[5][7] = {
    {10., -15., 6., 0, 0, 0, 0},
    {35., -84., 70., -20., 0, 0, 0},
    {126., -420., 540., -315., 70., 0, 0},
    {0, 0, 0, 0, 0, 0, 0},
    {1716., -9009., 20020., -24024., 16380., -6006., 924.0}};
static const double orc_dswcoef[5][7] = {
    {30., -60., 30., 0, 0, 0, 0},
    {140., -420., 420., -140., 0, 0, 0},
    {630., -2520., 3780., -2520., 630., 0, 0},
    {0, 0, 0, 0, 0, 0, 0},
    {12012., -72072., 180180., -240240., 180180., -72072., 12012.0}};

void orc_switch_init(orc_switch *s, int kind, double switchon, double cutoff,
                     double switchoff, double order) {
  memset(s, 0, sizeof(*s));
  s->kind = kind;
  switch (kind) {
  case ORC_SW_UNIVERSAL: /* UniversalSwitchingFunction.h:24-45 */
    s->use = 0; s->modify = 0;
    s->cutoff = 1.79769313486231570e+308; s->cutoff2 = s->cutoff;
    break;
  case ORC_SW_CUTOFF: /* CutoffSwitchingFunction.h:20-44 */
    s->use = 1; s->modify = 0;
    s->cutoff = cutoff; s->cutoff2 = cutoff * cutoff;
    break;
  case ORC_SW_C1: /* C1SwitchingFunction.cpp:12-18 */
    s->use = 1; s->modify = 1;
    s->cutoff = cutoff; s->cutoff2 = cutoff * cutoff;
    s->c1_15rc1 = 1.5 / cutoff;
    s->c1_05rc3 = 0.5 / (cutoff * cutoff * cutoff);
    s->c1_15rc3 = 1.5 / (cutoff * cutoff * cutoff);
    break;
  case ORC_SW_C2: { /* C2SwitchingFunction.cpp:12-18 */
    s->use = 1; s->modify = 1;
    s->switchon = switchon; s->switchon2 = switchon * switchon;
    s->cutoff = cutoff; s->cutoff2 = cutoff * cutoff;
    double d = cutoff * cutoff - switchon * switchon;
    s->s1 = 1.0 / (d * d * d); /* power<3> is d*d*d */
    s->s2 = cutoff * cutoff - 3.0 * switchon * switchon;
    s->s3 = 4.0 / (d * d * d);
    break;
  }
  case ORC_SW_CN: { /* CnSwitchingFunction.cpp:41-60 */
    s->use = 1; s->modify = 1;
    s->switchon = switchon; s->switchon2 = switchon * switchon;
    s->switchoff = switchoff;
    s->cutoff = switchoff != 0.0 ? switchoff : cutoff;
    s->cutoff2 = s->cutoff * s->cutoff;
    s->ord_int = (int)order;
    s->ord_idx = s->ord_int - 2;
    if (s->ord_int == 5) s->ord_int = 2;
    /* the range uses the constructor's cutoff argument (CnSwitchingFunction.cpp:56) */
    s->irange[0] = pow(switchon - cutoff, -(s->ord_int + 1));
    for (int i = 1; i < s->ord_int + 1; i++) s->irange[i] = s->irange[i - 1] / (switchon - cutoff);
    break;
  }
  case ORC_SW_CMPCNCN: { /* CmpCnCnSwitchingFunction.cpp:12-28 */
    s->use = 1; s->modify = 1;
    s->switchon = switchon; s->switchon2 = switchon * switchon;
    s->cutoff = cutoff; s->cutoff2 = cutoff * cutoff;
    s->switchoff = switchoff; s->switchoff2 = switchoff * switchoff;
    double ord = order == 5.0 ? 2.0 : order;
    s->ord_int = (int)ord;
    s->ord_idx = s->ord_int - 2;
    s->irange[0] = 1.0 / pow(switchon - switchoff, s->ord_int + 1);
    s->itrange[0] = 1.0 / pow(switchoff - cutoff, s->ord_int + 1);
    for (int i = 1; i < s->ord_int + 1; i++) {
      s->irange[i] = s->irange[i - 1] / (switchon - switchoff);
      s->itrange[i] = s->itrange[i - 1] / (switchoff - cutoff);
    }
    break;
  }
  }
}

void orc_switch_eval(const orc_switch *s, double r2, double *value, double *deriv) {
  switch (s->kind) {
  case ORC_SW_UNIVERSAL:
    *value = 1.0; *deriv = 0.0;
    return;
  case ORC_SW_CUTOFF: /* CutoffSwitchingFunction.h:41-44 */
    *deriv = 0.0;
    *value = r2 > s->cutoff2 ? 0.0 : 1.0;
    return;
  case ORC_SW_C1: { /* C1SwitchingFunction.h:41-50 */
    if (r2 > s->cutoff2) { *value = 0.0; *deriv = 0.0; return; }
    double d = sqrt(r2);
    *value = 1.0 - d * (s->c1_15rc1 - r2 * s->c1_05rc3);
    *deriv = -s->c1_15rc1 / d + d * s->c1_15rc3;
    return;
  }
  case ORC_SW_C2: /* C2SwitchingFunction.h:42-53 */
    *deriv = 0.0;
    if (r2 > s->cutoff2) *value = 0.0;
    else if (r2 >= s->switchon2) {
      double c2 = s->cutoff2 - r2;
      double c4 = c2 * (s->s2 + 2.0 * r2);
      *value = s->s1 * (c2 * c4);
      *deriv = s->s3 * (c2 * c2 - c4);
    } else *value = 1.0;
    return;
  case ORC_SW_CN: { /* CnSwitchingFunction.cpp:149-175 */
    double c[9], sw, sd;
    *deriv = 0.0; *value = 0.0;
    if (r2 > s->cutoff2) return;
    if (r2 < s->switchon2) { *value = 1.0; return; }
    sd = sqrt(r2);
    sw = sd - s->cutoff;
    c[0] = pow(sw, s->ord_int);
    for (int i = 1; i < s->ord_int + 2; i++) c[i] = c[i - 1] * sw;
    double v = 0.0, dv = 0.0;
    for (int i = 0; i < s->ord_int + 1; i++) {
      v += orc_swcoef[s->ord_idx][i] * c[i + 1] * s->irange[i];
      dv += orc_dswcoef[s->ord_idx][i] * c[i] * s->irange[i];
    }
    *value = v; *deriv = dv / sd;
    return;
  }
  case ORC_SW_CMPCNCN: { /* CmpCnCnSwitchingFunction.h:52-95 */
    double c[9], sw, sd;
    *deriv = 0.0;
    if (r2 > s->cutoff2) { *value = 0.0; return; }
    if (r2 > s->switchoff2) {
      sd = sqrt(r2);
      sw = sd - s->cutoff;
      c[0] = pow(sw, s->ord_int);
      for (int i = 1; i < s->ord_int + 2; i++) c[i] = c[i - 1] * sw;
      double v = 0.0, dv = 0.0;
      for (int i = 0; i < s->ord_int + 1; i++) {
        v += orc_swcoef[s->ord_idx][i] * c[i + 1] * s->itrange[i];
        dv += orc_dswcoef[s->ord_idx][i] * c[i] * s->itrange[i];
      }
      *value = v; *deriv = dv / sd;
      return;
    }
    if (r2 >= s->switchon2) {
      sd = sqrt(r2);
      sw = sd - s->switchoff;
      c[0] = pow(sw, s->ord_int);
      for (int i = 1; i < s->ord_int + 2; i++) c[i] = c[i - 1] * sw;
      double v = 0.0, dv = 0.0;
      for (int i = 0; i < s->ord_int + 1; i++) {
        v += orc_swcoef[s->ord_idx][i] * c[i + 1] * s->irange[i];
        dv -= orc_dswcoef[s->ord_idx][i] * c[i] * s->irange[i];
      }
      *value = 1.0 - v; *deriv = dv / sd;
      return;
    }
    *value = 0.0;
    return;
  }
  }
  *value = 1.0; *deriv = 0.0;
}

/* ------------------------------------------------------------- geometry */

typedef struct orc_pbc {
  double L[3], Lr[3], H[3], H2[3], D, min[3], max[3];
} orc_pbc;

/* PeriodicBoundaryConditions::set (PeriodicBoundaryConditions.cpp:32-77),
 * orthogonal box e1=(Lx,0,0), e2=(0,Ly,0), e3=(0,0,Lz). */
static void pbc_set(orc_pbc *p, const double box[3], const double origin[3]) {
  double Lx = box[0], Ly = box[1], Lz = box[2];
  p->L[0] = Lx; p->L[1] = Ly; p->L[2] = Lz;
  /* myE1r = (e2 x e3) / (e1 . (e2 x e3)) etc. */
  double a1 = Ly * Lz - 0.0 * 0.0, d1 = Lx * a1 + 0.0 * 0.0 + 0.0 * 0.0;
  double a2 = Lz * Lx - 0.0 * 0.0, d2 = 0.0 * 0.0 + Ly * a2 + 0.0 * 0.0;
  double a3 = Lx * Ly - 0.0 * 0.0, d3 = 0.0 * 0.0 + 0.0 * 0.0 + Lz * a3;
  p->Lr[0] = a1 / d1; p->Lr[1] = a2 / d2; p->Lr[2] = a3 / d3;
  for (int k = 0; k < 3; k++) {
    double a = origin[k] - box[k] * 0.5, b = origin[k] + box[k] * 0.5;
    p->min[k] = a < b ? a : b;
    p->max[k] = a > b ? a : b;
    p->H[k] = p->max[k] - p->min[k];
    p->H2[k] = p->H[k] * 0.5;
  }
  double dx = (Lx * 0.5) * (Lx * 0.5), dy = (Ly * 0.5) * (Ly * 0.5), dz = (Lz * 0.5) * (Lz * 0.5);
  p->D = dx < (dy < dz ? dy : dz) ? dx : (dy < dz ? dy : dz);
}

/* PeriodicBoundaryConditions::minimalDifference (PeriodicBoundaryConditions.h:72-110) */
static inline double pbc_min_diff(const orc_pbc *p, const double *c1, const double *c2, double *d) {
  d[0] = c2[0] - c1[0]; d[1] = c2[1] - c1[1]; d[2] = c2[2] - c1[2];
  double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
  if (r2 > p->D) {
    for (int k = 0; k < 3; k++) {
      if (d[k] < -p->H2[k]) d[k] += p->H[k];
      else if (d[k] > p->H2[k]) d[k] -= p->H[k];
    }
    r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    if (r2 > p->D) {
      for (int k = 0; k < 3; k++) d[k] -= p->L[k] * rint(p->Lr[k] * d[k]);
      r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    }
  }
  return r2;
}

/* PeriodicBoundaryConditions::minimalPosition (PeriodicBoundaryConditions.h:113-147) */
static inline void pbc_min_position(const orc_pbc *p, const double *origin, const double *c, double *out) {
  double d[3] = {c[0] - origin[0], c[1] - origin[1], c[2] - origin[2]};
  if (d[0] * d[0] + d[1] * d[1] + d[2] * d[2] <= p->D) { memcpy(out, d, sizeof d); return; }
  for (int k = 0; k < 3; k++) {
    if (d[k] < -p->H2[k]) d[k] += p->H[k];
    else if (d[k] > p->H2[k]) d[k] -= p->H[k];
  }
  if (d[0] * d[0] + d[1] * d[1] + d[2] * d[2] <= p->D) { memcpy(out, d, sizeof d); return; }
  for (int k = 0; k < 3; k++) out[k] = d[k] - p->L[k] * rint(p->Lr[k] * d[k]);
}

double orc_min_diff(const orc_system *sys, const double *c1, const double *c2, double *d) {
  if (!sys->periodic) {
    d[0] = c2[0] - c1[0]; d[1] = c2[1] - c1[1]; d[2] = c2[2] - c1[2];
    return d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
  }
  orc_pbc p;
  pbc_set(&p, sys->box, sys->origin);
  return pbc_min_diff(&p, c1, c2, d);
}

/* ----------------------------------------------------------- exclusions */

/* ExclusionTable (ExclusionTable.{h,cpp}): dense lowDeltas table of width
 * maxDelta+1, last add() of a pair wins; check() swaps to atom1 < atom2. */
typedef struct orc_excl {
  int width;          /* maxDelta + 1 */
  unsigned char *tab; /* n * width */
} orc_excl;

static int excl_build(orc_excl *t, const orc_system *sys) {
  int maxd = 0;
  for (int k = 0; k < sys->nexcl; k++) {
    int a = sys->excl[3 * k], b = sys->excl[3 * k + 1];
    int d = a < b ? b - a : a - b;
    if (d > maxd) maxd = d;
  }
  t->width = maxd + 1;
  t->tab = (unsigned char *)calloc((size_t)sys->n * (size_t)t->width, 1);
  if (!t->tab) return -1;
  for (int k = 0; k < sys->nexcl; k++) {
    int a = sys->excl[3 * k], b = sys->excl[3 * k + 1], c = sys->excl[3 * k + 2];
    if (a > b) { int s = a; a = b; b = s; }
    t->tab[(size_t)a * t->width + (b - a)] = (unsigned char)c;
  }
  return 0;
}

static inline int excl_check(const orc_excl *t, int a, int b) {
  if (b < a) { int s = a; a = b; b = s; }
  int d = b - a;
  if (d >= t->width) return ORC_EXCL_NONE;
  return t->tab[(size_t)a * t->width + d];
}

/* ------------------------------------------------------------ pair terms */

static inline void lj_params(const orc_system *sys, int i, int j, int excl, double *A, double *B) {
  const double *p = sys->lj + 4 * ((size_t)sys->type[i] * sys->ntypes + sys->type[j]);
  if (excl != ORC_EXCL_MODIFIED) { *A = p[0]; *B = p[1]; }
  else { *A = p[2]; *B = p[3]; }
}

/* LennardJonesForce::operator() (LennardJonesForce.h:28-60) */
static inline void term_lj(const orc_system *sys, int i, int j, int excl, double rr2,
                           double *energy, double *force) {
  double A, B;
  lj_params(sys, i, j, excl, &A, &B);
  double r6 = rr2 * rr2 * rr2;
  double r12 = r6 * r6;
  double r6B = B * r6;
  double r12A = A * r12;
  *energy = r12A - r6B;
  *force = 12.0 * r12A * rr2 - 6.0 * r6B * rr2;
}

/* CoulombForce::operator() (CoulombForce.h:29-40) */
static inline void term_coulomb(const orc_system *sys, int i, int j, int excl, double rr2,
                                double *energy, double *force) {
  double e = sys->charge[i] * sys->charge[j] * sqrt(rr2);
  if (excl == ORC_EXCL_MODIFIED) e *= sys->coulomb_scaling;
  *energy = e;
  *force = e * rr2;
}

static inline void add_virial(double *e, const double *f, const double *d) {
  /* ScalarStructure::addVirial (ScalarStructure.cpp:36-54) */
  double xy = f[0] * d[1], xz = f[0] * d[2], yz = f[1] * d[2];
  e[ORC_VIRIALXX] += f[0] * d[0];
  e[ORC_VIRIALXY] += xy;
  e[ORC_VIRIALXZ] += xz;
  e[ORC_VIRIALYX] += xy;
  e[ORC_VIRIALYY] += f[1] * d[1];
  e[ORC_VIRIALYZ] += yz;
  e[ORC_VIRIALZX] += xz;
  e[ORC_VIRIALZY] += yz;
  e[ORC_VIRIALZZ] += f[2] * d[2];
}

static inline void add_molvirial(double *e, const double *f, const double *d) {
  /* ScalarStructure::addMolVirial (ScalarStructure.cpp:56-75) */
  double xy = f[0] * d[1], xz = f[0] * d[2], yz = f[1] * d[2];
  e[ORC_MOLVIRIALXX] += f[0] * d[0];
  e[ORC_MOLVIRIALXY] += xy;
  e[ORC_MOLVIRIALXZ] += xz;
  e[ORC_MOLVIRIALYX] += xy;
  e[ORC_MOLVIRIALYY] += f[1] * d[1];
  e[ORC_MOLVIRIALYZ] += yz;
  e[ORC_MOLVIRIALZX] += xz;
  e[ORC_MOLVIRIALZY] += yz;
  e[ORC_MOLVIRIALZZ] += f[2] * d[2];
}

int orc_n_molecules(const orc_system *sys) {
  int m = 0;
  for (int i = 0; i < sys->n; i++)
    if (sys->molecule[i] + 1 > m) m = sys->molecule[i] + 1;
  return m;
}

void orc_molecule_com(const orc_system *sys, const double *x, double *com) {
  /* molecularCenterOfMass (TopologyUtilities.cpp:645-674) for every molecule
   * at once: the atoms of a molecule are visited in index order, the first
   * one starts the sum, the rest are added with Kahan's compensation */
  int nm = orc_n_molecules(sys);
  double *sm = (double *)calloc((size_t)(nm > 0 ? nm : 1), sizeof(double));
  double *tc = (double *)calloc((size_t)(nm > 0 ? nm : 1) * 3, sizeof(double));
  char *seen = (char *)calloc((size_t)(nm > 0 ? nm : 1), 1);
  for (int k = 0; k < 3 * nm; k++) com[k] = 0.0;
  for (int i = 0; i < sys->n; i++) {
    int m = sys->molecule[i];
    double mass = sys->mass[i];
    double *c = com + 3 * m, *t = tc + 3 * m;
    if (!seen[m]) {
      seen[m] = 1;
      sm[m] = mass;
      for (int k = 0; k < 3; k++) c[k] = x[3 * i + k] * mass;
      continue;
    }
    sm[m] += mass;
    for (int k = 0; k < 3; k++) {
      double tx = x[3 * i + k] * mass;
      double ty = tx - t[k];
      double tt = c[k] + ty;
      t[k] = (tt - c[k]) - ty;
      c[k] = tt;
    }
  }
  for (int m = 0; m < nm; m++)
    if (seen[m])
      for (int k = 0; k < 3; k++) com[3 * m + k] = com[3 * m + k] / sm[m];
  free(sm); free(tc); free(seen);
}

/* minimalDifference of two molecules' centres (the pair terms' comDiff) */
static inline void com_diff(const orc_pbc *p, const double *ci, const double *cj, double *d) {
  if (p) pbc_min_diff(p, ci, cj, d);
  else for (int k = 0; k < 3; k++) d[k] = cj[k] - ci[k];
}

typedef struct pair_ctx {
  const orc_system *sys;
  const orc_nonbonded *nb;
  const orc_pbc *pbc;
  const orc_excl *excl;
  const double *x;
  double *f;
  double *e;
  orc_census *census;
  double cut2;
  const double *com;   /* molecular centres of mass (molecular virial), 3 per molecule */
} pair_ctx;

/* OneAtomPair::doOneAtomPair (OneAtomPair.h:47-107) and
 * OneAtomPairTwo::doOneAtomPair (OneAtomPairTwo.h:30-100). */
static inline void one_atom_pair(pair_ctx *c, int i, int j) {
  const orc_nonbonded *nb = c->nb;
  const orc_system *sys = c->sys;
  double d[3];
  double r2 = c->pbc ? pbc_min_diff(c->pbc, c->x + 3 * i, c->x + 3 * j, d)
                     : (d[0] = c->x[3 * j] - c->x[3 * i], d[1] = c->x[3 * j + 1] - c->x[3 * i + 1],
                        d[2] = c->x[3 * j + 2] - c->x[3 * i + 2], d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  int two = nb->term_b != ORC_TERM_NONE;
  int use = nb->sw_a.use || (two && nb->sw_b.use);
  if (use && r2 > c->cut2) return;
  int same = sys->molecule[i] == sys->molecule[j];
  int ex = same ? excl_check(c->excl, i, j) : ORC_EXCL_NONE;
  if (ex == ORC_EXCL_FULL) return;
  census_add(c->census, i, j);
  double rr2 = 1.0 / r2;
  double e1 = 0, f1 = 0, e2 = 0, f2 = 0;
  if (nb->term_a == ORC_TERM_LJ) term_lj(sys, i, j, ex, rr2, &e1, &f1);
  else term_coulomb(sys, i, j, ex, rr2, &e1, &f1);
  if (two) {
    if (nb->term_b == ORC_TERM_LJ) term_lj(sys, i, j, ex, rr2, &e2, &f2);
    else term_coulomb(sys, i, j, ex, rr2, &e2, &f2);
  }
  if (nb->sw_a.modify || (two && nb->sw_b.modify)) {
    double s, ds;
    orc_switch_eval(&nb->sw_a, r2, &s, &ds);
    f1 = f1 * s - e1 * ds;
    e1 = e1 * s;
    if (two) {
      orc_switch_eval(&nb->sw_b, r2, &s, &ds);
      f2 = f2 * s - e2 * ds;
      e2 = e2 * s;
    }
  }
  c->e[nb->term_a == ORC_TERM_LJ ? ORC_LENNARDJONES : ORC_COULOMB] += e1;
  if (two) c->e[nb->term_b == ORC_TERM_LJ ? ORC_LENNARDJONES : ORC_COULOMB] += e2;
  double fs = two ? f1 + f2 : f1;
  double fij[3] = {d[0] * fs, d[1] * fs, d[2] * fs};
  c->f[3 * i] -= fij[0]; c->f[3 * i + 1] -= fij[1]; c->f[3 * i + 2] -= fij[2];
  c->f[3 * j] += fij[0]; c->f[3 * j + 1] += fij[1]; c->f[3 * j + 2] += fij[2];
  /* OneAtomPair.h:95-103: pairs of two molecules add the atomic and the
   * molecular virial when the molecular one is on, whatever virial() says */
  if (!same && nb->molvirial) {
    double cd[3];
    com_diff(c->pbc, c->com + 3 * sys->molecule[i], c->com + 3 * sys->molecule[j], cd);
    add_virial(c->e, fij, d);
    add_molvirial(c->e, fij, cd);
  } else if (nb->virial) add_virial(c->e, fij, d);
}

/* ------------------------------------------------------------ cell list */

typedef struct orc_cells {
  int nx, ny, nz, ncell;
  int *head;   /* ncell, -1 if empty */
  int *next;   /* n, cellListNext */
  int *ndelta; /* unused */
  int maxnx, maxny, maxnz, nx1, ny1, nz1, maxnx1, maxny1, maxnz1;
  double cs[3], csr[3];
  int (*delta)[3];
  int ndeltas;
} orc_cells;

static void cells_free(orc_cells *cl) {
  free(cl->head); free(cl->next); free(cl->delta);
  memset(cl, 0, sizeof(*cl));
}

/* Topology::updateCellLists (Topology.h:44-86) with CubicCellManager
 * (CubicCellManager.cpp:40-65) and ArrayCellListStructure::initialize
 * (ArrayCellListStructure.cpp:22-128), periodic case. */
static int cells_build(orc_cells *cl, const orc_system *sys, const orc_pbc *p,
                       double cellsize, const double *x) {
  memset(cl, 0, sizeof(*cl));
  double vdiff[3] = {p->max[0] - p->min[0], p->max[1] - p->min[1], p->max[2] - p->min[2]};
  for (int k = 0; k < 3; k++) {
    double nf = floor(vdiff[k] / cellsize);
    cl->cs[k] = vdiff[k] / (1.0 > nf ? 1.0 : nf);
    cl->csr[k] = 1.0 / cl->cs[k];
  }
  int n3[3];
  for (int k = 0; k < 3; k++) {
    int v = (int)floor(vdiff[k] / cl->cs[k] + ORC_EPSILON);
    n3[k] = v > 1 ? v : 1;
  }
  cl->nx = n3[0]; cl->ny = n3[1]; cl->nz = n3[2];
  cl->ncell = cl->nx * cl->ny * cl->nz;
  int max2 = ORC_MAX_INT_2;
  cl->maxnx = max2 - max2 % cl->nx; cl->maxny = max2 - max2 % cl->ny; cl->maxnz = max2 - max2 % cl->nz;
  cl->nx1 = -cl->nx / 2; cl->maxnx1 = cl->maxnx - cl->nx1;
  cl->ny1 = -cl->ny / 2; cl->maxny1 = cl->maxny - cl->ny1;
  cl->nz1 = -cl->nz / 2; cl->maxnz1 = cl->maxnz - cl->nz1;
  cl->head = (int *)malloc(sizeof(int) * (size_t)cl->ncell);
  cl->next = (int *)malloc(sizeof(int) * (size_t)(sys->n > 0 ? sys->n : 1));
  if (!cl->head || !cl->next) return -1;
  for (int c = 0; c < cl->ncell; c++) cl->head[c] = -1;
  double delta[3] = {sys->origin[0] - p->min[0], sys->origin[1] - p->min[1], sys->origin[2] - p->min[2]};
  for (int i = sys->n - 1; i >= 0; i--) {
    double mp[3];
    pbc_min_position(p, sys->origin, x + 3 * i, mp);
    int c3[3];
    for (int k = 0; k < 3; k++) {
      int v = (int)floor((delta[k] + mp[k]) * cl->csr[k]);
      /* an exact upper-face position maps one past the grid in the reference's
       * arithmetic; clamp it into the last cell */
      if (v >= n3[k]) v = n3[k] - 1;
      if (v < 0) v = 0;
      c3[k] = v;
    }
    int cid = (c3[0] * cl->ny + c3[1]) * cl->nz + c3[2];
    cl->next[i] = cl->head[cid];
    cl->head[cid] = i;
  }
  return 0;
}

static int cmp_cell(const void *a, const void *b) {
  const int *u = (const int *)a, *v = (const int *)b;
  for (int k = 0; k < 3; k++) {
    if (u[k] < v[k]) return -1;
    if (u[k] > v[k]) return 1;
  }
  return 0;
}

/* CellListEnumerator<PBC,Cubic>::initialize delta list
 * (CellListEnumeratorPeriodicBoundaries.h:25-86). */
static int cells_deltas(orc_cells *cl, double cutoff) {
  double cutoff2 = cutoff * cutoff;
  int nxx = (int)(cutoff / cl->cs[0] + 1.0 + ORC_EPSILON);
  int nyy = (int)(cutoff / cl->cs[1] + 1.0 + ORC_EPSILON);
  int nzz = (int)(cutoff / cl->cs[2] + 1.0 + ORC_EPSILON);
  int n0 = nxx < cl->nx ? nxx : cl->nx, n1 = nyy < cl->ny ? nyy : cl->ny, n2 = nzz < cl->nz ? nzz : cl->nz;
  double xx = cl->cs[0] * cl->cs[0], yy = cl->cs[1] * cl->cs[1], zz = cl->cs[2] * cl->cs[2];
  int cap = (2 * n0 + 1) * (2 * n1 + 1) * (2 * n2 + 1) + 1;
  cl->delta = (int (*)[3])malloc(sizeof(int[3]) * (size_t)cap);
  if (!cl->delta) return -1;
  int nd = 0;
  for (int k = -n0; k <= n0; k++) {
    int xk = abs(k) - 1; if (xk < 0) xk = 0;
    double d0 = xk * xk * xx;
    for (int l = -n1; l <= n1; l++) {
      int yl = abs(l) - 1; if (yl < 0) yl = 0;
      double d1 = d0 + yl * yl * yy;
      for (int m = -n2; m <= n2; m++) {
        int zm = abs(m) - 1; if (zm < 0) zm = 0;
        if (d1 + zm * zm * zz < cutoff2) {
          /* ArrayCellListStructure::basisCell (ArrayCellListStructure.h:109-113) */
          int b[3] = {(k + cl->maxnx1) % cl->nx + cl->nx1, (l + cl->maxny1) % cl->ny + cl->ny1,
                      (m + cl->maxnz1) % cl->nz + cl->nz1};
          if (b[0] == 0 && b[1] == 0 && b[2] == 0) continue; /* zero is pre-inserted */
          int dup = 0;
          for (int t = 0; t < nd; t++)
            if (cl->delta[t][0] == b[0] && cl->delta[t][1] == b[1] && cl->delta[t][2] == b[2]) { dup = 1; break; }
          if (!dup) { cl->delta[nd][0] = b[0]; cl->delta[nd][1] = b[1]; cl->delta[nd][2] = b[2]; nd++; }
        }
      }
    }
  }
  qsort(cl->delta, (size_t)nd, sizeof(int[3]), cmp_cell);
  cl->ndeltas = nd;
  return 0;
}

/* Enumerate the cell pairs in the order CellListEnumerator::next() yields them
 * (CellListEnumeratorPeriodicBoundaries.h:114-134): for each non-empty cell i
 * in address order, (i,i) first, then each sorted delta whose periodic
 * neighbour j is non-empty and has j > i.  Output pairs (i, j, icell_ordinal). */
typedef struct cell_pair { int i, j, ord; } cell_pair;

static int cells_pairs(const orc_cells *cl, cell_pair **out, int *npairs, int *nnonempty) {
  int cap = 1024, np = 0, ord = 0;
  cell_pair *v = (cell_pair *)malloc(sizeof(cell_pair) * (size_t)cap);
  if (!v) return -1;
  for (int cx = 0; cx < cl->nx; cx++)
    for (int cy = 0; cy < cl->ny; cy++)
      for (int cz = 0; cz < cl->nz; cz++) {
        int ci = (cx * cl->ny + cy) * cl->nz + cz;
        if (cl->head[ci] < 0) continue;
        if (np + cl->ndeltas + 1 > cap) {
          cap = 2 * (np + cl->ndeltas + 1);
          cell_pair *w = (cell_pair *)realloc(v, sizeof(cell_pair) * (size_t)cap);
          if (!w) { free(v); return -1; }
          v = w;
        }
        v[np].i = ci; v[np].j = ci; v[np].ord = ord; np++;
        for (int t = 0; t < cl->ndeltas; t++) {
          /* findPeriodic (ArrayCellListStructure.h:84-103) */
          int jx = (cx + cl->delta[t][0] + cl->maxnx) % cl->nx;
          int jy = (cy + cl->delta[t][1] + cl->maxny) % cl->ny;
          int jz = (cz + cl->delta[t][2] + cl->maxnz) % cl->nz;
          int cj = (jx * cl->ny + jy) * cl->nz + jz;
          if (cl->head[cj] < 0 || cj <= ci) continue;
          v[np].i = ci; v[np].j = cj; v[np].ord = ord; np++;
        }
        ord++;
      }
  *out = v; *npairs = np; *nnonempty = ord;
  return 0;
}

/* ---------------------------------------------------- threaded walking */

typedef void (*pair_fn)(void *ctx, int i, int j, int same_cell);

typedef struct walk_job {
  const orc_cells *cl;
  const cell_pair *pairs;
  int npairs, npackages, nthreads, tid, nnonempty;
  void *ctx;
  void (*cellpair)(void *ctx, const orc_cells *cl, int ci, int cj);
} walk_job;

static void *walk_thread(void *arg) {
  walk_job *w = (walk_job *)arg;
  /* packages = contiguous ranges of i-cells, distributed modulo the thread
   * count like Parallel::next() in STATIC mode (NonbondedCutoffSystemForce.h:57-77) */
  for (int t = 0; t < w->npairs; t++) {
    int ord = w->pairs[t].ord;
    long pkg = ((long)ord * w->npackages) / (w->nnonempty > 0 ? w->nnonempty : 1);
    if (pkg % w->nthreads != w->tid) continue;
    w->cellpair(w->ctx, w->cl, w->pairs[t].i, w->pairs[t].j);
  }
  return NULL;
}

/* NonbondedCutoffForce::doEvaluate inner loops (NonbondedCutoffForce.h:37-54) */
static void nb_cellpair(void *vctx, const orc_cells *cl, int ci, int cj) {
  pair_ctx *c = (pair_ctx *)vctx;
  int same = ci == cj;
  for (int i = cl->head[ci]; i != -1; i = cl->next[i])
    for (int j = same ? cl->next[i] : cl->head[cj]; j != -1; j = cl->next[j]) one_atom_pair(c, i, j);
}

static int run_walk(const orc_cells *cl, const cell_pair *pairs, int np, int nnonempty, int nthreads,
                    void **ctxs, void (*fn)(void *, const orc_cells *, int, int)) {
  if (nthreads <= 1) {
    for (int t = 0; t < np; t++) fn(ctxs[0], cl, pairs[t].i, pairs[t].j);
    return 0;
  }
  pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)nthreads);
  walk_job *jobs = (walk_job *)malloc(sizeof(walk_job) * (size_t)nthreads);
  if (!th || !jobs) { free(th); free(jobs); return -1; }
  int npk = nnonempty < 8 * nthreads ? nnonempty : 8 * nthreads;
  for (int t = 0; t < nthreads; t++) {
    jobs[t].cl = cl; jobs[t].pairs = pairs; jobs[t].npairs = np; jobs[t].npackages = npk;
    jobs[t].nthreads = nthreads; jobs[t].tid = t; jobs[t].nnonempty = nnonempty;
    jobs[t].ctx = ctxs[t]; jobs[t].cellpair = fn;
    pthread_create(&th[t], NULL, walk_thread, &jobs[t]);
  }
  for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
  free(th); free(jobs);
  return 0;
}

/* The vacuum case (VacuumBoundaryConditions: minimalDifference is x_j - x_i,
 * VacuumBoundaryConditions.h): the cell list with CellListEnumeratorStandard
 * yields every pair of atoms in neighbouring cells, and the pair function keeps
 * those within the cutoff, so the pair set is every pair within it.  The cell
 * walk only orders the FP64 sums; here the pairs go in index order. */
static int nonbonded_cutoff_vacuum(const orc_system *sys, const orc_nonbonded *nb, const double *x, double *f,
                                   double *e, orc_census *census, double cut2) {
  orc_excl ex;
  if (excl_build(&ex, sys)) return -1;
  double *com = NULL;
  if (nb->molvirial) {
    com = (double *)calloc((size_t)orc_n_molecules(sys) * 3 + 3, sizeof(double));
    if (!com) { free(ex.tab); return -1; }
    orc_molecule_com(sys, x, com);
  }
  pair_ctx c;
  memset(&c, 0, sizeof c);
  c.sys = sys; c.nb = nb; c.pbc = NULL; c.excl = &ex; c.x = x; c.f = f; c.e = e; c.census = census;
  c.cut2 = cut2; c.com = com;
  for (int i = 0; i < sys->n; i++)
    for (int j = i + 1; j < sys->n; j++) one_atom_pair(&c, i, j);
  free(com); free(ex.tab);
  return 0;
}

/* NonbondedCutoffSystemForce::evaluate (NonbondedCutoffSystemForce.h:45-55) */
int orc_nonbonded_cutoff(const orc_system *sys, const orc_nonbonded *nb, const double *x,
                         double *f, double *e, orc_census *census, int nthreads) {
  if (nb->term_a == ORC_TERM_NONE || nb->term_a == ORC_TERM_EWALD_REAL || nb->term_b == ORC_TERM_EWALD_REAL)
    return -3;
  if (!sys->periodic) {
    double c2 = nb->sw_a.cutoff2;
    if (nb->term_b != ORC_TERM_NONE && nb->sw_b.cutoff2 > c2) c2 = nb->sw_b.cutoff2;
    return nonbonded_cutoff_vacuum(sys, nb, x, f, e, census, c2);
  }
  if (nthreads < 1) nthreads = 1;
  orc_pbc pbc;
  pbc_set(&pbc, sys->box, sys->origin);
  orc_excl ex;
  if (excl_build(&ex, sys)) return -1;
  orc_cells cl;
  if (cells_build(&cl, sys, &pbc, nb->cellsize, x) || cells_deltas(&cl, nb->algorithm_cutoff)) {
    free(ex.tab); cells_free(&cl); return -1;
  }
  cell_pair *pairs = NULL;
  int np = 0, nne = 0;
  if (cells_pairs(&cl, &pairs, &np, &nne)) { free(ex.tab); cells_free(&cl); return -1; }
  /* OneAtomPair: mySquaredCutoff = switch cutoff^2; OneAtomPairTwo: max of both (OneAtomPairTwo.h:25-27) */
  double cut2 = nb->sw_a.cutoff2;
  if (nb->term_b != ORC_TERM_NONE && nb->sw_b.cutoff2 > cut2) cut2 = nb->sw_b.cutoff2;

  pair_ctx *ctx = (pair_ctx *)calloc((size_t)nthreads, sizeof(pair_ctx));
  void **ctxp = (void **)calloc((size_t)nthreads, sizeof(void *));
  double *fbuf = nthreads > 1 ? (double *)calloc((size_t)nthreads * 3 * (size_t)sys->n, sizeof(double)) : NULL;
  orc_acc *acc = acc_alloc(nthreads);
  double *com = NULL;
  int rc = 0;
  if (!ctx || !ctxp || !acc || (nthreads > 1 && !fbuf)) { rc = -1; goto done; }
  if (nb->molvirial) {
    com = (double *)calloc((size_t)orc_n_molecules(sys) * 3 + 3, sizeof(double));
    if (!com) { rc = -1; goto done; }
    orc_molecule_com(sys, x, com);
  }
  for (int t = 0; t < nthreads; t++) {
    ctx[t].com = com;
    ctx[t].sys = sys; ctx[t].nb = nb; ctx[t].pbc = &pbc; ctx[t].excl = &ex; ctx[t].x = x;
    ctx[t].f = nthreads > 1 ? fbuf + (size_t)t * 3 * sys->n : f;
    ctx[t].e = nthreads > 1 ? acc[t].e : e;
    ctx[t].census = census ? &acc[t].census : NULL;
    ctx[t].cut2 = cut2;
    ctxp[t] = &ctx[t];
  }
  rc = run_walk(&cl, pairs, np, nne, nthreads, ctxp, nb_cellpair);
  if (nthreads > 1) {
    for (int t = 0; t < nthreads; t++) {
      const double *ft = fbuf + (size_t)t * 3 * sys->n;
      for (size_t k = 0; k < 3 * (size_t)sys->n; k++) f[k] += ft[k];
      for (int s = 0; s < ORC_NSLOTS; s++) e[s] += acc[t].e[s];
    }
  }
  if (census)
    for (int t = 0; t < nthreads; t++) { census->count += acc[t].census.count; census->hash += acc[t].census.hash; }
done:
  free(ctx); free(ctxp); free(fbuf); free(acc); free(pairs); free(com);
  free(ex.tab); cells_free(&cl);
  return rc;
}

void orc_census_bruteforce(const orc_system *sys, const double *x, double cut2, orc_census *census) {
  orc_pbc pbc;
  if (sys->periodic) pbc_set(&pbc, sys->box, sys->origin);
  orc_excl ex;
  if (excl_build(&ex, sys)) return;
  for (int i = 0; i < sys->n; i++)
    for (int j = i + 1; j < sys->n; j++) {
      double d[3], r2;
      if (sys->periodic) r2 = pbc_min_diff(&pbc, x + 3 * i, x + 3 * j, d);
      else {
        d[0] = x[3 * j] - x[3 * i]; d[1] = x[3 * j + 1] - x[3 * i + 1]; d[2] = x[3 * j + 2] - x[3 * i + 2];
        r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
      }
      if (r2 > cut2) continue;
      if (sys->molecule[i] == sys->molecule[j] && excl_check(&ex, i, j) == ORC_EXCL_FULL) continue;
      census_add(census, i, j);
    }
  free(ex.tab);
}

/* ---------------------------------------------------------------- Ewald */

/* MathUtilities.cpp:39-49 (Abramowitz & Stegun 7.1.26 polynomial) */
double orc_poly5(double ar) {
  const double E1 = 0.254829592, E2 = -0.284496736, E3 = 1.421413741, E4 = -1.453152027,
               E5 = 1.061405429, PP = 0.3275911;
  double t = 1.0 / (1.0 + PP * ar);
  return (t) * (E1 + (t) * (E2 + (t) * (E3 + (t) * (E4 + (t) * E5))));
}

/* MathUtilities.cpp:12-36 (Numerical Recipes erfc, fractional error < 1.2e-7) */
int orc_ewald_init(orc_ewald *ew, double alpha, double accuracy, double cellsize, int correction,
                   int virial) {
  /* NonbondedFullEwaldSystemForce::initialize (NonbondedFullEwaldSystemForce.h:610-619) */
  memset(ew, 0, sizeof(*ew));
  ew->alpha = alpha;
  ew->accuracy = accuracy;
  ew->two_alpha_pi = 2.0 * alpha / sqrt(M_PI);
  double p = -log(accuracy);
  ew->rc = sqrt(p) / alpha;
  ew->rc2 = ew->rc * ew->rc;
  ew->cellsize = cellsize;
  ew->correction = correction;
  ew->virial = virial;
  return 0;
}

typedef struct ewald_ctx {
  const orc_system *sys;
  const orc_ewald *ew;
  const orc_pbc *pbc;
  const orc_excl *excl;
  const double *x;
  double *f;
  double *e;
  double *real_energy;
  orc_census *census;
  const double *lattice; /* nlat*3, lattice[0] = 0 */
  int nlat;
  const double *com;     /* molecular centres of mass (molecular virial) */
} ewald_ctx;

/* NonbondedFullEwaldSystemForce::realTerm inner loops (NonbondedFullEwaldSystemForce.h:787-850) */
static void ewald_cellpair(void *vctx, const orc_cells *cl, int ci, int cj) {
  ewald_ctx *c = (ewald_ctx *)vctx;
  const orc_system *sys = c->sys;
  int notsame = ci != cj;
  for (int i = cl->head[ci]; i != -1; i = cl->next[i]) {
    double qi = sys->charge[i];
    double fi[3] = {0, 0, 0};
    int mi = sys->molecule[i];
    for (int j = notsame ? cl->head[cj] : i; j != -1; j = cl->next[j]) {
      double rmin[3];
      pbc_min_diff(c->pbc, c->x + 3 * i, c->x + 3 * j, rmin);
      int mj = sys->molecule[j];
      double qj = sys->charge[j];
      int same = mi == mj;
      int ex = same ? excl_check(c->excl, i, j) : ORC_EXCL_NONE;
      if (i == j) ex = ORC_EXCL_FULL;
      for (int k = 0; k < c->nlat; k++) {
        if (ex == ORC_EXCL_FULL) { ex = ORC_EXCL_NONE; continue; }
        double rij[3] = {rmin[0] + c->lattice[3 * k], rmin[1] + c->lattice[3 * k + 1],
                         rmin[2] + c->lattice[3 * k + 2]};
        double r2 = rij[0] * rij[0] + rij[1] * rij[1] + rij[2] * rij[2];
        if (r2 > c->ew->rc2) continue;
        if (k == 0) census_add(c->census, i, j);
        double r = sqrt(r2);
        double qq = qi * qj;
        if (ex == ORC_EXCL_MODIFIED) qq *= sys->coulomb_scaling;
        double rr = 1.0 / r;
        double ar = c->ew->alpha * r;
        double ee = qq * exp(-ar * ar);
        double energy = orc_poly5(ar) * ee * rr;
        double force = ((energy + c->ew->two_alpha_pi * ee) * rr * rr);
        /* CutoffSwitchingFunction: value 1 inside Rc, derivative 0 */
        double sv = r2 > c->ew->rc2 ? 0.0 : 1.0;
        force = force * sv - energy * 0.0;
        energy = energy * sv;
        *c->real_energy += energy;
        double fij[3] = {rij[0] * force, rij[1] * force, rij[2] * force};
        fi[0] -= fij[0]; fi[1] -= fij[1]; fi[2] -= fij[2];
        c->f[3 * j] += fij[0]; c->f[3 * j + 1] += fij[1]; c->f[3 * j + 2] += fij[2];
        /* :839-848 */
        if (!same && c->ew->molvirial) {
          double cd[3];
          com_diff(c->pbc, c->com + 3 * mi, c->com + 3 * mj, cd);
          add_virial(c->e, fij, rij);
          add_molvirial(c->e, fij, cd);
        } else if (c->ew->virial) add_virial(c->e, fij, rij);
        ex = ORC_EXCL_NONE;
      }
    }
    c->f[3 * i] += fi[0]; c->f[3 * i + 1] += fi[1]; c->f[3 * i + 2] += fi[2];
  }
}

/* first[k] = 1 when exclusion record k is the first of its (unordered pair,
 * class) in table order: the correction term walks each such pair once. */
typedef struct excl_key { int a, b, cls, k; } excl_key;
static int cmp_excl_key(const void *x, const void *y) {
  const excl_key *u = (const excl_key *)x, *v = (const excl_key *)y;
  if (u->a != v->a) return u->a < v->a ? -1 : 1;
  if (u->b != v->b) return u->b < v->b ? -1 : 1;
  if (u->cls != v->cls) return u->cls < v->cls ? -1 : 1;
  return u->k < v->k ? -1 : (u->k > v->k ? 1 : 0);
}
static char *excl_first_occurrence(const orc_system *sys) {
  int n = sys->nexcl;
  char *first = (char *)calloc((size_t)(n > 0 ? n : 1), 1);
  excl_key *keys = (excl_key *)malloc(sizeof(excl_key) * (size_t)(n > 0 ? n : 1));
  if (!first || !keys) { free(first); free(keys); return NULL; }
  for (int k = 0; k < n; k++) {
    int a = sys->excl[3 * k], b = sys->excl[3 * k + 1];
    keys[k].a = a < b ? a : b;
    keys[k].b = a < b ? b : a;
    keys[k].cls = sys->excl[3 * k + 2];
    keys[k].k = k;
  }
  qsort(keys, (size_t)n, sizeof(excl_key), cmp_excl_key);
  for (int t = 0; t < n; t++)
    if (t == 0 || keys[t].a != keys[t - 1].a || keys[t].b != keys[t - 1].b || keys[t].cls != keys[t - 1].cls)
      first[keys[t].k] = 1;
  free(keys);
  return first;
}

static int cmp_norm(const void *a, const void *b) {
  const double *u = (const double *)a, *v = (const double *)b;
  double nu = u[0] * u[0] + u[1] * u[1] + u[2] * u[2], nv = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
  return nu < nv ? -1 : (nu > nv ? 1 : 0);
}

/* NonbondedFullEwaldSystemForce::evaluate (NonbondedFullEwaldSystemForce.h:329-398),
 * real + correction + point-self + charged-system terms, reciprocal off. */
int orc_ewald_real(const orc_system *sys, const orc_ewald *ew, const double *x, double *f, double *e,
                   orc_census *census, int nthreads) {
  if (!sys->periodic) return -2;
  if (nthreads < 1) nthreads = 1;
  orc_pbc pbc;
  pbc_set(&pbc, sys->box, sys->origin);
  orc_excl ex;
  if (excl_build(&ex, sys)) return -1;

  /* PeriodicBoundaryConditions::buildLatticeVectors (PeriodicBoundaryConditions.cpp:80-102) */
  int bk = (int)floor(ew->rc / fabs(sys->box[0]) + 1.0), bl = (int)floor(ew->rc / fabs(sys->box[1]) + 1.0),
      bm = (int)floor(ew->rc / fabs(sys->box[2]) + 1.0);
  int cap = (2 * bk + 1) * (2 * bl + 1) * (2 * bm + 1) + 1;
  double *lat = (double *)calloc((size_t)cap * 3, sizeof(double));
  int nlat = 1;
  for (int k = -bk; k <= bk; k++)
    for (int l = -bl; l <= bl; l++)
      for (int m = -bm; m <= bm; m++)
        if (k != 0 || l != 0 || m != 0) {
          double v[3] = {sys->box[0] * k, sys->box[1] * l, sys->box[2] * m};
          if (sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) <= 2 * ew->rc) {
            memcpy(lat + 3 * nlat, v, sizeof v);
            nlat++;
          }
        }
  qsort(lat + 3, (size_t)(nlat - 1), sizeof(double[3]), cmp_norm);

  double intra = 0.0, point = 0.0, charged = 0.0;
  if (ew->correction) {
    /* correctionTerm (NonbondedFullEwaldSystemForce.h:1097-1131) over the
     * ExclusionTable's pair table in add() order, duplicates dropped.  Off
     * Windows, MathUtilities.h:154-161 declares no ProtoMol::erf, so the
     * unqualified erf() there is the C library's. */
    char *first = excl_first_occurrence(sys);
    if (!first && sys->nexcl > 0) { free(ex.tab); free(lat); return -1; }
    for (int k = 0; k < sys->nexcl; k++) {
      int a1 = sys->excl[3 * k], a2 = sys->excl[3 * k + 1], cls = sys->excl[3 * k + 2];
      if (cls == ORC_EXCL_NONE || a1 == a2) continue;
      if (a1 > a2) { int s = a1; a1 = a2; a2 = s; }
      if (!first[k]) continue;
      double rij[3];
      double r2 = pbc_min_diff(&pbc, x + 3 * a1, x + 3 * a2, rij);
      double qq = sys->charge[a1] * sys->charge[a2];
      if (cls == ORC_EXCL_MODIFIED) qq *= 1 - sys->coulomb_scaling;
      double r = sqrt(r2);
      double rr = 1 / r;
      double ee = erf(ew->alpha * r) * rr;
      intra -= qq * ee;
      double fs = qq * (ew->two_alpha_pi * exp(-(ew->alpha * ew->alpha) * r2) - ee) * rr * rr;
      double fij[3] = {rij[0] * fs, rij[1] * fs, rij[2] * fs};
      f[3 * a1] -= fij[0]; f[3 * a1 + 1] -= fij[1]; f[3 * a1 + 2] -= fij[2];
      f[3 * a2] += fij[0]; f[3 * a2 + 1] += fij[1]; f[3 * a2 + 2] += fij[2];
      if (ew->virial) add_virial(e, fij, rij);
    }
    free(first);
    double q2 = 0.0, q = 0.0;
    for (int i = 0; i < sys->n; i++) q2 += sys->charge[i] * sys->charge[i];
    point = -q2 * ew->alpha / sqrt(M_PI);
    for (int i = 0; i < sys->n; i++) q += sys->charge[i];
    double V = fabs(sys->box[0] * sys->box[1] * sys->box[2]);
    if (fabs(q * 0.00268283) > 1.0e-5) charged = -M_PI / (2.0 * V * (ew->alpha * ew->alpha)) * q * q;
    if (ew->virial) { e[ORC_VIRIALXX] += charged; e[ORC_VIRIALYY] += charged; e[ORC_VIRIALZZ] += charged; }
  }

  orc_cells cl;
  cell_pair *pairs = NULL;
  int np = 0, nne = 0, rc = 0;
  if (cells_build(&cl, sys, &pbc, ew->cellsize, x) || cells_deltas(&cl, ew->rc) ||
      cells_pairs(&cl, &pairs, &np, &nne)) {
    free(ex.tab); free(lat); cells_free(&cl); free(pairs); return -1;
  }
  ewald_ctx *ctx = (ewald_ctx *)calloc((size_t)nthreads, sizeof(ewald_ctx));
  void **ctxp = (void **)calloc((size_t)nthreads, sizeof(void *));
  double *fbuf = nthreads > 1 ? (double *)calloc((size_t)nthreads * 3 * (size_t)sys->n, sizeof(double)) : NULL;
  orc_acc *acc = acc_alloc(nthreads);
  double *com = NULL;
  if (!ctx || !ctxp || !acc || (nthreads > 1 && !fbuf)) { rc = -1; goto done; }
  if (ew->molvirial) {
    com = (double *)calloc((size_t)orc_n_molecules(sys) * 3 + 3, sizeof(double));
    if (!com) { rc = -1; goto done; }
    orc_molecule_com(sys, x, com);
  }
  for (int t = 0; t < nthreads; t++) {
    ctx[t].com = com;
    ctx[t].sys = sys; ctx[t].ew = ew; ctx[t].pbc = &pbc; ctx[t].excl = &ex; ctx[t].x = x;
    ctx[t].f = nthreads > 1 ? fbuf + (size_t)t * 3 * sys->n : f;
    ctx[t].e = nthreads > 1 ? acc[t].e : e;
    ctx[t].census = census ? &acc[t].census : NULL;
    ctx[t].real_energy = &acc[t].real_energy;
    ctx[t].lattice = lat; ctx[t].nlat = nlat;
    ctxp[t] = &ctx[t];
  }
  rc = run_walk(&cl, pairs, np, nne, nthreads, ctxp, ewald_cellpair);
  double real = 0.0;
  for (int t = 0; t < nthreads; t++) real += acc[t].real_energy;
  if (nthreads > 1) {
    for (int t = 0; t < nthreads; t++) {
      const double *ft = fbuf + (size_t)t * 3 * sys->n;
      for (size_t k = 0; k < 3 * (size_t)sys->n; k++) f[k] += ft[k];
      for (int s = 0; s < ORC_NSLOTS; s++) e[s] += acc[t].e[s];
    }
  }
  if (census)
    for (int t = 0; t < nthreads; t++) { census->count += acc[t].census.count; census->hash += acc[t].census.hash; }
  e[ORC_COULOMB] += real + 0.0 + intra + 0.0 + point + charged;
done:
  free(ctx); free(ctxp); free(fbuf); free(acc); free(pairs); free(lat); free(com);
  free(ex.tab); cells_free(&cl);
  return rc;
}

/* ------------------------------------------------ Ewald reciprocal space */

/* The k-vector list of NonbondedFullEwaldSystemForce::initialize
 * (NonbondedFullEwaldSystemForce.h:604-667): h >= 0 half space, |k| < kc,
 * in (h, k, l) loop order.  Returns the count (or -1), fills *out (malloc'd,
 * 4 ints h, k, l, 0 per vector) and hkl[4] = {hmax, kmax, lmax, hklmax}. */
int orc_ewald_kvectors(const double box[3], double alpha, double accuracy, int **out, int hkl[4]) {
  orc_pbc p;
  double org[3] = {0, 0, 0};
  pbc_set(&p, box, org);
  double LX = box[0], LY = box[1], LZ = box[2];
  double pp = -log(accuracy);
  double kc = 2.0 * alpha * sqrt(pp);
  int hmax = (int)floor(kc / (2.0 * M_PI) * LX);
  int kmax = (int)floor(kc / (2.0 * M_PI) * LY);
  int lmax = (int)floor(kc / (2.0 * M_PI) * LZ);
  int m = kmax > lmax ? kmax : lmax;
  m = hmax > m ? hmax : m;
  int hklmax = 2 > m + 1 ? 2 : m + 1;
  hkl[0] = hmax; hkl[1] = kmax; hkl[2] = lmax; hkl[3] = hklmax;
  double kc2 = kc * kc;
  size_t cap = (size_t)(hmax + 1) * (2 * kmax + 1) * (2 * lmax + 1) + 1;
  int *v = (int *)malloc(cap * 4 * sizeof(int));
  if (!v) return -1;
  int nk = 0;
  for (int h = 0; h <= hmax; h++) {
    double kx = 2.0 * M_PI * h * p.Lr[0];
    for (int k = (h == 0 ? 0 : -kmax); k <= kmax; k++) {
      double ky = 2.0 * M_PI * k * p.Lr[1];
      for (int l = (h == 0 && k == 0 ? 1 : -lmax); l <= lmax; l++) {
        double kz = 2.0 * M_PI * l * p.Lr[2];
        if (kx * kx + ky * ky + kz * kz < kc2) {
          v[4 * nk] = h; v[4 * nk + 1] = k; v[4 * nk + 2] = l; v[4 * nk + 3] = 0;
          nk++;
        }
      }
    }
  }
  *out = v;
  return nk;
}

typedef struct recip_job {
  const orc_system *sys;
  const orc_ewald *ew;
  const orc_pbc *pbc;
  const int *kv;
  const double *sincos;  /* per atom 2 * hklmax entries of (x, y, z) */
  int hklmax, from, to;
  double *f;             /* 3n, this job's forces */
  double energy, vir[6];
} recip_job;

/* NonbondedFullEwaldSystemForce::reciprocalTerm k loop (NonbondedFullEwaldSystemForce.h:947-1051)
 * over k-vectors [from, to): the (h, k) products are cached while h and k
 * repeat, as the reference's myLastSinCos is. */
static void *recip_thread(void *arg) {
  recip_job *J = (recip_job *)arg;
  const int n = J->sys->n, hm = J->hklmax;
  const double *SC = J->sincos;
  const orc_pbc *p = J->pbc;
  double *last = (double *)malloc(2 * (size_t)n * sizeof(double));
  double *sca = (double *)malloc(2 * (size_t)n * sizeof(double));
  double alpha2r = 1.0 / (J->ew->alpha * J->ew->alpha);
  double fac = 1.0 / (4.0 * J->ew->alpha * J->ew->alpha);
  double V = p->L[0] * p->L[1] * p->L[2];
  double Vr = 1.0 / V;
  int lastH = INT_MAX, lastK = INT_MAX;
  J->energy = 0.0;
  memset(J->vir, 0, sizeof J->vir);
  for (int q = J->from; q < J->to; q++) {
    int ih = J->kv[4 * q], ik = J->kv[4 * q + 1], il = J->kv[4 * q + 2];
    double k0 = 2.0 * M_PI * ih * p->Lr[0], k1 = 2.0 * M_PI * ik * p->Lr[1], k2 = 2.0 * M_PI * il * p->Lr[2];
    double ksq = k0 * k0 + k1 * k1 + k2 * k2;
    int ika = abs(ik), iks = ik < 0 ? -1 : 1, ila = abs(il), ils = il < 0 ? -1 : 1;
    if (ih != lastH || ik != lastK) {
      for (int i = 0; i < n; i++) {
        const double *t = SC + (size_t)i * hm * 2 * 3;
        double xsin = t[(ih * 2) * 3 + 0], xcos = t[(ih * 2 + 1) * 3 + 0];
        double ysin = iks * t[(ika * 2) * 3 + 1], ycos = t[(ika * 2 + 1) * 3 + 1];
        last[2 * i] = xsin * ycos + xcos * ysin;
        last[2 * i + 1] = xcos * ycos - xsin * ysin;
      }
      lastH = ih;
      lastK = ik;
    }
    double sumSin = 0.0, sumCos = 0.0;
    for (int i = 0; i < n; i++) {
      const double *t = SC + (size_t)i * hm * 2 * 3;
      double xysin = last[2 * i], xycos = last[2 * i + 1];
      double zsin = ils * t[(ila * 2) * 3 + 2], zcos = t[(ila * 2 + 1) * 3 + 2];
      double sinA = xysin * zcos + xycos * zsin;
      double cosA = xycos * zcos - xysin * zsin;
      sca[2 * i] = sinA;
      sca[2 * i + 1] = cosA;
      sumSin += sinA;
      sumCos += cosA;
    }
    double b = 1.0 / ksq * exp(-ksq * alpha2r / 4.0);
    double e = b * (sumSin * sumSin + sumCos * sumCos);
    J->energy += e;
    if (J->ew->virial) {
      double c = 2.0 * (1.0 / ksq + fac);
      J->vir[0] += e * (1.0 - c * k0 * k0);
      J->vir[1] -= e * c * k0 * k1;
      J->vir[2] -= e * c * k0 * k2;
      J->vir[3] += e * (1.0 - c * k1 * k1);
      J->vir[4] -= e * c * k1 * k2;
      J->vir[5] += e * (1.0 - c * k2 * k2);
    }
    double a = 8.0 * M_PI * Vr * b;
    for (int i = 0; i < n; i++) {
      double s = a * (sca[2 * i] * sumCos - sca[2 * i + 1] * sumSin);
      J->f[3 * i] += k0 * s;
      J->f[3 * i + 1] += k1 * s;
      J->f[3 * i + 2] += k2 * s;
    }
  }
  free(last);
  free(sca);
  return NULL;
}

/* NonbondedFullEwaldSystemForce<PBC,Cubic,false,true,false,Cutoff>::evaluate:
 * the reciprocal-space term alone (NonbondedFullEwaldSystemForce.h:873-1083).
 * Adds forces into f, c * energy into e[COULOMB] and the virial when
 * ew->virial.  nthreads > 1 splits the k-vector range as parallelEvaluate
 * splits it across ranks (:473-481). */
int orc_ewald_reciprocal(const orc_system *sys, const orc_ewald *ew, const double *x, double *f, double *e,
                         int nthreads) {
  if (!sys->periodic) return -2;
  if (nthreads < 1) nthreads = 1;
  orc_pbc pbc;
  pbc_set(&pbc, sys->box, sys->origin);
  int *kv = NULL, hkl[4];
  int nk = orc_ewald_kvectors(sys->box, ew->alpha, ew->accuracy, &kv, hkl);
  if (nk < 0) return -1;
  const int n = sys->n, hm = hkl[3];
  double *SC = (double *)malloc((size_t)n * hm * 2 * 3 * sizeof(double));
  if (!SC) { free(kv); return -1; }
  /* sin/cos tables by the addition theorem (:900-945); x carries the charge */
  for (int j = 0; j < n; j++) {
    double r[3];
    pbc_min_position(&pbc, sys->origin, x + 3 * j, r);
    double qi = sys->charge[j];
    double ax = r[0] * 2.0 * M_PI * pbc.Lr[0], ay = r[1] * 2.0 * M_PI * pbc.Lr[1], az = r[2] * 2.0 * M_PI * pbc.Lr[2];
    double xs = sin(ax), ys = sin(ay), zs = sin(az), xc = cos(ax), yc = cos(ay), zc = cos(az);
    double *t = SC + (size_t)j * hm * 2 * 3;
    t[0] = 0.0; t[1] = 0.0; t[2] = 0.0;
    t[3] = qi * 1.0; t[4] = 1.0; t[5] = 1.0;
    t[6] = qi * xs; t[7] = ys; t[8] = zs;
    t[9] = qi * xc; t[10] = yc; t[11] = zc;
    for (int i = 4; i < 2 * hm; i += 2) {
      double *o = t + i * 3, *p1 = t + (i - 1) * 3, *p2 = t + (i - 2) * 3;
      o[0] = xs * p1[0] + xc * p2[0];
      o[1] = ys * p1[1] + yc * p2[1];
      o[2] = zs * p1[2] + zc * p2[2];
      o[3] = xc * p1[0] - xs * p2[0];
      o[4] = yc * p1[1] - ys * p2[1];
      o[5] = zc * p1[2] - zs * p2[2];
    }
  }
  recip_job *jobs = (recip_job *)calloc((size_t)nthreads, sizeof(recip_job));
  pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)nthreads);
  double *fb = (double *)calloc((size_t)nthreads * 3 * (size_t)n, sizeof(double));
  if (!jobs || !th || !fb) { free(jobs); free(th); free(fb); free(SC); free(kv); return -1; }
  for (int t = 0; t < nthreads; t++) {
    jobs[t].sys = sys; jobs[t].ew = ew; jobs[t].pbc = &pbc; jobs[t].kv = kv; jobs[t].sincos = SC;
    jobs[t].hklmax = hm;
    jobs[t].from = (int)((long long)nk * t / nthreads);
    jobs[t].to = (int)((long long)nk * (t + 1) / nthreads);
    jobs[t].f = fb + (size_t)t * 3 * n;
  }
  if (nthreads == 1) recip_thread(&jobs[0]);
  else {
    for (int t = 0; t < nthreads; t++) pthread_create(&th[t], NULL, recip_thread, &jobs[t]);
    for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
  }
  double c = 4.0 * M_PI * (1.0 / (pbc.L[0] * pbc.L[1] * pbc.L[2]));
  double recip = 0.0;
  for (int t = 0; t < nthreads; t++) {
    recip += c * jobs[t].energy;
    for (size_t k = 0; k < 3 * (size_t)n; k++) f[k] += jobs[t].f[k];
    /* :1055-1078: the tensor of the k sum goes to the atomic virial and, with
     * the molecular virial on, to that one too (it is summed only when the
     * atomic virial is on, :1017) */
    const double *v = jobs[t].vir;
    if (ew->virial) {
      e[ORC_VIRIALXX] += c * v[0]; e[ORC_VIRIALXY] += c * v[1]; e[ORC_VIRIALXZ] += c * v[2];
      e[ORC_VIRIALYX] += c * v[1]; e[ORC_VIRIALYY] += c * v[3]; e[ORC_VIRIALYZ] += c * v[4];
      e[ORC_VIRIALZX] += c * v[2]; e[ORC_VIRIALZY] += c * v[4]; e[ORC_VIRIALZZ] += c * v[5];
    }
    if (ew->molvirial && ew->virial) {
      e[ORC_MOLVIRIALXX] += c * v[0]; e[ORC_MOLVIRIALXY] += c * v[1]; e[ORC_MOLVIRIALXZ] += c * v[2];
      e[ORC_MOLVIRIALYX] += c * v[1]; e[ORC_MOLVIRIALYY] += c * v[3]; e[ORC_MOLVIRIALYZ] += c * v[4];
      e[ORC_MOLVIRIALZX] += c * v[2]; e[ORC_MOLVIRIALZY] += c * v[4]; e[ORC_MOLVIRIALZZ] += c * v[5];
    }
  }
  if (ew->molvirial) {
    /* :1039-1049: each atom's reciprocal force times the vector from its
     * minimal position to its molecule's centre; the reference adds it per k
     * vector, here once with the atom's force summed over k */
    double *com = (double *)calloc((size_t)orc_n_molecules(sys) * 3 + 3, sizeof(double));
    double *fr = (double *)calloc(3 * (size_t)n + 3, sizeof(double));
    if (!com || !fr) { free(com); free(fr); free(jobs); free(th); free(fb); free(SC); free(kv); return -1; }
    orc_molecule_com(sys, x, com);
    for (int t = 0; t < nthreads; t++)
      for (size_t k = 0; k < 3 * (size_t)n; k++) fr[k] += jobs[t].f[k];
    for (int i = 0; i < n; i++) {
      double ria[3], mri[3];
      pbc_min_position(&pbc, sys->origin, x + 3 * i, ria);
      pbc_min_diff(&pbc, ria, com + 3 * sys->molecule[i], mri);
      add_molvirial(e, fr + 3 * i, mri);
    }
    free(com); free(fr);
  }
  e[ORC_COULOMB] += recip;
  free(jobs); free(th); free(fb); free(SC); free(kv);
  return 0;
}

/* ------------------------------------------------------------------- GB */

/* The all-pairs order of NonbondedSimpleFullSystemForce::doEvaluate
 * (NonbondedSimpleFullSystemForce.h:125-150): blocks of block_size atoms,
 * i-block <= j-block, i < j. */
#define GB_FOR_PAIRS(n, bs, ...)                                                  \
  for (int bi = 0; bi < (n); bi += (bs))                                          \
    for (int bj = bi; bj < (n); bj += (bs)) {                                     \
      int ie = bi + (bs) < (n) ? bi + (bs) : (n);                                 \
      for (int i = bi; i < ie; i++) {                                             \
        int js = bj <= i ? i + 1 : bj;                                            \
        int je = bj + (bs) < (n) ? bj + (bs) : (n);                               \
        for (int j = js; j < je; j++) { __VA_ARGS__ }                                  \
      }                                                                           \
    }

/* Burial term contributions of atom j to atom i (GBBornRadii.cpp:59-91) */
static inline double gb_burial(double dist, double ori, double orj, double sj) {
  double L, U, C;
  if (ori >= dist + sj * orj) { L = 1.0; U = 1.0; }
  else {
    double a = fabs(dist - sj * orj);
    L = ori > a ? ori : a;
    U = dist + sj * orj;
  }
  if (ori < orj * sj - dist) C = 2.0 * (1.0 / ori - 1.0 / L);
  else C = 0;
  double iL = 1.0 / L, iU = 1.0 / U, iL2 = iL * iL, iU2 = iU * iU;
  double ratio = log(L / U);
  return (iL - iU) + (dist / 4.0) * (iU2 - iL2) + (1.0 / (2.0 * dist)) * ratio +
         ((sj * sj * orj * orj) / (4.0 * dist)) * (iL2 - iU2) + C;
}

/* d(burial_i)/d(r_ij) and the dR_i/dr_ij chain term (GBForce.cpp:55-127) */
static inline double gb_dR(double dist, double ri, double ori, double orj, double sj, double born,
                           double psi, double a, double b, double g) {
  double L, U;
  if (ori >= dist + sj * orj) { L = 1; U = 1; }
  else {
    double t = fabs(dist - sj * orj);
    L = ori > t ? ori : t;
    U = dist + sj * orj;
  }
  double iL = 1 / L;
  double dL = ori <= (dist - sj * orj) ? 1 : 0;
  double dU = ori < (dist + sj * orj) ? 1 : 0;
  double dC = ori <= (sj * orj - dist) ? 2 * iL * iL * dL : 0;
  double tanhp = a * psi - b * psi * psi + g * psi * psi * psi;
  double th = tanh(tanhp);
  double tderv = (a - 2 * b * psi + 3 * g * psi * psi);
  double st = (sj * orj) / dist;
  double dBT = -0.5 * dL * (1 / (L * L)) + 0.5 * dU * (1 / (U * U)) + 0.125 * ((1 / (U * U)) - (1 / (L * L))) +
               0.125 * dist * ((2 / (L * L * L)) * dL - (2 / (U * U * U)) * dU) -
               0.25 * (1 / (dist * dist)) * log(L / U) + (U / (4 * dist * L)) * ((1 / U) * dL - (L / (U * U)) * dU) -
               0.125 * (st * st) * ((1 / (L * L)) - (1 / (U * U))) +
               0.25 * ((sj * sj * orj * orj) / (dist * U * U * U)) * dU -
               0.25 * ((sj * sj * orj * orj) / (dist * L * L * L)) * dL + dC;
  return (born * born) * ori * (1 - th * th) * tderv * (1 / ri) * dBT;
}

int orc_gb_evaluate(const orc_system *sys, const orc_gb *gb, const double *x, double *f, double *e,
                    double *born_out) {
  int n = sys->n, bs = gb->block_size > 0 ? gb->block_size : 64;
  double *burial = (double *)calloc((size_t)n, sizeof(double));
  double *born = (double *)calloc((size_t)n, sizeof(double));
  double *psi = (double *)calloc((size_t)n, sizeof(double));
  double *partial = (double *)calloc((size_t)n, sizeof(double));
  double *self = (double *)calloc((size_t)n, sizeof(double));
  if (!burial || !born || !psi || !partial || !self) {
    free(burial); free(born); free(psi); free(partial); free(self); return -1;
  }
  const double *rho = gb->vdw_radius, *S = gb->scaling;
  double *com = NULL;
  if (gb->molvirial) {
    com = (double *)calloc((size_t)orc_n_molecules(sys) * 3 + 3, sizeof(double));
    if (!com) { free(burial); free(born); free(psi); free(partial); free(self); return -1; }
    orc_molecule_com(sys, x, com);
  }
  /* the virials of a pair, as OneAtomPairNoExclusion.h:73-81 adds them (the
   * Born-radius and partial-sum passes carry no force, so add nothing) */
#define GB_PAIR_VIRIAL(fij, d)                                                     \
  if (gb->molvirial && sys->molecule[i] != sys->molecule[j]) {                     \
    double cd[3];                                                                  \
    com_diff(NULL, com + 3 * sys->molecule[i], com + 3 * sys->molecule[j], cd);   \
    add_virial(e, fij, d);                                                         \
    add_molvirial(e, fij, cd);                                                     \
  } else if (gb->virial) add_virial(e, fij, d);
  /* pass 1: GBBornRadii::operator() (GBBornRadii.cpp:15-128) */
  GB_FOR_PAIRS(n, bs, {
    double d[3] = {x[3 * j] - x[3 * i], x[3 * j + 1] - x[3 * i + 1], x[3 * j + 2] - x[3 * i + 2]};
    double dist = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    double ori = rho[i] - gb->offset, orj = rho[j] - gb->offset;
    burial[i] += gb_burial(dist, ori, orj, S[j]);
    burial[j] += gb_burial(dist, orj, ori, S[i]);
  })
  /* GBBornRadii::postProcess (GBBornRadii.cpp:145-173) */
  for (int i = 0; i < n; i++) {
    double ori = rho[i] - gb->offset;
    psi[i] = 0.5 * burial[i] * ori;
    double p = psi[i];
    double tp = gb->alpha_obc * p - gb->beta_obc * p * p + gb->gamma_obc * p * p * p;
    double inv = (1 / ori) - (1 / rho[i]) * tanh(tp);
    born[i] = 1 / inv;
  }
  /* pass 2: GBPartialSum::operator() (GBPartialSum.cpp:11-42) */
  GB_FOR_PAIRS(n, bs, {
    double d[3] = {x[3 * j] - x[3 * i], x[3 * j + 1] - x[3 * i + 1], x[3 * j + 2] - x[3 * i + 2]};
    double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    double ril = sqrt(r2);
    double bi_ = born[i], bl = born[j];
    double expterm = exp(-(ril * ril) / (4.0 * bi_ * bl));
    double fil = sqrt(ril * ril + bi_ * bl * expterm);
    double part = sys->charge[i] * sys->charge[j] * (1 / (fil * fil)) * 0.5 * (1 / fil) * expterm;
    partial[i] += part * (bl + (ril * ril) / (4.0 * bi_));
    partial[j] += part * (bi_ + (ril * ril) / (4.0 * bl));
  })
  /* pass 3: GBForce::operator() (GBForce.cpp:14-175) */
  double fac = (1 / gb->solute_dielec) - (1 / gb->solvent_dielec);
  for (int i = 0; i < n; i++)
    self[i] = -0.5 * (sys->charge[i] * sys->charge[i]) * (1 / born[i]) * fac;
  double egb = 0.0;
  GB_FOR_PAIRS(n, bs, {
    double d[3] = {x[3 * j] - x[3 * i], x[3 * j + 1] - x[3 * i + 1], x[3 * j + 2] - x[3 * i + 2]};
    double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    double dist = sqrt(r2);
    double ori = rho[i] - gb->offset, orj = rho[j] - gb->offset;
    double bi_ = born[i], bj_ = born[j];
    double expterm = exp(-(dist * dist) / (4.0 * bi_ * bj_));
    double fGB = sqrt(dist * dist + bi_ * bj_ * expterm);
    double qi = sys->charge[i], qj = sys->charge[j];
    double energy = -(qi * qj) * (1 / fGB) * fac;
    double dRi = gb_dR(dist, rho[i], ori, orj, S[j], bi_, psi[i], gb->alpha_obc, gb->beta_obc, gb->gamma_obc);
    double dRj = gb_dR(dist, rho[j], orj, ori, S[i], bj_, psi[j], gb->alpha_obc, gb->beta_obc, gb->gamma_obc);
    double force = 0.0;
    force -= qi * qj * (1 / (fGB * fGB)) * 0.5 * (1 / fGB) *
             ((2 * dist - 0.5 * expterm * dist) + expterm * dRi * (bj_ + (dist * dist) / (4 * bi_)) +
              expterm * dRj * (bi_ + (dist * dist) / (4 * bj_))) *
             (1 / dist);
    force -= 0.5 * qi * qi * dRi / (bi_ * bi_) / dist;
    force -= 0.5 * qj * qj * dRj / (bj_ * bj_) / dist;
    /* partialTerm[i][j] = aTerm of the (i,j) pair, partialTerm[j][i] = bTerm */
    double part = qi * qj * (1 / (fGB * fGB)) * 0.5 * (1 / fGB) * expterm;
    double aT = part * (bj_ + (dist * dist) / (4.0 * bi_));
    double bT = part * (bi_ + (dist * dist) / (4.0 * bj_));
    force -= (partial[i] - aT) * (dRi / dist);
    force -= (partial[j] - bT) * (dRj / dist);
    force *= fac;
    egb += energy;
    double fij[3] = {d[0] * force, d[1] * force, d[2] * force};
    f[3 * i] -= fij[0]; f[3 * i + 1] -= fij[1]; f[3 * i + 2] -= fij[2];
    f[3 * j] += fij[0]; f[3 * j + 1] += fij[1]; f[3 * j + 2] += fij[2];
    GB_PAIR_VIRIAL(fij, d)
  })
  /* GBForce::postProcess (GBForce.cpp:230-242) */
  double tot = 0.0;
  for (int i = 0; i < n; i++) tot += self[i];
  e[ORC_COULOMB] += egb;
  e[ORC_COULOMB] += tot;
  if (gb->ace) {
    /* GBACEForce::operator() (GBACEForce.cpp:14-157), the same all-pairs walk;
     * PI is the reference's own constant 3.14169 (GBACEForce.cpp:7) */
    const double PIa = 3.14169, sg = gb->ace_sigma, rs = gb->ace_rho_s;
    const double c1 = 24 * PIa * sg;
    GB_FOR_PAIRS(n, bs, {
      double d[3] = {x[3 * j] - x[3 * i], x[3 * j + 1] - x[3 * i + 1], x[3 * j + 2] - x[3 * i + 2]};
      double dist = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
      double ori = rho[i] - gb->offset, orj = rho[j] - gb->offset;
      double dRi = gb_dR(dist, rho[i], ori, orj, S[j], born[i], psi[i], gb->alpha_obc, gb->beta_obc, gb->gamma_obc);
      double dRj = gb_dR(dist, rho[j], orj, ori, S[i], born[j], psi[j], gb->alpha_obc, gb->beta_obc, gb->gamma_obc);
      double ci = (rho[i] + rs) * (rho[i] + rs) * pow(rho[i], 6);
      double cj = (rho[j] + rs) * (rho[j] + rs) * pow(rho[j], 6);
      double bi7 = pow(born[i], 7), bj7 = pow(born[j], 7);
      double force = c1 * (ci * (1 / bi7) * dRi * (1 / dist) + cj * (1 / bj7) * dRj * (1 / dist));
      double fij[3] = {d[0] * force, d[1] * force, d[2] * force};
      f[3 * i] -= fij[0]; f[3 * i + 1] -= fij[1]; f[3 * i + 2] -= fij[2];
      f[3 * j] += fij[0]; f[3 * j + 1] += fij[1]; f[3 * j + 2] += fij[2];
      GB_PAIR_VIRIAL(fij, d)
    })
    /* GBACEForce::postProcess (GBACEForce.cpp:160-170): sum of the ACE
     * potentials 4 PI sigma (rho_i + rho_s)^2 (rho_i / R_i)^6 */
    double ace = 0.0;
    for (int i = 0; i < n; i++) {
      double p1 = rho[i] / born[i];
      ace += 4 * PIa * sg * (rho[i] + rs) * (rho[i] + rs) * pow(p1, 6);
    }
    e[ORC_COULOMB] += ace;
  }
#undef GB_PAIR_VIRIAL
  if (born_out) memcpy(born_out, born, sizeof(double) * (size_t)n);
  free(burial); free(born); free(psi); free(partial); free(self); free(com);
  return 0;
}

/* ------------------------------------------------------------- bonded */

void orc_bonded(const orc_system *sys, int nbonds, const int *bonds, const double *bond_k,
                const double *bond_r0, int nangles, const int *angles, const double *angle_k,
                const double *angle_t0, const double *x, double *f, double *e, int virial) {
  orc_pbc pbc;
  if (sys->periodic) pbc_set(&pbc, sys->box, sys->origin);
  /* BondSystemForce::calcBond (BondSystemForce.h:77-108) */
  for (int b = 0; b < nbonds; b++) {
    int a1 = bonds[2 * b], a2 = bonds[2 * b + 1];
    double r12[3];
    if (sys->periodic) pbc_min_diff(&pbc, x + 3 * a2, x + 3 * a1, r12);
    else for (int k = 0; k < 3; k++) r12[k] = x[3 * a1 + k] - x[3 * a2 + k];
    double r = sqrt(r12[0] * r12[0] + r12[1] * r12[1] + r12[2] * r12[2]);
    double dp = 2.0 * bond_k[b] * (r - bond_r0[b]);
    double f1[3] = {r12[0] * (-dp / r), r12[1] * (-dp / r), r12[2] * (-dp / r)};
    for (int k = 0; k < 3; k++) { f[3 * a1 + k] += f1[k]; f[3 * a2 + k] -= f1[k]; }
    e[ORC_BOND] += bond_k[b] * (r - bond_r0[b]) * (r - bond_r0[b]);
    if (virial) add_virial(e, f1, r12);  /* BondSystemForce.h:121-122 */
  }
  /* AngleSystemForce::calcAngle (AngleSystemForce.h:66-140), no Urey-Bradley */
  for (int a = 0; a < nangles; a++) {
    int a1 = angles[3 * a], a2 = angles[3 * a + 1], a3 = angles[3 * a + 2];
    double r12[3], r32[3];
    if (sys->periodic) {
      pbc_min_diff(&pbc, x + 3 * a2, x + 3 * a1, r12);
      pbc_min_diff(&pbc, x + 3 * a2, x + 3 * a3, r32);
    } else
      for (int k = 0; k < 3; k++) { r12[k] = x[3 * a1 + k] - x[3 * a2 + k]; r32[k] = x[3 * a3 + k] - x[3 * a2 + k]; }
    double d12 = sqrt(r12[0] * r12[0] + r12[1] * r12[1] + r12[2] * r12[2]);
    double d32 = sqrt(r32[0] * r32[0] + r32[1] * r32[1] + r32[2] * r32[2]);
    double cr[3] = {r12[1] * r32[2] - r12[2] * r32[1], r12[2] * r32[0] - r12[0] * r32[2],
                    r12[0] * r32[1] - r12[1] * r32[0]};
    double theta = atan2(sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]),
                         r12[0] * r32[0] + r12[1] * r32[1] + r12[2] * r32[2]);
    double st = sin(theta), ct = cos(theta);
    double dp = 2.0 * angle_k[a] * (theta - angle_t0[a]);
    double f1[3], f3[3];
    for (int k = 0; k < 3; k++) {
      double dr12 = r12[k] / d12, dr32 = r32[k] / d32;
      double dt1 = (dr12 * ct - dr32) / (st * d12);
      double dt3 = (dr32 * ct - dr12) / (st * d32);
      f1[k] = -dt1 * dp - 0.0;
      f3[k] = -dt3 * dp - 0.0;
    }
    for (int k = 0; k < 3; k++) {
      f[3 * a1 + k] += f1[k];
      f[3 * a2 + k] += -f1[k] - f3[k];
      f[3 * a3 + k] += f3[k];
    }
    e[ORC_ANGLE] += angle_k[a] * (theta - angle_t0[a]) * (theta - angle_t0[a]);
    if (virial) {  /* AngleSystemForce.h:155-172 */
      double xy = f1[0] * r12[1] + f3[0] * r32[1];
      double xz = f1[0] * r12[2] + f3[0] * r32[2];
      double yz = f1[1] * r12[2] + f3[1] * r32[2];
      e[ORC_VIRIALXX] += f1[0] * r12[0] + f3[0] * r32[0];
      e[ORC_VIRIALXY] += xy;
      e[ORC_VIRIALXZ] += xz;
      e[ORC_VIRIALYX] += xy;
      e[ORC_VIRIALYY] += f1[1] * r12[1] + f3[1] * r32[1];
      e[ORC_VIRIALYZ] += yz;
      e[ORC_VIRIALZX] += xz;
      e[ORC_VIRIALZY] += yz;
      e[ORC_VIRIALZZ] += f1[2] * r12[2] + f3[2] * r32[2];
    }
  }
}

/* ---------------------------------------------------------- RNG, velocity */

/* randomNumber(seed, randomtype) (MathUtilities.cpp:153-187): randomtype 1 is
 * ProtoMol's Random class (Random.cpp:33-68), the Park-Miller minimal standard
 * generator (Schrage's method) with a Bays-Durham shuffle of 32; randomtype 0
 * (the default, MainModule.cpp:53) is the C library's drand48 after
 * srand48(seed) on Linux -- here erand48 on a state of our own, the same
 * generator. */
#define RNG_IA 16807
#define RNG_IM 2147483647
#define RNG_IQ 127773
#define RNG_IR 2836
#define RNG_NTAB 32
#define RNG_NDIV (1 + (RNG_IM - 1) / RNG_NTAB)

void orc_rng_seed(orc_rng *r, uint32_t seed, int randomtype) {
  memset(r, 0, sizeof(*r));
  r->type = randomtype ? 1 : 0;
  /* srand48: X = seed << 16 | 0x330E (the low 32 bits of a long seed) */
  r->x48[0] = 0x330E;
  r->x48[1] = (unsigned short)(seed & 0xFFFFu);
  r->x48[2] = (unsigned short)(seed >> 16);
  int32_t idum = (int32_t)seed; /* Random::seed takes an unsigned, stores it in an int */
  if (idum < 1) idum = 1;
  for (int j = RNG_NTAB + 7; j >= 0; j--) {
    int32_t k = idum / RNG_IQ;
    idum = RNG_IA * (idum - k * RNG_IQ) - RNG_IR * k;
    if (idum < 0) idum += RNG_IM;
    if (j < RNG_NTAB) r->iv[j] = idum;
  }
  r->iy = r->iv[0];
  r->idum = idum;
}

double orc_rng_uniform(orc_rng *r) {
  if (r->type == 0) return erand48(r->x48);
  int32_t k = r->idum / RNG_IQ;
  r->idum = RNG_IA * (r->idum - k * RNG_IQ) - RNG_IR * k;
  if (r->idum < 0) r->idum += RNG_IM;
  int j = r->iy / RNG_NDIV;
  r->iy = r->iv[j];
  r->iv[j] = r->idum;
  const double inv = 1.0 / RNG_IM, less1 = 1.0 - 3.0e-16; /* Random.cpp:11-13 */
  double t = inv * r->iy;
  return t > less1 ? less1 : t;
}

/* randomGaussianNumber (MathUtilities.cpp:206-236): polar Box-Muller, the
 * second deviate of each pair kept for the next call */
double orc_rng_gaussian(orc_rng *r) {
  if (r->have_spare) { r->have_spare = 0; return r->spare; }
  double rr = 2.0, v1 = 0, v2 = 0;
  while (rr >= 1.0 || rr < 1.523e-8) {
    v1 = 2.0 * orc_rng_uniform(r) - 1.0;
    v2 = 2.0 * orc_rng_uniform(r) - 1.0;
    rr = v1 * v1 + v2 * v2;
  }
  double fac = sqrt(-2.0 * log(rr) / rr);
  r->spare = v1 * fac;
  r->have_spare = 1;
  return v2 * fac;
}

/* randomVelocity (TopologyUtilities.cpp:21-48): sqrt(kB T / m) times three
 * deviates per atom, drawn from r */
void orc_random_velocity(const orc_system *sys, double temperature, orc_rng *r, double *v) {
  const double kbT = temperature * 0.001987191; /* PMConstants.cpp:70 */
  for (int i = 0; i < sys->n; i++) {
    double s = sqrt(kbT / sys->mass[i]);
    v[3 * i] = s * orc_rng_gaussian(r);
    v[3 * i + 1] = s * orc_rng_gaussian(r);
    v[3 * i + 2] = s * orc_rng_gaussian(r);
  }
}

/* kineticEnergy (TopologyUtilities.cpp:163-171) */
double orc_kinetic_energy(const orc_system *sys, const double *v) {
  double ke = 0.0;
  for (int i = 0; i < sys->n; i++)
    ke += sys->mass[i] * (v[3 * i] * v[3 * i] + v[3 * i + 1] * v[3 * i + 1] + v[3 * i + 2] * v[3 * i + 2]);
  return ke * 0.5;
}
