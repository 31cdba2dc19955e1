// pm_core.h -- device data layout shared by the gfx950 kernels and the host
// side of the C ABI.
//
// Layout in HBM (per system; a context holds one or more systems and the
// kernels reach each through a device-resident table of SysDev records):
//   original atom order  x, v, xref (FP64 SoA), charge, inverse mass, slot
//   cluster slot order   xq (float4: cluster-local x, y, z and scaled charge),
//                        LJ type, original index, fixed-point force (3 x int64)
//   clusters of 8 slots  FP64 origin, FP32 bounding box
//   super-clusters       (8 z-consecutive clusters of a column) pair list of
//                        32-byte entries (j cluster, image shift, FP32 offset,
//                        i-cluster mask, exclusion-mask slot)
//   Ewald k space        per-atom records of e^{i n 2 pi x / L}, the
//                        k-vector rows, structure-factor partial sums
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pm {

constexpr int kCluster = 8;        // atoms per cluster (one 8x8 tile = one wave64)
constexpr int kScCl = 8;           // clusters per super-cluster (z-consecutive in one column)
constexpr int kMaskSlots = 32;     // least exclusion-mask slots per super-cluster list
constexpr int kPairSplitDefault = 12;  // waves sharing one super-cluster's list in the pair kernel
constexpr int kWave = 64;
// Ewald structure-factor tiling: waves per workgroup x 16-point column tiles
// per wave (the workgroup's points = waves x tiles x 16)
#ifndef PM_EW_SF_WAVES
#define PM_EW_SF_WAVES 4
#endif
#ifndef PM_EW_SF_TILES
#define PM_EW_SF_TILES 4
#endif
constexpr int kEwSFWaves = PM_EW_SF_WAVES;
constexpr int kEwSFTiles = PM_EW_SF_TILES;
constexpr int kEwSFPoints = kEwSFWaves * kEwSFTiles * 16;
constexpr int kSortCap = 4096;     // max atoms per x-y column (LDS bitonic sort)
constexpr double kFScale = 4294967296.0;      // 2^32 fixed point for forces
constexpr double kFScaleInv = 1.0 / 4294967296.0;
constexpr double kEScale = 1048576.0;         // 2^20 fixed point for energy/virial totals
constexpr double kEScaleInv = 1.0 / 1048576.0;
constexpr int kShiftCenter = 13;   // (0,0,0) image in the 27-shift encoding

// energy accumulator slots (int64 fixed point, per system)
enum { kAccLJ = 0, kAccCoul = 1, kAccVxx, kAccVxy, kAccVxz, kAccVyy, kAccVyz, kAccVzz,
       kAccBond, kAccAngle, kAccMVxx, kAccMVxy, kAccMVxz, kAccMVyy, kAccMVyz, kAccMVzz, kAccN = 16 };
// The accumulators exist in kAccCopies copies of one 128-byte line each; a
// wave adds into copy (wave index mod kAccCopies), k_gather_scalars sums the
// copies.  Integer sums are exact, so the totals do not depend on the split,
// and thousands of waves no longer queue for the atomics of one cache line.
constexpr int kAccCopies = 64;
constexpr int kDispSlots = 32;   // per-system slots of the step's largest displacement
__host__ __device__ inline long long *acc_copy(long long *acc, int w) {
  return acc + (size_t)(w & (kAccCopies - 1)) * kAccN;
}

// device-side status bits (per system)
enum { kErrColumnCap = 1, kErrListCap = 2, kErrNonFinite = 4, kErrClusterCap = 8, kErrVacuumRegion = 16 };

// kinds of the local terms of an atom (bonded terms, Ewald correction pairs)
enum { kLocBond = 0, kLocAngle = 1, kLocCorr = 2 };
enum { kLocalForces = 1, kLocalEnergies = 2 };
// The local terms of an atom in kLocShares shares by (kind, slot): share w
// takes the items slot, slot + nslots, ... of one kind of each atom (in the
// order of the atom's list of that kind), so a wave per share runs one kind's
// code.  Slots per kind: bonds 2, angles 2, Ewald corrections 4.
constexpr int kLocShares = 8;
__host__ __device__ inline void loc_share(int w, int &kind, int &slot, int &nslots) {
  if (w < 2) {
    kind = kLocBond, slot = w, nslots = 2;
  } else if (w < 4) {
    kind = kLocAngle, slot = w - 2, nslots = 2;
  } else {
    kind = kLocCorr, slot = w - 4, nslots = 4;
  }
}
// One local term of an atom, seen from that atom: the term's atoms, the
// atom's role in it (-1: no term) and its parameters -- bond (k, r0), angle
// (k, theta0), Ewald correction pair (q_a q_b, times 1 - csf for a modified
// pair).
struct LocRec {
  int a0, a1, a2, role;
  double p0, p1;
};
// host staging transfers (k_import / k_export)
enum { kXferX = 1, kXferV = 2, kXferF = 4 };
// packed per-system scalars (k_gather_scalars): acc[0..16), census[16..18), err,
// ctl[19..23), largest squared displacement of the run (float bits)
enum { kScalCensus = 16, kScalErr = 18, kScalCtl = 19, kScalMaxdispRun = 25, kScalN = 26 };
// what a pair-kernel launch accumulates; the MV kinds also sum the molecular
// virial (and the atomic one as the reference does with it on)
enum { kPairForces = 0, kPairForcesEnergies = 1, kPairCensus = 3, kPairForcesEnergiesMV = 4 };

// per-system control words: [0] rebuild the pair list in this step (0/1),
// [1] steps since the last rebuild, [2] rebuilds since the start of the run,
// [3] the next kick is the first of a run (a half kick), [4] entries (8x8
// cluster pairs) of the pair list at its last build, [5] force evaluations
// so far (the counter of the Langevin noise stream)
enum { kCtlRebuild = 0, kCtlSince = 1, kCtlCount = 2, kCtlFirst = 3, kCtlEntries = 4, kCtlNoise = 5, kCtlN = 6 };
// k_step_begin modes (bit set)
enum { kStepForce = 1, kStepAuto = 2 };

struct SwitchDev {
  int kind;                // PM_SWITCH_*
  float cut2, on2;         // C2 / cutoff: cutoff^2, switchon^2 (FP32 path)
  float s1, s2, s3;        // C2 constants (C2SwitchingFunction.cpp:12-18)
  // C2 in factored form (u = max(r^2, on^2), w = u - on^2, c2 = cut^2 - u):
  // S = a1 c2^2 (w + D/2), dS = -6 a1 c2 w, D = cut^2 - on^2, a1 = 2 s1
  float c2d, c2hd, c2a1, c2sc;  // D, D / 2, a1, 12 s1 (scale of pre-scaled LJ rows)
  float c15rc1, c05rc3, c15rc3;  // C1 constants (C1SwitchingFunction.cpp:12-18)
  // Cn / CmpCnCn are evaluated in FP64: their alternating-sign polynomials lose
  // too many digits in FP32
  double dcut, dcut2, don, don2, doff, doff2;
  int ord_int, ord_idx;
  double irange[8], itrange[8];
};

struct ForceDev {
  int term_a, term_b;      // PM_TERM_*
  SwitchDev sw_a, sw_b;
  float cutA2, cutB2;      // inclusion cutoff^2 per term (FP32 decision)
  double cutA2d, cutB2d;   // the same in FP64 (guard-band recheck)
  float guard;             // |r2 - cut2| <= guard  ->  FP64 recheck
  // the band's edges, cut2 -/+ guard per term: r2 <= lo is inside, r2 > hi
  // outside, anything between goes to the FP64 recheck
  float cutA2lo, cutA2hi, cutB2lo, cutB2hi;
  // the pair kernel finishes the force pass's step bookkeeping (step_finish_ctl);
  // 0 for a pair launch outside a force pass (the census of the MD state's list)
  int finish_step;
  // Ewald real space: alpha, 0.3275911 alpha (the A&S 7.1.26 t argument),
  // -alpha^2 log2(e) (exp2 argument per r^2), 2 alpha / sqrt(pi)
  float alpha, alpha_p, alpha2_log2e, two_alpha_pi;
  double dalpha, dtwo_alpha_pi;             // Ewald correction (FP64)
  float rlist2;
  unsigned flags;
  int pair_waves;          // waves per workgroup of the pair kernel (1, 2, 4 or 8)
  int pair_split;          // waves sharing one super-cluster's list
  // GB / OBC (ProtoMolApp.cpp:285-301, GBForce.cpp:11-12)
  double gb_offset, gb_a, gb_b, gb_g, gb_fac;  // fac = 1/eps_solute - 1/eps_solvent
  // GBACEForce (GBACEForce.cpp): sigma, rho_s; c1 = 24 PI sigma with the
  // reference's PI = 3.14169
  int gb_ace;
  double gb_ace_sigma, gb_ace_rho_s, gb_ace_c1;
  // all-pairs (vacuum) context: the GB passes, the LJ term, the Coulomb term
  int gb_on, lj_on, coul_on;
  int virial, molvirial;   // ScalarStructure::virial() / molecularVirial()
};

// One system as the kernels see it.  All pointers are device pointers owned
// by the host-side context; the struct itself lives in device memory.
struct SysDev {
  // geometry: PeriodicBoundaryConditions constants, reference-exact FP64
  double L[3], Lr[3], H[3], H2[3], D;
  double omin[3];          // origin - L/2: lower corner of the wrapped frame
  int periodic;
  int vac;                 // a vacuum system in an enclosing periodic cell: its atoms must stay
  double vac_lo[3], vac_hi[3];  // within [vac_lo, vac_hi) (checked at each rebuild)
  int ncx, ncy, ncol;      // x-y column grid
  int rx, ry;              // column search radius for the pair list
  double ax, ay;           // column widths
  int n_atoms, cl_cap, slot_cap, list_cap;
  int lj_mode, n_types;
  float csf;               // coulombScalingFactor
  float lj_scaling;
  // original-order arrays
  double *x, *v, *xref;    // 3 * n each, SoA (x[0..n), y[n..2n), z[2n..3n))
  double *xaos, *vaos, *faos;  // 3 * n AoS FP64 staging for host transfers
  const double *charge, *inv_mass;
  const int *molecule, *type;
  const float *sig_half, *eps_sqrt; // PM_LJ_PERATOM: sigma/2, sqrt(epsilon)
  const int *excl_start, *excl_partner, *excl_class;  // CSR of exclusions (both directions)
  const unsigned long long *excl_win;  // 2 per atom: FULL / MODIFIED bits over partners a-31..a+31; FULL bit 63: far partner
  int *slot;               // atom -> slot
  float4 *lbuild;          // local coordinates at the last rebuild (displacement check)
  // binning: per column a bucket of kSortCap (zkey << 32 | atom) keys and its count
  int *col_count, *col_cl_start, *col_ncl;
  int *col_sc_start, *col_nsc;  // per column: first super-cluster, count
  int *sc_col;             // super-cluster -> column
  int sc_cap;
  unsigned long long *bucket;
  // slot order
  float4 *xq;              // local coordinates + scaled charge
  int *ljt;                // LJ type (table mode) or -1 padding
  float2 *ljp;             // per-atom mode (sigma/2, sqrt(eps))
  int *orig;               // slot -> atom (-1 padding)
  int *smol;               // slot -> molecule (-1 padding)
  long long *fx, *fy, *fz; // fixed-point forces; fy = fx + f_stride, fz = fx + 2 f_stride
  long long *fl;           // fixed-point forces of the local terms (bonded, Ewald correction),
                           // original atom order, 3 * n SoA; every force pass of k_local
                           // writes all of them, and the readers of fx add them in
  int f_stride;
  // clusters
  double *cl_origin;       // 3 * cl_cap SoA
  float4 *cl_lo, *cl_hi;   // bounding boxes (w = unused)
  int *cl_col;             // cluster -> column
  int *cl_nvalid;
  // pair list per super-cluster: list_cap entries of two uint4,
  //   {cj | shift << 23 | (nvalid_j - 1) << 28 | has_masks << 31, FP32 offset of the j
  //    cluster's origin + image from the super-cluster origin (x, y, z)},
  //   {imask (i clusters of the super-cluster in range) | mask slot << 8, 0, 0, 0}
  // and mask_cap slots of 16 lane masks (FULL, MODIFIED for each i cluster)
  int *list_len;
  uint4 *list;
  unsigned long long *list_mask;
  int mask_cap;            // mask slots per super-cluster
  // LJ table (table mode): n_types^2 float4 (A, B, A14, B14), and the pair
  // kernel's rows (A, B, 12 A, 6 B)
  const float4 *lj_table;
  const float4 *lj_table12;
  // local terms: bonds, angles, Ewald correction pairs (unique exclusion
  // table pairs), as rows of records -- row loc_base[w] + k holds item k of
  // share w of every atom (loc_rows[w] rows), so a share's loads are
  // coalesced and need no index lookups
  int n_bonds, n_angles, n_corr;
  const LocRec *loc_rec;
  int loc_rows[kLocShares], loc_base[kLocShares];
  // GB
  const double *gb_rho, *gb_S;
  double *gb_born, *gb_drfac, *gb_partial, *gb_spare;
  int gb_nt;               // 64-atom tiles
  double *gb_fpart;        // per tile pair (it <= jt): [i side, j side][x, y, z][64] FP64 force partials
  // per-system scalars
  int *n_clusters, *n_sc;
  long long *acc;          // kAccCopies x kAccN fixed-point energy/virial accumulators
  unsigned long long *census; // count, hash
  unsigned *err;
  unsigned *maxdisp2;      // kDispSlots: float bits of the largest squared displacement from the last
                           // rebuild (this step), spread over slots so the kick's atomics do not queue on one
                           // address; the value is the maximum over them
  unsigned *maxdisp_run;   // the same, folded over every step of the current run
  int *ctl;                // kCtlN control words
  unsigned lv_seed;        // Langevin noise stream key
  // molecules (molecular virial): CSR of their atoms in index order, centres
  // of mass (SoA, 3 x n_mol) at the positions of the current evaluation
  int n_mol;
  const int *mol_start, *mol_atoms;
  const double *mass;      // n, the masses the centres are weighted by
  double *com;
  // Ewald reciprocal space (NonbondedFullEwaldSystemForce::reciprocalTerm):
  // k = 2 pi (h/Lx, k/Ly, l/Lz), h >= 0 half space, |k| < kc, in the
  // reference's (h, k, l) order; rows = runs of l at fixed (h, k)
  int ew_nk, ew_nrows;
  int ew_xrows, ew_yrows, ew_zrows;  // rows of the x, y, z tables (hmax+1 up to a multiple of 8,
                                     // max(kmax,1)+1, max(lmax,1)+1)
  int ew_chunk, ew_nchunks;          // structure factor: atoms per chunk, chunks
  int ew_ntiles, ew_rl;              // structure factor: (h block, kEwSFPoints points) tiles; record length
  int ew_nfb;                        // force pass: row blocks
  double ew_org[3];                  // box origin (minimalPosition)
  double ew_c, ew_a8, ew_fac;        // 4 pi / V, 8 pi / V, 1 / (4 alpha^2)
  double2 *ew_tab;                   // per atom (cos, sin)(n 2 pi r / L), n = 0.. of x, y, z in turn
                                     // (ew_rl entries); x carries the charge
  const int4 *ew_rows;               // h, k, first l, first k index (+ sentinel)
  const double4 *ew_kd;              // per k: kx, ky, kz, b = exp(-k^2 / (4 alpha^2)) / k^2
  const int4 *ew_tiles;              // first h, first point, end point of the h block, h count
  const short2 *ew_pts;              // (k, l) points of each h block
  const int *ew_kmap;                // per point and h of its block: k index or -1
  const int *ew_fb_rows;             // first row of each force block (+ end)
  double2 *ew_psum;                  // [chunk][k] partial sums of (cos, sin)
  double4 *ew_sk;                    // per k: (8 pi / V) b sum cos, (8 pi / V) b sum sin, kz, b |S|^2
  double *ew_pf;                     // [force block][3][atom] partial forces
};

}  // namespace pm
