/*
 * protomol_amd.h -- C ABI of the MI355X (gfx950) nonbonded force path.
 *
 * Plain pointers and sizes only; every entry point returns PM_OK (0) or a
 * negative PM_E* status, and pm_last_error() gives the message.  No C++
 * exception crosses this boundary: the host adapter in
 * protomol_amd_systemforce.hpp turns a status into ProtoMol's THROW.
 *
 * Reference interfaces each entry point replaces (paths relative to the
 * ProtoMol source tree):
 *   pm_ctx_create    Topology<PBC,CubicCellManager> + ForceFactory::make of
 *                    NonbondedCutoffSystemForce<CCM, OneAtomPair[Two]<...>>
 *                    (factory/ForceFactory.cpp:21, module/NonbondedCutoffForceModule.cpp:56-88),
 *                    NonbondedFullEwaldSystemForce<PBC,CCM,true,false,corr,Cutoff> and,
 *                    with PM_FLAG_EWALD_RECIPROCAL, <PBC,CCM,true,true,corr,Cutoff>
 *                    ("CoulombEwald", module/NonbondedFullElectrostaticForceModule.cpp:31-36), and the GB
 *                    NonbondedSimpleFullSystemForce<OneAtomPairNoExclusion<VBC,Universal,GB*>>
 *                    (module/NonbondedSimpleFullForceModule.cpp:203-206)
 *   pm_evaluate      SystemForce::evaluate(topo, positions, forces, energies)
 *                    (force/system/SystemForce.h:32-35) -- ADDS into forces and energies
 *   pm_md_*          LeapfrogIntegrator::run (integrator/leapfrog/LeapfrogIntegrator.cpp:89-104)
 *                    with the force group evaluated on the device every step
 *   pm_random_velocity  randomVelocity (topology/TopologyUtilities.cpp:21-48) after
 *                    MainModule's seeding (module/MainModule.cpp:88-102)
 *   pm_pair_census   no reference counterpart: order-independent digest of the
 *                    evaluated pair set, for bit-exact pair-set checks
 *   pm_ctx_reserve, pm_system_*, pm_*_systems, pm_md_begin/advance/end
 *                    no reference counterpart: systems that join and leave a
 *                    running context (admission by device-memory footprint),
 *                    and MD of any subset of its systems in one fused launch
 *                    sequence -- the serving and batching forms
 */
#ifndef PROTOMOL_AMD_H
#define PROTOMOL_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PM_ABI_VERSION 5

/* status codes */
#define PM_OK 0
#define PM_EINVAL -1    /* bad argument or unsupported configuration */
#define PM_ENOMEM -2    /* device or host allocation failed */
#define PM_EDEVICE -3   /* HIP runtime error */
#define PM_ECAPACITY -4 /* a device-side capacity was exceeded (column, list) */
#define PM_ENOTREADY -5 /* state not set yet */
#define PM_EFULL -6     /* the context's device reservation has no room for the system */

/* switching functions (protomol/switch/) */
#define PM_SWITCH_NONE 0    /* UniversalSwitchingFunction */
#define PM_SWITCH_CUTOFF 1  /* CutoffSwitchingFunction */
#define PM_SWITCH_C1 2      /* C1SwitchingFunction */
#define PM_SWITCH_C2 3      /* C2SwitchingFunction */
#define PM_SWITCH_CN 4      /* CnSwitchingFunction */
#define PM_SWITCH_CMPCNCN 5 /* CmpCnCnSwitchingFunction */

/* pair terms */
#define PM_TERM_NONE 0
#define PM_TERM_LJ 1         /* LennardJonesForce (force/LennardJonesForce.h:28-60) */
#define PM_TERM_COULOMB 2    /* CoulombForce (force/CoulombForce.h:29-40) */
#define PM_TERM_EWALD_REAL 3 /* NonbondedFullEwaldSystemForce::realTerm (:758-858) */

/* LJ parameter modes */
#define PM_LJ_TABLE 0   /* LennardJonesParameterTable: T x T of (A, B, A14, B14) */
#define PM_LJ_PERATOM 1 /* per-atom sigma/epsilon, GROMACS mixing (BuildTopology.cpp:650-659) */

/* exclusion classes (topology/ExclusionType.h) */
#define PM_EXCL_FULL 1
#define PM_EXCL_MODIFIED 2

/* flags */
#define PM_FLAG_VIRIAL 1u     /* ScalarStructure::virial() on */
#define PM_FLAG_GB 2u         /* GBBornRadii -> GBPartialSum -> GBForce, all pairs, vacuum */
#define PM_FLAG_BONDED 4u     /* harmonic Bond + Angle (force/bonded/{Bond,Angle}SystemForce.h) */
#define PM_FLAG_TIMING 8u     /* time every pair-kernel launch with HIP events (pm_stats [4], [5]) */
#define PM_FLAG_GB_ACE 32u   /* with PM_FLAG_GB: add GBACEForce, the ACE non-polar solvation term
                                 (force/GB/GBACEForce.cpp:14-170), parameters gb_ace_* */
#define PM_FLAG_EWALD_RECIPROCAL 16u /* with PM_TERM_EWALD_REAL: add the reciprocal-space sum over
                                        |k| < kc (NonbondedFullEwaldSystemForce::reciprocalTerm,
                                        force/nonbonded/NonbondedFullEwaldSystemForce.h:873-1083) */
#define PM_FLAG_MOLVIRIAL 64u /* ScalarStructure::molecularVirial() on: pairs of two molecules add
                                 f (x) (R_J - R_I) of the molecules' centres of mass to the
                                 MOLVIRIAL slots (force/OneAtomPair.h:95-103, OneAtomPairTwo.h:87-96,
                                 OneAtomPairNoExclusion.h:73-81, NonbondedFullEwaldSystemForce.h:
                                 839-848, 1039-1078) and f (x) d to the VIRIAL slots even with
                                 PM_FLAG_VIRIAL off, as the reference's addVirial does */

#define PM_FLAG_ALL_PAIRS 128u /* all-pairs vacuum context, NonbondedSimpleFullSystemForce
                                  (force/nonbonded/NonbondedSimpleFullSystemForce.h:125-150), with
                                  term_a = PM_TERM_LJ and/or term_b = PM_TERM_COULOMB (either may be
                                  PM_TERM_NONE) and, with PM_FLAG_GB as well, the GB passes;
                                  PM_FLAG_GB alone is GB + LJ + Coulomb */

/* ScalarStructure slots (type/ScalarStructure.h:26-67); energies[] has 36 */
#define PM_E_COULOMB 0
#define PM_E_LENNARDJONES 1
#define PM_E_BOND 2
#define PM_E_ANGLE 3
#define PM_E_VIRIALXX 7
#define PM_E_MOLVIRIALXX 16
#define PM_E_NSLOTS 36

typedef struct pm_switch_desc {
  int32_t kind;     /* PM_SWITCH_* */
  double switchon;  /* -switchon */
  double cutoff;    /* -cutoff */
  double switchoff; /* -switchoff (Cn, CmpCnCn) */
  double order;     /* -n (Cn, CmpCnCn) */
} pm_switch_desc;

typedef struct pm_system_desc {
  int32_t n_atoms;
  const double *scaled_charge; /* n; charge * 18.2226123264 (BuildTopology.cpp:92) */
  const double *mass;          /* n; amu */
  const int32_t *type;         /* n; LJ type (PM_LJ_TABLE) */
  const int32_t *molecule;     /* n; exclusions are only looked up within a molecule */
  int32_t lj_mode;
  int32_t n_types;
  const double *lj_table;      /* n_types^2 * 4: A, B, A14, B14 */
  const double *lj_sigma;      /* n (PM_LJ_PERATOM) */
  const double *lj_epsilon;    /* n (PM_LJ_PERATOM) */
  double lj_scaling;           /* LJScalingFactor for MODIFIED pairs (PM_LJ_PERATOM) */
  int32_t n_exclusions;
  const int32_t *exclusions;   /* n_exclusions * 3: atom a, atom b, PM_EXCL_* (ExclusionTable::add order) */
  double coulomb_scaling;      /* GenericTopology::coulombScalingFactor */
  int32_t periodic;            /* 1: orthogonal PeriodicBoundaryConditions, 0: vacuum */
  double box[3];               /* periodic: cellBasisVector1.x, 2.y, 3.z; vacuum on the cell-list
                                  path: the extent of the region the atoms stay in (an atom
                                  outside origin +- box / 2 is PM_ECAPACITY); vacuum all-pairs
                                  (PM_FLAG_GB, PM_FLAG_ALL_PAIRS): unused */
  double origin[3];            /* cellorigin; vacuum: the region's centre */
  int32_t n_bonds;
  const int32_t *bonds;        /* 2 per bond */
  const double *bond_k, *bond_r0;
  int32_t n_angles;
  const int32_t *angles;       /* 3 per angle, atom2 = vertex */
  const double *angle_k, *angle_theta0; /* theta0 in radians */
  const double *gb_radius;     /* n; vanDerWaalRadius (GB) */
  const double *gb_scaling;    /* n; scalingFactor (GB) */
} pm_system_desc;

typedef struct pm_force_desc {
  int32_t term_a;              /* ForceA of OneAtomPair[Two]: PM_TERM_LJ or PM_TERM_COULOMB or NONE */
  pm_switch_desc switch_a;
  int32_t term_b;              /* ForceB: PM_TERM_COULOMB, PM_TERM_EWALD_REAL or NONE */
  pm_switch_desc switch_b;     /* ignored for PM_TERM_EWALD_REAL (hard cutoff Rc) */
  double cutoff;               /* NonbondedCutoffForce -cutoff (cell search radius) */
  double ewald_alpha;          /* Ewald: alpha */
  double ewald_accuracy;       /* Ewald: Rc = sqrt(-ln(accuracy)) / alpha */
  int32_t ewald_correction;    /* add excluded-pair erf, point-self and charged-system terms */
  double gb_offset;            /* dielecOffset (0.09) */
  double gb_alpha, gb_beta, gb_gamma; /* OBC constants (ProtoMolApp.cpp:285-301) */
  double gb_solute_dielec, gb_solvent_dielec;
  double skin;                 /* pair-list buffer (Angstrom); 0 = choose */
  int32_t rebuild_interval;    /* MD: rebuild the pair list at least every this many steps
                                  (0 = only by the displacement test); it is always rebuilt
                                  when an atom has moved more than skin/2 since the build */
  uint32_t flags;              /* PM_FLAG_* */
  /* launch shapes (0 = default; an autotuner may choose them per workload) */
  double cell_scale;           /* x-y column width as a multiple of (8/density)^(1/3): 0.6 .. 2 */
  int32_t pair_waves;          /* waves per workgroup of the pair kernel: 1, 2, 4 (default) or 8 */
  int32_t pair_split;          /* waves sharing one super-cluster's pair list: 1 .. 64 (default 12) */
  double gb_ace_sigma;         /* GBACEForce -solvationparam (kcal/mol/A^2; 0 = 2.26/418.4) */
  double gb_ace_rho_s;         /* GBACEForce -watersphereradius (A; 0 = 1.4) */
} pm_force_desc;

typedef struct pm_ctx pm_ctx;

/* Library / device identity. */
int pm_abi_version(void);
int pm_device_info(char *name, int32_t name_len, int32_t *n_cu, double *hbm_gib);

/* Create a context holding n_systems independent systems (a batch of 1 for
 * the plain drop-in; 0 for a context that systems join later).  All systems
 * share the force description.  System ids are 0 .. n_systems-1, then the
 * ids pm_system_add returns (the lowest free one). */
int pm_ctx_create(const pm_force_desc *force, int32_t n_systems, const pm_system_desc *systems,
                  pm_ctx **out);
int pm_ctx_destroy(pm_ctx *ctx);
/* Reserve `bytes` of device memory for the context's systems (once, before
 * any system is added).  Afterwards every system is placed inside it and
 * pm_system_add returns PM_EFULL when it does not fit.  Admission leaves 1/16
 * of the reservation free for the pair lists of admitted systems, which grow
 * (double) when an MD step overflows them; a growth that still finds no room
 * makes that call return PM_EFULL. */
int pm_ctx_reserve(pm_ctx *ctx, uint64_t bytes);
/* Reserved and used bytes of the reservation (0, 0 without one). */
int pm_ctx_memory(pm_ctx *ctx, uint64_t *reserved, uint64_t *used);
/* Device bytes a system would take in this context (its admission price). */
int pm_system_footprint(pm_ctx *ctx, const pm_system_desc *system, uint64_t *bytes);
/* Add a system (copied; the description may be freed afterwards). */
int pm_system_add(pm_ctx *ctx, const pm_system_desc *system, int32_t *id);
/* Add n systems in order, as n pm_system_add calls would, for a serving loop
 * that admits many requests at once: the host tables are prepared on worker
 * threads and each system's tables reach the device in one staged copy.
 * ids[k] receives system k's id, or -1 when the reservation has no room for
 * it; after max_full (>= 1) such systems the rest are not tried (-1 as well).
 * PM_OK unless another failure occurs, which removes the systems this call
 * added.  (No counterpart in the reference, whose systems are fixed at
 * set-up; it batches what pm_system_add does one system at a time.) */
int pm_system_add_many(pm_ctx *ctx, int32_t n, const pm_system_desc *systems, int32_t max_full, int32_t *ids);
/* Remove a system and give its device memory back. */
int pm_system_remove(pm_ctx *ctx, int32_t id);
/* Last error message for ctx (or the calling thread when ctx is NULL). */
const char *pm_last_error(const pm_ctx *ctx);

/* SystemForce::evaluate: positions x (3n AoS FP64, host) of one system;
 * ADDS forces into f (3n AoS FP64, host) and energies into e[36]. */
int pm_evaluate(pm_ctx *ctx, int32_t system, const double *x, double *f, double *e);

/* The same over all systems at once (one fused launch sequence); x, f are the
 * systems' arrays concatenated in id order, e holds 36 slots per system. */
int pm_evaluate_batch(pm_ctx *ctx, const double *x, double *f, double *e);
/* The same over the listed systems, arrays concatenated in list order. */
int pm_evaluate_systems(pm_ctx *ctx, int32_t n, const int32_t *ids, const double *x, double *f, double *e);

/* Digest of the pair set the last evaluation of `system` used: count and
 * sum of splitmix64(lo | hi << 32) over pairs (lo < hi, original indices).
 * x == NULL: the pair set that the MD state's current pair list yields at
 * its positions, without a rebuild (the set the last MD step evaluated). */
int pm_pair_census(pm_ctx *ctx, int32_t system, const double *x, uint64_t *count, uint64_t *hash);

/* Device-resident leapfrog.  Set x, v (concatenated over systems), run
 * n_steps of LeapfrogIntegrator::run at dt_fs, read back state.  energies_every
 * > 0 evaluates energies and virial on every energies_every-th step (an
 * output frequency) and at the last step (pm_md_get_state reads the last);
 * 0: forces only.  Whole groups of 10 steps replay a captured HIP graph
 * unless PM_FLAG_TIMING is set or an energy step falls inside a group
 * (energies_every other than 10 or >= n_steps). */
int pm_md_set_state(pm_ctx *ctx, const double *x, const double *v);
int pm_md_run(pm_ctx *ctx, int32_t n_steps, double dt_fs, int32_t energies_every);
int pm_md_get_state(pm_ctx *ctx, double *x, double *v, double *f, double *e);

/* The same for a subset of systems.  pm_md_run = begin + advance + end:
 * begin makes the next kick of each listed system a half kick, advance runs
 * n_steps (kick + drift, forces) -- energies_last != 0 also sums energies and
 * virial at the last step --, end applies the closing half kick.  A system's
 * results depend only on its own calls, not on which systems share them. */
int pm_md_set_state_systems(pm_ctx *ctx, int32_t n, const int32_t *ids, const double *x, const double *v);
int pm_md_begin(pm_ctx *ctx, int32_t n, const int32_t *ids);
int pm_md_advance(pm_ctx *ctx, int32_t n, const int32_t *ids, int32_t n_steps, double dt_fs, int32_t energies_last);
int pm_md_end(pm_ctx *ctx, int32_t n, const int32_t *ids, double dt_fs);
int pm_md_get_state_systems(pm_ctx *ctx, int32_t n, const int32_t *ids, double *x, double *v, double *f, double *e);

/* LangevinLeapfrogIntegrator (integrator/base/LangevinLeapfrogIntegrator.cpp:
 * doHalfKick/doDrift, the STS step half kick, drift, forces, half kick) for
 * every later pm_md_* call of the context: temperature in K, gamma in 1/ps
 * (the reference's "gamma" parameter; 0 = plain leapfrog).  The noise of
 * atom a at the k-th half kick of its system is Philox-4x32-10 of (a, k) under
 * the system's seed, so a system's trajectory depends only on its own calls.
 * pm_system_set_seed gives one system its own seed (default: this seed). */
int pm_md_set_langevin(pm_ctx *ctx, double temperature, double gamma_per_ps, uint32_t seed);
int pm_system_set_seed(pm_ctx *ctx, int32_t id, uint32_t seed);

/* No counterpart in the reference: pm_evaluate for callers that cannot
 * allocate device memory (tight serving loops, harnesses that forbid
 * hipMalloc).  One system, no context: the force description and system as
 * for pm_ctx_create, positions x (3n AoS FP64, host); ADDS forces into f and
 * energies into e[36] as pm_evaluate does.  Every device array -- the
 * system's tables, cell and pair lists, the launch table -- is carved out of
 * a static device array of 64 MiB that each GPU holds until the process ends
 * (about 1.1 KB per atom of water at a 10 A cutoff: 60,000 atoms fit); a
 * system that needs more returns PM_ECAPACITY.  No reciprocal-space term
 * (PM_FLAG_EWALD_RECIPROCAL returns PM_EINVAL).  Runs on the null stream of
 * the current device and returns when done.  All calls on one GPU share its
 * array: calls that overlap on one GPU race (serialise them). */
int pm_evaluate_static(const pm_force_desc *force, const pm_system_desc *system, const double *x, double *f,
                       double *e);

/* The velocity start of a run, as ProtoMol draws it (host only, no device
 * work): randomNumber(seed, randomtype) seeds the generator and returns one
 * number (MainModule::configure, module/MainModule.cpp:88-102), then
 * randomVelocity(temperature, topology, velocities, seed)
 * (topology/TopologyUtilities.cpp:21-48) sets v = sqrt(kB T / m) times three
 * randomGaussianNumber deviates per atom (base/MathUtilities.cpp:206-236).
 * randomtype 0 (ProtoMol's default): the POSIX drand48 generator after
 * srand48(seed), which randomNumber uses on Linux (MathUtilities.cpp:175-183);
 * 1: ProtoMol's Random (base/Random.cpp:33-68).  v: 3n AoS FP64 (written). */
int pm_random_velocity(int32_t n, const double *mass, double temperature, uint32_t seed, int32_t randomtype,
                       double *v);

/* Statistics of the last call: stats[0] pair-list rebuilds (summed over
 * systems), [1] 1 if an MD advance was re-run after a pair-list capacity grew, [2] entries (8x8
 * cluster pairs) of the pair lists at their last build, summed over systems, [3] device ms of the last call,
 * [4] pair-kernel ms summed over the call's launches (PM_FLAG_TIMING, or the
 * single launch of an evaluation), [5] pair-kernel launches timed,
 * [6] largest displacement (A) of an atom from its position at the pair-list
 * build during the last MD call, [7] the pair-list skin (A) in use,
 * [8] reciprocal-space ms of the last evaluation (PM_FLAG_EWALD_RECIPROCAL),
 * [9] reciprocal-space k-vectors, summed over the context's systems. */
int pm_stats(pm_ctx *ctx, double *stats, int32_t n);

#ifdef __cplusplus
}
#endif
#endif
