// ref_driver.cpp -- runs ProtoMol's own force classes on a system read from a
// record file and writes forces, energies and the evaluated pair census (and,
// in mode 5, its velocity start: random number streams, randomVelocity and
// kineticEnergy).
//
// TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile against the ProtoMol
// sources where they lie (never copied); the binary lands in oracle/_ref/ and
// is used in this container to generate the golden fixtures under
// tests/golden/.  It never runs on the GPU box.
//
// Record file format (both directions): repeated
//   char name[16] (NUL padded) | int32 dtype (0 = f64, 1 = i32) | int64 count | data
#include "ref_common.hpp"

// ----------------------------------------------- census pair functor
// A force functor with the LennardJonesForce interface that records every
// pair OneAtomPair hands to it (after the cutoff test and exclusion check).
static unsigned long long g_count = 0, g_hash = 0;
static bool g_census = true;  // record "census" 0: skip the census pass (timing runs)
static unsigned long long pair_hash(unsigned lo, unsigned hi) {
  unsigned long long z = ((unsigned long long)hi << 32) | lo;
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
class CensusForce : public LennardJonesForce {
public:
  void operator()(Real &energy, Real &force, Real, Real, const Vector3D &, const GenericTopology *,
                  int a, int b, ExclusionClass) const {
    unsigned lo = (unsigned)(a < b ? a : b), hi = (unsigned)(a < b ? b : a);
    g_count++;
    g_hash += pair_hash(lo, hi);
    energy = 0;
    force = 0;
  }
  static CensusForce make(const std::vector<Value> &) { return CensusForce(); }
};

static C1SwitchingFunction mkC1(const double *s) { return C1SwitchingFunction(s[2]); }
static C2SwitchingFunction mkC2(const double *s) { return C2SwitchingFunction(s[1], s[2]); }
static CnSwitchingFunction mkCn(const double *s) { return CnSwitchingFunction(s[4], s[1], s[3], s[2]); }
static CmpCnCnSwitchingFunction mkCmp(const double *s) { return CmpCnCnSwitchingFunction(s[1], s[2], s[3], s[4]); }
static CutoffSwitchingFunction mkCut(const double *s) { return CutoffSwitchingFunction(s[2]); }

template <class Topo, class Pair>
static void run_cutoff(Topo &topo, Vector3DBlock &pos, Vector3DBlock &f, ScalarStructure &e, double cutoff,
                       const Pair &pair) {
  NonbondedCutoffSystemForce<CCM, Pair> force(cutoff, pair);
  force.preProcess(&topo, &pos);
  force.evaluate(&topo, &pos, &f, &e);
  force.postProcess(&topo, &e, &f);
  topo.uncacheCellList();
}

template <class BC, class S>
static void census_run(Topology<BC, CCM> &topo, Vector3DBlock &pos, double cutoff, const S &sw) {
  if (!g_census) return;
  typedef OneAtomPair<BC, S, CensusForce> CP;
  Vector3DBlock f(pos.size());
  f.zero();
  ScalarStructure e;
  e.clear();
  g_count = 0;
  g_hash = 0;
  run_cutoff(topo, pos, f, e, cutoff, CP(CensusForce(), sw));
}

// kinds: 1 C1, 2 C2, 3 Cn, 4 CmpCnCn, 5 Cutoff (switch vector: kind, switchon, cutoff, switchoff, order);
// BC: PeriodicBoundaryConditions (mode 1) or VacuumBoundaryConditions (mode 6,
// NonbondedCutoffForceModule.cpp:110-140)
template <class BC>
static int do_cutoff(Recs &r, Topology<BC, CCM> &topo, Vector3DBlock &pos, Vector3DBlock &f, ScalarStructure &e) {
  int ta = geti(r, "term_a", 1), tb = geti(r, "term_b", 0);
  const double *sa = &r["sw_a"].d[0];
  const double *sb = tb ? &r["sw_b"].d[0] : sa;
  int ka = (int)sa[0], kb = (int)sb[0];
  double cut = getd(r, "alg_cutoff", 10.0);
  typedef LennardJonesForce LJ;
  typedef CoulombForce CF;
  if (tb == 0) {
#define ONE(K, S, MK)                                                                               \
  if (ka == K) {                                                                                   \
    if (ta == 1) run_cutoff(topo, pos, f, e, cut, OneAtomPair<BC, S, LJ>(LJ(), MK(sa)));          \
    else run_cutoff(topo, pos, f, e, cut, OneAtomPair<BC, S, CF>(CF(), MK(sa)));                  \
    census_run(topo, pos, cut, MK(sa));                                                            \
    return 0;                                                                                      \
  }
    ONE(1, C1SwitchingFunction, mkC1)
    ONE(2, C2SwitchingFunction, mkC2)
    ONE(3, CnSwitchingFunction, mkCn)
    ONE(4, CmpCnCnSwitchingFunction, mkCmp)
    ONE(5, CutoffSwitchingFunction, mkCut)
#undef ONE
    return -1;
  }
  if (ta != 1 || tb != 2) return -1;
  // OneAtomPairTwo combinations registered in NonbondedCutoffForceModule.cpp:77-88
#define TWO(KA, KB, SA, SB, MKA, MKB)                                                              \
  if (ka == KA && kb == KB) {                                                                      \
    run_cutoff(topo, pos, f, e, cut, OneAtomPairTwo<BC, SA, LJ, SB, CF>(LJ(), MKA(sa), CF(), MKB(sb))); \
    double ca = MKA(sa).cutoff(), cb = MKB(sb).cutoff();                                           \
    if (ca >= cb) census_run(topo, pos, cut, MKA(sa));                                             \
    else census_run(topo, pos, cut, MKB(sb));                                                      \
    return 0;                                                                                      \
  }
  TWO(2, 1, C2SwitchingFunction, C1SwitchingFunction, mkC2, mkC1)
  TWO(2, 2, C2SwitchingFunction, C2SwitchingFunction, mkC2, mkC2)
  TWO(2, 3, C2SwitchingFunction, CnSwitchingFunction, mkC2, mkCn)
  TWO(3, 3, CnSwitchingFunction, CnSwitchingFunction, mkCn, mkCn)
  TWO(4, 4, CmpCnCnSwitchingFunction, CmpCnCnSwitchingFunction, mkCmp, mkCmp)
  // Cutoff switches (MODIFY = 0): not among the module's registrations, but
  // OneAtomPairTwo's own semantics for them (OneAtomPairTwo.h:66)
  TWO(5, 5, CutoffSwitchingFunction, CutoffSwitchingFunction, mkCut, mkCut)
  TWO(5, 2, CutoffSwitchingFunction, C2SwitchingFunction, mkCut, mkC2)
  TWO(2, 5, C2SwitchingFunction, CutoffSwitchingFunction, mkC2, mkCut)
#undef TWO
  return -1;
}

template <bool REAL, bool RECIP, bool CORR>
static void run_ewald(PTopo &topo, Vector3DBlock &pos, Vector3DBlock &f, ScalarStructure &e, double alpha,
                      double acc) {
  NonbondedFullEwaldSystemForce<PBC, CCM, REAL, RECIP, CORR, CutoffSwitchingFunction> force(alpha, acc, 0.0);
  force.preProcess(&topo, &pos);
  force.evaluate(&topo, &pos, &f, &e);
  force.postProcess(&topo, &e, &f);
}

static int do_ewald(Recs &r, PTopo &topo, Vector3DBlock &pos, Vector3DBlock &f, ScalarStructure &e) {
  double alpha = getd(r, "ew_alpha", 0.32), acc = getd(r, "ew_accuracy", 1e-5);
  int corr = geti(r, "ew_correction", 1);
  // terms: 1 real, 2 reciprocal, 4 correction
  int parts = geti(r, "ew_parts", corr ? 5 : 1);
  switch (parts) {
    case 1: run_ewald<true, false, false>(topo, pos, f, e, alpha, acc); break;
    case 5: run_ewald<true, false, true>(topo, pos, f, e, alpha, acc); break;
    case 2: run_ewald<false, true, false>(topo, pos, f, e, alpha, acc); break;
    case 3: run_ewald<true, true, false>(topo, pos, f, e, alpha, acc); break;
    case 7: run_ewald<true, true, true>(topo, pos, f, e, alpha, acc); break;
    default: return -1;
  }
  topo.uncacheCellList();
  // the real-term pair set: r2 <= Rc^2 and not FULL-excluded, the same test
  // OneAtomPair<PBC, Cutoff> applies (NonbondedFullEwaldSystemForce.h:808-820)
  double rc = sqrt(-log(acc)) / alpha;
  census_run(topo, pos, rc, CutoffSwitchingFunction(rc));
  return 0;
}

template <class P>
static void run_full(VTopo &topo, Vector3DBlock &pos, Vector3DBlock &f, ScalarStructure &e, const P &pair) {
  NonbondedSimpleFullSystemForce<P> force(pair);
  force.preProcess(&topo, &pos);
  force.evaluate(&topo, &pos, &f, &e);
  force.postProcess(&topo, &e, &f);
}

static int do_gb(Recs &r, Vector3DBlock &pos, Vector3DBlock &f, ScalarStructure &e, std::vector<double> &born) {
  int n = (int)pos.size();
  VTopo topo(1.0, ExclusionType(ExclusionType::ONE3), VBC(), CCM(geti(r, "cellsize", 5)), 16.0f);
  fill_common(topo, r, n);
  molecule_positions(topo, r, pos);
  const std::vector<double> &gp = r["gb_params"].d;  // offset, a, b, g, eps_in, eps_out
  fill_gb(topo, r, n);
  typedef UniversalSwitchingFunction U;
  run_full(topo, pos, f, e, OneAtomPairNoExclusion<VBC, U, GBBornRadii>(GBBornRadii(), U()));
  run_full(topo, pos, f, e, OneAtomPairNoExclusion<VBC, U, GBPartialSum>(GBPartialSum(), U()));
  run_full(topo, pos, f, e, OneAtomPairNoExclusion<VBC, U, GBForce>(GBForce(gp[4], gp[5]), U()));
  if (geti(r, "gb_ace", 0))
    run_full(topo, pos, f, e,
             OneAtomPairNoExclusion<VBC, U, GBACEForce>(GBACEForce(getd(r, "ace_sigma", 2.26 / 418.4),
                                                                   getd(r, "ace_rho_s", 1.4)), U()));
  born.resize(n);
  for (int i = 0; i < n; i++) born[i] = topo.atoms[i].myGBSA_T->bornRad;
  if (geti(r, "gb_ljcoul", 0)) {
    run_full(topo, pos, f, e, OneAtomPair<VBC, U, LennardJonesForce>(LennardJonesForce(), U()));
    run_full(topo, pos, f, e, OneAtomPair<VBC, U, CoulombForce>(CoulombForce(), U()));
  }
  return 0;
}

static int do_bonded(Recs &r, PTopo &topo, Vector3DBlock &pos, Vector3DBlock &f, ScalarStructure &e) {
  const std::vector<int> &b = r["bonds"].i;
  const std::vector<double> &bk = r["bond_k"].d, &br = r["bond_r0"].d;
  for (size_t k = 0; 2 * k + 1 < b.size(); k++) {
    Bond bd;
    bd.atom1 = b[2 * k]; bd.atom2 = b[2 * k + 1];
    bd.springConstant = bk[k]; bd.restLength = br[k];
    topo.bonds.push_back(bd);
  }
  const std::vector<int> &a = r["angles"].i;
  const std::vector<double> &ak = r["angle_k"].d, &at = r["angle_t0"].d;
  for (size_t k = 0; 3 * k + 2 < a.size(); k++) {
    Angle an;
    an.atom1 = a[3 * k]; an.atom2 = a[3 * k + 1]; an.atom3 = a[3 * k + 2];
    an.forceConstant = ak[k]; an.restAngle = at[k];
    an.ureyBradleyConstant = 0.0; an.ureyBradleyRestLength = 0.0;
    topo.angles.push_back(an);
  }
  BondSystemForce<PBC> bf;
  bf.evaluate(&topo, &pos, &f, &e);
  AngleSystemForce<PBC> af;
  af.evaluate(&topo, &pos, &f, &e);
  return 0;
}

// mode 7: a GROMACS/AMBER implicit-solvent regression input of the reference
// (test/tests/*_FFAMBER_GB_LL.conf), read by the reference's own parsers: the
// PDB positions, the .top with its force-field directory and optional GB
// parameter file (IOModule.cpp:148-170), the topology as ProtoMolApp::build
// makes it for doGBSAObc 2 (ProtoMolApp.cpp:283-313: OBC type 2, dielectric
// offset 0.09) with the run's exclusion rule, then each force of the GB group
// at those positions through NonbondedSimpleFull, one term at a time.  Writes
// the built topology (per-atom charge, mass, type, molecule, GB radius and
// scale, the LJ table, every excluded pair with its class, the harmonic bonds
// and angles) and the forces and energies of each term, the bonds and angles
// included.
static std::string gets(Recs &r, const char *k) {
  std::string s;
  Recs::iterator it = r.find(k);
  if (it == r.end()) return s;
  for (size_t i = 0; i < it->second.i.size(); i++) s += (char)it->second.i[i];
  return s;
}

template <class P>
static void term_full(VTopo &topo, Vector3DBlock &pos, const P &pair, std::vector<double> &fo, std::vector<double> &eo) {
  const int n = (int)pos.size();
  Vector3DBlock f(n);
  f.zero();
  ScalarStructure e;
  e.clear();
  e.virial(true);
  e.molecularVirial(false);
  NonbondedSimpleFullSystemForce<P> force(pair);
  force.preProcess(&topo, &pos);
  force.evaluate(&topo, &pos, &f, &e);
  force.postProcess(&topo, &e, &f);
  fo.resize(3 * (size_t)n);
  for (int i = 0; i < n; i++) { fo[3 * i] = f[i].c[0]; fo[3 * i + 1] = f[i].c[1]; fo[3 * i + 2] = f[i].c[2]; }
  eo.resize(ScalarStructure::LAST);
  for (int s = 0; s < ScalarStructure::LAST; s++) eo[s] = e[(ScalarStructure::Index)s];
}

static int do_gromacs_gb(Recs &r, const char *out) {
  PSF psf;
  PAR par;
  GromacsTopology gTopo;
  GromacsParameters gParams;
  PortGromacsParameters port;
  if (!port.Read_Basic_Gromacs_Parameters(psf, par, gTopo, gParams, gets(r, "path_top"), gets(r, "path_ff"))) return -1;
  const std::string gbfile = gets(r, "path_gb");
  if (!gbfile.empty() && !port.Read_Gromacs_GB_Parameters(gbfile)) return -1;
  PDBReader reader(gets(r, "path_pdb"));
  PDB pdb;
  if (!reader.read(pdb)) return -1;
  Vector3DBlock pos = pdb.coords;
  const int n = (int)pos.size();
  // "exclude scaled1-4" (the regressions' setting) is ONE4MODIFIED
  VTopo topo(1.0, ExclusionType(ExclusionType::ONE4MODIFIED), VBC(), CCM(getd(r, "cellsize", 5.0)), 16.0f);
  topo.doGBSAOpenMM = 1;
  topo.obcType = 2;
  topo.alphaObc = 1.0;
  topo.betaObc = 0.8;
  topo.gammaObc = 4.85;
  topo.dielecOffset = 0.09;
  topo.forceFieldFlag = GROMACS;
  buildTopology(&topo, psf, par, false, NULL);
  if ((int)topo.atoms.size() != n) return -1;
  typedef UniversalSwitchingFunction U;
  std::vector<double> f_gb, e_gb, f_lj, e_lj, f_cl, e_cl, tmpf, tmpe;
  // the GB chain in its force-group order; the first two passes only fill the
  // Born radii and partial sums (no force)
  term_full(topo, pos, OneAtomPairNoExclusion<VBC, U, GBBornRadii>(GBBornRadii(), U()), tmpf, tmpe);
  term_full(topo, pos, OneAtomPairNoExclusion<VBC, U, GBPartialSum>(GBPartialSum(), U()), tmpf, tmpe);
  term_full(topo, pos, OneAtomPairNoExclusion<VBC, U, GBForce>(GBForce(getd(r, "eps_in", 1.0), getd(r, "eps_out", 78.3)), U()),
            f_gb, e_gb);
  std::vector<double> f_ace, e_ace;
  term_full(topo, pos, OneAtomPairNoExclusion<VBC, U, GBACEForce>(GBACEForce(2.26 / 418.4, 1.4), U()), f_ace, e_ace);
  term_full(topo, pos, OneAtomPair<VBC, U, LennardJonesForce>(LennardJonesForce(), U()), f_lj, e_lj);
  term_full(topo, pos, OneAtomPair<VBC, U, CoulombForce>(CoulombForce(), U()), f_cl, e_cl);
  // the built topology
  std::vector<double> x(3 * (size_t)n), q(n), m(n), type(n), mol(n), rad(n), scl(n), off(n), born(n);
  for (int i = 0; i < n; i++) {
    for (int d = 0; d < 3; d++) x[3 * i + d] = pos[i].c[d];
    const Atom &a = topo.atoms[i];
    q[i] = a.scaledCharge;
    m[i] = a.scaledMass;
    type[i] = a.type;
    mol[i] = a.molecule;
    rad[i] = a.myGBSA_T->vanDerWaalRadius;
    scl[i] = a.myGBSA_T->scalingFactor;
    off[i] = a.myGBSA_T->offsetRadius;
    born[i] = a.myGBSA_T->bornRad;
  }
  const int T = (int)topo.atomTypes.size();
  std::vector<double> lj(4 * (size_t)T * T);
  for (int a = 0; a < T; a++)
    for (int b = 0; b < T; b++) {
      const LennardJonesParameters &p = topo.lennardJonesParameters(a, b);
      double *o = &lj[4 * ((size_t)a * T + b)];
      o[0] = p.A; o[1] = p.B; o[2] = p.A14; o[3] = p.B14;
    }
  std::vector<double> excl;
  for (int a = 0; a < n; a++)
    for (int b = a + 1; b < n; b++) {
      const ExclusionClass c = topo.exclusions.check(a, b);
      if (c == EXCLUSION_NONE) continue;
      excl.push_back(a);
      excl.push_back(b);
      excl.push_back(c == EXCLUSION_MODIFIED ? 2 : 1);
    }
  // the harmonic bonds and angles of the topology, and their forces
  // (BondSystemForce, AngleSystemForce; the dihedral terms stay out)
  std::vector<double> bonds, angles, f_bd, e_bd;
  for (size_t k = 0; k < topo.bonds.size(); k++) {
    const Bond &b = topo.bonds[k];
    bonds.push_back(b.atom1); bonds.push_back(b.atom2);
    bonds.push_back(b.springConstant); bonds.push_back(b.restLength);
  }
  for (size_t k = 0; k < topo.angles.size(); k++) {
    const Angle &a = topo.angles[k];
    angles.push_back(a.atom1); angles.push_back(a.atom2); angles.push_back(a.atom3);
    angles.push_back(a.forceConstant); angles.push_back(a.restAngle);
    angles.push_back(a.ureyBradleyConstant); angles.push_back(a.ureyBradleyRestLength);
  }
  {
    Vector3DBlock fb(n);
    fb.zero();
    ScalarStructure eb;
    eb.clear();
    eb.virial(true);
    eb.molecularVirial(false);
    BondSystemForce<VBC> bf;
    bf.evaluate(&topo, &pos, &fb, &eb);
    AngleSystemForce<VBC> af;
    af.evaluate(&topo, &pos, &fb, &eb);
    f_bd.resize(3 * (size_t)n);
    for (int i = 0; i < n; i++) { f_bd[3 * i] = fb[i].c[0]; f_bd[3 * i + 1] = fb[i].c[1]; f_bd[3 * i + 2] = fb[i].c[2]; }
    e_bd.resize(ScalarStructure::LAST);
    for (int q2 = 0; q2 < ScalarStructure::LAST; q2++) e_bd[q2] = eb[(ScalarStructure::Index)q2];
  }
  double scal[4] = {topo.coulombScalingFactor, topo.LJScalingFactor, (double)T, (double)n};
  FILE *fp = fopen(out, "wb");
  if (!fp) return -2;
  write_rec_d(fp, "x", &x[0], 3LL * n);
  write_rec_d(fp, "charge", &q[0], n);
  write_rec_d(fp, "mass", &m[0], n);
  write_rec_d(fp, "type", &type[0], n);
  write_rec_d(fp, "molecule", &mol[0], n);
  write_rec_d(fp, "gb_radius", &rad[0], n);
  write_rec_d(fp, "gb_scaling", &scl[0], n);
  write_rec_d(fp, "gb_offset", &off[0], n);
  write_rec_d(fp, "born", &born[0], n);
  write_rec_d(fp, "lj", &lj[0], (long long)lj.size());
  write_rec_d(fp, "excl", excl.empty() ? NULL : &excl[0], (long long)excl.size());
  write_rec_d(fp, "scalars", scal, 4);
  write_rec_d(fp, "f_gb", &f_gb[0], 3LL * n);
  write_rec_d(fp, "e_gb", &e_gb[0], (long long)e_gb.size());
  write_rec_d(fp, "f_ace", &f_ace[0], 3LL * n);
  write_rec_d(fp, "e_ace", &e_ace[0], (long long)e_ace.size());
  write_rec_d(fp, "f_lj", &f_lj[0], 3LL * n);
  write_rec_d(fp, "e_lj", &e_lj[0], (long long)e_lj.size());
  write_rec_d(fp, "f_coul", &f_cl[0], 3LL * n);
  write_rec_d(fp, "e_coul", &e_cl[0], (long long)e_cl.size());
  write_rec_d(fp, "bonds", bonds.empty() ? NULL : &bonds[0], (long long)bonds.size());
  write_rec_d(fp, "angles", angles.empty() ? NULL : &angles[0], (long long)angles.size());
  write_rec_d(fp, "f_bonded", &f_bd[0], 3LL * n);
  write_rec_d(fp, "e_bonded", &e_bd[0], (long long)e_bd.size());
  fclose(fp);
  return 0;
}

// mode 5: the velocity start of a run as MainModule draws it.  configure()
// seeds the generator with randomNumber(seed, randomtype), which also returns
// one number (MainModule.cpp:88-102); postBuild() then calls
// randomVelocity(temperature, topology, velocities, seed)
// (MainModule.cpp:143-147, TopologyUtilities.cpp:21-48).  Outputs: that first
// number, the velocities, kineticEnergy and temperature of them, and `ndraw`
// further randomNumber and then randomGaussianNumber values of the streams.
static int do_velocity(Recs &r, int n, Vector3DBlock &v, std::vector<double> &out_u0, std::vector<double> &out_ke,
                       std::vector<double> &out_u, std::vector<double> &out_g) {
  VTopo topo(1.0, ExclusionType(ExclusionType::ONE3), VBC(), CCM(5.0), 16.0f);
  topo.atoms.resize(n);
  for (int i = 0; i < n; i++) topo.atoms[i].scaledMass = r["mass"].d[i];
  topo.degreesOfFreedom = 3 * n - 3;  // BuildTopology.cpp:121
  const unsigned seed = (unsigned)geti(r, "seed", 1234);
  const unsigned rtype = (unsigned)geti(r, "randomtype", 0);
  out_u0.push_back(randomNumber(seed, rtype));
  randomVelocity(getd(r, "temperature", 300.0), &topo, &v, seed);
  out_ke.push_back(kineticEnergy(&topo, &v));
  out_ke.push_back(temperature(&topo, &v));
  const int nd = geti(r, "ndraw", 0);
  for (int k = 0; k < nd; k++) out_u.push_back(randomNumber(seed));
  for (int k = 0; k < nd; k++) out_g.push_back(randomGaussianNumber(seed));
  return 0;
}

int main(int argc, char **argv) {
  if (argc != 3) { fprintf(stderr, "usage: ref_driver in.rec out.rec\n"); return 2; }
  Report::report << Report::donthint;
  Recs r;
  if (!read_recs(argv[1], r)) { fprintf(stderr, "cannot read %s\n", argv[1]); return 2; }
  int n = geti(r, "n", 0);
  int mode = geti(r, "mode", 1);
  if (mode == 7) {
    const int rc7 = do_gromacs_gb(r, argv[2]);
    if (rc7) fprintf(stderr, "cannot build or evaluate the GROMACS GB input (%d)\n", rc7);
    return rc7 ? 3 : 0;
  }
  Vector3DBlock pos(n), f(n);
  for (int i = 0; i < n; i++) pos[i] = Vector3D(r["x"].d[3 * i], r["x"].d[3 * i + 1], r["x"].d[3 * i + 2]);
  f.zero();
  ScalarStructure e;
  e.clear();
  e.virial(geti(r, "virial", 0) != 0);
  // the molecular virial is on by default in ScalarStructure; the record says
  e.molecularVirial(geti(r, "molvirial", 0) != 0);
  std::vector<double> born, u0, ke, uu, gg;
  int rc = -1;
  // "repeat" > 1 evaluates the force that many times (the forces add up) and
  // "seconds" reports the wall time of those evaluations alone: the
  // reference's own CPU path, timed (single-threaded, as it runs)
  const int repeat = geti(r, "repeat", 1);
  g_census = geti(r, "census", 1) != 0;
  double seconds = 0.0;
  if (mode == 3) {
    timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    rc = do_gb(r, pos, f, e, born);
    clock_gettime(CLOCK_MONOTONIC, &t1);
    seconds = (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
  } else if (mode == 5) rc = do_velocity(r, n, f, u0, ke, uu, gg);
  else if (mode == 6) {
    VTopo topo(getd(r, "csf", 1.0), ExclusionType(ExclusionType::ONE3), VBC(), CCM(getd(r, "cellsize", 5.0)), 16.0f);
    fill_common(topo, r, n);
    molecule_positions(topo, r, pos);
    rc = do_cutoff(r, topo, pos, f, e);
  } else {
    const std::vector<double> &box = r["box"].d, &org = r["origin"].d;
    PBC bc(Vector3D(box[0], 0, 0), Vector3D(0, box[1], 0), Vector3D(0, 0, box[2]), Vector3D(org[0], org[1], org[2]));
    PTopo topo(getd(r, "csf", 1.0), ExclusionType(ExclusionType::ONE3), bc, CCM(getd(r, "cellsize", 5.0)), 16.0f);
    fill_common(topo, r, n);
    molecule_positions(topo, r, pos);
    timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    for (int k = 0; k < repeat; k++) {
      if (mode == 1) rc = do_cutoff(r, topo, pos, f, e);
      else if (mode == 2) rc = do_ewald(r, topo, pos, f, e);
      else if (mode == 4) rc = do_bonded(r, topo, pos, f, e);
      if (rc) break;
    }
    clock_gettime(CLOCK_MONOTONIC, &t1);
    seconds = (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
  }
  if (rc) { fprintf(stderr, "unsupported configuration\n"); return 3; }
  FILE *fp = fopen(argv[2], "wb");
  if (!fp) return 2;
  std::vector<double> fo(3 * (size_t)n), eo(ScalarStructure::LAST);
  for (int i = 0; i < n; i++) { fo[3 * i] = f[i].c[0]; fo[3 * i + 1] = f[i].c[1]; fo[3 * i + 2] = f[i].c[2]; }
  for (int s = 0; s < ScalarStructure::LAST; s++) eo[s] = e[(ScalarStructure::Index)s];
  write_rec_d(fp, "f", &fo[0], 3LL * n);
  write_rec_d(fp, "e", &eo[0], (long long)eo.size());
  double cen[2] = {(double)(g_count), 0.0};
  write_rec_d(fp, "census_count", cen, 1);
  // the 64-bit hash travels as two 32-bit halves stored exactly in doubles
  double hh[2] = {(double)(g_hash >> 32), (double)(g_hash & 0xffffffffull)};
  write_rec_d(fp, "census_hash", hh, 2);
  if (!born.empty()) write_rec_d(fp, "born", &born[0], (long long)born.size());
  write_rec_d(fp, "seconds", &seconds, 1);
  if (mode == 5) {
    write_rec_d(fp, "u0", &u0[0], 1);
    write_rec_d(fp, "ke_t", &ke[0], 2);
    write_rec_d(fp, "u", uu.empty() ? NULL : &uu[0], (long long)uu.size());
    write_rec_d(fp, "g", gg.empty() ? NULL : &gg[0], (long long)gg.size());
  }
  fclose(fp);
  return 0;
}
