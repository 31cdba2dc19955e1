// adapter_check.cpp -- the SystemForce adapters (protomol_amd_systemforce.hpp)
// made through ProtoMol's own ForceFactory from configuration strings, and
// evaluated beside ProtoMol's own forces made from the same strings (only the
// algorithm keyword differs), on a topology built with ProtoMol's classes.
//
// TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile (target adapter)
// against the ProtoMol sources where they lie: adapter_check_stub links a
// recording stand-in of the C ABI (pm_abi_stub.cpp), adapter_check links
// libprotomol_amd.so.
//
// usage: adapter_check in.rec out.rec
//   in.rec   the system records of ref_driver; mode 1: a periodic system and
//            the cutoff / Ewald cases, mode 3: a vacuum GB chain and the
//            all-pairs force group, mode 6: a vacuum system and the cutoff
//            cases, each at the given positions and again with the molecule
//            of the atom of largest x moved `exit` A further out in x
//   out.rec  per case k: f_ref<k>, e_ref<k> (ProtoMol's force) and
//            f_hip<k>, e_hip<k> (the adapter), and "cases" (their count)
#include "ref_common.hpp"

#include <protomol/factory/ForceFactory.h>

#include "protomol_amd_systemforce.hpp"

typedef LennardJonesForce LJ;
typedef CoulombForce CF;
typedef C1SwitchingFunction C1;
typedef C2SwitchingFunction C2;
typedef CnSwitchingFunction Cn;
typedef CutoffSwitchingFunction Cut;
typedef UniversalSwitchingFunction U;

// the periodic cases: (config with %s for the algorithm, reference algorithm, adapter algorithm)
static const char *const kPeriodic[][3] = {
    {"LennardJones Coulomb -algorithm %s -switchingFunction C2 -switchingFunction C1 -switchon 0.1 -cutoff 9 "
     "-cutoff 9 -cutoff 9", "NonbondedCutoff", "NonbondedCutoffHIP"},
    {"LennardJones Coulomb -algorithm %s -switchingFunction C2 -switchingFunction C2 -switchon 8 -cutoff 10 "
     "-switchon 8 -cutoff 10 -cutoff 10", "NonbondedCutoff", "NonbondedCutoffHIP"},
    {"LennardJones -algorithm %s -switchingFunction C2 -switchon 6 -cutoff 8 -cutoff 8", "NonbondedCutoff",
     "NonbondedCutoffHIP"},
    {"Coulomb -algorithm %s -switchingFunction C1 -cutoff 10 -cutoff 10", "NonbondedCutoff", "NonbondedCutoffHIP"},
    {"LennardJones Coulomb -algorithm %s -switchingFunction Cn -switchingFunction Cn -n 6 -switchon 8 "
     "-switchoff 0 -cutoff 10 -n 6 -switchon 8 -switchoff 0 -cutoff 10 -cutoff 10", "NonbondedCutoff",
     "NonbondedCutoffHIP"},
    {"Coulomb -algorithm %s -real -correction -alpha 0.32 -accuracy 3.5713e-05", "FullEwald", "FullEwaldHIP"},
    {"Coulomb -algorithm %s -real -reciprocal -correction -alpha 0.32 -accuracy 3.5713e-05", "FullEwald",
     "FullEwaldHIP"},
};

// the vacuum cutoff cases (NonbondedCutoffForceModule.cpp:110-140)
static const char *const kVacuum[] = {
    "LennardJones Coulomb -algorithm %s -switchingFunction C2 -switchingFunction C2 -switchon 6 -cutoff 8 "
    "-switchon 8 -cutoff 10 -cutoff 10",
    "Coulomb -algorithm %s -switchingFunction C1 -cutoff 10 -cutoff 10"};

// the all-pairs lines of the reference's GB regression tests (e.g.
// test/tests/www-fip35_FFAMBER_GB_LL.conf), one force group
static const char *const kFullGroup[] = {
    "LennardJones -algorithm %s", "Coulomb -algorithm %s", "GBBornRadii -algorithm %s",
    "GBPartialSum -algorithm %s", "GBForce -algorithm %s -soluteDielec 1.0 -solventDielec 78.3",
    "GBACEForce -algorithm %s"};

static std::string line(const char *fmt, const char *alg) {
  char buf[512];
  snprintf(buf, sizeof buf, fmt, alg);
  return buf;
}

static void register_periodic(ForceFactory &ff) {
  ff.reg(new NonbondedCutoffSystemForce<CCM, OneAtomPairTwo<PBC, C2, LJ, C1, CF> >());
  ff.reg(new NonbondedCutoffSystemForceHIP<OneAtomPairTwo<PBC, C2, LJ, C1, CF> >());
  ff.reg(new NonbondedCutoffSystemForce<CCM, OneAtomPairTwo<PBC, C2, LJ, C2, CF> >());
  ff.reg(new NonbondedCutoffSystemForceHIP<OneAtomPairTwo<PBC, C2, LJ, C2, CF> >());
  ff.reg(new NonbondedCutoffSystemForce<CCM, OneAtomPairTwo<PBC, Cn, LJ, Cn, CF> >());
  ff.reg(new NonbondedCutoffSystemForceHIP<OneAtomPairTwo<PBC, Cn, LJ, Cn, CF> >());
  ff.reg(new NonbondedCutoffSystemForce<CCM, OneAtomPair<PBC, C2, LJ> >());
  ff.reg(new NonbondedCutoffSystemForceHIP<OneAtomPair<PBC, C2, LJ> >());
  ff.reg(new NonbondedCutoffSystemForce<CCM, OneAtomPair<PBC, C1, CF> >());
  ff.reg(new NonbondedCutoffSystemForceHIP<OneAtomPair<PBC, C1, CF> >());
  ff.reg(new NonbondedFullEwaldSystemForce<PBC, CCM, true, false, true, Cut>());
  ff.reg(new NonbondedFullEwaldSystemForceHIP<true, false, true>());
  ff.reg(new NonbondedFullEwaldSystemForce<PBC, CCM, true, true, true, Cut>());
  ff.reg(new NonbondedFullEwaldSystemForceHIP<true, true, true>());
}

static void register_vacuum(ForceFactory &ff) {
  ff.reg(new NonbondedCutoffSystemForce<CCM, OneAtomPairTwo<VBC, C2, LJ, C2, CF> >());
  ff.reg(new NonbondedCutoffSystemForceHIP<OneAtomPairTwo<VBC, C2, LJ, C2, CF> >());
  ff.reg(new NonbondedCutoffSystemForce<CCM, OneAtomPair<VBC, C1, CF> >());
  ff.reg(new NonbondedCutoffSystemForceHIP<OneAtomPair<VBC, C1, CF> >());
}

static void register_full(ForceFactory &ff) {
  ff.reg(new NonbondedSimpleFullSystemForce<OneAtomPair<VBC, U, LJ> >());
  ff.reg(new NonbondedSimpleFullSystemForceHIP<LJ>());
  ff.reg(new NonbondedSimpleFullSystemForce<OneAtomPair<VBC, U, CF> >());
  ff.reg(new NonbondedSimpleFullSystemForceHIP<CF>());
  ff.reg(new NonbondedSimpleFullSystemForce<OneAtomPairNoExclusion<VBC, U, GBBornRadii> >());
  ff.reg(new NonbondedSimpleFullSystemForceHIP<GBBornRadii>());
  ff.reg(new NonbondedSimpleFullSystemForce<OneAtomPairNoExclusion<VBC, U, GBPartialSum> >());
  ff.reg(new NonbondedSimpleFullSystemForceHIP<GBPartialSum>());
  ff.reg(new NonbondedSimpleFullSystemForce<OneAtomPairNoExclusion<VBC, U, GBForce> >());
  ff.reg(new NonbondedSimpleFullSystemForceHIP<GBForce>());
  ff.reg(new NonbondedSimpleFullSystemForce<OneAtomPairNoExclusion<VBC, U, GBACEForce> >());
  ff.reg(new NonbondedSimpleFullSystemForceHIP<GBACEForce>());
}

// ForceGroup::evaluateSystemForces (force/ForceGroup.cpp:54-72) over the
// listed forces: uncache the cell list, then preProcess, evaluate and
// postProcess of each in order
template <class T>
static void evaluate_group(T &topo, const std::vector<SystemForce *> &fs, Vector3DBlock &pos, Vector3DBlock &f,
                           ScalarStructure &e) {
  topo.uncacheCellList();
  for (size_t k = 0; k < fs.size(); k++) {
    fs[k]->preProcess(&topo, &pos);
    fs[k]->evaluate(&topo, &pos, &f, &e);
    fs[k]->postProcess(&topo, &e, &f);
  }
}

static void put_result(FILE *fp, const char *tag, int k, const Vector3DBlock &f, const ScalarStructure &e) {
  char name[16];
  const int n = (int)f.size();
  std::vector<double> fo(3 * (size_t)n), eo(ScalarStructure::LAST);
  for (int i = 0; i < n; i++)
    for (int d = 0; d < 3; d++) fo[3 * i + d] = f[i].c[d];
  for (int s = 0; s < ScalarStructure::LAST; s++) eo[s] = e[(ScalarStructure::Index)s];
  snprintf(name, sizeof name, "f_%s%d", tag, k);
  write_rec_d(fp, name, n ? &fo[0] : 0, 3LL * n);
  snprintf(name, sizeof name, "e_%s%d", tag, k);
  write_rec_d(fp, name, &eo[0], (long long)eo.size());
}

// one side (reference algorithm or adapter) of a set of lines evaluated as
// one force group, at each of the positions in turn (the same force objects)
template <class T>
static void run_side(ForceFactory &ff, T &topo, const std::vector<std::string> &lines,
                     const std::vector<Vector3DBlock *> &positions, FILE *out, const char *tag, int k) {
  ForceGroup group;  // owns the forces
  std::vector<SystemForce *> fs;
  for (size_t i = 0; i < lines.size(); i++) {
    Force *f = ff.make(lines[i]);
    SystemForce *sf = dynamic_cast<SystemForce *>(f);
    if (!sf) THROW("not a system force: " + lines[i]);
    group.addForce(f);
    fs.push_back(sf);
  }
  for (size_t p = 0; p < positions.size(); p++) {
    Vector3DBlock f(positions[p]->size());
    f.zero();
    ScalarStructure e;
    e.clear();
    e.virial(true);
    e.molecularVirial(false);
    evaluate_group(topo, fs, *positions[p], f, e);
    put_result(out, tag, k + (int)p, f, e);
  }
}

int main(int argc, char **argv) {
  if (argc != 3) {
    fprintf(stderr, "usage: adapter_check in.rec out.rec\n");
    return 2;
  }
  Report::report << Report::donthint;
  Recs r;
  if (!read_recs(argv[1], r)) {
    fprintf(stderr, "cannot read %s\n", argv[1]);
    return 2;
  }
  const int n = geti(r, "n", 0), mode = geti(r, "mode", 1);
  Vector3DBlock pos(n);
  for (int i = 0; i < n; i++) pos[i] = Vector3D(r["x"].d[3 * i], r["x"].d[3 * i + 1], r["x"].d[3 * i + 2]);
  FILE *out = fopen(argv[2], "wb");
  if (!out) return 2;
  ForceFactory ff;
  int cases = 0;
  try {
    if (mode == 1) {
      register_periodic(ff);
      const std::vector<double> &box = r["box"].d, &org = r["origin"].d;
      PBC bc(Vector3D(box[0], 0, 0), Vector3D(0, box[1], 0), Vector3D(0, 0, box[2]), Vector3D(org[0], org[1], org[2]));
      PTopo topo(getd(r, "csf", 1.0), ExclusionType(ExclusionType::ONE3), bc, CCM(getd(r, "cellsize", 5.0)), 16.0f);
      fill_common(topo, r, n);
      const std::vector<Vector3DBlock *> at(1, &pos);
      for (size_t k = 0; k < sizeof kPeriodic / sizeof kPeriodic[0]; k++, cases++) {
        run_side(ff, topo, std::vector<std::string>(1, line(kPeriodic[k][0], kPeriodic[k][1])), at, out, "ref",
                 (int)k);
        run_side(ff, topo, std::vector<std::string>(1, line(kPeriodic[k][0], kPeriodic[k][2])), at, out, "hip",
                 (int)k);
      }
    } else if (mode == 3) {
      register_full(ff);
      VTopo topo(1.0, ExclusionType(ExclusionType::ONE3), VBC(), CCM(geti(r, "cellsize", 5)), 16.0f);
      fill_common(topo, r, n);
      fill_gb(topo, r, n);
      std::vector<std::string> ref, hip;
      for (size_t k = 0; k < sizeof kFullGroup / sizeof kFullGroup[0]; k++) {
        ref.push_back(line(kFullGroup[k], "NonbondedSimpleFull"));
        hip.push_back(line(kFullGroup[k], "NonbondedSimpleFullHIP"));
      }
      const std::vector<Vector3DBlock *> at(1, &pos);
      run_side(ff, topo, ref, at, out, "ref", 0);
      run_side(ff, topo, hip, at, out, "hip", 0);
      cases = 1;
    } else if (mode == 6) {
      register_vacuum(ff);
      VTopo topo(getd(r, "csf", 1.0), ExclusionType(ExclusionType::ONE3), VBC(), CCM(getd(r, "cellsize", 5.0)), 16.0f);
      fill_common(topo, r, n);
      int far = 0;
      for (int i = 1; i < n; i++)
        if (pos[i].c[0] > pos[far].c[0]) far = i;
      Vector3DBlock moved(pos);
      for (int i = 0; i < n; i++)
        if (topo.atoms[i].molecule == topo.atoms[far].molecule) moved[i].c[0] += getd(r, "exit", 15.0);
      std::vector<Vector3DBlock *> at;
      at.push_back(&pos);
      at.push_back(&moved);
      for (size_t k = 0; k < sizeof kVacuum / sizeof kVacuum[0]; k++, cases += 2) {
        run_side(ff, topo, std::vector<std::string>(1, line(kVacuum[k], "NonbondedCutoff")), at, out, "ref", cases);
        run_side(ff, topo, std::vector<std::string>(1, line(kVacuum[k], "NonbondedCutoffHIP")), at, out, "hip",
                 cases);
      }
    }
  } catch (const Exception &e) {
    fprintf(stderr, "adapter_check: %s\n", e.getMessage().c_str());
    fclose(out);
    return 3;
  }
  const double c = cases;
  write_rec_d(out, "cases", &c, 1);
  fclose(out);
  return 0;
}
