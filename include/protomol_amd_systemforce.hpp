// protomol_amd_systemforce.hpp -- ProtoMol SystemForce adapters over the C ABI.
//
// Header-only; compiled inside a ProtoMol tree (it includes ProtoMol's own
// headers).  Each class template is a drop-in prototype for a force the
// reference registers, with the reference's own id, parameter list and
// make() semantics, and its own algorithm keyword:
//
//   NonbondedCutoffSystemForceHIP<TOneAtomPair>
//       for NonbondedCutoffSystemForce<CubicCellManager, TOneAtomPair>
//       (force/nonbonded/NonbondedCutoffSystemForce.h:45-55,
//       NonbondedCutoffForce.h:70-89), TOneAtomPair = OneAtomPair<BC, S,
//       LennardJonesForce | CoulombForce> or OneAtomPairTwo<BC, SA,
//       LennardJonesForce, SB, CoulombForce>, BC periodic or vacuum
//       (module/NonbondedCutoffForceModule.cpp:56-88, 110-140); keyword
//       "NonbondedCutoffHIP".  A vacuum system keeps its atoms in a region
//       (origin +- box / 2) made from the positions of the first evaluation,
//       grown by kPmVacuumMargin on every side; an atom that leaves it makes
//       the evaluation build the device context again around the positions
//       of that time
//   NonbondedFullEwaldSystemForceHIP<real, reciprocal, correction>
//       for NonbondedFullEwaldSystemForce<PBC, Cubic, real, reciprocal,
//       correction, Cutoff> (force/nonbonded/NonbondedFullEwaldSystemForce.h:
//       116-150, 1205-1236; module/NonbondedFullElectrostaticForceModule.cpp:31-36)
//       with real on; keyword "FullEwaldHIP"
//   NonbondedSimpleFullSystemForceHIP<TForce>
//       for NonbondedSimpleFullSystemForce<OneAtomPair[NoExclusion]<VBC,
//       Universal, TForce>> (module/NonbondedSimpleFullForceModule.cpp:203-206),
//       TForce = GBBornRadii, GBPartialSum, GBForce, GBACEForce,
//       LennardJonesForce, CoulombForce; keyword "NonbondedSimpleFullHIP"
//
// A configuration line selects one by its algorithm, e.g.
//   force LennardJones Coulomb -algorithm NonbondedCutoffHIP
//         -switchingFunction C2 -switchingFunction C1
//         -switchon 0.1 -cutoff 6.5 -cutoff 6.5 -cutoff 6.5
// and ForceFactory::make parses the parameters into the prototype's
// doMake exactly as for the reference's force.
//
// evaluate() keeps the reference's semantics: it ADDS into forces and into
// the energy/virial slots of the ScalarStructure; a non-zero status from the
// C ABI becomes ProtoMol's THROW (base/Exception.h:258).
//
// The all-pairs forces of one ForceGroup (the GB passes, ACE, LJ and
// Coulomb) are one device launch sequence: the first of them added to the
// group evaluates the terms of all of them, the others add nothing.
#ifndef PROTOMOL_AMD_SYSTEMFORCE_HPP
#define PROTOMOL_AMD_SYSTEMFORCE_HPP

#include <protomol/base/Exception.h>
#include <protomol/config/Parameter.h>
#include <protomol/force/CoulombForce.h>
#include <protomol/force/ForceGroup.h>
#include <protomol/force/GB/GBACEForce.h>
#include <protomol/force/GB/GBBornRadii.h>
#include <protomol/force/GB/GBForce.h>
#include <protomol/force/GB/GBPartialSum.h>
#include <protomol/force/LennardJonesForce.h>
#include <protomol/force/OneAtomPair.h>
#include <protomol/force/OneAtomPairTwo.h>
#include <protomol/force/system/SystemForce.h>
#include <protomol/switch/CutoffSwitchingFunction.h>
#include <protomol/topology/GenericTopology.h>
#include <protomol/topology/PeriodicBoundaryConditions.h>
#include <protomol/topology/SemiGenericTopology.h>
#include <protomol/topology/VacuumBoundaryConditions.h>
#include <protomol/type/ScalarStructure.h>
#include <protomol/type/Vector3DBlock.h>

#include <cmath>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "protomol_amd.h"

namespace ProtoMol {

  // Reads what the nonbonded path needs out of a ProtoMol topology: scaled
  // charges, masses, LJ types and table (any number of types), molecules,
  // the exclusion table (in ExclusionTable::add order), the box (periodic,
  // orthogonal) or vacuum, and with `gb` the atoms' GBSA parameters.
  class PmTopologyImage {
  public:
    std::vector<double> charge, mass, lj, gb_radius, gb_scaling;
    std::vector<int32_t> type, molecule, excl;
    pm_system_desc desc;

    PmTopologyImage(const GenericTopology *topo, bool gb) {
      const int n = (int)topo->atoms.size();
      charge.resize(n); mass.resize(n); type.resize(n); molecule.resize(n);
      const int ntypes = (int)topo->atomTypes.size() > 0 ? (int)topo->atomTypes.size() : 1;
      for (int i = 0; i < n; i++) {
        charge[i] = topo->atoms[i].scaledCharge;
        mass[i] = topo->atoms[i].scaledMass;
        type[i] = topo->atoms[i].type;
        molecule[i] = topo->atoms[i].molecule;
        if (type[i] < 0 || type[i] >= ntypes) THROW("protomol_amd: atom type outside the topology's atom types");
      }
      // LennardJonesParameterTable, sized by the topology's atom types
      // (BuildTopology.cpp:582-586, LennardJonesParameterTable.h:34-36)
      lj.resize(4 * (size_t)ntypes * ntypes);
      for (int a = 0; a < ntypes; a++)
        for (int b = 0; b < ntypes; b++) {
          const LennardJonesParameters &p = topo->lennardJonesParameters(a, b);
          double *q = &lj[4 * ((size_t)a * ntypes + b)];
          q[0] = p.A; q[1] = p.B; q[2] = p.A14; q[3] = p.B14;
        }
      const std::vector<ExclusionPair> &t = topo->exclusions.getTable();
      for (size_t k = 0; k < t.size(); k++) {
        excl.push_back(t[k].a1);
        excl.push_back(t[k].a2);
        excl.push_back(t[k].excl == EXCLUSION_MODIFIED ? PM_EXCL_MODIFIED : PM_EXCL_FULL);
      }
      desc = pm_system_desc();
      desc.n_atoms = n;
      desc.scaled_charge = n ? &charge[0] : 0;
      desc.mass = n ? &mass[0] : 0;
      desc.type = n ? &type[0] : 0;
      desc.molecule = n ? &molecule[0] : 0;
      desc.lj_mode = PM_LJ_TABLE;
      desc.n_types = ntypes;
      desc.lj_table = &lj[0];
      desc.n_exclusions = (int32_t)(excl.size() / 3);
      desc.exclusions = excl.empty() ? 0 : &excl[0];
      desc.coulomb_scaling = topo->coulombScalingFactor;
      const SemiGenericTopology<PeriodicBoundaryConditions> *pt =
          dynamic_cast<const SemiGenericTopology<PeriodicBoundaryConditions> *>(topo);
      if (pt) {
        const PeriodicBoundaryConditions &bc = pt->boundaryConditions;
        if (!bc.isOrthogonal()) THROW("protomol_amd: orthogonal boxes only");
        desc.periodic = 1;
        desc.box[0] = bc.e1().c[0]; desc.box[1] = bc.e2().c[1]; desc.box[2] = bc.e3().c[2];
        desc.origin[0] = bc.origin().c[0]; desc.origin[1] = bc.origin().c[1]; desc.origin[2] = bc.origin().c[2];
      } else if (!dynamic_cast<const SemiGenericTopology<VacuumBoundaryConditions> *>(topo)) {
        THROW("protomol_amd: periodic or vacuum boundary conditions only");
      }
      if (gb) {
        gb_radius.resize(n);
        gb_scaling.resize(n);
        for (int i = 0; i < n; i++) {
          const GBSAAtomParameters *g = topo->atoms[i].myGBSA_T;
          if (!g) THROW("protomol_amd: GB forces need the atoms' GBSA parameters");
          gb_radius[i] = g->vanDerWaalRadius;
          gb_scaling[i] = g->scalingFactor;
        }
        desc.gb_radius = n ? &gb_radius[0] : 0;
        desc.gb_scaling = n ? &gb_scaling[0] : 0;
      }
    }
  };

  // The region a vacuum system's atoms stay in on the cell-list path: the
  // bounding box of the positions grown by this margin (A) on every side.
  static const double kPmVacuumMargin = 10.0;

  inline void pm_vacuum_region(const Vector3DBlock *positions, pm_system_desc &d) {
    double lo[3] = {0.0, 0.0, 0.0}, hi[3] = {0.0, 0.0, 0.0};
    for (size_t i = 0; i < positions->size(); i++)
      for (int k = 0; k < 3; k++) {
        const double v = (*positions)[i].c[k];
        if (i == 0 || v < lo[k]) lo[k] = v;
        if (i == 0 || v > hi[k]) hi[k] = v;
      }
    for (int k = 0; k < 3; k++) {
      d.box[k] = hi[k] - lo[k] + 2.0 * kPmVacuumMargin;
      d.origin[k] = 0.5 * (lo[k] + hi[k]);
    }
  }

  // The C ABI's switch record of a ProtoMol switching function, from its
  // getParameters() values (keywords -switchon, -cutoff, -switchoff, -n).
  template <class TSwitch>
  pm_switch_desc pm_switch_of(const TSwitch &sw) {
    std::vector<Parameter> p;
    sw.getParameters(p);
    pm_switch_desc d = pm_switch_desc();
    const std::string id = TSwitch::getId();
    d.kind = id == "C1" ? PM_SWITCH_C1 : id == "C2" ? PM_SWITCH_C2 : id == "Cn" ? PM_SWITCH_CN :
             id == "CmpCnCn" ? PM_SWITCH_CMPCNCN : id == "Cutoff" ? PM_SWITCH_CUTOFF : PM_SWITCH_NONE;
    for (size_t k = 0; k < p.size(); k++) {
      Real v = 0.0;
      p[k].value.get(v);
      if (p[k].keyword == "-switchon") d.switchon = v;
      else if (p[k].keyword == "-cutoff") d.cutoff = v;
      else if (p[k].keyword == "-switchoff") d.switchoff = v;
      else if (p[k].keyword == "-n") d.order = v;
    }
    return d;
  }

  template <class TForce> struct PmTerm;
  template <> struct PmTerm<LennardJonesForce> { enum { term = PM_TERM_LJ }; };
  template <> struct PmTerm<CoulombForce> { enum { term = PM_TERM_COULOMB }; };

  // the force terms and switches of a OneAtomPair[Two] into a force record
  template <class TPair> struct PmPairTerms;
  template <class BC, class S, class F>
  struct PmPairTerms<OneAtomPair<BC, S, F> > {
    typedef OneAtomPair<BC, S, F> Pair;
    static void describe(const Pair &, const S &sw, const S &, pm_force_desc &d) {
      d.term_a = PmTerm<F>::term;
      d.switch_a = pm_switch_of(sw);
      d.term_b = PM_TERM_NONE;
    }
    static Pair make(const S &sa, const S &) { return Pair(F(), sa); }
    static void split(const std::vector<Value> &v, S &sa, S &sb) {
      sa = S::make(std::vector<Value>(v.begin() + F::getParameterSize(), v.end()));
      sb = sa;
    }
    typedef S SwitchA;
    typedef S SwitchB;
  };
  template <class BC, class SA, class FA, class SB, class FB>
  struct PmPairTerms<OneAtomPairTwo<BC, SA, FA, SB, FB> > {
    typedef OneAtomPairTwo<BC, SA, FA, SB, FB> Pair;
    static void describe(const Pair &, const SA &sa, const SB &sb, pm_force_desc &d) {
      d.term_a = PmTerm<FA>::term;
      d.switch_a = pm_switch_of(sa);
      d.term_b = PmTerm<FB>::term;
      d.switch_b = pm_switch_of(sb);
    }
    static Pair make(const SA &sa, const SB &sb) { return Pair(FA(), sa, FB(), sb); }
    static void split(const std::vector<Value> &v, SA &sa, SB &sb) {
      const size_t a0 = FA::getParameterSize(), a1 = a0 + SA::getParameterSize();
      const size_t b0 = a1 + FB::getParameterSize();
      sa = SA::make(std::vector<Value>(v.begin() + a0, v.begin() + a1));
      sb = SB::make(std::vector<Value>(v.begin() + b0, v.end()));
    }
    typedef SA SwitchA;
    typedef SB SwitchB;
  };

  // A device context per topology; a change of topology or of the
  // ScalarStructure's virial() / molecularVirial() switches (PM_FLAG_VIRIAL,
  // PM_FLAG_MOLVIRIAL) or of the force record rebuilds it.
  class PmSystemForceBase : public SystemForce {
  public:
    explicit PmSystemForceBase(const pm_force_desc &d)
      : myDesc(d), myCtx(0), myTopo(0), myVirial(false), myMolVirial(false), myRegion(false) {}
    PmSystemForceBase(const PmSystemForceBase &o)
      : SystemForce(o), myDesc(o.myDesc), myCtx(0), myTopo(0), myVirial(false), myMolVirial(false),
        myRegion(false) {}
    virtual ~PmSystemForceBase() { if (myCtx) pm_ctx_destroy(myCtx); }

    virtual void evaluate(const GenericTopology *topo, const Vector3DBlock *positions,
                          Vector3DBlock *forces, ScalarStructure *energies) {
      const bool vir = energies->virial(), mol = energies->molecularVirial();
      ensure(topo, positions, vir, mol);
      std::vector<double> e(PM_E_NSLOTS, 0.0);
      const double *x = positions->size() ? (*positions)[0].c : 0;
      double *f = forces->size() ? (*forces)[0].c : 0;
      int rc = pm_evaluate(myCtx, 0, x, f, &e[0]);
      if (rc == PM_ECAPACITY && myRegion) {
        // an atom left the vacuum region the context was made for (a failed
        // evaluation adds nothing): once more, in a region around these positions
        uncache();
        ensure(topo, positions, vir, mol);
        rc = pm_evaluate(myCtx, 0, x, f, &e[0]);
      }
      if (rc != PM_OK) THROW(std::string("protomol_amd: ") + pm_last_error(myCtx));
      (*energies)[ScalarStructure::COULOMB] += e[PM_E_COULOMB];
      (*energies)[ScalarStructure::LENNARDJONES] += e[PM_E_LENNARDJONES];
      if (vir || mol)
        for (int k = 0; k < 9; k++)
          (*energies)[(ScalarStructure::Index)(ScalarStructure::VIRIALXX + k)] += e[PM_E_VIRIALXX + k];
      if (mol)
        for (int k = 0; k < 9; k++)
          (*energies)[(ScalarStructure::Index)(ScalarStructure::MOLVIRIALXX + k)] += e[PM_E_MOLVIRIALXX + k];
    }
    // one device evaluates everything; the MPI split has nothing to divide
    virtual void parallelEvaluate(const GenericTopology *topo, const Vector3DBlock *positions,
                                  Vector3DBlock *forces, ScalarStructure *energies) {
      evaluate(topo, positions, forces, energies);
    }
    virtual unsigned int numberOfBlocks(const GenericTopology *, const Vector3DBlock *) { return 1; }
    virtual void uncache() {
      if (myCtx) pm_ctx_destroy(myCtx);
      myCtx = 0;
      myTopo = 0;
    }
    const pm_force_desc &forceDesc() const { return myDesc; }

  protected:
    // the force record for this topology (the Ewald alpha default and the
    // GB constants come from it)
    virtual pm_force_desc describeFor(const GenericTopology *, const PmTopologyImage &) const { return myDesc; }
    virtual bool needsGB() const { return false; }
    // on the cell-list path a vacuum system needs a region (all-pairs does not)
    virtual bool cellPath() const { return false; }
    void ensure(const GenericTopology *topo, const Vector3DBlock *positions, bool vir, bool mol) {
      if (myCtx && myTopo == topo && myVirial == vir && myMolVirial == mol &&
          std::memcmp(&myBuilt, &myDesc, sizeof myDesc) == 0)
        return;
      uncache();
      PmTopologyImage img(topo, needsGB());
      myRegion = !img.desc.periodic && cellPath();
      if (myRegion) pm_vacuum_region(positions, img.desc);
      pm_force_desc d = describeFor(topo, img);
      if (vir) d.flags |= PM_FLAG_VIRIAL;
      if (mol) d.flags |= PM_FLAG_MOLVIRIAL;
      int rc = pm_ctx_create(&d, 1, &img.desc, &myCtx);
      if (rc != PM_OK) THROW(std::string("protomol_amd: ") + pm_last_error(0));
      myTopo = topo;
      myVirial = vir;
      myMolVirial = mol;
      myBuilt = myDesc;
    }
    pm_force_desc myDesc, myBuilt;
    pm_ctx *myCtx;
    const GenericTopology *myTopo;
    bool myVirial, myMolVirial, myRegion;
  };

  // ------------------------------------------------ NonbondedCutoffHIP

  template <class TOneAtomPair>
  class NonbondedCutoffSystemForceHIP : public PmSystemForceBase {
    typedef PmPairTerms<TOneAtomPair> Terms;
    typedef typename Terms::SwitchA SA;
    typedef typename Terms::SwitchB SB;

  public:
    NonbondedCutoffSystemForceHIP() : PmSystemForceBase(pm_force_desc()), myCutoff(0.0) {}
    NonbondedCutoffSystemForceHIP(Real cutoff, const SA &sa, const SB &sb)
      : PmSystemForceBase(describe(cutoff, sa, sb)), myCutoff(cutoff), mySwitchA(sa), mySwitchB(sb) {}

    virtual std::string getKeyword() const { return "NonbondedCutoffHIP"; }
    virtual std::string getIdNoAlias() const { return TOneAtomPair::getId() + " -algorithm " + getKeyword(); }
    // NonbondedCutoffForce::getParameters (NonbondedCutoffForce.h:83-89): the
    // pair's parameters, then -cutoff
    virtual void getParameters(std::vector<Parameter> &parameters) const {
      Terms::make(mySwitchA, mySwitchB).getParameters(parameters);
      parameters.push_back(Parameter("-cutoff", Value(myCutoff, ConstraintValueType::Positive()),
                                     Text("algorithm cutoff")));
    }

  protected:
    virtual bool cellPath() const { return true; }

  private:
    // NonbondedCutoffForce::doMake (NonbondedCutoffForce.h:70-74): the pair
    // from all values but the last (the switches validate their own), the
    // algorithm cutoff from the last
    virtual Force *doMake(const std::vector<Value> &values) const {
      std::vector<Value> pair(values.begin(), values.end() - 1);
      SA sa;
      SB sb;
      Terms::split(pair, sa, sb);
      return new NonbondedCutoffSystemForceHIP(values[values.size() - 1], sa, sb);
    }
    static pm_force_desc describe(Real cutoff, const SA &sa, const SB &sb) {
      pm_force_desc d = pm_force_desc();
      Terms::describe(Terms::make(sa, sb), sa, sb, d);
      d.cutoff = cutoff;
      d.rebuild_interval = 1;
      return d;
    }
    Real myCutoff;
    SA mySwitchA;
    SB mySwitchB;
  };

  // ------------------------------------------------ FullEwaldHIP

  template <bool REAL, bool RECIPROCAL, bool CORRECTION>
  class NonbondedFullEwaldSystemForceHIP : public PmSystemForceBase {
  public:
    NonbondedFullEwaldSystemForceHIP() : PmSystemForceBase(describe(-1.0, 0.00001)), myAlpha(-1.0), myAccuracy(0.00001) {}
    NonbondedFullEwaldSystemForceHIP(Real alpha, Real accuracy)
      : PmSystemForceBase(describe(alpha, accuracy)), myAlpha(alpha), myAccuracy(accuracy) {}

    virtual std::string getKeyword() const { return "FullEwaldHIP"; }
    // NonbondedFullEwaldSystemForce::getIdNoAlias (:1205-1218)
    virtual std::string getIdNoAlias() const {
      return std::string("Coulomb -algorithm ") + getKeyword() + (REAL ? " -real" : "") +
             (RECIPROCAL ? " -reciprocal" : "") + (CORRECTION ? " -correction" : "");
    }
    // :1231-1236 (periodic: no expansion factor)
    virtual void getParameters(std::vector<Parameter> &parameters) const {
      parameters.push_back(Parameter("-alpha", Value(myAlpha), -1.0, Text("splitting")));
      parameters.push_back(Parameter("-accuracy", Value(myAccuracy, ConstraintValueType::Positive()), 0.00001));
    }

  protected:
    // alpha <= 0: the reference's default, sqrt(pi) (5.5 N / V^2)^(1/6)
    // (NonbondedFullEwaldSystemForce.h:602-612)
    virtual pm_force_desc describeFor(const GenericTopology *, const PmTopologyImage &img) const {
      pm_force_desc d = myDesc;
      if (!(d.ewald_alpha > 0.0)) {
        const double V = img.desc.box[0] * img.desc.box[1] * img.desc.box[2];
        d.ewald_alpha = std::sqrt(M_PI) * std::pow(5.5 * img.desc.n_atoms / (V * V), 1.0 / 6.0);
      }
      return d;
    }

  private:
    // :120-140
    virtual Force *doMake(const std::vector<Value> &values) const {
      Real alpha = values[0], accuracy = values[1];
      if (!values[0].valid()) THROW("FullEwaldHIP: alpha '" + values[0].getString() + "' not valid.");
      if (!values[1].valid()) THROW("FullEwaldHIP: accuracy '" + values[1].getString() + "' not valid.");
      return new NonbondedFullEwaldSystemForceHIP(alpha, accuracy);
    }
    static pm_force_desc describe(Real alpha, Real accuracy) {
      pm_force_desc d = pm_force_desc();
      d.term_a = PM_TERM_NONE;
      d.term_b = PM_TERM_EWALD_REAL;
      d.ewald_alpha = alpha;
      d.ewald_accuracy = accuracy;
      d.ewald_correction = CORRECTION ? 1 : 0;
      d.rebuild_interval = 1;
      d.flags = RECIPROCAL ? PM_FLAG_EWALD_RECIPROCAL : 0u;
      return d;
    }
    Real myAlpha, myAccuracy;
  };

  // ------------------------------------------------ NonbondedSimpleFullHIP

  // The all-pairs terms one ForceGroup asks of the device
  struct PmFullGroup {
    bool gb, ace, lj, coulomb;
    double solute, solvent, ace_sigma, ace_rho_s;
    const void *owner;
    PmFullGroup() : gb(false), ace(false), lj(false), coulomb(false), solute(1.0), solvent(80.0),
                    ace_sigma(2.26 / 418.4), ace_rho_s(1.4), owner(0) {}
  };
  inline std::map<const void *, PmFullGroup> &pm_full_groups() {
    static std::map<const void *, PmFullGroup> groups;
    return groups;
  }

  // what each reference force class contributes to the group's launch
  template <class TForce> struct PmFullTerm;
  template <> struct PmFullTerm<GBBornRadii> {
    static void join(const GBBornRadii &, PmFullGroup &g) { g.gb = true; }
  };
  template <> struct PmFullTerm<GBPartialSum> {
    static void join(const GBPartialSum &, PmFullGroup &g) { g.gb = true; }
  };
  template <class TForce> double pm_param(const TForce &f, const char *key, double def) {
    std::vector<Parameter> p;
    f.getParameters(p);
    for (size_t k = 0; k < p.size(); k++)
      if (p[k].keyword == key) {
        Real v = def;
        p[k].value.get(v);
        return v;
      }
    return def;
  }
  template <> struct PmFullTerm<GBForce> {
    // GBForce -soluteDielec -solventDielec (GBForce.cpp:296-303)
    static void join(const GBForce &f, PmFullGroup &g) {
      g.gb = true;
      g.solute = pm_param(f, "-soluteDielec", 1.0);
      g.solvent = pm_param(f, "-solventDielec", 80.0);
    }
  };
  template <> struct PmFullTerm<GBACEForce> {
    // GBACEForce -solvationparam -watersphereradius (GBACEForce.cpp:222-230)
    static void join(const GBACEForce &f, PmFullGroup &g) {
      g.gb = true;
      g.ace = true;
      g.ace_sigma = pm_param(f, "-solvationparam", 2.26 / 418.4);
      g.ace_rho_s = pm_param(f, "-watersphereradius", 1.4);
    }
  };
  template <> struct PmFullTerm<LennardJonesForce> {
    static void join(const LennardJonesForce &, PmFullGroup &g) { g.lj = true; }
  };
  template <> struct PmFullTerm<CoulombForce> {
    static void join(const CoulombForce &, PmFullGroup &g) { g.coulomb = true; }
  };

  template <class TForce>
  class NonbondedSimpleFullSystemForceHIP : public PmSystemForceBase {
  public:
    NonbondedSimpleFullSystemForceHIP() : PmSystemForceBase(pm_force_desc()), myGroup(0), myNeedsGB(false) {}
    explicit NonbondedSimpleFullSystemForceHIP(const TForce &f)
      : PmSystemForceBase(pm_force_desc()), myForce(f), myGroup(0), myNeedsGB(false) {}
    virtual ~NonbondedSimpleFullSystemForceHIP() {
      std::map<const void *, PmFullGroup>::iterator it = pm_full_groups().find(myGroup);
      if (myGroup && it != pm_full_groups().end() && it->second.owner == this) pm_full_groups().erase(it);
    }

    virtual std::string getKeyword() const { return "NonbondedSimpleFullHIP"; }
    virtual std::string getIdNoAlias() const { return TForce::getId() + " -algorithm " + getKeyword(); }
    virtual void getParameters(std::vector<Parameter> &parameters) const { myForce.getParameters(parameters); }

    // SystemForce::addToForceGroup, and the term joins the group's launch
    virtual void addToForceGroup(ForceGroup *fg) {
      SystemForce::addToForceGroup(fg);
      myGroup = fg;
      PmFullGroup &g = pm_full_groups()[fg];
      if (!g.owner) g.owner = this;
      PmFullTerm<TForce>::join(myForce, g);
    }

    // the group's first member evaluates every term of the group; outside a
    // group a member evaluates its own term alone
    virtual void evaluate(const GenericTopology *topo, const Vector3DBlock *positions,
                          Vector3DBlock *forces, ScalarStructure *energies) {
      PmFullGroup alone;
      const PmFullGroup *g = &alone;
      if (myGroup) {
        g = &pm_full_groups()[myGroup];
        if (g->owner != this) return;
      } else {
        PmFullTerm<TForce>::join(myForce, alone);
      }
      myDesc = describe(*g);
      myNeedsGB = g->gb;
      PmSystemForceBase::evaluate(topo, positions, forces, energies);
    }

  protected:
    virtual bool needsGB() const { return myNeedsGB; }
    // the OBC constants come from the topology (ProtoMolApp.cpp:285-301)
    virtual pm_force_desc describeFor(const GenericTopology *topo, const PmTopologyImage &) const {
      pm_force_desc d = myDesc;
      d.gb_offset = topo->dielecOffset;
      d.gb_alpha = topo->alphaObc;
      d.gb_beta = topo->betaObc;
      d.gb_gamma = topo->gammaObc;
      return d;
    }

  private:
    static pm_force_desc describe(const PmFullGroup &g) {
      pm_force_desc d = pm_force_desc();
      d.flags = PM_FLAG_ALL_PAIRS | (g.gb ? PM_FLAG_GB : 0u) | (g.ace ? PM_FLAG_GB_ACE : 0u);
      d.term_a = g.lj ? PM_TERM_LJ : PM_TERM_NONE;
      d.term_b = g.coulomb ? PM_TERM_COULOMB : PM_TERM_NONE;
      d.gb_solute_dielec = g.solute;
      d.gb_solvent_dielec = g.solvent;
      d.gb_ace_sigma = g.ace_sigma;
      d.gb_ace_rho_s = g.ace_rho_s;
      return d;
    }
    virtual Force *doMake(const std::vector<Value> &values) const {
      return new NonbondedSimpleFullSystemForceHIP(TForce::make(values));
    }
    TForce myForce;
    const ForceGroup *myGroup;
    bool myNeedsGB;
  };

}  // namespace ProtoMol

#endif
