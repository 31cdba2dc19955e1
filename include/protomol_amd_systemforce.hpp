// protomol_amd_systemforce.hpp -- ProtoMol SystemForce adapters over the C ABI.
//
// Header-only; compiled inside a ProtoMol tree (it includes ProtoMol's own
// headers).  Each class is a drop-in SystemForce for the force it names:
//
//   NonbondedCutoffSystemForceHIP  for NonbondedCutoffSystemForce<CubicCellManager,
//       OneAtomPair<PBC, S, F>> and OneAtomPairTwo<PBC, SA, LennardJones, SB, Coulomb>
//       (force/nonbonded/NonbondedCutoffSystemForce.h:45-55)
//   NonbondedEwaldRealSystemForceHIP  for NonbondedFullEwaldSystemForce<PBC, Cubic,
//       true, false, correction, Cutoff> (force/nonbonded/NonbondedFullEwaldSystemForce.h:329-398)
//   NonbondedFullEwaldSystemForceHIP  for NonbondedFullEwaldSystemForce<PBC, Cubic,
//       true, true, correction, Cutoff> (the same, with reciprocalTerm :873-1083)
//
// evaluate() keeps the reference's semantics: it ADDS into forces and into
// the energy/virial slots of the ScalarStructure; a non-zero status from the
// C ABI becomes ProtoMol's THROW (base/Exception.h:258).
#ifndef PROTOMOL_AMD_SYSTEMFORCE_HPP
#define PROTOMOL_AMD_SYSTEMFORCE_HPP

#include <protomol/base/Exception.h>
#include <protomol/force/system/SystemForce.h>
#include <protomol/topology/GenericTopology.h>
#include <protomol/topology/PeriodicBoundaryConditions.h>
#include <protomol/topology/SemiGenericTopology.h>
#include <protomol/type/ScalarStructure.h>
#include <protomol/type/Vector3DBlock.h>

#include <string>
#include <vector>

#include "protomol_amd.h"

namespace ProtoMol {

  // Reads the fields the nonbonded path needs out of a ProtoMol topology:
  // scaled charges, masses, LJ types and table, molecules, the exclusion
  // table (in ExclusionTable::add order) and the periodic box.
  class PmTopologyImage {
  public:
    std::vector<double> charge, mass, lj;
    std::vector<int32_t> type, molecule, excl;
    pm_system_desc desc;

    explicit PmTopologyImage(const GenericTopology *topo) {
      const int n = (int)topo->atoms.size();
      charge.resize(n); mass.resize(n); type.resize(n); molecule.resize(n);
      int ntypes = 0;
      for (int i = 0; i < n; i++) {
        charge[i] = topo->atoms[i].scaledCharge;
        mass[i] = topo->atoms[i].scaledMass;
        type[i] = topo->atoms[i].type;
        molecule[i] = topo->atoms[i].molecule;
        if (type[i] + 1 > ntypes) ntypes = type[i] + 1;
      }
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
      desc.scaled_charge = &charge[0];
      desc.mass = &mass[0];
      desc.type = &type[0];
      desc.molecule = &molecule[0];
      desc.lj_mode = PM_LJ_TABLE;
      desc.n_types = ntypes;
      desc.lj_table = lj.empty() ? 0 : &lj[0];
      desc.n_exclusions = (int32_t)(excl.size() / 3);
      desc.exclusions = excl.empty() ? 0 : &excl[0];
      desc.coulomb_scaling = topo->coulombScalingFactor;
      const SemiGenericTopology<PeriodicBoundaryConditions> *pt =
          dynamic_cast<const SemiGenericTopology<PeriodicBoundaryConditions> *>(topo);
      if (!pt) THROW("protomol_amd: the cell-list force needs periodic boundary conditions");
      const PeriodicBoundaryConditions &bc = pt->boundaryConditions;
      if (!bc.isOrthogonal()) THROW("protomol_amd: orthogonal boxes only");
      desc.periodic = 1;
      desc.box[0] = bc.e1().c[0]; desc.box[1] = bc.e2().c[1]; desc.box[2] = bc.e3().c[2];
      desc.origin[0] = bc.origin().c[0]; desc.origin[1] = bc.origin().c[1]; desc.origin[2] = bc.origin().c[2];
    }
  };

  class PmSystemForceBase : public SystemForce {
  public:
    explicit PmSystemForceBase(const pm_force_desc &d)
      : myDesc(d), myCtx(0), myTopo(0), myVirial(false), myMolVirial(false) {}
    virtual ~PmSystemForceBase() { if (myCtx) pm_ctx_destroy(myCtx); }

    // the virial flags follow the ScalarStructure (ScalarStructure::virial(),
    // molecularVirial()); a change rebuilds the context
    virtual void evaluate(const GenericTopology *topo, const Vector3DBlock *positions,
                          Vector3DBlock *forces, ScalarStructure *energies) {
      const bool vir = energies->virial(), mol = energies->molecularVirial();
      ensure(topo, vir, mol);
      std::vector<double> e(PM_E_NSLOTS, 0.0);
      const double *x = positions->size() ? (*positions)[0].c : 0;
      double *f = forces->size() ? (*forces)[0].c : 0;
      int rc = pm_evaluate(myCtx, 0, x, f, &e[0]);
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

  protected:
    void ensure(const GenericTopology *topo, bool vir, bool mol) {
      if (myCtx && myTopo == topo && myVirial == vir && myMolVirial == mol) return;
      uncache();
      PmTopologyImage img(topo);
      pm_force_desc d = myDesc;
      if (vir) d.flags |= PM_FLAG_VIRIAL;
      if (mol) d.flags |= PM_FLAG_MOLVIRIAL;
      int rc = pm_ctx_create(&d, 1, &img.desc, &myCtx);
      if (rc != PM_OK) THROW(std::string("protomol_amd: ") + pm_last_error(0));
      myTopo = topo;
      myVirial = vir;
      myMolVirial = mol;
    }
    pm_force_desc myDesc;
    pm_ctx *myCtx;
    const GenericTopology *myTopo;
    bool myVirial, myMolVirial;
  };

  // LennardJones and/or Coulomb with a switching function, cell-list cutoff
  class NonbondedCutoffSystemForceHIP : public PmSystemForceBase {
  public:
    NonbondedCutoffSystemForceHIP(const pm_force_desc &d) : PmSystemForceBase(d) {}
    virtual std::string getKeyword() const { return "NonbondedCutoffHIP"; }
    virtual std::string getIdNoAlias() const { return "LennardJones Coulomb -algorithm NonbondedCutoffHIP"; }
    virtual void getParameters(std::vector<Parameter> &) const {}
  private:
    virtual Force *doMake(const std::vector<Value> &) const { return new NonbondedCutoffSystemForceHIP(myDesc); }
  };

  // Ewald real space + excluded-pair correction + self + charged-system terms
  class NonbondedEwaldRealSystemForceHIP : public PmSystemForceBase {
  public:
    NonbondedEwaldRealSystemForceHIP(double alpha, double accuracy, bool correction)
      : PmSystemForceBase(describe(alpha, accuracy, correction)) {}
    virtual std::string getKeyword() const { return "FullEwaldHIP"; }
    virtual std::string getIdNoAlias() const { return "CoulombEwaldReal -algorithm FullEwaldHIP"; }
    virtual void getParameters(std::vector<Parameter> &) const {}
  private:
    static pm_force_desc describe(double alpha, double accuracy, bool correction) {
      pm_force_desc d = pm_force_desc();
      d.term_a = PM_TERM_NONE;
      d.term_b = PM_TERM_EWALD_REAL;
      d.ewald_alpha = alpha;
      d.ewald_accuracy = accuracy;
      d.ewald_correction = correction ? 1 : 0;
      d.rebuild_interval = 1;
      return d;
    }
    virtual Force *doMake(const std::vector<Value> &) const { return new NonbondedEwaldRealSystemForceHIP(*this); }
  };

  // The full Ewald sum: real space + reciprocal space (+ correction, self and
  // charged-system terms), NonbondedFullEwaldSystemForce<PBC, Cubic, true,
  // true, correction, Cutoff> ("CoulombEwald")
  class NonbondedFullEwaldSystemForceHIP : public PmSystemForceBase {
  public:
    NonbondedFullEwaldSystemForceHIP(double alpha, double accuracy, bool correction)
      : PmSystemForceBase(describe(alpha, accuracy, correction)) {}
    virtual std::string getKeyword() const { return "FullEwaldHIP"; }
    virtual std::string getIdNoAlias() const { return "CoulombEwald -algorithm FullEwaldHIP"; }
    virtual void getParameters(std::vector<Parameter> &) const {}
  private:
    static pm_force_desc describe(double alpha, double accuracy, bool correction) {
      pm_force_desc d = pm_force_desc();
      d.term_a = PM_TERM_NONE;
      d.term_b = PM_TERM_EWALD_REAL;
      d.ewald_alpha = alpha;
      d.ewald_accuracy = accuracy;
      d.ewald_correction = correction ? 1 : 0;
      d.rebuild_interval = 1;
      d.flags = PM_FLAG_EWALD_RECIPROCAL;
      return d;
    }
    virtual Force *doMake(const std::vector<Value> &) const { return new NonbondedFullEwaldSystemForceHIP(*this); }
  };

}  // namespace ProtoMol

#endif
