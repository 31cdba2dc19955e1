// ref_common.hpp -- record-file I/O and topology set-up shared by the
// programs that drive ProtoMol's own classes (ref_driver, adapter_check).
//
// TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile against the ProtoMol
// sources where they lie; nothing here ships in the library.
//
// Record file format (both directions): repeated
//   char name[16] (NUL padded) | int32 dtype (0 = f64, 1 = i32) | int64 count | data
#ifndef PM_REF_COMMON_HPP
#define PM_REF_COMMON_HPP

#include <protomol/topology/Topology.h>
#include <protomol/topology/PeriodicBoundaryConditions.h>
#include <protomol/topology/VacuumBoundaryConditions.h>
#include <protomol/topology/CubicCellManager.h>
#include <protomol/topology/CellListEnumeratorPeriodicBoundaries.h>
#include <protomol/topology/CellListEnumeratorStandard.h>
#include <protomol/topology/TopologyUtilities.h>
#include <protomol/force/nonbonded/NonbondedCutoffSystemForce.h>
#include <protomol/force/nonbonded/NonbondedFullEwaldSystemForce.h>
#include <protomol/force/nonbonded/NonbondedSimpleFullSystemForce.h>
#include <protomol/force/OneAtomPair.h>
#include <protomol/force/OneAtomPairTwo.h>
#include <protomol/force/OneAtomPairNoExclusion.h>
#include <protomol/force/LennardJonesForce.h>
#include <protomol/force/CoulombForce.h>
#include <protomol/force/GB/GBBornRadii.h>
#include <protomol/force/GB/GBPartialSum.h>
#include <protomol/force/GB/GBForce.h>
#include <protomol/force/GB/GBACEForce.h>
#include <protomol/force/bonded/BondSystemForce.h>
#include <protomol/force/bonded/AngleSystemForce.h>
#include <protomol/switch/C1SwitchingFunction.h>
#include <protomol/switch/C2SwitchingFunction.h>
#include <protomol/switch/CnSwitchingFunction.h>
#include <protomol/switch/CmpCnCnSwitchingFunction.h>
#include <protomol/switch/CutoffSwitchingFunction.h>
#include <protomol/switch/UniversalSwitchingFunction.h>
#include <protomol/type/ScalarStructure.h>
#include <protomol/type/Vector3DBlock.h>
#include <protomol/base/Report.h>
#include <protomol/topology/BuildTopology.h>
#include <protomol/type/PSF.h>
#include <protomol/type/PAR.h>
#include <protomol/type/PDB.h>
#include <protomol/type/GromacsTopology.h>
#include <protomol/type/GromacsParameters.h>
#include <protomol/io/PDBReader.h>
#include <protomol/io/gromacs/PortGromacsParameters.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <map>
#include <string>
#include <vector>

using namespace ProtoMol;

typedef PeriodicBoundaryConditions PBC;
typedef VacuumBoundaryConditions VBC;
typedef CubicCellManager CCM;
typedef Topology<PBC, CCM> PTopo;
typedef Topology<VBC, CCM> VTopo;

// ------------------------------------------------------------- records
struct Rec {
  int dtype;
  std::vector<double> d;
  std::vector<int> i;
};
typedef std::map<std::string, Rec> Recs;

inline bool read_recs(const char *path, Recs &out) {
  FILE *fp = fopen(path, "rb");
  if (!fp) return false;
  for (;;) {
    char name[17];
    memset(name, 0, sizeof name);
    if (fread(name, 1, 16, fp) != 16) break;
    int dtype;
    long long count;
    if (fread(&dtype, 4, 1, fp) != 1 || fread(&count, 8, 1, fp) != 1) { fclose(fp); return false; }
    Rec r;
    r.dtype = dtype;
    if (dtype == 0) {
      r.d.resize((size_t)count);
      if (count && fread(&r.d[0], 8, (size_t)count, fp) != (size_t)count) { fclose(fp); return false; }
    } else {
      r.i.resize((size_t)count);
      if (count && fread(&r.i[0], 4, (size_t)count, fp) != (size_t)count) { fclose(fp); return false; }
    }
    out[std::string(name)] = r;
  }
  fclose(fp);
  return true;
}

inline void write_rec_d(FILE *fp, const char *name, const double *d, long long n) {
  char nm[16];
  memset(nm, 0, sizeof nm);
  strncpy(nm, name, 15);
  int dt = 0;
  fwrite(nm, 1, 16, fp);
  fwrite(&dt, 4, 1, fp);
  fwrite(&n, 8, 1, fp);
  if (n) fwrite(d, 8, (size_t)n, fp);
}

inline double getd(Recs &r, const char *k, double def) {
  Recs::iterator it = r.find(k);
  if (it == r.end()) return def;
  return it->second.dtype == 0 ? it->second.d[0] : (double)it->second.i[0];
}
inline int geti(Recs &r, const char *k, int def) {
  Recs::iterator it = r.find(k);
  if (it == r.end()) return def;
  return it->second.dtype == 1 ? it->second.i[0] : (int)it->second.d[0];
}

// --------------------------------------------------------- topology fill
template <class T>
inline void fill_common(T &topo, Recs &r, int n) {
  topo.atoms.resize(n);
  int ntypes = geti(r, "ntypes", 1);
  topo.atomTypes.resize(ntypes);
  topo.lennardJonesParameters.resize(ntypes);
  const std::vector<double> &lj = r["lj"].d;
  for (int a = 0; a < ntypes; a++)
    for (int b = 0; b < ntypes; b++) {
      LennardJonesParameters p;
      const double *q = &lj[4 * (a * ntypes + b)];
      p.A = q[0]; p.B = q[1]; p.A14 = q[2]; p.B14 = q[3];
      topo.lennardJonesParameters.set(a, b, p);
    }
  const std::vector<int> &mol = r["molecule"].i;
  int nmol = 0;
  for (int i = 0; i < n; i++) if (mol[i] + 1 > nmol) nmol = mol[i] + 1;
  topo.molecules.resize(nmol);
  for (int i = 0; i < n; i++) {
    topo.atoms[i].type = r["type"].i[i];
    topo.atoms[i].scaledCharge = r["charge"].d[i];
    topo.atoms[i].scaledMass = r.count("mass") ? r["mass"].d[i] : 1.0;
    topo.atoms[i].molecule = mol[i];
    topo.molecules[mol[i]].atoms.push_back(i);
  }
  topo.coulombScalingFactor = getd(r, "csf", 1.0);
  topo.exclusions.resize(n);
  if (r.count("excl")) {
    const std::vector<int> &ex = r["excl"].i;
    for (size_t k = 0; k + 2 < ex.size(); k += 3)
      topo.exclusions.add(ex[k], ex[k + 1], ex[k + 2] == 2 ? EXCLUSION_MODIFIED : EXCLUSION_FULL);
  }
  topo.exclusions.optimize();
}

// the molecules' centres of mass, as the integrators build them before each
// force evaluation (STSIntegrator.cpp:51); the molecular virial reads them
template <class T>
inline void molecule_positions(T &topo, Recs &r, Vector3DBlock &pos) {
  if (geti(r, "molvirial", 0)) buildMolecularCenterOfMass(&pos, &topo);
}

// the GB/OBC state of a vacuum topology from the records (gb_params:
// offset, alpha, beta, gamma, eps_in, eps_out; gb_rho, gb_S per atom), as
// the topology builder and ProtoMolApp set it (BuildTopology.cpp:825-885,
// ProtoMolApp.cpp:285-301)
inline void fill_gb(VTopo &topo, Recs &r, int n) {
  const std::vector<double> &gp = r["gb_params"].d;  // offset, a, b, g, eps_in, eps_out
  topo.doGBSAOpenMM = 1;
  topo.obcType = 2;
  topo.dielecOffset = gp[0];
  topo.alphaObc = gp[1];
  topo.betaObc = gp[2];
  topo.gammaObc = gp[3];
  for (int i = 0; i < n; i++) {
    GBSAAtomParameters *g = new GBSAAtomParameters();
    topo.atoms[i].myGBSA_T = g;
    g->offsetRadius = gp[0];
    g->scalingFactor = r["gb_S"].d[i];
    g->vanDerWaalRadius = r["gb_rho"].d[i];
    g->SetSpaceForBornRadiusDerivatives(n);
    g->SetSpaceLvalues(n);
    g->SetSpaceUvalues(n);
    g->SetSpaceDistij(n);
    g->expTerm.resize(n);
    g->filTerm.resize(n);
    g->partialTerm.resize(n);
  }
}

#endif
