/*
 * PmnbCutoffSystemForce.h -- a ProtoMol SystemForce backed by the MI355X C
 * ABI (include/pmnb.h).  Compiled inside a ProtoMol source tree (it includes
 * ProtoMol headers and needs -std=c++0x -fpermissive, as ProtoMol does);
 * the HIP library never includes ProtoMol headers.
 *
 * Drop-in for NonbondedCutoffSystemForce<CubicCellManager,
 *   OneAtomPairTwo<PeriodicBoundaryConditions, C2SwitchingFunction,
 *                  LennardJonesForce, TSwitchElec, TElec>>
 * (force/nonbonded/NonbondedCutoffSystemForce.h:11-78) with TElec =
 * CoulombForce (force/CoulombForce.h:13-60) or PmnbEwaldRealForce (below):
 * same keyword, id, parameters and make() arguments, so the same force line
 * in a .conf file resolves to it once it is registered in place of the CPU
 * exemplar (module/NonbondedCutoffForceModule.cpp:74-80).
 */
#ifndef PMNB_CUTOFF_SYSTEM_FORCE_H
#define PMNB_CUTOFF_SYSTEM_FORCE_H

#include <protomol/force/system/SystemForce.h>
#include <protomol/force/OneAtomPairTwo.h>
#include <protomol/force/LennardJonesForce.h>
#include <protomol/force/CoulombForce.h>
#include <protomol/switch/C2SwitchingFunction.h>
#include <protomol/switch/C1SwitchingFunction.h>
#include <protomol/topology/Topology.h>
#include <protomol/topology/PeriodicBoundaryConditions.h>
#include <protomol/topology/CubicCellManager.h>
#include <protomol/type/ScalarStructure.h>
#include <protomol/base/Report.h>
#include <protomol/base/MathUtilities.h>
#include <protomol/config/Parameter.h>

#include <vector>

#include "pmnb.h"

namespace ProtoMol {

// The Ewald real-space term qq erfc(alpha r) / r as a cutoff functor with
// the Abramowitz-Stegun erfc of NonbondedFullEwaldSystemForce.h:810-821
// (the reference computes this term only inside its full Ewald force).  It
// has CoulombForce's interface (force/CoulombForce.h:13-60) plus one
// parameter, -alpha, so OneAtomPairTwo reads it from the force line; the
// device computes the term, operator() is the same arithmetic on the host.
class PmnbEwaldRealForce {
public:
  enum {DIST_R2 = 1};
  enum {CUTOFF = 0};
  PmnbEwaldRealForce() : myAlpha(-1.0) {}
  explicit PmnbEwaldRealForce(Real alpha) : myAlpha(alpha) {}

  void operator()(Real &energy, Real &force, Real distSquared,
                  Real /*rDistSquared*/, const Vector3D &,
                  const GenericTopology *topo, int atom1, int atom2,
                  ExclusionClass excl) const {
    Real qq = topo->atoms[atom1].scaledCharge * topo->atoms[atom2].scaledCharge;
    if (excl == EXCLUSION_MODIFIED) qq *= topo->coulombScalingFactor;
    const Real r = sqrt(distSquared), rr = 1.0 / r, ar = myAlpha * r;
    const Real e = qq * exp(-ar * ar);
    energy = poly5(ar) * e * rr;
    force = (energy + 2.0 * myAlpha / sqrt(M_PI) * e) * rr * rr;
  }
  static void accumulateEnergy(ScalarStructure *energies, Real energy) {
    (*energies)[ScalarStructure::COULOMB] += energy;
  }
  static Real getEnergy(const ScalarStructure *energies) {
    return (*energies)[ScalarStructure::COULOMB];
  }
  static std::string getId() {return "EwaldReal";}
  static unsigned int getParameterSize() {return 1;}
  void getParameters(std::vector<Parameter> &parameters) const {
    parameters.push_back(Parameter("-alpha",
                                   Value(myAlpha, ConstraintValueType::Positive()),
                                   Text("Ewald splitting parameter (1/A)")));
  }
  static PmnbEwaldRealForce make(const std::vector<Value> &values) {
    Real a = -1.0;
    if (!values.empty()) values[0].get(a);
    return PmnbEwaldRealForce(a);
  }
  Real alpha() const {return myAlpha;}

private:
  Real myAlpha;
};

template<class TElec>
struct PmnbElecKind;
template<> struct PmnbElecKind<CoulombForce> {
  enum {VALUE = PMNB_ELEC_COULOMB};
};
template<> struct PmnbElecKind<PmnbEwaldRealForce> {
  enum {VALUE = PMNB_ELEC_EWALD_REAL};
};

template<class TSwitchElec>
struct PmnbSwitchKind;
template<> struct PmnbSwitchKind<C2SwitchingFunction> {
  enum {VALUE = PMNB_SWITCH_C2};
};
template<> struct PmnbSwitchKind<C1SwitchingFunction> {
  enum {VALUE = PMNB_SWITCH_C1};
};

template<class TSwitchElec, class TElec = CoulombForce>
class PmnbCutoffSystemForce : public SystemForce {
  typedef OneAtomPairTwo<PeriodicBoundaryConditions, C2SwitchingFunction,
                         LennardJonesForce, TSwitchElec, TElec> Pair;
  typedef Topology<PeriodicBoundaryConditions, CubicCellManager> RealTopology;

public:
  PmnbCutoffSystemForce() : myCutoff(0.0), myCtx(0), myTopo(0), myN(-1) {}

  // values: the OneAtomPairTwo parameters in the order of
  // OneAtomPairTwo::make (force/OneAtomPairTwo.h:141-154): LJ (none), its
  // C2 switchon and cutoff, the electrostatic functor's (-alpha for
  // PmnbEwaldRealForce), the electrostatic switch's; the algorithm -cutoff
  // comes last, as NonbondedCutoffForce::doMake reads it
  // (NonbondedCutoffForce.h:71-75).
  PmnbCutoffSystemForce(Real cutoff, const std::vector<Value> &pairValues,
                        int precision = PMNB_PREC_MIXED, int device = 0)
    : myCutoff(cutoff), myPairValues(pairValues), myPair(Pair::make(pairValues)),
      myPrecision(precision), myDevice(device), myCtx(0), myTopo(0), myN(-1) {}

  virtual ~PmnbCutoffSystemForce() {if (myCtx) pmnb_destroy(myCtx);}

  using SystemForce::evaluate;
  virtual void evaluate(const GenericTopology *topo,
                        const Vector3DBlock *positions, Vector3DBlock *forces,
                        ScalarStructure *energies) {
    const RealTopology *t = dynamic_cast<const RealTopology *>(topo);
    if (!t)
      Report::report << Report::error << "[PmnbCutoffSystemForce] needs a "
                     << "periodic cubic-cell topology." << Report::endr;
    setup(t);
    unsigned int flags = 0;
    if (energies->virial()) flags |= PMNB_FLAG_VIRIAL;
    std::vector<double> com;
    if (energies->molecularVirial()) {
      flags |= PMNB_FLAG_MOLVIRIAL;
      com.resize(3 * t->molecules.size());
      for (unsigned int m = 0; m < t->molecules.size(); m++)
        for (int k = 0; k < 3; k++) com[3 * m + k] = t->molecules[m].position.c[k];
    }
    double e[2] = {0.0, 0.0}, vir[9] = {0}, mvir[9] = {0};
    check(pmnb_evaluate(myCtx, positions->c, forces->c, e, vir, mvir,
                        com.empty() ? 0 : &com[0], flags));
    (*energies)[ScalarStructure::LENNARDJONES] += e[0];
    (*energies)[ScalarStructure::COULOMB] += e[1];
    for (int k = 0; k < 9 && flags; k++)
      (*energies)[(ScalarStructure::Index)(ScalarStructure::VIRIALXX + k)] += vir[k];
    for (int k = 0; k < 9 && (flags & PMNB_FLAG_MOLVIRIAL); k++)
      (*energies)[(ScalarStructure::Index)(ScalarStructure::MOLVIRIALXX + k)] += mvir[k];
  }

  // SystemForce::parallelEvaluate defaults to evaluate (SystemForce.h:91-96)
  virtual void parallelEvaluate(const GenericTopology *topo,
                                const Vector3DBlock *positions,
                                Vector3DBlock *forces,
                                ScalarStructure *energies) {
    evaluate(topo, positions, forces, energies);
  }

  virtual unsigned int numberOfBlocks(const GenericTopology *,
                                      const Vector3DBlock *) {return 1;}
  virtual std::string getKeyword() const {return "NonbondedCutoff";}
  virtual std::string getIdNoAlias() const {
    return Pair::getId() + " -algorithm " + getKeyword();
  }
  virtual void getParameters(std::vector<Parameter> &parameters) const {
    myPair.getParameters(parameters);
    parameters.push_back(Parameter("-cutoff",
                                   Value(myCutoff, ConstraintValueType::Positive()),
                                   Text("algorithm cutoff")));
  }

private:
  virtual Force *doMake(const std::vector<Value> &values) const {
    std::vector<Value> pairValues(values.begin(), values.end() - 1);
    return new PmnbCutoffSystemForce(values[values.size() - 1], pairValues,
                                     myPrecision, myDevice);
  }

  void check(int rc) const {
    if (rc != PMNB_OK)
      Report::report << Report::error << "[PmnbCutoffSystemForce] "
                     << pmnb_last_error(myCtx) << " (" << rc << ")"
                     << Report::endr;
  }

  void setup(const RealTopology *t) {
    if (!myCtx) {
      pmnb_params p;
      std::vector<Real> v(myPairValues.size(), 0.0);
      for (unsigned int k = 0; k < myPairValues.size(); k++)
        myPairValues[k].get(v[k]);
      // [switchon, cutoff] of LJ's C2, the electrostatic functor's values,
      // then its switch's: [switchon, cutoff] (C2) or [cutoff] (C1)
      const unsigned int e0 = 2, s0 = e0 + TElec::getParameterSize();
      if (v.size() < s0 + TSwitchElec::getParameterSize())
        Report::report << Report::error << "[PmnbCutoffSystemForce] "
                       << "missing force parameters." << Report::endr;
      p.elec = PmnbElecKind<TElec>::VALUE;
      p.elec_switch = PmnbSwitchKind<TSwitchElec>::VALUE;
      p.precision = myPrecision;
      p.device = myDevice;
      p.lj_switchon = v[0];
      p.lj_cutoff = v[1];
      if (p.elec_switch == PMNB_SWITCH_C2) {
        p.elec_switchon = v[s0];
        p.elec_cutoff = v[s0 + 1];
      } else {
        p.elec_switchon = 0.0;
        p.elec_cutoff = v[s0];
      }
      p.algorithm_cutoff = myCutoff;
      p.cell_size = t->cellManager.getCellSize();
      p.ewald_alpha = p.elec == PMNB_ELEC_EWALD_REAL ? v[e0] : 0.0;
      check(pmnb_create(&p, &myCtx));
    }
    const int n = (int)t->atoms.size();
    if (t != myTopo || n != myN) {
      std::vector<int32_t> type(n), mol(n);
      std::vector<double> q(n);
      for (int i = 0; i < n; i++) {
        type[i] = t->atoms[i].type;
        mol[i] = t->atoms[i].molecule;
        q[i] = t->atoms[i].scaledCharge;
      }
      const int T = (int)t->atomTypes.size();
      std::vector<double> lj(4 * T * T);
      for (int a = 0; a < T; a++)
        for (int b = 0; b < T; b++) {
          const LennardJonesParameters &l = t->lennardJonesParameters(a, b);
          lj[4 * (a * T + b)] = l.A;
          lj[4 * (a * T + b) + 1] = l.B;
          lj[4 * (a * T + b) + 2] = l.A14;
          lj[4 * (a * T + b) + 3] = l.B14;
        }
      // ExclusionTable::getTable rows with the class the table resolves
      // (a pair added twice keeps the last class)
      const std::vector<ExclusionPair> &tab = t->exclusions.getTable();
      std::vector<int32_t> ex(3 * tab.size());
      for (unsigned int k = 0; k < tab.size(); k++) {
        ex[3 * k] = tab[k].a1;
        ex[3 * k + 1] = tab[k].a2;
        ex[3 * k + 2] = (int32_t)t->exclusions.check(tab[k].a1, tab[k].a2);
      }
      check(pmnb_set_topology(myCtx, n, n ? &type[0] : 0, n ? &q[0] : 0,
                              n ? &mol[0] : 0, T, &lj[0],
                              t->coulombScalingFactor, (int32_t)tab.size(),
                              tab.empty() ? 0 : &ex[0]));
      myTopo = t;
      myN = n;
    }
    const PeriodicBoundaryConditions &bc = t->boundaryConditions;
    check(pmnb_set_box(myCtx, bc.e1().c, bc.e2().c, bc.e3().c, bc.origin().c));
  }

  Real myCutoff;
  std::vector<Value> myPairValues;
  Pair myPair;
  int myPrecision, myDevice;
  pmnb_ctx *myCtx;
  const RealTopology *myTopo;
  int myN;
};

}  // namespace ProtoMol

#endif  // PMNB_CUTOFF_SYSTEM_FORCE_H
