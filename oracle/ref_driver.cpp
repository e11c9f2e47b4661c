// Reference-build driver: instantiates ProtoMol's own short-range nonbonded
// cutoff force from the ProtoMol sources (compiled in place, never copied) and
// runs one evaluation on a system read from a binary file.  It exists only to
// pin the oracle in oracle/pmnb_oracle.c against the real reference code; it
// is test infrastructure and never part of the product path.
//
// Force instantiated (force/nonbonded/NonbondedCutoffSystemForce.h:11-78):
//   NonbondedCutoffSystemForce<CubicCellManager,
//     OneAtomPairTwo<PeriodicBoundaryConditions, SW1, LennardJonesForce,
//                    SW2, ELEC, PairRecorder>>
// with (SW2, ELEC) = (C2, CoulombForce), (C2, EwaldRealForce) or
// (C1, CoulombForce).  EwaldRealForce is a cutoff functor written here from
// the real-space term of force/nonbonded/NonbondedFullEwaldSystemForce.h:810-821
// (the reference has no cutoff functor for it).  PairRecorder is a
// OneAtomPairTwo constraint (force/OneAtomContraints.h:29-45 interface) whose
// POST_CHECK hook records every pair that passed the cutoff and exclusion
// tests, in the reference's own enumeration order.
//
// Input file (little endian), written by oracle/oracle.py:write_ref_input:
//   int32[8]  : magic 0x504d4e42, N, T, nExcl, kind, flags, nMol, 0
//   double[24]: e1[3] e2[3] e3[3] origin[3] ljSwitchon ljCutoff elSwitchon
//               elCutoff algCutoff cellSize csf alpha 0 0 0 0
//   double pos[3N], int32 type[N], double q[N], int32 mol[N],
//   double lj[T*T*4] (A,B,A14,B14), int32 excl[nExcl*3] (a1,a2,class),
//   double com[nMol*3]
// kind: 0 LJ(C2)+Coulomb(C2), 1 LJ(C2)+EwaldReal(C2), 2 LJ(C2)+Coulomb(C1)
// flags: bit0 virial, bit1 molecular virial
// With a third argument R (timing mode) the force runs R times without the
// pair recorder and the wall time of each evaluation is printed ("ms t");
// the output file then holds the last evaluation's results and no pairs.
// Output file: double f[3N], double e[2] (LJ, Coulomb), double vir[9],
//   double molvir[9], int64 nPairs, int32 pairs[nPairs*3] (i, j, class)
#include <protomol/topology/Topology.h>
#include <protomol/topology/PeriodicBoundaryConditions.h>
#include <protomol/topology/CubicCellManager.h>
#include <protomol/topology/CellListEnumeratorPeriodicBoundaries.h>
#include <protomol/force/nonbonded/NonbondedCutoffSystemForce.h>
#include <protomol/force/OneAtomPairTwo.h>
#include <protomol/force/LennardJonesForce.h>
#include <protomol/force/CoulombForce.h>
#include <protomol/switch/C2SwitchingFunction.h>
#include <protomol/switch/C1SwitchingFunction.h>
#include <protomol/base/MathUtilities.h>
#include <protomol/type/ScalarStructure.h>
#include <protomol/type/Vector3DBlock.h>

#include <ctime>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <cmath>
#include <vector>
#include <stdint.h>

using namespace ProtoMol;

typedef Topology<PeriodicBoundaryConditions, CubicCellManager> PbcTopology;

static std::vector<int32_t> g_pairs;
static double g_alpha = 0.0;

struct PairRecorder {
  enum {PRE_CHECK = 0, POST_CHECK = 1};
  static bool check(const GenericTopology *, int, int) {return true;}
  static bool check(const GenericTopology *, int) {return true;}
  static std::string getPrefixId() {return "";}
  static std::string getPostfixId() {return "Rec";}
  static void check(const GenericTopology *topo, int i, int j,
                    const Vector3D &, Real, const Vector3D &) {
    int a = i < j ? i : j, b = i < j ? j : i;
    int same = topo->atoms[a].molecule == topo->atoms[b].molecule;
    g_pairs.push_back(a);
    g_pairs.push_back(b);
    g_pairs.push_back(same ? (int)topo->exclusions.check(a, b) : 0);
  }
};

// Ewald real-space term as a cutoff functor; arithmetic restated from the
// A&S branch of NonbondedFullEwaldSystemForce.h:810-821.
struct EwaldRealForce {
  enum {DIST_R2 = 1};
  enum {CUTOFF = 0};
  void operator()(Real &energy, Real &force, Real distSquared,
                  Real /*rDistSquared*/, const Vector3D &,
                  const GenericTopology *topo, int atom1, int atom2,
                  ExclusionClass excl) const {
    Real r = sqrt(distSquared);
    Real qq = topo->atoms[atom1].scaledCharge * topo->atoms[atom2].scaledCharge;
    if (excl == EXCLUSION_MODIFIED) qq *= topo->coulombScalingFactor;
    Real rr = 1.0 / r;
    Real ar = g_alpha * r;
    Real e = qq * exp(-ar * ar);
    Real en = poly5(ar) * e * rr;
    Real my2AlphaPI = 2.0 * g_alpha / sqrt(M_PI);
    energy = en;
    force = ((en + my2AlphaPI * e) * rr * rr);
  }
  static void accumulateEnergy(ScalarStructure *energies, Real energy) {
    (*energies)[ScalarStructure::COULOMB] += energy;
  }
  static std::string getId() {return "EwaldReal";}
  static unsigned int getParameterSize() {return 0;}
  void getParameters(std::vector<Parameter> &) const {}
  static EwaldRealForce make(const std::vector<Value> &) {return EwaldRealForce();}
};

template<class F>
static void runForce(F &force, PbcTopology *topo, Vector3DBlock &pos,
                     Vector3DBlock &forces, ScalarStructure &en) {
  topo->uncacheCellList();
  force.evaluate(topo, &pos, &forces, &en);
}

static void die(const char *m);

// One evaluation of the force `kind` selects; returns its wall time in ms.
template <class C>
static double dispatch(int kind, double algCutoff,
                       const C2SwitchingFunction &swLJ, const double *par,
                       PbcTopology *topo, Vector3DBlock &pos,
                       Vector3DBlock &forces, ScalarStructure &en) {
  timespec t0, t1;
  clock_gettime(CLOCK_MONOTONIC, &t0);
  if (kind == 0) {
    typedef OneAtomPairTwo<PeriodicBoundaryConditions, C2SwitchingFunction,
      LennardJonesForce, C2SwitchingFunction, CoulombForce, C> P;
    NonbondedCutoffSystemForce<CubicCellManager, P>
      f(algCutoff, P(LennardJonesForce(), swLJ, CoulombForce(),
                     C2SwitchingFunction(par[14], par[15])));
    clock_gettime(CLOCK_MONOTONIC, &t0);
    runForce(f, topo, pos, forces, en);
  } else if (kind == 1) {
    typedef OneAtomPairTwo<PeriodicBoundaryConditions, C2SwitchingFunction,
      LennardJonesForce, C2SwitchingFunction, EwaldRealForce, C> P;
    NonbondedCutoffSystemForce<CubicCellManager, P>
      f(algCutoff, P(LennardJonesForce(), swLJ, EwaldRealForce(),
                     C2SwitchingFunction(par[14], par[15])));
    clock_gettime(CLOCK_MONOTONIC, &t0);
    runForce(f, topo, pos, forces, en);
  } else if (kind == 2) {
    typedef OneAtomPairTwo<PeriodicBoundaryConditions, C2SwitchingFunction,
      LennardJonesForce, C1SwitchingFunction, CoulombForce, C> P;
    NonbondedCutoffSystemForce<CubicCellManager, P>
      f(algCutoff, P(LennardJonesForce(), swLJ, CoulombForce(),
                     C1SwitchingFunction(par[15])));
    clock_gettime(CLOCK_MONOTONIC, &t0);
    runForce(f, topo, pos, forces, en);
  } else {
    die("unknown kind");
  }
  clock_gettime(CLOCK_MONOTONIC, &t1);
  return 1e3 * (t1.tv_sec - t0.tv_sec) + 1e-6 * (t1.tv_nsec - t0.tv_nsec);
}

static void die(const char *m) {fprintf(stderr, "ref_driver: %s\n", m); exit(2);}

int main(int argc, char **argv) {
  if (argc != 3 && argc != 4)
    die("usage: ref_driver in.bin out.bin [timed repeats]");
  FILE *fi = fopen(argv[1], "rb");
  if (!fi) die("cannot open input");
  int32_t hdr[8];
  double par[24];
  if (fread(hdr, 4, 8, fi) != 8 || hdr[0] != 0x504d4e42) die("bad header");
  if (fread(par, 8, 24, fi) != 24) die("bad params");
  const int N = hdr[1], T = hdr[2], nExcl = hdr[3], kind = hdr[4],
            flags = hdr[5], nMol = hdr[6];
  std::vector<double> p(3 * N), q(N), lj(4 * T * T), com(3 * nMol);
  std::vector<int32_t> type(N), mol(N), excl(3 * nExcl);
  if (fread(&p[0], 8, 3 * N, fi) != (size_t)(3 * N)) die("pos");
  if (fread(&type[0], 4, N, fi) != (size_t)N) die("type");
  if (fread(&q[0], 8, N, fi) != (size_t)N) die("q");
  if (fread(&mol[0], 4, N, fi) != (size_t)N) die("mol");
  if (fread(&lj[0], 8, 4 * T * T, fi) != (size_t)(4 * T * T)) die("lj");
  if (nExcl && fread(&excl[0], 4, 3 * nExcl, fi) != (size_t)(3 * nExcl)) die("excl");
  if (nMol && fread(&com[0], 8, 3 * nMol, fi) != (size_t)(3 * nMol)) die("com");
  fclose(fi);

  PeriodicBoundaryConditions bc(Vector3D(par[0], par[1], par[2]),
                                Vector3D(par[3], par[4], par[5]),
                                Vector3D(par[6], par[7], par[8]),
                                Vector3D(par[9], par[10], par[11]));
  CubicCellManager cm(par[17]);
  ExclusionType et("scaled1-4");
  PbcTopology *topo = new PbcTopology(par[18], et, bc, cm, 16.0f);
  g_alpha = par[19];

  topo->atoms.resize(N);
  for (int i = 0; i < N; i++) {
    topo->atoms[i].type = type[i];
    topo->atoms[i].scaledCharge = q[i];
    topo->atoms[i].molecule = mol[i];
    topo->atoms[i].cellListNext = -1;
  }
  topo->atomTypes.resize(T);
  topo->lennardJonesParameters.resize(T);
  for (int a = 0; a < T; a++)
    for (int b = a; b < T; b++) {
      const double *l = &lj[4 * (a * T + b)];
      topo->lennardJonesParameters.set(a, b,
                                       LennardJonesParameters(l[0], l[1], l[2], l[3]));
    }
  topo->exclusions.resize(N);
  for (int k = 0; k < nExcl; k++)
    topo->exclusions.add(excl[3 * k], excl[3 * k + 1],
                         (ExclusionClass)excl[3 * k + 2]);
  topo->exclusions.optimize();
  topo->molecules.resize(nMol);
  for (int m = 0; m < nMol; m++)
    topo->molecules[m].position = Vector3D(com[3 * m], com[3 * m + 1], com[3 * m + 2]);

  Vector3DBlock pos(N), forces(N);
  for (int i = 0; i < N; i++) {
    pos[i] = Vector3D(p[3 * i], p[3 * i + 1], p[3 * i + 2]);
    forces[i] = Vector3D(0.0, 0.0, 0.0);
  }
  ScalarStructure en;
  en.clear();
  en.virial((flags & 1) != 0);
  en.molecularVirial((flags & 2) != 0);

  C2SwitchingFunction swLJ(par[12], par[13]);
  const double algCutoff = par[16];
  const int repeats = argc == 4 ? atoi(argv[3]) : 0;
  std::vector<double> ms;
  if (repeats > 0) {
    // timing: the same force without the pair recorder, evaluated
    // `repeats` times (cell lists rebuilt each time, as every step does)
    for (int r = 0; r < repeats; r++) {
      for (int i = 0; i < N; i++) forces[i] = Vector3D(0.0, 0.0, 0.0);
      en.clear();
      ms.push_back(dispatch<NoConstraint>(kind, algCutoff, swLJ, par, topo,
                                          pos, forces, en));
    }
  } else {
    dispatch<PairRecorder>(kind, algCutoff, swLJ, par, topo, pos, forces, en);
  }

  FILE *fo = fopen(argv[2], "wb");
  if (!fo) die("cannot open output");
  std::vector<double> out(3 * N);
  for (int i = 0; i < N; i++) {
    out[3 * i] = forces[i].c[0];
    out[3 * i + 1] = forces[i].c[1];
    out[3 * i + 2] = forces[i].c[2];
  }
  fwrite(&out[0], 8, 3 * N, fo);
  double e2[2] = {en[ScalarStructure::LENNARDJONES], en[ScalarStructure::COULOMB]};
  fwrite(e2, 8, 2, fo);
  double v[18];
  for (int k = 0; k < 9; k++) {
    v[k] = en[(ScalarStructure::Index)(ScalarStructure::VIRIALXX + k)];
    v[9 + k] = en[(ScalarStructure::Index)(ScalarStructure::MOLVIRIALXX + k)];
  }
  fwrite(v, 8, 18, fo);
  int64_t np = (int64_t)(g_pairs.size() / 3);
  fwrite(&np, 8, 1, fo);
  if (np) fwrite(&g_pairs[0], 4, g_pairs.size(), fo);
  fclose(fo);
  for (size_t k = 0; k < ms.size(); k++) printf("ms %.6f\n", ms[k]);
  delete topo;
  return 0;
}
