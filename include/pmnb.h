/*
 * pmnb.h -- C ABI of the MI355X short-range nonbonded cutoff force.
 *
 * Replaces, for one GPU, ProtoMol's
 *   NonbondedCutoffSystemForce<CubicCellManager,
 *     OneAtomPairTwo<PeriodicBoundaryConditions, C2SwitchingFunction,
 *                    LennardJonesForce, C2SwitchingFunction,
 *                    CoulombForce | Ewald real-space>>
 * (force/nonbonded/NonbondedCutoffSystemForce.h:45-55), i.e. the virtual
 *   SystemForce::evaluate(const GenericTopology*, const Vector3DBlock*,
 *                         Vector3DBlock*, ScalarStructure*)
 * (force/system/SystemForce.h:32-35).  Paths are relative to the reference's
 * protomol/src/protomol.  Plain C types only: no HIP or torch types cross
 * this boundary.  Every call returns PMNB_OK (0) or an error code; the text
 * of the last error of a context is returned by pmnb_last_error().  No
 * exception crosses the ABI.  A context is used by one host thread at a time.
 */
#ifndef PMNB_H
#define PMNB_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. */
#define PMNB_OK 0
#define PMNB_ERR_ARG 1         /* invalid argument or size                */
#define PMNB_ERR_HIP 2         /* HIP runtime error                       */
#define PMNB_ERR_UNSUPPORTED 3 /* configuration outside what is built     */
#define PMNB_ERR_STATE 4       /* call order (e.g. evaluate before set_*)  */
#define PMNB_ERR_CAPACITY 5    /* output buffer too small                 */

/* Electrostatic functor paired with LennardJonesForce (OneAtomPairTwo's
   second force): CoulombForce (force/CoulombForce.h:27-38) or the Ewald
   real-space term qq*erfc(alpha r)/r with the Abramowitz-Stegun erfc of
   force/nonbonded/NonbondedFullEwaldSystemForce.h:810-821. */
#define PMNB_ELEC_NONE 0
#define PMNB_ELEC_COULOMB 1
#define PMNB_ELEC_EWALD_REAL 2

/* Switching functions (switch/C2SwitchingFunction.h:42-53,
   switch/C1SwitchingFunction.h:41-50).  LJ always uses C2; the
   electrostatic term may use C2 or C1.  C1 is FP64-only. */
#define PMNB_SWITCH_C2 0
#define PMNB_SWITCH_C1 1

/* Arithmetic of the pair kernel.  FP64: every pair computed in double with
   the reference's operation order for the minimum image and cutoff test.
   MIXED: FP32 pair arithmetic on cell-local coordinates, FP64 box wrap,
   FP64 re-test of pairs near the cutoff (so the pair set is the FP64 one),
   64-bit fixed-point force sums, FP64 energy and virial sums. */
#define PMNB_PREC_FP64 0
#define PMNB_PREC_MIXED 1

/* evaluate() flags: ScalarStructure::virial() / molecularVirial()
   (type/ScalarStructure.h:115-119). */
#define PMNB_FLAG_VIRIAL 1u
#define PMNB_FLAG_MOLVIRIAL 2u

/* Exclusion classes (topology/ExclusionTable.h:19-23). */
#define PMNB_EXCL_NONE 0
#define PMNB_EXCL_MODIFIED 1
#define PMNB_EXCL_FULL 2

typedef struct pmnb_ctx pmnb_ctx;

/* Force parameters: what NonbondedCutoffForce::doMake
   (force/nonbonded/NonbondedCutoffForce.h:71-75) and OneAtomPairTwo::make
   (force/OneAtomPairTwo.h:142-155) read from the force line, plus the cell
   manager's cell size (topology/CubicCellManager.cpp:75-83). */
typedef struct {
  int32_t elec;            /* PMNB_ELEC_*                                 */
  int32_t elec_switch;     /* PMNB_SWITCH_*                               */
  int32_t precision;       /* PMNB_PREC_*                                 */
  int32_t device;          /* HIP device ordinal                          */
  double lj_switchon;      /* C2 -switchon of the LJ term                 */
  double lj_cutoff;        /* C2 -cutoff of the LJ term                   */
  double elec_switchon;    /* C2 -switchon of the electrostatic term      */
  double elec_cutoff;      /* C2 / C1 -cutoff of the electrostatic term   */
  double algorithm_cutoff; /* NonbondedCutoff -cutoff (cell enumeration)  */
  double cell_size;        /* CubicCellManager cellsize                   */
  double ewald_alpha;      /* alpha of erfc(alpha r), 1/Angstrom          */
} pmnb_params;

/* Creates a context on params->device.  Replaces force construction via
   ForceFactory::make (factory/ForceFactory.cpp:21-242). */
int pmnb_create(const pmnb_params *params, pmnb_ctx **out);

/* Uploads the topology (once; call again when it changes).
   type/scaled_charge/molecule: GenericTopology::atoms[i].{type,
     scaledCharge, molecule} (topology/Atom.h:167-189); scaled charges are
     charge * 18.2226123264 (topology/BuildTopology.cpp:92).
   lj: n_types*n_types*4 doubles {A, B, A14, B14}, symmetric
     (topology/LennardJonesParameterTable.h:34-36).
   excl: n_excl triples (a1, a2, class) of ExclusionTable::getTable()
     (topology/ExclusionTable.h:87) in insertion order; a pair given twice
     takes the class given last (ExclusionTable.cpp:105-120).  Pairs of one
     molecule with |a1 - a2| > 64 are rejected with PMNB_ERR_UNSUPPORTED.
   coulomb_scaling_factor: GenericTopology::coulombScalingFactor. */
int pmnb_set_topology(pmnb_ctx *ctx, int32_t n_atoms, const int32_t *type,
                      const double *scaled_charge, const int32_t *molecule,
                      int32_t n_types, const double *lj,
                      double coulomb_scaling_factor, int32_t n_excl,
                      const int32_t *excl);

/* Periodic box: PeriodicBoundaryConditions::set(e1, e2, e3, origin)
   (topology/PeriodicBoundaryConditions.cpp:32-77).  Orthogonal boxes only. */
int pmnb_set_box(pmnb_ctx *ctx, const double e1[3], const double e2[3],
                 const double e3[3], const double origin[3]);

/* One force evaluation on host arrays (synchronous):
   NonbondedCutoffSystemForce::evaluate.  positions/forces are Vector3DBlock
   storage, 3*n_atoms doubles x0 y0 z0 x1 ... (type/Vector3DBlock.h:27-38).
   Accumulates (+=): forces, energies[0] += LJ energy
   (ScalarStructure::LENNARDJONES), energies[1] += electrostatic energy
   (ScalarStructure::COULOMB), virial[9] (VIRIALXX..VIRIALZZ) when
   PMNB_FLAG_VIRIAL or PMNB_FLAG_MOLVIRIAL is set, molvirial[9]
   (MOLVIRIALXX..) when PMNB_FLAG_MOLVIRIAL is set, which also needs
   mol_com (3 doubles per molecule id: Molecule::position).  Pointers that
   a flag does not use may be NULL. */
int pmnb_evaluate(pmnb_ctx *ctx, const double *positions, double *forces,
                  double *energies, double *virial, double *molvirial,
                  const double *mol_com, uint32_t flags);

/* Same computation on device arrays, enqueued on `stream` (a hipStream_t,
   NULL = the context's stream) without host synchronisation.  d_out holds
   20 doubles: [0] LJ, [1] electrostatic, [2..10] virial, [11..19] molecular
   virial; all device outputs accumulate (+=).  d_mol_com may be NULL unless
   PMNB_FLAG_MOLVIRIAL is set.  No counterpart in the reference (which has no
   device); it is the entry a device-resident integrator calls. */
int pmnb_evaluate_device(pmnb_ctx *ctx, const double *d_positions,
                         double *d_forces, double *d_out,
                         const double *d_mol_com, uint32_t flags,
                         void *stream);

/* The pair set of one evaluation: every unordered pair that is not FULL
   excluded, is taken at its minimum image and has d^2 <= rc^2, where rc is
   the larger switching cutoff (force/OneAtomPairTwo.h:57-80).  Writes
   n_pairs and, when capacity allows, triples (i, j, class) with i < j,
   sorted by (i, j).  With pairs == NULL only the count is returned. */
int pmnb_pairset(pmnb_ctx *ctx, const double *positions, int32_t *pairs,
                 int64_t capacity, int64_t *n_pairs);

/* HIP-event timing of the pair stage on the stream it is launched on.
   enable = 1 starts (and clears) timing; enable = 0 stops it.  Either call
   first waits for the recorded launches and returns the summed durations
   since timing was last enabled: ms[0] the neighbour-list build, ms[1] the
   pair-force kernels, and *launches the number of evaluations (either
   pointer may be NULL).  No counterpart in the reference; it feeds the
   benchmark's per-kernel roofline figure. */
int pmnb_kernel_timing(pmnb_ctx *ctx, int32_t enable, double *ms,
                       int64_t *launches);

/* Number of in-cutoff pairs counted by the last evaluate or pairset call
   (P_alg); waits for the device. */
int pmnb_last_pair_count(pmnb_ctx *ctx, int64_t *n_pairs);

/* Text of the last error on this context ("" if none). */
const char *pmnb_last_error(const pmnb_ctx *ctx);

/* Frees the context and its device memory. */
void pmnb_destroy(pmnb_ctx *ctx);

/* ABI version (major * 100 + minor). */
int pmnb_version(void);

#ifdef __cplusplus
}
#endif

#endif /* PMNB_H */
