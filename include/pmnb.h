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
typedef struct pmnb_ewald pmnb_ewald; /* Ewald k-space terms (below) */
typedef struct pmnb_pme pmnb_pme;     /* PME k-space terms (below) */

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
   a flag does not use may be NULL.  For 3 * n_atoms >= 65536 the copy of
   the positions into the context's staging buffer and the force add are
   shared by the calling thread and host threads that the context starts on
   its first such call (4 in all, or PMNB_HOST_THREADS from the environment
   at that call; 1 = the calling thread alone); they poll for 1 ms after
   each call, then sleep until the next, and end with pmnb_destroy. */
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

/* pmnb_evaluate_device without any allocation: same arguments, results
   and kernels (bit for bit), with every per-evaluation array (cell counts
   and starts, the sorted atom arrays, work items, partial sums, the
   fixed-point force array, the grid and box) in one static device array
   of 64 MiB per GPU instead of the context's buffers; nothing is
   allocated or freed and no host copy is made.  No counterpart in the
   reference: it is for callers that cannot allocate device memory (code
   under HIP graph capture, tight serving loops, harnesses that forbid
   allocation).  Covers one system on the tile path (algorithm cutoff >=
   the switching cutoff, at most 8 LJ types): PMNB_ERR_UNSUPPORTED
   otherwise; PMNB_ERR_CAPACITY when the system needs more than 64 MiB
   (100 bytes per atom mixed, 124 FP64, plus 128 per 16 atoms and, per
   cell, 8 and 4 per bucket slot, at least 4 times the mean occupancy:
   about 520,000 atoms of water mixed, 425,000 FP64).  Every call on one GPU
   shares that GPU's array, which
   keeps the last call's contents: calls that overlap on one GPU (other
   streams or threads) race.  pmnb_last_pair_count does not see these
   calls.  PMNB_ERR_STATE while kernel timing is on. */
int pmnb_evaluate_device_noalloc(pmnb_ctx *ctx, const double *d_positions,
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

/* The pairs the force's own pair kernel evaluates at these positions: one
   evaluation (its forces and energies discarded) with the kernel's pair
   record on, which appends every pair the kernel accepts at the point
   OneAtomPairTwo::doOneAtomPair would compute it (force/OneAtomPairTwo.h:
   57-80: minimum image, d^2 <= rc^2, not FULL excluded).  Same output
   format as pmnb_pairset.  pmnb_pairset decides the pairs with a separate
   FP64 kernel; this one reports what the production kernels (either
   precision, either path) decide.  No counterpart in the reference. */
int pmnb_force_pairs(pmnb_ctx *ctx, const double *positions, int32_t *pairs,
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

/* Pair path a context's evaluations take (known after pmnb_set_box):
   PMNB_PATH_TILE: the algorithm cutoff reaches every pair within the
     switching cutoff, so the pair set does not depend on the cell grid,
     and at most 8 LJ types: the path's own column grid and tile kernel;
   PMNB_PATH_LIST: the reference's cell grid and delta list
     (topology/CellListEnumeratorPeriodicBoundaries.h:46-83) drive a
     neighbour list per evaluation;
   PMNB_PATH_SCAN: the same cells scanned directly (more than 512 cell
     offsets). */
#define PMNB_PATH_TILE 0
#define PMNB_PATH_LIST 1
#define PMNB_PATH_SCAN 2
int pmnb_pair_path(const pmnb_ctx *ctx, int32_t *path);

/* Number of in-cutoff pairs counted by the last evaluate or pairset call
   (P_alg); waits for the device. */
int pmnb_last_pair_count(pmnb_ctx *ctx, int64_t *n_pairs);

/* Text of the last error on this context ("" if none). */
const char *pmnb_last_error(const pmnb_ctx *ctx);

/* Frees the context and its device memory. */
void pmnb_destroy(pmnb_ctx *ctx);

/* ABI version (major * 100 + minor). */
int pmnb_version(void);

/* ------------------------------------------------------------ derivatives
   Force matching differentiates the force itself.  No counterpart in the
   reference's force interface; the per-pair second derivatives it follows
   are ReducedHessLennardJones (force/hessian/ReducedHessLennardJones.cpp:
   7-42), ReducedHessCoulomb (force/hessian/ReducedHessCoulomb.cpp:6-35) and
   C2SwitchingFunction::hessian (switch/C2SwitchingFunction.cpp:43-52), as
   Hessian::evaluatePairsMatrix (integrator/hessian/Hessian.cpp:581-717)
   combines them. */

/* Replaces the LJ table (d_lj: n_types*n_types*4 doubles {A, B, A14, B14},
   symmetric) and/or the scaled charges (d_scaled_charge: n_atoms doubles)
   from device arrays, enqueued on `stream`; either may be NULL (kept).  The
   rest of the topology stays as uploaded. */
int pmnb_set_parameters_device(pmnb_ctx *ctx, const double *d_lj,
                               const double *d_scaled_charge, void *stream);

/* Adjoint of the forces of pmnb_evaluate_device: with d_grad_forces = dL/dF
   (3 doubles per atom), accumulates (+=)
     d_grad_positions (3 per atom): dL/dx = -H dL/dF, a Hessian-vector
       product with the Hessian of the pair energies (for the Ewald term,
       the Jacobian of the force expression the force computes);
     d_grad_charges (1 per atom): dL/d(scaled charge);
     d_grad_lj (n_types*n_types*4): dL/d{A, B, A14, B14}, symmetrised.
   Any output may be NULL.  The pairs are recomputed from the cell list, in
   the context's precision.  Needs C2 switching, at most 4 LJ types and the
   tile path (algorithm cutoff >= switching cutoff). */
int pmnb_adjoint_device(pmnb_ctx *ctx, const double *d_positions,
                        const double *d_grad_forces, double *d_grad_positions,
                        double *d_grad_charges, double *d_grad_lj,
                        void *stream);

/* ------------------------------------------------------------------ batch
   Batched replicas.  No counterpart in the reference: one ProtoMol process
   evaluates one system through NonbondedCutoffSystemForce::evaluate
   (force/nonbonded/NonbondedCutoffSystemForce.h:45-55).  A batch holds many
   independent periodic systems that share one force (params, LJ table,
   coulomb scaling factor) and evaluates them all with one launch of each
   kernel; replica r's result is what pmnb_evaluate_device gives for that
   replica alone.  Replicas use the tile path, which needs the algorithm
   cutoff >= the switching cutoff and at most 8 LJ types. */
typedef struct pmnb_batch pmnb_batch;

/* Creates an empty batch; lj and coulomb_scaling_factor as in
   pmnb_set_topology. */
int pmnb_batch_create(const pmnb_params *params, int32_t n_types,
                      const double *lj, double coulomb_scaling_factor,
                      pmnb_batch **out);

/* Appends a replica: per-atom arrays and exclusions as in
   pmnb_set_topology (indices local to the replica), its box as in
   pmnb_set_box.  In the batch's position and force arrays its atoms follow
   those of the replicas added before it.  *replica receives its index. */
int pmnb_batch_add(pmnb_batch *b, int32_t n_atoms, const int32_t *type,
                   const double *scaled_charge, const int32_t *molecule,
                   int32_t n_excl, const int32_t *excl, const double e1[3],
                   const double e2[3], const double e3[3],
                   const double origin[3], int32_t *replica);

/* Replicas and atoms added so far. */
int pmnb_batch_size(const pmnb_batch *b, int32_t *n_replicas,
                    int64_t *n_atoms);

/* One evaluation of every replica on device arrays, enqueued on `stream`:
   d_positions / d_forces hold 3 doubles per atom of the batch, d_out 20
   doubles per replica in the layout of pmnb_evaluate_device.  Accumulates.
   PMNB_FLAG_MOLVIRIAL is not supported. */
int pmnb_batch_evaluate_device(pmnb_batch *b, const double *d_positions,
                               double *d_forces, double *d_out,
                               uint32_t flags, void *stream);

/* n_steps evaluations as replays of one captured HIP graph: each replay
   zeroes d_forces and d_out, then evaluates every replica.  The graph is
   captured on the first call and again whenever the pointers, flags or
   kernel shape change; nothing is allocated during capture.  Enqueued on
   `stream` without host synchronisation. */
int pmnb_batch_run_graph(pmnb_batch *b, const double *d_positions,
                         double *d_forces, double *d_out, uint32_t flags,
                         int32_t n_steps, void *stream);

/* Autotuner: times every pair-kernel shape (the shell, and the buffered
   candidates per wave with the waves per SIMD its LDS allows: mixed 0 full
   shell 256 / 3, 1 full shell 128 / 4, 2 half shell 192 / 3; FP64 0 full
   shell 192 / 3, 1 half shell 128 / 2) and persistent grid size on these
   positions and keeps the fastest for later evaluations; returns the shape
   index,
   the grid cap in blocks (0 = one block per work group of items) and its
   time per evaluation in ms (any pointer may be NULL).  The outputs are
   overwritten with accumulated sums. */
int pmnb_batch_autotune(pmnb_batch *b, const double *d_positions,
                        double *d_forces, double *d_out, uint32_t flags,
                        int32_t *shape, int32_t *grid_blocks, double *ms);

/* Sets the pair-kernel shape (mixed 0 .. 2, FP64 0 .. 1, as the autotuner
   numbers them) and grid cap by hand.  The shape (autotuned or set) is the
   one of evaluations that form energies; a mixed evaluation whose energies
   nobody reads (a dynamics step before the last of a run) takes shape 0
   without energy sums, the fastest there. */
int pmnb_batch_set_shape(pmnb_batch *b, int32_t shape, int32_t grid_blocks);

/* pmnb_force_pairs for the whole batch at device positions d_positions:
   atom indices are the batch's (replicas in order of addition). */
int pmnb_batch_force_pairs(pmnb_batch *b, const double *d_positions,
                           int32_t *pairs, int64_t capacity,
                           int64_t *n_pairs);

/* In-cutoff pairs of all replicas in the last evaluation; waits for the
   device. */
int pmnb_batch_pair_count(pmnb_batch *b, int64_t *n_pairs);

/* Text of the last error on this batch ("" if none). */
const char *pmnb_batch_last_error(const pmnb_batch *b);

/* Frees the batch and its device memory. */
void pmnb_batch_destroy(pmnb_batch *b);

/* ------------------------------------------------------------ leapfrog
   Device-resident LeapfrogIntegrator::run (integrator/leapfrog/
   LeapfrogIntegrator.cpp:85-99) whose force group is this context's cutoff
   force, with the modifiers ModifierShake (attached post-drift,
   modifier/ModifierShake.cpp:16-89) and ModifierRattle (attached post-step,
   modifier/ModifierRattle.cpp:17-71) as module/ModifierModule.cpp:104-127
   attaches them.  The step sequence, the kicks and drifts
   (integrator/StandardIntegrator.cpp:91-116, STSIntegrator.cpp:44-48) and
   each SHAKE/RATTLE sweep follow the reference operation for operation in
   FP64; the sweep count is the reference's (mean error over all
   constraints <= epsilon, at most max_iter + 1 sweeps).  The constraint
   virial the modifiers add (ScalarStructure::addVirial) is not computed. */
typedef struct pmnb_leapfrog pmnb_leapfrog;

typedef struct {
  double timestep;        /* fs (STSIntegrator::myTimestep) */
  int32_t shake;          /* ModifierShake attached */
  int32_t rattle;         /* ModifierRattle attached */
  double shake_epsilon;   /* shakeEpsilon (default 1e-5) */
  int32_t shake_max_iter; /* shakeMaxIter (default 30) */
  int32_t rattle_max_iter;/* rattleMaxIter (default 30) */
  double rattle_epsilon;  /* rattleEpsilon (default 1e-5) */
} pmnb_lf_params;

/* Creates the integrator over ctx's system: mass[n_atoms] is
   Atom::scaledMass; constraints are (atom1, atom2) pairs with rest lengths
   in the order of GenericTopology::bondRattleShakeConstraints
   (topology/TopologyUtilities.cpp:694-775).  A connected set of
   constraints may span at most 4 atoms and 6 constraints
   (PMNB_ERR_UNSUPPORTED otherwise).  SHAKE's previous positions start as
   the positions given to the first run (ModifierMetaRattleShake::
   doInitialize). */
int pmnb_lf_create(pmnb_ctx *ctx, const pmnb_lf_params *params,
                   int32_t n_atoms, const double *mass, int32_t n_constraints,
                   const int32_t *constraint_atoms, const double *rest_length,
                   pmnb_leapfrog **out);

/* LeapfrogIntegrator::run(n_steps) on device arrays, enqueued on `stream`
   (NULL = the context's stream) without host synchronisation.
   d_positions, d_velocities: 3N doubles, updated.  d_forces: the forces at
   d_positions on entry (as after initializeForces), at the final positions
   on exit.  d_out: the 20-double block of pmnb_evaluate_device (22
   doubles when bonded terms are set: [20] bond, [21] angle energy),
   holding the last evaluation's energies on exit (calculateForces clears
   them before every evaluation).  flags: as pmnb_evaluate_device, without
   PMNB_FLAG_MOLVIRIAL; the virial, like the energies, is the last
   evaluation's (the cutoff force's energy sums and virial are formed for
   that evaluation only, the only one whose values reach d_out). */
int pmnb_lf_run_device(pmnb_leapfrog *lf, double *d_positions,
                       double *d_velocities, double *d_forces, double *d_out,
                       int32_t n_steps, uint32_t flags, void *stream);

/* Bonded terms evaluated with the cutoff force in every calculateForces of
   the integrator: BondSystemForce (force/bonded/BondSystemForce.h:89-122,
   E = k (r - r0)^2) over bond_atoms (2 per bond) and AngleSystemForce
   (force/bonded/AngleSystemForce.h:86-150, E = k (theta - theta0)^2 plus
   Urey-Bradley kub (d13 - r_ub)^2) over angle_atoms (3 per angle, theta0
   in radians; ub_k / ub_r0 may be NULL for none), with the minimum image of
   the context's box.  Replaces any earlier set; n = 0 for none.  Their
   virial is not computed. */
int pmnb_lf_set_bonded(pmnb_leapfrog *lf, int32_t n_bonds,
                       const int32_t *bond_atoms, const double *bond_k,
                       const double *bond_r0, int32_t n_angles,
                       const int32_t *angle_atoms, const double *angle_k,
                       const double *angle_theta0, const double *ub_k,
                       const double *ub_r0);

/* Adds k-space terms created on the integrator's context to its force
   group, evaluated after the cutoff force in every calculateForces: a PME
   (pmnb_pme_create) and/or a direct Ewald sum (pmnb_ewald_create); either
   may be NULL, and NULL for both removes them.  They must outlive the
   integrator's runs. */
int pmnb_lf_set_kspace(pmnb_leapfrog *lf, pmnb_pme *pme, pmnb_ewald *ewald);

/* StandardIntegrator::calculateForces of the integrator's force group:
   clears d_forces and d_out (20 or 22 doubles, as pmnb_lf_run_device),
   then evaluates the cutoff force and the bonded terms at d_positions.
   This is what initializeForces computes before the first run. */
int pmnb_lf_forces_device(pmnb_leapfrog *lf, const double *d_positions,
                          double *d_forces, double *d_out, uint32_t flags,
                          void *stream);

/* n_replays replays of one captured HIP graph of
   pmnb_lf_run_device(n_steps) on the same arrays: the result is that of
   n_replays consecutive pmnb_lf_run_device(n_steps) calls, bit for bit
   (LeapfrogIntegrator::run called n_replays times, as the reference's
   driver does between outputs).  The graph is captured on the first call
   and again whenever the arrays, flags, n_steps, the topology, the box or
   the kernel shape change; every allocation and upload happens before the
   capture.  Enqueued on `stream` without host synchronisation.  Not with
   k-space terms (PMNB_ERR_UNSUPPORTED) or kernel timing on
   (PMNB_ERR_STATE).  No counterpart in the reference: a launch-overhead
   form of run for callers that step many times. */
int pmnb_lf_run_graph(pmnb_leapfrog *lf, double *d_positions,
                      double *d_velocities, double *d_forces, double *d_out,
                      int32_t n_steps, int32_t n_replays, uint32_t flags,
                      void *stream);

/* The integrator over a batch of replicas (pmnb_batch): each replica runs
   LeapfrogIntegrator::run with its own SHAKE / RATTLE sweep counts (its
   mean error over its own constraints), so every replica's trajectory is,
   bit for bit, the one pmnb_lf_create + pmnb_lf_run_device give for that
   replica alone with the same force.  Arrays are the batch's (atoms of
   the replicas in order of addition; constraint atoms are batch indices
   and may not join two replicas); d_out holds 20 doubles per replica.
   pmnb_lf_run_device, pmnb_lf_run_graph, pmnb_lf_forces_device,
   pmnb_lf_iterations (n_replicas counts per call, [call * n_replicas +
   replica]) and pmnb_lf_destroy apply; bonded and k-space terms are not
   taken.  Adding replicas to the batch afterwards invalidates it.  No
   counterpart in the reference (one integrator per process there). */
int pmnb_batch_lf_create(pmnb_batch *b, const pmnb_lf_params *params,
                         int32_t n_atoms, const double *mass,
                         int32_t n_constraints,
                         const int32_t *constraint_atoms,
                         const double *rest_length, pmnb_leapfrog **out);

/* Starts (capacity > 0) or stops (0) recording the sweep count of every
   SHAKE and RATTLE call, and returns what was recorded since the last
   start: up to `capacity_out` counts into iters (may be NULL), the number
   of calls and the number that reached max_iter + 1 sweeps (the
   reference's "not converged" warning).  Waits for the device.  A batch's
   integrator records one count per replica and call. */
int pmnb_lf_iterations(pmnb_leapfrog *lf, int64_t capacity, int32_t *iters,
                       int64_t capacity_out, int64_t *n_calls,
                       int64_t *n_unconverged);

/* Releases the integrator; may be called before or after pmnb_destroy of
   its context (which must outlive every other pmnb_lf_* call). */
void pmnb_lf_destroy(pmnb_leapfrog *lf);

/* ------------------------------------------------------- Ewald k-space
   The reciprocal-space and correction terms of
   NonbondedFullEwaldSystemForce<PeriodicBoundaryConditions, ...,
   real = false, reciprocal = true, correction = true>
   (force/nonbonded/NonbondedFullEwaldSystemForce.h:324-398): the
   reciprocal sum over the k-vectors |k| < kc of the h >= 0 half space
   (:884-1086), the intramolecular correction over the exclusion table
   (:1092-1130), the point self energy and the charged-system energy
   (:680-700), all added to the electrostatic energy.  With the cutoff
   force's Ewald real-space functor it completes the Ewald sum.
   alpha <= 0 selects the reference's default alpha (:603-604); kc =
   2 alpha sqrt(-ln accuracy).  The box is the context's (orthogonal).
   scaled_charge and coulomb_scaling_factor are those given to
   pmnb_set_topology.  The correction walks the rows of
   ExclusionTable::getTable() (topology/ExclusionTable.cpp:103-110): of the
   rows given, those of an atom with itself and of class PMNB_EXCL_NONE
   are dropped and a pair given again keeps the class of its first row, so
   the rows given to pmnb_set_topology may be passed as they are.  The
   k-vectors, charges and rows are fixed at creation: after pmnb_set_box
   with another box or any pmnb_set_topology call, evaluate returns
   PMNB_ERR_STATE until the terms are created again.  The work is atoms x
   k-vectors, as in the reference. */

int pmnb_ewald_create(pmnb_ctx *ctx, double alpha, double accuracy,
                      int32_t n_atoms, const double *scaled_charge,
                      double coulomb_scaling_factor, int32_t n_excl,
                      const int32_t *excl, pmnb_ewald **out);

/* alpha, real-space cutoff sqrt(-ln accuracy)/alpha, kc and the number of
   k-vectors (pointers may be NULL). */
int pmnb_ewald_info(const pmnb_ewald *ew, double *alpha, double *rc,
                    double *kc, int32_t *n_kvectors);

/* Enqueued on `stream` (NULL = the context's stream): forces += ,
   d_out[1] (electrostatic energy) += , d_out[2..10] (atomic virial) +=
   with PMNB_FLAG_VIRIAL.  PMNB_FLAG_MOLVIRIAL is not supported. */
int pmnb_ewald_evaluate_device(pmnb_ewald *ew, const double *d_positions,
                               double *d_forces, double *d_out,
                               uint32_t flags, void *stream);

void pmnb_ewald_destroy(pmnb_ewald *ew);

/* ---------------------------------------------------------------- PME
   Smooth particle-mesh Ewald: the reciprocal term of
   NonbondedPMEwaldSystemForce<PeriodicBoundaryConditions, ..., real =
   false, reciprocal = true, correction = true, BSpline>
   (mdl/interface/force/nonbonded/NonbondedPMEwaldSystemForce.h:815-862,
   Grid.h, BSpline.cpp) with its correction, point self and charged-system
   terms (:878-912, :593-612), all added to the electrostatic energy; with
   the cutoff force's Ewald real-space functor it is the PME sum.
   Parameters as the force line's (:1000-1011): grid nx ny nz, B-spline
   order (even, 2..8 here, each size >= order), cutoff rc (sets the default
   alpha), accuracy, alpha (<= 0: the reference's default, erfc(alpha rc) /
   rc = accuracy by bisection).  The charge grid is summed as 64-bit
   fixed-point integers (2^-32), so results repeat bit for bit.  The
   correction rows follow pmnb_ewald_create's rule.  Fixed at creation like
   pmnb_ewald (PMNB_ERR_STATE after a box or topology change). */

int pmnb_pme_create(pmnb_ctx *ctx, int32_t nx, int32_t ny, int32_t nz,
                    int32_t order, double cutoff, double accuracy,
                    double alpha, int32_t n_atoms,
                    const double *scaled_charge,
                    double coulomb_scaling_factor, int32_t n_excl,
                    const int32_t *excl, pmnb_pme **out);

/* alpha and kc = 2 alpha sqrt(-ln accuracy) (pointers may be NULL). */
int pmnb_pme_info(const pmnb_pme *pme, double *alpha, double *kc);

/* Enqueued on `stream` (NULL = the context's stream): forces +=,
   d_out[1] +=, d_out[2..10] += atomic virial with PMNB_FLAG_VIRIAL. */
int pmnb_pme_evaluate_device(pmnb_pme *pme, const double *d_positions,
                             double *d_forces, double *d_out, uint32_t flags,
                             void *stream);

void pmnb_pme_destroy(pmnb_pme *pme);

#ifdef __cplusplus
}
#endif

#endif /* PMNB_H */
