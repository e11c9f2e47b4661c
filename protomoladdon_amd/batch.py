"""Batched replicas: many independent periodic systems sharing one force.

The reference evaluates one system per process through
``NonbondedCutoffSystemForce::evaluate``
(force/nonbonded/NonbondedCutoffSystemForce.h:45-55); a batch evaluates
hundreds of such systems with one launch of each kernel, and can replay the
whole evaluation as one captured HIP graph.  Replica r's forces, energies and
virial are what ``NonbondedCutoffSystemForce.evaluate_device`` gives for that
replica alone.
"""
import ctypes

import numpy as np

from . import _lib
from .force import NonbondedCutoffSystemForce

OUT_SLOTS = 20  # per replica: LJ, electrostatic, virial 3x3, mol. virial 3x3


class BatchedCutoffForce:
    """One NonbondedCutoff force over a batch of replicas.

    Every replica's topology must share the LJ table and the coulomb scaling
    factor; atoms, molecules, exclusions and box are per replica.  Positions
    and forces are device arrays of the concatenated replicas (in order of
    ``add``); outputs hold ``OUT_SLOTS`` doubles per replica.
    """

    def __init__(self, cutoff, one_atom_pair, precision="mixed", device=0):
        self._proto = NonbondedCutoffSystemForce(cutoff, one_atom_pair,
                                                 precision=precision,
                                                 device=device)
        self._lib = _lib.load()
        self._b = None
        self._lj = None
        self._csf = None
        self._cell_size = None
        self.offsets = [0]   # first atom of each replica, then the total
        self._keep = []

    @classmethod
    def from_workloads(cls, workloads, precision="mixed", device=0):
        w0 = workloads[0]
        f = cls(w0.cutoff, w0.one_atom_pair(), precision=precision,
                device=device)
        for w in workloads:
            f.add(w.topology)
        return f

    @property
    def n_replicas(self):
        return len(self.offsets) - 1

    @property
    def n_atoms(self):
        return self.offsets[-1]

    def add(self, topology):
        """Appends a replica; returns its index."""
        lib = self._lib
        t = topology
        lj = np.ascontiguousarray(t.lennard_jones_parameters.table,
                                  dtype=np.float64)
        cs = t.cell_manager.cell_size
        if self._b is None:
            prm = self._proto._params(cs)
            b = ctypes.c_void_p()
            _lib.check(lib, None, lib.pmnb_batch_create(
                ctypes.byref(prm), lj.shape[0], lj.ctypes.data,
                t.coulomb_scaling_factor, ctypes.byref(b)))
            self._b = b
            self._lj = lj
            self._csf = t.coulomb_scaling_factor
            self._cell_size = cs
        elif (lj.shape != self._lj.shape or not np.array_equal(lj, self._lj)
              or t.coulomb_scaling_factor != self._csf):
            raise ValueError("replicas of a batch share the LJ table and the "
                             "coulomb scaling factor")
        ex = np.ascontiguousarray(t.exclusions.table(), dtype=np.int32)
        bc = t.boundary_conditions
        box = [np.ascontiguousarray(v, dtype=np.float64)
               for v in (bc.e1, bc.e2, bc.e3, bc.origin)]
        r = ctypes.c_int32()
        _lib.check_batch(lib, self._b, lib.pmnb_batch_add(
            self._b, t.n_atoms, t.types.ctypes.data,
            t.scaled_charges.ctypes.data, t.molecules.ctypes.data, len(ex),
            ex.ctypes.data if len(ex) else None,
            *[a.ctypes.data for a in box], ctypes.byref(r)))
        self.offsets.append(self.offsets[-1] + t.n_atoms)
        return r.value

    def _stream(self, stream):
        return ctypes.c_void_p(stream) if stream is not None else None

    def evaluate_device(self, d_positions, d_forces, d_out, flags=0,
                        stream=None):
        """One evaluation of every replica (accumulates)."""
        _lib.check_batch(self._lib, self._b, self._lib.pmnb_batch_evaluate_device(
            self._b, d_positions.data_ptr(), d_forces.data_ptr(),
            d_out.data_ptr(), flags, self._stream(stream)))

    def run_graph(self, d_positions, d_forces, d_out, flags=0, n_steps=1,
                  stream=None):
        """n_steps replays of the captured evaluation (each zeroes d_forces
        and d_out first)."""
        _lib.check_batch(self._lib, self._b, self._lib.pmnb_batch_run_graph(
            self._b, d_positions.data_ptr(), d_forces.data_ptr(),
            d_out.data_ptr(), flags, int(n_steps), self._stream(stream)))

    def autotune(self, d_positions, d_forces, d_out, flags=0):
        """Times every kernel shape and grid size and keeps the fastest;
        returns (shape, grid_blocks, ms per evaluation)."""
        sh = ctypes.c_int32()
        gb = ctypes.c_int32()
        ms = ctypes.c_double()
        _lib.check_batch(self._lib, self._b, self._lib.pmnb_batch_autotune(
            self._b, d_positions.data_ptr(), d_forces.data_ptr(),
            d_out.data_ptr(), flags, ctypes.byref(sh), ctypes.byref(gb),
            ctypes.byref(ms)))
        return sh.value, gb.value, ms.value

    def set_shape(self, shape, grid_blocks=0):
        _lib.check_batch(self._lib, self._b, self._lib.pmnb_batch_set_shape(
            self._b, int(shape), int(grid_blocks)))

    def force_pairs(self, d_positions):
        """The pairs the batch's pair kernel evaluates at the device
        positions (one evaluation with its pair record on): sorted (P, 3)
        int32 rows (i, j, class), i < j, indices into the batch's atoms."""
        lib = self._lib
        n = ctypes.c_int64()
        _lib.check_batch(lib, self._b, lib.pmnb_batch_force_pairs(
            self._b, d_positions.data_ptr(), None, 0, ctypes.byref(n)))
        out = np.zeros((max(n.value, 1), 3), dtype=np.int32)
        _lib.check_batch(lib, self._b, lib.pmnb_batch_force_pairs(
            self._b, d_positions.data_ptr(), out.ctypes.data, n.value,
            ctypes.byref(n)))
        return out[:n.value]

    def pair_count(self):
        n = ctypes.c_int64()
        _lib.check_batch(self._lib, self._b, self._lib.pmnb_batch_pair_count(
            self._b, ctypes.byref(n)))
        return n.value

    def close(self):
        if self._b is not None:
            self._lib.pmnb_batch_destroy(self._b)
            self._b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BatchedLeapfrog:
    """LeapfrogIntegrator with SHAKE / RATTLE over every replica of a
    BatchedCutoffForce (``pmnb_batch_lf_*``): each replica keeps its own
    sweep counts, so its trajectory is the one a LeapfrogIntegrator gives
    for that replica alone with the same force, bit for bit.  Positions,
    velocities and forces are the batch's concatenated device arrays;
    ``d_out`` holds OUT_SLOTS doubles per replica.

    constraints: per replica (pairs (M, 2) with replica-local indices, rest
    lengths (M,)), e.g. ``integrator.topology_constraints`` of each
    replica's topology; masses: per replica arrays.
    """

    def __init__(self, timestep, batch, masses, constraints, shake=None,
                 rattle=None):
        from .integrator import LeapfrogIntegrator
        if len(masses) != batch.n_replicas or len(constraints) != batch.n_replicas:
            raise ValueError("one mass array and one constraint list per replica")
        self.batch = batch
        self._lib = batch._lib
        proto = LeapfrogIntegrator(timestep, None, shake=shake, rattle=rattle)
        prm = proto._params()
        mass = np.ascontiguousarray(np.concatenate(masses), dtype=np.float64)
        pairs, rest = [], []
        for r, (pr, rl) in enumerate(constraints):
            pairs.append(np.asarray(pr, dtype=np.int64).reshape(-1, 2) +
                         batch.offsets[r])
            rest.append(np.asarray(rl, dtype=np.float64))
        pairs = np.ascontiguousarray(np.concatenate(pairs), dtype=np.int32)
        rest = np.ascontiguousarray(np.concatenate(rest), dtype=np.float64)
        if len(mass) != batch.n_atoms:
            raise ValueError("masses do not cover the batch's atoms")
        lf = ctypes.c_void_p()
        _lib.check_batch(self._lib, batch._b, self._lib.pmnb_batch_lf_create(
            batch._b, ctypes.byref(prm), batch.n_atoms, mass.ctypes.data,
            len(pairs), pairs.ctypes.data if len(pairs) else None,
            rest.ctypes.data if len(pairs) else None, ctypes.byref(lf)))
        self._lf = lf

    def _check(self, code):
        _lib.check_batch(self._lib, self.batch._b, code)

    def initialize(self, d_positions, d_forces, d_out, flags=0, stream=None):
        """Forces and energies at the start positions (cleared first)."""
        self._check(self._lib.pmnb_lf_forces_device(
            self._lf, d_positions.data_ptr(), d_forces.data_ptr(),
            d_out.data_ptr(), int(flags), self.batch._stream(stream)))

    def run(self, d_positions, d_velocities, d_forces, d_out, n_steps,
            flags=0, stream=None):
        """LeapfrogIntegrator::run(n_steps) of every replica, eagerly."""
        self._check(self._lib.pmnb_lf_run_device(
            self._lf, d_positions.data_ptr(), d_velocities.data_ptr(),
            d_forces.data_ptr(), d_out.data_ptr(), int(n_steps), int(flags),
            self.batch._stream(stream)))

    def run_graph(self, d_positions, d_velocities, d_forces, d_out, n_steps,
                  n_replays=1, flags=0, stream=None):
        """n_replays calls of run(n_steps) as replays of one captured HIP
        graph."""
        self._check(self._lib.pmnb_lf_run_graph(
            self._lf, d_positions.data_ptr(), d_velocities.data_ptr(),
            d_forces.data_ptr(), d_out.data_ptr(), int(n_steps),
            int(n_replays), int(flags), self.batch._stream(stream)))

    def record_iterations(self, capacity):
        """Starts (capacity > 0) or stops (0) recording sweep counts;
        returns (counts (calls, replicas), n_calls, n_unconverged)."""
        R = self.batch.n_replicas
        cap_out = 1 << 22
        out = np.zeros(cap_out, dtype=np.int32)
        n_calls = ctypes.c_int64()
        n_unc = ctypes.c_int64()
        self._check(self._lib.pmnb_lf_iterations(
            self._lf, int(capacity), out.ctypes.data, cap_out,
            ctypes.byref(n_calls), ctypes.byref(n_unc)))
        n = min(n_calls.value * R, cap_out) // R
        return out[:n * R].reshape(n, R).copy(), n_calls.value, n_unc.value

    def close(self):
        if self._lf is not None:
            self._lib.pmnb_lf_destroy(self._lf)
            self._lf = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
