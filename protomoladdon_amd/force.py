"""NonbondedCutoffSystemForce on MI355X: the reference's force object, its
functors and switching functions, backed by the HIP library through the C
ABI in include/pmnb.h.

Reference interface mirrored (paths relative to protomol/src/protomol):
* ``NonbondedCutoffSystemForce(cutoff, one_atom_pair)`` and
  ``evaluate(topology, positions, forces, energies)``:
  force/nonbonded/NonbondedCutoffSystemForce.h:11-78,
  force/nonbonded/NonbondedCutoffForce.h:71-91 (``-cutoff`` last).
* ``OneAtomPairTwo(f1, sw1, f2, sw2)``: force/OneAtomPairTwo.h:33-41.
* ``LennardJonesForce``, ``CoulombForce``: force/LennardJonesForce.h,
  force/CoulombForce.h; ``EwaldRealForce(alpha)`` is the real-space term of
  force/nonbonded/NonbondedFullEwaldSystemForce.h:810-821 as a cutoff functor.
* ``C2SwitchingFunction(switchon, cutoff)``, ``C1SwitchingFunction(cutoff)``:
  switch/C2SwitchingFunction.cpp:30-41 and switch/C1SwitchingFunction.cpp.

Errors follow the reference's convention of refusing invalid parameters at
construction (THROW in the make functions) and reporting runtime failures
as exceptions (``PmnbError``).
"""
import ctypes

import numpy as np

from . import _lib
from .topology import ScalarStructure


class LennardJonesForce:
    keyword = "LennardJones"


class CoulombForce:
    keyword = "Coulomb"


class EwaldRealForce:
    keyword = "EwaldReal"

    def __init__(self, alpha):
        if not alpha > 0.0:
            raise ValueError("EwaldReal: alpha > 0")
        self.alpha = float(alpha)


class C2SwitchingFunction:
    keyword = "C2"

    def __init__(self, switchon, cutoff):
        # C2SwitchingFunction::make (switch/C2SwitchingFunction.cpp:30-41)
        if not (switchon >= 0.0 and cutoff > 0.0 and switchon < cutoff):
            raise ValueError("C2 switching function: 0 <= switchon (=%r) < "
                             "cutoff (=%r)." % (switchon, cutoff))
        self.switchon = float(switchon)
        self.cutoff = float(cutoff)


class C1SwitchingFunction:
    keyword = "C1"

    def __init__(self, cutoff):
        if not cutoff > 0.0:
            raise ValueError("C1 switching function: cutoff > 0")
        self.switchon = 0.0
        self.cutoff = float(cutoff)


class OneAtomPairTwo:
    """LennardJonesForce with C2 switching, paired with an electrostatic
    functor (CoulombForce, EwaldRealForce or None) and its switch."""

    def __init__(self, lj=None, lj_switch=None, elec=None, elec_switch=None):
        self.lj = lj if lj is not None else LennardJonesForce()
        if not isinstance(lj_switch, C2SwitchingFunction):
            raise ValueError("the LJ term takes a C2 switching function")
        self.lj_switch = lj_switch
        self.elec = elec
        self.elec_switch = elec_switch if elec_switch is not None else lj_switch

    @property
    def squared_cutoff(self):
        return max(self.lj_switch.cutoff, self.elec_switch.cutoff) ** 2


PRECISIONS = {"fp64": _lib.PMNB_PREC_FP64, "mixed": _lib.PMNB_PREC_MIXED}


class NonbondedCutoffSystemForce:
    """Short-range nonbonded force over the cubic cell list on one GPU.

    ``evaluate`` accumulates (+=) into ``forces`` and ``energies`` exactly as
    the reference's ``SystemForce::evaluate`` does; ``evaluate_device`` is
    the same computation on device tensors without host round trips.
    """

    keyword = "NonbondedCutoff"

    def __init__(self, cutoff, one_atom_pair, precision="fp64", device=0):
        if not cutoff > 0.0:
            raise ValueError("NonbondedCutoff: -cutoff > 0")
        if precision not in PRECISIONS:
            raise ValueError("precision must be one of %s" % list(PRECISIONS))
        self.cutoff = float(cutoff)
        self.pair = one_atom_pair
        self.precision = precision
        self.device = int(device)
        self._lib = _lib.load()
        self._ctx = None
        self._topo_key = None
        self._box_key = None
        self._cell_size = None

    # ------------------------------------------------------------ context
    def _params(self, cell_size):
        p = self.pair
        prm = _lib.PmnbParams()
        if p.elec is None:
            prm.elec = _lib.PMNB_ELEC_NONE
        elif isinstance(p.elec, CoulombForce):
            prm.elec = _lib.PMNB_ELEC_COULOMB
        elif isinstance(p.elec, EwaldRealForce):
            prm.elec = _lib.PMNB_ELEC_EWALD_REAL
            prm.ewald_alpha = p.elec.alpha
        else:
            raise ValueError("unsupported electrostatic functor %r" % p.elec)
        prm.elec_switch = (_lib.PMNB_SWITCH_C1
                           if isinstance(p.elec_switch, C1SwitchingFunction)
                           else _lib.PMNB_SWITCH_C2)
        prm.precision = PRECISIONS[self.precision]
        prm.device = self.device
        prm.lj_switchon = p.lj_switch.switchon
        prm.lj_cutoff = p.lj_switch.cutoff
        prm.elec_switchon = p.elec_switch.switchon
        prm.elec_cutoff = p.elec_switch.cutoff
        prm.algorithm_cutoff = self.cutoff
        prm.cell_size = cell_size
        return prm

    def _ensure(self, topology):
        lib = self._lib
        cs = topology.cell_manager.cell_size
        if self._ctx is None or cs != self._cell_size:
            self.close()
            prm = self._params(cs)
            ctx = ctypes.c_void_p()
            _lib.check(lib, None, lib.pmnb_create(ctypes.byref(prm),
                                                  ctypes.byref(ctx)))
            self._ctx = ctx
            self._cell_size = cs
            self._topo_key = None
            self._box_key = None
        key = (id(topology), topology.version)
        if key != self._topo_key:
            t = topology
            lj = np.ascontiguousarray(t.lennard_jones_parameters.table,
                                      dtype=np.float64)
            ex = np.ascontiguousarray(t.exclusions.table(), dtype=np.int32)
            self._keep = (t.types, t.scaled_charges, t.molecules, lj, ex)
            _lib.check(lib, self._ctx, lib.pmnb_set_topology(
                self._ctx, t.n_atoms, t.types.ctypes.data,
                t.scaled_charges.ctypes.data, t.molecules.ctypes.data,
                lj.shape[0], lj.ctypes.data, t.coulomb_scaling_factor,
                len(ex), ex.ctypes.data if len(ex) else None))
            self._topo_key = key
        bc = topology.boundary_conditions
        bkey = tuple(np.concatenate([bc.e1, bc.e2, bc.e3, bc.origin]).tolist())
        if bkey != self._box_key:
            arrs = [np.ascontiguousarray(v, dtype=np.float64)
                    for v in (bc.e1, bc.e2, bc.e3, bc.origin)]
            _lib.check(lib, self._ctx, lib.pmnb_set_box(
                self._ctx, *[a.ctypes.data for a in arrs]))
            self._box_key = bkey

    def close(self):
        if self._ctx is not None:
            self._lib.pmnb_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------ evaluate
    def evaluate(self, topology, positions, forces, energies):
        """Adds this force's contribution to forces (N x 3 float64, C order)
        and to the LENNARDJONES / COULOMB / virial slots of energies."""
        self._ensure(topology)
        pos = np.ascontiguousarray(positions, dtype=np.float64)
        if pos.shape != (topology.n_atoms, 3):
            raise ValueError("positions must be (%d, 3)" % topology.n_atoms)
        if (forces.dtype != np.float64 or not forces.flags.c_contiguous
                or forces.shape != pos.shape):
            raise ValueError("forces must be a C-contiguous float64 (N, 3)")
        flags = 0
        if energies.virial():
            flags |= _lib.PMNB_FLAG_VIRIAL
        com = None
        if energies.molecular_virial():
            flags |= _lib.PMNB_FLAG_MOLVIRIAL
            com = np.ascontiguousarray(topology.molecule_centers(pos))
        e2 = np.zeros(2)
        vir = np.zeros(9)
        mvir = np.zeros(9)
        lib = self._lib
        _lib.check(lib, self._ctx, lib.pmnb_evaluate(
            self._ctx, pos.ctypes.data, forces.ctypes.data, e2.ctypes.data,
            vir.ctypes.data, mvir.ctypes.data,
            com.ctypes.data if com is not None else None, flags))
        energies[ScalarStructure.LENNARDJONES] += e2[0]
        energies[ScalarStructure.COULOMB] += e2[1]
        if flags:
            energies.table[ScalarStructure.VIRIALXX:
                           ScalarStructure.VIRIALXX + 9] += vir
        if flags & _lib.PMNB_FLAG_MOLVIRIAL:
            energies.table[ScalarStructure.MOLVIRIALXX:
                           ScalarStructure.MOLVIRIALXX + 9] += mvir

    def parallel_evaluate(self, topology, positions, forces, energies):
        """SystemForce::parallelEvaluate defaults to evaluate on one device
        (force/system/SystemForce.h:91-96)."""
        self.evaluate(topology, positions, forces, energies)

    def evaluate_device(self, topology, d_positions, d_forces, d_out,
                        flags=0, d_mol_com=None, stream=None):
        """Device-array evaluation; the arguments are torch tensors (or any
        object with data_ptr()) on this force's device.  d_out has 20 float64
        slots; everything accumulates."""
        self._ensure(topology)
        lib = self._lib
        sp = None
        if stream is not None:
            sp = ctypes.c_void_p(stream)
        _lib.check(lib, self._ctx, lib.pmnb_evaluate_device(
            self._ctx, d_positions.data_ptr(), d_forces.data_ptr(),
            d_out.data_ptr(),
            d_mol_com.data_ptr() if d_mol_com is not None else None,
            flags, sp))

    def evaluate_device_noalloc(self, topology, d_positions, d_forces, d_out,
                                flags=0, d_mol_com=None, stream=None):
        """evaluate_device without allocating: the per-evaluation arrays
        live in the library's static 64 MiB scratch on this device (one
        system on the tile path; calls overlapping on one GPU race)."""
        self._ensure(topology)
        lib = self._lib
        sp = ctypes.c_void_p(stream) if stream is not None else None
        _lib.check(lib, self._ctx, lib.pmnb_evaluate_device_noalloc(
            self._ctx, d_positions.data_ptr(), d_forces.data_ptr(),
            d_out.data_ptr(),
            d_mol_com.data_ptr() if d_mol_com is not None else None,
            flags, sp))

    def set_parameters_device(self, topology, d_lj=None, d_scaled_charges=None,
                              stream=None):
        """Replaces the device copy of the LJ table (T x T x 4 float64
        {A, B, A14, B14}) and/or the scaled charges (N float64) from device
        tensors; the host topology is left as it is (a later change of the
        topology uploads it again)."""
        self._ensure(topology)
        lib = self._lib
        _lib.check(lib, self._ctx, lib.pmnb_set_parameters_device(
            self._ctx, d_lj.data_ptr() if d_lj is not None else None,
            d_scaled_charges.data_ptr() if d_scaled_charges is not None
            else None, ctypes.c_void_p(stream) if stream else None))

    def adjoint_device(self, topology, d_positions, d_grad_forces,
                       d_grad_positions=None, d_grad_charges=None,
                       d_grad_lj=None, stream=None):
        """Accumulates dL/dx (= -H dL/dF), dL/d(scaled charges) and
        dL/d{A, B, A14, B14} for the upstream gradient d_grad_forces = dL/dF;
        the pairs are recomputed.  Device tensors; any output may be None."""
        self._ensure(topology)
        lib = self._lib

        def ptr(t):
            return t.data_ptr() if t is not None else None
        _lib.check(lib, self._ctx, lib.pmnb_adjoint_device(
            self._ctx, d_positions.data_ptr(), d_grad_forces.data_ptr(),
            ptr(d_grad_positions), ptr(d_grad_charges), ptr(d_grad_lj),
            ctypes.c_void_p(stream) if stream else None))

    def pair_set(self, topology, positions):
        """Sorted (P, 3) int32 rows (i, j, exclusion class), i < j."""
        self._ensure(topology)
        pos = np.ascontiguousarray(positions, dtype=np.float64)
        lib = self._lib
        n = ctypes.c_int64()
        _lib.check(lib, self._ctx, lib.pmnb_pairset(
            self._ctx, pos.ctypes.data, None, 0, ctypes.byref(n)))
        out = np.zeros((max(n.value, 1), 3), dtype=np.int32)
        _lib.check(lib, self._ctx, lib.pmnb_pairset(
            self._ctx, pos.ctypes.data, out.ctypes.data, n.value,
            ctypes.byref(n)))
        return out[:n.value]

    def force_pairs(self, topology, positions):
        """The pairs the force's own pair kernel evaluates at `positions`
        (one evaluation with its pair record on): sorted (P, 3) int32 rows
        (i, j, exclusion class), i < j."""
        self._ensure(topology)
        pos = np.ascontiguousarray(positions, dtype=np.float64)
        lib = self._lib
        n = ctypes.c_int64()
        _lib.check(lib, self._ctx, lib.pmnb_force_pairs(
            self._ctx, pos.ctypes.data, None, 0, ctypes.byref(n)))
        out = np.zeros((max(n.value, 1), 3), dtype=np.int32)
        _lib.check(lib, self._ctx, lib.pmnb_force_pairs(
            self._ctx, pos.ctypes.data, out.ctypes.data, n.value,
            ctypes.byref(n)))
        return out[:n.value]

    def pair_path(self, topology):
        """'tile', 'list' or 'scan': the pair kernels this force's
        evaluations take for the topology (pmnb_pair_path)."""
        self._ensure(topology)
        p = ctypes.c_int32()
        _lib.check(self._lib, self._ctx,
                   self._lib.pmnb_pair_path(self._ctx, ctypes.byref(p)))
        return ("tile", "list", "scan")[p.value]

    def kernel_timing(self, enable):
        """Starts (True) or stops (False) HIP-event timing of the pair stage;
        returns (build_ms, pair_ms, evaluations) recorded since timing was
        enabled."""
        ms = np.zeros(2)
        n = ctypes.c_int64()
        _lib.check(self._lib, self._ctx, self._lib.pmnb_kernel_timing(
            self._ctx, 1 if enable else 0, ms.ctypes.data, ctypes.byref(n)))
        return float(ms[0]), float(ms[1]), n.value

    def last_pair_count(self):
        n = ctypes.c_int64()
        _lib.check(self._lib, self._ctx,
                   self._lib.pmnb_last_pair_count(self._ctx, ctypes.byref(n)))
        return n.value

    def get_id(self):
        """getIdNoAlias (NonbondedCutoffForce.h:79-81) of the built force."""
        p = self.pair
        parts = [p.lj.keyword]
        if p.elec is not None:
            parts.append(p.elec.keyword)
        s = " ".join(parts) + " -switchingFunction " + p.lj_switch.keyword
        if p.elec is not None:
            s += " -switchingFunction " + p.elec_switch.keyword
        return s + " -algorithm " + self.keyword
