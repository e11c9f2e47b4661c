"""Reciprocal-space and correction terms of ProtoMol's
NonbondedFullEwaldSystemForce on MI355X, through the C ABI in
include/pmnb.h (``pmnb_ewald_*``).

Reference: force/nonbonded/NonbondedFullEwaldSystemForce.h (paths relative
to protomol/src/protomol): ``alpha`` and ``accuracy`` are the force's
parameters (alpha <= 0 selects its default, :603-604); the k-vectors are
those with |k| < kc = 2 alpha sqrt(-ln accuracy) in the h >= 0 half space
(:626-652); the reciprocal sum (:884-1086), the intramolecular correction
over the exclusion table (:1092-1130), the point self energy and the
charged-system energy (:680-700) all add to the electrostatic energy.
Together with the cutoff force's Ewald real-space functor this is the
Ewald sum; the real-space part here is the cutoff force's (switched) one.
"""
import ctypes

import numpy as np

from . import _lib


class NonbondedFullEwaldKSpace:
    """The k-space half of the Ewald sum, on the GPU context of a
    NonbondedCutoffSystemForce (same device, box and atoms)."""

    def __init__(self, force, alpha=-1.0, accuracy=1e-5):
        if not 0.0 < accuracy < 1.0:
            raise ValueError("accuracy must lie in (0, 1)")
        self.force = force
        self.alpha = float(alpha)
        self.accuracy = float(accuracy)
        self._ew = None
        self._key = None

    def _ensure(self, topology):
        self.force._ensure(topology)
        key = (id(topology), topology.version, self.force._ctx.value,
               self.force._box_key)
        if self._ew is not None and key == self._key:
            return
        self.close()
        lib = self.force._lib
        q = np.ascontiguousarray(topology.scaled_charges, dtype=np.float64)
        ex = np.ascontiguousarray(topology.exclusions.get_table(),
                                dtype=np.int32)
        ew = ctypes.c_void_p()
        _lib.check(lib, self.force._ctx, lib.pmnb_ewald_create(
            self.force._ctx, self.alpha, self.accuracy, topology.n_atoms,
            q.ctypes.data, topology.coulomb_scaling_factor, len(ex),
            ex.ctypes.data if len(ex) else None, ctypes.byref(ew)))
        self._ew = ew
        self._key = key

    def info(self, topology):
        """(alpha, real-space cutoff, kc, number of k-vectors)."""
        self._ensure(topology)
        a, rc, kc = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        nk = ctypes.c_int32()
        self.force._lib.pmnb_ewald_info(self._ew, ctypes.byref(a),
                                        ctypes.byref(rc), ctypes.byref(kc),
                                        ctypes.byref(nk))
        return a.value, rc.value, kc.value, nk.value

    def evaluate_device(self, topology, d_positions, d_forces, d_out, flags=0,
                        stream=None):
        """forces += ; d_out[1] += electrostatic energy; d_out[2..10] +=
        atomic virial with flags & 1 (torch tensors on the force's
        device)."""
        self._ensure(topology)
        lib = self.force._lib
        _lib.check(lib, self.force._ctx, lib.pmnb_ewald_evaluate_device(
            self._ew, d_positions.data_ptr(), d_forces.data_ptr(),
            d_out.data_ptr(), int(flags),
            ctypes.c_void_p(stream) if stream is not None else None))

    def close(self):
        if self._ew is not None:
            self.force._lib.pmnb_ewald_destroy(self._ew)
            self._ew = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
