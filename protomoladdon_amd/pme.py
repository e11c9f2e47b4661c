"""Smooth particle-mesh Ewald (the reciprocal term of ProtoMol's
NonbondedPMEwaldSystemForce with its correction, point-self and
charged-system terms) on MI355X, through the C ABI in include/pmnb.h
(``pmnb_pme_*``).

Reference: mdl/interface/force/nonbonded/NonbondedPMEwaldSystemForce.h with
Grid.h and BSpline.cpp: the force line's -gridsize nx ny nz, -cutoff,
-order (default 4), -accuracy (default 1e-6) and -alpha (<= 0: the
reference's default, erfc(alpha rc)/rc = accuracy) (:1000-1058); the
reciprocal term (:815-862), the correction over the exclusion table
(:878-912) and the self and charged-system energies (:593-612) add to the
electrostatic energy.  With the cutoff force's Ewald real-space functor this
is the PME sum.
"""
import ctypes

import numpy as np

from . import _lib


class NonbondedPMEwaldKSpace:
    """The k-space half of the PME sum, on the GPU context of a
    NonbondedCutoffSystemForce (same device, box and atoms)."""

    def __init__(self, force, gridsize, cutoff, order=4, accuracy=1e-6,
                 alpha=-1.0):
        nx, ny, nz = (int(v) for v in gridsize)
        if order < 2 or order % 2 or min(nx, ny, nz) < order:
            raise ValueError("PME needs an even order >= 2 and grid sizes "
                             ">= order")
        self.force = force
        self.grid = (nx, ny, nz)
        self.order = int(order)
        self.cutoff = float(cutoff)
        self.accuracy = float(accuracy)
        self.alpha = float(alpha)
        self._pme = None
        self._key = None

    def _ensure(self, topology):
        self.force._ensure(topology)
        key = (id(topology), topology.version, self.force._ctx.value,
               self.force._box_key)
        if self._pme is not None and key == self._key:
            return
        self.close()
        lib = self.force._lib
        q = np.ascontiguousarray(topology.scaled_charges, dtype=np.float64)
        ex = np.ascontiguousarray(topology.exclusions.get_table(),
                                  dtype=np.int32)
        pm = ctypes.c_void_p()
        _lib.check(lib, self.force._ctx, lib.pmnb_pme_create(
            self.force._ctx, *self.grid, self.order, self.cutoff,
            self.accuracy, self.alpha, topology.n_atoms, q.ctypes.data,
            topology.coulomb_scaling_factor, len(ex),
            ex.ctypes.data if len(ex) else None, ctypes.byref(pm)))
        self._pme = pm
        self._key = key

    @property
    def handle(self):
        return self._pme

    def info(self, topology):
        """(alpha, kc)."""
        self._ensure(topology)
        a, kc = ctypes.c_double(), ctypes.c_double()
        self.force._lib.pmnb_pme_info(self._pme, ctypes.byref(a),
                                      ctypes.byref(kc))
        return a.value, kc.value

    def evaluate_device(self, topology, d_positions, d_forces, d_out, flags=0,
                        stream=None):
        """forces += ; d_out[1] += electrostatic energy; d_out[2..10] +=
        atomic virial with flags & 1 (torch tensors on the force's
        device)."""
        self._ensure(topology)
        lib = self.force._lib
        _lib.check(lib, self.force._ctx, lib.pmnb_pme_evaluate_device(
            self._pme, d_positions.data_ptr(), d_forces.data_ptr(),
            d_out.data_ptr(), int(flags),
            ctypes.c_void_p(stream) if stream is not None else None))

    def close(self):
        if self._pme is not None:
            self.force._lib.pmnb_pme_destroy(self._pme)
            self._pme = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
