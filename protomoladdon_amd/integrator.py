"""LeapfrogIntegrator with SHAKE / RATTLE on MI355X, the cutoff force as its
force group, through the C ABI in include/pmnb.h (``pmnb_lf_*``).

Reference interface mirrored (paths relative to protomol/src/protomol):
* ``LeapfrogIntegrator(timestep)``, ``initialize`` and ``run(n)``:
  integrator/leapfrog/LeapfrogIntegrator.cpp:18-99 (``initialize`` computes
  the forces at the start positions, StandardIntegrator.cpp:38-42).
* ``ModifierShake(epsilon, max_iter)`` attached post-drift and
  ``ModifierRattle(epsilon, max_iter)`` attached post-step, only when
  epsilon > 0 and max_iter > 0: module/ModifierModule.cpp:104-127; the
  configuration defaults are epsilon 1e-5 and 30 iterations
  (ModifierModule.cpp:42-47).
* ``build_rattle_shake_constraints``: the constraint list of
  topology/TopologyUtilities.cpp:694-775 (per H-O-H angle the H-H, H1-O and
  H2-O distances; with ``all`` every other bond that has a hydrogen).
* ``kinetic_energy``: topology/TopologyUtilities.cpp:160-170.

The whole loop runs on the device: positions, velocities and forces stay
in HBM, and ``run`` enqueues the steps without host synchronisation.
"""
import ctypes
import math

import numpy as np

from . import _lib

TIMEFACTOR = 48.88821290839616  # base/PMConstants.cpp:50


def degrees_to_radians(deg):
    """dtor (base/MathUtilities.h:126-128)."""
    return deg * math.pi / 180.0


def build_rattle_shake_constraints(atom_names, bonds, bond_rest, angles,
                                   angle_rest, all_bonds=True):
    """Returns (pairs (M, 2) int32, rest lengths (M,)) in the reference's
    order.  atom_names: per-atom names (only the first letter is read);
    bonds: (B, 2) atom pairs with rest lengths bond_rest; angles: (A, 3)
    with rest angles angle_rest in radians."""
    names = [str(n) for n in atom_names]
    bonds = np.asarray(bonds, dtype=np.int64).reshape(-1, 2)
    angles = np.asarray(angles, dtype=np.int64).reshape(-1, 3)
    mybonds = [[] for _ in names]   # BuildTopology.cpp:178-179
    for i, (a, b) in enumerate(bonds):
        mybonds[a].append(i)
        mybonds[b].append(i)
    pairs, rest, water_bonds = [], [], []
    for i, atoms in enumerate(angles):
        hyd, oxy, at, ato = 0, 0, [0, 0], 0
        for a in atoms:
            if names[a][0] == "H":
                if hyd < 2:
                    at[hyd] = int(a)
                hyd += 1
            if names[a][0] == "O":
                ato = int(a)
                oxy += 1
        if hyd == 2 and oxy == 1:
            b = float(bond_rest[mybonds[at[0]][0]])
            c = float(bond_rest[mybonds[at[1]][0]])
            alpha = float(angle_rest[i])
            pairs.append((at[0], at[1]))
            rest.append(math.sqrt(b * b + c * c - 2 * b * c * math.cos(alpha)))
            pairs.append((at[0], ato))
            rest.append(b)
            pairs.append((at[1], ato))
            rest.append(c)
            water_bonds += [mybonds[at[0]][0], mybonds[at[1]][0]]
    water_bonds.sort()
    if all_bonds:
        used = 0
        for i, (a, b) in enumerate(bonds):
            if water_bonds and water_bonds[used] == i:
                if used < len(water_bonds) - 1:
                    used += 1
            elif names[a][0] == "H" or names[b][0] == "H":
                pairs.append((int(a), int(b)))
                rest.append(float(bond_rest[i]))
    return (np.asarray(pairs, dtype=np.int32).reshape(-1, 2),
            np.asarray(rest, dtype=np.float64))


def topology_constraints(topology, all_bonds=True):
    """The constraint list of a topology that carries atom_names, bonds,
    bond_rest, angles and angle_rest (the synthetic water systems do)."""
    return build_rattle_shake_constraints(
        topology.atom_names, topology.bonds, topology.bond_rest,
        topology.angles, topology.angle_rest, all_bonds)


def kinetic_energy(masses, velocities):
    """sum 0.5 m v.v in kcal/mol (velocities in the integrator's units)."""
    v = np.asarray(velocities, dtype=np.float64).reshape(-1, 3)
    return 0.5 * float(np.sum(np.asarray(masses) * np.sum(v * v, axis=1)))


class LeapfrogIntegrator:
    """Leapfrog (velocity Verlet split) over one NonbondedCutoffSystemForce
    and, optionally, harmonic bonds and angles (the "force Bond, Angle"
    lines of a configuration).

    shake / rattle: (epsilon, max_iter) to attach the modifier, or None.
    bonded: dict with bonds (B, 2), bond_k, bond_r0, angles (A, 3), angle_k,
    angle_theta0 (radians) and optionally ub_k, ub_r0; or None.  With
    bonded terms the energy block has 22 slots ([20] bond, [21] angle).
    kspace: k-space terms on the same force (a NonbondedPMEwaldKSpace or a
    NonbondedFullEwaldKSpace) evaluated in every force evaluation, or None.
    """

    keyword = "Leapfrog"

    def __init__(self, timestep, force, shake=None, rattle=None, bonded=None,
                 kspace=None):
        if not timestep > 0:
            raise ValueError("Leapfrog: timestep > 0")
        self.timestep = float(timestep)
        self.force = force
        self.shake = shake
        self.rattle = rattle
        self.bonded = bonded
        self.kspace = kspace
        if kspace is not None and kspace.force is not force:
            raise ValueError("the k-space terms must belong to the same force")
        self._lf = None
        self._key = None

    @property
    def out_slots(self):
        return 22 if self.bonded is not None else 20

    def _params(self):
        p = _lib.PmnbLfParams()
        p.timestep = self.timestep
        if self.shake is not None:
            p.shake = 1
            p.shake_epsilon, p.shake_max_iter = self.shake
        if self.rattle is not None:
            p.rattle = 1
            p.rattle_epsilon, p.rattle_max_iter = self.rattle
        return p

    def _ensure(self, topology, constraints):
        self.force._ensure(topology)
        key = (id(topology), topology.version, id(constraints),
               self.force._ctx.value)
        if self._lf is not None and key == self._key:
            return
        self.close()
        lib = self.force._lib
        ctx = self.force._ctx
        if topology.masses is None:
            raise ValueError("the topology has no masses")
        mass = np.ascontiguousarray(topology.masses, dtype=np.float64)
        pairs, rest = constraints if constraints is not None else (
            np.zeros((0, 2), np.int32), np.zeros(0))
        pairs = np.ascontiguousarray(pairs, dtype=np.int32)
        rest = np.ascontiguousarray(rest, dtype=np.float64)
        prm = self._params()
        lf = ctypes.c_void_p()
        _lib.check(lib, ctx, lib.pmnb_lf_create(
            ctx, ctypes.byref(prm), topology.n_atoms, mass.ctypes.data,
            len(pairs), pairs.ctypes.data if len(pairs) else None,
            rest.ctypes.data if len(pairs) else None, ctypes.byref(lf)))
        self._lf = lf
        self._key = key
        self._constraints = constraints
        if self.kspace is not None:
            from .pme import NonbondedPMEwaldKSpace
            self.kspace._ensure(topology)
            if isinstance(self.kspace, NonbondedPMEwaldKSpace):
                pm, ew = self.kspace._pme, None
            else:
                pm, ew = None, self.kspace._ew
            _lib.check(lib, ctx, lib.pmnb_lf_set_kspace(lf, pm, ew))
        if self.bonded is not None:
            b = self.bonded
            arr = {k: np.ascontiguousarray(b[k], dtype=np.int32 if k in (
                "bonds", "angles") else np.float64) for k in b}
            self._bonded_arrays = arr
            ub = "ub_k" in arr
            _lib.check(lib, ctx, lib.pmnb_lf_set_bonded(
                lf, len(arr["bond_k"]), arr["bonds"].ctypes.data,
                arr["bond_k"].ctypes.data, arr["bond_r0"].ctypes.data,
                len(arr["angle_k"]), arr["angles"].ctypes.data,
                arr["angle_k"].ctypes.data, arr["angle_theta0"].ctypes.data,
                arr["ub_k"].ctypes.data if ub else None,
                arr["ub_r0"].ctypes.data if ub else None))

    def close(self):
        if self._lf is not None:
            self.force._lib.pmnb_lf_destroy(self._lf)
            self._lf = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def initialize(self, topology, d_positions, d_forces, d_out,
                   constraints=None, flags=0, stream=None):
        """StandardIntegrator::initializeForces: forces (and the energies in
        d_out) at the start positions; d_forces and d_out are cleared
        first (calculateForces)."""
        self._ensure(topology, constraints)
        lib = self.force._lib
        _lib.check(lib, self.force._ctx, lib.pmnb_lf_forces_device(
            self._lf, d_positions.data_ptr(), d_forces.data_ptr(),
            d_out.data_ptr(), int(flags),
            ctypes.c_void_p(stream) if stream is not None else None))

    def run(self, topology, d_positions, d_velocities, d_forces, d_out,
            n_steps, constraints=None, flags=0, stream=None):
        """LeapfrogIntegrator::run(n_steps) on device tensors (float64,
        (N, 3)); d_forces must hold the forces at d_positions."""
        self._ensure(topology, constraints)
        lib = self.force._lib
        _lib.check(lib, self.force._ctx, lib.pmnb_lf_run_device(
            self._lf, d_positions.data_ptr(), d_velocities.data_ptr(),
            d_forces.data_ptr(), d_out.data_ptr(), int(n_steps), int(flags),
            ctypes.c_void_p(stream) if stream is not None else None))

    def run_graph(self, topology, d_positions, d_velocities, d_forces, d_out,
                  n_steps, n_replays=1, constraints=None, flags=0,
                  stream=None):
        """n_replays calls of run(n_steps) as replays of one captured HIP
        graph (the same result, bit for bit, with one launch per replay)."""
        self._ensure(topology, constraints)
        lib = self.force._lib
        _lib.check(lib, self.force._ctx, lib.pmnb_lf_run_graph(
            self._lf, d_positions.data_ptr(), d_velocities.data_ptr(),
            d_forces.data_ptr(), d_out.data_ptr(), int(n_steps),
            int(n_replays), int(flags),
            ctypes.c_void_p(stream) if stream is not None else None))

    def record_iterations(self, capacity):
        """Starts (capacity > 0) or stops (0) recording SHAKE/RATTLE sweep
        counts; returns (counts, n_calls, n_unconverged) recorded since the
        previous call."""
        lib = self.force._lib
        cap_out = 1 << 20
        out = np.zeros(cap_out, dtype=np.int32)
        n_calls = ctypes.c_int64()
        n_unc = ctypes.c_int64()
        _lib.check(lib, self.force._ctx, lib.pmnb_lf_iterations(
            self._lf, int(capacity), out.ctypes.data, cap_out,
            ctypes.byref(n_calls), ctypes.byref(n_unc)))
        return out[:min(n_calls.value, cap_out)].copy(), n_calls.value, n_unc.value
