"""Host-side mirror of the parts of ProtoMol's topology the cutoff force reads.

Paths are relative to the reference's protomol/src/protomol.

* ``PeriodicBoundaryConditions``  topology/PeriodicBoundaryConditions.cpp:32-77
* ``CubicCellManager``            topology/CubicCellManager.cpp:19-29
* ``ExclusionTable``              topology/ExclusionTable.{h,cpp}
* ``LennardJonesParameterTable``  topology/LennardJonesParameterTable.h:34-36
* ``Topology``                    topology/Topology.h + GenericTopology.h
* ``build_exclusion_table``       topology/BuildTopology.cpp:1098-1155
* ``lorentz_berthelot_table``     topology/BuildTopology.cpp:650-661 (GROMACS)
* ``charmm_table``                topology/BuildTopology.cpp:663-673 (CHARMM)
* ``ScalarStructure``             type/ScalarStructure.h:26-137
"""
import numpy as np

SQRT_COULOMB_CONSTANT = 18.2226123264  # base/PMConstants.cpp:64

EXCLUSION_NONE = 0
EXCLUSION_MODIFIED = 1
EXCLUSION_FULL = 2

# ExclusionType values (topology/ExclusionType.h:15-24).
EXCLUDE_NONE = "none"
EXCLUDE_ONE2 = "1-2"
EXCLUDE_ONE3 = "1-3"
EXCLUDE_ONE4 = "1-4"
EXCLUDE_SCALED14 = "scaled1-4"


class PeriodicBoundaryConditions:
    """Orthogonal or general periodic box given by e1, e2, e3 and origin."""

    def __init__(self, e1, e2, e3, origin=(0.0, 0.0, 0.0)):
        self.set(e1, e2, e3, origin)

    def set(self, e1, e2, e3, origin):
        self.e1 = np.asarray(e1, dtype=np.float64).reshape(3)
        self.e2 = np.asarray(e2, dtype=np.float64).reshape(3)
        self.e3 = np.asarray(e3, dtype=np.float64).reshape(3)
        self.origin = np.asarray(origin, dtype=np.float64).reshape(3)
        vol = abs(np.dot(np.cross(self.e1, self.e2), self.e3))
        if vol < 1e-14:
            raise ValueError("[PeriodicBoundaryConditions::set] No volume")

    @classmethod
    def cubic(cls, length, origin=None):
        if origin is None:
            origin = (0.5 * length,) * 3
        return cls((length, 0, 0), (0, length, 0), (0, 0, length), origin)

    def is_orthogonal(self):
        return not (self.e1[1] or self.e1[2] or self.e2[0] or self.e2[2]
                    or self.e3[0] or self.e3[1])

    def lengths(self):
        return np.array([self.e1[0], self.e2[1], self.e3[2]])

    def volume(self):
        return float(abs(np.dot(np.cross(self.e1, self.e2), self.e3)))


class CubicCellManager:
    """Cell size of the linked-cell grid (``cellsize`` keyword)."""

    def __init__(self, cell_size):
        if not cell_size > 0.0:
            raise ValueError("Cubic cellmanager: cutoff > 0")
        self.cell_size = float(cell_size)


class ExclusionTable:
    """Ordered record of exclusion adds; a pair added twice keeps the class
    added last, as the reference's lookup table does."""

    def __init__(self, n_atoms=0):
        self.resize(n_atoms)

    def resize(self, n_atoms):
        self.n_atoms = int(n_atoms)
        self.clear()

    def clear(self):
        self._chunks = []
        self._lookup = None
        self.max_delta = 0

    def add(self, a1, a2, cls):
        self.add_many([(a1, a2)], cls)

    def add_many(self, pairs, cls):
        p = np.asarray(pairs, dtype=np.int64).reshape(-1, 2)
        if not len(p):
            return
        a = np.minimum(p[:, 0], p[:, 1])
        b = np.maximum(p[:, 0], p[:, 1])
        rows = np.stack([a, b, np.full(len(a), int(cls))], axis=1)
        self._chunks.append(rows.astype(np.int32))
        self._lookup = None
        self.max_delta = max(self.max_delta, int((b - a).max()))

    def check(self, a1, a2):
        if self._lookup is None:
            self._lookup = {}
            for x, y, c in self.table().tolist():
                if x != y:
                    self._lookup[(x, y)] = c
        if a1 > a2:
            a1, a2 = a2, a1
        return self._lookup.get((int(a1), int(a2)), EXCLUSION_NONE)

    def table(self):
        """(n, 3) int32 rows (a1, a2, class) of every add, in insertion
        order (the lookup resolves a repeated pair to its last class)."""
        if not self._chunks:
            return np.zeros((0, 3), dtype=np.int32)
        if len(self._chunks) > 1:
            self._chunks = [np.concatenate(self._chunks)]
        return self._chunks[0]

    def get_table(self):
        """ExclusionTable::getTable() (topology/ExclusionTable.h:87): the
        rows ExclusionTable::add stores (ExclusionTable.cpp:103-110), one
        per pair with the class of its FIRST add (its set compares atoms
        only), no rows of one atom with itself and no EXCLUSION_NONE rows,
        in insertion order.  The Ewald correction term iterates this view
        (force/nonbonded/NonbondedFullEwaldSystemForce.h:1105-1129)."""
        rows = self.table()
        keep = (rows[:, 0] != rows[:, 1]) & (rows[:, 2] != EXCLUSION_NONE)
        rows = rows[keep]
        if not len(rows):
            return np.zeros((0, 3), dtype=np.int32)
        key = rows[:, 0].astype(np.int64) * max(self.n_atoms, 1) + rows[:, 1]
        _, first = np.unique(key, return_index=True)
        return np.ascontiguousarray(rows[np.sort(first)], dtype=np.int32)

    @classmethod
    def from_array(cls, n_atoms, rows):
        t = cls(n_atoms)
        rows = np.asarray(rows, dtype=np.int64).reshape(-1, 3)
        if len(rows):
            a = np.minimum(rows[:, 0], rows[:, 1])
            b = np.maximum(rows[:, 0], rows[:, 1])
            t._chunks = [np.stack([a, b, rows[:, 2]], 1).astype(np.int32)]
            t.max_delta = int((b - a).max())
        return t


class LennardJonesParameterTable:
    """Symmetric table of {A, B, A14, B14} per type pair."""

    def __init__(self, n_types=0):
        self.resize(n_types)

    def resize(self, n_types):
        self.table = np.zeros((n_types, n_types, 4), dtype=np.float64)

    def set(self, i, j, A, B, A14, B14):
        self.table[i, j] = (A, B, A14, B14)
        self.table[j, i] = (A, B, A14, B14)

    @property
    def n_types(self):
        return self.table.shape[0]


def lorentz_berthelot_table(sigma, epsilon, lj_scaling_factor=1.0):
    """GROMACS branch of BuildTopology.cpp:650-661: sigma_ij arithmetic mean,
    eps_ij geometric mean, A = 4 eps sigma^12, B = 4 eps sigma^6,
    A14 = fudgeLJ * A, B14 = fudgeLJ * B."""
    sigma = np.asarray(sigma, dtype=np.float64)
    epsilon = np.asarray(epsilon, dtype=np.float64)
    t = LennardJonesParameterTable(len(sigma))
    for i in range(len(sigma)):
        for j in range(i, len(sigma)):
            r = 0.5 * (sigma[i] + sigma[j])
            e = np.sqrt(epsilon[i] * epsilon[j])
            r2 = r * r
            r3 = r2 * r
            r6 = r3 * r3
            r12 = r6 * r6
            A = r12 * e * 4.0
            B = r6 * e * 4.0
            t.set(i, j, A, B, lj_scaling_factor * A, lj_scaling_factor * B)
    return t


def charmm_table(rmin_half, epsilon, rmin14_half=None, epsilon14=None):
    """CHARMM branch of BuildTopology.cpp:663-673: r_ij = Rmin/2_i + Rmin/2_j,
    eps_ij = sqrt(eps_i eps_j), A = r^12 eps, B = 2 r^6 eps."""
    rmin_half = np.asarray(rmin_half, dtype=np.float64)
    epsilon = np.asarray(epsilon, dtype=np.float64)
    rmin14_half = rmin_half if rmin14_half is None else np.asarray(rmin14_half)
    epsilon14 = epsilon if epsilon14 is None else np.asarray(epsilon14)
    t = LennardJonesParameterTable(len(rmin_half))

    def p12(x):  # power<12>: square(power<6>) with power<6> = square(power<3>)
        x3 = (x * x) * x
        x6 = x3 * x3
        return x6 * x6

    def p6(x):
        x3 = (x * x) * x
        return x3 * x3

    for i in range(len(rmin_half)):
        for j in range(i, len(rmin_half)):
            r = rmin_half[i] + rmin_half[j]
            e = np.sqrt(epsilon[i] * epsilon[j])
            r14 = rmin14_half[i] + rmin14_half[j]
            e14 = np.sqrt(epsilon14[i] * epsilon14[j])
            t.set(i, j, p12(r) * e, 2 * p6(r) * e, p12(r14) * e14,
                  2 * p6(r14) * e14)
    return t


def build_exclusion_table(n_atoms, exclude, bonds=(), angles=(),
                          dihedrals=()):
    """buildExclusionTable (BuildTopology.cpp:1098-1155): bond ends FULL,
    then (unless 1-2) angle ends FULL, then (unless 1-2 or 1-3 exclusion)
    dihedral ends FULL for 1-4 or MODIFIED for scaled1-4, in that add
    order and unconditionally: in a ring a pair may be added more than
    once, the lookup then gives its last class and getTable() its first."""
    t = ExclusionTable(n_atoms)
    if exclude == EXCLUDE_NONE:
        return t
    bonds = np.asarray(bonds, dtype=np.int64).reshape(-1, 2)
    t.add_many(bonds, EXCLUSION_FULL)
    if exclude != EXCLUDE_ONE2:
        angles = np.asarray(angles, dtype=np.int64).reshape(-1, 3)
        t.add_many(angles[:, [0, 2]], EXCLUSION_FULL)
        if exclude != EXCLUDE_ONE3:
            dih = np.asarray(dihedrals, dtype=np.int64).reshape(-1, 4)
            cls = EXCLUSION_FULL if exclude == EXCLUDE_ONE4 else EXCLUSION_MODIFIED
            t.add_many(dih[:, [0, 3]], cls)
    return t


class Topology:
    """Atoms (type, scaled charge, molecule), LJ table, exclusions, box and
    cell manager: the inputs of Topology<PeriodicBoundaryConditions,
    CubicCellManager> that the cutoff force reads."""

    def __init__(self, types, charges, molecules, lj_table, exclusions,
                 boundary_conditions, cell_manager,
                 coulomb_scaling_factor=1.0, masses=None, scaled=False):
        self.types = np.ascontiguousarray(types, dtype=np.int32)
        q = np.asarray(charges, dtype=np.float64)
        # GenericTopology::atoms[i].scaledCharge = q * sqrt(332.0636)
        self.scaled_charges = np.ascontiguousarray(
            q if scaled else q * SQRT_COULOMB_CONSTANT)
        self.molecules = np.ascontiguousarray(molecules, dtype=np.int32)
        self.lennard_jones_parameters = lj_table
        self.exclusions = exclusions
        self.boundary_conditions = boundary_conditions
        self.cell_manager = cell_manager
        self.coulomb_scaling_factor = float(coulomb_scaling_factor)
        self.masses = None if masses is None else np.asarray(masses, np.float64)
        self.version = 0

    @property
    def n_atoms(self):
        return len(self.types)

    @property
    def n_molecules(self):
        return int(self.molecules.max()) + 1 if len(self.molecules) else 0

    def touch(self):
        """Marks the topology as changed (forces re-upload it)."""
        self.version += 1

    def molecule_centers(self, positions):
        """Centre of mass per molecule (Molecule::position), plain average
        when no masses are set; used for the molecular virial."""
        pos = np.asarray(positions, dtype=np.float64).reshape(-1, 3)
        m = self.masses if self.masses is not None else np.ones(len(pos))
        nm = self.n_molecules
        tot = np.zeros(nm)
        np.add.at(tot, self.molecules, m)
        c = np.zeros((nm, 3))
        for k in range(3):
            np.add.at(c[:, k], self.molecules, m * pos[:, k])
        return c / np.maximum(tot, 1e-300)[:, None]


class ScalarStructure:
    """Energy and virial slots (type/ScalarStructure.h:26-67)."""

    COULOMB = 0
    LENNARDJONES = 1
    VIRIALXX = 7
    MOLVIRIALXX = 16
    LAST = 36

    def __init__(self, virial=False, molecular_virial=False):
        self.table = np.zeros(self.LAST, dtype=np.float64)
        self._virial = bool(virial)
        self._molvirial = bool(molecular_virial)

    def clear(self):
        self.table[:] = 0.0

    def virial(self, flag=None):
        old = self._virial
        if flag is not None:
            self._virial = bool(flag)
        return old

    def molecular_virial(self, flag=None):
        old = self._molvirial
        if flag is not None:
            self._molvirial = bool(flag)
        return old

    def __getitem__(self, i):
        return self.table[i]

    def __setitem__(self, i, v):
        self.table[i] = v

    def virial_tensor(self):
        return self.table[self.VIRIALXX:self.VIRIALXX + 9].reshape(3, 3)

    def molecular_virial_tensor(self):
        return self.table[self.MOLVIRIALXX:self.MOLVIRIALXX + 9].reshape(3, 3)
