"""Input and output formats around the force path: the readers a ProtoMol
run starts from and the writers whose files the reference's regression
comparator reads.

Reference behaviour mirrored (paths relative to protomol/src/protomol):
* ``read_psf``: io/PSFReader.cpp:45-210 (!NATOM rows: number, segment,
  residue number, residue name, atom name, atom type, charge, mass;
  !NBOND pairs, !NTHETA triples, !NPHI / !NIMPHI quadruples; 1-based).
* ``read_pdb``: io/PDBReader.cpp (ATOM / HETATM records, coordinates in
  columns 31-54, atom name in columns 13-16).
* ``read_par``: io/PARReader.cpp:64-230, the CHARMM sections the cutoff
  force and the bonded terms read: BONDS (type1 type2 Kb b0), ANGLES
  (type1 type2 type3 Ktheta theta0 [Kub S0]) and NONBONDED (type, ignored,
  epsilon, Rmin/2 [, ignored, epsilon 1-4, Rmin/2 1-4]); a section starts
  at a line whose first four letters name it, '!' starts a comment.
* ``build_charmm_system``: topology/BuildTopology.cpp: scaled charges,
  CHARMM LJ table, bond and angle parameters by type (either direction),
  theta0 through dtor, exclusion table, molecules as bond-connected sets,
  degrees of freedom 3N - 3 (BuildTopology.cpp:121).
* ``write_xyz`` / ``read_xyz``: io/XYZWriter.cpp:34-60 (count line, then
  per atom the type name, a tab and three fields of width 24 with 15
  significant digits).
* ``EnergiesWriter``: output/OutputEnergies.cpp:41-105 (header file and
  one row per output step: time with 2 decimals, then 8 significant digits
  per column, width 14, the shadow energy with 16 digits in width 20).
* ``PDB_VELOCITY_SCALING``: base/PMConstants.cpp:68; velocities read from
  a PDB are divided by it (module/IOModule.cpp:119-124).
"""
import math
from dataclasses import dataclass, field

import numpy as np

from .topology import (EXCLUDE_SCALED14, CubicCellManager,
                       PeriodicBoundaryConditions, Topology,
                       build_exclusion_table, charmm_table)

PDB_VELOCITY_SCALING = 20.45482706   # base/PMConstants.cpp:68
BOLTZMANN = 0.001987191              # base/PMConstants.cpp:67, kcal/mol/K


# ------------------------------------------------------------------ PSF
@dataclass
class PSF:
    atom_names: list
    atom_types: list
    residue_names: list
    charges: np.ndarray
    masses: np.ndarray
    bonds: np.ndarray       # (B, 2), 0-based
    angles: np.ndarray      # (A, 3)
    dihedrals: np.ndarray   # (D, 4)
    impropers: np.ndarray   # (I, 4)

    @property
    def n_atoms(self):
        return len(self.atom_names)


def read_psf(path):
    tokens_by_line = [l.split() for l in open(path).read().splitlines()]
    atoms = None
    lists = {"!NBOND": (2, None), "!NTHETA": (3, None), "!NPHI": (4, None),
             "!NIMPHI": (4, None)}
    k = 0
    while k < len(tokens_by_line):
        t = tokens_by_line[k]
        key = next((w for w in t if w.startswith("!")), None)
        if key is None or not t or not t[0].lstrip("-").isdigit():
            k += 1
            continue
        count = int(t[0])
        key = key.rstrip(":").upper()
        if key.startswith("!NATOM"):
            atoms = tokens_by_line[k + 1:k + 1 + count]
            k += 1 + count
            continue
        for name, (width, _) in list(lists.items()):
            if key.startswith(name) and (key == name or key[len(name)] in ":"):
                vals = []
                j = k + 1
                while len(vals) < count * width and j < len(tokens_by_line):
                    vals += [int(v) for v in tokens_by_line[j]]
                    j += 1
                lists[name] = (width, np.array(vals[:count * width],
                                               dtype=np.int64).reshape(-1, width) - 1)
                k = j
                break
        else:
            k += 1
    if atoms is None:
        raise ValueError("%s: no !NATOM section" % path)

    def get(name, width):
        a = lists[name][1]
        return a if a is not None else np.zeros((0, width), dtype=np.int64)
    return PSF(atom_names=[a[4] for a in atoms],
               atom_types=[a[5] for a in atoms],
               residue_names=[a[3] for a in atoms],
               charges=np.array([float(a[6]) for a in atoms]),
               masses=np.array([float(a[7]) for a in atoms]),
               bonds=get("!NBOND", 2), angles=get("!NTHETA", 3),
               dihedrals=get("!NPHI", 4), impropers=get("!NIMPHI", 4))


# ------------------------------------------------------------------ PDB
def read_pdb(path):
    """(positions (N, 3), atom names) of the ATOM / HETATM records."""
    xyz, names = [], []
    for l in open(path):
        if l.startswith(("ATOM", "HETATM")):
            xyz.append([float(l[30:38]), float(l[38:46]), float(l[46:54])])
            names.append(l[12:16].strip())
    return np.array(xyz, dtype=np.float64).reshape(-1, 3), names


# ------------------------------------------------------------------ PAR
@dataclass
class CharmmParameters:
    bonds: dict = field(default_factory=dict)      # (t1, t2) -> (k, b0)
    angles: dict = field(default_factory=dict)     # (t1, t2, t3) -> (k, theta0 deg, kub, s0)
    nonbonded: dict = field(default_factory=dict)  # t -> (eps, rmin/2, eps14, rmin14/2)

    def bond(self, t1, t2):
        p = self.bonds.get((t1, t2)) or self.bonds.get((t2, t1))
        if p is None:
            raise KeyError("no bond parameters for %s-%s" % (t1, t2))
        return p

    def angle(self, t1, t2, t3):
        p = self.angles.get((t1, t2, t3)) or self.angles.get((t3, t2, t1))
        if p is None:
            raise KeyError("no angle parameters for %s-%s-%s" % (t1, t2, t3))
        return p


_SECTIONS = ("BOND", "ANGL", "DIHE", "IMPR", "NONB", "HBON", "NBFI", "CMAP",
             "END")


def read_par(path):
    par = CharmmParameters()
    section = None
    for line in open(path):
        s = line.split("!")[0].strip()
        if not s:
            continue
        w = s.split()
        head = w[0].upper()
        if head[:4] in _SECTIONS:
            section = head[:4]
            continue
        if head in ("REMARK", "SET", "READ") or head.startswith("*"):
            continue
        try:
            if section == "BOND" and len(w) >= 4:
                par.bonds.setdefault((w[0], w[1]), (float(w[2]), float(w[3])))
            elif section == "ANGL" and len(w) >= 5:
                kub, s0 = ((float(w[5]), float(w[6])) if len(w) >= 7 and
                           _is_number(w[5]) else (0.0, 0.0))
                par.angles.setdefault((w[0], w[1], w[2]),
                                      (float(w[3]), float(w[4]), kub, s0))
            elif section == "NONB" and len(w) >= 4 and _is_number(w[1]):
                eps, rmin = float(w[2]), float(w[3])
                if len(w) >= 7 and _is_number(w[5]):
                    e14, r14 = float(w[5]), float(w[6])
                else:
                    e14, r14 = eps, rmin
                par.nonbonded.setdefault(w[0], (eps, rmin, e14, r14))
        except ValueError:
            continue
    return par


def _is_number(s):
    try:
        float(s)
        return True
    except ValueError:
        return False


# ------------------------------------------------------------------ system
def degrees_to_radians(deg):
    """dtor (base/MathUtilities.h:126-128)."""
    return deg * math.pi / 180.0


def build_charmm_system(psf, par, box, origin=(0.0, 0.0, 0.0), cell_size=None,
                        exclude=EXCLUDE_SCALED14):
    """Topology of a PSF / PAR pair in an orthogonal periodic box (box: the
    three edge lengths), and the bonded terms as LeapfrogIntegrator takes
    them.  The topology carries atom_names, bonds, bond_rest, angles and
    angle_rest for the SHAKE / RATTLE constraint list."""
    types_by_name = sorted(set(psf.atom_types), key=psf.atom_types.index)
    tidx = np.array([types_by_name.index(t) for t in psf.atom_types],
                    dtype=np.int32)
    nb = np.array([par.nonbonded[t] for t in types_by_name])
    table = charmm_table(nb[:, 1], nb[:, 0], nb[:, 3], nb[:, 2])
    n = psf.n_atoms
    excl = build_exclusion_table(n, exclude, psf.bonds, psf.angles,
                                 psf.dihedrals)
    # molecules: bond-connected sets, numbered in order of first atom
    parent = list(range(n))

    def find(a):
        while parent[a] != a:
            parent[a] = parent[parent[a]]
            a = parent[a]
        return a
    for a, b in psf.bonds:
        parent[find(int(a))] = find(int(b))
    roots, mol = {}, np.zeros(n, dtype=np.int32)
    for i in range(n):
        mol[i] = roots.setdefault(find(i), len(roots))
    L = [float(v) for v in box]
    bc = PeriodicBoundaryConditions((L[0], 0, 0), (0, L[1], 0), (0, 0, L[2]),
                                    origin)
    topo = Topology(tidx, psf.charges, mol, table, excl, bc,
                    CubicCellManager(float(cell_size)), 1.0,
                    masses=psf.masses)
    at = psf.atom_types
    bp = np.array([par.bond(at[a], at[b]) for a, b in psf.bonds]).reshape(-1, 2)
    ap = np.array([par.angle(at[a], at[b], at[c]) for a, b, c in psf.angles]
                  ).reshape(-1, 4)
    theta0 = np.array([degrees_to_radians(t) for t in ap[:, 1]])
    topo.atom_names = list(psf.atom_names)
    topo.type_names = types_by_name
    topo.bonds = psf.bonds
    topo.bond_rest = bp[:, 1]
    topo.angles = psf.angles
    topo.angle_rest = theta0
    topo.degrees_of_freedom = 3 * n - 3
    bonded = dict(bonds=psf.bonds, bond_k=bp[:, 0], bond_r0=bp[:, 1],
                  angles=psf.angles, angle_k=ap[:, 0], angle_theta0=theta0,
                  ub_k=ap[:, 2], ub_r0=ap[:, 3])
    return topo, bonded


def temperature(kinetic, degrees_of_freedom):
    """TopologyUtilities temperature: 2 KE / (k_B DOF)."""
    return 2.0 * kinetic / (BOLTZMANN * degrees_of_freedom)


# ------------------------------------------------------------------ XYZ
def write_xyz(path, names, coords):
    c = np.asarray(coords, dtype=np.float64).reshape(-1, 3)
    if len(names) != len(c):
        raise ValueError("coordinate and atom name counts differ")
    with open(path, "w") as f:
        f.write("%d\n" % len(c))
        for name, (x, y, z) in zip(names, c):
            f.write("%s\t%24.15g%24.15g%24.15g\n" % (name, x, y, z))


def read_xyz(path):
    lines = open(path).read().splitlines()
    n = int(lines[0].split()[0])
    rows = [l.split() for l in lines[1:] if len(l.split()) == 4]
    if len(rows) != n:
        raise ValueError("%s: %d atoms announced, %d read" % (path, n, len(rows)))
    return [r[0] for r in rows], np.array([[float(v) for v in r[1:]]
                                           for r in rows])


# ------------------------------------------------------------------ energies
ENERGY_COLUMNS = ("Time(fs)", "E_potential", "E_kinetic", "E_total",
                  "Temperature", "E_bond", "E_angle", "E_dihedral",
                  "E_improper", "E_VdW", "E_coulomb", "E_other", "Volume(A^3)")


class EnergiesWriter:
    """The all-energies file of OutputEnergies (no pressure columns)."""

    def __init__(self, path):
        self.path = path
        with open(path + ".header", "w") as h:
            h.write(" ".join("%14s" % c for c in ENERGY_COLUMNS) +
                    " %20s\n" % "E_shadow")
        self._f = open(path, "w")

    def write(self, time_fs, potential, kinetic, total, temperature_k, bond,
              angle, dihedral, improper, lj, coulomb, other, volume,
              shadow=0.0):
        vals = (potential, kinetic, total, temperature_k, bond, angle,
                dihedral, improper, lj, coulomb, other, volume)
        self._f.write("%14.2f " % time_fs +
                      " ".join("%14.8g" % v for v in vals) +
                      " %20.16g\n" % shadow)

    def close(self):
        self._f.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
