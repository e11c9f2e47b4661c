"""Seeded synthetic systems for the cutoff-force workloads.

Each builder returns a ``Workload``: a ``Topology`` (types, scaled charges,
molecules, LJ table, exclusion table in the reference's add order, box,
cell manager), FP64 positions and the force parameters (switching,
electrostatics, cutoff).  Choices the workload descriptions leave open are
fixed here:

* water lattices use ceil(n^(1/3))^3 sites; when that exceeds the molecule
  count a seeded random subset of sites is filled (uniform vacancies);
* the box is [0, L)^3 (origin at L/2); O atoms sit at site centres jittered
  by U(-0.3, 0.3) per axis; each water gets a uniformly random orientation
  and orientations are redrawn until no H is within 1.5 A of an atom of
  another molecule (30 rounds at most);
* chains start on a seeded subset of a cubic lattice and grow with a fixed
  111 degree bend; each new atom takes, of 24 random dihedrals, the one
  farthest from the atoms already placed (chains grow in four interleaved
  groups per step so same-step atoms see each other); then, for up to 20
  rounds, each chain with an atom within 1.2 A of a non-neighbour has its
  tail from that atom rotated rigidly about the preceding bond to the best
  of its current and 64 random angles (bond lengths and bend angles stay
  exact);
* configs 2 and 3 switch with C2 from rc - 1.5 A;
* H atoms of water have epsilon 0, so Lorentz-Berthelot gives A = B = 0 for
  every pair involving H and LJ acts on O-O only.
"""
import math
from dataclasses import dataclass, field

import numpy as np

from .topology import (EXCLUDE_SCALED14, CubicCellManager,
                       PeriodicBoundaryConditions, Topology,
                       build_exclusion_table, lorentz_berthelot_table)

WATER_R_OH = 0.9572
WATER_ANGLE_HOH = 104.52
WATER_Q_O = -0.834
WATER_Q_H = 0.417
WATER_SIGMA_O = 3.1507
WATER_EPS_O = 0.1521
WATER_MASS_O = 15.9994
WATER_MASS_H = 1.008
WATER_DENSITY = 0.0334  # molecules per cubic Angstrom


@dataclass
class Workload:
    name: str
    topology: Topology
    positions: np.ndarray
    cutoff: float
    lj_switch: tuple            # (switchon, cutoff), C2
    elec: str                   # "coulomb" | "ewald" | "none"
    elec_switch: tuple          # (switchon, cutoff) C2, or ("C1", cutoff)
    alpha: float = 0.0
    meta: dict = field(default_factory=dict)

    @property
    def n_atoms(self):
        return self.topology.n_atoms

    def one_atom_pair(self):
        from . import force as F
        lj_sw = F.C2SwitchingFunction(*self.lj_switch)
        if self.elec_switch[0] == "C1":
            el_sw = F.C1SwitchingFunction(self.elec_switch[1])
        else:
            el_sw = F.C2SwitchingFunction(*self.elec_switch)
        if self.elec == "coulomb":
            el = F.CoulombForce()
        elif self.elec == "ewald":
            el = F.EwaldRealForce(self.alpha)
        else:
            el = None
        return F.OneAtomPairTwo(F.LennardJonesForce(), lj_sw, el, el_sw)

    def make_force(self, precision="fp64", device=0):
        from . import force as F
        return F.NonbondedCutoffSystemForce(self.cutoff, self.one_atom_pair(),
                                            precision=precision, device=device)


# ------------------------------------------------------------------ water
def _random_rotations(rng, n):
    """Uniform random rotation matrices from uniform unit quaternions."""
    u1, u2, u3 = rng.random(n), rng.random(n), rng.random(n)
    q0 = np.sqrt(1 - u1) * np.sin(2 * np.pi * u2)
    q1 = np.sqrt(1 - u1) * np.cos(2 * np.pi * u2)
    q2 = np.sqrt(u1) * np.sin(2 * np.pi * u3)
    q3 = np.sqrt(u1) * np.cos(2 * np.pi * u3)
    R = np.empty((n, 3, 3))
    R[:, 0, 0] = 1 - 2 * (q2 * q2 + q3 * q3)
    R[:, 0, 1] = 2 * (q1 * q2 - q0 * q3)
    R[:, 0, 2] = 2 * (q1 * q3 + q0 * q2)
    R[:, 1, 0] = 2 * (q1 * q2 + q0 * q3)
    R[:, 1, 1] = 1 - 2 * (q1 * q1 + q3 * q3)
    R[:, 1, 2] = 2 * (q2 * q3 - q0 * q1)
    R[:, 2, 0] = 2 * (q1 * q3 - q0 * q2)
    R[:, 2, 1] = 2 * (q2 * q3 + q0 * q1)
    R[:, 2, 2] = 1 - 2 * (q1 * q1 + q2 * q2)
    return R


def _water_hydrogens(O, R):
    half = math.radians(WATER_ANGLE_HOH) / 2
    local = np.array([[math.sin(half), 0.0, math.cos(half)],
                      [-math.sin(half), 0.0, math.cos(half)]]) * WATER_R_OH
    return O[:, None, :] + np.einsum("nij,hj->nhi", R, local)


def water_positions(n_mol, L, seed, jitter=0.3, min_h_dist=1.5, rounds=30):
    """(3 n_mol, 3) positions O, H, H per molecule in a box [0, L)^3."""
    from scipy.spatial import cKDTree
    rng = np.random.default_rng(seed)
    k = int(math.ceil(round(n_mol ** (1.0 / 3.0), 9)))
    while k ** 3 < n_mol:
        k += 1
    spacing = L / k
    idx = np.arange(k ** 3)
    if k ** 3 > n_mol:
        idx = np.sort(rng.choice(k ** 3, size=n_mol, replace=False))
    ix, iy, iz = idx // (k * k), (idx // k) % k, idx % k
    O = (np.stack([ix, iy, iz], 1) + 0.5) * spacing
    O = O + rng.uniform(-jitter, jitter, size=O.shape)
    R = _random_rotations(rng, n_mol)
    H = _water_hydrogens(O, R)
    mol = np.repeat(np.arange(n_mol), 3)
    for _ in range(rounds):
        pos = np.concatenate([O[:, None, :], H], axis=1).reshape(-1, 3)
        tree = cKDTree(np.mod(pos, L), boxsize=L)
        pairs = tree.query_pairs(min_h_dist, output_type="ndarray")
        if len(pairs):
            a, b = pairs[:, 0], pairs[:, 1]
            bad = (mol[a] != mol[b]) & ((a % 3 != 0) | (b % 3 != 0))
            pairs = pairs[bad]
        if not len(pairs):
            break
        redo = np.unique(mol[pairs.ravel()])
        H[redo] = _water_hydrogens(O[redo], _random_rotations(rng, len(redo)))
    pos = np.concatenate([O[:, None, :], H], axis=1).reshape(-1, 3)
    return np.ascontiguousarray(pos)


def water_topology(n_mol, L, cell_size):
    n = 3 * n_mol
    types = np.tile(np.array([0, 1, 1], dtype=np.int32), n_mol)
    q = np.tile(np.array([WATER_Q_O, WATER_Q_H, WATER_Q_H]), n_mol)
    mol = np.repeat(np.arange(n_mol, dtype=np.int32), 3)
    masses = np.tile(np.array([WATER_MASS_O, WATER_MASS_H, WATER_MASS_H]),
                     n_mol)
    base = 3 * np.arange(n_mol)
    bonds = np.concatenate([np.stack([base, base + 1], 1),
                            np.stack([base, base + 2], 1)])
    # bond order as a PSF lists them: per molecule O-H1, O-H2
    bonds = bonds.reshape(2, n_mol, 2).transpose(1, 0, 2).reshape(-1, 2)
    angles = np.stack([base + 1, base, base + 2], 1)
    excl = build_exclusion_table(n, EXCLUDE_SCALED14, bonds, angles, [])
    lj = lorentz_berthelot_table([WATER_SIGMA_O, 0.4], [WATER_EPS_O, 0.0])
    bc = PeriodicBoundaryConditions.cubic(L)
    topo = Topology(types, q, mol, lj, excl, bc, CubicCellManager(cell_size),
                    coulomb_scaling_factor=1.0, masses=masses)
    # TIP3 atom names and the bond / angle terms with their rest values, as
    # the SHAKE / RATTLE constraint list reads them
    topo.atom_names = np.tile(np.array(["OH2", "H1", "H2"]), n_mol)
    topo.bonds = bonds
    topo.bond_rest = np.full(len(bonds), WATER_R_OH)
    topo.angles = angles
    topo.angle_rest = np.full(n_mol, WATER_ANGLE_HOH * math.pi / 180.0)
    return topo


def water_workload(name, n_mol, L, seed, rc, rs, elec, alpha=0.0,
                   cell_size=None):
    cell_size = rc / 2 if cell_size is None else cell_size
    topo = water_topology(n_mol, L, cell_size)
    pos = water_positions(n_mol, L, seed)
    return Workload(name, topo, pos, rc, (rs, rc), elec, (rs, rc), alpha,
                    meta={"n_mol": n_mol, "box": L, "seed": seed})


def config0(seed=1234):
    """1,000 waters, 31.0 A box, LJ(O-O) + Coulomb, C2 8 -> 10."""
    return water_workload("water3k_coulomb", 1000, 31.0, seed, 10.0, 8.0,
                          "coulomb")


def config1(seed=1234):
    """32,000 waters (96,000 atoms), 98.6 A box, LJ + Ewald real space
    (alpha 0.32), C2 7.5 -> 9."""
    return water_workload("water96k_ewald", 32000, 98.6, seed, 9.0, 7.5,
                          "ewald", alpha=0.32)


def config3_sizes(seed=1234, n_replicas=512, lo=167, hi=2667):
    rng = np.random.default_rng(seed + 3)
    return rng.integers(lo, hi + 1, size=n_replicas)


def config3_replica(i, n_mol, seed=1234):
    """Replica i of the batched set: config-0 recipe at water density."""
    L = (n_mol / WATER_DENSITY) ** (1.0 / 3.0)
    return water_workload("replica%03d" % i, int(n_mol), L, seed + 1000 + i,
                          9.0, 7.5, "ewald", alpha=0.32)


# ------------------------------------------------------------------ chains
CHAIN_SIGMA = (3.0, 3.4, 3.7, 4.0)
CHAIN_EPS = (0.05, 0.1, 0.15, 0.2)


def chain_positions(n_chains, chain_len, L, seed, bond=1.53, bend=111.0,
                    candidates=24, batches=4):
    """Random-walk chains with fixed bond length and bend angle.  Each new
    atom takes, of `candidates` random dihedrals, the one farthest from every
    atom already placed (other chains, and its own chain two or more bonds
    back); chains grow one atom at a time in `batches` interleaved groups so
    every group sees the atoms the earlier groups just placed."""
    from scipy.spatial import cKDTree
    rng = np.random.default_rng(seed)
    k = int(math.ceil(n_chains ** (1.0 / 3.0)))
    while k ** 3 < n_chains:
        k += 1
    sites = np.sort(rng.choice(k ** 3, size=n_chains, replace=False))
    ix, iy, iz = sites // (k * k), (sites // k) % k, sites % k
    X = np.zeros((n_chains, chain_len, 3))
    X[:, 0] = (np.stack([ix, iy, iz], 1) + 0.5) * (L / k)
    d0 = rng.normal(size=(n_chains, 3))
    d0 /= np.linalg.norm(d0, axis=1, keepdims=True)
    X[:, 1] = X[:, 0] + bond * d0
    theta = math.radians(180.0 - bend)  # turn between consecutive bonds
    ref0 = rng.normal(size=(n_chains, 3))
    order = rng.permutation(n_chains)
    groups = [np.sort(order[g::batches]) for g in range(batches)]
    for j in range(2, chain_len):
        done = np.zeros(n_chains, dtype=bool)
        for idx in groups:
            # atoms placed so far: steps < j of every chain, step j of the
            # chains that already grew in this step
            pts = [X[:, :j].reshape(-1, 3), X[done, j]]
            own = [np.repeat(np.arange(n_chains), j), np.nonzero(done)[0]]
            dep = [np.tile(np.arange(j), n_chains),
                   np.full(int(done.sum()), j)]
            pts, own, dep = (np.concatenate(pts), np.concatenate(own),
                             np.concatenate(dep))
            tree = cKDTree(np.mod(pts, L), boxsize=L)
            b1 = X[idx, j - 1] - X[idx, j - 2]
            b1 /= np.linalg.norm(b1, axis=1, keepdims=True)
            b0 = X[idx, j - 2] - X[idx, j - 3] if j >= 3 else ref0[idx]
            n1 = np.cross(b0, b1)
            n1 /= np.linalg.norm(n1, axis=1, keepdims=True)
            m1 = np.cross(n1, b1)
            phi = rng.uniform(0, 2 * np.pi, size=(len(idx), candidates))
            d = (math.cos(theta) * b1[:, None, :]
                 + math.sin(theta) * (np.cos(phi)[..., None] * m1[:, None, :]
                                      + np.sin(phi)[..., None] * n1[:, None, :]))
            cand = X[idx, j - 1][:, None, :] + bond * d
            dist, nb = tree.query(np.mod(cand.reshape(-1, 3), L), k=4)
            dist = dist.reshape(len(idx), candidates, 4)
            nb = nb.reshape(len(idx), candidates, 4)
            bonded = (own[nb] == idx[:, None, None]) & (dep[nb] >= j - 2)
            score = np.where(bonded, np.inf, dist).min(axis=2)
            best = np.argmax(score, axis=1)
            X[idx, j] = cand[np.arange(len(idx)), best]
            done[idx] = True
    _repair_chains(X, L, bond, theta, rng)
    return np.ascontiguousarray(np.mod(X.reshape(-1, 3), L))


def _chain_clashes(X, L, min_dist):
    """Atoms (flat index) closer than min_dist to an atom that is not its
    1-2 or 1-3 neighbour in the same chain."""
    from scipy.spatial import cKDTree
    n_chains, chain_len = X.shape[:2]
    P = np.mod(X.reshape(-1, 3), L)
    pairs = cKDTree(P, boxsize=L).query_pairs(min_dist, output_type="ndarray")
    if not len(pairs):
        return pairs
    a, b = pairs[:, 0], pairs[:, 1]
    near = (a // chain_len == b // chain_len) & (np.abs(a - b) <= 2)
    return pairs[~near]


def _repair_chains(X, L, bond, theta, rng, min_dist=1.2, candidates=64,
                   rounds=20, batches=8):
    """Pivot moves: for each chain that still has an atom within min_dist of
    a non-neighbour, the tail from the earliest such atom is rotated rigidly
    about the preceding bond (bond lengths and bend angles are unchanged);
    of the current angle and `candidates` random ones, the one keeping the
    tail farthest from every other atom is taken.  Chains are moved in
    `batches` groups per round, each scored against all current positions."""
    from scipy.spatial import cKDTree
    n_chains, chain_len = X.shape[:2]
    for _ in range(rounds):
        pairs = _chain_clashes(X, L, min_dist)
        if not len(pairs):
            return
        pos = pairs % chain_len
        first = np.full(n_chains, chain_len)
        for k in (0, 1):
            np.minimum.at(first, pairs[:, k] // chain_len,
                          np.maximum(pos[:, k], 2))
        fix = rng.permutation(np.nonzero(first < chain_len)[0])
        for group in np.array_split(fix, batches):
            tree = cKDTree(np.mod(X.reshape(-1, 3), L), boxsize=L)
            for c in group:
                p = first[c]
                axis = X[c, p - 1] - X[c, p - 2]
                axis /= np.linalg.norm(axis)
                tail = X[c, p:] - X[c, p - 1]
                psi = np.concatenate([[0.0], rng.uniform(0, 2 * np.pi,
                                                         candidates)])
                cs, sn = np.cos(psi)[:, None, None], np.sin(psi)[:, None, None]
                kx = np.cross(axis[None, :], tail)[None]
                kd = (tail @ axis)[None, :, None] * axis[None, None, :]
                rot = tail[None] * cs + kx * sn + kd * (1 - cs)  # Rodrigues
                cand = X[c, p - 1][None, None, :] + rot
                dist, nb = tree.query(np.mod(cand.reshape(-1, 3), L), k=8)
                nb = np.minimum(nb, n_chains * chain_len - 1)
                tail_pos = np.tile(np.arange(p, chain_len),
                                   len(psi))[:, None]
                own = nb // chain_len == c
                skip = own & ((nb % chain_len >= p)
                              | (nb % chain_len >= tail_pos - 2))
                dist = np.where(skip, np.inf, dist).min(axis=1)
                score = dist.reshape(len(psi), -1).min(axis=1)
                X[c, p:] = cand[int(np.argmax(score))]


def chain_topology(n_chains, chain_len, L, seed, cell_size, sigma=CHAIN_SIGMA,
                   eps=CHAIN_EPS, charges=None):
    rng = np.random.default_rng(seed + 7)
    n = n_chains * chain_len
    types = rng.integers(0, len(sigma), size=n).astype(np.int32)
    if charges is None:
        q = rng.normal(0.0, 0.3, size=(n_chains, chain_len))
        q -= q.mean(axis=1, keepdims=True)
        charges = q.reshape(-1)
    mol = np.repeat(np.arange(n_chains, dtype=np.int32), chain_len)
    start = chain_len * np.arange(n_chains)[:, None, None]
    i = np.arange(chain_len)
    bonds = (start + np.stack([i[:-1], i[:-1] + 1], 1)[None]).reshape(-1, 2)
    angles = (start + np.stack([i[:-2], i[:-2] + 1, i[:-2] + 2], 1)[None]
              ).reshape(-1, 3)
    dih = (start + np.stack([i[:-3], i[:-3] + 1, i[:-3] + 2, i[:-3] + 3], 1)[None]
           ).reshape(-1, 4)
    excl = build_exclusion_table(n, EXCLUDE_SCALED14, bonds, angles, dih)
    lj = lorentz_berthelot_table(sigma, eps, lj_scaling_factor=0.5)
    bc = PeriodicBoundaryConditions.cubic(L)
    return Topology(types, charges, mol, lj, excl, bc,
                    CubicCellManager(cell_size), coulomb_scaling_factor=0.5,
                    masses=np.full(n, 14.027))


def chain_workload(name, n_chains, chain_len, L, seed, rc=10.0, rs=8.5,
                   alpha=0.34, cell_size=None, sigma=CHAIN_SIGMA,
                   eps=CHAIN_EPS):
    cell_size = rc / 2 if cell_size is None else cell_size
    topo = chain_topology(n_chains, chain_len, L, seed, cell_size, sigma, eps)
    pos = chain_positions(n_chains, chain_len, L, seed)
    return Workload(name, topo, pos, rc, (rs, rc), "ewald", (rs, rc), alpha,
                    meta={"n_chains": n_chains, "chain_len": chain_len,
                          "box": L, "seed": seed})


def config2(seed=1234):
    """12,500 chains of 20 atoms (250,000 atoms), 138 A box, LB mixing of
    4 types, neutral chains, 1-2/1-3 excluded, 1-4 scaled by 0.5, Ewald real
    space alpha 0.34, C2 8.5 -> 10."""
    return chain_workload("chains250k_ewald", 12500, 20, 138.0, seed)


def small_chains(seed=1234, n_chains=150, chain_len=20, L=None):
    """Config-2 recipe at a size the O(N^2) checks finish quickly."""
    if L is None:
        L = 138.0 * (n_chains / 12500.0) ** (1.0 / 3.0)
    return chain_workload("chains_small", n_chains, chain_len, L, seed)


def small_water(seed=1234, n_mol=300, elec="ewald", rc=9.0, rs=7.5,
                alpha=0.32):
    L = (n_mol / WATER_DENSITY) ** (1.0 / 3.0)
    return water_workload("water_small", n_mol, L, seed, rc, rs, elec, alpha)


def with_algorithm(w, cutoff, cell_size, name=None):
    """The workload with another NonbondedCutoff -cutoff and cell size.
    With the cutoff below the switching cutoff the pair set is the one the
    reference's cell grid and delta list reach
    (topology/CellListEnumeratorPeriodicBoundaries.h:46-83,114-134)."""
    w.cutoff = float(cutoff)
    w.topology.cell_manager = CubicCellManager(cell_size)
    w.topology.touch()
    if name:
        w.name = name
    return w


def many_type_chains(seed=1234, n_chains=60, n_types=12, chain_len=20):
    """small_chains with n_types LJ types (sigma 3.0 .. 4.0 A, epsilon
    0.05 .. 0.2 kcal/mol, Lorentz-Berthelot): more than the tile path
    holds, so the force takes the general cell-list path."""
    L = 138.0 * (n_chains / 12500.0) ** (1.0 / 3.0)
    sigma = tuple(np.linspace(3.0, 4.0, n_types))
    eps = tuple(np.linspace(0.05, 0.2, n_types))
    return chain_workload("chains_%dtypes" % n_types, n_chains, chain_len, L,
                          seed, sigma=sigma, eps=eps)


CONFIGS = {0: config0, 1: config1, 2: config2}
