"""Workloads built from the golden fixtures of the reference's own tests,
plus the harmonic bond/angle forces needed to separate the nonbonded part of
the reference's total forces (force/bonded/BondSystemForce.h,
force/bonded/AngleSystemForce.h: E = k (r - r0)^2, E = k (theta - theta0)^2)."""
import os

import numpy as np

from protomoladdon_amd.systems import Workload
from protomoladdon_amd.topology import (EXCLUDE_ONE4, EXCLUDE_SCALED14,
                                        CubicCellManager,
                                        PeriodicBoundaryConditions, Topology,
                                        build_exclusion_table, charmm_table)

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def water423():
    d = load("water423.npz")
    n = len(d["types"])
    lj = d["lj_eps_rmin"]
    table = charmm_table(lj[:, 1], lj[:, 0], lj[:, 3], lj[:, 2])
    excl = build_exclusion_table(n, EXCLUDE_SCALED14, d["bonds"], d["angles"])
    mol = np.zeros(n, dtype=np.int32)
    # molecules = connected components of the bond graph
    parent = list(range(n))

    def find(a):
        while parent[a] != a:
            parent[a] = parent[parent[a]]
            a = parent[a]
        return a
    for a, b in d["bonds"]:
        parent[find(a)] = find(b)
    roots = {}
    for i in range(n):
        mol[i] = roots.setdefault(find(i), len(roots))
    L = d["box"]
    bc = PeriodicBoundaryConditions((L[0], 0, 0), (0, L[1], 0), (0, 0, L[2]),
                                    d["origin"])
    topo = Topology(d["types"], d["charges"], mol, table, excl, bc,
                    CubicCellManager(float(d["cell_size"])), 1.0,
                    masses=d["masses"])
    sw = tuple(float(v) for v in d["lj_switch"])
    w = Workload("water423", topo, np.array(d["positions"]),
                 float(d["algorithm_cutoff"]), sw, "coulomb",
                 ("C1", float(d["coulomb_c1_cutoff"])))
    return w, d


def pentane320():
    d = load("pentane320.npz")
    n = len(d["types"])
    lj = d["lj_eps_rmin"]
    table = charmm_table(lj[:, 1], lj[:, 0], lj[:, 3], lj[:, 2])
    excl = build_exclusion_table(n, EXCLUDE_ONE4, d["bonds"], d["angles"],
                                 d["dihedrals"])
    mol = np.repeat(np.arange(n // 5, dtype=np.int32), 5)
    L = d["box"]
    bc = PeriodicBoundaryConditions((L[0], 0, 0), (0, L[1], 0), (0, 0, L[2]),
                                    d["origin"])
    topo = Topology(d["types"], d["charges"], mol, table, excl, bc,
                    CubicCellManager(float(d["cell_size"])), 1.0)
    sw = tuple(float(v) for v in d["lj_switch"])
    w = Workload("pentane320", topo, np.array(d["positions"]),
                 float(d["algorithm_cutoff"]), sw, "none", sw)
    return w, d


def min_image(d, L):
    return d - L * np.rint(d / L)


def harmonic_bonded_forces(pos, bonds, angles, kb, r0, ka, theta0_deg, L):
    f = np.zeros_like(pos)
    e_b = e_a = 0.0
    for a, b in bonds:
        d = min_image(pos[b] - pos[a], L)
        r = np.linalg.norm(d)
        e_b += kb * (r - r0) ** 2
        g = -2.0 * kb * (r - r0) * d / r  # force on b
        f[b] += g
        f[a] -= g
    th0 = np.radians(theta0_deg)
    for a, b, c in angles:
        u = min_image(pos[a] - pos[b], L)
        v = min_image(pos[c] - pos[b], L)
        ru, rv = np.linalg.norm(u), np.linalg.norm(v)
        cs = np.dot(u, v) / (ru * rv)
        th = np.arccos(np.clip(cs, -1.0, 1.0))
        e_a += ka * (th - th0) ** 2
        dE = 2.0 * ka * (th - th0)
        sn = np.sin(th)
        # d theta / d u = -(v/(ru rv) - cs u/ru^2) / sin
        dtu = -(v / (ru * rv) - cs * u / ru ** 2) / sn
        dtv = -(u / (ru * rv) - cs * v / rv ** 2) / sn
        f[a] -= dE * dtu
        f[c] -= dE * dtv
        f[b] += dE * (dtu + dtv)
    return f, e_b, e_a


DATA = os.path.join(GOLDEN, "data")


def comparator(expected_path, new_path, epsilon=1e-5, scalar=1.0):
    """The reference's test/comparator.py rule: the same number of fields
    per line; fields that parse as integers are skipped; other numbers
    must agree within an absolute epsilon (test/test.py:14 default 1e-5).
    Returns the number of differing lines or fields."""
    diffs = 0
    with open(expected_path) as fe, open(new_path) as fn:
        for le in fe:
            a, b = le.split(), fn.readline().split()
            if len(a) != len(b):
                diffs += 1
                continue
            for x, y in zip(a, b):
                try:
                    int(x)
                    int(y)
                    continue
                except ValueError:
                    pass
                try:
                    if abs(float(x) * scalar - float(y)) > epsilon:
                        diffs += 1
                except ValueError:
                    pass
    return diffs
