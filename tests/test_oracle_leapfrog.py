"""Pins the oracle's LeapfrogIntegrator / SHAKE / RATTLE restatement and the
host-side constraint list on the CPU.

The reference holds no fixture for this integrator with the cutoff force
alone (its SHAKE regression, protomol-test ConfigTests/data/ShakeTest.conf,
needs the bonded force field and a vacuum full-electrostatics force), so the
checks are what the mathematics fixes:
* the constraint list of a water is the law-of-cosines H-H distance and the
  two O-H bonds, in the reference's order; H-X bonds are added only with
  ``all``;
* free flight: with no force, x_n = x_0 + n h v_0 and v is unchanged, which
  fixes the drift scale h = timestep / TIMEFACTOR;
* SHAKE leaves v = (x_new - x_old) / h (its velocity update divides the
  position correction by the same step), conserves momentum and meets the
  constraint to epsilon;
* RATTLE changes velocities only along the bond, conserves momentum and
  leaves the bond velocity orthogonal to the bond;
* kicks: one step from rest of two interacting atoms gives
  v = (F(x_0) + F(x_1)) h / (2 m), the leapfrog composition;
* the sweep count is max_iter + 1 when epsilon cannot be met;
* total energy of rigid water with the switched cutoff force is conserved
  over 100 fs.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from protomoladdon_amd import systems
from protomoladdon_amd.integrator import (TIMEFACTOR,
                                          build_rattle_shake_constraints,
                                          kinetic_energy, topology_constraints)
from protomoladdon_amd.topology import (CubicCellManager,
                                        PeriodicBoundaryConditions, Topology,
                                        build_exclusion_table,
                                        lorentz_berthelot_table)

H = 1.0 / TIMEFACTOR  # drift scale of a 1 fs step
KB = 0.001987191      # base/PMConstants.cpp (BOLTZMANN, kcal/mol/K)


def maxwell(masses, seed, temperature=300.0):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((len(masses), 3)) * np.sqrt(
        KB * temperature / np.asarray(masses))[:, None]


def toy_system(pos, masses, L=60.0, bonds=(), angles=()):
    """Atoms of type 0 (LJ) in a large box; bonded pairs are excluded."""
    n = len(pos)
    types = np.zeros(n, np.int32)
    mol = np.zeros(n, np.int32)
    excl = build_exclusion_table(n, "scaled1-4", np.asarray(bonds, int).reshape(-1, 2),
                                 np.asarray(angles, int).reshape(-1, 3), [])
    lj = lorentz_berthelot_table([3.0], [0.2])
    topo = Topology(types, np.zeros(n), mol, lj, excl,
                    PeriodicBoundaryConditions.cubic(L), CubicCellManager(5.0),
                    masses=masses)
    w = systems.Workload("toy", topo, np.asarray(pos, float), 10.0, (8.0, 10.0),
                         "none", (8.0, 10.0))
    return w, O.System.from_workload(w)


def test_water_constraint_list():
    t = systems.small_water(n_mol=5).topology
    pairs, rest = topology_constraints(t)
    hh = 2 * 0.9572 * math.sin(math.radians(104.52) / 2)
    for m in range(5):
        o = 3 * m
        assert pairs[3 * m].tolist() == [o + 1, o + 2]
        assert pairs[3 * m + 1].tolist() == [o + 1, o]
        assert pairs[3 * m + 2].tolist() == [o + 2, o]
        assert abs(rest[3 * m] - hh) < 1e-12
        assert rest[3 * m + 1] == rest[3 * m + 2] == 0.9572


def test_hx_bond_constraints():
    # a water (atoms 0-2), a methyl-like C with three H (3-6) and a C-C bond
    names = ["OH2", "H1", "H2", "CT", "HA", "HB", "HC", "CX"]
    bonds = [(0, 1), (0, 2), (3, 4), (3, 5), (3, 6), (3, 7)]
    rest = [0.9572, 0.9572, 1.09, 1.10, 1.11, 1.53]
    angles = [(1, 0, 2), (4, 3, 5)]
    arest = [math.radians(104.52), math.radians(109.5)]
    p, r = build_rattle_shake_constraints(names, bonds, rest, angles, arest)
    assert p.tolist() == [[1, 2], [1, 0], [2, 0], [3, 4], [3, 5], [3, 6]]
    assert r[3:].tolist() == [1.09, 1.10, 1.11]
    p, r = build_rattle_shake_constraints(names, bonds, rest, angles, arest,
                                          all_bonds=False)
    assert p.tolist() == [[1, 2], [1, 0], [2, 0]]


def test_free_flight():
    pos = np.array([[5.0, 5.0, 5.0], [35.0, 35.0, 35.0]])
    m = np.array([12.0, 16.0])
    w, s = toy_system(pos, m)
    v0 = np.array([[0.1, -0.2, 0.3], [-0.05, 0.02, 0.07]])
    st = O.MDState(pos, v0, np.zeros((2, 3)))
    O.leapfrog_run(s, st, 25, m, np.zeros((0, 2)), np.zeros(0), 1.0)
    np.testing.assert_allclose(st.x, pos + 25 * H * v0, rtol=1e-13, atol=1e-12)
    np.testing.assert_array_equal(st.v, v0)
    assert not st.f.any()


def test_extra_force_of_the_group_is_kicked():
    """A further force of the force group (the hook the k-space terms use):
    a constant force F gives the constant-acceleration path x_n = x_0 +
    n h v_0 + n^2 h^2 F / (2 m), v_n = v_0 + n h F / m, and its energy joins
    the electrostatic slot of every evaluation."""
    pos = np.array([[5.0, 5.0, 5.0], [35.0, 35.0, 35.0]])
    m = np.array([12.0, 16.0])
    _, s = toy_system(pos, m)
    v0 = np.array([[0.1, -0.2, 0.3], [-0.05, 0.02, 0.07]])
    F = np.array([[0.3, -0.1, 0.2], [-0.4, 0.25, 0.05]])
    calls = []

    def extra(x):
        calls.append(x.copy())
        return F, 1.25
    f0, e0, _ = O.lf_forces(s, pos, m, extra=extra)
    np.testing.assert_array_equal(f0, F)
    assert e0[1] == 1.25
    st = O.MDState(pos, v0, f0, energies=e0)
    n = 12
    O.leapfrog_run(s, st, n, m, np.zeros((0, 2)), np.zeros(0), 1.0,
                   extra=extra)
    assert len(calls) == 1 + n
    np.testing.assert_allclose(st.v, v0 + n * H * F / m[:, None], rtol=1e-13)
    np.testing.assert_allclose(
        st.x, pos + n * H * v0 + (n * H) ** 2 * F / (2 * m[:, None]),
        rtol=1e-13)
    np.testing.assert_array_equal(calls[-1], st.x)
    assert st.energies[1] == 1.25


def test_shake_velocity_and_momentum():
    d0 = 1.2
    pos = np.array([[10.0, 10.0, 10.0], [10.0 + d0, 10.0, 10.0]])
    m = np.array([15.9994, 1.008])
    w, s = toy_system(pos, m, bonds=[(0, 1)])
    v0 = np.array([[0.01, 0.3, -0.1], [0.5, -2.0, 0.7]])
    st = O.MDState(pos, v0, np.zeros((2, 3)))
    p0 = (m[:, None] * v0).sum(0)
    iters = O.leapfrog_run(s, st, 1, m, [[0, 1]], [d0], 1.0, shake=(1e-12, 50))
    assert len(iters) == 1 and iters[0] >= 1
    np.testing.assert_allclose(st.v * H, st.x - pos, rtol=1e-10, atol=1e-15)
    d = np.linalg.norm(st.x[0] - st.x[1])
    assert abs(d * d - d0 * d0) / (2 * d0 * d0) <= 1e-12
    np.testing.assert_allclose((m[:, None] * st.v).sum(0), p0, atol=1e-12)


def test_rattle_projection():
    d0 = 1.0
    pos = np.array([[10.0, 10.0, 10.0], [10.6, 10.8, 10.0]])
    m = np.array([12.0, 1.008])
    w, s = toy_system(pos, m, bonds=[(0, 1)])
    v0 = np.array([[0.2, -0.1, 0.05], [-0.3, 0.9, 0.4]])
    st = O.MDState(pos, v0, np.zeros((2, 3)))
    # one step: drift moves the atoms, then RATTLE at the new positions
    O.leapfrog_run(s, st, 1, m, [[0, 1]], [d0], 1.0, rattle=(1e-14, 50))
    rab = st.x[0] - st.x[1]
    vab = st.v[0] - st.v[1]
    assert abs(rab @ vab) <= 1e-13
    dv = st.v[0] - v0[0]
    assert np.linalg.norm(np.cross(dv, rab)) <= 1e-12 * np.linalg.norm(dv) * np.linalg.norm(rab) + 1e-15
    np.testing.assert_allclose((m[:, None] * st.v).sum(0),
                               (m[:, None] * v0).sum(0), atol=1e-13)


def test_kick_composition():
    pos = np.array([[10.0, 10.0, 10.0], [13.4, 10.0, 10.0]])
    m = np.array([12.0, 14.0])
    w, s = toy_system(pos, m)
    f0 = O.evaluate(s).forces
    st = O.MDState(pos, np.zeros((2, 3)), f0)
    O.leapfrog_run(s, st, 1, m, np.zeros((0, 2)), np.zeros(0), 1.0)
    x1 = pos + H * (f0 * H * 0.5 / m[:, None])
    np.testing.assert_allclose(st.x, x1, rtol=1e-14)
    f1 = O.evaluate(O.System.from_workload(w, positions=x1)).forces
    np.testing.assert_allclose(st.v, (f0 + f1) * H / (2 * m[:, None]), rtol=1e-12)
    np.testing.assert_allclose(st.f, f1, rtol=1e-14)


def test_unconverged_sweep_count():
    w = systems.small_water(n_mol=300, elec="coulomb")
    t = w.topology
    pairs, rest = topology_constraints(t)
    s = O.System.from_workload(w)
    st = O.MDState(w.positions, maxwell(t.masses, 3), O.evaluate(s).forces)
    it = O.leapfrog_run(s, st, 3, t.masses, pairs, rest, 1.0,
                        shake=(1e-30, 2), rattle=(1e-30, 4))
    # calls: SHAKE, (RATTLE, SHAKE) x 2, RATTLE
    assert it.tolist() == [3, 5, 3, 5, 3, 5]


@pytest.mark.parametrize("elec", ["coulomb", "ewald"])
def test_energy_conservation(elec):
    w = systems.small_water(n_mol=300, elec=elec)
    t = w.topology
    pairs, rest = topology_constraints(t)
    s = O.System.from_workload(w)
    r0 = O.evaluate(s)
    st = O.MDState(w.positions, maxwell(t.masses, 11), r0.forces,
                   energies=r0.energies)
    O.leapfrog_run(s, st, 10, t.masses, pairs, rest, 1.0,
                   shake=(1e-5, 30), rattle=(1e-5, 30))
    energies = []
    for _ in range(10):
        it = O.leapfrog_run(s, st, 10, t.masses, pairs, rest, 1.0,
                            shake=(1e-5, 30), rattle=(1e-5, 30))
        assert it.max() <= 31
        energies.append(kinetic_energy(t.masses, st.v) + st.energies.sum())
        d = st.x[pairs[:, 0]] - st.x[pairs[:, 1]]
        err = np.mean(np.abs((d * d).sum(1) - rest ** 2) / (2 * rest ** 2))
        assert err <= 1e-5
    kin = kinetic_energy(t.masses, st.v)
    spread = (max(energies) - min(energies)) / kin
    print(elec, "E", energies[0], energies[-1], "spread/KE", spread)
    assert spread < 5e-3


PDB_VEL_SCALING = 20.45482706  # base/PMConstants.cpp:68 (PDBVELSCALINGFACTOR)


def water423_md():
    """The reference's regression test/tests/water_CHARMM_PERIODIC_L.conf:
    141 flexible TIP3 waters, Leapfrog 0.5 fs x 200, forces Bond, Angle,
    Coulomb (C1, 6.5) and LennardJones (C2 0.1 -> 6.5), periodic 28 A box."""
    from tests import fixtures_lib as FX
    w, d = FX.water423()
    kb, r0 = d["bond_k_r0"]
    ka, t0 = d["angle_k_theta0"]
    bonded = dict(bonds=d["bonds"], bond_k=np.full(len(d["bonds"]), kb),
                  bond_r0=np.full(len(d["bonds"]), r0), angles=d["angles"],
                  angle_k=np.full(len(d["angles"]), ka),
                  angle_theta0=np.full(len(d["angles"]), t0 * math.pi / 180.0))
    v0 = d["velocities_pdb"] / PDB_VEL_SCALING
    return w, d, bonded, v0


def check_water423_outputs(x, v, energies, bonded_e, masses, d, tol):
    """Final positions / velocities (expected/*.pos, *.vel) and the 100 fs
    energy row: potential, kinetic, total, bond, angle, LJ, Coulomb."""
    np.testing.assert_allclose(x, d["final_positions"], rtol=0, atol=tol)
    np.testing.assert_allclose(v, d["final_velocities"], rtol=0, atol=tol)
    row = d["energy_row100"]
    ke = kinetic_energy(masses, v)
    pe = bonded_e[0] + bonded_e[1] + energies[0] + energies[1]
    # columns (output/OutputEnergies.cpp): time, potential, kinetic, total,
    # temperature, bond, angle, dihedral, improper, LJ, Coulomb, ...
    got = [pe, ke, ke + pe, bonded_e[0], bonded_e[1], energies[0], energies[1]]
    want = [row[1], row[2], row[3], row[5], row[6], row[9], row[10]]
    print("100 fs energies", got, "expected", want)
    # the expected file prints 8 significant digits
    np.testing.assert_allclose(got, want, rtol=2e-7, atol=2e-4)


def test_water423_200_steps_match_reference_outputs():
    w, d, bonded, v0 = water423_md()
    s = O.System.from_workload(w)
    f0, e0, b0 = O.lf_forces(s, w.positions, w.topology.masses, bonded)
    # the step-0 row of the expected energies
    row0 = d["energy_row0"]
    np.testing.assert_allclose([b0[0], b0[1], e0[0], e0[1]],
                               [row0[5], row0[6], row0[9], row0[10]],
                               rtol=2e-7, atol=2e-4)
    st = O.MDState(w.positions, v0, f0, energies=e0, bonded=b0)
    O.leapfrog_run(s, st, int(d["n_steps"]), w.topology.masses,
                   np.zeros((0, 2)), np.zeros(0), float(d["timestep"]),
                   bonded=bonded)
    err_x = np.abs(st.x - d["final_positions"]).max()
    err_v = np.abs(st.v - d["final_velocities"]).max()
    print("max |x - expected| %.3g, max |v - expected| %.3g" % (err_x, err_v))
    check_water423_outputs(st.x, st.v, st.energies, st.bonded,
                           w.topology.masses, d, tol=1e-10)
