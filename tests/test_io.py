"""Input and output formats (protomoladdon_amd/io.py) against the data and
expected outputs of the reference's regression case
test/tests/water_CHARMM_PERIODIC_L.conf, committed under tests/golden/data/
(water_423.psf / .pos.pdb / .vel.pdb, par_all27_prot_lipid.inp and
expected/water_CHARMM_PERIODIC_L.{pos,vel,energy}).

* the readers give the topology of the committed numeric fixture;
* writing the expected XYZ and energy rows back reproduces the files byte
  for byte (the writers' layout is the reference's);
* the whole run (readers, oracle leapfrog, writers) passes the reference's
  comparator rule at its default epsilon and agrees within 1e-10."""
import math
import os

import numpy as np

from oracle import oracle as O
from protomoladdon_amd import io
from protomoladdon_amd.integrator import kinetic_energy
from tests import fixtures_lib as FX

D = FX.DATA
CONF = dict(box=(28.0, 28.0, 28.0), cell_size=6.5, timestep=0.5, n_steps=200,
            lj_switch=(0.1, 6.5), coulomb_c1=6.5)


def load_system():
    psf = io.read_psf(os.path.join(D, "water_423.psf"))
    par = io.read_par(os.path.join(D, "par_all27_prot_lipid.inp"))
    pos, _ = io.read_pdb(os.path.join(D, "water_423.pos.pdb"))
    vel, _ = io.read_pdb(os.path.join(D, "water_423.vel.pdb"))
    topo, bonded = io.build_charmm_system(psf, par, CONF["box"],
                                          cell_size=CONF["cell_size"])
    from protomoladdon_amd.systems import Workload
    w = Workload("water423", topo, pos, 6.5, CONF["lj_switch"], "coulomb",
                 ("C1", CONF["coulomb_c1"]))
    return w, bonded, vel / io.PDB_VELOCITY_SCALING


def test_readers_match_numeric_fixture():
    w, bonded, v0 = load_system()
    wf, d = FX.water423()
    t, tf = w.topology, wf.topology
    np.testing.assert_array_equal(w.positions, wf.positions)
    np.testing.assert_array_equal(t.types, tf.types)
    np.testing.assert_array_equal(t.scaled_charges, tf.scaled_charges)
    np.testing.assert_array_equal(t.molecules, tf.molecules)
    np.testing.assert_array_equal(t.masses, tf.masses)
    np.testing.assert_array_equal(t.lennard_jones_parameters.table,
                                  tf.lennard_jones_parameters.table)
    np.testing.assert_array_equal(t.exclusions.table(), tf.exclusions.table())
    np.testing.assert_array_equal(v0, d["velocities_pdb"] / io.PDB_VELOCITY_SCALING)
    assert set(bonded["bond_k"]) == {450.0} and set(bonded["bond_r0"]) == {0.9572}
    assert set(bonded["angle_k"]) == {55.0} and not bonded["ub_k"].any()
    assert np.all(bonded["angle_theta0"] == 104.52 * math.pi / 180.0)
    assert t.degrees_of_freedom == 3 * 423 - 3
    assert t.type_names == ["OT", "HT"]


def test_par_reader_sections():
    par = io.read_par(os.path.join(D, "par_all27_prot_lipid.inp"))
    # rows of par_all27_prot_lipid.inp: a bond, an angle with Urey-Bradley
    # terms, one without, and a NONBONDED row with 1-4 values
    assert par.bond("CT1", "C") == par.bond("C", "CT1")
    k, t0, kub, s0 = par.angle("HA", "CT1", "CT1")
    assert kub > 0 and s0 > 0
    assert par.angle("HT", "OT", "HT") == (55.0, 104.52, 0.0, 0.0)
    assert par.nonbonded["HT"] == (-0.046, 0.2245, -0.046, 0.2245)


def test_xyz_writer_reproduces_reference_files(tmp_path):
    for name in ("water_CHARMM_PERIODIC_L.pos", "water_CHARMM_PERIODIC_L.vel"):
        names, c = io.read_xyz(os.path.join(D, name))
        out = tmp_path / name
        io.write_xyz(str(out), names, c)
        assert out.read_text() == open(os.path.join(D, name)).read()


def test_energies_writer_reproduces_reference_rows(tmp_path):
    rows = [[float(v) for v in l.split()] for l in
            open(os.path.join(D, "water_CHARMM_PERIODIC_L.energy"))]
    out = str(tmp_path / "e.energy")
    with io.EnergiesWriter(out) as ew:
        for r in rows:
            ew.write(*r)
    assert open(out).read() == open(os.path.join(D, "water_CHARMM_PERIODIC_L.energy")).read()
    header = open(out + ".header").read().split()
    assert header[:3] == ["Time(fs)", "E_potential", "E_kinetic"]


def write_outputs(tmp, names, x, v, rows):
    io.write_xyz(os.path.join(tmp, "run.pos"), names, x)
    io.write_xyz(os.path.join(tmp, "run.vel"), names, v)
    with io.EnergiesWriter(os.path.join(tmp, "run.energy")) as ew:
        for r in rows:
            ew.write(*r)


def energy_row(time_fs, masses, v, nonbonded2, bonded2, dof, volume):
    ke = kinetic_energy(masses, v)
    pe = bonded2[0] + bonded2[1] + nonbonded2[0] + nonbonded2[1]
    return (time_fs, pe, ke, pe + ke, io.temperature(ke, dof), bonded2[0],
            bonded2[1], 0.0, 0.0, nonbonded2[0], nonbonded2[1], 0.0, volume)


def check_run_files(tmp):
    for ext in ("pos", "vel", "energy"):
        exp = os.path.join(D, "water_CHARMM_PERIODIC_L." + ext)
        assert FX.comparator(exp, os.path.join(tmp, "run." + ext)) == 0, ext


def test_run_through_readers_and_writers_passes_reference_comparator(tmp_path):
    w, bonded, v0 = load_system()
    t = w.topology
    s = O.System.from_workload(w)
    f0, e0, b0 = O.lf_forces(s, w.positions, t.masses, bonded)
    vol = float(np.prod(CONF["box"]))
    rows = [energy_row(0.0, t.masses, v0, e0, b0, t.degrees_of_freedom, vol)]
    st = O.MDState(w.positions, v0, f0, energies=e0, bonded=b0)
    O.leapfrog_run(s, st, CONF["n_steps"], t.masses, np.zeros((0, 2)),
                   np.zeros(0), CONF["timestep"], bonded=bonded)
    rows.append(energy_row(CONF["n_steps"] * CONF["timestep"], t.masses, st.v,
                           st.energies, st.bonded, t.degrees_of_freedom, vol))
    names = [t.type_names[k] for k in t.types]
    write_outputs(str(tmp_path), names, st.x, st.v, rows)
    check_run_files(str(tmp_path))
    # and far inside the comparator's epsilon
    _, xe = io.read_xyz(os.path.join(D, "water_CHARMM_PERIODIC_L.pos"))
    assert np.abs(st.x - xe).max() <= 1e-10
