"""ctypes binding of the FP64 CPU oracle (pmnb_oracle.c) and the file
interface of the reference-build driver (ref_driver.cpp, built into _ref/).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product never imports it.
"""
import ctypes
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_DRIVER = os.path.join(HERE, "_ref", "ref_driver")

ELEC = {"none": 0, "coulomb": 1, "ewald": 2}


class OrcSystem(ctypes.Structure):
    _fields_ = [
        ("n_atoms", ctypes.c_int32), ("n_types", ctypes.c_int32),
        ("n_excl", ctypes.c_int32), ("n_mol", ctypes.c_int32),
        ("pos", ctypes.c_void_p), ("type", ctypes.c_void_p),
        ("q", ctypes.c_void_p), ("mol", ctypes.c_void_p),
        ("lj", ctypes.c_void_p), ("excl", ctypes.c_void_p),
        ("com", ctypes.c_void_p),
        ("e1", ctypes.c_double * 3), ("e2", ctypes.c_double * 3),
        ("e3", ctypes.c_double * 3), ("origin", ctypes.c_double * 3),
        ("lj_switchon", ctypes.c_double), ("lj_cutoff", ctypes.c_double),
        ("el_switchon", ctypes.c_double), ("el_cutoff", ctypes.c_double),
        ("alg_cutoff", ctypes.c_double), ("cell_size", ctypes.c_double),
        ("csf", ctypes.c_double), ("alpha", ctypes.c_double),
        ("elec", ctypes.c_int32), ("el_switch", ctypes.c_int32),
        ("flags", ctypes.c_int32), ("pad", ctypes.c_int32),
    ]


_LIB = None


def build():
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)


def build_ref():
    """Compiles the reference's own path (needs the reference sources)."""
    subprocess.run(["make", "-s", "-C", HERE, "-j8", "ref"], check=True)


def ref_available():
    return os.path.exists(REF_DRIVER)


def load():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            build()
        lib = ctypes.CDLL(LIB_PATH)
        P = ctypes.POINTER(OrcSystem)
        vp = ctypes.c_void_p
        lib.orc_evaluate.argtypes = [P, vp, vp, vp, vp, vp, ctypes.c_int64,
                                     ctypes.POINTER(ctypes.c_int64)]
        lib.orc_evaluate_bruteforce.argtypes = lib.orc_evaluate.argtypes
        lib.orc_grid_info.argtypes = [P, vp, vp, vp, ctypes.c_int32]
        lib.orc_cells.argtypes = [P, vp]
        lib.orc_min_diff.argtypes = [P, vp, vp, vp, vp]
        lib.orc_poly5.argtypes = [ctypes.c_double]
        lib.orc_poly5.restype = ctypes.c_double
        lib.orc_switch_c2.argtypes = [ctypes.c_double, ctypes.c_double,
                                      ctypes.c_double, vp, vp]
        _LIB = lib
    return _LIB


class System:
    """Flat arrays of one evaluation's inputs (kept alive for ctypes)."""

    def __init__(self, topology, positions, cutoff, lj_switch, elec,
                 elec_switch, alpha=0.0, flags=0, com=None):
        t = topology
        self.pos = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1, 3)
        self.type = np.ascontiguousarray(t.types, dtype=np.int32)
        self.q = np.ascontiguousarray(t.scaled_charges, dtype=np.float64)
        self.mol = np.ascontiguousarray(t.molecules, dtype=np.int32)
        self.lj = np.ascontiguousarray(t.lennard_jones_parameters.table,
                                       dtype=np.float64)
        self.excl = np.ascontiguousarray(t.exclusions.table(), dtype=np.int32)
        # ExclusionTable::getTable() rows, which the Ewald correction walks
        self.excl_table = np.ascontiguousarray(t.exclusions.get_table(),
                                               dtype=np.int32)
        if com is None and (flags & 2):
            com = t.molecule_centers(self.pos)
        self.com = (np.ascontiguousarray(com, dtype=np.float64)
                    if com is not None else np.zeros((1, 3)))
        bc = t.boundary_conditions
        self.box = (bc.e1, bc.e2, bc.e3, bc.origin)
        self.cell_size = t.cell_manager.cell_size
        self.csf = t.coulomb_scaling_factor
        self.cutoff = float(cutoff)
        self.lj_switch = tuple(float(v) for v in lj_switch)
        self.elec = elec
        self.c1 = elec_switch[0] == "C1"
        self.elec_switch = ((0.0, float(elec_switch[1])) if self.c1
                            else tuple(float(v) for v in elec_switch))
        self.alpha = float(alpha)
        self.flags = int(flags)

    @classmethod
    def from_workload(cls, w, positions=None, flags=0, com=None):
        return cls(w.topology, w.positions if positions is None else positions,
                   w.cutoff, w.lj_switch, w.elec, w.elec_switch, w.alpha,
                   flags, com)

    @property
    def n(self):
        return len(self.pos)

    def struct(self):
        s = OrcSystem()
        s.n_atoms = self.n
        s.n_types = self.lj.shape[0]
        s.n_excl = len(self.excl)
        s.n_mol = len(self.com)
        s.pos = self.pos.ctypes.data
        s.type = self.type.ctypes.data
        s.q = self.q.ctypes.data
        s.mol = self.mol.ctypes.data
        s.lj = self.lj.ctypes.data
        s.excl = self.excl.ctypes.data if len(self.excl) else None
        s.com = self.com.ctypes.data
        for name, v in zip(("e1", "e2", "e3", "origin"), self.box):
            getattr(s, name)[:] = [float(x) for x in v]
        s.lj_switchon, s.lj_cutoff = self.lj_switch
        s.el_switchon, s.el_cutoff = self.elec_switch
        s.alg_cutoff = self.cutoff
        s.cell_size = self.cell_size
        s.csf = self.csf
        s.alpha = self.alpha
        s.elec = ELEC[self.elec]
        s.el_switch = 1 if self.c1 else 0
        s.flags = self.flags
        return s


class Result:
    def __init__(self, n):
        self.forces = np.zeros((n, 3))
        self.energies = np.zeros(2)  # LJ, electrostatic
        self.virial = np.zeros(9)
        self.molvirial = np.zeros(9)
        self.n_pairs = 0
        self.pairs = None


def _run(fn, system, record_pairs, pair_cap=None):
    lib = load()
    st = system.struct()
    r = Result(system.n)
    npairs = ctypes.c_int64()
    pairs = None
    cap = 0
    if record_pairs:
        cap = pair_cap if pair_cap is not None else max(
            1, system.n * 400)
        pairs = np.zeros((cap, 3), dtype=np.int32)
    rc = fn(ctypes.byref(st), r.forces.ctypes.data, r.energies.ctypes.data,
            r.virial.ctypes.data, r.molvirial.ctypes.data,
            pairs.ctypes.data if pairs is not None else None, cap,
            ctypes.byref(npairs))
    if rc != 0:
        raise RuntimeError("oracle evaluation failed (%d)" % rc)
    r.n_pairs = npairs.value
    if pairs is not None:
        if r.n_pairs > cap:
            return _run(fn, system, True, r.n_pairs)
        r.pairs = pairs[:r.n_pairs]
    return r


def evaluate(system, record_pairs=False):
    """Cell-list evaluation in the reference's enumeration order."""
    return _run(load().orc_evaluate, system, record_pairs)


def evaluate_threads(system, nthreads=0):
    """orc_evaluate over host threads (0 = all): (forces, energies[2],
    virial[9], pairs, threads used).  CPU-baseline timing only."""
    lib = load()
    if not hasattr(lib, "_thr_ready"):
        lib.orc_evaluate_threads.argtypes = [ctypes.POINTER(OrcSystem)] + [
            ctypes.c_void_p] * 3 + [ctypes.c_int, ctypes.POINTER(ctypes.c_int64)]
        lib.orc_evaluate_threads.restype = ctypes.c_int
        lib._thr_ready = True
    st = system.struct()
    f = np.zeros((system.n, 3))
    e = np.zeros(2)
    vir = np.zeros(9)
    n = ctypes.c_int64()
    used = lib.orc_evaluate_threads(ctypes.byref(st), f.ctypes.data,
                                    e.ctypes.data, vir.ctypes.data,
                                    int(nthreads), ctypes.byref(n))
    if used < 0:
        raise RuntimeError("threaded oracle evaluation unavailable")
    return f, e, vir, n.value, used


def evaluate_bruteforce(system, record_pairs=False):
    """All i<j pairs, same per-pair function."""
    return _run(load().orc_evaluate_bruteforce, system, record_pairs)


def grid_info(system):
    lib = load()
    st = system.struct()
    n3 = np.zeros(3, dtype=np.int32)
    cs3 = np.zeros(3)
    cap = 4096
    deltas = np.zeros((cap, 3), dtype=np.int32)
    nd = lib.orc_grid_info(ctypes.byref(st), n3.ctypes.data, cs3.ctypes.data,
                           deltas.ctypes.data, cap)
    return n3, cs3, deltas[:nd]


def cells(system):
    lib = load()
    st = system.struct()
    out = np.zeros((system.n, 3), dtype=np.int32)
    lib.orc_cells(ctypes.byref(st), out.ctypes.data)
    return out


def min_diff(system, c1, c2):
    lib = load()
    st = system.struct()
    a = np.ascontiguousarray(c1, dtype=np.float64)
    b = np.ascontiguousarray(c2, dtype=np.float64)
    d = np.zeros(3)
    d2 = np.zeros(1)
    lib.orc_min_diff(ctypes.byref(st), a.ctypes.data, b.ctypes.data,
                     d.ctypes.data, d2.ctypes.data)
    return d, d2[0]


def poly5(ar):
    return load().orc_poly5(float(ar))


def switch_c2(switchon, cutoff, d2):
    v = np.zeros(1)
    dv = np.zeros(1)
    load().orc_switch_c2(switchon, cutoff, d2, v.ctypes.data, dv.ctypes.data)
    return v[0], dv[0]


def sorted_pairs(pairs):
    """Rows (i, j, class) sorted by (i, j)."""
    p = np.asarray(pairs, dtype=np.int64)
    if not len(p):
        return p.reshape(0, 3)
    key = p[:, 0] * (p[:, 1].max() + 1) + p[:, 1]
    return p[np.argsort(key, kind="stable")]


# ----------------------------------------------------- reference driver IO
REF_KIND = {("coulomb", False): 0, ("ewald", False): 1, ("coulomb", True): 2}


def write_ref_input(system, path):
    kind = REF_KIND[(system.elec if system.elec != "none" else "coulomb",
                     system.c1)]
    hdr = np.array([0x504d4e42, system.n, system.lj.shape[0], len(system.excl),
                    kind, system.flags, len(system.com), 0], dtype=np.int32)
    par = np.zeros(24)
    par[0:3], par[3:6], par[6:9], par[9:12] = system.box
    par[12], par[13] = system.lj_switch
    par[14], par[15] = system.elec_switch
    par[16] = system.cutoff
    par[17] = system.cell_size
    par[18] = system.csf
    par[19] = system.alpha
    q = system.q if system.elec != "none" else np.zeros_like(system.q)
    with open(path, "wb") as f:
        f.write(hdr.tobytes())
        f.write(par.tobytes())
        f.write(system.pos.tobytes())
        f.write(system.type.tobytes())
        f.write(np.ascontiguousarray(q).tobytes())
        f.write(system.mol.tobytes())
        f.write(system.lj.tobytes())
        f.write(system.excl.tobytes())
        f.write(system.com.tobytes())


def run_ref(system):
    """Runs ProtoMol's own NonbondedCutoffSystemForce (built from the
    reference sources) on the system; returns a Result with pairs in the
    reference's enumeration order."""
    if not ref_available():
        raise FileNotFoundError(REF_DRIVER)
    with tempfile.TemporaryDirectory() as d:
        fin, fout = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
        write_ref_input(system, fin)
        subprocess.run([REF_DRIVER, fin, fout], check=True,
                       stdout=subprocess.DEVNULL)
        raw = open(fout, "rb").read()
    n = system.n
    r = Result(n)
    off = 0
    r.forces = np.frombuffer(raw, np.float64, 3 * n, off).reshape(n, 3).copy()
    off += 24 * n
    r.energies = np.frombuffer(raw, np.float64, 2, off).copy()
    off += 16
    v = np.frombuffer(raw, np.float64, 18, off)
    r.virial, r.molvirial = v[:9].copy(), v[9:].copy()
    off += 144
    r.n_pairs = int(np.frombuffer(raw, np.int64, 1, off)[0])
    off += 8
    r.pairs = np.frombuffer(raw, np.int32, 3 * r.n_pairs, off).reshape(-1, 3).copy()
    return r


# ------------------------------------------------------------------ leapfrog
class OrcLf(ctypes.Structure):
    _fields_ = [
        ("n_cons", ctypes.c_int32), ("cons", ctypes.c_void_p),
        ("rest", ctypes.c_void_p), ("mass", ctypes.c_void_p),
        ("timestep", ctypes.c_double), ("shake_eps", ctypes.c_double),
        ("rattle_eps", ctypes.c_double), ("shake_max_iter", ctypes.c_int32),
        ("rattle_max_iter", ctypes.c_int32), ("shake", ctypes.c_int32),
        ("rattle", ctypes.c_int32),
        ("n_bonds", ctypes.c_int32), ("n_angles", ctypes.c_int32),
        ("bonds", ctypes.c_void_p), ("bond_k", ctypes.c_void_p),
        ("bond_r0", ctypes.c_void_p), ("angles", ctypes.c_void_p),
        ("angle_k", ctypes.c_void_p), ("angle_t0", ctypes.c_void_p),
        ("ub_k", ctypes.c_void_p), ("ub_r0", ctypes.c_void_p),
        ("extra", ctypes.c_void_p), ("extra_user", ctypes.c_void_p),
    ]


EXTRA_FN = ctypes.CFUNCTYPE(None, ctypes.POINTER(ctypes.c_double),
                            ctypes.POINTER(ctypes.c_double),
                            ctypes.POINTER(ctypes.c_double), ctypes.c_void_p)


def _extra_callback(n, extra):
    """A C callback for orc_lf.extra from extra(x (n, 3)) -> (forces (n,
    3), electrostatic energy), added to the force group's sums."""
    def cb(px, pf, pe, _user):
        x = np.ctypeslib.as_array(px, shape=(n, 3))
        f = np.ctypeslib.as_array(pf, shape=(n, 3))
        fe, ee = extra(x.copy())
        f += fe
        pe[0] += ee
    return EXTRA_FN(cb)


class MDState:
    """Positions, velocities, forces at the positions and SHAKE's previous
    positions, as LeapfrogIntegrator and its modifiers keep them."""

    def __init__(self, x, v, f, last=None, energies=None, bonded=None):
        self.x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 3).copy()
        self.v = np.ascontiguousarray(v, dtype=np.float64).reshape(-1, 3).copy()
        self.f = np.ascontiguousarray(f, dtype=np.float64).reshape(-1, 3).copy()
        self.last = self.x.copy() if last is None else np.array(
            last, dtype=np.float64).reshape(-1, 3)
        self.energies = np.zeros(2) if energies is None else np.array(energies)
        self.bonded = np.zeros(2) if bonded is None else np.array(bonded)


def _lf_struct(masses, constraints, rest, timestep, shake, rattle, bonded,
               keep, extra=None):
    cons = np.ascontiguousarray(constraints, dtype=np.int32).reshape(-1, 2)
    rest = np.ascontiguousarray(rest, dtype=np.float64)
    mass = np.ascontiguousarray(masses, dtype=np.float64)
    keep += [cons, rest, mass]
    lf = OrcLf()
    lf.n_cons = len(cons)
    lf.cons = cons.ctypes.data if len(cons) else None
    lf.rest = rest.ctypes.data if len(cons) else None
    lf.mass = mass.ctypes.data
    lf.timestep = float(timestep)
    lf.shake = int(shake is not None)
    lf.rattle = int(rattle is not None)
    lf.shake_eps, lf.shake_max_iter = shake if shake else (0.0, 0)
    lf.rattle_eps, lf.rattle_max_iter = rattle if rattle else (0.0, 0)
    if bonded is not None:
        def arr(name, dt=np.float64):
            a = np.ascontiguousarray(bonded[name], dtype=dt)
            keep.append(a)
            return a.ctypes.data
        lf.n_bonds = len(bonded["bond_k"])
        lf.n_angles = len(bonded["angle_k"])
        lf.bonds = arr("bonds", np.int32)
        lf.bond_k, lf.bond_r0 = arr("bond_k"), arr("bond_r0")
        lf.angles = arr("angles", np.int32)
        lf.angle_k, lf.angle_t0 = arr("angle_k"), arr("angle_theta0")
        if "ub_k" in bonded:
            lf.ub_k, lf.ub_r0 = arr("ub_k"), arr("ub_r0")
    if extra is not None:
        cb = _extra_callback(len(mass), extra)
        keep.append(cb)
        lf.extra = ctypes.cast(cb, ctypes.c_void_p).value
    return lf


def lf_forces(system, positions, masses, bonded=None, extra=None):
    """calculateForces of the leapfrog's force group (cleared forces and
    energies, then the bonded terms, the cutoff force and `extra`, a
    callable x -> (forces, electrostatic energy)): returns an MDState-ready
    (forces, energies (LJ, electrostatic), bonded (bond, angle))."""
    lib = load()
    if not hasattr(lib, "_lff_ready"):
        lib.orc_lf_forces.argtypes = [ctypes.POINTER(OrcSystem),
                                      ctypes.POINTER(OrcLf)] + [ctypes.c_void_p] * 4
        lib._lff_ready = True
    keep = []
    lf = _lf_struct(masses, np.zeros((0, 2)), np.zeros(0), 1.0, None, None,
                    bonded, keep, extra)
    x = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1, 3).copy()
    f = np.zeros_like(x)
    e, b = np.zeros(2), np.zeros(2)
    st = system.struct()
    if lib.orc_lf_forces(ctypes.byref(st), ctypes.byref(lf), x.ctypes.data,
                         f.ctypes.data, e.ctypes.data, b.ctypes.data):
        raise RuntimeError("oracle force evaluation failed")
    return f, e, b


def leapfrog_run(system, state, n_steps, masses, constraints, rest, timestep,
                 shake=None, rattle=None, bonded=None, max_calls=1 << 16,
                 extra=None):
    """LeapfrogIntegrator::run(n_steps) on `state` (updated in place) with
    the cutoff force of `system` and the optional bonded terms as forces.
    shake / rattle are (epsilon, max_iter) or None; bonded is a dict with
    bonds, bond_k, bond_r0, angles, angle_k, angle_theta0 (radians) and
    optionally ub_k, ub_r0; extra (optional) is a further force of the
    group, a callable x -> (forces, electrostatic energy) such as the
    k-space terms.  Returns the sweep count of every SHAKE and
    RATTLE call in call order."""
    lib = load()
    if not hasattr(lib, "_lf_ready"):
        lib.orc_leapfrog_run.argtypes = [
            ctypes.POINTER(OrcSystem), ctypes.POINTER(OrcLf)] + \
            [ctypes.c_void_p] * 6 + [ctypes.c_int32, ctypes.c_void_p,
                                     ctypes.c_int64]
        lib.orc_leapfrog_run.restype = ctypes.c_int64
        lib._lf_ready = True
    st = system.struct()
    keep = []
    lf = _lf_struct(masses, constraints, rest, timestep, shake, rattle,
                    bonded, keep, extra)
    iters = np.zeros(max_calls, dtype=np.int32)
    n = lib.orc_leapfrog_run(ctypes.byref(st), ctypes.byref(lf),
                             state.x.ctypes.data, state.v.ctypes.data,
                             state.f.ctypes.data, state.last.ctypes.data,
                             state.energies.ctypes.data,
                             state.bonded.ctypes.data, int(n_steps),
                             iters.ctypes.data, max_calls)
    if n < 0:
        raise RuntimeError("oracle leapfrog failed")
    return iters[:min(n, max_calls)].copy()


# ------------------------------------------------------------------ Ewald
def ewald_kspace(system, alpha, accuracy, virial=True, extended=False):
    """Reciprocal + correction + self + charged-system terms of
    NonbondedFullEwaldSystemForce: returns (forces, electrostatic energy,
    virial[9], info) with info = (alpha, rc, kc, number of k-vectors).
    extended=True sums the structure factors and energies in long double
    (same terms, without the double running sum's rounding error)."""
    lib = load()
    if not hasattr(lib, "_ew_ready"):
        lib.orc_ewald_eval.argtypes = [ctypes.POINTER(OrcSystem), ctypes.c_double,
                                  ctypes.c_double] + [ctypes.c_void_p] * 4 + [
                                      ctypes.c_int]
        lib._ew_ready = True
    st = system.struct()
    # the correction term iterates ExclusionTable::getTable()
    # (NonbondedFullEwaldSystemForce.h:1105-1129)
    st.n_excl = len(system.excl_table)
    st.excl = system.excl_table.ctypes.data if len(system.excl_table) else None
    f = np.zeros((system.n, 3))
    e = np.zeros(2)
    vir = np.zeros(9)
    info = np.zeros(4)
    if lib.orc_ewald_eval(ctypes.byref(st), float(alpha), float(accuracy),
                     f.ctypes.data, e.ctypes.data,
                     vir.ctypes.data if virial else None, info.ctypes.data,
                     int(bool(extended))):
        raise RuntimeError("oracle Ewald evaluation failed")
    return f, e[1], vir, info
