"""Batched leapfrog dynamics (pmnb_batch_lf_create) and captured runs
(pmnb_lf_run_graph), through the C ABI.

Bars, all bit for bit:
* every replica of a batch integrated together (LeapfrogIntegrator::run
  with SHAKE and RATTLE, integrator/leapfrog/LeapfrogIntegrator.cpp:85-99)
  ends in the positions, velocities and forces, with the SHAKE / RATTLE
  sweep counts, that a single-system integrator gives for that replica
  alone with the same force;
* n replays of a captured run(k) equal n eager run(k) calls, checked after
  every replay with eager work and synchronisations between the replays;
  for one system and for a batch.
The single-system integrator is pinned to the oracle in test_gpu_leapfrog.
"""
import numpy as np
import pytest

from protomoladdon_amd import systems
from protomoladdon_amd.batch import BatchedCutoffForce, BatchedLeapfrog, OUT_SLOTS
from protomoladdon_amd.integrator import LeapfrogIntegrator, topology_constraints

pytestmark = pytest.mark.gpu

KB = 0.001987191
DEFAULT = (1e-5, 30)
# boxes 17.1 (one column per axis), 20.8, 27.6, 31.0, 35.5 A
SIZES = (167, 300, 700, 1000, 1500)


def maxwell(masses, seed, temperature=300.0):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((len(masses), 3)) * np.sqrt(
        KB * temperature / np.asarray(masses))[:, None]


@pytest.fixture(scope="module")
def replicas():
    return [systems.config3_replica(i, n) for i, n in enumerate(SIZES)]


def single_run(w, precision, v0, chunks, n_steps):
    """Single-system run: (x, v, f, out, sweep counts) after `chunks` calls
    of run(n_steps)."""
    import torch
    cons = topology_constraints(w.topology)
    force = w.make_force(precision)
    lf = LeapfrogIntegrator(1.0, force, shake=DEFAULT, rattle=DEFAULT)
    dev = torch.device("cuda:0")
    x = torch.tensor(w.positions, dtype=torch.float64, device=dev)
    v = torch.tensor(v0, dtype=torch.float64, device=dev)
    f = torch.zeros_like(x)
    out = torch.zeros(20, dtype=torch.float64, device=dev)
    lf.initialize(w.topology, x, f, out, cons)
    lf.record_iterations(1 << 12)
    for _ in range(chunks):
        lf.run(w.topology, x, v, f, out, n_steps, cons)
    torch.cuda.synchronize()
    its, n, _ = lf.record_iterations(0)
    res = (x.cpu().numpy(), v.cpu().numpy(), f.cpu().numpy(),
           out.cpu().numpy(), its[:n])
    lf.close()
    force.close()
    return res


class Batch:
    def __init__(self, ws, precision, v0s):
        import torch
        self.torch = torch
        self.ws = ws
        self.bf = BatchedCutoffForce.from_workloads(ws, precision=precision)
        cons = [topology_constraints(w.topology) for w in ws]
        self.lf = BatchedLeapfrog(1.0, self.bf, [w.topology.masses for w in ws],
                                  cons, shake=DEFAULT, rattle=DEFAULT)
        dev = torch.device("cuda:0")
        self.x = torch.tensor(np.concatenate([w.positions for w in ws]),
                              dtype=torch.float64, device=dev)
        self.v = torch.tensor(np.concatenate(v0s), dtype=torch.float64,
                              device=dev)
        self.f = torch.zeros_like(self.x)
        self.out = torch.zeros(len(ws) * OUT_SLOTS, dtype=torch.float64,
                               device=dev)
        self.lf.initialize(self.x, self.f, self.out)
        self.lf.record_iterations(1 << 16)

    def host(self, r):
        a, b = self.bf.offsets[r], self.bf.offsets[r + 1]
        return (self.x[a:b].cpu().numpy(), self.v[a:b].cpu().numpy(),
                self.f[a:b].cpu().numpy(),
                self.out[r * OUT_SLOTS:(r + 1) * OUT_SLOTS].cpu().numpy())

    def close(self):
        self.lf.close()
        self.bf.close()


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_batch_leapfrog_matches_single_replicas(replicas, precision):
    v0s = [maxwell(w.topology.masses, 50 + i) for i, w in enumerate(replicas)]
    chunks, n_steps = 2, 5
    b = Batch(replicas, precision, v0s)
    print("batch leapfrog", precision, "atoms", b.bf.n_atoms, flush=True)
    for _ in range(chunks):
        b.lf.run(b.x, b.v, b.f, b.out, n_steps)
    b.torch.cuda.synchronize()
    its, n_calls, unc = b.lf.record_iterations(0)
    assert unc == 0
    assert its.shape == (2 * chunks * n_steps, len(replicas))
    for r, w in enumerate(replicas):
        print("replica", r, w.n_atoms, flush=True)
        x, v, f, out, its_s = single_run(w, precision, v0s[r], chunks, n_steps)
        xb, vb, fb, ob = b.host(r)
        assert np.array_equal(xb, x) and np.array_equal(vb, v)
        assert np.array_equal(fb, f)
        assert np.array_equal(ob, out)
        assert list(its[:, r]) == list(its_s)
    b.close()


def test_batch_leapfrog_graph_replays_match_eager(replicas):
    v0s = [maxwell(w.topology.masses, 70 + i) for i, w in enumerate(replicas)]
    eager = Batch(replicas, "mixed", v0s)
    graph = Batch(replicas, "mixed", v0s)
    for rep in range(3):
        print("replay", rep, flush=True)
        eager.lf.run(eager.x, eager.v, eager.f, eager.out, 4)
        graph.lf.run_graph(graph.x, graph.v, graph.f, graph.out, 4, 1)
        eager.torch.cuda.synchronize()
        for t in ("x", "v", "f", "out"):
            assert eager.torch.equal(getattr(eager, t), getattr(graph, t)), t
    # several replays in one call
    eager.lf.run(eager.x, eager.v, eager.f, eager.out, 4)
    eager.lf.run(eager.x, eager.v, eager.f, eager.out, 4)
    graph.lf.run_graph(graph.x, graph.v, graph.f, graph.out, 4, 2)
    eager.torch.cuda.synchronize()
    for t in ("x", "v", "f", "out"):
        assert eager.torch.equal(getattr(eager, t), getattr(graph, t)), t
    its_e, _, unc_e = eager.lf.record_iterations(0)
    assert unc_e == 0
    eager.close()
    graph.close()


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_single_system_graph_replays_match_eager(precision):
    import torch
    w = systems.small_water(n_mol=1000, elec="ewald")
    cons = topology_constraints(w.topology)
    v0 = maxwell(w.topology.masses, 5)
    dev = torch.device("cuda:0")
    runs = []
    for mode in ("eager", "graph"):
        force = w.make_force(precision)
        lf = LeapfrogIntegrator(1.0, force, shake=DEFAULT, rattle=DEFAULT)
        x = torch.tensor(w.positions, dtype=torch.float64, device=dev)
        v = torch.tensor(v0, dtype=torch.float64, device=dev)
        f = torch.zeros_like(x)
        out = torch.zeros(20, dtype=torch.float64, device=dev)
        lf.initialize(w.topology, x, f, out, cons)
        runs.append((force, lf, x, v, f, out))
    (fe, le, xe, ve, ffe, oe), (fg, lg, xg, vg, ffg, og) = runs
    for rep in range(3):
        print("replay", rep, precision, flush=True)
        le.run(w.topology, xe, ve, ffe, oe, 3, cons)
        lg.run_graph(w.topology, xg, vg, ffg, og, 3, 1, cons)
        torch.cuda.synchronize()
        assert torch.equal(xe, xg) and torch.equal(ve, vg)
        assert torch.equal(ffe, ffg) and torch.equal(oe, og)
    for force, lf, *_ in runs:
        lf.close()
        force.close()


def test_batch_leapfrog_rejects_bad_input(replicas):
    bf = BatchedCutoffForce.from_workloads(replicas[:2], precision="mixed")
    cons = [topology_constraints(w.topology) for w in replicas[:2]]
    masses = [w.topology.masses for w in replicas[:2]]
    with pytest.raises(ValueError):
        BatchedLeapfrog(1.0, bf, masses[:1], cons, shake=DEFAULT)
    # a constraint joining the two replicas
    bad = list(cons)
    pairs, rest = cons[1]
    pairs = np.array(pairs, copy=True)
    pairs[0, 0] = -bf.offsets[1]  # batch index 0: an atom of replica 0
    bad[1] = (pairs, rest)
    from protomoladdon_amd._lib import PmnbError
    with pytest.raises(PmnbError):
        BatchedLeapfrog(1.0, bf, masses, bad, shake=DEFAULT)
    bf.close()
