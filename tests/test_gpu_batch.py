"""Batched replicas through the C ABI (pmnb_batch_*).

Replica r of a batch must give what a single context gives for that replica
alone, bit for bit (same kernels, same per-replica work order), and the
oracle's pair count and forces (mixed: RMS relative force error <= 1e-5,
FP64: within 1e-12 of the largest force).  The replica sizes cover boxes
narrower than three cutoffs (per-pair minimum image in x and y) and wider
ones.  Graph replays are checked one by one against an eager evaluation,
with eager work between the replays."""
import numpy as np
import pytest

from oracle import oracle as O
from protomoladdon_amd import systems
from protomoladdon_amd.batch import BatchedCutoffForce, OUT_SLOTS

pytestmark = pytest.mark.gpu

# boxes 17.1 (one column per axis), 20.8, 27.6 (three columns), 31.0, 35.5 A
SIZES = (167, 300, 700, 1000, 1500)


@pytest.fixture(scope="module")
def replicas():
    return [systems.config3_replica(i, n) for i, n in enumerate(SIZES)]


def rms_rel(f, ref):
    return float(np.sqrt(np.mean(np.sum((f - ref) ** 2, axis=1)) /
                         np.mean(np.sum(ref ** 2, axis=1))))


def batch_eval(bf, ws, flags=1):
    import torch
    pos = torch.from_numpy(np.concatenate([w.positions for w in ws])).cuda()
    frc = torch.zeros_like(pos)
    out = torch.zeros(len(ws) * OUT_SLOTS, dtype=torch.float64, device="cuda")
    bf.evaluate_device(pos, frc, out, flags)
    torch.cuda.synchronize()
    return pos, frc.cpu().numpy(), out.cpu().numpy().reshape(len(ws), OUT_SLOTS)


def single_eval(w, precision, flags=1):
    import torch
    f = w.make_force(precision)
    pos = torch.from_numpy(w.positions).cuda()
    frc = torch.zeros_like(pos)
    out = torch.zeros(OUT_SLOTS, dtype=torch.float64, device="cuda")
    f.evaluate_device(w.topology, pos, frc, out, flags)
    torch.cuda.synchronize()
    n = f.last_pair_count()
    f.close()
    return frc.cpu().numpy(), out.cpu().numpy(), n


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_batch_matches_single_and_oracle(replicas, precision):
    bf = BatchedCutoffForce.from_workloads(replicas, precision=precision)
    print("batch", precision, "atoms", bf.n_atoms, flush=True)
    pos, frc, out = batch_eval(bf, replicas)
    npairs = bf.pair_count()
    # the pairs the batch's pair kernel evaluated, split by replica
    bpairs = bf.force_pairs(pos)
    bf.close()
    total = 0
    for r, w in enumerate(replicas):
        a, b = bf.offsets[r], bf.offsets[r + 1]
        f1, o1, n1 = single_eval(w, precision)
        np.testing.assert_array_equal(frc[a:b], f1)
        np.testing.assert_array_equal(out[r], o1)
        ref = O.evaluate(O.System.from_workload(w, flags=1), record_pairs=True)
        total += ref.n_pairs
        assert n1 == ref.n_pairs
        mine = bpairs[(bpairs[:, 0] >= a) & (bpairs[:, 0] < b)].copy()
        assert ((mine[:, 1] >= a) & (mine[:, 1] < b)).all()
        mine[:, :2] -= a
        np.testing.assert_array_equal(mine, O.sorted_pairs(ref.pairs))
        if precision == "mixed":
            assert rms_rel(frc[a:b], ref.forces) <= 1e-5
            np.testing.assert_allclose(out[r, :2], ref.energies, rtol=2e-5)
        else:
            scale = max(1.0, np.abs(ref.forces).max())
            assert np.abs(frc[a:b] - ref.forces).max() <= 1e-11 * scale
            np.testing.assert_allclose(out[r, :2], ref.energies, rtol=1e-11)
            np.testing.assert_allclose(out[r, 2:11], ref.virial, rtol=1e-10,
                                       atol=1e-9 * np.abs(ref.virial).max())
    assert npairs == total


def test_graph_replays_match_eager(replicas):
    import torch
    bf = BatchedCutoffForce.from_workloads(replicas, precision="mixed")
    pos, frc0, out0 = batch_eval(bf, replicas)
    frc = torch.full_like(pos, 7.0)   # replays must zero it themselves
    out = torch.full((len(replicas) * OUT_SLOTS,), 3.0, dtype=torch.float64,
                     device="cuda")
    for k in range(3):
        bf.run_graph(pos, frc, out, flags=1, n_steps=1)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(frc.cpu().numpy(), frc0)
        np.testing.assert_array_equal(
            out.cpu().numpy().reshape(len(replicas), OUT_SLOTS), out0)
        # eager work on the same buffers between replays
        bf.evaluate_device(pos, frc, out, 1)
        torch.cuda.synchronize()
    bf.run_graph(pos, frc, out, flags=1, n_steps=5)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(frc.cpu().numpy(), frc0)
    bf.close()


def test_autotune_keeps_accuracy(replicas):
    import torch
    bf = BatchedCutoffForce.from_workloads(replicas, precision="mixed")
    pos = torch.from_numpy(np.concatenate([w.positions for w in replicas])).cuda()
    frc = torch.zeros_like(pos)
    out = torch.zeros(len(replicas) * OUT_SLOTS, dtype=torch.float64,
                      device="cuda")
    shape, grid, ms = bf.autotune(pos, frc, out, 1)
    print("autotune shape", shape, "grid", grid, "ms", ms, flush=True)
    assert 0 <= shape < 3 and grid >= 0 and ms > 0
    frc.zero_()
    out.zero_()
    bf.evaluate_device(pos, frc, out, 1)
    torch.cuda.synchronize()
    f = frc.cpu().numpy()
    for r, w in enumerate(replicas):
        ref = O.evaluate(O.System.from_workload(w, flags=1))
        a, b = bf.offsets[r], bf.offsets[r + 1]
        assert rms_rel(f[a:b], ref.forces) <= 1e-5
    bf.close()


def test_config3_full_size_512_replicas():
    """The 512-replica workload at its size (2.08 M atoms): graph replays
    equal the eager evaluation bit for bit; every replica's forces within
    RMS 1e-5 of the FP64 oracle and its energies within 2e-5; the pair
    count equals the oracle's over all replicas."""
    import torch
    sizes = systems.config3_sizes()
    ws = [systems.config3_replica(i, n) for i, n in enumerate(sizes)]
    bf = BatchedCutoffForce.from_workloads(ws, precision="mixed")
    print("512 replicas, atoms", bf.n_atoms, flush=True)
    pos, frc0, out0 = batch_eval(bf, ws)
    npairs = bf.pair_count()
    frc = torch.full_like(pos, 7.0)
    out = torch.full((len(ws) * OUT_SLOTS,), 3.0, dtype=torch.float64,
                     device="cuda")
    for k in range(3):
        bf.run_graph(pos, frc, out, flags=1, n_steps=1)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(frc.cpu().numpy(), frc0)
        np.testing.assert_array_equal(
            out.cpu().numpy().reshape(len(ws), OUT_SLOTS), out0)
        bf.evaluate_device(pos, frc, out, 1)
        torch.cuda.synchronize()
    bf.close()
    total, worst = 0, 0.0
    for r, w in enumerate(ws):
        ref = O.evaluate(O.System.from_workload(w, flags=1))
        a, b = bf.offsets[r], bf.offsets[r + 1]
        total += ref.n_pairs
        err = rms_rel(frc0[a:b], ref.forces)
        worst = max(worst, err)
        assert err <= 1e-5, (r, err)
        np.testing.assert_allclose(out0[r, :2], ref.energies, rtol=2e-5)
    print("worst replica rms rel force error", worst, flush=True)
    assert npairs == total


@pytest.mark.parametrize("precision,shape", [("mixed", 2), ("fp64", 1)])
def test_half_shell_shape_matches_oracle(replicas, precision, shape):
    """The half-shell pair kernel (each pair evaluated once, the partner's
    force added through fixed-point LDS slots and global integer atomics),
    selected as a batch shape: the pairs it evaluates are the oracle's, bit
    for bit, forces and energies within the bars of the full shell, and
    repeated evaluations are bitwise equal."""
    bf = BatchedCutoffForce.from_workloads(replicas, precision=precision)
    bf.set_shape(shape)
    pos, frc, out = batch_eval(bf, replicas)
    _, frc2, out2 = batch_eval(bf, replicas)
    np.testing.assert_array_equal(frc2, frc)
    np.testing.assert_array_equal(out2, out)
    bpairs = bf.force_pairs(pos)
    bf.close()
    for r, w in enumerate(replicas):
        a, b = bf.offsets[r], bf.offsets[r + 1]
        ref = O.evaluate(O.System.from_workload(w, flags=1), record_pairs=True)
        mine = bpairs[(bpairs[:, 0] >= a) & (bpairs[:, 0] < b)].copy()
        mine[:, :2] -= a
        np.testing.assert_array_equal(mine, O.sorted_pairs(ref.pairs))
        if precision == "mixed":
            assert rms_rel(frc[a:b], ref.forces) <= 1e-5
            np.testing.assert_allclose(out[r, :2], ref.energies, rtol=2e-5)
        else:
            scale = max(1.0, np.abs(ref.forces).max())
            assert np.abs(frc[a:b] - ref.forces).max() <= 1e-11 * scale
            np.testing.assert_allclose(out[r, :2], ref.energies, rtol=1e-11)
            np.testing.assert_allclose(out[r, 2:11], ref.virial, rtol=1e-10,
                                       atol=1e-9 * np.abs(ref.virial).max())
