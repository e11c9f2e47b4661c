"""The evaluation that allocates nothing (pmnb_evaluate_device_noalloc),
through the C ABI, against pmnb_evaluate_device.

Bar: bit for bit (same kernels, same order of every sum), forces and the
20-double output block, in both precisions, on two calls in a row on one
GPU with the second system smaller than the first (the shared static
scratch keeps the first call's contents), then the first system again.
Systems outside its scope or above its 64 MiB are refused with an error
before anything is launched."""
import numpy as np
import pytest

from protomoladdon_amd import systems
from protomoladdon_amd._lib import PmnbError

pytestmark = pytest.mark.gpu


def both(w, precision, flags=1):
    """(forces, out) of pmnb_evaluate_device and of the no-allocation
    entry, each on zeroed outputs."""
    import torch
    f = w.make_force(precision)
    pos = torch.from_numpy(np.ascontiguousarray(w.positions)).cuda()
    res = []
    for noalloc in (False, True):
        frc = torch.zeros_like(pos)
        out = torch.zeros(20, dtype=torch.float64, device="cuda")
        if noalloc:
            f.evaluate_device_noalloc(w.topology, pos, frc, out, flags)
        else:
            f.evaluate_device(w.topology, pos, frc, out, flags)
        torch.cuda.synchronize()
        res.append((frc.cpu().numpy(), out.cpu().numpy()))
    f.close()
    return res


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_noalloc_matches_evaluate_device(precision):
    big = systems.config0()                       # 3,000 atoms, 31.0 A
    small = systems.small_water(n_mol=300, elec="ewald")   # 900 atoms
    for w in (big, small, big):
        print("noalloc", precision, w.name, w.n_atoms, flush=True)
        (f0, o0), (f1, o1) = both(w, precision)
        assert np.abs(f0).max() > 0
        np.testing.assert_array_equal(f1, f0)
        np.testing.assert_array_equal(o1, o0)


def test_noalloc_accumulates():
    import torch
    w = systems.small_water(n_mol=300, elec="ewald")
    f = w.make_force("mixed")
    pos = torch.from_numpy(np.ascontiguousarray(w.positions)).cuda()
    frc = torch.zeros_like(pos)
    out = torch.zeros(20, dtype=torch.float64, device="cuda")
    f.evaluate_device_noalloc(w.topology, pos, frc, out, 1)
    torch.cuda.synchronize()
    once = frc.cpu().numpy().copy()
    f.evaluate_device_noalloc(w.topology, pos, frc, out, 1)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(frc.cpu().numpy(), 2.0 * once)
    f.close()


def test_noalloc_refuses_what_it_does_not_cover():
    import torch
    # algorithm cutoff below the switching cutoff: the general path
    w = systems.with_algorithm(systems.small_water(n_mol=400), 6.5, 3.0)
    f = w.make_force("fp64")
    pos = torch.from_numpy(np.ascontiguousarray(w.positions)).cuda()
    frc = torch.zeros_like(pos)
    out = torch.zeros(20, dtype=torch.float64, device="cuda")
    with pytest.raises(PmnbError) as e:
        f.evaluate_device_noalloc(w.topology, pos, frc, out, 1)
    assert e.value.code == 3  # PMNB_ERR_UNSUPPORTED
    f.close()
    # 600,000 atoms need more than the 64 MiB scratch
    n_mol = 200000
    big = systems.water_workload("water600k", n_mol,
                                 (n_mol / systems.WATER_DENSITY) ** (1 / 3),
                                 7, 9.0, 7.5, "ewald", alpha=0.32)
    f = big.make_force("mixed")
    pos = torch.from_numpy(np.ascontiguousarray(big.positions)).cuda()
    frc = torch.zeros_like(pos)
    with pytest.raises(PmnbError) as e:
        f.evaluate_device_noalloc(big.topology, pos, frc, out, 1)
    assert e.value.code == 5  # PMNB_ERR_CAPACITY
    assert not frc.abs().max().item()
    f.close()
