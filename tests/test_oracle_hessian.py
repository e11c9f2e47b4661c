"""Pins the pair-Hessian oracle (oracle/hessian.py), the check of the
adjoint's Hessian-vector product.

Central finite differences of the C force oracle (itself pinned to the
reference build) give columns of -H; the C2 switch Hessian is checked
against finite differences of the switch's gradient S~ x.  A planted
mistake in the oracle (a dropped switch term) must fail the same check."""
import numpy as np
import pytest

from oracle import hessian as OH
from oracle import oracle as O
from protomoladdon_amd import systems


@pytest.fixture(scope="module")
def water():
    # 60 waters at water density (box ~12.2 A < 2 cutoffs: minimum image
    # through the cell list), LJ + Coulomb, C2 5 -> 6
    return systems.water_workload("water_h", 60, (60 / 0.0334) ** (1 / 3),
                                  77, 6.0, 5.0, "coulomb")


def forces(w, pos):
    return O.evaluate(O.System.from_workload(w, positions=pos)).forces


def fd_columns(w, cols, h=1e-5):
    out = []
    for k in cols:
        p = w.positions.copy()
        p.reshape(-1)[k] += h
        fp = forces(w, p)
        p.reshape(-1)[k] -= 2 * h
        fm = forces(w, p)
        out.append(-(fp - fm).reshape(-1) / (2 * h))
    return np.stack(out, 1)


def test_hessian_matches_finite_differences(water):
    w = water
    ref = O.evaluate(O.System.from_workload(w), record_pairs=True)
    H = OH.hessian(w, ref.pairs)
    rng = np.random.default_rng(5)
    cols = rng.choice(3 * w.n_atoms, 12, replace=False)
    fd = fd_columns(w, cols)
    Hc = H[:, cols].toarray()
    scale = np.abs(fd).max()
    assert np.abs(Hc - fd).max() <= 2e-6 * scale
    # symmetric, and rows sum to zero (translation invariance)
    assert abs(H - H.T).max() <= 1e-12 * abs(H).max()
    tr = np.asarray(H.sum(axis=1)).reshape(-1, 3)
    assert np.abs(tr).max() <= 1e-9 * abs(H).max()


def test_planted_mistake_is_caught(water, monkeypatch):
    w = water
    ref = O.evaluate(O.System.from_workload(w), record_pairs=True)
    cols = np.array([0, 4, 11])
    fd = fd_columns(w, cols)
    orig = OH.c2_parts

    def no_switch_hessian(a, on, cut):
        S, Sd, cI, cR = orig(a, on, cut)
        return S, Sd, 0.0 * cI, 0.0 * cR

    monkeypatch.setattr(OH, "c2_parts", no_switch_hessian)
    Hc = OH.hessian(w, ref.pairs)[:, cols].toarray()
    assert np.abs(Hc - fd).max() > 1e-4 * np.abs(fd).max()


def test_c2_switch_hessian_finite_differences():
    on, cut = 7.5, 9.0
    rng = np.random.default_rng(3)
    for _ in range(20):
        x = rng.normal(size=3)
        x *= rng.uniform(on + 0.05, cut - 0.05) / np.linalg.norm(x)
        S, Sd, cI, cR = OH.c2_parts(np.array([x @ x]), on, cut)
        Hs = cI[0] * np.eye(3) + cR[0] * np.outer(x, x)
        h = 1e-6
        fd = np.zeros((3, 3))
        for k in range(3):
            e = np.zeros(3)
            e[k] = h
            gp = O.switch_c2(on, cut, (x + e) @ (x + e))[1] * (x + e)
            gm = O.switch_c2(on, cut, (x - e) @ (x - e))[1] * (x - e)
            fd[:, k] = (gp - gm) / (2 * h)
        assert np.abs(Hs - fd).max() <= 1e-6 * max(1.0, np.abs(fd).max())
        # the restated value and S~ are the oracle's
        v, dv = O.switch_c2(on, cut, x @ x)
        assert S[0] == pytest.approx(v, rel=1e-14)
        assert Sd[0] == pytest.approx(dv, rel=1e-12)
