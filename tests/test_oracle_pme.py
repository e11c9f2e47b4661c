"""Pins the oracle's numpy restatement of NonbondedPMEwaldSystemForce
(oracle/pme.py: B-spline spreading, FFT, influence-function sum with its
virial, force interpolation, plus the correction, self and charged-system
terms) by what the mathematics fixes:

* the B-spline weights of BSpline::set sum to one at every fraction, their
  derivatives are the w-derivative of the weights, and order 4 is the
  textbook cubic B-spline (1-w)^3/6, (3w^3-6w^2+4)/6, (-3w^3+3w^2+3w+1)/6,
  w^3/6;
* Grid::dftmod's B(m) for order 4 is the closed form 9 / (2 + cos(2 pi m /
  n))^2 of the Euler exponential spline;
* smooth PME converges to the Ewald sum: at order 10 on a fine grid the
  energy, forces and atomic virial equal those of the direct k-space sum
  (oracle/pmnb_oracle.c, itself pinned by the Madelung constant and alpha
  independence) to ~1e-10, with net charge, FULL and MODIFIED exclusions,
  rings (getTable rows) and an orthorhombic box of odd grid sizes; at
  lower order or coarser grids the error shrinks as the grid is refined;
* the PME forces are minus the finite-difference gradient of the PME
  energy (smooth PME differentiates its own interpolated energy).
A dropped term (correction, self, charged system), a wrong sign, a
transposed axis or a missed doubling of the half spectrum fails the
comparison with the direct sum.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from oracle import pme as OP
from tests.test_oracle_ewald import chains, make_system, rings, rock_salt


@pytest.mark.parametrize("order", [2, 4, 6, 8, 10])
def test_bspline_partition_of_unity_and_derivative(order):
    w = np.linspace(0.0, 0.999, 37)
    th, dth = OP.bspline(order, w)
    assert np.abs(th.sum(axis=1) - 1.0).max() < 1e-14
    assert np.abs(dth.sum(axis=1)).max() < 1e-13
    assert (th >= -1e-16).all()
    h = 1e-6
    tp, _ = OP.bspline(order, w[1:-1] + h)
    tm, _ = OP.bspline(order, w[1:-1] - h)
    np.testing.assert_allclose(dth[1:-1], (tp - tm) / (2 * h), atol=1e-8)


def test_bspline_order4_is_the_cubic_bspline():
    w = np.linspace(0.0, 0.999, 29)
    th, _ = OP.bspline(4, w)
    ref = np.stack([(1 - w) ** 3 / 6, (3 * w ** 3 - 6 * w ** 2 + 4) / 6,
                    (-3 * w ** 3 + 3 * w ** 2 + 3 * w + 1) / 6, w ** 3 / 6],
                   axis=1)
    np.testing.assert_allclose(th, ref, atol=1e-15)


@pytest.mark.parametrize("n", [16, 21, 32])
def test_dftmod_order4_closed_form(n):
    m = np.arange(n)
    ref = 9.0 / (2.0 + np.cos(2 * math.pi * m / n)) ** 2
    np.testing.assert_allclose(OP.dftmod(4, n), ref, rtol=1e-13)


def rel(a, b):
    return np.abs(np.asarray(a) - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("case", ["chains", "rings", "orthorhombic_odd",
                                  "rock_salt"])
def test_high_order_pme_equals_direct_ewald_sum(case):
    if case == "chains":
        pos, q, L, bonds, angles, dih, mol = chains()
        w, s = make_system(pos, q, L, bonds, angles, dih, mol, csf=0.5)
        grid, alpha = (64, 64, 64), 0.7
    elif case == "rings":
        pos, q, L, bonds, angles, dih, mol = rings()
        w, s = make_system(pos, q, L, bonds, angles, dih, mol, csf=0.5)
        grid, alpha = (64, 64, 64), 0.6
    elif case == "orthorhombic_odd":
        pos, q, _, bonds, angles, dih, mol = chains(seed=11)
        L = np.array([13.0, 14.5, 12.2])
        pos = np.mod(pos, L)
        w, s = make_system(pos, q, L, bonds, angles, dih, mol, csf=0.5)
        grid, alpha = (63, 71, 61), 0.7
    else:
        pos, q = rock_salt(2, 5.64)
        w, s = make_system(pos, q, 2 * 5.64)
        grid, alpha = (48, 48, 48), 0.8
    f0, e0, v0, _ = O.ewald_kspace(s, alpha, 1e-14)
    f1, e1, v1 = OP.pme_kspace(w, grid, 10, alpha)
    assert abs(e1 - e0) <= 1e-9 * abs(e0)
    if case == "rock_salt":  # every ion at a centre of symmetry
        assert np.abs(f1).max() < 1e-9
    else:
        assert rel(f1, f0) < 1e-9
    assert rel(v1, v0) < 1e-9
    assert np.abs(v1 - v1.reshape(3, 3).T.reshape(9)).max() <= \
        1e-12 * np.abs(v1).max()


def test_pme_error_falls_with_grid_and_order():
    pos, q, L, bonds, angles, dih, mol = chains()
    w, s = make_system(pos, q, L, bonds, angles, dih, mol, csf=0.5)
    f0, e0, _, _ = O.ewald_kspace(s, 0.7, 1e-14)
    errs = []
    for order, g in ((4, 16), (4, 32), (6, 32), (8, 48)):
        f1, e1, _ = OP.pme_kspace(w, (g, g, g), order, 0.7)
        errs.append(rel(f1, f0))
    assert errs[0] > 1e-2 and errs[-1] < 1e-6
    assert all(a > 10 * b for a, b in zip(errs, errs[1:]))


def test_pme_forces_are_its_energy_gradient():
    pos, q, L, bonds, angles, dih, mol = chains(seed=5)
    w, _ = make_system(pos, q, L, bonds, angles, dih, mol, csf=0.5)
    grid, order, alpha = (20, 20, 20), 4, 0.7
    f, _, _ = OP.pme_kspace(w, grid, order, alpha)
    h = 1e-5
    for i in (0, 7, 13):
        for k in range(3):
            ep = []
            for sgn in (1, -1):
                p = pos.copy()
                p[i, k] += sgn * h
                wp, _ = make_system(p, q, L, bonds, angles, dih, mol, csf=0.5)
                ep.append(OP.pme_kspace(wp, grid, order, alpha,
                                        virial=False)[1])
            fd = -(ep[0] - ep[1]) / (2 * h)
            assert abs(fd - f[i, k]) <= 1e-6 * max(1.0, abs(f[i, k]))


def test_default_alpha_meets_the_accuracy():
    from scipy.special import erfc
    for rc, acc in ((9.0, 1e-6), (10.0, 1e-5), (12.0, 1e-8)):
        a = OP.default_alpha(rc, acc)
        assert abs(erfc(a * rc) / rc - acc) <= 1e-12 * acc
