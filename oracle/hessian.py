"""Test oracle: the reference's pair Hessians, restated in numpy (FP64).

TEST INFRASTRUCTURE ONLY.  Imported by tests/ and nothing else; the product
never calls it.

Follows, for LJ + Coulomb with C2 switching:
  ReducedHessLennardJones::operator()
      (force/hessian/ReducedHessLennardJones.cpp:7-42)
  ReducedHessCoulomb::operator()
      (force/hessian/ReducedHessCoulomb.cpp:6-35)
  C2SwitchingFunction::hessian (switch/C2SwitchingFunction.cpp:43-52)
combined as Hessian::evaluatePairsMatrix does
(integrator/hessian/Hessian.cpp:581-688): the raw energy, raw force and
switch Hessian enter only inside the switching region (0 < S < 1); FULL
pairs give no block; MODIFIED pairs take A14/B14 and the coulomb scaling
factor.  The blocks are assembled as outputSparsePairMatrix
(Hessian.cpp:690-717) does: +H on (i,i) and (j,j), -H on (i,j) and (j,i).

Pinned by tests/test_oracle_hessian.py against central finite differences
of the (reference-pinned) C force oracle and against finite differences of
the C2 switch; the Ewald real-space term has no reference Hessian (the
adjoint's Ewald part is checked against finite differences instead).
"""
import numpy as np
import scipy.sparse as sp


def c2_parts(a, switchon, cutoff):
    """S, S~ = (dS/dr)/r, and the coefficients (cI, cR) of the switch
    Hessian cI I + cR d d^T (C2SwitchingFunction.h:42-53, .cpp:43-52)."""
    on2, cut2 = switchon * switchon, cutoff * cutoff
    D = cut2 - on2
    s1 = 1.0 / (D * D * D)
    s2 = cut2 - 3.0 * on2
    s3 = 4.0 / (D * D * D)
    S = np.ones_like(a)
    Sd = np.zeros_like(a)
    mid = (a >= on2) & (a <= cut2)
    c2 = cut2 - a
    c4 = c2 * (s2 + 2.0 * a)
    S = np.where(a > cut2, 0.0, np.where(mid, s1 * (c2 * c4), 1.0))
    Sd = np.where(mid, s3 * (c2 * c2 - c4), 0.0)
    tm2 = D * D * D
    cI = 12.0 * (a - cut2) * (a - on2) / tm2
    cR = 24.0 * (2.0 * a - cut2 - on2) / tm2
    return S, Sd, cI, cR


def pair_blocks(d, A, B, qq, lj_switch, el_switch, elec="coulomb"):
    """3x3 Hessian block of every pair: d (P,3) minimum-image x_j - x_i,
    A, B the (A14, B14 for MODIFIED) LJ parameters, qq the charge product
    (times the coulomb scaling factor for MODIFIED)."""
    a = np.einsum("pi,pi->p", d, d)
    rr = np.einsum("pi,pj->pij", d, d)
    eye = np.broadcast_to(np.eye(3), rr.shape)
    # ReducedHessLennardJones: (tm1 I + tm2 rr) S + mz rawE - rr 2 rawF S~
    S, Sd, cI, cR = c2_parts(a, *lj_switch)
    sw = (S > 0.0) & (S < 1.0)
    a4, a7 = a ** 4, a ** 7
    a5, a8 = a ** 5, a ** 8
    tm1 = 6.0 * B / a4 - 12.0 * A / a7
    tm2 = -48.0 * B / a5 + 168.0 * A / a8
    rawE = np.where(sw, A / a ** 6 - B / a ** 3, 0.0)
    rawF = np.where(sw, (12.0 * A / a ** 6 - 6.0 * B / a ** 3) / a, 0.0)
    cI = np.where(sw, cI, 0.0)
    cR = np.where(sw, cR, 0.0)
    H = (tm1[:, None, None] * eye + tm2[:, None, None] * rr) * S[:, None, None]
    H += (cI[:, None, None] * eye + cR[:, None, None] * rr) * rawE[:, None, None]
    H -= rr * (2.0 * rawF * Sd)[:, None, None]
    if elec == "coulomb":
        # ReducedHessCoulomb: (I - 3 rr / a) tm1 S + mz rawE - rr 2 rawF S~
        S, Sd, cI, cR = c2_parts(a, *el_switch)
        sw = (S > 0.0) & (S < 1.0)
        na = np.sqrt(a)
        t1 = -qq / a / na
        rawE = np.where(sw, qq / na, 0.0)
        rawF = np.where(sw, rawE / a, 0.0)
        cI = np.where(sw, cI, 0.0)
        cR = np.where(sw, cR, 0.0)
        H += (eye - rr * (3.0 / a)[:, None, None]) * (t1 * S)[:, None, None]
        H += (cI[:, None, None] * eye + cR[:, None, None] * rr) * rawE[:, None, None]
        H -= rr * (2.0 * rawF * Sd)[:, None, None]
    elif elec != "none":
        raise ValueError("the reference has pair Hessians for LJ and Coulomb")
    return H


def min_image(d, L):
    """Orthogonal-box minimum image (PeriodicBoundaryConditions.h:72-110
    up to ties at exactly half a box)."""
    return d - L * np.rint(d / L)


def hessian(workload, pairs, positions=None):
    """Sparse 3N x 3N Hessian of the pair energies over `pairs` (rows i, j,
    exclusion class) of a workload with C2 switching."""
    w = workload
    t = w.topology
    pos = w.positions if positions is None else positions
    L = t.boundary_conditions.lengths()
    i, j, cls = pairs[:, 0], pairs[:, 1], pairs[:, 2]
    keep = cls != 2
    i, j, cls = i[keep], j[keep], cls[keep]
    d = min_image(pos[j] - pos[i], L)
    lj = t.lennard_jones_parameters.table
    ti, tj = t.types[i], t.types[j]
    rec = lj[ti, tj]
    mod = cls == 1
    A = np.where(mod, rec[:, 2], rec[:, 0])
    B = np.where(mod, rec[:, 3], rec[:, 1])
    qq = t.scaled_charges[i] * t.scaled_charges[j]
    qq = np.where(mod, qq * t.coulomb_scaling_factor, qq)
    if w.elec_switch[0] == "C1":
        raise ValueError("C1 has no Hessian in this oracle")
    H = pair_blocks(d, A, B, qq, w.lj_switch, w.elec_switch, w.elec)
    n = t.n_atoms
    r3 = np.arange(3)
    rows, cols, vals = [], [], []
    for (p, q, sgn) in ((i, i, 1.0), (j, j, 1.0), (i, j, -1.0), (j, i, -1.0)):
        rows.append((3 * p[:, None, None] + r3[None, :, None]
                     + 0 * r3[None, None, :]).ravel())
        cols.append((3 * q[:, None, None] + 0 * r3[None, :, None]
                     + r3[None, None, :]).ravel())
        vals.append((sgn * H).ravel())
    M = sp.coo_matrix((np.concatenate(vals),
                       (np.concatenate(rows), np.concatenate(cols))),
                      shape=(3 * n, 3 * n))
    return M.tocsr()
