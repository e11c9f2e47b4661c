"""Smooth particle-mesh Ewald, restated in numpy from the reference's
NonbondedPMEwaldSystemForce (mdl/interface/force/nonbonded/
NonbondedPMEwaldSystemForce.h) with its Grid (Grid.h) and B-spline
(BSpline.cpp): the reciprocal term by B-spline charge spreading, a 3-D FFT,
the influence-function scalar sum with its atomic virial, and the force
interpolation; plus the correction, point-self and charged-system terms,
which are those of the direct Ewald sum.

TEST INFRASTRUCTURE ONLY: imported by tests/ and bench.py's cpu_baseline
leg, as the checker.  The product never imports it.

Conventions followed (reference lines in Grid.h unless said otherwise):
* grid coordinates x = position * n / L wrapped into [0, n) by adding or
  subtracting n (anterpolateCharge, :107-131); base point int(x) and the
  stencil of `order` points starting (order - 1) / 2 below it, the
  quotient truncated as C's integer division does (:509-511: 0 for order
  2, 1 for order 4);
* the B-spline weights and derivatives of BSpline::set (BSpline.cpp:58-95);
* unnormalised transforms (FFTW plans without scaling); the energy is half
  the sum over m != 0 of theta(m) |S(m)|^2 with theta(m) =
  B(m) exp(-pi^2 m^2 / alpha^2) / (pi V m^2) and m = (i0/Lx, j0/Ly, k0/Lz)
  (scalarSum, :237-341), B(m) = 1 / |sum_j M(j) e^{2 pi i m j / n}|^2
  (dftmod, :590-602);
* the virial slots XX..ZZ of 0.5 sum e (delta - v m m^T), v = 2 (1/m^2 +
  pi^2/alpha^2), summed as scalarSum sums them (:283-304): over the half
  spectrum k = 0 .. nz/2 with e doubled for 0 < k < nz/2 (and k = nz/2 when
  nz is odd), each term taking the signed frequencies of its own point.
  On the planes i = nx/2 and j = ny/2 of an even grid the doubled point's
  conjugate has the same i0 (or j0) but the opposite k0, so these
  off-diagonal terms differ from a sum over the full grid;
* the force on atom i, -q_i grad of the convolved grid times n / L per axis
  (interpolateForce, :171-199).
"""
import math

import numpy as np
from scipy.special import erf


def bspline(order, w):
    """BSpline::set(w) (BSpline.cpp:58-95) for an array of w: (theta,
    dtheta), each (len(w), order), in the reference's operation order."""
    w = np.asarray(w, dtype=np.float64)
    th = np.zeros(w.shape + (order,))
    dth = np.zeros(w.shape + (order,))
    if order == 1:
        th[..., 0] = 1.0
        return th, dth
    if order == 2:
        th[..., 1] = w
        th[..., 0] = 1.0 - w
        dth[..., 1] = 1.0
        dth[..., 0] = -1.0
        return th, dth
    th[..., order - 1] = 0.0
    th[..., 1] = w
    th[..., 0] = 1.0 - w
    for k in range(2, order - 1):
        div = 1.0 / k
        th[..., k] = div * w * th[..., k - 1]
        for j in range(1, k):
            th[..., k - j] = div * ((w + j) * th[..., k - j - 1] +
                                    (k + 1 - j - w) * th[..., k - j])
        th[..., 0] = div * (1 - w) * th[..., 0]
    dth[..., 0] = -th[..., 0]
    for j in range(1, order):
        dth[..., j] = th[..., j - 1] - th[..., j]
    div = 1.0 / (order - 1.0)
    th[..., order - 1] = div * w * th[..., order - 2]
    for j in range(1, order - 1):
        th[..., order - 1 - j] = div * ((w + j) * th[..., order - j - 2] +
                                        (order - j - w) * th[..., order - 1 - j])
    th[..., 0] = div * (1 - w) * th[..., 0]
    return th, dth


def dftmod(order, n):
    """Grid::dftmod (:590-602): B(m) = 1 / |b(m)|^2 per grid index."""
    th, _ = bspline(order, np.zeros(1))
    th = th[0]
    out = np.zeros(n)
    for i in range(n):
        sc = ss = 0.0
        for j in range(order):
            x = math.pi * 2.0 * i * j / n
            sc += th[j] * math.cos(x)
            ss += th[j] * math.sin(x)
        out[i] = 1.0 / (sc * sc + ss * ss)
    return out


def default_alpha(cutoff, accuracy):
    """The reference's default alpha (NonbondedPMEwaldSystemForce.h
    :547-562): the alpha with erfc(alpha rc) / rc = accuracy, by doubling
    and 100 bisection steps."""
    from scipy.special import erfc
    a = 1.0
    while erfc(a * cutoff) / cutoff >= accuracy:
        a *= 2.0
    lo, hi = 0.0, a
    for _ in range(100):
        a = 0.5 * (lo + hi)
        if erfc(a * cutoff) / cutoff >= accuracy:
            lo = a
        else:
            hi = a
    return a


def _grid_coord(pos, n, hr):
    """Scaled coordinate wrapped into [0, n) as anterpolateCharge does."""
    x = pos * hr
    x = np.where(np.isfinite(x), x, 0.0)
    # while (x < 0) x += n; while (x >= n) x -= n
    for _ in range(1000):
        lo = x < 0.0
        if not lo.any():
            break
        x = np.where(lo, x + n, x)
    for _ in range(1000):
        hi = x >= n
        if not hi.any():
            break
        x = np.where(hi, x - n, x)
    return x


def pme_reciprocal(pos, q, box, grid, order, alpha, virial=True):
    """Reciprocal term of NonbondedPMEwaldSystemForce (reciprocalTerm,
    NonbondedPMEwaldSystemForce.h:815-862): (forces (N, 3), energy,
    virial[9]) for charges q (scaled) in the orthogonal box (Lx, Ly, Lz) on
    a grid (nx, ny, nz) with B-splines of `order`."""
    pos = np.asarray(pos, dtype=np.float64).reshape(-1, 3)
    q = np.asarray(q, dtype=np.float64)
    L = np.asarray(box, dtype=np.float64)
    n = np.asarray(grid, dtype=np.int64)
    hr = n / L
    offset = -((order - 1) // 2)   # C: -(order - 1) / 2, truncated
    idx, th, dth = [], [], []
    for a in range(3):
        x = _grid_coord(pos[:, a], n[a], hr[a])
        ix = x.astype(np.int64)
        t, d = bspline(order, x - ix)
        base = ix + offset + n[a]
        idx.append((base[:, None] + np.arange(order)[None, :]) % n[a])
        th.append(t)
        dth.append(d)
    Q = np.zeros(tuple(n))
    # charge spreading: q thetaX thetaY thetaZ onto the stencil
    w = (q[:, None, None, None] * th[0][:, :, None, None] *
         th[1][:, None, :, None] * th[2][:, None, None, :])
    I = np.broadcast_to(idx[0][:, :, None, None], w.shape)
    J = np.broadcast_to(idx[1][:, None, :, None], w.shape)
    K = np.broadcast_to(idx[2][:, None, None, :], w.shape)
    np.add.at(Q, (I.ravel(), J.ravel(), K.ravel()), w.ravel())
    S = np.fft.fftn(Q)
    # influence function on the full grid
    m = [np.where(np.arange(n[a]) <= n[a] // 2, np.arange(n[a]),
                  np.arange(n[a]) - n[a]) / L[a] for a in range(3)]
    fac = math.pi * math.pi / (alpha * alpha)
    V = L[0] * L[1] * L[2]
    mods = [dftmod(order, int(n[a])) for a in range(3)]
    MX, MY, MZ = np.meshgrid(m[0], m[1], m[2], indexing="ij")
    m2 = MX * MX + MY * MY + MZ * MZ
    m2[0, 0, 0] = 1.0
    B = (mods[0][:, None, None] * mods[1][None, :, None] *
         mods[2][None, None, :])
    theta = B * np.exp(-fac * m2) / (math.pi * V * m2)
    theta[0, 0, 0] = 0.0
    e_m = theta * (S.real ** 2 + S.imag ** 2)
    # scalarSum's half spectrum: k = 0 .. nz/2, e doubled where the
    # conjugate point lies outside it
    kh = int(n[2]) // 2 + 1
    k_idx = np.arange(kh)
    dbl = np.where((k_idx > 0) & ((k_idx != int(n[2]) // 2) | (int(n[2]) & 1)),
                   2.0, 1.0)
    e_h = e_m[:, :, :kh] * dbl[None, None, :]
    energy = 0.5 * e_h.sum()
    vir = np.zeros(9)
    if virial:
        v = 2.0 * (1.0 / m2[:, :, :kh] + fac)
        comps = [MX[:, :, :kh], MY[:, :, :kh], MZ[:, :, :kh]]
        for a in range(3):
            for b in range(3):
                d = 1.0 if a == b else 0.0
                vir[3 * a + b] = 0.5 * np.sum(e_h * (d - v * comps[a] *
                                                     comps[b]))
    conv = np.fft.ifftn(theta * S).real * Q.size
    # force interpolation
    G = conv[I, J, K]
    fx = -np.sum(G * dth[0][:, :, None, None] * th[1][:, None, :, None] *
                 th[2][:, None, None, :], axis=(1, 2, 3))
    fy = -np.sum(G * th[0][:, :, None, None] * dth[1][:, None, :, None] *
                 th[2][:, None, None, :], axis=(1, 2, 3))
    fz = -np.sum(G * th[0][:, :, None, None] * th[1][:, None, :, None] *
                 dth[2][:, None, None, :], axis=(1, 2, 3))
    f = np.stack([fx * hr[0] * q, fy * hr[1] * q, fz * hr[2] * q], axis=1)
    return f, energy, vir


def correction_terms(pos, q, box, rows, csf, alpha, virial=True):
    """Correction (over ExclusionTable::getTable() rows), point-self and
    charged-system terms, as the direct Ewald sum has them
    (NonbondedPMEwaldSystemForce.h:878-912, :593-612): (forces, energy,
    virial[9])."""
    pos = np.asarray(pos, dtype=np.float64).reshape(-1, 3)
    L = np.asarray(box, dtype=np.float64)
    f = np.zeros_like(pos)
    energy = 0.0
    vir = np.zeros((3, 3))
    two_a_pi = 2.0 * alpha / math.sqrt(math.pi)
    for a1, a2, cls in np.asarray(rows).reshape(-1, 3).tolist():
        d = pos[a2] - pos[a1]
        d = d - L * np.rint(d / L)
        r2 = float(d @ d)
        qq = q[a1] * q[a2]
        if cls == 1:
            qq *= 1 - csf
        r = math.sqrt(r2)
        e = erf(alpha * r) / r
        energy -= qq * e
        g = qq * (two_a_pi * math.exp(-alpha * alpha * r2) - e) / r2
        fij = d * g
        f[a1] -= fij
        f[a2] += fij
        vir += np.outer(fij, d)
    energy += -np.sum(q * q) * alpha / math.sqrt(math.pi)
    qs = float(np.sum(q))
    charged = 0.0
    if abs(qs * 0.00268283) > 1.0e-5:
        charged = -math.pi / (2.0 * L.prod() * alpha * alpha) * qs * qs
    energy += charged
    if virial:
        vir += np.eye(3) * charged
    return f, energy, vir.reshape(9)


def pme_kspace(w, grid, order, alpha, virial=True):
    """Reciprocal + correction + self + charged terms for a workload."""
    t = w.topology
    L = t.boundary_conditions.lengths()
    q = t.scaled_charges
    f1, e1, v1 = pme_reciprocal(w.positions, q, L, grid, order, alpha, virial)
    f2, e2, v2 = correction_terms(w.positions, q, L, t.exclusions.get_table(),
                                  t.coulomb_scaling_factor, alpha, virial)
    return f1 + f2, e1 + e2, v1 + v2
