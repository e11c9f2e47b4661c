// pmnb_pme.h -- smooth particle-mesh Ewald: the reciprocal term of
// ProtoMol's NonbondedPMEwaldSystemForce (mdl/interface/force/nonbonded/
// NonbondedPMEwaldSystemForce.h with Grid.h and BSpline.cpp) for an
// orthogonal periodic box.  Its correction, point-self and charged-system
// terms are the direct Ewald sum's (pmnb_ewald.h), computed by the same
// kernels.
//
// One evaluation:
//   k_pme_spread     one thread per atom: grid coordinates x = r n / L
//                    wrapped into [0, n) as Grid::anterpolateCharge does
//                    (Grid.h:107-131), the B-spline weights of
//                    BSpline::set (BSpline.cpp:58-95), and q Mx My Mz onto
//                    the order^3 stencil as 64-bit fixed-point integers
//                    (2^-40) with atomics: integer sums do not depend on
//                    the order of the adds, so the grid is the same bits
//                    on every run;
//   k_pme_to_real    the integer grid to FP64, and cleared for the next run;
//   hipFFT D2Z       the unnormalised transform of the real grid (nx, ny,
//                    nz/2 + 1 half spectrum; the reference transforms the
//                    full complex grid, whose other half is the conjugate);
//   k_pme_scalar     one thread per spectrum point of the half spectrum,
//                    the loop of Grid::scalarSum (Grid.h:237-341): theta(m)
//                    = B(m) exp(-pi^2 m^2 / alpha^2) / (pi V m^2), energy
//                    and virial terms (doubled where the reference doubles
//                    them), the spectrum multiplied by theta; per-block sums
//                    in a fixed tree;
//   hipFFT Z2D       back to the real, convolved grid;
//   k_pme_gather     one thread per atom: the force -q grad (the convolved
//                    grid) n / L, interpolateForce's loop (Grid.h:171-199),
//                    plus the atom's correction pairs; no atomics;
//   k_pme_finish     the blocks' energy and virial sums in block order,
//                    half of them into the output (the reference returns
//                    energy * 0.5 and adds virial * 0.5).
#pragma once

#include "pmnb_ewald.h"

namespace pmnb {

constexpr int kPmeBlock = 256;
constexpr int kPmeSums = 7;  // energy, xx, xy, xz, yy, yz, zz
// Charge grid resolution 2^-40: with order 8 on a 32^3 grid the rounding of
// 2^-32 moves forces by 1e-9 of the largest, 2^-40 by 3e-12.  A grid value
// must stay below 2^23 (8.4e6, about 4.6e5 elementary charges on one grid
// point) in magnitude; each contribution is clamped to 2^22.
constexpr double kPmeFixScale = 1099511627776.0;       // 2^40
constexpr double kPmeFixInv = 1.0 / 1099511627776.0;
constexpr double kPmeFixClamp = 4194304.0;             // 2^22

struct PmeDev {
  int n;                   // atoms
  int nx, ny, nz, nzh;     // grid, nz / 2 + 1
  double hr[3];            // n / L per axis
  int off[3];              // -(order - 1) / 2 + n per axis (Grid.h:509-511)
  double Lr[3];            // 1 / L per axis
  double piVr;             // 1 / (pi V)
  double fac;              // pi^2 / alpha^2
  const double *q;         // scaled charges
  long long *qfix;         // nx ny nz fixed-point grid (zero between runs)
  double *qreal;           // nx ny nz real grid
  double2 *qk;             // nx ny nzh half spectrum
  const double *mod[3];    // B(m) per axis index (Grid::dftmod)
  const double *ex[3];     // exp(-fac m^2) per axis index (ex[2]: k <= nz/2)
  double *part;            // kPmeSums per scalar-sum block
  int n_part;              // scalar-sum blocks
};

// BSpline::set(w) (BSpline.cpp:58-95) with compile-time order P >= 2, in
// the reference's operation order.
template <int P>
__device__ __forceinline__ void pme_bspline(double w, double (&th)[P],
                                            double (&dth)[P]) {
  if constexpr (P == 2) {
    th[1] = w;
    th[0] = 1.0 - w;
    dth[1] = 1.0;
    dth[0] = -1.0;
  } else {
    th[P - 1] = 0.0;
    th[1] = w;
    th[0] = 1.0 - w;
#pragma unroll
    for (int k = 2; k < P - 1; k++) {
      const double div = 1.0 / k;
      th[k] = div * w * th[k - 1];
#pragma unroll
      for (int j = 1; j < k; j++)
        th[k - j] = div * ((w + j) * th[k - j - 1] + (k + 1 - j - w) * th[k - j]);
      th[0] = div * (1 - w) * th[0];
    }
    dth[0] = -th[0];
#pragma unroll
    for (int j = 1; j <= P - 1; j++) dth[j] = th[j - 1] - th[j];
    const double div = 1.0 / (P - 1.0);
    th[P - 1] = div * w * th[P - 2];
#pragma unroll
    for (int j = 1; j < P - 1; j++)
      th[P - 1 - j] = div * ((w + j) * th[P - j - 2] + (P - j - w) * th[P - 1 - j]);
    th[0] = div * (1 - w) * th[0];
  }
}

// Grid coordinate of position p on an axis of n points: p n / L, brought
// into [0, n) by adding or subtracting n as the reference's while loops do
// (Grid.h:111-116); positions more than 64 boxes away and non-finite ones
// are folded in one step, so every loop ends.
__device__ __forceinline__ double pme_coord(double p, double hr, int n) {
  double x = p * hr;
  if (!(x == x) || x > 1e15 || x < -1e15) return 0.0;
  const double dn = (double)n;
  for (int it = 0; it < 64 && x < 0.0; it++) x += dn;
  for (int it = 0; it < 64 && x >= dn; it++) x -= dn;
  if (x < 0.0 || x >= dn) {
    x -= dn * floor(x / dn);
    if (!(x >= 0.0 && x < dn)) x = 0.0;
  }
  return x;
}

template <int P>
__global__ __launch_bounds__(kPmeBlock) void k_pme_spread(PmeDev d,
                                                          const double *x) {
  const int i = blockIdx.x * kPmeBlock + threadIdx.x;
  if (i >= d.n) return;
  const double q = d.q[i];
  const int nn[3] = {d.nx, d.ny, d.nz};
  int b[3];
  double th[3][P], dth[3][P];
#pragma unroll
  for (int a = 0; a < 3; a++) {
    const double g = pme_coord(x[3 * i + a], d.hr[a], nn[a]);
    const int ig = (int)g;
    pme_bspline<P>(g - ig, th[a], dth[a]);
    b[a] = ig + d.off[a];
  }
#pragma unroll
  for (int u = 0; u < P; u++) {
    const double av = q * th[0][u];
    const int i1 = (u + b[0]) % d.nx;
#pragma unroll
    for (int v = 0; v < P; v++) {
      const double ab = av * th[1][v];
      const int j1 = (v + b[1]) % d.ny;
      long long *row = d.qfix + ((size_t)i1 * d.ny + j1) * d.nz;
#pragma unroll
      for (int w = 0; w < P; w++) {
        const int k1 = (w + b[2]) % d.nz;
        double val = ab * th[2][w];
        val = fmin(fmax(val, -kPmeFixClamp), kPmeFixClamp);
        atomicAdd(reinterpret_cast<unsigned long long *>(row + k1),
                  (unsigned long long)(long long)__builtin_rint(val * kPmeFixScale));
      }
    }
  }
}

__global__ __launch_bounds__(kPmeBlock) void k_pme_to_real(PmeDev d,
                                                           long long total) {
  const long long g = (long long)blockIdx.x * kPmeBlock + threadIdx.x;
  if (g >= total) return;
  d.qreal[g] = (double)d.qfix[g] * kPmeFixInv;
  d.qfix[g] = 0;
}

__device__ __forceinline__ double pme_block_sum(double v, double *red) {
  const int t = threadIdx.x;
  red[t] = v;
  __syncthreads();
  for (int o = kPmeBlock / 2; o > 0; o >>= 1) {
    if (t < o) red[t] += red[t + o];
    __syncthreads();
  }
  const double r = red[0];
  __syncthreads();
  return r;
}

__global__ __launch_bounds__(kPmeBlock) void k_pme_scalar(PmeDev d,
                                                          int virial) {
  __shared__ double red[kPmeBlock];
  const long long idx = (long long)blockIdx.x * kPmeBlock + threadIdx.x;
  const long long total = (long long)d.nx * d.ny * d.nzh;
  double acc[kPmeSums];
#pragma unroll
  for (int c = 0; c < kPmeSums; c++) acc[c] = 0.0;
  if (idx < total) {
    const int k = (int)(idx % d.nzh);
    const int j = (int)((idx / d.nzh) % d.ny);
    const int i = (int)(idx / ((long long)d.nzh * d.ny));
    double2 s = d.qk[idx];
    if (i == 0 && j == 0 && k == 0) {
      s = make_double2(0.0, 0.0);
    } else {
      const int i0 = i <= d.nx / 2 ? i : i - d.nx;
      const int j0 = j <= d.ny / 2 ? j : j - d.ny;
      const int k0 = k <= d.nz / 2 ? k : k - d.nz;
      const double mi = i0 * d.Lr[0], mj = j0 * d.Lr[1], mk = k0 * d.Lr[2];
      const double exv = d.ex[0][i] * d.piVr;
      const double modxy = d.mod[0][i] * d.mod[1][j];
      const double mij = mi * mi + mj * mj;
      const double exy = exv * d.ex[1][j];
      const double m2 = mij + mk * mk;
      const double theta = modxy * d.mod[2][k] * exy * d.ex[2][k] / m2;
      const double qq = s.x * s.x + s.y * s.y;
      double e = qq * theta;
      const double v = 2.0 * (1.0 / m2 + d.fac);
      if (k > 0 && ((k != d.nz / 2) || (d.nz & 1))) e *= 2.0;
      acc[0] = e;
      if (virial) {
        acc[1] = e * (1.0 - v * mi * mi);
        acc[2] = -(e * (v * mi * mj));
        acc[3] = -(e * (v * mi * mk));
        acc[4] = e * (1.0 - v * mj * mj);
        acc[5] = -(e * (v * mj * mk));
        acc[6] = e * (1.0 - v * mk * mk);
      }
      s.x *= theta;
      s.y *= theta;
    }
    d.qk[idx] = s;
  }
  double *out = d.part + (size_t)blockIdx.x * kPmeSums;
#pragma unroll
  for (int c = 0; c < kPmeSums; c++) {
    const double t = pme_block_sum(acc[c], red);
    if (threadIdx.x == 0) out[c] = t;
  }
}

// Force of the reciprocal term on each atom and the correction forces of
// its excluded pairs (ew_correction_force), added to f in a fixed order.
template <int P>
__global__ __launch_bounds__(kPmeBlock) void k_pme_gather(PmeDev d, EwDev e,
                                                          const double *x,
                                                          double *f) {
  const int i = blockIdx.x * kPmeBlock + threadIdx.x;
  if (i >= d.n) return;
  const int nn[3] = {d.nx, d.ny, d.nz};
  int b[3];
  double th[3][P], dth[3][P];
#pragma unroll
  for (int a = 0; a < 3; a++) {
    const double g = pme_coord(x[3 * i + a], d.hr[a], nn[a]);
    const int ig = (int)g;
    pme_bspline<P>(g - ig, th[a], dth[a]);
    b[a] = ig + d.off[a];
  }
  double fx = 0.0, fy = 0.0, fz = 0.0;
#pragma unroll
  for (int u = 0; u < P; u++) {
    const int i1 = (u + b[0]) % d.nx;
#pragma unroll
    for (int v = 0; v < P; v++) {
      const int j1 = (v + b[1]) % d.ny;
      const double xij = dth[0][u] * th[1][v];
      const double yij = th[0][u] * dth[1][v];
      const double zij = th[0][u] * th[1][v];
      const double *row = d.qreal + ((size_t)i1 * d.ny + j1) * d.nz;
#pragma unroll
      for (int w = 0; w < P; w++) {
        const double term = row[(w + b[2]) % d.nz];
        fx -= term * xij * th[2][w];
        fy -= term * yij * th[2][w];
        fz -= term * zij * dth[2][w];
      }
    }
  }
  const double q = d.q[i];
  double gx = fx * d.hr[0] * q, gy = fy * d.hr[1] * q, gz = fz * d.hr[2] * q;
  ew_correction_force(e, x, i, gx, gy, gz);
  f[3 * i] += gx;
  f[3 * i + 1] += gy;
  f[3 * i + 2] += gz;
}

// One wave: lane c < kPmeSums adds sum c of every block in block order;
// out[1] += energy / 2, and with the virial flag the virial slots XX..ZZ
// += the virial sums / 2 (Grid.h:316-326).
__global__ __launch_bounds__(64) void k_pme_finish(PmeDev d, double *out,
                                                   int virial) {
  __shared__ double tot[kPmeSums];
  const int t = threadIdx.x;
  if (t < kPmeSums) {
    double v = 0.0;
    for (int b = 0; b < d.n_part; b++) v += d.part[(size_t)b * kPmeSums + t];
    tot[t] = v;
  }
  __syncthreads();
  if (t == 0) {
    out[1] += tot[0] * 0.5;
    if (virial) {
      out[2] += tot[1] * 0.5; out[3] += tot[2] * 0.5; out[4] += tot[3] * 0.5;
      out[5] += tot[2] * 0.5; out[6] += tot[4] * 0.5; out[7] += tot[5] * 0.5;
      out[8] += tot[3] * 0.5; out[9] += tot[5] * 0.5; out[10] += tot[6] * 0.5;
    }
  }
}

}  // namespace pmnb
