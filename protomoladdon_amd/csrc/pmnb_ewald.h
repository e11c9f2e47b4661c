// pmnb_ewald.h -- reciprocal-space and correction terms of ProtoMol's
// NonbondedFullEwaldSystemForce for an orthogonal periodic box.
//
// Reference (force/nonbonded/NonbondedFullEwaldSystemForce.h):
//   :597-652   alpha, rc, kc and the k-vector list (h >= 0 half space),
//              built on the host in the same order
//   :680-700   point self energy and charged-system energy (host)
//   :884-1086  reciprocal term: per-atom sin/cos tables by the addition
//              theorem, structure factor, energy, atomic virial, forces
//   :1092-1130 intramolecular correction over the exclusion table
//
// The k-vector list is a sequence of runs: for each (h, k) the l values
// form one contiguous interval (the reference's loop order).
// Kernels: k_ew_tables (one thread per atom: the sin/cos tables of the
// three axes, charge folded into x as the reference does, stored
// [axis][order][sin|cos][atom] so a wave reads consecutive atoms);
// k_ew_sk (one block per chunk of at most kEwChunk values of |l| of a run,
// both signs: the structure factors, threads over atoms, the x-y product
// formed once per atom and chunk and exp(i |l| z) stepped from one table
// value, serving +l and -l; fixed-order reductions); k_ew_force (two atoms
// per lane, the block's waves on different runs: per run the x-y product
// once, then exp(i l z) for l = 0, 1, ... by the same addition-theorem step
// that built the tables, so the values are the tables' without reading
// them, and the terms of +l and -l, with per-k records (k-vector and force
// prefactors) written by k_ew_sk; then the
// correction of the atom's excluded pairs, each pair evaluated by both of
// its atoms so no atomics are needed); k_ew_partials and k_ew_finish
// (energy and virial sums: fixed slices over 128 blocks, then one wave adds
// the blocks' sums in block order).  Work is N x K (atoms x k-vectors), as
// in the reference.
#pragma once

namespace pmnb {

constexpr int kEwBlock = 256;
constexpr int kEwChunk = 8;   // |l| values per structure-factor block
// waves sharing an atom in k_ew_force: 4, or 16 when the system gives
// fewer blocks than the chip has CUs
constexpr int kEwLanesLarge = 4, kEwLanesSmall = 16;
constexpr int kEwAtomsLarge = 2;  // atoms per lane of the large launch

struct EwRun {
  int h, k, l_lo, l_hi, start;  // l in [l_lo, l_hi] at list index start + l - l_lo
};
struct EwChunk {
  int run, m0, m1;              // l = +-m, m in [m0, m1], of run
};
// What the force needs of one k-vector: the force on atom i is
// k * (sA_i ps - cA_i pc), ps = a0 b S_cos, pc = a0 b S_sin, a0 = 8 pi / V
struct EwKF {
  double kx, ky, kz, ps, pc;
};

struct EwDev {
  Box box;
  int n, nk, H;               // atoms, k-vectors, table orders
  const double4 *kv;          // kx, ky, kz, b = exp(-k^2/(4 alpha^2))/k^2
  const EwRun *runs;
  const EwChunk *chunks;
  int n_runs, n_chunks;
  const double *q;            // scaled charges
  double *tab;                // 3 * H * 2 * n
  double2 *sk;                // per k: (sum sin, sum cos)
  EwKF *kf;                   // per k: k-vector and force prefactors
  double *part;               // k_ew_partials' sums, 14 per block
  const int *xptr, *xent;     // per atom: exclusion rows (row << 1 | role)
  const int *xrows;           // 3 per row: a1, a2, class
  int n_excl;
  double alpha, alpha2, two_alpha_pi, csf, Vr, fac;
};

__device__ __forceinline__ double &ew_tab(const EwDev &d, int axis, int m,
                                          int sc, int i) {
  return d.tab[((size_t)(axis * d.H + m) * 2 + sc) * d.n + i];
}

// (s, c) <- (s, c) * exp(i z) for (s1, c1) = (sin z, cos z): the
// addition-theorem step of the tables, with fused multiply-adds.  k_ew_sk and
// k_ew_force take the same step, so their values of exp(i m z) are the
// tables' bit for bit; one step from (0, 1) gives (s1, c1) exactly
// (fma(s1, 1, c1 * 0) = s1), so m = 1 needs no case of its own.
__device__ __forceinline__ void ew_step(double s1, double c1, double &s,
                                        double &c) {
  const double s2 = fma(s1, c, c1 * s);
  const double c2 = fma(c1, c, -(s1 * s));
  s = s2;
  c = c2;
}

// Intramolecular correction force on atom i (:1105-1129): every excluded
// pair of i, in row order, each pair evaluated by both of its atoms (no
// atomics).
__device__ __forceinline__ void ew_correction_force(const EwDev &d,
                                                    const double *x, int i,
                                                    double &gx, double &gy,
                                                    double &gz) {
  for (int e = d.xptr[i]; e < d.xptr[i + 1]; e++) {
    const int row = d.xent[e] >> 1, role = d.xent[e] & 1;
    const int a1 = d.xrows[3 * row], a2 = d.xrows[3 * row + 1],
              cls = d.xrows[3 * row + 2];
    double rx, ry, rz, r2;
    min_diff_exact(d.box, x[3 * a1], x[3 * a1 + 1], x[3 * a1 + 2], x[3 * a2],
                   x[3 * a2 + 1], x[3 * a2 + 2], rx, ry, rz, r2);
    double qq = d.q[a1] * d.q[a2];
    if (cls == 1) qq *= 1 - d.csf;  // EXCLUSION_MODIFIED
    const double r = sqrt(r2), rr = 1 / r;
    const double e1 = erf(d.alpha * r) * rr;
    const double g = qq * (d.two_alpha_pi * exp(-d.alpha2 * r2) - e1) * rr * rr;
    if (role == 0) {
      gx -= rx * g;
      gy -= ry * g;
      gz -= rz * g;
    } else {
      gx += rx * g;
      gy += ry * g;
      gz += rz * g;
    }
  }
}

__global__ __launch_bounds__(kEwBlock) void k_ew_tables(EwDev d,
                                                        const double *x) {
  const int i = blockIdx.x * kEwBlock + threadIdx.x;
  if (i >= d.n) return;
  double r[3];
  min_pos_exact(d.box, x[3 * i], x[3 * i + 1], x[3 * i + 2], r[0], r[1], r[2]);
  const double qi = d.q[i];
  for (int a = 0; a < 3; a++) {
    const double xa = r[a] * 2.0 * M_PI * d.box.er[a];
    const double sn = sin(xa), cs = cos(xa);
    double s0 = 0.0, c0 = a == 0 ? qi * 1.0 : 1.0;
    double s1 = a == 0 ? qi * sn : sn, c1 = a == 0 ? qi * cs : cs;
    ew_tab(d, a, 0, 0, i) = s0;
    ew_tab(d, a, 0, 1, i) = c0;
    ew_tab(d, a, 1, 0, i) = s1;
    ew_tab(d, a, 1, 1, i) = c1;
    for (int m = 2; m < d.H; m++) {
      ew_step(sn, cs, s1, c1);
      ew_tab(d, a, m, 0, i) = s1;
      ew_tab(d, a, m, 1, i) = c1;
    }
  }
}

__device__ __forceinline__ double ew_wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// A chunk is an interval [m0, m1] of |l| (at most kEwChunk values) of one
// run, taken with both signs: slot jj of the accumulators holds l = +m and
// l = -m for m = m0 + jj, which share exp(i m z) (stepped from the table
// value at m0 by the table's own step) and the x-y product; sin(-m z) =
// -sin(m z).  A sign whose l is not in the run is computed and not stored.
__global__ __launch_bounds__(kEwBlock) void k_ew_sk(EwDev d) {
  __shared__ double red[kEwBlock / 64][4 * kEwChunk];
  const EwChunk c = d.chunks[blockIdx.x];
  const EwRun r = d.runs[c.run];
  const int ka = r.k < 0 ? -r.k : r.k;
  const int nm = c.m1 - c.m0 + 1;
  const int m0 = c.m0;
  double sp[kEwChunk], cp[kEwChunk], sn[kEwChunk], cn[kEwChunk];
#pragma unroll
  for (int j = 0; j < kEwChunk; j++) sp[j] = cp[j] = sn[j] = cn[j] = 0.0;
  for (int i = threadIdx.x; i < d.n; i += kEwBlock) {
    const double xs = ew_tab(d, 0, r.h, 0, i), xc = ew_tab(d, 0, r.h, 1, i);
    const double ys = (r.k < 0 ? -1 : 1) * ew_tab(d, 1, ka, 0, i);
    const double yc = ew_tab(d, 1, ka, 1, i);
    const double xys = fma(xs, yc, xc * ys), xyc = fma(xc, yc, -(xs * ys));
    const double zs1 = ew_tab(d, 2, 1, 0, i), zc1 = ew_tab(d, 2, 1, 1, i);
    double zs = ew_tab(d, 2, m0, 0, i), zc = ew_tab(d, 2, m0, 1, i);
#pragma unroll
    for (int jj = 0; jj < kEwChunk; jj++) {
      if (jj < nm) {
        if (jj > 0) ew_step(zs1, zc1, zs, zc);
        // sin/cos of (x + y +- m z)
        sp[jj] = fma(xyc, zs, fma(xys, zc, sp[jj]));
        cp[jj] = fma(-xys, zs, fma(xyc, zc, cp[jj]));
        sn[jj] = fma(-xyc, zs, fma(xys, zc, sn[jj]));
        cn[jj] = fma(xys, zs, fma(xyc, zc, cn[jj]));
      }
    }
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int j = 0; j < kEwChunk; j++) {
    const double a = ew_wave_sum(sp[j]), b = ew_wave_sum(cp[j]);
    const double e = ew_wave_sum(sn[j]), g = ew_wave_sum(cn[j]);
    if (lane == 0) {
      red[wave][4 * j] = a;
      red[wave][4 * j + 1] = b;
      red[wave][4 * j + 2] = e;
      red[wave][4 * j + 3] = g;
    }
  }
  __syncthreads();
  if ((int)threadIdx.x < 2 * nm) {
    const int jj = threadIdx.x >> 1, neg = threadIdx.x & 1;
    const int l = neg ? -(m0 + jj) : m0 + jj;
    if ((neg && l == 0) || l < r.l_lo || l > r.l_hi) return;
    double a = 0, b = 0;
#pragma unroll
    for (int w = 0; w < kEwBlock / 64; w++) {
      a += red[w][4 * jj + 2 * neg];
      b += red[w][4 * jj + 2 * neg + 1];
    }
    const int kk = r.start + (l - r.l_lo);
    d.sk[kk] = make_double2(a, b);
    const double4 k = d.kv[kk];
    const double ak = 8.0 * M_PI * d.Vr * k.w;
    d.kf[kk] = EwKF{k.x, k.y, k.z, ak * b, ak * a};
  }
}

// Force on atom i: every k-vector, then the correction of its excluded
// pairs (the pair force of row (a1, a2) acts as -f on a1 and +f on a2).
// A block holds kEwAtoms x 64 atoms (each lane kEwAtoms, which share the
// k-vector loads); its wave w takes runs w, w + kEwLanes, ... (every lane of
// a wave on the same run, so those loads are wave-uniform), and the kEwLanes
// partial forces of an atom are added in wave order through LDS, by wave u
// for the atoms of slot u (kEwLanes >= kEwAtoms).  Four waves per block
// and two atoms per lane on a 96k-atom system (750 blocks) measured faster
// than five or six waves, and than one atom per lane (more waves, but one
// dependent exp(i m z) chain per lane) or three; a small system, with fewer
// blocks than CUs, takes sixteen waves.
template <int kEwLanes, int kEwAtoms>
__global__ __launch_bounds__(64 * kEwLanes) void k_ew_force(EwDev d,
                                                            const double *x,
                                                            double *f) {
  __shared__ double part_f[kEwLanes][kEwAtoms][3][64];
  // part is the wave's index: uniform, so the run and k-vector data below
  // come through scalar loads
  const int lane = threadIdx.x & 63,
            part = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  int at[kEwAtoms];
  double fx[kEwAtoms], fy[kEwAtoms], fz[kEwAtoms];
  double zs1[kEwAtoms], zc1[kEwAtoms];
#pragma unroll
  for (int u = 0; u < kEwAtoms; u++) {
    const int i0 = blockIdx.x * (64 * kEwAtoms) + u * 64 + lane;
    at[u] = i0 < d.n ? i0 : d.n - 1;  // past the end: a real atom, not written
    fx[u] = fy[u] = fz[u] = 0.0;
    zs1[u] = ew_tab(d, 2, 1, 0, at[u]);
    zc1[u] = ew_tab(d, 2, 1, 1, at[u]);
  }
  for (int ri = part; ri < d.n_runs; ri += kEwLanes) {
    const EwRun r = d.runs[ri];
    const int ka = r.k < 0 ? -r.k : r.k;
    double xys[kEwAtoms], xyc[kEwAtoms], zs[kEwAtoms], zc[kEwAtoms];
#pragma unroll
    for (int u = 0; u < kEwAtoms; u++) {
      const int i = at[u];
      const double xs = ew_tab(d, 0, r.h, 0, i), xc = ew_tab(d, 0, r.h, 1, i);
      const double ys = (r.k < 0 ? -1 : 1) * ew_tab(d, 1, ka, 0, i);
      const double yc = ew_tab(d, 1, ka, 1, i);
      xys[u] = fma(xs, yc, xc * ys);
      xyc[u] = fma(xc, yc, -(xs * ys));
      zs[u] = 0.0;  // exp(i m z) for m = 0
      zc[u] = 1.0;
    }
    const EwKF *kr = d.kf + (r.start - r.l_lo);  // kr[l], l in [l_lo, l_hi]
    const int mmax = max(-r.l_lo, r.l_hi);
    for (int m = 0; m <= mmax; m++) {
      if (m > 0) {  // the step k_ew_tables takes
#pragma unroll
        for (int u = 0; u < kEwAtoms; u++) ew_step(zs1[u], zc1[u], zs[u], zc[u]);
      }
#pragma unroll
      for (int sg = 0; sg < 2; sg++) {
        const int l = sg == 0 ? m : -m;
        if ((sg == 1 && m == 0) || l < r.l_lo || l > r.l_hi) continue;
        const EwKF k = kr[l];
#pragma unroll
        for (int u = 0; u < kEwAtoms; u++) {
          const double zsl = sg == 0 ? zs[u] : -zs[u];  // sin(l z), l = +-m
          const double sA = fma(xyc[u], zsl, xys[u] * zc[u]);
          const double cA = fma(-xys[u], zsl, xyc[u] * zc[u]);
          const double w = fma(sA, k.ps, -(cA * k.pc));
          fx[u] = fma(k.kx, w, fx[u]);
          fy[u] = fma(k.ky, w, fy[u]);
          fz[u] = fma(k.kz, w, fz[u]);
        }
      }
    }
  }
#pragma unroll
  for (int u = 0; u < kEwAtoms; u++) {
    part_f[part][u][0][lane] = fx[u];
    part_f[part][u][1][lane] = fy[u];
    part_f[part][u][2][lane] = fz[u];
  }
  __syncthreads();
  if (part >= kEwAtoms) return;
  const int i0 = blockIdx.x * (64 * kEwAtoms) + part * 64 + lane;
  if (i0 >= d.n) return;
  const int i = i0;
  double gx = 0, gy = 0, gz = 0;
#pragma unroll
  for (int w = 0; w < kEwLanes; w++) {
    gx += part_f[w][part][0][lane];
    gy += part_f[w][part][1][lane];
    gz += part_f[w][part][2][lane];
  }
  ew_correction_force(d, x, i, gx, gy, gz);
  f[3 * i] += gx;
  f[3 * i + 1] += gy;
  f[3 * i + 2] += gz;
}

// Energy and virial: k_ew_partials spreads the k-vectors and the excluded
// pairs over kEwFinBlocks blocks (contiguous slices), each writing its 14
// sums (reciprocal energy, six reciprocal virial sums, correction energy,
// six correction virial sums) after a fixed-order tree; k_ew_finish adds
// the blocks' sums in block order and scales.  Both orders are fixed, so
// results repeat bit for bit.
constexpr int kEwFinBlocks = 128, kEwFinSums = 14;

__device__ __forceinline__ double ew_block_sum(double v, double *red) {
  const int t = threadIdx.x;
  red[t] = v;
  __syncthreads();
  for (int o = kEwBlock / 2; o > 0; o >>= 1) {
    if (t < o) red[t] += red[t + o];
    __syncthreads();
  }
  const double r = red[0];
  __syncthreads();
  return r;
}

__global__ __launch_bounds__(kEwBlock) void k_ew_partials(EwDev d,
                                                          const double *x) {
  __shared__ double red[kEwBlock];
  const int t = threadIdx.x, b = blockIdx.x;
  double acc[8];  // energy, xx, xy, xz, yy, yz, zz, correction energy
  for (int c = 0; c < 8; c++) acc[c] = 0;
  const int k0 = (int)((long long)d.nk * b / kEwFinBlocks),
            k1 = (int)((long long)d.nk * (b + 1) / kEwFinBlocks);
  for (int kk = k0 + t; kk < k1; kk += kEwBlock) {
    const double4 k = d.kv[kk];
    const double2 s = d.sk[kk];
    const double k2 = k.x * k.x + k.y * k.y + k.z * k.z;
    const double e = k.w * (s.x * s.x + s.y * s.y);
    acc[0] += e;
    const double c = 2.0 * (1.0 / k2 + d.fac);
    acc[1] += e * (1.0 - c * k.x * k.x);
    acc[2] -= e * c * k.x * k.y;
    acc[3] -= e * c * k.x * k.z;
    acc[4] += e * (1.0 - c * k.y * k.y);
    acc[5] -= e * c * k.y * k.z;
    acc[6] += e * (1.0 - c * k.z * k.z);
  }
  // correction: energy (acc[7]) and virial of each excluded pair
  double cv[6] = {0, 0, 0, 0, 0, 0};
  const int r0 = (int)((long long)d.n_excl * b / kEwFinBlocks),
            r1 = (int)((long long)d.n_excl * (b + 1) / kEwFinBlocks);
  for (int row = r0 + t; row < r1; row += kEwBlock) {
    const int a1 = d.xrows[3 * row], a2 = d.xrows[3 * row + 1],
              cls = d.xrows[3 * row + 2];
    double rx, ry, rz, r2;
    min_diff_exact(d.box, x[3 * a1], x[3 * a1 + 1], x[3 * a1 + 2], x[3 * a2],
                   x[3 * a2 + 1], x[3 * a2 + 2], rx, ry, rz, r2);
    double qq = d.q[a1] * d.q[a2];
    if (cls == 1) qq *= 1 - d.csf;
    const double r = sqrt(r2), rr = 1 / r;
    const double e1 = erf(d.alpha * r) * rr;
    acc[7] -= qq * e1;
    const double g = qq * (d.two_alpha_pi * exp(-d.alpha2 * r2) - e1) * rr * rr;
    const double f0 = rx * g, f1 = ry * g, f2 = rz * g;
    cv[0] += f0 * rx;
    cv[1] += f0 * ry;
    cv[2] += f0 * rz;
    cv[3] += f1 * ry;
    cv[4] += f1 * rz;
    cv[5] += f2 * rz;
  }
  double *out = d.part + (size_t)b * kEwFinSums;
  for (int c = 0; c < 8; c++) {
    const double v = ew_block_sum(acc[c], red);
    if (t == 0) out[c] = v;
  }
  for (int c = 0; c < 6; c++) {
    const double v = ew_block_sum(cv[c], red);
    if (t == 0) out[8 + c] = v;
  }
}

// One wave: lane c < 14 adds sum c of every block, in block order; then
// out[1] += reciprocal + correction + constant energies, and with the
// virial flag out[2..10] += reciprocal and correction virial +
// charged-system diagonal (e_const = self + charged energy).
__global__ __launch_bounds__(64) void k_ew_finish(EwDev d, double *out,
                                                  int virial, double e_const,
                                                  double e_charged) {
  __shared__ double tot[kEwFinSums];
  const int t = threadIdx.x;
  if (t < kEwFinSums) {
    double v = 0.0;
    for (int b = 0; b < kEwFinBlocks; b++) v += d.part[(size_t)b * kEwFinSums + t];
    tot[t] = v;
  }
  __syncthreads();
  if (t == 0) {
    const double c = 4.0 * M_PI * d.Vr;
    out[1] += c * tot[0] + tot[7] + e_const;
    if (virial) {
      // slots XX XY XZ YX YY YZ ZX ZY ZZ (ScalarStructure::addVirial)
      const double xx = c * tot[1] + tot[8] + e_charged;
      const double xy = c * tot[2] + tot[9];
      const double xz = c * tot[3] + tot[10];
      const double yy = c * tot[4] + tot[11] + e_charged;
      const double yz = c * tot[5] + tot[12];
      const double zz = c * tot[6] + tot[13] + e_charged;
      out[2] += xx; out[3] += xy; out[4] += xz;
      out[5] += xy; out[6] += yy; out[7] += yz;
      out[8] += xz; out[9] += yz; out[10] += zz;
    }
  }
}

}  // namespace pmnb
