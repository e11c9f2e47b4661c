// Pair kernels of the nonbonded cutoff force (full-shell, one lane per
// atom).  Each lane owns one atom of the sorted order, walks every cell of
// the shell around its own cell (the reference's delta list plus the cell
// itself) and sums the force on its atom in a fixed order, so no two lanes
// ever write the same force and every result is reproducible bit for bit.
// Every pair is visited from both ends; energies and virials take half of
// each visit.
//
// Per-pair arithmetic follows OneAtomPairTwo::doOneAtomPair
// (force/OneAtomPairTwo.h:57-125) with LennardJonesForce
// (force/LennardJonesForce.h:29-61), CoulombForce (force/CoulombForce.h:27-38)
// or the Ewald real-space term (force/nonbonded/NonbondedFullEwaldSystemForce.h
// :810-821) and C2/C1 switching.
#pragma once

#include "pmnb_common.h"

namespace pmnb {

struct Sorted {
  const double *x, *y, *z;       // raw FP64 positions in sorted order
  const float4 *loc;             // cell-local FP32 offsets + charge
  const int *atom;               // original index of each sorted slot
  const int *type;
  const double *q;
  const int *mol;
  const int *cell;               // linear cell of each sorted slot
  const int2 *rec;               // {type | molecule << 8, original index}
                                 // (tile path only; < 256 types and
                                 // < 2^24 atoms there)
  const int *cstart, *ccount;    // per-cell segment of the sorted order
  const int4 *delta;             // shell offsets, (0,0,0) first
  const unsigned long long *xfull, *xmod;  // exclusion masks (original idx)
  const unsigned long long *xown;  // per atom {forward FULL, forward MODIFIED,
                                   // reverse FULL, reverse MODIFIED}
  const double *lj;              // T*T*4 {A, B, A14, B14}
  const float4 *ljf;             // same in FP32
  const double *com;             // molecule centres (molecular virial)
  const Box *box;                // device copy of the box (read on demand)
};

// Wave-then-block reduction of ACC_N doubles in a fixed order; block b
// of NW waves writes partial[b * ACC_N + k].
template <int NW = kBlock / 64>
__device__ __forceinline__ void block_reduce_store(double (&acc)[ACC_N],
                                                   double *partial) {
  __shared__ double red[NW][ACC_N];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < ACC_N; k++) {
    double v = acc[k];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if (lane == 0) red[wave][k] = v;
  }
  __syncthreads();
  if (threadIdx.x < ACC_N) {
    double s = 0.0;
    for (int w = 0; w < NW; w++) s += red[w][threadIdx.x];
    partial[blockIdx.x * ACC_N + threadIdx.x] = s;
  }
}

template <int NW = kBlock / 64>
__device__ __forceinline__ unsigned long long block_count(unsigned c) {
  __shared__ unsigned long long cred[NW];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  unsigned long long v = c;
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  if (lane == 0) cred[wave] = v;
  __syncthreads();
  unsigned long long s = 0;
  if (threadIdx.x == 0)
    for (int w = 0; w < NW; w++) s += cred[w];
  return s;
}

__device__ __forceinline__ void shell_cell(const Grid &g, int cx, int cy,
                                           int cz, int4 d, int &cj) {
  int jx = cx + d.x, jy = cy + d.y, jz = cz + d.z;
  jx = jx < 0 ? jx + g.n[0] : (jx >= g.n[0] ? jx - g.n[0] : jx);
  jy = jy < 0 ? jy + g.n[1] : (jy >= g.n[1] ? jy - g.n[1] : jy);
  jz = jz < 0 ? jz + g.n[2] : (jz >= g.n[2] ? jz - g.n[2] : jz);
  cj = (jx * g.n[1] + jy) * g.n[2] + jz;
}

__device__ __forceinline__ void add_virial(double (&acc)[ACC_N], int base,
                                           double fx, double fy, double fz,
                                           double dx, double dy, double dz,
                                           double w) {
  // ScalarStructure::addVirial (type/ScalarStructure.cpp:36-54): slots
  // XX, XY=YX=f_x d_y, XZ=ZX=f_x d_z, YY, YZ=ZY=f_y d_z, ZZ.
  acc[base + 0] += w * (fx * dx);
  acc[base + 1] += w * (fx * dy);
  acc[base + 2] += w * (fx * dz);
  acc[base + 3] += w * (fy * dy);
  acc[base + 4] += w * (fy * dz);
  acc[base + 5] += w * (fz * dz);
}

// ------------------------------------------------------------------ FP64
template <int ELEC, int ELSW>
__global__ void __launch_bounds__(kBlock)
k_pair_fp64(ForceParams P, Sorted S, double *__restrict__ forces,
            double *__restrict__ partial,
            unsigned long long *__restrict__ npairs,
            const int *__restrict__ only, int cap, PairLog plog) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  double acc[ACC_N];
#pragma unroll
  for (int k = 0; k < ACC_N; k++) acc[k] = 0.0;
  unsigned cnt = 0;
  if (s < P.n_atoms && (!only || only[s] > cap)) {
    const double xi = S.x[s], yi = S.y[s], zi = S.z[s];
    const int a = S.atom[s], ti = S.type[s], mi = S.mol[s];
    const double qi = S.q[s];
    const int c = S.cell[s];
    const int cx = c / (P.grid.n[1] * P.grid.n[2]);
    const int cy = (c / P.grid.n[2]) % P.grid.n[1];
    const int cz = c % P.grid.n[2];
    const double *ljrow = S.lj + 4 * (ti * P.n_types);
    double fx = 0.0, fy = 0.0, fz = 0.0;
    for (int d = 0; d < P.ndelta; d++) {
      int cj;
      shell_cell(P.grid, cx, cy, cz, S.delta[d], cj);
      const int t0 = S.cstart[cj], t1 = t0 + S.ccount[cj];
      for (int t = t0; t < t1; t++) {
        if (t == s) continue;
        double dx, dy, dz, d2;
        min_diff_exact(P.box, xi, yi, zi, S.x[t], S.y[t], S.z[t], dx, dy, dz,
                       d2);
        if (d2 > P.cut2) continue;
        const int b = S.atom[t], mj = S.mol[t];
        const bool same = (mi == mj);
        const int excl =
            same ? excl_class(S.xfull, S.xmod, P.max_excl_delta, a, b) : 0;
        if (excl == 2) continue;
        const double rd2 = 1.0 / d2;
        // LennardJonesForce
        const double *lp = ljrow + 4 * S.type[t];
        const double A = excl != 1 ? lp[0] : lp[2];
        const double B = excl != 1 ? lp[1] : lp[3];
        const double r6 = rd2 * rd2 * rd2;
        const double r12 = r6 * r6;
        const double r6B = B * r6;
        const double r12A = A * r12;
        double e1 = r12A - r6B;
        double f1 = 12.0 * r12A * rd2 - 6.0 * r6B * rd2;
        double e2 = 0.0, f2 = 0.0;
        if (ELEC == 1) {  // CoulombForce
          e2 = qi * S.q[t] * sqrt(rd2);
          if (excl == 1) e2 *= P.csf;
          f2 = e2 * rd2;
        } else if (ELEC == 2) {  // Ewald real space, A&S erfc
          const double r = sqrt(d2);
          double qq = qi * S.q[t];
          if (excl == 1) qq *= P.csf;
          const double rr = 1.0 / r;
          const double ar = P.alpha * r;
          const double e = qq * exp(-ar * ar);
          e2 = poly5_exact(ar) * e * rr;
          f2 = ((e2 + P.two_alpha_pi * e) * rr * rr);
        }
        double sv, sd;
        c2_exact(P.sw_lj, d2, sv, sd);
        f1 = f1 * sv - e1 * sd;
        e1 = e1 * sv;
        if (ELEC != 0) {
          if (ELSW == 0) c2_exact(P.sw_el, d2, sv, sd);
          else c1_exact(P.sw_el, d2, sv, sd);
          f2 = f2 * sv - e2 * sd;
          e2 = e2 * sv;
        }
        acc[ACC_ELJ] += 0.5 * e1;
        acc[ACC_EEL] += 0.5 * e2;
        const double ff = f1 + f2;
        const double gx = dx * ff, gy = dy * ff, gz = dz * ff;
        fx -= gx;
        fy -= gy;
        fz -= gz;
        // OneAtomPairTwo.h:114-121: atomic + molecular virial for pairs of
        // different molecules when the molecular virial is on, else the
        // atomic virial alone when it is on.
        if ((P.flags & 2u) && !same) {
          add_virial(acc, ACC_VXX, gx, gy, gz, dx, dy, dz, 0.5);
          double cdx, cdy, cdz, cd2;
          min_diff_exact(P.box, S.com[3 * mi], S.com[3 * mi + 1],
                         S.com[3 * mi + 2], S.com[3 * mj], S.com[3 * mj + 1],
                         S.com[3 * mj + 2], cdx, cdy, cdz, cd2);
          add_virial(acc, ACC_MXX, gx, gy, gz, cdx, cdy, cdz, 0.5);
        } else if (P.flags & 1u) {
          add_virial(acc, ACC_VXX, gx, gy, gz, dx, dy, dz, 0.5);
        }
        if (a < b) {
          cnt++;
          if (plog.out && !only) log_pair(plog, a, b, excl);
        }
      }
    }
    forces[3 * a] += fx;
    forces[3 * a + 1] += fy;
    forces[3 * a + 2] += fz;
  }
  block_reduce_store(acc, partial);
  if (!only) {
    unsigned long long tot = block_count(cnt);
    if (threadIdx.x == 0 && tot) atomicAdd(npairs, tot);
  }
}

// ----------------------------------------------------------------- mixed
// 64-bit fixed point with 2^-32 resolution; values beyond +-2^31 (far
// outside any physical force in kcal/mol/A) saturate instead of wrapping.
__device__ __forceinline__ long long to_fixed(double v) {
  const double lim = 2147483647.0;
  v = fmin(fmax(v, -lim), lim);
  return (long long)__builtin_rint(v * kFixedScale);
}

// FP32 pair of OneAtomPairTwo: returns the switched F/r (ff) and the two
// switched energies.  A, B are already the MODIFIED (A14, B14) values when
// excl == 1; qq is the raw charge product.
template <int ELEC>
__device__ __forceinline__ void pair_f32(const F32Params &F, float r2,
                                         float A, float B, float qq,
                                         int excl, float &ff, float &e1,
                                         float &e2) {
  // with electrostatics 1/r^2 is the square of the 1/r the term needs (one
  // transcendental instead of two)
  const float rr = ELEC != 0 ? __builtin_amdgcn_rsqf(r2) : 0.f;
  const float rd2 = ELEC != 0 ? rr * rr : __builtin_amdgcn_rcpf(r2);
  const float r6 = rd2 * rd2 * rd2;
  const float r6B = B * r6;
  const float r12A = A * r6 * r6;
  e1 = r12A - r6B;
  float f1 = fmaf(12.f, r12A, -6.f * r6B) * rd2;
  float f2 = 0.f;
  e2 = 0.f;
  if (ELEC == 1) {
    if (excl == 1) qq *= F.csf;
    e2 = qq * rr;
    f2 = e2 * rd2;
  } else if (ELEC == 2) {
    if (excl == 1) qq *= F.csf;
    const float r = r2 * rr;
    const float ar = F.alpha * r;
    // Abramowitz & Stegun 7.1.26 (base/MathUtilities.cpp:39-49)
    const float tt = __builtin_amdgcn_rcpf(fmaf(0.3275911f, ar, 1.f));
    const float poly =
        tt * fmaf(tt, fmaf(tt, fmaf(tt, fmaf(tt, 1.061405429f, -1.453152027f),
                                    1.421413741f),
                           -0.284496736f),
                  0.254829592f);
    const float e = qq * __builtin_amdgcn_exp2f(-ar * ar * 1.4426950408889634f);
    e2 = poly * e * rr;
    f2 = fmaf(F.tapi, e, e2) * rd2;
  }
  // C2 switch without branches (most waves hold pairs on both sides of the
  // switch-on distance): c2 = 0 beyond the cutoff gives S = S~ = 0
  float c2 = fmaxf(F.cut2l - r2, 0.f);
  float c4 = c2 * fmaf(2.f, r2, F.s2l);
  const bool swl = r2 >= F.on2l;
  const float sv = swl ? F.s1l * (c2 * c4) : 1.f;
  const float sd = swl ? F.s3l * fmaf(c2, c2, -c4) : 0.f;
  f1 = fmaf(f1, sv, -e1 * sd);
  e1 = e1 * sv;
  if (ELEC != 0 && F.same_switch) {
    f2 = fmaf(f2, sv, -e2 * sd);
    e2 = e2 * sv;
  } else if (ELEC != 0) {
    c2 = fmaxf(F.cut2e - r2, 0.f);
    c4 = c2 * fmaf(2.f, r2, F.s2e);
    const bool swe = r2 >= F.on2e;
    const float se = swe ? F.s1e * (c2 * c4) : 1.f;
    const float sde = swe ? F.s3e * fmaf(c2, c2, -c4) : 0.f;
    f2 = fmaf(f2, se, -e2 * sde);
    e2 = e2 * se;
  }
  ff = f1 + f2;
}

// Direct scan of the shell cells for one atom per lane (mixed precision).
// With `only` non-null, only lanes whose neighbour count exceeds `cap` run:
// the overflow path of the neighbour-list kernels.
template <int ELEC>
__global__ void __launch_bounds__(kBlock)
k_pair_mixed(ForceParams P, Sorted S, double *__restrict__ forces,
             double *__restrict__ partial,
             unsigned long long *__restrict__ npairs,
             const int *__restrict__ only, int cap, PairLog plog) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  double acc[ACC_N];
#pragma unroll
  for (int k = 0; k < ACC_N; k++) acc[k] = 0.0;
  unsigned cnt = 0;
  const bool run = s < P.n_atoms && (!only || only[s] > cap);
  if (run) {
    const F32Params &F = P.f;
    const float4 li = S.loc[s];
    const int a = S.atom[s], ti = S.type[s], mi = S.mol[s];
    const int c = S.cell[s];
    const int cx = c / (P.grid.n[1] * P.grid.n[2]);
    const int cy = (c / P.grid.n[2]) % P.grid.n[1];
    const int cz = c % P.grid.n[2];
    const float4 *ljrow = S.ljf + ti * P.n_types;
    long long Fx = 0, Fy = 0, Fz = 0;
    for (int d = 0; d < P.ndelta; d++) {
      const int4 dd = S.delta[d];
      int cj;
      shell_cell(P.grid, cx, cy, cz, dd, cj);
      const float ox = (float)(dd.x * P.grid.cs[0]) - li.x;
      const float oy = (float)(dd.y * P.grid.cs[1]) - li.y;
      const float oz = (float)(dd.z * P.grid.cs[2]) - li.z;
      const int t0 = S.cstart[cj], t1 = t0 + S.ccount[cj];
      float px = 0.f, py = 0.f, pz = 0.f;
      float pelj = 0.f, peel = 0.f;
      float v0 = 0.f, v1 = 0.f, v2 = 0.f, v3 = 0.f, v4 = 0.f, v5 = 0.f;
      for (int t = t0; t < t1; t++) {
        if (t == s) continue;
        const float4 lj4 = S.loc[t];
        float dx = lj4.x + ox, dy = lj4.y + oy, dz = lj4.z + oz;
        if (P.per_pair_wrap) {
          dx = fmaf(-F.Lx, rintf(dx * F.iLx), dx);
          dy = fmaf(-F.Ly, rintf(dy * F.iLy), dy);
          dz = fmaf(-F.Lz, rintf(dz * F.iLz), dz);
        }
        const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
        if (r2 > F.cut2 + F.band) continue;
        if (r2 >= F.cut2 - F.band) {
          // Near the cutoff: decide with the reference's FP64 arithmetic.
          double ex, ey, ez, e2d;
          min_diff_exact(P.box, S.x[s], S.y[s], S.z[s], S.x[t], S.y[t],
                         S.z[t], ex, ey, ez, e2d);
          if (e2d > P.cut2) continue;
        }
        const int b = S.atom[t], mj = S.mol[t];
        const bool same = (mi == mj);
        const int excl =
            same ? excl_class(S.xfull, S.xmod, P.max_excl_delta, a, b) : 0;
        if (excl == 2) continue;
        const float4 lp = ljrow[S.type[t]];
        float ff, e1, e2;
        pair_f32<ELEC>(F, r2, excl != 1 ? lp.x : lp.z, excl != 1 ? lp.y : lp.w,
                       li.w * lj4.w, excl, ff, e1, e2);
        pelj += e1;
        peel += e2;
        const float gx = dx * ff, gy = dy * ff, gz = dz * ff;
        px -= gx;
        py -= gy;
        pz -= gz;
        const bool molv = (P.flags & 2u) && !same;
        if (molv || (P.flags & 1u)) {
          v0 = fmaf(gx, dx, v0);
          v1 = fmaf(gx, dy, v1);
          v2 = fmaf(gx, dz, v2);
          v3 = fmaf(gy, dy, v3);
          v4 = fmaf(gy, dz, v4);
          v5 = fmaf(gz, dz, v5);
          if (molv) {
            double cdx, cdy, cdz, cd2;
            min_diff_exact(P.box, S.com[3 * mi], S.com[3 * mi + 1],
                           S.com[3 * mi + 2], S.com[3 * mj],
                           S.com[3 * mj + 1], S.com[3 * mj + 2], cdx, cdy,
                           cdz, cd2);
            add_virial(acc, ACC_MXX, gx, gy, gz, cdx, cdy, cdz, 0.5);
          }
        }
        if (a < b) {
          cnt++;
          if (plog.out && !only) log_pair(plog, a, b, excl);
        }
      }
      // Flush this cell's FP32 partials: forces to 64-bit fixed point,
      // energies and virial to FP64.
      Fx += to_fixed((double)px);
      Fy += to_fixed((double)py);
      Fz += to_fixed((double)pz);
      acc[ACC_ELJ] += 0.5 * (double)pelj;
      acc[ACC_EEL] += 0.5 * (double)peel;
      acc[ACC_VXX] += 0.5 * (double)v0;
      acc[ACC_VXY] += 0.5 * (double)v1;
      acc[ACC_VXZ] += 0.5 * (double)v2;
      acc[ACC_VYY] += 0.5 * (double)v3;
      acc[ACC_VYZ] += 0.5 * (double)v4;
      acc[ACC_VZZ] += 0.5 * (double)v5;
    }
    forces[3 * a] += (double)Fx * kFixedInv;
    forces[3 * a + 1] += (double)Fy * kFixedInv;
    forces[3 * a + 2] += (double)Fz * kFixedInv;
  }
  block_reduce_store(acc, partial);
  if (!only) {
    unsigned long long tot = block_count(cnt);
    if (threadIdx.x == 0 && tot) atomicAdd(npairs, tot);
  }
}

// ------------------------------------------------------- pair-set extract
// FP64 pair decision exactly as the FP64 kernel; writes (i, j, class) for
// i < j into out (when below cap) and counts all.
__global__ void __launch_bounds__(kBlock)
k_pairset(ForceParams P, Sorted S, int *__restrict__ out, long long cap,
          unsigned long long *__restrict__ counter) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= P.n_atoms) return;
  const double xi = S.x[s], yi = S.y[s], zi = S.z[s];
  const int a = S.atom[s], mi = S.mol[s];
  const int c = S.cell[s];
  const int cx = c / (P.grid.n[1] * P.grid.n[2]);
  const int cy = (c / P.grid.n[2]) % P.grid.n[1];
  const int cz = c % P.grid.n[2];
  for (int d = 0; d < P.ndelta; d++) {
    int cj;
    shell_cell(P.grid, cx, cy, cz, S.delta[d], cj);
    const int t0 = S.cstart[cj], t1 = t0 + S.ccount[cj];
    for (int t = t0; t < t1; t++) {
      const int b = S.atom[t];
      if (b <= a) continue;
      double dx, dy, dz, d2;
      min_diff_exact(P.box, xi, yi, zi, S.x[t], S.y[t], S.z[t], dx, dy, dz,
                     d2);
      if (d2 > P.cut2) continue;
      const int excl = (mi == S.mol[t])
                           ? excl_class(S.xfull, S.xmod, P.max_excl_delta, a, b)
                           : 0;
      if (excl == 2) continue;
      unsigned long long k = atomicAdd(counter, 1ull);
      if ((long long)k < cap) {
        out[3 * k] = a;
        out[3 * k + 1] = b;
        out[3 * k + 2] = excl;
      }
    }
  }
}

// Sums the per-block partials of up to two kernels (one wave per slot,
// lanes striding over blocks, then a fixed shuffle tree: deterministic) and
// accumulates into out[] in the 20-slot layout of pmnb_evaluate_device.
__global__ void __launch_bounds__(1024)
k_finish(const double *__restrict__ p1, int nb1, const double *__restrict__ p2,
         int nb2, double *__restrict__ out) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int k = wave; k < ACC_N; k += 16) {
    double v = 0.0;
    for (int b = lane; b < nb1; b += 64) v += p1[b * ACC_N + k];
    for (int b = lane; b < nb2; b += 64) v += p2[b * ACC_N + k];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if (lane != 0) continue;
    // out: [0] LJ, [1] elec, [2..10] virial 3x3, [11..19] mol virial 3x3
    if (k < 2) {
      out[k] += v;
    } else {
      const int base = k < ACC_MXX ? 2 : 11;
      const int u = k < ACC_MXX ? k - ACC_VXX : k - ACC_MXX;
      // u: 0 xx, 1 xy, 2 xz, 3 yy, 4 yz, 5 zz
      switch (u) {
        case 0: out[base + 0] += v; break;
        case 1: out[base + 1] += v; out[base + 3] += v; break;
        case 2: out[base + 2] += v; out[base + 6] += v; break;
        case 3: out[base + 4] += v; break;
        case 4: out[base + 5] += v; out[base + 7] += v; break;
        default: out[base + 8] += v; break;
      }
    }
  }
}

}  // namespace pmnb
