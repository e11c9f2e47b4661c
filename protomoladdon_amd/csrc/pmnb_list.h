// Neighbour-list path of the cutoff force.
//
// Every evaluation rebuilds the list from the cell list (no skin, no reuse
// across calls), so the result is a function of the current positions only,
// as in the reference, which rebuilds its cell lists on every call
// (force/ForceGroup.cpp:57, topology/Topology.h:44-86).
//
// k_build_list: one lane per atom of the sorted order walks its shell cells,
//   skips cells whose nearest point is beyond the cutoff, tests every atom in
//   FP32 on cell-local coordinates and re-tests pairs within `band` of the
//   cutoff with the reference's FP64 minimum image, so the kept set is
//   exactly the FP64 pair set; FULL exclusions are dropped.  Each kept
//   neighbour is stored as (sorted index | shell offset << 23) in a
//   slot-major array (slot k of atom s at k * n + s) so the force kernel
//   reads it coalesced.
// k_force_list_*: one lane per atom sums its forces over its list in a fixed
//   order (no atomics); energies and virials take half of each visit.
//   Atoms with more neighbours than the list holds are computed by the
//   direct-scan kernels instead.
#pragma once

#include "pmnb_common.h"
#include "pmnb_pair.h"

namespace pmnb {

constexpr int kEntryAtomBits = 23;
constexpr unsigned kEntryAtomMask = (1u << kEntryAtomBits) - 1u;
constexpr int kMaxShell = 512;  // shell offsets held in LDS / entry bits
constexpr int kChunk = 16;      // pairs per FP32 partial before flushing

template <bool WRAP>
__global__ void __launch_bounds__(kBlock)
k_build_list(ForceParams P, Sorted S, unsigned *__restrict__ list,
             int *__restrict__ count, int cap,
             unsigned long long *__restrict__ npairs, PairLog plog) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  const int n = P.n_atoms;
  unsigned cnt = 0;
  if (s < n) {
    const float cut2 = (float)P.cut2, band = P.band;
    const float hi = cut2 + band, lo = cut2 - band;
    const float Lx = (float)P.box.e[0], Ly = (float)P.box.e[1],
                Lz = (float)P.box.e[2];
    const float iLx = (float)P.box.er[0], iLy = (float)P.box.er[1],
                iLz = (float)P.box.er[2];
    const float csx = (float)P.grid.cs[0], csy = (float)P.grid.cs[1],
                csz = (float)P.grid.cs[2];
    const float4 li = S.loc[s];
    const int a = S.atom[s], mi = S.mol[s];
    const int c = S.cell[s];
    const int cx = c / (P.grid.n[1] * P.grid.n[2]);
    const int cy = (c / P.grid.n[2]) % P.grid.n[1];
    const int cz = c % P.grid.n[2];
    int m = 0;
    for (int d = 0; d < P.ndelta; d++) {
      const int4 dd = S.delta[d];
      if (!WRAP) {
        // distance from the atom to the neighbour cell's box
        const float gx = fmaxf(0.f, fmaxf((float)dd.x * csx - li.x,
                                          li.x - (float)(dd.x + 1) * csx));
        const float gy = fmaxf(0.f, fmaxf((float)dd.y * csy - li.y,
                                          li.y - (float)(dd.y + 1) * csy));
        const float gz = fmaxf(0.f, fmaxf((float)dd.z * csz - li.z,
                                          li.z - (float)(dd.z + 1) * csz));
        if (fmaf(gx, gx, fmaf(gy, gy, gz * gz)) > hi) continue;
      }
      int cj;
      shell_cell(P.grid, cx, cy, cz, dd, cj);
      const float ox = (float)(dd.x * P.grid.cs[0]) - li.x;
      const float oy = (float)(dd.y * P.grid.cs[1]) - li.y;
      const float oz = (float)(dd.z * P.grid.cs[2]) - li.z;
      const int t0 = S.cstart[cj], t1 = t0 + S.ccount[cj];
      for (int t = t0; t < t1; t++) {
        if (t == s) continue;
        const float4 lt = S.loc[t];
        float dx = lt.x + ox, dy = lt.y + oy, dz = lt.z + oz;
        if (WRAP) {
          dx = fmaf(-Lx, rintf(dx * iLx), dx);
          dy = fmaf(-Ly, rintf(dy * iLy), dy);
          dz = fmaf(-Lz, rintf(dz * iLz), dz);
        }
        const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
        if (r2 > hi) continue;
        if (r2 >= lo) {
          double ex, ey, ez, e2d;
          min_diff_exact(P.box, S.x[s], S.y[s], S.z[s], S.x[t], S.y[t],
                         S.z[t], ex, ey, ez, e2d);
          if (e2d > P.cut2) continue;
        }
        const int cls = S.mol[t] == mi ? excl_class(S.xfull, S.xmod,
                                                    P.max_excl_delta, a,
                                                    S.atom[t])
                                       : 0;
        if (cls == 2) continue;
        if (m < cap) list[(size_t)m * n + s] = (unsigned)t | ((unsigned)d << kEntryAtomBits);
        m++;
        if (s < t) {
          cnt++;
          if (plog.out) log_pair(plog, a, S.atom[t], cls);
        }
      }
    }
    count[s] = m;
  }
  unsigned long long tot = block_count(cnt);
  if (threadIdx.x == 0 && tot) atomicAdd(npairs, tot);
}

template <int ELEC>
__global__ void __launch_bounds__(kBlock)
k_force_list_mixed(ForceParams P, Sorted S, const unsigned *__restrict__ list,
                   const int *__restrict__ count, int cap,
                   const float4 *__restrict__ shell_off,
                   double *__restrict__ forces, double *__restrict__ partial) {
  __shared__ float4 off[kMaxShell];
  for (int k = threadIdx.x; k < P.ndelta; k += blockDim.x) off[k] = shell_off[k];
  __syncthreads();
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  const int n = P.n_atoms;
  double acc[ACC_N];
#pragma unroll
  for (int k = 0; k < ACC_N; k++) acc[k] = 0.0;
  const int m = s < n ? count[s] : 0;
  if (s < n && m <= cap) {
    const F32Params &F = P.f;
    const float4 li = S.loc[s];
    const int a = S.atom[s], ti = S.type[s], mi = S.mol[s];
    const float4 *ljrow = S.ljf + ti * P.n_types;
    const bool vir = (P.flags & 3u) != 0;
    long long Fx = 0, Fy = 0, Fz = 0;
    for (int k0 = 0; k0 < m; k0 += kChunk) {
      const int k1 = min(m, k0 + kChunk);
      float px = 0.f, py = 0.f, pz = 0.f, pelj = 0.f, peel = 0.f;
      float v0 = 0.f, v1 = 0.f, v2 = 0.f, v3 = 0.f, v4 = 0.f, v5 = 0.f;
      for (int k = k0; k < k1; k++) {
        const unsigned e = list[(size_t)k * n + s];
        const int t = (int)(e & kEntryAtomMask);
        const float4 o = off[e >> kEntryAtomBits];
        const float4 lt = S.loc[t];
        float dx = lt.x + (o.x - li.x), dy = lt.y + (o.y - li.y),
              dz = lt.z + (o.z - li.z);
        if (P.per_pair_wrap) {
          dx = fmaf(-F.Lx, rintf(dx * F.iLx), dx);
          dy = fmaf(-F.Ly, rintf(dy * F.iLy), dy);
          dz = fmaf(-F.Lz, rintf(dz * F.iLz), dz);
        }
        const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
        const int mj = S.mol[t];
        const bool same = (mi == mj);
        const int excl =
            same ? excl_class(S.xfull, S.xmod, P.max_excl_delta, a, S.atom[t])
                 : 0;
        const float4 lp = ljrow[S.type[t]];
        float ff, e1, e2;
        pair_f32<ELEC>(F, r2, excl != 1 ? lp.x : lp.z, excl != 1 ? lp.y : lp.w,
                       li.w * lt.w, excl, ff, e1, e2);
        pelj += e1;
        peel += e2;
        const float gx = dx * ff, gy = dy * ff, gz = dz * ff;
        px -= gx;
        py -= gy;
        pz -= gz;
        if (vir) {
          const bool molv = (P.flags & 2u) && !same;
          if (molv || (P.flags & 1u)) {
            v0 = fmaf(gx, dx, v0);
            v1 = fmaf(gx, dy, v1);
            v2 = fmaf(gx, dz, v2);
            v3 = fmaf(gy, dy, v3);
            v4 = fmaf(gy, dz, v4);
            v5 = fmaf(gz, dz, v5);
          }
          if (molv) {
            double cdx, cdy, cdz, cd2;
            min_diff_exact(P.box, S.com[3 * mi], S.com[3 * mi + 1],
                           S.com[3 * mi + 2], S.com[3 * mj],
                           S.com[3 * mj + 1], S.com[3 * mj + 2], cdx, cdy,
                           cdz, cd2);
            add_virial(acc, ACC_MXX, gx, gy, gz, cdx, cdy, cdz, 0.5);
          }
        }
      }
      Fx += to_fixed((double)px);
      Fy += to_fixed((double)py);
      Fz += to_fixed((double)pz);
      acc[ACC_ELJ] += 0.5 * (double)pelj;
      acc[ACC_EEL] += 0.5 * (double)peel;
      if (vir) {
        acc[ACC_VXX] += 0.5 * (double)v0;
        acc[ACC_VXY] += 0.5 * (double)v1;
        acc[ACC_VXZ] += 0.5 * (double)v2;
        acc[ACC_VYY] += 0.5 * (double)v3;
        acc[ACC_VYZ] += 0.5 * (double)v4;
        acc[ACC_VZZ] += 0.5 * (double)v5;
      }
    }
    forces[3 * a] += (double)Fx * kFixedInv;
    forces[3 * a + 1] += (double)Fy * kFixedInv;
    forces[3 * a + 2] += (double)Fz * kFixedInv;
  }
  block_reduce_store(acc, partial);
}

// FP64 pair of OneAtomPairTwo with the reference's operation order
// (force/OneAtomPairTwo.h:82-106).
template <int ELEC, int ELSW>
__device__ __forceinline__ void pair_f64(const ForceParams &P, double d2,
                                         double A, double B, double qq,
                                         int excl, double &ff, double &e1,
                                         double &e2) {
  const double rd2 = 1.0 / d2;
  const double r6 = rd2 * rd2 * rd2;
  const double r12 = r6 * r6;
  const double r6B = B * r6;
  const double r12A = A * r12;
  e1 = r12A - r6B;
  double f1 = 12.0 * r12A * rd2 - 6.0 * r6B * rd2;
  double f2 = 0.0;
  e2 = 0.0;
  if (ELEC == 1) {
    e2 = qq * sqrt(rd2);
    if (excl == 1) e2 *= P.csf;
    f2 = e2 * rd2;
  } else if (ELEC == 2) {
    const double r = sqrt(d2);
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
  ff = f1 + f2;
}

template <int ELEC, int ELSW>
__global__ void __launch_bounds__(kBlock)
k_force_list_fp64(ForceParams P, Sorted S, const unsigned *__restrict__ list,
                  const int *__restrict__ count, int cap,
                  double *__restrict__ forces, double *__restrict__ partial) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  const int n = P.n_atoms;
  double acc[ACC_N];
#pragma unroll
  for (int k = 0; k < ACC_N; k++) acc[k] = 0.0;
  const int m = s < n ? count[s] : 0;
  if (s < n && m <= cap) {
    const double xi = S.x[s], yi = S.y[s], zi = S.z[s];
    const int a = S.atom[s], ti = S.type[s], mi = S.mol[s];
    const double qi = S.q[s];
    const double *ljrow = S.lj + 4 * (ti * P.n_types);
    double fx = 0.0, fy = 0.0, fz = 0.0;
    for (int k = 0; k < m; k++) {
      const int t = (int)(list[(size_t)k * n + s] & kEntryAtomMask);
      double dx, dy, dz, d2;
      min_diff_exact(P.box, xi, yi, zi, S.x[t], S.y[t], S.z[t], dx, dy, dz,
                     d2);
      const int mj = S.mol[t];
      const bool same = (mi == mj);
      const int excl =
          same ? excl_class(S.xfull, S.xmod, P.max_excl_delta, a, S.atom[t])
               : 0;
      const double *lp = ljrow + 4 * S.type[t];
      double ff, e1, e2;
      pair_f64<ELEC, ELSW>(P, d2, excl != 1 ? lp[0] : lp[2],
                           excl != 1 ? lp[1] : lp[3], qi * S.q[t], excl, ff,
                           e1, e2);
      acc[ACC_ELJ] += 0.5 * e1;
      acc[ACC_EEL] += 0.5 * e2;
      const double gx = dx * ff, gy = dy * ff, gz = dz * ff;
      fx -= gx;
      fy -= gy;
      fz -= gz;
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
    }
    forces[3 * a] += fx;
    forces[3 * a + 1] += fy;
    forces[3 * a + 2] += fz;
  }
  block_reduce_store(acc, partial);
}

}  // namespace pmnb
