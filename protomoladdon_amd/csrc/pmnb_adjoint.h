// Adjoint (reverse-mode derivative) of the cutoff force.
//
// Given an upstream gradient g = dL/dF (3 per atom), returns
//   dL/dx_k = sum_j dL/dF_j . dF_j/dx_k   (a Hessian-vector product, -H g)
//   dL/dq~_k for every scaled charge, and
//   dL/d{A, B, A14, B14} for every unordered LJ type pair,
// recomputing the pairs from the cell list instead of storing them.
//
// The force on atom i from pair (i, j) is F_i = -d ff(a), d = x_j - x_i
// (minimum image), a = |d|^2, ff = F/r of OneAtomPairTwo::doOneAtomPair
// (force/OneAtomPairTwo.h:96-111): ff = F S - E S~ for each of the two terms,
// with S the switch and S~ its derivative divided by r.  Its Jacobian is
// dF_i/dx_i = ff I + 2 ff' d d^T (ff' = dff/da), the negative of the pair
// Hessian block the reference builds in ReducedHessLennardJones
// (force/hessian/ReducedHessLennardJones.cpp:7-42) and ReducedHessCoulomb
// (force/hessian/ReducedHessCoulomb.cpp:6-35) with the C2 switch Hessian
// (switch/C2SwitchingFunction.cpp:43-52).  With w = g_j - g_i:
//   dL/dx_i  += -(ff w + 2 ff' (w . d) d)
//   dL/dq~_i += (w . d) q~_j h        (ff of the electrostatic term = qq h)
//   dL/dA    += (w . d) dff/dA / 2    (each unordered pair seen twice)
// For the Ewald term ff' differentiates the force expression as the
// reference writes it (the erfc of its energy is the A&S approximation,
// its force uses the exact erfc derivative), so the Jacobian is that of the
// forces actually computed.
//
// Work decomposition and candidate streaming are those of the tile-path
// pair kernels (one wave per work item of 16 sorted atoms, candidates of the
// 3 x 3 neighbour columns buffered in LDS after a bounding-box test); the
// test appends passing (atom, record) entries to an LDS pair queue at
// ballot + mbcnt positions, which the wave consumes one pair per lane.  Each
// atom's dL/dx and dL/dq~ are summed by its own
// wave in per-lane LDS slots (fixed order); the LJ-parameter gradients are
// summed per lane in registers (one per unordered type pair, selected by
// comparison), then per item in a fixed order.
#pragma once

#include "pmnb_queue.h"

namespace pmnb {

constexpr int kAdjSlots = 32;  // LDS slots per atom (two lanes each)

// Constants of the adjoint in the kernel's precision.
template <typename R>
struct AdjC {
  R on2l, cut2l, s1l, s2l, s3l;   // C2 switch of the LJ term
  R on2e, cut2e, s1e, s2e, s3e;   // C2 switch of the electrostatic term
  R alpha, c2ap, csf;             // alpha, 2 alpha / sqrt(pi), 1-4 factor
};

// C2 value S, S~ = (dS/dr) / r = 2 dS/da and S'' = d2S/da2
// (switch/C2SwitchingFunction.h:42-53, .cpp:43-52).
template <typename R>
__device__ __forceinline__ void c2_adj(R a, R on2, R cut2, R s1, R s2, R s3,
                                       R &S, R &Sd, R &S2) {
  S = R(1);
  Sd = R(0);
  S2 = R(0);
  if (a > cut2) {
    S = R(0);
  } else if (a >= on2) {
    const R c2 = cut2 - a;
    const R c4 = c2 * (s2 + R(2) * a);
    S = s1 * (c2 * c4);
    Sd = s3 * (c2 * c2 - c4);
    S2 = R(2) * s1 * (s2 + R(2) * a - R(4) * c2);
  }
}

// Per-pair terms of the adjoint: ff, ff' = dff/da, h = (electrostatic ff) /
// qq, and dff/dA, dff/dB.  A, B are the MODIFIED (A14, B14) values for a
// 1-4 pair; qq includes the 1-4 factor.
template <typename R, int ELEC>
__device__ __forceinline__ void pair_adj(const AdjC<R> &C, R a, R A, R B,
                                         R qq, R &ff, R &ffp, R &h, R &dA,
                                         R &dB) {
  const R ia = R(1) / a;
  const R ia3 = ia * ia * ia, ia4 = ia3 * ia, ia6 = ia3 * ia3;
  const R ia5 = ia4 * ia, ia7 = ia6 * ia, ia8 = ia7 * ia;
  // LennardJonesForce (force/LennardJonesForce.h:29-61): E = A/r^12 - B/r^6
  const R E = A * ia6 - B * ia3;
  const R Ep = -R(6) * A * ia7 + R(3) * B * ia4;
  const R F = R(12) * A * ia7 - R(6) * B * ia4;
  const R Fp = -R(84) * A * ia8 + R(24) * B * ia5;
  R S, Sd, S2;
  c2_adj(a, C.on2l, C.cut2l, C.s1l, C.s2l, C.s3l, S, Sd, S2);
  ff = F * S - E * Sd;
  ffp = Fp * S + F * (R(0.5) * Sd) - Ep * Sd - R(2) * E * S2;
  dA = R(12) * ia7 * S - ia6 * Sd;
  dB = -R(6) * ia4 * S + ia3 * Sd;
  h = R(0);
  if (ELEC == 0) return;
  R u, up, v, vp;
  const R rr = sqrt(ia);
  if (ELEC == 1) {
    // CoulombForce (force/CoulombForce.h:27-38): E = qq / r, F/r = E / r^2
    u = rr;
    up = -R(0.5) * rr * ia;
    v = u * ia;
    vp = -R(1.5) * v * ia;
  } else {
    // Ewald real space: E = qq erfc(alpha r) / r with the A&S erfc
    // (base/MathUtilities.cpp:39-49), F/r = (E + qq 2 alpha/sqrt(pi)
    // exp(-alpha^2 r^2)) / r^2 (force/nonbonded/NonbondedFullEwaldSystemForce
    // .h:810-821)
    const R r = a * rr;
    const R ar = C.alpha * r;
    const R t = R(1) / (R(1) + R(0.3275911) * ar);
    const R e1 = R(0.254829592), e2 = R(-0.284496736), e3 = R(1.421413741),
            e4 = R(-1.453152027), e5 = R(1.061405429);
    const R P = t * (e1 + t * (e2 + t * (e3 + t * (e4 + t * e5))));
    const R Pt = e1 + t * (R(2) * e2 + t * (R(3) * e3 + t * (R(4) * e4 +
                                                             t * (R(5) * e5))));
    const R x = exp(-(C.alpha * C.alpha) * a);
    const R tp = -R(0.5) * R(0.3275911) * C.alpha * t * t * rr;  // dt/da
    u = P * x * rr;
    up = x * rr * (Pt * tp - C.alpha * C.alpha * P - R(0.5) * P * ia);
    v = (u + C.c2ap * x) * ia;
    vp = (up - C.c2ap * C.alpha * C.alpha * x) * ia - v * ia;
  }
  R Se, Sde, S2e;
  c2_adj(a, C.on2e, C.cut2e, C.s1e, C.s2e, C.s3e, Se, Sde, S2e);
  h = v * Se - u * Sde;
  const R hp = vp * Se + v * (R(0.5) * Sde) - up * Sde - R(2) * u * S2e;
  ff += qq * h;
  ffp += qq * hp;
}

// Index of the unordered type pair (ti, tj), ti, tj < T.
__device__ __forceinline__ int type_pair_key(int ti, int tj, int T) {
  const int lo = min(ti, tj), hi = max(ti, tj);
  return lo * T - (lo * (lo - 1)) / 2 + (hi - lo);
}

// MK: number of unordered type pairs held per lane (T (T + 1) / 2 <= MK).
template <bool F64, int ELEC, int NW, int BUF, int MK>
__global__ void __launch_bounds__(64 * NW)
k_tq_adj(ForceParams P, AdjC<float> C32, AdjC<double> C64, Sorted S,
         const TileRep *__restrict__ reps, const int4 *__restrict__ items,
         const int *__restrict__ nitems, const double *__restrict__ grad_f,
         double *__restrict__ grad_x, double *__restrict__ grad_q,
         double *__restrict__ partial_lj, int pstride,
         unsigned long long *__restrict__ npairs) {
  using Real = typename std::conditional<F64, double, float>::type;
  constexpr int KS = kTileAtoms * kAdjSlots;
  __shared__ float4 recf[NW][BUF];       // home-relative x, y, z; charge
  __shared__ int2 reci[NW][BUF];         // sorted index | type, molecule
  __shared__ float4 recg[F64 ? 1 : NW][F64 ? 1 : BUF];  // g_j (FP32)
  __shared__ double4 recd[F64 ? NW : 1][F64 ? BUF : 1];  // FP64 x, y, z, q
  __shared__ double4 recgd[F64 ? NW : 1][F64 ? BUF : 1]; // g_j (FP64)
  __shared__ unsigned short que[NW][kQCap + 2];
  __shared__ float4 ipos[NW][kTileAtoms];
  __shared__ int2 iint[NW][kTileAtoms];
  __shared__ double4 ipd[F64 ? NW : 1][kTileAtoms];
  __shared__ double4 igd[NW][kTileAtoms];  // g_i
  __shared__ Real facc[NW][4 * KS];        // dL/dx (3), dL/dq~ per atom
  __shared__ double ljs64[4 * kQLj];
  AdjC<Real> C;
  if constexpr (F64) C = C64;
  else C = C32;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int T = P.n_types;
  for (int k = threadIdx.x; k < T * T; k += 64 * NW)
    for (int c = 0; c < 4; c++) ljs64[4 * k + c] = S.lj[4 * k + c];
  __syncthreads();
  const F32Params &F = P.f;
  const float hi = F.cut2 + F.band, lo = F.cut2 - F.band;
  const int n_items = *nitems;
  unsigned cnt = 0;
  for (int item_idx = blockIdx.x * NW + wave; item_idx < n_items;
       item_idx += gridDim.x * NW) {
    Real pa[MK][4];  // dL/d{A, B, A14, B14} per unordered type pair
#pragma unroll
    for (int k = 0; k < MK; k++)
      for (int c = 0; c < 4; c++) pa[k][c] = Real(0);
    const int4 item = items[item_idx];
    const int a0 = uni(item.x), a1 = uni(item.y), col = uni(item.z);
    const TileRep &R = reps[uni(item.w)];
    const TileGrid G = R.tg;
    const int coff = R.cell_off;
    const bool wx = R.wrap[0] != 0, wy = R.wrap[1] != 0;
    const float Lx = R.L[0], Ly = R.L[1], Lz = R.L[2];
    const float iLx = R.iL[0], iLy = R.iL[1], iLz = R.iL[2];
    const int na = a1 - a0;
    const int cx = col / G.ny, cy = col % G.ny;
    const bool own = (lane & (kTileAtoms - 1)) < na;
    const int ii = a0 + (own ? (lane & (kTileAtoms - 1)) : 0);
    const int zlo = uni((S.cell[a0] - coff) % G.nz);
    const int zhi = uni((S.cell[a1 - 1] - coff) % G.nz);
    const float4 li = S.loc[ii];
    const float zi = (float)((S.cell[ii] - coff) % G.nz - zlo) * G.csz + li.z;
    if (lane < kTileAtoms) {
      const int ai = S.atom[ii];
      ipos[wave][lane] = make_float4(li.x, li.y, zi, li.w);
      iint[wave][lane] = make_int2(ii | (S.type[ii] << kQTypeShift), S.mol[ii]);
      igd[wave][lane] = make_double4(grad_f[3 * ai], grad_f[3 * ai + 1],
                                     grad_f[3 * ai + 2], 0.0);
      if (F64)
        ipd[wave][lane] = make_double4(S.x[ii], S.y[ii], S.z[ii], S.q[ii]);
    }
    for (int k = lane; k < 4 * KS; k += 64) facc[wave][k] = Real(0);
    const float bx0 = uni(wave_min(own ? li.x : 1e30f)),
                bx1 = uni(wave_max(own ? li.x : -1e30f));
    const float by0 = uni(wave_min(own ? li.y : 1e30f)),
                by1 = uni(wave_max(own ? li.y : -1e30f));
    const float bz0 = uni(wave_min(own ? zi : 1e30f)),
                bz1 = uni(wave_max(own ? zi : -1e30f));
    int zA = zlo - G.rz, zB = zhi + G.rz;
    const bool wz = zB - zA + 1 > G.nz;
    if (wz) zB = zA + G.nz - 1;
    int nbuf = 0;

    auto one = [&](int s, int k) {
      const int2 iq = iint[wave][s];
      const int2 rq = reci[wave][k];
      const int i = iq.x & kQIndexMask, t = rq.x & kQIndexMask;
      if (t == i) return;
      const float4 ip = ipos[wave][s];
      const float4 r = recf[wave][k];
      float dx = r.x - ip.x, dy = r.y - ip.y, dz = r.z - ip.z;
      if (wx) dx = fmaf(-Lx, rintf(dx * iLx), dx);
      if (wy) dy = fmaf(-Ly, rintf(dy * iLy), dy);
      if (wz) dz = fmaf(-Lz, rintf(dz * iLz), dz);
      const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
      Real ex = dx, ey = dy, ez = dz, a = r2, qi, qj;
      double4 gj;
      if (F64) {
        const double4 id = ipd[wave][s], rd = recd[wave][k];
        double e2d;
        double exd, eyd, ezd;
        min_diff_exact(R.box, id.x, id.y, id.z, rd.x, rd.y, rd.z, exd, eyd,
                       ezd, e2d);
        if (e2d > P.cut2) return;
        ex = exd;
        ey = eyd;
        ez = ezd;
        a = e2d;
        qi = id.w;
        qj = rd.w;
        gj = recgd[wave][k];
      } else {
        if (r2 >= lo) {
          double exd, eyd, ezd, e2d;
          min_diff_exact(R.box, S.x[i], S.y[i], S.z[i], S.x[t], S.y[t],
                         S.z[t], exd, eyd, ezd, e2d);
          if (e2d > P.cut2) return;
        }
        qi = ip.w;
        qj = r.w;
        const float4 g4 = recg[wave][k];
        gj = make_double4(g4.x, g4.y, g4.z, 0.0);
      }
      const bool same = rq.y == iq.y;
      const int excl = same ? excl_class(S.xfull, S.xmod, P.max_excl_delta,
                                         S.atom[i], S.atom[t])
                            : 0;
      if (excl == 2) return;
      if (i < t) cnt++;
      const int ti = iq.x >> kQTypeShift, tj = rq.x >> kQTypeShift;
      const double *lp = ljs64 + 4 * (ti * T + tj);
      const Real A = (Real)(excl != 1 ? lp[0] : lp[2]);
      const Real B = (Real)(excl != 1 ? lp[1] : lp[3]);
      const Real fac = excl == 1 ? C.csf : Real(1);
      Real ff, ffp, h, dA, dB;
      pair_adj<Real, ELEC>(C, a, A, B, qi * qj * fac, ff, ffp, h, dA, dB);
      const double4 gi = igd[wave][s];
      const Real wx_ = (Real)(gj.x - gi.x), wy_ = (Real)(gj.y - gi.y),
                 wz_ = (Real)(gj.z - gi.z);
      const Real wd = wx_ * ex + wy_ * ey + wz_ * ez;
      const Real c2 = Real(2) * ffp * wd;
      const Real gxo = -(ff * wx_ + c2 * ex), gyo = -(ff * wy_ + c2 * ey),
                 gzo = -(ff * wz_ + c2 * ez);
      const Real gq = wd * qj * h * fac;
      // LJ parameter gradients (half: every pair is seen from both atoms)
      const int key = type_pair_key(ti, tj, T);
      const int key0 = excl == 1 ? -1 : key, key1 = excl == 1 ? key : -1;
      const Real hA = Real(0.5) * wd * dA, hB = Real(0.5) * wd * dB;
#pragma unroll
      for (int kk = 0; kk < MK; kk++) {
        const bool m0 = kk == key0, m1 = kk == key1;
        pa[kk][0] += m0 ? hA : Real(0);
        pa[kk][1] += m0 ? hB : Real(0);
        pa[kk][2] += m1 ? hA : Real(0);
        pa[kk][3] += m1 ? hB : Real(0);
      }
      Real *fa = &facc[wave][s * kAdjSlots + (lane >> 1)];
      auto add4 = [&]() {
        fa[0] += gxo;
        fa[KS] += gyo;
        fa[2 * KS] += gzo;
        fa[3 * KS] += gq;
      };
      if ((lane & 1) == 0) add4();
      __builtin_amdgcn_sched_barrier(0);
      __asm__ volatile("" ::: "memory");
      if (lane & 1) add4();
    };

    auto eval_queue = [&](int qn) {
      for (int e0 = 0; e0 < qn; e0 += 64) {
        const int e = e0 + lane;
        if (e < qn) {
          const unsigned en = que[wave][e];
          one((int)(en & (kTileAtoms - 1)), (int)(en >> 4));
        }
      }
    };

    const int slot = lane & (kTileAtoms - 1), sub = lane >> 4;
    auto test_group = [&](int c0, int nt, int &qn) {
#pragma unroll 4
      for (int j = 0; j < 16; j++) {
        const int k = c0 + 4 * j + sub;
        const float4 r = recf[wave][k];
        float dx = r.x - li.x, dy = r.y - li.y, dz = r.z - zi;
        if (wx) dx = fmaf(-Lx, rintf(dx * iLx), dx);
        if (wy) dy = fmaf(-Ly, rintf(dy * iLy), dy);
        if (wz) dz = fmaf(-Lz, rintf(dz * iLz), dz);
        const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
        const bool pass = own && k < nt && r2 <= hi;
        const unsigned long long b = __ballot(pass);
        const int pos = pass ? qn + lane_prefix(b) : kQCap;
        que[wave][pos] = (unsigned short)((k << 4) | slot);
        qn += __popcll(b);
      }
    };

    auto process = [&](int n, bool last) -> int {
      const int nt = last ? n : n & ~63;
      wave_sync_lds();
      int qn = 0;
      for (int c0 = 0; c0 < nt; c0 += 64) {
        test_group(c0, nt, qn);
        if (qn > kQCap - kTileAtoms * 64) {
          wave_sync_lds();
          eval_queue(qn);
          qn = 0;
          wave_sync_lds();
        }
      }
      if (qn > 0) {
        wave_sync_lds();
        eval_queue(qn);
      }
      const int rem = n - nt;
      wave_sync_lds();
      if (lane < rem) {
        recf[wave][lane] = recf[wave][nt + lane];
        reci[wave][lane] = reci[wave][nt + lane];
        if (F64) {
          recd[wave][lane] = recd[wave][nt + lane];
          recgd[wave][lane] = recgd[wave][nt + lane];
        } else {
          recg[wave][lane] = recg[wave][nt + lane];
        }
      }
      wave_sync_lds();
      return rem;
    };

    for (int oc = 0; oc < 9; oc++) {
      const int ox = oc / 3 - 1, oy = oc % 3 - 1;
      // an axis with a single column takes it once
      if (wx && ox != 0) continue;
      if (wy && oy != 0) continue;
      const float gx = wx ? 0.f : fmaxf(0.f, fmaxf((float)ox * G.csx - bx1,
                                                  bx0 - (float)(ox + 1) * G.csx));
      const float gy = wy ? 0.f : fmaxf(0.f, fmaxf((float)oy * G.csy - by1,
                                                  by0 - (float)(oy + 1) * G.csy));
      if (fmaf(gx, gx, gy * gy) > hi) continue;
      int tx = cx + ox, ty = cy + oy;
      tx = tx < 0 ? tx + G.nx : (tx >= G.nx ? tx - G.nx : tx);
      ty = ty < 0 ? ty + G.ny : (ty >= G.ny ? ty - G.ny : ty);
      const int tbase = coff + (tx * G.ny + ty) * G.nz;
      const float oxc = (float)ox * G.csx, oyc = (float)oy * G.csy;
      const int wA = ((zA % G.nz) + G.nz) % G.nz;
      const int len = zB - zA + 1;
      for (int part = 0; part < 2; part++) {
        int w0, w1, zshift;
        if (part == 0) {
          w0 = wA;
          w1 = min(G.nz - 1, wA + len - 1);
          zshift = zA - wA;
        } else {
          if (wA + len - 1 < G.nz) break;
          w0 = 0;
          w1 = wA + len - 1 - G.nz;
          zshift = zA - wA + G.nz;
        }
        const int r0 = uni(S.cstart[tbase + w0]);
        const int r1 = uni(S.cstart[tbase + w1] + S.ccount[tbase + w1]);
        for (int base = r0; base < r1; base += 64) {
          const int t = base + lane;
          bool ok = t < r1;
          float4 rf = make_float4(0.f, 0.f, 0.f, 0.f);
          if (ok) {
            const float4 lt = S.loc[t];
            const int zt = (S.cell[t] - coff) % G.nz + zshift;
            rf.x = oxc + lt.x;
            rf.y = oyc + lt.y;
            rf.z = (float)(zt - zlo) * G.csz + lt.z;
            rf.w = lt.w;
            const float gzv = wz ? 0.f : fmaxf(0.f, fmaxf(bz0 - rf.z, rf.z - bz1));
            const float gxv = wx ? 0.f : fmaxf(0.f, fmaxf(bx0 - rf.x, rf.x - bx1));
            const float gyv = wy ? 0.f : fmaxf(0.f, fmaxf(by0 - rf.y, rf.y - by1));
            ok = fmaf(gxv, gxv, fmaf(gyv, gyv, gzv * gzv)) <= hi;
          }
          const unsigned long long ball = __ballot(ok);
          if (ok) {
            const int pos = nbuf + lane_prefix(ball);
            const int at = S.atom[t];
            recf[wave][pos] = rf;
            reci[wave][pos] = make_int2(t | (S.type[t] << kQTypeShift), S.mol[t]);
            if (F64) {
              recd[wave][pos] = make_double4(S.x[t], S.y[t], S.z[t], S.q[t]);
              recgd[wave][pos] = make_double4(grad_f[3 * at], grad_f[3 * at + 1],
                                              grad_f[3 * at + 2], 0.0);
            } else {
              recg[wave][pos] = make_float4((float)grad_f[3 * at],
                                            (float)grad_f[3 * at + 1],
                                            (float)grad_f[3 * at + 2], 0.f);
            }
          }
          nbuf += __popcll(ball);
          if (nbuf > BUF - 64) nbuf = process(nbuf, false);
        }
      }
    }
    if (nbuf > 0) process(nbuf, true);
    // each atom's dL/dx and dL/dq~ from its slots, in a fixed order
    wave_sync_lds();
    if (lane < 4 * kTileAtoms) {
      const int c = lane >> 4, s = lane & (kTileAtoms - 1);
      if (s < na) {
        const Real *fa = &facc[wave][c * KS + s * kAdjSlots];
        const int ai = S.atom[a0 + s];
        double sum = 0.0;
        for (int k = 0; k < kAdjSlots; k++)
          sum += (double)fa[(k + s) & (kAdjSlots - 1)];
        if (c < 3) {
          if (grad_x) grad_x[3 * ai + c] += sum;
        } else if (grad_q) {
          grad_q[ai] += sum;
        }
      }
    }
    // this item's LJ-parameter gradients
#pragma unroll
    for (int k = 0; k < MK; k++)
      for (int c = 0; c < 4; c++) {
        double v = (double)pa[k][c];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
        if (lane == 0) partial_lj[(size_t)(k * 4 + c) * pstride + item_idx] = v;
      }
    wave_sync_lds();
  }
  unsigned long long tot = block_count<NW>(cnt);
  if (threadIdx.x == 0 && tot) atomicAdd(npairs, tot);
}

// Sums the item partials of the LJ-parameter gradients (accumulator kc of
// item b at partial_lj[kc * pstride + b]) with a fixed reduction tree, one
// block of 1024 threads per accumulator, and adds them to grad_lj (T x T x
// 4, {A, B, A14, B14}), half of an unordered off-diagonal pair's gradient to
// each of its two entries.
template <int MK>
__global__ void __launch_bounds__(1024)
k_adj_lj_finish(const double *__restrict__ partial_lj, int pstride,
                const int *__restrict__ nitems, int T,
                double *__restrict__ grad_lj) {
  __shared__ double red[16];
  const int kc = blockIdx.x;  // key * 4 + component
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int n = *nitems;
  const double *col = partial_lj + (size_t)kc * pstride;
  double v = 0.0;
  for (int it = t; it < n; it += 1024) v += col[it];
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  if (lane == 0) red[w] = v;
  __syncthreads();
  if (t != 0) return;
  double sum = 0.0;
  for (int k = 0; k < 16; k++) sum += red[k];
  const int key = kc / 4, c = kc % 4;
  // key -> (lo, hi)
  int lo = 0, rem = key;
  while (lo < T && rem >= T - lo) {
    rem -= T - lo;
    lo++;
  }
  if (lo >= T) return;
  const int hi = lo + rem;
  if (lo == hi) {
    grad_lj[4 * (lo * T + hi) + c] += sum;
  } else {
    grad_lj[4 * (lo * T + hi) + c] += 0.5 * sum;
    grad_lj[4 * (hi * T + lo) + c] += 0.5 * sum;
  }
}

}  // namespace pmnb
