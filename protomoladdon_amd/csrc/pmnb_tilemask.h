// Tile kernel with per-lane pair masks (k_tm), for one system or a batch of
// replicas (TileRep, pmnb_queue.h).
//
// One wave handles 16 consecutive sorted atoms of one column, 4 lanes per
// atom (lane = atom slot + 16 * split).  Candidates of the 3 x 3 neighbour
// columns stream into an LDS buffer one z level at a time (stream_levels),
// after a test of their distance to the wave's bounding box.  When the
// buffer is nearly full each lane tests every 4th buffered candidate against
// its own atom into a pass mask, then evaluates the set bits one by one,
// reading each candidate from LDS one pair ahead of its arithmetic.  The LJ
// table and the exclusion masks of the wave's own atoms are held in LDS.
// Forces stay in registers: FP32 partials per buffer, added into 64-bit
// fixed point (mixed) or FP64 sums, and the 4 lanes of an atom combine by
// shuffles at the end.  Pairs within `band` of the cutoff are re-decided
// with the reference's FP64 minimum image.  Every pair is visited from both
// atoms; energies and virials take half.  Each item writes one partial row
// of energies and virials.
//
// k_tm walks the work items of any number of replicas with a grid stride;
// k_tile is its single-system form.  Against k_tq this trades lane idling (a
// lane with few pairs waits for the lane of its wave with most) for no LDS
// traffic per pair and higher occupancy; the autotuner picks between them.
#pragma once

#include <type_traits>

#include "pmnb_queue.h"

namespace pmnb {

constexpr int kTmSplit = 4;                 // lanes per atom
constexpr int kTmWaves = kBlock / 64;       // waves per block
// Waves per SIMD the register allocation must allow (FP32 / FP64 kernels).
#ifndef PMNB_TM_OCC32
#define PMNB_TM_OCC32 3
#endif
#ifndef PMNB_TM_OCC64
#define PMNB_TM_OCC64 3
#endif
// Candidate records buffered per wave (FP32 / FP64 kernels): a larger
// buffer evens out the lanes' pair counts per buffer; 3 blocks per CU must
// still fit in LDS (32 / 64 bytes per record, 4 waves per block).
#ifndef PMNB_TM_BUF32
#define PMNB_TM_BUF32 256
#endif
#ifndef PMNB_TM_BUF64
#define PMNB_TM_BUF64 192
#endif
constexpr int kTmBuf32 = PMNB_TM_BUF32, kTmBuf64 = PMNB_TM_BUF64;

// Per-lane pass masks over the buffered records split + 4 j, j < jmax, in
// NW 32-bit words: word w holds j = 32 w .. 32 w + cw - 1 (cw = min(32,
// jmax - 32 w)) with record j at bit cw - 1 - (j - 32 w), so the test
// shifts its bits in.  Records n .. 4 jmax - 1 hold a far sentinel; the
// atom itself passes and is skipped by the pair code.
// (Both walk the words by compile-time recursion, so the masks stay in
// registers.)
template <int W, int NW, class Test>
__device__ __forceinline__ void mask_test(int jmax, unsigned (&m)[NW],
                                          Test &test) {
  unsigned mm = 0u;
  const int j1 = min(jmax, 32 * W + 32);
#pragma unroll 4
  for (int j = 32 * W; j < j1; j++) mm = (mm << 1) + test(j);
  m[W] = mm;
  if constexpr (W + 1 < NW) mask_test<W + 1, NW>(jmax, m, test);
}

// Record index of the lowest set bit of the first non-empty word, which is
// cleared.  At least one bit must be set, except in the two- and
// three-word SELECT forms, which leave an empty mask empty and return an
// unusable index.
template <int W, int NW, bool SELECT = true>
__device__ __forceinline__ int mask_pop(unsigned (&m)[NW], int split,
                                        int jmax) {
  if constexpr (SELECT && W == 0 && NW == 2) {
    // two words with selects instead of branches
    const bool first = m[0] != 0u;
    unsigned w = first ? m[0] : m[1];
    const int b = __builtin_ctz(w);
    w &= w - 1u;
    m[0] = first ? w : m[0];
    m[1] = first ? m[1] : w;
    return split + kTmSplit * ((first ? min(32, jmax) : jmax) - 1 - b);
  } else if constexpr (SELECT && W == 0 && NW == 3) {
    const bool w0 = m[0] != 0u, w1 = !w0 && m[1] != 0u, w2 = !w0 && !w1;
    unsigned w = w0 ? m[0] : (w1 ? m[1] : m[2]);
    const int b = __builtin_ctz(w);
    w &= w - 1u;
    m[0] = w0 ? w : m[0];
    m[1] = w1 ? w : m[1];
    m[2] = w2 ? w : m[2];
    const int top = w0 ? min(32, jmax) : (w1 ? min(64, jmax) : jmax);
    return split + kTmSplit * (top - 1 - b);
  } else {
    if constexpr (W + 1 < NW) {
      if (!m[W]) return mask_pop<W + 1, NW, SELECT>(m, split, jmax);
    }
    const int b = __builtin_ctz(m[W]);
    m[W] &= m[W] - 1u;
    return split + kTmSplit * (32 * W + min(32, jmax - 32 * W) - 1 - b);
  }
}

template <int NW>
__device__ __forceinline__ bool mask_any(const unsigned (&m)[NW]) {
  unsigned a = m[0];
  if constexpr (NW > 1) a |= m[1];
  if constexpr (NW > 2) a |= m[2];
  if constexpr (NW > 3) a |= m[3];
  static_assert(NW <= 4, "at most 4 mask words");
  return a != 0u;
}

// Tests the buffered records against one atom (mask_test layout).
// WMODE 0: no minimum image, 1: along z, 2: along x, y and z, 3: along the
// axes whose runtime flag wx, wy, wz is set.
template <int WMODE, int NW>
__device__ __forceinline__ void tm_test(const float4 *__restrict__ rec, int n,
                                        int split, float xi, float yi,
                                        float zi, float hi, const float *L,
                                        const float *iL, bool wx, bool wy,
                                        bool wz, unsigned (&m)[NW]) {
  const bool WX = WMODE == 2 || (WMODE == 3 && wx);
  const bool WY = WMODE == 2 || (WMODE == 3 && wy);
  const bool WZ = WMODE == 1 || WMODE == 2 || (WMODE == 3 && wz);
  auto test = [&](int j) -> unsigned {
    const float4 r = rec[split + kTmSplit * j];
    float dx = r.x - xi, dy = r.y - yi, dz = r.z - zi;
    if (WX) dx = fmaf(-L[0], rintf(dx * iL[0]), dx);
    if (WY) dy = fmaf(-L[1], rintf(dy * iL[1]), dy);
    if (WZ) dz = fmaf(-L[2], rintf(dz * iL[2]), dz);
    const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
    return (unsigned)(r2 <= hi);
  };
  mask_test<0, NW>((n + kTmSplit - 1) / kTmSplit, m, test);
}

// Streams the candidate records of a wave's neighbour columns one z level at
// a time, the wave's own levels first and then alternately one below and one
// above.  A buffer then holds candidates from every direction around the
// wave's atoms, so the atoms' pair counts per buffer are alike and the lanes
// of the mask loop wait less for one another (a buffer filled column by
// column holds the candidates of one side).  Lane l < 9 keeps neighbour
// column (l / 3 - 1, l % 3 - 1); a level's cells of all passing columns are
// one virtual run, walked 64 records per call of
//   emit(ok, t, ox, oy, z0): sorted record t (valid if ok) with column image
//   offsets ox, oy and level base z0, all relative to the wave's home cell.
// wx / wy: single column on that axis (taken once, per-pair image); wz: the
// levels zA..zB cover every z cell once (per-pair image along z).
template <class Emit>
__device__ __forceinline__ void stream_levels(
    const Sorted &S, const TileGrid &G, int coff, int cx, int cy, bool wx,
    bool wy, bool wz, int zlo, int zhi, int zA, int zB, float bx0, float bx1,
    float by0, float by1, float bz0, float bz1, float hi, int lane,
    Emit &&emit) {
  const int ox = lane / 3 - 1, oy = lane % 3 - 1;
  const float gx = wx ? 0.f
                      : fmaxf(0.f, fmaxf((float)ox * G.csx - bx1,
                                         bx0 - (float)(ox + 1) * G.csx));
  const float gy = wy ? 0.f
                      : fmaxf(0.f, fmaxf((float)oy * G.csy - by1,
                                         by0 - (float)(oy + 1) * G.csy));
  const float gxy2 = fmaf(gx, gx, gy * gy);
  const bool cok = lane < 9 && !(wx && ox != 0) && !(wy && oy != 0) &&
                   gxy2 <= hi;
  int tx = cx + ox, ty = cy + oy;
  tx = tx < 0 ? tx + G.nx : (tx >= G.nx ? tx - G.nx : tx);
  ty = ty < 0 ? ty + G.ny : (ty >= G.ny ? ty - G.ny : ty);
  const int tbase = coff + (tx * G.ny + ty) * G.nz;
  const int ncz = zhi - zlo + 1;
  const int nlev = ncz + 2 * G.rz;
  auto level_z = [&](int q) {
    if (q < ncz) return zlo + q;
    const int d = q - ncz, k = (d >> 1) + 1;
    return (d & 1) ? zhi + k : zlo - k;
  };
  // this lane's cell of level zc: record count and first sorted record
  auto level_cell = [&](int zc, int &cnt, int &start) {
    const int w = ((zc % G.nz) + G.nz) % G.nz;
    const float z0 = (float)(zc - zlo) * G.csz;
    // level gap to the wave box, less a margin for the rounding of the
    // cell-local offsets, so the per-record test alone decides
    float gz = wz ? 0.f : fmaxf(0.f, fmaxf(z0 - bz1, bz0 - (z0 + G.csz)));
    gz = fmaxf(0.f, gz - 1e-3f);
    const bool lok = cok && fmaf(gz, gz, gxy2) <= hi;
    cnt = lok ? S.ccount[tbase + w] : 0;
    start = lok ? S.cstart[tbase + w] : 0;
  };
  // level q = 0 (zlo) always lies in zA .. zB; the next level's cells are
  // read before the current level is walked
  int q = 0, zc = zlo, cnt, start;
  level_cell(zc, cnt, start);
  while (true) {
    int qn = q + 1, zn = 0;
    for (; qn < nlev; qn++) {
      zn = level_z(qn);
      if (zn >= zA && zn <= zB) break;
    }
    int cnt_n = 0, start_n = 0;
    if (qn < nlev) level_cell(zn, cnt_n, start_n);
    // inclusive prefix of the counts over lanes 0 .. 8 (row shifts within
    // the first 16 lanes)
    int inc = cnt;
    inc += __builtin_amdgcn_update_dpp(0, inc, 0x111, 0xf, 0xf, false);
    inc += __builtin_amdgcn_update_dpp(0, inc, 0x112, 0xf, 0xf, false);
    inc += __builtin_amdgcn_update_dpp(0, inc, 0x114, 0xf, 0xf, false);
    inc += __builtin_amdgcn_update_dpp(0, inc, 0x118, 0xf, 0xf, false);
    const int total = __builtin_amdgcn_readlane(inc, 8);
    if (total > 0) {
      const float z0 = (float)(zc - zlo) * G.csz;
      const int rel = start - (inc - cnt);  // record = run index + rel
      int ends[8];
#pragma unroll
      for (int j = 0; j < 8; j++) ends[j] = __builtin_amdgcn_readlane(inc, j);
      for (int base = 0; base < total; base += 64) {
        const int idx = base + lane;
        int k = 0;
#pragma unroll
        for (int j = 0; j < 8; j++) k += idx >= ends[j] ? 1 : 0;
        const int t = idx + __shfl(rel, k, 64);
        emit(idx < total, t, (float)(k / 3 - 1) * G.csx,
             (float)(k % 3 - 1) * G.csy, z0);
      }
    }
    if (qn >= nlev) break;
    q = qn;
    zc = zn;
    cnt = cnt_n;
    start = start_n;
  }
}

template <bool F64, int ELEC, int ELSW, int BUF>
__global__ void __launch_bounds__(kBlock)
__attribute__((amdgpu_waves_per_eu(F64 ? PMNB_TM_OCC64 : PMNB_TM_OCC32)))
k_tm(ForceParams P, Sorted S, const TileRep *__restrict__ reps,
     const int4 *__restrict__ items, const int *__restrict__ nitems,
     double *__restrict__ forces, double *__restrict__ partial, int pstride,
     unsigned long long *__restrict__ npairs) {
  constexpr int kNW = (BUF / kTmSplit + 31) / 32;  // mask words
  __shared__ float4 recf[kTmWaves][BUF];
  __shared__ int4 reci[kTmWaves][BUF];   // sorted index, type, mol, atom
  __shared__ double4 recd[F64 ? kTmWaves : 1][F64 ? BUF : 1];
  // exclusion masks of the wave's own atoms (excl_own layout)
  __shared__ unsigned long long xm[kTmWaves][kTileAtoms][4];
  __shared__ float4 ljs[F64 ? 1 : kQLj];  // FP32 LJ table (T * T <= kQLj)
  if (!F64) {
    for (int k = threadIdx.x; k < P.n_types * P.n_types; k += kBlock)
      ljs[k] = S.ljf[k];
    __syncthreads();
  }
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const F32Params &F = P.f;
  const float hi = F.cut2 + F.band, lo = F.cut2 - F.band;
  const bool vir = (P.flags & 3u) != 0;
  const int n_items = *nitems;
  unsigned cnt = 0;
  for (int item_idx = blockIdx.x * kTmWaves + wave; item_idx < n_items;
       item_idx += gridDim.x * kTmWaves) {
    double acc[ACC_N];
#pragma unroll
    for (int k = 0; k < ACC_N; k++) acc[k] = 0.0;
    const int4 item = items[item_idx];
    const int a0 = uni(item.x), a1 = uni(item.y), col = uni(item.z);
    const TileRep &R = reps[uni(item.w)];
    const TileGrid G = R.tg;
    const int coff = R.cell_off;
    const bool wx = R.wrap[0] != 0, wy = R.wrap[1] != 0;
    float L[3], iL[3];
    for (int k = 0; k < 3; k++) {
      L[k] = R.L[k];
      iL[k] = R.iL[k];
    }
    const int cx = col / G.ny, cy = col % G.ny;
    const int slot = lane & (kTileAtoms - 1), split = lane >> 4;
    const int i = a0 + slot;
    const bool active = i < a1;
    const int ii = active ? i : a0;
    const int zlo = uni((S.cell[a0] - coff) % G.nz);
    const int zhi = uni((S.cell[a1 - 1] - coff) % G.nz);
    const float4 li = S.loc[ii];
    const float xi = li.x, yi = li.y;
    const float zi = (float)((S.cell[ii] - coff) % G.nz - zlo) * G.csz + li.z;
    const int ti = S.type[ii], mi = S.mol[ii], ai = S.atom[ii];
    if (split == 0) {
#pragma unroll
      for (int k = 0; k < 4; k++) xm[wave][slot][k] = S.xown[4 * ai + k];
    }
    wave_sync_lds();
    const double qi = S.q[ii];
    const double xd = S.x[ii], yd = S.y[ii], zd = S.z[ii];
    const int ljbase = ti * P.n_types;  // row of this atom's type in ljs
    const double *ljrow64 = S.lj + 4 * (ti * P.n_types);
    // bounding box of the wave's atoms in home-relative coordinates
    const float bx0 = (wave_min(active ? xi : 1e30f)),
                bx1 = (wave_max(active ? xi : -1e30f));
    const float by0 = (wave_min(active ? yi : 1e30f)),
                by1 = (wave_max(active ? yi : -1e30f));
    const float bz0 = (wave_min(active ? zi : 1e30f)),
                bz1 = (wave_max(active ? zi : -1e30f));
    int zA = zlo - G.rz, zB = zhi + G.rz;
    const bool wz = zB - zA + 1 > G.nz;  // all z cells: per-pair image
    if (wz) zB = zA + G.nz - 1;
    long long Fx = 0, Fy = 0, Fz = 0;
    double fx64 = 0.0, fy64 = 0.0, fz64 = 0.0;
    int nbuf = 0;

    // FP32 partials of one buffer (flushed by process)
    float px = 0.f, py = 0.f, pz = 0.f, pelj = 0.f, peel = 0.f;
    float v0 = 0.f, v1 = 0.f, v2 = 0.f, v3 = 0.f, v4 = 0.f, v5 = 0.f;

    // One pair: candidate record k of this wave's buffer.
    // WM: minimum-image mode as in tm_test (a compile-time tag, so the
    // common modes carry no selects).
    auto eval = [&](const float4 &r, const int4 &q4, int k,
                    auto WM) __attribute__((always_inline)) {
      constexpr int M = decltype(WM)::value;
      const int t = q4.x;
      if (t == i) return;
      float dx = r.x - xi, dy = r.y - yi, dz = r.z - zi;
      if (M == 2 || (M == 3 && wx)) dx = fmaf(-L[0], rintf(dx * iL[0]), dx);
      if (M == 2 || (M == 3 && wy)) dy = fmaf(-L[1], rintf(dy * iL[1]), dy);
      if (M == 1 || M == 2 || (M == 3 && wz))
        dz = fmaf(-L[2], rintf(dz * iL[2]), dz);
      const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
      double ex = 0.0, ey = 0.0, ez = 0.0, e2d = 0.0;
      if (F64) {
        const double4 rd = recd[wave][k];
        min_diff_exact(R.box, xd, yd, zd, rd.x, rd.y, rd.z, ex, ey, ez, e2d);
        if (e2d > P.cut2) return;
      } else if (r2 >= lo) {
        // near the cutoff: decide with the reference's FP64 minimum image
        min_diff_exact(rare_box(&R.box), xd, yd, zd, S.x[t], S.y[t], S.z[t],
                       ex, ey, ez, e2d);
        if (e2d > P.cut2) return;
      }
      const bool same = q4.z == mi;
      const int excl =
          same ? excl_own(xm[wave][slot], q4.w - ai) : 0;
      if (excl == 2) return;
      if (i < t) cnt++;
      const bool molv = (P.flags & 2u) && !same;
      if (F64) {
        const double *lp = ljrow64 + 4 * q4.y;
        double ff, e1, e2;
        pair_f64<ELEC, ELSW>(P, e2d, excl != 1 ? lp[0] : lp[2],
                             excl != 1 ? lp[1] : lp[3], qi * recd[wave][k].w,
                             excl, ff, e1, e2);
        acc[ACC_ELJ] += 0.5 * e1;
        acc[ACC_EEL] += 0.5 * e2;
        const double gx = ex * ff, gy = ey * ff, gz = ez * ff;
        fx64 -= gx;
        fy64 -= gy;
        fz64 -= gz;
        if (molv || (P.flags & 1u))
          add_virial(acc, ACC_VXX, gx, gy, gz, ex, ey, ez, 0.5);
        if (molv) {
          double cdx, cdy, cdz, cd2;
          min_diff_exact(rare_box(&R.box), S.com[3 * mi], S.com[3 * mi + 1], S.com[3 * mi + 2], S.com[3 * q4.z], S.com[3 * q4.z + 1], S.com[3 * q4.z + 2],
                         cdx, cdy, cdz, cd2);
          add_virial(acc, ACC_MXX, gx, gy, gz, cdx, cdy, cdz, 0.5);
        }
      } else {
        const float4 lp = ljs[ljbase + q4.y];
        float ff, e1, e2;
        pair_f32<ELEC>(F, r2, excl != 1 ? lp.x : lp.z, excl != 1 ? lp.y : lp.w,
                       li.w * r.w, excl, ff, e1, e2);
        pelj += e1;
        peel += e2;
        const float gx = dx * ff, gy = dy * ff, gz = dz * ff;
        px -= gx;
        py -= gy;
        pz -= gz;
        if (vir) {
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
            min_diff_exact(rare_box(&R.box), S.com[3 * mi], S.com[3 * mi + 1], S.com[3 * mi + 2], S.com[3 * q4.z], S.com[3 * q4.z + 1], S.com[3 * q4.z + 2],
                           cdx, cdy, cdz, cd2);
            add_virial(acc, ACC_MXX, gx, gy, gz, cdx, cdy, cdz, 0.5);
          }
        }
      }
    };

    const int wmode = (wx ? 1 : 0) | (wy ? 2 : 0) | (wz ? 4 : 0);
    // Tests the buffered candidates against this lane's atom (the atom
    // itself has r2 == 0 and is left out), then evaluates the pairs that
    // pass; the next pair's records are read from LDS before the current
    // pair's arithmetic.
    auto process = [&](int n) __attribute__((always_inline)) {
      // records n .. n + 3 (the last group of 4) are made far away
      if (lane < kTmSplit && n + lane < BUF)
        recf[wave][n + lane] = make_float4(1e18f, 1e18f, 1e18f, 0.f);
      wave_sync_lds();
      unsigned m[kNW];
      const float4 *rb = recf[wave];
      if (wmode == 0)
        tm_test<0>(rb, n, split, xi, yi, zi, hi, L, iL, wx, wy, wz, m);
      else if (wmode == 4)
        tm_test<1>(rb, n, split, xi, yi, zi, hi, L, iL, wx, wy, wz, m);
      else if (wmode == 7)
        tm_test<2>(rb, n, split, xi, yi, zi, hi, L, iL, wx, wy, wz, m);
      else
        tm_test<3>(rb, n, split, xi, yi, zi, hi, L, iL, wx, wy, wz, m);
      if (!active) {
#pragma unroll
        for (int w = 0; w < kNW; w++) m[w] = 0u;
      }
      const int jmax = (n + kTmSplit - 1) / kTmSplit;
      px = py = pz = pelj = peel = 0.f;
      v0 = v1 = v2 = v3 = v4 = v5 = 0.f;
      auto run = [&](auto WM) __attribute__((always_inline)) {
        if (!mask_any(m)) return;
        if constexpr (F64) {
          int k = mask_pop<0, kNW, false>(m, split, jmax);
          float4 r = recf[wave][k];
          int4 q4 = reci[wave][k];
          while (true) {
            const bool more = mask_any(m);
            int kn = k;
            if (more) kn = mask_pop<0, kNW, false>(m, split, jmax);
            const float4 rn = recf[wave][kn];
            const int4 q4n = reci[wave][kn];
            eval(r, q4, k, WM);
            if (!more) break;
            r = rn;
            q4 = q4n;
            k = kn;
          }
        } else {
          // two record sets in turn (no register moves); the next pop is
          // taken unconditionally and its index replaced by 0 when the
          // mask was empty, so every LDS read stays in the buffer
          int k0 = mask_pop<0, kNW, true>(m, split, jmax);
          float4 r0 = recf[wave][k0];
          int4 q0 = reci[wave][k0];
          while (true) {
            const bool more1 = mask_any(m);
            int k1 = mask_pop<0, kNW, true>(m, split, jmax);
            k1 = more1 ? k1 : 0;
            const float4 r1 = recf[wave][k1];
            const int4 q1 = reci[wave][k1];
            eval(r0, q0, k0, WM);
            if (!more1) break;
            const bool more0 = mask_any(m);
            k0 = mask_pop<0, kNW, true>(m, split, jmax);
            k0 = more0 ? k0 : 0;
            r0 = recf[wave][k0];
            q0 = reci[wave][k0];
            eval(r1, q1, k1, WM);
            if (!more0) break;
          }
        }
      };
      // FP64 keeps one loop with runtime flags (fewer spilled registers)
      if (F64 || (wmode != 0 && wmode != 4 && wmode != 7))
        run(std::integral_constant<int, 3>{});
      else if (wmode == 0)
        run(std::integral_constant<int, 0>{});
      else if (wmode == 4)
        run(std::integral_constant<int, 1>{});
      else
        run(std::integral_constant<int, 2>{});
      if (!F64) {
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
      wave_sync_lds();  // buffer is refilled next
    };

    auto emit = [&](bool ok, int t, float oxc, float oyc,
                    float z0) __attribute__((always_inline)) {
      float4 rf = make_float4(0.f, 0.f, 0.f, 0.f);
      int4 ri = make_int4(0, 0, 0, 0);
      if (ok) {
        const float4 lt = S.loc[t];
        ri = S.rec[t];
        rf.x = oxc + lt.x;
        rf.y = oyc + lt.y;
        rf.z = z0 + lt.z;
        rf.w = lt.w;
        const float gzv = wz ? 0.f : fmaxf(0.f, fmaxf(bz0 - rf.z, rf.z - bz1));
        const float gxv = wx ? 0.f : fmaxf(0.f, fmaxf(bx0 - rf.x, rf.x - bx1));
        const float gyv = wy ? 0.f : fmaxf(0.f, fmaxf(by0 - rf.y, rf.y - by1));
        ok = fmaf(gxv, gxv, fmaf(gyv, gyv, gzv * gzv)) <= hi;
      }
      const unsigned long long ball = __ballot(ok);
      if (ok) {
        const int pos = nbuf + lane_prefix(ball);
        recf[wave][pos] = rf;
        reci[wave][pos] = ri;
        if (F64)
          recd[wave][pos] = make_double4(S.x[t], S.y[t], S.z[t], S.q[t]);
      }
      nbuf += __popcll(ball);
      if (nbuf > BUF - 64) {
        process(nbuf);
        nbuf = 0;
      }
    };
    stream_levels(S, G, coff, cx, cy, wx, wy, wz, zlo, zhi, zA, zB, bx0, bx1,
                  by0, by1, bz0, bz1, hi, lane, emit);
    if (nbuf > 0) process(nbuf);
    // the 4 lanes of an atom add their partial forces
    if (F64) {
      fx64 += __shfl_down(fx64, 32, 64);
      fy64 += __shfl_down(fy64, 32, 64);
      fz64 += __shfl_down(fz64, 32, 64);
      fx64 += __shfl_down(fx64, 16, 64);
      fy64 += __shfl_down(fy64, 16, 64);
      fz64 += __shfl_down(fz64, 16, 64);
      if (active && split == 0) {
        forces[3 * ai] += fx64;
        forces[3 * ai + 1] += fy64;
        forces[3 * ai + 2] += fz64;
      }
    } else {
      Fx += __shfl_down(Fx, 32, 64);
      Fy += __shfl_down(Fy, 32, 64);
      Fz += __shfl_down(Fz, 32, 64);
      Fx += __shfl_down(Fx, 16, 64);
      Fy += __shfl_down(Fy, 16, 64);
      Fz += __shfl_down(Fz, 16, 64);
      if (active && split == 0) {
        forces[3 * ai] += (double)Fx * kFixedInv;
        forces[3 * ai + 1] += (double)Fy * kFixedInv;
        forces[3 * ai + 2] += (double)Fz * kFixedInv;
      }
    }
    // this item's energies and virials
#pragma unroll
    for (int k = 0; k < ACC_N; k++) {
      double v = acc[k];
      for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
      if (lane == 0) partial[(size_t)k * pstride + item_idx] = v;
    }
  }
  unsigned long long tot = block_count<kTmWaves>(cnt);
  if (threadIdx.x == 0 && tot) atomicAdd(npairs, tot);
}

// Tests the buffered records against one atom (mask_test layout).  WRAPZ:
// minimum image along z.
template <bool WRAPZ, int NW>
__device__ __forceinline__ void tile_test(const float4 *__restrict__ rec,
                                          int n, int split, float xi,
                                          float yi, float zi, float hi,
                                          float Lz, float iLz,
                                          unsigned (&m)[NW]) {
  auto test = [&](int j) -> unsigned {
    const float4 r = rec[split + kTmSplit * j];
    const float dx = r.x - xi, dy = r.y - yi;
    float dz = r.z - zi;
    if (WRAPZ) dz = fmaf(-Lz, rintf(dz * iLz), dz);
    const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
    return (unsigned)(r2 <= hi);
  };
  mask_test<0, NW>((n + kTmSplit - 1) / kTmSplit, m, test);
}

// Single-system form of k_tm: one wave per item, the column grid as a
// kernel argument, minimum image along z only when the cutoff reaches every
// z cell (x and y need at least two columns).  Same per-lane work order and
// per-item partial rows as k_tm, so both give the same bits; without a
// replica table it needs fewer scalar registers, which measures faster for
// one large system.
template <bool F64, int ELEC, int ELSW, int BUF>
__global__ void __launch_bounds__(kBlock)
__attribute__((amdgpu_waves_per_eu(F64 ? PMNB_TM_OCC64 : PMNB_TM_OCC32)))
k_tile(ForceParams P, Sorted S, TileGrid G, const int4 *__restrict__ items,
       const int *__restrict__ nitems, double *__restrict__ forces,
       double *__restrict__ partial, int pstride,
       unsigned long long *__restrict__ npairs) {
  constexpr int kNW = (BUF / kTmSplit + 31) / 32;  // mask words
  __shared__ float4 recf[kTmWaves][BUF];
  __shared__ int4 reci[kTmWaves][BUF];
  __shared__ double4 recd[F64 ? kTmWaves : 1][F64 ? BUF : 1];
  // exclusion masks of the wave's own atoms (excl_own layout)
  __shared__ unsigned long long xm[kTmWaves][kTileAtoms][4];
  __shared__ float4 ljs[F64 ? 1 : kQLj];  // FP32 LJ table (T * T <= kQLj)
  if (!F64) {
    for (int k = threadIdx.x; k < P.n_types * P.n_types; k += kBlock)
      ljs[k] = S.ljf[k];
    __syncthreads();
  }
  // FP64: values equal on all lanes of the wave are marked uniform (uni), so
  // the work item, its z range and box live in scalar registers (in the
  // FP32 kernel the extra scalar registers spill and cost more than they
  // save)
  auto U = [](auto v) { return F64 ? uni(v) : v; };
  const int wave = U((int)(threadIdx.x >> 6)), lane = threadIdx.x & 63;
  // FP64: blocks go to the 8 XCDs in turn; block b takes the items of block
  // position (b % 8) * per + b / 8 (the grid is a multiple of 8), so each
  // XCD works through one contiguous range of items and the candidates that
  // neighbouring items share stay in that XCD's L2 (measured faster for the
  // FP64 kernel, slightly slower for the FP32 one).
  const int per = gridDim.x / 8;
  const int blk =
      F64 ? (int)(blockIdx.x % 8) * per + (int)(blockIdx.x / 8) : blockIdx.x;
  const int item_idx = blk * kTmWaves + wave;
  double acc[ACC_N];
#pragma unroll
  for (int k = 0; k < ACC_N; k++) acc[k] = 0.0;
  unsigned cnt = 0;
  if (item_idx < *nitems) {
    const int4 item = items[item_idx];
    const int a0 = U(item.x), a1 = U(item.y), col = U(item.z);
    const int cx = col / G.ny, cy = col % G.ny;
    const int slot = lane & (kTileAtoms - 1), split = lane >> 4;
    const int i = a0 + slot;
    const bool active = i < a1;
    const int ii = active ? i : a0;
    const F32Params &F = P.f;
    const float hi = F.cut2 + F.band, lo = F.cut2 - F.band;
    const int zlo = U(S.cell[a0] % G.nz), zhi = U(S.cell[a1 - 1] % G.nz);
    const float4 li = S.loc[ii];
    const int zci = S.cell[ii] % G.nz;
    const float xi = li.x, yi = li.y;
    const float zi = (float)(zci - zlo) * G.csz + li.z;
    const int ti = S.type[ii], mi = S.mol[ii], ai = S.atom[ii];
    if (split == 0) {
#pragma unroll
      for (int k = 0; k < 4; k++) xm[wave][slot][k] = S.xown[4 * ai + k];
    }
    wave_sync_lds();
    const double qi = S.q[ii];
    const double xd = S.x[ii], yd = S.y[ii], zd = S.z[ii];
    const int ljbase = ti * P.n_types;  // row of this atom's type in ljs
    const double *ljrow64 = S.lj + 4 * (ti * P.n_types);
    // bounding box of the wave's atoms in home-relative coordinates
    const float bx0 = U(wave_min(active ? xi : 1e30f)),
                bx1 = U(wave_max(active ? xi : -1e30f));
    const float by0 = U(wave_min(active ? yi : 1e30f)),
                by1 = U(wave_max(active ? yi : -1e30f));
    const float bz0 = U(wave_min(active ? zi : 1e30f)),
                bz1 = U(wave_max(active ? zi : -1e30f));
    int zA = zlo - G.rz, zB = zhi + G.rz;
    const bool clampz = zB - zA + 1 > G.nz;
    if (clampz) zB = zA + G.nz - 1;
    const float Lz = F.Lz, iLz = F.iLz;
    long long Fx = 0, Fy = 0, Fz = 0;
    double fx64 = 0.0, fy64 = 0.0, fz64 = 0.0;
    const bool vir = (P.flags & 3u) != 0;
    int nbuf = 0;

    // FP32 partials of one buffer (flushed by process)
    float px = 0.f, py = 0.f, pz = 0.f, pelj = 0.f, peel = 0.f;
    float v0 = 0.f, v1 = 0.f, v2 = 0.f, v3 = 0.f, v4 = 0.f, v5 = 0.f;

    // One pair: candidate record k of this wave's buffer.
    // WZ: minimum image along z, 0 no, 1 yes, 2 as the runtime flag (a
    // compile-time tag, so the common case carries no select).
    auto eval = [&](const float4 &r, const int4 &q4, int k,
                    auto WZ) __attribute__((always_inline)) {
      constexpr int M = decltype(WZ)::value;
      const int t = q4.x;
      if (t == i) return;
      float dx = r.x - xi, dy = r.y - yi, dz = r.z - zi;
      if (M == 1 || (M == 2 && clampz)) dz = fmaf(-Lz, rintf(dz * iLz), dz);
      const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
      double ex = 0.0, ey = 0.0, ez = 0.0, e2d = 0.0;
      if (F64) {
        const double4 rd = recd[wave][k];
        min_diff_exact(*S.box, xd, yd, zd, rd.x, rd.y, rd.z, ex, ey, ez, e2d);
        if (e2d > P.cut2) return;
      } else if (r2 >= lo) {
        // near the cutoff: decide with the reference's FP64 minimum image
        min_diff_exact(*S.box, xd, yd, zd, S.x[t], S.y[t], S.z[t], ex, ey, ez,
                       e2d);
        if (e2d > P.cut2) return;
      }
      const bool same = q4.z == mi;
      const int excl =
          same ? excl_own(xm[wave][slot], q4.w - ai) : 0;
      if (excl == 2) return;
      if (i < t) cnt++;
      const bool molv = (P.flags & 2u) && !same;
      if (F64) {
        const double *lp = ljrow64 + 4 * q4.y;
        double ff, e1, e2;
        pair_f64<ELEC, ELSW>(P, e2d, excl != 1 ? lp[0] : lp[2],
                             excl != 1 ? lp[1] : lp[3], qi * recd[wave][k].w,
                             excl, ff, e1, e2);
        acc[ACC_ELJ] += 0.5 * e1;
        acc[ACC_EEL] += 0.5 * e2;
        const double gx = ex * ff, gy = ey * ff, gz = ez * ff;
        fx64 -= gx;
        fy64 -= gy;
        fz64 -= gz;
        if (molv || (P.flags & 1u))
          add_virial(acc, ACC_VXX, gx, gy, gz, ex, ey, ez, 0.5);
        if (molv) {
          double cdx, cdy, cdz, cd2;
          min_diff_exact(*S.box, S.com[3 * mi], S.com[3 * mi + 1],
                         S.com[3 * mi + 2], S.com[3 * q4.z],
                         S.com[3 * q4.z + 1], S.com[3 * q4.z + 2], cdx, cdy,
                         cdz, cd2);
          add_virial(acc, ACC_MXX, gx, gy, gz, cdx, cdy, cdz, 0.5);
        }
      } else {
        const float4 lp = ljs[ljbase + q4.y];
        float ff, e1, e2;
        pair_f32<ELEC>(F, r2, excl != 1 ? lp.x : lp.z, excl != 1 ? lp.y : lp.w,
                       li.w * r.w, excl, ff, e1, e2);
        pelj += e1;
        peel += e2;
        const float gx = dx * ff, gy = dy * ff, gz = dz * ff;
        px -= gx;
        py -= gy;
        pz -= gz;
        if (vir) {
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
            min_diff_exact(*S.box, S.com[3 * mi], S.com[3 * mi + 1],
                           S.com[3 * mi + 2], S.com[3 * q4.z],
                           S.com[3 * q4.z + 1], S.com[3 * q4.z + 2], cdx, cdy,
                           cdz, cd2);
            add_virial(acc, ACC_MXX, gx, gy, gz, cdx, cdy, cdz, 0.5);
          }
        }
      }
    };

    // Tests the buffered candidates against this lane's atom (the atom
    // itself has r2 == 0 and is left out), then evaluates the pairs that
    // pass; the next pair's records are read from LDS before the current
    // pair's arithmetic.
    auto process = [&](int n) __attribute__((always_inline)) {
      // records n .. n + 3 (the last group of 4) are made far away
      if (lane < kTmSplit && n + lane < BUF)
        recf[wave][n + lane] = make_float4(1e18f, 1e18f, 1e18f, 0.f);
      wave_sync_lds();
      unsigned m[kNW];
      if (clampz)
        tile_test<true>(recf[wave], n, split, xi, yi, zi, hi, Lz, iLz, m);
      else
        tile_test<false>(recf[wave], n, split, xi, yi, zi, hi, Lz, iLz, m);
      if (!active) {
#pragma unroll
        for (int w = 0; w < kNW; w++) m[w] = 0u;
      }
      const int jmax = (n + kTmSplit - 1) / kTmSplit;
      px = py = pz = pelj = peel = 0.f;
      v0 = v1 = v2 = v3 = v4 = v5 = 0.f;
      auto run = [&](auto WZ) __attribute__((always_inline)) {
        if (!mask_any(m)) return;
        if constexpr (F64) {
          int k = mask_pop<0, kNW, false>(m, split, jmax);
          float4 r = recf[wave][k];
          int4 q4 = reci[wave][k];
          while (true) {
            const bool more = mask_any(m);
            int kn = k;
            if (more) kn = mask_pop<0, kNW, false>(m, split, jmax);
            const float4 rn = recf[wave][kn];
            const int4 q4n = reci[wave][kn];
            eval(r, q4, k, WZ);
            if (!more) break;
            r = rn;
            q4 = q4n;
            k = kn;
          }
        } else {
          // two record sets in turn (no register moves); the next pop is
          // taken unconditionally and its index replaced by 0 when the
          // mask was empty, so every LDS read stays in the buffer
          int k0 = mask_pop<0, kNW, true>(m, split, jmax);
          float4 r0 = recf[wave][k0];
          int4 q0 = reci[wave][k0];
          while (true) {
            const bool more1 = mask_any(m);
            int k1 = mask_pop<0, kNW, true>(m, split, jmax);
            k1 = more1 ? k1 : 0;
            const float4 r1 = recf[wave][k1];
            const int4 q1 = reci[wave][k1];
            eval(r0, q0, k0, WZ);
            if (!more1) break;
            const bool more0 = mask_any(m);
            k0 = mask_pop<0, kNW, true>(m, split, jmax);
            k0 = more0 ? k0 : 0;
            r0 = recf[wave][k0];
            q0 = reci[wave][k0];
            eval(r1, q1, k1, WZ);
            if (!more0) break;
          }
        }
      };
      // FP64 keeps one loop with a runtime flag (fewer spilled registers)
      if (F64)
        run(std::integral_constant<int, 2>{});
      else if (clampz)
        run(std::integral_constant<int, 1>{});
      else
        run(std::integral_constant<int, 0>{});
      if (!F64) {
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
      wave_sync_lds();  // buffer is refilled next
    };

    auto emit = [&](bool ok, int t, float oxc, float oyc,
                    float z0) __attribute__((always_inline)) {
      float4 rf = make_float4(0.f, 0.f, 0.f, 0.f);
      int4 ri = make_int4(0, 0, 0, 0);
      if (ok) {
        const float4 lt = S.loc[t];
        ri = S.rec[t];
        rf.x = oxc + lt.x;
        rf.y = oyc + lt.y;
        rf.z = z0 + lt.z;
        rf.w = lt.w;
        float gzv = fmaxf(0.f, fmaxf(bz0 - rf.z, rf.z - bz1));
        if (clampz) gzv = 0.f;
        const float gxv = fmaxf(0.f, fmaxf(bx0 - rf.x, rf.x - bx1));
        const float gyv = fmaxf(0.f, fmaxf(by0 - rf.y, rf.y - by1));
        ok = fmaf(gxv, gxv, fmaf(gyv, gyv, gzv * gzv)) <= hi;
      }
      const unsigned long long ball = __ballot(ok);
      if (ok) {
        const int pos = nbuf + lane_prefix(ball);
        recf[wave][pos] = rf;
        reci[wave][pos] = ri;
        if (F64) recd[wave][pos] = make_double4(S.x[t], S.y[t], S.z[t],
                                                S.q[t]);
      }
      nbuf += __popcll(ball);
      if (nbuf > BUF - 64) {
        process(nbuf);
        nbuf = 0;
      }
    };
    stream_levels(S, G, 0, cx, cy, false, false, clampz, zlo, zhi, zA, zB,
                  bx0, bx1, by0, by1, bz0, bz1, hi, lane, emit);
    if (nbuf > 0) process(nbuf);
    // the 4 lanes of an atom add their partial forces
    if (F64) {
      fx64 += __shfl_down(fx64, 32, 64);
      fy64 += __shfl_down(fy64, 32, 64);
      fz64 += __shfl_down(fz64, 32, 64);
      fx64 += __shfl_down(fx64, 16, 64);
      fy64 += __shfl_down(fy64, 16, 64);
      fz64 += __shfl_down(fz64, 16, 64);
      if (active && split == 0) {
        forces[3 * ai] += fx64;
        forces[3 * ai + 1] += fy64;
        forces[3 * ai + 2] += fz64;
      }
    } else {
      Fx += __shfl_down(Fx, 32, 64);
      Fy += __shfl_down(Fy, 32, 64);
      Fz += __shfl_down(Fz, 32, 64);
      Fx += __shfl_down(Fx, 16, 64);
      Fy += __shfl_down(Fy, 16, 64);
      Fz += __shfl_down(Fz, 16, 64);
      if (active && split == 0) {
        forces[3 * ai] += (double)Fx * kFixedInv;
        forces[3 * ai + 1] += (double)Fy * kFixedInv;
        forces[3 * ai + 2] += (double)Fz * kFixedInv;
      }
    }
  }
  // this item's energies and virials (one row per item, as k_tm)
  if (item_idx < *nitems) {
#pragma unroll
    for (int k = 0; k < ACC_N; k++) {
      double v = acc[k];
      for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
      if (lane == 0) partial[(size_t)k * pstride + item_idx] = v;
    }
  }
  unsigned long long tot = block_count(cnt);
  if (threadIdx.x == 0 && tot) atomicAdd(npairs, tot);
}

}  // namespace pmnb
