// Tile kernel with per-lane pair masks (k_half), for one system or a batch
// of replicas (TileRep, pmnb_queue.h), over the full or the half shell.
//
// One wave handles 16 consecutive sorted atoms of one column, 4 lanes per
// atom (lane = atom slot + 16 * split).  Candidates stream into an LDS
// buffer one z level at a time (stream_levels) after a test of their
// distance to the wave's bounding box.  When the buffer is nearly full each
// lane tests every 4th buffered candidate against its own atom into a pass
// mask, then evaluates the set bits one by one, reading each candidate from
// LDS one pair ahead of its arithmetic.  The LJ table and the exclusion
// masks of the wave's own atoms are held in LDS.  Pairs within `band` of
// the cutoff are re-decided with the reference's FP64 minimum image.  Each
// item writes one partial row of energies and virials.
//
// Full shell (HALF = false, the path every single system takes): the
// candidates are the 3 x 3 neighbour columns, every pair is visited from
// both of its atoms and each visit takes half the pair's energy and virial.
// An atom's force is summed by its own wave only (FP32 partials per buffer
// added in 64-bit fixed point, or FP64 sums) and stored once into the
// per-atom fixed-point array: no atomics.
//
// Half shell (HALF = true, a batch shape): every unordered pair is
// evaluated once (Newton's third law, as NonbondedCutoffForce::doEvaluate
// and OneAtomPairTwo::doOneAtomPair do: force/nonbonded/NonbondedCutoffForce.h
// :48-52, force/OneAtomPairTwo.h:106-111): a wave takes its own column's
// atoms of higher sorted index and four of the eight neighbour columns, the
// image offsets (0, +1), (+1, -1), (+1, 0), (+1, +1).  Of a pair's two
// views, i -> j at image offset d and j -> i at -d, exactly one is taken:
// for d != 0 one of d, -d is in that half, for d = 0 the lower sorted index
// takes it.  With two columns on an axis the offsets -1 and +1 reach the
// same column at images one box (>= 2 cutoffs) apart, so at most one image
// of a pair is within the cutoff and it is taken once.  The partner's force
// is added into the candidate's LDS slot as 64-bit fixed-point integers,
// and after each buffer the slots go to the per-atom fixed-point array with
// global integer atomics.  Integer sums do not depend on the order of the
// adds, so the result is bitwise reproducible.  On the 96,000-atom water
// this is slower than the full shell: the LDS integer atomics and the
// global atomics cost more than the pair visits they save.
//
// Fixed point: mixed x, y, z at 2^-32; FP64 each component as (hi, lo) with
// value hi 2^-24 + lo 2^-56, exact to 2^-57 per term.  k_half_finish
// converts the array into the caller's forces (+=) and clears it for the
// next evaluation.
#pragma once

#include <type_traits>

#include "pmnb_queue.h"

namespace pmnb {

constexpr int kTmSplit = 4;                 // lanes per atom
constexpr int kTmWaves = kBlock / 64;       // waves per block

// Fixed-point words per atom of the force array: mixed x, y, z; FP64
// (hi, lo) of x, y, z.
template <bool F64>
struct FixWords {
  static constexpr int W = F64 ? 6 : 3;
};

// FP64 value as hi 2^-24 + lo 2^-56 with lo in [0, 2^32]: hi and frac are
// exact, lo rounds once (|error| <= 2^-57).  Values are clamped to
// +-2^37 so the integers cannot overflow.
__device__ __forceinline__ void fix64_split(double v, long long &hi,
                                            long long &lo) {
  v = fmin(fmax(v, -137438953472.0), 137438953472.0);
  const double s = v * 16777216.0;
  const double h = floor(s);
  hi = (long long)h;
  lo = (long long)__builtin_rint((s - h) * 4294967296.0);
}

__device__ __forceinline__ double fix64_value(long long hi, long long lo) {
  return (double)hi * (1.0 / 16777216.0) +
         (double)lo * (1.0 / 72057594037927936.0);
}

__device__ __forceinline__ void lds_add_i64(long long *p, long long v) {
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ __forceinline__ void global_add_i64(long long *p, long long v) {
  atomicAdd(reinterpret_cast<unsigned long long *>(p), (unsigned long long)v);
}

// Per-lane pass masks over the buffered records split + 4 j, j < jmax, in
// NW 32-bit words: word w holds j = 32 w .. 32 w + cw - 1 (cw = min(32,
// jmax - 32 w)) with record j at bit cw - 1 - (j - 32 w), so the test
// shifts its bits in.  Records n .. 4 jmax - 1 hold a far sentinel.
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

// Streams the candidate records of a wave's half shell one z level at a
// time, the wave's own levels first and then alternately one below and one
// above.  A buffer then holds candidates from every direction around the
// wave's atoms, so the atoms' pair counts per buffer are alike and the
// lanes of the mask loop wait less for one another.  Lane l in 4 .. 8 keeps
// neighbour column (l / 3 - 1, l % 3 - 1): (0, 0), (0, +1), (+1, -1),
// (+1, 0), (+1, +1).  In its own column (lane 4) only records of sorted
// index above a0 (the wave's first atom) are taken.  A level's cells of all
// passing columns are one virtual run, walked 64 records per call of
//   emit(ok, t, ox, oy, z0): sorted record t (valid if ok) with column image
//   offsets ox, oy and level base z0, all relative to the wave's home cell.
// wx / wy: single column on that axis (taken once, per-pair image); wz: the
// levels zA..zB cover every z cell once (per-pair image along z).
template <bool HALF, class Emit>
__device__ __forceinline__ void stream_levels(
    const Sorted &S, const TileGrid &G, int coff, int cx, int cy, bool wx,
    bool wy, bool wz, int zlo, int zhi, int zA, int zB, int a0, float bx0,
    float bx1, float by0, float by1, float bz0, float bz1, float hi, int lane,
    Emit &&emit) {
  const int ox = lane / 3 - 1, oy = lane % 3 - 1;
  const bool own = HALF && lane == 4;
  const float gx = wx ? 0.f
                      : fmaxf(0.f, fmaxf((float)ox * G.csx - bx1,
                                         bx0 - (float)(ox + 1) * G.csx));
  const float gy = wy ? 0.f
                      : fmaxf(0.f, fmaxf((float)oy * G.csy - by1,
                                         by0 - (float)(oy + 1) * G.csy));
  const float gxy2 = fmaf(gx, gx, gy * gy);
  const bool cok = lane >= (HALF ? 4 : 0) && lane < 9 &&
                   !(wx && ox != 0) && !(wy && oy != 0) && gxy2 <= hi;
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
    if (own) {  // records above a0 only
      const int lo = max(start, a0 + 1);
      cnt = max(0, start + cnt - lo);
      start = lo;
    }
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
    // the first 16 lanes; in the half shell lanes 0 .. 3 count nothing)
    int inc = cnt;
    inc += __builtin_amdgcn_update_dpp(0, inc, 0x111, 0xf, 0xf, false);
    inc += __builtin_amdgcn_update_dpp(0, inc, 0x112, 0xf, 0xf, false);
    inc += __builtin_amdgcn_update_dpp(0, inc, 0x114, 0xf, 0xf, false);
    inc += __builtin_amdgcn_update_dpp(0, inc, 0x118, 0xf, 0xf, false);
    const int total = __builtin_amdgcn_readlane(inc, 8);
    if (total > 0) {
      const float z0 = (float)(zc - zlo) * G.csz;
      const int rel = start - (inc - cnt);  // record = run index + rel
      constexpr int K0 = HALF ? 4 : 0, NE = 8 - K0;
      int ends[NE];
#pragma unroll
      for (int j = 0; j < NE; j++) ends[j] = __builtin_amdgcn_readlane(inc, K0 + j);
      for (int base = 0; base < total; base += 64) {
        const int idx = base + lane;
        int k = K0;
#pragma unroll
        for (int j = 0; j < NE; j++) k += idx >= ends[j] ? 1 : 0;
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

// The pair kernel.  REP = false: one system, one item per wave, the column
// grid and box as arguments, x and y with at least two columns; REP = true:
// a batch of replicas (per-item TileRep, per-pair minimum image on axes
// with a single column), a persistent item loop with a grid stride.  OCC:
// waves per SIMD the register allocation must allow.  LOG: the pair record
// of pmnb_force_pairs is compiled in (the same code, with each evaluated
// pair appended to plog).  ENER = false: forces and the pair count only,
// no energy or virial rows.
template <bool F64, int ELEC, int ELSW, int BUF, int OCC, bool REP, bool HALF,
          bool LOG, bool ENER = true>
__global__ void __launch_bounds__(kBlock)
__attribute__((amdgpu_waves_per_eu(OCC)))
k_half(ForceParams P, Sorted S, const TileRep *__restrict__ reps,
       TileGrid G1, const int4 *__restrict__ items,
       const int *__restrict__ nitems, long long *__restrict__ ffix,
       double *__restrict__ partial, int pstride,
       unsigned long long *__restrict__ npairs, PairLog plog) {
  constexpr int kNW = (BUF / kTmSplit + 31) / 32;  // mask words
  constexpr int W = FixWords<F64>::W;
  __shared__ float4 recf[kTmWaves][BUF];
  __shared__ int4 reci[kTmWaves][BUF];   // sorted index, type, mol, atom
  __shared__ double4 recd[F64 ? kTmWaves : 1][F64 ? BUF : 1];
  // fixed-point force on each buffered record from this wave's atoms
  // (half shell only)
  __shared__ long long pfix[HALF ? kTmWaves : 1][HALF ? W : 1][HALF ? BUF : 1];
  // exclusion masks of the wave's own atoms (excl_own layout)
  __shared__ unsigned long long xm[kTmWaves][kTileAtoms][4];
  __shared__ float4 ljs[F64 ? 1 : kQLj];  // FP32 LJ table (T * T <= kQLj)
  // values equal on all lanes of the wave are marked uniform (scalar
  // registers) where that measured faster
  auto U = [](auto v) { return (F64 || REP) ? uni(v) : v; };
  const int wave = uni((int)(threadIdx.x >> 6)), lane = threadIdx.x & 63;
  if (HALF)
    for (int k = lane; k < W * BUF; k += 64) (&pfix[HALF ? wave : 0][0][0])[k] = 0;
  // each pair is seen from both atoms in the full shell: half its energy
  // and virial to each
  constexpr double kShare = HALF ? 1.0 : 0.5;
  if (!F64) {
    for (int k = threadIdx.x; k < P.n_types * P.n_types; k += kBlock)
      ljs[k] = S.ljf[k];
  }
  __syncthreads();
  const F32Params &F = P.f;
  const float hi = F.cut2 + F.band, lo = F.cut2 - F.band;
  const bool vir = ENER && (P.flags & 3u) != 0;
  const int n_items = *nitems;
  // FP64 single system: blocks go to the 8 XCDs in turn; block b takes the
  // items of block position (b % 8) * per + b / 8 (the grid is a multiple
  // of 8), so each XCD works through one contiguous range of items and the
  // candidates neighbouring items share stay in that XCD's L2.
  int item_idx;
  if (REP || !F64) {
    item_idx = blockIdx.x * kTmWaves + wave;
  } else {
    const int per = gridDim.x / 8;
    item_idx = ((int)(blockIdx.x % 8) * per + (int)(blockIdx.x / 8)) *
                   kTmWaves + wave;
  }
  unsigned cnt = 0;
  for (; item_idx < n_items; item_idx += gridDim.x * kTmWaves) {
    double acc[ACC_N];
#pragma unroll
    for (int k = 0; k < ACC_N; k++) acc[k] = 0.0;
    const int4 item = items[item_idx];
    const int a0 = U(item.x), a1 = U(item.y), col = U(item.z);
    const TileRep *Rp = REP ? &reps[uni(item.w)] : nullptr;
    const TileGrid G = REP ? Rp->tg : G1;
    const int coff = REP ? Rp->cell_off : 0;
    const bool wx = REP ? Rp->wrap[0] != 0 : false;
    const bool wy = REP ? Rp->wrap[1] != 0 : false;
    const Box *bptr = REP ? &Rp->box : S.box;
    float L[3], iL[3];
    if (REP) {
      for (int k = 0; k < 3; k++) {
        L[k] = Rp->L[k];
        iL[k] = Rp->iL[k];
      }
    } else {
      L[0] = F.Lx; L[1] = F.Ly; L[2] = F.Lz;
      iL[0] = F.iLx; iL[1] = F.iLy; iL[2] = F.iLz;
    }
    const int cx = col / G.ny, cy = col % G.ny;
    const int slot = lane & (kTileAtoms - 1), split = lane >> 4;
    const int i = a0 + slot;
    const bool active = i < a1;
    const int ii = active ? i : a0;
    const int zlo = U((S.cell[a0] - coff) % G.nz);
    const int zhi = U((S.cell[a1 - 1] - coff) % G.nz);
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
    const float bx0 = U(wave_min(active ? xi : 1e30f)),
                bx1 = U(wave_max(active ? xi : -1e30f));
    const float by0 = U(wave_min(active ? yi : 1e30f)),
                by1 = U(wave_max(active ? yi : -1e30f));
    const float bz0 = U(wave_min(active ? zi : 1e30f)),
                bz1 = U(wave_max(active ? zi : -1e30f));
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
      // half shell: the atom itself and the wave's own atoms below it
      // (records of the own column start above a0; other columns hold no
      // atom of this wave); full shell: the atom itself
      if (HALF ? (unsigned)(t - a0) <= (unsigned)slot : t == i) return;
      float dx = r.x - xi, dy = r.y - yi, dz = r.z - zi;
      if (M == 2 || (M == 3 && wx)) dx = fmaf(-L[0], rintf(dx * iL[0]), dx);
      if (M == 2 || (M == 3 && wy)) dy = fmaf(-L[1], rintf(dy * iL[1]), dy);
      if (M == 1 || M == 2 || (M == 3 && wz))
        dz = fmaf(-L[2], rintf(dz * iL[2]), dz);
      const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
      double ex = 0.0, ey = 0.0, ez = 0.0, e2d = 0.0;
      if (F64) {
        const double4 rd = recd[wave][k];
        min_diff_exact(*bptr, xd, yd, zd, rd.x, rd.y, rd.z, ex, ey, ez, e2d);
        if (e2d > P.cut2) return;
      } else if (r2 >= lo) {
        // near the cutoff: decide with the reference's FP64 minimum image
        min_diff_exact(rare_box(bptr), xd, yd, zd, S.x[t], S.y[t], S.z[t],
                       ex, ey, ez, e2d);
        if (e2d > P.cut2) return;
      }
      const bool same = q4.z == mi;
      const int excl = same ? excl_own(xm[wave][slot], q4.w - ai) : 0;
      if (excl == 2) return;
      cnt += (HALF || i < t) ? 1u : 0u;
      if (LOG && (HALF || i < t)) log_pair(plog, ai, q4.w, excl);
      const bool molv = (P.flags & 2u) && !same;
      if (F64) {
        const double *lp = ljrow64 + 4 * q4.y;
        double ff, e1, e2;
        pair_f64<ELEC, ELSW>(P, e2d, excl != 1 ? lp[0] : lp[2],
                             excl != 1 ? lp[1] : lp[3], qi * recd[wave][k].w,
                             excl, ff, e1, e2);
        if (ENER) {
          acc[ACC_ELJ] += kShare * e1;
          acc[ACC_EEL] += kShare * e2;
        }
        const double gx = ex * ff, gy = ey * ff, gz = ez * ff;
        fx64 -= gx;
        fy64 -= gy;
        fz64 -= gz;
        if constexpr (HALF) {
          long long h, l;
          fix64_split(gx, h, l);
          lds_add_i64(&pfix[wave][0][k], h);
          lds_add_i64(&pfix[wave][1][k], l);
          fix64_split(gy, h, l);
          lds_add_i64(&pfix[wave][2][k], h);
          lds_add_i64(&pfix[wave][3][k], l);
          fix64_split(gz, h, l);
          lds_add_i64(&pfix[wave][4][k], h);
          lds_add_i64(&pfix[wave][5][k], l);
        }
        if (ENER && (molv || (P.flags & 1u)))
          add_virial(acc, ACC_VXX, gx, gy, gz, ex, ey, ez, kShare);
        if (ENER && molv) {
          const Box &bm = rare_box(bptr);
          double cdx, cdy, cdz, cd2;
          min_diff_exact(bm, S.com[3 * mi], S.com[3 * mi + 1],
                         S.com[3 * mi + 2], S.com[3 * q4.z],
                         S.com[3 * q4.z + 1], S.com[3 * q4.z + 2], cdx, cdy,
                         cdz, cd2);
          add_virial(acc, ACC_MXX, gx, gy, gz, cdx, cdy, cdz, kShare);
        }
      } else {
        const float4 lp = ljs[ljbase + q4.y];
        float ff, e1, e2;
        pair_f32<ELEC>(F, r2, excl != 1 ? lp.x : lp.z, excl != 1 ? lp.y : lp.w,
                       li.w * r.w, excl, ff, e1, e2);
        if (ENER) {
          pelj += e1;
          peel += e2;
        }
        const float gx = dx * ff, gy = dy * ff, gz = dz * ff;
        px -= gx;
        py -= gy;
        pz -= gz;
        if constexpr (HALF) {
          lds_add_i64(&pfix[wave][0][k], to_fixed_f32(gx));
          lds_add_i64(&pfix[wave][1][k], to_fixed_f32(gy));
          lds_add_i64(&pfix[wave][2][k], to_fixed_f32(gz));
        }
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
            const Box &bm = rare_box(bptr);
            double cdx, cdy, cdz, cd2;
            min_diff_exact(bm, S.com[3 * mi], S.com[3 * mi + 1],
                           S.com[3 * mi + 2], S.com[3 * q4.z],
                           S.com[3 * q4.z + 1], S.com[3 * q4.z + 2], cdx, cdy,
                           cdz, cd2);
            add_virial(acc, ACC_MXX, gx, gy, gz, cdx, cdy, cdz, kShare);
          }
        }
      }
    };

    const int wmode = (wx ? 1 : 0) | (wy ? 2 : 0) | (wz ? 4 : 0);
    // Tests the buffered candidates against this lane's atom, evaluates the
    // pairs that pass (the next pair's records are read from LDS before the
    // current pair's arithmetic), then moves the candidates' fixed-point
    // forces to the force array and clears their slots.
    auto process = [&](int n) __attribute__((always_inline)) {
      // records n .. n + 3 (the last group of 4) get NaN coordinates, which
      // fail the distance test under any minimum image
      if (lane < kTmSplit && n + lane < BUF) {
        const float nan = __builtin_nanf("");
        recf[wave][n + lane] = make_float4(nan, nan, nan, 0.f);
      }
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
        Fx += to_fixed_f32(px);
        Fy += to_fixed_f32(py);
        Fz += to_fixed_f32(pz);
        if (ENER) {
          acc[ACC_ELJ] += kShare * (double)pelj;
          acc[ACC_EEL] += kShare * (double)peel;
        }
        if (vir) {
          acc[ACC_VXX] += kShare * (double)v0;
          acc[ACC_VXY] += kShare * (double)v1;
          acc[ACC_VXZ] += kShare * (double)v2;
          acc[ACC_VYY] += kShare * (double)v3;
          acc[ACC_VYZ] += kShare * (double)v4;
          acc[ACC_VZZ] += kShare * (double)v5;
        }
      }
      if (!HALF) {
        wave_sync_lds();  // the buffer is refilled next
        return;
      }
      wave_sync_lds();  // every lane's slot adds are in
      for (int k = lane; k < n; k += 64) {
        long long v[W];
        long long any = 0;
#pragma unroll
        for (int c = 0; c < W; c++) {
          v[c] = pfix[wave][c][k];
          any |= v[c];
          pfix[wave][c][k] = 0;
        }
        if (any) {
          long long *dst = ffix + (size_t)W * reci[wave][k].x;
#pragma unroll
          for (int c = 0; c < W; c++) global_add_i64(dst + c, v[c]);
        }
      }
      wave_sync_lds();  // buffer and slots are refilled next
    };

    auto emit = [&](bool ok, int t, float oxc, float oyc,
                    float z0) __attribute__((always_inline)) {
      float4 rf = make_float4(0.f, 0.f, 0.f, 0.f);
      int4 ri = make_int4(0, 0, 0, 0);
      if (ok) {
        const float4 lt = S.loc[t];
        {
          const int2 rr = S.rec[t];
          ri = make_int4(t, rr.x & 255, rr.x >> 8, rr.y);
        }
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
    stream_levels<HALF>(S, G, coff, cx, cy, wx, wy, wz, zlo, zhi, zA, zB, a0,
                        bx0, bx1, by0, by1, bz0, bz1, hi, lane, emit);
    if (nbuf > 0) process(nbuf);
    // the 4 lanes of an atom add their partial forces; in the half shell
    // the atom's force then joins the partner forces other waves add to
    // it, in the full shell it is the whole force (one writer per atom)
    if (F64) {
      fx64 += __shfl_down(fx64, 32, 64);
      fy64 += __shfl_down(fy64, 32, 64);
      fz64 += __shfl_down(fz64, 32, 64);
      fx64 += __shfl_down(fx64, 16, 64);
      fy64 += __shfl_down(fy64, 16, 64);
      fz64 += __shfl_down(fz64, 16, 64);
      if (active && split == 0) {
        long long *dst = ffix + (size_t)W * ii;
        long long h[3], l[3];
        fix64_split(fx64, h[0], l[0]);
        fix64_split(fy64, h[1], l[1]);
        fix64_split(fz64, h[2], l[2]);
#pragma unroll
        for (int c = 0; c < 3; c++) {
          if (HALF) {
            global_add_i64(dst + 2 * c, h[c]);
            global_add_i64(dst + 2 * c + 1, l[c]);
          } else {
            dst[2 * c] = h[c];
            dst[2 * c + 1] = l[c];
          }
        }
      }
    } else {
      Fx += __shfl_down(Fx, 32, 64);
      Fy += __shfl_down(Fy, 32, 64);
      Fz += __shfl_down(Fz, 32, 64);
      Fx += __shfl_down(Fx, 16, 64);
      Fy += __shfl_down(Fy, 16, 64);
      Fz += __shfl_down(Fz, 16, 64);
      if (active && split == 0) {
        long long *dst = ffix + (size_t)W * ii;
        if (HALF) {
          global_add_i64(dst, Fx);
          global_add_i64(dst + 1, Fy);
          global_add_i64(dst + 2, Fz);
        } else {
          dst[0] = Fx;
          dst[1] = Fy;
          dst[2] = Fz;
        }
      }
    }
    // this item's energies and virials
    if (ENER) {
#pragma unroll
      for (int k = 0; k < ACC_N; k++) {
        double v = acc[k];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
        if (lane == 0) partial[(size_t)k * pstride + item_idx] = v;
      }
    }
    if (!REP) break;  // one item per wave
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

// The pair kernel of one system (the path every single system takes): the
// full shell of k_half with REP = false, written for one replica only: the
// column grid and box are kernel arguments, x and y have at least two
// columns, z takes the minimum image only when the cutoff reaches every z
// cell, and each atom's force is added straight into the caller's array
// (+=, one writer per atom).  Fewer scalar registers than the
// replica-general form, measured faster on one large system.  LOG: the
// pair record of pmnb_force_pairs is compiled in.  ENER = false: forces
// and the pair count only, no energy or virial rows (for evaluations whose
// energies nobody reads, such as the steps of a dynamics run before its
// last).
template <bool F64, int ELEC, int ELSW, int BUF, bool LOG, bool ENER = true>
__global__ void __launch_bounds__(kBlock)
__attribute__((amdgpu_waves_per_eu(3)))
k_tile(ForceParams P, Sorted S, TileGrid G, const int4 *__restrict__ items,
       const int *__restrict__ nitems, double *__restrict__ forces,
       double *__restrict__ partial, int pstride,
       unsigned long long *__restrict__ npairs, PairLog plog) {
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
  // FP64: values equal on all lanes of the wave are marked uniform (uni), so
  // the work item, its z range and box live in scalar registers (in the
  // FP32 kernel the extra scalar registers spill and cost more than they
  // save)
  auto U = [](auto v) { return F64 ? uni(v) : v; };
  const int wave = U((int)(threadIdx.x >> 6)), lane = threadIdx.x & 63;
  // FP64 and the forces-only form: blocks go to the 8 XCDs in turn; block b
  // takes the items of block position (b % 8) * per + b / 8 (the grid is a
  // multiple of 8), so each XCD works through one contiguous range of items
  // and the candidates that neighbouring items share stay in that XCD's L2
  // (measured faster for those two, 1.5% for the forces-only form with its
  // four waves per SIMD, and slightly slower for the FP32 kernel with
  // energies).
  const int per = gridDim.x / 8;
  const int blk =
      (F64 || !ENER) ? (int)(blockIdx.x % 8) * per + (int)(blockIdx.x / 8)
                     : blockIdx.x;
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
    const bool vir = ENER && (P.flags & 3u) != 0;
    int nbuf = 0;

    // FP32 partials of one buffer (flushed by process)
    float px = 0.f, py = 0.f, pz = 0.f, pelj = 0.f, peel = 0.f;
    float v0 = 0.f, v1 = 0.f, v2 = 0.f, v3 = 0.f, v4 = 0.f, v5 = 0.f;

    // One pair: candidate record k of this wave's buffer, seen from atom i
    // (each pair is seen from both atoms: half its energy and virial here).
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
      const int excl = same ? excl_own(xm[wave][slot], q4.w - ai) : 0;
      if (excl == 2) return;
      if (i < t) {
        cnt++;
        if (LOG) log_pair(plog, ai, q4.w, excl);
      }
      const bool molv = (P.flags & 2u) && !same;
      if (F64) {
        const double *lp = ljrow64 + 4 * q4.y;
        double ff, e1, e2;
        pair_f64<ELEC, ELSW>(P, e2d, excl != 1 ? lp[0] : lp[2],
                             excl != 1 ? lp[1] : lp[3], qi * recd[wave][k].w,
                             excl, ff, e1, e2);
        if (ENER) {
          acc[ACC_ELJ] += 0.5 * e1;
          acc[ACC_EEL] += 0.5 * e2;
        }
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
        if (ENER) {
          pelj += e1;
          peel += e2;
        }
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

    // Tests the buffered candidates against this lane's atom, then
    // evaluates the pairs that pass; the next pair's records are read from
    // LDS before the current pair's arithmetic.
    auto process = [&](int n) __attribute__((always_inline)) {
      // records n .. n + 3 (the last group of 4) get NaN coordinates, which
      // fail the distance test
      if (lane < kTmSplit && n + lane < BUF) {
        const float nan = __builtin_nanf("");
        recf[wave][n + lane] = make_float4(nan, nan, nan, 0.f);
      }
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
        if (ENER) {
          acc[ACC_ELJ] += 0.5 * (double)pelj;
          acc[ACC_EEL] += 0.5 * (double)peel;
        }
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
        {
          const int2 rr = S.rec[t];
          ri = make_int4(t, rr.x & 255, rr.x >> 8, rr.y);
        }
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
    stream_levels<false>(S, G, 0, cx, cy, false, false, clampz, zlo, zhi, zA,
                         zB, a0, bx0, bx1, by0, by1, bz0, bz1, hi, lane, emit);
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
        if (P.flags & kFlagFresh) {
          forces[3 * ai] = fx64;
          forces[3 * ai + 1] = fy64;
          forces[3 * ai + 2] = fz64;
        } else {
          forces[3 * ai] += fx64;
          forces[3 * ai + 1] += fy64;
          forces[3 * ai + 2] += fz64;
        }
      }
    } else {
      Fx += __shfl_down(Fx, 32, 64);
      Fy += __shfl_down(Fy, 32, 64);
      Fz += __shfl_down(Fz, 32, 64);
      Fx += __shfl_down(Fx, 16, 64);
      Fy += __shfl_down(Fy, 16, 64);
      Fz += __shfl_down(Fz, 16, 64);
      if (active && split == 0) {
        if (P.flags & kFlagFresh) {
          forces[3 * ai] = (double)Fx * kFixedInv;
          forces[3 * ai + 1] = (double)Fy * kFixedInv;
          forces[3 * ai + 2] = (double)Fz * kFixedInv;
        } else {
          forces[3 * ai] += (double)Fx * kFixedInv;
          forces[3 * ai + 1] += (double)Fy * kFixedInv;
          forces[3 * ai + 2] += (double)Fz * kFixedInv;
        }
      }
    }
  }
  // this item's energies and virials (one row per item)
  if (ENER && item_idx < *nitems) {
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

// After the pair kernel: blocks below atom_blocks convert the fixed-point
// force of each sorted slot into the caller's forces (+= at the original
// index, = with fresh) and clear it; the next rep_blocks sum the item
// partials of one (replica, accumulator) each (finish_rep_block), the
// first of them also storing the pair count in np_out when given; the rest
// clear the cell counts for the next evaluation (nothing reads them after
// the pair kernel).
template <bool F64>
__global__ void __launch_bounds__(1024)
k_half_finish(long long *__restrict__ ffix, const int *__restrict__ atom,
              int n, double *__restrict__ forces, int atom_blocks,
              const double *__restrict__ partial, int pstride,
              const int *__restrict__ rep_first, double *__restrict__ out,
              int rep_blocks, int *__restrict__ count, int ncell, int fresh,
              const unsigned long long *__restrict__ npairs,
              unsigned long long *__restrict__ np_out) {
  constexpr int W = FixWords<F64>::W;
  if ((int)blockIdx.x < atom_blocks) {
    const int s = blockIdx.x * 1024 + threadIdx.x;
    if (s >= n) return;
    long long *src = ffix + (size_t)W * s;
    const int a = atom[s];
    for (int c = 0; c < 3; c++) {
      const double f = F64 ? fix64_value(src[2 * c], src[2 * c + 1])
                           : (double)src[c] * kFixedInv;
      forces[3 * a + c] = fresh ? f : forces[3 * a + c] + f;
      if (F64) {
        src[2 * c] = 0;
        src[2 * c + 1] = 0;
      } else {
        src[c] = 0;
      }
    }
    return;
  }
  const int b = blockIdx.x - atom_blocks;
  if (b >= rep_blocks) {
    const int c = (b - rep_blocks) * 1024 + threadIdx.x;
    if (c < ncell) count[c] = 0;
    return;
  }
  if (np_out && b == 0 && threadIdx.x == 0) *np_out = *npairs;
  finish_rep_block(partial, pstride, rep_first, out, b / ACC_N, b % ACC_N,
                   fresh);
}

}  // namespace pmnb
