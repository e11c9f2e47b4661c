// Pair-queue tile kernel: the production pair kernel of the cutoff force,
// for one system or a batch of independent periodic replicas.
//
// Work decomposition (grid in pmnb_tile.h): one wave per 16 consecutive
// sorted atoms of one x/y column of one replica, whose candidates stream from
// the neighbour columns into an LDS record buffer, 64 at a time, after a test
// of their distance to the wave's bounding box (ballot + mbcnt compaction).
// The pair phase is converged instead of per-lane:
//   test     lane = atom slot + 16 * sub holds its atom's position; in 16
//            steps the wave tests a group of 64 buffered records (record
//            4 j + sub at step j) against all 16 atoms and appends the
//            passing (atom, record) entries to an LDS pair queue at ballot +
//            mbcnt positions (failing lanes write a dummy entry, so the step
//            has no branch).  Only whole groups of 64 records are tested;
//            the rest stay buffered for the next round;
//   evaluate the queue is consumed 64 entries at a time, one pair per lane,
//            with the next group's entries and records read from LDS before
//            the current group's arithmetic, so no lane idles while another
//            atom of the wave has more pairs.
// A lane adds the force of its pair into its own LDS slot of that atom, so
// every floating-point sum has an order fixed by the queue; the slots of an
// atom are combined at the end in 64-bit fixed point (mixed) or FP64 in a
// fixed order.  Every pair is visited from both atoms; energies and virials
// take half.
//
// Each replica has its own column grid (TileRep).  An axis with a single
// column (box shorter than two cutoffs) takes it once and the minimum image
// of every pair along it (as z does when the cutoff reaches all z cells), so
// small boxes keep the reference's one-image pair set.
//
// The kernel is persistent: each wave walks the work items with a grid
// stride, and writes one partial row (energies, virials) per item.
//
// Records and atoms carry {sorted index | type << 24, molecule}; the original
// atom index (exclusion lookup) is read from global memory only for pairs of
// one molecule.  The LJ table is held in LDS (at most kQLj entries).
#pragma once

#include <type_traits>

#include "pmnb_cells.h"
#include "pmnb_list.h"
#include "pmnb_tile.h"

namespace pmnb {

constexpr int kQCap = 2048;   // pair-queue entries per wave
constexpr int kQLj = 64;      // LJ table entries held in LDS
constexpr int kQTypeShift = 24;
constexpr int kQIndexMask = (1 << kQTypeShift) - 1;

// Geometry of one replica on the tile path.
struct TileRep {
  Box box;            // FP64 box: minimum image of pairs near the cutoff
  Grid grid;          // FP64 column grid (cell assignment)
  TileGrid tg;        // FP32 cell sizes and the z reach of the cutoff
  float L[3], iL[3];  // box lengths and inverses
  int wrap[2];        // x, y: a single column (per-pair minimum image)
  int cell_off;       // first global cell of the replica
  int col_off;        // first global column
  int ncol;           // columns (tg.nx * tg.ny)
  int pad;
};

// 64-bit fixed point (2^-32) of an FP32 value, saturating beyond +-2^31;
// equal to to_fixed((double)v).
__device__ __forceinline__ long long to_fixed_f32(float v) {
  v = fminf(fmaxf(v, -2147483520.f), 2147483520.f);
  return (long long)rintf(v * 4294967296.f);
}

template <typename T>
__device__ __forceinline__ void lds_add(T *p, T v) {
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// The box of a rare path (pairs near the cutoff, molecule centres): the
// pointer passes through an empty asm statement, so the loads of the box
// fields stay in that branch instead of being hoisted into scalar
// registers held across the whole work item.  The pointer must be equal on
// all lanes (one replica per work item); it is read from the first lane, so
// it sits in scalar registers whatever the compiler infers.
__device__ __forceinline__ const Box &rare_box(const Box *b) {
  const unsigned long long v = (unsigned long long)b;
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
  b = (const Box *)(((unsigned long long)hi << 32) | lo);
  __asm__ volatile("" : "+s"(b));
  return *b;
}

// Marks a value equal on all lanes as wave-uniform (scalar register), so the
// loops and branches it controls compile to scalar control flow.
__device__ __forceinline__ int uni(int v) {
  return __builtin_amdgcn_readfirstlane(v);
}
__device__ __forceinline__ float uni(float v) {
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}

// All work items in one block: each thread takes a contiguous run of global
// columns, counts their waves (ceil(atoms / 16)), the block scans the counts,
// then every thread writes the items {first atom, end atom, local column,
// replica} of its columns.  rep_first[r] = first item of replica r;
// rep_first[nrep] and *nitems = the item count.
__device__ __forceinline__ void build_items(
    const TileRep *__restrict__ reps, const int *__restrict__ col_rep,
    const int *__restrict__ cstart, const int *__restrict__ ccount, int ncol,
    int nrep, int4 *__restrict__ items, int *__restrict__ nitems,
    int *__restrict__ rep_first) {
  const int t = threadIdx.x;
  const int chunk = (ncol + 1023) / 1024;
  const int b0 = min(ncol, t * chunk), b1 = min(ncol, b0 + chunk);
  auto span = [&](int gc, int &r, int &lc, int &a0, int &a1) {
    r = col_rep[gc];
    const TileRep &R = reps[r];
    const int nz = R.tg.nz;
    lc = gc - R.col_off;
    const int first = R.cell_off + lc * nz, last = first + nz - 1;
    a0 = cstart[first];
    a1 = cstart[last] + ccount[last];
  };
  int sum = 0;
  for (int gc = b0; gc < b1; gc++) {
    int r, lc, a0, a1;
    span(gc, r, lc, a0, a1);
    sum += (a1 - a0 + kTileAtoms - 1) / kTileAtoms;
  }
  int total;
  int o = block_scan_1024(sum, &total);  // first item of this thread's columns
  for (int gc = b0; gc < b1; gc++) {
    int r, lc, a0, a1;
    span(gc, r, lc, a0, a1);
    if (lc == 0) rep_first[r] = o;
    for (int a = a0; a < a1; a += kTileAtoms)
      items[o++] = make_int4(a, min(a + kTileAtoms, a1), lc, r);
  }
  if (t == 1023) {
    *nitems = total;
    rep_first[nrep] = total;
  }
}

// Cell starts (exclusive scan of the counts) and then the work items, in
// one block: the items read the starts this block has just written.
__global__ void __launch_bounds__(1024)
k_scan_items(const int *__restrict__ count, int *__restrict__ start,
             int ncell, const TileRep *__restrict__ reps,
             const int *__restrict__ col_rep, int ncol, int nrep,
             int4 *__restrict__ items, int *__restrict__ nitems,
             int *__restrict__ rep_first) {
  scan_counts(count, start, ncell);
  __syncthreads();
  build_items(reps, col_rep, start, count, ncol, nrep, items, nitems,
              rep_first);
}

// Per-replica sum of one accumulator of the item partials (accumulator k
// of item b at partial[k * pstride + b]) by one block of 1024 threads with
// a fixed reduction tree, added into out[20 r + ..]:
// [0] LJ, [1] electrostatic, [2..10] virial 3x3, [11..19] molecular virial.
__device__ __forceinline__ void finish_rep_block(
    const double *__restrict__ partial, int pstride,
    const int *__restrict__ rep_first, double *__restrict__ out, int r,
    int k) {
  __shared__ double red[16];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int i0 = rep_first[r], i1 = rep_first[r + 1];
  const double *pk = partial + (size_t)k * pstride;
  // four independent sums per thread, so the loads overlap
  double v0 = 0.0, v1 = 0.0, v2 = 0.0, v3 = 0.0;
  int b = i0 + (int)threadIdx.x;
  for (; b + 3072 < i1; b += 4096) {
    v0 += pk[b];
    v1 += pk[b + 1024];
    v2 += pk[b + 2048];
    v3 += pk[b + 3072];
  }
  for (; b < i1; b += 1024) v0 += pk[b];
  double v = (v0 + v1) + (v2 + v3);
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  if (lane == 0) red[wave] = v;
  __syncthreads();
  if (threadIdx.x != 0) return;
  v = 0.0;
  for (int w = 0; w < 16; w++) v += red[w];
  double *o = out + 20 * r;
  if (k < 2) {
    o[k] += v;
  } else {
    const int base = k < ACC_MXX ? 2 : 11;
    const int u = k < ACC_MXX ? k - ACC_VXX : k - ACC_MXX;
    // u: 0 xx, 1 xy, 2 xz, 3 yy, 4 yz, 5 zz
    switch (u) {
      case 0: o[base + 0] += v; break;
      case 1: o[base + 1] += v; o[base + 3] += v; break;
      case 2: o[base + 2] += v; o[base + 6] += v; break;
      case 3: o[base + 4] += v; break;
      case 4: o[base + 5] += v; o[base + 7] += v; break;
      default: o[base + 8] += v; break;
    }
  }
}

// One block per (replica, accumulator).
__global__ void __launch_bounds__(1024)
k_finish_rep(const double *__restrict__ partial, int pstride,
             const int *__restrict__ rep_first, double *__restrict__ out) {
  finish_rep_block(partial, pstride, rep_first, out, blockIdx.x, blockIdx.y);
}

// Tile-path kernel shape: waves per block, buffered records per wave, force
// slots per atom (64 / SLOTS lanes share a slot and add one after another)
// and whether the slots take LDS atomic adds or plain read-modify-writes.
template <int NW_, int BUF_, int SLOTS_, bool ATOM_>
struct TqShape {
  static constexpr int NW = NW_, BUF = BUF_, SLOTS = SLOTS_;
  static constexpr bool ATOM = ATOM_;
};

template <bool F64, int ELEC, int ELSW, class SH_>
__global__ void __launch_bounds__(64 * SH_::NW)
k_tq(ForceParams P, Sorted S, const TileRep *__restrict__ reps,
     const int4 *__restrict__ items, const int *__restrict__ nitems,
     double *__restrict__ forces, double *__restrict__ partial, int pstride,
     unsigned long long *__restrict__ npairs) {
  constexpr int NW = SH_::NW, BUF = SH_::BUF, SLOTS = SH_::SLOTS;
  constexpr bool ATOM = SH_::ATOM;
  using Real = typename std::conditional<F64, double, float>::type;
  constexpr int KS = kTileAtoms * SLOTS;   // component stride of facc
  constexpr int LPS = 64 / SLOTS;          // lanes per slot
  __shared__ float4 recf[NW][BUF];         // home-relative x, y, z; charge
  __shared__ int2 reci[NW][BUF];           // sorted index | type, molecule
  __shared__ double4 recd[F64 ? NW : 1][F64 ? BUF : 1];  // FP64 x, y, z, q
  // atom slot | record << 4; entry kQCap takes the writes of failed tests
  __shared__ unsigned short que[NW][kQCap + 2];
  __shared__ float4 ipos[NW][kTileAtoms];
  __shared__ int2 iint[NW][kTileAtoms];
  __shared__ double4 ipd[F64 ? NW : 1][kTileAtoms];
  __shared__ Real facc[NW][3 * KS];
  __shared__ float4 ljs[F64 ? 1 : kQLj];
  __shared__ double ljs64[F64 ? 4 * kQLj : 1];
  // wave index as a wave-uniform (scalar) value
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int nlj = P.n_types * P.n_types;  // <= kQLj (checked by the host)
  for (int k = threadIdx.x; k < nlj; k += 64 * NW) {
    if (F64) {
      for (int c = 0; c < 4; c++) ljs64[4 * k + c] = S.lj[4 * k + c];
    } else {
      ljs[k] = S.ljf[k];
    }
  }
  __syncthreads();
  const F32Params &F = P.f;
  const float hi = F.cut2 + F.band, lo = F.cut2 - F.band;
  const bool vir = (P.flags & 3u) != 0;
  const int n_items = *nitems;
  unsigned cnt = 0;
  for (int item_idx = blockIdx.x * NW + wave; item_idx < n_items;
       item_idx += gridDim.x * NW) {
    double acc[ACC_N];
#pragma unroll
    for (int k = 0; k < ACC_N; k++) acc[k] = 0.0;
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
    // every lane holds atom slot lane % 16 (the test phase uses all four
    // copies; the bounding box is unaffected by the repeats)
    const bool own = (lane & (kTileAtoms - 1)) < na;
    const int ii = a0 + (own ? (lane & (kTileAtoms - 1)) : 0);
    const int zlo = uni((S.cell[a0] - coff) % G.nz);
    const int zhi = uni((S.cell[a1 - 1] - coff) % G.nz);
    const float4 li = S.loc[ii];
    const float zi = (float)((S.cell[ii] - coff) % G.nz - zlo) * G.csz + li.z;
    if (lane < kTileAtoms) {
      ipos[wave][lane] = make_float4(li.x, li.y, zi, li.w);
      iint[wave][lane] = make_int2(ii | (S.type[ii] << kQTypeShift), S.mol[ii]);
      if (F64)
        ipd[wave][lane] = make_double4(S.x[ii], S.y[ii], S.z[ii], S.q[ii]);
    }
    for (int k = lane; k < 3 * KS; k += 64) facc[wave][k] = Real(0);
    // bounding box of the wave's atoms in home-relative coordinates
    const float bx0 = uni(wave_min(own ? li.x : 1e30f)),
                bx1 = uni(wave_max(own ? li.x : -1e30f));
    const float by0 = uni(wave_min(own ? li.y : 1e30f)),
                by1 = uni(wave_max(own ? li.y : -1e30f));
    const float bz0 = uni(wave_min(own ? zi : 1e30f)),
                bz1 = uni(wave_max(own ? zi : -1e30f));
    int zA = zlo - G.rz, zB = zhi + G.rz;
    const bool wz = zB - zA + 1 > G.nz;  // all z cells: per-pair image
    if (wz) zB = zA + G.nz - 1;
    int nbuf = 0;

    float pelj = 0.f, peel = 0.f;
    float v0 = 0.f, v1 = 0.f, v2 = 0.f, v3 = 0.f, v4 = 0.f, v5 = 0.f;
    // One queued pair: atom slot s (ip, iq, id) against a buffered record
    // (r, rq, rd).
    auto one = [&](int s, const float4 &ip, const int2 &iq, const double4 &id,
                   const float4 &r, const int2 &rq, const double4 &rd) {
      const int i = iq.x & kQIndexMask, t = rq.x & kQIndexMask;
      if (t == i) return;
      float dx = r.x - ip.x, dy = r.y - ip.y, dz = r.z - ip.z;
      if (wx) dx = fmaf(-Lx, rintf(dx * iLx), dx);
      if (wy) dy = fmaf(-Ly, rintf(dy * iLy), dy);
      if (wz) dz = fmaf(-Lz, rintf(dz * iLz), dz);
      const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
      double ex = 0.0, ey = 0.0, ez = 0.0, e2d = 0.0;
      if (F64) {
        min_diff_exact(R.box, id.x, id.y, id.z, rd.x, rd.y, rd.z, ex, ey, ez,
                       e2d);
        if (e2d > P.cut2) return;
      } else if (r2 >= lo) {
        // near the cutoff: decide with the reference's FP64 minimum image
        min_diff_exact(rare_box(&R.box), S.x[i], S.y[i], S.z[i], S.x[t],
                       S.y[t], S.z[t], ex, ey, ez, e2d);
        if (e2d > P.cut2) return;
      }
      const int mi = iq.y;
      const bool same = rq.y == mi;
      const int excl = same ? excl_class(S.xfull, S.xmod, P.max_excl_delta,
                                         S.atom[i], S.atom[t])
                            : 0;
      if (excl == 2) return;
      if (i < t) cnt++;
      const bool molv = (P.flags & 2u) && !same;
      const int lk = (iq.x >> kQTypeShift) * P.n_types + (rq.x >> kQTypeShift);
      Real gx, gy, gz;
      if (F64) {
        const double *lp = ljs64 + 4 * lk;
        double ff, e1, e2;
        pair_f64<ELEC, ELSW>(P, e2d, excl != 1 ? lp[0] : lp[2],
                             excl != 1 ? lp[1] : lp[3], id.w * rd.w, excl, ff,
                             e1, e2);
        acc[ACC_ELJ] += 0.5 * e1;
        acc[ACC_EEL] += 0.5 * e2;
        gx = ex * ff;
        gy = ey * ff;
        gz = ez * ff;
        if (molv || (P.flags & 1u))
          add_virial(acc, ACC_VXX, gx, gy, gz, ex, ey, ez, 0.5);
      } else {
        const float4 lp = ljs[lk];
        float ff, e1, e2;
        pair_f32<ELEC>(F, r2, excl != 1 ? lp.x : lp.z, excl != 1 ? lp.y : lp.w,
                       ip.w * r.w, excl, ff, e1, e2);
        pelj += e1;
        peel += e2;
        gx = dx * ff;
        gy = dy * ff;
        gz = dz * ff;
        if (vir && (molv || (P.flags & 1u))) {
          v0 = fmaf(gx, dx, v0);
          v1 = fmaf(gx, dy, v1);
          v2 = fmaf(gx, dz, v2);
          v3 = fmaf(gy, dy, v3);
          v4 = fmaf(gy, dz, v4);
          v5 = fmaf(gz, dz, v5);
        }
      }
      if (molv) {
        const int mj = rq.y;
        double cdx, cdy, cdz, cd2;
        min_diff_exact(rare_box(&R.box), S.com[3 * mi], S.com[3 * mi + 1],
                       S.com[3 * mi + 2], S.com[3 * mj], S.com[3 * mj + 1],
                       S.com[3 * mj + 2], cdx, cdy, cdz, cd2);
        add_virial(acc, ACC_MXX, (double)gx, (double)gy, (double)gz, cdx, cdy,
                   cdz, 0.5);
      }
      Real *fa = &facc[wave][s * SLOTS + lane / LPS];
      auto add3 = [&]() {
        if (ATOM) {
          lds_add(fa, -gx);
          lds_add(fa + KS, -gy);
          lds_add(fa + 2 * KS, -gz);
        } else {
          fa[0] -= gx;
          fa[KS] -= gy;
          fa[2 * KS] -= gz;
        }
      };
      // the LPS lanes of a slot add in LPS separate steps (lane order)
#pragma unroll
      for (int p = 0; p < LPS; p++) {
        if ((lane & (LPS - 1)) == p) add3();
        __builtin_amdgcn_sched_barrier(0);
        __asm__ volatile("" ::: "memory");
      }
    };

    // Evaluates queue entries 0 .. qn - 1 (qn >= 1), 64 at a time; group
    // g + 1's records and group g + 2's entries are read before group g's
    // arithmetic.
    auto eval_queue = [&](int qn) {
      pelj = peel = 0.f;
      v0 = v1 = v2 = v3 = v4 = v5 = 0.f;
      const double4 dz4 = make_double4(0.0, 0.0, 0.0, 0.0);
      unsigned en = que[wave][min(lane, qn - 1)];
      unsigned en1 = que[wave][min(64 + lane, qn - 1)];
      int s = en & (kTileAtoms - 1), k = en >> 4;
      float4 ip = ipos[wave][s], r = recf[wave][k];
      int2 iq = iint[wave][s], rq = reci[wave][k];
      double4 id = dz4, rd = dz4;
      if (F64) {
        id = ipd[wave][s];
        rd = recd[wave][k];
      }
      for (int e0 = 0; e0 < qn; e0 += 64) {
        const int sn = en1 & (kTileAtoms - 1), kn = en1 >> 4;
        const float4 ipn = ipos[wave][sn], rn = recf[wave][kn];
        const int2 iqn = iint[wave][sn], rqn = reci[wave][kn];
        double4 idn = dz4, rdn = dz4;
        if (F64) {
          idn = ipd[wave][sn];
          rdn = recd[wave][kn];
        }
        const unsigned en2 = que[wave][min(e0 + 128 + lane, qn - 1)];
        if (e0 + lane < qn) one(s, ip, iq, id, r, rq, rd);
        s = sn;
        ip = ipn;
        iq = iqn;
        r = rn;
        rq = rqn;
        id = idn;
        rd = rdn;
        en1 = en2;
      }
      if (!F64) {
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
    };

    // Tests records c0 .. c0 + 63 (those below nt) against the wave's atoms
    // and appends the passing entries to the queue.  WX, WY, WZ: per-pair
    // minimum image along that axis.
    const int slot = lane & (kTileAtoms - 1), sub = lane >> 4;
    // WMODE 0: no minimum image, 1: along z, 2: along x, y and z, 3: along
    // the axes whose flag wx, wy, wz is set.
    auto test_group = [&](auto mode, int c0, int nt, int &qn) {
      constexpr int WMODE = decltype(mode)::value;
      const bool WX = WMODE == 2 || (WMODE == 3 && wx);
      const bool WY = WMODE == 2 || (WMODE == 3 && wy);
      const bool WZ = WMODE == 1 || WMODE == 2 || (WMODE == 3 && wz);
#pragma unroll 4
      for (int j = 0; j < 16; j++) {
        const int k = c0 + 4 * j + sub;
        const float4 r = recf[wave][k];
        float dx = r.x - li.x, dy = r.y - li.y, dz = r.z - zi;
        if (WX) dx = fmaf(-Lx, rintf(dx * iLx), dx);
        if (WY) dy = fmaf(-Ly, rintf(dy * iLy), dy);
        if (WZ) dz = fmaf(-Lz, rintf(dz * iLz), dz);
        const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
        const bool pass = own && k < nt && r2 <= hi;
        const unsigned long long b = __ballot(pass);
        const int pos = pass ? qn + lane_prefix(b) : kQCap;
        que[wave][pos] = (unsigned short)((k << 4) | slot);
        qn += __popcll(b);
      }
    };

    // Tests buffered records against the wave's atoms, queueing and
    // evaluating the pairs that pass (the atom itself passes and is skipped
    // by the pair code).  Unless `last`, only whole groups of 64 records are
    // tested and the remaining n % 64 move to the front of the buffer;
    // returns the number of records left buffered.
    const int wmode = (wx ? 1 : 0) | (wy ? 2 : 0) | (wz ? 4 : 0);
    auto process = [&](int n, bool last) -> int {
      const int nt = last ? n : n & ~63;
      wave_sync_lds();
      int qn = 0;
      for (int c0 = 0; c0 < nt; c0 += 64) {
        if (wmode == 0)
          test_group(std::integral_constant<int, 0>{}, c0, nt, qn);
        else if (wmode == 4)
          test_group(std::integral_constant<int, 1>{}, c0, nt, qn);
        else if (wmode == 7)
          test_group(std::integral_constant<int, 2>{}, c0, nt, qn);
        else
          test_group(std::integral_constant<int, 3>{}, c0, nt, qn);
        if (qn > kQCap - kTileAtoms * 64) {
          wave_sync_lds();
          eval_queue(qn);
          qn = 0;
          wave_sync_lds();  // the queue is rewritten next
        }
      }
      if (qn > 0) {
        wave_sync_lds();
        eval_queue(qn);
      }
      const int rem = n - nt;
      wave_sync_lds();
      if (lane < rem) {  // sources nt.. (>= 64) never overlap 0 .. rem - 1
        recf[wave][lane] = recf[wave][nt + lane];
        reci[wave][lane] = reci[wave][nt + lane];
        if (F64) recd[wave][lane] = recd[wave][nt + lane];
      }
      wave_sync_lds();  // the record buffer is refilled next
      return rem;
    };

    for (int oc = 0; oc < 9; oc++) {
      const int ox = oc / 3 - 1, oy = oc % 3 - 1;
      // an axis with a single column takes it once
      if (wx && ox != 0) continue;
      if (wy && oy != 0) continue;
      // column prefilter: gap from the wave box to the neighbour column
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
      // z cells zA..zB (unwrapped) as at most two contiguous sorted runs
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
            recf[wave][pos] = rf;
            reci[wave][pos] = make_int2(t | (S.type[t] << kQTypeShift), S.mol[t]);
            if (F64)
              recd[wave][pos] = make_double4(S.x[t], S.y[t], S.z[t], S.q[t]);
          }
          nbuf += __popcll(ball);
          if (nbuf > BUF - 64) nbuf = process(nbuf, false);
        }
      }
    }
    if (nbuf > 0) process(nbuf, true);
    // combine the slots of each atom in a fixed order
    wave_sync_lds();
    if (lane < 3 * kTileAtoms) {
      const int c = lane >> 4, s = lane & (kTileAtoms - 1);
      if (s < na) {
        const Real *fa = &facc[wave][c * KS + s * SLOTS];
        const int ai = S.atom[a0 + s];
        if (F64) {
          double sum = 0.0;
          for (int k = 0; k < SLOTS; k++)
            sum += (double)fa[(k + s) & (SLOTS - 1)];
          forces[3 * ai + c] += sum;
        } else {
          long long sum = 0;
          for (int k = 0; k < SLOTS; k++)
            sum += to_fixed_f32((float)fa[(k + s) & (SLOTS - 1)]);
          forces[3 * ai + c] += (double)sum * kFixedInv;
        }
      }
    }
    // this item's energies and virials
#pragma unroll
    for (int k = 0; k < ACC_N; k++) {
      double v = acc[k];
      for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
      if (lane == 0) partial[(size_t)k * pstride + item_idx] = v;
    }
    wave_sync_lds();  // facc and the buffers are reused by the next item
  }
  unsigned long long tot = block_count<NW>(cnt);
  if (threadIdx.x == 0 && tot) atomicAdd(npairs, tot);
}

}  // namespace pmnb
