// Pieces the tile-path pair kernels (pmnb_tilemask.h) and the adjoint
// (pmnb_adjoint.h) share: the geometry of one replica (TileRep), the work
// items (one per 16 consecutive sorted atoms of one x/y column of one
// replica) built in the same block as the cell starts (k_scan_items), the
// per-replica reduction of the items' partial rows in a fixed order
// (finish_rep_block), fixed-point conversion, LDS adds and scalar-register
// helpers.
//
// Each replica has its own column grid (TileRep).  An axis with a single
// column (box shorter than two cutoffs) is taken once with the minimum image
// of every pair along it (as z is when the cutoff reaches all z cells), so
// small boxes keep the reference's one-image pair set.
//
// The adjoint's records carry {sorted index | type << 24, molecule}; the LJ
// table is held in LDS (at most kQLj entries).
#pragma once

#include <type_traits>

#include "pmnb_cells.h"
#include "pmnb_list.h"
#include "pmnb_tile.h"

namespace pmnb {

constexpr int kQCap = 2048;   // adjoint pair-queue entries per wave
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
// one block: the items read the starts this block has just written.  It
// also clears the pair counter of the evaluation.
__global__ void __launch_bounds__(1024)
k_scan_items(const int *__restrict__ count, int *__restrict__ start,
             int ncell, const TileRep *__restrict__ reps,
             const int *__restrict__ col_rep, int ncol, int nrep,
             int4 *__restrict__ items, int *__restrict__ nitems,
             int *__restrict__ rep_first,
             unsigned long long *__restrict__ npairs) {
  if (threadIdx.x == 0) *npairs = 0ull;
  scan_counts(count, start, ncell);
  __syncthreads();
  build_items(reps, col_rep, start, count, ncol, nrep, items, nitems,
              rep_first);
}

// The same starts and items for many cells, in three launches over the
// global columns (a column's cells are consecutive, and columns are
// numbered in cell order), so the loads spread over the chip instead of
// one block.  Columns first: the atoms of every column.
__device__ __forceinline__ void col_cells(const TileRep *__restrict__ reps,
                                          const int *__restrict__ col_rep,
                                          int gc, int &r, int &lc,
                                          int &first, int &nz) {
  r = col_rep[gc];
  const TileRep &R = reps[r];
  nz = R.tg.nz;
  lc = gc - R.col_off;
  first = R.cell_off + lc * nz;
}

__global__ void k_col_atoms(const int *__restrict__ count,
                            const TileRep *__restrict__ reps,
                            const int *__restrict__ col_rep, int ncol,
                            int *__restrict__ col_n) {
  const int gc = blockIdx.x * blockDim.x + threadIdx.x;
  if (gc >= ncol) return;
  int r, lc, first, nz;
  col_cells(reps, col_rep, gc, r, lc, first, nz);
  int a = 0;
  for (int z = 0; z < nz; z++) a += count[first + z];
  col_n[gc] = a;
}

// One block of 1024: each column's first atom (col_a0) and first work item
// (col_i0), exclusive scans in column order; the item count, the end of
// rep_first, and the pair counter cleared.
__global__ void __launch_bounds__(1024)
k_col_scan(const int *__restrict__ col_n, int ncol, int nrep,
           int *__restrict__ col_a0, int *__restrict__ col_i0,
           int *__restrict__ nitems, int *__restrict__ rep_first,
           unsigned long long *__restrict__ npairs) {
  const int t = threadIdx.x;
  if (t == 0) *npairs = 0ull;
  const int chunk = (ncol + 1023) / 1024;
  const int b0 = min(ncol, t * chunk), b1 = min(ncol, b0 + chunk);
  int sa = 0, si = 0;
  for (int gc = b0; gc < b1; gc++) {
    const int a = col_n[gc];
    sa += a;
    si += (a + kTileAtoms - 1) / kTileAtoms;
  }
  int ta, ti;
  int oa = block_scan_1024(sa, &ta);
  int oi = block_scan_1024(si, &ti);
  for (int gc = b0; gc < b1; gc++) {
    const int a = col_n[gc];
    col_a0[gc] = oa;
    col_i0[gc] = oi;
    oa += a;
    oi += (a + kTileAtoms - 1) / kTileAtoms;
  }
  if (t == 1023) {
    *nitems = ti;
    rep_first[nrep] = ti;
  }
}

// Per column: the starts of its cells and its work items (as build_items).
__global__ void k_col_fill(const int *__restrict__ count,
                           int *__restrict__ start,
                           const TileRep *__restrict__ reps,
                           const int *__restrict__ col_rep, int ncol,
                           const int *__restrict__ col_a0,
                           const int *__restrict__ col_i0,
                           int4 *__restrict__ items,
                           int *__restrict__ rep_first) {
  const int gc = blockIdx.x * blockDim.x + threadIdx.x;
  if (gc >= ncol) return;
  int r, lc, first, nz;
  col_cells(reps, col_rep, gc, r, lc, first, nz);
  const int a0 = col_a0[gc];
  int o = col_i0[gc];
  int run = a0;
  for (int z = 0; z < nz; z++) {
    start[first + z] = run;
    run += count[first + z];
  }
  if (lc == 0) rep_first[r] = o;
  for (int a = a0; a < run; a += kTileAtoms)
    items[o++] = make_int4(a, min(a + kTileAtoms, run), lc, r);
}

// Per-replica sum of one accumulator of the item partials (accumulator k
// of item b at partial[k * pstride + b]) by one block of 1024 threads with
// a fixed reduction tree, added into out[20 r + ..]:
// [0] LJ, [1] electrostatic, [2..10] virial 3x3, [11..19] molecular virial.
__device__ __forceinline__ void finish_rep_block(
    const double *__restrict__ partial, int pstride,
    const int *__restrict__ rep_first, double *__restrict__ out, int r,
    int k, int fresh = 0) {
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
  // fresh: every one of the 20 slots is written by exactly one accumulator
  auto put = [&](int slot) { o[slot] = fresh ? v : o[slot] + v; };
  if (k < 2) {
    put(k);
  } else {
    const int base = k < ACC_MXX ? 2 : 11;
    const int u = k < ACC_MXX ? k - ACC_VXX : k - ACC_MXX;
    // u: 0 xx, 1 xy, 2 xz, 3 yy, 4 yz, 5 zz
    switch (u) {
      case 0: put(base + 0); break;
      case 1: put(base + 1); put(base + 3); break;
      case 2: put(base + 2); put(base + 6); break;
      case 3: put(base + 4); break;
      case 4: put(base + 5); put(base + 7); break;
      default: put(base + 8); break;
    }
  }
}

}  // namespace pmnb
