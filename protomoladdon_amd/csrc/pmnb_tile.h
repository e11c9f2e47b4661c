// Column grid and wave helpers of the tile path (the pair kernels are in
// pmnb_queue.h and pmnb_adjoint.h).
//
// Cell grid: x/y columns at least one cutoff wide (a neighbour's column
// offset is -1, 0 or +1) and z cells of about a third (mixed) or a quarter
// (FP64) of the cutoff; atoms are
// sorted column by column, z cell by z cell, ascending atom index inside a
// cell, so a column's z-range of cells is one contiguous run of the sorted
// arrays (two when it wraps).  One wave handles 16 consecutive sorted atoms
// of one column.
#pragma once

#include "pmnb_common.h"
#include "pmnb_pair.h"

namespace pmnb {

constexpr int kTileAtoms = 16;   // atoms per wave

struct TileGrid {
  int nx, ny, nz, rz;   // grid and z-cell reach of the cutoff
  float csx, csy, csz;  // cell sizes
};

__device__ __forceinline__ float wave_min(float v) {
  for (int m = 32; m > 0; m >>= 1) v = fminf(v, __shfl_xor(v, m, 64));
  return v;
}
__device__ __forceinline__ float wave_max(float v) {
  for (int m = 32; m > 0; m >>= 1) v = fmaxf(v, __shfl_xor(v, m, 64));
  return v;
}

__device__ __forceinline__ void wave_sync_lds() {
  // LDS written by some lanes of this wave is read by other lanes next.
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ int lane_prefix(unsigned long long mask) {
  return __builtin_amdgcn_mbcnt_hi((unsigned)(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo((unsigned)mask, 0u));
}

}  // namespace pmnb
