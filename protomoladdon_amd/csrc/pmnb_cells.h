// GPU cell-list build: integer cell hashing, counting sort, per-cell
// ordering and the gather into sorted structure-of-arrays.
//
// Replaces Topology::updateCellLists (topology/Topology.h:44-86) with
// CubicCellManager::findCell (topology/CubicCellManager.cpp:40-44).  The
// reference builds singly linked lists in ascending atom order; here each
// cell is a contiguous segment of the sorted arrays, also in ascending atom
// order, so traversal order inside a cell is the reference's.
#pragma once

#include "pmnb_common.h"
#include "pmnb_tile.h"

namespace pmnb {

// Fills n ints with v (used instead of hipMemsetAsync so every clear is an
// ordinary kernel node if a caller captures the stream into a graph).
__global__ void k_fill_i32(int *p, int n, int v) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_fill_u64(unsigned long long *p, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0ull;
}

// Cell of atom i: findCell(origin - min + minimalPosition(x_i)).
__device__ __forceinline__ void cell_coords(const Box &b, const Grid &g,
                                            double x, double y, double z,
                                            int &cx, int &cy, int &cz,
                                            double &vx, double &vy,
                                            double &vz) {
  double px, py, pz;
  min_pos_exact(b, x, y, z, px, py, pz);
  vx = b.shift[0] + px;
  vy = b.shift[1] + py;
  vz = b.shift[2] + pz;
  double fx = floor(vx * g.inv[0]);
  double fy = floor(vy * g.inv[1]);
  double fz = floor(vz * g.inv[2]);
  // Clamp before the integer conversion so a NaN or huge value cannot
  // produce an undefined conversion.
  fx = (fx == fx) ? fmin(fmax(fx, -2.0), (double)g.n[0] + 1.0) : 0.0;
  fy = (fy == fy) ? fmin(fmax(fy, -2.0), (double)g.n[1] + 1.0) : 0.0;
  fz = (fz == fz) ? fmin(fmax(fz, -2.0), (double)g.n[2] + 1.0) : 0.0;
  int rx = (int)fx, ry = (int)fy, rz = (int)fz;
  cx = wrap_cell(rx, g.n[0]);
  cy = wrap_cell(ry, g.n[1]);
  cz = wrap_cell(rz, g.n[2]);
  // Keep the wrapped coordinate consistent with the wrapped cell.
  if (rx >= g.n[0]) vx -= b.H[0];
  else if (rx < 0) vx += b.H[0];
  if (ry >= g.n[1]) vy -= b.H[1];
  else if (ry < 0) vy += b.H[1];
  if (rz >= g.n[2]) vz -= b.H[2];
  else if (rz < 0) vz += b.H[2];
}

// Cell of every atom; rank[i] is the atom's arrival rank in its cell (any
// order: k_sort_cells restores ascending atom order afterwards).  With
// pos_copy the positions read (from host memory mapped into the device,
// say) are also stored there for the kernels after this one.
__global__ void k_assign(Box b, Grid g, const double *__restrict__ pos,
                         int n, int *__restrict__ cell_of,
                         int *__restrict__ count, int *__restrict__ rank,
                         double *__restrict__ pos_copy) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int cx, cy, cz;
  double vx, vy, vz;
  const double px = pos[3 * i], py = pos[3 * i + 1], pz = pos[3 * i + 2];
  if (pos_copy) {
    pos_copy[3 * i] = px;
    pos_copy[3 * i + 1] = py;
    pos_copy[3 * i + 2] = pz;
  }
  cell_coords(b, g, px, py, pz, cx, cy, cz, vx, vy, vz);
  int c = (cx * g.n[1] + cy) * g.n[2] + cz;
  cell_of[i] = c;
  rank[i] = atomicAdd(&count[c], 1);
}

// Exclusive prefix of v over the 1024 threads of a block (thread order);
// *total receives the block sum.  Wave scans by shuffles, then one wave
// scans the 16 wave sums.  Every thread of the block must call it.
__device__ __forceinline__ int block_scan_1024(int v, int *total) {
  __shared__ int wsum[16];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  int x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) wsum[w] = x;
  __syncthreads();
  if (w == 0) {
    int y = lane < 16 ? wsum[lane] : 0;
#pragma unroll
    for (int o = 1; o < 16; o <<= 1) {
      const int z = __shfl_up(y, o, 64);
      if (lane >= o) y += z;
    }
    if (lane < 16) wsum[lane] = y;
  }
  __syncthreads();
  *total = wsum[15];
  const int off = w > 0 ? wsum[w - 1] : 0;
  __syncthreads();  // wsum may be reused by the next call
  return off + x - v;
}

// Exclusive scan of count[0..n) into start[0..n) by one block of 1024.
__device__ __forceinline__ void scan_counts(const int *__restrict__ count,
                                            int *__restrict__ start, int n) {
  const int t = threadIdx.x;
  const int chunk = (n + 1023) / 1024;
  const int b0 = min(n, t * chunk), b1 = min(n, b0 + chunk);
  int s = 0;
#pragma unroll 4
  for (int i = b0; i < b1; i++) s += count[i];
  int total;
  int run = block_scan_1024(s, &total);  // exclusive prefix of this chunk
#pragma unroll 4
  for (int i = b0; i < b1; i++) {
    const int c = count[i];
    start[i] = run;
    run += c;
  }
}

__global__ void __launch_bounds__(1024)
k_scan(const int *__restrict__ count, int *__restrict__ start, int n) {
  scan_counts(count, start, n);
}

__global__ void k_scatter(const int *__restrict__ cell_of, int n,
                          const int *__restrict__ start,
                          const int *__restrict__ rank,
                          int *__restrict__ sorted) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  sorted[start[cell_of[i]] + rank[i]] = i;
}

constexpr int kSortLds = 512;  // cell entries a wave ranks in LDS

// Ascending atom order inside each cell (the reference's list order): one
// wave per cell sorts up to 64 entries with a bitonic network across lanes;
// up to kSortLds entries by rank (each lane counts, for its entries, the
// smaller keys of the cell held in LDS; keys are distinct atom indices);
// larger cells fall back to an insertion sort by one lane.
__global__ void __launch_bounds__(kBlock)
k_sort_cells(const int *__restrict__ start, const int *__restrict__ count,
             int ncell, int *__restrict__ sorted) {
  __shared__ int keys[kBlock / 64][kSortLds];
  const int w = threadIdx.x >> 6;
  const int c = blockIdx.x * (kBlock / 64) + w;
  const int lane = threadIdx.x & 63;
  if (c >= ncell) return;
  const int m = count[c];
  if (m <= 1) return;
  int *seg = sorted + start[c];
  if (m > 64 && m <= kSortLds) {
    for (int k = lane; k < m; k += 64) keys[w][k] = seg[k];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    for (int k = lane; k < m; k += 64) {
      const int v = keys[w][k];
      int rank = 0;
      for (int j = 0; j < m; j++) rank += keys[w][j] < v;
      seg[rank] = v;
    }
    return;
  }
  if (m <= 64) {
    int v = lane < m ? seg[lane] : 0x7fffffff;
    for (int k = 2; k <= 64; k <<= 1) {
      for (int j = k >> 1; j > 0; j >>= 1) {
        const int other = __shfl_xor(v, j, 64);
        const bool up = (lane & k) == 0;
        const bool lower = (lane & j) == 0;
        v = (lower == up) ? min(v, other) : max(v, other);
      }
    }
    if (lane < m) seg[lane] = v;
  } else if (lane == 0) {
    for (int i = 1; i < m; i++) {
      int v = seg[i];
      int j = i - 1;
      while (j >= 0 && seg[j] > v) {
        seg[j + 1] = seg[j];
        j--;
      }
      seg[j + 1] = v;
    }
  }
}

// Clears the per-cell counters and the pair counter in one launch.
__global__ void k_clear(int *__restrict__ count, int ncell,
                        unsigned long long *__restrict__ npairs) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < ncell) count[i] = 0;
  if (i == 0) *npairs = 0ull;
}

// Gathers per-atom data into sorted order.  loc holds, for the mixed path,
// the FP32 offset of the atom from its cell's lower corner (computed in
// FP64) and its scaled charge.
__global__ void k_gather(Box b, Grid g, const double *__restrict__ pos,
                         const int *__restrict__ sorted,
                         const int *__restrict__ cell_of,
                         const int *__restrict__ type,
                         const double *__restrict__ q,
                         const int *__restrict__ mol, int n,
                         double *__restrict__ sx, double *__restrict__ sy,
                         double *__restrict__ sz, float4 *__restrict__ sloc,
                         int *__restrict__ stype, double *__restrict__ sq,
                         int *__restrict__ smol, int *__restrict__ scell) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  int a = sorted[s];
  double x = pos[3 * a], y = pos[3 * a + 1], z = pos[3 * a + 2];
  sx[s] = x;
  sy[s] = y;
  sz[s] = z;
  int cx, cy, cz;
  double vx, vy, vz;
  cell_coords(b, g, x, y, z, cx, cy, cz, vx, vy, vz);
  float4 l;
  l.x = (float)(vx - (double)cx * g.cs[0]);
  l.y = (float)(vy - (double)cy * g.cs[1]);
  l.z = (float)(vz - (double)cz * g.cs[2]);
  l.w = (float)q[a];
  sloc[s] = l;
  stype[s] = type[a];
  sq[s] = q[a];
  smol[s] = mol[a];
  scell[s] = cell_of[a];
}

// Replica-aware cell assignment and gather of the tile path: atom i belongs
// to replica rep_of[i] (replica 0 when rep_of is null), whose box and
// column grid place it; cells are numbered globally from the replica's
// cell offset.  RepT provides box, grid and cell_off.
// Cells of a block's atoms are few (consecutive atoms are neighbours): the
// block counts its atoms per cell in an LDS table (open addressing on the
// cell index), takes one global atomic add per distinct cell, all issued
// together, and gives each atom the cell's base plus its rank in the block:
// its arrival rank in the cell, at which it is written into the cell's row
// of `bucket` (cap entries per cell; an atom ranked beyond cap is found
// again from cell_of by k_sort_gather_r).  pos_copy as in k_assign.
// Requires blockDim.x == kBlock.
template <class RepT>
__global__ void __launch_bounds__(kBlock)
k_assign_r(const RepT *__restrict__ reps, const int *__restrict__ rep_of,
           const double *__restrict__ pos, int n, int *__restrict__ cell_of,
           int *__restrict__ count, int *__restrict__ bucket, int cap,
           double *__restrict__ pos_copy) {
  __shared__ int key[kBlock], lcount[kBlock], base[kBlock];
  const int t = threadIdx.x;
  const int i = blockIdx.x * kBlock + t;
  key[t] = -1;
  lcount[t] = 0;
  __syncthreads();
  int slot = -1, lrank = 0, c = 0;
  if (i < n) {
    const RepT &R = reps[rep_of ? rep_of[i] : 0];
    int cx, cy, cz;
    double vx, vy, vz;
    const double px = pos[3 * i], py = pos[3 * i + 1], pz = pos[3 * i + 2];
    if (pos_copy) {
      pos_copy[3 * i] = px;
      pos_copy[3 * i + 1] = py;
      pos_copy[3 * i + 2] = pz;
    }
    cell_coords(R.box, R.grid, px, py, pz, cx, cy, cz, vx, vy, vz);
    c = R.cell_off + (cx * R.grid.n[1] + cy) * R.grid.n[2] + cz;
    cell_of[i] = c;
    // at most kBlock distinct cells in kBlock slots: the probe ends
    int h = c & (kBlock - 1);
    while (true) {
      const int old = atomicCAS(&key[h], -1, c);
      if (old == -1 || old == c) break;
      h = (h + 1) & (kBlock - 1);
    }
    slot = h;
    lrank = atomicAdd(&lcount[h], 1);
  }
  __syncthreads();
  if (key[t] >= 0) base[t] = atomicAdd(&count[key[t]], lcount[t]);
  __syncthreads();
  if (i < n) {
    const int r = base[slot] + lrank;
    if (r < cap) bucket[(size_t)c * cap + r] = i;
  }
}

// Per-cell ordering and gather of the tile path in one pass: one wave per
// cell takes the cell's atoms from its bucket row (or, for a cell of more
// than cap atoms, from cell_of), puts them in ascending atom order (as
// k_sort_cells) and writes, for each sorted slot, the atom index, its FP64
// position, the FP32 offset from its cell's lower corner (computed in FP64)
// with its charge, type, charge, molecule, cell and packed record.
template <class RepT>
__global__ void __launch_bounds__(kBlock)
k_sort_gather_r(const RepT *__restrict__ reps, const int *__restrict__ rep_of,
                const double *__restrict__ pos, const int *__restrict__ start,
                const int *__restrict__ count, int ncell,
                const int *__restrict__ bucket, int cap,
                const int *__restrict__ cell_of, int n,
                const int *__restrict__ type, const double *__restrict__ q,
                const int *__restrict__ mol, int *__restrict__ sorted,
                double *__restrict__ sx, double *__restrict__ sy,
                double *__restrict__ sz, float4 *__restrict__ sloc,
                int *__restrict__ stype, double *__restrict__ sq,
                int *__restrict__ smol, int *__restrict__ scell,
                int2 *__restrict__ srec) {
  __shared__ int keys[kBlock / 64][kSortLds];
  __shared__ int found[kBlock / 64];
  const int w = threadIdx.x >> 6;
  const int c = blockIdx.x * (kBlock / 64) + w;
  const int lane = threadIdx.x & 63;
  if (c >= ncell) return;
  const int m = count[c];
  if (m == 0) return;
  const int s0 = start[c];
  int *seg = sorted + s0;
  const int *row = bucket + (size_t)c * cap;  // the first min(m, cap) atoms
  // a cell of more than cap atoms: the wave finds them all in cell_of and
  // appends them to dst (any order; sorted below)
  auto collect = [&](int *dst) {
    if (lane == 0) found[w] = 0;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    for (int i = lane; i < n; i += 64)
      if (cell_of[i] == c) dst[atomicAdd(&found[w], 1)] = i;
  };
  // slot s0 + k receives atom a
  auto put = [&](int k, int a) {
    const int s = s0 + k;
    seg[k] = a;
    const RepT &R = reps[rep_of ? rep_of[a] : 0];
    const double x = pos[3 * a], y = pos[3 * a + 1], z = pos[3 * a + 2];
    sx[s] = x;
    sy[s] = y;
    sz[s] = z;
    int cx, cy, cz;
    double vx, vy, vz;
    cell_coords(R.box, R.grid, x, y, z, cx, cy, cz, vx, vy, vz);
    float4 l;
    l.x = (float)(vx - (double)cx * R.grid.cs[0]);
    l.y = (float)(vy - (double)cy * R.grid.cs[1]);
    l.z = (float)(vz - (double)cz * R.grid.cs[2]);
    l.w = (float)q[a];
    sloc[s] = l;
    stype[s] = type[a];
    sq[s] = q[a];
    smol[s] = mol[a];
    scell[s] = c;
    srec[s] = make_int2(type[a] | (mol[a] << 8), a);
  };
  if (m <= 64) {  // cap >= 64: every atom is in the row
    int v = lane < m ? row[lane] : 0x7fffffff;
    for (int k = 2; k <= 64; k <<= 1) {
      for (int j = k >> 1; j > 0; j >>= 1) {
        const int other = __shfl_xor(v, j, 64);
        const bool up = (lane & k) == 0;
        const bool lower = (lane & j) == 0;
        v = (lower == up) ? min(v, other) : max(v, other);
      }
    }
    if (lane < m) put(lane, v);
  } else if (m <= kSortLds) {
    if (m <= cap) {
      for (int k = lane; k < m; k += 64) keys[w][k] = row[k];
    } else {
      collect(keys[w]);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    for (int k = lane; k < m; k += 64) {
      const int v = keys[w][k];
      int rank = 0;
      for (int j = 0; j < m; j++) rank += keys[w][j] < v;
      put(rank, v);
    }
  } else {
    // m > kSortLds >= cap
    collect(seg);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    if (lane == 0) {
      for (int i = 1; i < m; i++) {
        int v = seg[i];
        int j = i - 1;
        while (j >= 0 && seg[j] > v) {
          seg[j + 1] = seg[j];
          j--;
        }
        seg[j + 1] = v;
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    for (int k = lane; k < m; k += 64) put(k, seg[k]);
  }
}

// Sets n 64-bit integers to zero.
__global__ void k_zero_i64(long long *p, long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// Sets n doubles to zero (a kernel node, not a memset, inside graphs).
__global__ void k_zero_f64(double *p, long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

}  // namespace pmnb
