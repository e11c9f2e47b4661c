// Shared device-side definitions for the nonbonded cutoff kernels.
//
// All code in this library is compiled with -ffp-contract=off so every FP64
// expression rounds exactly as the reference's x86-64 build does (which has
// no FMA); FP32 kernels request fused multiply-adds explicitly with fmaf().
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pmnb {

constexpr int kBlock = 256;        // threads per block of the per-atom kernels
constexpr int kMaxExclDelta = 64;  // exclusion deltas held in the per-atom masks
constexpr double kFixedScale = 4294967296.0;        // 2^32 fixed-point scale
constexpr double kFixedInv = 1.0 / 4294967296.0;

// Accumulator slots of the per-block partial sums.
enum Acc {
  ACC_ELJ = 0, ACC_EEL = 1,
  ACC_VXX = 2, ACC_VXY, ACC_VXZ, ACC_VYY, ACC_VYZ, ACC_VZZ,
  ACC_MXX = 8, ACC_MXY, ACC_MXZ, ACC_MYY, ACC_MYZ, ACC_MZZ,
  ACC_N = 14
};

// Orthogonal periodic box in the reference's representation
// (topology/PeriodicBoundaryConditions.cpp:32-77).
struct Box {
  double e[3];      // diagonal of e1, e2, e3
  double er[3];     // diagonal of e1r, e2r, e3r
  double origin[3];
  double mn[3];     // myMin
  double H[3];      // myMax - myMin
  double H2[3];     // H * 0.5
  double D;         // min((e_k/2)^2)
  double shift[3];  // origin - myMin (Topology.h:58)
};

// Cell grid (topology/CubicCellManager.cpp:46-65,
// topology/ArrayCellListStructure.cpp:40-42).
struct Grid {
  int n[3];
  int ncell;
  double cs[3];
  double inv[3];
};

// Switching constants (switch/C2SwitchingFunction.cpp:12-18,
// switch/C1SwitchingFunction.cpp:12-16).
struct Switch {
  double on2, cut2, s1, s2, s3;
  double c1_15c1, c1_05c3, c1_15c3;
};

// FP32 copies of the constants the mixed-precision kernels use (computed
// on the host so they arrive in scalar registers).
struct F32Params {
  float cut2, band, Lx, Ly, Lz, iLx, iLy, iLz;
  float on2l, cut2l, s1l, s2l, s3l, on2e, cut2e, s1e, s2e, s3e;
  float alpha, tapi, csf;
  int same_switch;  // LJ and electrostatic switches are the same C2
};

// Internal evaluation flag (never a caller's flag): the tile path writes
// the forces and the 20 output doubles instead of adding to them, so the
// caller need not clear them first.
constexpr uint32_t kFlagFresh = 1u << 31;

struct ForceParams {
  Box box;
  Grid grid;
  Switch sw_lj, sw_el;
  double cut2;           // max of the two squared switching cutoffs
  double alpha, two_alpha_pi, csf;
  int n_atoms, n_types, ndelta;
  int elec;              // PMNB_ELEC_*
  int el_switch;         // PMNB_SWITCH_*
  int max_excl_delta;
  int per_pair_wrap;     // mixed path: grid too small for cell-delta images
  uint32_t flags;
  float band;            // |d2_f32 - cut2| below which FP64 re-tests
  F32Params f;
};

// ---------------------------------------------------------------- FP64 exact
// PeriodicBoundaryConditions::minimalDifference(c1, c2, distSquared),
// topology/PeriodicBoundaryConditions.h:72-110 (orthogonal branch).
__device__ __forceinline__ void min_diff_exact(const Box &b, double c1x,
                                               double c1y, double c1z,
                                               double c2x, double c2y,
                                               double c2z, double &dx,
                                               double &dy, double &dz,
                                               double &d2) {
  dx = c2x - c1x;
  dy = c2y - c1y;
  dz = c2z - c1z;
  d2 = dx * dx + dy * dy + dz * dz;
  if (d2 > b.D) {
    if (dx < -b.H2[0]) dx += b.H[0];
    else if (dx > b.H2[0]) dx -= b.H[0];
    if (dy < -b.H2[1]) dy += b.H[1];
    else if (dy > b.H2[1]) dy -= b.H[1];
    if (dz < -b.H2[2]) dz += b.H[2];
    else if (dz > b.H2[2]) dz -= b.H[2];
    d2 = dx * dx + dy * dy + dz * dz;
    if (d2 > b.D) {
      dx -= b.e[0] * __builtin_rint(b.er[0] * dx);
      dy -= b.e[1] * __builtin_rint(b.er[1] * dy);
      dz -= b.e[2] * __builtin_rint(b.er[2] * dz);
      d2 = dx * dx + dy * dy + dz * dz;
    }
  }
}

// PeriodicBoundaryConditions::minimalPosition,
// topology/PeriodicBoundaryConditions.h:113-147 (orthogonal branch).
__device__ __forceinline__ void min_pos_exact(const Box &b, double x, double y,
                                              double z, double &ox, double &oy,
                                              double &oz) {
  double dx = x - b.origin[0], dy = y - b.origin[1], dz = z - b.origin[2];
  if (dx * dx + dy * dy + dz * dz <= b.D) {
    ox = dx; oy = dy; oz = dz;
    return;
  }
  if (dx < -b.H2[0]) dx += b.H[0];
  else if (dx > b.H2[0]) dx -= b.H[0];
  if (dy < -b.H2[1]) dy += b.H[1];
  else if (dy > b.H2[1]) dy -= b.H[1];
  if (dz < -b.H2[2]) dz += b.H[2];
  else if (dz > b.H2[2]) dz -= b.H[2];
  if (dx * dx + dy * dy + dz * dz <= b.D) {
    ox = dx; oy = dy; oz = dz;
    return;
  }
  ox = dx - b.e[0] * __builtin_rint(b.er[0] * dx);
  oy = dy - b.e[1] * __builtin_rint(b.er[1] * dy);
  oz = dz - b.e[2] * __builtin_rint(b.er[2] * dz);
}

__device__ __forceinline__ int wrap_cell(int c, int n) {
  // The reference leaves -1 and n unwrapped (CubicCellManager.cpp:40-44 then
  // an unchecked store, Topology.h:72-77); both are the periodic neighbours
  // of a face, so they are wrapped.  Anything else (NaN positions) is
  // clamped so no index leaves the grid.
  if (c >= 0 && c < n) return c;
  if (c == -1) return n - 1;
  if (c == n) return 0;
  int r = c % n;
  if (r < 0) r += n;
  return (r >= 0 && r < n) ? r : 0;
}

// C2SwitchingFunction::operator() (switch/C2SwitchingFunction.h:42-53).
__device__ __forceinline__ void c2_exact(const Switch &s, double d2,
                                         double &v, double &dv) {
  dv = 0.0;
  if (d2 > s.cut2) {
    v = 0.0;
  } else if (d2 >= s.on2) {
    double c2 = s.cut2 - d2;
    double c4 = c2 * (s.s2 + 2.0 * d2);
    v = s.s1 * (c2 * c4);
    dv = s.s3 * (c2 * c2 - c4);
  } else {
    v = 1.0;
  }
}

// C1SwitchingFunction::operator() (switch/C1SwitchingFunction.h:41-50).
__device__ __forceinline__ void c1_exact(const Switch &s, double d2,
                                         double &v, double &dv) {
  if (d2 > s.cut2) {
    v = 0.0;
    dv = 0.0;
    return;
  }
  double dist = sqrt(d2);
  v = 1.0 - dist * (s.c1_15c1 - d2 * s.c1_05c3);
  dv = -s.c1_15c1 / dist + dist * s.c1_15c3;
}

// poly5 (base/MathUtilities.cpp:39-49): Abramowitz & Stegun 7.1.26.
__device__ __forceinline__ double poly5_exact(double ar) {
  const double E1 = 0.254829592, E2 = -0.284496736, E3 = 1.421413741,
               E4 = -1.453152027, E5 = 1.061405429, PP = 0.3275911;
  double t = 1.0 / (1.0 + PP * ar);
  return (t) * (E1 + (t) * (E2 + (t) * (E3 + (t) * (E4 + (t) * E5))));
}

// Exclusion class of atoms a, b of the same molecule from the per-atom
// masks (bit d-1 of full[min]/mod[min] holds pair (min, min + d)).
// Exclusion class of the pair (a, a + d) from atom a's own masks m =
// {forward FULL, forward MODIFIED, reverse FULL, reverse MODIFIED} (bit
// |d| - 1; forward for d > 0, reverse for d < 0): the same answer as
// excl_class, without reading the partner's masks.
__device__ __forceinline__ int excl_own(const unsigned long long *m, int d) {
  const int ad = d < 0 ? -d : d;
  if (ad < 1 || ad > 64) return 0;
  const unsigned long long bit = 1ull << (ad - 1);
  const int o = d > 0 ? 0 : 2;
  if (m[o] & bit) return 2;
  if (m[o + 1] & bit) return 1;
  return 0;
}

__device__ __forceinline__ int excl_class(const unsigned long long *full,
                                          const unsigned long long *mod,
                                          int max_delta, int a, int b) {
  int lo = a < b ? a : b;
  int d = a < b ? b - a : a - b;
  if (d == 0 || d > max_delta) return 0;
  unsigned long long bit = 1ull << (d - 1);
  if (full[lo] & bit) return 2;
  if (mod[lo] & bit) return 1;
  return 0;
}

// Record of the pairs a force kernel evaluates (parity tests): when out is
// not null every evaluated pair appends (i, j, class) with i < j original
// indices; *count counts all of them, out holds the first cap.
struct PairLog {
  int *out;
  long long cap;
  unsigned long long *count;
};

__device__ __forceinline__ void log_pair(const PairLog &lg, int a, int b,
                                         int cls) {
  const unsigned long long k = atomicAdd(lg.count, 1ull);
  if ((long long)k < lg.cap) {
    lg.out[3 * k] = a < b ? a : b;
    lg.out[3 * k + 1] = a < b ? b : a;
    lg.out[3 * k + 2] = cls;
  }
}

}  // namespace pmnb
