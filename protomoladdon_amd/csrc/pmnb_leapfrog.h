// pmnb_leapfrog.h -- device-resident LeapfrogIntegrator::run with the SHAKE
// (post-drift) and RATTLE (post-step) modifiers.
//
// Reference (paths relative to protomol/src/protomol):
//   integrator/leapfrog/LeapfrogIntegrator.cpp:33-99  run, doHalfKickdoDrift,
//                                                     doKickdoDrift
//   integrator/StandardIntegrator.cpp:91-116          doHalfKick, doKick
//   integrator/STSIntegrator.cpp:28-51                doDrift, calculateForces
//   modifier/ModifierShake.cpp:16-89                  SHAKE sweep
//   modifier/ModifierRattle.cpp:17-71                 RATTLE sweep
//   modifier/ModifierMetaShake.cpp:11-33              SHAKE error
//   modifier/ModifierMetaRattle.cpp:11-30             RATTLE error
//
// Layout.  Constraints are grouped into connected components ("groups": a
// water is one group of 3 atoms and 3 constraints).  One thread owns one
// group and keeps its atoms' positions, velocities, previous positions and
// inverse masses in registers; the group's constraints are applied in the
// reference's list order, so each sweep is the reference's arithmetic for
// those atoms (constraints of different groups touch disjoint atoms, so the
// order across groups does not change any value).  Atoms in no constraint
// ("free" atoms) get one thread each in the blocks after the group blocks.
//
// Sweep count.  The reference sweeps every constraint until the mean error
// over all constraints is <= epsilon (at most max_iter + 1 sweeps), a
// decision over the whole system.  A "probe" records the summed group
// errors after 0, 1, ..., L-1 sweeps (one partial per block); the kernel
// that applies the stage reduces these partials in a fixed order (every
// block computes the same value), finds the reference's sweep count k, and
// sweeps k times from the stage input (or from the state the probe kept
// after the previous count, below).  L is the previous call's sweep count of
// the same stage plus a margin (every level on the first call); when the
// error is still above epsilon after L levels, the apply kernel's blocks
// compute the remaining levels themselves from the state the probe stored
// after L sweeps (lf_fallback).  No host synchronisation, no atomics on the
// decision.
//
// The global error is summed per group then per block rather than in the
// reference's constraint order, and each constraint's relative error is
// multiplied by a precomputed 1/(2 r^2): the sweep count can differ from
// the reference's only when the mean error is within rounding of epsilon.
// The positions and velocities of a given sweep count are the reference's
// arithmetic (FP contraction is off for the whole library).
//
// Replicas.  A batch of independent systems (pmnb_batch) runs as one
// integrator: each replica's groups occupy their own run of blocks (the
// replica's first group block gb0, gnb blocks), its errors are summed over
// those blocks only and its sweep count, hint and iteration record are its
// own, as when the replica runs alone.  A single system is one replica.
//
// Cells.  The kernel that clears the forces before an evaluation (ZERO) also
// does the evaluation's cell assignment on the tile path when given the
// cell arrays (as k_assign_r: cell of every atom, one global count atomic
// per distinct cell of a block, bucket slot), from the positions it has
// just written, which saves the evaluation that launch (single systems; a
// batch measured faster with the evaluation's own launch).
#pragma once

#include "pmnb_queue.h"

namespace pmnb {

constexpr int kLfBlock = 256;
constexpr int kLfWaves = kLfBlock / 64;
constexpr int kLfFirstLevels = 32;  // first probe's levels with no history
constexpr int kLfMargin = 1;        // levels past the previous sweep count
constexpr int kLfMaxAtoms = 4;      // largest group
constexpr int kLfMaxCons = 6;       // most constraints in a group

// Group shapes: the register layout (atoms, constraints) and whether the
// constraint pattern is fixed at compile time.  kShapeWater: atoms O < H1 <
// H2 with the constraints (H1,H2), (H1,O), (H2,O) of
// topology/TopologyUtilities.cpp:732-737.
enum LfShapeId { kShape33 = 0, kShape46 = 1, kShapeWater = 2 };
template <int S> struct LfShape;
template <> struct LfShape<kShape33> {
  static constexpr int NA = 3, NC = 3;
  static constexpr bool fixed = false;
  static constexpr int A[3] = {0, 0, 0}, B[3] = {0, 0, 0};
};
template <> struct LfShape<kShape46> {
  static constexpr int NA = 4, NC = 6;
  static constexpr bool fixed = false;
  static constexpr int A[6] = {0, 0, 0, 0, 0, 0}, B[6] = {0, 0, 0, 0, 0, 0};
};
template <> struct LfShape<kShapeWater> {
  static constexpr int NA = 3, NC = 3;
  static constexpr bool fixed = true;
  static constexpr int A[3] = {1, 1, 2}, B[3] = {2, 0, 0};
};

enum LfStage { kStNone = 0, kStShake = 1, kStRattle = 2 };
enum LfPre {
  kPreNone = 0,
  kPreFirst = 1,      // doHalfKickdoDrift: v += F*h*0.5/m; x += v*h
  kPreHalf = 2,       // doHalfKick: v += F*(h/2)/m
  kPreHalfDrift = 3,  // doHalfKick then doDrift
  kPreFullDrift = 4,  // doKickdoDrift without modifiers: v += F*h/m; x += v*h
};

// One replica's constraint groups: blocks [gb0, gb0 + gnb) hold its groups
// goff .. goff + ng - 1, 256 to a block; ncons constraints.
struct LfRep {
  int gb0, gnb, goff, ng, ncons;
};

struct LfDev {
  double *x, *v, *f, *out;  // positions, velocities, forces, energy blocks
  double *last;             // SHAKE's previous positions
  const double *mass, *rmass;
  const int *g_atoms;       // NA per group, -1 padded
  const int *g_cons;        // NC per group: local a | local b << 4, -1 padded
  const double *g_rest;     // NC per group: rest length
  const double *g_rinv;     // NC per group: 1 / (2 rest^2)
  const int *free_atoms;
  int n_groups, n_free, ngb;  // ngb: blocks of group threads (all replicas)
  int n_cons;
  int n_out;                // doubles of the energy blocks cleared with forces
  const LfRep *reps;        // per replica
  const int *blk_rep;       // replica of each group block
  int n_reps;
  double h, h2, dtc;        // ts/TIMEFACTOR-scaled steps (see host)
  double eps[3];            // by stage
  int maxit[3];
  double *partial[3];       // by stage: (maxit + 1) levels x ngb blocks
  int *hint;                // [8 rep + 2 stage + call parity]: sweep count
  double *xs, *vs;          // state after the first probe (per atom)
  double *xk, *vk, *vr;     // state after the predicted sweep count, stored
                            // by the probe: SHAKE x, v; RATTLE v (per atom)
  int *iters;               // sweep count per stage call and replica
                            // ([call * n_reps + rep], optional)
  long long iters_cap;
  int *nonconv;             // [3] stage calls that hit max_iter
  // cell assignment of the next evaluation (tile path; c_count null: none)
  const TileRep *c_reps;
  const int *c_rep_of;      // replica of each atom (null: one replica)
  int *c_cell_of, *c_count, *c_bucket;
  int c_cap;                // atom slots per cell of c_bucket
  int *z_count;             // cell counts to clear (null: none)
  int z_nc;
};

struct D3 {
  double x, y, z;
};

__device__ __forceinline__ D3 d3_sub(D3 a, D3 b) {
  return {a.x - b.x, a.y - b.y, a.z - b.z};
}
__device__ __forceinline__ double d3_dot(D3 a, D3 b) {
  return a.x * b.x + a.y * b.y + a.z * b.z;
}
__device__ __forceinline__ D3 ld3(const double *p, int i) {
  return {p[3 * i], p[3 * i + 1], p[3 * i + 2]};
}
__device__ __forceinline__ void st3(double *p, int i, D3 v) {
  p[3 * i] = v.x;
  p[3 * i + 1] = v.y;
  p[3 * i + 2] = v.z;
}

// Element i of a register array, and its update, written as recursive
// templates so that every array index is a constant from the start: an
// index chosen by a select or a branch lets the optimiser address the
// array through a pointer, which moves it to scratch memory.
template <int K, int NA>
__device__ __forceinline__ D3 pick3_r(const D3 (&a)[NA], int i) {
  if constexpr (K == 0) {
    return a[0];
  } else {
    const D3 r = pick3_r<K - 1>(a, i);
    const bool s = i == K;
    return {s ? a[K].x : r.x, s ? a[K].y : r.y, s ? a[K].z : r.z};
  }
}
template <int NA>
__device__ __forceinline__ D3 pick3(const D3 (&a)[NA], int i) {
  return pick3_r<NA - 1>(a, i);
}
template <int K, int NA>
__device__ __forceinline__ double pick1_r(const double (&a)[NA], int i) {
  if constexpr (K == 0) return a[0];
  else return i == K ? a[K] : pick1_r<K - 1>(a, i);
}
template <int NA>
__device__ __forceinline__ double pick1(const double (&a)[NA], int i) {
  return pick1_r<NA - 1>(a, i);
}
template <int K, int NA>
__device__ __forceinline__ void put3_r(D3 (&a)[NA], int i, D3 n) {
  const bool s = i == K;
  a[K] = {s ? n.x : a[K].x, s ? n.y : a[K].y, s ? n.z : a[K].z};
  if constexpr (K > 0) put3_r<K - 1>(a, i, n);
}
// a[i] += d * s (SGN = +1) or a[i] -= d * s (SGN = -1), per component
template <int SGN, int NA>
__device__ __forceinline__ void upd3(D3 (&a)[NA], int i, D3 d, double s) {
  const D3 c = pick3(a, i);
  const D3 n = SGN > 0 ? D3{c.x + d.x * s, c.y + d.y * s, c.z + d.z * s}
                       : D3{c.x - d.x * s, c.y - d.y * s, c.z - d.z * s};
  put3_r<NA - 1>(a, i, n);
}

template <class SH>
struct LfGroup {
  D3 x[SH::NA], v[SH::NA], l[SH::NA];
  double rm[SH::NA];
  int idx[SH::NA];
  int cn[SH::NC];
  double rest[SH::NC], rinv[SH::NC];
};

// Local atom indices of constraint C: compile-time constants for a fixed
// shape, otherwise from the group's table (-1: padding).
template <class SH, int C>
__device__ __forceinline__ bool cons_ab(const LfGroup<SH> &g, int &a, int &b) {
  if constexpr (SH::fixed) {
    a = SH::A[C];
    b = SH::B[C];
    return true;
  } else {
    a = g.cn[C] & 15;
    b = g.cn[C] >> 4;
    return g.cn[C] >= 0;
  }
}

// ModifierShake::doExecute for constraint C onwards (one sweep).  With VEL
// false the velocity updates are left out (they never feed back into
// positions or the error; the probe uses this).
template <bool VEL, class SH, int C = 0>
__device__ __forceinline__ void shake_sweep(LfGroup<SH> &g, double dt) {
  if constexpr (C < SH::NC) {
    int a, b;
    if (cons_ab<SH, C>(g, a, b)) {
      const D3 pab = d3_sub(pick3(g.x, a), pick3(g.x, b));
      const double pabsq = d3_dot(pab, pab);
      const double rabsq = g.rest[C] * g.rest[C];
      const double diffsq = rabsq - pabsq;
      const D3 rab = d3_sub(pick3(g.l, a), pick3(g.l, b));
      const double rpab = d3_dot(rab, pab);
      const double rM1 = pick1(g.rm, a), rM2 = pick1(g.rm, b);
      const double gab = diffsq / (2 * (rM1 + rM2) * rpab);
      D3 dp = {rab.x * gab, rab.y * gab, rab.z * gab};
      upd3<1>(g.x, a, dp, rM1);
      upd3<-1>(g.x, b, dp, rM2);
      if (VEL) {
        dp = {dp.x / dt, dp.y / dt, dp.z / dt};
        upd3<1>(g.v, a, dp, rM1);
        upd3<-1>(g.v, b, dp, rM2);
      }
    }
    shake_sweep<VEL, SH, C + 1>(g, dt);
  }
}

// ModifierRattle::doExecute for constraint C onwards (one sweep; positions
// are not changed).
template <class SH, int C = 0>
__device__ __forceinline__ void rattle_sweep(LfGroup<SH> &g, double dt) {
  if constexpr (C < SH::NC) {
    int a, b;
    if (cons_ab<SH, C>(g, a, b)) {
      const double rM1 = pick1(g.rm, a), rM2 = pick1(g.rm, b);
      const D3 rab = d3_sub(pick3(g.x, a), pick3(g.x, b));
      const double rabsq = d3_dot(rab, rab);
      const D3 vab = d3_sub(pick3(g.v, a), pick3(g.v, b));
      const double rvab = d3_dot(rab, vab);
      const double gab = -rvab / (dt * (rM1 + rM2) * rabsq);
      const D3 dp = {rab.x * gab, rab.y * gab, rab.z * gab};
      const D3 dpt = {dp.x * dt, dp.y * dt, dp.z * dt};
      upd3<1>(g.v, a, dpt, rM1);
      upd3<-1>(g.v, b, dpt, rM2);
    }
    rattle_sweep<SH, C + 1>(g, dt);
  }
}

// Sum over the group's constraints of the reference's per-constraint error
// (ModifierMetaShake::calcError / ModifierMetaRattle::calcError).
template <int STAGE, class SH, int C = 0>
__device__ __forceinline__ double group_error(const LfGroup<SH> &g) {
  if constexpr (C < SH::NC) {
    double e = 0;
    int a, b;
    if (cons_ab<SH, C>(g, a, b)) {
      if (STAGE == kStShake) {
        const D3 d = d3_sub(pick3(g.x, b), pick3(g.x, a));
        const double r2 = g.rest[C] * g.rest[C];
        e = fabs(d3_dot(d, d) - r2) * g.rinv[C];
      } else {
        const D3 rab = d3_sub(pick3(g.x, a), pick3(g.x, b));
        const D3 vab = d3_sub(pick3(g.v, a), pick3(g.v, b));
        e = fabs(d3_dot(rab, vab));
      }
    }
    return e + group_error<STAGE, SH, C + 1>(g);
  } else {
    return 0.0;
  }
}

template <int STAGE, class SH>
__device__ __forceinline__ void sweep(LfGroup<SH> &g, double dt, bool vel) {
  if (STAGE == kStShake) {
    if (vel) shake_sweep<true>(g, dt);
    else shake_sweep<false>(g, dt);
  } else {
    rattle_sweep(g, dt);
  }
}

// The per-atom part of the step between modifier calls.
template <int PRE>
__device__ __forceinline__ void pre_op(D3 &x, D3 &v, D3 f, double m,
                                       const LfDev &d) {
  if (PRE == kPreFirst) {
    v.x += f.x * d.h * 0.5 / m;
    v.y += f.y * d.h * 0.5 / m;
    v.z += f.z * d.h * 0.5 / m;
  } else if (PRE == kPreHalf || PRE == kPreHalfDrift) {
    v.x += f.x * d.h2 / m;
    v.y += f.y * d.h2 / m;
    v.z += f.z * d.h2 / m;
  } else if (PRE == kPreFullDrift) {
    v.x += f.x * d.h / m;
    v.y += f.y * d.h / m;
    v.z += f.z * d.h / m;
  }
  if (PRE == kPreFirst || PRE == kPreHalfDrift || PRE == kPreFullDrift) {
    x.x += v.x * d.h;
    x.y += v.y * d.h;
    x.z += v.z * d.h;
  }
}

// The previous call's sweep count of a stage (-1 before the first call):
// the same value in every kernel of call `scall` (the apply kernel of call c
// writes slot c & 1, calls read slot (c + 1) & 1).
__device__ __forceinline__ int lf_hint(const LfDev &d, int stage,
                                       long long scall, int rep) {
  return d.hint[8 * rep + stage * 2 + (int)((scall + 1) & 1)];
}

// Levels the first probe of stage call `scall` records: the previous call's
// sweep count plus a margin.
__device__ __forceinline__ int lf_first_levels(const LfDev &d, int stage,
                                               long long scall, int rep) {
  const int hint = lf_hint(d, stage, scall, rep);
  const int L = hint < 0 ? kLfFirstLevels : hint + kLfMargin + 1;
  return min(max(L, 1), d.maxit[stage] + 1);
}

// Predicted sweep count: the first probe of the call stores the state after
// this many sweeps (xk, vk or vr), and the apply starts from it when the
// decided count is at least as large; 0: nothing stored.  It never exceeds
// the first probe's levels, after the last of which the probe still sweeps.
// A batch of replicas keeps nothing: with many waves per SIMD the sweeps
// cost less than moving the kept state through memory (measured on 512
// water replicas); the results are the same bits either way.
__device__ __forceinline__ int lf_kept_level(const LfDev &d, int stage,
                                             long long scall, int rep) {
  if (d.n_reps > 1) return 0;
  return max(lf_hint(d, stage, scall, rep), 0);
}

// First level in [j0, j1) whose mean error over the replica's constraints
// is <= epsilon, or -1.  Block-wide; every block of the replica sums its
// per-block partials in the same order, so they all return the same level.
// A thread sums blocks sub, sub + 8, ... (relative to the replica's first)
// of one level; its loads are issued in batches before the (ordered)
// additions so that they overlap.  The first wave then tests the levels of
// the chunk in parallel (one lane per level) and takes the lowest passing.
__device__ __forceinline__ int lf_scan(const LfDev &d, int stage, int j0,
                                       int j1, const LfRep &R) {
  constexpr int kSub = 8, kChunk = kLfBlock / kSub, kBatch = 8;
  static_assert(kChunk <= 64, "one lane per level of a chunk");
  __shared__ double red[kSub][kChunk];
  __shared__ int hit;
  const int t = threadIdx.x;
  const double *P = d.partial[stage] + R.gb0;
  const int nb = R.gnb;
  for (int c0 = j0; c0 < j1; c0 += kChunk) {
    const int n = min(kChunk, j1 - c0);
    __syncthreads();
    {
      const int jj = t % kChunk, sub = t / kChunk;
      double s = 0;
      if (jj < n) {
        const double *row = P + (size_t)(c0 + jj) * d.ngb;
        for (int q0 = sub; q0 < nb; q0 += kSub * kBatch) {
          double v[kBatch];
#pragma unroll
          for (int u = 0; u < kBatch; u++) {
            const int q = q0 + u * kSub;
            v[u] = q < nb ? row[q] : 0.0;
          }
#pragma unroll
          for (int u = 0; u < kBatch; u++) s += v[u];
        }
      }
      red[sub][jj] = s;
    }
    __syncthreads();
    if (t < 64) {
      bool pass = false;
      if (t < n) {
        double s = 0;
#pragma unroll
        for (int k = 0; k < kSub; k++) s += red[k][t];
        pass = s / R.ncons <= d.eps[stage];
      }
      const unsigned long long m = __ballot(pass);
      if (t == 0) hit = m ? c0 + __builtin_ctzll(m) : -1;
    }
    __syncthreads();
    const int hv = hit;
    if (hv >= 0) return hv;
  }
  return -1;
}

// Records the error levels [j0, j1) of a stage for this block's groups from
// the state in g (active threads only): one partial per block and level.
// The levels run back to back, each thread leaving its group's error in
// LDS; each chunk of levels is then summed in a fixed order (16 runs of 16
// threads per level, then the 16 run sums).  With `store`, the state after
// j1 - j0 sweeps is left in xs / vs.  With kept > 0 (and j0 = 0), the full
// state after `kept` sweeps (velocities updated as the apply does) is left
// in xk, vk (SHAKE) or vr (RATTLE).  Block-wide (all threads call it).
constexpr int kLfLevels = 16;                  // levels per chunk
constexpr int kLfRun = kLfBlock / kLfLevels;    // threads per run
static_assert(kLfRun == 16, "reduction shape");

// LDS of the level reductions, one object per kernel shared by the probe
// and the apply kernel's fallback: each thread's group error per level of a
// chunk, and the run sums.
__device__ __forceinline__ double (&lf_elev())[kLfLevels][kLfBlock] {
  __shared__ double e[kLfLevels][kLfBlock];
  return e;
}
__device__ __forceinline__ double (&lf_runs())[kLfRun][kLfLevels] {
  __shared__ double r[kLfRun][kLfLevels];
  return r;
}

// The block sums of the first n levels in lf_elev() in a fixed order (16
// runs of 16 threads per level, then the run sums): thread t < n returns
// level t's.  Block-wide.
__device__ __forceinline__ double lf_level_sums(int n) {
  auto &elev = lf_elev();
  auto &runs = lf_runs();
  const int t = threadIdx.x;
  __syncthreads();
  {
    const int jj = t / kLfRun, r = t % kLfRun;
    double s = 0;
    if (jj < n) {
#pragma unroll
      for (int k = 0; k < kLfRun; k++) s += elev[jj][r + kLfRun * k];
    }
    runs[r][jj] = s;
  }
  __syncthreads();
  double s = 0;
  if (t < n) {
#pragma unroll
    for (int r = 0; r < kLfRun; r++) s += runs[r][t];
  }
  __syncthreads();  // the LDS is refilled next
  return s;
}

template <int STAGE, class SH>
__device__ __forceinline__ void lf_probe(const LfDev &d, LfGroup<SH> &g,
                                         bool act, int j0, int j1, bool store,
                                         int kept) {
  constexpr int kLevels = kLfLevels;
  auto &elev = lf_elev();
  const int t = threadIdx.x;
  double *P = d.partial[STAGE] + blockIdx.x;
  for (int c0 = j0; c0 < j1; c0 += kLevels) {
    const int n = min(kLevels, j1 - c0);
    // inactive threads hold a copy of a real group and only contribute 0.
    // The sweep after the last level is only used by `store` and `kept`.
    // SHAKE's velocity updates (which never feed back into positions or
    // errors) are made up to the kept level only.
    for (int jj = 0; jj < n; jj++) {
      const int lev = c0 + jj;
      const double e = group_error<STAGE>(g);
      elev[jj][t] = act ? e : 0.0;
      sweep<STAGE>(g, d.dtc, lev < kept);
      if (lev + 1 == kept && act) {
#pragma unroll
        for (int k = 0; k < SH::NA; k++) {
          if (g.idx[k] < 0) continue;
          if (STAGE == kStShake) {
            st3(d.xk, g.idx[k], g.x[k]);
            st3(d.vk, g.idx[k], g.v[k]);
          } else {
            st3(d.vr, g.idx[k], g.v[k]);
          }
        }
      }
    }
    const double s = lf_level_sums(n);
    if (t < n) P[(size_t)(c0 + t) * d.ngb] = s;
  }
  if (store && act) {
#pragma unroll
    for (int k = 0; k < SH::NA; k++) {
      if (g.idx[k] < 0) continue;
      if (STAGE == kStShake) st3(d.xs, g.idx[k], g.x[k]);
      else st3(d.vs, g.idx[k], g.v[k]);
    }
  }
}

template <class SH>
__device__ __forceinline__ void load_group(const LfDev &d, int gi,
                                           LfGroup<SH> &g, bool need_last) {
#pragma unroll
  for (int k = 0; k < SH::NA; k++) {
    const int i = d.g_atoms[(size_t)gi * SH::NA + k];
    g.idx[k] = i;
    const int j = i < 0 ? 0 : i;  // padding slots read atom 0, never written
    g.x[k] = ld3(d.x, j);
    g.v[k] = ld3(d.v, j);
    g.l[k] = need_last ? ld3(d.last, j) : D3{0, 0, 0};
    g.rm[k] = d.rmass[j];
  }
#pragma unroll
  for (int c = 0; c < SH::NC; c++) {
    g.cn[c] = d.g_cons[(size_t)gi * SH::NC + c];
    g.rest[c] = d.g_rest[(size_t)gi * SH::NC + c];
    g.rinv[c] = d.g_rinv[(size_t)gi * SH::NC + c];
  }
}

// The levels [L, maxit] of a stage when the first probe's L levels did not
// decide (the count rose by more than the margin): every block of the
// replica computes each of the replica's block partials itself, from the
// state the first probe stored after L sweeps (xs / vs), exactly as that
// block's probe would (same sweeps, same LDS order), and sums them per level
// in lf_scan's order (block q into the sum of q mod 8, in increasing q, then
// the eight sums in order).  All blocks reach the same count, the one the
// partials would give; slow (each block redoes the replica) but rare: the
// first call probes every level, and a steady count needs L = count + 2.
template <int STAGE, class SH>
__device__ int lf_fallback(const LfDev &d, int L, const LfRep &R) {
  constexpr int kSub = 8;
  __shared__ double acc[kLfLevels][kSub];
  __shared__ int hit;
  auto &elev = lf_elev();
  const int t = threadIdx.x;
  const int maxl = d.maxit[STAGE] + 1;
  for (int c0 = L; c0 < maxl; c0 += kLfLevels) {
    const int n = min(kLfLevels, maxl - c0);
    if (t < kLfLevels * kSub) acc[t / kSub][t % kSub] = 0.0;
    for (int vb = 0; vb < R.gnb; vb++) {
      const int local = vb * kLfBlock + t;
      const bool act = local < R.ng;
      LfGroup<SH> g;
      load_group(d, R.goff + (act ? local : R.ng - 1), g, STAGE == kStShake);
#pragma unroll
      for (int a = 0; a < SH::NA; a++) {
        if (g.idx[a] < 0) continue;
        if (STAGE == kStShake) g.x[a] = ld3(d.xs, g.idx[a]);
        else g.v[a] = ld3(d.vs, g.idx[a]);
      }
      for (int lev = L; lev < c0; lev++) sweep<STAGE>(g, d.dtc, false);
      for (int jj = 0; jj < n; jj++) {
        const double e = group_error<STAGE>(g);
        elev[jj][t] = act ? e : 0.0;
        sweep<STAGE>(g, d.dtc, false);
      }
      const double s = lf_level_sums(n);
      if (t < n) acc[t][vb % kSub] += s;
    }
    __syncthreads();
    if (t < 64) {
      bool pass = false;
      if (t < n) {
        double s = 0;
#pragma unroll
        for (int k = 0; k < kSub; k++) s += acc[t][k];
        pass = s / R.ncons <= d.eps[STAGE];
      }
      const unsigned long long m = __ballot(pass);
      if (t == 0) hit = m ? c0 + __builtin_ctzll(m) : -1;
    }
    __syncthreads();
    const int hv = hit;
    __syncthreads();  // acc and hit are rewritten by the next chunk
    if (hv >= 0) return hv;
  }
  return maxl;
}

// Sweep count of a stage call: the first level j (0..maxit) whose mean
// error is <= epsilon, else maxit + 1.  Block-wide.
template <int STAGE, class SH>
__device__ __forceinline__ int lf_decide(const LfDev &d, long long scall,
                                         const LfRep &R, int rep) {
  const int L = lf_first_levels(d, STAGE, scall, rep);
  const int k = lf_scan(d, STAGE, 0, L, R);
  if (k >= 0) return k;
  if (L > d.maxit[STAGE]) return d.maxit[STAGE] + 1;
  return lf_fallback<STAGE, SH>(d, L, R);
}

// Cell assignment of the atoms idx[0..NA) (negative: none) at positions x
// (k_assign_r's scheme for a block of up to kLfBlock * NA atoms): the block
// counts its atoms per cell in an LDS table, takes one global atomic add
// per distinct cell and writes each atom into its cell's bucket row at the
// cell's base plus its rank (below the row's capacity).  Block-wide (all
// threads call it).
constexpr int kLfCellSlots = 1024;
// LDS of lf_assign_cells, one object per kernel whatever NA
__device__ __forceinline__ int (&lf_cell_lds())[3][kLfCellSlots] {
  __shared__ int t[3][kLfCellSlots];
  return t;
}

template <int NA>
__device__ __forceinline__ void lf_assign_cells(const LfDev &d,
                                                const int (&idx)[NA],
                                                const D3 (&x)[NA]) {
  constexpr int NS = kLfCellSlots;
  static_assert(kLfBlock * NA <= NS, "one slot per atom of the block");
  auto &lds = lf_cell_lds();
  int *key = lds[0], *lcount = lds[1], *base = lds[2];
  const int t = threadIdx.x;
  for (int q = t; q < NS; q += kLfBlock) {
    key[q] = -1;
    lcount[q] = 0;
  }
  __syncthreads();
  int slot[NA], lrank[NA];
#pragma unroll
  for (int a = 0; a < NA; a++) {
    slot[a] = -1;
    lrank[a] = 0;
    const int i = idx[a];
    if (i < 0) continue;
    const TileRep &R = d.c_reps[d.c_rep_of ? d.c_rep_of[i] : 0];
    int cx, cy, cz;
    double vx, vy, vz;
    cell_coords(R.box, R.grid, x[a].x, x[a].y, x[a].z, cx, cy, cz, vx, vy, vz);
    const int c = R.cell_off + (cx * R.grid.n[1] + cy) * R.grid.n[2] + cz;
    d.c_cell_of[i] = c;
    // at most NS distinct cells in NS slots: the probe ends
    int h = c & (NS - 1);
    while (true) {
      const int old = atomicCAS(&key[h], -1, c);
      if (old == -1 || old == c) break;
      h = (h + 1) & (NS - 1);
    }
    slot[a] = h;
    lrank[a] = atomicAdd(&lcount[h], 1);
  }
  __syncthreads();
  for (int q = t; q < NS; q += kLfBlock)
    if (key[q] >= 0) base[q] = atomicAdd(&d.c_count[key[q]], lcount[q]);
  __syncthreads();
#pragma unroll
  for (int a = 0; a < NA; a++) {
    if (slot[a] < 0) continue;
    const int r = base[slot[a]] + lrank[a];
    if (r < d.c_cap) d.c_bucket[(size_t)key[slot[a]] * d.c_cap + r] = idx[a];
  }
}

// One fused launch: apply stage APPLY (sweep count from its probe), then the
// per-atom update PRE, store, then the first probe of stage PROBE; with
// ZERO, clear the forces and the energy block for the next evaluation.
// sc_apply / sc_probe: the per-stage call numbers; call: the global one.
template <int APPLY, int PRE, int PROBE, bool ZERO, int SHAPE>
__global__ __launch_bounds__(kLfBlock) void k_lf_step(LfDev d, long long sc_apply,
                                                      long long sc_probe,
                                                      long long call) {
  using SH = LfShape<SHAPE>;
  const bool gblock = (int)blockIdx.x < d.ngb;
  if (d.z_count) {
    for (int t = blockIdx.x * kLfBlock + threadIdx.x; t < d.z_nc;
         t += gridDim.x * kLfBlock)
      d.z_count[t] = 0;
  }
  if (ZERO) {
    for (int t = blockIdx.x * kLfBlock + threadIdx.x; t < d.n_out;
         t += gridDim.x * kLfBlock)
      d.out[t] = 0;
  }
  if (!gblock) {
    const int t = ((int)blockIdx.x - d.ngb) * kLfBlock + threadIdx.x;
    int ia[1] = {-1};
    D3 xa[1];
    if (t < d.n_free) {
      const int i = d.free_atoms[t];
      D3 x = ld3(d.x, i);
      if (PRE != kPreNone) {
        D3 v = ld3(d.v, i);
        pre_op<PRE>(x, v, ld3(d.f, i), d.mass[i], d);
        st3(d.x, i, x);
        st3(d.v, i, v);
      }
      if (ZERO) st3(d.f, i, D3{0, 0, 0});
      ia[0] = i;
      xa[0] = x;
    }
    if constexpr (ZERO) {
      if (d.c_count) lf_assign_cells(d, ia, xa);
    }
    return;
  }
  const int rep = d.blk_rep[blockIdx.x];
  const LfRep R = d.reps[rep];
  const int local = ((int)blockIdx.x - R.gb0) * kLfBlock + threadIdx.x;
  const bool act = local < R.ng;
  LfGroup<SH> g;
  // inactive threads hold a copy of the replica's last group (read only)
  load_group(d, R.goff + (act ? local : R.ng - 1), g,
             APPLY == kStShake || PROBE == kStShake);
  // what the apply and the per-atom update read, loaded before the sweep
  // count is decided so that the loads overlap the decision: the state this
  // call's probe kept, the forces and the masses
  const int kept =
      APPLY != kStNone ? lf_kept_level(d, APPLY, sc_apply, rep) : 0;
  D3 kx[SH::NA], kv[SH::NA], fa[SH::NA];
  double ma[SH::NA];
#pragma unroll
  for (int a = 0; a < SH::NA; a++) {
    const int i = g.idx[a] < 0 ? 0 : g.idx[a];  // padding: read, not used
    if (APPLY == kStShake && kept > 0) {
      kx[a] = ld3(d.xk, i);
      kv[a] = ld3(d.vk, i);
    }
    if (APPLY == kStRattle && kept > 0) kv[a] = ld3(d.vr, i);
    if (PRE != kPreNone) {
      fa[a] = ld3(d.f, i);
      ma[a] = d.mass[i];
    }
  }
  int k = 0;
  if constexpr (APPLY != kStNone) {
    k = lf_decide<APPLY, SH>(d, sc_apply, R, rep);
    if ((int)blockIdx.x == R.gb0 && threadIdx.x == 0) {
      d.hint[8 * rep + APPLY * 2 + (int)(sc_apply & 1)] = k;
      const long long slot = call * d.n_reps + rep;
      if (d.iters && slot < d.iters_cap) d.iters[slot] = k;
      if (k > d.maxit[APPLY]) atomicAdd(&d.nonconv[APPLY], 1);
    }
  }
  if (act) {
    int s0 = 0;
    // start from the state this call's probe kept, when it is on the way
    if (APPLY != kStNone && kept > 0 && k >= kept) {
#pragma unroll
      for (int a = 0; a < SH::NA; a++) {
        if (g.idx[a] < 0) continue;
        if (APPLY == kStShake) g.x[a] = kx[a];
        g.v[a] = kv[a];
      }
      s0 = kept;
    }
    for (int s = s0; s < k; s++) sweep<APPLY>(g, d.dtc, true);
#pragma unroll
    for (int a = 0; a < SH::NA; a++) {
      const int i = g.idx[a];
      if (i < 0) continue;
      if (APPLY == kStShake) {
        g.l[a] = g.x[a];
        st3(d.last, i, g.x[a]);
      }
      if (PRE != kPreNone) pre_op<PRE>(g.x[a], g.v[a], fa[a], ma[a], d);
      if (APPLY != kStNone || PRE != kPreNone) {
        st3(d.x, i, g.x[a]);
        st3(d.v, i, g.v[a]);
      }
      if (ZERO) st3(d.f, i, D3{0, 0, 0});
    }
  }
  if constexpr (ZERO) {
    if (d.c_count) {
      int ia[SH::NA];
#pragma unroll
      for (int a = 0; a < SH::NA; a++) ia[a] = act ? g.idx[a] : -1;
      lf_assign_cells(d, ia, g.x);
    }
  }
  if (PROBE != kStNone) {
    const int L = lf_first_levels(d, PROBE, sc_probe, rep);
    lf_probe<PROBE>(d, g, act, 0, L, L <= d.maxit[PROBE],
                    lf_kept_level(d, PROBE, sc_probe, rep));
  }
}

// ------------------------------------------------------------ bonded terms
// BondSystemForce::calcBond (force/bonded/BondSystemForce.h:89-122) and
// AngleSystemForce::calcAngle (force/bonded/AngleSystemForce.h:86-150) with
// Urey-Bradley, evaluated per atom: a thread sums, in list order, the force
// its atom receives from each term it belongs to (each term is recomputed by
// each of its atoms), so no atomics are needed and the sum order is fixed.
struct LfBonded {
  const int *bptr, *bent;  // per atom: entries bond << 1 | role (0 atom1)
  const int *aptr, *aent;  // per atom: entries angle << 2 | role (0, 1, 2)
  const int *bonds;        // 2 per bond
  const double *bk, *br0;
  const int *angles;       // 3 per angle
  const double *ak, *at0, *ubk, *ubr0;
  int n_atoms, n_bonds, n_angles;
};

__device__ __forceinline__ D3 min_diff3(const Box &b, D3 c1, D3 c2) {
  D3 r;
  double d2;
  min_diff_exact(b, c1.x, c1.y, c1.z, c2.x, c2.y, c2.z, r.x, r.y, r.z, d2);
  return r;
}
__device__ __forceinline__ double d3_norm(D3 a) {
  return sqrt(a.x * a.x + a.y * a.y + a.z * a.z);
}

// force on atom1 of bond i; energy k (r - r0)^2
__device__ __forceinline__ D3 bond_f1(const Box &box, const double *x,
                                      const LfBonded &B, int i, double &e) {
  const int a1 = B.bonds[2 * i], a2 = B.bonds[2 * i + 1];
  const double r0 = B.br0[i], k = B.bk[i];
  const D3 r12 = min_diff3(box, ld3(x, a2), ld3(x, a1));
  const double r = d3_norm(r12);
  const double dpotdr = 2.0 * k * (r - r0);
  e = k * (r - r0) * (r - r0);
  const double s = -dpotdr / r;
  return {r12.x * s, r12.y * s, r12.z * s};
}

// forces on atoms 1 and 3 of angle i; energy k (theta - theta0)^2 + UB
__device__ __forceinline__ void angle_f13(const Box &box, const double *x,
                                          const LfBonded &B, int i, D3 &f1,
                                          D3 &f3, double &e) {
  const int a1 = B.angles[3 * i], a2 = B.angles[3 * i + 1],
            a3 = B.angles[3 * i + 2];
  const double t0 = B.at0[i], ka = B.ak[i];
  const double kub = B.ubk[i], rub = B.ubr0[i];
  const D3 p1 = ld3(x, a1), p2 = ld3(x, a2), p3 = ld3(x, a3);
  const D3 r12 = min_diff3(box, p2, p1);
  const D3 r32 = min_diff3(box, p2, p3);
  const D3 r13 = min_diff3(box, p3, p1);
  const double d12 = d3_norm(r12), d32 = d3_norm(r32), d13 = d3_norm(r13);
  const D3 c = {r12.y * r32.z - r12.z * r32.y, r12.z * r32.x - r12.x * r32.z,
                r12.x * r32.y - r12.y * r32.x};
  const double theta = atan2(d3_norm(c), d3_dot(r12, r32));
  const double sn = sin(theta), cs = cos(theta);
  const double dpot = 2.0 * ka * (theta - t0);
  const D3 dr12 = {r12.x / d12, r12.y / d12, r12.z / d12};
  const D3 dr32 = {r32.x / d32, r32.y / d32, r32.z / d32};
  const D3 dr13 = {r13.x / d13, r13.y / d13, r13.z / d13};
  const double s1 = sn * d12, s3 = sn * d32;
  const D3 dt1 = {(dr12.x * cs - dr32.x) / s1, (dr12.y * cs - dr32.y) / s1,
                  (dr12.z * cs - dr32.z) / s1};
  const D3 dt3 = {(dr32.x * cs - dr12.x) / s3, (dr32.y * cs - dr12.y) / s3,
                  (dr32.z * cs - dr12.z) / s3};
  const double u = 2.0 * kub * (d13 - rub);
  const D3 ub1 = {dr13.x * u, dr13.y * u, dr13.z * u};
  f1 = {-dt1.x * dpot - ub1.x, -dt1.y * dpot - ub1.y, -dt1.z * dpot - ub1.z};
  f3 = {-dt3.x * dpot - -ub1.x, -dt3.y * dpot - -ub1.y, -dt3.z * dpot - -ub1.z};
  e = ka * (theta - t0) * (theta - t0) + kub * (d13 - rub) * (d13 - rub);
}

__global__ __launch_bounds__(kLfBlock) void k_lf_bonded(Box box,
                                                        const double *x,
                                                        double *f, LfBonded B) {
  const int i = blockIdx.x * kLfBlock + threadIdx.x;
  if (i >= B.n_atoms) return;
  D3 acc = {0, 0, 0};
  double e;
  for (int q = B.bptr[i]; q < B.bptr[i + 1]; q++) {
    const int ent = B.bent[q];
    const D3 f1 = bond_f1(box, x, B, ent >> 1, e);
    if (ent & 1) acc = {acc.x - f1.x, acc.y - f1.y, acc.z - f1.z};
    else acc = {acc.x + f1.x, acc.y + f1.y, acc.z + f1.z};
  }
  for (int q = B.aptr[i]; q < B.aptr[i + 1]; q++) {
    const int ent = B.aent[q], role = ent & 3;
    D3 f1, f3;
    angle_f13(box, x, B, ent >> 2, f1, f3, e);
    const D3 fr = role == 0 ? f1
                  : role == 2 ? f3
                              : D3{-f1.x - f3.x, -f1.y - f3.y, -f1.z - f3.z};
    acc = {acc.x + fr.x, acc.y + fr.y, acc.z + fr.z};
  }
  if (B.bptr[i + 1] > B.bptr[i] || B.aptr[i + 1] > B.aptr[i]) {
    const D3 fi = ld3(f, i);
    st3(f, i, D3{fi.x + acc.x, fi.y + acc.y, fi.z + acc.z});
  }
}

// Bond and angle energies into out[0], out[1] (+=): one block, terms
// strided over threads, a fixed-order tree across threads.
__global__ __launch_bounds__(kLfBlock) void k_lf_bonded_energy(
    Box box, const double *x, LfBonded B, double *out) {
  __shared__ double red[2][kLfBlock];
  const int t = threadIdx.x;
  double eb = 0, ea = 0, e;
  for (int i = t; i < B.n_bonds; i += kLfBlock) {
    bond_f1(box, x, B, i, e);
    eb += e;
  }
  for (int i = t; i < B.n_angles; i += kLfBlock) {
    D3 f1, f3;
    angle_f13(box, x, B, i, f1, f3, e);
    ea += e;
  }
  red[0][t] = eb;
  red[1][t] = ea;
  __syncthreads();
  for (int o = kLfBlock / 2; o > 0; o >>= 1) {
    if (t < o) {
      red[0][t] += red[0][t + o];
      red[1][t] += red[1][t + o];
    }
    __syncthreads();
  }
  if (t == 0) {
    out[0] += red[0][0];
    out[1] += red[1][0];
  }
}

}  // namespace pmnb
