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
// decision over the whole system.  A "probe" records, per block, the summed
// group errors after 0, 1, ..., W-1 sweeps without keeping the swept state;
// the kernel that applies the stage reduces these partials in a fixed order
// (every block computes the same value), finds the reference's sweep count
// k, and sweeps k times from the stage input.  When the error can stay
// above epsilon past W levels, further probe launches continue from the
// state stored after W sweeps; they return at once when an earlier window
// already decided.  No host synchronisation, no atomics on the decision.
//
// The global error is summed per group then per block rather than in the
// reference's constraint order, and each constraint's relative error is
// multiplied by a precomputed 1/(2 r^2): the sweep count can differ from
// the reference's only when the mean error is within rounding of epsilon.
// The positions and velocities of a given sweep count are the reference's
// arithmetic (FP contraction is off for the whole library).
#pragma once

namespace pmnb {

constexpr int kLfBlock = 256;
constexpr int kLfWindow = 8;     // error levels one probe launch records
constexpr int kLfMaxAtoms = 4;   // largest group
constexpr int kLfMaxCons = 6;    // most constraints in a group

enum LfStage { kStNone = 0, kStShake = 1, kStRattle = 2 };
enum LfPre {
  kPreNone = 0,
  kPreFirst = 1,      // doHalfKickdoDrift: v += F*h*0.5/m; x += v*h
  kPreHalf = 2,       // doHalfKick: v += F*(h/2)/m
  kPreHalfDrift = 3,  // doHalfKick then doDrift
  kPreFullDrift = 4,  // doKickdoDrift without modifiers: v += F*h/m; x += v*h
};

struct LfDev {
  double *x, *v, *f, *out;  // positions, velocities, forces, 20-double block
  double *last;             // SHAKE's previous positions
  const double *mass, *rmass;
  const int *g_atoms;       // NA per group, -1 padded
  const int *g_cons;        // NC per group: local a | local b << 4, -1 padded
  const double *g_rest;     // NC per group: rest length
  const double *g_rinv;     // NC per group: 1 / (2 rest^2)
  const int *free_atoms;
  int n_groups, n_free, ngb;  // ngb: blocks of group threads
  int n_cons;
  double h, h2, dtc;        // ts/TIMEFACTOR-scaled steps (see host)
  double eps[3];            // by stage
  int maxit[3];
  double *partial[3];       // by stage: (maxit + 1) levels x ngb blocks
  double *xs, *vs;          // state after a probe window (per atom)
  int *iters;               // sweep count per stage call (optional)
  long long iters_cap;
  int *nonconv;             // [3] stage calls that hit max_iter
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

template <int NA, int NC>
struct LfGroup {
  D3 x[NA], v[NA], l[NA];
  double rm[NA];
  int idx[NA];
  int cn[NC];
  double rest[NC], rinv[NC];
};

// ModifierShake::doExecute, one sweep over the group's constraints.  With
// VEL false the velocity updates are left out (they never feed back into
// positions or the error; the probe uses this).
template <bool VEL, int NA, int NC>
__device__ __forceinline__ void shake_sweep(LfGroup<NA, NC> &g, double dt) {
#pragma unroll
  for (int c = 0; c < NC; c++) {
    if (g.cn[c] < 0) continue;
    const int a = g.cn[c] & 15, b = g.cn[c] >> 4;
    const D3 pab = d3_sub(pick3(g.x, a), pick3(g.x, b));
    const double pabsq = d3_dot(pab, pab);
    const double rabsq = g.rest[c] * g.rest[c];
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
}

// ModifierRattle::doExecute, one sweep (positions are not changed).
template <int NA, int NC>
__device__ __forceinline__ void rattle_sweep(LfGroup<NA, NC> &g, double dt) {
#pragma unroll
  for (int c = 0; c < NC; c++) {
    if (g.cn[c] < 0) continue;
    const int a = g.cn[c] & 15, b = g.cn[c] >> 4;
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
}

// Sum over the group's constraints of the reference's per-constraint error
// (ModifierMetaShake::calcError / ModifierMetaRattle::calcError).
template <int STAGE, int NA, int NC>
__device__ __forceinline__ double group_error(const LfGroup<NA, NC> &g) {
  double e = 0;
#pragma unroll
  for (int c = 0; c < NC; c++) {
    if (g.cn[c] < 0) continue;
    const int a = g.cn[c] & 15, b = g.cn[c] >> 4;
    if (STAGE == kStShake) {
      const D3 d = d3_sub(pick3(g.x, b), pick3(g.x, a));
      const double r2 = g.rest[c] * g.rest[c];
      e += fabs(d3_dot(d, d) - r2) * g.rinv[c];
    } else {
      const D3 rab = d3_sub(pick3(g.x, a), pick3(g.x, b));
      const D3 vab = d3_sub(pick3(g.v, a), pick3(g.v, b));
      e += fabs(d3_dot(rab, vab));
    }
  }
  return e;
}

template <int STAGE, int NA, int NC>
__device__ __forceinline__ void sweep(LfGroup<NA, NC> &g, double dt, bool vel) {
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

// Sweep count of a stage: the first error level j (0..maxit) whose mean
// error over all constraints is <= epsilon, else maxit + 1.  Every block
// reduces the same partials in the same order.  Returns -1 when the
// windows before `upto` (exclusive) did not decide (only when upto is
// given, for the continuation probes).
__device__ __forceinline__ int lf_decide(const LfDev &d, int stage,
                                          int upto_window) {
  __shared__ double red[kLfWindow][16];
  __shared__ int hit;
  const int t = threadIdx.x;
  const int maxit = d.maxit[stage];
  const double *P = d.partial[stage];
  const int nwin = (maxit + 1 + kLfWindow - 1) / kLfWindow;
  const int last_window = upto_window >= 0 ? upto_window : nwin;
  for (int p = 0; p < last_window; p++) {
    const int j0 = p * kLfWindow;
    const int wn = min(kLfWindow, maxit + 1 - j0);
    __syncthreads();
    if (t < kLfWindow * 16) {
      const int jj = t / 16, sub = t % 16;
      double s = 0;
      if (jj < wn)
        for (int b = sub; b < d.ngb; b += 16) s += P[(size_t)(j0 + jj) * d.ngb + b];
      red[jj][sub] = s;
    }
    if (t == 0) hit = -1;
    __syncthreads();
    if (t == 0) {
      for (int jj = 0; jj < wn; jj++) {
        double s = 0;
        for (int k = 0; k < 16; k++) s += red[jj][k];
        if (s / d.n_cons <= d.eps[stage]) {
          hit = j0 + jj;
          break;
        }
      }
    }
    __syncthreads();
    const int hv = hit;
    if (hv >= 0) return hv;
  }
  return upto_window >= 0 ? -1 : maxit + 1;
}

// Records W error levels of a stage for this block's groups, starting at
// level j0 from the state in g (active threads only), and leaves the state
// after the window in xs / vs when a further window follows.
template <int STAGE, int NA, int NC>
__device__ __forceinline__ void lf_probe(const LfDev &d, LfGroup<NA, NC> &g,
                                         bool act, int j0) {
  __shared__ double wred[kLfBlock / 64][kLfWindow];
  const int maxit = d.maxit[STAGE];
  const int wn = min(kLfWindow, maxit + 1 - j0);
  const bool more = j0 + wn <= maxit;
  double e[kLfWindow];
#pragma unroll
  for (int jj = 0; jj < kLfWindow; jj++) {
    e[jj] = 0;
    if (jj < wn && act) {
      e[jj] = group_error<STAGE>(g);
      if (jj < wn - 1 || more) sweep<STAGE>(g, d.dtc, false);
    }
  }
  if (more && act) {
#pragma unroll
    for (int k = 0; k < NA; k++) {
      if (g.idx[k] < 0) continue;
      if (STAGE == kStShake) st3(d.xs, g.idx[k], g.x[k]);
      else st3(d.vs, g.idx[k], g.v[k]);
    }
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int jj = 0; jj < kLfWindow; jj++) {
    double s = e[jj];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (lane == 0) wred[wave][jj] = s;
  }
  __syncthreads();
  if (threadIdx.x < wn) {
    double s = 0;
#pragma unroll
    for (int w = 0; w < kLfBlock / 64; w++) s += wred[w][threadIdx.x];
    d.partial[STAGE][(size_t)(j0 + threadIdx.x) * d.ngb + blockIdx.x] = s;
  }
}

template <int NA, int NC>
__device__ __forceinline__ void load_group(const LfDev &d, int gi,
                                           LfGroup<NA, NC> &g, bool need_last) {
#pragma unroll
  for (int k = 0; k < NA; k++) {
    const int i = d.g_atoms[(size_t)gi * NA + k];
    g.idx[k] = i;
    const int j = i < 0 ? 0 : i;  // padding slots read atom 0, never written
    g.x[k] = ld3(d.x, j);
    g.v[k] = ld3(d.v, j);
    g.l[k] = need_last ? ld3(d.last, j) : D3{0, 0, 0};
    g.rm[k] = d.rmass[j];
  }
#pragma unroll
  for (int c = 0; c < NC; c++) {
    g.cn[c] = d.g_cons[(size_t)gi * NC + c];
    g.rest[c] = d.g_rest[(size_t)gi * NC + c];
    g.rinv[c] = d.g_rinv[(size_t)gi * NC + c];
  }
}

// One fused launch: apply stage APPLY (sweep count from its probe), then the
// per-atom update PRE, store, then probe stage PROBE's first window; with
// ZERO, clear the forces and the energy block for the next evaluation.
template <int APPLY, int PRE, int PROBE, bool ZERO, int NA, int NC>
__global__ __launch_bounds__(kLfBlock) void k_lf_step(LfDev d, long long call) {
  int k = 0;
  if (APPLY != kStNone) {
    k = lf_decide(d, APPLY, -1);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      if (d.iters && call < d.iters_cap) d.iters[call] = k;
      if (k > d.maxit[APPLY]) atomicAdd(&d.nonconv[APPLY], 1);
    }
  }
  if (ZERO && blockIdx.x == 0 && threadIdx.x < 20) d.out[threadIdx.x] = 0;
  if ((int)blockIdx.x >= d.ngb) {
    const int t = ((int)blockIdx.x - d.ngb) * kLfBlock + threadIdx.x;
    if (t >= d.n_free) return;
    const int i = d.free_atoms[t];
    if (PRE != kPreNone) {
      D3 x = ld3(d.x, i), v = ld3(d.v, i);
      pre_op<PRE>(x, v, ld3(d.f, i), d.mass[i], d);
      st3(d.x, i, x);
      st3(d.v, i, v);
    }
    if (ZERO) st3(d.f, i, D3{0, 0, 0});
    return;
  }
  const int gi = blockIdx.x * kLfBlock + threadIdx.x;
  const bool act = gi < d.n_groups;
  LfGroup<NA, NC> g;
  if (act) {
    load_group(d, gi, g, APPLY == kStShake || PROBE == kStShake);
    for (int s = 0; s < k; s++) sweep<APPLY>(g, d.dtc, true);
#pragma unroll
    for (int a = 0; a < NA; a++) {
      const int i = g.idx[a];
      if (i < 0) continue;
      if (APPLY == kStShake) {
        g.l[a] = g.x[a];
        st3(d.last, i, g.x[a]);
      }
      if (PRE != kPreNone) pre_op<PRE>(g.x[a], g.v[a], ld3(d.f, i), d.mass[i], d);
      if (APPLY != kStNone || PRE != kPreNone) {
        st3(d.x, i, g.x[a]);
        st3(d.v, i, g.v[a]);
      }
      if (ZERO) st3(d.f, i, D3{0, 0, 0});
    }
  }
  if (PROBE != kStNone) lf_probe<PROBE>(d, g, act, 0);
}

// Window p >= 1 of a stage's probe, from the state the previous window left.
template <int STAGE, int NA, int NC>
__global__ __launch_bounds__(kLfBlock) void k_lf_probe(LfDev d, int p) {
  if (lf_decide(d, STAGE, p) >= 0) return;
  const int gi = blockIdx.x * kLfBlock + threadIdx.x;
  const bool act = gi < d.n_groups;
  LfGroup<NA, NC> g;
  if (act) {
    load_group(d, gi, g, STAGE == kStShake);
#pragma unroll
    for (int a = 0; a < NA; a++) {
      if (g.idx[a] < 0) continue;
      if (STAGE == kStShake) g.x[a] = ld3(d.xs, g.idx[a]);
      else g.v[a] = ld3(d.vs, g.idx[a]);
    }
  }
  lf_probe<STAGE>(d, g, act, p * kLfWindow);
}

}  // namespace pmnb
