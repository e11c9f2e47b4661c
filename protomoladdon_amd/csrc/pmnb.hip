// C ABI of the MI355X nonbonded cutoff force (include/pmnb.h).
//
// One context = one force object on one GPU.  The host side mirrors what
// NonbondedCutoffSystemForce::evaluate does per call
// (force/nonbonded/NonbondedCutoffSystemForce.h:45-55): rebuild the cell
// list from the current positions, enumerate cell pairs, evaluate pairs,
// accumulate forces, energies and virials.  All of it runs on the device.
#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <type_traits>
#include <utility>
#include <vector>

#include "../../include/pmnb.h"
#include "pmnb_common.h"
#include "pmnb_cells.h"
#include "pmnb_pair.h"
#include "pmnb_list.h"
#include "pmnb_tile.h"
#include "pmnb_queue.h"
#include "pmnb_adjoint.h"
#include "pmnb_tilemask.h"
#include "pmnb_leapfrog.h"
#include "pmnb_ewald.h"
#include "pmnb_pme.h"

using namespace pmnb;

namespace {

constexpr double kEpsilon = 1.0e-14;                   // base/PMConstants.cpp:47
constexpr int kMaxInt2 = 2147483647 / 2;               // base/PMConstants.cpp:44
constexpr double kPi = 3.14159265358979323846;

struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  hipError_t ensure(size_t want) {
    if (want <= bytes && p) return hipSuccess;
    if (p) {
      hipError_t e = hipFree(p);
      p = nullptr;
      bytes = 0;
      if (e != hipSuccess) return e;
    }
    size_t b = want < 256 ? 256 : want;
    hipError_t e = hipMalloc(&p, b);
    if (e == hipSuccess) bytes = b;
    return e;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  template <class T> T *as() const { return static_cast<T *>(p); }
};

// Host threads that share pmnb_evaluate's staging copy and force add.  The
// caller's thread is member 0; the others wait for a task posted through a
// sequence number.  After a task a worker polls for the next one for
// kPollNs (the force add follows the copy by the kernels' time, and the
// next call's copy follows the add), then sleeps on a condition variable
// until a task or the end.
class HostCrew {
 public:
  ~HostCrew() {
    if (th_.empty()) return;
    {
      std::lock_guard<std::mutex> g(mu_);
      stop_ = true;
    }
    cv_.notify_all();
    for (std::thread &t : th_) t.join();
  }
  int size() const { return (int)th_.size() + 1; }
  // n members in all (n - 1 threads); once.  A thread that cannot be
  // started leaves the crew smaller (no exception leaves the C ABI).
  void start(int n) {
    const unsigned s0 = seq_.load(std::memory_order_relaxed);
    for (int k = 1; k < n; k++) {
      try {
        th_.emplace_back([this, k, s0] { loop(k, s0); });
      } catch (...) {
        break;
      }
    }
  }
  // f(k) for k = 0 .. size() - 1, f(0) on the calling thread; returns when
  // every member has returned
  template <class F>
  void run(F &&f) {
    if (th_.empty()) {
      f(0);
      return;
    }
    // the task as a plain function and argument: nothing allocated
    using Fn = typename std::remove_reference<F>::type;
    task_fn_ = [](void *p, int k) { (*static_cast<Fn *>(p))(k); };
    task_arg_ = (void *)&f;
    busy_.store((int)th_.size(), std::memory_order_relaxed);
    seq_.fetch_add(1, std::memory_order_seq_cst);
    if (sleepers_.load(std::memory_order_seq_cst) > 0) {
      std::lock_guard<std::mutex> g(mu_);
      cv_.notify_all();
    }
    f(0);
    while (busy_.load(std::memory_order_acquire) > 0) relax();
    task_fn_ = nullptr;
    task_arg_ = nullptr;
  }

 private:
  static constexpr long long kPollNs = 1000000;
  static void relax() { __builtin_ia32_pause(); }
  // seen: the sequence number before the first task
  void loop(int k, unsigned seen) {
    while (true) {
      const auto t0 = std::chrono::steady_clock::now();
      unsigned now = seq_.load(std::memory_order_acquire);
      while (now == seen && !stop_.load(std::memory_order_acquire)) {
        if (std::chrono::duration_cast<std::chrono::nanoseconds>(
                std::chrono::steady_clock::now() - t0).count() > kPollNs) {
          std::unique_lock<std::mutex> g(mu_);
          sleepers_.fetch_add(1, std::memory_order_seq_cst);
          cv_.wait(g, [&] {
            return stop_ || seq_.load(std::memory_order_seq_cst) != seen;
          });
          sleepers_.fetch_sub(1, std::memory_order_relaxed);
        } else {
          for (int i = 0; i < 64; i++) relax();
        }
        now = seq_.load(std::memory_order_acquire);
      }
      if (now == seen) return;  // stopped
      seen = now;
      task_fn_(task_arg_, k);
      busy_.fetch_sub(1, std::memory_order_release);
    }
  }
  std::vector<std::thread> th_;
  std::mutex mu_;
  std::condition_variable cv_;
  std::atomic<unsigned> seq_{0};
  std::atomic<int> busy_{0}, sleepers_{0};
  void (*task_fn_)(void *, int) = nullptr;  // the posted task
  void *task_arg_ = nullptr;
  std::atomic<bool> stop_{false};
};

// Members of pmnb_evaluate's host crew: PMNB_HOST_THREADS when set (1 = the
// caller's thread alone), else 4 (at most the hardware threads).
int host_crew_size() {
  if (const char *e = std::getenv("PMNB_HOST_THREADS")) {
    const int v = std::atoi(e);
    if (v >= 1) return std::min(v, 64);
  }
  const int hw = (int)std::thread::hardware_concurrency();
  return std::max(1, std::min(4, hw));
}

}  // namespace

// pmnb_evaluate shares its host copy and add over the crew for systems of
// at least kHostCrewMin coordinates
constexpr int kHostCrewMin = 1 << 16;
// Above this many cells the tile path scans the cell counts over the
// columns in three launches instead of one block (512 water replicas,
// about 63,000 cells in 5,500 columns: 146 us in one block)
constexpr int kScanCellsOneBlock = 16384;

struct pmnb_ctx {
  pmnb_params prm{};
  hipStream_t stream = nullptr;
  std::string err;
  // topology
  bool have_topo = false, have_box = false;
  int n = 0, ntypes = 0, nmol = 0, max_excl_delta = 0;
  double csf = 1.0;
  // box / grid
  Box box{};
  Grid grid{};
  std::vector<int4> shell;  // (0,0,0) then the reference's delta list
  int per_pair_wrap = 0;
  long long last_pairs = 0;
  // device counter of the last evaluation's pairs (pmnb_last_pair_count)
  const unsigned long long *last_np = nullptr;
  // device: topology (original order)
  DevBuf d_type, d_q, d_mol, d_xfull, d_xmod, d_xown, d_lj, d_ljf;
  // device: per evaluation
  DevBuf d_pos, d_force, d_out, d_com, d_cell_of, d_count, d_start, d_fill,
      d_sorted, d_sx, d_sy, d_sz, d_loc, d_stype, d_sq, d_smol, d_scell,
      d_srec, d_shell, d_partial, d_npairs, d_pairs, d_list, d_ncount, d_partial2,
      d_shell_off, d_box;
  int shell_uploaded = 0;
  std::vector<float4> shell_off;  // FP32 cell offsets of the shell entries
  bool use_list = false;          // neighbour-list path available
  bool use_tile = false;          // LDS-staged path (own column grid)
  TileGrid tile{};
  // tile path geometry: one replica for a single system, several for a
  // batch (atoms of replica r then carry r in d_rep_of)
  std::vector<TileRep> reps;
  std::vector<int> col_rep;       // replica of each global column
  int ncol_total = 0;
  bool batch = false;
  DevBuf d_rep_of;
  int rep_uploaded = 0;           // reps and col_rep are on the device
  DevBuf d_partial_lj;            // adjoint: LJ-parameter partials per item
  // tile path: per sorted slot the 64-bit fixed-point force (3 words mixed,
  // 6 FP64), all zero between evaluations (the finish kernel clears it)
  DevBuf d_ffix;
  const void *ffix_zeroed = nullptr;  // allocation last cleared
  DevBuf d_bucket;                    // tile path: cap atoms per cell
  int bucket_cap = 64;
  // d_count holds zeros at the start of the next evaluation on the tile path
  // (the finish kernel clears it after the pair kernel): no clearing launch
  bool counts_zero = false;
  const void *count_seen = nullptr;   // d_count allocation the flag is about
  // pair record of the force kernels (parity tests): on for one evaluation
  PairLog plog{nullptr, 0, nullptr};
  DevBuf d_plog, d_plog_count;
  // pinned host staging of pmnb_evaluate: positions, forces, out + count
  double *h_stage = nullptr;
  size_t h_stage_bytes = 0;
  hipEvent_t h_ev[1] = {};  // pmnb_evaluate: results are on the host
  HostCrew crew;             // pmnb_evaluate: host threads (started once)
  bool crew_started = false;
  int shape = 0;                  // tile-path kernel shape (mixed)
  int grid_cap = 0;               // persistent pair grid: max blocks (0: none)
  DevBuf d_items, d_nitems, d_reps, d_col_rep, d_rep_first, d_colscan;
  double box_key[12] = {0};       // e1, e2, e3, origin of the current box
  unsigned long long topo_gen = 0;  // pmnb_set_topology calls so far
  int cap = 0;                    // list slots per atom
  // pair-kernel timing
  bool timing = false;
  std::vector<hipEvent_t> ev_pool;  // 3 per launch: start, built, done
  size_t ev_used = 0;
};

struct pmnb_batch {
  pmnb_ctx *ctx = nullptr;  // the concatenated system, in batch mode
  int ntypes = 0;
  std::vector<double> lj;
  double csf = 1.0;
  // replicas in order of addition
  std::vector<int> type, mol, rep_of, excl;
  std::vector<double> q;
  std::vector<Box> boxes;
  int next_mol = 0;
  bool dirty = true;  // replicas added since the last upload
  // the captured evaluation and what it was captured for
  hipStream_t cap_stream = nullptr;  // non-blocking stream for capture
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  const void *gkey[3] = {nullptr, nullptr, nullptr};
  uint32_t gflags = 0;
  int gshape = 0, gcap = 0;
};

namespace {

int fail(pmnb_ctx *c, int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (c) c->err = buf;
  return code;
}

#define HIPCHK(call)                                                        \
  do {                                                                      \
    hipError_t e_ = (call);                                                 \
    if (e_ != hipSuccess)                                                   \
      return fail(ctx, PMNB_ERR_HIP, "%s failed: %s (%s:%d)", #call,        \
                  hipGetErrorString(e_), __FILE__, __LINE__);               \
  } while (0)

int blocks_for(long long n, int b = kBlock) {
  return (int)((n + b - 1) / b);
}

// Host restatement of the switching constants
// (switch/C2SwitchingFunction.cpp:12-18, switch/C1SwitchingFunction.cpp:12-16).
Switch make_c2(double on, double cut) {
  Switch s{};
  double d = cut * cut - on * on;
  s.on2 = on * on;
  s.cut2 = cut * cut;
  s.s1 = 1.0 / ((d * d) * d);
  s.s2 = cut * cut - 3.0 * on * on;
  s.s3 = 4.0 / ((d * d) * d);
  return s;
}

Switch make_c1(double cut) {
  Switch s{};
  s.on2 = 0.0;
  s.cut2 = cut * cut;
  s.c1_15c1 = 1.5 / cut;
  s.c1_05c3 = 0.5 / (cut * cut * cut);
  s.c1_15c3 = 1.5 / (cut * cut * cut);
  return s;
}

// Switching constants and the pair cutoff, from the force parameters alone.
ForceParams make_params_for(const pmnb_params &prm, double csf) {
  ForceParams P{};
  P.sw_lj = make_c2(prm.lj_switchon, prm.lj_cutoff);
  if (prm.elec_switch == PMNB_SWITCH_C1)
    P.sw_el = make_c1(prm.elec_cutoff);
  else
    P.sw_el = make_c2(prm.elec_switchon, prm.elec_cutoff);
  // OneAtomPairTwo::mySquaredCutoff: max of the two switching cutoffs
  // (force/OneAtomPairTwo.h:38-40).  With no electrostatic functor the
  // second switch still exists in the reference's two-force template, so
  // its cutoff still counts.
  P.cut2 = std::max(P.sw_lj.cut2, P.sw_el.cut2);
  P.alpha = prm.ewald_alpha;
  P.two_alpha_pi = 2.0 * prm.ewald_alpha / std::sqrt(kPi);
  P.csf = csf;
  P.elec = prm.elec;
  P.el_switch = prm.elec_switch;
  P.band = 1.0e-3f * (float)std::max(1.0, P.cut2 / 100.0);
  return P;
}

// z cell size of the tile path over the pair cutoff: the mixed kernel is as
// fast with rcp / 3 as with rcp / 4 and the cell build is faster with fewer
// cells; the FP64 kernel is faster with rcp / 4 (measured on the 96k water)
double tile_zfrac(const pmnb_params &prm) {
  return prm.precision == PMNB_PREC_MIXED ? 1.0 / 3.0 : 0.25;
}

// Column grid of the tile path for box b and pair cutoff rcp: x/y columns
// at least rcp wide, z cells of about zfrac * rcp.  An axis with a single column
// is marked for per-pair minimum image.  shell receives the distinct
// cell offsets within reach ((0,0,0) first), which the pair-set extraction
// walks.  Returns false when the grid is too large.
bool make_tile_rep(const Box &b, double rcp, double zfrac, TileRep &rep,
                   std::vector<int4> &shell) {
  const int tx = std::max(1, (int)std::floor(b.H[0] / rcp));
  const int ty = std::max(1, (int)std::floor(b.H[1] / rcp));
  const int tz = std::max(1, (int)std::floor(b.H[2] / (zfrac * rcp)));
  if ((long long)tx * ty * tz >= (1ll << 26)) return false;
  rep = TileRep{};
  const double csz = b.H[2] / tz;
  const int rz = (int)std::floor(rcp / csz) + 1;
  Grid &tg = rep.grid;
  tg.n[0] = tx;
  tg.n[1] = ty;
  tg.n[2] = tz;
  tg.ncell = tx * ty * tz;
  tg.cs[0] = b.H[0] / tx;
  tg.cs[1] = b.H[1] / ty;
  tg.cs[2] = csz;
  for (int k = 0; k < 3; k++) tg.inv[k] = 1.0 / tg.cs[k];
  rep.box = b;
  rep.tg.nx = tx;
  rep.tg.ny = ty;
  rep.tg.nz = tz;
  rep.tg.rz = rz;
  rep.tg.csx = (float)tg.cs[0];
  rep.tg.csy = (float)tg.cs[1];
  rep.tg.csz = (float)tg.cs[2];
  for (int k = 0; k < 3; k++) {
    rep.L[k] = (float)b.e[k];
    rep.iL[k] = (float)b.er[k];
  }
  // one column per axis (box < 2 rcp): per-pair minimum image; with two,
  // the column offsets -1 and +1 reach the same column at images one box
  // length (>= 2 rcp) apart, so at most one of them is within the cutoff
  rep.wrap[0] = tx == 1;
  rep.wrap[1] = ty == 1;
  rep.ncol = tx * ty;
  // distinct offsets per axis
  auto offsets = [](int n, int reach) {
    std::vector<int> o;
    if (n >= 2 * reach + 1) {
      for (int d = -reach; d <= reach; d++) o.push_back(d);
    } else {
      for (int d = 0; d < n; d++) o.push_back(d);
    }
    return o;
  };
  const std::vector<int> ox = offsets(tx, 1), oy = offsets(ty, 1),
                         oz = offsets(tz, rz);
  shell.clear();
  shell.push_back(make_int4(0, 0, 0, 0));
  for (int dx : ox)
    for (int dy : oy)
      for (int dz : oz)
        if (dx || dy || dz) shell.push_back(make_int4(dx, dy, dz, 0));
  return true;
}

// Box + grid + shell offsets exactly as PeriodicBoundaryConditions::set,
// CubicCellManager::initialize, ArrayCellListStructure::initialize and the
// periodic enumerator's delta list compute them.
int make_box(pmnb_ctx *ctx, const double *e1, const double *e2,
             const double *e3, const double *origin, Box &b) {
  if (e1[1] != 0.0 || e1[2] != 0.0 || e2[0] != 0.0 || e2[2] != 0.0 ||
      e3[0] != 0.0 || e3[1] != 0.0)
    return fail(ctx, PMNB_ERR_UNSUPPORTED, "only orthogonal boxes are built");
  if (!(e1[0] > 0.0) || !(e2[1] > 0.0) || !(e3[2] > 0.0))
    return fail(ctx, PMNB_ERR_ARG, "box vectors must be positive");
  b = Box{};
  const double E[3] = {e1[0], e2[1], e3[2]};
  // e_kr = cross / dot; for a diagonal box this is 1/e_k computed as the
  // reference does: (e2 x e3).x / (e1 . (e2 x e3)).
  const double a1 = e2[1] * e3[2] - e2[2] * e3[1];
  const double a2 = e3[2] * e1[0] - e3[0] * e1[2];
  const double a3 = e1[0] * e2[1] - e1[1] * e2[0];
  b.er[0] = a1 / (e1[0] * a1 + e1[1] * 0.0 + e1[2] * 0.0);
  b.er[1] = a2 / (e2[0] * 0.0 + e2[1] * a2 + e2[2] * 0.0);
  b.er[2] = a3 / (e3[0] * 0.0 + e3[1] * 0.0 + e3[2] * a3);
  for (int k = 0; k < 3; k++) {
    b.e[k] = E[k];
    b.origin[k] = origin[k];
    const double s = (k == 0 ? e1[0] + e2[0] + e3[0]
                             : (k == 1 ? e1[1] + e2[1] + e3[1]
                                       : e1[2] + e2[2] + e3[2]));
    const double lo = origin[k] - s * 0.5, hi = origin[k] + s * 0.5;
    const double mn = std::min(lo, hi), mx = std::max(lo, hi);
    b.mn[k] = mn;
    b.H[k] = mx - mn;
    b.H2[k] = b.H[k] * 0.5;
    b.shift[k] = origin[k] - mn;
  }
  const double dx = e1[0] * 0.5, dy = e2[1] * 0.5, dz = e3[2] * 0.5;
  const double DX = (dx * dx) * 1.0, DY = (dy * dy) * 1.0, DZ = (dz * dz) * 1.0;
  b.D = std::min(DX, std::min(DY, DZ));
  return PMNB_OK;
}

int build_box(pmnb_ctx *ctx, const double *e1, const double *e2,
              const double *e3, const double *origin) {
  Box b;
  const int brc = make_box(ctx, e1, e2, e3, origin, b);
  if (brc != PMNB_OK) return brc;

  const double cs0 = ctx->prm.cell_size;
  Grid g{};
  int maxN[3], n1[3], maxN1[3];
  for (int k = 0; k < 3; k++) {
    const double vd = b.H[k];
    const double f = std::floor(vd / cs0);
    const double cs = vd / std::max(1.0, f);
    g.cs[k] = cs;
    g.inv[k] = 1.0 / cs;
    const int n = (int)std::floor(vd / cs + kEpsilon);
    g.n[k] = std::max(1, n);
    maxN[k] = kMaxInt2 - kMaxInt2 % g.n[k];
    n1[k] = -g.n[k] / 2;
    maxN1[k] = maxN[k] - n1[k];
  }
  const long long ncell = (long long)g.n[0] * g.n[1] * g.n[2];
  if (ncell > (1ll << 26))
    return fail(ctx, PMNB_ERR_UNSUPPORTED,
                "cell grid %dx%dx%d too large; raise the cell size", g.n[0],
                g.n[1], g.n[2]);
  g.ncell = (int)ncell;

  // Delta list (CellListEnumeratorPeriodicBoundaries.h:46-83).
  const double cut = ctx->prm.algorithm_cutoff;
  const int nx = (int)(cut / g.cs[0] + 1.0 + kEpsilon);
  const int ny = (int)(cut / g.cs[1] + 1.0 + kEpsilon);
  const int nz = (int)(cut / g.cs[2] + 1.0 + kEpsilon);
  const int r0 = std::min(nx, g.n[0]), r1 = std::min(ny, g.n[1]),
            r2 = std::min(nz, g.n[2]);
  const double xx = g.cs[0] * g.cs[0], yy = g.cs[1] * g.cs[1],
               zz = g.cs[2] * g.cs[2], c2 = cut * cut;
  std::vector<int4> dl;
  auto seen = [&](int x, int y, int z) {
    if (x == 0 && y == 0 && z == 0) return true;
    for (const int4 &v : dl)
      if (v.x == x && v.y == y && v.z == z) return true;
    return false;
  };
  for (int k = -r0; k <= r0; k++) {
    int x = std::abs(k) - 1;
    if (x < 0) x = 0;
    const double d0 = x * x * xx;
    for (int l = -r1; l <= r1; l++) {
      int y = std::abs(l) - 1;
      if (y < 0) y = 0;
      const double d1 = d0 + y * y * yy;
      for (int m = -r2; m <= r2; m++) {
        int z = std::abs(m) - 1;
        if (z < 0) z = 0;
        if (d1 + z * z * zz < c2) {
          const int bx = (k + maxN1[0]) % g.n[0] + n1[0];
          const int by = (l + maxN1[1]) % g.n[1] + n1[1];
          const int bz = (m + maxN1[2]) % g.n[2] + n1[2];
          if (!seen(bx, by, bz)) dl.push_back(make_int4(bx, by, bz, 0));
        }
      }
    }
  }
  std::sort(dl.begin(), dl.end(), [](const int4 &p, const int4 &q) {
    if (p.x != q.x) return p.x < q.x;
    if (p.y != q.y) return p.y < q.y;
    return p.z < q.z;
  });
  ctx->shell.clear();
  ctx->shell.push_back(make_int4(0, 0, 0, 0));
  ctx->shell.insert(ctx->shell.end(), dl.begin(), dl.end());
  // The mixed kernel takes the image of a neighbour atom from its cell's
  // offset; that is the minimum image only when every cell offset a pair
  // within the cutoff can have, [-reach, reach] with reach taken before the
  // delta list folds offsets into one periodic class, is its own class.
  const int reach[3] = {nx, ny, nz};
  ctx->per_pair_wrap = 0;
  for (int k = 0; k < 3; k++)
    if (g.n[k] < 2 * reach[k] + 1) ctx->per_pair_wrap = 1;
  // Tile path: its own grid of x/y columns at least one cutoff wide and z
  // cells of about a quarter cutoff.  Used when the pair set does not depend
  // on the cell grid (algorithm cutoff >= switching cutoff).
  {
    const ForceParams P0 = make_params_for(ctx->prm, ctx->csf);
    const double rcp = std::sqrt(P0.cut2);
    // the pair kernel packs type and sorted index into 32 bits and holds
    // the LJ table in LDS
    const bool packs = ctx->ntypes * ctx->ntypes <= kQLj &&
                       ctx->n < (1 << kQTypeShift);
    std::vector<int4> tshell;
    TileRep rep{};
    ctx->use_tile = ctx->prm.algorithm_cutoff >= rcp && packs &&
                    make_tile_rep(b, rcp, tile_zfrac(ctx->prm), rep, tshell);
    if (ctx->use_tile) {
      ctx->tile = rep.tg;
      g = rep.grid;
      ctx->shell = tshell;  // pair-set extraction walks it
      ctx->reps.assign(1, rep);
      ctx->col_rep.assign(rep.ncol, 0);
      ctx->ncol_total = rep.ncol;
    }
  }
  ctx->shell_off.clear();
  for (const int4 &v : ctx->shell)
    ctx->shell_off.push_back(make_float4((float)(v.x * g.cs[0]),
                                         (float)(v.y * g.cs[1]),
                                         (float)(v.z * g.cs[2]), 0.f));
  ctx->use_list = !ctx->use_tile && (int)ctx->shell.size() <= kMaxShell;
  ctx->box = b;
  ctx->grid = g;
  ctx->shell_uploaded = 0;
  ctx->rep_uploaded = 0;
  ctx->have_box = true;
  return PMNB_OK;
}

// List slots per atom: 1.4x the mean neighbour count at the system's mean
// density plus 32, a multiple of 32, at most n - 1.  Atoms with more
// neighbours are computed by the direct-scan kernel.
int list_capacity(const pmnb_ctx *ctx, double cut2) {
  const double vol = ctx->box.H[0] * ctx->box.H[1] * ctx->box.H[2];
  const double rc = std::sqrt(cut2);
  const double mean = ctx->n / vol * 4.0 / 3.0 * kPi * rc * rc * rc;
  long long cap = (long long)(1.4 * mean) + 32;
  cap = (cap + 31) / 32 * 32;
  cap = std::min<long long>(cap, 4096);
  cap = std::min<long long>(cap, std::max(ctx->n - 1, 1));
  return (int)std::max<long long>(cap, 1);
}

ForceParams make_params(const pmnb_ctx *ctx, uint32_t flags) {
  ForceParams P = make_params_for(ctx->prm, ctx->csf);
  P.box = ctx->box;
  P.grid = ctx->grid;
  P.n_atoms = ctx->n;
  P.n_types = ctx->ntypes;
  P.ndelta = (int)ctx->shell.size();
  P.max_excl_delta = ctx->max_excl_delta;
  P.per_pair_wrap = ctx->per_pair_wrap;
  P.flags = flags;
  F32Params &F = P.f;
  F.cut2 = (float)P.cut2;
  F.band = P.band;
  for (int k = 0; k < 3; k++) {
    (&F.Lx)[k] = (float)P.box.e[k];
    (&F.iLx)[k] = (float)P.box.er[k];
  }
  F.on2l = (float)P.sw_lj.on2;
  F.cut2l = (float)P.sw_lj.cut2;
  F.s1l = (float)P.sw_lj.s1;
  F.s2l = (float)P.sw_lj.s2;
  F.s3l = (float)P.sw_lj.s3;
  F.on2e = (float)P.sw_el.on2;
  F.cut2e = (float)P.sw_el.cut2;
  F.s1e = (float)P.sw_el.s1;
  F.s2e = (float)P.sw_el.s2;
  F.s3e = (float)P.sw_el.s3;
  F.alpha = (float)P.alpha;
  F.tapi = (float)P.two_alpha_pi;
  F.csf = (float)P.csf;
  F.same_switch = P.sw_lj.on2 == P.sw_el.on2 && P.sw_lj.cut2 == P.sw_el.cut2 &&
                  ctx->prm.elec_switch == PMNB_SWITCH_C2;
  return P;
}

void launch_fill_i32(int *p, int n, int v, hipStream_t s) {
  if (n > 0)
    hipLaunchKernelGGL(k_fill_i32, dim3(blocks_for(n)), dim3(kBlock), 0, s, p,
                       n, v);
}

// Atom slots per cell of the tile path's buckets: at least 4 times the mean
// occupancy, a power of two in [64, kSortLds] (a fuller cell is read back
// from the cell index of every atom instead).
int tile_bucket_cap(int n, int nc) {
  const long long mean = ((long long)n + std::max(nc, 1) - 1) / std::max(nc, 1);
  int cap = 64;
  while (cap < 4 * mean && cap < kSortLds) cap *= 2;
  return cap;
}

int ensure_eval_buffers(pmnb_ctx *ctx) {
  const int n = ctx->n, nc = ctx->grid.ncell;
  const size_t N = (size_t)std::max(n, 1);
  HIPCHK(ctx->d_cell_of.ensure(N * sizeof(int)));
  HIPCHK(ctx->d_count.ensure((size_t)nc * sizeof(int)));
  if (ctx->count_seen != ctx->d_count.p) {
    ctx->counts_zero = false;
    ctx->count_seen = ctx->d_count.p;
  }
  HIPCHK(ctx->d_start.ensure((size_t)nc * sizeof(int)));
  HIPCHK(ctx->d_fill.ensure(N * sizeof(int)));  // per-atom rank in its cell
  HIPCHK(ctx->d_sorted.ensure(N * sizeof(int)));
  HIPCHK(ctx->d_sx.ensure(N * sizeof(double)));
  HIPCHK(ctx->d_sy.ensure(N * sizeof(double)));
  HIPCHK(ctx->d_sz.ensure(N * sizeof(double)));
  HIPCHK(ctx->d_loc.ensure(N * sizeof(float4)));
  HIPCHK(ctx->d_stype.ensure(N * sizeof(int)));
  HIPCHK(ctx->d_sq.ensure(N * sizeof(double)));
  HIPCHK(ctx->d_smol.ensure(N * sizeof(int)));
  HIPCHK(ctx->d_scell.ensure(N * sizeof(int)));
  if (ctx->use_tile) {
    HIPCHK(ctx->d_srec.ensure(N * sizeof(int2)));
    ctx->bucket_cap = tile_bucket_cap(n, nc);
    HIPCHK(ctx->d_bucket.ensure((size_t)std::max(nc, 1) * ctx->bucket_cap *
                                sizeof(int)));
    const int W = ctx->prm.precision == PMNB_PREC_FP64 ? FixWords<true>::W
                                                        : FixWords<false>::W;
    HIPCHK(ctx->d_ffix.ensure(N * W * sizeof(long long)));
    if (ctx->ffix_zeroed != ctx->d_ffix.p) {
      const long long nw = (long long)(ctx->d_ffix.bytes / sizeof(long long));
      hipLaunchKernelGGL(k_zero_i64, dim3((unsigned)((nw + kBlock - 1) / kBlock)),
                         dim3(kBlock), 0, ctx->stream, ctx->d_ffix.as<long long>(),
                         nw);
      HIPCHK(hipGetLastError());
      HIPCHK(hipStreamSynchronize(ctx->stream));
      ctx->ffix_zeroed = ctx->d_ffix.p;
    }
  }
  const int ncol_t = ctx->use_tile ? ctx->ncol_total : 0;
  const int tile_items = n / kTileAtoms + ncol_t + 1;  // partial row per item
  HIPCHK(ctx->d_partial.ensure((size_t)std::max(blocks_for(n), tile_items) *
                               ACC_N * sizeof(double)));
  HIPCHK(ctx->d_npairs.ensure(sizeof(unsigned long long)));
  if (!ctx->shell_uploaded) {
    HIPCHK(ctx->d_shell.ensure(ctx->shell.size() * sizeof(int4)));
    HIPCHK(hipMemcpy(ctx->d_shell.p, ctx->shell.data(),
                     ctx->shell.size() * sizeof(int4), hipMemcpyHostToDevice));
    HIPCHK(ctx->d_box.ensure(sizeof(Box)));
    HIPCHK(hipMemcpy(ctx->d_box.p, &ctx->box, sizeof(Box),
                     hipMemcpyHostToDevice));
    HIPCHK(ctx->d_shell_off.ensure(ctx->shell_off.size() * sizeof(float4)));
    HIPCHK(hipMemcpy(ctx->d_shell_off.p, ctx->shell_off.data(),
                     ctx->shell_off.size() * sizeof(float4),
                     hipMemcpyHostToDevice));
    ctx->shell_uploaded = 1;
  }
  if (ctx->use_tile) {
    const int ncol = ctx->ncol_total, nrep = (int)ctx->reps.size();
    HIPCHK(ctx->d_items.ensure((size_t)(n / kTileAtoms + ncol + 1) *
                               sizeof(int4)));
    HIPCHK(ctx->d_nitems.ensure(sizeof(int)));
    HIPCHK(ctx->d_rep_first.ensure((size_t)(nrep + 1) * sizeof(int)));
    HIPCHK(ctx->d_colscan.ensure((size_t)3 * std::max(ncol, 1) * sizeof(int)));
    if (!ctx->rep_uploaded) {
      HIPCHK(ctx->d_reps.ensure((size_t)nrep * sizeof(TileRep)));
      HIPCHK(hipMemcpy(ctx->d_reps.p, ctx->reps.data(),
                       (size_t)nrep * sizeof(TileRep), hipMemcpyHostToDevice));
      HIPCHK(ctx->d_col_rep.ensure((size_t)ncol * sizeof(int)));
      HIPCHK(hipMemcpy(ctx->d_col_rep.p, ctx->col_rep.data(),
                       (size_t)ncol * sizeof(int), hipMemcpyHostToDevice));
      ctx->rep_uploaded = 1;
    }
  }
  const bool list_ok = ctx->use_list && n < (1 << kEntryAtomBits);
  if (list_ok) {
    const ForceParams P = make_params(ctx, 0);
    ctx->cap = list_capacity(ctx, P.cut2);
    HIPCHK(ctx->d_list.ensure((size_t)ctx->cap * N * sizeof(unsigned)));
    HIPCHK(ctx->d_ncount.ensure(N * sizeof(int)));
    HIPCHK(ctx->d_partial2.ensure((size_t)blocks_for(n) * ACC_N *
                                  sizeof(double)));
  } else {
    ctx->cap = 0;
  }
  return PMNB_OK;
}

// Pair-kernel shapes of the tile path: {shell, buffered records per wave,
// waves per SIMD the register allocation must allow}.  LDS per record: 32
// bytes (mixed full shell), 56 (mixed half shell: + 3 fixed-point slots),
// 64 / 112 (FP64), four waves per block.  Mixed: 0 full 256 / 3, 1 full
// 128 / 4, 2 half 192 / 3; FP64: 0 full 192 / 3, 1 half 128 / 2.  A single
// system uses its context's shape (0 unless set); the batch autotuner
// chooses among them.
constexpr int kMixedShapes = 3, kF64Shapes = 2;

struct HalfArgs {
  int max_items;   // upper bound of the work items
  int grid_cap;    // persistent grid (batch): at most this many blocks
  hipStream_t s;
  const ForceParams *P;
  const Sorted *S;
  const TileRep *reps;
  const int4 *items;
  const int *nitems;
  long long *ffix;
  double *partial;
  unsigned long long *npairs;
  int pstride;     // partial row stride (items)
  bool single;     // one replica without wrapped x/y
  TileGrid G;      // its column grid
  PairLog log;
  double *forces;  // the caller's forces (k_tile adds into them)
  // no energy or virial rows are read: the mixed kernels leave them out
  bool forces_only = false;
};

// The single-system full-shell kernel adds its forces straight into the
// caller's array; every other shape goes through the fixed-point array.
bool direct_forces(const HalfArgs &a, int shape, bool mixed) {
  return a.single && (mixed ? shape != 2 : shape != 1);
}

template <bool F64, int EL, int SW, int BUF, int OCC, bool HALF, bool LOG>
void launch_half_log(const HalfArgs &a) {
  if (a.single) {
    // a multiple of 8 blocks (FP64 maps blocks to XCD-contiguous items)
    const int nb = 8 * blocks_for(std::max(1, blocks_for(a.max_items,
                                                         kTmWaves)), 8);
    if constexpr (!HALF && !F64 && !LOG) {
      if (a.forces_only)
        hipLaunchKernelGGL((k_tile<false, EL, SW, 256, false, false>),
                           dim3(nb), dim3(kBlock), 0, a.s, *a.P, *a.S, a.G,
                           a.items, a.nitems, a.forces, a.partial, a.pstride,
                           a.npairs, a.log);
      else
        hipLaunchKernelGGL((k_tile<false, EL, SW, 256, false>), dim3(nb),
                           dim3(kBlock), 0, a.s, *a.P, *a.S, a.G, a.items,
                           a.nitems, a.forces, a.partial, a.pstride,
                           a.npairs, a.log);
    } else if constexpr (!HALF)
      hipLaunchKernelGGL((k_tile<F64, EL, SW, F64 ? 192 : 256, LOG>),
                         dim3(nb), dim3(kBlock), 0, a.s, *a.P, *a.S, a.G,
                         a.items, a.nitems, a.forces, a.partial, a.pstride,
                         a.npairs, a.log);
    else
      hipLaunchKernelGGL((k_half<F64, EL, SW, BUF, OCC, false, HALF, LOG>),
                         dim3(nb), dim3(kBlock), 0, a.s, *a.P, *a.S, a.reps,
                         a.G, a.items, a.nitems, a.ffix, a.partial,
                         a.pstride, a.npairs, a.log);
    return;
  }
  int nb = blocks_for(a.max_items, kTmWaves);
  if (a.grid_cap > 0) nb = std::min(nb, a.grid_cap);
  nb = std::max(nb, 1);
  if constexpr (!F64 && !LOG) {
    if (a.forces_only) {
      hipLaunchKernelGGL((k_half<false, EL, SW, BUF, OCC, true, HALF, false,
                                 false>),
                         dim3(nb), dim3(kBlock), 0, a.s, *a.P, *a.S, a.reps,
                         a.G, a.items, a.nitems, a.ffix, a.partial, a.pstride,
                         a.npairs, a.log);
      return;
    }
  }
  hipLaunchKernelGGL((k_half<F64, EL, SW, BUF, OCC, true, HALF, LOG>),
                     dim3(nb), dim3(kBlock), 0, a.s, *a.P, *a.S, a.reps, a.G,
                     a.items, a.nitems, a.ffix, a.partial, a.pstride, a.npairs,
                     a.log);
}

// The pair record (a.log.out set) runs the same kernel code with the
// record compiled in.
template <bool F64, int EL, int SW, int BUF, int OCC, bool HALF>
void launch_half_shape(const HalfArgs &a) {
  if (a.log.out) launch_half_log<F64, EL, SW, BUF, OCC, HALF, true>(a);
  else launch_half_log<F64, EL, SW, BUF, OCC, HALF, false>(a);
}

template <bool F64, int EL, int SW>
void launch_half(int shape, const HalfArgs &a) {
  if constexpr (F64) {
    if (shape == 1) launch_half_shape<true, EL, SW, 128, 2, true>(a);
    else launch_half_shape<true, EL, SW, 192, 3, false>(a);
  } else {
    if (shape == 1) launch_half_shape<false, EL, SW, 128, 4, false>(a);
    else if (shape == 2) launch_half_shape<false, EL, SW, 192, 3, true>(a);
    else launch_half_shape<false, EL, SW, 256, 3, false>(a);
  }
}

// Pair kernel of the tile path for the force's precision and functors.
void launch_pair_half(bool mixed, int el, int sw, int shape,
                      const HalfArgs &a) {
  if (mixed) {
    if (el == PMNB_ELEC_NONE) launch_half<false, 0, 0>(shape, a);
    else if (el == PMNB_ELEC_COULOMB) launch_half<false, 1, 0>(shape, a);
    else launch_half<false, 2, 0>(shape, a);
  } else {
    if (el == PMNB_ELEC_NONE) launch_half<true, 0, 0>(shape, a);
    else if (el == PMNB_ELEC_COULOMB && sw == PMNB_SWITCH_C2)
      launch_half<true, 1, 0>(shape, a);
    else if (el == PMNB_ELEC_COULOMB) launch_half<true, 1, 1>(shape, a);
    else if (sw == PMNB_SWITCH_C2) launch_half<true, 2, 0>(shape, a);
    else launch_half<true, 2, 1>(shape, a);
  }
}

// Cell-list build on the device from d_pos (original order).
// FP32 copy of the LJ table (T*T records {A, B, A14, B14}).
__global__ void k_lj_to_f32(const double *__restrict__ lj, int n,
                            float4 *__restrict__ ljf) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n)
    ljf[k] = make_float4((float)lj[4 * k], (float)lj[4 * k + 1],
                         (float)lj[4 * k + 2], (float)lj[4 * k + 3]);
}

template <typename R>
AdjC<R> make_adj_consts(const ForceParams &P) {
  AdjC<R> C;
  C.on2l = (R)P.sw_lj.on2;
  C.cut2l = (R)P.sw_lj.cut2;
  C.s1l = (R)P.sw_lj.s1;
  C.s2l = (R)P.sw_lj.s2;
  C.s3l = (R)P.sw_lj.s3;
  C.on2e = (R)P.sw_el.on2;
  C.cut2e = (R)P.sw_el.cut2;
  C.s1e = (R)P.sw_el.s1;
  C.s2e = (R)P.sw_el.s2;
  C.s3e = (R)P.sw_el.s3;
  C.alpha = (R)P.alpha;
  C.c2ap = (R)P.two_alpha_pi;
  C.csf = (R)P.csf;
  return C;
}

template <bool F64, int EL, int MK>
void launch_adj(int max_items, hipStream_t s, const ForceParams &P,
                const Sorted &S, const TileRep *reps, const int4 *items,
                const int *nitems, const double *gf, double *gx, double *gq,
                double *plj, unsigned long long *np) {
  constexpr int NW = 1, BUF = 128;
  const int nb = std::max(1, blocks_for(max_items, NW));
  hipLaunchKernelGGL((k_tq_adj<F64, EL, NW, BUF, MK>), dim3(nb),
                     dim3(64 * NW), 0, s, P, make_adj_consts<float>(P),
                     make_adj_consts<double>(P), S, reps, items, nitems, gf,
                     gx, gq, plj, max_items, np);
}

template <bool F64, int EL>
void launch_adj_t(int T, int max_items, hipStream_t s, const ForceParams &P,
                  const Sorted &S, const TileRep *reps, const int4 *items,
                  const int *nitems, const double *gf, double *gx, double *gq,
                  double *plj, double *glj, unsigned long long *np) {
  if (T <= 2) {
    launch_adj<F64, EL, 3>(max_items, s, P, S, reps, items, nitems, gf, gx,
                           gq, plj, np);
    if (glj)
      hipLaunchKernelGGL(k_adj_lj_finish<3>, dim3(3 * 4), dim3(1024), 0, s,
                         plj, max_items, nitems, T, glj);
  } else {
    launch_adj<F64, EL, 10>(max_items, s, P, S, reps, items, nitems, gf, gx,
                            gq, plj, np);
    if (glj)
      hipLaunchKernelGGL(k_adj_lj_finish<10>, dim3(10 * 4), dim3(1024), 0, s,
                         plj, max_items, nitems, T, glj);
  }
}

// The per-evaluation device arrays of run_device: the context's own
// (ctx_bufs, allocated by ensure_eval_buffers) or, for
// pmnb_evaluate_device_noalloc, carved from the static scratch array.
struct EvalBufs {
  int *cell_of, *count, *start, *rank, *sorted;
  int *bucket;  // tile path: cap atom slots per cell, in arrival order
  int cap;
  double *sx, *sy, *sz, *sq;
  float4 *loc;
  int *stype, *smol, *scell;
  int2 *srec;
  long long *ffix;
  double *partial;
  unsigned long long *npairs;
  int4 *items;
  int *nitems, *rep_first;
  // 3 ints per global column (atoms, first atom, first item) for the
  // column scan of many cells; null: the one-block scan
  int *colscan;
  const TileRep *reps;
  const int *col_rep;
  const Box *box;
};

EvalBufs ctx_bufs(const pmnb_ctx *ctx) {
  EvalBufs B{};
  B.cell_of = ctx->d_cell_of.as<int>();
  B.count = ctx->d_count.as<int>();
  B.start = ctx->d_start.as<int>();
  B.rank = ctx->d_fill.as<int>();
  B.sorted = ctx->d_sorted.as<int>();
  B.bucket = ctx->d_bucket.as<int>();
  B.cap = ctx->bucket_cap;
  B.sx = ctx->d_sx.as<double>();
  B.sy = ctx->d_sy.as<double>();
  B.sz = ctx->d_sz.as<double>();
  B.sq = ctx->d_sq.as<double>();
  B.loc = ctx->d_loc.as<float4>();
  B.stype = ctx->d_stype.as<int>();
  B.smol = ctx->d_smol.as<int>();
  B.scell = ctx->d_scell.as<int>();
  B.srec = ctx->d_srec.as<int2>();
  B.ffix = ctx->d_ffix.as<long long>();
  B.partial = ctx->d_partial.as<double>();
  B.npairs = ctx->d_npairs.as<unsigned long long>();
  B.items = ctx->d_items.as<int4>();
  B.nitems = ctx->d_nitems.as<int>();
  B.rep_first = ctx->d_rep_first.as<int>();
  B.colscan = ctx->d_colscan.as<int>();
  B.reps = ctx->d_reps.as<TileRep>();
  B.col_rep = ctx->d_col_rep.as<int>();
  B.box = ctx->d_box.as<Box>();
  return B;
}

// On the tile path every replica's cells are numbered from its offset
// (ctx->grid.ncell is then the total).  counts_zero: B.count is known to
// hold zeros (tile path only), so no clearing launch is needed; the
// work-item kernel clears the pair counter.  assigned (tile path): the
// counts, cells and ranks of these positions are already in B (written by
// the integrator's kernel that produced them).  pos_copy (not with
// assigned): d_pos is read once, by the cell assignment, which stores the
// positions there for the kernels after it (d_pos may then be host memory
// mapped into the device).  Leaves ctx->counts_zero false: the counts are
// in use.
int build_cells(pmnb_ctx *ctx, const EvalBufs &B, const double *d_pos,
                hipStream_t s, bool counts_zero = false,
                bool assigned = false, double *pos_copy = nullptr) {
  const int n = ctx->n, nc = ctx->grid.ncell;
  const TileRep *reps = B.reps;
  const int *rep_of = ctx->batch ? ctx->d_rep_of.as<int>() : nullptr;
  int *rank = B.rank;
  ctx->counts_zero = false;
  assigned = assigned && ctx->use_tile;
  if (!((counts_zero || assigned) && ctx->use_tile))
    hipLaunchKernelGGL(k_clear, dim3(blocks_for(nc)), dim3(kBlock), 0, s,
                       B.count, nc, B.npairs);
  if (assigned) {
    // cells and buckets came with the positions (the integrator's kernel)
  } else if (n > 0 && ctx->use_tile) {
    hipLaunchKernelGGL(k_assign_r<TileRep>, dim3(blocks_for(n)), dim3(kBlock),
                       0, s, reps, rep_of, d_pos, n, B.cell_of,
                       B.count, B.bucket, B.cap, pos_copy);
  } else if (n > 0) {
    hipLaunchKernelGGL(k_assign, dim3(blocks_for(n)), dim3(kBlock), 0, s,
                       ctx->box, ctx->grid, d_pos, n, B.cell_of,
                       B.count, rank, pos_copy);
  }
  if (pos_copy && !assigned) d_pos = pos_copy;
  if (ctx->use_tile && B.colscan && nc > kScanCellsOneBlock) {
    // starts and work items over the columns (many cells: a batch)
    const int ncol = ctx->ncol_total;
    int *col_n = B.colscan, *col_a0 = col_n + ncol, *col_i0 = col_a0 + ncol;
    hipLaunchKernelGGL(k_col_atoms, dim3(blocks_for(ncol)), dim3(kBlock), 0,
                       s, B.count, reps, B.col_rep, ncol, col_n);
    hipLaunchKernelGGL(k_col_scan, dim3(1), dim3(1024), 0, s, col_n, ncol,
                       (int)ctx->reps.size(), col_a0, col_i0, B.nitems,
                       B.rep_first, B.npairs);
    hipLaunchKernelGGL(k_col_fill, dim3(blocks_for(ncol)), dim3(kBlock), 0,
                       s, B.count, B.start, reps, B.col_rep, ncol, col_a0,
                       col_i0, B.items, B.rep_first);
  } else if (ctx->use_tile) {
    // starts and the work items of the tile path in one block
    hipLaunchKernelGGL(k_scan_items, dim3(1), dim3(1024), 0, s,
                       B.count, B.start, nc,
                       reps, B.col_rep, ctx->ncol_total,
                       (int)ctx->reps.size(), B.items,
                       B.nitems, B.rep_first, B.npairs);
  } else if (!ctx->use_tile) {
    hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, s,
                       B.count, B.start, nc);
  }
  if (n > 0 && !ctx->use_tile) {
    hipLaunchKernelGGL(k_scatter, dim3(blocks_for(n)), dim3(kBlock), 0, s,
                       B.cell_of, n, B.start,
                       rank, B.sorted);
  }
  if (n > 0 && ctx->use_tile) {
    hipLaunchKernelGGL(k_sort_gather_r<TileRep>,
                       dim3(blocks_for(nc, kBlock / 64)), dim3(kBlock), 0, s,
                       reps, rep_of, d_pos, B.start,
                       B.count, nc, B.bucket, B.cap, B.cell_of, n,
                       ctx->d_type.as<int>(),
                       ctx->d_q.as<double>(), ctx->d_mol.as<int>(),
                       B.sorted, B.sx,
                       B.sy, B.sz,
                       B.loc, B.stype,
                       B.sq, B.smol,
                       B.scell, B.srec);
  } else if (n > 0) {
    hipLaunchKernelGGL(k_sort_cells, dim3(blocks_for(nc, kBlock / 64)),
                       dim3(kBlock), 0, s,
                       B.start, B.count, nc,
                       B.sorted);
    hipLaunchKernelGGL(k_gather, dim3(blocks_for(n)), dim3(kBlock), 0, s,
                       ctx->box, ctx->grid, d_pos, B.sorted,
                       B.cell_of, ctx->d_type.as<int>(),
                       ctx->d_q.as<double>(), ctx->d_mol.as<int>(), n,
                       B.sx, B.sy,
                       B.sz, B.loc,
                       B.stype, B.sq,
                       B.smol, B.scell);
  }
  HIPCHK(hipGetLastError());
  return PMNB_OK;
}

Sorted make_sorted(const pmnb_ctx *ctx, const EvalBufs &B,
                   const double *d_com) {
  Sorted S{};
  S.x = B.sx;
  S.y = B.sy;
  S.z = B.sz;
  S.loc = B.loc;
  S.atom = B.sorted;
  S.type = B.stype;
  S.q = B.sq;
  S.mol = B.smol;
  S.cell = B.scell;
  S.rec = B.srec;
  S.cstart = B.start;
  S.ccount = B.count;
  S.delta = ctx->d_shell.as<int4>();
  S.xfull = ctx->d_xfull.as<unsigned long long>();
  S.xmod = ctx->d_xmod.as<unsigned long long>();
  S.xown = ctx->d_xown.as<unsigned long long>();
  S.lj = ctx->d_lj.as<double>();
  S.ljf = ctx->d_ljf.as<float4>();
  S.com = d_com;
  S.box = B.box;
  return S;
}

int check_ready(pmnb_ctx *ctx) {
  if (!ctx) return PMNB_ERR_ARG;
  if (!ctx->have_topo)
    return fail(ctx, PMNB_ERR_STATE, "pmnb_set_topology has not been called");
  if (!ctx->have_box)
    return fail(ctx, PMNB_ERR_STATE, "pmnb_set_box has not been called");
  return PMNB_OK;
}

// own: B are the context's buffers (not the static scratch), whose cell
// counts the finish kernel leaves cleared for the next evaluation.  fresh:
// on the tile path the forces and the 20 output doubles are written, not
// added to (the caller's arrays need no clearing).  np_out (tile path):
// the finish kernel also stores the pair count there.  finish_skipped
// (given): no energies or virial are wanted in d_out; the mixed pair
// kernels form none, a single-system tile evaluation whose pair kernel
// wrote the forces itself launches no finish kernel (the caller clears the
// cell counts before the next evaluation), and a batch's finish kernel
// sums no energies; *finish_skipped tells whether the finish was left
// out.
int run_device_bufs(pmnb_ctx *ctx, const EvalBufs &B, const double *d_pos,
                    double *d_forces, double *d_out, const double *d_com,
                    uint32_t flags, hipStream_t s, bool own,
                    bool fresh = false,
                    unsigned long long *np_out = nullptr,
                    bool assigned = false, bool *finish_skipped = nullptr,
                    double *pos_copy = nullptr) {
  int rc = PMNB_OK;
  flags &= ~kFlagFresh;  // internal: never taken from a caller
  if (fresh && ctx->use_tile) flags |= kFlagFresh;
  if ((flags & PMNB_FLAG_MOLVIRIAL) && !d_com)
    return fail(ctx, PMNB_ERR_ARG, "molecular virial needs molecule centres");
  // timing: [start, built) = cell build (+ work items or neighbour lists),
  // [built, stop) = pair kernel
  hipEvent_t ev_built = nullptr, ev_stop = nullptr;
  if (ctx->timing) {
    while (ctx->ev_pool.size() < ctx->ev_used + 3) {
      hipEvent_t e;
      HIPCHK(hipEventCreate(&e));
      ctx->ev_pool.push_back(e);
    }
    HIPCHK(hipEventRecord(ctx->ev_pool[ctx->ev_used], s));
    ev_built = ctx->ev_pool[ctx->ev_used + 1];
    ev_stop = ctx->ev_pool[ctx->ev_used + 2];
    ctx->ev_used += 3;
  }
  rc = build_cells(ctx, B, d_pos, s, own && ctx->counts_zero, assigned,
                   pos_copy);
  if (rc) return rc;
  if (own) ctx->last_np = B.npairs;  // not the shared scratch's counter
  const ForceParams P = make_params(ctx, flags);
  const Sorted S = make_sorted(ctx, B, d_com);
  const int nb = std::max(1, blocks_for(ctx->n));
  double *part = B.partial;
  double *part2 = ctx->d_partial2.as<double>();
  unsigned long long *np = B.npairs;
  const bool mixed = ctx->prm.precision == PMNB_PREC_MIXED;
  const int el = ctx->prm.elec, sw = ctx->prm.elec_switch;
  const bool list = ctx->cap > 0;
  const int cap = ctx->cap;
  unsigned *lst = ctx->d_list.as<unsigned>();
  int *cnt = ctx->d_ncount.as<int>();
  if (ctx->use_tile) {
    const int ncol = ctx->ncol_total, nrep = (int)ctx->reps.size();
    const int max_items = ctx->n / kTileAtoms + ncol + 1;
    const TileRep *reps = B.reps;
    int4 *items = B.items;
    int *nitems = B.nitems;
    int *rep_first = B.rep_first;
    long long *ffix = B.ffix;
    bool direct = false;
    if (ev_built) HIPCHK(hipEventRecord(ev_built, s));
    if (ctx->n > 0) {
      const bool single = !ctx->batch && nrep == 1 &&
                          !ctx->reps[0].wrap[0] && !ctx->reps[0].wrap[1];
      HalfArgs a{max_items, ctx->grid_cap, s, &P, &S, reps, items, nitems,
                 ffix, part, np, max_items, single, ctx->tile, ctx->plog,
                 d_forces};
      // no energies wanted: the mixed kernels leave them out, in the full
      // shell of 256 records whatever shape the evaluations with energies
      // use (114 VGPRs, four waves per SIMD; the half shell's fixed-point
      // slots in LDS hold it to three: 3.8 against 4.8 ms per dynamics step
      // of the 512 water replicas).  A single system's kernel then needs
      // no finish; a batch's finish still converts the forces but sums no
      // energies.
      a.forces_only = finish_skipped != nullptr;
      const int shape = a.forces_only && mixed ? 0 : ctx->shape;
      direct = direct_forces(a, shape, mixed);
      launch_pair_half(mixed, el, sw, shape, a);
    }
    if (ev_stop) HIPCHK(hipEventRecord(ev_stop, s));
    if (finish_skipped) {
      *finish_skipped = direct;
      if (direct) return PMNB_OK;
    }
    // fixed-point forces into the caller's array (unless the pair kernel
    // added them itself), per-replica sums
    // and the cell counts cleared for the next evaluation
    const int ab = direct ? 0 : blocks_for(ctx->n, 1024);
    const int rb = finish_skipped ? 0 : nrep * ACC_N, nc = ctx->grid.ncell;
    const int fresh = (flags & kFlagFresh) ? 1 : 0;
    const dim3 fg((unsigned)(ab + rb + blocks_for(nc, 1024)));
    if (mixed)
      hipLaunchKernelGGL(k_half_finish<false>, fg, dim3(1024), 0, s, ffix,
                         S.atom, ctx->n, d_forces, ab, part, max_items,
                         rep_first, d_out, rb, B.count, nc, fresh, np,
                         np_out);
    else
      hipLaunchKernelGGL(k_half_finish<true>, fg, dim3(1024), 0, s, ffix,
                         S.atom, ctx->n, d_forces, ab, part, max_items,
                         rep_first, d_out, rb, B.count, nc, fresh, np,
                         np_out);
    HIPCHK(hipGetLastError());
    if (own) ctx->counts_zero = true;
    return PMNB_OK;
  }
  // only / cap of the direct-scan kernels: everything (no list) or the
  // atoms whose neighbours overflowed the list
  const int *only = list ? cnt : nullptr;
  if (list) {
    if (ctx->per_pair_wrap)
      hipLaunchKernelGGL(k_build_list<true>, dim3(nb), dim3(kBlock), 0, s, P,
                         S, lst, cnt, cap, np, ctx->plog);
    else
      hipLaunchKernelGGL(k_build_list<false>, dim3(nb), dim3(kBlock), 0, s, P,
                         S, lst, cnt, cap, np, ctx->plog);
  }
  if (ev_built) HIPCHK(hipEventRecord(ev_built, s));
  if (list && mixed) {
    const float4 *so = ctx->d_shell_off.as<float4>();
    if (el == PMNB_ELEC_NONE)
      hipLaunchKernelGGL(k_force_list_mixed<0>, dim3(nb), dim3(kBlock), 0, s,
                         P, S, lst, cnt, cap, so, d_forces, part);
    else if (el == PMNB_ELEC_COULOMB)
      hipLaunchKernelGGL(k_force_list_mixed<1>, dim3(nb), dim3(kBlock), 0, s,
                         P, S, lst, cnt, cap, so, d_forces, part);
    else
      hipLaunchKernelGGL(k_force_list_mixed<2>, dim3(nb), dim3(kBlock), 0, s,
                         P, S, lst, cnt, cap, so, d_forces, part);
  } else if (list) {
    if (el == PMNB_ELEC_NONE)
      hipLaunchKernelGGL((k_force_list_fp64<0, 0>), dim3(nb), dim3(kBlock), 0,
                         s, P, S, lst, cnt, cap, d_forces, part);
    else if (el == PMNB_ELEC_COULOMB && sw == PMNB_SWITCH_C2)
      hipLaunchKernelGGL((k_force_list_fp64<1, 0>), dim3(nb), dim3(kBlock), 0,
                         s, P, S, lst, cnt, cap, d_forces, part);
    else if (el == PMNB_ELEC_COULOMB)
      hipLaunchKernelGGL((k_force_list_fp64<1, 1>), dim3(nb), dim3(kBlock), 0,
                         s, P, S, lst, cnt, cap, d_forces, part);
    else if (sw == PMNB_SWITCH_C2)
      hipLaunchKernelGGL((k_force_list_fp64<2, 0>), dim3(nb), dim3(kBlock), 0,
                         s, P, S, lst, cnt, cap, d_forces, part);
    else
      hipLaunchKernelGGL((k_force_list_fp64<2, 1>), dim3(nb), dim3(kBlock), 0,
                         s, P, S, lst, cnt, cap, d_forces, part);
  }
  // direct scan: the whole force without a list, else the overflow atoms
  double *pd = list ? part2 : part;
  if (mixed) {
    if (el == PMNB_ELEC_NONE)
      hipLaunchKernelGGL(k_pair_mixed<0>, dim3(nb), dim3(kBlock), 0, s, P, S,
                         d_forces, pd, np, only, cap, ctx->plog);
    else if (el == PMNB_ELEC_COULOMB)
      hipLaunchKernelGGL(k_pair_mixed<1>, dim3(nb), dim3(kBlock), 0, s, P, S,
                         d_forces, pd, np, only, cap, ctx->plog);
    else
      hipLaunchKernelGGL(k_pair_mixed<2>, dim3(nb), dim3(kBlock), 0, s, P, S,
                         d_forces, pd, np, only, cap, ctx->plog);
  } else {
    if (el == PMNB_ELEC_NONE)
      hipLaunchKernelGGL((k_pair_fp64<0, 0>), dim3(nb), dim3(kBlock), 0, s, P,
                         S, d_forces, pd, np, only, cap, ctx->plog);
    else if (el == PMNB_ELEC_COULOMB && sw == PMNB_SWITCH_C2)
      hipLaunchKernelGGL((k_pair_fp64<1, 0>), dim3(nb), dim3(kBlock), 0, s, P,
                         S, d_forces, pd, np, only, cap, ctx->plog);
    else if (el == PMNB_ELEC_COULOMB)
      hipLaunchKernelGGL((k_pair_fp64<1, 1>), dim3(nb), dim3(kBlock), 0, s, P,
                         S, d_forces, pd, np, only, cap, ctx->plog);
    else if (sw == PMNB_SWITCH_C2)
      hipLaunchKernelGGL((k_pair_fp64<2, 0>), dim3(nb), dim3(kBlock), 0, s, P,
                         S, d_forces, pd, np, only, cap, ctx->plog);
    else
      hipLaunchKernelGGL((k_pair_fp64<2, 1>), dim3(nb), dim3(kBlock), 0, s, P,
                         S, d_forces, pd, np, only, cap, ctx->plog);
  }
  if (ev_stop) HIPCHK(hipEventRecord(ev_stop, s));
  hipLaunchKernelGGL(k_finish, dim3(1), dim3(1024), 0, s, part, nb,
                     list ? part2 : part, list ? nb : 0, d_out);
  HIPCHK(hipGetLastError());
  return PMNB_OK;
}

int run_device(pmnb_ctx *ctx, const double *d_pos, double *d_forces,
               double *d_out, const double *d_com, uint32_t flags,
               hipStream_t s, bool fresh = false, bool assigned = false,
               bool *finish_skipped = nullptr) {
  const int rc = ensure_eval_buffers(ctx);
  if (rc) return rc;
  return run_device_bufs(ctx, ctx_bufs(ctx), d_pos, d_forces, d_out, d_com,
                         flags, s, true, fresh, nullptr, assigned,
                         finish_skipped);
}

}  // namespace

namespace {
void scratch_init(int device);  // the evaluation without allocation, below
}  // namespace

extern "C" {

int pmnb_version(void) { return 100; }

int pmnb_create(const pmnb_params *params, pmnb_ctx **out) {
  pmnb_ctx *ctx = nullptr;
  if (!params || !out) return PMNB_ERR_ARG;
  *out = nullptr;
  if (params->elec < PMNB_ELEC_NONE || params->elec > PMNB_ELEC_EWALD_REAL)
    return PMNB_ERR_ARG;
  if (params->elec_switch != PMNB_SWITCH_C2 &&
      params->elec_switch != PMNB_SWITCH_C1)
    return PMNB_ERR_ARG;
  if (params->precision != PMNB_PREC_FP64 &&
      params->precision != PMNB_PREC_MIXED)
    return PMNB_ERR_ARG;
  if (params->precision == PMNB_PREC_MIXED &&
      params->elec_switch == PMNB_SWITCH_C1)
    return PMNB_ERR_UNSUPPORTED;
  // C2SwitchingFunction::make: 0 <= switchon < cutoff
  // (switch/C2SwitchingFunction.cpp:30-41).
  if (!(params->lj_switchon >= 0.0) ||
      !(params->lj_switchon < params->lj_cutoff))
    return PMNB_ERR_ARG;
  if (params->elec_switch == PMNB_SWITCH_C2 &&
      (!(params->elec_switchon >= 0.0) ||
       !(params->elec_switchon < params->elec_cutoff)))
    return PMNB_ERR_ARG;
  if (!(params->elec_cutoff > 0.0) || !(params->algorithm_cutoff > 0.0) ||
      !(params->cell_size > 0.0))
    return PMNB_ERR_ARG;
  if (params->elec == PMNB_ELEC_EWALD_REAL && !(params->ewald_alpha > 0.0))
    return PMNB_ERR_ARG;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return PMNB_ERR_HIP;
  if (params->device < 0 || params->device >= ndev) return PMNB_ERR_ARG;
  if (hipSetDevice(params->device) != hipSuccess) return PMNB_ERR_HIP;
  ctx = new pmnb_ctx();
  ctx->prm = *params;
  // a blocking stream: work enqueued on the legacy default stream (a
  // caller's stream handle 0) stays ordered with it
  if (hipStreamCreate(&ctx->stream) != hipSuccess) {
    delete ctx;
    return PMNB_ERR_HIP;
  }
  // the device address of the static scratch on this device (once per
  // device, with it current)
  scratch_init(params->device);
  *out = ctx;
  return PMNB_OK;
}

int pmnb_set_topology(pmnb_ctx *ctx, int32_t n_atoms, const int32_t *type,
                      const double *scaled_charge, const int32_t *molecule,
                      int32_t n_types, const double *lj,
                      double coulomb_scaling_factor, int32_t n_excl,
                      const int32_t *excl) {
  if (!ctx) return PMNB_ERR_ARG;
  if (n_atoms < 0 || n_types <= 0 || n_excl < 0)
    return fail(ctx, PMNB_ERR_ARG, "negative sizes");
  if (n_atoms > 0 && (!type || !scaled_charge || !molecule))
    return fail(ctx, PMNB_ERR_ARG, "null per-atom array");
  if (!lj) return fail(ctx, PMNB_ERR_ARG, "null LJ table");
  if (n_excl > 0 && !excl) return fail(ctx, PMNB_ERR_ARG, "null exclusions");
  int nmol = 0;
  for (int i = 0; i < n_atoms; i++) {
    if (type[i] < 0 || type[i] >= n_types)
      return fail(ctx, PMNB_ERR_ARG, "atom %d has type %d outside [0,%d)", i,
                  type[i], n_types);
    if (molecule[i] < 0)
      return fail(ctx, PMNB_ERR_ARG, "atom %d has molecule id %d", i,
                  molecule[i]);
    nmol = std::max(nmol, molecule[i] + 1);
  }
  HIPCHK(hipSetDevice(ctx->prm.device));
  // Exclusion masks: last writer wins per pair (ExclusionTable.cpp:105-120).
  std::vector<unsigned long long> full((size_t)std::max(n_atoms, 1), 0ull),
      mod((size_t)std::max(n_atoms, 1), 0ull);
  int maxd = 0;
  for (int k = 0; k < n_excl; k++) {
    int a = excl[3 * k], b = excl[3 * k + 1], cls = excl[3 * k + 2];
    if (a < 0 || b < 0 || a >= n_atoms || b >= n_atoms)
      return fail(ctx, PMNB_ERR_ARG, "exclusion %d (%d,%d) out of range", k, a,
                  b);
    if (cls < PMNB_EXCL_NONE || cls > PMNB_EXCL_FULL)
      return fail(ctx, PMNB_ERR_ARG, "exclusion %d has class %d", k, cls);
    if (a == b) continue;
    if (a > b) std::swap(a, b);
    const int d = b - a;
    if (molecule[a] != molecule[b]) continue;  // never consulted
    if (d > kMaxExclDelta)
      return fail(ctx, PMNB_ERR_UNSUPPORTED,
                  "exclusion (%d,%d) spans %d atoms; at most %d are held", a,
                  b, d, kMaxExclDelta);
    const unsigned long long bit = 1ull << (d - 1);
    full[a] &= ~bit;
    mod[a] &= ~bit;
    if (cls == PMNB_EXCL_FULL) full[a] |= bit;
    else if (cls == PMNB_EXCL_MODIFIED) mod[a] |= bit;
    maxd = std::max(maxd, d);
  }
  // Each atom's own view of its exclusions: forward masks (partners a + d)
  // and reverse masks (partners a - d, bit d - 1), so the tile kernels read
  // only the masks of their own atoms.
  std::vector<unsigned long long> xown((size_t)std::max(n_atoms, 1) * 4, 0ull);
  for (int a = 0; a < n_atoms; a++) {
    xown[4 * (size_t)a] = full[a];
    xown[4 * (size_t)a + 1] = mod[a];
    const unsigned long long fb = full[a], mb = mod[a];
    for (unsigned long long r = fb | mb; r; r &= r - 1) {
      const int d = __builtin_ctzll(r) + 1;
      const unsigned long long bit = 1ull << (d - 1);
      if (fb & bit) xown[4 * (size_t)(a + d) + 2] |= bit;
      if (mb & bit) xown[4 * (size_t)(a + d) + 3] |= bit;
    }
  }
  // LJ table: entries (i <= j) mirrored, as LennardJonesParameterTable::set
  // holds one record per unordered type pair.
  const int T = n_types;
  std::vector<double> ljd((size_t)T * T * 4);
  std::vector<float4> ljf((size_t)T * T);
  for (int i = 0; i < T; i++)
    for (int j = 0; j < T; j++) {
      const int lo = std::min(i, j), hi = std::max(i, j);
      const double *src = lj + 4 * ((size_t)lo * T + hi);
      for (int k = 0; k < 4; k++) ljd[4 * ((size_t)i * T + j) + k] = src[k];
      ljf[(size_t)i * T + j] = make_float4((float)src[0], (float)src[1],
                                           (float)src[2], (float)src[3]);
    }
  const size_t N = (size_t)std::max(n_atoms, 1);
  HIPCHK(ctx->d_type.ensure(N * sizeof(int)));
  HIPCHK(ctx->d_q.ensure(N * sizeof(double)));
  HIPCHK(ctx->d_mol.ensure(N * sizeof(int)));
  HIPCHK(ctx->d_xfull.ensure(N * sizeof(unsigned long long)));
  HIPCHK(ctx->d_xmod.ensure(N * sizeof(unsigned long long)));
  HIPCHK(ctx->d_xown.ensure(N * 4 * sizeof(unsigned long long)));
  HIPCHK(ctx->d_lj.ensure(ljd.size() * sizeof(double)));
  HIPCHK(ctx->d_ljf.ensure(ljf.size() * sizeof(float4)));
  if (n_atoms > 0) {
    HIPCHK(hipMemcpy(ctx->d_type.p, type, n_atoms * sizeof(int),
                     hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->d_q.p, scaled_charge, n_atoms * sizeof(double),
                     hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->d_mol.p, molecule, n_atoms * sizeof(int),
                     hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->d_xfull.p, full.data(), n_atoms * 8,
                     hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->d_xmod.p, mod.data(), n_atoms * 8,
                     hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->d_xown.p, xown.data(), (size_t)n_atoms * 32,
                     hipMemcpyHostToDevice));
  }
  HIPCHK(hipMemcpy(ctx->d_lj.p, ljd.data(), ljd.size() * sizeof(double),
                   hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(ctx->d_ljf.p, ljf.data(), ljf.size() * sizeof(float4),
                   hipMemcpyHostToDevice));
  ctx->n = n_atoms;
  ctx->ntypes = T;
  ctx->nmol = nmol;
  ctx->csf = coulomb_scaling_factor;
  ctx->max_excl_delta = maxd;
  ctx->have_topo = true;
  ctx->topo_gen++;
  // the choice of pair kernel depends on the topology: rebuild the grid
  if (ctx->have_box) {
    const double *v = ctx->box_key;
    const int rc = build_box(ctx, v, v + 3, v + 6, v + 9);
    if (rc != PMNB_OK) ctx->have_box = false;
    return rc;
  }
  return PMNB_OK;
}

int pmnb_set_box(pmnb_ctx *ctx, const double e1[3], const double e2[3],
                 const double e3[3], const double origin[3]) {
  if (!ctx || !e1 || !e2 || !e3 || !origin) return PMNB_ERR_ARG;
  double v[12];
  for (int k = 0; k < 3; k++) {
    v[k] = e1[k];
    v[3 + k] = e2[k];
    v[6 + k] = e3[k];
    v[9 + k] = origin[k];
  }
  if (ctx->have_box && std::memcmp(v, ctx->box_key, sizeof v) == 0)
    return PMNB_OK;  // unchanged: keep the grid, shell and uploads
  const int rc = build_box(ctx, e1, e2, e3, origin);
  if (rc == PMNB_OK) std::memcpy(ctx->box_key, v, sizeof v);
  return rc;
}

int pmnb_evaluate_device(pmnb_ctx *ctx, const double *d_positions,
                         double *d_forces, double *d_out,
                         const double *d_mol_com, uint32_t flags,
                         void *stream) {
  int rc = check_ready(ctx);
  if (rc) return rc;
  if (!d_positions || !d_forces || !d_out)
    return fail(ctx, PMNB_ERR_ARG, "null device pointer");
  HIPCHK(hipSetDevice(ctx->prm.device));
  hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
  return run_device(ctx, d_positions, d_forces, d_out, d_mol_com, flags, s);
}

int pmnb_evaluate(pmnb_ctx *ctx, const double *positions, double *forces,
                  double *energies, double *virial, double *molvirial,
                  const double *mol_com, uint32_t flags) {
  int rc = check_ready(ctx);
  if (rc) return rc;
  if (ctx->n > 0 && (!positions || !forces))
    return fail(ctx, PMNB_ERR_ARG, "null positions or forces");
  if (!energies) return fail(ctx, PMNB_ERR_ARG, "null energies");
  if ((flags & (PMNB_FLAG_VIRIAL | PMNB_FLAG_MOLVIRIAL)) && !virial)
    return fail(ctx, PMNB_ERR_ARG, "null virial");
  if ((flags & PMNB_FLAG_MOLVIRIAL) && (!molvirial || !mol_com))
    return fail(ctx, PMNB_ERR_ARG, "null molecular virial or centres");
  HIPCHK(hipSetDevice(ctx->prm.device));
  const size_t N = (size_t)std::max(ctx->n, 1);
  const size_t n3 = (size_t)ctx->n * 3;
  hipStream_t s = ctx->stream;
  // pinned staging, mapped into the device and coherent (allocated once per
  // size).  The host threads copy the positions in, and the cell assignment
  // reads them from there over the bus and keeps a device copy for the
  // kernels after it: no DMA transfer, whose hand-over between the copy
  // engine and the kernels took longer than the data.  On the tile path the
  // kernels write the forces, the 20 output doubles and the pair count
  // straight into the buffer (else: device results copied back at once).
  const size_t nres = n3 + 21;
  HIPCHK(ctx->d_pos.ensure(N * 3 * sizeof(double)));
  const size_t want = (2 * N * 3 + 24) * sizeof(double);
  if (ctx->h_stage_bytes < want) {
    if (ctx->h_stage) HIPCHK(hipHostFree(ctx->h_stage));
    ctx->h_stage = nullptr;
    ctx->h_stage_bytes = 0;
    HIPCHK(hipHostMalloc((void **)&ctx->h_stage, want,
                         hipHostMallocMapped | hipHostMallocCoherent));
    ctx->h_stage_bytes = want;
  }
  double *h_pos = ctx->h_stage, *h_frc = ctx->h_stage + N * 3;
  double *h_out = h_frc + n3;
  double *g_stage = nullptr;  // device address of h_stage
  HIPCHK(hipHostGetDevicePointer((void **)&g_stage, ctx->h_stage, 0));
  const double *g_pos = g_stage;
  double *g_frc = g_stage + N * 3;
  for (hipEvent_t &e : ctx->h_ev)
    if (!e) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  // large systems: the copy in and the force add are split over the crew
  const bool big = n3 >= (size_t)kHostCrewMin;
  if (big && !ctx->crew_started) {
    ctx->crew.start(host_crew_size());
    ctx->crew_started = true;
  }
  const int T = big ? ctx->crew.size() : 1;
  if (n3)
    ctx->crew.run([&](int k) {
      if (k >= T) return;
      const size_t a = n3 * k / T, b = n3 * (k + 1) / T;
      std::memcpy(h_pos + a, positions + a, (b - a) * sizeof(double));
    });
  const double *d_com = nullptr;
  if (flags & PMNB_FLAG_MOLVIRIAL) {
    // centres are indexed by molecule id
    const int nmol = std::max(ctx->nmol, 1);
    HIPCHK(ctx->d_com.ensure((size_t)nmol * 3 * sizeof(double)));
    HIPCHK(hipMemcpyAsync(ctx->d_com.p, mol_com, ctx->nmol * 3 * sizeof(double),
                          hipMemcpyHostToDevice, s));
    d_com = ctx->d_com.as<double>();
  }
  rc = ensure_eval_buffers(ctx);
  if (rc) return rc;
  if (ctx->use_tile) {
    rc = run_device_bufs(ctx, ctx_bufs(ctx), g_pos, g_frc, g_frc + n3, d_com,
                         flags, s, true, true,
                         reinterpret_cast<unsigned long long *>(g_frc + n3 + 20),
                         false, nullptr, ctx->d_pos.as<double>());
    if (rc) return rc;
  } else {
    HIPCHK(ctx->d_force.ensure((N * 3 + 24) * sizeof(double)));
    double *d_res = ctx->d_force.as<double>();
    launch_fill_i32(ctx->d_force.as<int>(), (int)(2 * nres), 0, s);
    EvalBufs B = ctx_bufs(ctx);
    B.npairs = reinterpret_cast<unsigned long long *>(d_res + n3 + 20);
    rc = run_device_bufs(ctx, B, g_pos, d_res, d_res + n3, d_com, flags, s,
                         true, false, nullptr, false, nullptr,
                         ctx->d_pos.as<double>());
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(h_frc, d_res, nres * sizeof(double),
                          hipMemcpyDeviceToHost, s));
  }
  HIPCHK(hipEventRecord(ctx->h_ev[0], s));
  hipError_t q;
  while ((q = hipEventQuery(ctx->h_ev[0])) == hipErrorNotReady) {
  }
  HIPCHK(q);
  ctx->crew.run([&](int k) {
    if (k >= T) return;
    const size_t a = n3 * k / T, b = n3 * (k + 1) / T;
    for (size_t i = a; i < b; i++) forces[i] += h_frc[i];
  });
  energies[0] += h_out[0];
  energies[1] += h_out[1];
  if (flags & (PMNB_FLAG_VIRIAL | PMNB_FLAG_MOLVIRIAL))
    for (int k = 0; k < 9; k++) virial[k] += h_out[2 + k];
  if (flags & PMNB_FLAG_MOLVIRIAL)
    for (int k = 0; k < 9; k++) molvirial[k] += h_out[11 + k];
  unsigned long long np = 0;
  std::memcpy(&np, h_out + 20, sizeof np);
  ctx->last_pairs = (long long)np;
  return PMNB_OK;
}

int pmnb_pairset(pmnb_ctx *ctx, const double *positions, int32_t *pairs,
                 int64_t capacity, int64_t *n_pairs) {
  int rc = check_ready(ctx);
  if (rc) return rc;
  if (!n_pairs || (ctx->n > 0 && !positions) || capacity < 0)
    return fail(ctx, PMNB_ERR_ARG, "bad pairset arguments");
  HIPCHK(hipSetDevice(ctx->prm.device));
  hipStream_t s = ctx->stream;
  const size_t N = (size_t)std::max(ctx->n, 1);
  HIPCHK(ctx->d_pos.ensure(N * 3 * sizeof(double)));
  if (ctx->n > 0)
    HIPCHK(hipMemcpyAsync(ctx->d_pos.p, positions, ctx->n * 3 * sizeof(double),
                          hipMemcpyHostToDevice, s));
  rc = ensure_eval_buffers(ctx);
  if (rc) return rc;
  rc = build_cells(ctx, ctx_bufs(ctx), ctx->d_pos.as<double>(), s);
  if (rc) return rc;
  const long long cap = pairs ? capacity : 0;
  HIPCHK(ctx->d_pairs.ensure((size_t)std::max<long long>(cap, 1) * 3 *
                             sizeof(int)));
  const ForceParams P = make_params(ctx, 0);
  const Sorted S = make_sorted(ctx, ctx_bufs(ctx), nullptr);
  if (ctx->n > 0)
    hipLaunchKernelGGL(k_pairset, dim3(blocks_for(ctx->n)), dim3(kBlock), 0, s,
                       P, S, ctx->d_pairs.as<int>(), cap,
                       ctx->d_npairs.as<unsigned long long>());
  HIPCHK(hipGetLastError());
  unsigned long long np = 0;
  HIPCHK(hipMemcpyAsync(&np, ctx->d_npairs.p, sizeof np, hipMemcpyDeviceToHost,
                        s));
  HIPCHK(hipStreamSynchronize(s));
  *n_pairs = (int64_t)np;
  if (!pairs) return PMNB_OK;
  if ((long long)np > cap)
    return fail(ctx, PMNB_ERR_CAPACITY, "pair set has %llu pairs, capacity %lld",
                np, cap);
  if (np > 0)
    HIPCHK(hipMemcpy(pairs, ctx->d_pairs.p, (size_t)np * 3 * sizeof(int),
                     hipMemcpyDeviceToHost));
  // sort by (i, j)
  std::vector<long long> key((size_t)np);
  std::vector<int32_t> cls((size_t)np);
  for (size_t k = 0; k < (size_t)np; k++) {
    key[k] = (long long)pairs[3 * k] * ctx->n + pairs[3 * k + 1];
    cls[k] = pairs[3 * k + 2];
  }
  std::vector<size_t> ord((size_t)np);
  for (size_t k = 0; k < ord.size(); k++) ord[k] = k;
  std::sort(ord.begin(), ord.end(),
            [&](size_t x, size_t y) { return key[x] < key[y]; });
  for (size_t k = 0; k < ord.size(); k++) {
    pairs[3 * k] = (int32_t)(key[ord[k]] / ctx->n);
    pairs[3 * k + 1] = (int32_t)(key[ord[k]] % ctx->n);
    pairs[3 * k + 2] = cls[ord[k]];
  }
  return PMNB_OK;
}

namespace {

// One evaluation at device positions d_pos with the force kernels' pair
// record on; the forces and energies go to scratch arrays.  Returns the
// recorded (i, j, class) triples sorted by (i, j) in pairs (when capacity
// allows) and their number.
int record_force_pairs(pmnb_ctx *ctx, const double *d_pos, int32_t *pairs,
                       int64_t capacity, int64_t *n_pairs) {
  hipStream_t s = ctx->stream;
  const size_t N = (size_t)std::max(ctx->n, 1);
  const int nout = 20 * std::max(1, (int)ctx->reps.size());
  HIPCHK(ctx->d_force.ensure(N * 3 * sizeof(double)));
  HIPCHK(ctx->d_out.ensure((size_t)nout * sizeof(double)));
  const long long cap = pairs ? capacity : 0;
  HIPCHK(ctx->d_plog.ensure((size_t)std::max<long long>(cap, 1) * 3 *
                            sizeof(int)));
  HIPCHK(ctx->d_plog_count.ensure(sizeof(unsigned long long)));
  launch_fill_i32(ctx->d_force.as<int>(), (int)(N * 6), 0, s);
  launch_fill_i32(ctx->d_out.as<int>(), 2 * nout, 0, s);
  launch_fill_i32(ctx->d_plog_count.as<int>(), 2, 0, s);
  ctx->plog = PairLog{ctx->d_plog.as<int>(), cap,
                      ctx->d_plog_count.as<unsigned long long>()};
  const int rc = run_device(ctx, d_pos, ctx->d_force.as<double>(),
                            ctx->d_out.as<double>(), nullptr, 0, s);
  ctx->plog = PairLog{nullptr, 0, nullptr};
  if (rc) return rc;
  unsigned long long np = 0;
  HIPCHK(hipMemcpyAsync(&np, ctx->d_plog_count.p, sizeof np,
                        hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  *n_pairs = (int64_t)np;
  if (!pairs) return PMNB_OK;
  if ((long long)np > cap)
    return fail(ctx, PMNB_ERR_CAPACITY, "pair set has %llu pairs, capacity %lld",
                np, cap);
  if (np > 0)
    HIPCHK(hipMemcpy(pairs, ctx->d_plog.p, (size_t)np * 3 * sizeof(int),
                     hipMemcpyDeviceToHost));
  std::vector<size_t> ord((size_t)np);
  for (size_t k = 0; k < ord.size(); k++) ord[k] = k;
  std::sort(ord.begin(), ord.end(), [&](size_t x, size_t y) {
    if (pairs[3 * x] != pairs[3 * y]) return pairs[3 * x] < pairs[3 * y];
    return pairs[3 * x + 1] < pairs[3 * y + 1];
  });
  std::vector<int32_t> tmp((size_t)np * 3);
  for (size_t k = 0; k < ord.size(); k++)
    for (int c = 0; c < 3; c++) tmp[3 * k + c] = pairs[3 * ord[k] + c];
  std::copy(tmp.begin(), tmp.end(), pairs);
  return PMNB_OK;
}

}  // namespace

int pmnb_force_pairs(pmnb_ctx *ctx, const double *positions, int32_t *pairs,
                     int64_t capacity, int64_t *n_pairs) {
  int rc = check_ready(ctx);
  if (rc) return rc;
  if (!n_pairs || (ctx->n > 0 && !positions) || capacity < 0)
    return fail(ctx, PMNB_ERR_ARG, "bad force-pairs arguments");
  HIPCHK(hipSetDevice(ctx->prm.device));
  const size_t N = (size_t)std::max(ctx->n, 1);
  HIPCHK(ctx->d_pos.ensure(N * 3 * sizeof(double)));
  if (ctx->n > 0)
    HIPCHK(hipMemcpyAsync(ctx->d_pos.p, positions, ctx->n * 3 * sizeof(double),
                          hipMemcpyHostToDevice, ctx->stream));
  return record_force_pairs(ctx, ctx->d_pos.as<double>(), pairs, capacity,
                            n_pairs);
}

int pmnb_kernel_timing(pmnb_ctx *ctx, int32_t enable, double *ms,
                       int64_t *launches) {
  if (!ctx) return PMNB_ERR_ARG;
  HIPCHK(hipSetDevice(ctx->prm.device));
  double build = 0.0, pair = 0.0;
  const size_t used = ctx->ev_used;
  if (used) HIPCHK(hipEventSynchronize(ctx->ev_pool[used - 1]));
  for (size_t k = 0; k + 2 < used; k += 3) {
    float a = 0.f, b = 0.f;
    HIPCHK(hipEventElapsedTime(&a, ctx->ev_pool[k], ctx->ev_pool[k + 1]));
    HIPCHK(hipEventElapsedTime(&b, ctx->ev_pool[k + 1], ctx->ev_pool[k + 2]));
    build += a;
    pair += b;
  }
  if (ms) {
    ms[0] = build;
    ms[1] = pair;
  }
  if (launches) *launches = (int64_t)(used / 3);
  ctx->ev_used = 0;
  ctx->timing = enable != 0;
  return PMNB_OK;
}

int pmnb_pair_path(const pmnb_ctx *ctx, int32_t *path) {
  if (!ctx || !path) return PMNB_ERR_ARG;
  if (!ctx->have_box || !ctx->have_topo) return PMNB_ERR_STATE;
  const bool list = ctx->use_list && ctx->n < (1 << kEntryAtomBits);
  *path = ctx->use_tile ? PMNB_PATH_TILE : (list ? PMNB_PATH_LIST
                                                 : PMNB_PATH_SCAN);
  return PMNB_OK;
}

int pmnb_last_pair_count(pmnb_ctx *ctx, int64_t *n_pairs) {
  if (!ctx || !n_pairs) return PMNB_ERR_ARG;
  if (ctx->last_np) {
    // the counter of the last launch, on whichever stream it ran
    HIPCHK(hipSetDevice(ctx->prm.device));
    HIPCHK(hipDeviceSynchronize());
    unsigned long long np = 0;
    HIPCHK(hipMemcpy(&np, ctx->last_np, sizeof np, hipMemcpyDeviceToHost));
    ctx->last_pairs = (long long)np;
  }
  *n_pairs = ctx->last_pairs;
  return PMNB_OK;
}

const char *pmnb_last_error(const pmnb_ctx *ctx) {
  return ctx ? ctx->err.c_str() : "null context";
}

void pmnb_destroy(pmnb_ctx *ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->prm.device);
  (void)hipStreamSynchronize(ctx->stream);
  DevBuf *bufs[] = {&ctx->d_type, &ctx->d_q, &ctx->d_mol, &ctx->d_xfull,
                    &ctx->d_xmod, &ctx->d_xown, &ctx->d_lj, &ctx->d_ljf, &ctx->d_pos,
                    &ctx->d_force, &ctx->d_out, &ctx->d_com, &ctx->d_cell_of,
                    &ctx->d_count, &ctx->d_start, &ctx->d_fill, &ctx->d_sorted,
                    &ctx->d_sx, &ctx->d_sy, &ctx->d_sz, &ctx->d_loc,
                    &ctx->d_stype, &ctx->d_sq, &ctx->d_smol, &ctx->d_scell,
                    &ctx->d_srec, &ctx->d_shell, &ctx->d_partial, &ctx->d_npairs,
                    &ctx->d_pairs, &ctx->d_list, &ctx->d_ncount,
                    &ctx->d_partial2, &ctx->d_shell_off, &ctx->d_box,
                    &ctx->d_items,
                    &ctx->d_nitems, &ctx->d_reps, &ctx->d_col_rep,
                    &ctx->d_rep_first, &ctx->d_rep_of, &ctx->d_partial_lj,
                    &ctx->d_ffix, &ctx->d_plog, &ctx->d_plog_count,
                    &ctx->d_bucket, &ctx->d_colscan};
  for (DevBuf *b : bufs) b->release();
  for (hipEvent_t e : ctx->ev_pool) (void)hipEventDestroy(e);
  for (hipEvent_t e : ctx->h_ev)
    if (e) (void)hipEventDestroy(e);
  if (ctx->h_stage) (void)hipHostFree(ctx->h_stage);
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

// ----------------------------------------------------------- derivatives

int pmnb_set_parameters_device(pmnb_ctx *ctx, const double *d_lj,
                               const double *d_scaled_charge, void *stream) {
  if (!ctx) return PMNB_ERR_ARG;
  if (!ctx->have_topo)
    return fail(ctx, PMNB_ERR_STATE, "pmnb_set_topology has not been called");
  HIPCHK(hipSetDevice(ctx->prm.device));
  hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
  const int nlj = ctx->ntypes * ctx->ntypes;
  if (d_lj) {
    HIPCHK(hipMemcpyAsync(ctx->d_lj.p, d_lj, (size_t)nlj * 4 * sizeof(double),
                          hipMemcpyDeviceToDevice, s));
    hipLaunchKernelGGL(k_lj_to_f32, dim3(blocks_for(nlj)), dim3(kBlock), 0, s,
                       ctx->d_lj.as<double>(), nlj, ctx->d_ljf.as<float4>());
  }
  if (d_scaled_charge && ctx->n > 0)
    HIPCHK(hipMemcpyAsync(ctx->d_q.p, d_scaled_charge,
                          (size_t)ctx->n * sizeof(double),
                          hipMemcpyDeviceToDevice, s));
  HIPCHK(hipGetLastError());
  return PMNB_OK;
}

int pmnb_adjoint_device(pmnb_ctx *ctx, const double *d_positions,
                        const double *d_grad_forces, double *d_grad_positions,
                        double *d_grad_charges, double *d_grad_lj,
                        void *stream) {
  int rc = check_ready(ctx);
  if (rc) return rc;
  if (!d_positions || !d_grad_forces)
    return fail(ctx, PMNB_ERR_ARG, "null device pointer");
  if (!ctx->use_tile)
    return fail(ctx, PMNB_ERR_UNSUPPORTED,
                "the adjoint needs the tile path (algorithm cutoff >= "
                "switching cutoff, at most 8 LJ types)");
  if (ctx->prm.elec_switch != PMNB_SWITCH_C2)
    return fail(ctx, PMNB_ERR_UNSUPPORTED, "the adjoint is built for C2 only");
  if (ctx->ntypes > 4)
    return fail(ctx, PMNB_ERR_UNSUPPORTED,
                "the adjoint holds at most 4 LJ types");
  HIPCHK(hipSetDevice(ctx->prm.device));
  hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
  rc = ensure_eval_buffers(ctx);
  if (rc) return rc;
  const int ncol = ctx->ncol_total, nrep = (int)ctx->reps.size();
  const int max_items = ctx->n / kTileAtoms + ncol + 1;
  HIPCHK(ctx->d_partial_lj.ensure((size_t)max_items * 10 * 4 *
                                  sizeof(double)));
  rc = build_cells(ctx, ctx_bufs(ctx), d_positions, s);
  if (rc) return rc;
  const ForceParams P = make_params(ctx, 0);
  const Sorted S = make_sorted(ctx, ctx_bufs(ctx), nullptr);
  const TileRep *reps = ctx->d_reps.as<TileRep>();
  int4 *items = ctx->d_items.as<int4>();
  int *nitems = ctx->d_nitems.as<int>();
  // work items: built with the cell starts
  if (ctx->n > 0) {
    const bool f64 = ctx->prm.precision == PMNB_PREC_FP64;
    const int el = ctx->prm.elec, T = ctx->ntypes;
    double *plj = ctx->d_partial_lj.as<double>();
    unsigned long long *np = ctx->d_npairs.as<unsigned long long>();
#define PMNB_ADJ(F, E)                                                        \
  launch_adj_t<F, E>(T, max_items, s, P, S, reps, items, nitems,              \
                     d_grad_forces, d_grad_positions, d_grad_charges, plj,    \
                     d_grad_lj, np)
    if (f64) {
      if (el == PMNB_ELEC_NONE) PMNB_ADJ(true, 0);
      else if (el == PMNB_ELEC_COULOMB) PMNB_ADJ(true, 1);
      else PMNB_ADJ(true, 2);
    } else {
      if (el == PMNB_ELEC_NONE) PMNB_ADJ(false, 0);
      else if (el == PMNB_ELEC_COULOMB) PMNB_ADJ(false, 1);
      else PMNB_ADJ(false, 2);
    }
#undef PMNB_ADJ
  }
  HIPCHK(hipGetLastError());
  return PMNB_OK;
}

// ------------------------------------------------------------------ batch
// A batch is a context in batch mode: the replicas' atoms concatenated in
// order of addition (molecule ids and exclusion indices offset), one tile
// grid per replica, cells and columns numbered globally.

int pmnb_batch_create(const pmnb_params *params, int32_t n_types,
                      const double *lj, double coulomb_scaling_factor,
                      pmnb_batch **out) {
  if (!params || !out || !lj || n_types <= 0) return PMNB_ERR_ARG;
  *out = nullptr;
  pmnb_ctx *ctx = nullptr;
  const int rc = pmnb_create(params, &ctx);
  if (rc) return rc;
  pmnb_batch *b = new pmnb_batch();
  b->ctx = ctx;
  b->ntypes = n_types;
  b->lj.assign(lj, lj + (size_t)n_types * n_types * 4);
  b->csf = coulomb_scaling_factor;
  *out = b;
  return PMNB_OK;
}

int pmnb_batch_add(pmnb_batch *b, int32_t n_atoms, const int32_t *type,
                   const double *scaled_charge, const int32_t *molecule,
                   int32_t n_excl, const int32_t *excl, const double e1[3],
                   const double e2[3], const double e3[3],
                   const double origin[3], int32_t *replica) {
  if (!b) return PMNB_ERR_ARG;
  pmnb_ctx *ctx = b->ctx;
  if (n_atoms <= 0 || n_excl < 0 || !type || !scaled_charge || !molecule ||
      (n_excl > 0 && !excl) || !e1 || !e2 || !e3 || !origin)
    return fail(ctx, PMNB_ERR_ARG, "replica: invalid sizes or null arrays");
  Box box;
  const int rc = make_box(ctx, e1, e2, e3, origin, box);
  if (rc) return rc;
  const int off = (int)b->type.size();
  if ((long long)off + n_atoms >= (1ll << kQTypeShift))
    return fail(ctx, PMNB_ERR_UNSUPPORTED, "batch exceeds %d atoms",
                (1 << kQTypeShift) - 1);
  int nmol = 0;
  for (int i = 0; i < n_atoms; i++) {
    if (molecule[i] < 0)
      return fail(ctx, PMNB_ERR_ARG, "atom %d has molecule id %d", i,
                  molecule[i]);
    nmol = std::max(nmol, molecule[i] + 1);
  }
  for (int k = 0; k < n_excl; k++) {
    const int a = excl[3 * k], c = excl[3 * k + 1];
    if (a < 0 || c < 0 || a >= n_atoms || c >= n_atoms)
      return fail(ctx, PMNB_ERR_ARG, "exclusion %d (%d,%d) out of range", k, a,
                  c);
  }
  const int r = (int)b->boxes.size();
  for (int i = 0; i < n_atoms; i++) {
    b->type.push_back(type[i]);
    b->q.push_back(scaled_charge[i]);
    b->mol.push_back(b->next_mol + molecule[i]);
    b->rep_of.push_back(r);
  }
  for (int k = 0; k < n_excl; k++) {
    b->excl.push_back(off + excl[3 * k]);
    b->excl.push_back(off + excl[3 * k + 1]);
    b->excl.push_back(excl[3 * k + 2]);
  }
  b->next_mol += nmol;
  b->boxes.push_back(box);
  b->dirty = true;
  if (replica) *replica = r;
  return PMNB_OK;
}

namespace {

void drop_graph(pmnb_batch *b) {
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->gexec = nullptr;
  b->graph = nullptr;
}

// Uploads the concatenated topology and the replicas' tile grids.
int batch_finalize(pmnb_batch *b) {
  if (!b->dirty) return PMNB_OK;
  pmnb_ctx *ctx = b->ctx;
  const int R = (int)b->boxes.size();
  if (R == 0) return fail(ctx, PMNB_ERR_STATE, "the batch has no replica");
  drop_graph(b);
  ctx->have_box = false;
  ctx->batch = false;
  int rc = pmnb_set_topology(ctx, (int)b->type.size(), b->type.data(),
                             b->q.data(), b->mol.data(), b->ntypes,
                             b->lj.data(), b->csf, (int)b->excl.size() / 3,
                             b->excl.data());
  if (rc) return rc;
  const ForceParams P0 = make_params_for(ctx->prm, ctx->csf);
  const double rcp = std::sqrt(P0.cut2);
  if (ctx->prm.algorithm_cutoff < rcp)
    return fail(ctx, PMNB_ERR_UNSUPPORTED,
                "a batch needs the algorithm cutoff >= the switching cutoff");
  if (ctx->ntypes * ctx->ntypes > kQLj)
    return fail(ctx, PMNB_ERR_UNSUPPORTED, "a batch holds at most %d LJ types",
                8);
  ctx->reps.clear();
  ctx->col_rep.clear();
  long long cells = 0;
  int cols = 0;
  std::vector<int4> shell;
  for (int r = 0; r < R; r++) {
    TileRep rep{};
    if (!make_tile_rep(b->boxes[r], rcp, tile_zfrac(ctx->prm), rep, shell))
      return fail(ctx, PMNB_ERR_UNSUPPORTED, "replica %d grid too large", r);
    rep.cell_off = (int)cells;
    rep.col_off = cols;
    cells += rep.grid.ncell;
    cols += rep.ncol;
    if (cells >= (1ll << 28))
      return fail(ctx, PMNB_ERR_UNSUPPORTED, "batch cell grid too large");
    for (int c = 0; c < rep.ncol; c++) ctx->col_rep.push_back(r);
    ctx->reps.push_back(rep);
  }
  ctx->use_tile = true;
  ctx->use_list = false;
  ctx->batch = true;
  ctx->tile = ctx->reps[0].tg;
  ctx->box = ctx->reps[0].box;
  ctx->grid = Grid{};
  ctx->grid.ncell = (int)cells;
  ctx->ncol_total = cols;
  ctx->shell.assign(1, make_int4(0, 0, 0, 0));
  ctx->shell_off.assign(1, make_float4(0.f, 0.f, 0.f, 0.f));
  ctx->shell_uploaded = 0;
  ctx->rep_uploaded = 0;
  ctx->have_box = true;
  HIPCHK(hipSetDevice(ctx->prm.device));
  HIPCHK(ctx->d_rep_of.ensure(b->rep_of.size() * sizeof(int)));
  HIPCHK(hipMemcpy(ctx->d_rep_of.p, b->rep_of.data(),
                   b->rep_of.size() * sizeof(int), hipMemcpyHostToDevice));
  rc = ensure_eval_buffers(ctx);
  if (rc) return rc;
  b->dirty = false;
  return PMNB_OK;
}

}  // namespace

int pmnb_batch_size(const pmnb_batch *b, int32_t *n_replicas,
                    int64_t *n_atoms) {
  if (!b) return PMNB_ERR_ARG;
  if (n_replicas) *n_replicas = (int32_t)b->boxes.size();
  if (n_atoms) *n_atoms = (int64_t)b->type.size();
  return PMNB_OK;
}

int pmnb_batch_evaluate_device(pmnb_batch *b, const double *d_positions,
                               double *d_forces, double *d_out,
                               uint32_t flags, void *stream) {
  if (!b) return PMNB_ERR_ARG;
  pmnb_ctx *ctx = b->ctx;
  if (flags & PMNB_FLAG_MOLVIRIAL)
    return fail(ctx, PMNB_ERR_UNSUPPORTED,
                "the batch computes no molecular virial");
  int rc = batch_finalize(b);
  if (rc) return rc;
  return pmnb_evaluate_device(ctx, d_positions, d_forces, d_out, nullptr,
                              flags, stream);
}

int pmnb_batch_run_graph(pmnb_batch *b, const double *d_positions,
                         double *d_forces, double *d_out, uint32_t flags,
                         int32_t n_steps, void *stream) {
  if (!b || n_steps < 0) return PMNB_ERR_ARG;
  pmnb_ctx *ctx = b->ctx;
  if (!d_positions || !d_forces || !d_out)
    return fail(ctx, PMNB_ERR_ARG, "null device pointer");
  if (flags & PMNB_FLAG_MOLVIRIAL)
    return fail(ctx, PMNB_ERR_UNSUPPORTED,
                "the batch computes no molecular virial");
  int rc = batch_finalize(b);
  if (rc) return rc;
  HIPCHK(hipSetDevice(ctx->prm.device));
  hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
  const bool same = b->gexec && b->gkey[0] == d_positions &&
                    b->gkey[1] == d_forces && b->gkey[2] == d_out &&
                    b->gflags == flags && b->gshape == ctx->shape &&
                    b->gcap == ctx->grid_cap;
  if (!same) {
    drop_graph(b);
    if (ctx->timing)
      return fail(ctx, PMNB_ERR_STATE, "kernel timing is on: no capture");
    // one step: clear forces and outputs, then the whole evaluation; the
    // buffers are allocated by batch_finalize, so nothing allocates here
    const long long n3 = 3ll * ctx->n, no = 20ll * (long long)ctx->reps.size();
    if (!b->cap_stream)
      HIPCHK(hipStreamCreateWithFlags(&b->cap_stream, hipStreamNonBlocking));
    hipStream_t cs = b->cap_stream;
    // a replay must not rely on what ran before it: the captured
    // evaluation clears the cell counts itself
    ctx->counts_zero = false;
    HIPCHK(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
    hipLaunchKernelGGL(k_zero_f64, dim3((unsigned)((n3 + kBlock - 1) / kBlock)),
                       dim3(kBlock), 0, cs, d_forces, n3);
    hipLaunchKernelGGL(k_zero_f64, dim3((unsigned)((no + kBlock - 1) / kBlock)),
                       dim3(kBlock), 0, cs, d_out, no);
    rc = run_device(ctx, d_positions, d_forces, d_out, nullptr, flags, cs);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(cs, &g);
    ctx->counts_zero = false;  // the captured work has not run
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    HIPCHK(e);
    b->graph = g;
    HIPCHK(hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0));
    b->gkey[0] = d_positions;
    b->gkey[1] = d_forces;
    b->gkey[2] = d_out;
    b->gflags = flags;
    b->gshape = ctx->shape;
    b->gcap = ctx->grid_cap;
  }
  for (int k = 0; k < n_steps; k++) HIPCHK(hipGraphLaunch(b->gexec, s));
  return PMNB_OK;
}

int pmnb_batch_set_shape(pmnb_batch *b, int32_t shape, int32_t grid_blocks) {
  if (!b) return PMNB_ERR_ARG;
  const int nshape = b->ctx->prm.precision == PMNB_PREC_MIXED ? kMixedShapes
                                                              : kF64Shapes;
  if (shape < 0 || shape >= nshape || grid_blocks < 0)
    return fail(b->ctx, PMNB_ERR_ARG, "shape %d / grid %d out of range", shape,
                grid_blocks);
  b->ctx->shape = shape;
  b->ctx->grid_cap = grid_blocks;
  return PMNB_OK;
}

int pmnb_batch_autotune(pmnb_batch *b, const double *d_positions,
                        double *d_forces, double *d_out, uint32_t flags,
                        int32_t *shape, int32_t *grid_blocks, double *ms) {
  if (!b) return PMNB_ERR_ARG;
  pmnb_ctx *ctx = b->ctx;
  int rc = batch_finalize(b);
  if (rc) return rc;
  HIPCHK(hipSetDevice(ctx->prm.device));
  int ncu = 0;
  HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount,
                               ctx->prm.device));
  const bool mixed = ctx->prm.precision == PMNB_PREC_MIXED;
  const int nshape = mixed ? kMixedShapes : kF64Shapes;
  const int caps[4] = {0, 2 * ncu, 4 * ncu, 8 * ncu};
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  double best = 1e300;
  int bs = 0, bc = 0;
  const bool timing = ctx->timing;
  ctx->timing = false;
  for (int sh = 0; sh < nshape && rc == PMNB_OK; sh++) {
    for (int ci = 0; ci < 4 && rc == PMNB_OK; ci++) {
      ctx->shape = sh;
      ctx->grid_cap = caps[ci];
      double t = 1e300;
      for (int rep = 0; rep < 4 && rc == PMNB_OK; rep++) {
        if (hipEventRecord(e0, ctx->stream) != hipSuccess) rc = PMNB_ERR_HIP;
        if (rc == PMNB_OK)
          rc = run_device(ctx, d_positions, d_forces, d_out, nullptr, flags,
                          ctx->stream);
        if (rc == PMNB_OK && (hipEventRecord(e1, ctx->stream) != hipSuccess ||
                              hipEventSynchronize(e1) != hipSuccess))
          rc = fail(ctx, PMNB_ERR_HIP, "autotune timing failed");
        float f = 0.f;
        if (rc == PMNB_OK && hipEventElapsedTime(&f, e0, e1) != hipSuccess)
          rc = fail(ctx, PMNB_ERR_HIP, "autotune timing failed");
        if (rep > 0) t = std::min(t, (double)f);  // first run warms up
      }
      if (t < best) {
        best = t;
        bs = sh;
        bc = caps[ci];
      }
    }
  }
  ctx->timing = timing;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  ctx->shape = bs;
  ctx->grid_cap = bc;
  if (shape) *shape = bs;
  if (grid_blocks) *grid_blocks = bc;
  if (ms) *ms = best;
  return PMNB_OK;
}

int pmnb_batch_force_pairs(pmnb_batch *b, const double *d_positions,
                           int32_t *pairs, int64_t capacity,
                           int64_t *n_pairs) {
  if (!b) return PMNB_ERR_ARG;
  pmnb_ctx *ctx = b->ctx;
  if (!n_pairs || !d_positions || capacity < 0)
    return fail(ctx, PMNB_ERR_ARG, "bad force-pairs arguments");
  int rc = batch_finalize(b);
  if (rc) return rc;
  HIPCHK(hipSetDevice(ctx->prm.device));
  return record_force_pairs(ctx, d_positions, pairs, capacity, n_pairs);
}

int pmnb_batch_pair_count(pmnb_batch *b, int64_t *n_pairs) {
  if (!b) return PMNB_ERR_ARG;
  return pmnb_last_pair_count(b->ctx, n_pairs);
}

const char *pmnb_batch_last_error(const pmnb_batch *b) {
  return b ? b->ctx->err.c_str() : "null batch";
}

void pmnb_batch_destroy(pmnb_batch *b) {
  if (!b) return;
  (void)hipSetDevice(b->ctx->prm.device);
  (void)hipStreamSynchronize(b->ctx->stream);
  drop_graph(b);
  if (b->cap_stream) (void)hipStreamDestroy(b->cap_stream);
  pmnb_destroy(b->ctx);
  delete b;
}

}  // extern "C"

// ------------------------------------------------------------------ leapfrog
struct pmnb_leapfrog {
  pmnb_ctx *ctx = nullptr;
  int device = 0;
  pmnb_lf_params p{};
  bool shake = false, rattle = false;  // attached and effective
  int n = 0, ng = 0, nf = 0, ngb = 0, nfb = 0, n_cons = 0, shape = 0;
  DevBuf mass, rmass, g_atoms, g_cons, g_rest, g_rinv, free_atoms, last, xs,
      vs, xk, vk, vr, partial_s, partial_r, iters, nonconv, hint;
  bool last_ready = false;
  bool zero_counts = false;  // the next step kernel clears the cell counts
  long long calls = 0, iters_cap = 0;
  long long scalls[3] = {0, 0, 0};  // calls per stage
  // bonded terms evaluated with the force (pmnb_lf_set_bonded)
  int n_bonds = 0, n_angles = 0;
  DevBuf bptr, bent, aptr, aent, bonds, bk, br0, angles, ak, at0, ubk, ubr0;
  // k-space terms evaluated with the force (pmnb_lf_set_kspace)
  pmnb_pme *pme = nullptr;
  pmnb_ewald *ewald = nullptr;
  // replicas (one for a single system): constraint-group segments, the
  // replica of each group block, energy-block doubles cleared per force
  int n_reps = 1, n_out = 20;
  bool batch = false;
  DevBuf reps, blk_rep;
  // captured run (pmnb_lf_run_graph) and what it was captured for
  hipStream_t cap_stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  const void *gkey[4] = {nullptr, nullptr, nullptr, nullptr};
  uint32_t gflags = 0;
  int gsteps = 0, gshape = 0, gcap = 0;
  unsigned long long gtopo = 0;
  double gbox[12] = {0};
};

namespace {

constexpr double kTimeFactor = 48.88821290839616;      // base/PMConstants.cpp:50
constexpr double kInvTimeFactor = 0.02045482828087295; // base/PMConstants.cpp:51

LfDev lf_dev(pmnb_leapfrog *lf, double *x, double *v, double *f,
             double *out) {
  LfDev d{};
  d.x = x;
  d.v = v;
  d.f = f;
  d.out = out;
  d.last = lf->last.as<double>();
  d.mass = lf->mass.as<double>();
  d.rmass = lf->rmass.as<double>();
  d.g_atoms = lf->g_atoms.as<int>();
  d.g_cons = lf->g_cons.as<int>();
  d.g_rest = lf->g_rest.as<double>();
  d.g_rinv = lf->g_rinv.as<double>();
  d.free_atoms = lf->free_atoms.as<int>();
  d.n_groups = lf->ng;
  d.n_free = lf->nf;
  d.ngb = lf->ngb;
  d.n_cons = lf->n_cons;
  d.n_out = lf->n_out + ((lf->n_bonds + lf->n_angles) > 0 ? 2 : 0);
  d.reps = lf->reps.as<LfRep>();
  d.blk_rep = lf->blk_rep.as<int>();
  d.n_reps = lf->n_reps;
  // LeapfrogIntegrator / StandardIntegrator use timestep * INV_TIMEFACTOR,
  // the modifiers timestep / TIMEFACTOR.
  d.h = lf->p.timestep * kInvTimeFactor;
  d.h2 = 0.5 * lf->p.timestep * kInvTimeFactor;
  d.dtc = lf->p.timestep / kTimeFactor;
  d.eps[kStShake] = lf->p.shake_epsilon;
  d.eps[kStRattle] = lf->p.rattle_epsilon;
  d.maxit[kStShake] = lf->p.shake_max_iter;
  d.maxit[kStRattle] = lf->p.rattle_max_iter;
  d.partial[kStShake] = lf->partial_s.as<double>();
  d.partial[kStRattle] = lf->partial_r.as<double>();
  d.hint = lf->hint.as<int>();
  d.xs = lf->xs.as<double>();
  d.vs = lf->vs.as<double>();
  d.xk = lf->xk.as<double>();
  d.vk = lf->vk.as<double>();
  d.vr = lf->vr.as<double>();
  d.iters = lf->iters_cap > 0 ? lf->iters.as<int>() : nullptr;
  d.iters_cap = lf->iters_cap;
  d.nonconv = lf->nonconv.as<int>();
  return d;
}

struct LfKernel {
  int apply = kStNone, pre = kPreNone, probe = kStNone;
  bool zero = false;
  bool empty() const {
    return apply == kStNone && pre == kPreNone && probe == kStNone && !zero;
  }
};

#define PMNB_LF_CASE(A, P, Q, Z)                                            \
  if (k.apply == A && k.pre == P && k.probe == Q && k.zero == Z) {          \
    hipLaunchKernelGGL((k_lf_step<A, P, Q, Z, SHAPE>), dim3(grid),           \
                       dim3(kLfBlock), 0, s, d, sc_apply, sc_probe, call);   \
    return hipGetLastError();                                                \
  }

template <int SHAPE>
hipError_t lf_launch_step(const LfKernel &k, const LfDev &d, int grid,
                          long long sc_apply, long long sc_probe,
                          long long call, hipStream_t s) {
  PMNB_LF_CASE(kStNone, kPreFirst, kStShake, false)
  PMNB_LF_CASE(kStShake, kPreNone, kStNone, true)
  PMNB_LF_CASE(kStNone, kPreFirst, kStNone, true)
  PMNB_LF_CASE(kStNone, kPreHalf, kStRattle, false)
  PMNB_LF_CASE(kStRattle, kPreHalfDrift, kStShake, false)
  PMNB_LF_CASE(kStRattle, kPreHalfDrift, kStNone, true)
  PMNB_LF_CASE(kStNone, kPreHalf, kStNone, false)
  PMNB_LF_CASE(kStNone, kPreHalfDrift, kStNone, true)
  PMNB_LF_CASE(kStNone, kPreFullDrift, kStShake, false)
  PMNB_LF_CASE(kStNone, kPreFullDrift, kStNone, true)
  PMNB_LF_CASE(kStRattle, kPreNone, kStNone, false)
  return hipErrorInvalidValue;  // not a sequence run() produces
}
#undef PMNB_LF_CASE

// Launches one fused kernel.
int lf_flush(pmnb_leapfrog *lf, LfKernel &k, const LfDev &d0, hipStream_t s) {
  pmnb_ctx *ctx = lf->ctx;
  if (k.empty()) return PMNB_OK;
  LfDev d = d0;
  const bool zc = lf->zero_counts;
  if (zc) {
    // the evaluation before launched no finish kernel: clear its counts
    d.z_count = ctx->d_count.as<int>();
    d.z_nc = ctx->grid.ncell;
  }
  const long long call = k.apply != kStNone ? lf->calls++ : -1;
  const long long sc_apply = lf->scalls[k.apply];
  const long long sc_probe = lf->scalls[k.probe];
  const int grid = lf->ngb + lf->nfb;
  if (grid > 0) {
    hipError_t e =
        lf->shape == kShapeWater
            ? lf_launch_step<kShapeWater>(k, d, grid, sc_apply, sc_probe, call, s)
        : lf->shape == kShape33
            ? lf_launch_step<kShape33>(k, d, grid, sc_apply, sc_probe, call, s)
            : lf_launch_step<kShape46>(k, d, grid, sc_apply, sc_probe, call, s);
    if (e != hipSuccess)
      return fail(ctx, PMNB_ERR_HIP, "leapfrog step launch failed: %s",
                  hipGetErrorString(e));
  }
  if (k.apply != kStNone) lf->scalls[k.apply]++;
  if (zc) {
    lf->zero_counts = false;
    ctx->counts_zero = true;
  }
  k = LfKernel{};
  return PMNB_OK;
}

LfBonded lf_bonded(pmnb_leapfrog *lf) {
  LfBonded B{};
  B.bptr = lf->bptr.as<int>();
  B.bent = lf->bent.as<int>();
  B.aptr = lf->aptr.as<int>();
  B.aent = lf->aent.as<int>();
  B.bonds = lf->bonds.as<int>();
  B.bk = lf->bk.as<double>();
  B.br0 = lf->br0.as<double>();
  B.angles = lf->angles.as<int>();
  B.ak = lf->ak.as<double>();
  B.at0 = lf->at0.as<double>();
  B.ubk = lf->ubk.as<double>();
  B.ubr0 = lf->ubr0.as<double>();
  B.n_atoms = lf->n;
  B.n_bonds = lf->n_bonds;
  B.n_angles = lf->n_angles;
  return B;
}

// calculateForces of the integrator's force group after the forces and the
// energy block were cleared: the cutoff force, then the bonded terms (their
// energies into out[20], out[21] when `energies`).  Only the last
// evaluation of a run leaves its energies in d_out (the next step clears
// them), so the others (energies false) skip the cutoff force's virial and,
// on a single system, its energy reduction; the cell counts are then
// cleared by the integrator's next kernel (lf->zero_counts).
int lf_calc_forces(pmnb_leapfrog *lf, const double *d_pos, double *d_force,
                   double *d_out, uint32_t flags, bool energies,
                   hipStream_t s, bool assigned = false) {
  pmnb_ctx *ctx = lf->ctx;
  bool skipped = false;
  const uint32_t fl =
      energies ? flags : flags & ~(PMNB_FLAG_VIRIAL | PMNB_FLAG_MOLVIRIAL);
  int rc = run_device(ctx, d_pos, d_force, d_out, nullptr, fl, s, false,
                      assigned, energies ? nullptr : &skipped);
  lf->zero_counts = skipped;
  if (rc) return rc;
  if (lf->pme) {
    rc = pmnb_pme_evaluate_device(lf->pme, d_pos, d_force, d_out,
                                  flags & PMNB_FLAG_VIRIAL, s);
    if (rc) return rc;
  }
  if (lf->ewald) {
    rc = pmnb_ewald_evaluate_device(lf->ewald, d_pos, d_force, d_out,
                                    flags & PMNB_FLAG_VIRIAL, s);
    if (rc) return rc;
  }
  if (lf->n_bonds + lf->n_angles == 0) return PMNB_OK;
  const LfBonded B = lf_bonded(lf);
  hipLaunchKernelGGL(k_lf_bonded, dim3(blocks_for(lf->n, kLfBlock)),
                     dim3(kLfBlock), 0, s, ctx->box, d_pos, d_force, B);
  HIPCHK(hipGetLastError());
  if (energies) {
    hipLaunchKernelGGL(k_lf_bonded_energy, dim3(1), dim3(kLfBlock), 0, s,
                       ctx->box, d_pos, B, d_out + 20);
    HIPCHK(hipGetLastError());
  }
  return PMNB_OK;
}

// The integrator over ctx's atoms, split into replicas at the atom indices
// rep_first (R + 1 ascending entries, rep_first[0] = 0, rep_first[R] =
// n_atoms); a constraint must join two atoms of one replica.
int lf_create_impl(pmnb_ctx *ctx, const pmnb_lf_params *params,
                   int32_t n_atoms, const double *mass, int32_t n_constraints,
                   const int32_t *constraint_atoms, const double *rest_length,
                   const std::vector<int> &rep_first, bool batch,
                   pmnb_leapfrog **out) {
  if (!ctx || !params || !out) return PMNB_ERR_ARG;
  *out = nullptr;
  int rc = check_ready(ctx);
  if (rc) return rc;
  if (n_atoms != ctx->n)
    return fail(ctx, PMNB_ERR_ARG, "n_atoms %d differs from the topology's %d",
                n_atoms, ctx->n);
  if (n_atoms > 0 && !mass) return fail(ctx, PMNB_ERR_ARG, "null mass");
  if (n_constraints < 0 ||
      (n_constraints > 0 && (!constraint_atoms || !rest_length)))
    return fail(ctx, PMNB_ERR_ARG, "bad constraint arrays");
  const pmnb_lf_params &p = *params;
  if (!(p.timestep > 0))
    return fail(ctx, PMNB_ERR_ARG, "timestep must be positive");
  for (int i = 0; i < n_atoms; i++)
    if (!(mass[i] > 0))
      return fail(ctx, PMNB_ERR_ARG, "mass of atom %d is not positive", i);
  // union-find over constraint atoms
  std::vector<int> parent(n_atoms);
  for (int i = 0; i < n_atoms; i++) parent[i] = i;
  auto root = [&](int a) {
    while (parent[a] != a) a = parent[a] = parent[parent[a]];
    return a;
  };
  for (int c = 0; c < n_constraints; c++) {
    const int a = constraint_atoms[2 * c], b = constraint_atoms[2 * c + 1];
    if (a < 0 || b < 0 || a >= n_atoms || b >= n_atoms || a == b)
      return fail(ctx, PMNB_ERR_ARG, "constraint %d has atoms (%d, %d)", c, a, b);
    if (!(rest_length[c] > 0))
      return fail(ctx, PMNB_ERR_ARG, "constraint %d has rest length <= 0", c);
    parent[root(a)] = root(b);
  }
  const int R = (int)rep_first.size() - 1;
  auto rep_of = [&](int a) {
    return (int)(std::upper_bound(rep_first.begin(), rep_first.end(), a) -
                 rep_first.begin()) - 1;
  };
  std::vector<int> rep_cons(std::max(R, 1), 0);
  for (int c = 0; c < n_constraints; c++) {
    const int ra = rep_of(constraint_atoms[2 * c]);
    if (ra != rep_of(constraint_atoms[2 * c + 1]))
      return fail(ctx, PMNB_ERR_ARG, "constraint %d joins two replicas", c);
    rep_cons[ra]++;
  }
  std::vector<char> constrained(n_atoms, 0);
  for (int c = 0; c < 2 * n_constraints; c++) constrained[constraint_atoms[c]] = 1;
  // groups in order of their lowest atom; atoms ascending within a group;
  // constraints in list order
  std::vector<int> gid(n_atoms, -1), first_of_root(n_atoms, -1);
  std::vector<std::vector<int>> gat;
  for (int i = 0; i < n_atoms; i++) {
    if (!constrained[i]) continue;
    const int r = root(i);
    if (first_of_root[r] < 0) {
      first_of_root[r] = (int)gat.size();
      gat.emplace_back();
    }
    gid[i] = first_of_root[r];
    gat[gid[i]].push_back(i);
  }
  const int ng = (int)gat.size();
  std::vector<std::vector<int>> gcon(ng);
  for (int c = 0; c < n_constraints; c++)
    gcon[gid[constraint_atoms[2 * c]]].push_back(c);
  int max_a = 0, max_c = 0;
  for (int g = 0; g < ng; g++) {
    max_a = std::max(max_a, (int)gat[g].size());
    max_c = std::max(max_c, (int)gcon[g].size());
  }
  if (max_a > kLfMaxAtoms || max_c > kLfMaxCons)
    return fail(ctx, PMNB_ERR_UNSUPPORTED,
                "a connected set of constraints spans %d atoms and %d "
                "constraints (at most %d and %d supported)",
                max_a, max_c, kLfMaxAtoms, kLfMaxCons);
  int shape = (max_a <= 3 && max_c <= 3) ? kShape33 : kShape46;
  const int NA = shape == kShape33 ? 3 : 4, NC = shape == kShape33 ? 3 : 6;
  std::vector<int> h_atoms((size_t)std::max(ng, 1) * NA, -1);
  std::vector<int> h_cons((size_t)std::max(ng, 1) * NC, -1);
  std::vector<double> h_rest((size_t)std::max(ng, 1) * NC, 1.0);
  std::vector<double> h_rinv((size_t)std::max(ng, 1) * NC, 0.0);
  for (int g = 0; g < ng; g++) {
    for (size_t k = 0; k < gat[g].size(); k++) h_atoms[(size_t)g * NA + k] = gat[g][k];
    for (size_t k = 0; k < gcon[g].size(); k++) {
      const int c = gcon[g][k];
      auto local = [&](int atom) {
        return (int)(std::find(gat[g].begin(), gat[g].end(), atom) - gat[g].begin());
      };
      h_cons[(size_t)g * NC + k] =
          local(constraint_atoms[2 * c]) | (local(constraint_atoms[2 * c + 1]) << 4);
      const double r = rest_length[c];
      h_rest[(size_t)g * NC + k] = r;
      h_rinv[(size_t)g * NC + k] = 1.0 / (2.0 * (r * r));
    }
  }
  // every group a water in the reference's constraint order (atoms O < H1 <
  // H2, constraints (H1,H2), (H1,O), (H2,O)): the compile-time shape
  if (shape == kShape33 && ng > 0) {
    bool water = true;
    const int want[3] = {1 | (2 << 4), 1 | (0 << 4), 2 | (0 << 4)};
    for (int g = 0; g < ng && water; g++) {
      water = gat[g].size() == 3 && gcon[g].size() == 3;
      for (int c = 0; c < 3 && water; c++) water = h_cons[(size_t)g * 3 + c] == want[c];
    }
    if (water) shape = kShapeWater;
  }
  std::vector<int> h_free;
  for (int i = 0; i < n_atoms; i++)
    if (!constrained[i]) h_free.push_back(i);
  std::vector<double> h_rmass(n_atoms);
  for (int i = 0; i < n_atoms; i++) h_rmass[i] = 1 / mass[i];
  // replica segments: groups are ordered by their lowest atom and replicas
  // are contiguous atom ranges, so each replica's groups are contiguous
  std::vector<LfRep> h_reps(std::max(R, 1), LfRep{0, 0, 0, 0, 0});
  {
    int g = 0, blk = 0;
    for (int r = 0; r < R; r++) {
      LfRep &q = h_reps[r];
      q.goff = g;
      while (g < ng && rep_of(gat[g][0]) == r) g++;
      q.ng = g - q.goff;
      q.gb0 = blk;
      q.gnb = blocks_for(q.ng, kLfBlock);
      q.ncons = rep_cons[r];
      blk += q.gnb;
    }
  }
  std::vector<int> h_blk_rep;
  for (int r = 0; r < R; r++)
    for (int k = 0; k < h_reps[r].gnb; k++) h_blk_rep.push_back(r);

  pmnb_leapfrog *lf = new pmnb_leapfrog;
  lf->ctx = ctx;
  lf->device = ctx->prm.device;
  lf->p = p;
  // module/ModifierModule.cpp:110-125: a modifier is attached only with a
  // positive epsilon and max_iter
  lf->shake = p.shake && p.shake_epsilon > 0.0 && p.shake_max_iter > 0;
  lf->rattle = p.rattle && p.rattle_epsilon > 0.0 && p.rattle_max_iter > 0;
  lf->n = n_atoms;
  lf->ng = ng;
  lf->nf = (int)h_free.size();
  lf->ngb = (int)h_blk_rep.size();
  lf->nfb = blocks_for(lf->nf, kLfBlock);
  lf->n_reps = std::max(R, 1);
  lf->n_out = batch ? 20 * R : 20;
  lf->batch = batch;
  lf->n_cons = n_constraints;
  lf->shape = shape;
  auto up = [&](DevBuf &b, const void *src, size_t bytes) -> hipError_t {
    hipError_t e = b.ensure(bytes);
    if (e == hipSuccess && bytes && src)
      e = hipMemcpy(b.p, src, bytes, hipMemcpyHostToDevice);
    return e;
  };
  const size_t N = (size_t)std::max(n_atoms, 1);
  const size_t ngb = (size_t)std::max(lf->ngb, 1);
  hipError_t e = hipSetDevice(ctx->prm.device);
  if (e == hipSuccess) e = up(lf->mass, mass, n_atoms * sizeof(double));
  if (e == hipSuccess) e = up(lf->rmass, h_rmass.data(), n_atoms * sizeof(double));
  if (e == hipSuccess) e = up(lf->g_atoms, h_atoms.data(), h_atoms.size() * sizeof(int));
  if (e == hipSuccess) e = up(lf->g_cons, h_cons.data(), h_cons.size() * sizeof(int));
  if (e == hipSuccess) e = up(lf->g_rest, h_rest.data(), h_rest.size() * sizeof(double));
  if (e == hipSuccess) e = up(lf->g_rinv, h_rinv.data(), h_rinv.size() * sizeof(double));
  if (e == hipSuccess)
    e = up(lf->free_atoms, h_free.data(), h_free.size() * sizeof(int));
  if (e == hipSuccess) e = lf->last.ensure(N * 3 * sizeof(double));
  if (e == hipSuccess) e = lf->xs.ensure(N * 3 * sizeof(double));
  if (e == hipSuccess) e = lf->vs.ensure(N * 3 * sizeof(double));
  if (e == hipSuccess) e = lf->xk.ensure(N * 3 * sizeof(double));
  if (e == hipSuccess) e = lf->vk.ensure(N * 3 * sizeof(double));
  if (e == hipSuccess) e = lf->vr.ensure(N * 3 * sizeof(double));
  if (e == hipSuccess)
    e = lf->partial_s.ensure((size_t)(std::max(p.shake_max_iter, 0) + 1) * ngb *
                             sizeof(double));
  if (e == hipSuccess)
    e = lf->partial_r.ensure((size_t)(std::max(p.rattle_max_iter, 0) + 1) * ngb *
                             sizeof(double));
  if (e == hipSuccess) e = lf->nonconv.ensure(4 * sizeof(int));
  if (e == hipSuccess) e = hipMemset(lf->nonconv.p, 0, 4 * sizeof(int));
  if (e == hipSuccess)
    e = up(lf->reps, h_reps.data(), h_reps.size() * sizeof(LfRep));
  if (e == hipSuccess)
    e = up(lf->blk_rep, h_blk_rep.data(), h_blk_rep.size() * sizeof(int));
  const size_t hint_bytes = (size_t)8 * lf->n_reps * sizeof(int);
  if (e == hipSuccess) e = lf->hint.ensure(hint_bytes);
  if (e == hipSuccess) e = hipMemset(lf->hint.p, 0xff, hint_bytes);  // -1
  if (e != hipSuccess) {
    pmnb_lf_destroy(lf);
    return fail(ctx, PMNB_ERR_HIP, "leapfrog set-up failed: %s",
                hipGetErrorString(e));
  }
  *out = lf;
  return PMNB_OK;
}

void lf_drop_graph(pmnb_leapfrog *lf) {
  if (lf->gexec) (void)hipGraphExecDestroy(lf->gexec);
  if (lf->graph) (void)hipGraphDestroy(lf->graph);
  lf->gexec = nullptr;
  lf->graph = nullptr;
}

}  // namespace

extern "C" {

int pmnb_lf_create(pmnb_ctx *ctx, const pmnb_lf_params *params,
                   int32_t n_atoms, const double *mass, int32_t n_constraints,
                   const int32_t *constraint_atoms, const double *rest_length,
                   pmnb_leapfrog **out) {
  if (ctx && ctx->batch)
    return fail(ctx, PMNB_ERR_ARG, "a batch's integrator is made by "
                                   "pmnb_batch_lf_create");
  return lf_create_impl(ctx, params, n_atoms, mass, n_constraints,
                        constraint_atoms, rest_length,
                        std::vector<int>{0, n_atoms}, false, out);
}

int pmnb_lf_run_device(pmnb_leapfrog *lf, double *d_positions,
                       double *d_velocities, double *d_forces, double *d_out,
                       int32_t n_steps, uint32_t flags, void *stream) {
  if (!lf) return PMNB_ERR_ARG;
  pmnb_ctx *ctx = lf->ctx;
  int rc = check_ready(ctx);
  if (rc) return rc;
  if (ctx->n != lf->n)
    return fail(ctx, PMNB_ERR_STATE, "topology changed since pmnb_lf_create");
  if (flags & PMNB_FLAG_MOLVIRIAL)
    return fail(ctx, PMNB_ERR_UNSUPPORTED,
                "the molecular virial needs molecule centres per step");
  if (n_steps < 1) return PMNB_OK;  // LeapfrogIntegrator::run returns
  if (lf->n > 0 && (!d_positions || !d_velocities || !d_forces || !d_out))
    return fail(ctx, PMNB_ERR_ARG, "null device pointer");
  HIPCHK(hipSetDevice(ctx->prm.device));
  hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
  if (!lf->last_ready) {
    if (lf->n > 0)
      HIPCHK(hipMemcpyAsync(lf->last.p, d_positions, (size_t)lf->n * 3 * sizeof(double),
                            hipMemcpyDeviceToDevice, s));
    lf->last_ready = true;
  }
  if (lf->n == 0) return PMNB_OK;
  const LfDev d = lf_dev(lf, d_positions, d_velocities, d_forces, d_out);
  const bool S = lf->shake && lf->n_cons > 0;
  const bool R = lf->rattle && lf->n_cons > 0;
  LfKernel k;
  auto pre = [&](int op) -> int {
    if (k.pre != kPreNone || k.probe != kStNone || k.zero) {
      int r = lf_flush(lf, k, d, s);
      if (r) return r;
    }
    k.pre = op;
    return PMNB_OK;
  };
  auto stage = [&](int st) -> int {
    k.probe = st;
    int r = lf_flush(lf, k, d, s);
    k.apply = st;
    return r;
  };
  // the kernel that clears the forces also assigns the cells of the
  // evaluation that follows (single system on the tile path; on zeroed
  // counts)
  const bool assign = ctx->use_tile && !lf->batch;
  LfDev da = d;
  if (assign) {
    rc = ensure_eval_buffers(ctx);
    if (rc) return rc;
    const EvalBufs B = ctx_bufs(ctx);
    da.c_reps = B.reps;
    da.c_rep_of = ctx->batch ? ctx->d_rep_of.as<int>() : nullptr;
    da.c_cell_of = B.cell_of;
    da.c_count = B.count;
    da.c_bucket = B.bucket;
    da.c_cap = B.cap;
  }
  auto force = [&](bool last) -> int {
    k.zero = true;
    if (assign && !ctx->counts_zero) {
      // (also what a clearing left to this kernel would have done: the
      // assignment's atomics must not race a clearing in the same kernel)
      hipLaunchKernelGGL(k_clear, dim3(blocks_for(ctx->grid.ncell)),
                         dim3(kBlock), 0, s, da.c_count, ctx->grid.ncell,
                         ctx->d_npairs.as<unsigned long long>());
      lf->zero_counts = false;
    }
    int r = lf_flush(lf, k, assign ? da : d, s);
    if (r) return r;
    return lf_calc_forces(lf, d_positions, d_forces, d_out, flags, last, s,
                          assign);
  };
  // doHalfKickdoDrift (LeapfrogIntegrator.cpp:33-58), SHAKE post-drift
  if ((rc = pre(kPreFirst))) return rc;
  if (S && (rc = stage(kStShake))) return rc;
  if ((rc = force(n_steps == 1))) return rc;
  for (int i = 1; i < n_steps; i++) {
    // doKickdoDrift (LeapfrogIntegrator.cpp:60-83)
    if (lf->rattle) {
      if ((rc = pre(kPreHalf))) return rc;
      if (R && (rc = stage(kStRattle))) return rc;
      if ((rc = pre(kPreHalfDrift))) return rc;
    } else {
      if ((rc = pre(kPreFullDrift))) return rc;
    }
    if (S && (rc = stage(kStShake))) return rc;
    if ((rc = force(i == n_steps - 1))) return rc;
  }
  // doHalfKick and postStepModify (LeapfrogIntegrator.cpp:96-97)
  if ((rc = pre(kPreHalf))) return rc;
  if (R && (rc = stage(kStRattle))) return rc;
  return lf_flush(lf, k, d, s);
}

int pmnb_lf_iterations(pmnb_leapfrog *lf, int64_t capacity, int32_t *iters,
                       int64_t capacity_out, int64_t *n_calls,
                       int64_t *n_unconverged) {
  if (!lf) return PMNB_ERR_ARG;
  pmnb_ctx *ctx = lf->ctx;
  HIPCHK(hipSetDevice(ctx->prm.device));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  HIPCHK(hipDeviceSynchronize());
  if (n_calls) *n_calls = lf->calls;
  const long long have = std::min<long long>(lf->calls * lf->n_reps,
                                             lf->iters_cap);
  const long long n_out = std::min<long long>(have, std::max<int64_t>(capacity_out, 0));
  if (iters && n_out > 0)
    HIPCHK(hipMemcpy(iters, lf->iters.p, n_out * sizeof(int), hipMemcpyDeviceToHost));
  int nc[4] = {0, 0, 0, 0};
  HIPCHK(hipMemcpy(nc, lf->nonconv.p, sizeof nc, hipMemcpyDeviceToHost));
  if (n_unconverged) *n_unconverged = (int64_t)nc[kStShake] + nc[kStRattle];
  // a captured run holds the record's address and capacity
  lf_drop_graph(lf);
  if (capacity > 0) HIPCHK(lf->iters.ensure((size_t)capacity * sizeof(int)));
  lf->iters_cap = capacity > 0 ? capacity : 0;
  lf->calls = 0;
  HIPCHK(hipMemset(lf->nonconv.p, 0, 4 * sizeof(int)));
  return PMNB_OK;
}

void pmnb_lf_destroy(pmnb_leapfrog *lf) {
  if (!lf) return;
  // the context may already be destroyed: only the device is used here
  (void)hipSetDevice(lf->device);
  (void)hipDeviceSynchronize();
  lf_drop_graph(lf);
  if (lf->cap_stream) (void)hipStreamDestroy(lf->cap_stream);
  DevBuf *bufs[] = {&lf->mass, &lf->rmass, &lf->g_atoms, &lf->g_cons,
                    &lf->g_rest, &lf->g_rinv, &lf->free_atoms, &lf->last,
                    &lf->xs, &lf->vs, &lf->xk, &lf->vk, &lf->vr,
                    &lf->partial_s, &lf->partial_r,
                    &lf->iters, &lf->nonconv, &lf->hint, &lf->bptr,
                    &lf->bent, &lf->aptr, &lf->aent, &lf->bonds, &lf->bk,
                    &lf->br0, &lf->angles, &lf->ak, &lf->at0, &lf->ubk,
                    &lf->ubr0, &lf->reps, &lf->blk_rep};
  for (DevBuf *b : bufs) b->release();
  delete lf;
}

int pmnb_lf_set_bonded(pmnb_leapfrog *lf, int32_t n_bonds,
                       const int32_t *bond_atoms, const double *bond_k,
                       const double *bond_r0, int32_t n_angles,
                       const int32_t *angle_atoms, const double *angle_k,
                       const double *angle_theta0, const double *ub_k,
                       const double *ub_r0) {
  if (!lf) return PMNB_ERR_ARG;
  pmnb_ctx *ctx = lf->ctx;
  const int n = lf->n;
  if (lf->batch && (n_bonds > 0 || n_angles > 0))
    return fail(ctx, PMNB_ERR_UNSUPPORTED,
                "a batch's integrator takes no bonded terms");
  lf_drop_graph(lf);
  if (n_bonds < 0 || n_angles < 0 ||
      (n_bonds > 0 && (!bond_atoms || !bond_k || !bond_r0)) ||
      (n_angles > 0 && (!angle_atoms || !angle_k || !angle_theta0)))
    return fail(ctx, PMNB_ERR_ARG, "bad bonded term arrays");
  if (n_bonds > (1 << 29) || n_angles > (1 << 28))
    return fail(ctx, PMNB_ERR_ARG, "too many bonded terms");
  for (int i = 0; i < 2 * n_bonds; i++)
    if (bond_atoms[i] < 0 || bond_atoms[i] >= n)
      return fail(ctx, PMNB_ERR_ARG, "bond atom %d out of range", bond_atoms[i]);
  for (int i = 0; i < 3 * n_angles; i++)
    if (angle_atoms[i] < 0 || angle_atoms[i] >= n)
      return fail(ctx, PMNB_ERR_ARG, "angle atom %d out of range", angle_atoms[i]);
  // per-atom entry lists in term order
  std::vector<int> bcnt(n + 1, 0), acnt(n + 1, 0);
  for (int i = 0; i < 2 * n_bonds; i++) bcnt[bond_atoms[i] + 1]++;
  for (int i = 0; i < 3 * n_angles; i++) acnt[angle_atoms[i] + 1]++;
  for (int i = 0; i < n; i++) {
    bcnt[i + 1] += bcnt[i];
    acnt[i + 1] += acnt[i];
  }
  std::vector<int> bent(std::max(2 * n_bonds, 1)), aent(std::max(3 * n_angles, 1));
  {
    std::vector<int> fill(bcnt.begin(), bcnt.end() - 1);
    for (int i = 0; i < n_bonds; i++)
      for (int r = 0; r < 2; r++) bent[fill[bond_atoms[2 * i + r]]++] = (i << 1) | r;
    std::vector<int> filla(acnt.begin(), acnt.end() - 1);
    for (int i = 0; i < n_angles; i++)
      for (int r = 0; r < 3; r++) aent[filla[angle_atoms[3 * i + r]]++] = (i << 2) | r;
  }
  std::vector<double> zeros(std::max(n_angles, 1), 0.0);
  auto up = [&](DevBuf &b, const void *src, size_t bytes) -> hipError_t {
    hipError_t e = b.ensure(std::max<size_t>(bytes, 8));
    if (e == hipSuccess && bytes && src)
      e = hipMemcpy(b.p, src, bytes, hipMemcpyHostToDevice);
    return e;
  };
  HIPCHK(hipSetDevice(ctx->prm.device));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  HIPCHK(up(lf->bptr, bcnt.data(), (n + 1) * sizeof(int)));
  HIPCHK(up(lf->aptr, acnt.data(), (n + 1) * sizeof(int)));
  HIPCHK(up(lf->bent, bent.data(), 2 * (size_t)n_bonds * sizeof(int)));
  HIPCHK(up(lf->aent, aent.data(), 3 * (size_t)n_angles * sizeof(int)));
  HIPCHK(up(lf->bonds, bond_atoms, 2 * (size_t)n_bonds * sizeof(int)));
  HIPCHK(up(lf->bk, bond_k, (size_t)n_bonds * sizeof(double)));
  HIPCHK(up(lf->br0, bond_r0, (size_t)n_bonds * sizeof(double)));
  HIPCHK(up(lf->angles, angle_atoms, 3 * (size_t)n_angles * sizeof(int)));
  HIPCHK(up(lf->ak, angle_k, (size_t)n_angles * sizeof(double)));
  HIPCHK(up(lf->at0, angle_theta0, (size_t)n_angles * sizeof(double)));
  HIPCHK(up(lf->ubk, ub_k ? ub_k : zeros.data(), (size_t)n_angles * sizeof(double)));
  HIPCHK(up(lf->ubr0, ub_r0 ? ub_r0 : zeros.data(), (size_t)n_angles * sizeof(double)));
  lf->n_bonds = n_bonds;
  lf->n_angles = n_angles;
  return PMNB_OK;
}

int pmnb_lf_forces_device(pmnb_leapfrog *lf, const double *d_positions,
                          double *d_forces, double *d_out, uint32_t flags,
                          void *stream) {
  if (!lf) return PMNB_ERR_ARG;
  pmnb_ctx *ctx = lf->ctx;
  int rc = check_ready(ctx);
  if (rc) return rc;
  if (lf->n > 0 && (!d_positions || !d_forces || !d_out))
    return fail(ctx, PMNB_ERR_ARG, "null device pointer");
  if (flags & PMNB_FLAG_MOLVIRIAL)
    return fail(ctx, PMNB_ERR_UNSUPPORTED,
                "the molecular virial needs molecule centres per step");
  HIPCHK(hipSetDevice(ctx->prm.device));
  hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
  if (lf->n == 0) return PMNB_OK;
  launch_fill_i32(reinterpret_cast<int *>(d_forces), lf->n * 6, 0, s);
  launch_fill_i32(reinterpret_cast<int *>(d_out),
                  2 * (lf->n_out + ((lf->n_bonds + lf->n_angles) > 0 ? 2 : 0)),
                  0, s);
  return lf_calc_forces(lf, d_positions, d_forces, d_out, flags, true, s);
}

int pmnb_lf_run_graph(pmnb_leapfrog *lf, double *d_positions,
                      double *d_velocities, double *d_forces, double *d_out,
                      int32_t n_steps, int32_t n_replays, uint32_t flags,
                      void *stream) {
  if (!lf || n_steps < 0 || n_replays < 0) return PMNB_ERR_ARG;
  pmnb_ctx *ctx = lf->ctx;
  int rc = check_ready(ctx);
  if (rc) return rc;
  if (ctx->n != lf->n)
    return fail(ctx, PMNB_ERR_STATE, "topology changed since the integrator "
                                     "was created");
  if (n_steps == 0 || n_replays == 0 || lf->n == 0) return PMNB_OK;
  if (!d_positions || !d_velocities || !d_forces || !d_out)
    return fail(ctx, PMNB_ERR_ARG, "null device pointer");
  if (flags & PMNB_FLAG_MOLVIRIAL)
    return fail(ctx, PMNB_ERR_UNSUPPORTED,
                "the molecular virial needs molecule centres per step");
  if (lf->pme || lf->ewald)
    return fail(ctx, PMNB_ERR_UNSUPPORTED,
                "k-space terms are not captured; use pmnb_lf_run_device");
  if (ctx->timing)
    return fail(ctx, PMNB_ERR_STATE, "kernel timing is on: no capture");
  HIPCHK(hipSetDevice(ctx->prm.device));
  hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
  const bool same =
      lf->gexec && lf->gkey[0] == d_positions && lf->gkey[1] == d_velocities &&
      lf->gkey[2] == d_forces && lf->gkey[3] == d_out && lf->gflags == flags &&
      lf->gsteps == n_steps && lf->gtopo == ctx->topo_gen &&
      std::equal(lf->gbox, lf->gbox + 12, ctx->box_key) &&
      lf->gshape == ctx->shape && lf->gcap == ctx->grid_cap;
  if (!same) {
    lf_drop_graph(lf);
    // every allocation, upload and first-use clear of the run happens
    // before the capture, and SHAKE's previous positions are set here
    HIPCHK(hipStreamSynchronize(s));
    rc = ensure_eval_buffers(ctx);
    if (rc) return rc;
    if (!lf->last_ready) {
      HIPCHK(hipMemcpyAsync(lf->last.p, d_positions,
                            (size_t)lf->n * 3 * sizeof(double),
                            hipMemcpyDeviceToDevice, s));
      lf->last_ready = true;
    }
    HIPCHK(hipStreamSynchronize(s));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (!lf->cap_stream)
      HIPCHK(hipStreamCreateWithFlags(&lf->cap_stream, hipStreamNonBlocking));
    hipStream_t cs = lf->cap_stream;
    // a replay must not rely on what ran before it: its first evaluation
    // clears the cell counts itself
    ctx->counts_zero = false;
    HIPCHK(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
    rc = pmnb_lf_run_device(lf, d_positions, d_velocities, d_forces, d_out,
                            n_steps, flags, cs);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(cs, &g);
    ctx->counts_zero = false;  // the captured work has not run
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    HIPCHK(e);
    lf->graph = g;
    HIPCHK(hipGraphInstantiate(&lf->gexec, g, nullptr, nullptr, 0));
    lf->gkey[0] = d_positions;
    lf->gkey[1] = d_velocities;
    lf->gkey[2] = d_forces;
    lf->gkey[3] = d_out;
    lf->gflags = flags;
    lf->gsteps = n_steps;
    lf->gtopo = ctx->topo_gen;
    std::copy(ctx->box_key, ctx->box_key + 12, lf->gbox);
    lf->gshape = ctx->shape;
    lf->gcap = ctx->grid_cap;
  }
  for (int k = 0; k < n_replays; k++) HIPCHK(hipGraphLaunch(lf->gexec, s));
  return PMNB_OK;
}

int pmnb_batch_lf_create(pmnb_batch *b, const pmnb_lf_params *params,
                         int32_t n_atoms, const double *mass,
                         int32_t n_constraints,
                         const int32_t *constraint_atoms,
                         const double *rest_length, pmnb_leapfrog **out) {
  if (!b || !out) return PMNB_ERR_ARG;
  *out = nullptr;
  int rc = batch_finalize(b);
  if (rc) return rc;
  std::vector<int> first{0};
  const std::vector<int> &ro = b->rep_of;
  for (size_t i = 1; i < ro.size(); i++)
    if (ro[i] != ro[i - 1]) first.push_back((int)i);
  first.push_back((int)ro.size());
  return lf_create_impl(b->ctx, params, n_atoms, mass, n_constraints,
                        constraint_atoms, rest_length, first, true, out);
}

}  // extern "C"

// ------------------------------------------------------------------ Ewald
namespace {

// The rows ExclusionTable::add keeps in getTable()
// (topology/ExclusionTable.cpp:103-110) from rows given in add order: no
// row of an atom with itself, no EXCLUSION_NONE row, one row per pair with
// the class of its first add, in the order given.  The correction terms of
// the Ewald sums walk these.  Rows are (a1 < a2, class).
void correction_rows(int n_atoms, int n_excl, const int32_t *excl,
                     std::vector<int32_t> &rows) {
  rows.clear();
  rows.reserve(3 * (size_t)n_excl);
  std::vector<long long> keys((size_t)n_excl, -1);
  for (int i = 0; i < n_excl; i++) {
    int a = excl[3 * i], c = excl[3 * i + 1];
    if (a == c || excl[3 * i + 2] == PMNB_EXCL_NONE) continue;
    if (a > c) std::swap(a, c);
    keys[i] = (long long)a * n_atoms + c;
  }
  std::vector<long long> sorted_keys;
  for (long long k : keys)
    if (k >= 0) sorted_keys.push_back(k);
  std::sort(sorted_keys.begin(), sorted_keys.end());
  std::vector<char> taken(sorted_keys.size(), 0);
  for (int i = 0; i < n_excl; i++) {
    if (keys[i] < 0) continue;
    const size_t pos = std::lower_bound(sorted_keys.begin(), sorted_keys.end(),
                                        keys[i]) -
                       sorted_keys.begin();
    if (taken[pos]) continue;  // a later add of a pair already held
    taken[pos] = 1;
    rows.push_back((int32_t)(keys[i] / n_atoms));
    rows.push_back((int32_t)(keys[i] % n_atoms));
    rows.push_back(excl[3 * i + 2]);
  }
}

// Per atom, its correction rows in row order (CSR: cnt[n + 1], entries
// row << 1 | role, role 0 for a1 and 1 for a2).
void correction_csr(int n_atoms, int n_excl, const int32_t *excl,
                    std::vector<int> &cnt, std::vector<int> &ent) {
  cnt.assign(n_atoms + 1, 0);
  for (int i = 0; i < n_excl; i++) {
    cnt[excl[3 * i] + 1]++;
    cnt[excl[3 * i + 1] + 1]++;
  }
  for (int i = 0; i < n_atoms; i++) cnt[i + 1] += cnt[i];
  ent.assign(std::max(2 * n_excl, 1), 0);
  std::vector<int> fill(cnt.begin(), cnt.end() - 1);
  for (int i = 0; i < n_excl; i++) {
    ent[fill[excl[3 * i]]++] = i << 1;
    ent[fill[excl[3 * i + 1]]++] = (i << 1) | 1;
  }
}

}  // namespace

struct pmnb_ewald {
  pmnb_ctx *ctx = nullptr;
  int device = 0, n = 0, nk = 0, H = 0, n_excl = 0;
  double alpha = 0, alpha2 = 0, rc = 0, kc = 0, fac = 0, V = 0, Vr = 0;
  double csf = 1.0, e_self = 0, e_charged = 0;
  int n_runs = 0, n_chunks = 0;
  DevBuf kv, q, tab, sk, kf, part, xptr, xent, xrows, runs, chunks;
  // the box and topology the k-vectors, charges and rows were built for
  double box_key[12] = {0};
  unsigned long long topo_gen = 0;
};

extern "C" {

int pmnb_ewald_create(pmnb_ctx *ctx, double alpha, double accuracy,
                      int32_t n_atoms, const double *scaled_charge,
                      double coulomb_scaling_factor, int32_t n_excl,
                      const int32_t *excl, pmnb_ewald **out) {
  if (!ctx || !out) return PMNB_ERR_ARG;
  *out = nullptr;
  int rc = check_ready(ctx);
  if (rc) return rc;
  if (n_atoms != ctx->n)
    return fail(ctx, PMNB_ERR_ARG, "n_atoms %d differs from the topology's %d",
                n_atoms, ctx->n);
  if (n_atoms > 0 && !scaled_charge)
    return fail(ctx, PMNB_ERR_ARG, "null charges");
  if (!(accuracy > 0 && accuracy < 1))
    return fail(ctx, PMNB_ERR_ARG, "accuracy must lie in (0, 1)");
  if (n_excl < 0 || (n_excl > 0 && !excl))
    return fail(ctx, PMNB_ERR_ARG, "bad exclusion table");
  for (int i = 0; i < n_excl; i++) {
    for (int k = 0; k < 2; k++)
      if (excl[3 * i + k] < 0 || excl[3 * i + k] >= n_atoms)
        return fail(ctx, PMNB_ERR_ARG, "exclusion row %d out of range", i);
    if (excl[3 * i + 2] < PMNB_EXCL_NONE || excl[3 * i + 2] > PMNB_EXCL_FULL)
      return fail(ctx, PMNB_ERR_ARG, "exclusion row %d has class %d", i,
                  excl[3 * i + 2]);
  }
  std::vector<int32_t> rows;
  correction_rows(n_atoms, n_excl, excl, rows);
  n_excl = (int32_t)(rows.size() / 3);
  excl = rows.empty() ? nullptr : rows.data();
  const Box &b = ctx->box;
  pmnb_ewald *ew = new pmnb_ewald;
  ew->ctx = ctx;
  ew->device = ctx->prm.device;
  ew->n = n_atoms;
  ew->n_excl = n_excl;
  std::memcpy(ew->box_key, ctx->box_key, sizeof ew->box_key);
  ew->topo_gen = ctx->topo_gen;
  ew->csf = coulomb_scaling_factor;
  // NonbondedFullEwaldSystemForce::initialize (:597-700), orthogonal box
  const double L[3] = {b.e[0], b.e[1], b.e[2]};
  ew->V = std::fabs((L[0] * L[1]) * L[2]);
  ew->Vr = 1.0 / ew->V;
  ew->alpha = alpha > 0 ? alpha
                        : std::sqrt(M_PI) * std::pow(5.5 * n_atoms / (ew->V * ew->V),
                                                     1.0 / 6.0);
  ew->alpha2 = ew->alpha * ew->alpha;
  const double alpha2r = 1.0 / ew->alpha2;
  const double p = -std::log(accuracy);
  ew->rc = std::sqrt(p) / ew->alpha;
  ew->kc = 2.0 * ew->alpha * std::sqrt(p);
  ew->fac = 1.0 / (4.0 * ew->alpha * ew->alpha);
  const int hmax = (int)std::floor(ew->kc / (2.0 * M_PI) * L[0]);
  const int kmax = (int)std::floor(ew->kc / (2.0 * M_PI) * L[1]);
  const int lmax = (int)std::floor(ew->kc / (2.0 * M_PI) * L[2]);
  ew->H = std::max(2, std::max(hmax, std::max(kmax, lmax)) + 1);
  std::vector<double4> kv;
  std::vector<int4> hkl;
  const double kc2 = ew->kc * ew->kc;
  for (int h = 0; h <= hmax; h++) {
    const double kx = 2.0 * M_PI * h * b.er[0];
    for (int k = (h == 0 ? 0 : -kmax); k <= kmax; k++) {
      const double ky = 2.0 * M_PI * k * b.er[1];
      for (int l = (h == 0 && k == 0 ? 1 : -lmax); l <= lmax; l++) {
        const double kz = 2.0 * M_PI * l * b.er[2];
        const double k2 = kx * kx + ky * ky + kz * kz;
        if (k2 < kc2) {
          const double bk = 1.0 / k2 * std::exp(-k2 * alpha2r / 4.0);
          kv.push_back(make_double4(kx, ky, kz, bk));
          hkl.push_back(make_int4(h, k, l, 0));
        }
      }
    }
  }
  ew->nk = (int)kv.size();
  // runs of equal (h, k) (contiguous l) and their chunks of l values
  std::vector<EwRun> runs;
  std::vector<EwChunk> chunks;
  for (int i = 0; i < ew->nk; i++) {
    if (runs.empty() || runs.back().h != hkl[i].x || runs.back().k != hkl[i].y)
      runs.push_back(EwRun{hkl[i].x, hkl[i].y, hkl[i].z, hkl[i].z, i});
    else
      runs.back().l_hi = hkl[i].z;
  }
  // chunks of |l| values, both signs of l in one chunk: |l| runs over
  // [m_lo, m_hi], the |l| of the run's l interval
  for (int r = 0; r < (int)runs.size(); r++) {
    const int lo = runs[r].l_lo, hi = runs[r].l_hi;
    const int m_hi = std::max(std::abs(lo), std::abs(hi));
    const int m_lo = (lo <= 0 && hi >= 0) ? 0 : std::min(std::abs(lo), std::abs(hi));
    for (int m0 = m_lo; m0 <= m_hi; m0 += kEwChunk)
      chunks.push_back(EwChunk{r, m0, std::min(m0 + kEwChunk - 1, m_hi)});
  }
  ew->n_runs = (int)runs.size();
  ew->n_chunks = (int)chunks.size();
  double q2 = 0.0, qs = 0.0;
  for (int i = 0; i < n_atoms; i++) q2 += scaled_charge[i] * scaled_charge[i];
  for (int i = 0; i < n_atoms; i++) qs += scaled_charge[i];
  ew->e_self = -q2 * ew->alpha / std::sqrt(M_PI);
  ew->e_charged = 0.0;
  if (std::fabs(qs * 0.00268283) > 1.0e-5)
    ew->e_charged = -M_PI / (2.0 * ew->V * ew->alpha2) * qs * qs;
  // per-atom exclusion rows in row order
  std::vector<int> cnt, ent;
  correction_csr(n_atoms, n_excl, excl, cnt, ent);
  auto up = [&](DevBuf &d, const void *src, size_t bytes) -> hipError_t {
    hipError_t e = d.ensure(std::max<size_t>(bytes, 16));
    if (e == hipSuccess && bytes && src)
      e = hipMemcpy(d.p, src, bytes, hipMemcpyHostToDevice);
    return e;
  };
  hipError_t e = hipSetDevice(ew->device);
  if (e == hipSuccess) e = up(ew->kv, kv.data(), kv.size() * sizeof(double4));
  if (e == hipSuccess) e = up(ew->q, scaled_charge, (size_t)n_atoms * sizeof(double));
  if (e == hipSuccess)
    e = ew->tab.ensure((size_t)3 * ew->H * 2 * std::max(n_atoms, 1) * sizeof(double));
  if (e == hipSuccess) e = ew->sk.ensure((size_t)std::max(ew->nk, 1) * sizeof(double2));
  if (e == hipSuccess) e = ew->kf.ensure((size_t)std::max(ew->nk, 1) * sizeof(EwKF));
  if (e == hipSuccess) e = ew->part.ensure((size_t)kEwFinBlocks * kEwFinSums * sizeof(double));
  if (e == hipSuccess) e = up(ew->xptr, cnt.data(), (n_atoms + 1) * sizeof(int));
  if (e == hipSuccess) e = up(ew->xent, ent.data(), 2 * (size_t)n_excl * sizeof(int));
  if (e == hipSuccess) e = up(ew->xrows, excl, 3 * (size_t)n_excl * sizeof(int));
  if (e == hipSuccess) e = up(ew->runs, runs.data(), runs.size() * sizeof(EwRun));
  if (e == hipSuccess) e = up(ew->chunks, chunks.data(), chunks.size() * sizeof(EwChunk));
  if (e != hipSuccess) {
    pmnb_ewald_destroy(ew);
    return fail(ctx, PMNB_ERR_HIP, "Ewald set-up failed: %s", hipGetErrorString(e));
  }
  *out = ew;
  return PMNB_OK;
}

int pmnb_ewald_info(const pmnb_ewald *ew, double *alpha, double *rc,
                    double *kc, int32_t *n_kvectors) {
  if (!ew) return PMNB_ERR_ARG;
  if (alpha) *alpha = ew->alpha;
  if (rc) *rc = ew->rc;
  if (kc) *kc = ew->kc;
  if (n_kvectors) *n_kvectors = ew->nk;
  return PMNB_OK;
}

int pmnb_ewald_evaluate_device(pmnb_ewald *ew, const double *d_positions,
                               double *d_forces, double *d_out, uint32_t flags,
                               void *stream) {
  if (!ew) return PMNB_ERR_ARG;
  pmnb_ctx *ctx = ew->ctx;
  int rc = check_ready(ctx);
  if (rc) return rc;
  if (ctx->n != ew->n || ctx->topo_gen != ew->topo_gen ||
      std::memcmp(ctx->box_key, ew->box_key, sizeof ew->box_key) != 0)
    return fail(ctx, PMNB_ERR_STATE,
                "the box or topology changed since pmnb_ewald_create; "
                "create the k-space terms again");
  if (flags & PMNB_FLAG_MOLVIRIAL)
    return fail(ctx, PMNB_ERR_UNSUPPORTED,
                "the molecular virial of the k-space term is not built");
  if (ew->n > 0 && (!d_positions || !d_forces || !d_out))
    return fail(ctx, PMNB_ERR_ARG, "null device pointer");
  HIPCHK(hipSetDevice(ew->device));
  hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
  if (ew->n == 0) return PMNB_OK;
  EwDev d{};
  d.box = ctx->box;
  d.n = ew->n;
  d.nk = ew->nk;
  d.H = ew->H;
  d.kv = ew->kv.as<double4>();
  d.runs = ew->runs.as<EwRun>();
  d.chunks = ew->chunks.as<EwChunk>();
  d.n_runs = ew->n_runs;
  d.n_chunks = ew->n_chunks;
  d.q = ew->q.as<double>();
  d.tab = ew->tab.as<double>();
  d.sk = ew->sk.as<double2>();
  d.kf = ew->kf.as<EwKF>();
  d.part = ew->part.as<double>();
  d.xptr = ew->xptr.as<int>();
  d.xent = ew->xent.as<int>();
  d.xrows = ew->xrows.as<int>();
  d.n_excl = ew->n_excl;
  d.alpha = ew->alpha;
  d.alpha2 = ew->alpha2;
  d.two_alpha_pi = 2.0 * ew->alpha / std::sqrt(M_PI);
  d.csf = ew->csf;
  d.Vr = ew->Vr;
  d.fac = ew->fac;
  const int nb = blocks_for(ew->n, kEwBlock);
  hipLaunchKernelGGL(k_ew_tables, dim3(nb), dim3(kEwBlock), 0, s, d, d_positions);
  if (ew->n_chunks > 0)
    hipLaunchKernelGGL(k_ew_sk, dim3(ew->n_chunks), dim3(kEwBlock), 0, s, d);
  const int nbf = blocks_for(ew->n, 64 * kEwAtomsLarge);
  if (nbf >= 256)
    hipLaunchKernelGGL((k_ew_force<kEwLanesLarge, kEwAtomsLarge>), dim3(nbf),
                       dim3(64 * kEwLanesLarge), 0, s, d, d_positions, d_forces);
  else
    hipLaunchKernelGGL((k_ew_force<kEwLanesSmall, 2>), dim3(blocks_for(ew->n, 128)),
                       dim3(64 * kEwLanesSmall), 0, s, d, d_positions, d_forces);
  hipLaunchKernelGGL(k_ew_partials, dim3(kEwFinBlocks), dim3(kEwBlock), 0, s, d,
                     d_positions);
  hipLaunchKernelGGL(k_ew_finish, dim3(1), dim3(64), 0, s, d, d_out,
                     (int)((flags & PMNB_FLAG_VIRIAL) != 0),
                     ew->e_self + ew->e_charged, ew->e_charged);
  HIPCHK(hipGetLastError());
  return PMNB_OK;
}

void pmnb_ewald_destroy(pmnb_ewald *ew) {
  if (!ew) return;
  (void)hipSetDevice(ew->device);
  (void)hipDeviceSynchronize();
  DevBuf *bufs[] = {&ew->kv, &ew->q, &ew->tab, &ew->sk, &ew->kf, &ew->part,
                    &ew->xptr, &ew->xent, &ew->xrows, &ew->runs, &ew->chunks};
  for (DevBuf *b : bufs) b->release();
  delete ew;
}

}  // extern "C"

// ------------------------------------------------------------------- PME
struct pmnb_pme {
  pmnb_ctx *ctx = nullptr;
  int device = 0, n = 0, n_excl = 0, order = 0;
  int nx = 0, ny = 0, nz = 0;
  double alpha = 0, kc = 0, V = 0;
  double csf = 1.0, e_self = 0, e_charged = 0;
  DevBuf q, qfix, qreal, qk, mod, ex, part, ew_part, xptr, xent, xrows;
  int n_part = 0;
  hipfftHandle fwd = 0, inv = 0;
  bool have_fwd = false, have_inv = false;
  double box_key[12] = {0};
  unsigned long long topo_gen = 0;
};

namespace {

// Host restatement of BSpline::set at w (BSpline.cpp:58-95), for dftmod.
void host_bspline(int order, double w, std::vector<double> &th) {
  th.assign(order, 0.0);
  if (order == 1) {
    th[0] = 1;
    return;
  }
  if (order == 2) {
    th[1] = w;
    th[0] = 1.0 - w;
    return;
  }
  th[order - 1] = 0.0;
  th[1] = w;
  th[0] = 1.0 - w;
  for (int k = 2; k < order - 1; k++) {
    const double div = 1.0 / k;
    th[k] = div * w * th[k - 1];
    for (int j = 1; j < k; j++)
      th[k - j] = div * ((w + j) * th[k - j - 1] + (k + 1 - j - w) * th[k - j]);
    th[0] = div * (1 - w) * th[0];
  }
  const double div = 1.0 / (order - 1.0);
  th[order - 1] = div * w * th[order - 2];
  for (int j = 1; j < order - 1; j++)
    th[order - 1 - j] =
        div * ((w + j) * th[order - j - 2] + (order - j - w) * th[order - 1 - j]);
  th[0] = div * (1 - w) * th[0];
}

// Grid::dftmod (Grid.h:590-602): B(m) = 1 / |sum_j M(j) e^{2 pi i m j / n}|^2.
void host_dftmod(int order, int n, std::vector<double> &out) {
  std::vector<double> th;
  host_bspline(order, 0.0, th);
  out.assign(n, 0.0);
  for (int i = 0; i < n; i++) {
    double sc = 0.0, ss = 0.0;
    for (int j = 0; j < order; j++) {
      const double x = M_PI * 2.0 * i * j / (double)n;
      sc += th[j] * std::cos(x);
      ss += th[j] * std::sin(x);
    }
    out[i] = 1.0 / (sc * sc + ss * ss);
  }
}

PmeDev pme_dev(pmnb_pme *pm) {
  PmeDev d{};
  d.n = pm->n;
  d.nx = pm->nx;
  d.ny = pm->ny;
  d.nz = pm->nz;
  d.nzh = pm->nz / 2 + 1;
  const Box &b = pm->ctx->box;
  const int nn[3] = {pm->nx, pm->ny, pm->nz};
  for (int a = 0; a < 3; a++) {
    // Grid::initialize (Grid.h:507-528): h^-1 = n / L, offsets
    d.hr[a] = nn[a] / b.e[a];
    d.off[a] = -(pm->order - 1) / 2 + nn[a];
    d.Lr[a] = 1.0 / b.e[a];
  }
  d.piVr = 1.0 / (M_PI * pm->V);
  d.fac = M_PI * M_PI / (pm->alpha * pm->alpha);
  d.q = pm->q.as<double>();
  d.qfix = pm->qfix.as<long long>();
  d.qreal = pm->qreal.as<double>();
  d.qk = pm->qk.as<double2>();
  const double *mod = pm->mod.as<double>(), *ex = pm->ex.as<double>();
  d.mod[0] = mod;
  d.mod[1] = mod + pm->nx;
  d.mod[2] = mod + pm->nx + pm->ny;
  d.ex[0] = ex;
  d.ex[1] = ex + pm->nx;
  d.ex[2] = ex + pm->nx + pm->ny;
  d.part = pm->part.as<double>();
  d.n_part = pm->n_part;
  return d;
}

EwDev pme_corr_dev(pmnb_pme *pm) {
  EwDev e{};
  e.box = pm->ctx->box;
  e.n = pm->n;
  e.nk = 0;  // correction terms only
  e.q = pm->q.as<double>();
  e.part = pm->ew_part.as<double>();
  e.xptr = pm->xptr.as<int>();
  e.xent = pm->xent.as<int>();
  e.xrows = pm->xrows.as<int>();
  e.n_excl = pm->n_excl;
  e.alpha = pm->alpha;
  e.alpha2 = pm->alpha * pm->alpha;
  e.two_alpha_pi = 2.0 * pm->alpha / std::sqrt(M_PI);
  e.csf = pm->csf;
  e.Vr = 1.0 / pm->V;
  e.fac = 1.0 / (4.0 * pm->alpha * pm->alpha);
  return e;
}

template <int P>
void pme_spread(pmnb_pme *pm, const PmeDev &d, const double *x,
                hipStream_t s) {
  hipLaunchKernelGGL(k_pme_spread<P>, dim3(blocks_for(pm->n, kPmeBlock)),
                     dim3(kPmeBlock), 0, s, d, x);
  const long long total = (long long)pm->nx * pm->ny * pm->nz;
  hipLaunchKernelGGL(k_pme_to_real,
                     dim3((unsigned)((total + kPmeBlock - 1) / kPmeBlock)),
                     dim3(kPmeBlock), 0, s, d, total);
}

template <int P>
void pme_launch_gather(pmnb_pme *pm, const PmeDev &d, const EwDev &e,
                       const double *x, double *f, hipStream_t s) {
  hipLaunchKernelGGL(k_pme_gather<P>, dim3(blocks_for(pm->n, kPmeBlock)),
                     dim3(kPmeBlock), 0, s, d, e, x, f);
}

}  // namespace

extern "C" {

int pmnb_pme_create(pmnb_ctx *ctx, int32_t nx, int32_t ny, int32_t nz,
                    int32_t order, double cutoff, double accuracy,
                    double alpha, int32_t n_atoms,
                    const double *scaled_charge,
                    double coulomb_scaling_factor, int32_t n_excl,
                    const int32_t *excl, pmnb_pme **out) {
  if (!ctx || !out) return PMNB_ERR_ARG;
  *out = nullptr;
  int rc = check_ready(ctx);
  if (rc) return rc;
  // NonbondedPMEwaldSystemForce::doMake (NonbondedPMEwaldSystemForce.h
  // :1029-1058): even order >= 2, every grid size >= order, cutoff and
  // accuracy positive
  if (order < 2 || (order % 2) != 0 || nx < order || ny < order ||
      nz < order || !(cutoff > 0.0) || !(accuracy > 0.0))
    return fail(ctx, PMNB_ERR_ARG,
                "PME needs an even order >= 2, grid sizes >= order, cutoff "
                "and accuracy > 0");
  if (order > 8)
    return fail(ctx, PMNB_ERR_UNSUPPORTED, "B-spline orders 2 to 8 are built");
  if ((long long)nx * ny * nz > (1ll << 28))
    return fail(ctx, PMNB_ERR_UNSUPPORTED, "PME grid too large");
  if (n_atoms != ctx->n)
    return fail(ctx, PMNB_ERR_ARG, "n_atoms %d differs from the topology's %d",
                n_atoms, ctx->n);
  if (n_atoms > 0 && !scaled_charge)
    return fail(ctx, PMNB_ERR_ARG, "null charges");
  if (n_excl < 0 || (n_excl > 0 && !excl))
    return fail(ctx, PMNB_ERR_ARG, "bad exclusion table");
  for (int i = 0; i < n_excl; i++) {
    for (int k = 0; k < 2; k++)
      if (excl[3 * i + k] < 0 || excl[3 * i + k] >= n_atoms)
        return fail(ctx, PMNB_ERR_ARG, "exclusion row %d out of range", i);
    if (excl[3 * i + 2] < PMNB_EXCL_NONE || excl[3 * i + 2] > PMNB_EXCL_FULL)
      return fail(ctx, PMNB_ERR_ARG, "exclusion row %d has class %d", i,
                  excl[3 * i + 2]);
  }
  std::vector<int32_t> rows;
  correction_rows(n_atoms, n_excl, excl, rows);
  n_excl = (int32_t)(rows.size() / 3);
  const Box &b = ctx->box;
  pmnb_pme *pm = new pmnb_pme;
  pm->ctx = ctx;
  pm->device = ctx->prm.device;
  pm->n = n_atoms;
  pm->n_excl = n_excl;
  pm->order = order;
  pm->nx = nx;
  pm->ny = ny;
  pm->nz = nz;
  pm->csf = coulomb_scaling_factor;
  std::memcpy(pm->box_key, ctx->box_key, sizeof pm->box_key);
  pm->topo_gen = ctx->topo_gen;
  // NonbondedPMEwaldSystemForce::initialize (:525-612), orthogonal box
  pm->V = std::fabs((b.e[0] * b.e[1]) * b.e[2]);
  if (alpha > 0.0) {
    pm->alpha = alpha;
  } else {
    // erfc(alpha rc) / rc = accuracy by doubling, then 100 bisection steps
    double a = 1.0;
    while (std::erfc(a * cutoff) / cutoff >= accuracy) a *= 2.0;
    double lo = 0.0, hi = a;
    for (int i = 0; i < 100; i++) {
      a = 0.5 * (lo + hi);
      if (std::erfc(a * cutoff) / cutoff >= accuracy) lo = a;
      else hi = a;
    }
    pm->alpha = a;
  }
  pm->kc = 2.0 * pm->alpha * std::sqrt(-std::log(accuracy));
  double q2 = 0.0, qs = 0.0;
  for (int i = 0; i < n_atoms; i++) q2 += scaled_charge[i] * scaled_charge[i];
  for (int i = 0; i < n_atoms; i++) qs += scaled_charge[i];
  pm->e_self = -q2 * pm->alpha / std::sqrt(M_PI);
  pm->e_charged = 0.0;
  if (std::fabs(qs * 0.00268283) > 1.0e-5)
    pm->e_charged = -M_PI / (2.0 * pm->V * pm->alpha * pm->alpha) * qs * qs;
  // per-axis tables: B(m) and exp(-pi^2 m^2 / alpha^2), m = i0 / L
  const double fac = M_PI * M_PI / (pm->alpha * pm->alpha);
  std::vector<double> mod, ex, tmp;
  const int nn[3] = {nx, ny, nz};
  for (int a = 0; a < 3; a++) {
    host_dftmod(order, nn[a], tmp);
    mod.insert(mod.end(), tmp.begin(), tmp.end());
    const double Lr = 1.0 / b.e[a];
    for (int i = 0; i < nn[a]; i++) {
      const int i0 = i <= nn[a] / 2 ? i : i - nn[a];
      const double m = i0 * Lr;
      ex.push_back(std::exp(-fac * (m * m)));
    }
  }
  std::vector<int> cnt, ent;
  correction_csr(n_atoms, n_excl, rows.empty() ? nullptr : rows.data(), cnt,
                 ent);
  const long long total = (long long)nx * ny * nz;
  const long long half = (long long)nx * ny * (nz / 2 + 1);
  pm->n_part = (int)((half + kPmeBlock - 1) / kPmeBlock);
  auto up = [&](DevBuf &d, const void *src, size_t bytes) -> hipError_t {
    hipError_t e = d.ensure(std::max<size_t>(bytes, 16));
    if (e == hipSuccess && bytes && src)
      e = hipMemcpy(d.p, src, bytes, hipMemcpyHostToDevice);
    return e;
  };
  hipError_t e = hipSetDevice(pm->device);
  if (e == hipSuccess) e = up(pm->q, scaled_charge, (size_t)n_atoms * sizeof(double));
  if (e == hipSuccess) e = pm->qfix.ensure((size_t)total * sizeof(long long));
  if (e == hipSuccess) e = hipMemset(pm->qfix.p, 0, (size_t)total * sizeof(long long));
  if (e == hipSuccess) e = pm->qreal.ensure((size_t)total * sizeof(double));
  if (e == hipSuccess) e = pm->qk.ensure((size_t)half * sizeof(double2));
  if (e == hipSuccess) e = up(pm->mod, mod.data(), mod.size() * sizeof(double));
  if (e == hipSuccess) e = up(pm->ex, ex.data(), ex.size() * sizeof(double));
  if (e == hipSuccess)
    e = pm->part.ensure((size_t)pm->n_part * kPmeSums * sizeof(double));
  if (e == hipSuccess)
    e = pm->ew_part.ensure((size_t)kEwFinBlocks * kEwFinSums * sizeof(double));
  if (e == hipSuccess) e = up(pm->xptr, cnt.data(), (n_atoms + 1) * sizeof(int));
  if (e == hipSuccess) e = up(pm->xent, ent.data(), 2 * (size_t)n_excl * sizeof(int));
  if (e == hipSuccess) e = up(pm->xrows, rows.data(), rows.size() * sizeof(int));
  if (e != hipSuccess) {
    pmnb_pme_destroy(pm);
    return fail(ctx, PMNB_ERR_HIP, "PME set-up failed: %s", hipGetErrorString(e));
  }
  if (hipfftPlan3d(&pm->fwd, nx, ny, nz, HIPFFT_D2Z) != HIPFFT_SUCCESS) {
    pmnb_pme_destroy(pm);
    return fail(ctx, PMNB_ERR_HIP, "hipFFT D2Z plan %dx%dx%d failed", nx, ny, nz);
  }
  pm->have_fwd = true;
  if (hipfftPlan3d(&pm->inv, nx, ny, nz, HIPFFT_Z2D) != HIPFFT_SUCCESS) {
    pmnb_pme_destroy(pm);
    return fail(ctx, PMNB_ERR_HIP, "hipFFT Z2D plan %dx%dx%d failed", nx, ny, nz);
  }
  pm->have_inv = true;
  *out = pm;
  return PMNB_OK;
}

int pmnb_pme_info(const pmnb_pme *pm, double *alpha, double *kc) {
  if (!pm) return PMNB_ERR_ARG;
  if (alpha) *alpha = pm->alpha;
  if (kc) *kc = pm->kc;
  return PMNB_OK;
}

int pmnb_pme_evaluate_device(pmnb_pme *pm, const double *d_positions,
                             double *d_forces, double *d_out, uint32_t flags,
                             void *stream) {
  if (!pm) return PMNB_ERR_ARG;
  pmnb_ctx *ctx = pm->ctx;
  int rc = check_ready(ctx);
  if (rc) return rc;
  if (ctx->n != pm->n || ctx->topo_gen != pm->topo_gen ||
      std::memcmp(ctx->box_key, pm->box_key, sizeof pm->box_key) != 0)
    return fail(ctx, PMNB_ERR_STATE,
                "the box or topology changed since pmnb_pme_create; create "
                "the PME terms again");
  if (flags & PMNB_FLAG_MOLVIRIAL)
    return fail(ctx, PMNB_ERR_UNSUPPORTED,
                "the molecular virial of the PME terms is not built");
  if (pm->n > 0 && (!d_positions || !d_forces || !d_out))
    return fail(ctx, PMNB_ERR_ARG, "null device pointer");
  HIPCHK(hipSetDevice(pm->device));
  hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
  if (pm->n == 0) return PMNB_OK;
  const PmeDev d = pme_dev(pm);
  const EwDev e = pme_corr_dev(pm);
  const int virial = (flags & PMNB_FLAG_VIRIAL) != 0;
  switch (pm->order) {
    case 2: pme_spread<2>(pm, d, d_positions, s); break;
    case 4: pme_spread<4>(pm, d, d_positions, s); break;
    case 6: pme_spread<6>(pm, d, d_positions, s); break;
    default: pme_spread<8>(pm, d, d_positions, s); break;
  }
  HIPCHK(hipGetLastError());
  if (hipfftSetStream(pm->fwd, s) != HIPFFT_SUCCESS ||
      hipfftExecD2Z(pm->fwd, d.qreal, (hipfftDoubleComplex *)d.qk) !=
          HIPFFT_SUCCESS)
    return fail(ctx, PMNB_ERR_HIP, "hipFFT D2Z failed");
  hipLaunchKernelGGL(k_pme_scalar, dim3(pm->n_part), dim3(kPmeBlock), 0, s, d,
                     virial);
  HIPCHK(hipGetLastError());
  if (hipfftSetStream(pm->inv, s) != HIPFFT_SUCCESS ||
      hipfftExecZ2D(pm->inv, (hipfftDoubleComplex *)d.qk, d.qreal) !=
          HIPFFT_SUCCESS)
    return fail(ctx, PMNB_ERR_HIP, "hipFFT Z2D failed");
  switch (pm->order) {
    case 2: pme_launch_gather<2>(pm, d, e, d_positions, d_forces, s); break;
    case 4: pme_launch_gather<4>(pm, d, e, d_positions, d_forces, s); break;
    case 6: pme_launch_gather<6>(pm, d, e, d_positions, d_forces, s); break;
    default: pme_launch_gather<8>(pm, d, e, d_positions, d_forces, s); break;
  }
  hipLaunchKernelGGL(k_pme_finish, dim3(1), dim3(64), 0, s, d, d_out, virial);
  hipLaunchKernelGGL(k_ew_partials, dim3(kEwFinBlocks), dim3(kEwBlock), 0, s,
                     e, d_positions);
  hipLaunchKernelGGL(k_ew_finish, dim3(1), dim3(64), 0, s, e, d_out, virial,
                     pm->e_self + pm->e_charged, pm->e_charged);
  HIPCHK(hipGetLastError());
  return PMNB_OK;
}

void pmnb_pme_destroy(pmnb_pme *pm) {
  if (!pm) return;
  (void)hipSetDevice(pm->device);
  (void)hipDeviceSynchronize();
  if (pm->have_fwd) (void)hipfftDestroy(pm->fwd);
  if (pm->have_inv) (void)hipfftDestroy(pm->inv);
  DevBuf *bufs[] = {&pm->q, &pm->qfix, &pm->qreal, &pm->qk, &pm->mod, &pm->ex,
                    &pm->part, &pm->ew_part, &pm->xptr, &pm->xent, &pm->xrows};
  for (DevBuf *b : bufs) b->release();
  delete pm;
}

}  // extern "C"

extern "C" {

int pmnb_lf_set_kspace(pmnb_leapfrog *lf, pmnb_pme *pme, pmnb_ewald *ewald) {
  if (!lf) return PMNB_ERR_ARG;
  if ((pme && pme->ctx != lf->ctx) || (ewald && ewald->ctx != lf->ctx))
    return fail(lf->ctx, PMNB_ERR_ARG,
                "the k-space terms belong to another context");
  lf_drop_graph(lf);
  lf->pme = pme;
  lf->ewald = ewald;
  return PMNB_OK;
}

}  // extern "C"

// --------------------------------------------- evaluation without allocation
// pmnb_evaluate_device_noalloc keeps every per-evaluation array of the tile
// path (cell counts and starts, the sorted structure of arrays, work items,
// partial sums, the fixed-point force array, the system's grid and box) in
// `scratch`, a file-scope device array of kScratchBytes that each GPU
// holds for the life of the process.  Its device address on each GPU is
// looked up once, by pmnb_create with that device current; host code uses
// only that address (the array's own name is a host-side stand-in).
namespace {

constexpr size_t kScratchBytes = (size_t)64 << 20;  // 64 MiB per GPU
constexpr int kScratchDevices = 64;

__device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scratch_mu;
unsigned char *g_scratch_ptr[kScratchDevices] = {};
hipError_t g_scratch_err[kScratchDevices] = {};
bool g_scratch_done[kScratchDevices] = {};

// With `device` current: records the device address of `scratch` on it.
void scratch_init(int device) {
  if (device < 0 || device >= kScratchDevices) return;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_done[device]) return;
  void *ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  g_scratch_err[device] = e;
  g_scratch_ptr[device] = e == hipSuccess ? static_cast<unsigned char *>(ptr)
                                          : nullptr;
  g_scratch_done[device] = true;
}

// Carves 256-byte aligned arrays from a base address (or only counts the
// bytes when base is null).
struct Carve {
  unsigned char *base;
  size_t off = 0;
  template <class T> T *take(size_t count) {
    off = (off + 255) & ~(size_t)255;
    T *p = base ? reinterpret_cast<T *>(base + off) : nullptr;
    off += std::max<size_t>(count, 1) * sizeof(T);
    return p;
  }
};

// The tile path's arrays for one system of ctx, carved from `base`.
EvalBufs carve_bufs(const pmnb_ctx *ctx, unsigned char *base, size_t &bytes,
                    long long &nfix) {
  Carve c{base};
  const size_t N = (size_t)std::max(ctx->n, 1);
  const size_t nc = (size_t)std::max(ctx->grid.ncell, 1);
  const int W = ctx->prm.precision == PMNB_PREC_FP64 ? FixWords<true>::W
                                                      : FixWords<false>::W;
  const size_t items = (size_t)(ctx->n / kTileAtoms + ctx->ncol_total + 1);
  EvalBufs B{};
  B.ffix = c.take<long long>(N * W);
  B.cell_of = c.take<int>(N);
  B.count = c.take<int>(nc);
  B.start = c.take<int>(nc);
  B.rank = nullptr;  // the tile path fills bucket rows instead
  B.sorted = c.take<int>(N);
  B.cap = tile_bucket_cap(ctx->n, (int)nc);
  B.bucket = c.take<int>(nc * (size_t)B.cap);
  B.sx = c.take<double>(N);
  B.sy = c.take<double>(N);
  B.sz = c.take<double>(N);
  B.sq = c.take<double>(N);
  B.loc = c.take<float4>(N);
  B.stype = c.take<int>(N);
  B.smol = c.take<int>(N);
  B.scell = c.take<int>(N);
  B.srec = c.take<int2>(N);
  B.partial = c.take<double>(items * ACC_N);
  B.npairs = c.take<unsigned long long>(1);
  B.items = c.take<int4>(items);
  B.nitems = c.take<int>(1);
  B.rep_first = c.take<int>(2);
  B.reps = c.take<TileRep>(1);
  B.col_rep = c.take<int>((size_t)std::max(ctx->ncol_total, 1));
  B.box = c.take<Box>(1);
  bytes = c.off;
  nfix = (long long)(N * W);
  return B;
}

// Writes the system's grid and box into the scratch, marks every column as
// replica 0 and clears the fixed-point force array (the previous call on
// this GPU may have been another system's, or may have stopped early).
__global__ void k_scratch_setup(TileRep rep, Box box, TileRep *d_rep,
                                Box *d_box, int *col_rep, int ncol,
                                long long *ffix, long long nfix) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long stride = (long long)gridDim.x * blockDim.x;
  if (t == 0) {
    *d_rep = rep;
    *d_box = box;
  }
  for (long long k = t; k < ncol; k += stride) col_rep[k] = 0;
  for (long long k = t; k < nfix; k += stride) ffix[k] = 0;
}

}  // namespace

extern "C" {

int pmnb_evaluate_device_noalloc(pmnb_ctx *ctx, const double *d_positions,
                                 double *d_forces, double *d_out,
                                 const double *d_mol_com, uint32_t flags,
                                 void *stream) {
  int rc = check_ready(ctx);
  if (rc) return rc;
  if (ctx->batch || !ctx->use_tile || ctx->reps.size() != 1)
    return fail(ctx, PMNB_ERR_UNSUPPORTED,
                "the allocation-free evaluation covers one system on the tile "
                "path (algorithm cutoff >= switching cutoff, at most 8 LJ "
                "types)");
  if (ctx->timing)
    return fail(ctx, PMNB_ERR_STATE, "kernel timing is on");
  if (ctx->n > 0 && (!d_positions || !d_forces || !d_out))
    return fail(ctx, PMNB_ERR_ARG, "null device pointer");
  if ((flags & PMNB_FLAG_MOLVIRIAL) && !d_mol_com)
    return fail(ctx, PMNB_ERR_ARG, "molecular virial needs molecule centres");
  const int dev = ctx->prm.device;
  unsigned char *base = nullptr;
  hipError_t serr = hipErrorInvalidDevice;
  {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (dev >= 0 && dev < kScratchDevices && g_scratch_done[dev]) {
      base = g_scratch_ptr[dev];
      serr = g_scratch_err[dev];
    }
  }
  if (!base)
    return fail(ctx, PMNB_ERR_HIP, "no scratch address on device %d: %s", dev,
                hipGetErrorString(serr));
  size_t need = 0;
  long long nfix = 0;
  carve_bufs(ctx, nullptr, need, nfix);
  if (need > kScratchBytes)
    return fail(ctx, PMNB_ERR_CAPACITY,
                "the evaluation needs %zu bytes of scratch; the static scratch "
                "holds %zu (64 MiB)", need, kScratchBytes);
  const EvalBufs B = carve_bufs(ctx, base, need, nfix);
  HIPCHK(hipSetDevice(dev));
  hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
  if (ctx->n == 0) return PMNB_OK;
  const long long work = std::max<long long>(nfix, ctx->ncol_total);
  const unsigned grid = (unsigned)std::min<long long>(
      std::max<long long>(1, (work + kBlock - 1) / kBlock), 4096);
  hipLaunchKernelGGL(k_scratch_setup, dim3(grid), dim3(kBlock), 0, s,
                     ctx->reps[0], ctx->box, const_cast<TileRep *>(B.reps),
                     const_cast<Box *>(B.box), const_cast<int *>(B.col_rep),
                     ctx->ncol_total, B.ffix, nfix);
  HIPCHK(hipGetLastError());
  return run_device_bufs(ctx, B, d_positions, d_forces, d_out, d_mol_com,
                         flags, s, false);
}

}  // extern "C"
