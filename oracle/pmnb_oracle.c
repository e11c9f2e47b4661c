/*
 * pmnb_oracle.c -- FP64 CPU restatement of ProtoMol's short-range nonbonded
 * cutoff force path.  TEST INFRASTRUCTURE ONLY: it is the checker for the
 * HIP product path (tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load it); the product never links or calls it.
 *
 * Restated from (paths relative to the reference's protomol/src/protomol):
 *   topology/PeriodicBoundaryConditions.cpp:32-77   pbc_set
 *   topology/PeriodicBoundaryConditions.h:72-110    pbc_min_diff
 *   topology/PeriodicBoundaryConditions.h:113-147   pbc_min_pos
 *   topology/CubicCellManager.cpp:40-65             cell size / findCell
 *   topology/ArrayCellListStructure.cpp:22-152      grid dims, non-empty cells
 *   topology/ArrayCellListStructure.h:94-121        findPeriodic, basisCell
 *   topology/Topology.h:44-86                       linked cell lists
 *   topology/CellListEnumeratorPeriodicBoundaries.h:25-134  delta list + walk
 *   force/nonbonded/NonbondedCutoffForce.h:37-54    cell-pair loop
 *   force/OneAtomPairTwo.h:57-125                   per-pair sequence
 *   force/LennardJonesForce.h:29-61                 LJ functor
 *   force/CoulombForce.h:27-38                      Coulomb functor
 *   force/nonbonded/NonbondedFullEwaldSystemForce.h:810-821  Ewald real (A&S)
 *   base/MathUtilities.cpp:39-49                    poly5
 *   switch/C2SwitchingFunction.{h:42-53,cpp:12-18}  C2
 *   switch/C1SwitchingFunction.{h:41-50,cpp:12-16}  C1
 *   topology/ExclusionTable.{h:116-131,cpp:105-120} exclusion lookup
 *   type/ScalarStructure.cpp:36-112                 addVirial
 *   force/nonbonded/NonbondedSimpleFullSystemForce.h:40-44  brute force i<j
 *
 * Deliberate difference: an atom whose cell index comes out as -1 or n
 * (position within an ulp of the box face, CubicCellManager.cpp:40-44) is
 * wrapped into [0, n) instead of indexing out of bounds (Topology.h:72-77);
 * the pair set is cell independent, so this changes no result.
 *
 * Built with gcc -O2 and no FMA contraction (-ffp-contract=off), matching
 * the reference's own x86-64 build, so sums follow the same rounding.
 */
#include <math.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_EPSILON 1.0e-14         /* base/PMConstants.cpp:47 */
#define ORC_MAX_INT_2 (2147483647 / 2) /* base/PMConstants.cpp:44 */
#define ORC_PI 3.14159265358979323846     /* glibc M_PI */

enum { ORC_ELEC_NONE = 0, ORC_ELEC_COULOMB = 1, ORC_ELEC_EWALD = 2 };
enum { ORC_SW_C2 = 0, ORC_SW_C1 = 1 };
enum { ORC_NONE = 0, ORC_MODIFIED = 1, ORC_FULL = 2 };

typedef struct {
  int32_t n_atoms, n_types, n_excl, n_mol;
  const double *pos;    /* 3N, AoS x0 y0 z0 x1 ... */
  const int32_t *type;  /* N */
  const double *q;      /* N, scaled charges (q * 18.2226123264) */
  const int32_t *mol;   /* N */
  const double *lj;     /* T*T*4: A, B, A14, B14 (symmetric) */
  const int32_t *excl;  /* n_excl*3: a1, a2, class (applied in order) */
  const double *com;    /* n_mol*3 molecule centres (molecular virial) */
  double e1[3], e2[3], e3[3], origin[3];
  double lj_switchon, lj_cutoff, el_switchon, el_cutoff;
  double alg_cutoff, cell_size, csf, alpha;
  int32_t elec, el_switch, flags; /* flags bit0 virial, bit1 mol virial */
  int32_t pad;
} orc_system;

/* ------------------------------------------------------------------ PBC */
typedef struct {
  double e1[3], e2[3], e3[3], e1r[3], e2r[3], e3r[3];
  double origin[3], mn[3], mx[3], H[3], H2[3];
  double D, V;
  int orth;
} orc_pbc;

static void cross3(const double *a, const double *b, double *r) {
  r[0] = a[1] * b[2] - a[2] * b[1];
  r[1] = a[2] * b[0] - a[0] * b[2];
  r[2] = a[0] * b[1] - a[1] * b[0];
}
static double dot3(const double *a, const double *b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

void orc_pbc_set(orc_pbc *p, const double *e1, const double *e2,
                 const double *e3, const double *origin) {
  double a1[3], a2[3], a3[3], c12[3], s[3], a[3], b[3];
  int k;
  memcpy(p->e1, e1, 24); memcpy(p->e2, e2, 24); memcpy(p->e3, e3, 24);
  memcpy(p->origin, origin, 24);
  cross3(e1, e2, c12);
  p->V = fabs(dot3(c12, e3));
  p->orth = !(e1[1] != 0.0 || e1[2] != 0.0 || e2[0] != 0.0 || e2[2] != 0.0 ||
              e3[0] != 0.0 || e3[1] != 0.0);
  cross3(e2, e3, a1);
  { double d = dot3(e1, a1); for (k = 0; k < 3; k++) p->e1r[k] = a1[k] / d; }
  cross3(e3, e1, a2);
  { double d = dot3(e2, a2); for (k = 0; k < 3; k++) p->e2r[k] = a2[k] / d; }
  cross3(e1, e2, a3);
  { double d = dot3(e3, a3); for (k = 0; k < 3; k++) p->e3r[k] = a3[k] / d; }
  for (k = 0; k < 3; k++) s[k] = e1[k] + e2[k] + e3[k];
  for (k = 0; k < 3; k++) {
    a[k] = origin[k] - s[k] * 0.5;
    b[k] = origin[k] + s[k] * 0.5;
    p->mn[k] = a[k] < b[k] ? a[k] : b[k];
    p->mx[k] = a[k] > b[k] ? a[k] : b[k];
  }
  {
    double dx = e1[0] * 0.5, dy = e2[1] * 0.5, dz = e3[2] * 0.5;
    double DX = (dx * dx) * 1.0, DY = (dy * dy) * 1.0, DZ = (dz * dz) * 1.0;
    double m = DY < DZ ? DY : DZ;
    p->D = DX < m ? DX : m;
  }
  for (k = 0; k < 3; k++) {
    p->H[k] = p->mx[k] - p->mn[k];
    p->H2[k] = p->H[k] * 0.5;
  }
}

static double nsq(const double *d) {
  return d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
}

/* minimalDifference(c1, c2, distSquared): returns c2 - c1 min image. */
void orc_pbc_min_diff(const orc_pbc *p, const double *c1, const double *c2,
                      double *diff, double *dist2) {
  int k;
  for (k = 0; k < 3; k++) diff[k] = c2[k] - c1[k];
  if (p->orth) {
    *dist2 = nsq(diff);
    if (*dist2 > p->D) {
      for (k = 0; k < 3; k++) {
        if (diff[k] < -p->H2[k]) diff[k] += p->H[k];
        else if (diff[k] > p->H2[k]) diff[k] -= p->H[k];
      }
      *dist2 = nsq(diff);
      if (*dist2 > p->D) {
        diff[0] -= p->e1[0] * rint(p->e1r[0] * diff[0]);
        diff[1] -= p->e2[1] * rint(p->e2r[1] * diff[1]);
        diff[2] -= p->e3[2] * rint(p->e3r[2] * diff[2]);
        *dist2 = nsq(diff);
      }
    }
  } else {
    double r1 = rint(dot3(p->e1r, diff)), r2 = rint(dot3(p->e2r, diff)),
           r3 = rint(dot3(p->e3r, diff));
    double t[3];
    for (k = 0; k < 3; k++) t[k] = p->e1[k] * r1 + p->e2[k] * r2;
    for (k = 0; k < 3; k++) t[k] = t[k] + p->e3[k] * r3;
    for (k = 0; k < 3; k++) diff[k] -= t[k];
    *dist2 = nsq(diff);
  }
}

/* minimalPosition(c): position relative to origin in the basis image. */
void orc_pbc_min_pos(const orc_pbc *p, const double *c, double *out) {
  double d[3];
  int k;
  for (k = 0; k < 3; k++) d[k] = c[k] - p->origin[k];
  if (p->orth) {
    if (nsq(d) <= p->D) { memcpy(out, d, 24); return; }
    for (k = 0; k < 3; k++) {
      if (d[k] < -p->H2[k]) d[k] += p->H[k];
      else if (d[k] > p->H2[k]) d[k] -= p->H[k];
    }
    if (nsq(d) <= p->D) { memcpy(out, d, 24); return; }
    {
      double t0 = p->e1[0] * rint(p->e1r[0] * d[0]);
      double t1 = p->e2[1] * rint(p->e2r[1] * d[1]);
      double t2 = p->e3[2] * rint(p->e3r[2] * d[2]);
      out[0] = d[0] - t0; out[1] = d[1] - t1; out[2] = d[2] - t2;
    }
  } else {
    double r1 = rint(dot3(p->e1r, d)), r2 = rint(dot3(p->e2r, d)),
           r3 = rint(dot3(p->e3r, d));
    for (k = 0; k < 3; k++)
      out[k] = ((d[k] - p->e1[k] * r1) - p->e2[k] * r2) - p->e3[k] * r3;
  }
}

/* ---------------------------------------------------------- cell grid */
typedef struct {
  int n[3];
  double cs[3], inv[3];
  int maxN[3], n1[3], maxN1[3];
} orc_grid;

void orc_grid_init(orc_grid *g, const orc_pbc *p, double cell_size) {
  int k;
  for (k = 0; k < 3; k++) {
    double vd = p->mx[k] - p->mn[k];
    double f = floor(vd / cell_size);
    double cs = vd / (f > 1.0 ? f : 1.0);
    int n;
    g->cs[k] = cs;
    g->inv[k] = 1.0 / cs;
    n = (int)floor(vd / cs + ORC_EPSILON);
    g->n[k] = n > 1 ? n : 1;
    g->maxN[k] = ORC_MAX_INT_2 - ORC_MAX_INT_2 % g->n[k];
    g->n1[k] = -g->n[k] / 2;
    g->maxN1[k] = g->maxN[k] - g->n1[k];
  }
}

static int wrap_index(int c, int n) {
  if (c < 0 || c >= n) {
    c %= n;
    if (c < 0) c += n;
  }
  return c;
}

/* Cell of a position: findCell(origin - min + minimalPosition(x)). */
void orc_cell_of(const orc_pbc *p, const orc_grid *g, const double *x,
                 int *cell3) {
  double mp[3], v;
  int k;
  orc_pbc_min_pos(p, x, mp);
  for (k = 0; k < 3; k++) {
    double delta = p->origin[k] - p->mn[k];
    v = delta + mp[k];
    cell3[k] = wrap_index((int)floor(v * g->inv[k]), g->n[k]);
  }
}

/* Delta list: CellListEnumeratorPeriodicBoundaries.h:46-83 (sorted, no 0). */
int orc_delta_list(const orc_grid *g, double cutoff, int *out, int cap) {
  int nx = (int)(cutoff / g->cs[0] + 1.0 + ORC_EPSILON);
  int ny = (int)(cutoff / g->cs[1] + 1.0 + ORC_EPSILON);
  int nz = (int)(cutoff / g->cs[2] + 1.0 + ORC_EPSILON);
  int n0 = nx < g->n[0] ? nx : g->n[0];
  int n1 = ny < g->n[1] ? ny : g->n[1];
  int n2 = nz < g->n[2] ? nz : g->n[2];
  double xx = g->cs[0] * g->cs[0], yy = g->cs[1] * g->cs[1],
         zz = g->cs[2] * g->cs[2], c2 = cutoff * cutoff;
  int cnt = 0, k, l, m, i, j;
  for (k = -n0; k <= n0; k++) {
    int x = abs(k) - 1;
    double d0;
    if (x < 0) x = 0;
    d0 = x * x * xx;
    for (l = -n1; l <= n1; l++) {
      int y = abs(l) - 1;
      double d1;
      if (y < 0) y = 0;
      d1 = d0 + y * y * yy;
      for (m = -n2; m <= n2; m++) {
        int z = abs(m) - 1;
        if (z < 0) z = 0;
        if (d1 + z * z * zz < c2) {
          int b0 = (k + g->maxN1[0]) % g->n[0] + g->n1[0];
          int b1 = (l + g->maxN1[1]) % g->n[1] + g->n1[1];
          int b2 = (m + g->maxN1[2]) % g->n[2] + g->n1[2];
          int dup = (b0 == 0 && b1 == 0 && b2 == 0);
          for (i = 0; i < cnt && !dup; i++)
            if (out[3 * i] == b0 && out[3 * i + 1] == b1 && out[3 * i + 2] == b2)
              dup = 1;
          if (!dup) {
            if (cnt >= cap) return -1;
            out[3 * cnt] = b0; out[3 * cnt + 1] = b1; out[3 * cnt + 2] = b2;
            cnt++;
          }
        }
      }
    }
  }
  /* sort lexicographically (CubicCellLocation::operator<) */
  for (i = 1; i < cnt; i++) {
    int t0 = out[3 * i], t1 = out[3 * i + 1], t2 = out[3 * i + 2];
    j = i - 1;
    while (j >= 0 && (out[3 * j] > t0 ||
                      (out[3 * j] == t0 && (out[3 * j + 1] > t1 ||
                                            (out[3 * j + 1] == t1 && out[3 * j + 2] > t2))))) {
      out[3 * j + 3] = out[3 * j]; out[3 * j + 4] = out[3 * j + 1];
      out[3 * j + 5] = out[3 * j + 2];
      j--;
    }
    out[3 * j + 3] = t0; out[3 * j + 4] = t1; out[3 * j + 5] = t2;
  }
  return cnt;
}

/* ---------------------------------------------------------- exclusions */
typedef struct {
  int64_t n;
  int64_t *key;  /* a1 * N + a2 (a1 < a2), sorted */
  int8_t *cls;
  int32_t max_delta;
} orc_excl;

static int cmp_rec(const void *a, const void *b) {
  const int64_t *x = (const int64_t *)a, *y = (const int64_t *)b;
  if (x[0] != y[0]) return x[0] < y[0] ? -1 : 1;
  return x[1] < y[1] ? -1 : x[1] > y[1];
}

/* Builds the lookup with last-writer-wins per pair (ExclusionTable.cpp:105-120):
   records are sorted by (pair key, insertion order) and the last record of
   each key gives the class. */
static int excl_build(orc_excl *e, const int32_t *ex, int32_t n_excl, int32_t N) {
  int64_t *rec = (int64_t *)malloc(sizeof(int64_t) * 2 * (n_excl ? n_excl : 1));
  int64_t i, m = 0;
  e->max_delta = 0;
  e->key = (int64_t *)malloc(sizeof(int64_t) * (n_excl ? n_excl : 1));
  e->cls = (int8_t *)malloc(n_excl ? n_excl : 1);
  if (!rec || !e->key || !e->cls) return -1;
  for (i = 0; i < n_excl; i++) {
    int32_t a = ex[3 * i], b = ex[3 * i + 1];
    if (a > b) { int32_t t = a; a = b; b = t; }
    if (b - a > e->max_delta) e->max_delta = b - a;
    rec[2 * i] = (int64_t)a * N + b;
    rec[2 * i + 1] = i;
  }
  qsort(rec, (size_t)n_excl, 2 * sizeof(int64_t), cmp_rec);
  for (i = 0; i < n_excl; i++) {
    int64_t k = rec[2 * i], src = rec[2 * i + 1];
    if (ex[3 * src] == ex[3 * src + 1]) continue; /* self pairs never checked */
    if (m > 0 && e->key[m - 1] == k) {
      e->cls[m - 1] = (int8_t)ex[3 * src + 2];
    } else {
      e->key[m] = k;
      e->cls[m] = (int8_t)ex[3 * src + 2];
      m++;
    }
  }
  e->n = m;
  free(rec);
  return 0;
}

static int excl_check(const orc_excl *e, int32_t a, int32_t b, int32_t N) {
  int64_t lo = 0, hi = e->n - 1, k;
  if (a > b) { int32_t t = a; a = b; b = t; }
  if (b - a > e->max_delta) return ORC_NONE;
  k = (int64_t)a * N + b;
  while (lo <= hi) {
    int64_t mid = (lo + hi) >> 1;
    if (e->key[mid] == k) return e->cls[mid];
    if (e->key[mid] < k) lo = mid + 1; else hi = mid - 1;
  }
  return ORC_NONE;
}

/* ----------------------------------------------------------- functors */
static double poly5(double ar) {
  const double E1 = 0.254829592, E2 = -0.284496736, E3 = 1.421413741,
               E4 = -1.453152027, E5 = 1.061405429, PP = 0.3275911;
  double t = 1.0 / (1.0 + PP * ar);
  return (t) * (E1 + (t) * (E2 + (t) * (E3 + (t) * (E4 + (t) * E5))));
}

typedef struct {
  double switchon2, cutoff2, s1, s2, s3;       /* C2 */
  double c1cut2, c1_15c1, c1_05c3, c1_15c3;     /* C1 */
  int kind;
} orc_switch;

static void sw_init_c2(orc_switch *s, double switchon, double cutoff) {
  double d = cutoff * cutoff - switchon * switchon;
  s->kind = ORC_SW_C2;
  s->switchon2 = switchon * switchon;
  s->cutoff2 = cutoff * cutoff;
  s->s1 = 1.0 / ((d * d) * d);
  s->s2 = cutoff * cutoff - 3.0 * switchon * switchon;
  s->s3 = 4.0 / ((d * d) * d);
}

static void sw_init_c1(orc_switch *s, double cutoff) {
  s->kind = ORC_SW_C1;
  s->cutoff2 = cutoff * cutoff;
  s->c1cut2 = cutoff * cutoff;
  s->c1_15c1 = 1.5 / cutoff;
  s->c1_05c3 = 0.5 / (cutoff * cutoff * cutoff);
  s->c1_15c3 = 1.5 / (cutoff * cutoff * cutoff);
}

static void sw_eval(const orc_switch *s, double d2, double *value, double *deriv) {
  if (s->kind == ORC_SW_C2) {
    *deriv = 0.0;
    if (d2 > s->cutoff2) *value = 0.0;
    else if (d2 >= s->switchon2) {
      double c2 = s->cutoff2 - d2;
      double c4 = c2 * (s->s2 + 2.0 * d2);
      *value = s->s1 * (c2 * c4);
      *deriv = s->s3 * (c2 * c2 - c4);
    } else *value = 1.0;
  } else {
    double dist;
    if (d2 > s->c1cut2) { *value = 0.0; *deriv = 0.0; return; }
    dist = sqrt(d2);
    *value = 1.0 - dist * (s->c1_15c1 - d2 * s->c1_05c3);
    *deriv = -s->c1_15c1 / dist + dist * s->c1_15c3;
  }
}

/* ----------------------------------------------------------- evaluator */
typedef struct {
  const orc_system *sys;
  orc_pbc pbc;
  orc_excl ex;
  orc_switch sw_lj, sw_el;
  double cut2;        /* OneAtomPairTwo::mySquaredCutoff */
  double two_alpha_pi;
  double *f;          /* forces += */
  double e_lj, e_el;
  double vir[9], molvir[9];
  int32_t *pairs;     /* optional record (i, j, class) */
  int64_t npairs, pair_cap;
} orc_eval;

static void one_pair(orc_eval *ev, int i, int j) {
  const orc_system *s = ev->sys;
  double diff[3], d2, rd2, e1 = 0.0, f1 = 0.0, e2 = 0.0, f2 = 0.0, fij[3];
  int mi, mj, same, excl = ORC_NONE, k;
  orc_pbc_min_diff(&ev->pbc, s->pos + 3 * i, s->pos + 3 * j, diff, &d2);
  if (d2 > ev->cut2) return;
  mi = s->mol[i]; mj = s->mol[j];
  same = (mi == mj);
  if (same) excl = excl_check(&ev->ex, i, j, s->n_atoms);
  if (excl == ORC_FULL) return;
  rd2 = 1.0 / d2;
  { /* LennardJonesForce::operator() */
    const double *P = s->lj + 4 * (s->type[i] * s->n_types + s->type[j]);
    double A = excl != ORC_MODIFIED ? P[0] : P[2];
    double B = excl != ORC_MODIFIED ? P[1] : P[3];
    double r6 = rd2 * rd2 * rd2, r12 = r6 * r6, r6B = B * r6, r12A = A * r12;
    e1 = r12A - r6B;
    f1 = 12.0 * r12A * rd2 - 6.0 * r6B * rd2;
  }
  if (s->elec == ORC_ELEC_COULOMB) { /* CoulombForce::operator() */
    e2 = s->q[i] * s->q[j] * sqrt(rd2);
    if (excl == ORC_MODIFIED) e2 *= s->csf;
    f2 = e2 * rd2;
  } else if (s->elec == ORC_ELEC_EWALD) { /* real-space Ewald, A&S */
    double r = sqrt(d2), qq = s->q[i] * s->q[j], rr, ar, e;
    if (excl == ORC_MODIFIED) qq *= s->csf;
    rr = 1.0 / r;
    ar = s->alpha * r;
    e = qq * exp(-ar * ar);
    e2 = poly5(ar) * e * rr;
    f2 = ((e2 + ev->two_alpha_pi * e) * rr * rr);
  }
  {
    double sv, sd;
    sw_eval(&ev->sw_lj, d2, &sv, &sd);
    f1 = f1 * sv - e1 * sd;
    e1 = e1 * sv;
    sw_eval(&ev->sw_el, d2, &sv, &sd);
    f2 = f2 * sv - e2 * sd;
    e2 = e2 * sv;
  }
  ev->e_lj += e1;
  ev->e_el += e2;
  {
    double ff = f1 + f2;
    for (k = 0; k < 3; k++) fij[k] = diff[k] * ff;
  }
  for (k = 0; k < 3; k++) ev->f[3 * i + k] -= fij[k];
  for (k = 0; k < 3; k++) ev->f[3 * j + k] += fij[k];
  if (!same && (s->flags & 2)) {
    double cd[3], cd2, xy, xz, yz;
    orc_pbc_min_diff(&ev->pbc, s->com + 3 * mi, s->com + 3 * mj, cd, &cd2);
    xy = fij[0] * diff[1]; xz = fij[0] * diff[2]; yz = fij[1] * diff[2];
    ev->vir[0] += fij[0] * diff[0]; ev->vir[1] += xy; ev->vir[2] += xz;
    ev->vir[3] += xy; ev->vir[4] += fij[1] * diff[1]; ev->vir[5] += yz;
    ev->vir[6] += xz; ev->vir[7] += yz; ev->vir[8] += fij[2] * diff[2];
    xy = fij[0] * cd[1]; xz = fij[0] * cd[2]; yz = fij[1] * cd[2];
    ev->molvir[0] += fij[0] * cd[0]; ev->molvir[1] += xy; ev->molvir[2] += xz;
    ev->molvir[3] += xy; ev->molvir[4] += fij[1] * cd[1]; ev->molvir[5] += yz;
    ev->molvir[6] += xz; ev->molvir[7] += yz; ev->molvir[8] += fij[2] * cd[2];
  } else if (s->flags & 1) {
    double xy = fij[0] * diff[1], xz = fij[0] * diff[2], yz = fij[1] * diff[2];
    ev->vir[0] += fij[0] * diff[0]; ev->vir[1] += xy; ev->vir[2] += xz;
    ev->vir[3] += xy; ev->vir[4] += fij[1] * diff[1]; ev->vir[5] += yz;
    ev->vir[6] += xz; ev->vir[7] += yz; ev->vir[8] += fij[2] * diff[2];
  }
  if (ev->pairs) {
    if (ev->npairs < ev->pair_cap) {
      int a = i < j ? i : j, b = i < j ? j : i;
      ev->pairs[3 * ev->npairs] = a;
      ev->pairs[3 * ev->npairs + 1] = b;
      ev->pairs[3 * ev->npairs + 2] = excl;
    }
  }
  ev->npairs++;
}

static int eval_init(orc_eval *ev, const orc_system *s, double *forces,
                     int32_t *pairs, int64_t pair_cap) {
  memset(ev, 0, sizeof(*ev));
  ev->sys = s;
  ev->f = forces;
  ev->pairs = pairs;
  ev->pair_cap = pair_cap;
  orc_pbc_set(&ev->pbc, s->e1, s->e2, s->e3, s->origin);
  if (excl_build(&ev->ex, s->excl, s->n_excl, s->n_atoms)) return -1;
  sw_init_c2(&ev->sw_lj, s->lj_switchon, s->lj_cutoff);
  if (s->el_switch == ORC_SW_C1) sw_init_c1(&ev->sw_el, s->el_cutoff);
  else sw_init_c2(&ev->sw_el, s->el_switchon, s->el_cutoff);
  ev->cut2 = ev->sw_lj.cutoff2 > ev->sw_el.cutoff2 ? ev->sw_lj.cutoff2
                                                   : ev->sw_el.cutoff2;
  ev->two_alpha_pi = 2.0 * s->alpha / sqrt(ORC_PI);
  return 0;
}

static void eval_finish(orc_eval *ev, double *energies2, double *vir9,
                        double *molvir9, int64_t *npairs) {
  int k;
  if (energies2) { energies2[0] += ev->e_lj; energies2[1] += ev->e_el; }
  if (vir9) for (k = 0; k < 9; k++) vir9[k] += ev->vir[k];
  if (molvir9) for (k = 0; k < 9; k++) molvir9[k] += ev->molvir[k];
  if (npairs) *npairs = ev->npairs;
  free(ev->ex.key);
  free(ev->ex.cls);
}

/*
 * Cell-list evaluation in the reference's enumeration order:
 * NonbondedCutoffSystemForce::evaluate -> updateCellLists -> enumerator ->
 * doEvaluate.  forces/energies/virials are accumulated (+=).
 * Returns 0 on success, -1 on allocation failure.
 */
int orc_evaluate(const orc_system *s, double *forces, double *energies2,
                 double *vir9, double *molvir9, int32_t *pairs,
                 int64_t pair_cap, int64_t *npairs) {
  orc_eval ev;
  orc_grid g;
  int N = s->n_atoms, ncell, i, c, nd, *delta, *head, *next;
  int *cell_lin;
  if (eval_init(&ev, s, forces, pairs, pair_cap)) return -1;
  orc_grid_init(&g, &ev.pbc, s->cell_size);
  ncell = g.n[0] * g.n[1] * g.n[2];
  head = (int *)malloc(sizeof(int) * ncell);
  next = (int *)malloc(sizeof(int) * (N ? N : 1));
  cell_lin = (int *)malloc(sizeof(int) * (N ? N : 1));
  delta = (int *)malloc(sizeof(int) * 3 * (2 * g.n[0] + 1) * (2 * g.n[1] + 1) *
                        (2 * g.n[2] + 1));
  if (!head || !next || !cell_lin || !delta) return -1;
  for (c = 0; c < ncell; c++) head[c] = -1;
  /* Topology.h:67-82: descending i, push front -> ascending lists */
  for (i = N - 1; i >= 0; i--) {
    int c3[3], lin;
    orc_cell_of(&ev.pbc, &g, s->pos + 3 * i, c3);
    lin = (c3[0] * g.n[1] + c3[1]) * g.n[2] + c3[2];
    cell_lin[i] = lin;
    next[i] = head[lin];
    head[lin] = i;
  }
  nd = orc_delta_list(&g, s->alg_cutoff, delta,
                      (2 * g.n[0] + 1) * (2 * g.n[1] + 1) * (2 * g.n[2] + 1));
  /* enumerator walk: every non-empty cell in linear order, self first,
     then each delta (sorted) whose periodic cell is non-empty and whose
     linear index is not below this one */
  for (c = 0; c < ncell; c++) {
    int cx, cy, cz, d, a, b;
    if (head[c] < 0) continue;
    cx = c / (g.n[1] * g.n[2]);
    cy = (c / g.n[2]) % g.n[1];
    cz = c % g.n[2];
    for (a = head[c]; a != -1; a = next[a])
      for (b = next[a]; b != -1; b = next[b]) one_pair(&ev, a, b);
    for (d = 0; d < nd; d++) {
      int px = (cx + delta[3 * d] + g.maxN[0]) % g.n[0];
      int py = (cy + delta[3 * d + 1] + g.maxN[1]) % g.n[1];
      int pz = (cz + delta[3 * d + 2] + g.maxN[2]) % g.n[2];
      int cj = (px * g.n[1] + py) * g.n[2] + pz;
      if (head[cj] < 0 || cj < c) continue;
      for (a = head[c]; a != -1; a = next[a])
        for (b = head[cj]; b != -1; b = next[b]) one_pair(&ev, a, b);
    }
  }
  free(head); free(next); free(cell_lin); free(delta);
  eval_finish(&ev, energies2, vir9, molvir9, npairs);
  return 0;
}

/*
 * orc_evaluate spread over host threads, for the CPU baseline's all-core
 * figure (the reference itself has no threading): the same cell lists,
 * delta list and per-pair function; thread t walks the cells c with
 * c % nthreads == t (c and its delta partners, as the serial walk does)
 * into a private force array and energy/virial sums, then the arrays are
 * added in thread order.  Pair records are not kept.  Returns the number
 * of threads used, -1 on allocation failure.
 */
int orc_evaluate_threads(const orc_system *s, double *forces,
                         double *energies2, double *vir9, int nthreads,
                         int64_t *npairs) {
#ifdef _OPENMP
  orc_eval ev0;
  orc_grid g;
  int N = s->n_atoms, ncell, i, c, nd, *delta, *head, *next, T;
  double *fbuf, *sums;
  int64_t np = 0;
  if (eval_init(&ev0, s, forces, NULL, 0)) return -1;
  orc_grid_init(&g, &ev0.pbc, s->cell_size);
  ncell = g.n[0] * g.n[1] * g.n[2];
  T = nthreads > 0 ? nthreads : omp_get_max_threads();
  head = (int *)malloc(sizeof(int) * ncell);
  next = (int *)malloc(sizeof(int) * (N ? N : 1));
  delta = (int *)malloc(sizeof(int) * 3 * (2 * g.n[0] + 1) * (2 * g.n[1] + 1) *
                        (2 * g.n[2] + 1));
  fbuf = (double *)calloc((size_t)T * 3 * (N ? N : 1), sizeof(double));
  sums = (double *)calloc((size_t)T * 11, sizeof(double));
  if (!head || !next || !delta || !fbuf || !sums) return -1;
  for (c = 0; c < ncell; c++) head[c] = -1;
  for (i = N - 1; i >= 0; i--) {
    int c3[3], lin;
    orc_cell_of(&ev0.pbc, &g, s->pos + 3 * i, c3);
    lin = (c3[0] * g.n[1] + c3[1]) * g.n[2] + c3[2];
    next[i] = head[lin];
    head[lin] = i;
  }
  nd = orc_delta_list(&g, s->alg_cutoff, delta,
                      (2 * g.n[0] + 1) * (2 * g.n[1] + 1) * (2 * g.n[2] + 1));
#pragma omp parallel num_threads(T) reduction(+ : np)
  {
    const int t = omp_get_thread_num();
    orc_eval ev = ev0;
    int cc, k;
    ev.f = fbuf + (size_t)t * 3 * (N ? N : 1);
    ev.e_lj = ev.e_el = 0.0;
    for (k = 0; k < 9; k++) ev.vir[k] = ev.molvir[k] = 0.0;
    ev.npairs = 0;
    for (cc = t; cc < ncell; cc += T) {
      int cx, cy, cz, d, a, b;
      if (head[cc] < 0) continue;
      cx = cc / (g.n[1] * g.n[2]);
      cy = (cc / g.n[2]) % g.n[1];
      cz = cc % g.n[2];
      for (a = head[cc]; a != -1; a = next[a])
        for (b = next[a]; b != -1; b = next[b]) one_pair(&ev, a, b);
      for (d = 0; d < nd; d++) {
        int px = (cx + delta[3 * d] + g.maxN[0]) % g.n[0];
        int py = (cy + delta[3 * d + 1] + g.maxN[1]) % g.n[1];
        int pz = (cz + delta[3 * d + 2] + g.maxN[2]) % g.n[2];
        int cj = (px * g.n[1] + py) * g.n[2] + pz;
        if (head[cj] < 0 || cj < cc) continue;
        for (a = head[cc]; a != -1; a = next[a])
          for (b = head[cj]; b != -1; b = next[b]) one_pair(&ev, a, b);
      }
    }
    sums[11 * t] = ev.e_lj;
    sums[11 * t + 1] = ev.e_el;
    for (k = 0; k < 9; k++) sums[11 * t + 2 + k] = ev.vir[k];
    np += ev.npairs;
  }
  for (c = 0; c < T; c++) {
    const double *fb = fbuf + (size_t)c * 3 * (N ? N : 1);
    for (i = 0; i < 3 * N; i++) forces[i] += fb[i];
    if (energies2) {
      energies2[0] += sums[11 * c];
      energies2[1] += sums[11 * c + 1];
    }
    if (vir9)
      for (i = 0; i < 9; i++) vir9[i] += sums[11 * c + 2 + i];
  }
  if (npairs) *npairs = np;
  free(head); free(next); free(delta); free(fbuf); free(sums);
  free(ev0.ex.key);
  free(ev0.ex.cls);
  return T;
#else
  (void)s; (void)forces; (void)energies2; (void)vir9; (void)nthreads;
  (void)npairs;
  return -1;
#endif
}

/*
 * O(N^2) evaluation over all i<j (NonbondedSimpleFullSystemForce.h:40-44
 * semantics) with the same per-pair function: defines the pair set
 * independently of any cell structure.
 */
int orc_evaluate_bruteforce(const orc_system *s, double *forces,
                            double *energies2, double *vir9, double *molvir9,
                            int32_t *pairs, int64_t pair_cap, int64_t *npairs) {
  orc_eval ev;
  int i, j;
  if (eval_init(&ev, s, forces, pairs, pair_cap)) return -1;
  for (i = 0; i < s->n_atoms; i++)
    for (j = i + 1; j < s->n_atoms; j++) one_pair(&ev, i, j);
  eval_finish(&ev, energies2, vir9, molvir9, npairs);
  return 0;
}

/* Grid and delta list as the reference enumerator would build them. */
int orc_grid_info(const orc_system *s, int32_t *n3, double *cs3,
                  int32_t *deltas, int32_t cap) {
  orc_pbc p;
  orc_grid g;
  orc_pbc_set(&p, s->e1, s->e2, s->e3, s->origin);
  orc_grid_init(&g, &p, s->cell_size);
  n3[0] = g.n[0]; n3[1] = g.n[1]; n3[2] = g.n[2];
  cs3[0] = g.cs[0]; cs3[1] = g.cs[1]; cs3[2] = g.cs[2];
  return orc_delta_list(&g, s->alg_cutoff, deltas, cap);
}

/* Cell (x, y, z) of each atom as the reference's updateCellLists computes it
   (with the face wrap described above). */
void orc_cells(const orc_system *s, int32_t *cell3) {
  orc_pbc p;
  orc_grid g;
  int i;
  orc_pbc_set(&p, s->e1, s->e2, s->e3, s->origin);
  orc_grid_init(&g, &p, s->cell_size);
  for (i = 0; i < s->n_atoms; i++) orc_cell_of(&p, &g, s->pos + 3 * i, cell3 + 3 * i);
}

/* Single minimal difference (exposed for property tests). */
void orc_min_diff(const orc_system *s, const double *c1, const double *c2,
                  double *diff, double *d2) {
  orc_pbc p;
  orc_pbc_set(&p, s->e1, s->e2, s->e3, s->origin);
  orc_pbc_min_diff(&p, c1, c2, diff, d2);
}

double orc_poly5(double ar) { return poly5(ar); }

void orc_switch_c2(double switchon, double cutoff, double d2, double *v, double *dv) {
  orc_switch s;
  sw_init_c2(&s, switchon, cutoff);
  sw_eval(&s, d2, v, dv);
}

/* ------------------------------------------------------------ leapfrog
 * LeapfrogIntegrator::run with the SHAKE (post-drift) and RATTLE
 * (post-step) modifiers, restated operation for operation:
 *   integrator/leapfrog/LeapfrogIntegrator.cpp:33-99   run, doHalfKickdoDrift,
 *                                                      doKickdoDrift
 *   integrator/StandardIntegrator.cpp:91-116           doHalfKick, doKick
 *   integrator/STSIntegrator.cpp:28-51                 doDrift, calculateForces
 *   modifier/ModifierShake.cpp:16-89                   SHAKE sweeps
 *   modifier/ModifierRattle.cpp:17-71                  RATTLE sweeps
 *   modifier/ModifierMetaShake.cpp:11-33, ModifierMetaRattle.cpp:11-30  errors
 *   module/ModifierModule.cpp:104-127                  which hook each uses
 *   force/bonded/BondSystemForce.h:89-122              harmonic bond
 *   force/bonded/AngleSystemForce.h:86-150             angle + Urey-Bradley
 * The constraint virial (ScalarStructure::addVirial in both modifiers) is
 * not accumulated: calculateForces clears the energies after every SHAKE
 * and every RATTLE but the last one.
 */
#define ORC_TIMEFACTOR 48.88821290839616       /* base/PMConstants.cpp:50 */
#define ORC_INV_TIMEFACTOR 0.02045482828087295 /* base/PMConstants.cpp:51 */

typedef struct {
  int32_t n_cons;
  const int32_t *cons;  /* 2 per constraint: atom1, atom2, in list order */
  const double *rest;   /* rest length per constraint */
  const double *mass;   /* scaledMass per atom */
  double timestep;      /* fs */
  double shake_eps, rattle_eps;
  int32_t shake_max_iter, rattle_max_iter;
  int32_t shake, rattle; /* modifier present (1) or not (0) */
  /* optional bonded terms evaluated with the cutoff force */
  int32_t n_bonds, n_angles;
  const int32_t *bonds;       /* 2 per bond */
  const double *bond_k, *bond_r0;
  const int32_t *angles;      /* 3 per angle */
  const double *angle_k, *angle_t0, *ub_k, *ub_r0;  /* t0 in radians */
  /* optional further force of the force group (k-space terms in the
     tests), called after the cutoff force: forces += , *e_el += */
  void (*extra)(const double *x, double *f, double *e_el, void *user);
  void *extra_user;
} orc_lf;

typedef struct {
  const orc_system *s;
  const orc_lf *lf;
  orc_pbc pbc;
  double *x, *v, *f, *last, *energies2, *bonded2;
  int32_t *iters;
  int64_t iters_cap, n_calls;
} orc_md;

static void md_record(orc_md *m, int iters) {
  if (m->iters && m->n_calls < m->iters_cap) m->iters[m->n_calls] = iters;
  m->n_calls++;
}

/* ModifierMetaShake::calcError */
static double shake_error(const orc_md *m) {
  const orc_lf *lf = m->lf;
  double error = 0;
  int i, k;
  for (i = 0; i < lf->n_cons; i++) {
    int a1 = lf->cons[2 * i], a2 = lf->cons[2 * i + 1];
    double r2 = lf->rest[i] * lf->rest[i], d[3], n2;
    for (k = 0; k < 3; k++) d[k] = m->x[3 * a2 + k] - m->x[3 * a1 + k];
    n2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    error += fabs(n2 - r2) / (2.0 * r2);
  }
  return error / lf->n_cons;
}

/* ModifierMetaRattle::calcError */
static double rattle_error(const orc_md *m) {
  const orc_lf *lf = m->lf;
  double error = 0;
  int i, k;
  for (i = 0; i < lf->n_cons; i++) {
    int a1 = lf->cons[2 * i], a2 = lf->cons[2 * i + 1];
    double vab[3], rab[3];
    for (k = 0; k < 3; k++) {
      vab[k] = m->v[3 * a1 + k] - m->v[3 * a2 + k];
      rab[k] = m->x[3 * a1 + k] - m->x[3 * a2 + k];
    }
    error += fabs(rab[0] * vab[0] + rab[1] * vab[1] + rab[2] * vab[2]);
  }
  return error / lf->n_cons;
}

/* ModifierShake::doExecute (virial term left out, see above) */
static void md_shake(orc_md *m) {
  const orc_lf *lf = m->lf;
  double error, dt = lf->timestep / ORC_TIMEFACTOR;
  int iter = 0, i, k;
  if (!lf->shake || lf->n_cons == 0) return;
  error = shake_error(m);
  while (error > lf->shake_eps) {
    for (i = 0; i < lf->n_cons; i++) {
      int a1 = lf->cons[2 * i], a2 = lf->cons[2 * i + 1];
      double pab[3], rab[3], dp[3], pabsq, rabsq, diffsq, rpab, rM1, rM2, gab;
      for (k = 0; k < 3; k++) pab[k] = m->x[3 * a1 + k] - m->x[3 * a2 + k];
      pabsq = pab[0] * pab[0] + pab[1] * pab[1] + pab[2] * pab[2];
      rabsq = lf->rest[i] * lf->rest[i];
      diffsq = rabsq - pabsq;
      for (k = 0; k < 3; k++) rab[k] = m->last[3 * a1 + k] - m->last[3 * a2 + k];
      rpab = rab[0] * pab[0] + rab[1] * pab[1] + rab[2] * pab[2];
      rM1 = 1 / lf->mass[a1];
      rM2 = 1 / lf->mass[a2];
      gab = diffsq / (2 * (rM1 + rM2) * rpab);
      for (k = 0; k < 3; k++) {
        dp[k] = rab[k] * gab;
        m->x[3 * a1 + k] += dp[k] * rM1;
        m->x[3 * a2 + k] -= dp[k] * rM2;
        dp[k] /= dt;
        m->v[3 * a1 + k] += dp[k] * rM1;
        m->v[3 * a2 + k] -= dp[k] * rM2;
      }
    }
    error = shake_error(m);
    iter++;
    if (iter > lf->shake_max_iter) break;
  }
  md_record(m, iter);
  memcpy(m->last, m->x, sizeof(double) * 3 * m->s->n_atoms);
}

/* ModifierRattle::doExecute (virial term left out, see above) */
static void md_rattle(orc_md *m) {
  const orc_lf *lf = m->lf;
  double error, dt = lf->timestep / ORC_TIMEFACTOR;
  int iter = 0, i, k;
  if (!lf->rattle || lf->n_cons == 0) return;
  error = rattle_error(m);
  while (error > lf->rattle_eps) {
    for (i = 0; i < lf->n_cons; i++) {
      int a1 = lf->cons[2 * i], a2 = lf->cons[2 * i + 1];
      double rab[3], vab[3], dp[3], rabsq, rvab, rM1, rM2, gab;
      rM1 = 1 / lf->mass[a1];
      rM2 = 1 / lf->mass[a2];
      for (k = 0; k < 3; k++) rab[k] = m->x[3 * a1 + k] - m->x[3 * a2 + k];
      rabsq = rab[0] * rab[0] + rab[1] * rab[1] + rab[2] * rab[2];
      for (k = 0; k < 3; k++) vab[k] = m->v[3 * a1 + k] - m->v[3 * a2 + k];
      rvab = rab[0] * vab[0] + rab[1] * vab[1] + rab[2] * vab[2];
      gab = -rvab / (dt * (rM1 + rM2) * rabsq);
      for (k = 0; k < 3; k++) {
        dp[k] = rab[k] * gab;
        m->v[3 * a1 + k] += dp[k] * dt * rM1;
        m->v[3 * a2 + k] -= dp[k] * dt * rM2;
      }
    }
    error = rattle_error(m);
    iter++;
    if (iter > lf->rattle_max_iter) break;
  }
  md_record(m, iter);
}

/* v += F * w / m  (w = h/2 for a half kick, h for a full kick) */
static void md_kick(orc_md *m, double w) {
  int i, k;
  for (i = 0; i < m->s->n_atoms; i++)
    for (k = 0; k < 3; k++)
      m->v[3 * i + k] += m->f[3 * i + k] * w / m->lf->mass[i];
}

/* x += v * h */
static void md_drift(orc_md *m) {
  int i;
  double h = m->lf->timestep * ORC_INV_TIMEFACTOR;
  for (i = 0; i < 3 * m->s->n_atoms; i++) m->x[i] += m->v[i] * h;
}

static double norm3(const double *a) {
  return sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
}

/* BondSystemForce::calcBond and AngleSystemForce::calcAngle over all terms
   (forces += , bonded2[0] += bond energy, bonded2[1] += angle energy). */
static void md_bonded(orc_md *m) {
  const orc_lf *lf = m->lf;
  double *x = m->x, *f = m->f;
  int i, k;
  for (i = 0; i < lf->n_bonds; i++) {
    int a1 = lf->bonds[2 * i], a2 = lf->bonds[2 * i + 1];
    double r12[3], d2, r, dpotdr, f1[3];
    orc_pbc_min_diff(&m->pbc, x + 3 * a2, x + 3 * a1, r12, &d2);
    r = norm3(r12);
    dpotdr = 2.0 * lf->bond_k[i] * (r - lf->bond_r0[i]);
    for (k = 0; k < 3; k++) {
      f1[k] = r12[k] * (-dpotdr / r);
      f[3 * a1 + k] += f1[k];
      f[3 * a2 + k] -= f1[k];
    }
    m->bonded2[0] += lf->bond_k[i] * (r - lf->bond_r0[i]) * (r - lf->bond_r0[i]);
  }
  for (i = 0; i < lf->n_angles; i++) {
    int a1 = lf->angles[3 * i], a2 = lf->angles[3 * i + 1],
        a3 = lf->angles[3 * i + 2];
    double t0 = lf->angle_t0[i], ka = lf->angle_k[i];
    double kub = lf->ub_k ? lf->ub_k[i] : 0.0, rub = lf->ub_r0 ? lf->ub_r0[i] : 0.0;
    double r12[3], r32[3], r13[3], d2, d12, d32, d13, c[3], theta, sn, cs,
        dpot, dr12[3], dr32[3], dr13[3], dt1[3], dt3[3], ub1[3], f1[3], f3[3];
    orc_pbc_min_diff(&m->pbc, x + 3 * a2, x + 3 * a1, r12, &d2);
    orc_pbc_min_diff(&m->pbc, x + 3 * a2, x + 3 * a3, r32, &d2);
    orc_pbc_min_diff(&m->pbc, x + 3 * a3, x + 3 * a1, r13, &d2);
    d12 = norm3(r12);
    d32 = norm3(r32);
    d13 = norm3(r13);
    cross3(r12, r32, c);
    theta = atan2(norm3(c), dot3(r12, r32));
    sn = sin(theta);
    cs = cos(theta);
    dpot = 2.0 * ka * (theta - t0);
    for (k = 0; k < 3; k++) {
      dr12[k] = r12[k] / d12;
      dr32[k] = r32[k] / d32;
      dr13[k] = r13[k] / d13;
    }
    for (k = 0; k < 3; k++) {
      dt1[k] = (dr12[k] * cs - dr32[k]) / (sn * d12);
      dt3[k] = (dr32[k] * cs - dr12[k]) / (sn * d32);
      ub1[k] = dr13[k] * (2.0 * kub * (d13 - rub));
      f1[k] = -dt1[k] * dpot - ub1[k];
      f3[k] = -dt3[k] * dpot - (-ub1[k]);
      f[3 * a1 + k] += f1[k];
      f[3 * a2 + k] += -f1[k] - f3[k];
      f[3 * a3 + k] += f3[k];
    }
    m->bonded2[1] += ka * (theta - t0) * (theta - t0) +
                     kub * (d13 - rub) * (d13 - rub);
  }
}

/* StandardIntegrator::calculateForces + STSIntegrator::calculateForces:
   energies and forces cleared, then the bonded terms and this force. */
static int md_forces(orc_md *m) {
  orc_system sys = *m->s;
  memset(m->f, 0, sizeof(double) * 3 * sys.n_atoms);
  m->energies2[0] = m->energies2[1] = 0;
  m->bonded2[0] = m->bonded2[1] = 0;
  md_bonded(m);
  sys.pos = m->x;
  if (orc_evaluate(&sys, m->f, m->energies2, NULL, NULL, NULL, 0, NULL))
    return -1;
  if (m->lf->extra)
    m->lf->extra(m->x, m->f, m->energies2 + 1, m->lf->extra_user);
  return 0;
}

/*
 * LeapfrogIntegrator::run(n_steps).  x, v: positions and velocities (in
 * and out); f: the forces at x on entry (initializeForces), at the final x
 * on exit; last: SHAKE's previous positions (ModifierMetaRattleShake keeps
 * them across calls; initialise them to x before the first call);
 * energies2: LJ and electrostatic energy of the last force evaluation;
 * bonded2: its bond and angle energies (the terms in lf, may be none).
 * iters (optional): the sweep count of every SHAKE and RATTLE call in call
 * order.  Returns the number of modifier calls, or -1 on failure.
 */
int64_t orc_leapfrog_run(const orc_system *s, const orc_lf *lf, double *x,
                         double *v, double *f, double *last, double *energies2,
                         double *bonded2, int32_t n_steps, int32_t *iters,
                         int64_t iters_cap) {
  orc_md m;
  double h = lf->timestep * ORC_INV_TIMEFACTOR;
  double h2 = 0.5 * lf->timestep * ORC_INV_TIMEFACTOR;
  int i;
  m.s = s; m.lf = lf; m.x = x; m.v = v; m.f = f; m.last = last;
  m.energies2 = energies2; m.bonded2 = bonded2; m.iters = iters;
  m.iters_cap = iters_cap; m.n_calls = 0;
  orc_pbc_set(&m.pbc, s->e1, s->e2, s->e3, s->origin);
  if (n_steps < 1) return 0;
  /* doHalfKickdoDrift (no pre-drift modifier): v += F*h*0.5/m, x += v*h */
  {
    int a, k;
    for (a = 0; a < s->n_atoms; a++)
      for (k = 0; k < 3; k++)
        v[3 * a + k] += f[3 * a + k] * h * 0.5 / lf->mass[a];
  }
  md_drift(&m);
  md_shake(&m);
  if (md_forces(&m)) return -1;
  for (i = 1; i < n_steps; i++) {
    if (lf->rattle) {   /* doKickdoDrift with a post-step modifier */
      md_kick(&m, h2);
      md_rattle(&m);
      md_kick(&m, h2);
    } else {
      md_kick(&m, h);
    }
    md_drift(&m);
    md_shake(&m);
    if (md_forces(&m)) return -1;
  }
  md_kick(&m, h2);
  md_rattle(&m);
  return m.n_calls;
}

/* calculateForces alone (forces and energies cleared, then the bonded terms
   of lf and the cutoff force at x): the forces initializeForces gives. */
int orc_lf_forces(const orc_system *s, const orc_lf *lf, double *x, double *f,
                  double *energies2, double *bonded2) {
  orc_md m;
  memset(&m, 0, sizeof m);
  m.s = s; m.lf = lf; m.x = x; m.f = f;
  m.energies2 = energies2; m.bonded2 = bonded2;
  orc_pbc_set(&m.pbc, s->e1, s->e2, s->e3, s->origin);
  return md_forces(&m);
}

/* ------------------------------------------------------- Ewald k-space
 * Reciprocal and correction terms of NonbondedFullEwaldSystemForce
 * (force/nonbonded/NonbondedFullEwaldSystemForce.h) for an orthogonal
 * periodic box, restated in its operation order:
 *   :597-652   alpha default, rc, kc, k-vector list (h >= 0 half space)
 *   :680-700   point self energy, charged system energy
 *   :884-1086  reciprocal term with the sin/cos tables built by the
 *              addition theorem, energy, atomic virial, forces
 *   :1092-1130 intramolecular correction over the exclusion table
 * All terms go into the electrostatic energy, as evaluate does (:376-384).
 */
typedef struct {
  double alpha, alpha2, alpha2r, two_alpha_pi, rc, kc, fac, V, Vr;
  double L[3], Lr[3];
  int hmax, kmax, lmax, hklmax, nk;
  int *hkl;         /* 3 per k-vector */
  double *k, *k2;   /* 3 per k-vector, |k|^2 */
  double self_energy, charged_energy;
} orc_ewald;

/* alpha <= 0 selects the default alpha of initialize (:603-604). */
int orc_ewald_init(orc_ewald *ew, const orc_system *s, double alpha,
                   double accuracy) {
  orc_pbc p;
  double t_rate = 5.5, pp, kc2, q, q2;
  int h, k, l, i, cap;
  orc_pbc_set(&p, s->e1, s->e2, s->e3, s->origin);
  memset(ew, 0, sizeof(*ew));
  ew->L[0] = s->e1[0]; ew->L[1] = s->e2[1]; ew->L[2] = s->e3[2];
  ew->V = p.V;
  ew->Lr[0] = p.e1r[0]; ew->Lr[1] = p.e2r[1]; ew->Lr[2] = p.e3r[2];
  ew->Vr = 1.0 / ew->V;
  ew->alpha = alpha > 0 ? alpha
                        : sqrt(ORC_PI) * pow(t_rate * s->n_atoms / (ew->V * ew->V),
                                             1.0 / 6.0);
  ew->alpha2 = ew->alpha * ew->alpha;
  ew->alpha2r = 1.0 / ew->alpha2;
  ew->two_alpha_pi = 2.0 * ew->alpha / sqrt(ORC_PI);
  pp = -log(accuracy);
  ew->rc = sqrt(pp) / ew->alpha;
  ew->kc = 2.0 * ew->alpha * sqrt(pp);
  ew->fac = 1.0 / (4.0 * ew->alpha * ew->alpha);
  ew->hmax = (int)floor(ew->kc / (2.0 * ORC_PI) * ew->L[0]);
  ew->kmax = (int)floor(ew->kc / (2.0 * ORC_PI) * ew->L[1]);
  ew->lmax = (int)floor(ew->kc / (2.0 * ORC_PI) * ew->L[2]);
  ew->hklmax = ew->hmax > ew->kmax ? ew->hmax : ew->kmax;
  if (ew->lmax > ew->hklmax) ew->hklmax = ew->lmax;
  ew->hklmax += 1;
  if (ew->hklmax < 2) ew->hklmax = 2;
  cap = (ew->hmax + 1) * (2 * ew->kmax + 1) * (2 * ew->lmax + 1);
  ew->hkl = (int *)malloc(sizeof(int) * 3 * (cap > 0 ? cap : 1));
  ew->k = (double *)malloc(sizeof(double) * 3 * (cap > 0 ? cap : 1));
  ew->k2 = (double *)malloc(sizeof(double) * (cap > 0 ? cap : 1));
  if (!ew->hkl || !ew->k || !ew->k2) return -1;
  kc2 = ew->kc * ew->kc;
  for (h = 0; h <= ew->hmax; h++) {
    double kx = 2.0 * ORC_PI * h * ew->Lr[0];
    for (k = (h == 0 ? 0 : -ew->kmax); k <= ew->kmax; k++) {
      double ky = 2.0 * ORC_PI * k * ew->Lr[1];
      for (l = (h == 0 && k == 0 ? 1 : -ew->lmax); l <= ew->lmax; l++) {
        double kz = 2.0 * ORC_PI * l * ew->Lr[2];
        if (kx * kx + ky * ky + kz * kz < kc2) {
          int n = ew->nk++;
          ew->hkl[3 * n] = h; ew->hkl[3 * n + 1] = k; ew->hkl[3 * n + 2] = l;
          ew->k[3 * n] = kx; ew->k[3 * n + 1] = ky; ew->k[3 * n + 2] = kz;
          ew->k2[n] = kx * kx + ky * ky + kz * kz;
        }
      }
    }
  }
  q2 = 0.0;
  for (i = 0; i < s->n_atoms; i++) q2 += s->q[i] * s->q[i];
  ew->self_energy = -q2 * ew->alpha / sqrt(ORC_PI);
  q = 0.0;
  for (i = 0; i < s->n_atoms; i++) q += s->q[i];
  ew->charged_energy = 0.0;
  if (fabs(q * 0.00268283) > 1.0e-5)
    ew->charged_energy = -ORC_PI / (2.0 * ew->V * ew->alpha2) * q * q;
  return 0;
}

void orc_ewald_free(orc_ewald *ew) {
  free(ew->hkl); free(ew->k); free(ew->k2);
}

/* forces +=, energies2[1] += reciprocal + correction + self + charged,
   vir9 += atomic virial (when not NULL).  extended != 0 keeps the structure
   factor, reciprocal energy and correction energy sums in long double
   (x87 extended) instead of the reference's double: the same terms, without
   the rounding error of a double running sum over N atoms, which grows with
   N (about 1e-12 of the self energy at 33,000 atoms); a checker of that
   size compares energies with this mode. */
int orc_ewald_kspace_x(const orc_system *s, const orc_ewald *ew,
                       double *forces, double *energies2, double *vir9,
                       int extended) {
  orc_pbc p;
  int N = s->n_atoms, H = ew->hklmax, i, n, j;
  double *tab, *sca, *last, energy = 0.0, intra = 0.0, c;
  long double energy_x = 0.0L, intra_x = 0.0L;
  double vxx = 0, vxy = 0, vxz = 0, vyy = 0, vyz = 0, vzz = 0;
  int lastH = 2147483647, lastK = 2147483647;
  orc_pbc_set(&p, s->e1, s->e2, s->e3, s->origin);
  tab = (double *)malloc(sizeof(double) * 3 * 2 * (size_t)H * (N ? N : 1));
  sca = (double *)malloc(sizeof(double) * 2 * (N ? N : 1));
  last = (double *)malloc(sizeof(double) * 2 * (N ? N : 1));
  if (!tab || !sca || !last) return -1;
  /* tables: tab[((j*H + m)*2 + sc)*3 + dim], sc 0 sin, 1 cos (:887-934) */
#define TAB(j, m, sc, d) tab[((((size_t)(j) * H + (m)) * 2 + (sc)) * 3) + (d)]
  for (j = 0; j < N; j++) {
    double r[3], x[3], sn[3], cs[3], qi = s->q[j];
    int d, m;
    orc_pbc_min_pos(&p, s->pos + 3 * j, r);
    for (d = 0; d < 3; d++) {
      x[d] = r[d] * 2.0 * ORC_PI * ew->Lr[d];
      sn[d] = sin(x[d]);
      cs[d] = cos(x[d]);
    }
    for (d = 0; d < 3; d++) {
      TAB(j, 0, 0, d) = 0.0;
      TAB(j, 0, 1, d) = d == 0 ? qi * 1.0 : 1.0;
      TAB(j, 1, 0, d) = d == 0 ? qi * sn[d] : sn[d];
      TAB(j, 1, 1, d) = d == 0 ? qi * cs[d] : cs[d];
    }
    for (m = 2; m < H; m++)
      for (d = 0; d < 3; d++) {
        TAB(j, m, 0, d) = sn[d] * TAB(j, m - 1, 1, d) + cs[d] * TAB(j, m - 1, 0, d);
        TAB(j, m, 1, d) = cs[d] * TAB(j, m - 1, 1, d) - sn[d] * TAB(j, m - 1, 0, d);
      }
  }
  for (n = 0; n < ew->nk; n++) {
    const double *kv = ew->k + 3 * n;
    double sumSin = 0.0, sumCos = 0.0, b, e, a;
    long double sumSin_x = 0.0L, sumCos_x = 0.0L;
    int ih = ew->hkl[3 * n], ik = ew->hkl[3 * n + 1], il = ew->hkl[3 * n + 2];
    int ika = ik < 0 ? -ik : ik, iks = ik < 0 ? -1 : 1;
    int ila = il < 0 ? -il : il, ils = il < 0 ? -1 : 1;
    if (ih != lastH || ik != lastK) {
      for (i = 0; i < N; i++) {
        double xs = TAB(i, ih, 0, 0), xc = TAB(i, ih, 1, 0);
        double ys = iks * TAB(i, ika, 0, 1), yc = TAB(i, ika, 1, 1);
        last[2 * i] = xs * yc + xc * ys;
        last[2 * i + 1] = xc * yc - xs * ys;
      }
      lastH = ih; lastK = ik;
    }
    for (i = 0; i < N; i++) {
      double xys = last[2 * i], xyc = last[2 * i + 1];
      double zs = ils * TAB(i, ila, 0, 2), zc = TAB(i, ila, 1, 2);
      double sA = xys * zc + xyc * zs, cA = xyc * zc - xys * zs;
      sca[2 * i] = sA;
      sca[2 * i + 1] = cA;
      if (extended) {
        sumSin_x += sA;
        sumCos_x += cA;
      } else {
        sumSin += sA;
        sumCos += cA;
      }
    }
    if (extended) {
      sumSin = (double)sumSin_x;
      sumCos = (double)sumCos_x;
    }
    b = 1.0 / ew->k2[n] * exp(-ew->k2[n] * ew->alpha2r / 4.0);
    e = b * (sumSin * sumSin + sumCos * sumCos);
    if (extended)
      energy_x += (long double)b * (sumSin_x * sumSin_x + sumCos_x * sumCos_x);
    else
      energy += e;
    if (vir9) {
      double cc = 2.0 * (1.0 / ew->k2[n] + ew->fac);
      vxx += e * (1.0 - cc * kv[0] * kv[0]);
      vxy -= e * cc * kv[0] * kv[1];
      vxz -= e * cc * kv[0] * kv[2];
      vyy += e * (1.0 - cc * kv[1] * kv[1]);
      vyz -= e * cc * kv[1] * kv[2];
      vzz += e * (1.0 - cc * kv[2] * kv[2]);
    }
    a = 8.0 * ORC_PI * ew->Vr * b;
    for (i = 0; i < N; i++) {
      double w = a * (sca[2 * i] * sumCos - sca[2 * i + 1] * sumSin);
      forces[3 * i] += kv[0] * w;
      forces[3 * i + 1] += kv[1] * w;
      forces[3 * i + 2] += kv[2] * w;
    }
  }
#undef TAB
  c = 4.0 * ORC_PI * ew->Vr;
  if (vir9) {
    vir9[0] += c * vxx; vir9[1] += c * vxy; vir9[2] += c * vxz;
    vir9[3] += c * vxy; vir9[4] += c * vyy; vir9[5] += c * vyz;
    vir9[6] += c * vxz; vir9[7] += c * vyz; vir9[8] += c * vzz;
    vir9[0] += ew->charged_energy;
    vir9[4] += ew->charged_energy;
    vir9[8] += ew->charged_energy;
  }
  /* intramolecular correction over the exclusion table rows (:1105-1129) */
  for (i = 0; i < s->n_excl; i++) {
    int a1 = s->excl[3 * i], a2 = s->excl[3 * i + 1], cls = s->excl[3 * i + 2];
    double rij[3], r2, qq, r, rr, e, g;
    orc_pbc_min_diff(&p, s->pos + 3 * a1, s->pos + 3 * a2, rij, &r2);
    qq = s->q[a1] * s->q[a2];
    if (cls == ORC_MODIFIED) qq *= 1 - s->csf;
    r = sqrt(r2);
    rr = 1 / r;
    e = erf(ew->alpha * r) * rr;
    if (extended)
      intra_x -= (long double)qq * e;
    else
      intra -= qq * e;
    g = qq * (ew->two_alpha_pi * exp(-ew->alpha2 * r2) - e) * rr * rr;
    for (n = 0; n < 3; n++) {
      forces[3 * a1 + n] -= rij[n] * g;
      forces[3 * a2 + n] += rij[n] * g;
    }
    if (vir9) { /* ScalarStructure::addVirial(fij, rij) */
      double f0 = rij[0] * g, f1 = rij[1] * g, f2 = rij[2] * g;
      double xy = f0 * rij[1], xz = f0 * rij[2], yz = f1 * rij[2];
      vir9[0] += f0 * rij[0]; vir9[1] += xy; vir9[2] += xz;
      vir9[3] += xy; vir9[4] += f1 * rij[1]; vir9[5] += yz;
      vir9[6] += xz; vir9[7] += yz; vir9[8] += f2 * rij[2];
    }
  }
  if (extended) {
    energies2[1] += (double)(c * energy_x + intra_x + ew->self_energy +
                             ew->charged_energy);
  } else {
    energies2[1] += c * energy + intra + ew->self_energy + ew->charged_energy;
  }
  free(tab); free(sca); free(last);
  return 0;
}

int orc_ewald_kspace(const orc_system *s, const orc_ewald *ew, double *forces,
                     double *energies2, double *vir9) {
  return orc_ewald_kspace_x(s, ew, forces, energies2, vir9, 0);
}

/* One k-space evaluation: info4 = (alpha, rc, kc, number of k-vectors);
   extended as for orc_ewald_kspace_x. */
int orc_ewald_eval(const orc_system *s, double alpha, double accuracy,
                   double *forces, double *energies2, double *vir9,
                   double *info4, int extended) {
  orc_ewald ew;
  int rc;
  if (orc_ewald_init(&ew, s, alpha, accuracy)) return -1;
  rc = orc_ewald_kspace_x(s, &ew, forces, energies2, vir9, extended);
  if (info4) {
    info4[0] = ew.alpha; info4[1] = ew.rc; info4[2] = ew.kc; info4[3] = ew.nk;
  }
  orc_ewald_free(&ew);
  return rc;
}
