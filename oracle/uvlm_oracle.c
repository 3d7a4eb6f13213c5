/*
 * uvlm_oracle.c -- FP64 CPU restatement of the UVLM per-timestep hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker the HIP path is compared against.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * it; the product never does.
 *
 * Each function cites the reference file:line it restates (paths relative to
 * the xeno14/UVLM source tree).  Parallelism: OpenMP over targets at the same
 * loop sites as the reference (advect.h:33-35, simple_simulator.cpp:86-88,
 * 110-112, 204-206); every per-target sum runs sequentially in the reference's
 * (n,i,j) ring order, so results do not depend on the thread count.
 */
#include "uvlm_oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct {
  double x, y, z;
} v3;

static inline v3 mk(double x, double y, double z) {
  v3 r = {x, y, z};
  return r;
}
static inline v3 ld(const double* p) { return mk(p[0], p[1], p[2]); }
static inline void st(double* p, v3 a) {
  p[0] = a.x;
  p[1] = a.y;
  p[2] = a.z;
}
static inline v3 add(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 sub(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 scl(v3 a, double s) { return mk(a.x * s, a.y * s, a.z * s); }
static inline v3 dvs(v3 a, double s) { return mk(a.x / s, a.y / s, a.z / s); }
static inline v3 neg(v3 a) { return mk(-a.x, -a.y, -a.z); }
static inline double dot(v3 a, v3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
static inline double sqn(v3 a) { return (a.x * a.x + a.y * a.y) + a.z * a.z; }
static inline double nrm(v3 a) { return sqrt(sqn(a)); }
static inline v3 crs(v3 a, v3 b) {
  return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
/* Eigen 3.3 normalize(): divide by the norm when the squared norm is > 0 */
static inline v3 normalized(v3 a) {
  double z = sqn(a);
  if (z > 0) {
    double n = sqrt(z);
    a.x /= n;
    a.y /= n;
    a.z /= n;
  }
  return a;
}

int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* ------------------------------------------------------------------------ */
/* vortex.cpp:21-39 */
/* vortex.cpp:31 cut off length (1e-10 in the reference; settable so that a
 * configured core cut-off runs with the same rule shape) */
static double BS_EPS = 1e-10;

void orc_set_cutoff(double eps) { BS_EPS = eps; }
double orc_get_cutoff(void) { return BS_EPS; }

static inline v3 bs_law(v3 start, v3 end, v3 pos) {
  v3 r0 = sub(end, start);
  v3 r1 = sub(pos, start);
  v3 r2 = sub(pos, end);
  v3 d = crs(r1, r2);
  double d_len = nrm(d);
  double r1_len = nrm(r1);
  double r2_len = nrm(r2);
  if (d_len * d_len < BS_EPS || r1_len < BS_EPS || r2_len < BS_EPS) {
    return mk(0, 0, 0);
  }
  /* 1. / 4. / M_PI / d.squaredNorm() * r0.dot(r1 / |r1| - r2 / |r2|) */
  const double K =
      1. / 4. / M_PI / sqn(d) * dot(r0, sub(dvs(r1, r1_len), dvs(r2, r2_len)));
  return scl(d, K);
}

void orc_biot_savart(const double start[3], const double end[3],
                     const double pos[3], double out[3]) {
  st(out, bs_law(ld(start), ld(end), ld(pos)));
}

#define NODE(nodes, num, rows, cols, n, i, j) \
  ld((nodes) + 3 * (((n) * ((rows) + 1) + (i)) * ((cols) + 1) + (j)))

/* vortex.cpp:145-160: four VORTEX(x, p_k, p_k+1, g) terms (vortex.h:173-182),
 * Gamma applied per segment before summing. */
static inline v3 voring(v3 x, const double* nodes, const double* gamma,
                        size_t num, size_t rows, size_t cols, size_t n,
                        size_t i, size_t j) {
  (void)num;
  const double g = gamma[(n * rows + i) * cols + j];
  v3 p0 = NODE(nodes, num, rows, cols, n, i, j);
  v3 p1 = NODE(nodes, num, rows, cols, n, i, j + 1);
  v3 p2 = NODE(nodes, num, rows, cols, n, i + 1, j + 1);
  v3 p3 = NODE(nodes, num, rows, cols, n, i + 1, j);
  v3 u = mk(0, 0, 0);
  u = add(u, scl(bs_law(p0, p1, x), g));
  u = add(u, scl(bs_law(p1, p2, x), g));
  u = add(u, scl(bs_law(p2, p3, x), g));
  u = add(u, scl(bs_law(p3, p0, x), g));
  return u;
}

void orc_voring(const double x[3], const double* nodes, const double* gamma,
                size_t num, size_t rows, size_t cols, size_t n, size_t i,
                size_t j, double out[3]) {
  st(out, voring(ld(x), nodes, gamma, num, rows, cols, n, i, j));
}

/* velocity.h:14-24: sum over gamma.list_index() in (n,i,j) order from zero */
static inline v3 induced(v3 x, const double* nodes, const double* gamma,
                         size_t num, size_t rows, size_t cols) {
  v3 res = mk(0, 0, 0);
  for (size_t n = 0; n < num; n++)
    for (size_t i = 0; i < rows; i++)
      for (size_t j = 0; j < cols; j++)
        res = add(res, voring(x, nodes, gamma, num, rows, cols, n, i, j));
  return res;
}

void orc_induced_velocity(const double* targets, size_t n_targets,
                          const double* nodes, const double* gamma, size_t num,
                          size_t rows, size_t cols, double* out) {
#pragma omp parallel for schedule(dynamic, 16)
  for (size_t k = 0; k < n_targets; k++) {
    st(out + 3 * k, induced(ld(targets + 3 * k), nodes, gamma, num, rows, cols));
  }
}

/* vortex.cpp:41-65 (VortexRing::BiotSavartLaw / ChordwiseBiotSavartLaw with
 * ForEachSegment order vortex.h:132-140) summed per shed.inl:22-33,66-77 */
static inline v3 ring_bs(const double* ring, double gamma, v3 pos,
                         int chordwise) {
  v3 res = mk(0, 0, 0);
  for (int k = 0; k < 4; k++) {
    if (chordwise && (k == 1 || k == 3)) continue;
    v3 start = ld(ring + 3 * ((k + 1) % 4));
    v3 end = ld(ring + 3 * k);
    res = add(res, bs_law(start, end, pos));
  }
  return scl(res, gamma);
}

void orc_ring_induced_velocity(const double* targets, size_t n_targets,
                               const double* rings, const double* gamma,
                               size_t n_rings, int chordwise, double* out) {
#pragma omp parallel for schedule(dynamic, 16)
  for (size_t k = 0; k < n_targets; k++) {
    v3 x = ld(targets + 3 * k);
    v3 res = mk(0, 0, 0);
    for (size_t r = 0; r < n_rings; r++)
      res = add(res, ring_bs(rings + 12 * r, gamma[r], x, chordwise));
    st(out + 3 * k, res);
  }
}

/* vortex.cpp:82-87 */
void orc_ring_normal(const double* ring, double out[3]) {
  v3 n0 = ld(ring), n1 = ld(ring + 3), n2 = ld(ring + 6), n3 = ld(ring + 9);
  st(out, normalized(crs(sub(n1, n3), sub(n2, n0))));
}
/* vortex.cpp:89-94 */
void orc_ring_tangent(const double* ring, double out[3]) {
  v3 n0 = ld(ring), n1 = ld(ring + 3), n2 = ld(ring + 6), n3 = ld(ring + 9);
  st(out, normalized(add(sub(n1, n3), sub(n2, n0))));
}
/* vortex.cpp:98-100 */
void orc_ring_centroid(const double* ring, double out[3]) {
  v3 n0 = ld(ring), n1 = ld(ring + 3), n2 = ld(ring + 6), n3 = ld(ring + 9);
  st(out, dvs(add(add(add(n0, n1), n2), n3), 4));
}
/* vortex.cpp:132-138 */
double orc_ring_area(const double* ring) {
  v3 n0 = ld(ring), n1 = ld(ring + 3), n2 = ld(ring + 6), n3 = ld(ring + 9);
  v3 u1 = sub(n1, n0), u2 = sub(n3, n0), w1 = sub(n3, n2), w2 = sub(n1, n2);
  return (nrm(crs(u1, u2)) + nrm(crs(w1, w2))) / 2;
}
/* vortex.cpp:140-143 */
void orc_ring_impulse(const double* ring, double gamma, double out[3]) {
  double n[3];
  orc_ring_normal(ring, n);
  st(out, scl(scl(scl(ld(n), gamma), orc_ring_area(ring)), -1));
}

/* ------------------------------------------------------------------------ */
/* simple_simulator.cpp:19-34 */
void orc_collocation_points(const double* nodes, size_t num, size_t rows,
                            size_t cols, double* out) {
  size_t K = 0;
  for (size_t n = 0; n < num; n++)
    for (size_t i = 0; i < rows; i++)
      for (size_t j = 0; j < cols; j++, K++) {
        v3 s = add(add(add(NODE(nodes, num, rows, cols, n, i, j),
                           NODE(nodes, num, rows, cols, n, i + 1, j)),
                       NODE(nodes, num, rows, cols, n, i + 1, j + 1)),
                   NODE(nodes, num, rows, cols, n, i, j + 1));
        st(out + 3 * K, dvs(s, 4));
      }
}

/* simple_simulator.cpp:36-49 */
void orc_normals(const double* nodes, size_t num, size_t rows, size_t cols,
                 double* out) {
  size_t K = 0;
  for (size_t n = 0; n < num; n++)
    for (size_t i = 0; i < rows; i++)
      for (size_t j = 0; j < cols; j++, K++) {
        v3 a = sub(NODE(nodes, num, rows, cols, n, i + 1, j + 1),
                   NODE(nodes, num, rows, cols, n, i, j));
        v3 b = sub(NODE(nodes, num, rows, cols, n, i, j + 1),
                   NODE(nodes, num, rows, cols, n, i + 1, j));
        st(out + 3 * K, normalized(crs(a, b)));
      }
}

/* joukowski.cpp:55-67 (all sheets) */
void orc_panel_area(const double* nodes, size_t num, size_t rows, size_t cols,
                    double* out) {
  size_t K = 0;
  for (size_t n = 0; n < num; n++)
    for (size_t i = 0; i < rows; i++)
      for (size_t j = 0; j < cols; j++, K++) {
        v3 p = NODE(nodes, num, rows, cols, n, i, j);
        v3 a = sub(NODE(nodes, num, rows, cols, n, i + 1, j), p);
        v3 b = sub(NODE(nodes, num, rows, cols, n, i, j + 1), p);
        out[K] = nrm(crs(a, b));
      }
}

/* ------------------------------------------------------------------------ */
/* simple_simulator.cpp:76-101 */
void orc_influence_matrix(const double* nodes, size_t num, size_t rows,
                          size_t cols, const double* cpos, const double* normals,
                          double* A) {
  const size_t N = num * rows * cols;
  double* ones = (double*)malloc(sizeof(double) * (N ? N : 1));
  for (size_t k = 0; k < N; k++) ones[k] = 1;
#pragma omp parallel for schedule(dynamic, 4)
  for (size_t K = 0; K < N; K++) {
    v3 cp = ld(cpos + 3 * K);
    v3 nl = ld(normals + 3 * K);
    size_t L = 0;
    for (size_t n = 0; n < num; n++)
      for (size_t i = 0; i < rows; i++)
        for (size_t j = 0; j < cols; j++, L++)
          A[K * N + L] = dot(voring(cp, nodes, ones, num, rows, cols, n, i, j), nl);
  }
  free(ones);
}

/* Eigen::FullPivLU::computeInPlace + _solve_impl (Eigen 3.3), restated:
 * pivot = first maximal |a| of the remaining corner in column-major visiting
 * order; row and column swap; column divided by the pivot; rank-1 update.
 * Rank: pivots with |u_kk| > maxpivot * eps * n (FullPivLU::rank()).
 * Triangular solves are column-oriented substitutions. */
size_t orc_fullpiv_solve(size_t n, const double* A, const double* b, double* x) {
  if (n == 0) return 0;
  double* m = (double*)malloc(sizeof(double) * n * n); /* column-major */
  size_t* rt = (size_t*)malloc(sizeof(size_t) * n);
  size_t* ct = (size_t*)malloc(sizeof(size_t) * n);
  double* c = (double*)malloc(sizeof(double) * n);
  for (size_t i = 0; i < n; i++)
    for (size_t j = 0; j < n; j++) m[i + j * n] = A[i * n + j];
  double maxpivot = 0;
  size_t nonzero = n;
  for (size_t k = 0; k < n; k++) {
    double best = -1;
    size_t br = k, bc = k;
    for (size_t cc = k; cc < n; cc++)
      for (size_t r = k; r < n; r++) {
        double a = fabs(m[r + cc * n]);
        if (a > best) {
          best = a;
          br = r;
          bc = cc;
        }
      }
    if (best == 0) {
      nonzero = k;
      for (size_t i = k; i < n; i++) rt[i] = ct[i] = i;
      break;
    }
    if (best > maxpivot) maxpivot = best;
    rt[k] = br;
    ct[k] = bc;
    if (br != k)
      for (size_t cc = 0; cc < n; cc++) {
        double t = m[k + cc * n];
        m[k + cc * n] = m[br + cc * n];
        m[br + cc * n] = t;
      }
    if (bc != k)
      for (size_t r = 0; r < n; r++) {
        double t = m[r + k * n];
        m[r + k * n] = m[r + bc * n];
        m[r + bc * n] = t;
      }
    const double piv = m[k + k * n];
    for (size_t r = k + 1; r < n; r++) m[r + k * n] /= piv;
#pragma omp parallel for schedule(static) if (n - k > 128)
    for (size_t cc = k + 1; cc < n; cc++) {
      const double u = m[k + cc * n];
      double* col = m + cc * n;
      const double* l = m + k * n;
      for (size_t r = k + 1; r < n; r++) col[r] -= l[r] * u;
    }
  }
  /* rank with the default threshold */
  const double thr = maxpivot * (2.220446049250313e-16 * (double)n);
  size_t rank = 0;
  for (size_t k = 0; k < nonzero; k++)
    if (fabs(m[k + k * n]) > thr) rank++;
  /* Step 1: c = P b */
  for (size_t i = 0; i < n; i++) c[i] = b[i];
  for (size_t k = 0; k < n; k++) {
    double t = c[k];
    c[k] = c[rt[k]];
    c[rt[k]] = t;
  }
  /* Step 2: unit lower */
  for (size_t k = 0; k < n; k++)
    for (size_t r = k + 1; r < n; r++) c[r] -= m[r + k * n] * c[k];
  /* Step 3: upper on the rank x rank block */
  for (size_t kk = rank; kk-- > 0;) {
    c[kk] /= m[kk + kk * n];
    for (size_t r = 0; r < kk; r++) c[r] -= m[r + kk * n] * c[kk];
  }
  for (size_t k = rank; k < n; k++) c[k] = 0;
  /* Step 4: x = Q c */
  for (size_t k = n; k-- > 0;) {
    double t = c[k];
    c[k] = c[ct[k]];
    c[ct[k]] = t;
  }
  for (size_t i = 0; i < n; i++) x[i] = c[i];
  free(m);
  free(rt);
  free(ct);
  free(c);
  return rank;
}

int orc_partialpiv_solve(size_t n, const double* A, const double* b, double* x) {
  if (n == 0) return 0;
  double* m = (double*)malloc(sizeof(double) * n * n);
  double* c = (double*)malloc(sizeof(double) * n);
  int singular = 0;
  for (size_t i = 0; i < n; i++)
    for (size_t j = 0; j < n; j++) m[i + j * n] = A[i * n + j];
  for (size_t i = 0; i < n; i++) c[i] = b[i];
  for (size_t k = 0; k < n; k++) {
    double best = -1;
    size_t br = k;
    for (size_t r = k; r < n; r++) {
      double a = fabs(m[r + k * n]);
      if (a > best) {
        best = a;
        br = r;
      }
    }
    if (best == 0) {
      singular = 1;
      break;
    }
    if (br != k) {
      for (size_t cc = 0; cc < n; cc++) {
        double t = m[k + cc * n];
        m[k + cc * n] = m[br + cc * n];
        m[br + cc * n] = t;
      }
      double t = c[k];
      c[k] = c[br];
      c[br] = t;
    }
    const double piv = m[k + k * n];
    for (size_t r = k + 1; r < n; r++) m[r + k * n] /= piv;
#pragma omp parallel for schedule(static) if (n - k > 128)
    for (size_t cc = k + 1; cc < n; cc++) {
      const double u = m[k + cc * n];
      double* col = m + cc * n;
      const double* l = m + k * n;
      for (size_t r = k + 1; r < n; r++) col[r] -= l[r] * u;
    }
  }
  if (!singular) {
    for (size_t k = 0; k < n; k++)
      for (size_t r = k + 1; r < n; r++) c[r] -= m[r + k * n] * c[k];
    for (size_t kk = n; kk-- > 0;) {
      c[kk] /= m[kk + kk * n];
      for (size_t r = 0; r < kk; r++) c[r] -= m[r + kk * n] * c[kk];
    }
    for (size_t i = 0; i < n; i++) x[i] = c[i];
  }
  free(m);
  free(c);
  return singular ? -1 : 0;
}

/* ------------------------------------------------------------------------ */
/* util.h:36-43 */
void orc_linspace(double a, double b, size_t n, double* out) {
  double dx = (b - a) / (double)(n - 1);
  for (size_t i = 0; i < n; i++) out[i] = a + (double)i * dx;
}

/* naca4digit.cpp:12-25 */
double orc_naca4digit(double x, double c, int xx) {
  int first = xx / 10;
  int second = xx % 10;
  const double m = first / 100.0;
  const double p = second / 10.0;
  double res = 0;
  if (0 <= x && x <= p * c) {
    res = m * x / p / p * (2 * p - x / c);
  } else if (p * c <= x && x <= c) {
    res = m * (c - x) / (1 - p) / (1 - p) * (1 + x / c - 2 * p);
  }
  return res;
}

/* naca00XX.cpp:15-20 */
double orc_naca00xx(double x, double c, int xx) {
  double t = xx * 0.01;
  double z = x / c;
  return 5 * t * c *
         (0.2969 * sqrt(z) + (-0.1260) * (z) + (-0.3516) * z * z +
          0.2843 * z * z * z + (-0.1015) * z * z * z * z);
}

/* rect.cpp:18-44 (uniform spacing) + naca*.cpp z-mapping, then wing.cpp:12-50
 * (mirror: reversed negated-y row, then the y>0 part), then
 * simple_simulator.cpp:160-168 (+origin). */
int orc_build_wing(int generator, int digit, double chord, double span,
                   size_t rows, size_t cols, const double origin[3],
                   double* nodes) {
  if (cols & 1) return -1;
  const size_t hc = cols / 2;
  double* xs = (double*)malloc(sizeof(double) * (rows + 1));
  double* ys = (double*)malloc(sizeof(double) * (hc + 1));
  orc_linspace(0, chord, rows + 1, xs);
  orc_linspace(0, span / 2, hc + 1, ys);
  const size_t C1 = cols + 1;
  for (size_t i = 0; i <= rows; i++) {
    const double x = xs[i];
    double z = 0;
    if (generator == ORC_GEN_NACA4) z = orc_naca4digit(x, chord, digit);
    if (generator == ORC_GEN_NACA00) z = orc_naca00xx(x, chord, digit);
    double* row = nodes + 3 * i * C1;
    size_t j = 0;
    for (size_t h = hc + 1; h-- > 0; j++) { /* mirrored, reversed */
      double y = ys[h];
      y *= -1;
      row[3 * j + 0] = x;
      row[3 * j + 1] = y;
      row[3 * j + 2] = z;
    }
    for (size_t h = 1; h <= hc; h++, j++) {
      row[3 * j + 0] = x;
      row[3 * j + 1] = ys[h];
      row[3 * j + 2] = z;
    }
  }
  for (size_t k = 0; k < (rows + 1) * C1; k++) {
    double* p = nodes + 3 * k;
    for (int d = 0; d < 3; d++) p[d] = p[d] + 0.0; /* half.origin() = 0 */
    for (int d = 0; d < 3; d++) p[d] = p[d] + origin[d];
  }
  free(xs);
  free(ys);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Motion laws (see header) */
static double law_plug(const orc_wing_spec* w, double t) {
  return w->plug_amp * cos(w->plug_omega * t + w->plug_phase);
}
static double law_flap(const orc_wing_spec* w, double t) {
  return w->flap_amp * cos(w->flap_omega * t + w->flap_phase);
}
static double law_twist(const orc_wing_spec* w, v3 x, double t) {
  const double eta = x.y / (w->span / 2);
  double s = 0;
  for (int m = 0; m < 4; m++)
    for (int q = 0; q < 2; q++)
      s = s + w->twist_amp[m][q] * sin((m + 1) * M_PI_2 * eta) *
                  cos((q + 1) * w->fourier_omega * t + w->twist_phase[m][q]);
  return s;
}
static double law_bend(const orc_wing_spec* w, v3 x, double t) {
  const double eta = x.y / (w->span / 2);
  const double xi = x.x / w->chord;
  double s = 0;
  for (int m = 0; m < 4; m++)
    for (int q = 0; q < 2; q++)
      s = s + w->bend_amp[m][q] * sin(M_PI * xi) * sin((m + 1) * M_PI_2 * eta) *
                  cos((q + 1) * w->fourier_omega * t + w->bend_phase[m][q]);
  return s;
}

typedef struct {
  double a[3][3];
} m3;

static m3 m3mul(const m3* A, const m3* B) {
  m3 r;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      r.a[i][j] = (A->a[i][0] * B->a[0][j] + A->a[i][1] * B->a[1][j]) +
                  A->a[i][2] * B->a[2][j];
  return r;
}
static v3 m3v(const m3* A, v3 v) {
  return mk((A->a[0][0] * v.x + A->a[0][1] * v.y) + A->a[0][2] * v.z,
            (A->a[1][0] * v.x + A->a[1][1] * v.y) + A->a[1][2] * v.z,
            (A->a[2][0] * v.x + A->a[2][1] * v.y) + A->a[2][2] * v.z);
}

/* morphing.cpp:54-76  attack * flap * twist */
static m3 prepare_matrix(const orc_wing_spec* w, v3 x0, double t) {
  const double phi = law_flap(w, t);
  m3 flap = {{{1, 0, 0}, {0, cos(phi), -sin(phi)}, {0, sin(phi), cos(phi)}}};
  const double alpha = w->alpha;
  m3 attack = {{{cos(alpha), 0, sin(alpha)}, {0, 1, 0}, {-sin(alpha), 0, cos(alpha)}}};
  const double beta = law_twist(w, x0, t);
  m3 twist = {{{cos(beta), 0, sin(beta)}, {0, 1, 0}, {-sin(beta), 0, cos(beta)}}};
  m3 af = m3mul(&attack, &flap);
  return m3mul(&af, &twist);
}

/* morphing.cpp:23-43 */
static v3 perform(const orc_wing_spec* w, v3 x0, double t) {
  v3 origin = ld(w->origin);
  v3 x_ref = sub(x0, origin);
  int is_negative = x_ref.y < 0;
  if (is_negative) x_ref.y *= -1;
  m3 T = prepare_matrix(w, x_ref, t);
  const double gamma_z = law_plug(w, t);
  const double gamma_b = law_bend(w, x_ref, t);
  v3 uz = mk(0, 0, 1);
  v3 x = add(scl(uz, gamma_z), m3v(&T, add(x_ref, scl(uz, gamma_b))));
  if (x_ref.y == 0) x.y = 0;
  if (is_negative) x.y *= -1;
  x = add(x, origin);
  return x;
}

/* morphing.cpp:45-52 */
static v3 morph_velocity(const orc_wing_spec* w, v3 x0, double t) {
  const double dt = 1e-6;
  v3 x1 = perform(w, x0, t - dt);
  v3 x2 = perform(w, x0, t + dt);
  return dvs(sub(x2, x1), 2 * dt);
}

void orc_morph_perform(const orc_wing_spec* w, const double x0[3], double t,
                       double out[3]) {
  st(out, perform(w, ld(x0), t));
}
void orc_morph_velocity(const orc_wing_spec* w, const double x0[3], double t,
                        double out[3]) {
  st(out, morph_velocity(w, ld(x0), t));
}

/* ------------------------------------------------------------------------ */
/* SimpleSimulator state (simple_simulator.h:65-80).  Sheets are sheet-major,
 * row-major (MultipleSheet layout K = j + cols*(i + rows*n)). */
struct orc_sim {
  orc_sim_config cfg;
  size_t num, R, C, N;
  double* wing_init; /* [num][R+1][C+1][3] */
  double* wing;
  double* gamma;      /* [num][R][C] */
  double* gamma_prev;
  double* wake;       /* [num][wr+1][C+1][3] (wr+1 node rows) */
  double* wake_gamma; /* [num][wg][C] */
  size_t wake_rows;   /* node rows */
  size_t wake_grows;  /* gamma rows */
  double* v_prev;     /* AB2 history [num][rows_prev][C+1][3] */
  size_t v_prev_rows;
  double* loads;      /* [num][2] */
  double pairs;
};

static void* xcalloc(size_t n, size_t sz) { return calloc(n ? n : 1, sz); }

orc_sim* orc_sim_create(const orc_sim_config* cfg) {
  if (cfg->num_wings < 1 || cfg->num_wings > ORC_MAX_WINGS) return NULL;
  if (cfg->cols & 1) return NULL;
  orc_sim* s = (orc_sim*)xcalloc(1, sizeof(orc_sim));
  s->cfg = *cfg;
  s->num = (size_t)cfg->num_wings;
  s->R = cfg->rows;
  s->C = cfg->cols;
  s->N = s->num * s->R * s->C;
  const size_t nn = s->num * (s->R + 1) * (s->C + 1);
  s->wing_init = (double*)xcalloc(3 * nn, sizeof(double));
  s->wing = (double*)xcalloc(3 * nn, sizeof(double));
  s->gamma = (double*)xcalloc(s->N, sizeof(double));
  s->gamma_prev = (double*)xcalloc(s->N, sizeof(double));
  s->loads = (double*)xcalloc(2 * s->num, sizeof(double));
  for (size_t n = 0; n < s->num; n++) {
    const orc_wing_spec* w = &cfg->wings[n];
    orc_build_wing(w->generator, w->digit, w->chord, w->span, s->R, s->C,
                   w->origin, s->wing_init + 3 * n * (s->R + 1) * (s->C + 1));
  }
  s->wake = (double*)xcalloc(1, sizeof(double));
  s->wake_gamma = (double*)xcalloc(1, sizeof(double));
  s->wake_rows = 0;
  s->wake_grows = 0;
  return s;
}

void orc_sim_destroy(orc_sim* s) {
  if (!s) return;
  free(s->wing_init);
  free(s->wing);
  free(s->gamma);
  free(s->gamma_prev);
  free(s->wake);
  free(s->wake_gamma);
  free(s->v_prev);
  free(s->loads);
  free(s);
}

size_t orc_sim_num_bound(const orc_sim* s) { return s->N; }
size_t orc_sim_wake_node_rows(const orc_sim* s) { return s->wake_rows; }
size_t orc_sim_wake_gamma_rows(const orc_sim* s) { return s->wake_grows; }
double orc_sim_pairs(const orc_sim* s) { return s->pairs; }
void orc_sim_get_wing_nodes(const orc_sim* s, double* out) {
  memcpy(out, s->wing, sizeof(double) * 3 * s->num * (s->R + 1) * (s->C + 1));
}
void orc_sim_get_wing_gamma(const orc_sim* s, double* out) {
  memcpy(out, s->gamma, sizeof(double) * s->N);
}
void orc_sim_get_wake(const orc_sim* s, double* nodes, double* gamma) {
  if (nodes)
    memcpy(nodes, s->wake, sizeof(double) * 3 * s->num * s->wake_rows * (s->C + 1));
  if (gamma)
    memcpy(gamma, s->wake_gamma, sizeof(double) * s->num * s->wake_grows * s->C);
}
void orc_sim_get_loads(const orc_sim* s, double* out) {
  memcpy(out, s->loads, sizeof(double) * 2 * s->num);
}

int orc_sim_set_state(orc_sim* s, size_t node_rows, size_t gamma_rows,
                      const double* wake_nodes, const double* wake_gamma,
                      const double* wing_gamma) {
  if (!(gamma_rows + 1 == node_rows || (node_rows == 0 && gamma_rows == 0)))
    return -1;
  free(s->wake);
  free(s->wake_gamma);
  s->wake = (double*)xcalloc(3 * s->num * node_rows * (s->C + 1), sizeof(double));
  s->wake_gamma = (double*)xcalloc(s->num * gamma_rows * s->C, sizeof(double));
  memcpy(s->wake, wake_nodes, sizeof(double) * 3 * s->num * node_rows * (s->C + 1));
  memcpy(s->wake_gamma, wake_gamma, sizeof(double) * s->num * gamma_rows * s->C);
  s->wake_rows = node_rows;
  s->wake_grows = gamma_rows;
  if (wing_gamma) memcpy(s->gamma, wing_gamma, sizeof(double) * s->N);
  free(s->v_prev);
  s->v_prev = NULL;
  s->v_prev_rows = 0;
  return 0;
}

/* MultipleSheet::prepend_row: insert one row per sheet as row 0. */
static double* prepend_rows(const double* old, size_t num, size_t rows,
                            size_t width, const double* row_data) {
  double* nw = (double*)xcalloc(num * (rows + 1) * width, sizeof(double));
  for (size_t n = 0; n < num; n++) {
    memcpy(nw + n * (rows + 1) * width, row_data + n * width,
           sizeof(double) * width);
    memcpy(nw + (n * (rows + 1) + 1) * width, old + n * rows * width,
           sizeof(double) * rows * width);
  }
  return nw;
}

/* MultipleSheet::erase_last_row */
static double* erase_last(const double* old, size_t num, size_t rows,
                          size_t width) {
  double* nw = (double*)xcalloc(num * (rows - 1) * width, sizeof(double));
  for (size_t n = 0; n < num; n++)
    memcpy(nw + n * (rows - 1) * width, old + n * rows * width,
           sizeof(double) * (rows - 1) * width);
  return nw;
}

/* advect.h:16-24  ((-ff + bound) + wake) */
static inline v3 flow_velocity(const orc_sim* s, v3 x, const double* wake,
                               size_t wake_grows) {
  v3 ff = ld(s->cfg.ff);
  v3 b = induced(x, s->wing, s->gamma, s->num, s->R, s->C);
  v3 w = induced(x, wake, s->wake_gamma, s->num, wake_grows, s->C);
  return add(add(neg(ff), b), w);
}

/* advect.h:26-46 on a wake-shaped node sheet 'pos' (also the wake sources) */
static void sheet_velocity(orc_sim* s, const double* pos, double* out) {
  const size_t nodes = s->num * s->wake_rows * (s->C + 1);
  const size_t wg = s->wake_grows;
#pragma omp parallel for schedule(dynamic, 8)
  for (size_t K = 0; K < nodes; K++)
    st(out + 3 * K, flow_velocity(s, ld(pos + 3 * K), pos, wg));
  v3 mff = neg(ld(s->cfg.ff));
  for (size_t n = 0; n < s->num; n++)
    for (size_t j = 0; j <= s->C; j++)
      st(out + 3 * (n * s->wake_rows * (s->C + 1) + j), mff);
  s->pairs += 4.0 * (double)nodes * (double)(s->N + s->num * wg * s->C);
}

void orc_sim_wake_velocity(const orc_sim* s, double* out) {
  sheet_velocity((orc_sim*)s, s->wake, out);
}

/* advect.cpp:11-216 */
static void advect(orc_sim* s, double dt) {
  const size_t nodes = s->num * s->wake_rows * (s->C + 1);
  const size_t C1 = s->C + 1;
  const size_t ign = s->cfg.ignore_steps;
  double* wake = s->wake;
  if (nodes == 0) return;
  if (s->cfg.scheme == ORC_EULER) {
    double* vel = (double*)xcalloc(3 * nodes, sizeof(double));
    sheet_velocity(s, wake, vel);
    for (size_t K = 0; K < nodes; K++)
      st(wake + 3 * K, add(ld(wake + 3 * K), scl(ld(vel + 3 * K), dt)));
    free(vel);
  } else if (s->cfg.scheme == ORC_RK2) {
    double* k1 = (double*)xcalloc(3 * nodes, sizeof(double));
    double* k2 = (double*)xcalloc(3 * nodes, sizeof(double));
    double* p2 = (double*)xcalloc(3 * nodes, sizeof(double));
    sheet_velocity(s, wake, k1);
    for (size_t K = 0; K < nodes; K++)
      st(p2 + 3 * K, add(ld(wake + 3 * K), scl(ld(k1 + 3 * K), dt)));
    sheet_velocity(s, p2, k2);
    for (size_t K = 0; K < nodes; K++) {
      const size_t i = (K / C1) % s->wake_rows;
      v3 w = ld(wake + 3 * K);
      if (i > ign)
        st(wake + 3 * K,
           add(w, dvs(scl(add(ld(k1 + 3 * K), ld(k2 + 3 * K)), dt), 2)));
      else
        st(wake + 3 * K, add(w, scl(ld(k1 + 3 * K), dt)));
    }
    free(k1);
    free(k2);
    free(p2);
  } else if (s->cfg.scheme == ORC_RK4 || s->cfg.scheme == ORC_RK4_CLASSIC) {
    const int classic = s->cfg.scheme == ORC_RK4_CLASSIC;
    double* k1 = (double*)xcalloc(3 * nodes, sizeof(double));
    double* k2 = (double*)xcalloc(3 * nodes, sizeof(double));
    double* k3 = (double*)xcalloc(3 * nodes, sizeof(double));
    double* k4 = classic ? (double*)xcalloc(3 * nodes, sizeof(double)) : NULL;
    double* p = (double*)xcalloc(3 * nodes, sizeof(double));
    sheet_velocity(s, wake, k1);
    for (size_t K = 0; K < nodes; K++)
      st(p + 3 * K, add(ld(wake + 3 * K), dvs(scl(ld(k1 + 3 * K), dt), 2)));
    sheet_velocity(s, p, k2);
    for (size_t K = 0; K < nodes; K++)
      st(p + 3 * K, add(ld(wake + 3 * K), dvs(scl(ld(k2 + 3 * K), dt), 2)));
    sheet_velocity(s, p, k3);
    /* k4 of the reference formula is computed and never used
     * (advect.cpp:150-164 vs :172-173); it is skipped here. */
    if (classic) {
      for (size_t K = 0; K < nodes; K++)
        st(p + 3 * K, add(ld(wake + 3 * K), scl(ld(k3 + 3 * K), dt)));
      sheet_velocity(s, p, k4);
    }
    for (size_t K = 0; K < nodes; K++) {
      const size_t i = (K / C1) % s->wake_rows;
      v3 w = ld(wake + 3 * K);
      v3 a = ld(k1 + 3 * K), b = ld(k2 + 3 * K), c = ld(k3 + 3 * K);
      if (i > ign) {
        v3 last = classic ? ld(k4 + 3 * K) : c;
        v3 sum = add(add(add(a, scl(b, 2)), scl(c, 2)), last);
        st(wake + 3 * K, add(w, dvs(scl(sum, dt), 6)));
      } else {
        st(wake + 3 * K, add(w, scl(a, dt)));
      }
    }
    free(k1);
    free(k2);
    free(k3);
    free(k4);
    free(p);
  } else if (s->cfg.scheme == ORC_AB2) {
    double* v = (double*)xcalloc(3 * nodes, sizeof(double));
    sheet_velocity(s, wake, v);
    if (!s->v_prev) {
      for (size_t K = 0; K < nodes; K++)
        st(wake + 3 * K, add(ld(wake + 3 * K), scl(ld(v + 3 * K), dt)));
    } else {
      const size_t pr = s->v_prev_rows;
      for (size_t n = 0; n < s->num; n++)
        for (size_t i = 0; i < s->wake_rows; i++)
          for (size_t j = 0; j < C1; j++) {
            const size_t K = (n * s->wake_rows + i) * C1 + j;
            if (i >= 1) {
              v3 vp = ld(s->v_prev + 3 * ((n * pr + (i - 1)) * C1 + j));
              v3 dv = sub(scl(ld(v + 3 * K), 3), vp);
              st(wake + 3 * K, add(ld(wake + 3 * K), dvs(scl(dv, dt), 2)));
            } else {
              st(wake + 3 * K, add(ld(wake + 3 * K), scl(ld(v + 3 * K), dt)));
            }
          }
    }
    free(s->v_prev);
    s->v_prev = v;
    s->v_prev_rows = s->wake_rows;
  }
}

/* ------------------------------------------------------------------------ */
/* Legacy loads and the Lamb vector (rings as [4][3] nodes, legacy order). */

static v3 rings_velocity(v3 x, const double* rings, const double* gamma, size_t n,
                         int chordwise) {
  v3 res = mk(0, 0, 0);
  for (size_t r = 0; r < n; r++)
    res = add(res, ring_bs(rings + 12 * r, gamma[r], x, chordwise));
  return res;
}

/* calc_impulse.h:14-30: sum over the segments node[(k+1)%4] -> node[k] of
 * (v_node[k] + u(mid) - U) x dl * gamma, u from the source rings. */
void orc_lamb_vector(const double ring[12], double gamma, const double v_nodes[12],
                     const double U[3], const double* rings, const double* ring_gamma,
                     size_t n, double out[3]) {
  v3 res = mk(0, 0, 0);
  for (int k = 0; k < 4; k++) {
    v3 start = ld(ring + 3 * ((k + 1) % 4)), end = ld(ring + 3 * k);
    v3 dl = sub(end, start);
    v3 pos = dvs(add(start, end), 2);
    v3 u = rings_velocity(pos, rings, ring_gamma, n, 0);
    v3 ue = sub(add(ld(v_nodes + 3 * k), u), ld(U));
    res = add(res, scl(crs(ue, dl), gamma));
  }
  st(out, res);
}

static double ring_b(const double* r) { return nrm(sub(ld(r), ld(r + 9))); } /* CalcB */
static double ring_c(const double* r) { return nrm(sub(ld(r), ld(r + 3))); } /* CalcC */

/* katz_plotkin.cpp:16-68 on one container (R x C rings, row-major) whose
 * wake and chordwise-bound velocities are zero (the loop's ring set is
 * never filled, simulator.cpp:37); F starts from zero.  out = F, Pin, Pout. */
void orc_katz_plotkin_load(const double* rings, const double* gamma,
                           const double* gamma_prev, size_t R, size_t C,
                           const orc_wing_spec* w, const double U[3], double rho,
                           double t, double dt, double out[5]) {
  v3 F = mk(0, 0, 0);
  double Pin = 0;
  for (size_t i = 0; i < R; i++)
    for (size_t j = 0; j < C; j++) {
      const size_t K = i * C + j;
      const double* r = rings + 12 * K;
      double cc[3], nn[3], tt[3];
      orc_ring_centroid(r, cc);
      orc_ring_normal(r, nn);
      orc_ring_tangent(r, tt);
      const v3 centroid = ld(cc), n = ld(nn), tg = ld(tt);
      const double db = ring_b(r), dc = ring_c(r);
      const v3 Um = sub(ld(U), morph_velocity(w, centroid, t)); /* CalcUm */
      const v3 u = normalized(Um);                            /* projection */
      const double P[3][3] = {{1 - u.x * u.x, 0 - u.x * u.y, 0 - u.x * u.z},
                              {0 - u.y * u.x, 1 - u.y * u.y, 0 - u.y * u.z},
                              {0 - u.z * u.x, 0 - u.z * u.y, 1 - u.z * u.z}};
      const double alpha = atan2(dot(Um, n), dot(Um, tg));
      const double dg = gamma[K] - (i == 0 ? 0 : gamma[K - C]); /* DeltaGamma */
      const v3 grad = dvs(scl(tg, dg), dc);                      /* Grad */
      const double Lst = rho * dot(Um, grad) * db * dc * cos(alpha);
      const v3 Pn = mk((P[0][0] * n.x + P[0][1] * n.y) + P[0][2] * n.z,
                       (P[1][0] * n.x + P[1][1] * n.y) + P[1][2] * n.z,
                       (P[2][0] * n.x + P[2][1] * n.y) + P[2][2] * n.z);
      const double Dst = rho * (-dot(Um, Pn) * dg * db);
      double dg_dt = (gamma[K] - gamma_prev[K]) / dt;
      if (i > 0) dg_dt += (gamma[K - C] - gamma_prev[K - C]) / dt;
      dg_dt /= 2;
      const double Lunst = rho * dg_dt * db * dc * cos(alpha);
      const double Dunst = rho * dg_dt * db * dc * sin(alpha);
      const v3 e_drag = dvs(Um, nrm(Um));
      const v3 e_lift = Pn;
      const v3 dF = add(scl(e_drag, Dst + Dunst), scl(e_lift, Lst + Lunst));
      F = add(F, dF);
      Pin += (-1) * dot(Um, n) * dot(dF, n);
    }
  out[0] = F.x;
  out[1] = F.y;
  out[2] = F.z;
  out[3] = Pin;
  out[4] = 0;
}

/* joukowski.h:63-99 on one container: per ring, U x (end - start) rho gamma
 * over the segments node[(k+1)%4] -> node[k] (k = 1 of the last row
 * skipped), U = velocity of the source rings at the midpoint + freestream -
 * surface velocity at the reference midpoint; plus normal rho dG/dt CalcB
 * CalcC; Pin = sum dF . (surface velocity at the reference centroid),
 * Pout = -F.x |U|.  rings0: reference nodes.  out = F, Pin, Pout. */
void orc_joukowski_container_load(const double* rings, const double* rings0,
                                  const double* gamma, const double* gamma_prev,
                                  size_t R, size_t C, const double* src,
                                  const double* src_gamma, size_t n_src,
                                  const orc_wing_spec* w, const double U[3], double rho,
                                  double t, double dt, double out[5]) {
  const size_t N = R * C;
  double* dF = (double*)xcalloc(3 * N, sizeof(double));
  double Fx = 0, Fy = 0, Fz = 0, Pin = 0;
#pragma omp parallel for schedule(dynamic, 4)
  for (size_t K = 0; K < N; K++) {
    const size_t i = K / C;
    const double* r = rings + 12 * K;
    const double* r0 = rings0 + 12 * K;
    v3 st_ = mk(0, 0, 0);
    for (int k = 0; k < 4; k++) {
      if (i + 1 == R && k == 1) continue;
      v3 start = ld(r + 3 * ((k + 1) % 4)), end = ld(r + 3 * k);
      v3 start0 = ld(r0 + 3 * ((k + 1) % 4)), end0 = ld(r0 + 3 * k);
      v3 mid = dvs(add(start, end), 2), mid0 = dvs(add(start0, end0), 2);
      v3 Uind = rings_velocity(mid, src, src_gamma, n_src, 0);
      v3 Um = sub(ld(U), morph_velocity(w, mid0, t));
      v3 Uv = add(Uind, Um);
      st_ = add(st_, scl(scl(crs(Uv, sub(end, start)), rho), gamma[K]));
    }
    double nn[3];
    orc_ring_normal(r, nn);
    const double dGdt = (gamma[K] - gamma_prev[K]) / dt;
    v3 un = scl(scl(scl(scl(ld(nn), rho), dGdt), ring_b(r)), ring_c(r));
    st(dF + 3 * K, add(st_, un));
  }
  for (size_t K = 0; K < N; K++) {
    const double* r0 = rings0 + 12 * K;
    double c0[3];
    orc_ring_centroid(r0, c0);
    v3 d = ld(dF + 3 * K);
    Fx += d.x;
    Fy += d.y;
    Fz += d.z;
    Pin += dot(d, morph_velocity(w, ld(c0), t));
  }
  free(dF);
  out[0] = Fx;
  out[1] = Fy;
  out[2] = Fz;
  out[3] = Pin;
  out[4] = Fx * nrm(ld(U)) * (-1);
}

/* One Advect call alone (advect.cpp:11-216) on the current wake, the
 * current wing and its circulations as sources. */
void orc_sim_advect(orc_sim* s, double dt) { advect(s, dt); }

/* Replace the wake (any shape) and keep the AB2 velocity history. */
int orc_sim_replace_wake(orc_sim* s, size_t node_rows, size_t gamma_rows,
                         const double* wake_nodes, const double* wake_gamma) {
  if (!(gamma_rows + 1 == node_rows || (node_rows == 0 && gamma_rows == 0)))
    return -1;
  free(s->wake);
  free(s->wake_gamma);
  s->wake = (double*)xcalloc(3 * s->num * node_rows * (s->C + 1), sizeof(double));
  s->wake_gamma = (double*)xcalloc(s->num * gamma_rows * s->C, sizeof(double));
  memcpy(s->wake, wake_nodes, sizeof(double) * 3 * s->num * node_rows * (s->C + 1));
  memcpy(s->wake_gamma, wake_gamma, sizeof(double) * s->num * gamma_rows * s->C);
  s->wake_rows = node_rows;
  s->wake_grows = gamma_rows;
  return 0;
}

/* joukowski.cpp:12-92 + simple_simulator.cpp:195-233 */
static void loads_impl(const double* wing, const double* wing_init,
                       const double* gamma, const double* gamma_prev,
                       size_t num, size_t R, size_t C, const double* wake,
                       const double* wake_gamma, size_t wake_grows,
                       const orc_wing_spec* wings, const double* ffp, double t,
                       double dt, double* out, double* pairs) {
  const size_t NR = num * R * C;
  v3 ff = ld(ffp);
  double* normal = (double*)xcalloc(3 * NR, sizeof(double));
  double* area = (double*)xcalloc(NR, sizeof(double));
  orc_normals(wing, num, R, C, normal);
  orc_panel_area(wing, num, R, C, area);
  const size_t nlines = 4 * R * C - C;
  double* lines = (double*)xcalloc(nlines * 13, sizeof(double));
  double* U = (double*)xcalloc(3 * nlines, sizeof(double));
  for (size_t n = 0; n < num; n++) {
    size_t l = 0;
    for (size_t i = 0; i < R; i++)
      for (size_t j = 0; j < C; j++) {
        v3 c[4] = {NODE(wing, num, R, C, n, i, j), NODE(wing, num, R, C, n, i, j + 1),
                   NODE(wing, num, R, C, n, i + 1, j + 1),
                   NODE(wing, num, R, C, n, i + 1, j)};
        v3 c0[4] = {NODE(wing_init, num, R, C, n, i, j),
                    NODE(wing_init, num, R, C, n, i, j + 1),
                    NODE(wing_init, num, R, C, n, i + 1, j + 1),
                    NODE(wing_init, num, R, C, n, i + 1, j)};
        for (int k = 0; k < 4; k++) {
          if (i == R - 1 && k == 2) continue; /* skip T.E. */
          double* L = lines + 13 * l++;
          st(L, c[k]);
          st(L + 3, c[(k + 1) % 4]);
          st(L + 6, c0[k]);
          st(L + 9, c0[(k + 1) % 4]);
          L[12] = gamma[(n * R + i) * C + j];
        }
      }
    const orc_wing_spec* w = &wings[n];
#pragma omp parallel for schedule(dynamic, 8)
    for (size_t q = 0; q < nlines; q++) {
      const double* L = lines + 13 * q;
      v3 mp = dvs(add(ld(L), ld(L + 3)), 2);
      v3 mp0 = dvs(add(ld(L + 6), ld(L + 9)), 2);
      v3 b = induced(mp, wing, gamma, num, R, C);
      v3 wv = induced(mp, wake, wake_gamma, num, wake_grows, C);
      v3 vel = add(add(neg(ff), b), wv);
      st(U + 3 * q, sub(vel, morph_velocity(w, mp0, t)));
    }
    if (pairs) *pairs += 4.0 * (double)nlines * (double)(NR + num * wake_grows * C);
    /* JoukowskiSteady (joukowski.cpp:36-53) */
    double Fx = 0, Fy = 0, Fz = 0;
    for (size_t q = 0; q < nlines; q++) {
      const double* L = lines + 13 * q;
      v3 df = scl(crs(ld(U + 3 * q), sub(ld(L + 3), ld(L))), L[12]);
      Fx += df.x;
      Fy += df.y;
      Fz += df.z;
    }
    v3 Fst = mk(Fx, Fy, Fz);
    /* JoukowskiUnsteady (joukowski.h:133-149) */
    Fx = 0, Fy = 0, Fz = 0;
    for (size_t k = 0; k < R * C; k++) {
      const size_t K = n * R * C + k;
      const double dg_dt = (gamma[K] - gamma_prev[K]) / dt;
      v3 df = scl(scl(ld(normal + 3 * K), dg_dt), area[K]);
      Fx += df.x;
      Fy += df.y;
      Fz += df.z;
    }
    v3 F = add(Fst, mk(Fx, Fy, Fz));
    const double Q = nrm(ff);
    const double S = w->chord * w->span;
    v3 Cf = dvs(F, 0.5 * Q * Q * S);
    out[2 * n + 0] = Cf.x;
    out[2 * n + 1] = Cf.z;
  }
  free(normal);
  free(area);
  free(lines);
  free(U);
}

void orc_loads(const double* wing_nodes, const double* wing_nodes_init,
               const double* wing_gamma, const double* wing_gamma_prev,
               size_t num, size_t rows, size_t cols, const double* wake_nodes,
               const double* wake_gamma, size_t wake_rows,
               const orc_wing_spec* wings, const double ff[3], double t,
               double dt, double* out) {
  loads_impl(wing_nodes, wing_nodes_init, wing_gamma, wing_gamma_prev, num,
             rows, cols, wake_nodes, wake_gamma, wake_rows, wings, ff, t, dt,
             out, NULL);
}

/* simple_simulator.cpp:240-286 (MainLoop) and :307-311 (erase) */
int orc_sim_step(orc_sim* s, size_t step) {
  const double dt = s->cfg.dt;
  const double t = (double)step * dt;
  const size_t num = s->num, R = s->R, C = s->C, N = s->N;
  const size_t C1 = C + 1;
  memcpy(s->gamma_prev, s->gamma, sizeof(double) * N);
  /* Morphing */
  for (size_t n = 0; n < num; n++) {
    const orc_wing_spec* w = &s->cfg.wings[n];
    const size_t off = n * (R + 1) * C1;
    for (size_t k = 0; k < (R + 1) * C1; k++)
      st(s->wing + 3 * (off + k), perform(w, ld(s->wing_init + 3 * (off + k)), t));
  }
  /* Shed (:171-188) */
  {
    double* te = (double*)xcalloc(3 * num * C1, sizeof(double));
    for (size_t n = 0; n < num; n++)
      memcpy(te + 3 * n * C1, s->wing + 3 * ((n * (R + 1) + R) * C1),
             sizeof(double) * 3 * C1);
    double* nw = prepend_rows(s->wake, num, s->wake_rows, 3 * C1, te);
    free(s->wake);
    s->wake = nw;
    s->wake_rows += 1;
    free(te);
    if (step > 0) {
      double* tg = (double*)xcalloc(num * C, sizeof(double));
      for (size_t n = 0; n < num; n++)
        memcpy(tg + n * C, s->gamma + (n * R + (R - 1)) * C, sizeof(double) * C);
      double* ng = prepend_rows(s->wake_gamma, num, s->wake_grows, C, tg);
      free(s->wake_gamma);
      s->wake_gamma = ng;
      s->wake_grows += 1;
      free(tg);
    }
  }
  double* cpos = (double*)xcalloc(3 * N, sizeof(double));
  double* cpos0 = (double*)xcalloc(3 * N, sizeof(double));
  double* normal = (double*)xcalloc(3 * N, sizeof(double));
  double* A = (double*)xcalloc(N * N, sizeof(double));
  double* rhs = (double*)xcalloc(N, sizeof(double));
  orc_collocation_points(s->wing, num, R, C, cpos);
  orc_collocation_points(s->wing_init, num, R, C, cpos0);
  orc_normals(s->wing, num, R, C, normal);
  /* CalcMatrix (:76-101) */
  orc_influence_matrix(s->wing, num, R, C, cpos, normal, A);
  s->pairs += 4.0 * (double)N * (double)N;
  /* CalcRhs (:103-124) */
  {
    v3 ff = ld(s->cfg.ff);
#pragma omp parallel for schedule(dynamic, 8)
    for (size_t K = 0; K < N; K++) {
      const size_t n = K / (R * C);
      v3 Uw = induced(ld(cpos + 3 * K), s->wake, s->wake_gamma, num, s->wake_grows, C);
      v3 Uls = morph_velocity(&s->cfg.wings[n], ld(cpos0 + 3 * K), t);
      v3 u = sub(Uw, add(Uls, ff));
      rhs[K] = -dot(u, ld(normal + 3 * K));
    }
    s->pairs += 4.0 * (double)N * (double)(num * s->wake_grows * C);
  }
  /* FullPivLU (:263-268) */
  if (s->cfg.lu == ORC_LU_PARTIALPIV)
    orc_partialpiv_solve(N, A, rhs, s->gamma);
  else
    orc_fullpiv_solve(N, A, rhs, s->gamma);
  /* CalcLoad (:195-233) */
  if (s->cfg.loads) {
    loads_impl(s->wing, s->wing_init, s->gamma, s->gamma_prev, num, R, C,
               s->wake, s->wake_gamma, s->wake_grows, s->cfg.wings, s->cfg.ff,
               t, dt, s->loads, &s->pairs);
  }
  /* Advect (:190-193) */
  advect(s, dt);
  free(cpos);
  free(cpos0);
  free(normal);
  free(A);
  free(rhs);
  /* EraseOldestWake (:235-238, :307-311) */
  if (s->cfg.erase_wake_after > 0 && step >= s->cfg.erase_wake_after) {
    if (s->wake_rows > 0) {
      double* nw = erase_last(s->wake, num, s->wake_rows, 3 * C1);
      free(s->wake);
      s->wake = nw;
      s->wake_rows -= 1;
    }
    if (s->wake_grows > 0) {
      double* ng = erase_last(s->wake_gamma, num, s->wake_grows, C);
      free(s->wake_gamma);
      s->wake_gamma = ng;
      s->wake_grows -= 1;
    }
  }
  return 0;
}
