/*
 * FP64 CPU oracle: plain-C restatement of the reference calculator's hot path.
 * TEST INFRASTRUCTURE ONLY (see oracle.h for who may load it and what pins
 * each part). Compile with -ffp-contract=off so every product and sum is the
 * IEEE operation written here, in the order written here.
 */
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* ------------------------------------------------------------------------ */
/* thread pool: dynamic parallel-for over configurations                     */
/* (restates std::for_each(std::execution::par, configs...) on a oneTBB      */
/*  arena sized to the host cores, src/optimization/forcesmith_functor.cpp:  */
/*  28-51)                                                                   */
/* ------------------------------------------------------------------------ */
static int g_threads = 1;
void or_set_threads(int n) { g_threads = n < 1 ? 1 : n; }
int or_get_threads(void) { return g_threads; }

typedef void (*or_task_fn)(void *arg, int64_t item);
typedef struct {
  or_task_fn fn;
  void *arg;
  int64_t n;
  _Atomic int64_t next;
} or_pfor;

static void *pfor_worker(void *p) {
  or_pfor *pf = (or_pfor *)p;
  for (;;) {
    const int64_t i = atomic_fetch_add(&pf->next, 1);
    if (i >= pf->n) break;
    pf->fn(pf->arg, i);
  }
  return NULL;
}

static void parallel_for(int64_t n, or_task_fn fn, void *arg) {
  or_pfor pf;
  pf.fn = fn;
  pf.arg = arg;
  pf.n = n;
  atomic_init(&pf.next, 0);
  int nt = g_threads;
  if (nt > n) nt = (int)n;
  if (nt <= 1) {
    for (int64_t i = 0; i < n; ++i) fn(arg, i);
    return;
  }
  pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)(nt - 1));
  for (int t = 0; t < nt - 1; ++t) pthread_create(&th[t], NULL, pfor_worker, &pf);
  pfor_worker(&pf);
  for (int t = 0; t < nt - 1; ++t) pthread_join(th[t], NULL);
  free(th);
}

/* ------------------------------------------------------------------------ */
/* periodic box: include/forcesmith/core/boundary_conditions.hpp:9-50        */
/* ------------------------------------------------------------------------ */
typedef struct {
  double b[9];   /* b[3r+c] = box(r,c); column c = lattice vector c */
  double inv[9]; /* inverse, or identity if singular */
  double vol;    /* |det| (0 if det non-finite) */
} or_box;

static void box_from_cell(const double *cell, or_box *bx) {
  for (int k = 0; k < 3; ++k)
    for (int a = 0; a < 3; ++a) bx->b[3 * a + k] = cell[3 * k + a];
  const double *m = bx->b;
  /* cofactor expansion along the first column */
  const double c00 = m[4] * m[8] - m[5] * m[7];
  const double c10 = m[2] * m[7] - m[1] * m[8];
  const double c20 = m[1] * m[5] - m[2] * m[4];
  const double det = m[0] * c00 + m[3] * c10 + m[6] * c20;
  int finite = isfinite(det);
  for (int i = 0; i < 9; ++i) finite = finite && isfinite(m[i]);
  if (finite && det != 0.0) {
    const double id = 1.0 / det;
    bx->inv[0] = c00 * id;
    bx->inv[1] = c10 * id;
    bx->inv[2] = c20 * id;
    bx->inv[3] = (m[5] * m[6] - m[3] * m[8]) * id;
    bx->inv[4] = (m[0] * m[8] - m[2] * m[6]) * id;
    bx->inv[5] = (m[2] * m[3] - m[0] * m[5]) * id;
    bx->inv[6] = (m[3] * m[7] - m[4] * m[6]) * id;
    bx->inv[7] = (m[1] * m[6] - m[0] * m[7]) * id;
    bx->inv[8] = (m[0] * m[4] - m[1] * m[3]) * id;
  } else {
    for (int i = 0; i < 9; ++i) bx->inv[i] = (i % 4 == 0) ? 1.0 : 0.0;
  }
  bx->vol = isfinite(det) ? fabs(det) : 0.0;
}

/* wrap = box * (f - floor(f)), f = inv * r  (boundary_conditions.hpp:11-16) */
static void box_wrap(const or_box *bx, const double *r, double *w) {
  double f[3];
  for (int a = 0; a < 3; ++a) {
    const double *row = bx->inv + 3 * a;
    f[a] = row[0] * r[0] + row[1] * r[1] + row[2] * r[2];
    f[a] = f[a] - floor(f[a]);
  }
  for (int a = 0; a < 3; ++a) {
    const double *row = bx->b + 3 * a;
    w[a] = row[0] * f[0] + row[1] * f[1] + row[2] * f[2];
  }
}

/* ------------------------------------------------------------------------ */
/* makima spline: src/potentials/spline.cpp, include/.../spline.hpp         */
/* ------------------------------------------------------------------------ */
typedef struct {
  int n;
  double *x, *y, *s; /* s has n entries when n >= 4 (else unused) */
  uint8_t *fixed;
  int has_s;
} or_spline;

/* knot grid x_k = rmin + k*h, h = (rmax-rmin)/(n-1)
 * (src/io/potential_reader.cpp:176-182) */
void or_knot_grid(int n, double rmin, double rmax, double *x) {
  const double h = (rmax - rmin) / (double)(n - 1);
  for (int k = 0; k < n; ++k) x[k] = rmin + (double)k * h;
}

/* recompute_slopes_ (src/potentials/spline.cpp:43-103) */
static void slopes(int n, const double *x, const double *y, double *s) {
  double m2 = (y[3] - y[2]) / (x[3] - x[2]);
  double m1 = (y[2] - y[1]) / (x[2] - x[1]);
  double m0 = (y[1] - y[0]) / (x[1] - x[0]);
  double mm1 = 2 * m0 - m1;
  double mm2 = 2 * mm1 - m0;
  double w1 = fabs(m1 - m0) + fabs(m1 + m0) / 2;
  double w2 = fabs(mm1 - mm2) + fabs(mm1 + mm2) / 2;
  s[0] = (w1 * mm1 + w2 * m0) / (w1 + w2);
  if (isnan(s[0])) s[0] = 0;
  w1 = fabs(m2 - m1) + fabs(m2 + m1) / 2;
  w2 = fabs(m0 - mm1) + fabs(m0 + mm1) / 2;
  s[1] = (w1 * m0 + w2 * m1) / (w1 + w2);
  if (isnan(s[1])) s[1] = 0;
  for (int i = 2; i < n - 2; ++i) {
    double mim2 = (y[i - 1] - y[i - 2]) / (x[i - 1] - x[i - 2]);
    double mim1 = (y[i] - y[i - 1]) / (x[i] - x[i - 1]);
    double mi = (y[i + 1] - y[i]) / (x[i + 1] - x[i]);
    double mip1 = (y[i + 2] - y[i + 1]) / (x[i + 2] - x[i + 1]);
    w1 = fabs(mip1 - mi) + fabs(mip1 + mi) / 2;
    w2 = fabs(mim1 - mim2) + fabs(mim1 + mim2) / 2;
    s[i] = (w1 * mim1 + w2 * mi) / (w1 + w2);
    if (isnan(s[i])) s[i] = 0;
  }
  double mnm4 = (y[n - 3] - y[n - 4]) / (x[n - 3] - x[n - 4]);
  double mnm3 = (y[n - 2] - y[n - 3]) / (x[n - 2] - x[n - 3]);
  double mnm2 = (y[n - 1] - y[n - 2]) / (x[n - 1] - x[n - 2]);
  double mnm1 = 2 * mnm2 - mnm3;
  double mn = 2 * mnm1 - mnm2;
  w1 = fabs(mnm1 - mnm2) + fabs(mnm1 + mnm2) / 2;
  w2 = fabs(mnm3 - mnm4) + fabs(mnm3 + mnm4) / 2;
  s[n - 2] = (w1 * mnm3 + w2 * mnm2) / (w1 + w2);
  if (isnan(s[n - 2])) s[n - 2] = 0;
  w1 = fabs(mn - mnm1) + fabs(mn + mnm1) / 2;
  w2 = fabs(mnm2 - mnm3) + fabs(mnm2 + mnm3) / 2;
  s[n - 1] = (w1 * mnm2 + w2 * mnm1) / (w1 + w2);
  if (isnan(s[n - 1])) s[n - 1] = 0;
}

int or_spline_slopes(int n, const double *x, const double *y, double *s) {
  if (n < 4) return 0;
  slopes(n, x, y, s);
  return 1;
}

static void spline_update(or_spline *sp) {
  sp->has_s = sp->n >= 4;
  if (sp->has_s) slopes(sp->n, sp->x, sp->y, sp->s);
}

/* locate_interval: upper_bound - 1 clamped to [0, n-2] (spline.cpp:18-30) */
static int locate(const double *x, int n, double r) {
  int lo = 0, hi = n; /* first index with x > r */
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (x[mid] > r) hi = mid; else lo = mid + 1;
  }
  int i = lo;
  if (i > 0) --i;
  if (i > n - 2) i = n - 2;
  return i;
}

/* Hermite value/derivative (spline.cpp:130-135, 183-189) */
static double herm_value(const double *x, const double *y, const double *s,
                         int i, double r) {
  const double x0 = x[i], dx = x[i + 1] - x[i];
  const double t = (r - x0) / dx;
  const double omt = 1.0 - t;
  return omt * omt * (y[i] * (1.0 + 2.0 * t) + s[i] * (r - x0)) +
         t * t * (y[i + 1] * (3.0 - 2.0 * t) + dx * s[i + 1] * (t - 1.0));
}
static double herm_deriv(const double *x, const double *y, const double *s,
                         int i, double r) {
  const double x0 = x[i], dx = x[i + 1] - x[i];
  const double t = (r - x0) / dx;
  const double omt = 1.0 - t;
  /* d/dr of the value expression, product rule with dt/dr = 1/dx */
  const double a = y[i] * (1.0 + 2.0 * t) + s[i] * (r - x0);
  const double da = y[i] * 2.0 / dx + s[i];
  const double b = y[i + 1] * (3.0 - 2.0 * t) + dx * s[i + 1] * (t - 1.0);
  const double db = -y[i + 1] * 2.0 / dx + s[i + 1];
  return -2.0 * omt / dx * a + omt * omt * da + 2.0 * t / dx * b + t * t * db;
}

/* basis of the value / r-derivative / second r-derivative w.r.t.
 * (y_i, y_{i+1}, s_i, s_{i+1}) (the ddx jacobian of spline.cpp:163-179) */
typedef struct { double c[4]; } basis4;
static void herm_basis(double x0, double dx, double r, basis4 *v, basis4 *d,
                       basis4 *d2) {
  const double t = (r - x0) / dx;
  const double omt = 1.0 - t;
  if (v) {
    v->c[0] = omt * omt * (1.0 + 2.0 * t);
    v->c[1] = t * t * (3.0 - 2.0 * t);
    v->c[2] = omt * omt * (r - x0);
    v->c[3] = t * t * (dx * (t - 1.0));
  }
  if (d) {
    d->c[0] = -6.0 * t * omt / dx;
    d->c[1] = 6.0 * t * omt / dx;
    d->c[2] = omt * (1.0 - 3.0 * t);
    d->c[3] = t * (3.0 * t - 2.0);
  }
  if (d2) {
    d2->c[0] = -6.0 * (1.0 - 2.0 * t) / (dx * dx);
    d2->c[1] = 6.0 * (1.0 - 2.0 * t) / (dx * dx);
    d2->c[2] = (6.0 * t - 4.0) / dx;
    d2->c[3] = (6.0 * t - 2.0) / dx;
  }
}

/* uncached eval / deriv (spline.cpp:215-279) */
static void sp_eval_uncached(const or_spline *sp, double r, double *v, double *d) {
  const int n = sp->n;
  const double *x = sp->x, *y = sp->y;
  if (r <= x[0]) {
    const double slope = (y[1] - y[0]) / (x[1] - x[0]);
    *v = y[0] + slope * (r - x[0]);
    *d = slope;
  } else if (r >= x[n - 1]) {
    const double slope = (y[n - 1] - y[n - 2]) / (x[n - 1] - x[n - 2]);
    *v = y[n - 1] + slope * (r - x[n - 1]);
    *d = slope;
  } else if (!sp->has_s) {
    const int i = locate(x, n, r);
    const double t = (r - x[i]) / (x[i + 1] - x[i]);
    *v = y[i] + t * (y[i + 1] - y[i]);
    *d = (y[i + 1] - y[i]) / (x[i + 1] - x[i]);
  } else {
    const int i = locate(x, n, r);
    *v = herm_value(x, y, sp->s, i, r);
    *d = herm_deriv(x, y, sp->s, i, r);
  }
}

/* cached site (EvalSite, spline.hpp:111-125; prepare_site spline.cpp:281-324) */
enum { SITE_NONE = 0, SITE_CUBIC = 1, SITE_LINEAR = 2 };
typedef struct {
  int32_t kind, i;
  double off, inv_dx;
  double v[4], d[4];
} or_site;

static void sp_prepare_site(const or_spline *sp, double r, or_site *st) {
  const int n = sp->n;
  const double *x = sp->x;
  if (r <= x[0]) {
    st->kind = SITE_LINEAR;
    st->i = 0;
  } else if (r >= x[n - 1]) {
    st->kind = SITE_LINEAR;
    st->i = n - 2;
  } else if (!sp->has_s) {
    st->kind = SITE_LINEAR;
    st->i = locate(x, n, r);
  } else {
    st->kind = SITE_CUBIC;
    st->i = locate(x, n, r);
  }
  const int i = st->i;
  const double dx = x[i + 1] - x[i];
  st->inv_dx = 1.0 / dx;
  st->off = r - x[i];
  memset(st->v, 0, sizeof st->v);
  memset(st->d, 0, sizeof st->d);
  if (st->kind == SITE_CUBIC) {
    basis4 bv, bd;
    herm_basis(x[i], dx, r, &bv, &bd, NULL);
    memcpy(st->v, bv.c, sizeof st->v);
    memcpy(st->d, bd.c, sizeof st->d);
  }
}

/* eval_and_deriv_at (spline.hpp:80-98) */
static void sp_eval_site(const or_spline *sp, const or_site *st, double *v,
                         double *d) {
  const int i = st->i;
  const double *y = sp->y;
  if (st->kind == SITE_LINEAR) {
    *v = y[i] + st->off * st->inv_dx * (y[i + 1] - y[i]);
    *d = (y[i + 1] - y[i]) * st->inv_dx;
  } else {
    const double yi = y[i], yi1 = y[i + 1], si = sp->s[i], si1 = sp->s[i + 1];
    *v = st->v[0] * yi + st->v[1] * yi1 + st->v[2] * si + st->v[3] * si1;
    *d = st->d[0] * yi + st->d[1] * yi1 + st->d[2] * si + st->d[3] * si1;
  }
}

static int sp_in_range(const or_spline *sp, double r) {
  return r >= sp->x[0] && r <= sp->x[sp->n - 1];
}

/* d s_k / d y_m (dense n x n) by the chain rule through the makima weights.
 * Secants M[-2..n] (extended by the quadratic extrapolation of
 * spline.cpp:56-57,88-89); s_k = (w1 M[k-1] + w2 M[k])/(w1+w2) with
 * w1 = |M[k+1]-M[k]| + |M[k+1]+M[k]|/2, w2 = |M[k-1]-M[k-2]| + |M[k-1]+M[k-2]|/2.
 * Convention: d|u|/du = sign(u) with sign(0) = 0; w1+w2 == 0 gives zero rows
 * (the NaN -> 0 branch of spline.cpp:61). */
static double sgn(double u) { return (u > 0) - (u < 0); }

static void slope_jacobian(int n, const double *x, const double *y, double *S) {
  memset(S, 0, sizeof(double) * (size_t)n * (size_t)n);
  if (n < 4) return;
  const int nm = n + 3; /* M index q in [-2, n] stored at q+2 */
  double *M = (double *)calloc((size_t)nm, sizeof(double));
  double *dM = (double *)calloc((size_t)nm * (size_t)n, sizeof(double));
#define MI(q) M[(q) + 2]
#define DM(q, m) dM[(size_t)((q) + 2) * (size_t)n + (size_t)(m)]
  for (int k = 0; k <= n - 2; ++k) {
    const double h = x[k + 1] - x[k];
    MI(k) = (y[k + 1] - y[k]) / h;
    DM(k, k) = -1.0 / h;
    DM(k, k + 1) = 1.0 / h;
  }
  MI(-1) = 2 * MI(0) - MI(1);
  MI(-2) = 2 * MI(-1) - MI(0);
  MI(n - 1) = 2 * MI(n - 2) - MI(n - 3);
  MI(n) = 2 * MI(n - 1) - MI(n - 2);
  for (int m = 0; m < n; ++m) {
    DM(-1, m) = 2 * DM(0, m) - DM(1, m);
    DM(-2, m) = 2 * DM(-1, m) - DM(0, m);
    DM(n - 1, m) = 2 * DM(n - 2, m) - DM(n - 3, m);
    DM(n, m) = 2 * DM(n - 1, m) - DM(n - 2, m);
  }
  for (int k = 0; k < n; ++k) {
    const double A = MI(k - 1), B = MI(k), Cn = MI(k + 1), P = MI(k - 2);
    const double w1 = fabs(Cn - B) + fabs(Cn + B) / 2;
    const double w2 = fabs(A - P) + fabs(A + P) / 2;
    const double W = w1 + w2;
    if (!(W > 0)) continue;
    const double s = (w1 * A + w2 * B) / W;
    const double dw1_dC = sgn(Cn - B) + sgn(Cn + B) / 2;
    const double dw1_dB = -sgn(Cn - B) + sgn(Cn + B) / 2;
    const double dw2_dA = sgn(A - P) + sgn(A + P) / 2;
    const double dw2_dP = -sgn(A - P) + sgn(A + P) / 2;
    const double ds_dw1 = (A - s) / W, ds_dw2 = (B - s) / W;
    const double gP = ds_dw2 * dw2_dP;
    const double gA = w1 / W + ds_dw2 * dw2_dA;
    const double gB = w2 / W + ds_dw1 * dw1_dB;
    const double gC = ds_dw1 * dw1_dC;
    for (int m = 0; m < n; ++m)
      S[(size_t)k * n + m] = gP * DM(k - 2, m) + gA * DM(k - 1, m) +
                             gB * DM(k, m) + gC * DM(k + 1, m);
  }
#undef MI
#undef DM
  free(M);
  free(dM);
}

void or_spline_slope_jacobian(int n, const double *x, const double *y,
                              double *S) {
  slope_jacobian(n, x, y, S);
}

/* Partials of value, r-derivative and second r-derivative w.r.t. every knot
 * value y_m at r, uncached semantics (the spline counterpart of
 * param_grad / dderiv_dparam / deriv2 used by the analytic chain of
 * src/force/eam_force.cpp:108-279). S = slope jacobian of the table. */
static void sp_param_partials(const or_spline *sp, const double *S, double r,
                              double *dv, double *dd, double *d2) {
  const int n = sp->n;
  const double *x = sp->x;
  for (int m = 0; m < n; ++m) {
    if (dv) dv[m] = 0;
    if (dd) dd[m] = 0;
  }
  if (d2) *d2 = 0;
  if (r <= x[0] || r >= x[n - 1] || !sp->has_s) {
    int i;
    if (r <= x[0]) i = 0;
    else if (r >= x[n - 1]) i = n - 2;
    else i = locate(x, n, r);
    const double h = x[i + 1] - x[i];
    const double anchor = (r >= x[n - 1] && r > x[0]) ? x[n - 1] : x[i];
    /* value = y_a + slope * (r - x_a) with a = i (or n-1 at the top edge) */
    const double slope_di = -1.0 / h, slope_di1 = 1.0 / h;
    if (dv) {
      if (r >= x[n - 1] && r > x[0]) {
        dv[n - 1] += 1.0;
      } else {
        dv[i] += 1.0;
      }
      dv[i] += slope_di * (r - anchor);
      dv[i + 1] += slope_di1 * (r - anchor);
    }
    if (dd) {
      dd[i] += slope_di;
      dd[i + 1] += slope_di1;
    }
    return;
  }
  const int i = locate(x, n, r);
  basis4 bv, bd, b2;
  herm_basis(x[i], x[i + 1] - x[i], r, &bv, &bd, &b2);
  for (int m = 0; m < n; ++m) {
    const double si = S[(size_t)i * n + m], si1 = S[(size_t)(i + 1) * n + m];
    double a = bv.c[2] * si + bv.c[3] * si1;
    double b = bd.c[2] * si + bd.c[3] * si1;
    if (m == i) { a += bv.c[0]; b += bd.c[0]; }
    if (m == i + 1) { a += bv.c[1]; b += bd.c[1]; }
    if (dv) dv[m] = a;
    if (dd) dd[m] = b;
  }
  if (d2) {
    const double *y = sp->y, *s = sp->s;
    *d2 = b2.c[0] * y[i] + b2.c[1] * y[i + 1] + b2.c[2] * s[i] +
          b2.c[3] * s[i + 1];
  }
}

/* Public wrappers for unit tests */
static void tmp_spline(or_spline *sp, int n, const double *x, const double *y,
                       const double *s) {
  sp->n = n;
  sp->x = (double *)x;
  sp->y = (double *)y;
  sp->s = (double *)s;
  sp->has_s = n >= 4;
  sp->fixed = NULL;
}
void or_spline_eval(int n, const double *x, const double *y, const double *s,
                    double r, double *v, double *d) {
  or_spline sp;
  tmp_spline(&sp, n, x, y, s);
  sp_eval_uncached(&sp, r, v, d);
}
void or_spline_eval_cached(int n, const double *x, const double *y,
                           const double *s, double r, double *v, double *d) {
  or_spline sp;
  tmp_spline(&sp, n, x, y, s);
  or_site st;
  sp_prepare_site(&sp, r, &st);
  sp_eval_site(&sp, &st, v, d);
}
void or_spline_param_grad(int n, const double *x, const double *y, double r,
                          double *dv, double *dd, double *d2) {
  double *s = (double *)calloc((size_t)n, sizeof(double));
  double *S = (double *)calloc((size_t)n * (size_t)n, sizeof(double));
  or_spline sp;
  tmp_spline(&sp, n, x, y, s);
  if (n >= 4) slopes(n, x, y, s);
  slope_jacobian(n, x, y, S);
  sp_param_partials(&sp, S, r, dv, dd, d2);
  free(s);
  free(S);
}

/* ------------------------------------------------------------------------ */
/* context                                                                   */
/* ------------------------------------------------------------------------ */
#define NROLE_MAX 5
enum { ROLE_PHI = 0, ROLE_GI = 1, ROLE_GJ = 2, ROLE_DIP = 3, ROLE_QUAD = 4 };

typedef struct {
  double A, B, p, q, delta, a1, gamma, a2;
} or_sw;

#define OR_MAXT 8

struct or_ctx {
  int kind, ntypes, npair;
  int C;
  int64_t N;
  int64_t *atom_off;
  int32_t *types;
  int32_t *cfg_of; /* config of each atom */
  double *pos, *ref_force, *ref_energy, *ref_stress;
  or_box *box;
  /* tables */
  int ntab;
  or_spline *tab;
  double *Sjac; /* per table n*n slope jacobian, refreshed with the params */
  int64_t *Soff;
  /* SW */
  or_sw *sw;
  uint8_t *sw_fixed; /* npair * 8 */
  double *lam;
  uint8_t *lam_fixed;
  int nlam;
  /* weights */
  double ew, sw_w;
  double smooth_w; /* curvature-row weight, 0 = no rows */
  int64_t *row_off;
  /* neighbour list */
  double nl_rc;
  int64_t npairs;
  int64_t *pair_off;
  int32_t *nj;
  int32_t *img;
  double *nd;
  or_site *sites; /* npairs * nrole, kind SITE_NONE when not cacheable */
  int nrole;
  /* outputs / scratch */
  double *force, *energy, *stress, *limit;
  double *rho, *gradF, *mu, *lam9;
  /* EAM gauge: density scale, per-type embedding slope and intercept
   * (identity: gA = 1, gS = gF0 = 0) */
  int gauge;
  double gA, gS[OR_MAXT], gF0[OR_MAXT];
};

or_ctx *or_create(int kind, int ntypes, int n_configs, const int64_t *atom_off,
                  const int32_t *types, const double *pos,
                  const double *ref_force, const double *cell,
                  const double *ref_energy, const double *ref_stress) {
  or_ctx *c = (or_ctx *)calloc(1, sizeof(or_ctx));
  c->kind = kind;
  c->ntypes = ntypes;
  c->npair = ntypes * (ntypes + 1) / 2;
  c->C = n_configs;
  c->N = atom_off[n_configs];
  c->atom_off = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n_configs + 1));
  memcpy(c->atom_off, atom_off, sizeof(int64_t) * (size_t)(n_configs + 1));
  const size_t N = (size_t)c->N;
  c->types = (int32_t *)malloc(sizeof(int32_t) * (N ? N : 1));
  memcpy(c->types, types, sizeof(int32_t) * N);
  c->cfg_of = (int32_t *)malloc(sizeof(int32_t) * (N ? N : 1));
  for (int k = 0; k < n_configs; ++k)
    for (int64_t a = atom_off[k]; a < atom_off[k + 1]; ++a) c->cfg_of[a] = k;
  c->pos = (double *)malloc(sizeof(double) * 3 * (N ? N : 1));
  memcpy(c->pos, pos, sizeof(double) * 3 * N);
  c->ref_force = (double *)malloc(sizeof(double) * 3 * (N ? N : 1));
  memcpy(c->ref_force, ref_force, sizeof(double) * 3 * N);
  c->ref_energy = (double *)malloc(sizeof(double) * (size_t)(n_configs + 1));
  memcpy(c->ref_energy, ref_energy, sizeof(double) * (size_t)n_configs);
  c->ref_stress = (double *)malloc(sizeof(double) * 6 * (size_t)(n_configs + 1));
  memcpy(c->ref_stress, ref_stress, sizeof(double) * 6 * (size_t)n_configs);
  c->box = (or_box *)malloc(sizeof(or_box) * (size_t)(n_configs + 1));
  for (int k = 0; k < n_configs; ++k) box_from_cell(cell + 9 * k, &c->box[k]);
  c->force = (double *)calloc(3 * (N ? N : 1), sizeof(double));
  c->energy = (double *)calloc((size_t)n_configs + 1, sizeof(double));
  c->stress = (double *)calloc(9 * ((size_t)n_configs + 1), sizeof(double));
  c->limit = (double *)calloc((size_t)n_configs + 1, sizeof(double));
  c->rho = (double *)calloc(N ? N : 1, sizeof(double));
  c->gradF = (double *)calloc(N ? N : 1, sizeof(double));
  c->mu = (double *)calloc(3 * (N ? N : 1), sizeof(double));
  c->lam9 = (double *)calloc(9 * (N ? N : 1), sizeof(double));
  c->ew = 1.0;
  c->sw_w = 0.0;
  c->gA = 1.0;
  c->nl_rc = -1.0;
  c->nrole = kind == OR_ADP ? 5 : (kind == OR_EAM ? 3 : 0);
  c->row_off = (int64_t *)calloc((size_t)n_configs + 1, sizeof(int64_t));
  or_set_weights(c, 1.0, 0.0);
  return c;
}

static void free_nl(or_ctx *c) {
  free(c->pair_off);
  free(c->nj);
  free(c->img);
  free(c->nd);
  free(c->sites);
  c->pair_off = NULL;
  c->nj = NULL;
  c->img = NULL;
  c->nd = NULL;
  c->sites = NULL;
}

void or_destroy(or_ctx *c) {
  if (!c) return;
  free_nl(c);
  for (int t = 0; t < c->ntab; ++t) {
    free(c->tab[t].x);
    free(c->tab[t].y);
    free(c->tab[t].s);
    free(c->tab[t].fixed);
  }
  free(c->tab);
  free(c->Sjac);
  free(c->Soff);
  free(c->sw);
  free(c->sw_fixed);
  free(c->lam);
  free(c->lam_fixed);
  free(c->atom_off);
  free(c->types);
  free(c->cfg_of);
  free(c->pos);
  free(c->ref_force);
  free(c->ref_energy);
  free(c->ref_stress);
  free(c->box);
  free(c->force);
  free(c->energy);
  free(c->stress);
  free(c->limit);
  free(c->rho);
  free(c->gradF);
  free(c->mu);
  free(c->lam9);
  free(c->row_off);
  free(c);
}

/* index helpers (include/forcesmith/force/potential_table.hpp:31-37) */
static int pair_ordinal(int a, int b, int nt) {
  if (a > b) { int t = a; a = b; b = t; }
  return a * nt - a * (a - 1) / 2 + (b - a);
}

static int tab_pair(const or_ctx *c, int ti, int tj) {
  return pair_ordinal(ti, tj, c->ntypes);
}
static int tab_dens(const or_ctx *c, int t) { return c->npair + t; }
static int tab_emb(const or_ctx *c, int t) { return c->npair + c->ntypes + t; }
static int tab_dip(const or_ctx *c, int ti, int tj) {
  return c->npair + 2 * c->ntypes + pair_ordinal(ti, tj, c->ntypes);
}
static int tab_quad(const or_ctx *c, int ti, int tj) {
  return 2 * c->npair + 2 * c->ntypes + pair_ordinal(ti, tj, c->ntypes);
}

static void refresh_tables(or_ctx *c) {
  for (int t = 0; t < c->ntab; ++t) {
    spline_update(&c->tab[t]);
    slope_jacobian(c->tab[t].n, c->tab[t].x, c->tab[t].y, c->Sjac + c->Soff[t]);
  }
}

int or_set_tables(or_ctx *c, int ntab, const int32_t *n, const double *rmin,
                  const double *rmax, const double *y, const uint8_t *fixed) {
  const int want = c->kind == OR_ADP ? 3 * c->npair + 2 * c->ntypes
                                     : c->npair + 2 * c->ntypes;
  if (c->kind == OR_SW || ntab != want) return -1;
  c->ntab = ntab;
  c->tab = (or_spline *)calloc((size_t)ntab, sizeof(or_spline));
  c->Soff = (int64_t *)calloc((size_t)ntab + 1, sizeof(int64_t));
  int64_t off = 0, soff = 0;
  for (int t = 0; t < ntab; ++t) {
    or_spline *sp = &c->tab[t];
    if (n[t] < 2) return -2;
    sp->n = n[t];
    sp->x = (double *)malloc(sizeof(double) * (size_t)n[t]);
    sp->y = (double *)malloc(sizeof(double) * (size_t)n[t]);
    sp->s = (double *)calloc((size_t)n[t], sizeof(double));
    sp->fixed = (uint8_t *)malloc((size_t)n[t]);
    or_knot_grid(n[t], rmin[t], rmax[t], sp->x);
    memcpy(sp->y, y + off, sizeof(double) * (size_t)n[t]);
    memcpy(sp->fixed, fixed + off, (size_t)n[t]);
    off += n[t];
    c->Soff[t] = soff;
    soff += (int64_t)n[t] * n[t];
  }
  c->Soff[ntab] = soff;
  c->Sjac = (double *)calloc((size_t)soff, sizeof(double));
  refresh_tables(c);
  return 0;
}

int or_set_sw(or_ctx *c, const double *pp, const uint8_t *pf, const double *lam,
              const uint8_t *lf) {
  if (c->kind != OR_SW) return -1;
  c->sw = (or_sw *)malloc(sizeof(or_sw) * (size_t)c->npair);
  c->sw_fixed = (uint8_t *)malloc(8 * (size_t)c->npair);
  for (int s = 0; s < c->npair; ++s) {
    const double *q = pp + 8 * s;
    or_sw v = {q[0], q[1], q[2], q[3], q[4], q[5], q[6], q[7]};
    c->sw[s] = v;
  }
  memcpy(c->sw_fixed, pf, 8 * (size_t)c->npair);
  c->nlam = c->ntypes * c->npair;
  c->lam = (double *)malloc(sizeof(double) * (size_t)c->nlam);
  c->lam_fixed = (uint8_t *)malloc((size_t)c->nlam);
  memcpy(c->lam, lam, sizeof(double) * (size_t)c->nlam);
  memcpy(c->lam_fixed, lf, (size_t)c->nlam);
  return 0;
}

static double *sw_field(or_sw *p, int f) {
  switch (f) {
  case 0: return &p->A;
  case 1: return &p->B;
  case 2: return &p->p;
  case 3: return &p->q;
  case 4: return &p->delta;
  case 5: return &p->a1;
  case 6: return &p->gamma;
  default: return &p->a2;
  }
}

/* TableParams / StiwebForceCalculator::param_count
 * (force_calculator_concept.hpp:171-208, stiweb_force.cpp:169-211) */
int or_param_count(const or_ctx *c) {
  int n = 0;
  if (c->kind == OR_SW) {
    for (int i = 0; i < 8 * c->npair; ++i) n += !c->sw_fixed[i];
    for (int i = 0; i < c->nlam; ++i) n += !c->lam_fixed[i];
    return n;
  }
  for (int t = 0; t < c->ntab; ++t)
    for (int k = 0; k < c->tab[t].n; ++k) n += !c->tab[t].fixed[k];
  return n;
}

void or_gather_params(const or_ctx *c, double *x) {
  int o = 0;
  if (c->kind == OR_SW) {
    for (int s = 0; s < c->npair; ++s)
      for (int f = 0; f < 8; ++f)
        if (!c->sw_fixed[8 * s + f]) x[o++] = *sw_field(&c->sw[s], f);
    for (int i = 0; i < c->nlam; ++i)
      if (!c->lam_fixed[i]) x[o++] = c->lam[i];
    return;
  }
  for (int t = 0; t < c->ntab; ++t)
    for (int k = 0; k < c->tab[t].n; ++k)
      if (!c->tab[t].fixed[k]) x[o++] = c->tab[t].y[k];
}

void or_scatter_params(or_ctx *c, const double *x) {
  int o = 0;
  if (c->kind == OR_SW) {
    for (int s = 0; s < c->npair; ++s)
      for (int f = 0; f < 8; ++f)
        if (!c->sw_fixed[8 * s + f]) *sw_field(&c->sw[s], f) = x[o++];
    for (int i = 0; i < c->nlam; ++i)
      if (!c->lam_fixed[i]) c->lam[i] = x[o++];
    return;
  }
  for (int t = 0; t < c->ntab; ++t)
    for (int k = 0; k < c->tab[t].n; ++k)
      if (!c->tab[t].fixed[k]) c->tab[t].y[k] = x[o++];
  refresh_tables(c);
}

/* max_cutoff over radial tables (force_calculator_concept.hpp:210-221);
 * SW: max(a1, a2) over pair slots (stiweb_force.cpp:234-239) */
double or_max_cutoff(const or_ctx *c) {
  double m = 0.0;
  if (c->kind == OR_SW) {
    for (int s = 0; s < c->npair; ++s) {
      const double v = c->sw[s].a1 > c->sw[s].a2 ? c->sw[s].a1 : c->sw[s].a2;
      if (v > m) m = v;
    }
    return m;
  }
  for (int t = 0; t < c->ntab; ++t) {
    const int is_emb = t >= c->npair + c->ntypes && t < c->npair + 2 * c->ntypes;
    if (is_emb) continue;
    const double v = c->tab[t].x[c->tab[t].n - 1];
    if (v > m) m = v;
  }
  return m;
}

/* config_residual_count / row offsets (residual_layout.hpp:18-30,
 * forcesmith_functor.cpp:208-215) */
void or_set_weights(or_ctx *c, double ew, double sw) {
  c->ew = ew;
  c->sw_w = sw;
  c->row_off[0] = 0;
  for (int k = 0; k < c->C; ++k) {
    const int64_t na = c->atom_off[k + 1] - c->atom_off[k];
    c->row_off[k + 1] = c->row_off[k] + 3 * na + 2 + (sw > 0.0 ? 6 : 0);
  }
}
/* curvature rows: SplinePotential::curvature_count (spline.hpp:104-106) is
 * n - 2 for a table with at least one free knot and n >= 3, summed over the
 * tables in declaration order (smoothness.hpp:17-23); analytic SW slots have
 * none. The functor adds them only when smooth_weight > 0
 * (forcesmith_functor.cpp:202-207). */
int64_t or_smooth_count(const or_ctx *c) {
  if (!(c->smooth_w > 0.0) || c->kind == OR_SW) return 0;
  int64_t n = 0;
  for (int t = 0; t < c->ntab; ++t) {
    int nfree = 0;
    for (int k = 0; k < c->tab[t].n; ++k) nfree += !c->tab[t].fixed[k];
    if (nfree > 0 && c->tab[t].n >= 3) n += c->tab[t].n - 2;
  }
  return n;
}
void or_set_smooth_weight(or_ctx *c, double w) { c->smooth_w = w; }
int64_t or_residual_count(const or_ctx *c) { return c->row_off[c->C] + or_smooth_count(c); }
void or_row_offsets(const or_ctx *c, int64_t *r) {
  memcpy(r, c->row_off, sizeof(int64_t) * (size_t)(c->C + 1));
}

/* ------------------------------------------------------------------------ */
/* neighbour list: build_periodic (src/core/neighbor_list.cpp:59-92),        */
/* build_impl (:105-131)                                                     */
/* ------------------------------------------------------------------------ */
typedef struct {
  or_ctx *c;
  double rc;
  int64_t *count; /* per atom */
  int fill;
} nl_job;

static int shells(const or_box *bx, int axis, double rcut) {
  const double *row = bx->inv + 3 * axis;
  const double w = sqrt(row[0] * row[0] + row[1] * row[1] + row[2] * row[2]);
  if (!isfinite(w) || w <= 0.0) return 0;
  return (int)ceil(rcut * w);
}

static void nl_config(void *arg, int64_t k) {
  nl_job *jb = (nl_job *)arg;
  or_ctx *c = jb->c;
  const double rc = jb->rc;
  const int64_t a0 = c->atom_off[k], na = c->atom_off[k + 1] - a0;
  if (!(rc > 0.0 && isfinite(rc))) {
    if (!jb->fill)
      for (int64_t i = 0; i < na; ++i) jb->count[a0 + i] = 0;
    return;
  }
  const double rc2 = rc * rc;
  const or_box *bx = &c->box[k];
  const int sx = shells(bx, 0, rc), sy = shells(bx, 1, rc), sz = shells(bx, 2, rc);
  double *w = (double *)malloc(sizeof(double) * 3 * (size_t)(na ? na : 1));
  for (int64_t i = 0; i < na; ++i) box_wrap(bx, c->pos + 3 * (a0 + i), w + 3 * i);
  const double *A = bx->b; /* columns are lattice vectors */
  for (int64_t i = 0; i < na; ++i) {
    int64_t cnt = 0;
    int64_t o = jb->fill ? c->pair_off[a0 + i] : 0;
    for (int64_t j = 0; j < na; ++j) {
      double base[3];
      for (int a = 0; a < 3; ++a) base[a] = w[3 * j + a] - w[3 * i + a];
      for (int ix = -sx; ix <= sx; ++ix)
        for (int iy = -sy; iy <= sy; ++iy)
          for (int iz = -sz; iz <= sz; ++iz) {
            if (i == j && ix == 0 && iy == 0 && iz == 0) continue;
            double d[3];
            for (int a = 0; a < 3; ++a) {
              d[a] = base[a] + (double)ix * A[3 * a + 0];
              d[a] = d[a] + (double)iy * A[3 * a + 1];
              d[a] = d[a] + (double)iz * A[3 * a + 2];
            }
            const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            if (r2 > rc2) continue;
            if (jb->fill) {
              c->nj[o] = (int32_t)(a0 + j);
              c->img[3 * o + 0] = ix;
              c->img[3 * o + 1] = iy;
              c->img[3 * o + 2] = iz;
              memcpy(c->nd + 3 * o, d, sizeof d);
              ++o;
            }
            ++cnt;
          }
    }
    if (!jb->fill) jb->count[a0 + i] = cnt;
  }
  free(w);
}

static void prime_sites(or_ctx *c);

int64_t or_build_nl(or_ctx *c, double rcut) {
  free_nl(c);
  const size_t N = (size_t)c->N;
  int64_t *cnt = (int64_t *)calloc(N + 1, sizeof(int64_t));
  nl_job jb = {c, rcut, cnt, 0};
  parallel_for(c->C, nl_config, &jb);
  c->pair_off = (int64_t *)malloc(sizeof(int64_t) * (N + 1));
  c->pair_off[0] = 0;
  for (size_t a = 0; a < N; ++a) c->pair_off[a + 1] = c->pair_off[a] + cnt[a];
  c->npairs = c->pair_off[N];
  const size_t P = (size_t)c->npairs;
  c->nj = (int32_t *)malloc(sizeof(int32_t) * (P ? P : 1));
  c->img = (int32_t *)malloc(sizeof(int32_t) * 3 * (P ? P : 1));
  c->nd = (double *)malloc(sizeof(double) * 3 * (P ? P : 1));
  jb.fill = 1;
  parallel_for(c->C, nl_config, &jb);
  free(cnt);
  c->nl_rc = rcut;
  return c->npairs;
}

void or_get_nl(const or_ctx *c, int64_t *pair_off, int32_t *j, int32_t *img,
               double *d) {
  const size_t N = (size_t)c->N, P = (size_t)c->npairs;
  if (pair_off) memcpy(pair_off, c->pair_off, sizeof(int64_t) * (N + 1));
  if (j) memcpy(j, c->nj, sizeof(int32_t) * P);
  if (img) memcpy(img, c->img, sizeof(int32_t) * 3 * P);
  if (d) memcpy(d, c->nd, sizeof(double) * 3 * P);
}

static double pair_r(const or_ctx *c, int64_t p) {
  const double *d = c->nd + 3 * p;
  return sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
}

/* prime_site for each role: EAMForceCalculator::prepare (eam_force.cpp:
 * 281-296), ADPForceCalculator::prepare (adp_force.cpp:387-406);
 * prime_site = in_range ? prepare_site : none (kernels.hpp:14-16) */
static int role_table(const or_ctx *c, int role, int ti, int tj) {
  switch (role) {
  case ROLE_PHI: return tab_pair(c, ti, tj);
  case ROLE_GI: return tab_dens(c, ti);
  case ROLE_GJ: return tab_dens(c, tj);
  case ROLE_DIP: return tab_dip(c, ti, tj);
  default: return tab_quad(c, ti, tj);
  }
}

static void prime_sites(or_ctx *c) {
  free(c->sites);
  c->sites = NULL;
  if (c->nrole == 0) return;
  const size_t P = (size_t)c->npairs;
  c->sites = (or_site *)calloc((P ? P : 1) * (size_t)c->nrole, sizeof(or_site));
  for (int64_t i = 0; i < c->N; ++i) {
    const int ti = c->types[i];
    for (int64_t p = c->pair_off[i]; p < c->pair_off[i + 1]; ++p) {
      const int tj = c->types[c->nj[p]];
      const double r = pair_r(c, p);
      for (int role = 0; role < c->nrole; ++role) {
        const or_spline *sp = &c->tab[role_table(c, role, ti, tj)];
        or_site *st = &c->sites[(size_t)p * c->nrole + role];
        if (sp_in_range(sp, r)) sp_prepare_site(sp, r, st);
        else st->kind = SITE_NONE;
      }
    }
  }
}

int64_t or_prepare(or_ctx *c) {
  or_build_nl(c, or_max_cutoff(c));
  prime_sites(c);
  return c->npairs;
}

/* eval_gated (eam_force.cpp:19-43): cached site -> eval_at; otherwise the
 * pair is out of range of the table and contributes 0 */
static void gated(const or_ctx *c, int64_t p, int role, int tab, double r,
                  double *v, double *d) {
  const or_site *st = &c->sites[(size_t)p * c->nrole + role];
  if (st->kind != SITE_NONE) {
    sp_eval_site(&c->tab[tab], st, v, d);
  } else if (sp_in_range(&c->tab[tab], r)) {
    sp_eval_uncached(&c->tab[tab], r, v, d);
  } else {
    *v = 0.0;
    *d = 0.0;
  }
}

/* ------------------------------------------------------------------------ */
/* eval_forces                                                               */
/* ------------------------------------------------------------------------ */
#define KDUMMY 100.0

/* clamp_embedding (kernels.hpp:19-31) */
static void clamp_rho(const or_spline *emb, double *rho, double *limit) {
  const double lo = emb->x[0], hi = emb->x[emb->n - 1];
  if (*rho > hi) {
    const double d = *rho - hi;
    *limit += KDUMMY * 10.0 * d * d;
    *rho = hi;
  } else if (*rho < lo) {
    const double d = lo - *rho;
    *limit += KDUMMY * 10.0 * d * d;
    *rho = lo;
  }
}

/* accumulate_pair (kernels.hpp:33-38) */
static void acc_pair(double *F, double *E, double *sig, const double *d,
                     const double *f, double phi) {
  F[0] += f[0];
  F[1] += f[1];
  F[2] += f[2];
  *E += 0.5 * phi;
  for (int a = 0; a < 3; ++a) {
    const double hd = 0.5 * d[a];
    for (int b = 0; b < 3; ++b) sig[3 * a + b] -= hd * f[b];
  }
}

/* EAMForceCalculator::eval_forces (eam_force.cpp:45-102) inside
 * with_eval_scope (eval_scope.hpp:14-31) */
static void eam_config(or_ctx *c, int k) {
  const int64_t a0 = c->atom_off[k], a1 = c->atom_off[k + 1];
  double E = 0.0, lim = 0.0, sig[9] = {0};
  for (int64_t i = a0; i < a1; ++i) {
    c->force[3 * i] = c->force[3 * i + 1] = c->force[3 * i + 2] = 0.0;
    c->rho[i] = 0.0;
    c->gradF[i] = 0.0;
  }
  for (int64_t i = a0; i < a1; ++i) {
    for (int64_t p = c->pair_off[i]; p < c->pair_off[i + 1]; ++p) {
      const double r = pair_r(c, p);
      const int tj = c->types[c->nj[p]];
      double v, d;
      gated(c, p, ROLE_GJ, tab_dens(c, tj), r, &v, &d);
      c->rho[i] += v;
    }
  }
  for (int64_t i = a0; i < a1; ++i) {
    const or_spline *emb = &c->tab[tab_emb(c, c->types[i])];
    clamp_rho(emb, &c->rho[i], &lim);
    double F, dF;
    sp_eval_uncached(emb, c->rho[i], &F, &dF);
    E += F;
    c->gradF[i] = dF;
  }
  for (int64_t i = a0; i < a1; ++i) {
    const int ti = c->types[i];
    for (int64_t p = c->pair_off[i]; p < c->pair_off[i + 1]; ++p) {
      const int64_t j = c->nj[p];
      const int tj = c->types[j];
      const double r = pair_r(c, p);
      if (r < 1e-14) continue;
      const double inv_r = 1.0 / r;
      double phi, dphi, rj, drho_j, ri, drho_i;
      gated(c, p, ROLE_PHI, tab_pair(c, ti, tj), r, &phi, &dphi);
      gated(c, p, ROLE_GJ, tab_dens(c, tj), r, &rj, &drho_j);
      gated(c, p, ROLE_GI, tab_dens(c, ti), r, &ri, &drho_i);
      const double fscale =
          (dphi + c->gradF[i] * drho_j + c->gradF[j] * drho_i) * inv_r;
      const double *d = c->nd + 3 * p;
      const double f[3] = {fscale * d[0], fscale * d[1], fscale * d[2]};
      acc_pair(c->force + 3 * i, &E, sig, d, f, phi);
    }
  }
  const double V = c->box[k].vol;
  for (int q = 0; q < 9; ++q) c->stress[9 * k + q] = sig[q] / V;
  c->energy[k] = E;
  c->limit[k] = lim;
}

/* ------------------------------------------------------------------------ */
/* EAM gauge                                                                 */
/* ------------------------------------------------------------------------ */
/* The tables after rescale_eam (src/core/rescale.cpp:82-158) as the wrappers
 * of include/forcesmith/core/scale_potentials.hpp compose them, held as one
 * density scale A and per-type slope S_t and intercept F0_t:
 *   density     A g_t(r)                         ScaledOutputPotential (:48-60)
 *   embedding   F_t(rho / A) - S_t rho - F0_t,   ScaledArgPotential (:62-79)
 *               derivative F'_t(rho / A) / A - S_t, span A [lo, hi]
 *               under LinearAdjustedPotential (:34-46)
 *   pair        phi_ij(r) + S_i A g_j(r) + S_j A g_i(r) over the span of
 *               phi_ij, g evaluated ungated (CompensatedPairPotential :81-100)
 * Wrapped potentials have no site cache (DelegatingPotential :11-15), so
 * eval_gated takes its uncached branch: span check, then eval (eam_force.cpp:
 * 19-43). The density is wrapped when A != 1, a pair when S_i or S_j != 0.
 * Composition: a later rho-axis rescale by a maps (A, S) to (A a, S / a);
 * a later embed_shift by s adds s to S; a later intercept adds to F0. */
static int dens_wrapped(const or_ctx *c) { return c->gA != 1.0; }

/* value of density t at r, ungated (spline eval extrapolates) */
static void dens_value(const or_ctx *c, int t, double r, double *v, double *d) {
  sp_eval_uncached(&c->tab[tab_dens(c, t)], r, v, d);
  if (dens_wrapped(c)) {
    *v = c->gA * *v;
    *d = c->gA * *d;
  }
}

/* eval_gated of density t for pair p in role GJ / GI */
static void dens_gated(const or_ctx *c, int64_t p, int role, int t, double r, double *v,
                       double *d) {
  if (!dens_wrapped(c)) {
    gated(c, p, role, tab_dens(c, t), r, v, d);
  } else if (sp_in_range(&c->tab[tab_dens(c, t)], r)) {
    dens_value(c, t, r, v, d);
  } else {
    *v = *d = 0.0;
  }
}

/* eval_deriv_gated of pair (ti, tj) */
static void pair_gated(const or_ctx *c, int64_t p, int ti, int tj, double r, double *v,
                       double *d) {
  const int tp = tab_pair(c, ti, tj);
  if (c->gS[ti] == 0.0 && c->gS[tj] == 0.0) {
    gated(c, p, ROLE_PHI, tp, r, v, d);
    return;
  }
  if (!sp_in_range(&c->tab[tp], r)) {
    *v = *d = 0.0;
    return;
  }
  double pv, pd, bv, bd, av, ad;
  sp_eval_uncached(&c->tab[tp], r, &pv, &pd);
  dens_value(c, tj, r, &bv, &bd); /* g_beta: density of tj */
  dens_value(c, ti, r, &av, &ad); /* g_alpha: density of ti */
  *v = pv + c->gS[ti] * bv + c->gS[tj] * av;
  *d = pd + c->gS[ti] * bd + c->gS[tj] * ad;
}

/* the gauged embedding of type t: value and derivative at rho */
static void emb_gauged(const or_ctx *c, int t, double rho, double *F, double *dF) {
  const or_spline *e = &c->tab[tab_emb(c, t)];
  double v, d;
  if (dens_wrapped(c)) {
    sp_eval_uncached(e, rho / c->gA, &v, &d);
    d = d / c->gA;
  } else {
    sp_eval_uncached(e, rho, &v, &d);
  }
  *F = v - c->gS[t] * rho - c->gF0[t];
  *dF = d - c->gS[t];
}

/* span of the gauged embedding of type t (ScaledArgPotential::span) */
static void emb_span(const or_ctx *c, int t, double *lo, double *hi) {
  const or_spline *e = &c->tab[tab_emb(c, t)];
  *lo = c->gA * e->x[0];
  *hi = c->gA * e->x[e->n - 1];
}

/* eval_forces of the gauged tables (eam_force.cpp:45-102) */
static void eam_config_gauge(or_ctx *c, int k) {
  const int64_t a0 = c->atom_off[k], a1 = c->atom_off[k + 1];
  double E = 0.0, lim = 0.0, sig[9] = {0};
  for (int64_t i = a0; i < a1; ++i) {
    c->force[3 * i] = c->force[3 * i + 1] = c->force[3 * i + 2] = 0.0;
    c->rho[i] = 0.0;
    c->gradF[i] = 0.0;
  }
  for (int64_t i = a0; i < a1; ++i)
    for (int64_t p = c->pair_off[i]; p < c->pair_off[i + 1]; ++p) {
      double v, d;
      dens_gated(c, p, ROLE_GJ, c->types[c->nj[p]], pair_r(c, p), &v, &d);
      c->rho[i] += v;
    }
  for (int64_t i = a0; i < a1; ++i) {
    const int t = c->types[i];
    double begin, end; /* clamp_embedding (kernels.hpp:19-31) on the span */
    emb_span(c, t, &begin, &end);
    if (c->rho[i] > end) {
      const double dd = c->rho[i] - end;
      lim += KDUMMY * 10.0 * dd * dd;
      c->rho[i] = end;
    } else if (c->rho[i] < begin) {
      const double dd = begin - c->rho[i];
      lim += KDUMMY * 10.0 * dd * dd;
      c->rho[i] = begin;
    }
    double F, dF;
    emb_gauged(c, t, c->rho[i], &F, &dF);
    E += F;
    c->gradF[i] = dF;
  }
  for (int64_t i = a0; i < a1; ++i) {
    const int ti = c->types[i];
    for (int64_t p = c->pair_off[i]; p < c->pair_off[i + 1]; ++p) {
      const int64_t j = c->nj[p];
      const int tj = c->types[j];
      const double r = pair_r(c, p);
      if (r < 1e-14) continue;
      double phi, dphi, rj, drho_j, ri, drho_i;
      pair_gated(c, p, ti, tj, r, &phi, &dphi);
      dens_gated(c, p, ROLE_GJ, tj, r, &rj, &drho_j);
      dens_gated(c, p, ROLE_GI, ti, r, &ri, &drho_i);
      const double fscale = (dphi + c->gradF[i] * drho_j + c->gradF[j] * drho_i) * (1.0 / r);
      const double *d = c->nd + 3 * p;
      const double f[3] = {fscale * d[0], fscale * d[1], fscale * d[2]};
      acc_pair(c->force + 3 * i, &E, sig, d, f, phi);
    }
  }
  const double V = c->box[k].vol;
  for (int q = 0; q < 9; ++q) c->stress[9 * k + q] = sig[q] / V;
  c->energy[k] = E;
  c->limit[k] = lim;
}

int or_set_gauge(or_ctx *c, double A, const double *S, const double *F0) {
  if (c->kind != OR_EAM || c->ntypes > OR_MAXT) return -1;
  if (!isfinite(A) || A == 0.0) return -2;
  c->gA = A;
  int id = A == 1.0;
  for (int t = 0; t < c->ntypes; ++t) {
    c->gS[t] = S ? S[t] : 0.0;
    c->gF0[t] = F0 ? F0[t] : 0.0;
    id = id && c->gS[t] == 0.0 && c->gF0[t] == 0.0;
  }
  c->gauge = !id;
  return 0;
}

void or_get_gauge(const or_ctx *c, double *A, double *S, double *F0) {
  if (A) *A = c->gA;
  for (int t = 0; t < c->ntypes; ++t) {
    if (S) S[t] = c->gS[t];
    if (F0) F0[t] = c->gF0[t];
  }
}

void or_emb_gauged(const or_ctx *c, int t, double rho, double *F, double *dF) {
  emb_gauged(c, t, rho, F, dF);
}

/* compute_rho_ref (rescale.cpp:11-33) with the extremes rescale_rho_axis
 * takes (:39-52): every configuration evaluated, the clamped densities per
 * type; count 0 leaves min = +inf, max = -inf, mean = 0 */
void or_rho_stats(or_ctx *c, double *rmin, double *rmax, double *rmean, int64_t *count) {
  or_eval_forces(c, NULL, NULL, NULL, NULL);
  for (int t = 0; t < c->ntypes; ++t) {
    double lo = INFINITY, hi = -INFINITY, sum = 0.0;
    int64_t n = 0;
    for (int64_t i = 0; i < c->N; ++i) {
      if (c->types[i] != t) continue;
      sum += c->rho[i];
      if (c->rho[i] > hi) hi = c->rho[i];
      if (c->rho[i] < lo) lo = c->rho[i];
      ++n;
    }
    if (rmin) rmin[t] = lo;
    if (rmax) rmax[t] = hi;
    if (rmean) rmean[t] = n > 0 ? sum / (double)n : 0.0;
    if (count) count[t] = n;
  }
}

/* rescale_rho_axis (rescale.cpp:35-101): the scale a that takes the
 * dominant type's extreme density (plus 0.003 of its embedding span) to
 * the end of that span; 1 (no change) without data, for |a| within
 * [0.95, 1.05] while every type's densities lie inside its span, or for a
 * non-finite / vanishing a. Applies a to the gauge and returns it. */
double or_rescale_rho_axis(or_ctx *c) {
  const int n = c->ntypes;
  double mx[OR_MAXT], mn[OR_MAXT];
  or_rho_stats(c, mn, mx, NULL, NULL);
  int dom = 0, sign_pos = 1;
  double best = -1.0;
  for (int t = 0; t < n; ++t) {
    if (!isfinite(mx[t])) continue;
    const double ext = fmax(fabs(mx[t]), fabs(mn[t]));
    if (ext > best) {
      best = ext;
      dom = t;
      sign_pos = mx[t] >= -mn[t];
    }
  }
  if (best < 0.0) return 1.0;
  double elo, ehi;
  emb_span(c, dom, &elo, &ehi);
  const double pad = 0.003 * (ehi - elo);
  const double upper = sign_pos ? ehi : elo;
  const double right = sign_pos ? mx[dom] + pad : mn[dom] - pad;
  if (fabs(right) < 1e-30) return 1.0;
  const double a = upper / right;
  int violation = 0;
  for (int t = 0; t < n; ++t) {
    if (!isfinite(mx[t])) continue;
    double lo, hi;
    emb_span(c, t, &lo, &hi);
    if (mn[t] < lo || mx[t] > hi) violation = 1;
  }
  if (!isfinite(a) || fabs(a) < 1e-30 ||
      (!violation && fabs(a) >= 0.95 && fabs(a) <= 1.05))
    return 1.0;
  double S[OR_MAXT];
  for (int t = 0; t < n; ++t) S[t] = c->gS[t] / a;
  or_set_gauge(c, c->gA * a, S, c->gF0);
  return a;
}

/* embed_shift (rescale.cpp:103-136): per type the slope s_t = F'(rho_ref_t)
 * of the current embedding (types with rho_ref_t > 0), subtracted from the
 * embedding and compensated in the pairs. Slopes below 1e-14 are dropped
 * (the reference skips them for the embedding; its pair compensation keeps
 * such a slope while the partner's is larger, a term below 1e-14 g(r)). */
void or_embed_shift(or_ctx *c, const double *rho_ref) {
  double S[OR_MAXT];
  for (int t = 0; t < c->ntypes; ++t) {
    S[t] = c->gS[t];
    if (!(rho_ref[t] > 0.0)) continue;
    double F, dF;
    emb_gauged(c, t, rho_ref[t], &F, &dF);
    if (fabs(dF) >= 1e-14) S[t] += dF;
  }
  or_set_gauge(c, c->gA, S, c->gF0);
}

/* rescale_eam (rescale.cpp:138-158): rho axis, reference densities, slope
 * shift, then the intercept F0_t = F(rho_ref_t) of the shifted embedding
 * (types with rho_ref_t > 0 and |F0_t| >= 1e-14). Returns a. */
double or_rescale_eam(or_ctx *c, double *rho_ref_out) {
  const double a = or_rescale_rho_axis(c);
  double ref[OR_MAXT], F0[OR_MAXT];
  or_rho_stats(c, NULL, NULL, ref, NULL);
  or_embed_shift(c, ref);
  for (int t = 0; t < c->ntypes; ++t) {
    F0[t] = c->gF0[t];
    if (!(ref[t] > 0.0)) continue;
    double F, dF;
    emb_gauged(c, t, ref[t], &F, &dF);
    if (fabs(F) >= 1e-14) F0[t] += F;
  }
  or_set_gauge(c, c->gA, c->gS, F0);
  if (rho_ref_out)
    for (int t = 0; t < c->ntypes; ++t) rho_ref_out[t] = ref[t];
  return a;
}

/* ADPForceCalculator::eval_forces (adp_force.cpp:73-135); quad_nu/quad_xi
 * (adp_force.hpp:63-68) */
static double quad_nu(const double *M, const double *d) {
  double Md[3];
  for (int a = 0; a < 3; ++a) Md[a] = M[3 * a] * d[0] + M[3 * a + 1] * d[1] + M[3 * a + 2] * d[2];
  const double dd = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
  const double tr = M[0] + M[4] + M[8];
  return (d[0] * Md[0] + d[1] * Md[1] + d[2] * Md[2]) - dd / 3.0 * tr;
}
static void quad_xi(const double *M, const double *d, double *xi) {
  const double tr3 = (M[0] + M[4] + M[8]) / 3.0;
  for (int a = 0; a < 3; ++a)
    xi[a] = (M[3 * a] * d[0] + M[3 * a + 1] * d[1] + M[3 * a + 2] * d[2]) - tr3 * d[a];
}

static void adp_config(or_ctx *c, int k) {
  const int64_t a0 = c->atom_off[k], a1 = c->atom_off[k + 1];
  double E = 0.0, lim = 0.0, sig[9] = {0};
  for (int64_t i = a0; i < a1; ++i) {
    c->force[3 * i] = c->force[3 * i + 1] = c->force[3 * i + 2] = 0.0;
    c->rho[i] = 0.0;
    c->gradF[i] = 0.0;
    for (int q = 0; q < 3; ++q) c->mu[3 * i + q] = 0.0;
    for (int q = 0; q < 9; ++q) c->lam9[9 * i + q] = 0.0;
  }
  for (int64_t i = a0; i < a1; ++i) {
    const int ti = c->types[i];
    double rho = 0.0, mu[3] = {0}, lam[9] = {0};
    for (int64_t p = c->pair_off[i]; p < c->pair_off[i + 1]; ++p) {
      const int tj = c->types[c->nj[p]];
      const double r = pair_r(c, p);
      if (r < 1e-14) continue;
      const double *d = c->nd + 3 * p;
      double v, dv;
      gated(c, p, ROLE_GJ, tab_dens(c, tj), r, &v, &dv);
      rho += v;
      gated(c, p, ROLE_DIP, tab_dip(c, ti, tj), r, &v, &dv);
      for (int a = 0; a < 3; ++a) mu[a] += v * d[a];
      gated(c, p, ROLE_QUAD, tab_quad(c, ti, tj), r, &v, &dv);
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) lam[3 * a + b] += v * (d[a] * d[b]);
    }
    c->rho[i] += rho;
    for (int a = 0; a < 3; ++a) c->mu[3 * i + a] += mu[a];
    for (int q = 0; q < 9; ++q) c->lam9[9 * i + q] += lam[q];
  }
  double energy = 0.0;
  for (int64_t i = a0; i < a1; ++i) {
    const or_spline *emb = &c->tab[tab_emb(c, c->types[i])];
    clamp_rho(emb, &c->rho[i], &lim);
    double F, dF;
    sp_eval_uncached(emb, c->rho[i], &F, &dF);
    c->gradF[i] = dF;
    double e = F;
    const double *m = c->mu + 3 * i;
    e += 0.5 * (m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
    const double *L = c->lam9 + 9 * i;
    const double tr = L[0] + L[4] + L[8];
    double fro = 0.0;
    for (int q = 0; q < 9; ++q) fro += L[q] * L[q];
    e += 0.5 * (fro - tr * tr / 3.0);
    energy += e;
  }
  E += energy;
  for (int64_t i = a0; i < a1; ++i) {
    const int ti = c->types[i];
    for (int64_t p = c->pair_off[i]; p < c->pair_off[i + 1]; ++p) {
      const int64_t j = c->nj[p];
      const int tj = c->types[j];
      const double r = pair_r(c, p);
      if (r < 1e-14) continue;
      const double inv_r = 1.0 / r;
      const double *d = c->nd + 3 * p;
      double f[3] = {0, 0, 0};
      double phi, dphi, tmp, drho_j, drho_i;
      gated(c, p, ROLE_PHI, tab_pair(c, ti, tj), r, &phi, &dphi);
      gated(c, p, ROLE_GJ, tab_dens(c, tj), r, &tmp, &drho_j);
      gated(c, p, ROLE_GI, tab_dens(c, ti), r, &tmp, &drho_i);
      const double se = (dphi + c->gradF[i] * drho_j + c->gradF[j] * drho_i) * inv_r;
      for (int a = 0; a < 3; ++a) f[a] += se * d[a];
      double u, du;
      gated(c, p, ROLE_DIP, tab_dip(c, ti, tj), r, &u, &du);
      const double *mi = c->mu + 3 * i, *mj = c->mu + 3 * j;
      const double dot_i = mi[0] * d[0] + mi[1] * d[1] + mi[2] * d[2];
      const double dot_j = mj[0] * d[0] + mj[1] * d[1] + mj[2] * d[2];
      const double sd = du * inv_r * (dot_i - dot_j);
      for (int a = 0; a < 3; ++a) f[a] += sd * d[a] + u * (mi[a] - mj[a]);
      double w, dw;
      gated(c, p, ROLE_QUAD, tab_quad(c, ti, tj), r, &w, &dw);
      const double *Li = c->lam9 + 9 * i, *Lj = c->lam9 + 9 * j;
      const double nu = quad_nu(Li, d) + quad_nu(Lj, d);
      double xi_i[3], xi_j[3];
      quad_xi(Li, d, xi_i);
      quad_xi(Lj, d, xi_j);
      const double sq = dw * inv_r * nu;
      for (int a = 0; a < 3; ++a) f[a] += sq * d[a] + 2.0 * w * (xi_i[a] + xi_j[a]);
      acc_pair(c->force + 3 * i, &E, sig, d, f, phi);
    }
  }
  const double V = c->box[k].vol;
  for (int q = 0; q < 9; ++q) c->stress[9 * k + q] = sig[q] / V;
  c->energy[k] = E;
  c->limit[k] = lim;
}

/* Stillinger-Weber v2_dv2 / h_dh (stiweb_force.cpp:24-54) */
static void sw_v2(double r, const or_sw *p, double *v2, double *dv2) {
  *v2 = 0.0;
  *dv2 = 0.0;
  if (r >= p->a1) return;
  const double d = r - p->a1;
  const double e = exp(p->delta / d);
  if (e == 0.0) return;
  const double rp = pow(r, -p->p);
  const double rq = pow(r, -p->q);
  const double poly = p->A * rp - p->B * rq;
  const double dpoly = -p->A * p->p * pow(r, -p->p - 1.0) + p->B * p->q * pow(r, -p->q - 1.0);
  *v2 = poly * e;
  *dv2 = dpoly * e + poly * e * (-p->delta / (d * d));
}
static void sw_h(double r, const or_sw *p, double *h, double *dh) {
  *h = 0.0;
  *dh = 0.0;
  if (r >= p->a2) return;
  const double d = r - p->a2;
  const double e = exp(p->gamma / d);
  if (e == 0.0) return;
  *h = e;
  *dh = e * (-p->gamma / (d * d));
}

/* three_body_forces (kernels.hpp:40-56) */
static void three_body(const double *d1, const double *d2, double inv_r1,
                       double inv_r2, double c, double f1, double df1,
                       double f2, double df2, double g, double dg, double *fi,
                       double *fj, double *fk) {
  const double inv_r1r2 = inv_r1 * inv_r2;
  const double A = c * inv_r1 * inv_r1 - inv_r1r2;
  const double B = c * inv_r2 * inv_r2 - inv_r1r2;
  const double s1 = df1 * inv_r1 * f2 * g, s2 = f1 * df2 * inv_r2 * g, s3 = f1 * f2 * dg;
  for (int a = 0; a < 3; ++a) {
    fi[a] = s1 * d1[a] + s2 * d2[a] - s3 * (A * d1[a] + B * d2[a]);
    fj[a] = -s1 * d1[a] - s3 * (inv_r1r2 * d2[a] - c * inv_r1 * inv_r1 * d1[a]);
    fk[a] = -s2 * d2[a] - s3 * (inv_r1r2 * d1[a] - c * inv_r2 * inv_r2 * d2[a]);
  }
}

static int lambda_index(const or_ctx *c, int ti, int tj, int tk) {
  return ti * c->npair + pair_ordinal(tj, tk, c->ntypes);
}

/* StiwebForceCalculator::eval_forces (stiweb_force.cpp:241-271) */
static void sw_config(or_ctx *c, int k) {
  const int64_t a0 = c->atom_off[k], a1 = c->atom_off[k + 1];
  double E = 0.0, sig[9] = {0};
  for (int64_t i = a0; i < a1; ++i)
    c->force[3 * i] = c->force[3 * i + 1] = c->force[3 * i + 2] = 0.0;
  for (int64_t i = a0; i < a1; ++i) {
    const int ti = c->types[i];
    for (int64_t p = c->pair_off[i]; p < c->pair_off[i + 1]; ++p) {
      const int tj = c->types[c->nj[p]];
      const double r = pair_r(c, p);
      if (r < 1e-14) continue;
      double v2, dv2;
      sw_v2(r, &c->sw[pair_ordinal(ti, tj, c->ntypes)], &v2, &dv2);
      const double *d = c->nd + 3 * p;
      const double s = dv2 / r;
      const double f[3] = {s * d[0], s * d[1], s * d[2]};
      acc_pair(c->force + 3 * i, &E, sig, d, f, v2);
    }
  }
  for (int64_t i = a0; i < a1; ++i) {
    const int ti = c->types[i];
    const int64_t p0 = c->pair_off[i], nn = c->pair_off[i + 1] - p0;
    for (int64_t jj = 0; jj < nn; ++jj) {
      const int64_t pj = p0 + jj;
      const int64_t j = c->nj[pj];
      const int tj = c->types[j];
      const double r1 = pair_r(c, pj);
      if (r1 < 1e-14) continue;
      double h1, dh1;
      sw_h(r1, &c->sw[pair_ordinal(ti, tj, c->ntypes)], &h1, &dh1);
      if (h1 == 0.0 && dh1 == 0.0) continue;
      const double inv_r1 = 1.0 / r1;
      const double *d1 = c->nd + 3 * pj;
      for (int64_t kk = jj + 1; kk < nn; ++kk) {
        const int64_t pk = p0 + kk;
        const int64_t kx = c->nj[pk];
        const int tk = c->types[kx];
        const double r2 = pair_r(c, pk);
        if (r2 < 1e-14) continue;
        double h2, dh2;
        sw_h(r2, &c->sw[pair_ordinal(ti, tk, c->ntypes)], &h2, &dh2);
        if (h2 == 0.0 && dh2 == 0.0) continue;
        const double inv_r2 = 1.0 / r2;
        const double *d2 = c->nd + 3 * pk;
        const double lam = c->lam[lambda_index(c, ti, tj, tk)];
        const double cs = (d1[0] * d2[0] + d1[1] * d2[1] + d1[2] * d2[2]) * inv_r1 * inv_r2;
        const double cp13 = cs + 1.0 / 3.0;
        const double w = lam * cp13 * cp13;
        const double dw = 2.0 * lam * cp13;
        E += h1 * h2 * w;
        double fi[3], fj[3], fk[3];
        three_body(d1, d2, inv_r1, inv_r2, cs, h1, dh1, h2, dh2, w, dw, fi, fj, fk);
        for (int a = 0; a < 3; ++a) {
          c->force[3 * i + a] += fi[a];
          c->force[3 * j + a] += fj[a];
          c->force[3 * kx + a] += fk[a];
        }
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b) sig[3 * a + b] += d1[a] * fj[b];
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b) sig[3 * a + b] += d2[a] * fk[b];
      }
    }
  }
  const double V = c->box[k].vol;
  for (int q = 0; q < 9; ++q) c->stress[9 * k + q] = sig[q] / V;
  c->energy[k] = E;
  c->limit[k] = 0.0;
}

static void ensure_nl(or_ctx *c) {
  const double rc = or_max_cutoff(c);
  if (c->pair_off == NULL || rc != c->nl_rc) {
    or_build_nl(c, rc);
    prime_sites(c);
  }
}

static void eval_config_task(void *arg, int64_t k) {
  or_ctx *c = (or_ctx *)arg;
  if (c->kind == OR_EAM) {
    if (c->gauge) eam_config_gauge(c, (int)k);
    else eam_config(c, (int)k);
  } else if (c->kind == OR_ADP) adp_config(c, (int)k);
  else sw_config(c, (int)k);
}

void or_eval_forces(or_ctx *c, double *forces, double *energy, double *stress9,
                    double *limit) {
  ensure_nl(c);
  parallel_for(c->C, eval_config_task, c);
  if (forces) memcpy(forces, c->force, sizeof(double) * 3 * (size_t)c->N);
  if (energy) memcpy(energy, c->energy, sizeof(double) * (size_t)c->C);
  if (stress9) memcpy(stress9, c->stress, sizeof(double) * 9 * (size_t)c->C);
  if (limit) memcpy(limit, c->limit, sizeof(double) * (size_t)c->C);
}

void or_get_scratch(const or_ctx *c, double *rho, double *gradF, double *mu,
                    double *lambda9) {
  const size_t N = (size_t)c->N;
  if (rho) memcpy(rho, c->rho, sizeof(double) * N);
  if (gradF) memcpy(gradF, c->gradF, sizeof(double) * N);
  if (mu) memcpy(mu, c->mu, sizeof(double) * 3 * N);
  if (lambda9) memcpy(lambda9, c->lam9, sizeof(double) * 9 * N);
}

/* write_config_residuals (residual_layout.hpp:37-54) */
static const int V6[6][2] = {{0, 0}, {1, 1}, {2, 2}, {0, 1}, {0, 2}, {1, 2}};
static void write_rows(const or_ctx *c, int k, double *F) {
  int64_t row = c->row_off[k];
  for (int64_t i = c->atom_off[k]; i < c->atom_off[k + 1]; ++i) {
    F[row++] = c->force[3 * i] - c->ref_force[3 * i];
    F[row++] = c->force[3 * i + 1] - c->ref_force[3 * i + 1];
    F[row++] = c->force[3 * i + 2] - c->ref_force[3 * i + 2];
  }
  F[row++] = c->ew * (c->energy[k] - c->ref_energy[k]);
  if (c->sw_w > 0.0) {
    for (int v = 0; v < 6; ++v)
      F[row++] = c->sw_w * (c->stress[9 * k + 3 * V6[v][0] + V6[v][1]] - c->ref_stress[6 * k + v]);
  }
  F[row++] = c->limit[k];
}

typedef struct { or_ctx *c; double *F; } res_job;
static void residual_task(void *arg, int64_t k) {
  res_job *rj = (res_job *)arg;
  eval_config_task(rj->c, k);
  write_rows(rj->c, (int)k, rj->F);
}

/* rows w (y[k-1] - 2 y[k] + y[k+1]), k = 1 .. n-2, of every table that
 * has curvature rows, after the configuration rows
 * (SplinePotential::write_curvature, spline.cpp:205-213;
 * forcesmith_functor.cpp:53-55) */
static void write_smooth_rows(const or_ctx *c, double *F) {
  if (or_smooth_count(c) == 0) return;
  int64_t row = c->row_off[c->C];
  for (int t = 0; t < c->ntab; ++t) {
    const or_spline *sp = &c->tab[t];
    int nfree = 0;
    for (int k = 0; k < sp->n; ++k) nfree += !sp->fixed[k];
    if (nfree == 0 || sp->n < 3) continue;
    for (int k = 1; k + 1 < sp->n; ++k)
      F[row++] = c->smooth_w * (sp->y[k - 1] - 2.0 * sp->y[k] + sp->y[k + 1]);
  }
}

/* ForcesmithFunctor eval_into (forcesmith_functor.cpp:35-57) */
void or_residual(or_ctx *c, const double *x, double *F) {
  or_scatter_params(c, x);
  ensure_nl(c);
  res_job rj = {c, F};
  parallel_for(c->C, residual_task, &rj);
  write_smooth_rows(c, F);
}

/* opt::fd_jacobian (fd_jacobian.hpp:37-49), h = 1e-5 */
void or_jacobian_fd_cols(or_ctx *c, const double *x, int col0, int ncols,
                         double *Jc) {
  const int D = or_param_count(c);
  const int64_t R = or_residual_count(c);
  const double h = 1e-5;
  double *xp = (double *)malloc(sizeof(double) * (size_t)D);
  double *fp = (double *)malloc(sizeof(double) * (size_t)R);
  double *fm = (double *)malloc(sizeof(double) * (size_t)R);
  memcpy(xp, x, sizeof(double) * (size_t)D);
  for (int jc = 0; jc < ncols; ++jc) {
    const int j = col0 + jc;
    xp[j] += h;
    or_residual(c, xp, fp);
    xp[j] -= 2.0 * h;
    or_residual(c, xp, fm);
    xp[j] += h;
    for (int64_t r = 0; r < R; ++r) Jc[r * ncols + jc] = (fp[r] - fm[r]) / (2.0 * h);
  }
  or_scatter_params(c, x);
  free(xp);
  free(fp);
  free(fm);
}

void or_jacobian_fd(or_ctx *c, const double *x, double *J) {
  or_jacobian_fd_cols(c, x, 0, or_param_count(c), J);
}

/* ------------------------------------------------------------------------ */
/* analytic per-knot Jacobian                                                */
/* ------------------------------------------------------------------------ */
/* Column layout (force/param_jacobian.hpp:18-41): each table's free knots in
 * index order, tables in declaration order. colmap[t][m] = column of knot m
 * of table t, or -1 if fixed. */
typedef struct {
  int *col;    /* per table offset into col_of */
  int *col_of; /* knot -> column or -1 */
} colmap;

static colmap make_colmap(const or_ctx *c) {
  colmap cm;
  cm.col = (int *)calloc((size_t)c->ntab + 1, sizeof(int));
  int tot = 0;
  for (int t = 0; t < c->ntab; ++t) {
    cm.col[t] = tot;
    tot += c->tab[t].n;
  }
  cm.col[c->ntab] = tot;
  cm.col_of = (int *)malloc(sizeof(int) * (size_t)(tot ? tot : 1));
  int o = 0;
  for (int t = 0; t < c->ntab; ++t)
    for (int k = 0; k < c->tab[t].n; ++k)
      cm.col_of[cm.col[t] + k] = c->tab[t].fixed[k] ? -1 : o++;
  return cm;
}
static void free_colmap(colmap *cm) {
  free(cm->col);
  free(cm->col_of);
}

typedef struct {
  or_ctx *c;
  colmap cm;
  double *J;
  int D;
} jac_job;

/* add_force_column (param_jacobian.hpp:91-109) */
static void add_col(double *J, int D, int64_t atom_row, int64_t e_row,
                    int64_t s_row, int col, const double *d, const double *df,
                    double denergy, double ew, double sw, double inv_vol) {
  for (int a = 0; a < 3; ++a) J[(atom_row + a) * D + col] += df[a];
  if (denergy != 0.0) J[e_row * D + col] += ew * denergy;
  if (s_row >= 0)
    for (int v = 0; v < 6; ++v)
      J[(s_row + v) * D + col] -= sw * 0.5 * d[V6[v][0]] * df[V6[v][1]] * inv_vol;
}

/* add_radial_bond (param_jacobian.hpp:111-121) over the knots of a spline
 * table at r: df = (scale * dpart_k * inv_r) d, energy 0.5 * vpart_k */
static void add_bond(const jac_job *jj, int tab, double r, const double *d,
                     double inv_r, double scale, int with_energy,
                     int64_t atom_row, int64_t e_row, int64_t s_row,
                     double inv_vol, double *dv, double *dd) {
  const or_ctx *c = jj->c;
  const or_spline *sp = &c->tab[tab];
  sp_param_partials(sp, c->Sjac + c->Soff[tab], r, dv, dd, NULL);
  for (int m = 0; m < sp->n; ++m) {
    const int col = jj->cm.col_of[jj->cm.col[tab] + m];
    if (col < 0) continue;
    if (dv[m] == 0.0 && dd[m] == 0.0) continue;
    const double sc = scale * dd[m] * inv_r;
    const double df[3] = {sc * d[0], sc * d[1], sc * d[2]};
    add_col(jj->J, jj->D, atom_row, e_row, s_row, col, d, df,
            with_energy ? 0.5 * dv[m] : 0.0, c->ew, c->sw_w, inv_vol);
  }
}

/* EAM (eam_force.cpp:108-279) and ADP (adp_force.cpp:141-385) chains with
 * spline partials in place of the analytic param_grad/dderiv_dparam. */
static void spline_jac_config(void *arg, int64_t kk) {
  jac_job *jj = (jac_job *)arg;
  or_ctx *c = jj->c;
  const int k = (int)kk;
  const int D = jj->D;
  double *J = jj->J;
  const int adp = c->kind == OR_ADP;
  const int64_t a0 = c->atom_off[k], na = c->atom_off[k + 1] - a0;
  const int64_t force0 = c->row_off[k];
  const int64_t e_row = force0 + 3 * na;
  const int64_t s_row = c->sw_w > 0.0 ? e_row + 1 : -1;
  const int64_t l_row = e_row + 1 + (c->sw_w > 0.0 ? 6 : 0);
  const double inv_vol = 1.0 / c->box[k].vol;
  int nmax = 0;
  for (int t = 0; t < c->ntab; ++t) if (c->tab[t].n > nmax) nmax = c->tab[t].n;
  double *dv = (double *)calloc((size_t)nmax, sizeof(double));
  double *dd = (double *)calloc((size_t)nmax, sizeof(double));
  double *dv2 = (double *)calloc((size_t)nmax, sizeof(double));
  double *dd2 = (double *)calloc((size_t)nmax, sizeof(double));
  /* columns of the density / embedding / dipole / quadrupole blocks */
  const int dens_t0 = c->npair, emb_t0 = c->npair + c->ntypes;
  const int dip_t0 = c->npair + 2 * c->ntypes, quad_t0 = 2 * c->npair + 2 * c->ntypes;
  /* per-atom column-indexed chains, full D wide for simplicity */
  double *rho = (double *)calloc((size_t)na, sizeof(double));
  double *drho = (double *)calloc((size_t)na * D, sizeof(double));
  double *mu = adp ? (double *)calloc((size_t)na * 3, sizeof(double)) : NULL;
  double *dmu = adp ? (double *)calloc((size_t)na * D * 3, sizeof(double)) : NULL;
  double *lam = adp ? (double *)calloc((size_t)na * 9, sizeof(double)) : NULL;
  double *dlam = adp ? (double *)calloc((size_t)na * D * 9, sizeof(double)) : NULL;
  double *gradF = (double *)calloc((size_t)na, sizeof(double));
  double *curvF = (double *)calloc((size_t)na, sizeof(double));
  double *demb = (double *)calloc((size_t)na * D, sizeof(double));

  for (int64_t li = 0; li < na; ++li) {
    const int64_t i = a0 + li;
    const int ti = c->types[i];
    for (int64_t p = c->pair_off[i]; p < c->pair_off[i + 1]; ++p) {
      const int tj = c->types[c->nj[p]];
      const double r = pair_r(c, p);
      const double *d = c->nd + 3 * p;
      if (adp && r < 1e-14) continue;
      const int tg = tab_dens(c, tj);
      if (sp_in_range(&c->tab[tg], r)) {
        double v, dr;
        sp_eval_uncached(&c->tab[tg], r, &v, &dr);
        rho[li] += v;
        sp_param_partials(&c->tab[tg], c->Sjac + c->Soff[tg], r, dv, NULL, NULL);
        for (int m = 0; m < c->tab[tg].n; ++m) {
          const int col = jj->cm.col_of[jj->cm.col[tg] + m];
          if (col >= 0) drho[li * D + col] += dv[m];
        }
      }
      if (!adp) continue;
      const int tu = tab_dip(c, ti, tj), tw = tab_quad(c, ti, tj);
      if (sp_in_range(&c->tab[tu], r)) {
        double v, dr;
        sp_eval_uncached(&c->tab[tu], r, &v, &dr);
        for (int a = 0; a < 3; ++a) mu[3 * li + a] += v * d[a];
        sp_param_partials(&c->tab[tu], c->Sjac + c->Soff[tu], r, dv, NULL, NULL);
        for (int m = 0; m < c->tab[tu].n; ++m) {
          const int col = jj->cm.col_of[jj->cm.col[tu] + m];
          if (col >= 0)
            for (int a = 0; a < 3; ++a) dmu[(li * D + col) * 3 + a] += dv[m] * d[a];
        }
      }
      if (sp_in_range(&c->tab[tw], r)) {
        double v, dr;
        sp_eval_uncached(&c->tab[tw], r, &v, &dr);
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b) lam[9 * li + 3 * a + b] += v * (d[a] * d[b]);
        sp_param_partials(&c->tab[tw], c->Sjac + c->Soff[tw], r, dv, NULL, NULL);
        for (int m = 0; m < c->tab[tw].n; ++m) {
          const int col = jj->cm.col_of[jj->cm.col[tw] + m];
          if (col >= 0)
            for (int a = 0; a < 3; ++a)
              for (int b = 0; b < 3; ++b)
                dlam[(li * D + col) * 9 + 3 * a + b] += dv[m] * (d[a] * d[b]);
        }
      }
    }
  }

  for (int64_t li = 0; li < na; ++li) {
    const int64_t i = a0 + li;
    const int te = tab_emb(c, c->types[i]);
    const or_spline *emb = &c->tab[te];
    /* the gauge (identity outside EAM): clamp on A rho against A [lo, hi];
     * a clamped atom's energy keeps S (A rho - clamped) from the
     * compensated pairs, whose sum over the neighbours is S A rho when the
     * pair and density spans agree (the only case taken here) */
    const double gA = c->gA, gS = c->gS[c->types[i]];
    const double rr = gA * rho[li];
    const double lo = gA * emb->x[0], hi = gA * emb->x[emb->n - 1];
    double rb = rho[li], over = 0.0;
    if (rr > hi) {
      over = 2.0 * KDUMMY * 10.0 * (rr - hi) * gA;
      rb = hi / gA;
    } else if (rr < lo) {
      over = -2.0 * KDUMMY * 10.0 * (lo - rr) * gA;
      rb = lo / gA;
    }
    const int clamped = over != 0.0;
    double Fv, Fd, F2;
    sp_eval_uncached(emb, rb, &Fv, &Fd);
    sp_param_partials(emb, c->Sjac + c->Soff[te], rb, dv, dd, &F2);
    gradF[li] = Fd;
    curvF[li] = F2;
    for (int m = 0; m < emb->n; ++m) {
      const int col = jj->cm.col_of[jj->cm.col[te] + m];
      if (col < 0) continue;
      J[e_row * D + col] += c->ew * dv[m];
      demb[li * D + col] = dd[m];
    }
    for (int t = dens_t0; t < dens_t0 + c->ntypes; ++t)
      for (int m = 0; m < c->tab[t].n; ++m) {
        const int col = jj->cm.col_of[jj->cm.col[t] + m];
        if (col < 0) continue;
        const double dr = drho[li * D + col];
        if (dr == 0.0) continue;
        if (clamped) {
          J[l_row * D + col] += over * dr;
          if (gS != 0.0) J[e_row * D + col] += c->ew * (gS * gA) * dr;
        } else {
          J[e_row * D + col] += c->ew * gradF[li] * dr;
        }
      }
    if (clamped)
      for (int t = dens_t0; t < dens_t0 + c->ntypes; ++t)
        for (int m = 0; m < c->tab[t].n; ++m) {
          const int col = jj->cm.col_of[jj->cm.col[t] + m];
          if (col >= 0) drho[li * D + col] = 0.0;
        }
    rho[li] = rb;
    if (adp) {
      for (int t = dip_t0; t < dip_t0 + c->npair; ++t)
        for (int m = 0; m < c->tab[t].n; ++m) {
          const int col = jj->cm.col_of[jj->cm.col[t] + m];
          if (col < 0) continue;
          const double *dm = dmu + (li * D + col) * 3;
          J[e_row * D + col] += c->ew * (mu[3 * li] * dm[0] + mu[3 * li + 1] * dm[1] + mu[3 * li + 2] * dm[2]);
        }
      for (int t = quad_t0; t < quad_t0 + c->npair; ++t)
        for (int m = 0; m < c->tab[t].n; ++m) {
          const int col = jj->cm.col_of[jj->cm.col[t] + m];
          if (col < 0) continue;
          const double *dl = dlam + (li * D + col) * 9;
          const double *L = lam + 9 * li;
          double s = 0.0;
          for (int q = 0; q < 9; ++q) s += L[q] * dl[q];
          const double trL = L[0] + L[4] + L[8], trd = dl[0] + dl[4] + dl[8];
          J[e_row * D + col] += c->ew * (s - trL * trd / 3.0);
        }
    }
  }
  for (int64_t li = 0; li < na; ++li) {
    const int64_t i = a0 + li;
    const int ti = c->types[i];
    const int64_t atom_row = force0 + 3 * li;
    for (int64_t p = c->pair_off[i]; p < c->pair_off[i + 1]; ++p) {
      const double r = pair_r(c, p);
      if (r < 1e-14) continue;
      const int64_t j = c->nj[p];
      const int64_t lj = j - a0;
      const int tj = c->types[j];
      const double *d = c->nd + 3 * p;
      const double inv_r = 1.0 / r;
      const int tp = tab_pair(c, ti, tj);
      if (sp_in_range(&c->tab[tp], r))
        add_bond(jj, tp, r, d, inv_r, 1.0, 1, atom_row, e_row, s_row, inv_vol, dv, dd);
      const int tgj = tab_dens(c, tj), tgi = tab_dens(c, ti);
      const int gj_in = sp_in_range(&c->tab[tgj], r), gi_in = sp_in_range(&c->tab[tgi], r);
      double tmp, dg_j = 0.0, dg_i = 0.0;
      if (gj_in) sp_eval_uncached(&c->tab[tgj], r, &tmp, &dg_j);
      if (gi_in) sp_eval_uncached(&c->tab[tgi], r, &tmp, &dg_i);
      if (gj_in)
        add_bond(jj, tgj, r, d, inv_r, gradF[li], 0, atom_row, e_row, s_row, inv_vol, dv, dd);
      if (gi_in)
        add_bond(jj, tgi, r, d, inv_r, gradF[lj], 0, atom_row, e_row, s_row, inv_vol, dv2, dd2);
      for (int t = dens_t0; t < dens_t0 + c->ntypes; ++t)
        for (int m = 0; m < c->tab[t].n; ++m) {
          const int col = jj->cm.col_of[jj->cm.col[t] + m];
          if (col < 0) continue;
          const double chain = curvF[li] * drho[li * D + col] * dg_j +
                               curvF[lj] * drho[lj * D + col] * dg_i;
          if (chain == 0.0) continue;
          const double sc = chain * inv_r;
          const double df[3] = {sc * d[0], sc * d[1], sc * d[2]};
          add_col(J, D, atom_row, e_row, s_row, col, d, df, 0.0, c->ew, c->sw_w, inv_vol);
        }
      for (int t = emb_t0; t < emb_t0 + c->ntypes; ++t)
        for (int m = 0; m < c->tab[t].n; ++m) {
          const int col = jj->cm.col_of[jj->cm.col[t] + m];
          if (col < 0) continue;
          const double chain = demb[li * D + col] * dg_j + demb[lj * D + col] * dg_i;
          if (chain == 0.0) continue;
          const double sc = chain * inv_r;
          const double df[3] = {sc * d[0], sc * d[1], sc * d[2]};
          add_col(J, D, atom_row, e_row, s_row, col, d, df, 0.0, c->ew, c->sw_w, inv_vol);
        }
      if (!adp) continue;
      const int tu = tab_dip(c, ti, tj);
      if (sp_in_range(&c->tab[tu], r)) {
        double u, du;
        sp_eval_uncached(&c->tab[tu], r, &u, &du);
        const double *mi = mu + 3 * li, *mj = mu + 3 * lj;
        const double dot = (mi[0] * d[0] + mi[1] * d[1] + mi[2] * d[2]) -
                           (mj[0] * d[0] + mj[1] * d[1] + mj[2] * d[2]);
        sp_param_partials(&c->tab[tu], c->Sjac + c->Soff[tu], r, dv, dd, NULL);
        for (int m = 0; m < c->tab[tu].n; ++m) {
          const int col = jj->cm.col_of[jj->cm.col[tu] + m];
          if (col < 0) continue;
          const double sc = dd[m] * inv_r * dot;
          double df[3];
          for (int a = 0; a < 3; ++a) df[a] = sc * d[a] + dv[m] * (mi[a] - mj[a]);
          add_col(J, D, atom_row, e_row, s_row, col, d, df, 0.0, c->ew, c->sw_w, inv_vol);
        }
        for (int t = dip_t0; t < dip_t0 + c->npair; ++t)
          for (int m = 0; m < c->tab[t].n; ++m) {
            const int col = jj->cm.col_of[jj->cm.col[t] + m];
            if (col < 0) continue;
            double dm[3];
            for (int a = 0; a < 3; ++a)
              dm[a] = dmu[(li * D + col) * 3 + a] - dmu[(lj * D + col) * 3 + a];
            if (dm[0] == 0.0 && dm[1] == 0.0 && dm[2] == 0.0) continue;
            const double sc = du * inv_r * (dm[0] * d[0] + dm[1] * d[1] + dm[2] * d[2]);
            double df[3];
            for (int a = 0; a < 3; ++a) df[a] = sc * d[a] + u * dm[a];
            add_col(J, D, atom_row, e_row, s_row, col, d, df, 0.0, c->ew, c->sw_w, inv_vol);
          }
      }
      const int tw = tab_quad(c, ti, tj);
      if (sp_in_range(&c->tab[tw], r)) {
        double w, dw;
        sp_eval_uncached(&c->tab[tw], r, &w, &dw);
        const double *Li = lam + 9 * li, *Lj = lam + 9 * lj;
        const double nu = quad_nu(Li, d) + quad_nu(Lj, d);
        double xi[3], xj[3];
        quad_xi(Li, d, xi);
        quad_xi(Lj, d, xj);
        sp_param_partials(&c->tab[tw], c->Sjac + c->Soff[tw], r, dv, dd, NULL);
        for (int m = 0; m < c->tab[tw].n; ++m) {
          const int col = jj->cm.col_of[jj->cm.col[tw] + m];
          if (col < 0) continue;
          const double sc = dd[m] * inv_r * nu;
          double df[3];
          for (int a = 0; a < 3; ++a) df[a] = sc * d[a] + (2.0 * dv[m]) * (xi[a] + xj[a]);
          add_col(J, D, atom_row, e_row, s_row, col, d, df, 0.0, c->ew, c->sw_w, inv_vol);
        }
        for (int t = quad_t0; t < quad_t0 + c->npair; ++t)
          for (int m = 0; m < c->tab[t].n; ++m) {
            const int col = jj->cm.col_of[jj->cm.col[t] + m];
            if (col < 0) continue;
            double dl[9];
            int nz = 0;
            for (int q = 0; q < 9; ++q) {
              dl[q] = dlam[(li * D + col) * 9 + q] + dlam[(lj * D + col) * 9 + q];
              nz |= dl[q] != 0.0;
            }
            if (!nz) continue;
            double xq[3];
            quad_xi(dl, d, xq);
            const double sc = dw * inv_r * quad_nu(dl, d);
            double df[3];
            for (int a = 0; a < 3; ++a) df[a] = sc * d[a] + (2.0 * w) * xq[a];
            add_col(J, D, atom_row, e_row, s_row, col, d, df, 0.0, c->ew, c->sw_w, inv_vol);
          }
      }
    }
  }
  free(dv);
  free(dd);
  free(dv2);
  free(dd2);
  free(rho);
  free(drho);
  free(mu);
  free(dmu);
  free(lam);
  free(dlam);
  free(gradF);
  free(curvF);
  free(demb);
}

/* Stillinger-Weber analytic parameter derivatives (the reference has only
 * the FD Jacobian for SW; this chain is pinned against it). */
typedef struct { double v, dv; } dual2;

/* d(v2, dv2)/d(field f) for f in A,B,p,q,delta,a1 */
static void sw_v2_partial(double r, const or_sw *p, int f, double *pv, double *pdv) {
  *pv = 0.0;
  *pdv = 0.0;
  if (r >= p->a1) return;
  const double dd = r - p->a1;
  const double e = exp(p->delta / dd);
  if (e == 0.0) return;
  const double rp = pow(r, -p->p), rq = pow(r, -p->q);
  const double rp1 = pow(r, -p->p - 1.0), rq1 = pow(r, -p->q - 1.0);
  const double lr = log(r);
  const double poly = p->A * rp - p->B * rq;
  const double dpoly = -p->A * p->p * rp1 + p->B * p->q * rq1;
  const double g = -p->delta / (dd * dd); /* d(ln e)/dr */
  double P = 0.0, dP = 0.0, E = 0.0, G = 0.0; /* tangents of poly, dpoly, e, g */
  switch (f) {
  case 0: P = rp; dP = -p->p * rp1; break;
  case 1: P = -rq; dP = p->q * rq1; break;
  case 2: P = -p->A * lr * rp; dP = p->A * rp1 * (p->p * lr - 1.0); break;
  case 3: P = p->B * lr * rq; dP = p->B * rq1 * (1.0 - p->q * lr); break;
  case 4: E = e / dd; G = -1.0 / (dd * dd); break;
  case 5: E = e * p->delta / (dd * dd); G = -2.0 * p->delta / (dd * dd * dd); break;
  default: break;
  }
  /* v2 = poly e ; dv2 = dpoly e + poly e g */
  *pv = P * e + poly * E;
  *pdv = dP * e + dpoly * E + P * e * g + poly * E * g + poly * e * G;
}

/* d(h, dh)/d(field f) for f in gamma (6), a2 (7) */
static void sw_h_partial(double r, const or_sw *p, int f, double *ph, double *pdh) {
  *ph = 0.0;
  *pdh = 0.0;
  if (r >= p->a2) return;
  const double dd = r - p->a2;
  const double h = exp(p->gamma / dd);
  if (h == 0.0) return;
  const double g = -p->gamma / (dd * dd);
  double H = 0.0, G = 0.0;
  if (f == 6) { H = h / dd; G = -1.0 / (dd * dd); }
  else if (f == 7) { H = h * p->gamma / (dd * dd); G = -2.0 * p->gamma / (dd * dd * dd); }
  *ph = H;
  *pdh = H * g + h * G;
}

/* tangent of three_body_forces: every term is a product of one factor from
 * each of (f1|df1), (f2|df2), (g|dg); differentiate term by term. */
static void three_body_tangent(const double *d1, const double *d2, double inv_r1,
                               double inv_r2, double cs, double f1, double df1,
                               double f2, double df2, double g, double dg,
                               double tf1, double tdf1, double tf2, double tdf2,
                               double tg, double tdg, double *fi, double *fj,
                               double *fk) {
  double a[3], b[3], cc[3];
  three_body(d1, d2, inv_r1, inv_r2, cs, tf1, tdf1, f2, df2, g, dg, fi, fj, fk);
  three_body(d1, d2, inv_r1, inv_r2, cs, f1, df1, tf2, tdf2, g, dg, a, b, cc);
  for (int q = 0; q < 3; ++q) { fi[q] += a[q]; fj[q] += b[q]; fk[q] += cc[q]; }
  three_body(d1, d2, inv_r1, inv_r2, cs, f1, df1, f2, df2, tg, tdg, a, b, cc);
  for (int q = 0; q < 3; ++q) { fi[q] += a[q]; fj[q] += b[q]; fk[q] += cc[q]; }
}

typedef struct {
  or_ctx *c;
  double *J;
  int D;
  int *col_sw;  /* npair*8 -> column or -1 */
  int *col_lam; /* nlam -> column or -1 */
} swj_job;

static void sw_jac_config(void *arg, int64_t kk) {
  swj_job *jb = (swj_job *)arg;
  or_ctx *c = jb->c;
  const int k = (int)kk, D = jb->D;
  double *J = jb->J;
  const int64_t a0 = c->atom_off[k], na = c->atom_off[k + 1] - a0;
  const int64_t force0 = c->row_off[k];
  const int64_t e_row = force0 + 3 * na;
  const int64_t s_row = c->sw_w > 0.0 ? e_row + 1 : -1;
  const double inv_vol = 1.0 / c->box[k].vol;
  const double ew = c->ew, sww = c->sw_w;
  for (int64_t li = 0; li < na; ++li) {
    const int64_t i = a0 + li;
    const int ti = c->types[i];
    for (int64_t p = c->pair_off[i]; p < c->pair_off[i + 1]; ++p) {
      const int tj = c->types[c->nj[p]];
      const double r = pair_r(c, p);
      if (r < 1e-14) continue;
      const int slot = pair_ordinal(ti, tj, c->ntypes);
      const double *d = c->nd + 3 * p;
      for (int f = 0; f < 6; ++f) {
        const int col = jb->col_sw[8 * slot + f];
        if (col < 0) continue;
        double pv, pdv;
        sw_v2_partial(r, &c->sw[slot], f, &pv, &pdv);
        if (pv == 0.0 && pdv == 0.0) continue;
        const double s = pdv / r;
        const double df[3] = {s * d[0], s * d[1], s * d[2]};
        add_col(J, D, force0 + 3 * li, e_row, s_row, col, d, df, 0.5 * pv, ew, sww, inv_vol);
      }
    }
  }
  for (int64_t li = 0; li < na; ++li) {
    const int64_t i = a0 + li;
    const int ti = c->types[i];
    const int64_t p0 = c->pair_off[i], nn = c->pair_off[i + 1] - p0;
    for (int64_t jx = 0; jx < nn; ++jx) {
      const int64_t pj = p0 + jx;
      const int64_t j = c->nj[pj];
      const int tj = c->types[j];
      const double r1 = pair_r(c, pj);
      if (r1 < 1e-14) continue;
      const int s1 = pair_ordinal(ti, tj, c->ntypes);
      double h1, dh1;
      sw_h(r1, &c->sw[s1], &h1, &dh1);
      if (h1 == 0.0 && dh1 == 0.0) continue;
      const double inv_r1 = 1.0 / r1;
      const double *d1 = c->nd + 3 * pj;
      for (int64_t kx = jx + 1; kx < nn; ++kx) {
        const int64_t pk = p0 + kx;
        const int64_t kat = c->nj[pk];
        const int tk = c->types[kat];
        const double r2 = pair_r(c, pk);
        if (r2 < 1e-14) continue;
        const int s2 = pair_ordinal(ti, tk, c->ntypes);
        double h2, dh2;
        sw_h(r2, &c->sw[s2], &h2, &dh2);
        if (h2 == 0.0 && dh2 == 0.0) continue;
        const double inv_r2 = 1.0 / r2;
        const double *d2 = c->nd + 3 * pk;
        const int li_ = lambda_index(c, ti, tj, tk);
        const double lam = c->lam[li_];
        const double cs = (d1[0] * d2[0] + d1[1] * d2[1] + d1[2] * d2[2]) * inv_r1 * inv_r2;
        const double cp13 = cs + 1.0 / 3.0;
        const double w = lam * cp13 * cp13, dw = 2.0 * lam * cp13;
        /* parameters touching this triplet: gamma/a2 of slot s1 and s2, lambda */
        int cols[5];
        double tf1[5], tdf1[5], tf2[5], tdf2[5], tg[5], tdg[5];
        int nt = 0;
        for (int f = 6; f < 8; ++f) {
          for (int which = 0; which < 2; ++which) {
            const int slot = which == 0 ? s1 : s2;
            if (which == 1 && s2 == s1) break;
            const int col = jb->col_sw[8 * slot + f];
            if (col < 0) continue;
            double a1t = 0, b1t = 0, a2t = 0, b2t = 0;
            if (slot == s1) sw_h_partial(r1, &c->sw[s1], f, &a1t, &b1t);
            if (slot == s2) sw_h_partial(r2, &c->sw[s2], f, &a2t, &b2t);
            cols[nt] = col;
            tf1[nt] = a1t; tdf1[nt] = b1t; tf2[nt] = a2t; tdf2[nt] = b2t;
            tg[nt] = 0.0; tdg[nt] = 0.0;
            ++nt;
          }
        }
        if (jb->col_lam[li_] >= 0) {
          cols[nt] = jb->col_lam[li_];
          tf1[nt] = tdf1[nt] = tf2[nt] = tdf2[nt] = 0.0;
          tg[nt] = cp13 * cp13;
          tdg[nt] = 2.0 * cp13;
          ++nt;
        }
        for (int q = 0; q < nt; ++q) {
          double fi[3], fj[3], fk[3];
          three_body_tangent(d1, d2, inv_r1, inv_r2, cs, h1, dh1, h2, dh2, w, dw,
                             tf1[q], tdf1[q], tf2[q], tdf2[q], tg[q], tdg[q], fi, fj, fk);
          const double dE = tf1[q] * h2 * w + h1 * tf2[q] * w + h1 * h2 * tg[q];
          const int col = cols[q];
          const int64_t li_j = j - a0, li_k = kat - a0;
          for (int a = 0; a < 3; ++a) {
            J[(force0 + 3 * li + a) * D + col] += fi[a];
            J[(force0 + 3 * li_j + a) * D + col] += fj[a];
            J[(force0 + 3 * li_k + a) * D + col] += fk[a];
          }
          J[e_row * D + col] += ew * dE;
          if (s_row >= 0)
            for (int v = 0; v < 6; ++v) {
              const int A_ = V6[v][0], B_ = V6[v][1];
              J[(s_row + v) * D + col] += sww * (d1[A_] * fj[B_] + d2[A_] * fk[B_]) * inv_vol;
            }
        }
      }
    }
  }
}

/* the analytic chain of a gauge takes pair and density spans that agree
 * wherever a slope compensates a pair */
static int gauge_spans_match(const or_ctx *c) {
  for (int ti = 0; ti < c->ntypes; ++ti)
    for (int tj = 0; tj < c->ntypes; ++tj) {
      if (c->gS[ti] == 0.0) continue;
      const or_spline *p = &c->tab[tab_pair(c, ti, tj)], *g = &c->tab[tab_dens(c, tj)];
      if (p->x[0] != g->x[0] || p->x[p->n - 1] != g->x[g->n - 1]) return 0;
    }
  return 1;
}

int or_gauge_spans_match(const or_ctx *c) { return gauge_spans_match(c); }

int or_jacobian_analytic(or_ctx *c, const double *x, double *J) {
  if (c->gauge && !gauge_spans_match(c)) return -3;
  or_scatter_params(c, x);
  ensure_nl(c);
  const int D = or_param_count(c);
  const int64_t R = or_residual_count(c);
  memset(J, 0, sizeof(double) * (size_t)R * (size_t)D);
  /* eval_forces first so the per-config state is at x (used by nothing in the
   * chain below, but keeps scratch arrays consistent for callers) */
  if (c->kind == OR_SW) {
    swj_job jb;
    jb.c = c;
    jb.J = J;
    jb.D = D;
    jb.col_sw = (int *)malloc(sizeof(int) * 8 * (size_t)c->npair);
    jb.col_lam = (int *)malloc(sizeof(int) * (size_t)(c->nlam ? c->nlam : 1));
    int o = 0;
    for (int s = 0; s < c->npair; ++s)
      for (int f = 0; f < 8; ++f) jb.col_sw[8 * s + f] = c->sw_fixed[8 * s + f] ? -1 : o++;
    for (int q = 0; q < c->nlam; ++q) jb.col_lam[q] = c->lam_fixed[q] ? -1 : o++;
    parallel_for(c->C, sw_jac_config, &jb);
    free(jb.col_sw);
    free(jb.col_lam);
    return 0;
  }
  jac_job jj;
  jj.c = c;
  jj.cm = make_colmap(c);
  jj.J = J;
  jj.D = D;
  parallel_for(c->C, spline_jac_config, &jj);
  /* curvature rows: constant w (1, -2, 1) on the free knots k-1, k, k+1 */
  if (or_smooth_count(c) > 0) {
    int64_t row = c->row_off[c->C];
    for (int t = 0; t < c->ntab; ++t) {
      const or_spline *sp = &c->tab[t];
      int nfree = 0;
      for (int k = 0; k < sp->n; ++k) nfree += !sp->fixed[k];
      if (nfree == 0 || sp->n < 3) continue;
      for (int k = 1; k + 1 < sp->n; ++k, ++row) {
        const double w3[3] = {1.0, -2.0, 1.0};
        for (int d = 0; d < 3; ++d) {
          const int col = jj.cm.col_of[jj.cm.col[t] + k - 1 + d];
          if (col >= 0) J[row * D + col] = c->smooth_w * w3[d];
        }
      }
    }
  }
  free_colmap(&jj.cm);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Levenberg-Marquardt on the normal equations (FP64)                         */
/* The reference drives Eigen's MINPACK LM (src/optimization/solver.cpp:      */
/* 67-85); its iterates are not reproducible without Eigen, so this is the   */
/* algorithm the GPU driver shares: one Jacobian per outer iteration, inner  */
/* damping loop until the gain ratio exceeds 1e-4 (as MINPACK's lmpar loop), */
/* Marquardt diagonal scaling, Nielsen damping update.                       */
/* ------------------------------------------------------------------------ */
static int cholesky_solve(int n, double *A, double *b) {
  for (int j = 0; j < n; ++j) {
    double s = A[(size_t)j * n + j];
    for (int k = 0; k < j; ++k) s -= A[(size_t)j * n + k] * A[(size_t)j * n + k];
    if (!(s > 0.0)) return -1;
    const double l = sqrt(s);
    A[(size_t)j * n + j] = l;
    for (int i = j + 1; i < n; ++i) {
      double t = A[(size_t)i * n + j];
      for (int k = 0; k < j; ++k) t -= A[(size_t)i * n + k] * A[(size_t)j * n + k];
      A[(size_t)i * n + j] = t / l;
    }
  }
  for (int i = 0; i < n; ++i) {
    double t = b[i];
    for (int k = 0; k < i; ++k) t -= A[(size_t)i * n + k] * b[k];
    b[i] = t / A[(size_t)i * n + i];
  }
  for (int i = n - 1; i >= 0; --i) {
    double t = b[i];
    for (int k = i + 1; k < n; ++k) t -= A[(size_t)k * n + i] * b[k];
    b[i] = t / A[(size_t)i * n + i];
  }
  return 0;
}

static double dot(const double *a, const double *b, int64_t n) {
  double s = 0.0;
  for (int64_t i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}

static void normal_eq(const double *J, const double *F, int64_t R, int D,
                      double *A, double *g) {
  memset(A, 0, sizeof(double) * (size_t)D * D);
  memset(g, 0, sizeof(double) * (size_t)D);
  for (int64_t r = 0; r < R; ++r) {
    const double *row = J + r * D;
    for (int a = 0; a < D; ++a) {
      const double ja = row[a];
      if (ja == 0.0) continue;
      g[a] += ja * F[r];
      for (int b = 0; b <= a; ++b) A[(size_t)a * D + b] += ja * row[b];
    }
  }
  for (int a = 0; a < D; ++a)
    for (int b = a + 1; b < D; ++b) A[(size_t)a * D + b] = A[(size_t)b * D + a];
}

int or_lm(or_ctx *c, double *x, int max_iter, int use_fd, double *cost_hist) {
  const int D = or_param_count(c);
  const int64_t R = or_residual_count(c);
  double *F = (double *)malloc(sizeof(double) * (size_t)R);
  double *Fn = (double *)malloc(sizeof(double) * (size_t)R);
  double *J = (double *)malloc(sizeof(double) * (size_t)R * D);
  double *A = (double *)malloc(sizeof(double) * (size_t)D * D);
  double *M = (double *)malloc(sizeof(double) * (size_t)D * D);
  double *g = (double *)malloc(sizeof(double) * (size_t)D);
  double *dx = (double *)malloc(sizeof(double) * (size_t)D);
  double *xn = (double *)malloc(sizeof(double) * (size_t)D);
  double *dg = (double *)malloc(sizeof(double) * (size_t)D);
  or_residual(c, x, F);
  double cost = dot(F, F, R);
  if (use_fd) or_jacobian_fd(c, x, J); else or_jacobian_analytic(c, x, J);
  normal_eq(J, F, R, D, A, g);
  double maxd = 0.0;
  for (int a = 0; a < D; ++a) if (A[(size_t)a * D + a] > maxd) maxd = A[(size_t)a * D + a];
  double mu = 1e-3 * maxd, nu = 2.0;
  int accepted = 0;
  for (int it = 0; it < max_iter; ++it) {
    int ok = 0;
    for (int inner = 0; inner < 16; ++inner) {
      double dmax = 0.0;
      for (int a = 0; a < D; ++a) if (A[(size_t)a * D + a] > dmax) dmax = A[(size_t)a * D + a];
      for (int a = 0; a < D; ++a) {
        const double da = A[(size_t)a * D + a];
        dg[a] = da > 1e-12 * dmax ? da : 1e-12 * dmax;
        if (!(dg[a] > 0.0)) dg[a] = 1.0;
      }
      memcpy(M, A, sizeof(double) * (size_t)D * D);
      for (int a = 0; a < D; ++a) {
        M[(size_t)a * D + a] += mu * dg[a];
        dx[a] = -g[a];
      }
      if (cholesky_solve(D, M, dx) != 0) {
        mu *= nu;
        nu *= 2.0;
        continue;
      }
      for (int a = 0; a < D; ++a) xn[a] = x[a] + dx[a];
      or_residual(c, xn, Fn);
      const double cn = dot(Fn, Fn, R);
      double pred = 0.0;
      for (int a = 0; a < D; ++a) pred += dx[a] * (mu * dg[a] * dx[a] - g[a]);
      const double rho = (cost - cn) / pred;
      if (isfinite(cn) && pred > 0.0 && rho > 1e-4) {
        memcpy(x, xn, sizeof(double) * (size_t)D);
        memcpy(F, Fn, sizeof(double) * (size_t)R);
        cost = cn;
        const double t = 2.0 * rho - 1.0;
        const double f = 1.0 - t * t * t;
        mu *= f > 1.0 / 3.0 ? f : 1.0 / 3.0;
        nu = 2.0;
        ok = 1;
        break;
      }
      mu *= nu;
      nu *= 2.0;
    }
    if (cost_hist) cost_hist[it] = cost;
    if (!ok) {
      for (int r = it + 1; r < max_iter && cost_hist; ++r) cost_hist[r] = cost;
      break;
    }
    ++accepted;
    if (it + 1 < max_iter) {
      if (use_fd) or_jacobian_fd(c, x, J); else or_jacobian_analytic(c, x, J);
      normal_eq(J, F, R, D, A, g);
    }
  }
  or_scatter_params(c, x);
  free(F); free(Fn); free(J); free(A); free(M); free(g); free(dx); free(xn); free(dg);
  return accepted;
}

/* ------------------------------------------------------------------------ */
/* Levenberg-Marquardt with the reference's stopping semantics               */
/* ------------------------------------------------------------------------ */
/* EigenLMSolver::minimize (src/optimization/solver.cpp:67-85) runs Eigen's   */
/* LevenbergMarquardt of the NonLinearOptimization module (Eigen 3.4,         */
/* cmake/Dependencies.cmake:1; the dependency is not vendored). That class    */
/* is the MINPACK lmder algorithm published by More (1978), "The Levenberg-   */
/* Marquardt algorithm: implementation and theory", restated here: QR with    */
/* column pivoting of J, the lmpar search for the damping parameter that puts */
/* the scaled step on the trust-region boundary, the actual/predicted         */
/* reduction ratio with its step-bound update, and the termination tests      */
/* with Eigen's LevenbergMarquardtSpace status codes (1 relative reduction,   */
/* 2 relative error, 3 both, 4 cosine, 5 too many evaluations, 6-8 tolerances */
/* too small; 0 improper input). solver.cpp:73-84 sets maxfev =               */
/* max(max_iter, 1) * 100 + 100, xtol, ftol, calls minimizeOneStep at most    */
/* max_iter times and reports 2 when the solver is still running.             */
/* ------------------------------------------------------------------------ */
#define OR_EPS 2.220446049250313e-16
#define OR_DWARF 2.2250738585072014e-308

static double nrm2(const double *a, int64_t n) { return sqrt(dot(a, a, n)); }

/* Householder QR of the m x n column-major a with column pivoting (largest
 * remaining column norm, first on ties); a is overwritten. R (n x n,
 * column-major, upper) and qtf = (Q^T f)[0..n) on return; perm[k] = original
 * column at position k. Rank: |r_kk| > eps n max|r_jj| (ColPivHouseholderQR's
 * default threshold). */
static int qr_pivot(int64_t m, int n, double *a, const double *f, int *perm,
                    double *R, double *qtf) {
  double *fw = (double *)malloc(sizeof(double) * (size_t)m);
  memcpy(fw, f, sizeof(double) * (size_t)m);
  for (int j = 0; j < n; ++j) perm[j] = j;
  double maxpiv = 0.0;
  for (int k = 0; k < n; ++k) {
    int p = k;
    double best = -1.0;
    for (int j = k; j < n; ++j) {
      const double s = dot(a + (size_t)j * m + k, a + (size_t)j * m + k, m - k);
      if (s > best) {
        best = s;
        p = j;
      }
    }
    if (p != k) {
      for (int64_t i = 0; i < m; ++i) {
        const double t = a[(size_t)k * m + i];
        a[(size_t)k * m + i] = a[(size_t)p * m + i];
        a[(size_t)p * m + i] = t;
      }
      const int t = perm[k];
      perm[k] = perm[p];
      perm[p] = t;
    }
    double *v = a + (size_t)k * m + k; /* the column below the diagonal */
    const double norm = sqrt(best);
    if (norm > 0.0) {
      const double alpha = v[0] >= 0.0 ? -norm : norm;
      v[0] -= alpha; /* v = x - alpha e1 */
      const double vtv = dot(v, v, m - k);
      if (vtv > 0.0) {
        for (int j = k + 1; j < n; ++j) {
          double *aj = a + (size_t)j * m + k;
          const double s = 2.0 * dot(v, aj, m - k) / vtv;
          for (int64_t i = 0; i < m - k; ++i) aj[i] -= s * v[i];
        }
        const double s = 2.0 * dot(v, fw + k, m - k) / vtv;
        for (int64_t i = 0; i < m - k; ++i) fw[k + i] -= s * v[i];
      }
      v[0] = alpha;
      for (int64_t i = 1; i < m - k; ++i) v[i] = 0.0;
    }
    if (fabs(v[0]) > maxpiv) maxpiv = fabs(v[0]);
  }
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) R[(size_t)j * n + i] = i <= j ? a[(size_t)j * m + i] : 0.0;
  for (int i = 0; i < n; ++i) qtf[i] = fw[i];
  free(fw);
  int rank = 0;
  for (int k = 0; k < n; ++k) rank += fabs(R[(size_t)k * n + k]) > OR_EPS * n * maxpiv;
  return rank;
}

/* least-squares solution of (R P^T; sqrt(par) D) x = (qtf; 0) by a Householder
 * QR of the stacked 2n x n system (no pivoting): S (n x n, column-major upper,
 * S^T S = P^T (J^T J + par D^2) P) and z = P^T x on return */
static void damped_solve(int n, const double *R, const int *perm, const double *diag,
                         double par, const double *qtf, double *S, double *z) {
  const int m = 2 * n;
  double *a = (double *)calloc((size_t)m * n, sizeof(double));
  double *b = (double *)calloc((size_t)m, sizeof(double));
  const double sp = sqrt(par);
  for (int j = 0; j < n; ++j) {
    for (int i = 0; i <= j; ++i) a[(size_t)j * m + i] = R[(size_t)j * n + i];
    a[(size_t)j * m + n + j] = sp * diag[perm[j]];
    b[j] = qtf[j];
  }
  for (int k = 0; k < n; ++k) {
    double *v = a + (size_t)k * m + k;
    const double norm = nrm2(v, m - k);
    if (norm > 0.0) {
      const double alpha = v[0] >= 0.0 ? -norm : norm;
      v[0] -= alpha;
      const double vtv = dot(v, v, m - k);
      if (vtv > 0.0) {
        for (int j = k + 1; j < n; ++j) {
          double *aj = a + (size_t)j * m + k;
          const double s = 2.0 * dot(v, aj, m - k) / vtv;
          for (int i = 0; i < m - k; ++i) aj[i] -= s * v[i];
        }
        const double s = 2.0 * dot(v, b + k, m - k) / vtv;
        for (int i = 0; i < m - k; ++i) b[k + i] -= s * v[i];
      }
      v[0] = alpha;
    }
  }
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) S[(size_t)j * n + i] = i <= j ? a[(size_t)j * m + i] : 0.0;
  for (int i = n - 1; i >= 0; --i) {
    double t = b[i];
    for (int j = i + 1; j < n; ++j) t -= S[(size_t)j * n + i] * z[j];
    z[i] = t / S[(size_t)i * n + i];
  }
  free(a);
  free(b);
}

/* w := U^-T w for an upper-triangular column-major U (forward substitution) */
static void solve_upper_t(int n, const double *U, double *w) {
  for (int j = 0; j < n; ++j) {
    double t = w[j];
    for (int i = 0; i < j; ++i) t -= U[(size_t)j * n + i] * w[i];
    w[j] = t / U[(size_t)j * n + j];
  }
}

/* lmpar (More 1978, sections 5-7): par >= 0 and x with (J^T J + par D^2) x =
 * J^T f such that | |D x| - delta | <= 0.1 delta, or par = 0 when the
 * Gauss-Newton step (minimum-norm over the rank of R) is within 1.1 delta.
 * Returns the number of damped solves. */
static int lmpar_qr(int n, const double *R, const int *perm, int rank, const double *diag,
                    const double *qtf, double delta, double *par, double *x) {
  double *z = (double *)calloc((size_t)n, sizeof(double));
  double *w = (double *)malloc(sizeof(double) * (size_t)n);
  double *S = (double *)malloc(sizeof(double) * (size_t)n * n);
  /* Gauss-Newton direction over the leading rank columns */
  for (int i = rank - 1; i >= 0; --i) {
    double t = qtf[i];
    for (int j = i + 1; j < rank; ++j) t -= R[(size_t)j * n + i] * z[j];
    z[i] = t / R[(size_t)i * n + i];
  }
  for (int j = 0; j < n; ++j) x[perm[j]] = z[j];
  double dxnorm = 0.0;
  for (int j = 0; j < n; ++j) dxnorm += diag[j] * x[j] * diag[j] * x[j];
  dxnorm = sqrt(dxnorm);
  double fp = dxnorm - delta;
  int solves = 0;
  if (fp <= 0.1 * delta) {
    *par = 0.0;
    goto done;
  }
  {
    /* lower bound from the Newton step at par = 0 (full rank only) */
    double parl = 0.0;
    if (rank == n) {
      for (int j = 0; j < n; ++j) w[j] = diag[perm[j]] * diag[perm[j]] * z[j] / dxnorm;
      solve_upper_t(n, R, w);
      const double t = nrm2(w, n);
      parl = fp / delta / t / t;
    }
    /* upper bound |D^-1 J^T f| / delta */
    for (int j = 0; j < n; ++j) {
      double s = 0.0;
      for (int i = 0; i <= j; ++i) s += R[(size_t)j * n + i] * qtf[i];
      w[j] = s / diag[perm[j]];
    }
    const double gnorm = nrm2(w, n);
    double paru = gnorm / delta;
    if (paru == 0.0) paru = OR_DWARF / (delta < 0.1 ? delta : 0.1);
    double p = *par;
    if (p < parl) p = parl;
    if (p > paru) p = paru;
    if (p == 0.0) p = gnorm / dxnorm;
    for (int it = 1;; ++it) {
      if (p == 0.0) p = OR_DWARF > 0.001 * paru ? OR_DWARF : 0.001 * paru;
      damped_solve(n, R, perm, diag, p, qtf, S, z);
      ++solves;
      for (int j = 0; j < n; ++j) x[perm[j]] = z[j];
      dxnorm = 0.0;
      for (int j = 0; j < n; ++j) dxnorm += diag[j] * x[j] * diag[j] * x[j];
      dxnorm = sqrt(dxnorm);
      const double prev = fp;
      fp = dxnorm - delta;
      if (fabs(fp) <= 0.1 * delta || (parl == 0.0 && fp <= prev && prev < 0.0) || it == 10) break;
      /* Newton correction */
      for (int j = 0; j < n; ++j) w[j] = diag[perm[j]] * diag[perm[j]] * z[j] / dxnorm;
      solve_upper_t(n, S, w);
      const double t = nrm2(w, n);
      const double parc = fp / delta / t / t;
      if (fp > 0.0 && p > parl) parl = p;
      if (fp < 0.0 && p < paru) paru = p;
      p = p + parc > parl ? p + parc : parl;
    }
    *par = p;
  }
done:
  free(z);
  free(w);
  free(S);
  return solves;
}

int or_lmpar(int n, const double *R, const int32_t *perm, int rank, const double *diag,
             const double *qtf, double delta, double *par, double *x) {
  int *pm = (int *)malloc(sizeof(int) * (size_t)n);
  for (int j = 0; j < n; ++j) pm[j] = perm[j];
  const int s = lmpar_qr(n, R, pm, rank, diag, qtf, delta, par, x);
  free(pm);
  return s;
}

int or_qr_pivot(int64_t m, int n, const double *J, const double *f, int32_t *perm,
                double *R, double *qtf) {
  double *a = (double *)malloc(sizeof(double) * (size_t)m * n);
  for (int64_t i = 0; i < m; ++i)
    for (int j = 0; j < n; ++j) a[(size_t)j * m + i] = J[i * n + j];
  int *pm = (int *)malloc(sizeof(int) * (size_t)n);
  const int rank = qr_pivot(m, n, a, f, pm, R, qtf);
  for (int j = 0; j < n; ++j) perm[j] = pm[j];
  free(pm);
  free(a);
  return rank;
}

int or_lmder(int64_t m, int n, or_fcn fcn, or_jac jac, void *user, double *x,
             const or_lm_opts *o, int32_t *counts, double *cost_hist) {
  const int maxfev = o->maxfev > 0 ? o->maxfev : (o->max_iter > 1 ? o->max_iter : 1) * 100 + 100;
  if (counts) counts[0] = counts[1] = counts[2] = 0;
  if (n <= 0 || m < n || o->ftol < 0.0 || o->xtol < 0.0 || o->gtol < 0.0 || maxfev <= 0 ||
      o->factor <= 0.0)
    return 0;
  double *f = (double *)malloc(sizeof(double) * (size_t)m);
  double *fn = (double *)malloc(sizeof(double) * (size_t)m);
  double *J = (double *)malloc(sizeof(double) * (size_t)m * n);
  double *a = (double *)malloc(sizeof(double) * (size_t)m * n);
  double *R = (double *)malloc(sizeof(double) * (size_t)n * n);
  double *qtf = (double *)malloc(sizeof(double) * (size_t)n);
  double *diag = (double *)malloc(sizeof(double) * (size_t)n);
  double *cn = (double *)malloc(sizeof(double) * (size_t)n);
  double *step = (double *)malloc(sizeof(double) * (size_t)n);
  double *xt = (double *)malloc(sizeof(double) * (size_t)n);
  double *t3 = (double *)malloc(sizeof(double) * (size_t)n);
  int *perm = (int *)malloc(sizeof(int) * (size_t)n);
  fcn(user, x, f);
  int nfev = 1, njev = 0, iter = 1, steps = 0;
  double fnorm = nrm2(f, m), par = 0.0, delta = 0.0, xnorm = 0.0;
  const int RUNNING = -1;
  int status = RUNNING;
  for (int call = 0; call < o->max_iter && status == RUNNING; ++call) {
    /* minimizeOneStep */
    jac(user, x, J);
    ++njev;
    for (int j = 0; j < n; ++j) {
      double s = 0.0;
      for (int64_t i = 0; i < m; ++i) {
        a[(size_t)j * m + i] = J[i * n + j];
        s += J[i * n + j] * J[i * n + j];
      }
      cn[j] = sqrt(s);
    }
    const int rank = qr_pivot(m, n, a, f, perm, R, qtf);
    if (iter == 1) {
      for (int j = 0; j < n; ++j) diag[j] = cn[j] != 0.0 ? cn[j] : 1.0;
      xnorm = 0.0;
      for (int j = 0; j < n; ++j) xnorm += diag[j] * x[j] * diag[j] * x[j];
      xnorm = sqrt(xnorm);
      delta = o->factor * xnorm;
      if (delta == 0.0) delta = o->factor;
    }
    /* scaled gradient cosine */
    double gnorm = 0.0;
    if (fnorm != 0.0)
      for (int j = 0; j < n; ++j) {
        const int l = perm[j];
        if (cn[l] == 0.0) continue;
        double s = 0.0;
        for (int i = 0; i <= j; ++i) s += R[(size_t)j * n + i] * (qtf[i] / fnorm);
        const double v = fabs(s / cn[l]);
        if (v > gnorm) gnorm = v;
      }
    if (gnorm <= o->gtol) {
      status = 4;
      break;
    }
    for (int j = 0; j < n; ++j)
      if (cn[j] > diag[j]) diag[j] = cn[j];
    double ratio = 0.0;
    do {
      lmpar_qr(n, R, perm, rank, diag, qtf, delta, &par, step);
      for (int j = 0; j < n; ++j) {
        step[j] = -step[j];
        xt[j] = x[j] + step[j];
      }
      double pnorm = 0.0;
      for (int j = 0; j < n; ++j) pnorm += diag[j] * step[j] * diag[j] * step[j];
      pnorm = sqrt(pnorm);
      if (iter == 1 && pnorm < delta) delta = pnorm;
      fcn(user, xt, fn);
      ++nfev;
      const double fnorm1 = nrm2(fn, m);
      double actred = -1.0;
      if (0.1 * fnorm1 < fnorm) actred = 1.0 - (fnorm1 / fnorm) * (fnorm1 / fnorm);
      /* |J p| = |R P^T p| */
      for (int i = 0; i < n; ++i) {
        double s = 0.0;
        for (int j = i; j < n; ++j) s += R[(size_t)j * n + i] * step[perm[j]];
        t3[i] = s;
      }
      const double temp1 = nrm2(t3, n) / fnorm, temp2 = sqrt(par) * pnorm / fnorm;
      const double prered = temp1 * temp1 + temp2 * temp2 / 0.5;
      const double dirder = -(temp1 * temp1 + temp2 * temp2);
      ratio = prered != 0.0 ? actred / prered : 0.0;
      if (ratio <= 0.25) {
        double temp = actred >= 0.0 ? 0.5 : 0.5 * dirder / (dirder + 0.5 * actred);
        if (0.1 * fnorm1 >= fnorm || temp < 0.1) temp = 0.1;
        delta = temp * (delta < pnorm / 0.1 ? delta : pnorm / 0.1);
        par /= temp;
      } else if (par == 0.0 || ratio >= 0.75) {
        delta = pnorm / 0.5;
        par = 0.5 * par;
      }
      if (ratio >= 1e-4) {
        memcpy(x, xt, sizeof(double) * (size_t)n);
        memcpy(f, fn, sizeof(double) * (size_t)m);
        xnorm = 0.0;
        for (int j = 0; j < n; ++j) xnorm += diag[j] * x[j] * diag[j] * x[j];
        xnorm = sqrt(xnorm);
        fnorm = fnorm1;
        ++iter;
        ++steps;
      }
      const int ftol_ok = fabs(actred) <= o->ftol && prered <= o->ftol && 0.5 * ratio <= 1.0;
      if (ftol_ok && delta <= o->xtol * xnorm) { status = 3; break; }
      if (ftol_ok) { status = 1; break; }
      if (delta <= o->xtol * xnorm) { status = 2; break; }
      if (nfev >= maxfev) { status = 5; break; }
      if (fabs(actred) <= OR_EPS && prered <= OR_EPS && 0.5 * ratio <= 1.0) { status = 6; break; }
      if (delta <= OR_EPS * xnorm) { status = 7; break; }
      if (gnorm <= OR_EPS) { status = 8; break; }
    } while (ratio < 1e-4);
    if (cost_hist) cost_hist[call] = fnorm * fnorm;
  }
  if (counts) {
    counts[0] = nfev;
    counts[1] = njev;
    counts[2] = steps;
  }
  free(f); free(fn); free(J); free(a); free(R); free(qtf); free(diag); free(cn);
  free(step); free(xt); free(t3); free(perm);
  return status == RUNNING ? 2 : status;
}

/* force fitting through or_lmder: residual and analytic Jacobian of the
 * context */
static int ctx_fcn(void *u, const double *x, double *f) {
  or_residual((or_ctx *)u, x, f);
  return 0;
}
static int ctx_jac(void *u, const double *x, double *J) {
  return or_jacobian_analytic((or_ctx *)u, x, J);
}
int or_lm_minpack(or_ctx *c, double *x, const or_lm_opts *o, int32_t *counts,
                  double *cost_hist) {
  const int rc = or_lmder(or_residual_count(c), or_param_count(c), ctx_fcn, ctx_jac, c, x, o,
                          counts, cost_hist);
  or_scatter_params(c, x);
  return rc;
}

/* NormalEquationsSolver::minimize (src/optimization/solver.cpp:242-266): one
 * step x -= (J^T J + lambda I)^-1 J^T F(x), lambda = ridge if > 0, else
 * 1e-8 max(mean diag(J^T J), 1); the reference's LDLT equals the Cholesky
 * solve of the SPD matrix. Returns -1 when the matrix is not positive
 * definite. */
int or_normal_equations_solve(or_ctx *c, double ridge, double *x) {
  const int D = or_param_count(c);
  const int64_t R = or_residual_count(c);
  if (D == 0 || R == 0) return 0;
  double *F = (double *)malloc(sizeof(double) * (size_t)R);
  double *J = (double *)malloc(sizeof(double) * (size_t)R * D);
  double *A = (double *)malloc(sizeof(double) * (size_t)D * D);
  double *g = (double *)malloc(sizeof(double) * (size_t)D);
  or_residual(c, x, F);
  or_jacobian_analytic(c, x, J);
  normal_eq(J, F, R, D, A, g);
  double lambda = ridge;
  if (lambda <= 0.0) {
    double mean = 0.0;
    for (int k = 0; k < D; ++k) mean += A[(size_t)k * D + k];
    mean /= D;
    lambda = 1e-8 * (mean > 1.0 ? mean : 1.0);
  }
  for (int k = 0; k < D; ++k) A[(size_t)k * D + k] += lambda;
  const int rc = cholesky_solve(D, A, g);
  if (rc == 0)
    for (int k = 0; k < D; ++k) x[k] -= g[k];
  or_scatter_params(c, x);
  free(F);
  free(J);
  free(A);
  free(g);
  return rc;
}

/* ------------------------------------------------------------------------ */
/* population cost and differential evolution                                */
/* ------------------------------------------------------------------------ */
void or_population_cost(or_ctx *c, const double *X, int np, double *cost) {
  const int D = or_param_count(c);
  const int64_t R = or_residual_count(c);
  double *F = (double *)malloc(sizeof(double) * (size_t)R);
  double *x0 = (double *)malloc(sizeof(double) * (size_t)D);
  or_gather_params(c, x0);
  for (int m = 0; m < np; ++m) {
    or_residual(c, X + (size_t)m * D, F);
    cost[m] = dot(F, F, R);
  }
  or_scatter_params(c, x0);
  free(F);
  free(x0);
}

static uint64_t splitmix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
uint64_t or_rng_u64(uint64_t seed, uint64_t stream, uint64_t index) {
  return splitmix64(splitmix64(seed ^ splitmix64(stream)) + index);
}
double or_rng_unit(uint64_t seed, uint64_t stream, uint64_t index) {
  return (double)(or_rng_u64(seed, stream, index) >> 11) * 0x1.0p-53;
}

/* DE/rand/1/bin (Storn & Price, J. Global Optim. 11 (1997) 341-359) with
 * synchronous generations: the BoostDESolver semantics of
 * src/optimization/solver.cpp:113-155 with a documented counter RNG.
 *   init:   member 0 = clamp(x) (:136-140), member m > 0 component k =
 *           lo + u (hi - lo), u = unit(seed, 0, m*D + k)
 *   trial:  r1, r2, r3 distinct and != m by rejection (stream 2g, index
 *           m*1024 + t); jrand and crossover draws (stream 2g+1, index
 *           m*(D+1) + k, jrand at k = D); component k takes the mutant
 *           clamp(x_r1 + F (x_r2 - x_r3)) when its draw is < CR or k = jrand,
 *           else the parent's value
 *   select: the trial replaces the member when its cost is <= the member's */
static int de_pick(uint64_t seed, uint64_t stream, int m, int np, int *t,
                   int e0, int e1, int e2) {
  for (;;) {
    const int r = (int)(or_rng_unit(seed, stream, (uint64_t)m * 1024u + (uint64_t)(*t)) * np);
    ++*t;
    if (r != e0 && r != e1 && r != e2 && r < np) return r;
  }
}

static double clampd(double v, double lo, double hi) {
  if (v < lo) v = lo;
  if (v > hi) v = hi;
  return v;
}

void or_de_init(const double *x, const double *lo, const double *hi, int np, int D,
                uint64_t seed, double *pop) {
  for (int k = 0; k < D; ++k) pop[k] = clampd(x[k], lo[k], hi[k]);
  for (int m = 1; m < np; ++m)
    for (int k = 0; k < D; ++k)
      pop[(size_t)m * D + k] = lo[k] + or_rng_unit(seed, 0, (uint64_t)m * D + k) * (hi[k] - lo[k]);
}

void or_de_trial(const double *pop, int np, int D, const double *lo, const double *hi,
                 double Fm, double CR, uint64_t seed, int g, int m, double *out,
                 int *picks) {
  int t = 0;
  const int r1 = de_pick(seed, 2u * (uint64_t)g, m, np, &t, m, -1, -1);
  const int r2 = de_pick(seed, 2u * (uint64_t)g, m, np, &t, m, r1, -1);
  const int r3 = de_pick(seed, 2u * (uint64_t)g, m, np, &t, m, r1, r2);
  const uint64_t cs = 2u * (uint64_t)g + 1u;
  const int jr = (int)(or_rng_unit(seed, cs, (uint64_t)m * (D + 1) + D) * D);
  for (int k = 0; k < D; ++k) {
    if (or_rng_unit(seed, cs, (uint64_t)m * (D + 1) + k) < CR || k == jr)
      out[k] = clampd(pop[(size_t)r1 * D + k] + Fm * (pop[(size_t)r2 * D + k] - pop[(size_t)r3 * D + k]),
                      lo[k], hi[k]);
    else
      out[k] = pop[(size_t)m * D + k];
  }
  if (picks) {
    picks[0] = r1;
    picks[1] = r2;
    picks[2] = r3;
    picks[3] = jr;
  }
}

int or_de_select(int np, int D, double *pop, double *cost, const double *trial,
                 const double *tcost, unsigned char *kept) {
  int bm = 0;
  for (int m = 0; m < np; ++m) {
    const int take = tcost[m] <= cost[m];
    if (take) {
      cost[m] = tcost[m];
      memcpy(pop + (size_t)m * D, trial + (size_t)m * D, sizeof(double) * (size_t)D);
    }
    if (kept) kept[m] = (unsigned char)take;
    if (cost[m] < cost[bm]) bm = m;
  }
  return bm;
}

double or_de_fn(or_popcost fn, void *user, int D, double *x, const double *lo,
                const double *hi, int np, int generations, double Fm, double CR,
                uint64_t seed, double *best_hist) {
  double *pop = (double *)malloc(sizeof(double) * (size_t)np * D);
  double *trial = (double *)malloc(sizeof(double) * (size_t)np * D);
  double *cost = (double *)malloc(sizeof(double) * (size_t)np);
  double *tcost = (double *)malloc(sizeof(double) * (size_t)np);
  or_de_init(x, lo, hi, np, D, seed, pop);
  fn(user, pop, np, cost);
  int bm = 0;
  for (int m = 1; m < np; ++m) if (cost[m] < cost[bm]) bm = m;
  for (int g = 1; g <= generations; ++g) {
    for (int m = 0; m < np; ++m)
      or_de_trial(pop, np, D, lo, hi, Fm, CR, seed, g, m, trial + (size_t)m * D, NULL);
    fn(user, trial, np, tcost);
    bm = or_de_select(np, D, pop, cost, trial, tcost, NULL);
    if (best_hist) best_hist[g - 1] = cost[bm];
  }
  memcpy(x, pop + (size_t)bm * D, sizeof(double) * (size_t)D);
  const double b = cost[bm];
  free(pop); free(trial); free(cost); free(tcost);
  return b;
}

static void de_ctx_cost(void *user, const double *X, int np, double *cost) {
  or_population_cost((or_ctx *)user, X, np, cost);
}

double or_de(or_ctx *c, double *x, const double *lo, const double *hi, int np,
             int generations, double Fm, double CR, uint64_t seed,
             double *best_hist) {
  const double b = or_de_fn(de_ctx_cost, c, or_param_count(c), x, lo, hi, np, generations, Fm,
                            CR, seed, best_hist);
  or_scatter_params(c, x);
  return b;
}

/* ------------------------------------------------------------------------ */
/* MLP surrogate: FP64 forward / loss gradient (CPU check of the GPU kernels) */
/* y = W3 . gelu(W2 gelu(W1 x + b1) + b2) + b3, gelu tanh form,              */
/* loss = mean over samples of 0.5 (y - t)^2                                 */
/* ------------------------------------------------------------------------ */
static double gelu(double u) {
  const double k = 0.7978845608028654; /* sqrt(2/pi) */
  return 0.5 * u * (1.0 + tanh(k * (u + 0.044715 * u * u * u)));
}
static double gelu_d(double u) {
  const double k = 0.7978845608028654;
  const double z = k * (u + 0.044715 * u * u * u);
  const double th = tanh(z);
  return 0.5 * (1.0 + th) + 0.5 * u * (1.0 - th * th) * k * (1.0 + 3.0 * 0.044715 * u * u);
}

void or_mlp_forward(int n, int d, int h, const double *X, const double *W1,
                    const double *b1, const double *W2, const double *b2,
                    const double *W3, double b3, double *y) {
  double *h1 = (double *)malloc(sizeof(double) * (size_t)h);
  double *h2 = (double *)malloc(sizeof(double) * (size_t)h);
  for (int s = 0; s < n; ++s) {
    const double *x = X + (size_t)s * d;
    for (int a = 0; a < h; ++a) h1[a] = gelu(dot(W1 + (size_t)a * d, x, d) + b1[a]);
    for (int a = 0; a < h; ++a) h2[a] = gelu(dot(W2 + (size_t)a * h, h1, h) + b2[a]);
    y[s] = dot(W3, h2, h) + b3;
  }
  free(h1);
  free(h2);
}

double or_mlp_loss_grad(int n, int d, int h, const double *X, const double *t,
                        const double *W1, const double *b1, const double *W2,
                        const double *b2, const double *W3, double b3,
                        double *gW1, double *gb1, double *gW2, double *gb2,
                        double *gW3, double *gb3) {
  double *z1 = (double *)malloc(sizeof(double) * (size_t)h);
  double *h1 = (double *)malloc(sizeof(double) * (size_t)h);
  double *z2 = (double *)malloc(sizeof(double) * (size_t)h);
  double *h2 = (double *)malloc(sizeof(double) * (size_t)h);
  double *e2 = (double *)malloc(sizeof(double) * (size_t)h);
  double *e1 = (double *)malloc(sizeof(double) * (size_t)h);
  memset(gW1, 0, sizeof(double) * (size_t)h * d);
  memset(gb1, 0, sizeof(double) * (size_t)h);
  memset(gW2, 0, sizeof(double) * (size_t)h * h);
  memset(gb2, 0, sizeof(double) * (size_t)h);
  memset(gW3, 0, sizeof(double) * (size_t)h);
  *gb3 = 0.0;
  double loss = 0.0;
  for (int s = 0; s < n; ++s) {
    const double *x = X + (size_t)s * d;
    for (int a = 0; a < h; ++a) { z1[a] = dot(W1 + (size_t)a * d, x, d) + b1[a]; h1[a] = gelu(z1[a]); }
    for (int a = 0; a < h; ++a) { z2[a] = dot(W2 + (size_t)a * h, h1, h) + b2[a]; h2[a] = gelu(z2[a]); }
    const double y = dot(W3, h2, h) + b3;
    const double r = y - t[s];
    loss += 0.5 * r * r;
    const double gy = r / n;
    *gb3 += gy;
    for (int a = 0; a < h; ++a) {
      gW3[a] += gy * h2[a];
      e2[a] = gy * W3[a] * gelu_d(z2[a]);
      gb2[a] += e2[a];
    }
    for (int a = 0; a < h; ++a)
      for (int b = 0; b < h; ++b) gW2[(size_t)a * h + b] += e2[a] * h1[b];
    for (int b = 0; b < h; ++b) {
      double s2 = 0.0;
      for (int a = 0; a < h; ++a) s2 += e2[a] * W2[(size_t)a * h + b];
      e1[b] = s2 * gelu_d(z1[b]);
      gb1[b] += e1[b];
    }
    for (int a = 0; a < h; ++a)
      for (int q = 0; q < d; ++q) gW1[(size_t)a * d + q] += e1[a] * x[q];
  }
  free(z1); free(h1); free(z2); free(h2); free(e2); free(e1);
  return loss / n;
}

/* ------------------------------------------------------------------------ */
/* wall-clock CPU baseline                                                   */
/* ------------------------------------------------------------------------ */
static double now_s(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

double or_time_fd_jacobian(or_ctx *c, const double *x, int ncols, double *J) {
  const int64_t R = or_residual_count(c);
  double *F = (double *)malloc(sizeof(double) * (size_t)R);
  const double t0 = now_s();
  or_residual(c, x, F);
  or_jacobian_fd_cols(c, x, 0, ncols, J);
  const double t1 = now_s();
  free(F);
  return t1 - t0;
}

double or_time_residual(or_ctx *c, const double *x, int reps, double *F) {
  const double t0 = now_s();
  for (int r = 0; r < reps; ++r) or_residual(c, x, F);
  return now_s() - t0;
}
