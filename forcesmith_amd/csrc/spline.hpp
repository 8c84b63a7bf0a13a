// Device-side makima spline evaluation on a uniform knot grid.
//
// Two forms, both restating src/potentials/spline.cpp / spline.hpp:
//  * FP64 "site" form for pair tables (cached-site semantics, spline.hpp:
//    47-98 with prepare_site spline.cpp:281-324): upper_bound interval,
//    linear site at r <= x0 or r >= xn, Hermite basis dotted with
//    (y_i, y_i+1, s_i, s_i+1).
//  * FP64 uncached form for the embedding F(rho) (spline.cpp:215-279),
//    extrapolating with the end secant outside the span.
//  * FP32 form for the mixed-precision bulk: interval from the uniform grid,
//    power-form coefficients {a, b, c, d} of the interval (formed in FP64,
//    held as FP32 hi/lo pairs) read from LDS; linear intervals carry c = d = 0.
#pragma once

#include "common.hpp"

namespace fsa {

__host__ __device__ inline int spl_locate(const double *x, int n, double r) {
  int lo = 0, hi = n; // first index with x > r
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (x[mid] > r) hi = mid;
    else lo = mid + 1;
  }
  int i = lo;
  if (i > 0) --i;
  if (i > n - 2) i = n - 2;
  return i;
}

// Hermite basis of value and first/second r-derivative w.r.t.
// (y_i, y_i+1, s_i, s_i+1).
__device__ inline void herm_basis64(double x0, double dx, double r, double v[4], double d[4]) {
  const double t = (r - x0) / dx;
  const double omt = 1.0 - t;
  v[0] = omt * omt * (1.0 + 2.0 * t);
  v[1] = t * t * (3.0 - 2.0 * t);
  v[2] = omt * omt * (r - x0);
  v[3] = t * t * (dx * (t - 1.0));
  d[0] = -6.0 * t * omt / dx;
  d[1] = 6.0 * t * omt / dx;
  d[2] = omt * (1.0 - 3.0 * t);
  d[3] = t * (3.0 * t - 2.0);
}

__device__ inline void herm_basis2_64(double x0, double dx, double r, double d2[4]) {
  const double t = (r - x0) / dx;
  d2[0] = -6.0 * (1.0 - 2.0 * t) / (dx * dx);
  d2[1] = 6.0 * (1.0 - 2.0 * t) / (dx * dx);
  d2[2] = (6.0 * t - 4.0) / dx;
  d2[3] = (6.0 * t - 2.0) / dx;
}

// cached-site evaluation (r already known to be inside [x0, xn])
__device__ inline void spl_site64(const TableDesc &T, const double *x, const double *y,
                                  const double *s, double r, double &v, double &d) {
  const int n = T.n;
  int i;
  bool lin;
  if (r <= x[0]) {
    i = 0;
    lin = true;
  } else if (r >= x[n - 1]) {
    i = n - 2;
    lin = true;
  } else {
    i = spl_locate(x, n, r);
    lin = T.linear != 0;
  }
  const double dx = x[i + 1] - x[i];
  if (lin) {
    const double inv_dx = 1.0 / dx;
    const double off = r - x[i];
    v = y[i] + off * inv_dx * (y[i + 1] - y[i]);
    d = (y[i + 1] - y[i]) * inv_dx;
    return;
  }
  double bv[4], bd[4];
  herm_basis64(x[i], dx, r, bv, bd);
  const double yi = y[i], yi1 = y[i + 1], si = s[i], si1 = s[i + 1];
  v = bv[0] * yi + bv[1] * yi1 + bv[2] * si + bv[3] * si1;
  d = bd[0] * yi + bd[1] * yi1 + bd[2] * si + bd[3] * si1;
}

// uncached evaluation with linear extrapolation (embedding F(rho)); also
// the host's evaluation of the embedding for the gauge. KV reads the knot
// values and slopes: y(k), s(k).
template <class KV>
__host__ __device__ inline void spl_uncached64_kv(const TableDesc &T, const double *x, const KV &kv,
                                                  double r, double &v, double &d) {
  const int n = T.n;
  if (r <= x[0]) {
    const double slope = (kv.y(1) - kv.y(0)) / (x[1] - x[0]);
    v = kv.y(0) + slope * (r - x[0]);
    d = slope;
  } else if (r >= x[n - 1]) {
    const double slope = (kv.y(n - 1) - kv.y(n - 2)) / (x[n - 1] - x[n - 2]);
    v = kv.y(n - 1) + slope * (r - x[n - 1]);
    d = slope;
  } else if (T.linear) {
    const int i = spl_locate(x, n, r);
    const double t = (r - x[i]) / (x[i + 1] - x[i]);
    v = kv.y(i) + t * (kv.y(i + 1) - kv.y(i));
    d = (kv.y(i + 1) - kv.y(i)) / (x[i + 1] - x[i]);
  } else {
    const int i = spl_locate(x, n, r);
    const double x0 = x[i], dx = x[i + 1] - x[i];
    const double t = (r - x0) / dx;
    const double omt = 1.0 - t;
    const double yi = kv.y(i), yi1 = kv.y(i + 1), si = kv.s(i), si1 = kv.s(i + 1);
    const double a = yi * (1.0 + 2.0 * t) + si * (r - x0);
    const double da = yi * 2.0 / dx + si;
    const double b = yi1 * (3.0 - 2.0 * t) + dx * si1 * (t - 1.0);
    const double db = -yi1 * 2.0 / dx + si1;
    v = omt * omt * a + t * t * b;
    d = -2.0 * omt / dx * a + omt * omt * da + 2.0 * t / dx * b + t * t * db;
  }
}

struct KnotArrays {
  const double *ya, *sa;
  __host__ __device__ double y(int k) const { return ya[k]; }
  __host__ __device__ double s(int k) const { return sa[k]; }
};

__host__ __device__ inline void spl_uncached64(const TableDesc &T, const double *x, const double *y,
                                      const double *s, double r, double &v, double &d) {
  spl_uncached64_kv(T, x, KnotArrays{y, s}, r, v, d);
}

// FP32 bulk evaluation. Interval k of table T holds the power form
// v = a + b t + c t^2 + d t^3, t = (r - x_k) / h, split as FP32 hi + lo
// pairs (I[2k], I[2k+1]) so that the coefficients are exact to ~1e-14 while
// every operation is FP32: rounding the coefficients themselves would bias all
// pairs of an interval alike, a systematic ~1e-7 relative force error.
__device__ inline int spl_interval32(const TableDesc &T, float r, float &t) {
  const float u = (r - T.x0f) * T.inv_hf;
  int k = int(floorf(u));
  k = k < 0 ? 0 : (k > T.n - 2 ? T.n - 2 : k);
  t = u - float(k);
  return k;
}

// interval from the FP64 distance: the FP32 offset t carries no rounding of
// x_0 or 1/h into the pair's position on the grid
__device__ inline int spl_interval64(const TableDesc &T, double r, double &t) {
  const double u = (r - T.x0) * T.inv_h;
  int k = int(floor(u));
  k = k < 0 ? 0 : (k > T.n - 2 ? T.n - 2 : k);
  t = u - double(k);
  return k;
}

__device__ inline float spl_dscale(const TableDesc &T, float dt) {
  return fmaf(dt, T.inv_hf, dt * T.inv_hlo);
}

__device__ inline void spl_eval32(const TableDesc &T, const float4 *I, int k, float t, float &v,
                                  float &d) {
  const float4 q = I[2 * (T.ivl + k)], e = I[2 * (T.ivl + k) + 1];
  v = (q.x + t * (q.y + t * (q.z + t * q.w))) + (e.x + t * (e.y + t * (e.z + t * e.w)));
  const float dt = (q.y + t * (2.0f * q.z + 3.0f * t * q.w)) + (e.y + t * (2.0f * e.z + 3.0f * t * e.w));
  d = spl_dscale(T, dt);
}

__device__ inline void spl_deriv32(const TableDesc &T, const float4 *I, int k, float t, float &d) {
  const float4 q = I[2 * (T.ivl + k)], e = I[2 * (T.ivl + k) + 1];
  const float dt = (q.y + t * (2.0f * q.z + 3.0f * t * q.w)) + (e.y + t * (2.0f * e.z + 3.0f * t * e.w));
  d = spl_dscale(T, dt);
}

// FP64 value on the same power form (energy and density terms)
__device__ inline double spl_value64(const TableDesc &T, const double *P, int k, double t) {
  const double *q = P + 4 * size_t(T.ivl + k);
  return q[0] + t * (q[1] + t * (q[2] + t * q[3]));
}

__host__ __device__ inline int ord_pair(int a, int b, int nt) {
  if (a > b) {
    const int t = a;
    a = b;
    b = t;
  }
  return a * nt - a * (a - 1) / 2 + (b - a);
}

} // namespace fsa
