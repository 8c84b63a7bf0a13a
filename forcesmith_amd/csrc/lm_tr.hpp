// Levenberg-Marquardt with the reference's stopping semantics, host side.
// EigenLMSolver (src/optimization/solver.cpp:67-85) runs Eigen 3.4's
// LevenbergMarquardt, the MINPACK lmder algorithm (More 1978, "The
// Levenberg-Marquardt algorithm: implementation and theory"): a trust region
// on the scaled step |D p| <= delta, the lmpar search for the damping
// parameter, the ratio of actual to predicted reduction with its step-bound
// update, and termination tests reported as Eigen's LevenbergMarquardtSpace
// status codes. MINPACK factors J (QR with column pivoting); this restatement
// takes every quantity it uses from the normal equations J^T J and J^T F,
// which give the same values in exact arithmetic:
//   R^T R = P^T J^T J P   (pivoted Cholesky, largest remaining pivot first,
//                          the order of the QR's largest remaining column),
//   R^T Q^T F = P^T J^T F, |R P^T p| = sqrt(p^T J^T J p),
//   |S^-T q| = sqrt(q^T (J^T J + par D^2)^-1 q)  (the Newton correction).
// Rank: ColPivHouseholderQR's rule |r_kk| > eps n max|r| becomes pivot >
// (eps n)^2 max pivot (a pivot at or below 0 also ends it). The formed normal
// equations carry rounding near n eps max pivot, so directions whose pivots
// lie between the two levels are kept, as the QR keeps them, but resolved
// only to that rounding (forming J^T J squares the condition number). Pure
// C++: the caller brings the residual norm and the normal equations (the GPU
// forms them).
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <utility>
#include <vector>

namespace fsa {
namespace lmtr {

constexpr double kEps = 2.220446049250313e-16;
constexpr double kDwarf = 2.2250738585072014e-308;
constexpr int kRunning = -1;

struct Opts {
  int max_iter, maxfev; // maxfev <= 0: max(max_iter, 1) * 100 + 100
  double xtol, ftol, gtol, factor;
};

// four lanes of doubles (GCC/Clang vector extension: AVX2 where the build
// targets it, two SSE2 halves otherwise)
typedef double v4d __attribute__((vector_size(32)));

inline v4d load4(const double *p) {
  v4d v;
  __builtin_memcpy(&v, p, sizeof(v));
  return v;
}

// dot products of a with four rows b0..b3 over [0, len): four-wide partial
// sums, then the scalar tail (a fixed order: the result does not depend on
// alignment). Measured against two partial sums per row and against eight
// rows at a time: this form is the fastest at D = 240 (the loads bound it)
inline void dot4(const double *a, const double *b0, const double *b1, const double *b2,
                 const double *b3, int len, double out[4]) {
  v4d s0 = {0, 0, 0, 0}, s1 = s0, s2 = s0, s3 = s0;
  int k = 0;
  for (; k + 4 <= len; k += 4) {
    const v4d va = load4(a + k);
    s0 += va * load4(b0 + k);
    s1 += va * load4(b1 + k);
    s2 += va * load4(b2 + k);
    s3 += va * load4(b3 + k);
  }
  double r[4] = {(s0[0] + s0[1]) + (s0[2] + s0[3]), (s1[0] + s1[1]) + (s1[2] + s1[3]),
                 (s2[0] + s2[1]) + (s2[2] + s2[3]), (s3[0] + s3[1]) + (s3[2] + s3[3])};
  for (; k < len; ++k) {
    r[0] += a[k] * b0[k];
    r[1] += a[k] * b1[k];
    r[2] += a[k] * b2[k];
    r[3] += a[k] * b3[k];
  }
  for (int q = 0; q < 4; ++q) out[q] = r[q];
}

inline double dot1(const double *a, const double *b, int len) {
  v4d s = {0, 0, 0, 0};
  int k = 0;
  for (; k + 4 <= len; k += 4) s += load4(a + k) * load4(b + k);
  double r = (s[0] + s[1]) + (s[2] + s[3]);
  for (; k < len; ++k) r += a[k] * b[k];
  return r;
}

// symmetric-pivoted Cholesky P^T M P = L L^T of a D x D row-major matrix
// (plus an optional diagonal), factor in the lower triangle of W (row p holds
// L's row of pivot position p); stops at the first pivot not above tol_rel
// times the first (rank = that position). Left-looking: column j of L is
// (M_(perm i, perm j) - L_i,<j . L_j,<j) / L_jj, so every step is a set of
// contiguous dot products over rows already computed (four rows at a time),
// and the remaining diagonal, kept current, picks the pivot. M is read in
// place (symmetric: row perm[j] holds column perm[j]).
struct PivChol {
  int n = 0, rank = 0;
  std::vector<double> W, col;
  std::vector<int> perm;

  void factor(const double *M, int n_, const double *add_diag, double tol_rel) {
    n = n_;
    W.resize(size_t(n) * n);
    perm.resize(n);
    col.resize(n); // the remaining diagonal of each position
    for (int k = 0; k < n; ++k) {
      perm[k] = k;
      col[k] = M[size_t(k) * n + k] + (add_diag ? add_diag[k] : 0.0);
    }
    double d0 = 0.0;
    rank = n;
    for (int j = 0; j < n; ++j) {
      int p = j;
      for (int q = j + 1; q < n; ++q)
        if (col[q] > col[p]) p = q;
      if (p != j) {
        // positions j and p trade places: their computed rows of L, their
        // indices and remaining diagonals
        double *rj = &W[size_t(j) * n], *rp = &W[size_t(p) * n];
        for (int k = 0; k < j; ++k) std::swap(rj[k], rp[k]);
        std::swap(perm[j], perm[p]);
        std::swap(col[j], col[p]);
      }
      const double d = col[j];
      if (j == 0) d0 = d;
      if (!(d > 0.0) || !(d > tol_rel * d0)) {
        rank = j;
        break;
      }
      const double l = std::sqrt(d);
      double *Lj = &W[size_t(j) * n];
      Lj[j] = l;
      const double *Mj = M + size_t(perm[j]) * n;
      int i = j + 1;
      for (; i + 4 <= n; i += 4) {
        double s[4];
        dot4(Lj, &W[size_t(i) * n], &W[size_t(i + 1) * n], &W[size_t(i + 2) * n],
             &W[size_t(i + 3) * n], j, s);
        for (int q = 0; q < 4; ++q) {
          const double v = (Mj[perm[i + q]] - s[q]) / l;
          W[size_t(i + q) * n + j] = v;
          col[i + q] -= v * v;
        }
      }
      for (; i < n; ++i) {
        const double v = (Mj[perm[i]] - dot1(Lj, &W[size_t(i) * n], j)) / l;
        W[size_t(i) * n + j] = v;
        col[i] -= v * v;
      }
    }
  }
  // y := L11^-1 y over the leading rank positions
  void lower_solve(double *y) const {
    for (int i = 0; i < rank; ++i) {
      const double *wi = &W[size_t(i) * n];
      double t = y[i];
      for (int k = 0; k < i; ++k) t -= wi[k] * y[k];
      y[i] = t / wi[i];
    }
  }
  // y := L11^-T y
  void upper_solve(double *y) const {
    for (int i = rank - 1; i >= 0; --i) {
      double t = y[i];
      for (int k = i + 1; k < rank; ++k) t -= W[size_t(k) * n + i] * y[k];
      y[i] = t / W[size_t(i) * n + i];
    }
  }
  // x = P [L11^-T L11^-1 (P^T b)_1 ; 0]
  void solve(const double *b, double *x) const {
    std::vector<double> z(n, 0.0);
    for (int p = 0; p < rank; ++p) z[p] = b[perm[p]];
    lower_solve(z.data());
    upper_solve(z.data());
    for (int p = 0; p < n; ++p) x[perm[p]] = p < rank ? z[p] : 0.0;
  }
  // |L11^-1 P^T q|^2 with q_l = dg_l^2 x_l / dxnorm
  double newton_norm2(const double *dg, const double *x, double dxnorm) const {
    std::vector<double> w(n, 0.0);
    for (int p = 0; p < rank; ++p) {
      const int l = perm[p];
      w[p] = dg[l] * dg[l] * x[l] / dxnorm;
    }
    lower_solve(w.data());
    double s = 0.0;
    for (int p = 0; p < rank; ++p) s += w[p] * w[p];
    return s;
  }
};

inline double scaled_norm(const double *dg, const double *v, int n) {
  double s = 0.0;
  for (int k = 0; k < n; ++k) s += (dg[k] * v[k]) * (dg[k] * v[k]);
  return std::sqrt(s);
}

// lmpar (More 1978, sections 5-7): par >= 0 and x with (A + par D^2) x = g
// such that | |D x| - delta | <= 0.1 delta, or par = 0 with the Gauss-Newton
// step (over the rank of the pivoted factor gn of A) when |D x| <= 1.1 delta.
// Returns the number of damped factorizations.
inline int lmpar(int n, const double *A, const double *g, const PivChol &gn, const double *dg,
                 double delta, double &par, double *x, PivChol &damp) {
  gn.solve(g, x);
  double dxnorm = scaled_norm(dg, x, n);
  double fp = dxnorm - delta;
  if (fp <= 0.1 * delta) {
    par = 0.0;
    return 0;
  }
  double parl = 0.0;
  if (gn.rank == n) parl = fp / delta / gn.newton_norm2(dg, x, dxnorm);
  double gnorm = 0.0;
  for (int k = 0; k < n; ++k) gnorm += (g[k] / dg[k]) * (g[k] / dg[k]);
  gnorm = std::sqrt(gnorm);
  double paru = gnorm / delta;
  if (paru == 0.0) paru = kDwarf / std::min(delta, 0.1);
  par = std::min(std::max(par, parl), paru);
  if (par == 0.0) par = gnorm / dxnorm;
  std::vector<double> pd(n);
  int it = 1;
  for (;; ++it) {
    if (par == 0.0) par = std::max(kDwarf, 0.001 * paru);
    for (int k = 0; k < n; ++k) pd[k] = par * dg[k] * dg[k];
    damp.factor(A, n, pd.data(), 0.0);
    damp.solve(g, x);
    dxnorm = scaled_norm(dg, x, n);
    const double prev = fp;
    fp = dxnorm - delta;
    if (std::fabs(fp) <= 0.1 * delta || (parl == 0.0 && fp <= prev && prev < 0.0) || it == 10) break;
    const double parc = fp / delta / damp.newton_norm2(dg, x, dxnorm);
    if (fp > 0.0) parl = std::max(parl, par);
    if (fp < 0.0) paru = std::min(paru, par);
    par = std::max(parl, par + parc);
  }
  return it;
}

// The driver of EigenLMSolver::minimize: minimizeInit, then at most max_iter
// calls of minimizeOneStep; returns the status (2 when still running,
// solver.cpp:84). cost(x) evaluates |F(x)|^2 (and leaves the caller's state
// at x); normal(A, g) fills J^T J and J^T F at the last evaluated accepted x.
// counts: residual evaluations, Jacobian evaluations, accepted steps;
// hist[call] = |F|^2 after each call (may be null).
template <class Cost, class Normal>
int minimize(int n, int64_t m, const Opts &o, double *x, Cost &&cost, Normal &&normal,
             int *counts, double *hist) {
  const int maxfev = o.maxfev > 0 ? o.maxfev : std::max(o.max_iter, 1) * 100 + 100;
  counts[0] = counts[1] = counts[2] = 0;
  if (n <= 0 || m < n || o.ftol < 0.0 || o.xtol < 0.0 || o.gtol < 0.0 || o.factor <= 0.0)
    return 0; // ImproperInputParameters
  std::vector<double> A(size_t(n) * n), g(n), cn(n), dg(n), step(n), xt(n);
  PivChol gn, damp;
  double fnorm = std::sqrt(cost(x));
  int nfev = 1, njev = 0, iter = 1, steps = 0;
  double par = 0.0, delta = 0.0, xnorm = 0.0;
  int st = kRunning;
  for (int call = 0; call < o.max_iter && st == kRunning; ++call) {
    normal(A.data(), g.data());
    ++njev;
    for (int k = 0; k < n; ++k) cn[k] = std::sqrt(std::max(A[size_t(k) * n + k], 0.0));
    gn.factor(A.data(), n, nullptr, (kEps * n) * (kEps * n));
    if (iter == 1) {
      for (int k = 0; k < n; ++k) dg[k] = cn[k] != 0.0 ? cn[k] : 1.0;
      xnorm = scaled_norm(dg.data(), x, n);
      delta = o.factor * xnorm;
      if (delta == 0.0) delta = o.factor;
    }
    // scaled gradient cosine max |(J^T F)_l| / (|F| |J_l|)
    double gnorm = 0.0;
    if (fnorm != 0.0)
      for (int k = 0; k < n; ++k)
        if (cn[k] != 0.0) gnorm = std::max(gnorm, std::fabs(g[k] / fnorm) / cn[k]);
    if (gnorm <= o.gtol) {
      st = 4; // CosinusTooSmall
      break;
    }
    for (int k = 0; k < n; ++k) dg[k] = std::max(dg[k], cn[k]);
    double ratio = 0.0;
    do {
      lmpar(n, A.data(), g.data(), gn, dg.data(), delta, par, step.data(), damp);
      for (int k = 0; k < n; ++k) {
        step[k] = -step[k];
        xt[k] = x[k] + step[k];
      }
      const double pnorm = scaled_norm(dg.data(), step.data(), n);
      if (iter == 1) delta = std::min(delta, pnorm);
      const double fnorm1 = std::sqrt(cost(xt.data()));
      ++nfev;
      double actred = -1.0;
      if (0.1 * fnorm1 < fnorm) actred = 1.0 - (fnorm1 / fnorm) * (fnorm1 / fnorm);
      double pAp = 0.0;
      for (int i = 0; i < n; ++i) {
        double s = 0.0;
        const double *ai = &A[size_t(i) * n];
        for (int k = 0; k < n; ++k) s += ai[k] * step[k];
        pAp += step[i] * s;
      }
      const double temp1 = std::sqrt(std::max(pAp, 0.0)) / fnorm;
      const double temp2 = std::sqrt(par) * pnorm / fnorm;
      const double prered = temp1 * temp1 + temp2 * temp2 / 0.5;
      const double dirder = -(temp1 * temp1 + temp2 * temp2);
      ratio = prered != 0.0 ? actred / prered : 0.0;
      if (ratio <= 0.25) {
        double temp = actred >= 0.0 ? 0.5 : 0.5 * dirder / (dirder + 0.5 * actred);
        if (0.1 * fnorm1 >= fnorm || temp < 0.1) temp = 0.1;
        delta = temp * std::min(delta, pnorm / 0.1);
        par /= temp;
      } else if (par == 0.0 || ratio >= 0.75) {
        delta = pnorm / 0.5;
        par = 0.5 * par;
      }
      if (ratio >= 1e-4) {
        std::copy(xt.begin(), xt.end(), x);
        xnorm = scaled_norm(dg.data(), x, n);
        fnorm = fnorm1;
        ++iter;
        ++steps;
      }
      const bool ftol_ok = std::fabs(actred) <= o.ftol && prered <= o.ftol && 0.5 * ratio <= 1.0;
      if (ftol_ok && delta <= o.xtol * xnorm) st = 3;
      else if (ftol_ok) st = 1;
      else if (delta <= o.xtol * xnorm) st = 2;
      else if (nfev >= maxfev) st = 5;
      else if (std::fabs(actred) <= kEps && prered <= kEps && 0.5 * ratio <= 1.0) st = 6;
      else if (delta <= kEps * xnorm) st = 7;
      else if (gnorm <= kEps) st = 8;
    } while (st == kRunning && ratio < 1e-4);
    if (hist) hist[call] = fnorm * fnorm;
  }
  counts[0] = nfev;
  counts[1] = njev;
  counts[2] = steps;
  return st == kRunning ? 2 : st;
}

} // namespace lmtr
} // namespace fsa
