// C wrapper over forcesmith_amd/csrc/lm_tr.hpp for CPU tests: the host
// trust-region algebra of fs_lm_solve driven by callbacks (FP64 residual norm
// and normal equations supplied by the test), and lmpar alone.
#include "lm_tr.hpp"

extern "C" {
typedef double (*lmh_cost)(const double *x);
typedef void (*lmh_normal)(double *A, double *g);

int lmh_minimize(int n, int64_t m, int max_iter, int maxfev, double xtol, double ftol,
                 double gtol, double factor, double *x, lmh_cost cost, lmh_normal normal,
                 int32_t *counts, double *hist) {
  const fsa::lmtr::Opts o{max_iter, maxfev, xtol, ftol, gtol, factor};
  int cnt[3];
  const int st = fsa::lmtr::minimize(
      n, m, o, x, [&](const double *v) { return cost(v); },
      [&](double *A, double *g) { normal(A, g); }, cnt, hist);
  for (int k = 0; k < 3; ++k) counts[k] = cnt[k];
  return st;
}

int lmh_lmpar(int n, const double *A, const double *g, const double *dg, double delta,
              double *par, double *x) {
  fsa::lmtr::PivChol gn, damp;
  gn.factor(A, n, nullptr, (fsa::lmtr::kEps * n) * (fsa::lmtr::kEps * n));
  return fsa::lmtr::lmpar(n, A, g, gn, dg, delta, *par, x, damp);
}
}
