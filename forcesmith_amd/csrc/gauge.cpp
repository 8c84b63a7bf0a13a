// EAM gauge transformation: rescale_eam and its steps (src/core/rescale.cpp)
// on the densities the residual kernels compute, held as per-type scalars
// that the kernels apply.
//
// The reference rescales by wrapping its potentials
// (include/forcesmith/core/scale_potentials.hpp): the density by
// ScaledOutputPotential (A g), the embedding by ScaledArgPotential (F(rho / A),
// span A [lo, hi]) under LinearAdjustedPotential (- S rho - F0), each pair by
// CompensatedPairPotential (phi + S_i A g_j + S_j A g_i). Those compose
// into one scale A and per-type S_t, F0_t: a later rho-axis rescale by a
// maps (A, S) to (A a, S / a), a later slope shift by s adds s to S, a later
// intercept adds to F0. The knot values stay the fitted parameters, as
// through the wrappers' delegated gather/scatter.
//
// The kernels apply the gauge per atom (eval.hip gauge_clamp, the Jacobian's
// clamped-atom energy term): with pair and density tables on one span the
// pair compensation sums to S_t rho per atom and cancels the slope in the
// forces, so the only effects are the clamp and penalty in the gauge's
// density units, the intercept, and S_t (rho - clamped rho) at clamped atoms.
// A gauge whose slope would compensate a pair table of another span than the
// density table it pairs with is refused (FS_ERR_UNSUPPORTED).
//
// Deviation kept on purpose: slopes below 1e-14 are dropped from both the
// embedding and the pairs (the reference drops them from the embedding only;
// its pair term with such a slope is below 1e-14 g(r)).
#include "common.hpp"
#include "spline.hpp"

#include <cmath>
#include <limits>
#include <string>
#include <vector>

namespace fsa {

void evaluate_for_solver(Context &c);
void params_to_device(Context &c);

void upload_gauge(Context &c) {
  c.gauge.alloc(kGaugeLen);
  double *img = c.h_gauge_img;
  img[0] = c.gA;
  for (int t = 0; t < kMaxTypes; ++t) {
    img[1 + t] = t < c.ntypes ? c.h_gS[t] : 0.0;
    img[1 + kMaxTypes + t] = t < c.ntypes ? c.h_gF0[t] : 0.0;
  }
  FS_HIP(hipMemcpyAsync(c.gauge.p, img, sizeof(double) * kGaugeLen, hipMemcpyHostToDevice,
                        c.stream));
}

void reset_gauge(Context &c) {
  c.gA = 1.0;
  for (int t = 0; t < kMaxTypes; ++t) c.h_gS[t] = c.h_gF0[t] = 0.0;
  upload_gauge(c);
}

namespace {

int tab_pair(const Context &c, int a, int b) { return ord_pair(a, b, c.ntypes); }
int tab_dens(const Context &c, int t) { return c.npair + t; }
int tab_emb(const Context &c, int t) { return c.npair + c.ntypes + t; }

void require_eam(const Context &c) {
  c.fail_if_no_model();
  FS_REQUIRE(c.kind == FS_EAM, FS_ERR_UNSUPPORTED, "the gauge transformation applies to EAM models");
}

// a slope S_ti compensates pair (ti, tj) with density tj: both on one span
void check_spans(const Context &c, const double *S) {
  for (int ti = 0; ti < c.ntypes; ++ti) {
    if (S[ti] == 0.0) continue;
    for (int tj = 0; tj < c.ntypes; ++tj) {
      const TableDesc &P = c.h_tabs[tab_pair(c, ti, tj)], &G = c.h_tabs[tab_dens(c, tj)];
      FS_REQUIRE(P.x0 == G.x0 && P.xn == G.xn, FS_ERR_UNSUPPORTED,
                 "gauge slope of type " + std::to_string(ti) + " compensates pair table (" +
                     std::to_string(ti) + ", " + std::to_string(tj) +
                     ") on another span than density table " + std::to_string(tj));
    }
  }
}

void set_gauge(Context &c, double A, const double *S, const double *F0) {
  FS_REQUIRE(std::isfinite(A) && A != 0.0, FS_ERR_ARG, "gauge scale must be finite and nonzero");
  for (int t = 0; t < c.ntypes; ++t)
    FS_REQUIRE(std::isfinite(S[t]) && std::isfinite(F0[t]), FS_ERR_ARG,
               "gauge slopes and intercepts must be finite");
  check_spans(c, S);
  c.gA = A;
  for (int t = 0; t < c.ntypes; ++t) {
    c.h_gS[t] = S[t];
    c.h_gF0[t] = F0[t];
  }
  upload_gauge(c);
  c.forces_valid = false;
}

// the current embedding tables (knots from the device: the fitted values)
struct EmbTables {
  std::vector<double> y, s;
};

EmbTables embedding_tables(Context &c) {
  params_to_device(c); // knots and slopes of the current parameters
  EmbTables e;
  e.y.resize(c.nknots);
  e.s.resize(c.nknots);
  FS_HIP(hipMemcpyAsync(e.y.data(), c.tab_y.p, sizeof(double) * c.nknots, hipMemcpyDeviceToHost,
                        c.stream));
  FS_HIP(hipMemcpyAsync(e.s.data(), c.tab_s.p, sizeof(double) * c.nknots, hipMemcpyDeviceToHost,
                        c.stream));
  FS_HIP(hipStreamSynchronize(c.stream));
  return e;
}

// the gauged embedding of type t at rho: F(rho / A) - S rho - F0 and
// F'(rho / A) / A - S (scale_potentials.hpp:34-79)
void emb_gauged(const Context &c, const EmbTables &e, int t, double rho, double &F, double &dF) {
  const TableDesc &T = c.h_tabs[tab_emb(c, t)];
  double v, d;
  if (c.gA != 1.0) {
    spl_uncached64(T, c.h_x.data() + T.off, e.y.data() + T.off, e.s.data() + T.off, rho / c.gA, v,
                   d);
    d = d / c.gA;
  } else {
    spl_uncached64(T, c.h_x.data() + T.off, e.y.data() + T.off, e.s.data() + T.off, rho, v, d);
  }
  F = v - c.h_gS[t] * rho - c.h_gF0[t];
  dF = d - c.h_gS[t];
}

} // namespace

// compute_rho_ref (rescale.cpp:11-33) and the extremes rescale_rho_axis takes
// (:39-52): every configuration evaluated, the clamped densities per type in
// the gauge's units (A times the embedding argument the kernels keep)
void gauge_rho_stats(Context &c, double *rmin, double *rmax, double *rmean, int64_t *count) {
  require_eam(c);
  c.fail_if_no_dataset();
  evaluate_for_solver(c);
  std::vector<double> u(size_t(c.N));
  if (c.N > 0)
    FS_HIP(hipMemcpyAsync(u.data(), c.rho.p, sizeof(double) * c.N, hipMemcpyDeviceToHost, c.stream));
  FS_HIP(hipStreamSynchronize(c.stream));
  for (int t = 0; t < c.ntypes; ++t) {
    double lo = std::numeric_limits<double>::infinity(), hi = -lo, sum = 0.0;
    int64_t n = 0;
    for (int64_t i = 0; i < c.N; ++i) {
      if (c.h_types[i] != t) continue;
      const double r = c.gA * u[i];
      sum += r;
      lo = std::min(lo, r);
      hi = std::max(hi, r);
      ++n;
    }
    if (rmin) rmin[t] = lo;
    if (rmax) rmax[t] = hi;
    if (rmean) rmean[t] = n > 0 ? sum / double(n) : 0.0;
    if (count) count[t] = n;
  }
}

// rescale_rho_axis (rescale.cpp:35-101): the scale that takes the dominant
// type's extreme density, padded by 0.003 of its embedding span, to the end
// of that span; 1 without data, for a non-finite or vanishing scale, or for
// |a| within [0.95, 1.05] while every type's densities lie inside its span
double gauge_rescale_rho_axis(Context &c) {
  require_eam(c);
  const int n = c.ntypes;
  double mx[kMaxTypes], mn[kMaxTypes];
  gauge_rho_stats(c, mn, mx, nullptr, nullptr);
  auto span = [&](int t, double &lo, double &hi) {
    const TableDesc &T = c.h_tabs[tab_emb(c, t)];
    lo = c.gA * T.x0;
    hi = c.gA * T.xn;
  };
  int dom = 0;
  bool sign_pos = true;
  double best = -1.0;
  for (int t = 0; t < n; ++t) {
    if (!std::isfinite(mx[t])) continue;
    const double ext = std::max(std::fabs(mx[t]), std::fabs(mn[t]));
    if (ext > best) {
      best = ext;
      dom = t;
      sign_pos = mx[t] >= -mn[t];
    }
  }
  if (best < 0.0) return 1.0;
  double elo, ehi;
  span(dom, elo, ehi);
  const double pad = 0.003 * (ehi - elo);
  const double upper = sign_pos ? ehi : elo;
  const double right = sign_pos ? mx[dom] + pad : mn[dom] - pad;
  if (std::fabs(right) < 1e-30) return 1.0;
  const double a = upper / right;
  bool violation = false;
  for (int t = 0; t < n; ++t) {
    if (!std::isfinite(mx[t])) continue;
    double lo, hi;
    span(t, lo, hi);
    if (mn[t] < lo || mx[t] > hi) violation = true;
  }
  if (!std::isfinite(a) || std::fabs(a) < 1e-30 ||
      (!violation && std::fabs(a) >= 0.95 && std::fabs(a) <= 1.05))
    return 1.0;
  double S[kMaxTypes];
  for (int t = 0; t < n; ++t) S[t] = c.h_gS[t] / a;
  set_gauge(c, c.gA * a, S, c.h_gF0);
  return a;
}

// embed_shift (rescale.cpp:103-136): the slope F'(rho_ref_t) of the current
// embedding, for types with rho_ref_t > 0, subtracted from the embedding
// and compensated in the pairs
void gauge_embed_shift(Context &c, const double *rho_ref) {
  require_eam(c);
  FS_REQUIRE(rho_ref, FS_ERR_ARG, "null reference densities");
  const EmbTables e = embedding_tables(c);
  double S[kMaxTypes];
  for (int t = 0; t < c.ntypes; ++t) {
    S[t] = c.h_gS[t];
    FS_REQUIRE(std::isfinite(rho_ref[t]), FS_ERR_ARG, "reference densities must be finite");
    if (!(rho_ref[t] > 0.0)) continue;
    double F, dF;
    emb_gauged(c, e, t, rho_ref[t], F, dF);
    if (std::fabs(dF) >= 1e-14) S[t] += dF;
  }
  set_gauge(c, c.gA, S, c.h_gF0);
}

// rescale_eam (rescale.cpp:138-158): rho axis, reference densities, slope
// shift, then the intercept F0_t = F(rho_ref_t) of the shifted embedding
// (types with rho_ref_t > 0 and |F0_t| >= 1e-14)
double gauge_rescale_eam(Context &c, double *rho_ref_out) {
  require_eam(c);
  const double a = gauge_rescale_rho_axis(c);
  double ref[kMaxTypes];
  gauge_rho_stats(c, nullptr, nullptr, ref, nullptr);
  gauge_embed_shift(c, ref);
  const EmbTables e = embedding_tables(c);
  double F0[kMaxTypes];
  for (int t = 0; t < c.ntypes; ++t) {
    F0[t] = c.h_gF0[t];
    if (!(ref[t] > 0.0)) continue;
    double F, dF;
    emb_gauged(c, e, t, ref[t], F, dF);
    if (std::fabs(F) >= 1e-14) F0[t] += F;
  }
  set_gauge(c, c.gA, c.h_gS, F0);
  if (rho_ref_out)
    for (int t = 0; t < c.ntypes; ++t) rho_ref_out[t] = ref[t];
  return a;
}

void gauge_set(Context &c, double A, const double *S, const double *F0) {
  require_eam(c);
  double s[kMaxTypes], f[kMaxTypes];
  for (int t = 0; t < c.ntypes; ++t) {
    s[t] = S ? S[t] : 0.0;
    f[t] = F0 ? F0[t] : 0.0;
  }
  set_gauge(c, A, s, f);
}

} // namespace fsa
