// Host drivers above the kernels: Jacobian / normal equations, the
// Levenberg-Marquardt loop, population costs and differential evolution,
// kernel timing with HIP events on the context stream, and the launch-shape
// autotuner. Each step's arithmetic runs on the GPU; the host only carries
// scalars (costs, the LM damping, DE selection) between launches.
#include "common.hpp"

#include <chrono>
#include "lm_tr.hpp"

#include <algorithm>
#include <cmath>
#include <limits>
#include <vector>

namespace fsa {

void evaluate_for_solver(Context &c);
void rebuild_tiles(Context &c);
void ensure_ready(Context &c);
void launch_diag_scale(Context &c);
void launch_damped_solve(Context &c, double mu);
int de_run(Context &c, const fs_de_opts *o, const double *lo, const double *hi, double *x,
           double *best, double *hist, int *gens);

namespace {

void scatter_host(Context &c, const double *x) {
  int o = 0;
  if (c.kind == FS_SW) {
    for (size_t k = 0; k < c.h_sw.size(); ++k)
      if (!c.h_sw_fixed[k]) c.h_sw[k] = x[o++];
    for (size_t k = 0; k < c.h_lam.size(); ++k)
      if (!c.h_lam_fixed[k]) c.h_lam[k] = x[o++];
  } else {
    for (int k = 0; k < c.D; ++k) c.h_y[c.h_knot_of[k]] = x[k];
  }
  c.params_dirty = true;
  c.forces_valid = false;
}

void gather_host(const Context &c, double *x) {
  int o = 0;
  if (c.kind == FS_SW) {
    for (size_t k = 0; k < c.h_sw.size(); ++k)
      if (!c.h_sw_fixed[k]) x[o++] = c.h_sw[k];
    for (size_t k = 0; k < c.h_lam.size(); ++k)
      if (!c.h_lam_fixed[k]) x[o++] = c.h_lam[k];
  } else {
    for (int k = 0; k < c.D; ++k) x[k] = c.h_y[c.h_knot_of[k]];
  }
}

double cost_now(Context &c) {
  evaluate_for_solver(c);
  launch_cost(c, c.cost.p);
  double v = 0.0;
  FS_HIP(hipMemcpyAsync(&v, c.cost.p, sizeof(double), hipMemcpyDeviceToHost, c.stream));
  FS_HIP(hipStreamSynchronize(c.stream));
  return v;
}

} // namespace

int fs_jacobian_impl(Context &c, float *J) {
  evaluate_for_solver(c);
  launch_jacobian(c);
  if (J)
    FS_HIP(hipMemcpyAsync(J, c.J.p, sizeof(float) * size_t(c.R) * c.D, hipMemcpyDeviceToHost,
                          c.stream));
  FS_HIP(hipStreamSynchronize(c.stream));
  return FS_OK;
}

int fs_normal_impl(Context &c, double *JtJ, double *Jtr, double *cost) {
  evaluate_for_solver(c);
  launch_jacobian(c);
  launch_normal_equations(c);
  if (JtJ)
    FS_HIP(hipMemcpyAsync(JtJ, c.JtJ.p, sizeof(double) * size_t(c.D) * c.D, hipMemcpyDeviceToHost,
                          c.stream));
  if (Jtr)
    FS_HIP(hipMemcpyAsync(Jtr, c.Jtr.p, sizeof(double) * c.D, hipMemcpyDeviceToHost, c.stream));
  if (cost) {
    launch_cost(c, c.cost.p);
    FS_HIP(hipMemcpyAsync(cost, c.cost.p, sizeof(double), hipMemcpyDeviceToHost, c.stream));
  }
  FS_HIP(hipStreamSynchronize(c.stream));
  return FS_OK;
}

// Levenberg-Marquardt on the normal equations: one Jacobian per outer
// iteration, inner damping loop until the gain ratio exceeds 1e-4 (as
// MINPACK's lmpar loop inside Eigen's minimizeOneStep, solver.cpp:67-85),
// Marquardt diagonal scaling, Nielsen damping update. The same algorithm is
// restated by the CPU oracle.
int fs_lm_impl(Context &c, double *x, int max_iter, double *hist, int *accepted) {
  const int D = c.D;
  *accepted = 0;
  if (D == 0) return FS_OK;
  std::vector<double> xn(D), dx(D);
  scatter_host(c, x);
  double cost = cost_now(c);
  launch_jacobian(c);
  launch_normal_equations(c);
  launch_diag_scale(c);
  double scal[3];
  FS_HIP(hipMemcpyAsync(scal, c.lm_scal.p, sizeof(double), hipMemcpyDeviceToHost, c.stream));
  FS_HIP(hipStreamSynchronize(c.stream));
  double mu = 1e-3 * scal[0], nu = 2.0;
  for (int it = 0; it < max_iter; ++it) {
    bool ok = false;
    for (int inner = 0; inner < 16; ++inner) {
      launch_damped_solve(c, mu);
      FS_HIP(hipMemcpyAsync(scal, c.lm_scal.p, sizeof(double) * 3, hipMemcpyDeviceToHost, c.stream));
      FS_HIP(hipMemcpyAsync(dx.data(), c.lm_dx.p, sizeof(double) * D, hipMemcpyDeviceToHost,
                            c.stream));
      FS_HIP(hipStreamSynchronize(c.stream));
      if (scal[1] != 0.0) {
        mu *= nu;
        nu *= 2.0;
        continue;
      }
      for (int k = 0; k < D; ++k) xn[k] = x[k] + dx[k];
      scatter_host(c, xn.data());
      const double cn = cost_now(c);
      const double pred = scal[2];
      const double rho = (cost - cn) / pred;
      if (std::isfinite(cn) && pred > 0.0 && rho > 1e-4) {
        std::copy(xn.begin(), xn.end(), x);
        cost = cn;
        const double t = 2.0 * rho - 1.0;
        const double f = 1.0 - t * t * t;
        mu *= f > 1.0 / 3.0 ? f : 1.0 / 3.0;
        nu = 2.0;
        ok = true;
        break;
      }
      mu *= nu;
      nu *= 2.0;
    }
    if (hist) hist[it] = cost;
    if (!ok) {
      for (int r = it + 1; r < max_iter && hist; ++r) hist[r] = cost;
      break;
    }
    ++*accepted;
    if (it + 1 < max_iter) {
      // the residual at x was just evaluated, the per-atom state matches x
      launch_jacobian(c);
      launch_normal_equations(c);
      launch_diag_scale(c);
    }
  }
  scatter_host(c, x);
  evaluate_for_solver(c);
  FS_HIP(hipStreamSynchronize(c.stream));
  return FS_OK;
}

// Levenberg-Marquardt with the reference's stopping semantics (lm_tr.hpp):
// the residual norm and the FP64 normal equations come from the GPU, the
// D x D trust-region algebra runs on the host
int fs_lm_solve_impl(Context &c, const fs_lm_opts &o, double *x, int *status, int *counts,
                     double *hist) {
  const lmtr::Opts opts{o.max_iter, o.maxfev, o.xtol, o.ftol, o.gtol, o.factor};
  const int n = c.D;
  // wall time of the device calls (each ends in a stream synchronisation)
  // against the whole solve: the rest is the host's D x D algebra
  using clk = std::chrono::steady_clock;
  double dev_s = 0.0;
  auto timed = [&](auto &&f) {
    const auto t0 = clk::now();
    f();
    dev_s += std::chrono::duration<double>(clk::now() - t0).count();
  };
  auto cost = [&](const double *xv) {
    double v = 0.0;
    timed([&] {
      scatter_host(c, xv);
      v = cost_now(c);
    });
    return v;
  };
  auto normal = [&](double *A, double *g) {
    timed([&] {
      launch_jacobian(c);
      launch_normal_equations(c);
      FS_HIP(hipMemcpyAsync(A, c.JtJ.p, sizeof(double) * size_t(n) * n, hipMemcpyDeviceToHost,
                            c.stream));
      FS_HIP(hipMemcpyAsync(g, c.Jtr.p, sizeof(double) * n, hipMemcpyDeviceToHost, c.stream));
      FS_HIP(hipStreamSynchronize(c.stream));
    });
  };
  // lmtr::minimize calls normal() right after the cost of the accepted x
  // (the initial point or the trial it just accepted), so the device
  // residual and per-atom state are at x
  const auto t_all = clk::now();
  *status = lmtr::minimize(n, c.R, opts, x, cost, normal, counts, hist);
  const double all_s = std::chrono::duration<double>(clk::now() - t_all).count();
  c.lm_timing[0] = 1e3 * all_s;
  c.lm_timing[1] = 1e3 * dev_s;
  c.lm_timing[2] = 1e3 * (all_s - dev_s);
  // leave the device state at the returned parameters
  scatter_host(c, x);
  evaluate_for_solver(c);
  FS_HIP(hipStreamSynchronize(c.stream));
  return FS_OK;
}

// NormalEquationsSolver::minimize (src/optimization/solver.cpp:242-266): one
// ridge-regularised Gauss-Newton step x -= (J^T J + lambda I)^-1 J^T F(x),
// lambda = ridge if > 0, else 1e-8 max(mean diag(J^T J), 1). J^T J and J^T F
// are formed on the GPU; the D x D solve (Cholesky of the SPD matrix, which
// the reference's LDLT equals in exact arithmetic) runs on the host.
int fs_ne_solve_impl(Context &c, double ridge, double *x) {
  const int n = c.D;
  if (n == 0 || c.R == 0) return FS_OK;
  scatter_host(c, x);
  evaluate_for_solver(c);
  launch_jacobian(c);
  launch_normal_equations(c);
  std::vector<double> A(size_t(n) * n), g(n), dx(n);
  FS_HIP(hipMemcpyAsync(A.data(), c.JtJ.p, sizeof(double) * A.size(), hipMemcpyDeviceToHost,
                        c.stream));
  FS_HIP(hipMemcpyAsync(g.data(), c.Jtr.p, sizeof(double) * n, hipMemcpyDeviceToHost, c.stream));
  FS_HIP(hipStreamSynchronize(c.stream));
  double lambda = ridge;
  if (lambda <= 0.0) {
    double mean = 0.0;
    for (int k = 0; k < n; ++k) mean += A[size_t(k) * n + k];
    lambda = 1e-8 * std::max(mean / n, 1.0);
  }
  std::vector<double> add(n, lambda);
  lmtr::PivChol ch;
  ch.factor(A.data(), n, add.data(), 0.0);
  FS_REQUIRE(ch.rank == n, FS_ERR_NUMERIC, "normal equations are not positive definite");
  ch.solve(g.data(), dx.data());
  for (int k = 0; k < n; ++k) x[k] -= dx[k];
  scatter_host(c, x);
  evaluate_for_solver(c);
  FS_HIP(hipStreamSynchronize(c.stream));
  return FS_OK;
}

// one damped Gauss-Newton step at the current parameters: dx = -(J^T J +
// lambda I)^-1 J^T F, parameters unchanged
int fs_lm_step_impl(Context &c, double lambda, double *dx) {
  const int n = c.D;
  if (n == 0) return FS_OK;
  evaluate_for_solver(c);
  launch_jacobian(c);
  launch_normal_equations(c);
  std::vector<double> A(size_t(n) * n), g(n);
  FS_HIP(hipMemcpyAsync(A.data(), c.JtJ.p, sizeof(double) * A.size(), hipMemcpyDeviceToHost,
                        c.stream));
  FS_HIP(hipMemcpyAsync(g.data(), c.Jtr.p, sizeof(double) * n, hipMemcpyDeviceToHost, c.stream));
  FS_HIP(hipStreamSynchronize(c.stream));
  std::vector<double> add(n, lambda);
  lmtr::PivChol ch;
  ch.factor(A.data(), n, add.data(), 0.0);
  FS_REQUIRE(ch.rank == n, FS_ERR_NUMERIC, "damped normal equations are not positive definite");
  ch.solve(g.data(), dx);
  for (int k = 0; k < n; ++k) dx[k] = -dx[k];
  return FS_OK;
}

int fs_population_impl(Context &c, const double *X, int np, double *cost) {
  const int D = c.D;
  ensure_ready(c);
  if (population_cost_batched(c, X, np, cost)) return FS_OK;
  std::vector<double> x0(D);
  gather_host(c, x0.data());
  for (int m = 0; m < np; ++m) {
    scatter_host(c, X + size_t(m) * D);
    cost[m] = cost_now(c);
  }
  scatter_host(c, x0.data());
  return FS_OK;
}

// BoostDESolver's box (src/optimization/solver.cpp:118-125): a finite bound
// is kept, an infinite one becomes v -+ max(2 |v|, 5e-4) around the initial
// value v; NP <= 0 means NP_factor * D = 15 D (solver.hpp:78-81)
int fs_de_impl(Context &c, const fs_de_opts *o, const double *lo, const double *hi, double *x,
               double *best, double *hist, int *gens) {
  const int D = c.D;
  std::vector<double> lb(D), ub(D);
  for (int k = 0; k < D; ++k) {
    const double v = x[k], half = std::max(2.0 * std::fabs(v), 5e-4);
    lb[k] = std::isfinite(lo[k]) ? lo[k] : v - half;
    ub[k] = std::isfinite(hi[k]) ? hi[k] : v + half;
    FS_REQUIRE(lb[k] <= ub[k], FS_ERR_ARG, "DE box has lower > upper");
  }
  fs_de_opts oo = *o;
  if (oo.np <= 0) oo.np = 15 * D;
  return de_run(c, &oo, lb.data(), ub.data(), x, best, hist, gens);
}

// what: 0 = residual (density + force + finalize), 1 = Jacobian (all its
// kernels), 2 = normal equations, 3 = residual + Jacobian, 4 = the
// density/embedding pass of the Jacobian (ADP: with the dipole/quadrupole passes),
// 5 = the pair-column pass, 6 = the ADP quadrupole-column force pass (events
// around those launches only)
double time_step_impl(Context &c, int what, int reps) {
  FS_REQUIRE(what >= 0 && what <= 6, FS_ERR_ARG, "unknown timing stage");
  ensure_ready(c);
  evaluate_for_solver(c);
  if (what >= 4) {
    c.time_which = what - 4;
    if (!c.jev0) FS_HIP(hipEventCreate(&c.jev0));
    if (!c.jev1) FS_HIP(hipEventCreate(&c.jev1));
    launch_jacobian(c); // warm-up
    c.time_jac_main = true;
    c.jac_main_ms = 0.0;
    c.jac_main_n = 0;
    try {
      for (int r = 0; r < reps; ++r) launch_jacobian(c);
    } catch (...) {
      c.time_jac_main = false;
      throw;
    }
    c.time_jac_main = false;
    return c.jac_main_n ? c.jac_main_ms / c.jac_main_n : 0.0;
  }
  hipEvent_t e0, e1;
  FS_HIP(hipEventCreate(&e0));
  FS_HIP(hipEventCreate(&e1));
  auto body = [&]() {
    if (what == 0) launch_eval(c, true);
    else if (what == 1) launch_jacobian(c);
    else if (what == 2) launch_normal_equations(c);
    else {
      launch_eval(c, true);
      launch_jacobian(c);
    }
  };
  body(); // warm-up
  FS_HIP(hipEventRecord(e0, c.stream));
  for (int r = 0; r < reps; ++r) body();
  FS_HIP(hipEventRecord(e1, c.stream));
  FS_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  FS_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  c.forces_valid = true;
  return double(ms) / reps;
}

int autotune_impl(Context &c) {
  ensure_ready(c);
  const int cand[] = {4, 8, 16, 32, 64};
  int best = c.lanes_per_atom;
  double best_ms = std::numeric_limits<double>::infinity();
  for (int lpa : cand) {
    c.lanes_per_atom = lpa;
    rebuild_tiles(c);
    c.forces_valid = false;
    const double ms = time_step_impl(c, 0, 5);
    if (ms < best_ms) {
      best_ms = ms;
      best = lpa;
    }
  }
  c.lanes_per_atom = best;
  rebuild_tiles(c);
  c.forces_valid = false;
  // persistent-grid size of the residual kernels (blocks per CU; the grid
  // strides over the tiles, so the result does not depend on it)
  {
    int best_b = c.blocks_per_cu;
    double best_b_ms = std::numeric_limits<double>::infinity();
    for (int bpc : {4, 8, 16, 32}) {
      c.blocks_per_cu = bpc;
      c.forces_valid = false;
      const double ms = time_step_impl(c, 0, 5);
      if (ms < best_b_ms) {
        best_b_ms = ms;
        best_b = bpc;
      }
    }
    c.blocks_per_cu = best_b;
    c.forces_valid = false;
  }
  // Jacobian tile: powers of two from 4 to 64 atoms, timed over the whole
  // Jacobian; candidates that would cut the dataset into more than 64 tiles
  // per CU are skipped (their per-tile partials and row reductions grow with
  // the tile count, and the chip is full long before that); a tile fixed by
  // fs_set_jacobian_tile is kept
  if (!c.jac_tile_fixed && c.N > 0 && c.D > 0) {
    const int keep = c.jac_tile;
    int best_jt = -1;
    double best_jt_ms = std::numeric_limits<double>::infinity();
    for (int jt = 4; jt <= 64; jt *= 2) {
      if (jt < 64 && c.N / jt > int64_t(64) * c.cu_count) continue;
      c.jac_tile = jt;
      c.njt = 0;
      const double ms = time_step_impl(c, 1, 5);
      if (ms < best_jt_ms) {
        best_jt_ms = ms;
        best_jt = jt;
      }
    }
    c.jac_tile = best_jt > 0 ? best_jt : keep;
    c.njt = 0;
  }
  // persistent-grid size of the Jacobian passes
  if (c.N > 0 && c.D > 0) {
    int best_b = c.jac_blocks_per_cu;
    double best_b_ms = std::numeric_limits<double>::infinity();
    for (int bpc : {4, 8, 16, 32}) {
      c.jac_blocks_per_cu = bpc;
      const double ms = time_step_impl(c, 1, 5);
      if (ms < best_b_ms) {
        best_b_ms = ms;
        best_b = bpc;
      }
    }
    c.jac_blocks_per_cu = best_b;
  }
  return FS_OK;
}

} // namespace fsa
