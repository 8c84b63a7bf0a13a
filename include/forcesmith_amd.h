/*
 * forcesmith_amd: MI355X-native force-matching evaluator, C ABI.
 *
 * The drop-in boundary for the reference's force-calculator customisation
 * point (docs/extension.md:252-321; erased surface CForceCalcTE,
 * include/forcesmith/force/force_calculator.hpp:62-81). A reference-side
 * adapter type holds an fs_ctx* (shared, ref-counted: the reference deep-copies
 * calculators per FD column and per thread, forcesmith_functor.cpp:73,146,178)
 * and forwards each erased method to the entry point named below.
 *
 * Conventions
 *  - Every function returns 0 on success or a negative FS_ERR_* code; the
 *    message is available from fs_last_error(ctx). No exception crosses the
 *    ABI. Output buffers are caller-owned host memory unless stated; device
 *    memory is owned by the context.
 *  - Calls on one context are serialised by an internal mutex, so the
 *    reference's concurrent per-configuration calls
 *    (forcesmith_functor.cpp:43-51,127-131) are safe. Each context owns one
 *    HIP stream on its device.
 *  - cell[9c + 3k + a]: component a of lattice vector k of configuration c
 *    (the reference's "X","Y","Z", stored as box columns,
 *    src/io/config_reader.cpp:58-66). Stress in and out of the residual is in
 *    the order xx, yy, zz, xy, xz, yz (include/forcesmith/core/voigt.hpp:9-13);
 *    full 3x3 stresses are row-major.
 */
#ifndef FORCESMITH_AMD_H
#define FORCESMITH_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FS_ABI_VERSION 1

enum fs_status {
  FS_OK = 0,
  FS_ERR_ARG = -1,         /* invalid argument or shape */
  FS_ERR_STATE = -2,       /* call out of order (no dataset / model / prepare) */
  FS_ERR_HIP = -3,         /* HIP runtime error */
  FS_ERR_NOMEM = -4,       /* device allocation failed */
  FS_ERR_UNSUPPORTED = -5, /* feature not available for this model kind */
  FS_ERR_NUMERIC = -6      /* non-finite result or failed factorisation */
};

enum fs_model_kind { FS_EAM = 0, FS_ADP = 1, FS_SW = 2 };

/* FS_MIXED: FP32 pair/density/spline arithmetic, FP64 per-configuration
 * energy/stress accumulation, FP64 normal equations and LM solve.
 * FS_FP64: every step in FP64 (parity mode). */
enum fs_precision { FS_MIXED = 0, FS_FP64 = 1 };

typedef struct fs_opts {
  int32_t device;    /* HIP device ordinal */
  int32_t precision; /* enum fs_precision */
  int32_t reserved[6];
} fs_opts;

typedef struct fs_ctx fs_ctx;

int fs_abi_version(void);

/* Context lifetime. fs_ctx_retain/fs_ctx_release implement the shared handle
 * the reference's value-semantics calculator copies need
 * (force_calculator.hpp:84-104: copy constructor/assignment). */
int fs_ctx_create(const fs_opts *opts, fs_ctx **out);
int fs_ctx_retain(fs_ctx *ctx);
int fs_ctx_release(fs_ctx *ctx);
const char *fs_last_error(const fs_ctx *ctx);

/* Dataset: replaces the std::span<Configuration> handed to prepare()
 * (CForceCalcTE has_prepare, force_calculator.hpp:74) and the per-config
 * reference data read by write_config_residuals (residual_layout.hpp:37-54). */
int fs_upload_dataset(fs_ctx *ctx, int32_t n_configs, const int64_t *atom_off,
                      const int32_t *types, const double *pos,
                      const double *ref_force, const double *cell,
                      const double *ref_energy, const double *ref_stress);

/* Tabulated EAM/ADP model: tables in the reference declaration order
 * (eam_force.hpp:24-32, adp_force.hpp:37-51), knot grid of make_tabulated
 * (src/io/potential_reader.cpp:162-191). Replaces constructing the
 * EAMForceCalculator / ADPForceCalculator value. */
int fs_set_tabulated_model(fs_ctx *ctx, int32_t kind, int32_t ntypes,
                           int32_t ntab, const int32_t *n_knots,
                           const double *rmin, const double *rmax,
                           const double *y, const uint8_t *fixed);
/* Stillinger-Weber: pair slots of (A,B,p,q,delta,a1,gamma,a2) and lambda[T*P]
 * (stiweb_force.hpp:19-62). */
int fs_set_sw_model(fs_ctx *ctx, int32_t ntypes, const double *pair_params,
                    const uint8_t *pair_fixed, const double *lambda,
                    const uint8_t *lambda_fixed);
/* OptimizerOptions energy_weight / stress_weight (optimizer.hpp:11-15). */
int fs_set_weights(fs_ctx *ctx, double energy_weight, double stress_weight);
/* OptimizerOptions smooth_weight (optimizer.hpp:14): when > 0, the residual
 * gains the curvature rows w (y[k-1] - 2 y[k] + y[k+1]), k = 1..n-2, of every
 * table with a free knot and n >= 3, tables in declaration order, after the
 * configuration rows (forcesmith_functor.cpp:53-55,198-207; spline.cpp:
 * 205-213); their Jacobian rows are the constants w (1, -2, 1) on the free
 * knots. fs_residual_count includes them; fs_row_offsets[C] is their first
 * row. Stillinger-Weber models have none. */
int fs_set_smooth_weight(fs_ctx *ctx, double weight);
/* model_smoothness_count / model_write_smoothness (smoothness.hpp:17-31): the
 * number of curvature rows whatever the weight, and the rows themselves,
 * weight * (y[k-1] - 2 y[k] + y[k+1]), at the current parameters (dst[count];
 * nothing written for Stillinger-Weber). */
int fs_smoothness_count(fs_ctx *ctx, int64_t *count);
int fs_write_smoothness(fs_ctx *ctx, double weight, double *dst);

/* CFittable surface (fit_params.hpp:24-30; TableParams
 * force_calculator_concept.hpp:171-208). */
int fs_param_count(fs_ctx *ctx, int32_t *count);
int fs_gather_params(fs_ctx *ctx, double *x);
int fs_scatter_params(fs_ctx *ctx, const double *x);
int fs_gather_bounds(fs_ctx *ctx, double *lo, double *hi);
int fs_max_cutoff(fs_ctx *ctx, double *rcut);     /* max_cutoff() */
int fs_residual_count(fs_ctx *ctx, int64_t *rows); /* count_residuals */
int fs_row_offsets(fs_ctx *ctx, int64_t *row_off); /* functor row_offset_ */

/* prepare(): GPU cell-list neighbour build whose pair set, order and FP64
 * displacements equal the reference brute force (neighbor_list.cpp:59-92). */
int fs_prepare(fs_ctx *ctx);
int fs_neighbor_count(fs_ctx *ctx, int64_t *pairs);
int fs_get_neighbor_list(fs_ctx *ctx, int64_t *pair_off, int32_t *j,
                         int32_t *img, double *d);

/* eval_forces over every configuration (EAMForceCalculator::eval_forces
 * eam_force.cpp:45-102 and siblings): forces[3N], energy[C], stress9[9C]
 * (virial / volume, eval_scope.hpp:14-31), limit[C]. Any pointer may be NULL. */
int fs_eval_forces(fs_ctx *ctx, double *forces, double *energy,
                   double *stress9, double *limit);
/* eval_cached(cache_index, forces, energy, stress) for one configuration
 * (force_calculator.hpp:75-76) from the last fs_eval_forces/residual. */
int fs_eval_config(fs_ctx *ctx, int32_t config, double *forces,
                   double *energy, double *stress9, double *limit);
/* Residual vector F(x) at the current parameters (ForcesmithFunctor::
 * operator(), forcesmith_functor.cpp:221-231); F may be NULL (device only).
 * cost (may be NULL) receives |F|^2. */
int fs_eval_residual(fs_ctx *ctx, double *F, double *cost);
/* Jacobian dF/dx, analytic per-knot (per-parameter for SW), FP32, row-major
 * R x D. Replaces the central-difference fallback of
 * ForcesmithFunctor::df (forcesmith_functor.cpp:252-292; fd_jacobian.hpp:
 * 37-49) and eval_cached_jacobian (force_calculator.hpp:77-78). J may be
 * NULL (device only). */
int fs_eval_jacobian(fs_ctx *ctx, float *J);
/* One residual + Jacobian step at the parameter vector x[D]: the
 * functor's operator() and df at one x (forcesmith_functor.cpp:221-292)
 * in one call. The first call at a layout runs eagerly and captures the step
 * (parameter scatter, slopes, residual, Jacobian, cost) as a hipGraph that
 * reads x from a device buffer; later calls copy x there and replay the
 * graph (Stillinger-Weber models run eagerly). F[R], J[R*D] (FP32) and cost
 * may each be NULL (device only); the device buffers hold the step's results
 * as after fs_eval_residual and fs_eval_jacobian. With all three NULL the
 * step stays queued on the context stream (fs_synchronize waits for it). */
int fs_eval_step(fs_ctx *ctx, const double *x, double *F, float *J, double *cost);

/* FP64 normal equations of the last residual/Jacobian: JtJ[D*D], Jtr[D]. */
int fs_normal_equations(fs_ctx *ctx, double *JtJ, double *Jtr, double *cost);

/* FP64 normal equations of caller-supplied device arrays, allocating
 * nothing. No counterpart in the reference (its NormalEquationsSolver forms
 * J^T J on the CPU, solver.cpp:242-266); it is for callers that cannot
 * allocate device memory: code under HIP graph capture, tight serving loops.
 * J: float[R * D] row-major, F: double[R]; out: JtJ double[D * D], Jtr
 * double[D], cost double[1] (may be NULL: F . F); every pointer is device
 * memory of the context's device. The work is queued on the context's stream
 * (fs_synchronize waits for it); the products equal fs_normal_equations'
 * for the same J and F. Partial sums live in one static device array of
 * 64 MiB per device, set up by fs_ctx_create: D <= 256 (else
 * FS_ERR_UNSUPPORTED). Calls on one device share that array, so calls that
 * overlap on one device (other contexts or host threads) race. */
int fs_normal_equations_noalloc(fs_ctx *ctx, const float *J, const double *F,
                                int64_t R, int32_t D, double *JtJ, double *Jtr,
                                double *cost);

/* One damped Gauss-Newton step at the current parameters, for callers that
 * drive their own LM loop: dx = -(J^T J + lambda I)^-1 J^T F (parameters
 * unchanged; lambda >= 0). No reference counterpart: the reference's LM keeps
 * its steps inside Eigen. */
int fs_lm_step(fs_ctx *ctx, double lambda, double *dx);

/* One ridge-regularised normal-equation step, NormalEquationsSolver::minimize
 * (solver.cpp:242-266; solver.hpp:138-145): x -= (J^T J + lambda I)^-1 J^T F(x)
 * with lambda = ridge if > 0, else 1e-8 * max(mean diag(J^T J), 1). x is read
 * and written; the device state is left at the new x. */
int fs_normal_equations_solve(fs_ctx *ctx, double ridge, double *x);

/* Levenberg-Marquardt on the GPU (FP64 normal equations, damped Cholesky);
 * replaces Solver::minimize with EigenLMSolver (solver.cpp:67-85).
 * x is read and written; cost_hist[max_iter] (may be NULL). */
int fs_lm_minimize(fs_ctx *ctx, double *x, int32_t max_iter,
                   double *cost_hist, int32_t *accepted);

/* Levenberg-Marquardt with the reference's stopping semantics: replaces
 * Solver::minimize with EigenLMSolver (solver.cpp:67-85; solver.hpp:65-73),
 * i.e. Eigen's LevenbergMarquardt (MINPACK lmder, More 1978): trust region
 * on |D p|, the lmpar damping search, termination on xtol / ftol / gtol /
 * maxfev. J, J^T J and J^T F are formed on the GPU (FP64 products of the FP32
 * analytic Jacobian); the D x D trust-region algebra runs on the host from
 * the normal equations. */
typedef struct fs_lm_opts {
  int32_t max_iter; /* calls of one step (EigenLMSolver::max_iter; 500) */
  int32_t maxfev;   /* <= 0: max(max_iter, 1) * 100 + 100 (solver.cpp:75) */
  double xtol;      /* 1e-7 */
  double ftol;      /* 1e-7 */
  double gtol;      /* 0 (Eigen's default) */
  double factor;    /* initial step bound factor, 100 (Eigen's default) */
} fs_lm_opts;
/* opts NULL: the defaults above. x is read and written (the last accepted
 * point). status receives Eigen's LevenbergMarquardtSpace code as
 * EigenLMSolver::minimize returns it: 0 improper input, 1 relative reduction,
 * 2 relative error (also: still running after max_iter calls, solver.cpp:84),
 * 3 both, 4 gradient cosine, 5 too many evaluations, 6-8 tolerances too small.
 * counts[3] (may be NULL): residual evaluations, Jacobian evaluations,
 * accepted steps. cost_hist[max_iter] (may be NULL): |F|^2 after each call. */
int fs_lm_solve(fs_ctx *ctx, const fs_lm_opts *opts, double *x, int32_t *status,
                int32_t *counts, double *cost_hist);

/* Exact cost |F(x_m)|^2 of np parameter vectors X[np*D] (row-major), the
 * per-member cost of BoostDESolver (solver.cpp:143-146), batched. */
int fs_population_cost(fs_ctx *ctx, const double *X, int32_t np,
                       double *cost);

typedef struct fs_de_opts {
  int32_t np;          /* population size (reference: 15 * D) */
  int32_t generations; /* max generations */
  double F;            /* mutation factor */
  double CR;           /* crossover probability */
  uint64_t seed;
  int32_t surrogate;   /* 0 = plain exact-judged DE, 1 = MLP-screened */
  int32_t oversample;  /* trial candidates per member when screened (>= 1) */
  int32_t retrain_every;
  int32_t hidden;      /* MLP width */
  double target_cost;  /* stop when best <= target (<= 0: run all) */
  double rank_tau_min; /* fallback when Kendall tau of the surrogate drops */
  int32_t progress_every; /* print generation, best cost and elapsed time to
                             stderr every n generations (0: silent) */
  int32_t reserved[3];
} fs_de_opts;
/* Differential evolution with the whole population evaluated per batched
 * launch; replaces BoostDESolver::minimize (solver.cpp:113-155). lo/hi: the
 * bounds; an infinite bound becomes x -+ max(2 |x|, 5e-4) around the initial
 * value, as BoostDESolver derives its box (:118-125; fs_gather_bounds gives
 * +-inf for knots). np <= 0: 15 * D (NP_factor, solver.hpp:78-81). x: initial
 * guess in (clamped into the box as member 0, :138-141), best member out;
 * hist[generations] best cost. */
int fs_de_minimize(fs_ctx *ctx, const fs_de_opts *opts, const double *lo,
                   const double *hi, double *x, double *best_cost,
                   double *hist, int32_t *generations_run);

/* Statistics of the last fs_de_minimize call on this context (no reference
 * counterpart: the surrogate is new functionality). generations: generations
 * run; retrains: surrogate trainings; screened: generations whose trials were
 * picked by the surrogate; fallbacks: screened generations after which the
 * Kendall tau of predicted against exact costs fell below rank_tau_min (the
 * driver then proposes plain trials until the next retraining); first_loss /
 * last_loss: the training loss (mean 0.5 (y - t)^2 of standardised log costs)
 * at the first step of the first retraining and the last step of the last
 * (0 when nothing was trained). */
typedef struct fs_de_stats {
  int32_t generations;
  int32_t retrains;
  int32_t screened;
  int32_t fallbacks;
  float first_loss;
  float last_loss;
  int32_t reserved[2];
} fs_de_stats;
int fs_de_last_stats(fs_ctx *ctx, fs_de_stats *out);

/* EAM gauge transformation (src/core/rescale.cpp, include/forcesmith/core/
 * rescale.hpp:11-19): the reference wraps its tables
 * (core/scale_potentials.hpp) as density A g_t (ScaledOutputPotential),
 * embedding F_t(rho / A) - S_t rho - F0_t over the span A [lo, hi]
 * (ScaledArgPotential under LinearAdjustedPotential) and pair
 * phi_ij + S_i A g_j + S_j A g_i (CompensatedPairPotential). Here the gauge is
 * that scale A and the per-type slopes S and intercepts F0 (ntypes entries),
 * applied by every later evaluation, Jacobian, solver and population call;
 * the knot values stay the parameters. A new model resets it to the
 * identity (A = 1, S = F0 = 0). EAM only (FS_ERR_UNSUPPORTED otherwise); a
 * slope that would compensate a pair table of another span than the density
 * table it pairs with is refused with FS_ERR_UNSUPPORTED.
 *   fs_eam_set_gauge / fs_eam_get_gauge: the gauge itself (NULL arrays: 0).
 *   fs_eam_rho_stats: compute_rho_ref (rescale.cpp:11-33) and the extremes
 *     of rescale_rho_axis (:39-52): per type the min, max and mean of the
 *     clamped densities over every configuration, in the gauge's units, and
 *     the atom count (min = +inf, max = -inf, mean = 0 for a type with no
 *     atoms); any output may be NULL.
 *   fs_eam_rescale_rho_axis: rescale_rho_axis (:35-101); *a receives the
 *     scale applied (1 when it leaves the gauge unchanged).
 *   fs_eam_embed_shift: embed_shift (:103-136) at the given reference
 *     densities.
 *   fs_eam_rescale: rescale_eam (:138-158); *a and rho_ref[ntypes] (either
 *     may be NULL) receive the rho-axis scale and the reference densities. */
int fs_eam_set_gauge(fs_ctx *ctx, double scale, const double *slope, const double *intercept);
int fs_eam_get_gauge(fs_ctx *ctx, double *scale, double *slope, double *intercept);
int fs_eam_rho_stats(fs_ctx *ctx, double *rho_min, double *rho_max, double *rho_mean,
                     int64_t *count);
int fs_eam_rescale_rho_axis(fs_ctx *ctx, double *a);
int fs_eam_embed_shift(fs_ctx *ctx, const double *rho_ref);
int fs_eam_rescale(fs_ctx *ctx, double *a, double *rho_ref);

/* Device views and tuning (no reference counterpart: measurement hooks). */
enum fs_buffer {
  FS_BUF_RESIDUAL = 0, /* double[R] */
  FS_BUF_JACOBIAN = 1, /* float[R*D] */
  FS_BUF_JTJ = 2,      /* double[D*D] */
  FS_BUF_JTR = 3       /* double[D] */
};
int fs_device_buffer(fs_ctx *ctx, int32_t which, void **ptr, int64_t *bytes);
int fs_synchronize(fs_ctx *ctx);
/* Launch shape: lanes per atom (4..64), blocks per CU for persistent grids.
 * fs_autotune times the candidates on the resident dataset and keeps the
 * fastest (lanes per atom and blocks per CU over the residual, then the
 * Jacobian tile size and the Jacobian passes' blocks per CU). */
int fs_set_launch(fs_ctx *ctx, int32_t lanes_per_atom, int32_t blocks_per_cu);
/* Wall time of the last fs_lm_solve in ms: ms[0] the whole solve, ms[1] its
 * device calls (residual norms, Jacobians, normal equations, each ending in
 * a stream synchronisation, copies included), ms[2] the rest: the host's
 * D x D trust-region algebra (QR-free lmpar on the normal equations). */
int fs_lm_last_timing(fs_ctx *ctx, double *ms);
/* Atoms per Jacobian tile: 0 (the default) picks it from the atom count, else
 * 4, 8, 16, 32 or 64. Unless a size is fixed here (non-zero), fs_autotune
 * also times the tile sizes over the whole Jacobian and keeps the fastest.
 * fs_get_launch reports the current shape
 * (jacobian_tile: the tile size of the last Jacobian, or the requested one
 * before the first; jacobian_blocks_per_cu: the Jacobian passes' grid). No reference counterpart (launch tuning). */
int fs_set_jacobian_tile(fs_ctx *ctx, int32_t atoms);
/* Form of the Jacobian's pair-table columns and density rows: 0 (default)
 * the knot-sum form while it fits (per-atom sums of the pair geometry, built
 * once per neighbour list and model grid and folded with the current slope
 * band at every Jacobian; at most 128 pair-table and 128 density knots per
 * atom), else the window passes; 1 the window passes; 2 the knot-sum form
 * (FS_ERR_UNSUPPORTED from the Jacobian when it does not fit). Both forms
 * give the same columns to FP32 rounding. No reference counterpart (kernel
 * choice). */
int fs_set_jacobian_form(fs_ctx *ctx, int32_t form);
/* Form of the batched population cost (fs_population_cost, fs_de_minimize):
 * 0 (default) the member-lane passes where they apply (EAM models whose pair
 * and density tables share one knot grid, with a block's 64 members' knot
 * rows within the LDS): one member per lane, FP64 arithmetic with the cached
 * sites of the FP64 path in either precision mode; otherwise, and with 1, the
 * member-block passes (1, 2 or 4 members per block, the context's precision).
 * No reference counterpart (kernel choice). */
int fs_set_population_form(fs_ctx *ctx, int32_t form);
int fs_get_launch(fs_ctx *ctx, int32_t *lanes_per_atom, int32_t *blocks_per_cu,
                  int32_t *jacobian_tile, int32_t *jacobian_blocks_per_cu);
int fs_autotune(fs_ctx *ctx, int32_t *lanes_per_atom);
/* Average time (ms) of one stage over reps launches, measured with HIP events
 * on the context stream. what: 0 residual, 1 Jacobian (all its kernels),
 * 2 normal equations, 3 residual + Jacobian, 4 the density/embedding-column
 * pass of the Jacobian (ADP: with the dipole/quadrupole-column kernels that
 * follow it), 5 the pair-column pass, 6 the ADP quadrupole-column force
 * pass. No reference counterpart (measurement only). */
int fs_time_step(fs_ctx *ctx, int32_t what, int32_t reps, double *ms_per_rep);

/* MLP surrogate kernels on host buffers (no reference counterpart; the
 * surrogate is new functionality that screens DE trial vectors). params is
 * one FP32 buffer [W1 (h x d) | b1 (h) | W2 (h x h) | b2 (h) | W3 (h) | b3];
 * products run as BF16 matrix-core operations with FP32 accumulation, GELU in
 * its tanh form. fs_mlp_loss_grad returns the gradient of
 * mean 0.5 (y - t)^2 in the same layout. */
int fs_mlp_forward(fs_ctx *ctx, int32_t n, int32_t d, int32_t h,
                   const float *params, const float *X, float *y);
int fs_mlp_loss_grad(fs_ctx *ctx, int32_t n, int32_t d, int32_t h,
                     const float *params, const float *X, const float *t,
                     float *grad, float *loss);

#ifdef __cplusplus
}
#endif
#endif
