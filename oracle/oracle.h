/*
 * FP64 CPU oracle for the force-matching hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is a plain-C restatement of the
 * reference calculator's algorithm; it exists to check the HIP product path.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it, and only as the checker or as the timed CPU baseline. The product
 * (libforcesmith_amd.so) never links, loads or calls it.
 *
 * Every function cites the reference file:line it restates. Parity pins:
 *   neighbour list  - reference brute force restated exactly; pinned by the
 *                     FCC 12-neighbour / minimum-image known answers and by
 *                     brute-force enumeration on tiny inputs.
 *   makima slopes   - pinned by scipy Akima1DInterpolator(method="makima")
 *                     golden vectors (interior) and the reference's spline
 *                     known answers (knots, linear exact, extrapolation).
 *   EAM/ADP/SW      - pinned by the reference's closed forms (dimer density,
 *                     embedding, SW dimer energy), invariants (Newton's third
 *                     law, ADP(u=w=0)==EAM) and F = -dE/dx finite differences.
 *   analytic J      - pinned against the reference's own central-difference
 *                     Jacobian (h = 1e-5) with its test metric.
 *   LM (MINPACK)    - pinned by the published minima of the More-Garbow-
 *                     Hillstrom test problems and the defining properties of
 *                     the pivoted QR and lmpar (Eigen's LevenbergMarquardt is
 *                     not vendored by the reference).
 *   DE              - pinned by the invariants of Storn-Price rand/1/bin
 *                     (distinct picks, forced jrand component, CR = 0 / 1
 *                     limits, box clamping, <= selection, crossover rate) and
 *                     by the known minima of the sphere and Rosenbrock
 *                     functions; the RNG stream is this build's own (Boost.Math
 *                     and mt19937_64 trajectories are not reproducible here).
 *   MLP             - gradients pinned by central differences of the loss.
 */
#ifndef FS_ORACLE_H
#define FS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { OR_EAM = 0, OR_ADP = 1, OR_SW = 2 };

typedef struct or_ctx or_ctx;

/* Dataset: cell[9c + 3k + a] = component a of lattice vector k of config c
 * (the reference stores the vectors as the COLUMNS of its box matrix,
 * src/io/config_reader.cpp:58-66). ref_stress is 6 per config in the residual
 * order xx, yy, zz, xy, xz, yz (include/forcesmith/core/voigt.hpp:9-13). */
or_ctx *or_create(int kind, int ntypes, int n_configs, const int64_t *atom_off,
                  const int32_t *types, const double *pos,
                  const double *ref_force, const double *cell,
                  const double *ref_energy, const double *ref_stress);
void or_destroy(or_ctx *c);
void or_set_threads(int n);
int or_get_threads(void);

/* Tabulated model for EAM / ADP. Tables in the reference declaration order
 * (include/forcesmith/force/eam_force.hpp:24-32, adp_force.hpp:37-51):
 * pair[P] (pair_ordinal order), density[T], embedding[T] and for ADP
 * dipole[P], quadrupole[P]; P = T(T+1)/2. */
int or_set_tables(or_ctx *c, int ntab, const int32_t *n, const double *rmin,
                  const double *rmax, const double *y, const uint8_t *fixed);
/* Stillinger-Weber: pair slots (pair_ordinal order) of 8 params
 * A,B,p,q,delta,a1,gamma,a2 (include/forcesmith/force/stiweb_force.hpp:19-29)
 * and lambda[T * P] (lambda_index ti*P + slot(tj,tk)). */
int or_set_sw(or_ctx *c, const double *pair_params, const uint8_t *pair_fixed,
              const double *lambda, const uint8_t *lambda_fixed);

int or_param_count(const or_ctx *c);
void or_gather_params(const or_ctx *c, double *x);
void or_scatter_params(or_ctx *c, const double *x);
double or_max_cutoff(const or_ctx *c);
void or_set_weights(or_ctx *c, double energy_weight, double stress_weight);
/* curvature (smoothness) rows appended after the configuration rows when
 * the weight is > 0 (0 rows for SW) */
void or_set_smooth_weight(or_ctx *c, double weight);
int64_t or_smooth_count(const or_ctx *c);
int64_t or_residual_count(const or_ctx *c);
void or_row_offsets(const or_ctx *c, int64_t *row_off);

/* Neighbour list with the model cutoff (plus site priming), or an explicit
 * cutoff (site priming skipped). Returns the pair count. */
int64_t or_prepare(or_ctx *c);
int64_t or_build_nl(or_ctx *c, double rcut);
void or_get_nl(const or_ctx *c, int64_t *pair_off, int32_t *j, int32_t *img,
               double *d);

/* eval_forces over all configs at the current parameters. */
void or_eval_forces(or_ctx *c, double *forces, double *energy, double *stress9,
                    double *limit);
void or_get_scratch(const or_ctx *c, double *rho, double *gradF, double *mu,
                    double *lambda9);
void or_residual(or_ctx *c, const double *x, double *F);
void or_jacobian_fd(or_ctx *c, const double *x, double *J);
void or_jacobian_fd_cols(or_ctx *c, const double *x, int col0, int ncols,
                         double *Jcols);
int or_jacobian_analytic(or_ctx *c, const double *x, double *J);

/* EAM gauge (src/core/rescale.cpp, include/forcesmith/core/
 * scale_potentials.hpp): density scale A, per-type embedding slope S and
 * intercept F0 (ntypes each; NULL = zeros). or_set_gauge: -1 for a model
 * other than EAM, -2 for A not finite or 0. or_jacobian_analytic returns -3
 * under a gauge whose slopes compensate pairs of another span. */
int or_set_gauge(or_ctx *c, double A, const double *S, const double *F0);
void or_get_gauge(const or_ctx *c, double *A, double *S, double *F0);
int or_gauge_spans_match(const or_ctx *c);
void or_emb_gauged(const or_ctx *c, int t, double rho, double *F, double *dF);
void or_rho_stats(or_ctx *c, double *rmin, double *rmax, double *rmean, int64_t *count);
double or_rescale_rho_axis(or_ctx *c);
void or_embed_shift(or_ctx *c, const double *rho_ref);
double or_rescale_eam(or_ctx *c, double *rho_ref);

/* Spline primitives (SplinePotential), exposed for unit tests. */
int or_spline_slopes(int n, const double *x, const double *y, double *s);
void or_spline_eval(int n, const double *x, const double *y, const double *s,
                    double r, double *v, double *d);
void or_spline_eval_cached(int n, const double *x, const double *y,
                           const double *s, double r, double *v, double *d);
void or_spline_slope_jacobian(int n, const double *x, const double *y,
                              double *S);
void or_spline_param_grad(int n, const double *x, const double *y, double r,
                          double *dv, double *dd, double *d2);
void or_knot_grid(int n, double rmin, double rmax, double *x);

/* Levenberg-Marquardt on the normal equations (the algorithm the GPU path
 * runs). Returns the number of accepted steps; cost_hist[it] = |F|^2. */
int or_lm(or_ctx *c, double *x, int max_iter, int use_fd, double *cost_hist);

/* Exact population cost |F(x_m)|^2 for NP members (X row-major NP x D). */
void or_population_cost(or_ctx *c, const double *X, int np, double *cost);

/* Counter-based RNG shared with the GPU DE (splitmix64 of a key). */
uint64_t or_rng_u64(uint64_t seed, uint64_t stream, uint64_t index);
double or_rng_unit(uint64_t seed, uint64_t stream, uint64_t index);

/* Plain DE/rand/1/bin with synchronous generations (algorithm shared with
 * the GPU driver). Returns the best cost; x <- best member. */
double or_de(or_ctx *c, double *x, const double *lo, const double *hi, int np,
             int generations, double F, double CR, uint64_t seed,
             double *best_hist);
/* The pieces of or_de, exposed so tests can check each one:
 * or_de_init: member 0 = clamp(x), the others uniform in the box;
 * or_de_trial: trial of member m in generation g; picks (may be NULL)
 *   receives r1, r2, r3 and jrand;
 * or_de_select: greedy selection (trial kept when its cost is <= the
 *   member's); kept[m] (may be NULL) flags the replaced members; returns the
 *   index of the first member of lowest cost;
 * or_de_fn: the whole driver on any population cost function. */
typedef void (*or_popcost)(void *user, const double *X, int np, double *cost);
void or_de_init(const double *x, const double *lo, const double *hi, int np, int D,
                uint64_t seed, double *pop);
void or_de_trial(const double *pop, int np, int D, const double *lo, const double *hi,
                 double F, double CR, uint64_t seed, int g, int m, double *out,
                 int *picks);
int or_de_select(int np, int D, double *pop, double *cost, const double *trial,
                 const double *tcost, unsigned char *kept);
double or_de_fn(or_popcost fn, void *user, int D, double *x, const double *lo,
                const double *hi, int np, int generations, double F, double CR,
                uint64_t seed, double *best_hist);

/* FP64 MLP (D -> H -> H -> 1, GELU tanh form) forward and backward, the CPU
 * check for the BF16-MFMA surrogate kernels. W1[H x D], W2[H x H], W3[H]. */
void or_mlp_forward(int n, int d, int h, const double *X, const double *W1,
                    const double *b1, const double *W2, const double *b2,
                    const double *W3, double b3, double *y);
double or_mlp_loss_grad(int n, int d, int h, const double *X, const double *t,
                        const double *W1, const double *b1, const double *W2,
                        const double *b2, const double *W3, double b3,
                        double *gW1, double *gb1, double *gW2, double *gb2,
                        double *gW3, double *gb3);

/* Wall-clock CPU baseline: FD Jacobian over ncols columns + 1 residual,
 * config-parallel on the current thread count. Returns seconds. */
double or_time_fd_jacobian(or_ctx *c, const double *x, int ncols, double *J);
double or_time_residual(or_ctx *c, const double *x, int reps, double *F);

/* Levenberg-Marquardt with the reference's stopping semantics
 * (EigenLMSolver, solver.cpp:67-85): MINPACK lmder (More 1978) with QR and
 * column pivoting. Returns the Eigen status (0 improper input, 1..8, and 2
 * when still running after max_iter calls of one step); counts (may be NULL)
 * receives {nfev, njev, accepted steps}; cost_hist[max_iter] (may be NULL)
 * the cost |f|^2 after each call. */
typedef int (*or_fcn)(void *user, const double *x, double *f);
typedef int (*or_jac)(void *user, const double *x, double *J); /* m x n row-major */
typedef struct {
  int32_t max_iter, maxfev; /* maxfev <= 0: max(max_iter, 1) * 100 + 100 */
  double xtol, ftol, gtol, factor;
} or_lm_opts;
int or_lmder(int64_t m, int n, or_fcn fcn, or_jac jac, void *user, double *x,
             const or_lm_opts *o, int32_t *counts, double *cost_hist);
int or_lm_minpack(or_ctx *c, double *x, const or_lm_opts *o, int32_t *counts,
                  double *cost_hist);
/* NormalEquationsSolver (solver.cpp:242-266): one ridge step, x in/out */
int or_normal_equations_solve(or_ctx *c, double ridge, double *x);
/* the pieces, for tests: pivoted QR (returns the rank; R column-major n x n)
 * and lmpar (returns the number of damped solves) */
int or_qr_pivot(int64_t m, int n, const double *J, const double *f, int32_t *perm,
                double *R, double *qtf);
int or_lmpar(int n, const double *R, const int32_t *perm, int rank, const double *diag,
             const double *qtf, double delta, double *par, double *x);

#ifdef __cplusplus
}
#endif
#endif
