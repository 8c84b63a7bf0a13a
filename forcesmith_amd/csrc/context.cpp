// C-ABI entry points of libforcesmith_amd (declared in include/forcesmith_amd.h).
// Host-side state management; every compute step is a HIP kernel launched on
// the context's stream. No exception crosses the ABI: each entry point catches
// and converts to a status code plus a message.
#include "common.hpp"
#include "mlp.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <new>

namespace fsa {

void destroy_surrogate(SurrogateState *s);
void destroy_step_graph(Context &c);

Context::~Context() {
  destroy_step_graph(*this);
  delete mlp;
  destroy_surrogate(surr);
  if (jev0) (void)hipEventDestroy(jev0);
  if (jev1) (void)hipEventDestroy(jev1);
  if (stream) (void)hipStreamDestroy(stream);
}

namespace {

thread_local std::string g_err;

// box matrix (columns = lattice vectors), cofactor inverse, |det|; a singular
// or non-finite box gives the identity inverse and volume |det| or 0, as
// PeriodicBC::set_box (include/forcesmith/core/boundary_conditions.hpp:21-30)
void box_from_cell(const double *cell, double *b, double *inv, double *vol) {
  for (int k = 0; k < 3; ++k)
    for (int a = 0; a < 3; ++a) b[3 * a + k] = cell[3 * k + a];
  const double *m = b;
  const double c00 = m[4] * m[8] - m[5] * m[7];
  const double c10 = m[2] * m[7] - m[1] * m[8];
  const double c20 = m[1] * m[5] - m[2] * m[4];
  const double det = m[0] * c00 + m[3] * c10 + m[6] * c20;
  bool finite = std::isfinite(det);
  for (int i = 0; i < 9; ++i) finite = finite && std::isfinite(m[i]);
  if (finite && det != 0.0) {
    const double id = 1.0 / det;
    inv[0] = c00 * id;
    inv[1] = c10 * id;
    inv[2] = c20 * id;
    inv[3] = (m[5] * m[6] - m[3] * m[8]) * id;
    inv[4] = (m[0] * m[8] - m[2] * m[6]) * id;
    inv[5] = (m[2] * m[3] - m[0] * m[5]) * id;
    inv[6] = (m[3] * m[7] - m[4] * m[6]) * id;
    inv[7] = (m[1] * m[6] - m[0] * m[7]) * id;
    inv[8] = (m[0] * m[4] - m[1] * m[3]) * id;
  } else {
    for (int i = 0; i < 9; ++i) inv[i] = (i % 4 == 0) ? 1.0 : 0.0;
  }
  *vol = std::isfinite(det) ? std::fabs(det) : 0.0;
}

template <class T> void upload(Context &c, DBuf<T> &d, const std::vector<T> &h) {
  d.alloc(h.empty() ? 1 : h.size());
  if (!h.empty())
    FS_HIP(hipMemcpyAsync(d.p, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice, c.stream));
}

void build_tiles(Context &c) {
  const int per = 256 / c.lanes_per_atom;
  c.tile_atoms = per;
  c.h_tile_cfg.clear();
  c.h_tile_a0.clear();
  c.h_tile_na.clear();
  c.h_cfg_tile0.assign(c.C + 1, 0);
  for (int k = 0; k < c.C; ++k) {
    c.h_cfg_tile0[k] = int32_t(c.h_tile_cfg.size());
    for (int64_t a = c.h_atom_off[k]; a < c.h_atom_off[k + 1]; a += per) {
      c.h_tile_cfg.push_back(k);
      c.h_tile_a0.push_back(int32_t(a));
      c.h_tile_na.push_back(int32_t(std::min<int64_t>(per, c.h_atom_off[k + 1] - a)));
    }
  }
  c.h_cfg_tile0[c.C] = int32_t(c.h_tile_cfg.size());
  c.ntiles = int(c.h_tile_cfg.size());
  upload(c, c.tile_cfg, c.h_tile_cfg);
  upload(c, c.tile_a0, c.h_tile_a0);
  upload(c, c.tile_na, c.h_tile_na);
  upload(c, c.cfg_tile0, c.h_cfg_tile0);
  c.tile_part.alloc(size_t(std::max(c.ntiles, 1)) * 11);
}

} // namespace

bool stage_intervals(const Context &c) {
  return size_t(c.ntab) * sizeof(TableDesc) + 2 * size_t(c.nivl) * sizeof(float4) +
             4 * size_t(c.nivl) * sizeof(double) <=
         kTableStageBytes;
}

namespace {

// curvature rows of the tabulated model: n - 2 per table with a free knot and
// n >= 3, tables in declaration order (SplinePotential::curvature_count,
// spline.hpp:104-106; model_smoothness_count, smoothness.hpp:17-23), only when
// the weight is positive (forcesmith_functor.cpp:202-207); SW has none
void smooth_layout(Context &c) {
  std::vector<int4> rows;
  if (c.smooth_weight > 0.0 && c.have_model && c.kind != FS_SW) {
    for (int t = 0; t < c.ntab; ++t) {
      const TableDesc &T = c.h_tabs[t];
      bool any_free = false;
      for (int k = 0; k < T.n; ++k) any_free |= c.h_col_of[T.off + k] >= 0;
      if (!any_free || T.n < 3) continue;
      for (int k = 1; k + 1 < T.n; ++k)
        rows.push_back(int4{T.off + k - 1, c.h_col_of[T.off + k - 1], c.h_col_of[T.off + k],
                            c.h_col_of[T.off + k + 1]});
    }
  }
  c.n_smooth = int(rows.size());
  if (c.n_smooth > 0) upload(c, c.smooth_rows, rows);
  c.R = c.R_cfg + c.n_smooth;
  c.residual.alloc(size_t(std::max<int64_t>(c.R, 1)));
  c.forces_valid = false;
}

void compute_rows(Context &c) {
  c.h_row_off.assign(c.C + 1, 0);
  for (int k = 0; k < c.C; ++k) {
    const int64_t na = c.h_atom_off[k + 1] - c.h_atom_off[k];
    c.h_row_off[k + 1] = c.h_row_off[k] + 3 * na + 2 + (c.stress_weight > 0.0 ? 6 : 0);
  }
  c.R_cfg = c.h_row_off[c.C];
  upload(c, c.row_off, c.h_row_off);
  smooth_layout(c);
}

double max_cutoff(const Context &c) {
  double m = 0.0;
  if (c.kind == FS_SW) {
    for (int s = 0; s < c.npair; ++s)
      m = std::max(m, std::max(c.h_sw[8 * s + 5], c.h_sw[8 * s + 7]));
    return m;
  }
  for (int t = 0; t < c.ntab; ++t) {
    const bool emb = t >= c.npair + c.ntypes && t < c.npair + 2 * c.ntypes;
    if (!emb) m = std::max(m, c.h_tabs[t].xn);
  }
  return m;
}

// SW neighbour lists are built with a 5 % skin so parameter updates of a1/a2
// rarely force a rebuild; pairs beyond the current cutoff contribute exactly
// zero (v2 = 0 for r >= a1, h = 0 for r >= a2, stiweb_force.cpp:24-54).
constexpr double kSwSkin = 1.05;

void ensure_nl(Context &c) {
  c.fail_if_no_dataset();
  c.fail_if_no_model();
  const double rc = max_cutoff(c);
  bool rebuild = !c.have_nl;
  if (c.kind == FS_SW) {
    if (rc > c.nl_rc) rebuild = true;
    if (rebuild) c.nl_rc = rc * kSwSkin;
  } else {
    if (rc != c.nl_rc) rebuild = true;
    c.nl_rc = rc;
  }
  if (rebuild) {
    ++c.sw_gen;
    launch_neighbor_build(c);
    launch_pair_flags(c);
    if (c.kind == FS_SW) launch_sw_rev(c);
    c.forces_valid = false;
  }
}

void refresh_params(Context &c) {
  if (!c.params_dirty) return;
  if (c.kind == FS_SW) {
    c.sw.alloc(c.h_sw.size());
    FS_HIP(hipMemcpyAsync(c.sw.p, c.h_sw.data(), sizeof(double) * c.h_sw.size(),
                          hipMemcpyHostToDevice, c.stream));
    ++c.sw_gen;
    c.lam.alloc(std::max<size_t>(c.h_lam.size(), 1));
    FS_HIP(hipMemcpyAsync(c.lam.p, c.h_lam.data(), sizeof(double) * c.h_lam.size(),
                          hipMemcpyHostToDevice, c.stream));
  } else {
    FS_HIP(hipMemcpyAsync(c.tab_y.p, c.h_y.data(), sizeof(double) * c.h_y.size(),
                          hipMemcpyHostToDevice, c.stream));
    launch_slopes(c);
  }
  c.params_dirty = false;
  c.forces_valid = false;
}

void alloc_eval(Context &c) {
  const size_t N = size_t(std::max<int64_t>(c.N, 1));
  c.rho.alloc(N);
  c.clamp_over.alloc(N);
  c.Fp.alloc(N);
  c.Fp32.alloc(N);
  c.Fpp.alloc(N);
  c.e_atom.alloc(N);
  c.lim_atom.alloc(N);
  c.mu.alloc(3 * N);
  c.lamb.alloc(6 * N);
  c.mu32.alloc(N);
  c.lamb32.alloc(2 * N);
  c.forces.alloc(3 * N);
  c.cfg_E.alloc(std::max(c.C, 1));
  c.cfg_S.alloc(9 * size_t(std::max(c.C, 1)));
  c.cfg_L.alloc(std::max(c.C, 1));
  c.cost.alloc(4);
}

void evaluate(Context &c) {
  ensure_nl(c);
  refresh_params(c);
  if (c.forces_valid) return;
  alloc_eval(c);
  launch_eval(c, true);
  c.forces_valid = true;
}

template <class F> int guarded(fs_ctx *ctx, F &&f) {
  if (!ctx) {
    g_err = "null context";
    return FS_ERR_ARG;
  }
  Context &c = *reinterpret_cast<Context *>(ctx);
  std::lock_guard<std::mutex> lk(c.lock);
  try {
    FS_HIP(hipSetDevice(c.device));
    f(c);
    return FS_OK;
  } catch (const Error &e) {
    c.err = e.what();
    return e.code;
  } catch (const std::bad_alloc &) {
    c.err = "host allocation failed";
    return FS_ERR_NOMEM;
  } catch (const std::exception &e) {
    c.err = e.what();
    return FS_ERR_ARG;
  }
}

} // namespace

// these are defined in the solver translation unit
int fs_jacobian_impl(Context &c, float *J);
int fs_normal_impl(Context &c, double *JtJ, double *Jtr, double *cost);
int fs_lm_impl(Context &c, double *x, int max_iter, double *hist, int *accepted);
int fs_lm_solve_impl(Context &c, const fs_lm_opts &o, double *x, int *status, int *counts,
                     double *hist);
int fs_ne_solve_impl(Context &c, double ridge, double *x);
int fs_lm_step_impl(Context &c, double lambda, double *dx);
int fs_population_impl(Context &c, const double *X, int np, double *cost);
int fs_step_impl(Context &c, const double *x, double *F, float *J, double *cost);
int fs_ne_noalloc_impl(Context &c, const float *J, const double *F, int64_t R, int D, double *JtJ,
                       double *Jtr, double *cost);
int fs_de_impl(Context &c, const fs_de_opts *o, const double *lo, const double *hi, double *x,
               double *best, double *hist, int *gens);
double time_step_impl(Context &c, int what, int reps);
int autotune_impl(Context &c);
void evaluate_for_solver(Context &c) { evaluate(c); }
void params_to_device(Context &c) { refresh_params(c); }
void rebuild_tiles(Context &c) { build_tiles(c); }
void ensure_ready(Context &c) {
  ensure_nl(c);
  refresh_params(c);
  alloc_eval(c);
}

} // namespace fsa

using namespace fsa;

extern "C" {

int fs_abi_version(void) { return FS_ABI_VERSION; }

const char *fs_last_error(const fs_ctx *ctx) {
  if (!ctx) return g_err.c_str();
  return reinterpret_cast<const Context *>(ctx)->err.c_str();
}

int fs_ctx_create(const fs_opts *opts, fs_ctx **out) {
  if (!out) {
    g_err = "null output pointer";
    return FS_ERR_ARG;
  }
  *out = nullptr;
  Context *c = nullptr;
  try {
    c = new Context();
    c->device = opts ? opts->device : 0;
    c->precision = opts ? opts->precision : FS_MIXED;
    if (c->precision != FS_MIXED && c->precision != FS_FP64) throw Error(FS_ERR_ARG, "bad precision");
    int ndev = 0;
    FS_HIP(hipGetDeviceCount(&ndev));
    FS_REQUIRE(c->device >= 0 && c->device < ndev, FS_ERR_ARG, "no such HIP device");
    FS_HIP(hipSetDevice(c->device));
    FS_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    int cus = 0;
    FS_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device));
    c->cu_count = cus > 0 ? cus : 256;
    noalloc_scratch_init(c->device);
    *out = reinterpret_cast<fs_ctx *>(c);
    return FS_OK;
  } catch (const Error &e) {
    g_err = e.what();
    delete c;
    return e.code;
  } catch (const std::exception &e) {
    g_err = e.what();
    delete c;
    return FS_ERR_ARG;
  }
}

int fs_ctx_retain(fs_ctx *ctx) {
  return guarded(ctx, [](Context &c) { ++c.refs; });
}

int fs_ctx_release(fs_ctx *ctx) {
  if (!ctx) return FS_ERR_ARG;
  Context *c = reinterpret_cast<Context *>(ctx);
  bool last = false;
  {
    std::lock_guard<std::mutex> lk(c->lock);
    last = --c->refs == 0;
  }
  if (last) {
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    delete c;
  }
  return FS_OK;
}

int fs_upload_dataset(fs_ctx *ctx, int32_t n_configs, const int64_t *atom_off,
                      const int32_t *types, const double *pos, const double *ref_force,
                      const double *cell, const double *ref_energy, const double *ref_stress) {
  return guarded(ctx, [&](Context &c) {
    FS_REQUIRE(n_configs > 0 && atom_off && types && pos && ref_force && cell && ref_energy &&
                   ref_stress,
               FS_ERR_ARG, "dataset arrays must be non-null and n_configs > 0");
    FS_REQUIRE(atom_off[0] == 0, FS_ERR_ARG, "atom_off[0] must be 0");
    for (int k = 0; k < n_configs; ++k) {
      FS_REQUIRE(atom_off[k + 1] >= atom_off[k], FS_ERR_ARG, "atom_off must be non-decreasing");
      FS_REQUIRE(atom_off[k + 1] - atom_off[k] <= int64_t(kJMask), FS_ERR_UNSUPPORTED,
                 "configuration larger than 4194303 atoms");
    }
    const int64_t N = atom_off[n_configs];
    FS_REQUIRE(N < (int64_t(1) << 31), FS_ERR_UNSUPPORTED, "more than 2^31 atoms");
    for (int64_t a = 0; a < 3 * N; ++a)
      FS_REQUIRE(std::isfinite(pos[a]), FS_ERR_ARG, "non-finite atom position");
    for (int64_t a = 0; a < N; ++a)
      FS_REQUIRE(types[a] >= 0 && types[a] < kMaxTypes, FS_ERR_ARG, "atom type out of range");
    c.C = n_configs;
    c.N = N;
    c.h_atom_off.assign(atom_off, atom_off + n_configs + 1);
    c.h_types.assign(types, types + N);
    c.h_pos.assign(pos, pos + 3 * N);
    c.h_ref_force.assign(ref_force, ref_force + 3 * N);
    c.h_cell.assign(cell, cell + 9 * n_configs);
    c.h_ref_energy.assign(ref_energy, ref_energy + n_configs);
    c.h_ref_stress.assign(ref_stress, ref_stress + 6 * n_configs);
    c.h_cfg_of.resize(N);
    for (int k = 0; k < n_configs; ++k)
      for (int64_t a = atom_off[k]; a < atom_off[k + 1]; ++a) c.h_cfg_of[a] = k;
    c.h_box.resize(9 * n_configs);
    c.h_inv.resize(9 * n_configs);
    c.h_vol.resize(n_configs);
    for (int k = 0; k < n_configs; ++k)
      box_from_cell(cell + 9 * k, &c.h_box[9 * k], &c.h_inv[9 * k], &c.h_vol[k]);
    upload(c, c.atom_off, c.h_atom_off);
    upload(c, c.types, c.h_types);
    upload(c, c.cfg_of, c.h_cfg_of);
    upload(c, c.pos, c.h_pos);
    upload(c, c.ref_force, c.h_ref_force);
    upload(c, c.ref_energy, c.h_ref_energy);
    upload(c, c.ref_stress, c.h_ref_stress);
    upload(c, c.box, c.h_box);
    upload(c, c.inv, c.h_inv);
    upload(c, c.vol, c.h_vol);
    build_tiles(c);
    compute_rows(c);
    c.njt = 0;
    c.have_nl = false;
    c.nl_rc = -1.0;
    c.forces_valid = false;
    FS_HIP(hipStreamSynchronize(c.stream));
  });
}

int fs_set_tabulated_model(fs_ctx *ctx, int32_t kind, int32_t ntypes, int32_t ntab,
                           const int32_t *n_knots, const double *rmin, const double *rmax,
                           const double *y, const uint8_t *fixed) {
  return guarded(ctx, [&](Context &c) {
    FS_REQUIRE(kind == FS_EAM || kind == FS_ADP, FS_ERR_ARG, "kind must be FS_EAM or FS_ADP");
    FS_REQUIRE(ntypes >= 1 && ntypes <= kMaxTypes, FS_ERR_ARG, "ntypes must be in [1, 8]");
    const int np = ntypes * (ntypes + 1) / 2;
    const int want = kind == FS_ADP ? 3 * np + 2 * ntypes : np + 2 * ntypes;
    FS_REQUIRE(ntab == want, FS_ERR_ARG,
               "expected " + std::to_string(want) + " tables for this kind and ntypes");
    FS_REQUIRE(n_knots && rmin && rmax && y && fixed, FS_ERR_ARG, "null table array");
    c.kind = kind;
    c.ntypes = ntypes;
    c.npair = np;
    c.ntab = ntab;
    c.h_tabs.assign(ntab, TableDesc{});
    c.h_x.clear();
    c.h_y.clear();
    c.h_fixed.clear();
    int off = 0, ivl = 0;
    for (int t = 0; t < ntab; ++t) {
      const int n = n_knots[t];
      FS_REQUIRE(n >= 2 && n <= 4096, FS_ERR_ARG, "a table needs 2..4096 knots");
      FS_REQUIRE(std::isfinite(rmin[t]) && std::isfinite(rmax[t]) && rmax[t] > rmin[t],
                 FS_ERR_ARG, "tabulated potential requires rmax > rmin");
      const double h = (rmax[t] - rmin[t]) / double(n - 1);
      TableDesc &T = c.h_tabs[t];
      T.n = n;
      T.off = off;
      T.ivl = ivl;
      T.linear = n < 4;
      for (int k = 0; k < n; ++k) c.h_x.push_back(rmin[t] + double(k) * h);
      T.x0 = c.h_x[off];
      T.xn = c.h_x[off + n - 1];
      T.h = h;
      T.inv_h = 1.0 / h;
      T.x0f = float(T.x0);
      T.inv_hf = float(T.inv_h);
      T.hf = float(h);
      T.inv_hlo = float(T.inv_h - double(T.inv_hf));
      for (int k = 0; k < n; ++k) {
        c.h_y.push_back(y[off + k]);
        c.h_fixed.push_back(fixed[off + k] ? 1 : 0);
      }
      off += n;
      ivl += n - 1;
    }
    c.nknots = off;
    c.nivl = ivl;
    c.h_col_of.assign(off, -1);
    c.h_knot_of.clear();
    for (int k = 0; k < off; ++k)
      if (!c.h_fixed[k]) {
        c.h_col_of[k] = int32_t(c.h_knot_of.size());
        c.h_knot_of.push_back(k);
      }
    c.D = int(c.h_knot_of.size());
    upload(c, c.tabs, c.h_tabs);
    upload(c, c.tab_x, c.h_x);
    c.tab_y.alloc(off);
    c.tab_s.alloc(off);
    c.tab_S.alloc(size_t(off) * 5);
    c.ivl32.alloc(2 * size_t(std::max(ivl, 1)));
    c.ivl64.alloc(4 * size_t(std::max(ivl, 1)));
    upload(c, c.col_of, c.h_col_of);
    upload(c, c.knot_of, c.h_knot_of);
    reset_gauge(c);
    c.have_model = true;
    c.params_dirty = true;
    c.jl.valid = false;
    smooth_layout(c);
    if (c.have_nl && c.C > 0) {
      if (max_cutoff(c) == c.nl_rc) launch_pair_flags(c);
      else c.have_nl = false;
    }
    c.forces_valid = false;
    FS_HIP(hipStreamSynchronize(c.stream));
  });
}

int fs_set_sw_model(fs_ctx *ctx, int32_t ntypes, const double *pair_params,
                    const uint8_t *pair_fixed, const double *lambda, const uint8_t *lambda_fixed) {
  return guarded(ctx, [&](Context &c) {
    FS_REQUIRE(ntypes >= 1 && ntypes <= kMaxTypes, FS_ERR_ARG, "ntypes must be in [1, 8]");
    FS_REQUIRE(pair_params && pair_fixed && lambda && lambda_fixed, FS_ERR_ARG, "null SW array");
    c.kind = FS_SW;
    c.ntypes = ntypes;
    c.npair = ntypes * (ntypes + 1) / 2;
    c.ntab = 0;
    c.h_sw.assign(pair_params, pair_params + 8 * c.npair);
    c.h_sw_fixed.assign(pair_fixed, pair_fixed + 8 * c.npair);
    const int nl = ntypes * c.npair;
    c.h_lam.assign(lambda, lambda + nl);
    c.h_lam_fixed.assign(lambda_fixed, lambda_fixed + nl);
    c.D = 0;
    for (auto f : c.h_sw_fixed) c.D += !f;
    for (auto f : c.h_lam_fixed) c.D += !f;
    c.have_model = true;
    c.params_dirty = true;
    c.have_nl = false;
    c.jl.valid = false;
    smooth_layout(c);
    c.forces_valid = false;
    FS_HIP(hipStreamSynchronize(c.stream));
  });
}

int fs_set_weights(fs_ctx *ctx, double energy_weight, double stress_weight) {
  return guarded(ctx, [&](Context &c) {
    FS_REQUIRE(std::isfinite(energy_weight) && std::isfinite(stress_weight), FS_ERR_ARG,
               "weights must be finite");
    c.energy_weight = energy_weight;
    c.stress_weight = stress_weight;
    if (c.C > 0) compute_rows(c);
    FS_HIP(hipStreamSynchronize(c.stream));
  });
}

// curvature rows of every table with a free knot and n >= 3, whatever the
// weight (model_smoothness_count, smoothness.hpp:17-23)
int fs_smoothness_count(fs_ctx *ctx, int64_t *count) {
  return guarded(ctx, [&](Context &c) {
    c.fail_if_no_model();
    FS_REQUIRE(count, FS_ERR_ARG, "null output");
    int64_t n = 0;
    if (c.kind != FS_SW)
      for (int t = 0; t < c.ntab; ++t) {
        const TableDesc &T = c.h_tabs[t];
        bool any_free = false;
        for (int k = 0; k < T.n; ++k) any_free |= c.h_col_of[T.off + k] >= 0;
        if (any_free && T.n >= 3) n += T.n - 2;
      }
    *count = n;
  });
}

// model_write_smoothness (smoothness.hpp:26-31; spline.cpp:205-213) at the
// current parameters
int fs_write_smoothness(fs_ctx *ctx, double weight, double *dst) {
  return guarded(ctx, [&](Context &c) {
    c.fail_if_no_model();
    if (c.kind == FS_SW) return;
    int64_t row = 0;
    for (int t = 0; t < c.ntab; ++t) {
      const TableDesc &T = c.h_tabs[t];
      bool any_free = false;
      for (int k = 0; k < T.n; ++k) any_free |= c.h_col_of[T.off + k] >= 0;
      if (!any_free || T.n < 3) continue;
      FS_REQUIRE(dst, FS_ERR_ARG, "null output");
      const double *y = c.h_y.data() + T.off;
      for (int k = 1; k + 1 < T.n; ++k) dst[row++] = weight * (y[k - 1] - 2.0 * y[k] + y[k + 1]);
    }
  });
}

int fs_set_smooth_weight(fs_ctx *ctx, double weight) {
  return guarded(ctx, [&](Context &c) {
    FS_REQUIRE(std::isfinite(weight) && weight >= 0.0, FS_ERR_ARG,
               "smooth weight must be finite and >= 0");
    c.smooth_weight = weight;
    smooth_layout(c);
    FS_HIP(hipStreamSynchronize(c.stream));
  });
}

int fs_param_count(fs_ctx *ctx, int32_t *count) {
  return guarded(ctx, [&](Context &c) {
    c.fail_if_no_model();
    FS_REQUIRE(count, FS_ERR_ARG, "null output");
    *count = c.D;
  });
}

int fs_gather_params(fs_ctx *ctx, double *x) {
  return guarded(ctx, [&](Context &c) {
    c.fail_if_no_model();
    FS_REQUIRE(x || c.D == 0, FS_ERR_ARG, "null output");
    int o = 0;
    if (c.kind == FS_SW) {
      for (size_t k = 0; k < c.h_sw.size(); ++k)
        if (!c.h_sw_fixed[k]) x[o++] = c.h_sw[k];
      for (size_t k = 0; k < c.h_lam.size(); ++k)
        if (!c.h_lam_fixed[k]) x[o++] = c.h_lam[k];
      return;
    }
    for (int k = 0; k < c.D; ++k) x[k] = c.h_y[c.h_knot_of[k]];
  });
}

int fs_scatter_params(fs_ctx *ctx, const double *x) {
  return guarded(ctx, [&](Context &c) {
    c.fail_if_no_model();
    FS_REQUIRE(x || c.D == 0, FS_ERR_ARG, "null input");
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
  });
}

int fs_gather_bounds(fs_ctx *ctx, double *lo, double *hi) {
  return guarded(ctx, [&](Context &c) {
    c.fail_if_no_model();
    FS_REQUIRE((lo && hi) || c.D == 0, FS_ERR_ARG, "null output");
    // spline knots and SW fields carry no bounds (NoBounds, fit_params.hpp:32-41)
    for (int k = 0; k < c.D; ++k) {
      lo[k] = -INFINITY;
      hi[k] = INFINITY;
    }
  });
}

int fs_max_cutoff(fs_ctx *ctx, double *rcut) {
  return guarded(ctx, [&](Context &c) {
    c.fail_if_no_model();
    FS_REQUIRE(rcut, FS_ERR_ARG, "null output");
    *rcut = max_cutoff(c);
  });
}

int fs_residual_count(fs_ctx *ctx, int64_t *rows) {
  return guarded(ctx, [&](Context &c) {
    c.fail_if_no_dataset();
    FS_REQUIRE(rows, FS_ERR_ARG, "null output");
    *rows = c.R;
  });
}

int fs_row_offsets(fs_ctx *ctx, int64_t *row_off) {
  return guarded(ctx, [&](Context &c) {
    c.fail_if_no_dataset();
    FS_REQUIRE(row_off, FS_ERR_ARG, "null output");
    std::copy(c.h_row_off.begin(), c.h_row_off.end(), row_off);
  });
}

int fs_prepare(fs_ctx *ctx) {
  return guarded(ctx, [&](Context &c) {
    c.have_nl = false;
    ensure_nl(c);
    refresh_params(c);
    FS_HIP(hipStreamSynchronize(c.stream));
  });
}

int fs_neighbor_count(fs_ctx *ctx, int64_t *pairs) {
  return guarded(ctx, [&](Context &c) {
    FS_REQUIRE(c.have_nl, FS_ERR_STATE, "call fs_prepare first");
    FS_REQUIRE(pairs, FS_ERR_ARG, "null output");
    *pairs = c.P;
  });
}

int fs_get_neighbor_list(fs_ctx *ctx, int64_t *pair_off, int32_t *j, int32_t *img, double *d) {
  return guarded(ctx, [&](Context &c) {
    FS_REQUIRE(c.have_nl, FS_ERR_STATE, "call fs_prepare first");
    if (pair_off)
      FS_HIP(hipMemcpyAsync(pair_off, c.pair_off.p, sizeof(int64_t) * (c.N + 1),
                            hipMemcpyDeviceToHost, c.stream));
    std::vector<uint32_t> bits;
    std::vector<int32_t> im;
    if (c.P > 0) {
      if (d)
        FS_HIP(hipMemcpyAsync(d, c.pair_d.p, sizeof(double) * 3 * c.P, hipMemcpyDeviceToHost,
                              c.stream));
      if (j) {
        bits.resize(c.P);
        FS_HIP(hipMemcpyAsync(bits.data(), c.pair_bits.p, sizeof(uint32_t) * c.P,
                              hipMemcpyDeviceToHost, c.stream));
      }
      if (img) {
        im.resize(c.P);
        FS_HIP(hipMemcpyAsync(im.data(), c.pair_img.p, sizeof(int32_t) * c.P,
                              hipMemcpyDeviceToHost, c.stream));
      }
    }
    std::vector<int64_t> po(c.N + 1);
    FS_HIP(hipMemcpyAsync(po.data(), c.pair_off.p, sizeof(int64_t) * (c.N + 1),
                          hipMemcpyDeviceToHost, c.stream));
    FS_HIP(hipStreamSynchronize(c.stream));
    for (int64_t a = 0; a < c.N; ++a) {
      const int64_t a0 = c.h_atom_off[c.h_cfg_of[a]];
      for (int64_t p = po[a]; p < po[a + 1]; ++p) {
        if (j) j[p] = int32_t(a0 + pj_local(bits[p]));
        if (img) {
          img[3 * p] = (im[p] & 1023) - 512;
          img[3 * p + 1] = ((im[p] >> 10) & 1023) - 512;
          img[3 * p + 2] = ((im[p] >> 20) & 1023) - 512;
        }
      }
    }
  });
}

int fs_eval_forces(fs_ctx *ctx, double *forces, double *energy, double *stress9, double *limit) {
  return guarded(ctx, [&](Context &c) {
    evaluate(c);
    if (forces)
      FS_HIP(hipMemcpyAsync(forces, c.forces.p, sizeof(double) * 3 * c.N, hipMemcpyDeviceToHost,
                            c.stream));
    if (energy)
      FS_HIP(hipMemcpyAsync(energy, c.cfg_E.p, sizeof(double) * c.C, hipMemcpyDeviceToHost,
                            c.stream));
    if (stress9)
      FS_HIP(hipMemcpyAsync(stress9, c.cfg_S.p, sizeof(double) * 9 * c.C, hipMemcpyDeviceToHost,
                            c.stream));
    if (limit)
      FS_HIP(hipMemcpyAsync(limit, c.cfg_L.p, sizeof(double) * c.C, hipMemcpyDeviceToHost,
                            c.stream));
    FS_HIP(hipStreamSynchronize(c.stream));
  });
}

int fs_eval_config(fs_ctx *ctx, int32_t config, double *forces, double *energy, double *stress9,
                   double *limit) {
  return guarded(ctx, [&](Context &c) {
    FS_REQUIRE(config >= 0 && config < c.C, FS_ERR_ARG, "configuration index out of range");
    evaluate(c);
    const int64_t a0 = c.h_atom_off[config], na = c.h_atom_off[config + 1] - a0;
    if (forces)
      FS_HIP(hipMemcpyAsync(forces, c.forces.p + 3 * a0, sizeof(double) * 3 * na,
                            hipMemcpyDeviceToHost, c.stream));
    if (energy)
      FS_HIP(hipMemcpyAsync(energy, c.cfg_E.p + config, sizeof(double), hipMemcpyDeviceToHost,
                            c.stream));
    if (stress9)
      FS_HIP(hipMemcpyAsync(stress9, c.cfg_S.p + 9 * config, sizeof(double) * 9,
                            hipMemcpyDeviceToHost, c.stream));
    if (limit)
      FS_HIP(hipMemcpyAsync(limit, c.cfg_L.p + config, sizeof(double), hipMemcpyDeviceToHost,
                            c.stream));
    FS_HIP(hipStreamSynchronize(c.stream));
  });
}

int fs_eval_residual(fs_ctx *ctx, double *F, double *cost) {
  return guarded(ctx, [&](Context &c) {
    evaluate(c);
    if (F)
      FS_HIP(hipMemcpyAsync(F, c.residual.p, sizeof(double) * c.R, hipMemcpyDeviceToHost,
                            c.stream));
    if (cost) {
      launch_cost(c, c.cost.p);
      FS_HIP(hipMemcpyAsync(cost, c.cost.p, sizeof(double), hipMemcpyDeviceToHost, c.stream));
    }
    FS_HIP(hipStreamSynchronize(c.stream));
  });
}

int fs_eval_jacobian(fs_ctx *ctx, float *J) {
  int rc = FS_OK;
  const int g = guarded(ctx, [&](Context &c) { rc = fs_jacobian_impl(c, J); });
  return g != FS_OK ? g : rc;
}

int fs_eval_step(fs_ctx *ctx, const double *x, double *F, float *J, double *cost) {
  int rc = FS_OK;
  const int g = guarded(ctx, [&](Context &c) { rc = fs_step_impl(c, x, F, J, cost); });
  return g != FS_OK ? g : rc;
}

int fs_normal_equations(fs_ctx *ctx, double *JtJ, double *Jtr, double *cost) {
  int rc = FS_OK;
  const int g = guarded(ctx, [&](Context &c) { rc = fs_normal_impl(c, JtJ, Jtr, cost); });
  return g != FS_OK ? g : rc;
}

int fs_normal_equations_noalloc(fs_ctx *ctx, const float *J, const double *F, int64_t R,
                                int32_t D, double *JtJ, double *Jtr, double *cost) {
  int rc = FS_OK;
  const int g =
      guarded(ctx, [&](Context &c) { rc = fs_ne_noalloc_impl(c, J, F, R, D, JtJ, Jtr, cost); });
  return g != FS_OK ? g : rc;
}

int fs_lm_solve(fs_ctx *ctx, const fs_lm_opts *opts, double *x, int32_t *status,
                int32_t *counts, double *cost_hist) {
  return guarded(ctx, [&](Context &c) {
    FS_REQUIRE(x && status, FS_ERR_ARG, "null LM argument");
    c.fail_if_no_dataset();
    c.fail_if_no_model();
    fs_lm_opts o{500, 0, 1e-7, 1e-7, 0.0, 100.0};
    if (opts) o = *opts;
    FS_REQUIRE(o.max_iter >= 0, FS_ERR_ARG, "max_iter must be >= 0");
    ensure_ready(c);
    int st = 0, cnt[3] = {0, 0, 0};
    fs_lm_solve_impl(c, o, x, &st, cnt, cost_hist);
    *status = st;
    if (counts)
      for (int k = 0; k < 3; ++k) counts[k] = cnt[k];
  });
}

int fs_lm_step(fs_ctx *ctx, double lambda, double *dx) {
  return guarded(ctx, [&](Context &c) {
    FS_REQUIRE(dx || c.D == 0, FS_ERR_ARG, "null step");
    FS_REQUIRE(std::isfinite(lambda) && lambda >= 0.0, FS_ERR_ARG, "lambda must be finite and >= 0");
    c.fail_if_no_dataset();
    c.fail_if_no_model();
    ensure_ready(c);
    fs_lm_step_impl(c, lambda, dx);
  });
}

int fs_normal_equations_solve(fs_ctx *ctx, double ridge, double *x) {
  return guarded(ctx, [&](Context &c) {
    FS_REQUIRE(x || c.D == 0, FS_ERR_ARG, "null parameter vector");
    FS_REQUIRE(std::isfinite(ridge), FS_ERR_ARG, "ridge must be finite");
    c.fail_if_no_dataset();
    c.fail_if_no_model();
    ensure_ready(c);
    fs_ne_solve_impl(c, ridge, x);
  });
}

int fs_lm_minimize(fs_ctx *ctx, double *x, int32_t max_iter, double *cost_hist,
                   int32_t *accepted) {
  int rc = FS_OK;
  const int g = guarded(ctx, [&](Context &c) {
    FS_REQUIRE(x && max_iter >= 0, FS_ERR_ARG, "bad LM arguments");
    int acc = 0;
    rc = fs_lm_impl(c, x, max_iter, cost_hist, &acc);
    if (accepted) *accepted = acc;
  });
  return g != FS_OK ? g : rc;
}

int fs_population_cost(fs_ctx *ctx, const double *X, int32_t np, double *cost) {
  int rc = FS_OK;
  const int g = guarded(ctx, [&](Context &c) {
    FS_REQUIRE(X && cost && np > 0, FS_ERR_ARG, "bad population arguments");
    rc = fs_population_impl(c, X, np, cost);
  });
  return g != FS_OK ? g : rc;
}

int fs_de_minimize(fs_ctx *ctx, const fs_de_opts *opts, const double *lo, const double *hi,
                   double *x, double *best_cost, double *hist, int32_t *generations_run) {
  int rc = FS_OK;
  const int g = guarded(ctx, [&](Context &c) {
    FS_REQUIRE(opts && lo && hi && x, FS_ERR_ARG, "bad DE arguments");
    int gens = 0;
    rc = fs_de_impl(c, opts, lo, hi, x, best_cost, hist, &gens);
    if (generations_run) *generations_run = gens;
  });
  return g != FS_OK ? g : rc;
}

int fs_eam_set_gauge(fs_ctx *ctx, double scale, const double *slope, const double *intercept) {
  return guarded(ctx, [&](Context &c) {
    gauge_set(c, scale, slope, intercept);
    FS_HIP(hipStreamSynchronize(c.stream));
  });
}

int fs_eam_get_gauge(fs_ctx *ctx, double *scale, double *slope, double *intercept) {
  return guarded(ctx, [&](Context &c) {
    c.fail_if_no_model();
    if (scale) *scale = c.gA;
    for (int t = 0; t < c.ntypes; ++t) {
      if (slope) slope[t] = c.h_gS[t];
      if (intercept) intercept[t] = c.h_gF0[t];
    }
  });
}

int fs_eam_rho_stats(fs_ctx *ctx, double *rho_min, double *rho_max, double *rho_mean,
                     int64_t *count) {
  return guarded(ctx, [&](Context &c) { gauge_rho_stats(c, rho_min, rho_max, rho_mean, count); });
}

int fs_eam_rescale_rho_axis(fs_ctx *ctx, double *a) {
  return guarded(ctx, [&](Context &c) {
    const double v = gauge_rescale_rho_axis(c);
    if (a) *a = v;
    FS_HIP(hipStreamSynchronize(c.stream));
  });
}

int fs_eam_embed_shift(fs_ctx *ctx, const double *rho_ref) {
  return guarded(ctx, [&](Context &c) {
    gauge_embed_shift(c, rho_ref);
    FS_HIP(hipStreamSynchronize(c.stream));
  });
}

int fs_eam_rescale(fs_ctx *ctx, double *a, double *rho_ref) {
  return guarded(ctx, [&](Context &c) {
    const double v = gauge_rescale_eam(c, rho_ref);
    if (a) *a = v;
    FS_HIP(hipStreamSynchronize(c.stream));
  });
}

int fs_lm_last_timing(fs_ctx *ctx, double *ms) {
  return guarded(ctx, [&](Context &c) {
    FS_REQUIRE(ms, FS_ERR_ARG, "null output");
    for (int k = 0; k < 3; ++k) ms[k] = c.lm_timing[k];
  });
}

int fs_de_last_stats(fs_ctx *ctx, fs_de_stats *out) {
  return guarded(ctx, [&](Context &c) {
    FS_REQUIRE(out, FS_ERR_ARG, "null output");
    *out = c.de_stats;
  });
}

int fs_device_buffer(fs_ctx *ctx, int32_t which, void **ptr, int64_t *bytes) {
  return guarded(ctx, [&](Context &c) {
    FS_REQUIRE(ptr && bytes, FS_ERR_ARG, "null output");
    switch (which) {
    case FS_BUF_RESIDUAL:
      *ptr = c.residual.p;
      *bytes = int64_t(c.residual.bytes());
      break;
    case FS_BUF_JACOBIAN:
      *ptr = c.J.p;
      *bytes = int64_t(c.J.bytes());
      break;
    case FS_BUF_JTJ:
      *ptr = c.JtJ.p;
      *bytes = int64_t(c.JtJ.bytes());
      break;
    case FS_BUF_JTR:
      *ptr = c.Jtr.p;
      *bytes = int64_t(c.Jtr.bytes());
      break;
    default:
      throw Error(FS_ERR_ARG, "unknown buffer");
    }
  });
}

int fs_synchronize(fs_ctx *ctx) {
  return guarded(ctx, [&](Context &c) { FS_HIP(hipStreamSynchronize(c.stream)); });
}

int fs_set_launch(fs_ctx *ctx, int32_t lanes_per_atom, int32_t blocks_per_cu) {
  return guarded(ctx, [&](Context &c) {
    FS_REQUIRE(lanes_per_atom == 4 || lanes_per_atom == 8 || lanes_per_atom == 16 ||
                   lanes_per_atom == 32 || lanes_per_atom == 64,
               FS_ERR_ARG, "lanes_per_atom must be 4, 8, 16, 32 or 64");
    FS_REQUIRE(blocks_per_cu >= 1 && blocks_per_cu <= 64, FS_ERR_ARG, "blocks_per_cu in [1, 64]");
    c.lanes_per_atom = lanes_per_atom;
    c.blocks_per_cu = blocks_per_cu;
    if (c.C > 0) build_tiles(c);
    c.forces_valid = false;
    FS_HIP(hipStreamSynchronize(c.stream));
  });
}

int fs_set_jacobian_tile(fs_ctx *ctx, int32_t atoms) {
  return guarded(ctx, [&](Context &c) {
    FS_REQUIRE(atoms == 0 || atoms == 4 || atoms == 8 || atoms == 16 || atoms == 32 || atoms == 64,
               FS_ERR_ARG, "Jacobian tile must be 0 (from the atom count) or 4, 8, 16, 32 or 64 atoms");
    c.jac_tile = atoms;
    c.jac_tile_fixed = atoms != 0;
    c.njt = 0;
  });
}

int fs_set_jacobian_form(fs_ctx *ctx, int32_t form) {
  return guarded(ctx, [&](Context &c) {
    FS_REQUIRE(form >= 0 && form <= 2, FS_ERR_ARG, "Jacobian form must be 0, 1 or 2");
    c.jac_form = form;
  });
}

int fs_set_population_form(fs_ctx *ctx, int32_t form) {
  return guarded(ctx, [&](Context &c) {
    FS_REQUIRE(form == 0 || form == 1, FS_ERR_ARG, "population form must be 0 or 1");
    c.pop_ml_allowed = form == 0;
  });
}

int fs_get_launch(fs_ctx *ctx, int32_t *lanes_per_atom, int32_t *blocks_per_cu,
                  int32_t *jacobian_tile, int32_t *jacobian_blocks_per_cu) {
  return guarded(ctx, [&](Context &c) {
    if (lanes_per_atom) *lanes_per_atom = c.lanes_per_atom;
    if (blocks_per_cu) *blocks_per_cu = c.blocks_per_cu;
    if (jacobian_tile) *jacobian_tile = c.njt > 0 ? c.jt_atoms : c.jac_tile;
    if (jacobian_blocks_per_cu) *jacobian_blocks_per_cu = c.jac_blocks_per_cu;
  });
}

int fs_autotune(fs_ctx *ctx, int32_t *lanes_per_atom) {
  int rc = FS_OK;
  const int g = guarded(ctx, [&](Context &c) {
    rc = autotune_impl(c);
    if (lanes_per_atom) *lanes_per_atom = c.lanes_per_atom;
  });
  return g != FS_OK ? g : rc;
}

int fs_mlp_forward(fs_ctx *ctx, int32_t n, int32_t d, int32_t h, const float *params,
                   const float *X, float *y) {
  return guarded(ctx, [&](Context &c) {
    FS_REQUIRE(n > 0 && d > 0 && h > 0 && params && X && y, FS_ERR_ARG, "bad MLP arguments");
    Mlp m;
    mlp_init(m, d, h);
    FS_HIP(hipMemcpyAsync(m.P.p, params, sizeof(float) * m.n_params(), hipMemcpyHostToDevice,
                          c.stream));
    DBuf<float> dX, dy;
    dX.alloc(size_t(n) * d);
    dy.alloc(n);
    FS_HIP(hipMemcpyAsync(dX.p, X, sizeof(float) * size_t(n) * d, hipMemcpyHostToDevice, c.stream));
    mlp_forward(m, c.stream, n, dX.p, dy.p);
    FS_HIP(hipMemcpyAsync(y, dy.p, sizeof(float) * n, hipMemcpyDeviceToHost, c.stream));
    FS_HIP(hipStreamSynchronize(c.stream));
  });
}

int fs_mlp_loss_grad(fs_ctx *ctx, int32_t n, int32_t d, int32_t h, const float *params,
                     const float *X, const float *t, float *grad, float *loss) {
  return guarded(ctx, [&](Context &c) {
    FS_REQUIRE(n > 0 && d > 0 && h > 0 && params && X && t && grad && loss, FS_ERR_ARG,
               "bad MLP arguments");
    Mlp m;
    mlp_init(m, d, h);
    FS_HIP(hipMemcpyAsync(m.P.p, params, sizeof(float) * m.n_params(), hipMemcpyHostToDevice,
                          c.stream));
    DBuf<float> dX, dt;
    dX.alloc(size_t(n) * d);
    dt.alloc(n);
    FS_HIP(hipMemcpyAsync(dX.p, X, sizeof(float) * size_t(n) * d, hipMemcpyHostToDevice, c.stream));
    FS_HIP(hipMemcpyAsync(dt.p, t, sizeof(float) * n, hipMemcpyHostToDevice, c.stream));
    mlp_backward(m, c.stream, n, dX.p, dt.p);
    FS_HIP(hipMemcpyAsync(grad, m.G.p, sizeof(float) * m.n_params(), hipMemcpyDeviceToHost,
                          c.stream));
    FS_HIP(hipMemcpyAsync(loss, m.loss.p, sizeof(float), hipMemcpyDeviceToHost, c.stream));
    FS_HIP(hipStreamSynchronize(c.stream));
  });
}

int fs_time_step(fs_ctx *ctx, int32_t what, int32_t reps, double *ms_per_rep) {
  return guarded(ctx, [&](Context &c) {
    FS_REQUIRE(reps > 0 && ms_per_rep, FS_ERR_ARG, "bad timing arguments");
    *ms_per_rep = time_step_impl(c, what, reps);
  });
}

} // extern "C"
