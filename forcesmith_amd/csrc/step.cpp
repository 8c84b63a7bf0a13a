// One residual + Jacobian step replayed from a captured hipGraph
// (fs_eval_step): ForcesmithFunctor::operator() and df at one parameter
// vector (src/optimization/forcesmith_functor.cpp:221-292), with the
// parameters taken from a device buffer inside the graph.
//
// The first call at a given layout runs the step eagerly (it sizes every
// buffer and builds the cached column layouts, tiles and pair order), then
// captures scatter + slopes -> residual -> Jacobian -> cost on the context
// stream. Later calls copy x into the device buffer and launch the graph. The
// graph's kernels hold buffer addresses and launch shapes by value, so the
// capture is keyed on both: any reallocation or shape change (dataset, model,
// weights, launch shape, tile size) recaptures. Stillinger-Weber runs eagerly
// (its Jacobian uploads the column maps of the current parameter set).
#include "common.hpp"

#include <cstdint>
#include <vector>

namespace fsa {

void evaluate_for_solver(Context &c);
void ensure_ready(Context &c);

namespace {

template <class T> uintptr_t ad(const DBuf<T> &b) { return reinterpret_cast<uintptr_t>(b.p); }

// everything the captured launches take by value
std::vector<uintptr_t> step_key(const Context &c) {
  std::vector<uintptr_t> k = {
      uintptr_t(c.C), uintptr_t(c.N), uintptr_t(c.P), uintptr_t(c.D), uintptr_t(c.R),
      uintptr_t(c.R_cfg), uintptr_t(c.n_smooth), uintptr_t(c.ntab), uintptr_t(c.nknots),
      uintptr_t(c.nivl), uintptr_t(c.ntiles), uintptr_t(c.njt), uintptr_t(c.jt_atoms),
      uintptr_t(c.lanes_per_atom), uintptr_t(c.blocks_per_cu), uintptr_t(c.jac_blocks_per_cu),
      uintptr_t(c.precision), uintptr_t(c.kind), uintptr_t(c.ntypes), uintptr_t(c.jac_form),
      // cached layouts and sorted pairs the eager step rebuilds after a model
      // upload or a neighbour build (the captured launches do not)
      uintptr_t(c.jl.valid), uintptr_t(c.jsort_valid), uintptr_t(c.ks_valid)};
  for (double v : {c.energy_weight, c.stress_weight, c.smooth_weight}) {
    uint64_t u;
    std::memcpy(&u, &v, 8);
    k.push_back(uintptr_t(u));
  }
  const uintptr_t ptrs[] = {
      ad(c.atom_off), ad(c.row_off), ad(c.types), ad(c.cfg_of), ad(c.pos), ad(c.ref_force),
      ad(c.ref_energy), ad(c.ref_stress), ad(c.box), ad(c.inv), ad(c.vol), ad(c.tile_cfg),
      ad(c.tile_a0), ad(c.tile_na), ad(c.cfg_tile0), ad(c.tabs), ad(c.tab_x), ad(c.tab_y),
      ad(c.tab_s), ad(c.tab_S), ad(c.ivl32), ad(c.ivl64), ad(c.col_of), ad(c.knot_of),
      ad(c.smooth_rows), ad(c.pair_off), ad(c.pair_rec), ad(c.pair_bits), ad(c.pair_d),
      ad(c.pair_r64), ad(c.rho), ad(c.Fp), ad(c.Fpp), ad(c.e_atom), ad(c.lim_atom),
      ad(c.clamp_over), ad(c.Fp32), ad(c.mu), ad(c.lamb), ad(c.mu32), ad(c.lamb32), ad(c.forces),
      ad(c.tile_part), ad(c.cfg_E), ad(c.cfg_S), ad(c.cfg_L), ad(c.residual), ad(c.cost),
      ad(c.cost_part), ad(c.J), ad(c.H), ad(c.jac_part), ad(c.jt_cfg), ad(c.jt_a0), ad(c.jt_na),
      ad(c.jt_ct0), ad(c.jcols), ad(c.jslots), ad(c.jrec), ad(c.jr), ad(c.Hs), ad(c.emb_d),
      ad(c.emb_v), ad(c.emb_k), ad(c.adp_tan), ad(c.step_x), ad(c.gauge), ad(c.ks_pair),
      ad(c.ks_dens), ad(c.ks_cfg)};
  k.insert(k.end(), std::begin(ptrs), std::end(ptrs));
  return k;
}

void drop_graph(Context &c) {
  if (c.step_exec) (void)hipGraphExecDestroy(c.step_exec);
  if (c.step_graph) (void)hipGraphDestroy(c.step_graph);
  c.step_exec = nullptr;
  c.step_graph = nullptr;
  c.step_key.clear();
}

void step_launches(Context &c) {
  launch_scatter_slopes(c, c.step_x.p);
  launch_eval(c, true);
  launch_jacobian(c);
  launch_cost(c, c.cost.p);
}

} // namespace

void destroy_step_graph(Context &c) { drop_graph(c); }

int fs_step_impl(Context &c, const double *x, double *F, float *J, double *cost) {
  c.fail_if_no_dataset();
  c.fail_if_no_model();
  FS_REQUIRE(x || c.D == 0, FS_ERR_ARG, "null parameter vector");
  // host mirror of the parameters (gather_params, later eager evaluations)
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
  if (c.kind == FS_SW || c.D == 0) {
    // the SW pair-term pass of the residual also forms the Jacobian's tangents
    c.sw_tan_next = c.kind == FS_SW && c.D > 0;
    try {
      evaluate_for_solver(c);
    } catch (...) {
      c.sw_tan_next = false;
      throw;
    }
    c.sw_tan_next = false;
    launch_jacobian(c);
    launch_cost(c, c.cost.p);
  } else {
    c.step_x.alloc(size_t(c.D));
    if (!c.step_exec || c.step_key != step_key(c)) {
      drop_graph(c);
      // eager step: sizes every buffer, builds the cached layouts
      ensure_ready(c);
      evaluate_for_solver(c);
      launch_jacobian(c);
      launch_cost(c, c.cost.p);
      FS_HIP(hipStreamSynchronize(c.stream));
      FS_HIP(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
      try {
        step_launches(c);
      } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(c.stream, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
      }
      FS_HIP(hipStreamEndCapture(c.stream, &c.step_graph));
      FS_HIP(hipGraphInstantiate(&c.step_exec, c.step_graph, nullptr, nullptr, 0));
      c.step_key = step_key(c);
    }
    FS_HIP(hipMemcpyAsync(c.step_x.p, x, sizeof(double) * c.D, hipMemcpyHostToDevice, c.stream));
    FS_HIP(hipGraphLaunch(c.step_exec, c.stream));
    c.params_dirty = false;
    c.forces_valid = true;
  }
  if (F)
    FS_HIP(hipMemcpyAsync(F, c.residual.p, sizeof(double) * c.R, hipMemcpyDeviceToHost, c.stream));
  if (J)
    FS_HIP(hipMemcpyAsync(J, c.J.p, sizeof(float) * size_t(c.R) * c.D, hipMemcpyDeviceToHost,
                          c.stream));
  if (cost)
    FS_HIP(hipMemcpyAsync(cost, c.cost.p, sizeof(double), hipMemcpyDeviceToHost, c.stream));
  // x is copied from pageable memory, so it has been read when the call
  // returns; with no host output the step stays queued (fs_synchronize)
  if (F || J || cost) FS_HIP(hipStreamSynchronize(c.stream));
  return FS_OK;
}

} // namespace fsa
