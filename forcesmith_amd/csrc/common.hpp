// Internal types shared by the host runtime and the HIP kernels of
// libforcesmith_amd. Host code and kernels are compiled for gfx950 only.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/forcesmith_amd.h"

namespace fsa {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define FS_HIP(call)                                                            \
  do {                                                                          \
    hipError_t e_ = (call);                                                     \
    if (e_ != hipSuccess)                                                       \
      throw ::fsa::Error(FS_ERR_HIP, std::string(#call) + ": " +               \
                                         hipGetErrorString(e_));                \
  } while (0)

#define FS_REQUIRE(cond, code, msg)                                            \
  do {                                                                          \
    if (!(cond)) throw ::fsa::Error((code), (msg));                             \
  } while (0)

// Device buffer owned by a context; grows, never shrinks.
template <class T> struct DBuf {
  T *p = nullptr;
  size_t cap = 0;
  size_t n = 0;
  DBuf() = default;
  DBuf(const DBuf &) = delete;
  DBuf &operator=(const DBuf &) = delete;
  ~DBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = n = 0;
  }
  T *alloc(size_t count) {
    if (count > cap) {
      release();
      if (count) {
        hipError_t e = hipMalloc(reinterpret_cast<void **>(&p), count * sizeof(T));
        if (e != hipSuccess) {
          p = nullptr;
          throw Error(FS_ERR_NOMEM, "hipMalloc of " + std::to_string(count * sizeof(T)) +
                                        " bytes failed: " + hipGetErrorString(e));
        }
      }
      cap = count;
    }
    n = count;
    return p;
  }
  size_t bytes() const { return n * sizeof(T); }
};

// ---------------------------------------------------------------------------
// Pair record flags. A pair is stored as float4 {dx, dy, dz, bits} (mixed
// precision) plus a uint32 word 'bits' and double3 d (exact path):
//   bits[0:22)  neighbour index local to its configuration
//   bits[22:25) neighbour type (ntypes <= 8)
//   bits[25..]  gates computed once from the FP64 distance at prepare time,
//               the counterpart of the reference's site priming
//               (prime_site = in_range ? prepare_site : none, kernels.hpp:14-16)
// ---------------------------------------------------------------------------
constexpr uint32_t kJBits = 22;
constexpr uint32_t kJMask = (1u << kJBits) - 1u;
constexpr uint32_t kTShift = 22;
constexpr uint32_t kTMask = 7u;
constexpr uint32_t kFlagPhi = 1u << 25;   // r inside span of pair table (ti,tj)
constexpr uint32_t kFlagGj = 1u << 26;    // r inside span of density table tj
constexpr uint32_t kFlagGi = 1u << 27;    // r inside span of density table ti
constexpr uint32_t kFlagDip = 1u << 28;   // ADP dipole table
constexpr uint32_t kFlagQuad = 1u << 29;  // ADP quadrupole table
constexpr uint32_t kFlagEdge = 1u << 30;  // some role has r exactly at a span end
constexpr uint32_t kFlagTiny = 1u << 31;  // r < 1e-14: skipped by the force pass

__host__ __device__ inline uint32_t pj_local(uint32_t b) { return b & kJMask; }
__host__ __device__ inline int pj_type(uint32_t b) { return int((b >> kTShift) & kTMask); }

// Uniform-grid table descriptor. x_k = x0 + k*h evaluated on the host the way
// the reference builds its grid; the exact x array lives in tab_x.
struct TableDesc {
  int32_t n;      // knots
  int32_t off;    // first knot in the flat knot arrays
  int32_t ivl;    // first interval in the packed interval arrays
  int32_t linear; // n < 4: piecewise linear (no slopes)
  double x0, xn, h, inv_h;
  float x0f, inv_hf, hf, inv_hlo; // FP32 grid; inv_hf + inv_hlo carries 1/h to ~1e-14
};

struct CfgGrid {
  int32_t nc[3];  // cells per axis
  int32_t K[3];   // cell search half-width per axis
  int32_t s[3];   // reference image shells per axis
  int32_t cell0;  // first cell of this configuration
};

constexpr int kMaxTypes = 8;

// Knot tables are staged in LDS while they fit this budget per block; larger
// tables (hundreds of knots, as the reference's model writer resamples to)
// are read from global memory, where they stay L2-resident.
constexpr size_t kTableStageBytes = 48 * 1024;

struct Context;
struct Mlp;
struct SurrogateState;

// ---- kernel launchers (defined in the .hip translation units) -------------
void launch_neighbor_build(Context &c);
void launch_pair_flags(Context &c);
void launch_slopes(Context &c);
// the free knots from the device vector d_x, then the slopes (one launch)
void launch_scatter_slopes(Context &c, const double *d_x);
void launch_scatter_params(Context &c, const double *d_x);
void launch_eval(Context &c, bool want_forces);
void launch_cost(Context &c, double *d_out);
// sum of squares of v[n] into d_out[0] through part[kSumsqBlocks] block
// partials (every partial written, fixed order)
constexpr int kSumsqBlocks = 512;
void launch_sumsq(hipStream_t st, int64_t n, const double *v, double *part, double *d_out);
// the no-allocation scratch of the normal equations: per-device address of
// the static device array, looked up once per device by fs_ctx_create
void noalloc_scratch_init(int device);
void launch_smooth_residual(Context &c);
void launch_smooth_jacobian(Context &c);
void launch_jacobian(Context &c);
void launch_normal_equations(Context &c);
int lm_minimize_device(Context &c, double *x, int max_iter, double *hist);
// batched |F(x_m)|^2 over host parameter vectors X[np][D]; false when the
// model / precision has no batched kernels (the caller then loops members)
bool population_cost_batched(Context &c, const double *X, int np, double *cost);
bool population_supported(const Context &c);
void population_prepare(Context &c, int np);
void population_launch(Context &c, const double *d_X, int np, double *d_cost);
double population_tune_launch(Context &c, const double *d_X, int np, double *d_cost);
void launch_sw_rev(Context &c);
void launch_sw_eval(Context &c);
// per-pair SW scalars {v2, v2', h, h', dh/dgamma, dh'/dgamma, dh/da2, dh'/da2}
void launch_sw_pair_terms(Context &c, bool tangents);
// the gauge's device copy from gA / h_gS / h_gF0 (allocated on first use)
constexpr int kGaugeLen = 1 + 2 * kMaxTypes;
void upload_gauge(Context &c);
void reset_gauge(Context &c);
// rescale_eam and its steps (gauge.cpp); per-type arrays of ntypes entries
void gauge_set(Context &c, double A, const double *S, const double *F0);
void gauge_rho_stats(Context &c, double *rmin, double *rmax, double *rmean, int64_t *count);
double gauge_rescale_rho_axis(Context &c);
void gauge_embed_shift(Context &c, const double *rho_ref);
double gauge_rescale_eam(Context &c, double *rho_ref);

struct Context;
// the residual kernels' descriptors plus both interval forms fit the budget
bool stage_intervals(const Context &c);

struct Context {
  std::mutex lock;
  std::string err;
  int refs = 1;
  int device = 0;
  int precision = FS_MIXED;
  hipStream_t stream = nullptr;
  int cu_count = 256;
  int lanes_per_atom = 16;
  int blocks_per_cu = 8;
  // atoms per Jacobian tile: 0 picks it from the atom count, else a power of
  // two in [4, 64] (fs_set_jacobian_tile, fs_autotune)
  int jac_tile = 0;
  bool jac_tile_fixed = false; // set by hand: fs_autotune keeps it
  // form of the density/embedding-column pass (fs_set_jacobian_form):
  // 0 knot bins while they fit, else windows; 1 windows; 2 knot bins
  int jac_form = 0;
  // persistent-grid size of the Jacobian passes (blocks per CU; each tile is
  // one block's work whatever the grid, so the result does not depend on it)
  int jac_blocks_per_cu = 8;

  // ---- dataset (host mirror + device) ------------------------------------
  int C = 0;
  int64_t N = 0;
  std::vector<int64_t> h_atom_off;
  std::vector<int32_t> h_types, h_cfg_of;
  std::vector<double> h_pos, h_ref_force, h_cell, h_ref_energy, h_ref_stress;
  std::vector<double> h_box, h_inv, h_vol;
  DBuf<int64_t> atom_off, row_off;
  DBuf<int32_t> types, cfg_of;
  DBuf<double> pos, ref_force, ref_energy, ref_stress, box, inv, vol;

  // tiles: a tile is up to atoms_per_tile consecutive atoms of one config
  std::vector<int32_t> h_tile_cfg, h_tile_a0, h_tile_na, h_cfg_tile0;
  DBuf<int32_t> tile_cfg, tile_a0, tile_na, cfg_tile0;
  int ntiles = 0;
  int tile_atoms = 0;

  // ---- model --------------------------------------------------------------
  bool have_model = false;
  int kind = FS_EAM;
  int ntypes = 1, npair = 1, ntab = 0;
  std::vector<TableDesc> h_tabs;
  std::vector<double> h_x, h_y; // knot abscissae / values (all tables)
  std::vector<uint8_t> h_fixed;
  std::vector<int32_t> h_col_of;   // knot -> free column or -1
  std::vector<int32_t> h_knot_of;  // free column -> knot
  int D = 0, nknots = 0, nivl = 0;
  DBuf<TableDesc> tabs;
  DBuf<double> tab_x, tab_y, tab_s, tab_S; // S: n*5 band of ds/dy per table
  DBuf<float> tab_S32; // FP32 copy of tab_S for Jacobian passes that do not stage it
  DBuf<float4> ivl32; // per interval k: power form {a, b, c, d} as FP32 hi at 2k, FP32 lo at 2k+1
  DBuf<double> ivl64; // per interval k: the same power form in FP64 at 4k .. 4k+3
  DBuf<int32_t> col_of, knot_of;
  // EAM gauge of rescale_eam (src/core/rescale.cpp:82-158): density scale A,
  // per-type embedding slope S_t and intercept F0_t; the device copy is
  // {A, S[kMaxTypes], F0[kMaxTypes]}, the identity {1, 0.., 0..} by default
  // and for every other model kind (gauge.cpp)
  double gA = 1.0;
  double h_gS[kMaxTypes] = {}, h_gF0[kMaxTypes] = {};
  double h_gauge_img[1 + 2 * kMaxTypes] = {}; // source of the last upload
  DBuf<double> gauge;
  // SW
  std::vector<double> h_sw, h_lam;
  std::vector<uint8_t> h_sw_fixed, h_lam_fixed;
  DBuf<double> sw, lam;
  double sw_nl_rc = -1.0;
  DBuf<double> x_dev;
  bool params_dirty = true;
  double energy_weight = 1.0, stress_weight = 0.0;
  int64_t R = 0; // residual rows: configuration rows, then curvature rows
  // curvature (smoothness) rows w (y[k-1] - 2 y[k] + y[k+1]) after the
  // configuration rows: {first knot, free column of knots k-1, k, k+1 or -1}
  double smooth_weight = 0.0;
  int64_t R_cfg = 0;
  int n_smooth = 0;
  DBuf<int4> smooth_rows;
  std::vector<int64_t> h_row_off;

  // ---- neighbour list ------------------------------------------------------
  bool have_nl = false;
  double nl_rc = -1.0;
  int64_t P = 0;
  std::vector<CfgGrid> h_grid;
  DBuf<CfgGrid> grid;
  DBuf<int64_t> pair_off;
  DBuf<float4> pair_rec;
  DBuf<uint32_t> pair_bits;
  DBuf<double> pair_d;
  DBuf<double> pair_r64; // FP64 distance per pair (energy values, density)
  DBuf<int32_t> pair_img;
  DBuf<int32_t> rev;         // reverse pair index (SW gather)
  DBuf<double> wpos;
  DBuf<int32_t> cell_of, cell_count, cell_start, cell_atoms, counts;
  DBuf<uint64_t> keys;
  DBuf<uint8_t> scan_tmp;

  // ---- evaluation state ----------------------------------------------------
  DBuf<double> rho, Fp, Fpp, e_atom, lim_atom, clamp_over;
  DBuf<float> Fp32;
  DBuf<double> mu, lamb; // ADP per-atom dipole (3) and quadrupole (6)
  // their FP32 copies for the mixed force pass: mu32[N] (x, y, z, 0),
  // lamb32[2N] (xx, yy, zz, xy | xz, yz, 0, 0)
  DBuf<float4> mu32, lamb32;
  DBuf<double> forces, tile_part, cfg_E, cfg_S, cfg_L;
  DBuf<double> residual, cost, cost_part;
  DBuf<double> pairbuf; // SW triplet forces per pair
  DBuf<double> sw_pv;   // SW per-pair scalars, 8 per pair
  // SW triplet records, 8 per pair: unit vector, 1/r, h, h', dh/dgamma, dh/da2
  DBuf<double> sw_tri;
  // SW parameter / pair-list generation (bumped on every upload of the SW
  // parameters and every neighbour build) and the one sw_pv was computed for
  uint64_t sw_gen = 1, sw_pv_gen = 0, sw_pt2_gen = 0;
  bool sw_tan_next = false; // the next SW pair-term pass also forms the two-body tangents
  DBuf<double2> sw_pt2; // SW two-body tangents, [6 fields][pairs] of (tv, tdv / r)
  bool forces_valid = false;
  // Jacobian
  DBuf<float> J;
  DBuf<float> H;        // per-atom d rho / d knot (density columns) scaled
  DBuf<double> jac_part; // per-tile energy/stress/limit partial columns
  DBuf<double> JtJ, Jtr, jtj_part;
  // Jacobian work buffers: tiles of <= 64 atoms, column descriptors, per-atom
  // density histograms and embedding windows, SW tangents
  DBuf<int32_t> jt_cfg, jt_a0, jt_na, jt_ct0;
  int njt = 0;
  int jt_atoms = 0; // atoms per tile of the current Jacobian tiles
  DBuf<int4> jcols, jslots;
  // column descriptors / lane-slot maps derived from the model (jcols, jslots
  // hold them on the device); rebuilt after every model change
  struct JacLayout {
    bool valid = false;
    int Dd = 0, dens_c0 = 0, gi_nf = 0, NSP = 0, NSD = 0, NSU = 0, NSW = 0, Ddip = 0, Dquad = 0;
    // slots of each kind over all launches (a multiple of the per-launch width)
    int TSP = 0, TSD = 0, TSU = 0, TSW = 0;
  } jl;
  DBuf<float4> jrec; // pair records sorted by (partner type, r) per atom
  DBuf<double> jr;   // their FP64 distances
  bool jsort_valid = false;
  // knot sums of the geometry (jac.hip, k_knot_sums): per-atom pair-table
  // force sums and density value sums, per-configuration pair-table energy
  // and virial sums; built once per neighbour list and model grid, folded
  // with the current slope band by every Jacobian
  bool ks_valid = false;
  int ks_lp = 0, ks_ld = 0, ks_kp = 0; // entries per atom (pair, density), pair-table knots
  DBuf<float> ks_pair, ks_dens;
  DBuf<double> ks_tile, ks_cfg;
  DBuf<float> Hs, emb_d, emb_v;
  DBuf<int32_t> emb_k, sw_csw, sw_clam, sw_tmap;
  DBuf<double> sw_own, sw_ptan, sw_etan;
  DBuf<float> adp_tan; // ADP Jacobian tangents [N][10][D]
  DBuf<double> lm_M, lm_dx, lm_scal;
  DBuf<float> ne_W;   // normal equations of > 256 columns: two column groups of J
  DBuf<double> ne_M;  // and their product
  // batched population cost: member tables, embedding derivatives, partials
  DBuf<double> pop_X, pop_y, pop_s, pop_S, pop_i64, pop_Fp, pop_partE, pop_partF, pop_cmc, pop_cost;
  DBuf<float4> pop_i32;
  DBuf<float> pop_Fp32;
  DBuf<int32_t> pop_grp;
  std::vector<int32_t> h_pop_grp;
  int pop_grp_tiles = -1;
  // members per block of the population passes (1, 2 or 4); 0 until the DE
  // driver or fs_autotune has timed them (1 meanwhile)
  int pop_members = 0;
  // member-lane population passes (pair and density tables on one grid):
  // wave tiles (up to 32 atoms of one type in one configuration, type-major),
  // ranges of them per work item, the configurations' tile lists; members
  // packed 64 to a group, member-minor
  bool pop_ml = false, pop_ml_allowed = true;
  uint64_t pop_ml_key[4] = {~0ull, ~0ull, ~0ull, ~0ull};
  int pop_ml_nwt = 0, pop_ml_nrange = 0, pop_ml_n = 0, pop_ml_xoff = 0, pop_ml_nemb = 0;
  int pop_ml_ng = 0; // member groups allocated
  size_t pop_ml_lds = 0; // LDS of the member-lane density pass (0: the passes do not apply)
  DBuf<int32_t> pop_ml_wt_atom, pop_ml_wt_off, pop_ml_wt_cfg, pop_ml_rng_type, pop_ml_rng_wt0;
  DBuf<int32_t> pop_ml_cfg_off, pop_ml_cfg_wt;
  DBuf<double2> pop_ml_ys;
  DBuf<double> pop_ml_Fp, pop_ml_pE, pop_ml_pF;
  uint64_t nl_gen = 0; // bumped by every neighbour build
  fs_de_stats de_stats{};          // of the last fs_de_minimize call
  // of the last fs_lm_solve: whole solve, device calls, host algebra (ms)
  double lm_timing[3] = {0.0, 0.0, 0.0};
  // the residual + Jacobian step captured as a graph (step.cpp): parameters
  // from step_x, keyed on the buffer addresses and launch shapes it holds
  hipGraph_t step_graph = nullptr;
  hipGraphExec_t step_exec = nullptr;
  std::vector<uintptr_t> step_key;
  DBuf<double> step_x;
  Mlp *mlp = nullptr;              // owned; see mlp.hpp
  SurrogateState *surr = nullptr;  // owned; see de.cpp
  // per-kernel timing of one Jacobian stage (HIP events, context stream):
  // stage 0 = the density/embedding pass (EAM) or the tangent kernels (ADP),
  // stage 1 = the pair-column pass (EAM)
  hipEvent_t jev0 = nullptr, jev1 = nullptr;
  bool time_jac_main = false;
  int time_which = 0;
  double jac_main_ms = 0.0;
  int jac_main_n = 0;

  ~Context();
  void time_main_begin(int stage = 0) {
    if (time_jac_main && stage == time_which) FS_HIP(hipEventRecord(jev0, stream));
  }
  void time_main_end(int stage = 0) {
    if (!time_jac_main || stage != time_which) return;
    FS_HIP(hipEventRecord(jev1, stream));
    FS_HIP(hipEventSynchronize(jev1));
    float ms = 0.f;
    FS_HIP(hipEventElapsedTime(&ms, jev0, jev1));
    jac_main_ms += ms;
    ++jac_main_n;
  }
  void fail_if_no_dataset() const {
    FS_REQUIRE(C > 0, FS_ERR_STATE, "no dataset uploaded");
  }
  void fail_if_no_model() const { FS_REQUIRE(have_model, FS_ERR_STATE, "no model set"); }
};

} // namespace fsa
