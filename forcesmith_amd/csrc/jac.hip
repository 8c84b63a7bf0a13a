// Analytic Jacobian dF/dx on the GPU.
//
// Tabulated EAM: per-knot columns through the makima slope chain. The chain
// rule is the one of the reference's analytic-table Jacobian
// (EAMForceCalculator::write_param_jacobian, src/force/eam_force.cpp:108-279;
// add_force_column / add_radial_bond, include/forcesmith/force/
// param_jacobian.hpp:91-121) with spline partials in place of param_grad /
// dderiv_dparam:
//   d(value, deriv)(r)/dy_m = b0[m==k] + b1[m==k+1] + bs0 dS_k/dy_m + bs1 dS_k+1/dy_m
// where b are the Hermite basis coefficients of the interval k holding r and
// dS/dy is the band of the makima slope jacobian refreshed with the knots.
// The reference itself has no such columns for tabulated tables: it falls
// back to central differences (fd_jacobian.hpp:37-49).
//
// Kernels:
//   k_jac_sort   once per neighbour list: each atom's pairs sorted by
//                (partner type, FP32 r), so that every column's pairs form one
//                contiguous window per partner type.
//   k_jac_dens   one wave per atom: H_i[c] = d rho_i / d y_c over the density
//                columns, Hs_i = F''(rho_i) H_i (0 when the density is clamped),
//                and the embedding window of atom i.
//   k_jac_pair,  one wave per atom, lanes own columns (grouped by kind into
//   k_jac_de     slots; pair-table columns, then density and embedding
//                columns): each 64-pair chunk stages per-pair band
//                coefficients in LDS, then each column sums its window; FP32
//                force rows are written, energy / virial / limit partials are
//                reduced per tile in FP64.
//   k_jac_rows   per configuration: energy, stress and limit rows from the
//                tile partials, in tile order (deterministic, no atomics).
// ADP: the pair-table, density and embedding columns run through the same
// passes; k_adp_mom and k_adp_dq add the dipole and quadrupole columns.
// Stillinger-Weber: forward-mode tangents of the reference's pair and triplet
// formulas (stiweb_force.cpp:24-54, kernels.hpp:40-68), one column per grid
// row, with the triplet j/k forces gathered through the reverse pair index.
#include "common.hpp"
#include "spline.hpp"

#include <algorithm>

namespace fsa {

namespace {

constexpr int kCPL = 4;        // columns per lane (D <= 256)
constexpr int kJacTile = 64;   // most atoms per Jacobian tile
constexpr int kJP = 8;         // partial rows: energy, 6 virial, limit
constexpr int kXcds = 8;       // MI355X accelerator complex dies (one L2 each)
enum { CK_PAIR = 0, CK_DENS = 1, CK_EMB = 2, CK_DIP = 3, CK_QUAD = 4, CK_NONE = 7 };

struct ColDesc {
  int32_t tab;  // table
  int32_t m;    // knot index within the table
  int32_t kind; // CK_*
  int32_t dc;   // density-column index (kind == CK_DENS) or -1
};

__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Wave-uniform values the compiler cannot prove uniform (the wave index, and
// loads through pointers it must assume alias the kernel's own stores): taken
// from lane 0 they live in scalar registers, so loops and bounds on them
// branch on the scalar unit instead of masking lanes, and the loads and LDS
// reads of a batch issue back to back.
__device__ inline int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ inline int64_t uni64(int64_t v) {
  const uint64_t u = uint64_t(v);
  const uint32_t lo = uint32_t(__builtin_amdgcn_readfirstlane(int(uint32_t(u))));
  const uint32_t hi = uint32_t(__builtin_amdgcn_readfirstlane(int(uint32_t(u >> 32))));
  return int64_t((uint64_t(hi) << 32) | lo);
}

// one DPP move of a float across lanes (bound_ctrl off: every source lane
// is valid for the patterns used here)
template <int CTRL> __device__ inline float dpp_mov(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}

// sum over the wave, in every lane: within rows of 16 lanes by DPP (swap
// neighbours, swap pairs, mirror the half-rows, mirror the row), then across
// the four rows by two lane permutes
__device__ inline float wave_sum(float v) {
  v += dpp_mov<0xB1>(v);  // quad_perm [1, 0, 3, 2]
  v += dpp_mov<0x4E>(v);  // quad_perm [2, 3, 0, 1]
  v += dpp_mov<0x141>(v); // row_half_mirror
  v += dpp_mov<0x140>(v); // row_mirror
  v += __shfl_xor(v, 16, 64);
  v += __shfl_xor(v, 32, 64);
  return v;
}

// S band of table T at (k, m): d s_k / d y_m, zero outside |m - k| <= 2
__device__ inline float sband(const float *sS, const TableDesc &T, int k, int m) {
  const int q = m - k + 2;
  return (q >= 0 && q < 5 && k >= 0 && k < T.n) ? sS[(T.off + k) * 5 + q] : 0.0f;
}

__device__ inline float band(const float *sS, const TableDesc &T, int k, int m, const float c[4]) {
  float v = c[2] * sband(sS, T, k, m) + c[3] * sband(sS, T, k + 1, m);
  if (m == k) v += c[0];
  if (m == k + 1) v += c[1];
  return v;
}

// interval of a pair table at the FP64 distance and the FP32 offset in it:
// an FP32 distance and 1/h would place every pair of a fine grid (knots a few
// hundredths of an angstrom apart) off by up to u * 6e-8 intervals, a shift
// the derivative basis (slope ~6/h in t) turns into a systematic error that
// survives the cancelling sum over neighbours
__device__ inline int pair_interval(const TableDesc &T, double r, float &t) {
  double td;
  const int k = spl_interval64(T, r, td);
  t = float(td);
  return k;
}

// value / derivative basis of a pair table at offset t of its interval
// (cached-site semantics: linear site only at the span ends or for n < 4
// tables)
__device__ inline void basis_at(const TableDesc &T, float t, float bv[4], float bd[4]) {
  if (T.linear) {
    bv[0] = 1.0f - t;
    bv[1] = t;
    bv[2] = bv[3] = 0.0f;
    bd[0] = -T.inv_hf;
    bd[1] = T.inv_hf;
    bd[2] = bd[3] = 0.0f;
    return;
  }
  const float h = T.hf, omt = 1.0f - t;
  bv[0] = omt * omt * (1.0f + 2.0f * t);
  bv[1] = t * t * (3.0f - 2.0f * t);
  bv[2] = omt * omt * (t * h);
  bv[3] = t * t * (h * (t - 1.0f));
  bd[0] = -6.0f * t * omt * T.inv_hf;
  bd[1] = 6.0f * t * omt * T.inv_hf;
  bd[2] = omt * (1.0f - 3.0f * t);
  bd[3] = t * (3.0f * t - 2.0f);
}

// interval + value / derivative basis of a pair table at r
__device__ inline int pair_basis(const TableDesc &T, double r, float bv[4], float bd[4]) {
  float t;
  const int k = pair_interval(T, r, t);
  basis_at(T, t, bv, bd);
  return k;
}

struct JacArgs {
  int64_t N;
  int C, D, Dd, dens_c0;
  int ntypes, npair, ntab;
  int ntiles;
  const int32_t *tile_cfg, *tile_a0, *tile_na, *cfg_tile0;
  int gi_nf; // free knots per density table when all are equal and fit one slot, else 0
  const int64_t *atom_off, *pair_off, *row_off;
  const int32_t *types;
  const float4 *rec;
  const uint32_t *bits;
  const float4 *jrec; // pairs sorted by (partner type, r) per atom
  const double *jr;   // their FP64 distances (the sort key)
  const TableDesc *tabs;
  const double *tx, *ty, *ts, *tS;
  const float *tS32;      // FP32 copy of tS, read in place when it is not staged
  int stage_S;            // FP32 slope band staged in LDS
  const float4 *ivl32;
  const ColDesc *cols;
  const double *rho, *over, *Fp;
  const float *Fp32;
  float *H, *Hs;          // [N x Dd]
  int32_t *emb_k;         // [N]
  float *emb_d, *emb_v;   // [N x 4]
  double *Fpp;            // [N]
  const double *mu, *lamb; // ADP per-atom dipole (3) and quadrupole (6)
  int Ddip, Dquad;        // ADP dipole / quadrupole free columns
  float *dmu, *dlam;      // ADP moment tangents [N][Ddip][3], [N][Dquad][6]
  float *J;               // [R x D]
  double *part;           // [ntiles x kJP x D]
  const double *vol;
  double ew, sw;
  const double *gauge;    // EAM gauge {A, S[kMaxTypes], F0[kMaxTypes]}
  // knot stride of the pair pass's window tables (0: binary searches)
  int win_pair;
  // the density tables' intervals [ivl_lo, ivl_lo + nivl_st), their FP32
  // forms staged in LDS when stage_ivl
  int ivl_lo, nivl_st, stage_ivl;
};

// LDS staging of the descriptors and, by TS, of the tables the lookups read:
// TS 0 reads the FP32 slope bands (their FP32 copy) and the interval forms in
// place from global memory, TS 1 stages the slope bands, TS 2 the slope bands
// and the density tables' FP32 interval forms. The choice is a template
// argument, not a run-time flag: a pointer that may address either LDS or
// global memory compiles to flat loads, measurably slower than LDS reads on
// every table lookup. Returns the start of the per-wave staging area that
// follows.
template <int TS>
__device__ inline void *stage(const JacArgs &a, float4 *smem, TableDesc *&sT, const float *&sS,
                              int nk, const float4 **sI = nullptr) {
  sT = reinterpret_cast<TableDesc *>(smem);
  const float4 *src = reinterpret_cast<const float4 *>(a.tabs);
  const int nq = a.ntab * int(sizeof(TableDesc) / sizeof(float4));
  for (int k = threadIdx.x; k < nq; k += blockDim.x) smem[k] = src[k];
  float *after = reinterpret_cast<float *>(smem + nq);
  if constexpr (TS >= 1) {
    for (int k = threadIdx.x; k < nk * 5; k += blockDim.x) after[k] = float(a.tS[k]);
    sS = after;
    after += (nk * 5 + 3) & ~3;
  } else {
    sS = a.tS32;
  }
  // the FP32 interval forms of the density tables (the staging's spline
  // derivatives), intervals [ivl_lo, ivl_lo + nivl_st), while they fit; *sI
  // is indexed by interval like the global array
  if constexpr (TS >= 2) {
    float4 *d = reinterpret_cast<float4 *>(after);
    for (int k = threadIdx.x; k < 2 * a.nivl_st; k += blockDim.x) d[k] = a.ivl32[2 * a.ivl_lo + k];
    if (sI) *sI = d - 2 * a.ivl_lo;
    after += 8 * a.nivl_st;
  } else if (sI) {
    *sI = a.ivl32;
  }
  __syncthreads();
  return after;
}

// ---------------------------------------------------------------------------
// Jacobian pair order. Every column's pairs are those whose interval k on the
// column's table lies in [m - 3, m + 2]; with an atom's pairs sorted by
// (partner type, r) they form one contiguous run per partner type, found by
// binary search on the interval index (monotone in r). The FP64 distance
// stored here is the one every Jacobian kernel locates intervals from, so the
// search and the evaluation agree exactly.
__global__ void k_jac_sort(int64_t N, const int64_t *__restrict__ pair_off,
                           const float4 *__restrict__ rec, const double *__restrict__ r64,
                           float4 *__restrict__ jrec, double *__restrict__ jr) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const int64_t p0 = pair_off[i], p1 = pair_off[i + 1];
  for (int64_t p = p0; p < p1; ++p) {
    const float4 q = rec[p];
    const double r = r64[p];
    const int t = pj_type(__float_as_uint(q.w));
    int64_t s = p;
    while (s > p0) {
      const int ts = pj_type(__float_as_uint(jrec[s - 1].w));
      if (ts < t || (ts == t && jr[s - 1] <= r)) break;
      jrec[s] = jrec[s - 1];
      jr[s] = jr[s - 1];
      --s;
    }
    jrec[s] = q;
    jr[s] = r;
  }
}

// Band coefficients of one spline lookup: c[d] is the derivative of
// b . (y_k, y_k+1, s_k, s_k+1) with respect to knot y_(k-2+d), d = 0..5,
// through the makima slope band dS/dy (zero for linear tables).
__device__ inline void band6(const float *sS, const TableDesc &T, int k, const float b[4], float w,
                             float c[6]) {
  const float *S0 = sS + (T.off + k) * 5;
  const float *S1 = S0 + 5;
  c[0] = w * (b[2] * S0[0]);
  c[1] = w * (b[2] * S0[1] + b[3] * S1[0]);
  c[2] = w * (b[0] + b[2] * S0[2] + b[3] * S1[1]);
  c[3] = w * (b[1] + b[2] * S0[3] + b[3] * S1[2]);
  c[4] = w * (b[2] * S0[4] + b[3] * S1[3]);
  c[5] = w * (b[3] * S1[4]);
}

// first index in [lo, hi) of a non-decreasing interval array with k >= kmin
__device__ inline int first_ge(const int32_t *sk, int lo, int hi, int kmin) {
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (sk[mid] >= kmin) hi = mid;
    else lo = mid + 1;
  }
  return lo;
}

// Window starts of the staged chunk, the table form of first_ge: w[t nk + x]
// = the first pair q of partner segment t whose interval k[q] >= x, or the
// segment's end, for x in [0, nk) (nk: the knot stride, at least every
// table's knot count, so x = nk - 1 exceeds every interval index). The
// intervals rise along a segment (pairs sorted by r), so lane q owns
// x in (k[q - 1], k[q]] and the writes do not overlap. A column then reads
// its window bounds with two LDS accesses instead of two binary searches.
// Used by the pair pass (3 % faster there); in the density/embedding pass
// the same tables measured no faster at equal occupancy and cost registers.
constexpr int kWin = 64; // entries: ntypes * nk must fit
__device__ inline void window_init(int ntypes, int nk, int lane, const int32_t *seg, int32_t *w) {
  for (int e = lane; e < ntypes * nk; e += 64) w[e] = seg[e / nk + 1];
}
__device__ inline void window_fill(const int32_t *k, const int32_t *seg, int nk, int lane, int np,
                                   int ty, int32_t *w) {
  if (lane >= np) return;
  const int kq = k[lane];
  const int prev = lane > seg[ty] ? k[lane - 1] : -1;
  const int x1 = min(kq, nk - 1);
  for (int x = prev + 1; x <= x1; ++x) w[ty * nk + x] = lane;
}
// [lo, hi) of the pairs of segment t within three intervals of knot m
__device__ inline void window_of(const int32_t *w, int nk, int t, int m, int &lo, int &hi) {
  lo = w[t * nk + max(m - 3, 0)];
  hi = w[t * nk + min(m + 3, nk - 1)];
}

// segment starts of the partner types in the staged chunk, written to the
// wave's LDS: seg[t] = staged pairs of type < t, t = 0..ntypes (invalid lanes
// carry type kMaxTypes and sort last). Lanes then read their own column's
// bounds with one LDS access instead of selecting among registers.
constexpr int kSegPad = 16; // seg entries per wave (>= kMaxTypes + 1; keeps 64-B alignment)
static_assert(kSegPad >= kMaxTypes + 1, "segment table");
__device__ inline void stage_segments(int ty, int ntypes, int lane, int32_t *seg) {
  int v = 0;
  for (int u = 1; u <= ntypes; ++u) {
    const int cnt = int(__popcll(__ballot(ty < u)));
    v = lane == u ? cnt : v;
  }
  if (lane <= ntypes) seg[lane] = v;
}

// Embedding window of atom i at its clamped density: interval, FP32 value and
// derivative basis (for the embedding columns) and F''(rho) (FP64), with the
// cached-site rules of the embedding lookup (linear outside the span).
__device__ double emb_window(const JacArgs &a, const TableDesc *sT, int64_t i) {
  const int ti = a.types[i];
  const TableDesc Te = sT[a.npair + a.ntypes + ti];
  const double rb = a.rho[i];
  const double *x = a.tx + Te.off, *y = a.ty + Te.off, *s = a.ts + Te.off;
  const int n = Te.n;
  int k;
  bool lin;
  if (rb <= x[0]) {
    k = 0;
    lin = true;
  } else if (rb >= x[n - 1]) {
    k = n - 2;
    lin = true;
  } else {
    k = spl_locate(x, n, rb);
    lin = Te.linear != 0;
  }
  const double dx = x[k + 1] - x[k];
  float dcoef[4], vcoef[4];
  double F2 = 0.0;
  if (lin) {
    const double anchor = (rb >= x[n - 1] && rb > x[0]) ? x[n - 1] : x[k];
    const double f = (rb - anchor) / dx;
    if (rb >= x[n - 1] && rb > x[0]) {
      vcoef[0] = float(-f);
      vcoef[1] = float(1.0 + f);
    } else {
      vcoef[0] = float(1.0 - f);
      vcoef[1] = float(f);
    }
    vcoef[2] = vcoef[3] = 0.0f;
    dcoef[0] = float(-1.0 / dx);
    dcoef[1] = float(1.0 / dx);
    dcoef[2] = dcoef[3] = 0.0f;
  } else {
    double bv[4], bd[4], b2[4];
    herm_basis64(x[k], dx, rb, bv, bd);
    herm_basis2_64(x[k], dx, rb, b2);
    for (int q = 0; q < 4; ++q) {
      vcoef[q] = float(bv[q]);
      dcoef[q] = float(bd[q]);
    }
    F2 = b2[0] * y[k] + b2[1] * y[k + 1] + b2[2] * s[k] + b2[3] * s[k + 1];
  }
  a.emb_k[i] = k;
  for (int q = 0; q < 4; ++q) {
    a.emb_d[4 * i + q] = dcoef[q];
    a.emb_v[4 * i + q] = vcoef[q];
  }
  a.Fpp[i] = F2;
  return F2;
}

// ---------------------------------------------------------------------------
// H_i = d rho_i / d y over the density columns, from staged band coefficients
// band coefficients are read as [m - k + 2][q]: lanes of one slot often read
// the same pair q at different offsets d, so rows are 65 floats apart to put
// them in different LDS banks
constexpr int kBandPitch = 65;

struct DensStage {
  int32_t seg[kSegPad]; // partner-type segment starts of the chunk
  int32_t k[64];
  float c[6][kBandPitch];
};

// embedding window of every atom, one thread per atom (emb_k, emb_d, emb_v,
// Fpp), before the passes that read them
__global__ __launch_bounds__(256) void k_emb_sites(JacArgs a) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < a.N) (void)emb_window(a, a.tabs, i);
}

// density columns [dc0, dc0 + 256) of H and Hs, lane + 64 s per lane
template <int TS>
__global__ __launch_bounds__(256) void k_jac_dens(JacArgs a, int nknots, int dc0) {
  extern __shared__ float4 smem[];
  TableDesc *sT;
  const float *sS;
  DensStage *sAll = reinterpret_cast<DensStage *>(stage<TS>(a, smem, sT, sS, nknots));
  const int wave = uni(threadIdx.x >> 6), lane = threadIdx.x & 63;
  DensStage &P = sAll[wave];
  const int64_t gw = int64_t(blockIdx.x) * 4 + wave;
  const int64_t nw = int64_t(gridDim.x) * 4;
  // my density columns
  int ctab[kCPL], cm[kCPL];
  for (int s = 0; s < kCPL; ++s) {
    const int dc = dc0 + lane + 64 * s;
    if (dc < a.Dd) {
      const ColDesc cd = a.cols[a.dens_c0 + dc];
      ctab[s] = cd.tab;
      cm[s] = cd.m;
    } else {
      ctab[s] = -1;
      cm[s] = 0;
    }
  }
  for (int64_t i = gw; i < a.N; i += nw) {
    const int64_t p0 = uni64(a.pair_off[i]), p1 = uni64(a.pair_off[i + 1]);
    float H[kCPL] = {0, 0, 0, 0};
    for (int64_t pc = p0; pc < p1; pc += 64) {
      const int64_t p = pc + lane;
      int ty = kMaxTypes;
      if (p < p1) {
        const uint32_t b = __float_as_uint(a.jrec[p].w);
        const double r = a.jr[p];
        ty = pj_type(b);
        const TableDesc T = sT[a.npair + ty];
        float bv[4], bd[4];
        const int k = pair_basis(T, r, bv, bd);
        float c[6];
        band6(sS, T, k, bv, ((b & kFlagGj) && !(b & kFlagTiny)) ? 1.0f : 0.0f, c);
        P.k[lane] = k;
#pragma unroll
        for (int d = 0; d < 6; ++d) P.c[d][lane] = c[d];
      }
      stage_segments(ty, a.ntypes, lane, P.seg);
      const int32_t *seg = P.seg;
      wave_sync();
#pragma unroll
      for (int s = 0; s < kCPL; ++s) {
        if (ctab[s] < 0) continue;
        const int t = ctab[s] - a.npair;
        const int m = cm[s];
        const int lo = first_ge(P.k, seg[t], seg[t + 1], m - 3);
        const int hi = first_ge(P.k, lo, seg[t + 1], m + 3);
        for (int q = lo; q < hi; ++q) H[s] += P.c[m - P.k[q] + 2][q];
      }
      wave_sync();
    }
    // F''(rho_i) from k_emb_sites; a clamped atom carries no density chain
    const bool clamped = a.over[i] != 0.0;
    const float hs = clamped ? 0.0f : float(a.Fpp[i]);
#pragma unroll
    for (int s = 0; s < kCPL; ++s) {
      const int dc = dc0 + lane + 64 * s;
      if (dc < a.Dd) {
        a.H[i * a.Dd + dc] = H[s];
        a.Hs[i * a.Dd + dc] = hs * H[s];
      }
    }
  }
}

// ---------------------------------------------------------------------------
// The Jacobian passes: one wave per atom, lanes own columns, grouped by kind
// into slots (one slot holds pair, density or embedding columns only, so a
// slot's loop does not diverge). Per 64-pair chunk every lane first stages
// its pair: unit vector, distance, and for each role of the pair (pair table,
// density of j weighted by F'_i, density of i weighted by F'_j, embedding of
// j weighted by g'_ti(r)) the interval index and the six band coefficients.
// A column then walks only its window of pairs and reads one coefficient per
// pair.
// Virial rows only need per-configuration sums, and the full list holds every
// pair from both ends with the same distance and table: the F'_j g'_ti term of
// pair (i,j) equals the F'_i g'_tj term of (j,i), and the H and embedding
// terms of pair (i,j) equal atom j's per-atom terms through
// Gs_j = sum_i g'_ti(r) r u u. Those pair loops therefore accumulate forces
// only, and the matching per-atom / F'_i terms carry their virial twice.
// Two passes over the same tiles, each with the lane slots of its own columns
// and only the pair staging those columns read, which halves the registers
// and LDS a wave holds (and so doubles the waves a CU keeps in flight):
//   JP_PAIR: pair-table columns (the phi band of every pair);
//   JP_DE:   density and embedding columns (the g_tj, g_ti and F_tj bands,
//            the density-chain and per-atom terms).
enum { JP_PAIR = 1, JP_DE = 2 };
constexpr int kHBatch = 8; // Hs rows gathered per batch by the density-chain term
constexpr int kWavesPair = 5, kWavesDE = 4; // waves per SIMD asked of the two passes

struct JStagePair {
  int32_t seg[kSegPad]; // partner-type segment starts of the chunk
  float r[64], u0[64], u1[64], u2[64];
  int32_t kphi[64];
  int32_t wphi[kWin]; // window starts of the phi intervals (a.win_pair)
  float cphi[6][kBandPitch], ephi[6][kBandPitch];
};

struct JStageDE {
  int32_t seg[kSegPad]; // partner-type segment starts of the chunk
  float r[64], u0[64], u1[64], u2[64];
  float hw0[64], hw1[64], hw2[64]; // g'_ti(r) u: the density-chain weights
  int32_t jrow[64], kgj[64], kgi[64], kfj[64]; // jrow: j * Dd (row of Hs)
  float cgj[6][kBandPitch], cgi[6][kBandPitch], cemb[6][kBandPitch];
};

template <int PASS> struct JStageOf;
template <> struct JStageOf<JP_PAIR> { using T = JStagePair; };
template <> struct JStageOf<JP_DE> { using T = JStageDE; };

struct SlotDesc {
  int32_t tab, m, kind, dc; // as ColDesc; dc = density column or pair-table types
  int32_t col;              // Jacobian column
  int32_t qc;               // ADP dipole / quadrupole column: index among its kind
  int32_t pad0, pad1;
};

// per-atom force accumulator of one column; the column's energy, virial and
// limit sums go straight into the tile partial t[kJP] (e, xx, yy, zz, xy, xz,
// yz, limit), which keeps the register count down
struct ColAcc {
  float f[3];
};

__device__ inline void add_force(ColAcc &A, float c, float u0, float u1, float u2) {
  A.f[0] += c * u0;
  A.f[1] += c * u1;
  A.f[2] += c * u2;
}

__device__ inline void add_virial(float *t, float c, float r, float u0, float u1, float u2) {
  const float cr = c * r;
  t[1] += cr * u0 * u0;
  t[2] += cr * u1 * u1;
  t[3] += cr * u2 * u2;
  t[4] += cr * u0 * u1;
  t[5] += cr * u0 * u2;
  t[6] += cr * u1 * u2;
}

__device__ inline void add_pair(ColAcc &A, float *t, float c, float r, float u0, float u1, float u2) {
  add_force(A, c, u0, u1, u2);
  add_virial(t, c, r, u0, u1, u2);
}

template <int PASS, int NS, int TS>
__device__ __forceinline__ void jac_main_body(const JacArgs &a, int nknots, const SlotDesc *slots) {
  using Stage = typename JStageOf<PASS>::T;
  constexpr bool kPairPass = PASS == JP_PAIR;
  extern __shared__ float4 smem[];
  TableDesc *sT;
  const float *sS;
  const float4 *sI;
  Stage *sAll = reinterpret_cast<Stage *>(stage<TS>(a, smem, sT, sS, nknots, &sI));
  // the tile-end reduction reuses the staging area of the four waves: one
  // entry per partial row and slot column (s, lane)
  static_assert(4 * sizeof(Stage) >= sizeof(double) * kJP * 64 * NS, "reduction buffer");
  double(*sPart)[64 * NS] = reinterpret_cast<double(*)[64 * NS]>(sAll);
  const int wave = uni(threadIdx.x >> 6), lane = threadIdx.x & 63;
  Stage &P = sAll[wave];
  int ccol[NS], ctab[NS], cm[NS], ckind[NS], cdc[NS];
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    const SlotDesc sd = slots[s * 64 + lane];
    ccol[s] = sd.col;
    ctab[s] = sd.tab;
    cm[s] = sd.m;
    ckind[s] = sd.kind;
    cdc[s] = sd.dc;
  }
  // XCD-aware tiles: blocks are dispatched round-robin over the 8 XCDs, so
  // block b runs on XCD b % 8; each XCD takes one contiguous eighth of the
  // tiles (whole configurations), which keeps the neighbours' Hs rows that the
  // density-chain term gathers in that XCD's L2
  const int ng = min(kXcds, int(gridDim.x)); // tile ranges; every range gets a block
  const int xcd = blockIdx.x % ng, slot = blockIdx.x / ng;
  const int nslot = (int(gridDim.x) - xcd + ng - 1) / ng;
  const int t_end = int(int64_t(xcd + 1) * a.ntiles / ng);
  for (int tile = int(int64_t(xcd) * a.ntiles / ng) + slot; tile < t_end; tile += nslot) {
    const int na = uni(a.tile_na[tile]);
    const int c = uni(a.tile_cfg[tile]);
    const int64_t cbase = uni64(a.atom_off[c]);
    float tp[NS][kJP];
#pragma unroll
    for (int s = 0; s < NS; ++s)
#pragma unroll
      for (int q = 0; q < kJP; ++q) tp[s][q] = 0.0f;
    for (int li = wave; li < na; li += 4) {
      const int64_t i = int64_t(uni(a.tile_a0[tile])) + li;
      const int ti = uni(a.types[i]);
      const int64_t p0 = uni64(a.pair_off[i]), p1 = uni64(a.pair_off[i + 1]);
      ColAcc A[NS];
#pragma unroll
      for (int s = 0; s < NS; ++s) {
        for (int q = 0; q < 3; ++q) A[s].f[q] = 0.0f;
      }
      float G[3] = {0, 0, 0}, Gs[6] = {0, 0, 0, 0, 0, 0};
      // partner type of each pair column for this centre type (-1: not ti's table)
      int ptt[NS];
#pragma unroll
      for (int s = 0; s < NS; ++s) {
        ptt[s] = -1;
        if (kPairPass && ckind[s] == CK_PAIR) {
          const int ta = cdc[s] & 0xff, tb = (cdc[s] >> 8) & 0xff;
          ptt[s] = ti == ta ? tb : (ti == tb ? ta : -1);
        }
      }
      for (int64_t pc = p0; pc < p1; pc += 64) {
        const int64_t p = pc + lane;
        int ty = kMaxTypes;
        if (p < p1) {
          const float4 q = a.jrec[p];
          const uint32_t b = __float_as_uint(q.w);
          const double r64 = a.jr[p];
          const float rr = float(r64);
          const bool tiny = (b & kFlagTiny) != 0;
          const float inv_r = tiny ? 0.0f : 1.0f / rr;
          const int tj = pj_type(b);
          ty = tj;
          const float u0 = q.x * inv_r, u1 = q.y * inv_r, u2 = q.z * inv_r;
          P.r[lane] = rr;
          P.u0[lane] = u0;
          P.u1[lane] = u1;
          P.u2[lane] = u2;
          float bv[4], bd[4], cc[6];
          if constexpr (kPairPass) {
            // pair table phi_(ti,tj)
            float ce[6];
            const TableDesc T = sT[ord_pair(ti, tj, a.ntypes)];
            const int k = pair_basis(T, r64, bv, bd);
            const float w = ((b & kFlagPhi) && !tiny) ? 1.0f : 0.0f;
            band6(sS, T, k, bd, w, cc);
            band6(sS, T, k, bv, 0.5f * w, ce);
            P.kphi[lane] = k;
#pragma unroll
            for (int d = 0; d < 6; ++d) {
              P.cphi[d][lane] = cc[d];
              P.ephi[d][lane] = ce[d];
            }
          } else {
            const int32_t jg = int32_t(cbase + pj_local(b));
            P.jrow[lane] = jg * a.Dd;
            const float Fpi = a.Fp32[i];
            const float fpj = a.Fp32[jg];
            // density of j, weight F'_i; also g'_tj(r) for G_i
            float dgj = 0.0f;
            {
              const TableDesc T = sT[a.npair + tj];
              float tt;
              const int k = pair_interval(T, r64, tt);
              basis_at(T, tt, bv, bd);
              const bool on = (b & kFlagGj) && !tiny;
              band6(sS, T, k, bd, on ? Fpi : 0.0f, cc);
              P.kgj[lane] = k;
#pragma unroll
              for (int d = 0; d < 6; ++d) P.cgj[d][lane] = cc[d];
              if (on) {
                float v;
                spl_eval32(T, sI, k, tt, v, dgj);
              }
            }
            // density of i, weight F'_j; g'_ti(r) weights the embedding / H terms
            float dgi = 0.0f;
            {
              const TableDesc T = sT[a.npair + ti];
              float tt;
              const int k = pair_interval(T, r64, tt);
              basis_at(T, tt, bv, bd);
              const bool on = (b & kFlagGi) && !tiny;
              band6(sS, T, k, bd, on ? fpj : 0.0f, cc);
              P.kgi[lane] = k;
#pragma unroll
              for (int d = 0; d < 6; ++d) P.cgi[d][lane] = cc[d];
              if (on) {
                float v;
                spl_eval32(T, sI, k, tt, v, dgi);
              }
            }
            P.hw0[lane] = dgi * u0;
            P.hw1[lane] = dgi * u1;
            P.hw2[lane] = dgi * u2;
            // embedding of j at its density, weight g'_ti(r)
            {
              const TableDesc T = sT[a.npair + a.ntypes + tj];
              const int kf = a.emb_k[jg];
              const float4 e = reinterpret_cast<const float4 *>(a.emb_d)[jg];
              const float eb[4] = {e.x, e.y, e.z, e.w};
              band6(sS, T, kf, eb, dgi, cc);
              P.kfj[lane] = kf;
#pragma unroll
              for (int d = 0; d < 6; ++d) P.cemb[d][lane] = cc[d];
            }
            // G_i = sum_j g'_tj(r) u, virial counterpart sum_j g'_tj r u u
            G[0] += dgj * u0;
            G[1] += dgj * u1;
            G[2] += dgj * u2;
            const float gr = dgj * rr;
            Gs[0] += gr * u0 * u0;
            Gs[1] += gr * u1 * u1;
            Gs[2] += gr * u2 * u2;
            Gs[3] += gr * u0 * u1;
            Gs[4] += gr * u0 * u2;
            Gs[5] += gr * u1 * u2;
          }
        } else if constexpr (!kPairPass) {
          // past the chunk's pairs: a zero-weight gather of row 0, so that the
          // density-chain gathers run in whole batches
          P.jrow[lane] = 0;
          P.hw0[lane] = P.hw1[lane] = P.hw2[lane] = 0.0f;
        }
        stage_segments(ty, a.ntypes, lane, P.seg);
        const int32_t *seg = P.seg;
        wave_sync();
        const int np = int(min<int64_t>(64, p1 - pc));
        if constexpr (kPairPass) {
          if (a.win_pair) {
            window_init(a.ntypes, a.win_pair, lane, seg, P.wphi);
            wave_sync();
            window_fill(P.kphi, seg, a.win_pair, lane, np, ty, P.wphi);
            wave_sync();
          }
        }
#pragma unroll
        for (int s = 0; s < NS; ++s) {
          const int kind = ckind[s];
          if (kind == CK_NONE) continue;
          const int m = cm[s];
          if constexpr (kPairPass) {
            if (ptt[s] < 0) continue;
            const int s1 = seg[ptt[s] + 1];
            int lo, hi;
            if (a.win_pair) {
              window_of(P.wphi, a.win_pair, ptt[s], m, lo, hi);
            } else {
              lo = first_ge(P.kphi, seg[ptt[s]], s1, m - 3);
              hi = first_ge(P.kphi, lo, s1, m + 3);
            }
            for (int q = lo; q < hi; ++q) {
              const int d = m - P.kphi[q] + 2;
              add_pair(A[s], tp[s], P.cphi[d][q], P.r[q], P.u0[q], P.u1[q], P.u2[q]);
              tp[s][0] += P.ephi[d][q];
            }
          } else if (kind == CK_DENS) {
            const int t = ctab[s] - a.npair;
            // F'_i g'_tj(r), tj == t
            {
              const int s1 = seg[t + 1];
              const int lo = first_ge(P.kgj, seg[t], s1, m - 3);
              const int hi = first_ge(P.kgj, lo, s1, m + 3);
              for (int q = lo; q < hi; ++q) {
                const float cq = P.cgj[m - P.kgj[q] + 2][q];
                add_force(A[s], cq, P.u0[q], P.u1[q], P.u2[q]);
                add_virial(tp[s], 2.0f * cq, P.r[q], P.u0[q], P.u1[q], P.u2[q]);
              }
            }
            // F'_j g'_ti(r) of column (ti, j), every partner type. With equal
            // free-knot counts nf per density table (lane = t nf + j), lane
            // (t, j) takes partner segment t of column (ti, j) and the owner
            // lane (ti, j) collects the ntypes partials by shuffles
            if (a.gi_nf > 0) {
              const int nf = a.gi_nf;
              const int j = lane - t * nf;
              const int mg = __shfl(m, ti * nf + j, 64);
              float g0 = 0.0f, g1 = 0.0f, g2 = 0.0f;
              {
                const int s1 = seg[t + 1];
                const int lo = first_ge(P.kgi, seg[t], s1, mg - 3);
                const int hi = first_ge(P.kgi, lo, s1, mg + 3);
                for (int q = lo; q < hi; ++q) {
                  const float cq = P.cgi[mg - P.kgi[q] + 2][q];
                  g0 += cq * P.u0[q];
                  g1 += cq * P.u1[q];
                  g2 += cq * P.u2[q];
                }
              }
              for (int tt = 0; tt < a.ntypes; ++tt) {
                const int src = tt * nf + j;
                const float v0 = __shfl(g0, src, 64), v1 = __shfl(g1, src, 64), v2 = __shfl(g2, src, 64);
                if (t == ti) {
                  A[s].f[0] += v0;
                  A[s].f[1] += v1;
                  A[s].f[2] += v2;
                }
              }
            } else if (t == ti) {
              for (int tt = 0; tt < a.ntypes; ++tt) {
                const int s1 = seg[tt + 1];
                const int lo = first_ge(P.kgi, seg[tt], s1, m - 3);
                const int hi = first_ge(P.kgi, lo, s1, m + 3);
                for (int q = lo; q < hi; ++q)
                  add_force(A[s], P.cgi[m - P.kgi[q] + 2][q], P.u0[q], P.u1[q], P.u2[q]);
              }
            }
            // F''_j H_j g'_ti(r): every pair, rows of Hs gathered kHBatch at a time
            // 32-bit byte offsets from the uniform table base (the table is
            // < 4 GiB): one VALU op per address
            const char *Hb = reinterpret_cast<const char *>(a.Hs);
            const uint32_t colb = uint32_t(cdc[s]) << 2;
            auto hs = [&](int q) {
              return *reinterpret_cast<const float *>(Hb + ((uint32_t(P.jrow[q]) << 2) + colb));
            };
            // two batches in flight: batch b + 1's loads are issued before
            // batch b is consumed (0.8 % faster than one batch at a time);
            // the last batch is padded with the chunk's zero-weight lanes, so
            // no pair waits alone for its row
            const int nbat = (np + kHBatch - 1) / kHBatch;
            float hn[kHBatch];
            if (nbat > 0) {
#pragma unroll
              for (int v = 0; v < kHBatch; ++v) hn[v] = hs(v);
            } else {
#pragma unroll
              for (int v = 0; v < kHBatch; ++v) hn[v] = 0.0f;
            }
            for (int b = 0; b < nbat; ++b) {
              float h[kHBatch];
#pragma unroll
              for (int v = 0; v < kHBatch; ++v) h[v] = hn[v];
              if (b + 1 < nbat) {
#pragma unroll
                for (int v = 0; v < kHBatch; ++v) hn[v] = hs((b + 1) * kHBatch + v);
              }
              const int q0 = b * kHBatch;
#pragma unroll
              for (int v = 0; v < kHBatch; ++v)
                add_force(A[s], h[v], P.hw0[q0 + v], P.hw1[q0 + v], P.hw2[q0 + v]);
            }
          } else if (kind == CK_EMB) {
            // F'_j through the embedding window of j, tj == t
            const int t = ctab[s] - a.npair - a.ntypes;
            for (int q = seg[t]; q < seg[t + 1]; ++q) {
              const int d = m - P.kfj[q] + 2;
              if (d < 0 || d > 5) continue;
              add_force(A[s], P.cemb[d][q], P.u0[q], P.u1[q], P.u2[q]);
            }
          }
        }
        wave_sync();
      }
      const int64_t row = a.row_off[c] + 3 * (i - cbase);
      if constexpr (kPairPass) {
#pragma unroll
        for (int s = 0; s < NS; ++s) {
          if (ckind[s] == CK_NONE) continue;
          const int col = ccol[s];
          for (int q = 0; q < 3; ++q) a.J[(row + q) * a.D + col] = A[s].f[q];
        }
      } else {
        for (int q = 0; q < 3; ++q) G[q] = wave_sum(G[q]);
        for (int q = 0; q < 6; ++q) Gs[q] = wave_sum(Gs[q]);
        // per-atom terms
        // a clamped atom's energy keeps S_t A (rho - clamped rho) under the
        // gauge (the slope compensation of its pairs), d / d sigma = S_t A
        const bool clamped = a.over[i] != 0.0;
        const double Fpi64 = a.Fp[i], over = a.over[i];
        const double SA = a.gauge[1 + ti] * a.gauge[0];
        const int kfi = a.emb_k[i];
        float ebi[4], evi[4];
        for (int q = 0; q < 4; ++q) {
          ebi[q] = a.emb_d[4 * i + q];
          evi[q] = a.emb_v[4 * i + q];
        }
        const int temb_i = a.npair + a.ntypes + ti;
#pragma unroll
        for (int s = 0; s < NS; ++s) {
          if (ckind[s] == CK_NONE) continue;
          float g = 0.0f, e_extra = 0.0f, l_extra = 0.0f;
          if (ckind[s] == CK_DENS) {
            g = a.Hs[i * a.Dd + cdc[s]];
            const double h = a.H[i * a.Dd + cdc[s]];
            if (clamped) {
              l_extra = float(over * h);
              e_extra = float(SA * h);
            } else {
              e_extra = float(Fpi64 * h);
            }
          } else if (ckind[s] == CK_EMB && ctab[s] == temb_i) {
            const TableDesc D = sT[ctab[s]];
            g = band(sS, D, kfi, cm[s], ebi);
            e_extra = band(sS, D, kfi, cm[s], evi);
          }
          for (int q = 0; q < 3; ++q) A[s].f[q] += g * G[q];
          // twice: the virial of the mirrored pair terms (see the pair loop)
          for (int q = 0; q < 6; ++q) tp[s][1 + q] += 2.0f * g * Gs[q];
          const int col = ccol[s];
          for (int q = 0; q < 3; ++q) a.J[(row + q) * a.D + col] = A[s].f[q];
          tp[s][0] += e_extra;
          tp[s][7] += l_extra;
        }
      }
    }
    // reduce the four waves' partials in wave order (FP64), write the tile
    // partial of this pass's columns
    __syncthreads(); // every wave is done with its staging area
    for (int w = 0; w < 4; ++w) {
      if (wave == w) {
#pragma unroll
        for (int s = 0; s < NS; ++s) {
          if (ckind[s] == CK_NONE) continue;
          const int lc = 64 * s + lane;
          for (int q = 0; q < kJP; ++q)
            sPart[q][lc] = (w == 0 ? 0.0 : sPart[q][lc]) + double(tp[s][q]);
        }
      }
      __syncthreads();
    }
    if (wave == 0) {
#pragma unroll
      for (int s = 0; s < NS; ++s) {
        if (ckind[s] == CK_NONE) continue;
        const int col = ccol[s], lc = 64 * s + lane;
        for (int q = 0; q < kJP; ++q) a.part[(int64_t(tile) * kJP + q) * a.D + col] = sPart[q][lc];
      }
    }
    __syncthreads();
  }
}

// The pass kernels. The window and gather loops are latency-bound, so each
// pass asks for W waves per SIMD (the register budget follows): 4 for the
// density/embedding pass (128 VGPRs, no scratch; with the density tables'
// interval forms staged in LDS the compiler's own choice, 131 VGPRs, allows
// only 3 and measured 0.93 -> 1.12 ms on the 500-configuration ternary
// workload), and 5 for the lighter pair pass (4 and 6 measured the same).
template <int NS, int W, int TS>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(W))) void k_jac_pair(
    JacArgs a, int nknots, const SlotDesc *slots) {
  jac_main_body<JP_PAIR, NS, TS>(a, nknots, slots);
}

template <int NS, int W, int TS>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(W))) void k_jac_de(
    JacArgs a, int nknots, const SlotDesc *slots) {
  jac_main_body<JP_DE, NS, TS>(a, nknots, slots);
}

// ---------------------------------------------------------------------------
// Knot sums. A pair-table column of the Jacobian, and a density column's rows
// H_i, are linear in per-atom sums that depend on the geometry and the table
// grids only, not on the knot values. For a table T, knot m and the atom's
// pairs on T:
//   Y[m] = sum of b0 w over the pairs in interval m, plus b1 w over those in
//          interval m - 1,
//   S[m] = the same with b2 and b3,
// where (b0, b1, b2, b3) is the Hermite basis of the pair's place in its
// interval (the derivative basis for forces and the virial, the value basis
// for the energy and the density) and w the pair's weight (its direction u
// for forces, r u u for the virial, 1/2 for the energy, 1 for the density).
// The column of knot m is then
//   Y[m] + sum_{k = m-2..m+2} S[k] ds_k/dy_m,
// identity on the knot values and the makima slope band on the slopes:
// band6's coefficients summed over the pairs, regrouped by knot.
// k_knot_sums forms the sums once per neighbour list and model grid (lanes
// own knots and walk the pairs of their knot's two intervals in the sorted
// pair order); k_jac_pair_ks and k_jac_dens_ks fold them with the current
// slope band at every Jacobian, which reads the sums and writes the columns.
//   ks_pair  per atom: force sums of the tables phi(ti, t), t in type order,
//            [Y: 3 lp floats][S: 3 lp floats]
//   ks_dens  per atom: density value sums of the tables g_t, t in type order,
//            [Y: ld floats][S: ld floats]
//   ks_cfg   per configuration (FP64): pair-table energy and virial sums,
//            [14][kp] as (Y_e, S_e, Y_w[6], S_w[6]) by flat pair-table knot
// ---------------------------------------------------------------------------
struct KsArgs {
  float *pair, *dens;
  double *tile, *cfg;
  int lp, ld, kp;
};

// first entry of table phi(ti, t) among the pair sums of an atom of type ti
__device__ inline int ks_pair_off(const TableDesc *sT, int ntypes, int ti, int t) {
  int o = 0;
  for (int u = 0; u < t; ++u) o += sT[ord_pair(ti, u, ntypes)].n;
  return o;
}

struct KsStage {
  int32_t seg[kSegPad]; // partner-type segment starts of the chunk
  float r[64], u0[64], u1[64], u2[64];
  int32_t kp[64], kg[64];                     // intervals on phi(ti, tj) and g_tj
  float bdp[4][64], bvp[4][64], bvg[4][64];   // weighted bases
};

__host__ __device__ constexpr size_t ks_wave_bytes(int kp) {
  return (sizeof(KsStage) + sizeof(float) * 14 * size_t(kp) + 15) & ~size_t(15);
}

// one wave per atom; lanes own knots: two slots of the atom's pair-table
// knots (entry L = lane + 64 s of its pair sums), two of the density knots.
// Energy and virial sums go into the wave's FP32 accumulators by flat
// pair-table knot (distinct per lane for one atom), then per tile in FP64.
__global__ __launch_bounds__(256) void k_knot_sums(JacArgs a, KsArgs k) {
  extern __shared__ float4 smem[];
  TableDesc *sT = reinterpret_cast<TableDesc *>(smem);
  const int ndq = a.ntab * int(sizeof(TableDesc) / sizeof(float4));
  for (int e = threadIdx.x; e < ndq; e += blockDim.x)
    smem[e] = reinterpret_cast<const float4 *>(a.tabs)[e];
  const int wave = uni(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const size_t wb = ks_wave_bytes(k.kp);
  char *base = reinterpret_cast<char *>(smem + ndq);
  KsStage &P = *reinterpret_cast<KsStage *>(base + wave * wb);
  float *acc = reinterpret_cast<float *>(base + wave * wb + sizeof(KsStage)); // [14][kp]
  __syncthreads();
  const int nk = 14 * k.kp;
  for (int tile = blockIdx.x; tile < a.ntiles; tile += gridDim.x) {
    for (int e = lane; e < nk; e += 64) acc[e] = 0.0f;
    wave_sync();
    const int na = uni(a.tile_na[tile]);
    for (int li = wave; li < na; li += 4) {
      const int64_t i = int64_t(uni(a.tile_a0[tile])) + li;
      const int ti = uni(a.types[i]);
      const int64_t p0 = uni64(a.pair_off[i]), p1 = uni64(a.pair_off[i + 1]);
      // this lane's knots: (partner type, knot) of a pair table, of a density table
      int pt[2], pm[2], pK[2], dt[2], dm[2];
#pragma unroll
      for (int s = 0; s < 2; ++s) {
        const int L = lane + 64 * s;
        pt[s] = dt[s] = -1;
        pm[s] = dm[s] = pK[s] = 0;
        int o = 0;
        for (int t = 0; t < a.ntypes; ++t) {
          const TableDesc &T = sT[ord_pair(ti, t, a.ntypes)];
          if (pt[s] < 0 && L >= o && L < o + T.n) {
            pt[s] = t;
            pm[s] = L - o;
            pK[s] = T.off + L - o;
          }
          o += T.n;
        }
        o = 0;
        for (int t = 0; t < a.ntypes; ++t) {
          const int n = sT[a.npair + t].n;
          if (dt[s] < 0 && L >= o && L < o + n) {
            dt[s] = t;
            dm[s] = L - o;
          }
          o += n;
        }
      }
      float f[2][6], e[2][14], v[2][2];
#pragma unroll
      for (int s = 0; s < 2; ++s) {
#pragma unroll
        for (int q = 0; q < 6; ++q) f[s][q] = 0.0f;
#pragma unroll
        for (int q = 0; q < 14; ++q) e[s][q] = 0.0f;
        v[s][0] = v[s][1] = 0.0f;
      }
      for (int64_t pc = p0; pc < p1; pc += 64) {
        const int64_t p = pc + lane;
        int ty = kMaxTypes;
        if (p < p1) {
          const float4 q = a.jrec[p];
          const uint32_t b = __float_as_uint(q.w);
          const double r64 = a.jr[p];
          const float rr = float(r64);
          const bool tiny = (b & kFlagTiny) != 0;
          const float inv_r = tiny ? 0.0f : 1.0f / rr;
          const int tj = pj_type(b);
          ty = tj;
          P.r[lane] = rr;
          P.u0[lane] = q.x * inv_r;
          P.u1[lane] = q.y * inv_r;
          P.u2[lane] = q.z * inv_r;
          float bv[4], bd[4];
          {
            const TableDesc &T = sT[ord_pair(ti, tj, a.ntypes)];
            P.kp[lane] = pair_basis(T, r64, bv, bd);
            const float w = ((b & kFlagPhi) && !tiny) ? 1.0f : 0.0f;
#pragma unroll
            for (int c = 0; c < 4; ++c) {
              P.bdp[c][lane] = w * bd[c];
              P.bvp[c][lane] = w * bv[c];
            }
          }
          {
            const TableDesc &T = sT[a.npair + tj];
            P.kg[lane] = pair_basis(T, r64, bv, bd);
            const float w = ((b & kFlagGj) && !tiny) ? 1.0f : 0.0f;
#pragma unroll
            for (int c = 0; c < 4; ++c) P.bvg[c][lane] = w * bv[c];
          }
        }
        stage_segments(ty, a.ntypes, lane, P.seg);
        wave_sync();
#pragma unroll
        for (int s = 0; s < 2; ++s) {
          if (pt[s] >= 0) {
            const int m = pm[s], s1 = P.seg[pt[s] + 1];
            const int lo = first_ge(P.kp, P.seg[pt[s]], s1, m - 1);
            const int hi = first_ge(P.kp, lo, s1, m + 1);
            for (int q = lo; q < hi; ++q) {
              const bool at = P.kp[q] == m; // interval m (b0, b2) or m - 1 (b1, b3)
              const float cy = at ? P.bdp[0][q] : P.bdp[1][q], cs = at ? P.bdp[2][q] : P.bdp[3][q];
              const float ey = at ? P.bvp[0][q] : P.bvp[1][q], es = at ? P.bvp[2][q] : P.bvp[3][q];
              const float u0 = P.u0[q], u1 = P.u1[q], u2 = P.u2[q], r = P.r[q];
              f[s][0] += cy * u0;
              f[s][1] += cy * u1;
              f[s][2] += cy * u2;
              f[s][3] += cs * u0;
              f[s][4] += cs * u1;
              f[s][5] += cs * u2;
              e[s][0] += 0.5f * ey;
              e[s][1] += 0.5f * es;
              const float ruu[6] = {r * u0 * u0, r * u1 * u1, r * u2 * u2,
                                    r * u0 * u1, r * u0 * u2, r * u1 * u2};
#pragma unroll
              for (int c = 0; c < 6; ++c) {
                e[s][2 + c] += cy * ruu[c];
                e[s][8 + c] += cs * ruu[c];
              }
            }
          }
          if (dt[s] >= 0) {
            const int m = dm[s], s1 = P.seg[dt[s] + 1];
            const int lo = first_ge(P.kg, P.seg[dt[s]], s1, m - 1);
            const int hi = first_ge(P.kg, lo, s1, m + 1);
            for (int q = lo; q < hi; ++q) {
              const bool at = P.kg[q] == m;
              v[s][0] += at ? P.bvg[0][q] : P.bvg[1][q];
              v[s][1] += at ? P.bvg[2][q] : P.bvg[3][q];
            }
          }
        }
        wave_sync();
      }
      float *fp = k.pair + size_t(i) * 6 * k.lp;
      float *fd = k.dens + size_t(i) * 2 * k.ld;
#pragma unroll
      for (int s = 0; s < 2; ++s) {
        const int L = lane + 64 * s;
        if (pt[s] >= 0) {
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            fp[3 * L + c] = f[s][c];
            fp[3 * k.lp + 3 * L + c] = f[s][3 + c];
          }
#pragma unroll
          for (int q = 0; q < 14; ++q) acc[q * k.kp + pK[s]] += e[s][q];
        }
        if (dt[s] >= 0) {
          fd[L] = v[s][0];
          fd[k.ld + L] = v[s][1];
        }
      }
    }
    // tile sums in FP64, waves in order
    __syncthreads();
    for (int e = threadIdx.x; e < nk; e += blockDim.x) {
      double s = 0.0;
      for (int w = 0; w < 4; ++w)
        s += double(reinterpret_cast<const float *>(base + w * wb + sizeof(KsStage))[e]);
      k.tile[size_t(tile) * nk + e] = s;
    }
    __syncthreads();
  }
}

// per configuration: its tiles' energy and virial sums in tile order
__global__ void k_knot_sums_cfg(JacArgs a, KsArgs k) {
  const int c = blockIdx.x;
  const int t0 = a.cfg_tile0[c], t1 = a.cfg_tile0[c + 1];
  const int nk = 14 * k.kp;
  for (int e = threadIdx.x; e < nk; e += blockDim.x) {
    double s = 0.0;
    for (int t = t0; t < t1; ++t) s += k.tile[size_t(t) * nk + e];
    k.cfg[size_t(c) * nk + e] = s;
  }
}

// slope weights of the column of knot m of table T: ds_k/dy_m for
// k = m - 2 + q, q = 0..4 (zero outside the table)
__device__ inline void ks_slope_weights(const JacArgs &a, const TableDesc &T, int m, float Sw[5]) {
#pragma unroll
  for (int q = 0; q < 5; ++q) {
    const int kk = m - 2 + q;
    Sw[q] = (kk >= 0 && kk < T.n) ? float(a.tS[size_t(T.off + kk) * 5 + (4 - q)]) : 0.0f;
  }
}

// pair-table columns from the knot sums: the force rows of every atom, and
// the energy and virial partials of a configuration carried by its first
// tile (its other tiles carry zeros, so the per-configuration sums over tiles
// are the configuration's)
// (each wave copies its atom's sums into LDS with whole-wave float4 loads,
// one pass over the block, and the columns fold them from there)
template <int NS>
__global__ __launch_bounds__(256) void k_jac_pair_ks(JacArgs a, KsArgs k, const SlotDesc *slots) {
  extern __shared__ float4 smem[];
  TableDesc *sT = reinterpret_cast<TableDesc *>(smem);
  const int ndq = a.ntab * int(sizeof(TableDesc) / sizeof(float4));
  for (int e = threadIdx.x; e < ndq; e += blockDim.x)
    smem[e] = reinterpret_cast<const float4 *>(a.tabs)[e];
  __syncthreads();
  const int wave = uni(threadIdx.x >> 6), lane = threadIdx.x & 63;
  float *sK = reinterpret_cast<float *>(smem + ndq) + wave * 6 * k.lp; // the atom's sums
  const int nq = (6 * k.lp) >> 2;
  int ccol[NS], cm[NS], cn[NS], cta[NS], ctb[NS], coffa[NS], coffb[NS];
  float Sw[NS][5];
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    const SlotDesc sd = slots[s * 64 + lane];
    ccol[s] = sd.kind == CK_PAIR ? sd.col : -1;
    cm[s] = sd.m;
    cn[s] = 1;
    cta[s] = ctb[s] = -1;
    coffa[s] = coffb[s] = 0;
#pragma unroll
    for (int q = 0; q < 5; ++q) Sw[s][q] = 0.0f;
    if (sd.kind == CK_PAIR) {
      const TableDesc &T = sT[sd.tab];
      cn[s] = T.n;
      cta[s] = sd.dc & 0xff;
      ctb[s] = (sd.dc >> 8) & 0xff;
      coffa[s] = ks_pair_off(sT, a.ntypes, cta[s], ctb[s]); // centre of type ta
      coffb[s] = ks_pair_off(sT, a.ntypes, ctb[s], cta[s]); // centre of type tb
      ks_slope_weights(a, T, sd.m, Sw[s]);
    }
  }
  const int ng = min(kXcds, int(gridDim.x));
  const int xcd = blockIdx.x % ng, slot = blockIdx.x / ng;
  const int nslot = (int(gridDim.x) - xcd + ng - 1) / ng;
  const int t_end = int(int64_t(xcd + 1) * a.ntiles / ng);
  for (int tile = int(int64_t(xcd) * a.ntiles / ng) + slot; tile < t_end; tile += nslot) {
    const int na = uni(a.tile_na[tile]);
    const int c = uni(a.tile_cfg[tile]);
    const int64_t cbase = uni64(a.atom_off[c]);
    const int64_t row0 = a.row_off[c];
    for (int li = wave; li < na; li += 4) {
      const int64_t i = int64_t(uni(a.tile_a0[tile])) + li;
      const int ti = uni(a.types[i]);
      const float4 *src = reinterpret_cast<const float4 *>(k.pair + size_t(i) * 6 * k.lp);
      for (int e = lane; e < nq; e += 64) reinterpret_cast<float4 *>(sK)[e] = src[e];
      wave_sync();
      const float *Yb = sK, *Sb = sK + 3 * k.lp;
      const int64_t row = row0 + 3 * (i - cbase);
#pragma unroll
      for (int s = 0; s < NS; ++s) {
        if (ccol[s] < 0) continue;
        float f0 = 0.0f, f1 = 0.0f, f2 = 0.0f;
        const int off = ti == cta[s] ? coffa[s] : (ti == ctb[s] ? coffb[s] : -1);
        if (off >= 0) {
          const int m = cm[s], L = off + m;
          f0 = Yb[3 * L];
          f1 = Yb[3 * L + 1];
          f2 = Yb[3 * L + 2];
#pragma unroll
          for (int q = 0; q < 5; ++q) {
            const int kk = off + min(max(m - 2 + q, 0), cn[s] - 1);
            f0 += Sw[s][q] * Sb[3 * kk];
            f1 += Sw[s][q] * Sb[3 * kk + 1];
            f2 += Sw[s][q] * Sb[3 * kk + 2];
          }
        }
        float *Jr = a.J + row * a.D + ccol[s];
        Jr[0] = f0;
        Jr[a.D] = f1;
        Jr[2 * a.D] = f2;
      }
      wave_sync(); // every lane has folded before the next atom's copy
    }
  }
}

// energy and virial partials of the pair-table columns: a configuration's
// sums are carried by its first tile, its other tiles carry zeros (the rows
// sum a configuration's tiles); one block per configuration, one thread per
// lane-slot column
__global__ __launch_bounds__(256) void k_jac_pair_ks_rows(JacArgs a, KsArgs k, const SlotDesc *slots,
                                                          int nslot) {
  const int c = blockIdx.x;
  const int t0 = a.cfg_tile0[c], t1 = a.cfg_tile0[c + 1];
  const double *E = k.cfg + size_t(c) * 14 * k.kp;
  for (int e = threadIdx.x; e < 64 * nslot; e += blockDim.x) {
    const SlotDesc sd = slots[e];
    if (sd.kind != CK_PAIR) continue;
    const TableDesc T = a.tabs[sd.tab];
    float Sw[5];
    ks_slope_weights(a, T, sd.m, Sw);
    double val[kJP];
    const int K = T.off + sd.m;
    val[0] = E[K];
    for (int q = 0; q < 6; ++q) val[1 + q] = E[(2 + q) * k.kp + K];
    val[7] = 0.0;
    for (int q = 0; q < 5; ++q) {
      const int kk = T.off + min(max(sd.m - 2 + q, 0), T.n - 1);
      const double w = Sw[q];
      val[0] += w * E[k.kp + kk];
      for (int v = 0; v < 6; ++v) val[1 + v] += w * E[(8 + v) * k.kp + kk];
    }
    for (int t = t0; t < t1; ++t)
      for (int q = 0; q < kJP; ++q) a.part[(int64_t(t) * kJP + q) * a.D + sd.col] = t == t0 ? val[q] : 0.0;
  }
}

// H_i and Hs_i = F''(rho_i) H_i (0 for a clamped density) over the density
// columns [dc0, dc0 + 256), from the density value sums; one wave per atom
__global__ __launch_bounds__(256) void k_jac_dens_ks(JacArgs a, KsArgs k, int dc0) {
  extern __shared__ float4 smem[];
  const int wave = uni(threadIdx.x >> 6), lane = threadIdx.x & 63;
  float *sK = reinterpret_cast<float *>(smem) + wave * 2 * k.ld; // the atom's sums
  const int nq = (2 * k.ld) >> 2;
  int cdc[kCPL], cm[kCPL], cn[kCPL], coff[kCPL];
  float Sw[kCPL][5];
#pragma unroll
  for (int s = 0; s < kCPL; ++s) {
    const int dc = dc0 + lane + 64 * s;
    cdc[s] = -1;
    cm[s] = coff[s] = 0;
    cn[s] = 1;
#pragma unroll
    for (int q = 0; q < 5; ++q) Sw[s][q] = 0.0f;
    if (dc < a.Dd) {
      const ColDesc cd = a.cols[a.dens_c0 + dc];
      const TableDesc T = a.tabs[cd.tab];
      cdc[s] = dc;
      cm[s] = cd.m;
      cn[s] = T.n;
      int o = 0;
      for (int t = 0; t < cd.tab - a.npair; ++t) o += a.tabs[a.npair + t].n;
      coff[s] = o;
      ks_slope_weights(a, T, cd.m, Sw[s]);
    }
  }
  const int64_t gw = int64_t(blockIdx.x) * 4 + wave;
  const int64_t nw = int64_t(gridDim.x) * 4;
  for (int64_t i = uni64(gw); i < a.N; i += nw) {
    const float4 *src = reinterpret_cast<const float4 *>(k.dens + size_t(i) * 2 * k.ld);
    for (int e = lane; e < nq; e += 64) reinterpret_cast<float4 *>(sK)[e] = src[e];
    wave_sync();
    const float *Yb = sK, *Sb = sK + k.ld;
    const bool clamped = a.over[i] != 0.0;
    const float hs = clamped ? 0.0f : float(a.Fpp[i]);
#pragma unroll
    for (int s = 0; s < kCPL; ++s) {
      if (cdc[s] < 0) continue;
      const int m = cm[s], o = coff[s];
      float h = Yb[o + m];
#pragma unroll
      for (int q = 0; q < 5; ++q) h += Sw[s][q] * Sb[o + min(max(m - 2 + q, 0), cn[s] - 1)];
      a.H[i * a.Dd + cdc[s]] = h;
      a.Hs[i * a.Dd + cdc[s]] = hs * h;
    }
    wave_sync();
  }
}

// ---------------------------------------------------------------------------
// ADP Jacobian. The pair-table, density and embedding columns are those of the
// EAM passes unchanged (the dipole and quadrupole terms do not depend on the
// phi, rho and F tables). The dipole (u) and quadrupole (w) columns follow the
// chain of the reference's ADP parameter Jacobian
// (ADPForceCalculator::write_param_jacobian, src/force/adp_force.cpp:141-385)
// with spline band coefficients, in two kernels per kind:
//   k_adp_mom: per atom and column, d mu_i = sum_j du(r) d (dipole) or
//     d lambda_i = sum_j dw(r) d d (quadrupole), over the column's window of
//     the sorted pairs;
//   k_adp_dq: per atom and column, the force tangent of every pair term of
//     f_ij = (u'/r)(dmu . d) d + u dmu + (w'/r) nu(dLambda, d) d
//            + 2 w xi(dLambda, d)
//     (dmu = mu_i - mu_j, Lambda = lambda_i + lambda_j; adp_force.cpp:73-135):
//     the band terms of the column's own table over its window, and the
//     moment-tangent terms over every pair, with the neighbours' d mu_j /
//     d lambda_j gathered in batches; energy (mu_i . d mu_i and
//     lambda_i : d lambda_i - tr lambda_i tr d lambda_i / 3) and virial.
// ---------------------------------------------------------------------------
enum { JP_DIP = 3, JP_QUAD = 4 };
constexpr int kQBatch = 4; // neighbour moment tangents gathered per batch
// moment-tangent rows: [atom][column][comp], dipole as a float3 (12 B, one
// three-dword load), quadrupole as three float2, so that one lane reads its
// column's tangent with one or three loads and a wave's gather is one
// contiguous span

template <bool DIP> __device__ inline void store_moment(float *base, int64_t row, const float *v) {
  if constexpr (DIP) {
    reinterpret_cast<float3 *>(base)[row] = make_float3(v[0], v[1], v[2]);
  } else {
    float2 *d = reinterpret_cast<float2 *>(base) + 3 * row;
    d[0] = make_float2(v[0], v[1]);
    d[1] = make_float2(v[2], v[3]);
    d[2] = make_float2(v[4], v[5]);
  }
}

// bytes per moment-tangent entry (atom, column)
template <bool DIP> constexpr uint32_t moment_bytes() { return DIP ? 12u : 24u; }

// the same load from a 32-bit byte offset off the uniform table base (the
// table is < 4 GiB): one VALU op per address in the gather loops
template <bool DIP> __device__ inline void load_moment_at(const float *base, uint32_t off, float *v) {
  const char *b = reinterpret_cast<const char *>(base) + off;
  if constexpr (DIP) {
    const float3 q = *reinterpret_cast<const float3 *>(b);
    v[0] = q.x;
    v[1] = q.y;
    v[2] = q.z;
  } else {
    const float2 *d = reinterpret_cast<const float2 *>(b);
    const float2 a = d[0], e = d[1], c = d[2];
    v[0] = a.x;
    v[1] = a.y;
    v[2] = e.x;
    v[3] = e.y;
    v[4] = c.x;
    v[5] = c.y;
  }
}

template <bool DIP> __device__ inline void load_moment(const float *base, int64_t row, float *v) {
  if constexpr (DIP) {
    const float3 q = reinterpret_cast<const float3 *>(base)[row];
    v[0] = q.x;
    v[1] = q.y;
    v[2] = q.z;
  } else {
    const float2 *d = reinterpret_cast<const float2 *>(base) + 3 * row;
    const float2 a = d[0], b = d[1], c = d[2];
    v[0] = a.x;
    v[1] = a.y;
    v[2] = b.x;
    v[3] = b.y;
    v[4] = c.x;
    v[5] = c.y;
  }
}

struct MomStage {
  int32_t seg[kSegPad]; // partner-type segment starts of the chunk
  float d0[64], d1[64], d2[64];
  int32_t k[64];
  float c[6][kBandPitch];
};

// table and pair flag of the dipole (PASS == JP_DIP) or quadrupole pass
template <int PASS> __device__ inline int dq_tab0(const JacArgs &a) {
  return PASS == JP_DIP ? a.npair + 2 * a.ntypes : 2 * a.npair + 2 * a.ntypes;
}
template <int PASS> constexpr uint32_t dq_flag() { return PASS == JP_DIP ? kFlagDip : kFlagQuad; }

// partner type of a (ta, tb) pair-table column for centre type ti (-1: not ti's table)
__device__ inline int partner_type(int pt, int ti) {
  const int ta = pt & 0xff, tb = (pt >> 8) & 0xff;
  return ti == ta ? tb : (ti == tb ? ta : -1);
}

template <int PASS, int NS, int TS>
__global__ __launch_bounds__(256) void k_adp_mom(JacArgs a, int nknots, const SlotDesc *slots) {
  constexpr bool kDip = PASS == JP_DIP;
  constexpr int NC = kDip ? 3 : 6;
  extern __shared__ float4 smem[];
  TableDesc *sT;
  const float *sS;
  MomStage *sAll = reinterpret_cast<MomStage *>(stage<TS>(a, smem, sT, sS, nknots));
  const int wave = uni(threadIdx.x >> 6), lane = threadIdx.x & 63;
  MomStage &P = sAll[wave];
  int cm[NS], cqc[NS], cpt[NS];
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    const SlotDesc sd = slots[s * 64 + lane];
    cm[s] = sd.m;
    cqc[s] = sd.kind == CK_NONE ? -1 : sd.qc;
    cpt[s] = sd.dc;
  }
  const int tab0 = dq_tab0<PASS>(a);
  float *out = kDip ? a.dmu : a.dlam;
  const int Dq = kDip ? a.Ddip : a.Dquad;
  const int64_t gw = int64_t(blockIdx.x) * 4 + wave;
  const int64_t nw = int64_t(gridDim.x) * 4;
  for (int64_t i = gw; i < a.N; i += nw) {
    const int ti = uni(a.types[i]);
    const int64_t p0 = uni64(a.pair_off[i]), p1 = uni64(a.pair_off[i + 1]);
    int ptt[NS];
    float T[NS][NC];
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      ptt[s] = cqc[s] < 0 ? -1 : partner_type(cpt[s], ti);
#pragma unroll
      for (int q = 0; q < NC; ++q) T[s][q] = 0.0f;
    }
    for (int64_t pc = p0; pc < p1; pc += 64) {
      const int64_t p = pc + lane;
      int ty = kMaxTypes;
      if (p < p1) {
        const float4 q = a.jrec[p];
        const uint32_t b = __float_as_uint(q.w);
        const double r64 = a.jr[p];
        const int tj = pj_type(b);
        ty = tj;
        const TableDesc Tq = sT[tab0 + ord_pair(ti, tj, a.ntypes)];
        float bv[4], bd[4], c[6];
        const int k = pair_basis(Tq, r64, bv, bd);
        band6(sS, Tq, k, bv, ((b & dq_flag<PASS>()) && !(b & kFlagTiny)) ? 1.0f : 0.0f, c);
        P.k[lane] = k;
        P.d0[lane] = q.x;
        P.d1[lane] = q.y;
        P.d2[lane] = q.z;
#pragma unroll
        for (int d = 0; d < 6; ++d) P.c[d][lane] = c[d];
      }
      stage_segments(ty, a.ntypes, lane, P.seg);
      const int32_t *seg = P.seg;
      wave_sync();
#pragma unroll
      for (int s = 0; s < NS; ++s) {
        if (ptt[s] < 0) continue;
        const int m = cm[s], s1 = seg[ptt[s] + 1];
        const int lo = first_ge(P.k, seg[ptt[s]], s1, m - 3);
        const int hi = first_ge(P.k, lo, s1, m + 3);
        for (int q = lo; q < hi; ++q) {
          const float v = P.c[m - P.k[q] + 2][q], x = P.d0[q], y = P.d1[q], z = P.d2[q];
          if constexpr (kDip) {
            T[s][0] += v * x;
            T[s][1] += v * y;
            T[s][2] += v * z;
          } else {
            T[s][0] += v * (x * x);
            T[s][1] += v * (y * y);
            T[s][2] += v * (z * z);
            T[s][3] += v * (x * y);
            T[s][4] += v * (x * z);
            T[s][5] += v * (y * z);
          }
        }
      }
      wave_sync();
    }
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      if (cqc[s] < 0) continue;
      store_moment<kDip>(out, i * Dq + cqc[s], T[s]);
    }
  }
}

struct DQStage {
  int32_t seg[kSegPad]; // partner-type segment starts of the chunk
  float x[64], y[64], z[64];
  float g0[64], g1[64], g2[64]; // dipole: mu_i - mu_j; quadrupole: xi(Lambda, d)
  float sg[64];                 // dipole: (mu_i - mu_j) . d / r; quadrupole: nu(Lambda, d) / r
  float v[64], dv[64];          // u (w) and u'/r (w'/r); 0 when the role is off
  uint32_t jb[64];              // byte offset of atom j's moment-tangent row
  int32_t k[64];
  float cd[6][kBandPitch], cv[6][kBandPitch]; // derivative and value bands
};

// nu(L, d) and xi(L, d) of a symmetric L stored as (xx, yy, zz, xy, xz, yz)
// (quad_nu / quad_xi, adp_force.hpp:63-68)
__device__ inline float qnu(const float *L, float x, float y, float z, float xi[3]) {
  const float Mx = L[0] * x + L[3] * y + L[4] * z;
  const float My = L[3] * x + L[1] * y + L[5] * z;
  const float Mz = L[4] * x + L[5] * y + L[2] * z;
  const float tr3 = (L[0] + L[1] + L[2]) / 3.0f;
  xi[0] = Mx - tr3 * x;
  xi[1] = My - tr3 * y;
  xi[2] = Mz - tr3 * z;
  return (x * Mx + y * My + z * Mz) - (x * x + y * y + z * z) * tr3;
}

__device__ inline void add_dvirial(float *t, float x, float y, float z, float f0, float f1, float f2) {
  t[1] += x * f0;
  t[2] += y * f1;
  t[3] += z * f2;
  t[4] += x * f1;
  t[5] += x * f2;
  t[6] += y * f2;
}

// waves per SIMD asked of the moment-column force passes: 6 for the dipole
// pass (80 VGPRs, a few values in scratch) and 5 for the quadrupole pass (96,
// none); the compiler's own choices (83 and 113 VGPRs) allow 5 and 4, and
// measured 0.67-0.71 against 0.75 ms for the quadrupole pass of the binary
// 200-configuration workload
template <int PASS, int NS, int TS>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(PASS == JP_DIP ? 6 : 5))) void
k_adp_dq(JacArgs a, int nknots, const SlotDesc *slots) {
  constexpr bool kDip = PASS == JP_DIP;
  constexpr int NC = kDip ? 3 : 6;
  extern __shared__ float4 smem[];
  TableDesc *sT;
  const float *sS;
  DQStage *sAll = reinterpret_cast<DQStage *>(stage<TS>(a, smem, sT, sS, nknots));
  static_assert(4 * sizeof(DQStage) >= sizeof(double) * kJP * 64 * NS, "reduction buffer");
  double(*sPart)[64 * NS] = reinterpret_cast<double(*)[64 * NS]>(sAll);
  const int wave = uni(threadIdx.x >> 6), lane = threadIdx.x & 63;
  DQStage &P = sAll[wave];
  int ccol[NS], cm[NS], cqc[NS], cpt[NS];
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    const SlotDesc sd = slots[s * 64 + lane];
    ccol[s] = sd.col;
    cm[s] = sd.m;
    cqc[s] = sd.kind == CK_NONE ? -1 : sd.qc;
    cpt[s] = sd.dc;
  }
  const int tab0 = dq_tab0<PASS>(a);
  const float *mom = kDip ? a.dmu : a.dlam;
  const int Dq = kDip ? a.Ddip : a.Dquad;
  // XCD-aware tiles (as in the EAM passes): the neighbours' moment tangents
  // stay in the XCD's L2
  const int ng = min(kXcds, int(gridDim.x));
  const int xcd = blockIdx.x % ng, bslot = blockIdx.x / ng;
  const int nslot = (int(gridDim.x) - xcd + ng - 1) / ng;
  const int t_end = int(int64_t(xcd + 1) * a.ntiles / ng);
  for (int tile = int(int64_t(xcd) * a.ntiles / ng) + bslot; tile < t_end; tile += nslot) {
    const int na = uni(a.tile_na[tile]);
    const int c = uni(a.tile_cfg[tile]);
    const int64_t cbase = uni64(a.atom_off[c]);
    float tp[NS][kJP];
#pragma unroll
    for (int s = 0; s < NS; ++s)
#pragma unroll
      for (int q = 0; q < kJP; ++q) tp[s][q] = 0.0f;
    for (int li = wave; li < na; li += 4) {
      const int64_t i = int64_t(uni(a.tile_a0[tile])) + li;
      const int ti = uni(a.types[i]);
      const int64_t p0 = uni64(a.pair_off[i]), p1 = uni64(a.pair_off[i + 1]);
      // the atom's own moment (mu_i or lambda_i) and moment tangents
      float Mi[NC];
#pragma unroll
      for (int q = 0; q < NC; ++q) Mi[q] = float(kDip ? a.mu[3 * i + q] : a.lamb[6 * i + q]);
      int ptt[NS];
      float own[NS][NC], F[NS][3];
#pragma unroll
      for (int s = 0; s < NS; ++s) {
        ptt[s] = cqc[s] < 0 ? -1 : partner_type(cpt[s], ti);
        if (cqc[s] >= 0) load_moment<kDip>(mom, i * Dq + cqc[s], own[s]);
        else
#pragma unroll
          for (int q = 0; q < NC; ++q) own[s][q] = 0.0f;
        F[s][0] = F[s][1] = F[s][2] = 0.0f;
        // energy tangent: mu_i . d mu_i, or lambda_i : d lambda_i - tr tr / 3
        float e;
        if constexpr (kDip) {
          e = Mi[0] * own[s][0] + Mi[1] * own[s][1] + Mi[2] * own[s][2];
        } else {
          const float *dl = own[s];
          const float dot = Mi[0] * dl[0] + Mi[1] * dl[1] + Mi[2] * dl[2] +
                            2.0f * (Mi[3] * dl[3] + Mi[4] * dl[4] + Mi[5] * dl[5]);
          e = dot - (Mi[0] + Mi[1] + Mi[2]) * (dl[0] + dl[1] + dl[2]) / 3.0f;
        }
        tp[s][0] += e;
      }
      for (int64_t pc = p0; pc < p1; pc += 64) {
        const int64_t p = pc + lane;
        int ty = kMaxTypes;
        if (p < p1) {
          const float4 q = a.jrec[p];
          const uint32_t b = __float_as_uint(q.w);
          const double r64 = a.jr[p];
          const float rr = float(r64);
          const bool on = (b & dq_flag<PASS>()) && !(b & kFlagTiny);
          const int tj = pj_type(b);
          ty = tj;
          const int32_t jg = int32_t(cbase + pj_local(b));
          const TableDesc Tq = sT[tab0 + ord_pair(ti, tj, a.ntypes)];
          float bv[4], bd[4], cc[6], ce[6];
          const int k = pair_basis(Tq, r64, bv, bd);
          band6(sS, Tq, k, bd, on ? 1.0f : 0.0f, cc);
          band6(sS, Tq, k, bv, on ? 1.0f : 0.0f, ce);
          float v = 0.0f, dv = 0.0f;
          if (on) {
            float tt;
            const int kk = pair_interval(Tq, r64, tt);
            spl_eval32(Tq, a.ivl32, kk, tt, v, dv);
          }
          const float ir = on ? 1.0f / rr : 0.0f;
          P.x[lane] = q.x;
          P.y[lane] = q.y;
          P.z[lane] = q.z;
          P.jb[lane] = uint32_t(jg) * uint32_t(Dq) * moment_bytes<kDip>();
          P.k[lane] = k;
          P.v[lane] = v;
          P.dv[lane] = dv * ir;
#pragma unroll
          for (int d = 0; d < 6; ++d) {
            P.cd[d][lane] = cc[d];
            P.cv[d][lane] = ce[d];
          }
          if constexpr (kDip) {
            const float g0 = Mi[0] - float(a.mu[3 * int64_t(jg)]);
            const float g1 = Mi[1] - float(a.mu[3 * int64_t(jg) + 1]);
            const float g2 = Mi[2] - float(a.mu[3 * int64_t(jg) + 2]);
            P.g0[lane] = g0;
            P.g1[lane] = g1;
            P.g2[lane] = g2;
            P.sg[lane] = (g0 * q.x + g1 * q.y + g2 * q.z) * ir;
          } else {
            float Lj[6], xa[3], xb[3];
#pragma unroll
            for (int t = 0; t < 6; ++t) Lj[t] = float(a.lamb[6 * int64_t(jg) + t]);
            const float nu = qnu(Mi, q.x, q.y, q.z, xa) + qnu(Lj, q.x, q.y, q.z, xb);
            P.g0[lane] = xa[0] + xb[0];
            P.g1[lane] = xa[1] + xb[1];
            P.g2[lane] = xa[2] + xb[2];
            P.sg[lane] = nu * ir;
          }
        }
        stage_segments(ty, a.ntypes, lane, P.seg);
        const int32_t *seg = P.seg;
        wave_sync();
        const int np = int(min<int64_t>(64, p1 - pc));
#pragma unroll
        for (int s = 0; s < NS; ++s) {
          if (cqc[s] < 0) continue;
          // band terms of the column's own table: its window of the pairs
          // with the matching partner type
          if (ptt[s] >= 0) {
            const int m = cm[s], s1 = seg[ptt[s] + 1];
            const int lo = first_ge(P.k, seg[ptt[s]], s1, m - 3);
            const int hi = first_ge(P.k, lo, s1, m + 3);
            for (int q = lo; q < hi; ++q) {
              const int d = m - P.k[q] + 2;
              const float sc = P.cd[d][q] * P.sg[q];
              const float vv = (kDip ? 1.0f : 2.0f) * P.cv[d][q];
              const float x = P.x[q], y = P.y[q], z = P.z[q];
              const float f0 = sc * x + vv * P.g0[q], f1 = sc * y + vv * P.g1[q],
                          f2 = sc * z + vv * P.g2[q];
              F[s][0] += f0;
              F[s][1] += f1;
              F[s][2] += f2;
              add_dvirial(tp[s], x, y, z, f0, f1, f2);
            }
          }
          // moment-tangent terms: every pair, the neighbours' tangents
          // gathered kQBatch pairs at a time
          auto pair_term = [&](int q, const float *hq) {
            const float x = P.x[q], y = P.y[q], z = P.z[q];
            float f0, f1, f2;
            if constexpr (kDip) {
              const float gx = own[s][0] - hq[0], gy = own[s][1] - hq[1], gz = own[s][2] - hq[2];
              const float sc = P.dv[q] * (gx * x + gy * y + gz * z), vv = P.v[q];
              f0 = sc * x + vv * gx;
              f1 = sc * y + vv * gy;
              f2 = sc * z + vv * gz;
            } else {
              float dl[6], xq[3];
#pragma unroll
              for (int t = 0; t < 6; ++t) dl[t] = own[s][t] + hq[t];
              const float sc = P.dv[q] * qnu(dl, x, y, z, xq), w2 = 2.0f * P.v[q];
              f0 = sc * x + w2 * xq[0];
              f1 = sc * y + w2 * xq[1];
              f2 = sc * z + w2 * xq[2];
            }
            F[s][0] += f0;
            F[s][1] += f1;
            F[s][2] += f2;
            add_dvirial(tp[s], x, y, z, f0, f1, f2);
          };
          const uint32_t colb = uint32_t(cqc[s]) * moment_bytes<kDip>();
          int q0 = 0;
          for (; q0 + kQBatch <= np; q0 += kQBatch) {
            float h[kQBatch][NC];
#pragma unroll
            for (int u = 0; u < kQBatch; ++u) load_moment_at<kDip>(mom, P.jb[q0 + u] + colb, h[u]);
#pragma unroll
            for (int u = 0; u < kQBatch; ++u) pair_term(q0 + u, h[u]);
          }
          for (; q0 < np; ++q0) {
            float h[NC];
            load_moment_at<kDip>(mom, P.jb[q0] + colb, h);
            pair_term(q0, h);
          }
        }
        wave_sync();
      }
      const int64_t row = a.row_off[c] + 3 * (i - cbase);
#pragma unroll
      for (int s = 0; s < NS; ++s) {
        if (cqc[s] < 0) continue;
        for (int q = 0; q < 3; ++q) a.J[(row + q) * a.D + ccol[s]] = F[s][q];
      }
    }
    // reduce the four waves' partials in wave order (FP64), write the tile
    // partial of this pass's columns
    __syncthreads();
    for (int w = 0; w < 4; ++w) {
      if (wave == w) {
#pragma unroll
        for (int s = 0; s < NS; ++s) {
          if (cqc[s] < 0) continue;
          const int lc = 64 * s + lane;
          for (int q = 0; q < kJP; ++q) sPart[q][lc] = (w == 0 ? 0.0 : sPart[q][lc]) + double(tp[s][q]);
        }
      }
      __syncthreads();
    }
    if (wave == 0) {
#pragma unroll
      for (int s = 0; s < NS; ++s) {
        if (cqc[s] < 0) continue;
        const int col = ccol[s], lc = 64 * s + lane;
        for (int q = 0; q < kJP; ++q) a.part[(int64_t(tile) * kJP + q) * a.D + col] = sPart[q][lc];
      }
    }
    __syncthreads();
  }
}

// energy / stress / limit rows of each configuration
__global__ void k_jac_rows(JacArgs a) {
  const int c = blockIdx.x;
  const int64_t na = a.atom_off[c + 1] - a.atom_off[c];
  const int64_t e_row = a.row_off[c] + 3 * na;
  const bool ws = a.sw > 0.0;
  const double iv = 1.0 / a.vol[c];
  // block (c, q) sums partial row q of configuration c, one thread per
  // column: a wave reads consecutive doubles of each tile, four tiles' loads
  // are in flight before their sums, which are added in tile order
  const int t0 = a.cfg_tile0[c], t1 = a.cfg_tile0[c + 1];
  const int nqc = kJP * a.D;
  const int q = blockIdx.y;
  if (!ws && q >= 1 && q <= 6) return;
  for (int col = threadIdx.x; col < a.D; col += blockDim.x) {
    const int e = q * a.D + col;
    const double *p = a.part + int64_t(t0) * nqc + e;
    double acc = 0.0;
    int t = t0;
    for (; t + 4 <= t1; t += 4, p += 4 * int64_t(nqc)) {
      const double v0 = p[0], v1 = p[nqc], v2 = p[2 * int64_t(nqc)], v3 = p[3 * int64_t(nqc)];
      acc += v0;
      acc += v1;
      acc += v2;
      acc += v3;
    }
    for (; t < t1; ++t, p += nqc) acc += p[0];
    float v;
    int64_t row = e_row;
    if (q == 0) v = float(a.ew * acc);
    else if (q == 7) {
      v = float(acc);
      row += ws ? 7 : 1;
    } else {
      v = float(-0.5 * a.sw * acc * iv);
      row += q;
    }
    a.J[row * a.D + col] = v;
  }
}

// ---------------------------------------------------------------------------
// Stillinger-Weber tangents
// ---------------------------------------------------------------------------
struct SwJArgs {
  int64_t N;
  int D, ntypes, npair;
  const int64_t *pair_off;
  const int32_t *types;
  const double *pd;
  const uint32_t *bits;
  const int32_t *rev;
  const double *sw, *lam;
  const double *pv;       // 8 per pair: v2, v2', h, h' and the gamma / a2 tangents of h, h'
  const double *tri;      // 8 per pair: unit vector, 1/r, h, h', dh/dgamma, dh/da2
  double2 *pt2;           // [6][P]: (tv, tdv / r) of v2 w.r.t. each two-body field (k_sw_pair_terms)
  const int32_t *col_sw;  // npair*8 -> column or -1
  const int32_t *col_lam; // nlam -> column or -1
  const int32_t *tidx;    // [D]: the column's triplet pass, or -1 for a two-body column
  const int32_t *tcol;    // [nTri]: the column of each triplet pass
  const int32_t *tsf;     // [D]: 8 slot + field of a two-body column, -1 for a triplet column
  double *own;            // [nTri][N][3]
  double *ptan;           // [nTri][P][3]
  double *etan;           // [nTri][N][10]: energy + 9 virial
  int64_t P;
  // tiles
  int ntiles;
  const int32_t *tile_cfg, *tile_a0, *tile_na;
  const int64_t *atom_off, *row_off;
  float *J;
  double *part;
};

constexpr int kSwJLanes = 16; // lanes per centre atom (a divisor of 64)
constexpr int kSwStage = 32;  // neighbour records per centre staged in LDS

__device__ inline double sw_group_sum(double v) {
#pragma unroll
  for (int o = kSwJLanes / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Tangents of the SW forces, energy and virial of one centre atom w.r.t. one
// column (blockIdx.y). A group of kSwJLanes lanes takes the centre; the lane of
// neighbour j forms the tangent of the pair term of (i, j) and of the triplet
// force on j summed over every other neighbour k, T_j, which it writes into
// the per-pair slot of (i -> j). The centre's tangent force is the pair terms
// minus sum_j T_j (a triplet's forces sum to zero for every parameter value,
// so their tangents do too) and the triplet virial tangent sum_j d_j (x) T_j.
//
// The j force of a triplet (three_body_forces, kernels.hpp:40-56) in unit
// vectors u1, u2 and cosine c is -f1' f2 g u1 - f1 f2 g' (u2 - c u1) / r_ij,
// multilinear in the factor groups (f1, f1'), f2 and (g, g'): its tangent is
// the sum of the three forms with one group replaced by its tangent, whose
// u1 and (u2 - c u1) coefficients are summed over k as scalars (Ss, Sqc) and
// one vector (V), giving T_j once per bond. FP64 throughout; the bond's unit
// vector and 1/r come from the per-pair records, so the loop over k has no
// square root or division.
__global__ __launch_bounds__(256) void k_sw_jac_centre(SwJArgs a) {
  constexpr int kGroups = 256 / kSwJLanes;
  // per centre, the first kSwStage neighbours' records: unit vector and h,
  // the column's tangent of h, and the neighbour's type (-1: no triplets);
  // the lambda table and which lambda is the column
  __shared__ double4 s_u[kGroups][kSwStage];
  __shared__ double s_t[kGroups][kSwStage];
  __shared__ int s_ty[kGroups][kSwStage];
  __shared__ double s_lam[kMaxTypes * kMaxTypes * (kMaxTypes + 1) / 2];
  __shared__ int s_lamc[kMaxTypes * kMaxTypes * (kMaxTypes + 1) / 2];
  const int64_t g = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t i = g / kSwJLanes;
  const int l = int(g % kSwJLanes);
  const int grp = int(threadIdx.x) / kSwJLanes;
  const int tp = blockIdx.y, col = a.tcol[tp]; // a triplet column: gamma or a2 of a slot, or a lambda
  const bool live = i < a.N; // every lane takes part in the group sums
  double F[3] = {0, 0, 0}, E = 0.0, sg[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  double *pt = a.ptan + int64_t(tp) * a.P * 3;
  const int ti = live ? a.types[i] : 0;
  const int64_t p0 = live ? a.pair_off[i] : 0, p1 = live ? a.pair_off[i + 1] : 0;
  // record of neighbour pk of this centre: unit vector, h, the column's
  // tangent of h, type (-1 when the pair takes no part in triplets)
  auto record = [&](int64_t pk, double *u, double &h, double &t, int &ty) {
    const uint32_t bk = a.bits[pk];
    const double *r = a.tri + 8 * pk;
    u[0] = r[0];
    u[1] = r[1];
    u[2] = r[2];
    h = r[4];
    ty = (bk & kFlagTiny) || (r[4] == 0.0 && r[5] == 0.0) ? -1 : pj_type(bk);
    t = 0.0;
    if (ty >= 0) {
      const int s2 = ord_pair(ti, ty, a.ntypes);
      if (a.col_sw[8 * s2 + 6] == col) t += r[6];
      if (a.col_sw[8 * s2 + 7] == col) t += r[7];
    }
  };
  for (int q = threadIdx.x; q < a.ntypes * a.npair; q += blockDim.x) {
    s_lam[q] = a.lam[q];
    s_lamc[q] = a.col_lam[q] == col;
  }
  const int ns = int(std::min<int64_t>(p1 - p0, kSwStage));
  for (int k = l; k < ns; k += kSwJLanes) {
    double u[3], h, t;
    int ty;
    record(p0 + k, u, h, t, ty);
    s_u[grp][k] = make_double4(u[0], u[1], u[2], h);
    s_t[grp][k] = t;
    s_ty[grp][k] = ty;
  }
  __syncthreads();
  if (live) {
    for (int64_t pj = p0 + l; pj < p1; pj += kSwJLanes) {
      double T[3] = {0, 0, 0};
      const uint32_t bj = a.bits[pj];
      if (!(bj & kFlagTiny)) {
        const int tj = pj_type(bj);
        const int s1 = ord_pair(ti, tj, a.ntypes);
        const double *d1 = a.pd + 3 * pj;
        const double *r1p = a.tri + 8 * pj; // unit vector, 1/r, h, h', ...
        const double *q1 = a.pv + 8 * pj;
        const double h1 = q1[2], dh1 = q1[3];
        if (h1 != 0.0 || dh1 != 0.0) {
          // tangent of (h1, h1') when the column is gamma or a2 of slot s1
          double t1 = 0.0, td1 = 0.0;
          if (a.col_sw[8 * s1 + 6] == col) {
            t1 += q1[4];
            td1 += q1[5];
          }
          if (a.col_sw[8 * s1 + 7] == col) {
            t1 += q1[6];
            td1 += q1[7];
          }
          const double u1[3] = {r1p[0], r1p[1], r1p[2]}, i1 = r1p[3];
          double Ss = 0.0, Sqc = 0.0, Se = 0.0, V[3] = {0.0, 0.0, 0.0};
          for (int64_t pk = p0; pk < p1; ++pk) {
            if (pk == pj) continue;
            const int k = int(pk - p0);
            double u2[3], h2, t2;
            int tk;
            if (k < kSwStage) {
              const double4 U = s_u[grp][k];
              u2[0] = U.x;
              u2[1] = U.y;
              u2[2] = U.z;
              h2 = U.w;
              t2 = s_t[grp][k];
              tk = s_ty[grp][k];
            } else {
              record(pk, u2, h2, t2, tk);
            }
            if (tk < 0) continue;
            const int li = ti * a.npair + ord_pair(tj, tk, a.ntypes);
            const bool lam_col = s_lamc[li] != 0;
            if (t1 == 0.0 && td1 == 0.0 && t2 == 0.0 && !lam_col) continue;
            const double lam = s_lam[li];
            const double cs = u1[0] * u2[0] + u1[1] * u2[1] + u1[2] * u2[2];
            const double cp = cs + 1.0 / 3.0;
            const double w = lam * cp * cp, dw = 2.0 * lam * cp;
            const double tg = lam_col ? cp * cp : 0.0, tdg = lam_col ? 2.0 * cp : 0.0;
            // groups (t1, td1) for (f1, f1'), t2 for f2, (tg, tdg) for (g, g')
            const double sk = td1 * h2 * w + dh1 * (t2 * w + h2 * tg);
            const double qk = t1 * h2 * dw + h1 * (t2 * dw + h2 * tdg);
            Ss += sk;
            Sqc += qk * cs;
            for (int x = 0; x < 3; ++x) V[x] += qk * u2[x];
            Se += t1 * h2 * w + h1 * (t2 * w + h2 * tg);
          }
          const double cu = Ss - i1 * Sqc;
          for (int x = 0; x < 3; ++x) T[x] = -cu * u1[x] - i1 * V[x];
          E += 0.5 * Se; // each triplet is met from both of its bonds
          for (int x = 0; x < 3; ++x) F[x] -= T[x];
          for (int x = 0; x < 3; ++x)
            for (int y = 0; y < 3; ++y) sg[3 * x + y] += d1[x] * T[y];
        }
      }
      for (int x = 0; x < 3; ++x) pt[3 * pj + x] = T[x];
    }
  }
  for (int x = 0; x < 3; ++x) F[x] = sw_group_sum(F[x]);
  E = sw_group_sum(E);
  for (int k = 0; k < 9; ++k) sg[k] = sw_group_sum(sg[k]);
  if (live && l == 0) {
    double *o = a.own + (int64_t(tp) * a.N + i) * 3;
    for (int x = 0; x < 3; ++x) o[x] = F[x];
    double *e = a.etan + (int64_t(tp) * a.N + i) * 10;
    e[0] = E;
    for (int k = 0; k < 9; ++k) e[1 + k] = sg[k];
  }
}

// one block per tile, x = atoms of the tile, y = column
// one block per tile, x = atoms of the tile, y = column. A triplet column
// gathers, per atom, its own tangent force and the triplet slots of its
// reverse pairs, and takes the energy and virial tangents of the centre pass;
// a two-body column (a field of v2, no triplet term) forms the pair terms of
// the atom's own pairs here: force (v2' tangent / r) d, energy half the v2
// tangent, virial -d (x) force / 2.
__global__ __launch_bounds__(64) void k_sw_jac_tiles(SwJArgs a) {
  __shared__ double red[64][kJP];
  const int tile = blockIdx.x, col = blockIdx.y, t = threadIdx.x;
  const int na = a.tile_na[tile];
  const int tp = a.tidx[col];
  // kLpt lanes per atom (the same for every tile size, so an atom's sum does
  // not depend on the tiling): lane sub of atom li takes every kLpt-th pair
  // or reverse-pair slot, and the lanes' sums are added by shuffles within
  // the aligned group; 64 / kLpt atoms per round
  constexpr int kLpt = 4;
  const int sub = t % kLpt;
  const int vo[6] = {0, 4, 8, 1, 2, 5}; // xx yy zz xy xz yz
  for (int l0 = 0; l0 < na; l0 += 64 / kLpt) { // uniform trip count: every lane shuffles
    const int li = l0 + t / kLpt;
    const int64_t i = li < na ? int64_t(a.tile_a0[tile]) + li : 0;
    double F[3] = {0, 0, 0}, E = 0.0, W[6] = {0, 0, 0, 0, 0, 0};
    if (tp >= 0) {
      const double *pt = a.ptan + int64_t(tp) * a.P * 3;
      if (li < na) {
        for (int64_t p = a.pair_off[i] + sub; p < a.pair_off[i + 1]; p += kLpt) {
          const int32_t q = a.rev[p];
          if (q < 0) continue;
          for (int k = 0; k < 3; ++k) F[k] += pt[3 * int64_t(q) + k];
        }
      }
    } else if (li < na) {
      // the column is field f of slot s: the pairs of that slot carry it
      const int sf = a.tsf[col], slot = sf >> 3;
      const double2 *tf = a.pt2 + int64_t(sf & 7) * a.P;
      const int ti = a.types[i];
      for (int64_t p = a.pair_off[i] + sub; p < a.pair_off[i + 1]; p += kLpt) {
        const uint32_t b = a.bits[p];
        if ((b & kFlagTiny) || ord_pair(ti, pj_type(b), a.ntypes) != slot) continue;
        const double2 tt = tf[p];
        const double d[3] = {a.pd[3 * p], a.pd[3 * p + 1], a.pd[3 * p + 2]};
        const double df[3] = {tt.y * d[0], tt.y * d[1], tt.y * d[2]};
        for (int x = 0; x < 3; ++x) F[x] += df[x];
        E += 0.5 * tt.x;
        for (int v = 0; v < 6; ++v) W[v] += d[vo[v] / 3] * df[vo[v] % 3];
      }
    }
    for (int o = kLpt / 2; o > 0; o >>= 1) {
      for (int k = 0; k < 3; ++k) F[k] += __shfl_xor(F[k], o, 64);
      E += __shfl_xor(E, o, 64);
      for (int v = 0; v < 6; ++v) W[v] += __shfl_xor(W[v], o, 64);
    }
    if (li < na && sub == 0) {
      const int c = a.tile_cfg[tile];
      if (tp >= 0) {
        const double *o = a.own + (int64_t(tp) * a.N + i) * 3;
        for (int x = 0; x < 3; ++x) F[x] += o[x];
        const double *e = a.etan + (int64_t(tp) * a.N + i) * 10;
        E = e[0];
        // stored as -2 x virial so the row kernel's -w_S/(2V) scaling yields w_S * virial / V
        for (int v = 0; v < 6; ++v) W[v] = -2.0 * e[1 + vo[v]];
      }
      const int64_t row = a.row_off[c] + 3 * (i - a.atom_off[c]);
      for (int x = 0; x < 3; ++x) a.J[(row + x) * a.D + col] = float(F[x]);
      red[li][0] = E;
      for (int v = 0; v < 6; ++v) red[li][1 + v] = W[v];
      red[li][7] = 0.0;
    }
  }
  __syncthreads();
  if (t < kJP) {
    double s = 0.0;
    for (int q = 0; q < na; ++q) s += red[q][t];
    a.part[(int64_t(tile) * kJP + t) * a.D + col] = s;
  }
}

inline unsigned nb(int64_t n, int b) { return unsigned((n + b - 1) / b); }

// curvature rows: constant w (1, -2, 1) on the free knots k-1, k, k+1, zero
// elsewhere; one block per row
__global__ void k_smooth_jac(int D, const int4 *__restrict__ rows, float w, float *__restrict__ J) {
  const int4 rw = rows[blockIdx.x];
  float *row = J + int64_t(blockIdx.x) * D;
  for (int col = threadIdx.x; col < D; col += blockDim.x) {
    float v = 0.0f;
    if (col == rw.y) v = w;
    if (col == rw.z) v = -2.0f * w;
    if (col == rw.w) v = w;
    row[col] = v;
  }
}

} // namespace

void launch_smooth_jacobian(Context &c) {
  if (c.n_smooth == 0 || c.D == 0) return;
  hipLaunchKernelGGL(k_smooth_jac, dim3(unsigned(c.n_smooth)), dim3(256), 0, c.stream, c.D,
                     c.smooth_rows.p, float(c.smooth_weight), c.J.p + size_t(c.R_cfg) * c.D);
  FS_HIP(hipGetLastError());
}

// FP32 copy of the slope bands for passes that read them in place
__global__ void k_band_f32(int64_t n, const double *__restrict__ S, float *__restrict__ S32) {
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e < n) S32[e] = float(S[e]);
}

// slots padded to whole launches of w slots
inline int round_slots(int n, int w) { return w == 0 ? 0 : (n + w - 1) / w * w; }

void launch_jacobian(Context &c) {
  const int64_t R = c.R;
  c.J.alloc(size_t(std::max<int64_t>(R * c.D, 1)));
  // Jacobian tiles: <= 64 consecutive atoms of one configuration
  if (c.njt == 0) {
    // tile size: 64 atoms when there are at least 32 per wave slot of the chip
    // (16 waves per CU), else the power of two (>= 4) that gives about 16
    // waves per CU (the ADP and SW tile kernels run one wave per tile)
    const int64_t want = c.N / std::max<int64_t>(1, int64_t(c.cu_count) * 16);
    int jt = 4;
    while (jt < kJacTile && jt < want) jt *= 2;
    if (want >= kJacTile / 2) jt = kJacTile;
    if (c.jac_tile > 0) jt = c.jac_tile;
    std::vector<int32_t> tcfg, ta0, tna, ct0(c.C + 1, 0);
    for (int k = 0; k < c.C; ++k) {
      ct0[k] = int32_t(tcfg.size());
      for (int64_t a = c.h_atom_off[k]; a < c.h_atom_off[k + 1]; a += jt) {
        tcfg.push_back(k);
        ta0.push_back(int32_t(a));
        tna.push_back(int32_t(std::min<int64_t>(jt, c.h_atom_off[k + 1] - a)));
      }
    }
    ct0[c.C] = int32_t(tcfg.size());
    const int n = int(tcfg.size());
    c.jt_cfg.alloc(std::max(n, 1));
    c.jt_a0.alloc(std::max(n, 1));
    c.jt_na.alloc(std::max(n, 1));
    c.jt_ct0.alloc(c.C + 1);
    FS_HIP(hipMemcpyAsync(c.jt_cfg.p, tcfg.data(), 4 * n, hipMemcpyHostToDevice, c.stream));
    FS_HIP(hipMemcpyAsync(c.jt_a0.p, ta0.data(), 4 * n, hipMemcpyHostToDevice, c.stream));
    FS_HIP(hipMemcpyAsync(c.jt_na.p, tna.data(), 4 * n, hipMemcpyHostToDevice, c.stream));
    FS_HIP(hipMemcpyAsync(c.jt_ct0.p, ct0.data(), 4 * (c.C + 1), hipMemcpyHostToDevice, c.stream));
    FS_HIP(hipStreamSynchronize(c.stream));
    c.njt = n;
    c.jt_atoms = jt;
  }
  const int nt = c.njt;
  c.jac_part.alloc(size_t(std::max(nt, 1)) * kJP * std::max(c.D, 1));

  JacArgs a{};
  a.N = c.N;
  a.C = c.C;
  a.D = c.D;
  a.ntypes = c.ntypes;
  a.npair = c.npair;
  a.ntab = c.ntab;
  a.ntiles = nt;
  a.tile_cfg = c.jt_cfg.p;
  a.tile_a0 = c.jt_a0.p;
  a.tile_na = c.jt_na.p;
  a.cfg_tile0 = c.jt_ct0.p;
  a.atom_off = c.atom_off.p;
  a.pair_off = c.pair_off.p;
  a.row_off = c.row_off.p;
  a.types = c.types.p;
  a.rec = c.pair_rec.p;
  a.bits = c.pair_bits.p;
  a.J = c.J.p;
  a.part = c.jac_part.p;
  a.vol = c.vol.p;
  a.ew = c.energy_weight;
  a.sw = c.stress_weight;

  if (c.kind == FS_SW) {
    SwJArgs s{};
    s.N = c.N;
    s.D = c.D;
    s.ntypes = c.ntypes;
    s.npair = c.npair;
    s.pair_off = c.pair_off.p;
    s.types = c.types.p;
    s.pd = c.pair_d.p;
    s.bits = c.pair_bits.p;
    s.rev = c.rev.p;
    s.sw = c.sw.p;
    s.lam = c.lam.p;
    launch_sw_pair_terms(c, true);
    s.pv = c.sw_pv.p;
    s.tri = c.sw_tri.p;
    s.pt2 = c.sw_pt2.p;
    std::vector<int32_t> csw(8 * c.npair), clam(c.h_lam.size());
    int o = 0;
    for (size_t k = 0; k < csw.size(); ++k) csw[k] = c.h_sw_fixed[k] ? -1 : o++;
    for (size_t k = 0; k < clam.size(); ++k) clam[k] = c.h_lam_fixed[k] ? -1 : o++;
    c.sw_csw.alloc(csw.size());
    c.sw_clam.alloc(std::max<size_t>(clam.size(), 1));
    FS_HIP(hipMemcpyAsync(c.sw_csw.p, csw.data(), 4 * csw.size(), hipMemcpyHostToDevice, c.stream));
    FS_HIP(hipMemcpyAsync(c.sw_clam.p, clam.data(), 4 * clam.size(), hipMemcpyHostToDevice, c.stream));
    s.col_sw = c.sw_csw.p;
    s.col_lam = c.sw_clam.p;
    // triplet columns (gamma or a2 of a slot, or a lambda) take the centre
    // pass; the others are two-body fields, formed in the tile pass
    // [D] triplet pass of each column, [D] column of each triplet pass,
    // [D] 8 slot + field of each two-body column
    const size_t Dm = size_t(std::max(c.D, 1));
    std::vector<int32_t> tmap(3 * Dm, -1);
    int ntri = 0;
    for (int col = 0; col < c.D; ++col) {
      bool t = false;
      for (int q = 0; q < c.npair; ++q) t = t || csw[8 * q + 6] == col || csw[8 * q + 7] == col;
      for (int32_t v : clam) t = t || v == col;
      if (t) {
        tmap[col] = ntri;
        tmap[Dm + ntri] = col;
        ++ntri;
      }
      for (int q = 0; q < c.npair; ++q)
        for (int f = 0; f < 6; ++f)
          if (csw[8 * q + f] == col) tmap[2 * Dm + col] = 8 * q + f;
    }
    c.sw_tmap.alloc(tmap.size());
    FS_HIP(hipMemcpyAsync(c.sw_tmap.p, tmap.data(), 4 * tmap.size(), hipMemcpyHostToDevice, c.stream));
    s.tidx = c.sw_tmap.p;
    s.tcol = c.sw_tmap.p + Dm;
    s.tsf = c.sw_tmap.p + 2 * Dm;
    s.P = c.P;
    c.sw_own.alloc(size_t(ntri) * c.N * 3 + 1);
    c.sw_ptan.alloc(size_t(ntri) * std::max<int64_t>(c.P, 1) * 3);
    c.sw_etan.alloc(size_t(ntri) * c.N * 10 + 1);
    s.own = c.sw_own.p;
    s.ptan = c.sw_ptan.p;
    s.etan = c.sw_etan.p;
    s.ntiles = nt;
    s.tile_cfg = c.jt_cfg.p;
    s.tile_a0 = c.jt_a0.p;
    s.tile_na = c.jt_na.p;
    s.atom_off = c.atom_off.p;
    s.row_off = c.row_off.p;
    s.J = c.J.p;
    s.part = c.jac_part.p;
    if (c.N > 0 && c.D > 0) {
      c.time_main_begin();
      if (ntri > 0) {
        hipLaunchKernelGGL(k_sw_jac_centre, dim3(nb(c.N * kSwJLanes, 256), ntri), dim3(256), 0,
                           c.stream, s);
        FS_HIP(hipGetLastError());
      }
      c.time_main_end();
      hipLaunchKernelGGL(k_sw_jac_tiles, dim3(std::max(nt, 1), c.D), dim3(64), 0, c.stream, s);
      FS_HIP(hipGetLastError());
    }
  } else {
    // column descriptors and lane-slot maps: built and uploaded once per model
    const bool adp = c.kind == FS_ADP;
    if (!c.jl.valid || !c.jsort_valid) c.ks_valid = false; // model or neighbour list changed
    int Dd = 0, dens_c0 = 0, gi_nf = 0, NSP = 0, NSD = 0, NSU = 0, NSW = 0, Ddip = 0, Dquad = 0;
    if (!c.jl.valid) {
      // column descriptors
      std::vector<ColDesc> cols(std::max(c.D, 1));
      std::vector<std::pair<int, int>> pair_types(std::max(c.npair, 1));
      for (int ta = 0; ta < c.ntypes; ++ta)
        for (int tb = ta; tb < c.ntypes; ++tb) pair_types[ord_pair(ta, tb, c.ntypes)] = {ta, tb};
      dens_c0 = -1;
      Dd = 0;
      for (int t = 0; t < c.ntab; ++t) {
        const TableDesc &T = c.h_tabs[t];
        int kind = CK_PAIR;
        if (t >= c.npair && t < c.npair + c.ntypes) kind = CK_DENS;
        else if (t >= c.npair + c.ntypes && t < c.npair + 2 * c.ntypes) kind = CK_EMB;
        else if (t >= c.npair + 2 * c.ntypes && t < 2 * c.npair + 2 * c.ntypes) kind = CK_DIP;
        else if (t >= 2 * c.npair + 2 * c.ntypes) kind = CK_QUAD;
        for (int m = 0; m < T.n; ++m) {
          const int col = c.h_col_of[T.off + m];
          if (col < 0) continue;
          ColDesc cd{t, m, kind, -1};
          const int po = kind == CK_DIP ? t - c.npair - 2 * c.ntypes
                                        : (kind == CK_QUAD ? t - 2 * c.npair - 2 * c.ntypes : t);
          if (kind == CK_PAIR || kind == CK_DIP || kind == CK_QUAD)
            cd.dc = pair_types[po].first | (pair_types[po].second << 8);
          if (kind == CK_DIP) ++Ddip;
          if (kind == CK_QUAD) ++Dquad;
          if (kind == CK_DENS) {
            if (dens_c0 < 0) dens_c0 = col;
            cd.dc = Dd++;
          }
          cols[col] = cd;
        }
      }
      if (dens_c0 < 0) dens_c0 = 0;
      static_assert(sizeof(ColDesc) == sizeof(int4), "column descriptor layout");
      c.jcols.alloc(cols.size());
      FS_HIP(hipMemcpyAsync(c.jcols.p, cols.data(), sizeof(ColDesc) * cols.size(), hipMemcpyHostToDevice,
                            c.stream));
      FS_HIP(hipStreamSynchronize(c.stream));
      // the balanced g_ti windows need every density table to contribute the same
      // number of free columns, laid out table by table in one lane slot
      {
        std::vector<int> nfree(c.ntypes, 0);
        for (int col = 0; col < c.D; ++col)
          if (cols[col].kind == CK_DENS) ++nfree[cols[col].tab - c.npair];
        bool equal = c.ntypes > 0;
        for (int t = 1; t < c.ntypes; ++t) equal = equal && nfree[t] == nfree[0];
        gi_nf = (equal && nfree[0] > 0 && nfree[0] * c.ntypes <= 64) ? nfree[0] : 0;
      }
      // lane slots: pair, density and embedding columns each start a new slot.
      // Maps in order: pair-table columns (k_jac_pair), density and embedding
      // columns (k_jac_de), then for ADP the dipole and the quadrupole columns
      // (k_adp_mom / k_adp_dq, one map each); qc numbers a kind's columns.
      auto build_slots = [&](std::initializer_list<int> kinds) {
        std::vector<SlotDesc> sl;
        for (int kind : kinds) {
          int n = 0;
          for (int col = 0; col < c.D; ++col) {
            if (cols[col].kind != kind) continue;
            if (n % 64 == 0) sl.resize(sl.size() + 64, SlotDesc{-1, 0, CK_NONE, -1, 0, 0, 0, 0});
            const ColDesc &cd = cols[col];
            sl[sl.size() - 64 + n % 64] = SlotDesc{cd.tab, cd.m, cd.kind, cd.dc, col, n, 0, 0};
            ++n;
          }
        }
        return sl;
      };
      auto pad_to = [](std::vector<SlotDesc> &sl, int ns) {
        sl.resize(size_t(ns) * 64, SlotDesc{-1, 0, CK_NONE, -1, 0, 0, 0, 0});
      };
      // slots per launch: 2 or 4 pair slots, 2, 4 or 6 density/embedding
      // slots, 1, 2 or 4 dipole/quadrupole slots; wider column sets run as
      // several launches of the widest form over consecutive slot groups
      std::vector<SlotDesc> slots = build_slots({CK_PAIR}), sde = build_slots({CK_DENS, CK_EMB});
      const int nsp = int(slots.size() / 64), nsd = int(sde.size() / 64);
      NSP = nsp == 0 ? 0 : (nsp <= 2 ? 2 : 4);
      NSD = nsd == 0 ? 0 : (nsd <= 2 ? 2 : (nsd <= 4 ? 4 : 6));
      const int TSP = round_slots(nsp, NSP), TSD = round_slots(nsd, NSD);
      pad_to(slots, TSP);
      pad_to(sde, TSD);
      slots.insert(slots.end(), sde.begin(), sde.end());
      int TSU = 0, TSW = 0;
      if (adp) {
        std::vector<SlotDesc> su = build_slots({CK_DIP}), sw = build_slots({CK_QUAD});
        const int nsu = int(su.size() / 64), nsw = int(sw.size() / 64);
        NSU = nsu <= 1 ? nsu : (nsu <= 2 ? 2 : 4);
        NSW = nsw <= 1 ? nsw : (nsw <= 2 ? 2 : 4);
        TSU = round_slots(nsu, NSU);
        TSW = round_slots(nsw, NSW);
        pad_to(su, TSU);
        pad_to(sw, TSW);
        slots.insert(slots.end(), su.begin(), su.end());
        slots.insert(slots.end(), sw.begin(), sw.end());
      }
      static_assert(sizeof(SlotDesc) == 2 * sizeof(int4), "slot descriptor layout");
      c.jslots.alloc(std::max<size_t>(slots.size(), 1) * 2);
      FS_HIP(hipMemcpyAsync(c.jslots.p, slots.data(), sizeof(SlotDesc) * slots.size(),
                            hipMemcpyHostToDevice, c.stream));
      FS_HIP(hipStreamSynchronize(c.stream));

      c.jl = Context::JacLayout{true, Dd, dens_c0, gi_nf, NSP, NSD, NSU, NSW, Ddip, Dquad,
                                TSP, TSD, TSU, TSW};
    }
    Dd = c.jl.Dd;
    dens_c0 = c.jl.dens_c0;
    gi_nf = c.jl.gi_nf;
    NSP = c.jl.NSP;
    NSD = c.jl.NSD;
    NSU = c.jl.NSU;
    NSW = c.jl.NSW;
    Ddip = c.jl.Ddip;
    Dquad = c.jl.Dquad;
    const int TSP = c.jl.TSP, TSD = c.jl.TSD, TSU = c.jl.TSU, TSW = c.jl.TSW;
    a.Dd = Dd;
    a.gi_nf = gi_nf;
    FS_REQUIRE(int64_t(c.N) * std::max(Dd, 1) < (int64_t(1) << 30), FS_ERR_UNSUPPORTED,
               "density-column table exceeds 2^30 entries");
    a.dens_c0 = dens_c0;
    a.cols = reinterpret_cast<const ColDesc *>(c.jcols.p);
    a.tabs = c.tabs.p;
    a.tx = c.tab_x.p;
    a.ty = c.tab_y.p;
    a.ts = c.tab_s.p;
    a.tS = c.tab_S.p;
    a.ivl32 = c.ivl32.p;
    a.rho = c.rho.p;
    a.over = c.clamp_over.p;
    a.Fp = c.Fp.p;
    a.Fp32 = c.Fp32.p;
    a.gauge = c.gauge.p;
    {
      // window tables while ntypes x the pair tables' knots fit kWin entries
      int np_max = 0;
      for (int t = 0; t < c.npair; ++t) np_max = std::max(np_max, c.h_tabs[t].n);
      a.win_pair = c.ntypes * np_max <= kWin ? np_max : 0;
    }
    c.H.alloc(size_t(std::max<int64_t>(c.N * std::max(Dd, 1), 1)));
    c.Hs.alloc(size_t(std::max<int64_t>(c.N * std::max(Dd, 1), 1)));
    c.emb_d.alloc(4 * size_t(std::max<int64_t>(c.N, 1)));
    c.emb_v.alloc(4 * size_t(std::max<int64_t>(c.N, 1)));
    c.emb_k.alloc(size_t(std::max<int64_t>(c.N, 1)));
    a.H = c.H.p;
    a.Hs = c.Hs.p;
    a.emb_d = c.emb_d.p;
    a.emb_v = c.emb_v.p;
    a.emb_k = c.emb_k.p;
    a.Fpp = c.Fpp.p;
    // the slope bands are staged in LDS while every pass stays within the
    // staging budget; larger tables are read from their FP32 copy in place
    const size_t desc = size_t(c.ntab) * sizeof(TableDesc);
    const size_t band = ((size_t(c.nknots) * 5 + 3) & ~size_t(3)) * 4;
    const size_t widest = std::max({sizeof(DensStage), sizeof(JStagePair), sizeof(JStageDE),
                                    sizeof(MomStage), sizeof(DQStage)});
    a.stage_S = desc + band + 4 * widest <= kTableStageBytes ? 1 : 0;
    // the density tables' FP32 interval forms too, while everything still
    // fits and the density/embedding pass keeps four blocks per CU (40 KiB)
    a.ivl_lo = c.h_tabs[c.npair].ivl;
    a.nivl_st = (c.npair + c.ntypes < c.ntab ? c.h_tabs[c.npair + c.ntypes].ivl : c.nivl) - a.ivl_lo;
    const size_t ivl = size_t(a.nivl_st) * 2 * sizeof(float4);
    a.stage_ivl = a.stage_S && desc + band + ivl + 4 * widest <= kTableStageBytes &&
                          desc + band + ivl + 4 * sizeof(JStageDE) <= 40 * 1024
                      ? 1
                      : 0;
    a.tS32 = nullptr;
    if (!a.stage_S) {
      c.tab_S32.alloc(size_t(c.nknots) * 5);
      hipLaunchKernelGGL(k_band_f32, dim3(nb(int64_t(c.nknots) * 5, 256)), dim3(256), 0, c.stream,
                         int64_t(c.nknots) * 5, c.tab_S.p, c.tab_S32.p);
      FS_HIP(hipGetLastError());
      a.tS32 = c.tab_S32.p;
    }
    const size_t base = desc + (a.stage_S ? band : 0) + (a.stage_ivl ? ivl : 0);
    // table staging as a template argument of the passes (see stage); the
    // passes without interval lookups stage at most the slope bands
    const int ts = a.stage_S ? (a.stage_ivl ? 2 : 1) : 0;
    auto with_ts = [](int v, auto &&f) {
      if (v == 2) f(std::integral_constant<int, 2>{});
      else if (v == 1) f(std::integral_constant<int, 1>{});
      else f(std::integral_constant<int, 0>{});
    };
    const size_t lds_dens = base + 4 * sizeof(DensStage);
    const size_t lds_pair = base + 4 * sizeof(JStagePair), lds_de = base + 4 * sizeof(JStageDE);
    FS_REQUIRE(lds_de <= 64 * 1024 && lds_pair <= 64 * 1024, FS_ERR_UNSUPPORTED,
               "Jacobian LDS budget exceeded");
    const SlotDesc *d_slots = reinterpret_cast<const SlotDesc *>(c.jslots.p);
    if (!c.jsort_valid && c.N > 0) {
      c.jrec.alloc(size_t(std::max<int64_t>(c.P, 1)));
      c.jr.alloc(size_t(std::max<int64_t>(c.P, 1)));
      hipLaunchKernelGGL(k_jac_sort, dim3(nb(c.N, 128)), dim3(128), 0, c.stream, c.N, c.pair_off.p,
                         c.pair_rec.p, c.pair_r64.p, c.jrec.p, c.jr.p);
      FS_HIP(hipGetLastError());
      c.jsort_valid = true;
    }
    a.jrec = c.jrec.p;
    a.jr = c.jr.p;
    // knot sums: the pair-table columns and H_i fold per-atom sums of the
    // geometry, built once per neighbour list and model grid, while they fit
    // (at most 128 pair-table and 128 density knots per atom, the build's LDS
    // accumulators, 4 GiB of sums); else the window passes
    KsArgs ks{};
    bool use_ks = false;
    {
      int lp = 0, ld = 0, kp = 0;
      for (int ti = 0; ti < c.ntypes; ++ti) {
        int o = 0;
        for (int t = 0; t < c.ntypes; ++t) o += c.h_tabs[ord_pair(ti, t, c.ntypes)].n;
        lp = std::max(lp, o);
      }
      for (int t = 0; t < c.ntypes; ++t) ld += c.h_tabs[c.npair + t].n;
      for (int t = 0; t < c.npair; ++t) kp += c.h_tabs[t].n;
      // even entry counts: an atom's block is a whole number of float4
      lp = (lp + 1) & ~1;
      ld = (ld + 1) & ~1;
      const size_t lds_ks = desc + 4 * ks_wave_bytes(kp);
      const size_t bytes = size_t(c.N) * (6 * size_t(lp) + 2 * size_t(ld)) * sizeof(float);
      const bool fits = lp <= 128 && ld <= 128 && lds_ks <= 64 * 1024 && bytes <= (size_t(4) << 30) &&
                        (TSP == 0 || NSP == 2 || NSP == 4);
      FS_REQUIRE(c.jac_form != 2 || fits, FS_ERR_UNSUPPORTED,
                 "the knot-sum form of the Jacobian does not fit this model");
      use_ks = fits && c.jac_form != 1 && c.N > 0;
      if (use_ks) {
        const bool build = !c.ks_valid || c.ks_lp != lp || c.ks_ld != ld || c.ks_kp != kp;
        if (build) {
          c.ks_pair.alloc(size_t(c.N) * 6 * lp);
          c.ks_dens.alloc(size_t(c.N) * 2 * ld);
          c.ks_tile.alloc(size_t(std::max(nt, 1)) * 14 * kp);
          c.ks_cfg.alloc(size_t(c.C) * 14 * kp);
        }
        ks.pair = c.ks_pair.p;
        ks.dens = c.ks_dens.p;
        ks.tile = c.ks_tile.p;
        ks.cfg = c.ks_cfg.p;
        ks.lp = lp;
        ks.ld = ld;
        ks.kp = kp;
        if (build) {
          const unsigned gk = unsigned(std::max(1, std::min(nt, c.cu_count * 4)));
          hipLaunchKernelGGL(k_knot_sums, dim3(gk), dim3(256), lds_ks, c.stream, a, ks);
          FS_HIP(hipGetLastError());
          hipLaunchKernelGGL(k_knot_sums_cfg, dim3(c.C), dim3(256), 0, c.stream, a, ks);
          FS_HIP(hipGetLastError());
          c.ks_valid = true;
          c.ks_lp = lp;
          c.ks_ld = ld;
          c.ks_kp = kp;
        }
      }
    }
    if (c.N > 0) {
      // (the pair pass depends on neither kernel below; run beside them on a
      // second stream it measured no faster and stretched both)
      const unsigned g2 = unsigned(std::max(1, std::min(nt, c.cu_count * c.jac_blocks_per_cu)));
      const SlotDesc *d_de = d_slots + size_t(TSP) * 64;
      c.time_main_begin(1);
      // wider maps (more than 128 columns of a pass) keep the compiler's
      // occupancy rather than spill their larger slot state
      for (int g = 0; g < TSP; g += NSP) {
        if (use_ks) {
          const size_t lds = desc + 4 * 6 * sizeof(float) * size_t(ks.lp);
          if (NSP == 2)
            hipLaunchKernelGGL((k_jac_pair_ks<2>), dim3(g2), dim3(256), lds, c.stream, a, ks,
                               d_slots + size_t(g) * 64);
          else
            hipLaunchKernelGGL((k_jac_pair_ks<4>), dim3(g2), dim3(256), lds, c.stream, a, ks,
                               d_slots + size_t(g) * 64);
        } else {
          with_ts(ts, [&](auto tsc) {
            constexpr int TS = decltype(tsc)::value;
            if (NSP == 2)
              hipLaunchKernelGGL((k_jac_pair<2, kWavesPair, TS>), dim3(g2), dim3(256), lds_pair,
                                 c.stream, a, c.nknots, d_slots + size_t(g) * 64);
            else
              hipLaunchKernelGGL((k_jac_pair<4, 2, TS>), dim3(g2), dim3(256), lds_pair, c.stream, a,
                                 c.nknots, d_slots + size_t(g) * 64);
          });
        }
        FS_HIP(hipGetLastError());
      }
      if (use_ks && TSP > 0) {
        hipLaunchKernelGGL(k_jac_pair_ks_rows, dim3(c.C), dim3(256), 0, c.stream, a, ks, d_slots, TSP);
        FS_HIP(hipGetLastError());
      }
      c.time_main_end(1);
      hipLaunchKernelGGL(k_emb_sites, dim3(nb(c.N, 256)), dim3(256), 0, c.stream, a);
      FS_HIP(hipGetLastError());
      const unsigned g1 = unsigned(std::min<int64_t>((c.N + 3) / 4, int64_t(c.cu_count) * 16));
      for (int dc0 = 0; dc0 < Dd; dc0 += 64 * kCPL) {
        if (use_ks)
          hipLaunchKernelGGL(k_jac_dens_ks, dim3(g1), dim3(256), 4 * 2 * sizeof(float) * size_t(ks.ld),
                             c.stream, a, ks, dc0);
        else if (ts > 0)
          hipLaunchKernelGGL(k_jac_dens<1>, dim3(g1), dim3(256), lds_dens, c.stream, a, c.nknots, dc0);
        else
          hipLaunchKernelGGL(k_jac_dens<0>, dim3(g1), dim3(256), lds_dens, c.stream, a, c.nknots, dc0);
        FS_HIP(hipGetLastError());
      }
      c.time_main_begin(0);
      for (int g = 0; g < TSD; g += NSD) {
        const SlotDesc *sl = d_de + size_t(g) * 64;
        with_ts(ts, [&](auto tsc) {
          constexpr int TS = decltype(tsc)::value;
          if (NSD == 2)
            hipLaunchKernelGGL((k_jac_de<2, kWavesDE, TS>), dim3(g2), dim3(256), lds_de, c.stream, a,
                               c.nknots, sl);
          else if (NSD == 4)
            hipLaunchKernelGGL((k_jac_de<4, 2, TS>), dim3(g2), dim3(256), lds_de, c.stream, a,
                               c.nknots, sl);
          else
            hipLaunchKernelGGL((k_jac_de<6, 1, TS>), dim3(g2), dim3(256), lds_de, c.stream, a,
                               c.nknots, sl);
        });
        FS_HIP(hipGetLastError());
      }
      if (adp) {
        // dipole and quadrupole columns: moment tangents, then the force pass
        const size_t nmu = size_t(c.N) * 4 * Ddip, nla = size_t(c.N) * 6 * Dquad;
        FS_REQUIRE(4 * nmu < (size_t(1) << 32) && 4 * nla < (size_t(1) << 32), FS_ERR_UNSUPPORTED,
                   "ADP moment-tangent table exceeds 4 GiB");
        c.adp_tan.alloc(std::max<size_t>(nmu + nla, 1));
        a.dmu = c.adp_tan.p;
        a.dlam = c.adp_tan.p + nmu;
        a.Ddip = Ddip;
        a.Dquad = Dquad;
        a.mu = c.mu.p;
        a.lamb = c.lamb.p;
        const size_t lds_mom = base + 4 * sizeof(MomStage), lds_dq = base + 4 * sizeof(DQStage);
        FS_REQUIRE(lds_dq <= 64 * 1024, FS_ERR_UNSUPPORTED, "ADP Jacobian LDS budget exceeded");
        const SlotDesc *d_u = d_de + size_t(TSD) * 64, *d_w = d_u + size_t(TSU) * 64;
        // every moment-tangent group before any force pass reads the tangents
        auto pass = [&](auto kind, auto tsc, int ns, int total, const SlotDesc *slots0) {
          constexpr int K = decltype(kind)::value, TS = decltype(tsc)::value;
          if (ns == 0) return;
          for (int g = 0; g < total; g += ns) {
            const SlotDesc *sl = slots0 + size_t(g) * 64;
            if (ns == 1) hipLaunchKernelGGL((k_adp_mom<K, 1, TS>), dim3(g1), dim3(256), lds_mom, c.stream, a, c.nknots, sl);
            else if (ns == 2) hipLaunchKernelGGL((k_adp_mom<K, 2, TS>), dim3(g1), dim3(256), lds_mom, c.stream, a, c.nknots, sl);
            else hipLaunchKernelGGL((k_adp_mom<K, 4, TS>), dim3(g1), dim3(256), lds_mom, c.stream, a, c.nknots, sl);
            FS_HIP(hipGetLastError());
          }
          if (K == JP_QUAD) c.time_main_begin(2);
          for (int g = 0; g < total; g += ns) {
            const SlotDesc *sl = slots0 + size_t(g) * 64;
            if (ns == 1) hipLaunchKernelGGL((k_adp_dq<K, 1, TS>), dim3(g2), dim3(256), lds_dq, c.stream, a, c.nknots, sl);
            else if (ns == 2) hipLaunchKernelGGL((k_adp_dq<K, 2, TS>), dim3(g2), dim3(256), lds_dq, c.stream, a, c.nknots, sl);
            else hipLaunchKernelGGL((k_adp_dq<K, 4, TS>), dim3(g2), dim3(256), lds_dq, c.stream, a, c.nknots, sl);
            FS_HIP(hipGetLastError());
          }
          if (K == JP_QUAD) c.time_main_end(2);
        };
        with_ts(std::min(ts, 1), [&](auto tsc) {
          pass(std::integral_constant<int, JP_DIP>{}, tsc, NSU, TSU, d_u);
          pass(std::integral_constant<int, JP_QUAD>{}, tsc, NSW, TSW, d_w);
        });
      }
      c.time_main_end();
    }
  }
  hipLaunchKernelGGL(k_jac_rows, dim3(c.C, kJP), dim3(256), 0, c.stream, a);
  FS_HIP(hipGetLastError());
  launch_smooth_jacobian(c);
}

} // namespace fsa
