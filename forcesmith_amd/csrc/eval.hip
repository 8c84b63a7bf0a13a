// Residual path: makima slope refresh, EAM/ADP two-pass kernels
// (density -> clamp + embedding -> force/stress), per-configuration FP64
// reduction and residual assembly.
//
// Reference: EAMForceCalculator::eval_forces (src/force/eam_force.cpp:45-102),
// ADPForceCalculator::eval_forces (src/force/adp_force.cpp:73-135),
// clamp_embedding / accumulate_pair (include/forcesmith/force/kernels.hpp:
// 19-38), with_eval_scope (eval_scope.hpp:14-31), write_config_residuals
// (include/forcesmith/optimization/residual_layout.hpp:37-54).
//
// Layout: atoms are processed in tiles of consecutive atoms of one
// configuration; a block handles one tile per iteration of a persistent loop,
// LPA lanes per atom stride over the atom's CSR neighbour segment (coalesced
// 16-byte pair records), the atom's sums are reduced across its lanes with
// cross-lane shuffles, and the tile's energy/stress/limit are reduced in LDS
// in atom order and written as one FP64 partial per tile. Forces are
// gathered per atom: no atomics anywhere, every result is deterministic.
#include "common.hpp"
#include "spline.hpp"

#include <algorithm>
#include <type_traits>

namespace fsa {

namespace {

constexpr double kPenalty = 100.0 * 10.0; // kDummyWeight * 10 (kernels.hpp:24)

// ---------------------------------------------------------------------------
// makima slopes + band of d s / d y per table (spline.cpp:43-103)
// ---------------------------------------------------------------------------
__device__ inline double dmb(const double *x, int q, int m) {
  const double h = x[q + 1] - x[q];
  return m == q ? -1.0 / h : (m == q + 1 ? 1.0 / h : 0.0);
}
__device__ inline double dM(const double *x, int n, int q, int m) {
  if (q >= 0 && q <= n - 2) return dmb(x, q, m);
  if (q == -1) return 2.0 * dmb(x, 0, m) - dmb(x, 1, m);
  if (q == -2) return 2.0 * (2.0 * dmb(x, 0, m) - dmb(x, 1, m)) - dmb(x, 0, m);
  if (q == n - 1) return 2.0 * dmb(x, n - 2, m) - dmb(x, n - 3, m);
  return 2.0 * (2.0 * dmb(x, n - 2, m) - dmb(x, n - 3, m)) - dmb(x, n - 2, m);
}
__device__ inline double Mext(const double *x, const double *y, int n, int q) {
  auto sec = [&](int k) { return (y[k + 1] - y[k]) / (x[k + 1] - x[k]); };
  if (q >= 0 && q <= n - 2) return sec(q);
  if (q == -1) return 2 * sec(0) - sec(1);
  if (q == -2) {
    const double m1 = 2 * sec(0) - sec(1);
    return 2 * m1 - sec(0);
  }
  if (q == n - 1) return 2 * sec(n - 2) - sec(n - 3);
  const double m1 = 2 * sec(n - 2) - sec(n - 3);
  return 2 * m1 - sec(n - 2);
}
__device__ inline double sgnd(double u) { return double((u > 0) - (u < 0)); }

// makima slope of knot k and its row of the slope band dS_k/dy_(k-2..k+2),
// through the extended secants (quadratic extrapolation at both ends), so
// every knot follows one formula and the knots of a table run in parallel
__device__ void knot_slope(int n, int k, const double *x, const double *y, double &sk, double *Sk) {
  const double A = Mext(x, y, n, k - 1), B = Mext(x, y, n, k), Cn = Mext(x, y, n, k + 1),
               P = Mext(x, y, n, k - 2);
  const double W1 = fabs(Cn - B) + fabs(Cn + B) / 2;
  const double W2 = fabs(A - P) + fabs(A + P) / 2;
  const double W = W1 + W2;
  sk = (W1 * A + W2 * B) / W;
  if (isnan(sk)) sk = 0;
  for (int q = 0; q < 5; ++q) Sk[q] = 0.0;
  if (!(W > 0)) return;
  const double sv = sk;
  const double dw1C = sgnd(Cn - B) + sgnd(Cn + B) / 2, dw1B = -sgnd(Cn - B) + sgnd(Cn + B) / 2;
  const double dw2A = sgnd(A - P) + sgnd(A + P) / 2, dw2P = -sgnd(A - P) + sgnd(A + P) / 2;
  const double ds1 = (A - sv) / W, ds2 = (B - sv) / W;
  const double gP = ds2 * dw2P, gA = W1 / W + ds2 * dw2A, gB = W2 / W + ds1 * dw1B, gC = ds1 * dw1C;
  for (int q = 0; q < 5; ++q) {
    const int m = k - 2 + q;
    if (m < 0 || m >= n) continue;
    Sk[q] = gP * dM(x, n, k - 2, m) + gA * dM(x, n, k - 1, m) + gB * dM(x, n, k, m) +
            gC * dM(x, n, k + 1, m);
  }
}

// power form a + b t + c t^2 + d t^3 of interval k in t = (r - x_k) / h,
// formed in FP64; the FP32 copy is a hi/lo pair per coefficient
__device__ void interval_form(const TableDesc &T, int k, const double *x, const double *y,
                              const double *s, float4 *ivl, double *p64) {
  const double y0 = y[k], y1 = y[k + 1];
  double q[4];
  if (T.linear) {
    q[0] = y0;
    q[1] = y1 - y0;
    q[2] = q[3] = 0.0;
  } else {
    const double dx = x[k + 1] - x[k];
    const double b = dx * s[k], e = dx * s[k + 1];
    q[0] = y0;
    q[1] = b;
    q[2] = 3.0 * (y1 - y0) - 2.0 * b - e;
    q[3] = 2.0 * (y0 - y1) + b + e;
  }
  float hi[4], lo[4];
  for (int m = 0; m < 4; ++m) {
    hi[m] = float(q[m]);
    lo[m] = float(q[m] - double(hi[m]));
    p64[4 * (T.ivl + k) + m] = q[m];
  }
  ivl[2 * (T.ivl + k)] = make_float4(hi[0], hi[1], hi[2], hi[3]);
  ivl[2 * (T.ivl + k) + 1] = make_float4(lo[0], lo[1], lo[2], lo[3]);
}

// every table of one parameter set, by the threads of one block: slopes and
// band per knot, then the interval forms (n < 4: piecewise linear, no slopes)
__device__ __forceinline__ void tables_block(int ntab, const TableDesc *tabs, int nknots,
                                             const double *x, const double *y, double *s,
                                             double *S, float4 *ivl, double *p64) {
  for (int e = threadIdx.x; e < nknots; e += blockDim.x) {
    int t = 0;
    while (t + 1 < ntab && tabs[t + 1].off <= e) ++t;
    const TableDesc T = tabs[t];
    const int k = e - T.off;
    if (T.n >= 4) {
      knot_slope(T.n, k, x + T.off, y + T.off, s[e], S + 5 * size_t(e));
    } else {
      s[e] = 0.0;
      for (int q = 0; q < 5; ++q) S[5 * size_t(e) + q] = 0.0;
    }
  }
  __syncthreads();
  for (int e = threadIdx.x; e < nknots; e += blockDim.x) {
    int t = 0;
    while (t + 1 < ntab && tabs[t + 1].off <= e) ++t;
    const TableDesc T = tabs[t];
    const int k = e - T.off;
    if (k + 1 < T.n) interval_form(T, k, x + T.off, y + T.off, s + T.off, ivl, p64);
  }
}

// tables_block over LDS copies: the descriptors always (every knot searches
// them for its table) and, with ST, the knots' abscissae and values (each
// slope reads ten of them). One block does all of it, so without the copies
// the knots' dependent global reads set the kernel's time. ST is a template
// argument so that the reads compile to LDS loads (a pointer that may address
// either memory compiles to flat loads); the launch passes
// tables_lds_bytes(c, ST) of dynamic LDS.
template <bool ST>
__device__ __forceinline__ void tables_staged(int ntab, const TableDesc *tabs, int nknots,
                                              const double *x, const double *y, double *s,
                                              double *S, float4 *ivl, double *p64) {
  extern __shared__ float4 tlds[];
  const int nq = ntab * int(sizeof(TableDesc) / sizeof(float4));
  for (int k = threadIdx.x; k < nq; k += blockDim.x) tlds[k] = reinterpret_cast<const float4 *>(tabs)[k];
  const TableDesc *sT = reinterpret_cast<const TableDesc *>(tlds);
  if constexpr (ST) {
    double *sx = reinterpret_cast<double *>(tlds + nq);
    for (int k = threadIdx.x; k < nknots; k += blockDim.x) {
      sx[k] = x[k];
      sx[nknots + k] = y[k];
    }
    __syncthreads();
    tables_block(ntab, sT, nknots, sx, sx + nknots, s, S, ivl, p64);
  } else {
    __syncthreads();
    tables_block(ntab, sT, nknots, x, y, s, S, ivl, p64);
  }
}

template <bool ST>
__global__ void k_slopes(int ntab, int nknots, const TableDesc *__restrict__ tabs,
                         const double *__restrict__ x, const double *__restrict__ y,
                         double *__restrict__ s, double *__restrict__ S, float4 *__restrict__ ivl,
                         double *__restrict__ p64) {
  tables_staged<ST>(ntab, tabs, nknots, x, y, s, S, ivl, p64);
}

// the free knots from a device parameter vector, then every table's slopes
// and interval forms (the parameter scatter of the captured step folded into
// the slopes launch; one block, so the knot writes are visible after the
// barrier)
template <bool ST>
__global__ void k_scatter_slopes(int D, const double *__restrict__ xv,
                                 const int32_t *__restrict__ knot_of, int ntab, int nknots,
                                 const TableDesc *__restrict__ tabs, const double *__restrict__ x,
                                 double *__restrict__ y, double *__restrict__ s,
                                 double *__restrict__ S, float4 *__restrict__ ivl,
                                 double *__restrict__ p64) {
  for (int c = threadIdx.x; c < D; c += blockDim.x) y[knot_of[c]] = xv[c];
  __syncthreads();
  tables_staged<ST>(ntab, tabs, nknots, x, y, s, S, ivl, p64);
}

__global__ void k_scatter(int D, const double *__restrict__ xv, const int32_t *__restrict__ knot_of,
                          double *__restrict__ y) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < D) y[knot_of[c]] = xv[c];
}

// ---------------------------------------------------------------------------
// two-pass EAM / ADP
// ---------------------------------------------------------------------------
struct EvalArgs {
  int64_t N;
  int ntiles;
  const int32_t *tile_cfg, *tile_a0, *tile_na;
  const int64_t *atom_off, *pair_off, *row_off;
  const int32_t *types;
  const float4 *rec;
  const uint32_t *bits;
  const double *pd;
  const double *r64;
  const TableDesc *tabs;
  int ntab, nivl;
  int stage_ivl; // interval forms staged in LDS (else read from global memory)
  const float4 *ivl32;
  const double *ivl64;
  const double *tx, *ty, *ts;
  int ntypes, npair;
  const double *ref_force;
  double *rho, *over, *Fp;
  float *Fp32;
  double *e_atom, *lim, *mu, *lamb;
  float4 *mu32, *lamb32; // FP32 copies of mu, lamb (mixed ADP force pass)
  double *forces, *residual, *tile_part;
  const double *gauge; // EAM gauge {A, S[kMaxTypes], F0[kMaxTypes]} (identity: 1, 0, 0)
};

template <int LPA, class T> __device__ inline T sub_sum(T v) {
#pragma unroll
  for (int m = LPA / 2; m > 0; m >>= 1) v += __shfl_xor(v, m, LPA);
  return v;
}

// descriptors into LDS, and with ST the FP32 hi/lo and FP64 interval forms
// after them (while they fit the staging budget, a.stage_ivl); without ST the
// lookups read the interval forms from global memory (a few KB per table,
// L2-resident). ST is a template argument: a pointer that may address either
// LDS or global memory compiles to flat loads, measurably slower than LDS
// reads on every lookup.
template <bool ST>
__device__ inline void stage_tables(const EvalArgs &a, TableDesc *sT, float4 *smem_ivl,
                                    const float4 *&sI, const double *&sP) {
  const float4 *src = reinterpret_cast<const float4 *>(a.tabs);
  float4 *dst = reinterpret_cast<float4 *>(sT);
  const int nq = a.ntab * int(sizeof(TableDesc) / sizeof(float4));
  for (int k = threadIdx.x; k < nq; k += blockDim.x) dst[k] = src[k];
  if constexpr (ST) {
    for (int k = threadIdx.x; k < 2 * a.nivl; k += blockDim.x) smem_ivl[k] = a.ivl32[k];
    double *p64 = reinterpret_cast<double *>(smem_ivl + 2 * a.nivl);
    for (int k = threadIdx.x; k < 4 * a.nivl; k += blockDim.x) p64[k] = a.ivl64[k];
    sI = smem_ivl;
    sP = p64;
  } else {
    sI = a.ivl32;
    sP = a.ivl64;
  }
  __syncthreads();
}

// The EAM gauge (rescale_eam, src/core/rescale.cpp:82-158) on the embedding
// of one atom. sigma is the density sum of the unscaled tables; the gauge's
// density is rho = A sigma (ScaledOutputPotential), clamped against the
// embedding's span A [x0, xn] (ScaledArgPotential::span, clamp_embedding
// kernels.hpp:19-31, the reference's branch order for either sign of A). The
// embedding is evaluated at u = rho / A: sigma itself when no clamp applies,
// (A end) / A at a clamped atom, the reference's argument. The returned
// energy is F(u) - F0 + S (rho - clamped rho): the gauge's embedding
// F(u) - S rho_c - F0 plus the slope compensation of the pairs,
// sum_j S A g_j(r) = S rho, exact when the pair and density tables share one
// span (the host accepts no other gauge). Forces need no change: F'(u) / A - S
// times A g'(r) plus the compensation S A g'(r) is F'(u) g'(r). u is returned
// for the embedding lookup; lim and over (d lim / d sigma) for the limit row.
struct GaugedRho {
  double u, lim, over, e_shift;
};
__device__ inline GaugedRho gauge_clamp(const double *gauge, int ti, const TableDesc &Te,
                                        double sigma) {
  const double A = gauge[0], S = gauge[1 + ti], F0 = gauge[1 + kMaxTypes + ti];
  const double rho = A * sigma, begin = A * Te.x0, end = A * Te.xn;
  GaugedRho g{sigma, 0.0, 0.0, 0.0};
  double rc = rho;
  if (rho > end) {
    const double dd = rho - end;
    g.lim = kPenalty * dd * dd;
    g.over = 2.0 * kPenalty * dd * A;
    rc = end;
    g.u = rc / A;
  } else if (rho < begin) {
    const double dd = begin - rho;
    g.lim = kPenalty * dd * dd;
    g.over = -2.0 * kPenalty * dd * A;
    rc = begin;
    g.u = rc / A;
  }
  g.e_shift = S * (rho - rc) - F0;
  return g;
}

// table lookups -------------------------------------------------------------
struct Tabs {
  const TableDesc *T;
  const float4 *I;
  const double *x, *y, *s, *P;
  // FP64 cached-site evaluation (parity mode and exact edge pairs)
  __device__ void eval64(int t, double r, double &v, double &d) const {
    const TableDesc D = T[t];
    spl_site64(D, x + D.off, y + D.off, s + D.off, r, v, d);
  }
  // FP32 value and derivative, located from the FP64 distance
  __device__ void eval32(int t, double r, float &v, float &d) const {
    const TableDesc D = T[t];
    double u;
    const int k = spl_interval64(D, r, u);
    spl_eval32(D, I, k, float(u), v, d);
  }
  __device__ void deriv32(int t, double r, float &d) const {
    const TableDesc D = T[t];
    double u;
    const int k = spl_interval64(D, r, u);
    spl_deriv32(D, I, k, float(u), d);
  }
  // FP32 derivative with the FP64 value of the same interval
  __device__ void deriv32_value64(int t, double r, float &d, double &v) const {
    const TableDesc D = T[t];
    double u;
    const int k = spl_interval64(D, r, u);
    spl_deriv32(D, I, k, float(u), d);
    v = spl_value64(D, P, k, u);
  }
  __device__ double value64(int t, double r) const {
    const TableDesc D = T[t];
    double u;
    const int k = spl_interval64(D, r, u);
    return spl_value64(D, P, k, u);
  }
};

template <class R> struct Pair {
  uint32_t b;
  R d0, d1, d2, r;
};

template <bool F64> __device__ inline Pair<double> load_pair64(const EvalArgs &a, int64_t p) {
  Pair<double> q;
  q.b = a.bits[p];
  q.d0 = a.pd[3 * p];
  q.d1 = a.pd[3 * p + 1];
  q.d2 = a.pd[3 * p + 2];
  double r2 = q.d0 * q.d0 + q.d1 * q.d1;
  r2 = r2 + q.d2 * q.d2;
  q.r = sqrt(r2);
  return q;
}

// --- pass 1: density (+ ADP moments) ---------------------------------------
template <class R> struct DensAcc {
  R rho = 0, m0 = 0, m1 = 0, m2 = 0, l[6] = {0, 0, 0, 0, 0, 0};
};

template <int KIND, bool EXACT, class R>
__device__ inline void density_pair(const Tabs &tb, const EvalArgs &a, int ti, uint32_t b, R d0, R d1,
                                    R d2, R r, DensAcc<R> &acc) {
  if (KIND == FS_ADP && (b & kFlagTiny)) return;
  const int tj = pj_type(b);
  if (b & kFlagGj) {
    R v, d;
    if (EXACT) {
      double vv, dd;
      tb.eval64(a.npair + tj, double(r), vv, dd);
      v = R(vv);
    } else {
      v = R(tb.value64(a.npair + tj, double(r)));
    }
    acc.rho += v;
    (void)d;
  }
  if (KIND == FS_ADP) {
    const int o = ord_pair(ti, tj, a.ntypes);
    if (b & kFlagDip) {
      R u;
      if (EXACT) {
        double vv, dd;
        tb.eval64(a.npair + 2 * a.ntypes + o, double(r), vv, dd);
        u = R(vv);
      } else {
        u = R(tb.value64(a.npair + 2 * a.ntypes + o, double(r)));
      }
      acc.m0 += u * d0;
      acc.m1 += u * d1;
      acc.m2 += u * d2;
    }
    if (b & kFlagQuad) {
      R w;
      if (EXACT) {
        double vv, dd;
        tb.eval64(2 * a.npair + 2 * a.ntypes + o, double(r), vv, dd);
        w = R(vv);
      } else {
        w = R(tb.value64(2 * a.npair + 2 * a.ntypes + o, double(r)));
      }
      acc.l[0] += w * (d0 * d0);
      acc.l[1] += w * (d1 * d1);
      acc.l[2] += w * (d2 * d2);
      acc.l[3] += w * (d0 * d1);
      acc.l[4] += w * (d0 * d2);
      acc.l[5] += w * (d1 * d2);
    }
  }
}

// the mixed density passes ask for 8 (EAM, 54 VGPRs) and 5 (ADP, 92) waves per
// SIMD, one more than the compiler's own register choices allow (68 and 106,
// no scratch either way): residual 0.202 -> 0.200 ms (EAM, 500 configurations)
// and 0.208 -> 0.204 ms (binary ADP, 200 configurations)
template <int KIND, bool F64, int LPA, bool ST>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(F64 ? 1 : (KIND == FS_ADP ? 5 : 8)))) void
k_density(EvalArgs a) {
  // Densities (and ADP moments) feed the energy rows directly, whose 1e-6
  // cost bar is dominated by |E| >> |E - E_ref|: the per-pair values are
  // evaluated in FP64 in both precision modes, from the FP64 distance.
  extern __shared__ float4 smem[];
  TableDesc *sT = reinterpret_cast<TableDesc *>(smem);
  const float4 *sI;
  const double *sP;
  stage_tables<ST>(a, sT, smem + a.ntab * int(sizeof(TableDesc) / sizeof(float4)), sI, sP);
  const Tabs tb{sT, sI, a.tx, a.ty, a.ts, sP};
  const int sub = threadIdx.x / LPA, lane = threadIdx.x % LPA;
  for (int tile = blockIdx.x; tile < a.ntiles; tile += gridDim.x) {
    if (sub >= a.tile_na[tile]) continue;
    const int64_t i = int64_t(a.tile_a0[tile]) + sub;
    const int ti = a.types[i];
    const int64_t p0 = a.pair_off[i], p1 = a.pair_off[i + 1];
    DensAcc<double> acc;
    for (int64_t p = p0 + lane; p < p1; p += LPA) {
      if constexpr (KIND == FS_ADP) {
        // FP64 mode: the cached-site form; mixed: the FP64 interval power
        // form of the same interval (values only), as for the EAM densities
        const Pair<double> q = load_pair64<true>(a, p);
        density_pair<KIND, F64, double>(tb, a, ti, q.b, q.d0, q.d1, q.d2, q.r, acc);
      } else {
        const uint32_t b = a.bits[p];
        if (b & kFlagGj) {
          if constexpr (F64) {
            double v, d;
            tb.eval64(a.npair + pj_type(b), a.r64[p], v, d);
            acc.rho += v;
          } else {
            acc.rho += tb.value64(a.npair + pj_type(b), a.r64[p]);
          }
        }
      }
    }
    acc.rho = sub_sum<LPA>(acc.rho);
    if (KIND == FS_ADP) {
      acc.m0 = sub_sum<LPA>(acc.m0);
      acc.m1 = sub_sum<LPA>(acc.m1);
      acc.m2 = sub_sum<LPA>(acc.m2);
      for (int k = 0; k < 6; ++k) acc.l[k] = sub_sum<LPA>(acc.l[k]);
    }
    if (lane == 0) {
      // clamp_embedding + F(rho) in FP64 (kernels.hpp:19-31, eam_force.cpp:68-73)
      // under the gauge (the identity for ADP: the operations reduce exactly
      // to the plain clamp)
      const TableDesc Te = sT[a.npair + a.ntypes + ti];
      const GaugedRho gr = gauge_clamp(a.gauge, ti, Te, double(acc.rho));
      const double rh = gr.u, lim = gr.lim, over = gr.over;
      double F, dF;
      spl_uncached64(Te, a.tx + Te.off, a.ty + Te.off, a.ts + Te.off, rh, F, dF);
      double e = F + gr.e_shift;
      if (KIND == FS_ADP) {
        const double m0 = acc.m0, m1 = acc.m1, m2 = acc.m2;
        e += 0.5 * (m0 * m0 + m1 * m1 + m2 * m2);
        const double L0 = acc.l[0], L1 = acc.l[1], L2 = acc.l[2], L3 = acc.l[3], L4 = acc.l[4],
                     L5 = acc.l[5];
        const double tr = L0 + L1 + L2;
        const double fro = L0 * L0 + L1 * L1 + L2 * L2 + 2.0 * (L3 * L3 + L4 * L4 + L5 * L5);
        e += 0.5 * (fro - tr * tr / 3.0);
        a.mu[3 * i] = m0;
        a.mu[3 * i + 1] = m1;
        a.mu[3 * i + 2] = m2;
        for (int k = 0; k < 6; ++k) a.lamb[6 * i + k] = double(acc.l[k]);
        a.mu32[i] = float4{float(m0), float(m1), float(m2), 0.0f};
        a.lamb32[2 * i] = float4{float(L0), float(L1), float(L2), float(L3)};
        a.lamb32[2 * i + 1] = float4{float(L4), float(L5), 0.0f, 0.0f};
      }
      a.rho[i] = rh;
      a.over[i] = over;
      a.Fp[i] = dF;
      a.Fp32[i] = float(dF);
      a.e_atom[i] = e;
      a.lim[i] = lim;
    }
  }
}

// --- pass 2: forces, pair energy, virial -------------------------------------
template <class R> struct AtomState {
  R Fp;
  R mu[3];
  R l[6];
};

template <class R> __device__ inline void quad_terms(const R *L, R d0, R d1, R d2, R &nu, R xi[3]) {
  // L = (xx, yy, zz, xy, xz, yz); quad_nu / quad_xi (adp_force.hpp:63-68)
  const R Md0 = L[0] * d0 + L[3] * d1 + L[4] * d2;
  const R Md1 = L[3] * d0 + L[1] * d1 + L[5] * d2;
  const R Md2 = L[4] * d0 + L[5] * d1 + L[2] * d2;
  const R tr = L[0] + L[1] + L[2];
  const R dd = d0 * d0 + d1 * d1 + d2 * d2;
  nu += (d0 * Md0 + d1 * Md1 + d2 * Md2) - dd / R(3) * tr;
  const R t3 = tr / R(3);
  xi[0] += Md0 - t3 * d0;
  xi[1] += Md1 - t3 * d1;
  xi[2] += Md2 - t3 * d2;
}

template <class R> struct ForceAcc {
  R f0 = 0, f1 = 0, f2 = 0;
  double ep = 0.0;
  double sg[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
};

template <int KIND, bool EXACT, class R>
__device__ inline void force_pair(const Tabs &tb, const EvalArgs &a, int ti, int64_t cbase,
                                  const AtomState<R> &si, uint32_t b, R d0, R d1, R d2, R r,
                                  ForceAcc<R> &acc, double r64 = 0.0) {
  if (b & kFlagTiny) return;
  const int tj = pj_type(b);
  const int64_t jg = cbase + pj_local(b);
  const R inv_r = R(1) / r;
  R phi = 0, dphi = 0, dgj = 0, dgi = 0;
  const int o = ord_pair(ti, tj, a.ntypes);
  auto ev = [&](int t, R &v, R &d) {
    if (EXACT) {
      double vv, dd;
      tb.eval64(t, double(r), vv, dd);
      v = R(vv);
      d = R(dd);
    } else {
      float vv, dd;
      tb.eval32(t, r64, vv, dd);
      v = R(vv);
      d = R(dd);
    }
  };
  auto dg = [&](int t, R &d) {
    if (EXACT) {
      double vv, dd;
      tb.eval64(t, double(r), vv, dd);
      d = R(dd);
    } else {
      float dd;
      tb.deriv32(t, r64, dd);
      d = R(dd);
    }
  };
  double phi64 = 0.0;
  if (b & kFlagPhi) {
    if (EXACT) ev(o, phi, dphi);
    else {
      float dd;
      tb.deriv32_value64(o, r64, dd, phi64);
      dphi = R(dd);
    }
  }
  if (b & kFlagGj) dg(a.npair + tj, dgj);
  if (b & kFlagGi) dg(a.npair + ti, dgi);
  const R Fpj = EXACT ? R(a.Fp[jg]) : R(a.Fp32[jg]);
  const R fs = (dphi + si.Fp * dgj + Fpj * dgi) * inv_r;
  R f0 = fs * d0, f1 = fs * d1, f2 = fs * d2;
  if (KIND == FS_ADP) {
    R muj[3], lj[6];
    if constexpr (std::is_same<R, float>::value) {
      // the FP32 copies: the same roundings as R(double), in three loads
      const float4 m = a.mu32[jg], l0 = a.lamb32[2 * jg], l1 = a.lamb32[2 * jg + 1];
      muj[0] = m.x;
      muj[1] = m.y;
      muj[2] = m.z;
      lj[0] = l0.x;
      lj[1] = l0.y;
      lj[2] = l0.z;
      lj[3] = l0.w;
      lj[4] = l1.x;
      lj[5] = l1.y;
    } else {
      for (int k = 0; k < 3; ++k) muj[k] = R(a.mu[3 * jg + k]);
      for (int k = 0; k < 6; ++k) lj[k] = R(a.lamb[6 * jg + k]);
    }
    if (b & kFlagDip) {
      R u, du;
      ev(a.npair + 2 * a.ntypes + o, u, du);
      const R doti = si.mu[0] * d0 + si.mu[1] * d1 + si.mu[2] * d2;
      const R dotj = muj[0] * d0 + muj[1] * d1 + muj[2] * d2;
      const R sd = du * inv_r * (doti - dotj);
      f0 += sd * d0 + u * (si.mu[0] - muj[0]);
      f1 += sd * d1 + u * (si.mu[1] - muj[1]);
      f2 += sd * d2 + u * (si.mu[2] - muj[2]);
    }
    if (b & kFlagQuad) {
      R w, dw;
      ev(2 * a.npair + 2 * a.ntypes + o, w, dw);
      R nu = 0, xi[3] = {0, 0, 0};
      quad_terms(si.l, d0, d1, d2, nu, xi);
      quad_terms(lj, d0, d1, d2, nu, xi);
      const R sq = dw * inv_r * nu;
      f0 += sq * d0 + R(2) * w * xi[0];
      f1 += sq * d1 + R(2) * w * xi[1];
      f2 += sq * d2 + R(2) * w * xi[2];
    }
  }
  acc.f0 += f0;
  acc.f1 += f1;
  acc.f2 += f2;
  // pair energy value in FP64 from the FP64 distance in both modes
  acc.ep += EXACT ? 0.5 * double(phi) : 0.5 * phi64;
  const double hd[3] = {0.5 * double(d0), 0.5 * double(d1), 0.5 * double(d2)};
  const double fd[3] = {double(f0), double(f1), double(f2)};
#pragma unroll
  for (int p = 0; p < 3; ++p)
#pragma unroll
    for (int q = 0; q < 3; ++q) acc.sg[3 * p + q] -= hd[p] * fd[q];
}

constexpr int kPart = 11; // energy, limit, 9 virial components

// the mixed ADP force pass asks for 4 waves per SIMD (128 VGPRs, some values
// in scratch; the compiler's own 158 allow 3): 0.231 -> 0.206 ms for the
// residual of the binary 200-configuration workload. 5 measured slower, and
// 6 for the EAM pass (80 VGPRs) measured no faster.
template <int KIND, bool F64, int LPA, bool ST>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(KIND == FS_ADP && !F64 ? 4 : 1))) void
k_force(EvalArgs a) {
  using R = typename std::conditional<F64, double, float>::type;
  constexpr int APB = 256 / LPA;
  extern __shared__ float4 smem[];
  TableDesc *sT = reinterpret_cast<TableDesc *>(smem);
  __shared__ double sRed[APB][kPart];
  const float4 *sI;
  const double *sP;
  stage_tables<ST>(a, sT, smem + a.ntab * int(sizeof(TableDesc) / sizeof(float4)), sI, sP);
  const Tabs tb{sT, sI, a.tx, a.ty, a.ts, sP};
  const int sub = threadIdx.x / LPA, lane = threadIdx.x % LPA;
  for (int tile = blockIdx.x; tile < a.ntiles; tile += gridDim.x) {
    const int na = a.tile_na[tile];
    if (sub < na) {
      const int64_t i = int64_t(a.tile_a0[tile]) + sub;
      const int c = a.tile_cfg[tile];
      const int64_t cbase = a.atom_off[c];
      const int ti = a.types[i];
      AtomState<R> si;
      si.Fp = F64 ? R(a.Fp[i]) : R(a.Fp32[i]);
      if (KIND == FS_ADP) {
        for (int k = 0; k < 3; ++k) si.mu[k] = R(a.mu[3 * i + k]);
        for (int k = 0; k < 6; ++k) si.l[k] = R(a.lamb[6 * i + k]);
      }
      ForceAcc<R> acc;
      const int64_t p0 = a.pair_off[i], p1 = a.pair_off[i + 1];
      for (int64_t p = p0 + lane; p < p1; p += LPA) {
        if constexpr (F64) {
          const Pair<double> q = load_pair64<true>(a, p);
          force_pair<KIND, true, double>(tb, a, ti, cbase, si, q.b, q.d0, q.d1, q.d2, q.r, acc);
        } else {
          const float4 q = a.rec[p];
          const uint32_t b = __float_as_uint(q.w);
          if (b & kFlagEdge) {
            const Pair<double> e = load_pair64<true>(a, p);
            AtomState<double> sd;
            sd.Fp = a.Fp[i];
            if (KIND == FS_ADP) {
              for (int k = 0; k < 3; ++k) sd.mu[k] = a.mu[3 * i + k];
              for (int k = 0; k < 6; ++k) sd.l[k] = a.lamb[6 * i + k];
            }
            ForceAcc<double> t;
            force_pair<KIND, true, double>(tb, a, ti, cbase, sd, e.b, e.d0, e.d1, e.d2, e.r, t);
            acc.f0 += R(t.f0);
            acc.f1 += R(t.f1);
            acc.f2 += R(t.f2);
            acc.ep += t.ep;
            for (int k = 0; k < 9; ++k) acc.sg[k] += t.sg[k];
          } else {
            const float r = sqrtf(q.x * q.x + q.y * q.y + q.z * q.z);
            force_pair<KIND, false, float>(tb, a, ti, cbase, si, b, q.x, q.y, q.z, r, acc, a.r64[p]);
          }
        }
      }
      acc.f0 = sub_sum<LPA>(acc.f0);
      acc.f1 = sub_sum<LPA>(acc.f1);
      acc.f2 = sub_sum<LPA>(acc.f2);
      acc.ep = sub_sum<LPA>(acc.ep);
      for (int k = 0; k < 9; ++k) acc.sg[k] = sub_sum<LPA>(acc.sg[k]);
      if (lane == 0) {
        const double F[3] = {double(acc.f0), double(acc.f1), double(acc.f2)};
        const int64_t row = a.row_off[c] + 3 * (i - cbase);
        for (int k = 0; k < 3; ++k) {
          a.forces[3 * i + k] = F[k];
          a.residual[row + k] = F[k] - a.ref_force[3 * i + k];
        }
        sRed[sub][0] = a.e_atom[i] + acc.ep;
        sRed[sub][1] = a.lim[i];
        for (int k = 0; k < 9; ++k) sRed[sub][2 + k] = acc.sg[k];
      }
    }
    __syncthreads();
    if (threadIdx.x < kPart) {
      double s = 0.0;
      for (int q = 0; q < na && q < APB; ++q) s += sRed[q][threadIdx.x];
      a.tile_part[int64_t(tile) * kPart + threadIdx.x] = s;
    }
    __syncthreads();
  }
}

// --- per-configuration finalisation ------------------------------------------
__global__ void k_finalize(int C, const int32_t *__restrict__ cfg_tile0,
                           const double *__restrict__ part, const double *__restrict__ vol,
                           const int64_t *__restrict__ row_off,
                           const int64_t *__restrict__ atom_off,
                           const double *__restrict__ ref_energy,
                           const double *__restrict__ ref_stress, double ew, double sw,
                           double *__restrict__ residual, double *__restrict__ E,
                           double *__restrict__ S, double *__restrict__ L) {
  // one 64-thread block per configuration: thread t sums tiles t, t + 64, ...
  // then a fixed-shape tree combines the 64 partial sums (deterministic)
  const int c = blockIdx.x;
  if (c >= C) return;
  __shared__ double sred[64][kPart];
  double acc[kPart];
  for (int k = 0; k < kPart; ++k) acc[k] = 0.0;
  for (int t = cfg_tile0[c] + int(threadIdx.x); t < cfg_tile0[c + 1]; t += 64)
    for (int k = 0; k < kPart; ++k) acc[k] += part[int64_t(t) * kPart + k];
  for (int k = 0; k < kPart; ++k) sred[threadIdx.x][k] = acc[k];
  __syncthreads();
  for (int w = 32; w > 0; w >>= 1) {
    if (int(threadIdx.x) < w)
      for (int k = 0; k < kPart; ++k) sred[threadIdx.x][k] += sred[threadIdx.x + w][k];
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  for (int k = 0; k < kPart; ++k) acc[k] = sred[0][k];
  const double V = vol[c];
  double sig[9];
  for (int k = 0; k < 9; ++k) sig[k] = acc[2 + k] / V;
  int64_t row = row_off[c] + 3 * (atom_off[c + 1] - atom_off[c]);
  residual[row++] = ew * (acc[0] - ref_energy[c]);
  if (sw > 0.0) {
    const int vo[6] = {0, 4, 8, 1, 2, 5};
    for (int v = 0; v < 6; ++v) residual[row++] = sw * (sig[vo[v]] - ref_stress[6 * c + v]);
  }
  residual[row] = acc[1];
  E[c] = acc[0];
  L[c] = acc[1];
  for (int k = 0; k < 9; ++k) S[9 * c + k] = sig[k];
}

// --- |F|^2, deterministic two-level sum ---------------------------------------
// curvature rows w (y[k-1] - 2 y[k] + y[k+1]) (SplinePotential::
// write_curvature, src/potentials/spline.cpp:205-213)
__device__ inline double curvature_row(const int4 rw, const double *y, double w) {
  return w * (y[rw.x] - 2.0 * y[rw.x + 1] + y[rw.x + 2]);
}

__global__ void k_smooth_res(int n, const int4 *__restrict__ rows, const double *__restrict__ y,
                             double w, double *__restrict__ out) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < n) out[r] = curvature_row(rows[r], y, w);
}

__global__ void k_sumsq(int64_t n, const double *__restrict__ v, double *__restrict__ part) {
  __shared__ double s[256];
  double acc = 0.0;
  for (int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; k < n;
       k += int64_t(gridDim.x) * blockDim.x)
    acc += v[k] * v[k];
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}
__global__ void k_sum_final(int n, const double *__restrict__ part, double *__restrict__ out) {
  __shared__ double s[256];
  double acc = 0.0;
  for (int k = threadIdx.x; k < n; k += blockDim.x) acc += part[k];
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = s[0];
}

// ---------------------------------------------------------------------------
// Batched population cost (EAM, mixed or FP64): |F(x_m)|^2 for a batch of
// parameter vectors in a few launches instead of one residual per member.
// Block b works on M members over tile group b % G; the G = 8 groups are
// contiguous tile ranges, one per XCD under the round-robin block dispatch,
// so the blocks of the members in flight on an XCD share that group's pair
// records in its L2. Each block keeps its members' coefficients in LDS for
// its whole tile range.
constexpr int kPopG = 8;
constexpr int kPopE = 2;  // density pass: embedding energy, limit
constexpr int kPopF = 8;  // force pass: |F - F_ref|^2, pair energy, 6 virial

struct PopArgs {
  int nb;                            // members in this batch
  int nknots, nivl;
  int stage_ivl;                     // member interval forms staged in LDS (else global)
  const double *y, *s;               // [nb][nknots] member knots and slopes
  const float4 *ivl32;               // [nb][2 nivl] FP32 hi/lo interval forms
  const double *ivl64;               // [nb][4 nivl] FP64 interval forms
  double *Fp;                        // [nb][N] F'(rho) FP64
  float *Fp32;                       // [nb][N] F'(rho) FP32
  const int32_t *grp_t0;             // [G+1] tile range of each group
  double *partE;                     // [nb][C][G][kPopE]
  double *partF;                     // [nb][C][G][kPopF]
  int C;
};

// per-member tables: knots = base knots with the free ones from X[m]
template <bool ST>
__global__ void k_pop_tables(int ntab, int D, int nknots, const TableDesc *__restrict__ tabs,
                             const double *__restrict__ x, const double *__restrict__ y0,
                             const int32_t *__restrict__ knot_of, const double *__restrict__ X,
                             double *__restrict__ y, double *__restrict__ s, double *__restrict__ S,
                             float4 *__restrict__ ivl, double *__restrict__ p64, int nivl) {
  const int m = blockIdx.x;
  double *ym = y + size_t(m) * nknots;
  for (int k = threadIdx.x; k < nknots; k += blockDim.x) ym[k] = y0[k];
  __syncthreads();
  for (int c = threadIdx.x; c < D; c += blockDim.x) ym[knot_of[c]] = X[size_t(m) * D + c];
  __syncthreads();
  tables_staged<ST>(ntab, tabs, nknots, x, ym, s + size_t(m) * nknots, S + size_t(m) * 5 * nknots,
               ivl + size_t(m) * 2 * nivl, p64 + size_t(m) * 4 * nivl);
}

// The population passes take M members per block (block b: tile group
// b % kPopG, members M (b / kPopG) .. + M). Per pair the record, the distance
// and the interval location are formed once and applied to every member's
// coefficients staged in LDS: in FP32 mode the intervals' power forms located
// from the FP64 distance (the residual kernels' location), in FP64 mode the
// cached-site interval and Hermite basis of spl_site64. Per member the sums
// keep the single-member order (per atom over its lanes, per tile in atom
// order, per configuration in tile order).

// cached-site lookup split into its member-independent part (interval,
// linear-site flag and Hermite basis: spl_site64 with the same operations)
// and the member's values
struct Site64 {
  int i;
  bool lin;
  double inv_dx, off, bv[4], bd[4];
};

__device__ inline Site64 site64(const TableDesc &T, const double *x, double r) {
  Site64 s;
  const int n = T.n;
  if (r <= x[0]) {
    s.i = 0;
    s.lin = true;
  } else if (r >= x[n - 1]) {
    s.i = n - 2;
    s.lin = true;
  } else {
    s.i = spl_locate(x, n, r);
    s.lin = T.linear != 0;
  }
  const double dx = x[s.i + 1] - x[s.i];
  if (s.lin) {
    s.inv_dx = 1.0 / dx;
    s.off = r - x[s.i];
  } else {
    herm_basis64(x[s.i], dx, r, s.bv, s.bd);
  }
  return s;
}

__device__ inline void site64_apply(const Site64 &s, const double *y, const double *sl, double &v,
                                    double &d) {
  const int i = s.i;
  if (s.lin) {
    v = y[i] + s.off * s.inv_dx * (y[i + 1] - y[i]);
    d = (y[i + 1] - y[i]) * s.inv_dx;
    return;
  }
  const double yi = y[i], yi1 = y[i + 1], si = sl[i], si1 = sl[i + 1];
  v = s.bv[0] * yi + s.bv[1] * yi1 + s.bv[2] * si + s.bv[3] * si1;
  d = s.bd[0] * yi + s.bd[1] * yi1 + s.bd[2] * si + s.bd[3] * si1;
}

// FP64 value of interval k at offset u from its FP32 hi/lo power form (the
// hi + lo coefficients carry the FP64 form to ~1e-15)
__device__ inline double hilo_value64(const float4 &q, const float4 &e, double u) {
  const double c0 = double(q.x) + double(e.x), c1 = double(q.y) + double(e.y);
  const double c2 = double(q.z) + double(e.z), c3 = double(q.w) + double(e.w);
  return c0 + u * (c1 + u * (c2 + u * c3));
}

__device__ inline float hilo_deriv32(const TableDesc &T, const float4 &q, const float4 &e, float t) {
  const float dt = (q.y + t * (2.0f * q.z + 3.0f * t * q.w)) + (e.y + t * (2.0f * e.z + 3.0f * t * e.w));
  return spl_dscale(T, dt);
}

template <int LPA, int M>
__global__ __launch_bounds__(256) void k_pop_density(EvalArgs a, PopArgs p) {
  constexpr int APB = 256 / LPA;
  extern __shared__ float4 smem[];
  TableDesc *sT = reinterpret_cast<TableDesc *>(smem);
  __shared__ double sRed[APB][M * kPopE];
  const int g = blockIdx.x % kPopG, m0 = (blockIdx.x / kPopG) * M;
  const int nm = min(M, p.nb - m0);
  for (int k = threadIdx.x; k < a.ntab; k += blockDim.x) sT[k] = a.tabs[k];
  const size_t per = size_t(4) * p.nivl; // FP64 interval forms of one member
  const double *P64 = p.ivl64 + size_t(m0) * per;
  if (p.stage_ivl) {
    double *l64 = reinterpret_cast<double *>(smem + a.ntab * int(sizeof(TableDesc) / sizeof(float4)));
    for (size_t k = threadIdx.x; k < size_t(nm) * per; k += blockDim.x) l64[k] = P64[k];
    P64 = l64;
  }
  __syncthreads();
  const int sub = threadIdx.x / LPA, lane = threadIdx.x % LPA;
  // sRed[sub]: the running sums of atom slot sub over the tiles of the
  // current configuration, kept by the slot's lane 0; the block reduces them
  // when the configuration changes (the tiles of a configuration are
  // consecutive)
  if (lane == 0)
    for (int k = 0; k < M * kPopE; ++k) sRed[sub][k] = 0.0;
  int run_c = -1;
  auto flush = [&]() {
    __syncthreads();
    if (int(threadIdx.x) < nm * kPopE) {
      const int m = threadIdx.x / kPopE, comp = threadIdx.x % kPopE;
      double v = 0.0;
      for (int q = 0; q < APB; ++q) v += sRed[q][threadIdx.x];
      p.partE[((size_t(m0 + m) * p.C + run_c) * kPopG + g) * kPopE + comp] = v;
    }
    __syncthreads();
    if (lane == 0)
      for (int k = 0; k < M * kPopE; ++k) sRed[sub][k] = 0.0;
  };
  for (int tile = p.grp_t0[g]; tile < p.grp_t0[g + 1]; ++tile) {
    const int na = a.tile_na[tile];
    const int c = a.tile_cfg[tile];
    if (c != run_c) {
      if (run_c >= 0) flush();
      run_c = c;
    }
    if (sub < na) {
      const int64_t i = int64_t(a.tile_a0[tile]) + sub;
      const int ti = a.types[i];
      double rho[M];
#pragma unroll
      for (int m = 0; m < M; ++m) rho[m] = 0.0;
      for (int64_t q = a.pair_off[i] + lane; q < a.pair_off[i + 1]; q += LPA) {
        const uint32_t b = a.bits[q];
        if (b & kFlagGj) {
          const TableDesc D = sT[a.npair + pj_type(b)];
          double u;
          const int k = spl_interval64(D, a.r64[q], u);
#pragma unroll
          for (int m = 0; m < M; ++m)
            if (m < nm) rho[m] += spl_value64(D, P64 + m * per, k, u);
        }
      }
#pragma unroll
      for (int m = 0; m < M; ++m) rho[m] = sub_sum<LPA>(rho[m]);
      if (lane == 0) {
        const TableDesc Te = sT[a.npair + a.ntypes + ti];
#pragma unroll
        for (int m = 0; m < M; ++m) {
          if (m >= nm) break;
          const GaugedRho gr = gauge_clamp(a.gauge, ti, Te, rho[m]);
          const size_t mm = size_t(m0 + m);
          double F, dF;
          spl_uncached64(Te, a.tx + Te.off, p.y + mm * p.nknots + Te.off, p.s + mm * p.nknots + Te.off,
                         gr.u, F, dF);
          p.Fp[mm * a.N + i] = dF;
          p.Fp32[mm * a.N + i] = float(dF);
          sRed[sub][m * kPopE] += F + gr.e_shift;
          sRed[sub][m * kPopE + 1] += gr.lim;
        }
      }
    }
  }
  if (run_c >= 0) flush();
}

// per member and atom: force (FP32 or FP64), pair energy, virial in the six
// residual components xx, yy, zz, xy, xz, yz (-1/2 d_p f_q, p <= q)
template <class AT> struct PopAcc {
  AT f0 = 0, f1 = 0, f2 = 0;
  double ep = 0.0;
  double sg[6] = {0, 0, 0, 0, 0, 0};
};

template <class AT, class FT>
__device__ inline void pop_add(PopAcc<AT> &A, FT f0, FT f1, FT f2, double ep, double d0, double d1,
                               double d2) {
  A.f0 += AT(f0);
  A.f1 += AT(f1);
  A.f2 += AT(f2);
  A.ep += ep;
  const double h0 = 0.5 * d0, h1 = 0.5 * d1, h2 = 0.5 * d2;
  const double g0 = double(f0), g1 = double(f1), g2 = double(f2);
  A.sg[0] -= h0 * g0;
  A.sg[1] -= h1 * g1;
  A.sg[2] -= h2 * g2;
  A.sg[3] -= h0 * g1;
  A.sg[4] -= h0 * g2;
  A.sg[5] -= h1 * g2;
}

// F64: every pair through the FP64 cached-site path with FP64 sums; otherwise
// FP32 pairs with the FP64 path only for pairs at an exact span edge, as the
// residual kernels. LDS: the members' FP32 interval forms (FP32 mode) or knot
// values and slopes (FP64 mode), when they fit the staging budget.
// Mixed precision asks for 3 waves per SIMD: at 4 members per block the
// compiler's own pick (191 VGPRs, 2 waves) made generations 7 % slower at
// C5's shape; the FP64 form measured 13 % slower at 3 waves (more spills),
// so it keeps the compiler's choice.
template <int LPA, bool F64, int M>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(F64 ? 1 : 3))) void k_pop_force(
    EvalArgs a, PopArgs p) {
  using AT = typename std::conditional<F64, double, float>::type;
  constexpr int APB = 256 / LPA;
  extern __shared__ float4 smem[];
  TableDesc *sT = reinterpret_cast<TableDesc *>(smem);
  __shared__ double sRed[APB][M * kPopF];
  const int g = blockIdx.x % kPopG, m0 = (blockIdx.x / kPopG) * M;
  const int nm = min(M, p.nb - m0);
  const int nq = a.ntab * int(sizeof(TableDesc) / sizeof(float4));
  for (int k = threadIdx.x; k < nq; k += blockDim.x) smem[k] = reinterpret_cast<const float4 *>(a.tabs)[k];
  const size_t per32 = size_t(2) * p.nivl; // float4s of one member's FP32 interval forms
  const size_t perk = size_t(p.nknots);
  const float4 *I32 = p.ivl32 + size_t(m0) * per32;
  const double *Y = p.y + size_t(m0) * perk, *S = p.s + size_t(m0) * perk;
  if (p.stage_ivl) {
    if constexpr (F64) {
      double *ly = reinterpret_cast<double *>(smem + nq);
      for (size_t k = threadIdx.x; k < size_t(nm) * perk; k += blockDim.x) {
        ly[k] = Y[k];
        ly[size_t(M) * perk + k] = S[k];
      }
      Y = ly;
      S = ly + size_t(M) * perk;
    } else {
      float4 *l32 = smem + nq;
      for (size_t k = threadIdx.x; k < size_t(nm) * per32; k += blockDim.x) l32[k] = I32[k];
      I32 = l32;
    }
  }
  __syncthreads();
  const int sub = threadIdx.x / LPA, lane = threadIdx.x % LPA;
  // sRed[sub]: running sums of atom slot sub over the tiles of the current
  // configuration (its lane 0), reduced by the block when the configuration
  // changes
  if (lane == 0)
    for (int k = 0; k < M * kPopF; ++k) sRed[sub][k] = 0.0;
  int run_c = -1;
  auto flush = [&]() {
    __syncthreads();
    if (int(threadIdx.x) < nm * kPopF) {
      const int m = threadIdx.x / kPopF, comp = threadIdx.x % kPopF;
      double v = 0.0;
      for (int q = 0; q < APB; ++q) v += sRed[q][threadIdx.x];
      p.partF[((size_t(m0 + m) * p.C + run_c) * kPopG + g) * kPopF + comp] = v;
    }
    __syncthreads();
    if (lane == 0)
      for (int k = 0; k < M * kPopF; ++k) sRed[sub][k] = 0.0;
  };
  for (int tile = p.grp_t0[g]; tile < p.grp_t0[g + 1]; ++tile) {
    const int na = a.tile_na[tile];
    const int c = a.tile_cfg[tile];
    if (c != run_c) {
      if (run_c >= 0) flush();
      run_c = c;
    }
    if (sub < na) {
      const int64_t i = int64_t(a.tile_a0[tile]) + sub;
      const int64_t cbase = a.atom_off[c];
      const int ti = a.types[i];
      AT Fpi[M];
      PopAcc<AT> acc[M];
#pragma unroll
      for (int m = 0; m < M; ++m) {
        const size_t mm = size_t(m0 + (m < nm ? m : 0)) * a.N + i;
        Fpi[m] = F64 ? AT(p.Fp[mm]) : AT(p.Fp32[mm]);
      }
      // FP64 mode: every pair through the FP64 sites; FP32 mode: the FP32
      // pairs first, then (their own loop, so their FP64 bases do not add to
      // the main loop's registers) the few pairs at an exact span edge
      for (int pass = F64 ? 1 : 0; pass < 2; ++pass)
      for (int64_t q = a.pair_off[i] + lane; q < a.pair_off[i + 1]; q += LPA) {
        const float4 rec = a.rec[q];
        const uint32_t b = __float_as_uint(rec.w);
        if (b & kFlagTiny) continue;
        const bool exact = F64 || (b & kFlagEdge);
        if (exact != (pass == 1)) continue;
        const int tj = pj_type(b);
        const int64_t jg = cbase + pj_local(b);
        const int o = ord_pair(ti, tj, a.ntypes);
        if (exact) {
          // FP64 cached sites: interval and basis once, the members' values
          const Pair<double> e = load_pair64<true>(a, q);
          const double inv_r = 1.0 / e.r;
          // one table's site at a time, applied to every member's knots and
          // slopes (staged in FP64 mode, else read in place)
          double phi[M], dphi[M], dgj[M], dgi[M], vv;
#pragma unroll
          for (int m = 0; m < M; ++m) phi[m] = dphi[m] = dgj[m] = dgi[m] = 0.0;
          auto each = [&](int t, double *v, double *d) {
            const TableDesc &T = sT[t];
            const Site64 st = site64(T, a.tx + T.off, e.r);
#pragma unroll
            for (int m = 0; m < M; ++m)
              if (m < nm) site64_apply(st, Y + size_t(m) * perk + T.off, S + size_t(m) * perk + T.off,
                                       v ? v[m] : vv, d[m]);
          };
          if (b & kFlagPhi) each(o, phi, dphi);
          if (b & kFlagGj) each(a.npair + tj, nullptr, dgj);
          if (b & kFlagGi) each(a.npair + ti, nullptr, dgi);
#pragma unroll
          for (int m = 0; m < M; ++m) {
            if (m >= nm) break;
            const double fpi = F64 ? double(Fpi[m]) : p.Fp[size_t(m0 + m) * a.N + i];
            const double fpj = p.Fp[size_t(m0 + m) * a.N + jg];
            const double fs = (dphi[m] + fpi * dgj[m] + fpj * dgi[m]) * inv_r;
            pop_add(acc[m], fs * e.d0, fs * e.d1, fs * e.d2, 0.5 * phi[m], e.d0, e.d1, e.d2);
          }
        } else if constexpr (!F64) {
          const float r = sqrtf(rec.x * rec.x + rec.y * rec.y + rec.z * rec.z);
          const float inv_r = 1.0f / r;
          const double r64 = a.r64[q];
          int kp = -1, kj = -1, ki = -1;
          double up = 0.0, u;
          float tp = 0.f, tjf = 0.f, tif = 0.f;
          if (b & kFlagPhi) kp = spl_interval64(sT[o], r64, up), tp = float(up);
          if (b & kFlagGj) kj = spl_interval64(sT[a.npair + tj], r64, u), tjf = float(u);
          if (b & kFlagGi) ki = spl_interval64(sT[a.npair + ti], r64, u), tif = float(u);
          const int ip = kp >= 0 ? 2 * (sT[o].ivl + kp) : 0;
          const int ij = kj >= 0 ? 2 * (sT[a.npair + tj].ivl + kj) : 0;
          const int ii = ki >= 0 ? 2 * (sT[a.npair + ti].ivl + ki) : 0;
#pragma unroll
          for (int m = 0; m < M; ++m) {
            if (m >= nm) break;
            const float4 *Im = I32 + size_t(m) * per32;
            float dphi = 0.f, dgj = 0.f, dgi = 0.f;
            double phi64 = 0.0;
            if (kp >= 0) {
              const float4 q4 = Im[ip], e4 = Im[ip + 1];
              dphi = hilo_deriv32(sT[o], q4, e4, tp);
              phi64 = hilo_value64(q4, e4, up);
            }
            if (kj >= 0) dgj = hilo_deriv32(sT[a.npair + tj], Im[ij], Im[ij + 1], tjf);
            if (ki >= 0) dgi = hilo_deriv32(sT[a.npair + ti], Im[ii], Im[ii + 1], tif);
            const float fpj = p.Fp32[size_t(m0 + m) * a.N + jg];
            const float fs = (dphi + float(Fpi[m]) * dgj + fpj * dgi) * inv_r;
            pop_add(acc[m], fs * rec.x, fs * rec.y, fs * rec.z, 0.5 * phi64, double(rec.x),
                    double(rec.y), double(rec.z));
          }
        }
      }
#pragma unroll
      for (int m = 0; m < M; ++m) {
        acc[m].f0 = sub_sum<LPA>(acc[m].f0);
        acc[m].f1 = sub_sum<LPA>(acc[m].f1);
        acc[m].f2 = sub_sum<LPA>(acc[m].f2);
        acc[m].ep = sub_sum<LPA>(acc[m].ep);
#pragma unroll
        for (int k = 0; k < 6; ++k) acc[m].sg[k] = sub_sum<LPA>(acc[m].sg[k]);
      }
      if (lane == 0) {
#pragma unroll
        for (int m = 0; m < M; ++m) {
          if (m >= nm) break;
          const double F[3] = {double(acc[m].f0), double(acc[m].f1), double(acc[m].f2)};
          double r2 = 0.0;
          for (int k = 0; k < 3; ++k) {
            const double d = F[k] - a.ref_force[3 * i + k];
            r2 += d * d;
          }
          double *row = &sRed[sub][m * kPopF];
          row[0] += r2;
          row[1] += acc[m].ep;
#pragma unroll
          for (int k = 0; k < 6; ++k) row[2 + k] += acc[m].sg[k];
        }
      }
    }
  }
  if (run_c >= 0) flush();
}

// cost of every (member, configuration): the residual rows of
// write_config_residuals (residual_layout.hpp:37-54), squared and summed
__global__ void k_pop_finalize(int nb, int C, const int32_t *__restrict__ grp_t0,
                               const int32_t *__restrict__ cfg_tile0, const double *__restrict__ partE,
                               const double *__restrict__ partF, const double *__restrict__ vol,
                               const double *__restrict__ ref_energy,
                               const double *__restrict__ ref_stress, double ew, double sw,
                               double *__restrict__ cost_mc) {
  const int64_t id = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (id >= int64_t(nb) * C) return;
  const int m = int(id / C), c = int(id % C);
  double e = 0.0, lim = 0.0, r2 = 0.0, vir[6] = {0, 0, 0, 0, 0, 0};
  const int t0 = cfg_tile0[c], t1 = cfg_tile0[c + 1];
  for (int g = 0; g < kPopG; ++g) {
    if (!(grp_t0[g] < t1 && grp_t0[g + 1] > t0)) continue; // group holds no tile of c
    const double *pe = partE + ((size_t(m) * C + c) * kPopG + g) * kPopE;
    const double *pf = partF + ((size_t(m) * C + c) * kPopG + g) * kPopF;
    e += pe[0] + pf[1];
    lim += pe[1];
    r2 += pf[0];
    for (int k = 0; k < 6; ++k) vir[k] += pf[2 + k];
  }
  const double er = ew * (e - ref_energy[c]);
  double cost = r2 + er * er + lim * lim;
  if (sw > 0.0) {
    for (int v = 0; v < 6; ++v) { // xx, yy, zz, xy, xz, yz
      const double sr = sw * (vir[v] / vol[c] - ref_stress[6 * c + v]);
      cost += sr * sr;
    }
  }
  cost_mc[id] = cost;
}

// member cost: its configurations' costs, then its curvature rows squared
__global__ void k_pop_sum(int nb, int C, const double *__restrict__ cost_mc, int n_smooth,
                          const int4 *__restrict__ rows, double w, const double *__restrict__ y,
                          int nknots, double *__restrict__ cost) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= nb) return;
  double s = 0.0;
  for (int c = 0; c < C; ++c) s += cost_mc[size_t(m) * C + c];
  const double *ym = y + size_t(m) * nknots;
  for (int r = 0; r < n_smooth; ++r) {
    const double v = curvature_row(rows[r], ym, w);
    s += v * v;
  }
  cost[m] = s;
}

// ---------------------------------------------------------------------------
// Member-lane population passes. When every pair and density table shares one
// knot grid (the same abscissae, hence the same span), a pair's interval and
// Hermite basis are the same for its three table roles and for every member.
// A wave then holds 64 members, one per lane. It forms the sites of 64 pairs at
// a time with lanes as pairs, then walks those pairs with each site broadcast
// from its staging lane (v_readlane into scalar registers), every lane applying
// its own member's knot values and slopes. Member data is member-minor, so
// each member access is one contiguous 1 KB row: the knot values and slopes
// ([group][knot][64] (y, s), staged in LDS), the embedding derivatives
// ([group][atom][64], read for the partner of every pair).
// The arithmetic is FP64 throughout, with the cached-site semantics of the
// FP64 path (spl_locate, linear sites at the span ends, the Hermite basis
// dotted with (y_i, y_i+1, s_i, s_i+1)). On gfx950, FP64 FMA issues at the
// rate of unpacked FP32 FMA; the passes are bound by instruction issue and the
// member rows read from LDS, which FP32 would halve only with hi/lo pairs
// (the coefficient rounding of single FP32 values biases every pair of an
// interval alike).
// A block (16 waves, one per CU: the member rows fill the LDS) holds one
// member group and the tables of one centre type: slots 0..nt-1 the pair
// tables (ti, t), nt..2nt-1 the density tables, then (density pass) the
// embedding table of ti. A work item is (member group, range of wave tiles of
// one type); a wave tile is up to kMlTile atoms of one type in one
// configuration, and writes its own partial sums per member, which
// k_pop_ml_finalize sums per configuration in a fixed order. Waves take the
// item's tiles from an LDS counter; the sums do not depend on which wave takes
// which tile.
// ---------------------------------------------------------------------------
constexpr int kMlWaves = 16;
constexpr int kMlTile = 32;
constexpr int kMlE = 2; // density pass partials: energy (embedding, gauge shift, pair), limit
constexpr int kMlF = 7; // force pass partials: |F - F_ref|^2, virial xx, yy, zz, xy, xz, yz
constexpr int kMlRanges = 64; // wave-tile ranges per member group (split by pair counts)

struct MlArgs {
  int ng, nb;   // member groups, members
  int n;        // knots of the shared grid
  int linear;   // its tables are piecewise linear
  int nknots;   // all tables' knots (rows of a member group)
  int nwt, nrange;
  const double *x;    // abscissae of the shared grid
  int xrows;          // LDS entries (double2) before the member rows: abscissae
  const double2 *ys;  // [ng][nknots][64] knot value and slope of each member
  double *Fp;         // [ng][N][64] F'(rho) of each member
  double *pE;         // [ng][nwt][kMlE][64]
  double *pF;         // [ng][nwt][kMlF][64]
  const int32_t *wt_atom, *wt_off, *wt_cfg; // wave tiles: atoms [wt_off[w], wt_off[w+1])
  const int32_t *rng_type, *rng_wt0;        // range r: type, tiles [rng_wt0[r], rng_wt0[r+1])
  const int32_t *cfg_off, *cfg_wt;          // a configuration's tiles
};

__device__ inline int ml_first(int v) { return __builtin_amdgcn_readfirstlane(v); }

// lane q's value in every lane
__device__ inline int ml_lane(int v, int q) { return __builtin_amdgcn_readlane(v, q); }
__device__ inline double ml_lane(double v, int q) {
  const uint64_t u = __builtin_bit_cast(uint64_t, v);
  const uint32_t lo = uint32_t(__builtin_amdgcn_readlane(int(uint32_t(u)), q));
  const uint32_t hi = uint32_t(__builtin_amdgcn_readlane(int(uint32_t(u >> 32)), q));
  return __builtin_bit_cast(double, (uint64_t(hi) << 32) | lo);
}

// the cached site on the shared grid (site64's interval and basis); a linear
// site as the basis (1 - w, w, 0, 0) and derivative (-1/dx, 1/dx, 0, 0)
__device__ inline int ml_site(const MlArgs &p, const double *x, double r, double bv[4], double bd[4]) {
  const int n = p.n;
  int i;
  bool lin;
  if (r <= x[0]) {
    i = 0;
    lin = true;
  } else if (r >= x[n - 1]) {
    i = n - 2;
    lin = true;
  } else {
    i = spl_locate(x, n, r);
    lin = p.linear != 0;
  }
  const double dx = x[i + 1] - x[i];
  if (lin) {
    const double inv_dx = 1.0 / dx;
    const double w = (r - x[i]) * inv_dx;
    bv[0] = 1.0 - w;
    bv[1] = w;
    bv[2] = bv[3] = 0.0;
    bd[0] = -inv_dx;
    bd[1] = inv_dx;
    bd[2] = bd[3] = 0.0;
  } else {
    herm_basis64(x[i], dx, r, bv, bd);
  }
  return i;
}

// basis . (y_k, y_k+1, s_k, s_k+1) of the member rows at r (knot k) and r + 64
__device__ inline double ml_dot(const double2 *r, double b0, double b1, double b2, double b3) {
  const double2 u = r[0], w = r[64];
  return b0 * u.x + b1 * w.x + b2 * u.y + b3 * w.y;
}

// a member's embedding knots in LDS: row k at Y[64 k]
struct MlKnots {
  const double2 *Y;
  __device__ double y(int k) const { return Y[64 * k].x; }
  __device__ double s(int k) const { return Y[64 * k].y; }
};

// stage the member rows of item (g, type t): slots of 2 nt tables of n knots,
// then nemb rows of the embedding table of t (density pass); every wave must
// be past the previous item (the caller's barrier)
__device__ inline void ml_stage(const EvalArgs &a, const MlArgs &p, double2 *sY, int g, int t,
                                int nemb, int wave, int lane) {
  const int nt = a.ntypes, n = p.n, nsl = 2 * nt * n;
  const int eoff = a.tabs[a.npair + nt + t].off;
  for (int row = wave; row < nsl + nemb; row += kMlWaves) {
    int knot;
    if (row < nsl) {
      const int slot = row / n, k = row - slot * n;
      const int tab = slot < nt ? ord_pair(t, slot, nt) : a.npair + slot - nt;
      knot = a.tabs[tab].off + k;
    } else {
      knot = eoff + row - nsl;
    }
    sY[row * 64 + lane] = p.ys[(size_t(g) * p.nknots + knot) * 64 + lane];
  }
}

__global__ __launch_bounds__(1024) void k_pop_ml_density(EvalArgs a, MlArgs p) {
  extern __shared__ double2 sY[];
  __shared__ int s_next;
  // abscissae first (the shared grid, then the item's embedding table), then
  // the member rows
  double *sX = reinterpret_cast<double *>(sY);
  double2 *sR = sY + p.xrows;
  const int wave = ml_first(int(threadIdx.x >> 6)), lane = threadIdx.x & 63;
  const int nt = a.ntypes, n = p.n;
  const int nitem = p.ng * p.nrange;
  for (int item = blockIdx.x; item < nitem; item += gridDim.x) {
    const int g = item / p.nrange, rg = item - g * p.nrange;
    const int t = p.rng_type[rg];
    const TableDesc Te = a.tabs[a.npair + nt + t];
    __syncthreads(); // every wave is done with the previous item's rows
    ml_stage(a, p, sR, g, t, Te.n, wave, lane);
    for (int k = threadIdx.x; k < n + Te.n; k += blockDim.x)
      sX[k] = k < n ? p.x[k] : a.tx[Te.off + k - n];
    if (threadIdx.x == 0) s_next = p.rng_wt0[rg];
    __syncthreads();
    const MlKnots emb{sR + 2 * nt * n * 64 + lane};
    const double *xe = sX + n;
    const int w_end = p.rng_wt0[rg + 1];
    for (;;) {
      int w = 0;
      if (lane == 0) w = atomicAdd(&s_next, 1);
      w = ml_first(w);
      if (w >= w_end) break;
      double e_acc = 0.0, l_acc = 0.0;
      const int a1 = p.wt_off[w + 1];
      for (int ai = p.wt_off[w]; ai < a1; ++ai) {
        const int64_t i = p.wt_atom[ai];
        const int64_t p0 = a.pair_off[i], p1 = a.pair_off[i + 1];
        double rho = 0.0, ep = 0.0;
        for (int64_t pc = p0; pc < p1; pc += 64) {
          const int np = int(min<int64_t>(64, p1 - pc));
          // sites, lanes as pairs: meta = k | tj << 16 | pair-energy bit 24 |
          // density bit 25 (0 for a pair with no table role)
          int meta = 0;
          double b1 = 0.0, b2 = 0.0, b3 = 0.0;
          if (lane < np) {
            const int64_t q = pc + lane;
            const uint32_t b = a.bits[q];
            if (b & kFlagGj) {
              double bv[4], bd[4];
              const int k = ml_site(p, sX, a.r64[q], bv, bd);
              const bool phi = (b & kFlagPhi) && !(b & kFlagTiny);
              meta = k | (pj_type(b) << 16) | (phi ? (1 << 24) : 0) | (1 << 25);
              b1 = bv[1];
              b2 = bv[2];
              b3 = bv[3];
            }
          }
          // one pair ahead: pair q + 1's member rows are read while pair q is
          // summed (two register sets, the walk unrolled by two)
          auto rows = [&](int q, double2 (&R)[4]) {
            const int m = ml_lane(meta, q);
            const int k = m & 0xffff, tj = (m >> 16) & 0xff;
            const double2 *rj = sR + ((nt + tj) * n + k) * 64 + lane;
            const double2 *rp = sR + (tj * n + k) * 64 + lane;
            R[0] = rj[0];
            R[1] = rj[64];
            R[2] = rp[0];
            R[3] = rp[64];
          };
          // the value basis sums to one on the knot values (Hermite and
          // linear sites alike): y_k + b1 (y_k+1 - y_k) + b2 s_k + b3 s_k+1;
          // a pair with no table role (bit 25 clear: off the span, or a
          // padding lane) adds nothing
          auto sum = [&](int q, const double2 (&R)[4]) {
            const int m = ml_lane(meta, q);
            if (!(m & (1 << 25))) return;
            const double B1 = ml_lane(b1, q), B2 = ml_lane(b2, q), B3 = ml_lane(b3, q);
            rho += R[0].x + B1 * (R[1].x - R[0].x) + B2 * R[0].y + B3 * R[1].y;
            if (m & (1 << 24)) ep += R[2].x + B1 * (R[3].x - R[2].x) + B2 * R[2].y + B3 * R[3].y;
          };
          // (an odd count walks the zero-basis lane np as well: the reads and
          // sums are unconditional; scheduling barriers keep each read set
          // ahead of the sums it does not feed)
          const int npe = (np + 1) & ~1;
          double2 RA[4], RB[4];
          rows(0, RA);
          for (int qq = 0; qq < npe; qq += 2) {
            rows(qq + 1, RB);
            __builtin_amdgcn_sched_barrier(0);
            sum(qq, RA);
            __builtin_amdgcn_sched_barrier(0);
            rows(min(qq + 2, npe - 1), RA);
            __builtin_amdgcn_sched_barrier(0);
            sum(qq + 1, RB);
            __builtin_amdgcn_sched_barrier(0);
          }
        }
        // clamp and F(rho) under the gauge, as the single-member density pass
        const GaugedRho gr = gauge_clamp(a.gauge, t, Te, rho);
        double F, dF;
        spl_uncached64_kv(Te, xe, emb, gr.u, F, dF);
        p.Fp[(size_t(g) * a.N + i) * 64 + lane] = dF;
        e_acc += F + gr.e_shift + 0.5 * ep;
        l_acc += gr.lim;
      }
      double *pe = p.pE + (size_t(g) * p.nwt + w) * kMlE * 64 + lane;
      pe[0] = e_acc;
      pe[64] = l_acc;
    }
  }
}

__global__ __launch_bounds__(1024) void k_pop_ml_force(EvalArgs a, MlArgs p) {
  extern __shared__ double2 sY[];
  __shared__ int s_next;
  // abscissae first (the shared grid, then the item's embedding table), then
  // the member rows
  double *sX = reinterpret_cast<double *>(sY);
  double2 *sR = sY + p.xrows;
  const int wave = ml_first(int(threadIdx.x >> 6)), lane = threadIdx.x & 63;
  const int nt = a.ntypes, n = p.n;
  const int nitem = p.ng * p.nrange;
  for (int item = blockIdx.x; item < nitem; item += gridDim.x) {
    const int g = item / p.nrange, rg = item - g * p.nrange;
    const int t = p.rng_type[rg];
    __syncthreads();
    ml_stage(a, p, sR, g, t, 0, wave, lane);
    for (int k = threadIdx.x; k < n; k += blockDim.x) sX[k] = p.x[k];
    if (threadIdx.x == 0) s_next = p.rng_wt0[rg];
    __syncthreads();
    const double *Fpg = p.Fp + size_t(g) * a.N * 64 + lane;
    const double2 *rowi = sR + (nt + t) * n * 64 + lane; // density table of the centre type
    const int w_end = p.rng_wt0[rg + 1];
    for (;;) {
      int w = 0;
      if (lane == 0) w = atomicAdd(&s_next, 1);
      w = ml_first(w);
      if (w >= w_end) break;
      const int64_t cbase = a.atom_off[p.wt_cfg[w]];
      double r2 = 0.0, v0 = 0.0, v1 = 0.0, v2 = 0.0, v3 = 0.0, v4 = 0.0, v5 = 0.0;
      const int a1 = p.wt_off[w + 1];
      for (int ai = p.wt_off[w]; ai < a1; ++ai) {
        const int64_t i = p.wt_atom[ai];
        const int64_t p0 = a.pair_off[i], p1 = a.pair_off[i + 1];
        const double fpi = Fpg[i * 64];
        double f0 = 0.0, f1 = 0.0, f2 = 0.0;
        for (int64_t pc = p0; pc < p1; pc += 64) {
          const int np = int(min<int64_t>(64, p1 - pc));
          // sites, lanes as pairs: meta = k | tj << 7 | (the partner's index
          // in its configuration) << 10, the derivative basis over r and the
          // displacement. The basis is (b0, -b0, b2, b3): the Hermite and the
          // linear derivative bases both weigh the interval's two knot values
          // by exact negatives. The tables share one span, so the pair,
          // partner-density and centre-density gates agree; a pair outside it
          // (or at r < 1e-14) keeps a zero basis and displacement and adds
          // exact zeros.
          int meta = 0;
          double b0 = 0.0, b2 = 0.0, b3 = 0.0, e0 = 0.0, e1 = 0.0, e2 = 0.0;
          if (lane < np) {
            const int64_t q = pc + lane;
            const uint32_t b = a.bits[q];
            if ((b & kFlagPhi) && !(b & kFlagTiny)) {
              e0 = a.pd[3 * q];
              e1 = a.pd[3 * q + 1];
              e2 = a.pd[3 * q + 2];
              double rr = e0 * e0 + e1 * e1;
              rr = rr + e2 * e2;
              const double r = sqrt(rr);
              double bv[4], bd[4];
              const int k = ml_site(p, sX, r, bv, bd);
              const double inv_r = 1.0 / r;
              meta = k | (pj_type(b) << 7) | int(pj_local(b) << 10);
              b0 = bd[0] * inv_r;
              b2 = bd[2] * inv_r;
              b3 = bd[3] * inv_r;
            }
          }
          // one pair ahead: pair q + 1's member rows and partner F' are read
          // while pair q is summed (two register sets, the walk unrolled by two)
          auto rows = [&](int q, double2 (&R)[6], double &fj) {
            const uint32_t m = uint32_t(ml_lane(meta, q));
            const int k = int(m & 127u), tj = int((m >> 7) & 7u);
            const double2 *rp = sR + (tj * n + k) * 64 + lane;
            const double2 *rj = sR + ((nt + tj) * n + k) * 64 + lane;
            R[0] = rp[0];
            R[1] = rp[64];
            R[2] = rj[0];
            R[3] = rj[64];
            R[4] = rowi[k * 64];
            R[5] = rowi[k * 64 + 64];
            fj = Fpg[(cbase + int64_t(m >> 10)) * 64];
          };
          // the three roles' rows weighted first (pair + F'_i density of j +
          // F'_j density of i), then one basis dot
          auto sum = [&](int q, const double2 (&R)[6], double fpj) {
            const double B0 = ml_lane(b0, q), B2 = ml_lane(b2, q), B3 = ml_lane(b3, q);
            const double y0 = R[0].x + fpi * R[2].x + fpj * R[4].x;
            const double y1 = R[1].x + fpi * R[3].x + fpj * R[5].x;
            const double s0 = R[0].y + fpi * R[2].y + fpj * R[4].y;
            const double s1 = R[1].y + fpi * R[3].y + fpj * R[5].y;
            const double fs = B0 * (y0 - y1) + B2 * s0 + B3 * s1;
            const double D0 = ml_lane(e0, q), D1 = ml_lane(e1, q), D2 = ml_lane(e2, q);
            const double fx = fs * D0, fy = fs * D1, fz = fs * D2;
            f0 += fx;
            f1 += fy;
            f2 += fz;
            v0 += D0 * fx;
            v1 += D1 * fy;
            v2 += D2 * fz;
            v3 += D0 * fy;
            v4 += D0 * fz;
            v5 += D1 * fz;
          };
          // (an odd count walks the zero-basis lane np as well: the reads and
          // sums are unconditional; scheduling barriers keep each read set
          // ahead of the sums it does not feed)
          const int npe = (np + 1) & ~1;
          double2 RA[6], RB[6];
          double fa = 0.0, fb = 0.0;
          rows(0, RA, fa);
          for (int qq = 0; qq < npe; qq += 2) {
            rows(qq + 1, RB, fb);
            __builtin_amdgcn_sched_barrier(0);
            sum(qq, RA, fa);
            __builtin_amdgcn_sched_barrier(0);
            rows(min(qq + 2, npe - 1), RA, fa);
            __builtin_amdgcn_sched_barrier(0);
            sum(qq + 1, RB, fb);
            __builtin_amdgcn_sched_barrier(0);
          }
        }
        const double x0 = f0 - a.ref_force[3 * i], x1 = f1 - a.ref_force[3 * i + 1],
                     x2 = f2 - a.ref_force[3 * i + 2];
        r2 += x0 * x0 + x1 * x1 + x2 * x2;
      }
      // the virial rows take -1/2 d_p f_q per pair
      double *pf = p.pF + (size_t(g) * p.nwt + w) * kMlF * 64 + lane;
      pf[0] = r2;
      pf[64] = -0.5 * v0;
      pf[128] = -0.5 * v1;
      pf[192] = -0.5 * v2;
      pf[256] = -0.5 * v3;
      pf[320] = -0.5 * v4;
      pf[384] = -0.5 * v5;
    }
  }
}

// members into groups of 64, member-minor: ys[g][k][lane] = (y, s) of member
// 64 g + lane at knot k (zeros past the batch)
__global__ void k_pop_ml_pack(int nb, int nknots, const double *__restrict__ y,
                              const double *__restrict__ s, double2 *__restrict__ ys) {
  const int g = blockIdx.y, k = blockIdx.x * 4 + int(threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (k >= nknots) return;
  const int m = g * 64 + lane;
  double2 v{0.0, 0.0};
  if (m < nb) v = double2{y[size_t(m) * nknots + k], s[size_t(m) * nknots + k]};
  ys[(size_t(g) * nknots + k) * 64 + lane] = v;
}

// cost of every (member, configuration) from its wave tiles' partials, in the
// configuration's tile order; the rows as k_pop_finalize. Block (c, g), lane =
// member of group g.
__global__ void k_pop_ml_finalize(MlArgs p, int C, const double *__restrict__ vol,
                                  const double *__restrict__ ref_energy,
                                  const double *__restrict__ ref_stress, double ew, double sw,
                                  double *__restrict__ cost_mc) {
  const int c = blockIdx.x, g = blockIdx.y, lane = threadIdx.x;
  const int m = g * 64 + lane;
  if (m >= p.nb) return;
  double e = 0.0, lim = 0.0, r2 = 0.0, vir[6] = {0, 0, 0, 0, 0, 0};
  for (int q = p.cfg_off[c]; q < p.cfg_off[c + 1]; ++q) {
    const size_t w = size_t(p.cfg_wt[q]);
    const double *pe = p.pE + (size_t(g) * p.nwt + w) * kMlE * 64 + lane;
    const double *pf = p.pF + (size_t(g) * p.nwt + w) * kMlF * 64 + lane;
    e += pe[0];
    lim += pe[64];
    r2 += pf[0];
    for (int k = 0; k < 6; ++k) vir[k] += pf[64 * (k + 1)];
  }
  const double er = ew * (e - ref_energy[c]);
  double cost = r2 + er * er + lim * lim;
  if (sw > 0.0) {
    for (int v = 0; v < 6; ++v) {
      const double sr = sw * (vir[v] / vol[c] - ref_stress[6 * c + v]);
      cost += sr * sr;
    }
  }
  cost_mc[size_t(m) * C + c] = cost;
}

EvalArgs make_args(Context &c) {
  EvalArgs a;
  a.N = c.N;
  a.ntiles = c.ntiles;
  a.tile_cfg = c.tile_cfg.p;
  a.tile_a0 = c.tile_a0.p;
  a.tile_na = c.tile_na.p;
  a.atom_off = c.atom_off.p;
  a.pair_off = c.pair_off.p;
  a.row_off = c.row_off.p;
  a.types = c.types.p;
  a.rec = c.pair_rec.p;
  a.bits = c.pair_bits.p;
  a.pd = c.pair_d.p;
  a.r64 = c.pair_r64.p;
  a.tabs = c.tabs.p;
  a.ntab = c.ntab;
  a.nivl = c.nivl;
  a.stage_ivl = stage_intervals(c) ? 1 : 0;
  a.ivl32 = c.ivl32.p;
  a.ivl64 = c.ivl64.p;
  a.tx = c.tab_x.p;
  a.ty = c.tab_y.p;
  a.ts = c.tab_s.p;
  a.ntypes = c.ntypes;
  a.npair = c.npair;
  a.ref_force = c.ref_force.p;
  a.rho = c.rho.p;
  a.over = c.clamp_over.p;
  a.Fp = c.Fp.p;
  a.Fp32 = c.Fp32.p;
  a.e_atom = c.e_atom.p;
  a.lim = c.lim_atom.p;
  a.mu = c.mu.p;
  a.lamb = c.lamb.p;
  a.mu32 = c.mu32.p;
  a.lamb32 = c.lamb32.p;
  a.forces = c.forces.p;
  a.residual = c.residual.p;
  a.tile_part = c.tile_part.p;
  a.gauge = c.gauge.p;
  return a;
}

template <int KIND, bool F64, int LPA> void run_eval(Context &c, const EvalArgs &a, size_t lds,
                                                      int grid) {
  if (a.stage_ivl) {
    hipLaunchKernelGGL((k_density<KIND, F64, LPA, true>), dim3(grid), dim3(256), lds, c.stream, a);
    FS_HIP(hipGetLastError());
    hipLaunchKernelGGL((k_force<KIND, F64, LPA, true>), dim3(grid), dim3(256), lds, c.stream, a);
  } else {
    hipLaunchKernelGGL((k_density<KIND, F64, LPA, false>), dim3(grid), dim3(256), lds, c.stream, a);
    FS_HIP(hipGetLastError());
    hipLaunchKernelGGL((k_force<KIND, F64, LPA, false>), dim3(grid), dim3(256), lds, c.stream, a);
  }
  FS_HIP(hipGetLastError());
}

template <int KIND, bool F64> void dispatch_lpa(Context &c, const EvalArgs &a, size_t lds, int grid) {
  switch (c.lanes_per_atom) {
  case 4: run_eval<KIND, F64, 4>(c, a, lds, grid); break;
  case 8: run_eval<KIND, F64, 8>(c, a, lds, grid); break;
  case 16: run_eval<KIND, F64, 16>(c, a, lds, grid); break;
  case 32: run_eval<KIND, F64, 32>(c, a, lds, grid); break;
  default: run_eval<KIND, F64, 64>(c, a, lds, grid); break;
  }
}

} // namespace

void launch_scatter_params(Context &c, const double *d_x) {
  if (c.D > 0)
    hipLaunchKernelGGL(k_scatter, dim3((c.D + 255) / 256), dim3(256), 0, c.stream, c.D, d_x,
                       c.knot_of.p, c.tab_y.p);
  FS_HIP(hipGetLastError());
}

// the slope kernels stage the knots' abscissae and values while they fit the
// staging budget beside the descriptors
static bool tables_st(const Context &c) {
  return size_t(c.ntab) * sizeof(TableDesc) + 2 * sizeof(double) * size_t(c.nknots) <= kTableStageBytes;
}
static size_t tables_lds_bytes(const Context &c, bool st) {
  return size_t(c.ntab) * sizeof(TableDesc) + (st ? 2 * sizeof(double) * size_t(c.nknots) : 0);
}

void launch_scatter_slopes(Context &c, const double *d_x) {
  if (c.kind == FS_SW || c.ntab == 0) return;
  const bool st = tables_st(c);
  hipLaunchKernelGGL(st ? k_scatter_slopes<true> : k_scatter_slopes<false>, dim3(1), dim3(256),
                     tables_lds_bytes(c, st), c.stream, c.D, d_x, c.knot_of.p, c.ntab, c.nknots,
                     c.tabs.p, c.tab_x.p, c.tab_y.p, c.tab_s.p, c.tab_S.p, c.ivl32.p, c.ivl64.p);
  FS_HIP(hipGetLastError());
}

void launch_slopes(Context &c) {
  if (c.kind == FS_SW || c.ntab == 0) return;
  const bool st = tables_st(c);
  hipLaunchKernelGGL(st ? k_slopes<true> : k_slopes<false>, dim3(1), dim3(256),
                     tables_lds_bytes(c, st), c.stream, c.ntab, c.nknots, c.tabs.p, c.tab_x.p,
                     c.tab_y.p, c.tab_s.p, c.tab_S.p, c.ivl32.p, c.ivl64.p);
  FS_HIP(hipGetLastError());
}

void launch_eval(Context &c, bool) {
  if (c.kind == FS_SW) {
    launch_sw_eval(c);
  } else {
    const EvalArgs a = make_args(c);
    const size_t lds = size_t(c.ntab) * sizeof(TableDesc) +
                       (a.stage_ivl ? 2 * size_t(c.nivl) * sizeof(float4) +
                                          4 * size_t(c.nivl) * sizeof(double)
                                    : 0);
    const int grid = std::max(1, std::min(c.ntiles, c.blocks_per_cu * c.cu_count));
    if (c.ntiles > 0) {
      const bool f64 = c.precision == FS_FP64;
      if (c.kind == FS_EAM) {
        if (f64) dispatch_lpa<FS_EAM, true>(c, a, lds, grid);
        else dispatch_lpa<FS_EAM, false>(c, a, lds, grid);
      } else {
        if (f64) dispatch_lpa<FS_ADP, true>(c, a, lds, grid);
        else dispatch_lpa<FS_ADP, false>(c, a, lds, grid);
      }
    }
  }
  hipLaunchKernelGGL(k_finalize, dim3(std::max(c.C, 1)), dim3(64), 0, c.stream, c.C,
                     c.cfg_tile0.p, c.tile_part.p, c.vol.p, c.row_off.p, c.atom_off.p,
                     c.ref_energy.p, c.ref_stress.p, c.energy_weight, c.stress_weight,
                     c.residual.p, c.cfg_E.p, c.cfg_S.p, c.cfg_L.p);
  FS_HIP(hipGetLastError());
  launch_smooth_residual(c);
}

void launch_smooth_residual(Context &c) {
  if (c.n_smooth == 0) return;
  hipLaunchKernelGGL(k_smooth_res, dim3(unsigned((c.n_smooth + 127) / 128)), dim3(128), 0, c.stream,
                     c.n_smooth, c.smooth_rows.p, c.tab_y.p, c.smooth_weight,
                     c.residual.p + c.R_cfg);
  FS_HIP(hipGetLastError());
}

bool population_supported(const Context &c) {
  return c.kind == FS_EAM && c.ntiles > 0 && c.D > 0;
}

// LDS entries (double2) of the abscissae ahead of the member rows: the shared
// grid's n and an embedding table's up to nemb, in whole 1 KB rows
static size_t ml_xrows(int n, int nemb) { return (size_t(n + nemb + 127) / 128) * 64; }

// LDS of the member-lane density pass (the force pass stages its first
// 2 nt n rows), or 0 where those passes do not apply: they need an EAM model
// whose pair and density tables share one grid (the same abscissae, so the
// same span and interval for every table role of a pair), a neighbour list,
// and a block's member rows within the device's LDS
static size_t ml_lds_bytes(const Context &c) {
  if (c.kind != FS_EAM || c.ntab == 0 || c.P == 0 || c.ntypes > kMaxTypes) return 0;
  const int nt = c.ntypes, nsh = c.npair + nt;
  const TableDesc &T0 = c.h_tabs[0];
  if (T0.n < 2 || T0.n > 128) return 0; // the force pass packs k in 7 bits
  for (int t = 1; t < nsh; ++t) {
    const TableDesc &T = c.h_tabs[t];
    if (T.n != T0.n || T.linear != T0.linear) return 0;
    for (int k = 0; k < T.n; ++k)
      if (c.h_x[size_t(T.off) + k] != c.h_x[size_t(T0.off) + k]) return 0;
  }
  int nemb = 0;
  for (int t = 0; t < nt; ++t) nemb = std::max(nemb, c.h_tabs[c.npair + nt + t].n);
  const size_t lds = (ml_xrows(T0.n, nemb) + (size_t(2 * nt) * T0.n + nemb) * 64) * sizeof(double2);
  int lmax = 0;
  FS_HIP(hipDeviceGetAttribute(&lmax, hipDeviceAttributeMaxSharedMemoryPerBlock, c.device));
  return lds + 64 <= size_t(lmax) ? lds : 0;
}

// members per batch: per member 12 N bytes of embedding derivatives, the
// per-(config, group) partials and its tables; about 4 GiB per batch. The
// member-lane passes hold 8 N bytes of embedding derivatives and 72 bytes per
// wave tile per member, in whole groups of 64, up to 16 GiB per batch.
int population_batch_cap(const Context &c, int np) {
  const size_t tabs = size_t(c.nknots) * 8 * 7 + size_t(c.nivl) * 64 + size_t(c.C) * 8 + 8 * c.D;
  if (c.pop_ml) {
    const size_t per = 8 * size_t(c.N) + size_t(c.pop_ml_nwt) * 8 * (kMlE + kMlF) +
                       size_t(c.nknots) * 16 + tabs;
    size_t cap = std::max<size_t>(64, ((size_t(16) << 30) / per) & ~size_t(63));
    return int(std::min<size_t>(size_t(std::max(np, 1)), cap));
  }
  const size_t per = 12 * size_t(c.N) + size_t(c.C) * kPopG * 8 * (kPopE + kPopF) + tabs;
  return int(std::max<size_t>(1, std::min<size_t>(size_t(std::max(np, 1)), (size_t(4) << 30) / per)));
}

// wave tiles of the member-lane passes (once per neighbour list and dataset):
// the atoms of each type, per configuration in Morton order, split into tiles
// of up to kMlTile; per type, ranges of tiles with about equal pair counts (about
// kMlRanges in all); per configuration, its tiles in type order
static void ml_layout(Context &c) {
  const uint64_t key[4] = {c.nl_gen, uint64_t(c.N), uint64_t(c.ntypes), uint64_t(c.C)};
  if (std::equal(key, key + 4, c.pop_ml_key)) return;
  const int64_t N = c.N;
  const int nt = c.ntypes;
  std::vector<int64_t> off(size_t(N) + 1);
  FS_HIP(hipMemcpyAsync(off.data(), c.pair_off.p, sizeof(int64_t) * (N + 1), hipMemcpyDeviceToHost,
                        c.stream));
  FS_HIP(hipStreamSynchronize(c.stream));
  std::vector<int32_t> wt_atom, wt_off{0}, wt_cfg, rng_type, rng_wt0{0};
  std::vector<int64_t> wt_pairs;
  std::vector<std::vector<int32_t>> per_cfg(c.C);
  std::vector<std::pair<uint64_t, int32_t>> sel;
  const int64_t P = std::max<int64_t>(off[N], 1);
  for (int t = 0; t < nt; ++t) {
    const size_t w0 = wt_cfg.size();
    int64_t tpairs = 0;
    for (int k = 0; k < c.C; ++k) {
      int cnt = 0;
      int64_t pairs = 0;
      auto close = [&]() {
        wt_off.push_back(int32_t(wt_atom.size()));
        per_cfg[k].push_back(int32_t(wt_cfg.size()));
        wt_cfg.push_back(k);
        wt_pairs.push_back(pairs);
        tpairs += pairs;
        cnt = 0;
        pairs = 0;
      };
      // the configuration's atoms of type t in Morton order of their cell
      // (fractional) coordinates: a tile is then a compact cluster, whose
      // pairs share most partners, so the partners' F' rows are re-read from
      // cache
      sel.clear();
      const double *I = c.h_inv.data() + 9 * size_t(k);
      for (int64_t i = c.h_atom_off[k]; i < c.h_atom_off[k + 1]; ++i) {
        if (c.h_types[size_t(i)] != t) continue;
        const double *r = c.h_pos.data() + 3 * size_t(i);
        uint64_t code = 0;
        for (int d = 0; d < 3; ++d) {
          double v = I[3 * d] * r[0] + I[3 * d + 1] * r[1];
          v = v + I[3 * d + 2] * r[2];
          v -= std::floor(v);
          const uint64_t q = std::min<uint64_t>(1023, uint64_t(std::max(0.0, v) * 1024.0));
          for (int b = 0; b < 10; ++b) code |= ((q >> b) & 1u) << (3 * b + d);
        }
        sel.push_back({code, int32_t(i)});
      }
      std::sort(sel.begin(), sel.end());
      for (const auto &e : sel) {
        const int32_t i = e.second;
        wt_atom.push_back(i);
        pairs += off[size_t(i) + 1] - off[size_t(i)];
        if (++cnt == kMlTile) close();
      }
      if (cnt > 0) close();
    }
    const size_t w1 = wt_cfg.size();
    if (w1 == w0) continue;
    const int nr = int(std::max<int64_t>(1, std::min<int64_t>(int64_t(w1 - w0),
                                                              (kMlRanges * tpairs + P - 1) / P)));
    int64_t acc = 0;
    int r = 1;
    for (size_t w = w0; w < w1; ++w) {
      acc += wt_pairs[w];
      if (r < nr && acc * nr >= tpairs * r && w + 1 < w1) {
        rng_type.push_back(t);
        rng_wt0.push_back(int32_t(w + 1));
        ++r;
      }
    }
    rng_type.push_back(t);
    rng_wt0.push_back(int32_t(w1));
  }
  std::vector<int32_t> cfg_off{0}, cfg_wt;
  for (int k = 0; k < c.C; ++k) {
    cfg_wt.insert(cfg_wt.end(), per_cfg[k].begin(), per_cfg[k].end());
    cfg_off.push_back(int32_t(cfg_wt.size()));
  }
  auto up = [&](DBuf<int32_t> &d, const std::vector<int32_t> &h) {
    d.alloc(std::max<size_t>(h.size(), 1));
    if (!h.empty())
      FS_HIP(hipMemcpyAsync(d.p, h.data(), 4 * h.size(), hipMemcpyHostToDevice, c.stream));
  };
  up(c.pop_ml_wt_atom, wt_atom);
  up(c.pop_ml_wt_off, wt_off);
  up(c.pop_ml_wt_cfg, wt_cfg);
  up(c.pop_ml_rng_type, rng_type);
  up(c.pop_ml_rng_wt0, rng_wt0);
  up(c.pop_ml_cfg_off, cfg_off);
  up(c.pop_ml_cfg_wt, cfg_wt);
  FS_HIP(hipStreamSynchronize(c.stream));
  c.pop_ml_nwt = int(wt_cfg.size());
  c.pop_ml_nrange = int(rng_type.size());
  std::copy(key, key + 4, c.pop_ml_key);
}

// every allocation of a population evaluation of np members; no launch
void population_prepare(Context &c, int np) {
  const int G = kPopG;
  if (c.h_pop_grp.size() != size_t(G + 1) || c.pop_grp_tiles != c.ntiles) {
    c.h_pop_grp.assign(G + 1, 0);
    for (int g = 0; g <= G; ++g) c.h_pop_grp[g] = int32_t(int64_t(g) * c.ntiles / G);
    c.pop_grp.alloc(G + 1);
    FS_HIP(hipMemcpyAsync(c.pop_grp.p, c.h_pop_grp.data(), 4 * (G + 1), hipMemcpyHostToDevice,
                          c.stream));
    FS_HIP(hipStreamSynchronize(c.stream));
    c.pop_grp_tiles = c.ntiles;
  }
  c.pop_ml_lds = c.pop_ml_allowed ? ml_lds_bytes(c) : 0;
  c.pop_ml = c.pop_ml_lds > 0;
  if (c.pop_ml) {
    ml_layout(c);
    const TableDesc &T0 = c.h_tabs[0];
    c.pop_ml_n = T0.n;
    c.pop_ml_xoff = T0.off;
  }
  const int cap = population_batch_cap(c, np);
  c.pop_y.alloc(size_t(cap) * c.nknots);
  c.pop_s.alloc(size_t(cap) * c.nknots);
  c.pop_S.alloc(size_t(cap) * 5 * c.nknots);
  c.pop_i32.alloc(size_t(cap) * 2 * c.nivl);
  c.pop_i64.alloc(size_t(cap) * 4 * c.nivl);
  c.pop_cmc.alloc(size_t(cap) * c.C);
  if (c.pop_ml) {
    const size_t ng = (size_t(cap) + 63) / 64;
    c.pop_ml_ng = int(ng);
    c.pop_ml_ys.alloc(ng * c.nknots * 64);
    c.pop_ml_Fp.alloc(ng * size_t(c.N) * 64);
    c.pop_ml_pE.alloc(ng * size_t(c.pop_ml_nwt) * kMlE * 64);
    c.pop_ml_pF.alloc(ng * size_t(c.pop_ml_nwt) * kMlF * 64);
  } else {
    c.pop_Fp.alloc(size_t(cap) * c.N);
    c.pop_Fp32.alloc(size_t(cap) * c.N);
    c.pop_partE.alloc(size_t(cap) * c.C * G * kPopE);
    c.pop_partF.alloc(size_t(cap) * c.C * G * kPopF);
  }
}

// the member-lane passes over np members (population_prepare has chosen them)
static void population_ml_launch(Context &c, const double *d_X, int np, double *d_cost) {
  const int cap = population_batch_cap(c, np);
  EvalArgs a = make_args(c);
  const size_t lds_d = c.pop_ml_lds; // measured by population_prepare
  int nemb = 0;
  for (int t = 0; t < c.ntypes; ++t) nemb = std::max(nemb, c.h_tabs[c.npair + c.ntypes + t].n);
  const size_t xrows = ml_xrows(c.pop_ml_n, nemb);
  const size_t lds_f = (xrows + size_t(2 * c.ntypes) * c.pop_ml_n * 64) * sizeof(double2);
  for (int m0 = 0; m0 < np; m0 += cap) {
    const int nb = std::min(cap, np - m0);
    const int ng = (nb + 63) / 64;
    hipLaunchKernelGGL(tables_st(c) ? k_pop_tables<true> : k_pop_tables<false>, dim3(nb), dim3(64),
                       tables_lds_bytes(c, tables_st(c)), c.stream, c.ntab, c.D, c.nknots,
                       c.tabs.p, c.tab_x.p, c.tab_y.p, c.knot_of.p, d_X + size_t(m0) * c.D, c.pop_y.p,
                       c.pop_s.p, c.pop_S.p, c.pop_i32.p, c.pop_i64.p, c.nivl);
    FS_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_pop_ml_pack, dim3(unsigned((c.nknots + 3) / 4), unsigned(ng)), dim3(256), 0,
                       c.stream, nb, c.nknots, c.pop_y.p, c.pop_s.p, c.pop_ml_ys.p);
    FS_HIP(hipGetLastError());
    MlArgs p{};
    p.ng = ng;
    p.nb = nb;
    p.n = c.pop_ml_n;
    p.linear = c.h_tabs[0].linear;
    p.nknots = c.nknots;
    p.nwt = c.pop_ml_nwt;
    p.nrange = c.pop_ml_nrange;
    p.x = c.tab_x.p + c.pop_ml_xoff;
    p.xrows = int(xrows);
    p.ys = c.pop_ml_ys.p;
    p.Fp = c.pop_ml_Fp.p;
    p.pE = c.pop_ml_pE.p;
    p.pF = c.pop_ml_pF.p;
    p.wt_atom = c.pop_ml_wt_atom.p;
    p.wt_off = c.pop_ml_wt_off.p;
    p.wt_cfg = c.pop_ml_wt_cfg.p;
    p.rng_type = c.pop_ml_rng_type.p;
    p.rng_wt0 = c.pop_ml_rng_wt0.p;
    p.cfg_off = c.pop_ml_cfg_off.p;
    p.cfg_wt = c.pop_ml_cfg_wt.p;
    const int grid = std::max(1, std::min(c.cu_count, ng * c.pop_ml_nrange));
    hipLaunchKernelGGL(k_pop_ml_density, dim3(grid), dim3(64 * kMlWaves), lds_d, c.stream, a, p);
    FS_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_pop_ml_force, dim3(grid), dim3(64 * kMlWaves), lds_f, c.stream, a, p);
    FS_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_pop_ml_finalize, dim3(unsigned(c.C), unsigned(ng)), dim3(64), 0, c.stream, p,
                       c.C, c.vol.p, c.ref_energy.p, c.ref_stress.p, c.energy_weight, c.stress_weight,
                       c.pop_cmc.p);
    FS_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_pop_sum, dim3(unsigned((nb + 127) / 128)), dim3(128), 0, c.stream, nb, c.C,
                       c.pop_cmc.p, c.n_smooth, c.smooth_rows.p, c.smooth_weight, c.pop_y.p,
                       c.nknots, d_cost + m0);
    FS_HIP(hipGetLastError());
  }
}

// |F(x_m)|^2 of the np device vectors d_X[np][D] into d_cost[np]: launches
// only (no allocation, copy or synchronisation, so a stream capture can
// record it); population_prepare(c, np) must have run
template <int M> void population_launch_m(Context &c, const double *d_X, int np, double *d_cost);

// members per block: c.pop_members (1, 2 or 4; the DE driver times them on its
// first population)
void population_launch(Context &c, const double *d_X, int np, double *d_cost) {
  if (c.pop_ml) {
    population_ml_launch(c, d_X, np, d_cost);
    return;
  }
  switch (c.pop_members) {
  case 2: population_launch_m<2>(c, d_X, np, d_cost); break;
  case 4: population_launch_m<4>(c, d_X, np, d_cost); break;
  default: population_launch_m<1>(c, d_X, np, d_cost); break;
  }
}

template <int M> void population_launch_m(Context &c, const double *d_X, int np, double *d_cost) {
  const int G = kPopG;
  const int cap = population_batch_cap(c, np);
  EvalArgs a = make_args(c);
  // the M members' coefficients of a block are staged while they fit the
  // budget: FP64 interval forms (density pass), FP32 hi/lo interval forms or,
  // in FP64 mode, knot values and slopes (force pass)
  const bool f64 = c.precision == FS_FP64;
  const size_t lds_t = size_t(c.ntab) * sizeof(TableDesc);
  const size_t mem_d = size_t(M) * 4 * c.nivl * sizeof(double);
  const size_t mem_f = f64 ? size_t(M) * 2 * c.nknots * sizeof(double)
                           : size_t(M) * 2 * c.nivl * sizeof(float4);
  const bool stage_d = lds_t + mem_d <= kTableStageBytes, stage_f = lds_t + mem_f <= kTableStageBytes;
  const size_t lds_d = lds_t + (stage_d ? mem_d : 0), lds_f = lds_t + (stage_f ? mem_f : 0);
  for (int m0 = 0; m0 < np; m0 += cap) {
    const int nb = std::min(cap, np - m0);
    hipLaunchKernelGGL(tables_st(c) ? k_pop_tables<true> : k_pop_tables<false>, dim3(nb), dim3(64),
                       tables_lds_bytes(c, tables_st(c)), c.stream, c.ntab, c.D, c.nknots,
                       c.tabs.p, c.tab_x.p, c.tab_y.p, c.knot_of.p, d_X + size_t(m0) * c.D, c.pop_y.p,
                       c.pop_s.p, c.pop_S.p, c.pop_i32.p, c.pop_i64.p, c.nivl);
    FS_HIP(hipGetLastError());
    PopArgs p{};
    p.nb = nb;
    p.nknots = c.nknots;
    p.nivl = c.nivl;
    p.stage_ivl = stage_d ? 1 : 0;
    p.y = c.pop_y.p;
    p.s = c.pop_s.p;
    p.ivl32 = c.pop_i32.p;
    p.ivl64 = c.pop_i64.p;
    p.Fp = c.pop_Fp.p;
    p.Fp32 = c.pop_Fp32.p;
    p.grp_t0 = c.pop_grp.p;
    p.partE = c.pop_partE.p;
    p.partF = c.pop_partF.p;
    p.C = c.C;
    PopArgs pf = p;
    pf.stage_ivl = stage_f ? 1 : 0;
    const dim3 grid(unsigned((nb + M - 1) / M) * G);
    switch (c.lanes_per_atom) {
#define FS_POP(L)                                                                                   \
  case L:                                                                                           \
    hipLaunchKernelGGL((k_pop_density<L, M>), grid, dim3(256), lds_d, c.stream, a, p);              \
    FS_HIP(hipGetLastError());                                                                      \
    if (f64)                                                                                        \
      hipLaunchKernelGGL((k_pop_force<L, true, M>), grid, dim3(256), lds_f, c.stream, a, pf);       \
    else                                                                                            \
      hipLaunchKernelGGL((k_pop_force<L, false, M>), grid, dim3(256), lds_f, c.stream, a, pf);      \
    break;
      FS_POP(4)
      FS_POP(8)
      FS_POP(16)
      FS_POP(32)
      FS_POP(64)
#undef FS_POP
    default: throw Error(FS_ERR_STATE, "unsupported lanes per atom");
    }
    FS_HIP(hipGetLastError());
    const int64_t nmc = int64_t(nb) * c.C;
    hipLaunchKernelGGL(k_pop_finalize, dim3(unsigned((nmc + 127) / 128)), dim3(128), 0, c.stream, nb,
                       c.C, c.pop_grp.p, c.cfg_tile0.p, c.pop_partE.p, c.pop_partF.p, c.vol.p,
                       c.ref_energy.p, c.ref_stress.p, c.energy_weight, c.stress_weight,
                       c.pop_cmc.p);
    FS_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_pop_sum, dim3(unsigned((nb + 127) / 128)), dim3(128), 0, c.stream, nb, c.C,
                       c.pop_cmc.p, c.n_smooth, c.smooth_rows.p, c.smooth_weight, c.pop_y.p,
                       c.nknots, d_cost + m0);
    FS_HIP(hipGetLastError());
  }
}

// the population cost of np members with c.pop_members picked by timing 1, 2
// and 4 members per block on this population (once per context; every
// variant gives the same costs: a member's sums do not depend on the others
// in its block); returns the time of the chosen variant in ms
double population_tune_launch(Context &c, const double *d_X, int np, double *d_cost) {
  if (c.pop_members != 0 || c.pop_ml) {
    population_launch(c, d_X, np, d_cost);
    return 0.0;
  }
  hipEvent_t e0, e1;
  FS_HIP(hipEventCreate(&e0));
  FS_HIP(hipEventCreate(&e1));
  double best = 1e300;
  int pick = 1;
  try {
    for (int m : {1, 2, 4}) {
      c.pop_members = m;
      FS_HIP(hipEventRecord(e0, c.stream));
      population_launch(c, d_X, np, d_cost);
      FS_HIP(hipEventRecord(e1, c.stream));
      FS_HIP(hipEventSynchronize(e1));
      float ms = 0.f;
      FS_HIP(hipEventElapsedTime(&ms, e0, e1));
      if (ms < best) best = ms, pick = m;
    }
  } catch (...) {
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    c.pop_members = 0;
    throw;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  c.pop_members = pick;
  return best;
}

bool population_cost_batched(Context &c, const double *X, int np, double *cost) {
  if (!population_supported(c) || np <= 0) return false;
  population_prepare(c, np);
  c.pop_X.alloc(size_t(np) * c.D);
  c.pop_cost.alloc(size_t(np));
  FS_HIP(hipMemcpyAsync(c.pop_X.p, X, sizeof(double) * size_t(np) * c.D, hipMemcpyHostToDevice,
                        c.stream));
  population_tune_launch(c, c.pop_X.p, np, c.pop_cost.p);
  FS_HIP(hipMemcpyAsync(cost, c.pop_cost.p, sizeof(double) * np, hipMemcpyDeviceToHost, c.stream));
  FS_HIP(hipStreamSynchronize(c.stream));
  return true;
}

void launch_sumsq(hipStream_t st, int64_t n, const double *v, double *part, double *d_out) {
  hipLaunchKernelGGL(k_sumsq, dim3(kSumsqBlocks), dim3(256), 0, st, n, v, part);
  FS_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(256), 0, st, kSumsqBlocks, part, d_out);
  FS_HIP(hipGetLastError());
}

void launch_cost(Context &c, double *d_out) {
  c.cost_part.alloc(kSumsqBlocks);
  launch_sumsq(c.stream, c.R, c.residual.p, c.cost_part.p, d_out);
}

} // namespace fsa
