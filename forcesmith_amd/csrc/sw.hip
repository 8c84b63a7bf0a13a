// Stillinger-Weber residual path: pair term plus triplets centred on each
// atom, with the reference's scatter of the triplet forces onto j and k
// (include/forcesmith/force/kernels.hpp:58-66) recast as a deterministic
// gather. k_sw_pair_terms evaluates the radial factors once per pair; in
// k_sw_centre a group of kSwLanes lanes takes one centre atom i, and the lane
// of neighbour j sums over every other neighbour k the triplet force on j,
//   T_j = sum_k f_j(i; j, k),
// and writes it into the per-pair slot of (i -> j). A triplet's three forces
// sum to zero (the term depends on the two bond vectors only), so the centre
// force is the pair forces minus sum_j T_j, and the triplet virial
// sum (d_j (x) f_j + d_k (x) f_k) over triplets is sum_j d_j (x) T_j. Every
// atom then gathers the slots of the reverse pairs (k -> i for each of its
// pairs i -> k).
//
// Reference: StiwebForceCalculator::eval_forces (src/force/stiweb_force.cpp:
// 241-271), v2_dv2 / h_dh (:24-54), three_body_forces (kernels.hpp:40-56).
// Everything runs in FP64, in both precision modes. The radial factors (pow,
// exp), the pair terms, the bond's unit vector and 1/r are formed once per
// pair, so the triplet loop has no square root or division: per triplet it
// forms the cosine and the angular factor and adds four scalar sums, from
// which the lane's T_j follows once per bond (the j-constant factors h(r_ij),
// h'(r_ij), 1/r_ij are taken out of the sum over k).
#include "common.hpp"
#include "spline.hpp"

namespace fsa {

namespace {

struct SwP {
  double A, B, p, q, delta, a1, gamma, a2;
};

__device__ inline SwP load_sw(const double *sw, int s) {
  const double *q = sw + 8 * s;
  return SwP{q[0], q[1], q[2], q[3], q[4], q[5], q[6], q[7]};
}

// Radial factors of one pair at distance r (ir = 1 / r), v2_dv2 and h_dh
// (stiweb_force.cpp:24-54): o = {v2, v2', h, h', dh/dgamma, dh'/dgamma,
// dh/da2, dh'/da2}; with TAN also the tangents (tv, tdv / r) of (v2, v2')
// w.r.t. each two-body field f = 0..5 (A, B, p, q, delta, a1) into t. The
// powers come from one logarithm, r^-p = exp(-p ln r) and r^(-p-1) =
// r^-p / r, shared by the values and the tangents.
template <bool TAN>
__device__ inline void sw_radial(double r, double ir, const SwP &P, double *o, double2 *t) {
  for (int k = 0; k < 8; ++k) o[k] = 0.0;
  if (TAN)
    for (int f = 0; f < 6; ++f) t[f] = make_double2(0.0, 0.0);
  if (r < P.a1) {
    const double dd = r - P.a1;
    const double e = exp(P.delta / dd);
    if (e != 0.0) {
      const double lr = log(r);
      const double rp = exp(-P.p * lr), rq = exp(-P.q * lr);
      const double rp1 = rp * ir, rq1 = rq * ir;
      const double poly = P.A * rp - P.B * rq;
      const double dpoly = -P.A * P.p * rp1 + P.B * P.q * rq1;
      const double g = -P.delta / (dd * dd);
      o[0] = poly * e;
      o[1] = dpoly * e + poly * e * g;
      if (TAN) {
        for (int f = 0; f < 6; ++f) {
          double Pf = 0.0, dPf = 0.0, Ef = 0.0, Gt = 0.0;
          switch (f) {
          case 0: Pf = rp; dPf = -P.p * rp1; break;
          case 1: Pf = -rq; dPf = P.q * rq1; break;
          case 2: Pf = -P.A * lr * rp; dPf = P.A * rp1 * (P.p * lr - 1.0); break;
          case 3: Pf = P.B * lr * rq; dPf = P.B * rq1 * (1.0 - P.q * lr); break;
          case 4: Ef = e / dd; Gt = -1.0 / (dd * dd); break;
          default: Ef = e * P.delta / (dd * dd); Gt = -2.0 * P.delta / (dd * dd * dd); break;
          }
          const double tv = Pf * e + poly * Ef;
          const double tdv = dPf * e + dpoly * Ef + Pf * e * g + poly * Ef * g + poly * e * Gt;
          t[f] = make_double2(tv, tdv * ir);
        }
      }
    }
  }
  if (r < P.a2) {
    const double dd = r - P.a2;
    const double e = exp(P.gamma / dd);
    if (e != 0.0) {
      const double g1 = -P.gamma / (dd * dd);
      o[2] = e;
      o[3] = e * g1;
      // tangents of (h, h') w.r.t. gamma and a2 (h = exp(gamma / (r - a2)))
      const double Hg = e / dd, Gg = -1.0 / (dd * dd);
      o[4] = Hg;
      o[5] = Hg * g1 + e * Gg;
      const double Ha = e * P.gamma / (dd * dd), Ga = -2.0 * P.gamma / (dd * dd * dd);
      o[6] = Ha;
      o[7] = Ha * g1 + e * Ga;
    }
  }
}

// reverse pair: for p = (i -> j, image n) the entry (j -> i, image -n)
__global__ void k_sw_rev(int64_t N, const int64_t *__restrict__ pair_off,
                         const int32_t *__restrict__ cfg_of, const int64_t *__restrict__ atom_off,
                         const uint32_t *__restrict__ bits, const int32_t *__restrict__ img,
                         int32_t *__restrict__ rev) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const int64_t a0 = atom_off[cfg_of[i]];
  const uint32_t il = uint32_t(i - a0);
  for (int64_t p = pair_off[i]; p < pair_off[i + 1]; ++p) {
    const int64_t j = a0 + pj_local(bits[p]);
    const int im = img[p];
    const int nx = (im & 1023) - 512, ny = ((im >> 10) & 1023) - 512, nz = ((im >> 20) & 1023) - 512;
    const int want = (-nx + 512) | ((-ny + 512) << 10) | ((-nz + 512) << 20);
    int32_t found = -1;
    for (int64_t q = pair_off[j]; q < pair_off[j + 1]; ++q)
      if (pj_local(bits[q]) == il && img[q] == want) {
        found = int32_t(q);
        break;
      }
    rev[p] = found;
  }
}

constexpr int kSwLanes = 16; // lanes per centre atom (a divisor of 64)
constexpr int kSwStage = 32; // neighbour records per centre staged in LDS

// sum over the kSwLanes lanes of one centre's group (all lanes of the wave
// take part)
__device__ inline double group_sum(double v) {
#pragma unroll
  for (int o = kSwLanes / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// per pair (i -> j) of the pair's parameter slot (ti, tj), zeros for an
// r < 1e-14 pair: pv = {v2, v2', h, h', dh/dgamma, dh'/dgamma, dh/da2,
// dh'/da2}, the triplet record tri = {unit vector, 1/r, h, h', dh/dgamma,
// dh/da2}, and with TAN the two-body tangents pt2, field-major ([6][P]) so a
// Jacobian column's pass reads 16 contiguous bytes per pair
template <bool TAN>
__global__ __launch_bounds__(256) void k_sw_pair_terms(int64_t N, int64_t P,
                                                       const int64_t *__restrict__ pair_off,
                                                       const int32_t *__restrict__ types,
                                                       const double *__restrict__ pd,
                                                       const uint32_t *__restrict__ bits,
                                                       const double *__restrict__ sw, int ntypes,
                                                       double *__restrict__ pv,
                                                       double *__restrict__ tri,
                                                       double2 *__restrict__ pt2) {
  const int64_t g = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t i = g / kSwLanes;
  const int l = int(g % kSwLanes);
  if (i >= N) return;
  const int ti = types[i];
  for (int64_t p = pair_off[i] + l; p < pair_off[i + 1]; p += kSwLanes) {
    double o[8], u[4] = {0, 0, 0, 0};
    double2 t[6];
    const uint32_t b = bits[p];
    if (!(b & kFlagTiny)) {
      const double d0 = pd[3 * p], d1 = pd[3 * p + 1], d2 = pd[3 * p + 2];
      double r2 = d0 * d0 + d1 * d1;
      r2 = r2 + d2 * d2;
      const double r = sqrt(r2);
      u[3] = 1.0 / r;
      u[0] = d0 * u[3];
      u[1] = d1 * u[3];
      u[2] = d2 * u[3];
      sw_radial<TAN>(r, u[3], load_sw(sw, ord_pair(ti, pj_type(b), ntypes)), o, t);
    } else {
      for (int k = 0; k < 8; ++k) o[k] = 0.0;
      for (int f = 0; f < 6; ++f) t[f] = make_double2(0.0, 0.0);
    }
    double *q = pv + 8 * p;
    for (int k = 0; k < 8; ++k) q[k] = o[k];
    double *w = tri + 8 * p;
    for (int k = 0; k < 4; ++k) w[k] = u[k];
    w[4] = o[2];
    w[5] = o[3];
    w[6] = o[4];
    w[7] = o[6];
    if (TAN)
      for (int f = 0; f < 6; ++f) pt2[int64_t(f) * P + p] = t[f];
  }
}

struct SwArgs {
  int64_t N;
  const int64_t *pair_off;
  const int32_t *types;
  const double *pd;
  const uint32_t *bits;
  const int32_t *rev;
  const double *sw, *lam;
  int ntypes, npair;
  const double *pv; // 8 per pair (k_sw_pair_terms)
  const double *tri; // 8 per pair: unit vector, 1/r, h, h', dh/dgamma, dh/da2
  double *pairbuf;  // 3 per pair: force on the neighbour from triplets centred here
  double *fcentre;  // 3 per atom
  double *e_atom;   // per atom energy
  double *sig_atom; // 9 per atom virial
  const int32_t *tile_cfg, *tile_a0, *tile_na;
  int ntiles;
  const int64_t *atom_off, *row_off;
  const double *ref_force;
  double *forces, *residual, *tile_part, *lim_atom;
};

// The force on j of the triplet (i; j, k) (three_body_forces,
// kernels.hpp:40-56), with unit vectors u1, u2 of d_ij, d_ik, cosine c, radial
// factors f1 = h(r_ij), f2 = h(r_ik) and angular factor g = w(c), is
//   f_j = -f1' f2 g u1 - f1 f2 g' (u2 - c u1) / r_ij,
// so T_j = sum_k f_j = -(h1' Sa - h1 Sbc / r_ij) u1 - h1 V / r_ij with
// Sa = sum h2 w, Sbc = sum h2 w' c, V = sum h2 w' u2; the triplet energy seen
// from bond j is h1 Sa / 2 (each triplet is met from both of its bonds).
// The first kSwStage neighbours' records of each centre (unit vector, h and
// type, -1 for a pair that takes no part in triplets) and the lambda table
// are staged in LDS.
// 5 waves per SIMD asked (96 VGPRs, two values in scratch; the compiler's own
// 98 allow 4): residual 0.068 -> 0.065 ms on the 256-configuration Si workload.
// The same ask of the pair-term and Jacobian centre passes spilled more and
// measured slower.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5))) void k_sw_centre(SwArgs a) {
  constexpr int kGroups = 256 / kSwLanes;
  __shared__ double4 s_u[kGroups][kSwStage];
  __shared__ int s_ty[kGroups][kSwStage];
  __shared__ double s_lam[kMaxTypes * kMaxTypes * (kMaxTypes + 1) / 2];
  const int64_t g = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t i = g / kSwLanes;
  const int l = int(g % kSwLanes);
  const int grp = int(threadIdx.x) / kSwLanes;
  const bool live = i < a.N; // every lane takes part in the group sums
  double F[3] = {0, 0, 0}, E = 0.0, sg[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  const int64_t p0 = live ? a.pair_off[i] : 0, p1 = live ? a.pair_off[i + 1] : 0;
  auto record = [&](int64_t pk, double4 &u, int &ty) {
    const uint32_t bk = a.bits[pk];
    const double *r = a.tri + 8 * pk;
    u = make_double4(r[0], r[1], r[2], r[4]);
    ty = (bk & kFlagTiny) || (r[4] == 0.0 && r[5] == 0.0) ? -1 : pj_type(bk);
  };
  for (int q = threadIdx.x; q < a.ntypes * a.npair; q += blockDim.x) s_lam[q] = a.lam[q];
  for (int k = l; k < int(std::min<int64_t>(p1 - p0, kSwStage)); k += kSwLanes)
    record(p0 + k, s_u[grp][k], s_ty[grp][k]);
  __syncthreads();
  if (live) {
    const int ti = a.types[i];
    for (int64_t pj = p0 + l; pj < p1; pj += kSwLanes) {
      double T[3] = {0, 0, 0};
      const uint32_t bj = a.bits[pj];
      if (!(bj & kFlagTiny)) {
        const double *d1 = a.pd + 3 * pj;
        const double *q1 = a.pv + 8 * pj;
        const double *t1 = a.tri + 8 * pj;
        const double v2 = q1[0], dv2 = q1[1], h1 = q1[2], dh1 = q1[3];
        // pair term
        const double s = dv2 * t1[3];
        const double f[3] = {s * d1[0], s * d1[1], s * d1[2]};
        for (int k = 0; k < 3; ++k) F[k] += f[k];
        E += 0.5 * v2;
        for (int x = 0; x < 3; ++x)
          for (int y = 0; y < 3; ++y) sg[3 * x + y] -= 0.5 * d1[x] * f[y];
        // triplets (i; j, k) over every other neighbour k
        if (h1 != 0.0 || dh1 != 0.0) {
          const int tj = pj_type(bj);
          const double u1[3] = {t1[0], t1[1], t1[2]}, i1 = t1[3];
          double Sa = 0.0, Sbc = 0.0, V[3] = {0.0, 0.0, 0.0};
          for (int64_t pk = p0; pk < p1; ++pk) {
            if (pk == pj) continue;
            const int k = int(pk - p0);
            double4 u2;
            int tk;
            if (k < kSwStage) {
              u2 = s_u[grp][k];
              tk = s_ty[grp][k];
            } else {
              record(pk, u2, tk);
            }
            if (tk < 0) continue;
            const double lam = s_lam[ti * a.npair + ord_pair(tj, tk, a.ntypes)];
            const double cs = u1[0] * u2.x + u1[1] * u2.y + u1[2] * u2.z;
            const double cp13 = cs + 1.0 / 3.0;
            const double hw = u2.w * (lam * cp13 * cp13), hdw = u2.w * (2.0 * lam * cp13);
            Sa += hw;
            Sbc += hdw * cs;
            V[0] += hdw * u2.x;
            V[1] += hdw * u2.y;
            V[2] += hdw * u2.z;
          }
          const double hi = h1 * i1, cu = dh1 * Sa - hi * Sbc;
          for (int k = 0; k < 3; ++k) T[k] = -cu * u1[k] - hi * V[k];
          E += 0.5 * (h1 * Sa);
          for (int k = 0; k < 3; ++k) F[k] -= T[k];
          for (int x = 0; x < 3; ++x)
            for (int y = 0; y < 3; ++y) sg[3 * x + y] += d1[x] * T[y];
        }
      }
      for (int k = 0; k < 3; ++k) a.pairbuf[3 * pj + k] = T[k];
    }
  }
  for (int k = 0; k < 3; ++k) F[k] = group_sum(F[k]);
  E = group_sum(E);
  for (int k = 0; k < 9; ++k) sg[k] = group_sum(sg[k]);
  if (live && l == 0) {
    for (int k = 0; k < 3; ++k) a.fcentre[3 * i + k] = F[k];
    a.e_atom[i] = E;
    for (int k = 0; k < 9; ++k) a.sig_atom[9 * i + k] = sg[k];
  }
}

// one block per tile: gather the reverse-pair slots, write force rows, reduce
// the tile's energy and virial in atom order
__global__ __launch_bounds__(64) void k_sw_tiles(SwArgs a) {
  __shared__ double red[64][11];
  const int tile = blockIdx.x;
  const int na = a.tile_na[tile];
  const int t = threadIdx.x;
  if (t < na) {
    const int64_t i = int64_t(a.tile_a0[tile]) + t;
    const int c = a.tile_cfg[tile];
    double F[3] = {a.fcentre[3 * i], a.fcentre[3 * i + 1], a.fcentre[3 * i + 2]};
    for (int64_t p = a.pair_off[i]; p < a.pair_off[i + 1]; ++p) {
      const int32_t q = a.rev[p];
      if (q < 0) continue;
      for (int k = 0; k < 3; ++k) F[k] += a.pairbuf[3 * int64_t(q) + k];
    }
    const int64_t row = a.row_off[c] + 3 * (i - a.atom_off[c]);
    for (int k = 0; k < 3; ++k) {
      a.forces[3 * i + k] = F[k];
      a.residual[row + k] = F[k] - a.ref_force[3 * i + k];
    }
    red[t][0] = a.e_atom[i];
    red[t][1] = 0.0;
    for (int k = 0; k < 9; ++k) red[t][2 + k] = a.sig_atom[9 * i + k];
  }
  __syncthreads();
  if (t < 11) {
    double s = 0.0;
    for (int q = 0; q < na; ++q) s += red[q][t];
    a.tile_part[int64_t(tile) * 11 + t] = s;
  }
}

} // namespace

void launch_sw_rev(Context &c) {
  c.rev.alloc(size_t(std::max<int64_t>(c.P, 1)));
  if (c.P == 0) return;
  hipLaunchKernelGGL(k_sw_rev, dim3(unsigned((c.N + 127) / 128)), dim3(128), 0, c.stream, c.N,
                     c.pair_off.p, c.cfg_of.p, c.atom_off.p, c.pair_bits.p, c.pair_img.p, c.rev.p);
  FS_HIP(hipGetLastError());
}

SwArgs make_sw_args(Context &c) {
  c.pairbuf.alloc(3 * size_t(std::max<int64_t>(c.P, 1)));
  c.Fpp.alloc(3 * size_t(std::max<int64_t>(c.N, 1))); // centre forces
  c.lamb.alloc(9 * size_t(std::max<int64_t>(c.N, 1))); // per-atom virial
  SwArgs a;
  a.N = c.N;
  a.pair_off = c.pair_off.p;
  a.types = c.types.p;
  a.pd = c.pair_d.p;
  a.bits = c.pair_bits.p;
  a.rev = c.rev.p;
  a.sw = c.sw.p;
  a.lam = c.lam.p;
  a.ntypes = c.ntypes;
  a.npair = c.npair;
  a.pv = c.sw_pv.p;
  a.tri = c.sw_tri.p;
  a.pairbuf = c.pairbuf.p;
  a.fcentre = c.Fpp.p;
  a.e_atom = c.e_atom.p;
  a.sig_atom = c.lamb.p;
  a.tile_cfg = c.tile_cfg.p;
  a.tile_a0 = c.tile_a0.p;
  a.tile_na = c.tile_na.p;
  a.ntiles = c.ntiles;
  a.atom_off = c.atom_off.p;
  a.row_off = c.row_off.p;
  a.ref_force = c.ref_force.p;
  a.forces = c.forces.p;
  a.residual = c.residual.p;
  a.tile_part = c.tile_part.p;
  a.lim_atom = c.lim_atom.p;
  return a;
}

void launch_sw_pair_terms(Context &c, bool tangents) {
  const size_t P1 = size_t(std::max<int64_t>(c.P, 1));
  c.sw_pv.alloc(8 * P1);
  c.sw_tri.alloc(8 * P1);
  tangents = tangents || c.sw_tan_next;
  if (tangents) c.sw_pt2.alloc(6 * P1);
  if (c.N == 0) return;
  // the residual and the Jacobian at one parameter vector share the factors;
  // the Jacobian's two-body tangents come from the same pass when asked for
  if (c.sw_pv_gen == c.sw_gen && (!tangents || c.sw_pt2_gen == c.sw_gen)) return;
  c.sw_pv_gen = c.sw_gen;
  const dim3 grid(unsigned((c.N * kSwLanes + 255) / 256));
  if (tangents) {
    c.sw_pt2_gen = c.sw_gen;
    hipLaunchKernelGGL(k_sw_pair_terms<true>, grid, dim3(256), 0, c.stream, c.N, c.P, c.pair_off.p,
                       c.types.p, c.pair_d.p, c.pair_bits.p, c.sw.p, c.ntypes, c.sw_pv.p, c.sw_tri.p,
                       c.sw_pt2.p);
  } else {
    hipLaunchKernelGGL(k_sw_pair_terms<false>, grid, dim3(256), 0, c.stream, c.N, c.P, c.pair_off.p,
                       c.types.p, c.pair_d.p, c.pair_bits.p, c.sw.p, c.ntypes, c.sw_pv.p, c.sw_tri.p,
                       nullptr);
  }
  FS_HIP(hipGetLastError());
}

void launch_sw_eval(Context &c) {
  FS_REQUIRE(c.tile_atoms <= 64, FS_ERR_UNSUPPORTED, "SW tiles hold at most 64 atoms");
  launch_sw_pair_terms(c, false);
  const SwArgs a = make_sw_args(c);
  if (c.N > 0) {
    const dim3 grid(unsigned((c.N * kSwLanes + 255) / 256));
    hipLaunchKernelGGL(k_sw_centre, grid, dim3(256), 0, c.stream, a);
  }
  FS_HIP(hipGetLastError());
  if (c.ntiles > 0)
    hipLaunchKernelGGL(k_sw_tiles, dim3(c.ntiles), dim3(64), 0, c.stream, a);
  FS_HIP(hipGetLastError());
}

} // namespace fsa
