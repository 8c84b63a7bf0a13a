// Stillinger-Weber residual path: pair term plus triplets centred on each
// atom, with the reference's scatter of the triplet forces onto j and k
// (include/forcesmith/force/kernels.hpp:58-66) recast as a deterministic
// gather: the centre's thread writes the j- and k-forces into per-pair slots
// of its own neighbour segment, and every atom then gathers the slots of the
// reverse pairs (k -> i for each of its pairs i -> k).
//
// Reference: StiwebForceCalculator::eval_forces (src/force/stiweb_force.cpp:
// 241-271), v2_dv2 / h_dh (:24-54), three_body_forces (kernels.hpp:40-56).
// The SW arithmetic (pow, exp) runs in FP64.
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

__device__ inline void v2_dv2(double r, const SwP &p, double &v2, double &dv2) {
  v2 = 0.0;
  dv2 = 0.0;
  if (r >= p.a1) return;
  const double d = r - p.a1;
  const double e = exp(p.delta / d);
  if (e == 0.0) return;
  const double rp = pow(r, -p.p);
  const double rq = pow(r, -p.q);
  const double poly = p.A * rp - p.B * rq;
  const double dpoly = -p.A * p.p * pow(r, -p.p - 1.0) + p.B * p.q * pow(r, -p.q - 1.0);
  v2 = poly * e;
  dv2 = dpoly * e + poly * e * (-p.delta / (d * d));
}

__device__ inline void h_dh(double r, const SwP &p, double &h, double &dh) {
  h = 0.0;
  dh = 0.0;
  if (r >= p.a2) return;
  const double d = r - p.a2;
  const double e = exp(p.gamma / d);
  if (e == 0.0) return;
  h = e;
  dh = e * (-p.gamma / (d * d));
}

__device__ inline void three_body(const double *d1, const double *d2, double inv_r1, double inv_r2,
                                  double c, double f1, double df1, double f2, double df2, double g,
                                  double dg, double *fi, double *fj, double *fk) {
  const double inv_r1r2 = inv_r1 * inv_r2;
  const double A = c * inv_r1 * inv_r1 - inv_r1r2;
  const double B = c * inv_r2 * inv_r2 - inv_r1r2;
  const double s1 = df1 * inv_r1 * f2 * g, s2 = f1 * df2 * inv_r2 * g, s3 = f1 * f2 * dg;
  for (int a = 0; a < 3; ++a) {
    fi[a] = s1 * d1[a] + s2 * d2[a] - s3 * (A * d1[a] + B * d2[a]);
    fj[a] = -s1 * d1[a] - s3 * (inv_r1r2 * d2[a] - c * inv_r1 * inv_r1 * d1[a]);
    fk[a] = -s2 * d2[a] - s3 * (inv_r1r2 * d1[a] - c * inv_r2 * inv_r2 * d2[a]);
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

struct SwArgs {
  int64_t N;
  const int64_t *pair_off;
  const int32_t *types;
  const double *pd;
  const uint32_t *bits;
  const int32_t *rev;
  const double *sw, *lam;
  int ntypes, npair;
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

__device__ inline double pr(const double *pd, int64_t p) {
  const double d0 = pd[3 * p], d1 = pd[3 * p + 1], d2 = pd[3 * p + 2];
  double r2 = d0 * d0 + d1 * d1;
  r2 = r2 + d2 * d2;
  return sqrt(r2);
}

__global__ __launch_bounds__(128) void k_sw_centre(SwArgs a) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= a.N) return;
  const int ti = a.types[i];
  const int64_t p0 = a.pair_off[i], p1 = a.pair_off[i + 1];
  double F[3] = {0, 0, 0}, E = 0.0, sg[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int64_t p = p0; p < p1; ++p) {
    a.pairbuf[3 * p] = 0.0;
    a.pairbuf[3 * p + 1] = 0.0;
    a.pairbuf[3 * p + 2] = 0.0;
  }
  for (int64_t p = p0; p < p1; ++p) {
    const uint32_t b = a.bits[p];
    if (b & kFlagTiny) continue;
    const double r = pr(a.pd, p);
    const SwP P = load_sw(a.sw, ord_pair(ti, pj_type(b), a.ntypes));
    double v2, dv2;
    v2_dv2(r, P, v2, dv2);
    const double s = dv2 / r;
    const double *d = a.pd + 3 * p;
    const double f[3] = {s * d[0], s * d[1], s * d[2]};
    for (int k = 0; k < 3; ++k) F[k] += f[k];
    E += 0.5 * v2;
    for (int x = 0; x < 3; ++x)
      for (int y = 0; y < 3; ++y) sg[3 * x + y] -= 0.5 * d[x] * f[y];
  }
  for (int64_t pj = p0; pj < p1; ++pj) {
    const uint32_t bj = a.bits[pj];
    if (bj & kFlagTiny) continue;
    const int tj = pj_type(bj);
    const double r1 = pr(a.pd, pj);
    double h1, dh1;
    h_dh(r1, load_sw(a.sw, ord_pair(ti, tj, a.ntypes)), h1, dh1);
    if (h1 == 0.0 && dh1 == 0.0) continue;
    const double inv_r1 = 1.0 / r1;
    const double *d1 = a.pd + 3 * pj;
    for (int64_t pk = pj + 1; pk < p1; ++pk) {
      const uint32_t bk = a.bits[pk];
      if (bk & kFlagTiny) continue;
      const int tk = pj_type(bk);
      const double r2 = pr(a.pd, pk);
      double h2, dh2;
      h_dh(r2, load_sw(a.sw, ord_pair(ti, tk, a.ntypes)), h2, dh2);
      if (h2 == 0.0 && dh2 == 0.0) continue;
      const double inv_r2 = 1.0 / r2;
      const double *d2 = a.pd + 3 * pk;
      const double lam = a.lam[ti * a.npair + ord_pair(tj, tk, a.ntypes)];
      const double cs = (d1[0] * d2[0] + d1[1] * d2[1] + d1[2] * d2[2]) * inv_r1 * inv_r2;
      const double cp13 = cs + 1.0 / 3.0;
      const double w = lam * cp13 * cp13, dw = 2.0 * lam * cp13;
      E += h1 * h2 * w;
      double fi[3], fj[3], fk[3];
      three_body(d1, d2, inv_r1, inv_r2, cs, h1, dh1, h2, dh2, w, dw, fi, fj, fk);
      for (int k = 0; k < 3; ++k) {
        F[k] += fi[k];
        a.pairbuf[3 * pj + k] += fj[k];
        a.pairbuf[3 * pk + k] += fk[k];
      }
      for (int x = 0; x < 3; ++x)
        for (int y = 0; y < 3; ++y) sg[3 * x + y] += d1[x] * fj[y] + d2[x] * fk[y];
    }
  }
  for (int k = 0; k < 3; ++k) a.fcentre[3 * i + k] = F[k];
  a.e_atom[i] = E;
  for (int k = 0; k < 9; ++k) a.sig_atom[9 * i + k] = sg[k];
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

void launch_sw_eval(Context &c) {
  FS_REQUIRE(c.tile_atoms <= 64, FS_ERR_UNSUPPORTED, "SW tiles hold at most 64 atoms");
  const SwArgs a = make_sw_args(c);
  if (c.N > 0)
    hipLaunchKernelGGL(k_sw_centre, dim3(unsigned((c.N + 127) / 128)), dim3(128), 0, c.stream, a);
  FS_HIP(hipGetLastError());
  if (c.ntiles > 0)
    hipLaunchKernelGGL(k_sw_tiles, dim3(c.ntiles), dim3(64), 0, c.stream, a);
  FS_HIP(hipGetLastError());
}

} // namespace fsa
