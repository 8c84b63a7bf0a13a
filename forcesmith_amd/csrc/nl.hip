// GPU neighbour-list build: per-configuration cell lists in fractional
// coordinates, FP64 geometry with no floating-point contraction, so the pair
// set, the per-atom order and every displacement bit equal the reference's
// brute-force build (src/core/neighbor_list.cpp:59-92):
//   wrap = box*(f - floor(f)), f = inv*r            (boundary_conditions.hpp:11-16)
//   d = ((w_j - w_i) + ix*a) + iy*b) + iz*c, keep iff !(|d|^2 > rc^2)
//   skip (i == j, image 0); |ix| <= s_x = ceil(rc*|row_x(inv)|), ...
//   per atom: j ascending, then (ix, iy, iz) lexicographic.
// Cells are >= rc wide in fractional units, so a 3x3x3 stencil (wider for
// boxes thinner than rc) finds every candidate; the per-atom segment is then
// sorted into the reference order.
#pragma clang fp contract(off)

#include "common.hpp"

#include <hipcub/hipcub.hpp>

#include <cmath>

namespace fsa {

namespace {

__device__ inline int floordiv(int u, int n) {
  int q = u / n;
  if ((u % n != 0) && ((u < 0) != (n < 0))) --q;
  return q;
}

__global__ void k_wrap_bin(int64_t N, const int32_t *__restrict__ cfg_of,
                           const double *__restrict__ pos, const double *__restrict__ box,
                           const double *__restrict__ inv, const CfgGrid *__restrict__ grid,
                           double *__restrict__ wpos, int32_t *__restrict__ cell_of,
                           int32_t *__restrict__ cell_count) {
  const int64_t a = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (a >= N) return;
  const int c = cfg_of[a];
  const double *B = box + 9 * c;
  const double *I = inv + 9 * c;
  const double r0 = pos[3 * a], r1 = pos[3 * a + 1], r2 = pos[3 * a + 2];
  double f[3];
  for (int k = 0; k < 3; ++k) {
    double v = I[3 * k] * r0 + I[3 * k + 1] * r1;
    v = v + I[3 * k + 2] * r2;
    f[k] = v - floor(v);
  }
  for (int k = 0; k < 3; ++k) {
    double v = B[3 * k] * f[0] + B[3 * k + 1] * f[1];
    wpos[3 * a + k] = v + B[3 * k + 2] * f[2];
  }
  const CfgGrid g = grid[c];
  int ci[3];
  for (int k = 0; k < 3; ++k) {
    int q = 0;
    if (f[k] > 0.0) q = int(f[k] * double(g.nc[k]));
    if (q >= g.nc[k]) q = g.nc[k] - 1;
    if (q < 0) q = 0;
    ci[k] = q;
  }
  const int cell = g.cell0 + ci[0] + g.nc[0] * (ci[1] + g.nc[1] * ci[2]);
  cell_of[a] = cell;
  atomicAdd(&cell_count[cell], 1);
}

__global__ void k_cell_scatter(int64_t N, const int32_t *__restrict__ cell_of,
                               int32_t *__restrict__ cursor, int32_t *__restrict__ cell_atoms) {
  const int64_t a = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (a >= N) return;
  const int slot = atomicAdd(&cursor[cell_of[a]], 1);
  cell_atoms[slot] = int32_t(a);
}

__device__ inline uint64_t pair_key(int jl, int ix, int iy, int iz) {
  return (uint64_t(uint32_t(jl)) << 15) | (uint64_t(ix + 16) << 10) |
         (uint64_t(iy + 16) << 5) | uint64_t(iz + 16);
}

template <bool FILL>
__global__ void k_nl_pairs(int64_t N, const int32_t *__restrict__ cfg_of,
                           const int64_t *__restrict__ atom_off,
                           const CfgGrid *__restrict__ grid, const double *__restrict__ box,
                           const double *__restrict__ wpos, const int32_t *__restrict__ cell_of,
                           const int32_t *__restrict__ cell_start,
                           const int32_t *__restrict__ cell_atoms, double rc2,
                           int64_t *__restrict__ counts, const int64_t *__restrict__ pair_off,
                           uint64_t *__restrict__ keys, double *__restrict__ pd) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const int c = cfg_of[i];
  const CfgGrid g = grid[c];
  const double *B = box + 9 * c;
  const int64_t a0 = atom_off[c];
  const double wi0 = wpos[3 * i], wi1 = wpos[3 * i + 1], wi2 = wpos[3 * i + 2];
  int cl = cell_of[i] - g.cell0;
  const int cx = cl % g.nc[0];
  cl /= g.nc[0];
  const int cy = cl % g.nc[1];
  const int cz = cl / g.nc[1];
  int64_t cnt = 0;
  int64_t o = FILL ? pair_off[i] : 0;
  for (int dz = -g.K[2]; dz <= g.K[2]; ++dz) {
    const int uz = cz + dz;
    const int nz = floordiv(uz, g.nc[2]);
    if (nz < -g.s[2] || nz > g.s[2]) continue;
    const int mz = uz - nz * g.nc[2];
    for (int dy = -g.K[1]; dy <= g.K[1]; ++dy) {
      const int uy = cy + dy;
      const int ny = floordiv(uy, g.nc[1]);
      if (ny < -g.s[1] || ny > g.s[1]) continue;
      const int my = uy - ny * g.nc[1];
      for (int dx = -g.K[0]; dx <= g.K[0]; ++dx) {
        const int ux = cx + dx;
        const int nx = floordiv(ux, g.nc[0]);
        if (nx < -g.s[0] || nx > g.s[0]) continue;
        const int mx = ux - nx * g.nc[0];
        const int m = g.cell0 + mx + g.nc[0] * (my + g.nc[1] * mz);
        const double fx = double(nx), fy = double(ny), fz = double(nz);
        for (int q = cell_start[m]; q < cell_start[m + 1]; ++q) {
          const int64_t j = cell_atoms[q];
          if (j == i && nx == 0 && ny == 0 && nz == 0) continue;
          double d[3];
          const double base[3] = {wpos[3 * j] - wi0, wpos[3 * j + 1] - wi1, wpos[3 * j + 2] - wi2};
          for (int k = 0; k < 3; ++k) {
            double v = base[k] + fx * B[3 * k + 0];
            v = v + fy * B[3 * k + 1];
            d[k] = v + fz * B[3 * k + 2];
          }
          double r2 = d[0] * d[0] + d[1] * d[1];
          r2 = r2 + d[2] * d[2];
          if (r2 > rc2) continue;
          if (FILL) {
            keys[o] = pair_key(int(j - a0), nx, ny, nz);
            pd[3 * o] = d[0];
            pd[3 * o + 1] = d[1];
            pd[3 * o + 2] = d[2];
            ++o;
          }
          ++cnt;
        }
      }
    }
  }
  if (!FILL) counts[i] = cnt;
}

// Insertion sort of each atom's segment into the reference order.
__global__ void k_nl_sort(int64_t N, const int64_t *__restrict__ pair_off,
                          uint64_t *__restrict__ keys, double *__restrict__ pd) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const int64_t p0 = pair_off[i], p1 = pair_off[i + 1];
  for (int64_t p = p0 + 1; p < p1; ++p) {
    const uint64_t k = keys[p];
    if (keys[p - 1] <= k) continue;
    const double d0 = pd[3 * p], d1 = pd[3 * p + 1], d2 = pd[3 * p + 2];
    int64_t q = p - 1;
    while (q >= p0 && keys[q] > k) {
      keys[q + 1] = keys[q];
      pd[3 * (q + 1)] = pd[3 * q];
      pd[3 * (q + 1) + 1] = pd[3 * q + 1];
      pd[3 * (q + 1) + 2] = pd[3 * q + 2];
      --q;
    }
    keys[q + 1] = k;
    pd[3 * (q + 1)] = d0;
    pd[3 * (q + 1) + 1] = d1;
    pd[3 * (q + 1) + 2] = d2;
  }
}

__global__ void k_nl_finalize(int64_t N, const int64_t *__restrict__ pair_off,
                              const int32_t *__restrict__ cfg_of,
                              const int64_t *__restrict__ atom_off,
                              const int32_t *__restrict__ types,
                              const uint64_t *__restrict__ keys, uint32_t *__restrict__ bits,
                              int32_t *__restrict__ img) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const int64_t a0 = atom_off[cfg_of[i]];
  for (int64_t p = pair_off[i]; p < pair_off[i + 1]; ++p) {
    const uint64_t k = keys[p];
    const uint32_t jl = uint32_t(k >> 15);
    const int iz = int(k & 31u) - 16, iy = int((k >> 5) & 31u) - 16, ix = int((k >> 10) & 31u) - 16;
    const int tj = types[a0 + jl];
    bits[p] = jl | (uint32_t(tj) << kTShift);
    img[p] = (ix + 512) | ((iy + 512) << 10) | ((iz + 512) << 20);
  }
}

struct PairFlagArgs {
  int64_t N;
  const int64_t *pair_off;
  const int32_t *cfg_of;
  const int32_t *types;
  const double *pd;
  const TableDesc *tabs;
  int kind, ntypes, npair;
  uint32_t *bits;
  float4 *rec;
  double *r64;
};

__device__ inline int ord(int a, int b, int nt) {
  if (a > b) {
    const int t = a;
    a = b;
    b = t;
  }
  return a * nt - a * (a - 1) / 2 + (b - a);
}

// Gates from the FP64 distance: in_range(table, r) = r >= x0 && r <= xn
// (force_calculator_concept.hpp:94-98), plus the exact-edge and r < 1e-14
// markers. The float4 record carries the rounded displacement.
__global__ void k_pair_flags(PairFlagArgs a) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= a.N) return;
  const int ti = a.types[i];
  for (int64_t p = a.pair_off[i]; p < a.pair_off[i + 1]; ++p) {
    const double d0 = a.pd[3 * p], d1 = a.pd[3 * p + 1], d2 = a.pd[3 * p + 2];
    double r2 = d0 * d0 + d1 * d1;
    r2 = r2 + d2 * d2;
    const double r = sqrt(r2);
    uint32_t b = a.bits[p] & (kJMask | (kTMask << kTShift));
    const int tj = pj_type(b);
    if (r < 1e-14) b |= kFlagTiny;
    if (a.kind != FS_SW) {
      int tab[5];
      uint32_t fl[5] = {kFlagPhi, kFlagGj, kFlagGi, kFlagDip, kFlagQuad};
      tab[0] = ord(ti, tj, a.ntypes);
      tab[1] = a.npair + tj;
      tab[2] = a.npair + ti;
      tab[3] = a.npair + 2 * a.ntypes + ord(ti, tj, a.ntypes);
      tab[4] = 2 * a.npair + 2 * a.ntypes + ord(ti, tj, a.ntypes);
      const int nrole = a.kind == FS_ADP ? 5 : 3;
      for (int q = 0; q < nrole; ++q) {
        const TableDesc T = a.tabs[tab[q]];
        if (r >= T.x0 && r <= T.xn) {
          b |= fl[q];
          if (r == T.x0 || r == T.xn) b |= kFlagEdge;
        }
      }
    }
    a.bits[p] = b;
    a.r64[p] = r;
    a.rec[p] = make_float4(float(d0), float(d1), float(d2), __uint_as_float(b));
  }
}

inline unsigned nblk(int64_t n, int bs) { return unsigned((n + bs - 1) / bs); }

template <class T>
void inclusive_sum(Context &c, const T *in, T *out, int64_t n) {
  size_t tmp = 0;
  FS_HIP(hipcub::DeviceScan::InclusiveSum(nullptr, tmp, in, out, int(n), c.stream));
  c.scan_tmp.alloc(tmp + 16);
  FS_HIP(hipcub::DeviceScan::InclusiveSum(c.scan_tmp.p, tmp, in, out, int(n), c.stream));
}

} // namespace

void launch_neighbor_build(Context &c) {
  const double rc = c.nl_rc;
  const int64_t N = c.N;
  c.pair_off.alloc(size_t(N) + 1);
  FS_HIP(hipMemsetAsync(c.pair_off.p, 0, sizeof(int64_t) * (N + 1), c.stream));
  c.P = 0;
  ++c.nl_gen;
  if (!(rc > 0.0 && std::isfinite(rc)) || N == 0) {
    c.have_nl = true;
    return;
  }
  const double rc2 = rc * rc;
  // per-configuration cell grids
  c.h_grid.resize(c.C);
  int64_t ncell = 0;
  for (int k = 0; k < c.C; ++k) {
    CfgGrid g{};
    const double *I = c.h_inv.data() + 9 * k;
    for (int a = 0; a < 3; ++a) {
      const double w = std::sqrt(I[3 * a] * I[3 * a] + I[3 * a + 1] * I[3 * a + 1] +
                                 I[3 * a + 2] * I[3 * a + 2]);
      int s = 0;
      if (std::isfinite(w) && w > 0.0) s = int(std::ceil(rc * w));
      FS_REQUIRE(s <= 15, FS_ERR_UNSUPPORTED,
                 "cutoff spans more than 15 periodic images of config " + std::to_string(k));
      int nc = 1;
      if (std::isfinite(w) && w > 0.0) {
        const double q = std::floor(1.0 / (rc * w));
        nc = q < 1.0 ? 1 : (q > 64.0 ? 64 : int(q));
      }
      const double kappa = rc * (std::isfinite(w) ? w : 1.0) * nc;
      int K = int(std::ceil(kappa * (1.0 + 1e-12) + 1e-12));
      if (K < 1) K = 1;
      g.nc[a] = nc;
      g.K[a] = K;
      g.s[a] = s;
    }
    g.cell0 = int32_t(ncell);
    ncell += int64_t(g.nc[0]) * g.nc[1] * g.nc[2];
    c.h_grid[k] = g;
  }
  FS_REQUIRE(ncell < (int64_t(1) << 30), FS_ERR_UNSUPPORTED, "too many cells");
  c.grid.alloc(c.C);
  FS_HIP(hipMemcpyAsync(c.grid.p, c.h_grid.data(), sizeof(CfgGrid) * c.C, hipMemcpyHostToDevice,
                        c.stream));
  c.wpos.alloc(3 * N);
  c.cell_of.alloc(N);
  c.cell_count.alloc(ncell);
  c.cell_start.alloc(ncell + 1);
  c.cell_atoms.alloc(N);
  FS_HIP(hipMemsetAsync(c.cell_count.p, 0, sizeof(int32_t) * ncell, c.stream));
  const int bs = 256;
  hipLaunchKernelGGL(k_wrap_bin, dim3(nblk(N, bs)), dim3(bs), 0, c.stream, N, c.cfg_of.p, c.pos.p,
                     c.box.p, c.inv.p, c.grid.p, c.wpos.p, c.cell_of.p, c.cell_count.p);
  FS_HIP(hipGetLastError());
  FS_HIP(hipMemsetAsync(c.cell_start.p, 0, sizeof(int32_t), c.stream));
  inclusive_sum(c, c.cell_count.p, c.cell_start.p + 1, ncell);
  // cursor = cell_start (reuse cell_count as the cursor)
  FS_HIP(hipMemcpyAsync(c.cell_count.p, c.cell_start.p, sizeof(int32_t) * ncell,
                        hipMemcpyDeviceToDevice, c.stream));
  hipLaunchKernelGGL(k_cell_scatter, dim3(nblk(N, bs)), dim3(bs), 0, c.stream, N, c.cell_of.p,
                     c.cell_count.p, c.cell_atoms.p);
  FS_HIP(hipGetLastError());
  DBuf<int64_t> cnt64;
  cnt64.alloc(N);
  hipLaunchKernelGGL(k_nl_pairs<false>, dim3(nblk(N, bs)), dim3(bs), 0, c.stream, N, c.cfg_of.p,
                     c.atom_off.p, c.grid.p, c.box.p, c.wpos.p, c.cell_of.p, c.cell_start.p,
                     c.cell_atoms.p, rc2, cnt64.p, (const int64_t *)nullptr, (uint64_t *)nullptr,
                     (double *)nullptr);
  FS_HIP(hipGetLastError());
  inclusive_sum(c, cnt64.p, c.pair_off.p + 1, N);
  FS_HIP(hipMemcpyAsync(&c.P, c.pair_off.p + N, sizeof(int64_t), hipMemcpyDeviceToHost, c.stream));
  FS_HIP(hipStreamSynchronize(c.stream));
  const int64_t P = c.P;
  c.keys.alloc(P ? P : 1);
  c.pair_d.alloc(3 * (P ? P : 1));
  c.pair_bits.alloc(P ? P : 1);
  c.pair_img.alloc(P ? P : 1);
  c.pair_rec.alloc(P ? P : 1);
  if (P > 0) {
    hipLaunchKernelGGL(k_nl_pairs<true>, dim3(nblk(N, bs)), dim3(bs), 0, c.stream, N, c.cfg_of.p,
                       c.atom_off.p, c.grid.p, c.box.p, c.wpos.p, c.cell_of.p, c.cell_start.p,
                       c.cell_atoms.p, rc2, (int64_t *)nullptr, c.pair_off.p, c.keys.p,
                       c.pair_d.p);
    FS_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_nl_sort, dim3(nblk(N, 64)), dim3(64), 0, c.stream, N, c.pair_off.p,
                       c.keys.p, c.pair_d.p);
    FS_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_nl_finalize, dim3(nblk(N, bs)), dim3(bs), 0, c.stream, N, c.pair_off.p,
                       c.cfg_of.p, c.atom_off.p, c.types.p, c.keys.p, c.pair_bits.p, c.pair_img.p);
    FS_HIP(hipGetLastError());
  }
  FS_HIP(hipStreamSynchronize(c.stream));
  c.have_nl = true;
}

void launch_pair_flags(Context &c) {
  if (c.P == 0) return;
  PairFlagArgs a;
  a.N = c.N;
  a.pair_off = c.pair_off.p;
  a.cfg_of = c.cfg_of.p;
  a.types = c.types.p;
  a.pd = c.pair_d.p;
  a.tabs = c.tabs.p;
  a.kind = c.kind;
  a.ntypes = c.ntypes;
  a.npair = c.npair;
  a.bits = c.pair_bits.p;
  a.rec = c.pair_rec.p;
  c.pair_r64.alloc(size_t(std::max<int64_t>(c.P, 1)));
  a.r64 = c.pair_r64.p;
  hipLaunchKernelGGL(k_pair_flags, dim3(nblk(c.N, 128)), dim3(128), 0, c.stream, a);
  FS_HIP(hipGetLastError());
  c.jsort_valid = false;
}

} // namespace fsa
