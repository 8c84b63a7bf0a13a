// FP64 normal equations and the damped LM solve on the GPU.
//
//   JtJ = J^T J and g = J^T F from the FP32 Jacobian and the FP64 residual,
//   with every product and sum in FP64 on the matrix cores (one pass over J,
//   split-K over row chunks reduced in a fixed chunk order: deterministic).
//   (JtJ + mu * diag(d)) dx = -g by an in-place Cholesky factorisation and two
//   triangular solves in one workgroup, plus the predicted reduction
//   dx . (mu d dx - g) the LM gain ratio needs.
// The reference's normal-equation step (NormalEquationsSolver::minimize,
// src/optimization/solver.cpp:242-266) forms the same matrices on the CPU.
#include "common.hpp"

#include <algorithm>
#include <cmath>
#include <mutex>
#include <type_traits>

namespace fsa {

namespace {

typedef double d4v __attribute__((ext_vector_type(4)));

constexpr int kMT = 16; // MFMA tile edge (v_mfma_f64_16x16x4f64)

// Upper-triangle 16x16 tile pairs (I <= J) of an nt x nt tile grid, row-major.
__host__ __device__ inline int tile_pair(int I, int J, int nt) { return I * nt - I * (I - 1) / 2 + (J - I); }

// One pass over a chunk of Jacobian rows on the FP64 matrix cores. Wave w
// owns tile rows I0 = w and I1 = NT - 1 - w of the upper triangle of J^T J
// (NT + 1 tiles per wave, so the waves are balanced; for odd NT the middle
// wave owns row I0 = I1 alone, (NT + 1) / 2 tiles), four rows per step, plus
// the Jt r entries of its columns. Operand layout of v_mfma_f64_16x16x4f64
// (measured on gfx950): lane l holds A[l % 16][l / 16] and B[l / 16][l % 16],
// and after the product D[4 v + l / 16][l % 16] in element v. A is the
// transposed column tile I of four rows and B the column tile J of the same
// rows: the FP32 entries are widened to FP64 and every product and sum is
// FP64. Partials per chunk are reduced in chunk order (deterministic).
//
// NT (tiles per row of the grid) is a template parameter, so a wave's tile
// list is a fixed, fully unrolled sequence of independent accumulators: a
// step reads all its B operands from LDS first and then issues its MFMAs back
// to back. The 32-row slabs are double-buffered in LDS: the next slab's
// global loads are in flight in registers while the current one is
// multiplied, and one barrier per slab separates the buffers.
// 32-row slabs: 16 floats per thread in flight. (64-row slabs, 32 floats in
// flight, spill more registers and measured 3 % slower; 64-row slabs staged
// without double buffering were 0.2 % slower on 240 columns and 12 % on 140.)
constexpr int kNeRows = 32; // Jacobian rows per LDS stage
constexpr int kNePre = 16;  // floats per thread of a slab in flight (see launch)

// LDS row pitch: D rounded up to 32 floats plus 8, so the four rows one MFMA
// step reads start 8 banks apart and every bank serves exactly two lanes
__host__ __device__ inline int ne_pitch(int D) { return ((D + 31) / 32) * 32 + 8; }

// The first K of the KA tile slots: slot j < n0 is tile (I0, I0 + j), slot
// j >= n0 is tile (I1, I1 + j - n0); n0 is wave-uniform
template <int K, int KA>
__device__ __forceinline__ void ne_step(const float *row, double f, int c0, int c1, int cl,
                                        int I0, int I1, int n0, d4v (&acc)[KA], double &g0,
                                        double &g1) {
  const double a0 = double(row[c0]);
  const double a1 = double(row[c1]);
  g0 += a0 * f;
  g1 += a1 * f;
  float b[K];
#pragma unroll
  for (int j = 0; j < K; ++j) {
    const int jt = j < n0 ? I0 + j : I1 + (j - n0);
    b[j] = row[kMT * jt + cl];
  }
#pragma unroll
  for (int j = 0; j < K; ++j)
    acc[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(j < n0 ? a0 : a1, double(b[j]), acc[j], 0, 0, 0);
}

template <int K, int KA>
__device__ __forceinline__ void ne_store(const d4v (&acc)[KA], double *o, int I0, int I1, int n0,
                                         int NT, int kr, int cl) {
#pragma unroll
  for (int j = 0; j < K; ++j) {
    const int I = j < n0 ? I0 : I1;
    const int Jt = j < n0 ? I0 + j : I1 + (j - n0);
    double *t = o + int64_t(tile_pair(I, Jt, NT)) * 256;
#pragma unroll
    for (int v = 0; v < 4; ++v) t[(4 * v + kr) * kMT + cl] = acc[j][v];
  }
}

template <int NT>
__global__ __launch_bounds__(512) void k_ne_mfma(const float *__restrict__ J,
                                                 const double *__restrict__ F, int64_t R, int D,
                                                 int nchunk, double *__restrict__ part) {
  constexpr int KW = NT + 1, KM = (NT + 1) / 2; // tile slots: paired rows, middle row
  extern __shared__ float sJ[];                 // [2][kNeRows][pitch]
  __shared__ double sF[2][kNeRows];
  const int pitch = ne_pitch(D);
  const int w = __builtin_amdgcn_readfirstlane(int(threadIdx.x >> 6)), l = threadIdx.x & 63;
  const int I0 = w, I1 = NT - 1 - w;
  const bool mid = I0 == I1; // odd NT: the middle wave owns one tile row
  const int n0 = NT - I0;
  const int ntp = NT * (NT + 1) / 2;
  const int64_t rows_per = (((R + nchunk - 1) / nchunk) + 3) & ~int64_t(3);
  const int64_t r0 = int64_t(blockIdx.x) * rows_per;
  const int64_t r1 = r0 + rows_per < R ? r0 + rows_per : R;
  const int kr = l >> 4, cl = l & 15;
  const int c0 = kMT * I0 + cl, c1 = kMT * I1 + cl;
  const int nthr = blockDim.x;
  d4v acc[KW];
#pragma unroll
  for (int j = 0; j < KW; ++j) acc[j] = d4v{0.0, 0.0, 0.0, 0.0};
  double g0 = 0.0, g1 = 0.0;
  // pad columns D .. pitch of both buffers stay zero for the whole pass
  for (int e = threadIdx.x; e < 2 * kNeRows * (pitch - D); e += nthr)
    sJ[(e / (pitch - D)) * pitch + D + e % (pitch - D)] = 0.0f;
  // a slab in flight: float4 elements (rows 16-byte aligned) or single floats
  const bool vec = (D & 3) == 0;
  const int q = vec ? D >> 2 : D; // elements per row
  float pre[kNePre];
  double fpre = 0.0;
  auto load = [&](int64_t sb) {
    const int nr = int(r1 - sb < kNeRows ? r1 - sb : kNeRows);
    const int ne = nr * q;
    if (vec) {
      const float4 *src = reinterpret_cast<const float4 *>(J + sb * D);
#pragma unroll
      for (int k = 0; k < kNePre / 4; ++k) {
        const int e = threadIdx.x + k * nthr;
        const float4 v = e < ne ? src[e] : float4{0.0f, 0.0f, 0.0f, 0.0f};
        pre[4 * k] = v.x;
        pre[4 * k + 1] = v.y;
        pre[4 * k + 2] = v.z;
        pre[4 * k + 3] = v.w;
      }
    } else {
      const float *src = J + sb * D;
#pragma unroll
      for (int k = 0; k < kNePre; ++k) {
        const int e = threadIdx.x + k * nthr;
        pre[k] = e < ne ? src[e] : 0.0f;
      }
    }
    fpre = int(threadIdx.x) < nr ? F[sb + threadIdx.x] : 0.0;
  };
  // rows past the slab's end are written as zeros (a short last slab)
  auto store = [&](int buf) {
    float *dst = sJ + buf * kNeRows * pitch;
    const int ne = kNeRows * q;
    if (vec) {
#pragma unroll
      for (int k = 0; k < kNePre / 4; ++k) {
        const int e = threadIdx.x + k * nthr;
        if (e < ne) {
          float *d = dst + (e / q) * pitch + 4 * (e % q);
          d[0] = pre[4 * k];
          d[1] = pre[4 * k + 1];
          d[2] = pre[4 * k + 2];
          d[3] = pre[4 * k + 3];
        }
      }
    } else {
#pragma unroll
      for (int k = 0; k < kNePre; ++k) {
        const int e = threadIdx.x + k * nthr;
        if (e < ne) dst[(e / q) * pitch + e % q] = pre[k];
      }
    }
    if (int(threadIdx.x) < kNeRows) sF[buf][threadIdx.x] = fpre;
  };
  if (r0 < r1) {
    load(r0);
    store(0);
  }
  __syncthreads();
  int buf = 0;
  for (int64_t sb = r0; sb < r1; sb += kNeRows, buf ^= 1) {
    const int nr = int(r1 - sb < kNeRows ? r1 - sb : kNeRows);
    const bool more = sb + kNeRows < r1;
    if (more) load(sb + kNeRows); // in flight during the products below
    const float *base = sJ + buf * kNeRows * pitch;
    if (mid) {
      for (int rb = 0; rb < nr; rb += 4)
        ne_step<KM>(base + (rb + kr) * pitch, sF[buf][rb + kr], c0, c0, cl, I0, I0, n0, acc, g0, g1);
    } else {
      for (int rb = 0; rb < nr; rb += 4)
        ne_step<KW>(base + (rb + kr) * pitch, sF[buf][rb + kr], c0, c1, cl, I0, I1, n0, acc, g0, g1);
    }
    if (more) store(buf ^ 1);
    __syncthreads();
  }
  // Jt r: the four lanes of one column hold its four row-phase sums
  g0 += __shfl_xor(g0, 16, 64);
  g0 += __shfl_xor(g0, 32, 64);
  g1 += __shfl_xor(g1, 16, 64);
  g1 += __shfl_xor(g1, 32, 64);
  double *o = part + int64_t(blockIdx.x) * (int64_t(ntp) * 256 + D);
  if (mid) ne_store<KM>(acc, o, I0, I0, n0, NT, kr, cl);
  else ne_store<KW>(acc, o, I0, I1, n0, NT, kr, cl);
  if (kr == 0) {
    if (c0 < D) o[int64_t(ntp) * 256 + c0] = g0;
    if (!mid && c1 < D) o[int64_t(ntp) * 256 + c1] = g1;
  }
}

// First level of the chunk reduction: group gi sums chunks [gi*per_g, ...)
// in order; the second level (k_ne_reduce) sums the groups in order.
constexpr int kNeGroups = 16;
__global__ void k_ne_reduce1(int64_t per, int nchunk, const double *__restrict__ part,
                             double *__restrict__ part2) {
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const int gi = blockIdx.y;
  if (e >= per) return;
  const int cpg = (nchunk + kNeGroups - 1) / kNeGroups;
  const int k0 = gi * cpg, k1 = k0 + cpg < nchunk ? k0 + cpg : nchunk;
  double s = 0.0;
  for (int k = k0; k < k1; ++k) s += part[int64_t(k) * per + e];
  part2[int64_t(gi) * per + e] = s;
}

// Sum the chunk partials in chunk order; both triangles of J^T J are written
// from the upper one (diagonal tiles from their i <= j half), so the matrix is
// exactly symmetric.
__global__ void k_ne_reduce(int D, int nt, int nchunk, const double *__restrict__ part,
                            double *__restrict__ A, double *__restrict__ g) {
  const int ntp = nt * (nt + 1) / 2;
  const int64_t per = int64_t(ntp) * 256 + D;
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= per) return;
  double s = 0.0;
  for (int k = 0; k < nchunk; ++k) s += part[int64_t(k) * per + e];
  if (e >= int64_t(ntp) * 256) {
    g[e - int64_t(ntp) * 256] = s;
    return;
  }
  const int tp = int(e / 256), w = int(e % 256), i = w / kMT, j = w % kMT;
  int I = 0;
  while (tile_pair(I, nt - 1, nt) < tp) ++I; // tile row of pair tp
  const int Jt = I + (tp - tile_pair(I, I, nt));
  const int a = kMT * I + i, b = kMT * Jt + j;
  if (a >= D || b >= D) return;
  if (I == Jt && i > j) return;
  A[int64_t(a) * D + b] = s;
  A[int64_t(b) * D + a] = s;
}

// Marquardt scaling d_k = max(A_kk, 1e-12 max_k A_kk) (1 when all zero);
// scal[0] = max diagonal
__global__ void k_diag_scale(int D, const double *__restrict__ A, double *__restrict__ dg,
                             double *__restrict__ scal) {
  __shared__ double smax[256];
  double m = 0.0;
  for (int k = threadIdx.x; k < D; k += blockDim.x) m = fmax(m, A[int64_t(k) * D + k]);
  smax[threadIdx.x] = m;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + w]);
    __syncthreads();
  }
  const double mx = smax[0];
  for (int k = threadIdx.x; k < D; k += blockDim.x) {
    const double a = A[int64_t(k) * D + k];
    double d = a > 1e-12 * mx ? a : 1e-12 * mx;
    if (!(d > 0.0)) d = 1.0;
    dg[k] = d;
  }
  if (threadIdx.x == 0) scal[0] = mx;
}

// (A + mu diag(dg)) dx = -g by a blocked Cholesky in one workgroup (D <= 256);
// scal[1] = info (0 ok), scal[2] = dx.(mu dg dx - g). The lower triangle of M
// (global, L2-resident) holds the factor. Per 16-column block: the panel
// (rows j0..D) is staged in LDS and factored column by column; the trailing
// lower triangle takes one rank-16 update, each thread keeping one panel row
// in registers; the triangular solves run block by block.
constexpr int kCB = 16;

__global__ __launch_bounds__(1024) void k_chol_solve(int D, const double *__restrict__ A,
                                                     const double *__restrict__ g,
                                                     const double *__restrict__ dg, double mu,
                                                     double *__restrict__ M, double *__restrict__ dx,
                                                     double *__restrict__ scal) {
  __shared__ double sP[256][kCB + 1];
  __shared__ double red[1024 / 64];
  __shared__ int bad;
  const int t = threadIdx.x, nt = blockDim.x;
  if (t == 0) bad = 0;
  for (int e = t; e < D * D; e += nt) M[e] = A[e];
  __syncthreads();
  for (int k = t; k < D; k += nt) {
    M[int64_t(k) * D + k] += mu * dg[k];
    dx[k] = -g[k];
  }
  __syncthreads();
  for (int j0 = 0; j0 < D; j0 += kCB) {
    const int w = min(kCB, D - j0), n2 = D - j0;
    for (int e = t; e < n2 * w; e += nt) {
      const int r = e / w, cc = e % w;
      sP[r][cc] = M[int64_t(j0 + r) * D + j0 + cc];
    }
    __syncthreads();
    // factor the panel (rows 0..n2 of it, columns 0..w)
    for (int jj = 0; jj < w; ++jj) {
      if (t == 0) {
        const double d = sP[jj][jj];
        if (!(d > 0.0) || !isfinite(d)) bad = 1;
        sP[jj][jj] = d > 0.0 ? sqrt(d) : 1.0;
      }
      __syncthreads();
      const double l = sP[jj][jj];
      for (int r = jj + 1 + t; r < n2; r += nt) sP[r][jj] /= l;
      __syncthreads();
      const int nc = w - jj - 1;
      if (nc > 0)
        for (int e = t; e < (n2 - jj - 1) * nc; e += nt) {
          const int r = jj + 1 + e / nc, cc = jj + 1 + e % nc;
          if (cc <= r) sP[r][cc] -= sP[r][jj] * sP[cc][jj];
        }
      __syncthreads();
    }
    for (int e = t; e < n2 * w; e += nt) {
      const int r = e / w, cc = e % w;
      if (cc <= r) M[int64_t(j0 + r) * D + j0 + cc] = sP[r][cc];
    }
    // trailing update: M[i][k] -= sum_p P[i][p] P[k][p] for w <= k <= i < n2
    const int m2 = n2 - w;
    if (m2 > 0) {
      const int G = nt / m2; // threads per trailing row (>= 4 for D <= 256)
      const int ri = t / G, sub = t % G;
      if (ri < m2) {
        double li[kCB];
#pragma unroll
        for (int p = 0; p < kCB; ++p) li[p] = p < w ? sP[w + ri][p] : 0.0;
        double *Mi = M + int64_t(j0 + w + ri) * D + j0 + w;
        for (int rk = sub; rk <= ri; rk += G) {
          double s = 0.0;
#pragma unroll
          for (int p = 0; p < kCB; ++p)
            if (p < w) s += li[p] * sP[w + rk][p];
          Mi[rk] -= s;
        }
      }
    }
    __syncthreads();
  }
  // L y = b, then L^T x = y, block by block with the right-hand side and each
  // diagonal block in LDS (the serial part is a 16 x 16 solve on thread 0)
  __shared__ double sx[256];
  for (int k = t; k < D; k += nt) sx[k] = dx[k];
  __syncthreads();
  for (int j0 = 0; j0 < D; j0 += kCB) {
    const int w = min(kCB, D - j0);
    for (int e = t; e < w * w; e += nt) {
      const int r = e / w, cc = e % w;
      sP[r][cc] = cc <= r ? M[int64_t(j0 + r) * D + j0 + cc] : 0.0;
    }
    __syncthreads();
    if (t == 0)
      for (int k = 0; k < w; ++k) {
        double v = sx[j0 + k];
        for (int p = 0; p < k; ++p) v -= sP[k][p] * sx[j0 + p];
        sx[j0 + k] = v / sP[k][k];
      }
    __syncthreads();
    for (int i = j0 + w + t; i < D; i += nt) {
      double v = 0.0;
      for (int p = 0; p < w; ++p) v += M[int64_t(i) * D + j0 + p] * sx[j0 + p];
      sx[i] -= v;
    }
    __syncthreads();
  }
  for (int j1 = D; j1 > 0; j1 -= kCB) {
    const int j0 = max(0, j1 - kCB), w = j1 - j0;
    for (int e = t; e < w * w; e += nt) {
      const int r = e / w, cc = e % w;
      sP[r][cc] = cc <= r ? M[int64_t(j0 + r) * D + j0 + cc] : 0.0;
    }
    __syncthreads();
    if (t == 0)
      for (int k = w - 1; k >= 0; --k) {
        double v = sx[j0 + k];
        for (int i = k + 1; i < w; ++i) v -= sP[i][k] * sx[j0 + i];
        sx[j0 + k] = v / sP[k][k];
      }
    __syncthreads();
    for (int k = t; k < j0; k += nt) {
      double v = 0.0;
      for (int i = j0; i < j1; ++i) v += M[int64_t(i) * D + k] * sx[i];
      sx[k] -= v;
    }
    __syncthreads();
  }
  for (int k = t; k < D; k += nt) dx[k] = sx[k];
  double s = 0.0;
  for (int k = t; k < D; k += nt) s += sx[k] * (mu * dg[k] * sx[k] - g[k]);
  for (int m = 32; m > 0; m >>= 1) s += __shfl_xor(s, m, 64);
  if ((t & 63) == 0) red[t >> 6] = s;
  __syncthreads();
  if (t == 0) {
    double tot = 0.0;
    for (int w = 0; w < nt / 64; ++w) tot += red[w];
    scal[1] = bad ? 1.0 : 0.0;
    scal[2] = tot;
  }
}

} // namespace

// the partial sums of one normal-equation evaluation: nchunk chunk partials
// and kNeGroups group partials of per = ntp 256 + D doubles each
struct NeShape {
  int nt, nchunk;
  int64_t per;
};

// one chunk of rows per CU (at least 64 rows per chunk): one 8-wave block
// fits a CU at a time; 1, 2, 3 and 4 chunks per CU measured 1.15, 1.18, 1.21
// and 1.23 ms at 240 columns. max_chunks caps the count (a fixed-size
// scratch); the chunking fixes the summation order, so equal chunk counts
// give bit-identical products.
NeShape ne_shape(int cu_count, int64_t R, int D, int max_chunks) {
  FS_REQUIRE(D <= 16 * kMT, FS_ERR_UNSUPPORTED, "normal equations support at most 256 parameters");
  NeShape sh;
  sh.nt = (D + kMT - 1) / kMT;
  sh.per = int64_t(sh.nt * (sh.nt + 1) / 2) * 256 + D;
  sh.nchunk = int(std::max<int64_t>(1, std::min<int64_t>({int64_t(cu_count), (R + 63) / 64,
                                                          int64_t(max_chunks)})));
  return sh;
}

// JtJ[D x D] = J^T J, Jtr[D] = J^T F of J (FP32, R x D row-major) and F
// (FP64) through part[(nchunk + kNeGroups) * per]; every partial is written
// before it is read, so nothing needs clearing
void ne_launch(hipStream_t st, const NeShape &sh, const float *J, const double *F, int64_t R, int D,
               double *part, double *JtJ, double *Jtr) {
  const int nt = sh.nt, nchunk = sh.nchunk;
  const int64_t per = sh.per;
  double *part2 = part + size_t(nchunk) * per;
  // two slab buffers; a thread holds at most kNePre floats of a slab in
  // flight (kNeRows rows of D floats over 64 (nt + 1) / 2 threads: D <= 16 nt)
  const size_t lds = 2 * size_t(kNeRows) * ne_pitch(D) * sizeof(float);
  const int nthr = 64 * ((nt + 1) / 2);
  FS_REQUIRE((kNeRows * D + nthr - 1) / nthr <= kNePre, FS_ERR_UNSUPPORTED,
             "normal-equation slab does not fit the per-thread staging");
  auto go = [&](auto ntc) {
    constexpr int NT = decltype(ntc)::value;
    hipLaunchKernelGGL((k_ne_mfma<NT>), dim3(nchunk), dim3(nthr), lds, st, J, F, R, D, nchunk, part);
  };
  switch (nt) {
  case 1: go(std::integral_constant<int, 1>{}); break;
  case 2: go(std::integral_constant<int, 2>{}); break;
  case 3: go(std::integral_constant<int, 3>{}); break;
  case 4: go(std::integral_constant<int, 4>{}); break;
  case 5: go(std::integral_constant<int, 5>{}); break;
  case 6: go(std::integral_constant<int, 6>{}); break;
  case 7: go(std::integral_constant<int, 7>{}); break;
  case 8: go(std::integral_constant<int, 8>{}); break;
  case 9: go(std::integral_constant<int, 9>{}); break;
  case 10: go(std::integral_constant<int, 10>{}); break;
  case 11: go(std::integral_constant<int, 11>{}); break;
  case 12: go(std::integral_constant<int, 12>{}); break;
  case 13: go(std::integral_constant<int, 13>{}); break;
  case 14: go(std::integral_constant<int, 14>{}); break;
  case 15: go(std::integral_constant<int, 15>{}); break;
  default: go(std::integral_constant<int, 16>{}); break;
  }
  FS_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_ne_reduce1, dim3(unsigned((per + 255) / 256), kNeGroups), dim3(256), 0, st, per,
                     nchunk, part, part2);
  FS_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_ne_reduce, dim3(unsigned((per + 255) / 256)), dim3(256), 0, st, D, nt,
                     kNeGroups, part2, JtJ, Jtr);
  FS_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Normal equations that allocate nothing (fs_normal_equations_noalloc): the
// chunk and group partials and the cost partials live in one file-scope
// device array. Every device holds its own copy of it at its own address,
// looked up once per device, with that device current, by fs_ctx_create
// (noalloc_scratch_init); host code uses only those addresses, never the
// array's name (in host code that name is a host-side stand-in).
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20; // 64 MiB per device
constexpr int kMaxDevices = 64;
static __device__ __align__(256) double scratch[kNoAllocScratchBytes / sizeof(double)];

namespace {
std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxDevices] = {};

double *scratch_on(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return device >= 0 && device < kMaxDevices ? g_scratch_ptr[device] : nullptr;
}

// a device pointer of this device (not host memory, not another device's)
void require_device_ptr(const void *p, int device, const char *what) {
  hipPointerAttribute_t at{};
  const hipError_t e = hipPointerGetAttributes(&at, p);
  if (e != hipSuccess) (void)hipGetLastError(); // host memory unknown to HIP
  FS_REQUIRE(e == hipSuccess && at.device == device &&
                 (at.type == hipMemoryTypeDevice || at.type == hipMemoryTypeManaged),
             FS_ERR_ARG, std::string(what) + " must be device memory of the context's device");
}
} // namespace

void noalloc_scratch_init(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  FS_REQUIRE(device >= 0 && device < kMaxDevices, FS_ERR_UNSUPPORTED, "device ordinal above 63");
  if (g_scratch_ptr[device]) return;
  void *p = nullptr;
  FS_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  g_scratch_ptr[device] = static_cast<double *>(p);
}

int fs_ne_noalloc_impl(Context &c, const float *J, const double *F, int64_t R, int D, double *JtJ,
                       double *Jtr, double *cost) {
  FS_REQUIRE(J && F && JtJ && Jtr, FS_ERR_ARG, "null normal-equation argument");
  FS_REQUIRE(R >= 1 && D >= 1, FS_ERR_ARG, "R and D must be positive");
  FS_REQUIRE(D <= 16 * kMT, FS_ERR_UNSUPPORTED,
             "normal equations without allocation support at most 256 parameters");
  double *base = scratch_on(c.device);
  FS_REQUIRE(base, FS_ERR_STATE, "no-allocation scratch not set up on this device");
  require_device_ptr(J, c.device, "J");
  require_device_ptr(F, c.device, "F");
  require_device_ptr(JtJ, c.device, "JtJ");
  require_device_ptr(Jtr, c.device, "Jtr");
  if (cost) require_device_ptr(cost, c.device, "cost");
  // cost partials first, then the chunk and group partials: as many chunks as
  // the array holds, up to the allocating entry point's one per CU
  const int64_t avail = int64_t(kNoAllocScratchBytes / sizeof(double)) - kSumsqBlocks;
  const int nt = (D + kMT - 1) / kMT;
  const int64_t per = int64_t(nt * (nt + 1) / 2) * 256 + D;
  const int64_t max_chunks = avail / per - kNeGroups;
  FS_REQUIRE(max_chunks >= 1, FS_ERR_UNSUPPORTED, "normal-equation partials exceed the scratch");
  const NeShape sh = ne_shape(c.cu_count, R, D, int(std::min<int64_t>(max_chunks, 1 << 30)));
  ne_launch(c.stream, sh, J, F, R, D, base + kSumsqBlocks, JtJ, Jtr);
  if (cost) launch_sumsq(c.stream, R, F, base, cost);
  return FS_OK;
}

// columns [a0, a0 + na) and [b0, b0 + nb) of J (R x D) side by side into W
// (R x (na + nb))
__global__ void k_gather_cols(int64_t R, int D, const float *__restrict__ J, int a0, int na, int b0,
                              int nb, float *__restrict__ W) {
  const int w = na + nb;
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= R * w) return;
  const int64_t r = e / w;
  const int k = int(e % w);
  W[e] = J[r * D + (k < na ? a0 + k : b0 + k - na)];
}

// the blocks of M = W^T W (w = na + nb) into J^T J: the diagonal block of
// group a when put_a, of group b when put_b, and the off-diagonal (a, b)
// block with its transpose; J^T F of the groups likewise
__global__ void k_put_blocks(int D, int a0, int na, int b0, int nb, int put_a, int put_b,
                             const double *__restrict__ M, const double *__restrict__ gw,
                             double *__restrict__ A, double *__restrict__ g) {
  const int w = na + nb;
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= int64_t(w) * w) return;
  const int i = int(e / w), j = int(e % w);
  const bool ia = i < na, ja = j < na;
  const int gi = ia ? a0 + i : b0 + i - na, gj = ja ? a0 + j : b0 + j - na;
  if ((ia && ja && put_a) || (!ia && !ja && put_b) || (ia != ja))
    A[int64_t(gi) * D + gj] = M[e];
  if (j == 0 && ((ia && put_a) || (!ia && put_b))) g[gi] = gw[i];
}

// Normal equations of more than 256 columns: the columns in groups of 128;
// for every pair of groups the two side by side are one <= 256-column product
// of the kernels above, whose off-diagonal block is that pair's block of J^T J
// (the diagonal blocks are taken from the first pair that holds them)
void normal_equations_wide(Context &c) {
  const int D = c.D;
  constexpr int kGc = 128;
  const int G = (D + kGc - 1) / kGc;
  const int64_t R = c.R;
  c.ne_W.alloc(size_t(std::max<int64_t>(R, 1)) * 2 * kGc);
  c.ne_M.alloc(size_t(2 * kGc) * 2 * kGc + 2 * kGc);
  const NeShape big = ne_shape(c.cu_count, R, 2 * kGc, 1 << 30);
  c.jtj_part.alloc(size_t(big.nchunk + kNeGroups) * big.per);
  for (int ga = 0; ga < G; ++ga)
    for (int gb = ga + 1; gb < G; ++gb) {
      const int a0 = ga * kGc, na = std::min(kGc, D - a0), b0 = gb * kGc, nb = std::min(kGc, D - b0);
      const int w = na + nb;
      hipLaunchKernelGGL(k_gather_cols, dim3(unsigned((R * w + 255) / 256)), dim3(256), 0, c.stream, R,
                         D, c.J.p, a0, na, b0, nb, c.ne_W.p);
      FS_HIP(hipGetLastError());
      const NeShape sh = ne_shape(c.cu_count, R, w, 1 << 30);
      double *M = c.ne_M.p, *gw = c.ne_M.p + size_t(w) * w;
      ne_launch(c.stream, sh, c.ne_W.p, c.residual.p, R, w, c.jtj_part.p, M, gw);
      const int put_a = gb == ga + 1, put_b = gb == G - 1 && ga == G - 2;
      hipLaunchKernelGGL(k_put_blocks, dim3(unsigned((int64_t(w) * w + 255) / 256)), dim3(256), 0,
                         c.stream, D, a0, na, b0, nb, put_a, put_b, M, gw, c.JtJ.p, c.Jtr.p);
      FS_HIP(hipGetLastError());
    }
}

void launch_normal_equations(Context &c) {
  const int D = c.D;
  c.JtJ.alloc(size_t(std::max(D, 1)) * std::max(D, 1));
  c.Jtr.alloc(std::max(D, 1));
  if (D == 0) return;
  if (D > 16 * kMT) {
    normal_equations_wide(c);
    return;
  }
  const NeShape sh = ne_shape(c.cu_count, c.R, D, 1 << 30);
  c.jtj_part.alloc(size_t(sh.nchunk + kNeGroups) * sh.per);
  ne_launch(c.stream, sh, c.J.p, c.residual.p, c.R, D, c.jtj_part.p, c.JtJ.p, c.Jtr.p);
}

// scal layout: [0] max diag, [1] info, [2] predicted reduction
void launch_diag_scale(Context &c) {
  c.lm_scal.alloc(8);
  c.lm_dx.alloc(2 * size_t(std::max(c.D, 1)));
  hipLaunchKernelGGL(k_diag_scale, dim3(1), dim3(256), 0, c.stream, c.D, c.JtJ.p,
                     c.lm_dx.p + c.D, c.lm_scal.p);
  FS_HIP(hipGetLastError());
}

// the damped solve of k_chol_solve on the host, for more than 256 columns
// (one workgroup's LDS panels hold 256): the same outputs in lm_dx / lm_scal
void damped_solve_host(Context &c, double mu) {
  const int D = c.D;
  std::vector<double> M(size_t(D) * D), g(D), dg(D), dx(D, 0.0);
  FS_HIP(hipMemcpyAsync(M.data(), c.JtJ.p, sizeof(double) * M.size(), hipMemcpyDeviceToHost, c.stream));
  FS_HIP(hipMemcpyAsync(g.data(), c.Jtr.p, sizeof(double) * D, hipMemcpyDeviceToHost, c.stream));
  FS_HIP(hipMemcpyAsync(dg.data(), c.lm_dx.p + D, sizeof(double) * D, hipMemcpyDeviceToHost, c.stream));
  FS_HIP(hipStreamSynchronize(c.stream));
  for (int k = 0; k < D; ++k) M[size_t(k) * D + k] += mu * dg[k];
  // lower Cholesky factor in place, row by row
  double info = 0.0;
  for (int i = 0; i < D && info == 0.0; ++i) {
    double *Ri = &M[size_t(i) * D];
    for (int j = 0; j <= i; ++j) {
      const double *Rj = &M[size_t(j) * D];
      double v = Ri[j];
      for (int k = 0; k < j; ++k) v -= Ri[k] * Rj[k];
      if (j < i) {
        Ri[j] = v / Rj[j];
      } else if (v > 0.0 && std::isfinite(v)) {
        Ri[i] = std::sqrt(v);
      } else {
        info = 1.0;
      }
    }
  }
  double pred = 0.0;
  if (info == 0.0) {
    for (int i = 0; i < D; ++i) { // L y = -g
      double v = -g[i];
      for (int k = 0; k < i; ++k) v -= M[size_t(i) * D + k] * dx[k];
      dx[i] = v / M[size_t(i) * D + i];
    }
    for (int i = D - 1; i >= 0; --i) { // L^T dx = y
      double v = dx[i];
      for (int k = i + 1; k < D; ++k) v -= M[size_t(k) * D + i] * dx[k];
      dx[i] = v / M[size_t(i) * D + i];
    }
    for (int k = 0; k < D; ++k) pred += dx[k] * (mu * dg[k] * dx[k] - g[k]);
  }
  const double sc[2] = {info, pred};
  FS_HIP(hipMemcpyAsync(c.lm_dx.p, dx.data(), sizeof(double) * D, hipMemcpyHostToDevice, c.stream));
  FS_HIP(hipMemcpyAsync(c.lm_scal.p + 1, sc, sizeof(sc), hipMemcpyHostToDevice, c.stream));
  FS_HIP(hipStreamSynchronize(c.stream)); // the host vectors go out of scope
}

void launch_damped_solve(Context &c, double mu) {
  if (c.D > 256) {
    damped_solve_host(c, mu);
    return;
  }
  c.lm_M.alloc(size_t(std::max(c.D, 1)) * std::max(c.D, 1));
  hipLaunchKernelGGL(k_chol_solve, dim3(1), dim3(1024), 0, c.stream, c.D, c.JtJ.p, c.Jtr.p,
                     c.lm_dx.p + c.D, mu, c.lm_M.p, c.lm_dx.p, c.lm_scal.p);
  FS_HIP(hipGetLastError());
}

} // namespace fsa
