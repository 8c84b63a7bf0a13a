// FP64 normal equations and the damped LM solve on the GPU.
//
//   JtJ = J^T J and g = J^T F from the FP32 Jacobian and the FP64 residual,
//   with every product and sum in FP64: split-K partial tiles reduced in a
//   fixed chunk order (deterministic).
//   (JtJ + mu * diag(d)) dx = -g by an in-place Cholesky factorisation and two
//   triangular solves in one workgroup, plus the predicted reduction
//   dx . (mu d dx - g) the LM gain ratio needs.
// The reference's normal-equation step (NormalEquationsSolver::minimize,
// src/optimization/solver.cpp:242-266) forms the same matrices on the CPU.
#include "common.hpp"

#include <algorithm>

namespace fsa {

namespace {

constexpr int kT = 64;     // output tile edge
constexpr int kRB = 32;    // rows per LDS stage

__global__ __launch_bounds__(256) void k_jtj_partial(const float *__restrict__ J, int64_t R, int D,
                                                     int nchunk, const int2 *__restrict__ tiles,
                                                     double *__restrict__ part) {
  __shared__ float sA[kRB][kT];
  __shared__ float sB[kRB][kT];
  const int chunk = blockIdx.x, tp = blockIdx.y;
  const int2 tt = tiles[tp];
  const int64_t rows_per = (R + nchunk - 1) / nchunk;
  const int64_t r0 = int64_t(chunk) * rows_per;
  const int64_t r1 = std::min<int64_t>(R, r0 + rows_per);
  const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
  double acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = 0.0;
  const int ca = tt.x * kT, cb = tt.y * kT;
  for (int64_t rb = r0; rb < r1; rb += kRB) {
    for (int k = threadIdx.x; k < kRB * kT; k += 256) {
      const int rr = k / kT, cc = k % kT;
      const int64_t row = rb + rr;
      const bool ok = row < r1;
      sA[rr][cc] = (ok && ca + cc < D) ? J[row * D + ca + cc] : 0.0f;
      sB[rr][cc] = (ok && cb + cc < D) ? J[row * D + cb + cc] : 0.0f;
    }
    __syncthreads();
#pragma unroll 4
    for (int k = 0; k < kRB; ++k) {
      double av[4], bv[4];
#pragma unroll
      for (int a = 0; a < 4; ++a) av[a] = double(sA[k][4 * ty + a]);
#pragma unroll
      for (int b = 0; b < 4; ++b) bv[b] = double(sB[k][4 * tx + b]);
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] += av[a] * bv[b];
    }
    __syncthreads();
  }
  double *o = part + (int64_t(tp) * nchunk + chunk) * (kT * kT);
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) o[(4 * ty + a) * kT + 4 * tx + b] = acc[a][b];
}

__global__ void k_jtj_reduce(int D, int nchunk, const int2 *__restrict__ tiles,
                             const double *__restrict__ part, double *__restrict__ A) {
  const int tp = blockIdx.x;
  const int2 tt = tiles[tp];
  for (int e = threadIdx.x; e < kT * kT; e += blockDim.x) {
    const int a = e / kT, b = e % kT;
    const int ia = tt.x * kT + a, ib = tt.y * kT + b;
    if (ia >= D || ib >= D) continue;
    double s = 0.0;
    const double *p = part + int64_t(tp) * nchunk * (kT * kT) + e;
    for (int k = 0; k < nchunk; ++k) s += p[int64_t(k) * (kT * kT)];
    A[int64_t(ia) * D + ib] = s;
    A[int64_t(ib) * D + ia] = s;
  }
}

__global__ void k_jtr_partial(const float *__restrict__ J, const double *__restrict__ F, int64_t R,
                              int D, int nchunk, double *__restrict__ part) {
  const int chunk = blockIdx.x;
  const int64_t rows_per = (R + nchunk - 1) / nchunk;
  const int64_t r0 = int64_t(chunk) * rows_per;
  const int64_t r1 = std::min<int64_t>(R, r0 + rows_per);
  for (int col = threadIdx.x; col < D; col += blockDim.x) {
    double s = 0.0;
    for (int64_t r = r0; r < r1; ++r) s += double(J[r * D + col]) * F[r];
    part[int64_t(chunk) * D + col] = s;
  }
}

__global__ void k_jtr_reduce(int D, int nchunk, const double *__restrict__ part,
                             double *__restrict__ g) {
  for (int col = blockIdx.x * blockDim.x + threadIdx.x; col < D; col += gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int k = 0; k < nchunk; ++k) s += part[int64_t(k) * D + col];
    g[col] = s;
  }
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

// (A + mu diag(dg)) dx = -g; scal[1] = info (0 ok), scal[2] = dx.(mu dg dx - g)
__global__ __launch_bounds__(256) void k_chol_solve(int D, const double *__restrict__ A,
                                                    const double *__restrict__ g,
                                                    const double *__restrict__ dg, double mu,
                                                    double *__restrict__ M, double *__restrict__ dx,
                                                    double *__restrict__ scal) {
  __shared__ int bad;
  __shared__ double red[256];
  const int t = threadIdx.x;
  if (t == 0) bad = 0;
  for (int64_t e = t; e < int64_t(D) * D; e += blockDim.x) M[e] = A[e];
  __syncthreads();
  for (int k = t; k < D; k += blockDim.x) {
    M[int64_t(k) * D + k] += mu * dg[k];
    dx[k] = -g[k];
  }
  __syncthreads();
  for (int j = 0; j < D; ++j) {
    if (t == 0) {
      const double v = M[int64_t(j) * D + j];
      if (!(v > 0.0) || !isfinite(v)) bad = 1;
      M[int64_t(j) * D + j] = v > 0.0 ? sqrt(v) : 1.0;
    }
    __syncthreads();
    const double ljj = M[int64_t(j) * D + j];
    for (int i = j + 1 + t; i < D; i += blockDim.x) M[int64_t(i) * D + j] /= ljj;
    __syncthreads();
    // trailing update of the lower triangle
    const int n = D - j - 1;
    for (int64_t e = t; e < int64_t(n) * n; e += blockDim.x) {
      const int ii = j + 1 + int(e / n), kk = j + 1 + int(e % n);
      if (kk <= ii) M[int64_t(ii) * D + kk] -= M[int64_t(ii) * D + j] * M[int64_t(kk) * D + j];
    }
    __syncthreads();
  }
  // L y = b (column sweep)
  for (int j = 0; j < D; ++j) {
    if (t == 0) dx[j] /= M[int64_t(j) * D + j];
    __syncthreads();
    const double yj = dx[j];
    for (int i = j + 1 + t; i < D; i += blockDim.x) dx[i] -= M[int64_t(i) * D + j] * yj;
    __syncthreads();
  }
  // L^T x = y (row sweep from the bottom)
  for (int j = D - 1; j >= 0; --j) {
    if (t == 0) dx[j] /= M[int64_t(j) * D + j];
    __syncthreads();
    const double xj = dx[j];
    for (int i = t; i < j; i += blockDim.x) dx[i] -= M[int64_t(j) * D + i] * xj;
    __syncthreads();
  }
  double s = 0.0;
  for (int k = t; k < D; k += blockDim.x) s += dx[k] * (mu * dg[k] * dx[k] - g[k]);
  red[t] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) red[t] += red[t + w];
    __syncthreads();
  }
  if (t == 0) {
    scal[1] = bad ? 1.0 : 0.0;
    scal[2] = red[0];
  }
}

} // namespace

void launch_normal_equations(Context &c) {
  const int D = c.D;
  const int64_t R = c.R;
  c.JtJ.alloc(size_t(std::max(D, 1)) * std::max(D, 1));
  c.Jtr.alloc(std::max(D, 1));
  if (D == 0) return;
  const int nt = (D + kT - 1) / kT;
  std::vector<int2> tiles;
  for (int a = 0; a < nt; ++a)
    for (int b = a; b < nt; ++b) tiles.push_back(make_int2(a, b));
  const int ntp = int(tiles.size());
  const int nchunk = int(std::max<int64_t>(1, std::min<int64_t>(512, (R + 4095) / 4096)));
  static_assert(sizeof(int2) == 8, "int2");
  c.jtj_part.alloc(size_t(ntp) * nchunk * kT * kT + size_t(nchunk) * D + 2 * ntp);
  int2 *d_tiles = reinterpret_cast<int2 *>(c.jtj_part.p + size_t(ntp) * nchunk * kT * kT +
                                           size_t(nchunk) * D);
  FS_HIP(hipMemcpyAsync(d_tiles, tiles.data(), sizeof(int2) * ntp, hipMemcpyHostToDevice, c.stream));
  hipLaunchKernelGGL(k_jtj_partial, dim3(nchunk, ntp), dim3(256), 0, c.stream, c.J.p, R, D, nchunk,
                     d_tiles, c.jtj_part.p);
  FS_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_jtj_reduce, dim3(ntp), dim3(256), 0, c.stream, D, nchunk, d_tiles,
                     c.jtj_part.p, c.JtJ.p);
  FS_HIP(hipGetLastError());
  double *gpart = c.jtj_part.p + size_t(ntp) * nchunk * kT * kT;
  hipLaunchKernelGGL(k_jtr_partial, dim3(nchunk), dim3(256), 0, c.stream, c.J.p, c.residual.p, R,
                     D, nchunk, gpart);
  FS_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_jtr_reduce, dim3((D + 255) / 256), dim3(256), 0, c.stream, D, nchunk, gpart,
                     c.Jtr.p);
  FS_HIP(hipGetLastError());
}

// scal layout: [0] max diag, [1] info, [2] predicted reduction
void launch_diag_scale(Context &c) {
  c.lm_scal.alloc(8);
  c.lm_dx.alloc(2 * size_t(std::max(c.D, 1)));
  hipLaunchKernelGGL(k_diag_scale, dim3(1), dim3(256), 0, c.stream, c.D, c.JtJ.p,
                     c.lm_dx.p + c.D, c.lm_scal.p);
  FS_HIP(hipGetLastError());
}

void launch_damped_solve(Context &c, double mu) {
  c.lm_M.alloc(size_t(std::max(c.D, 1)) * std::max(c.D, 1));
  hipLaunchKernelGGL(k_chol_solve, dim3(1), dim3(256), 0, c.stream, c.D, c.JtJ.p, c.Jtr.p,
                     c.lm_dx.p + c.D, mu, c.lm_M.p, c.lm_dx.p, c.lm_scal.p);
  FS_HIP(hipGetLastError());
}

} // namespace fsa
