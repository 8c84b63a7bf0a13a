// MLP surrogate for differential evolution: D -> H -> H -> 1, GELU (tanh
// form), BF16 matrix-core products with FP32 accumulation, FP32 master
// weights and Adam. No counterpart in the reference: the surrogate screens DE
// trial vectors; the exact GPU cost still judges every accepted member.
//
// GEMM: C[M x N] = op(A)[M x K] * op(B)[K x N]; a 256-thread block owns a
// 64 x 64 tile of C, each wave a 32 x 32 quarter as 2 x 2 tiles of
// v_mfma_f32_16x16x32_bf16. A and B tiles are converted to BF16 while staged
// in LDS (A row-major [64][32], B transposed [64][32]) so every lane reads its
// eight consecutive K values with one 16-byte LDS load.
#include "common.hpp"
#include "mlp.hpp"

#include <algorithm>
#include <cmath>

namespace fsa {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

namespace {

constexpr int kBM = 64, kBN = 64, kBK = 32;
constexpr int EPI_STORE = MLP_EPI_STORE, EPI_BIAS_GELU = MLP_EPI_BIAS_GELU;

__device__ inline float gelu_f(float u) {
  const float k = 0.7978845608028654f;
  return 0.5f * u * (1.0f + tanhf(k * (u + 0.044715f * u * u * u)));
}
__device__ inline float gelu_df(float u) {
  const float k = 0.7978845608028654f;
  const float th = tanhf(k * (u + 0.044715f * u * u * u));
  return 0.5f * (1.0f + th) + 0.5f * u * (1.0f - th * th) * k * (1.0f + 3.0f * 0.044715f * u * u);
}

template <bool TA, bool TB, int EPI>
__global__ __launch_bounds__(256) void k_gemm_bf16(int M, int N, int K, const float *__restrict__ A,
                                                   int lda, const float *__restrict__ B, int ldb,
                                                   float *__restrict__ C, int ldc,
                                                   const float *__restrict__ bias,
                                                   float *__restrict__ H, int kc,
                                                   int64_t zstride) {
  __shared__ __attribute__((aligned(16))) __bf16 sA[kBM][kBK + 8];
  __shared__ __attribute__((aligned(16))) __bf16 sB[kBN][kBK + 8];
  const int m0 = blockIdx.y * kBM, n0 = blockIdx.x * kBN;
  // split K: block z sums k in [z kc, (z + 1) kc) into the partial tile z
  const int kb = int(blockIdx.z) * kc, ke = min(K, kb + kc);
  C += int64_t(blockIdx.z) * zstride;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  f32x4 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int k0 = kb; k0 < ke; k0 += kBK) {
    for (int e = threadIdx.x; e < kBM * kBK; e += 256) {
      const int r = e / kBK, kk = e % kBK;
      const int m = m0 + r, k = k0 + kk;
      float v = 0.f;
      if (m < M && k < ke) v = TA ? A[int64_t(k) * lda + m] : A[int64_t(m) * lda + k];
      sA[r][kk] = __bf16(v);
    }
    for (int e = threadIdx.x; e < kBN * kBK; e += 256) {
      const int c = e / kBK, kk = e % kBK;
      const int n = n0 + c, k = k0 + kk;
      float v = 0.f;
      if (n < N && k < ke) v = TB ? B[int64_t(n) * ldb + k] : B[int64_t(k) * ldb + n];
      sB[c][kk] = __bf16(v);
    }
    __syncthreads();
    const int kq = 8 * (lane >> 4);
#pragma unroll
    for (int a = 0; a < 2; ++a) {
      const bf16x8 fa = *reinterpret_cast<const bf16x8 *>(&sA[wm + 16 * a + (lane & 15)][kq]);
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        const bf16x8 fb = *reinterpret_cast<const bf16x8 *>(&sB[wn + 16 * b + (lane & 15)][kq]);
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa, fb, acc[a][b], 0, 0, 0);
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int m = m0 + wm + 16 * a + 4 * (lane >> 4) + i;
        const int n = n0 + wn + 16 * b + (lane & 15);
        if (m >= M || n >= N) continue;
        float v = acc[a][b][i];
        if (EPI == EPI_BIAS_GELU) {
          v += bias[n];
          C[int64_t(m) * ldc + n] = v;
          H[int64_t(m) * ldc + n] = gelu_f(v);
        } else {
          C[int64_t(m) * ldc + n] = v;
        }
      }
}

// C = the sum of the nz split-K partial tiles, in partial order
__global__ void k_splitk_sum(int M, int N, int nz, const float *__restrict__ P, int64_t zstride,
                             float *__restrict__ C, int ldc) {
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= int64_t(M) * N) return;
  const int m = int(e / N), n = int(e % N);
  float s = 0.f;
  for (int z = 0; z < nz; ++z) s += P[z * zstride + int64_t(m) * N + n];
  C[int64_t(m) * ldc + n] = s;
}

// y[b] = H2[b] . W3 + b3  (one wave per row)
__global__ void k_head(int B, int Hn, const float *__restrict__ H2, const float *__restrict__ W3,
                       const float *__restrict__ b3, float *__restrict__ y) {
  const int row = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (row >= B) return;
  float s = 0.f;
  for (int k = lane; k < Hn; k += 64) s += H2[int64_t(row) * Hn + k] * W3[k];
  for (int m = 32; m > 0; m >>= 1) s += __shfl_xor(s, m, 64);
  if (lane == 0) y[row] = s + b3[0];
}

// gy = (y - t)/B; dZ2 = gy * W3 * gelu'(Z2); loss partial
__global__ void k_head_back(int B, int Hn, const float *__restrict__ y, const float *__restrict__ t,
                            const float *__restrict__ W3, const float *__restrict__ Z2,
                            float *__restrict__ gy, float *__restrict__ dZ2) {
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= int64_t(B) * Hn) return;
  const int b = int(e / Hn), k = int(e % Hn);
  const float g = (y[b] - t[b]) / float(B);
  if (k == 0) gy[b] = g;
  dZ2[e] = g * W3[k] * gelu_df(Z2[e]);
}

__global__ void k_gelu_back(int64_t n, const float *__restrict__ Z, float *__restrict__ dH) {
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e < n) dH[e] *= gelu_df(Z[e]);
}

// column sums over rows: out[c] = sum_b A[b*ld + c] (* w[b] if w). A block
// takes 64 columns; its 16 row groups sum rows g, g + 16, ... in order, and
// the group partials are added in group order: a fixed order, 16 independent
// load streams per column instead of one serial walk of the rows
constexpr int kColGroups = 16;
__global__ __launch_bounds__(64 * kColGroups) void k_colsum(int B, int Ncol,
                                                            const float *__restrict__ A,
                                                            const float *__restrict__ w,
                                                            float *__restrict__ out) {
  __shared__ float part[kColGroups][64];
  const int cl = threadIdx.x & 63, g = threadIdx.x >> 6;
  const int c = blockIdx.x * 64 + cl;
  float s = 0.f;
  if (c < Ncol)
    for (int b = g; b < B; b += kColGroups) s += A[int64_t(b) * Ncol + c] * (w ? w[b] : 1.0f);
  part[g][cl] = s;
  __syncthreads();
  if (g == 0 && c < Ncol) {
    float t = 0.f;
    for (int q = 0; q < kColGroups; ++q) t += part[q][cl];
    out[c] = t;
  }
}

__global__ void k_loss(int B, const float *__restrict__ y, const float *__restrict__ t,
                       float *__restrict__ out) {
  __shared__ float s[256];
  float a = 0.f;
  for (int b = threadIdx.x; b < B; b += blockDim.x) {
    const float r = y[b] - t[b];
    a += 0.5f * r * r;
  }
  s[threadIdx.x] = a;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = s[0] / float(B);
}

__global__ void k_adam(int64_t n, float *__restrict__ p, const float *__restrict__ g,
                       float *__restrict__ m, float *__restrict__ v, float lr, float b1, float b2,
                       float c1, float c2, float eps) {
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const float gg = g[e];
  m[e] = b1 * m[e] + (1.f - b1) * gg;
  v[e] = b2 * v[e] + (1.f - b2) * gg * gg;
  p[e] -= lr * (m[e] / c1) / (sqrtf(v[e] / c2) + eps);
}

inline unsigned cdiv(int64_t a, int b) { return unsigned((a + b - 1) / b); }

// the kernels behind the step sequence of mlp_seq.hpp, on one stream
struct GpuOps {
  hipStream_t s;
  DBuf<float> *ws; // split-K partial tiles
  // a product with few output tiles and a long K (the weight gradients:
  // K = the batch) is split along K over enough blocks to fill the GPU; the
  // partial tiles are summed in a fixed order
  template <bool TA, bool TB, int EPI>
  void gemm(int M, int N, int K, const float *A, int lda, const float *B, int ldb, float *C, int ldc,
            const float *bias, float *H) {
    if (M <= 0 || N <= 0) return;
    dim3 grid(cdiv(N, kBN), cdiv(M, kBM));
    const int tiles = int(grid.x * grid.y);
    int nz = 1;
    if (EPI == EPI_STORE && ws && tiles < 128)
      nz = std::max(1, std::min(256 / tiles, K / (4 * kBK)));
    if (nz <= 1) {
      hipLaunchKernelGGL((k_gemm_bf16<TA, TB, EPI>), grid, dim3(256), 0, s, M, N, K, A, lda, B, ldb, C,
                         ldc, bias, H, K, int64_t(0));
      FS_HIP(hipGetLastError());
      return;
    }
    const int kc = int(cdiv(cdiv(K, nz), kBK)) * kBK;
    nz = int(cdiv(K, kc));
    const int64_t zs = int64_t(M) * N;
    ws->alloc(size_t(nz) * size_t(zs));
    grid.z = unsigned(nz);
    hipLaunchKernelGGL((k_gemm_bf16<TA, TB, EPI>), grid, dim3(256), 0, s, M, N, K, A, lda, B, ldb,
                       ws->p, N, bias, H, kc, zs);
    FS_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_splitk_sum, dim3(cdiv(zs, 256)), dim3(256), 0, s, M, N, nz, ws->p, zs, C, ldc);
    FS_HIP(hipGetLastError());
  }
  void head(int B, int Hn, const float *H2, const float *W3, const float *b3, float *y) {
    hipLaunchKernelGGL(k_head, dim3(cdiv(B, 4)), dim3(256), 0, s, B, Hn, H2, W3, b3, y);
    FS_HIP(hipGetLastError());
  }
  void loss(int B, const float *y, const float *t, float *out) {
    hipLaunchKernelGGL(k_loss, dim3(1), dim3(256), 0, s, B, y, t, out);
    FS_HIP(hipGetLastError());
  }
  void head_back(int B, int Hn, const float *y, const float *t, const float *W3, const float *Z2,
                 float *gy, float *dZ2) {
    hipLaunchKernelGGL(k_head_back, dim3(cdiv(int64_t(B) * Hn, 256)), dim3(256), 0, s, B, Hn, y, t, W3,
                       Z2, gy, dZ2);
    FS_HIP(hipGetLastError());
  }
  void colsum(int B, int Ncol, const float *A, const float *w, float *out) {
    hipLaunchKernelGGL(k_colsum, dim3(cdiv(Ncol, 64)), dim3(64 * kColGroups), 0, s, B, Ncol, A, w,
                       out);
    FS_HIP(hipGetLastError());
  }
  void gelu_back(int64_t n, const float *Z, float *dH) {
    hipLaunchKernelGGL(k_gelu_back, dim3(cdiv(n, 256)), dim3(256), 0, s, n, Z, dH);
    FS_HIP(hipGetLastError());
  }
  void adam(int64_t n, float *p, const float *g, float *m, float *v, float lr, float b1, float b2,
            float c1, float c2, float eps) {
    hipLaunchKernelGGL(k_adam, dim3(cdiv(n, 256)), dim3(256), 0, s, n, p, g, m, v, lr, b1, b2, c1, c2,
                       eps);
    FS_HIP(hipGetLastError());
  }
};

} // namespace

void mlp_init(Mlp &m, int D, int Hn) { mlp_state_init(m, D, Hn); }

void mlp_forward(Mlp &m, hipStream_t s, int B, const float *X, float *y_out) {
  GpuOps ops{s, &m.ws};
  mlp_forward_seq(ops, m, B, X, y_out);
}

void mlp_backward(Mlp &m, hipStream_t s, int B, const float *X, const float *t) {
  GpuOps ops{s, &m.ws};
  mlp_backward_seq(ops, m, B, X, t);
}

void mlp_adam(Mlp &m, hipStream_t s, float lr) {
  GpuOps ops{s, &m.ws};
  mlp_adam_seq(ops, m, lr);
}

} // namespace fsa
