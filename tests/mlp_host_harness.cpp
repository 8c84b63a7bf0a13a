// CPU replay of the surrogate training schedule of the screened DE driver,
// through the step sequence the device runs (forcesmith_amd/csrc/mlp_seq.hpp),
// with every buffer on the host heap: built with AddressSanitizer, a pointer
// read before a workspace reallocation and written afterwards is reported as
// a heap-use-after-free. Test code only (tests/test_mlp_host.py).
//
//   mlp_host_harness schedule fixed   the product order (exit 0, no report)
//   mlp_host_harness schedule prefix  the former order, in which the backward
//                                     pass read the workspace's output pointer
//                                     before the forward pass sized the
//                                     workspace (ASan reports the write)
//   mlp_host_harness grad             backward against central differences
//
// The schedule is the one of the faulting DE run: D = 42 parameters,
// population 168, 4 candidates per member, hidden width 64, retraining every
// 5 generations on min(samples, 2048)-row minibatches (1008, 1848, then 2048
// rows), screening 672 candidates between retrains.
#include "../forcesmith_amd/csrc/mlp_seq.hpp"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace {

template <class T> struct HostBuf {
  T *p = nullptr;
  size_t cap = 0, n = 0;
  HostBuf() = default;
  HostBuf(const HostBuf &) = delete;
  HostBuf &operator=(const HostBuf &) = delete;
  ~HostBuf() { std::free(p); }
  T *alloc(size_t count) { // same growth rule as the device buffer
    if (count > cap) {
      std::free(p);
      p = nullptr;
      if (count) p = static_cast<T *>(std::malloc(count * sizeof(T)));
      cap = count;
    }
    n = count;
    return p;
  }
};

float gelu(float u) {
  const float k = 0.7978845608028654f;
  return 0.5f * u * (1.0f + std::tanh(k * (u + 0.044715f * u * u * u)));
}
float gelu_d(float u) {
  const float k = 0.7978845608028654f;
  const float th = std::tanh(k * (u + 0.044715f * u * u * u));
  return 0.5f * (1.0f + th) + 0.5f * u * (1.0f - th * th) * k * (1.0f + 3.0f * 0.044715f * u * u);
}

struct CpuOps {
  template <bool TA, bool TB, int EPI>
  void gemm(int M, int N, int K, const float *A, int lda, const float *B, int ldb, float *C, int ldc,
            const float *bias, float *H) {
    for (int m = 0; m < M; ++m)
      for (int n = 0; n < N; ++n) {
        double s = 0.0;
        for (int k = 0; k < K; ++k) {
          const float a = TA ? A[size_t(k) * lda + m] : A[size_t(m) * lda + k];
          const float b = TB ? B[size_t(n) * ldb + k] : B[size_t(k) * ldb + n];
          s += double(a) * b;
        }
        float v = float(s);
        if (EPI == fsa::MLP_EPI_BIAS_GELU) {
          v += bias[n];
          C[size_t(m) * ldc + n] = v;
          H[size_t(m) * ldc + n] = gelu(v);
        } else {
          C[size_t(m) * ldc + n] = v;
        }
      }
  }
  void head(int B, int Hn, const float *H2, const float *W3, const float *b3, float *y) {
    for (int b = 0; b < B; ++b) {
      double s = 0.0;
      for (int k = 0; k < Hn; ++k) s += double(H2[size_t(b) * Hn + k]) * W3[k];
      y[b] = float(s) + b3[0];
    }
  }
  void loss(int B, const float *y, const float *t, float *out) {
    double s = 0.0;
    for (int b = 0; b < B; ++b) s += 0.5 * double(y[b] - t[b]) * (y[b] - t[b]);
    out[0] = float(s / B);
  }
  void head_back(int B, int Hn, const float *y, const float *t, const float *W3, const float *Z2,
                 float *gy, float *dZ2) {
    for (int b = 0; b < B; ++b) {
      const float g = (y[b] - t[b]) / float(B);
      gy[b] = g;
      for (int k = 0; k < Hn; ++k) dZ2[size_t(b) * Hn + k] = g * W3[k] * gelu_d(Z2[size_t(b) * Hn + k]);
    }
  }
  void colsum(int B, int Ncol, const float *A, const float *w, float *out) {
    for (int c = 0; c < Ncol; ++c) {
      double s = 0.0;
      for (int b = 0; b < B; ++b) s += double(A[size_t(b) * Ncol + c]) * (w ? w[b] : 1.0f);
      out[c] = float(s);
    }
  }
  void gelu_back(int64_t n, const float *Z, float *dH) {
    for (int64_t e = 0; e < n; ++e) dH[e] *= gelu_d(Z[e]);
  }
  void adam(int64_t n, float *p, const float *g, float *m, float *v, float lr, float b1, float b2,
            float c1, float c2, float eps) {
    for (int64_t e = 0; e < n; ++e) {
      m[e] = b1 * m[e] + (1.f - b1) * g[e];
      v[e] = b2 * v[e] + (1.f - b2) * g[e] * g[e];
      p[e] -= lr * (m[e] / c1) / (std::sqrt(v[e] / c2) + eps);
    }
  }
};

using HostMlp = fsa::MlpState<HostBuf<float>>;

// the former order of the backward pass: the output pointer taken as an
// argument before the forward pass sizes the workspace
void backward_prefix(CpuOps &ops, HostMlp &m, int B, const float *X, const float *t) {
  fsa::mlp_forward_seq(ops, m, B, X, m.y.p);
  fsa::mlp_backward_tail_seq(ops, m, B, X, t);
}

uint64_t lcg(uint64_t &s) {
  s = s * 6364136223846793005ull + 1442695040888963407ull;
  return s >> 11;
}
float unit(uint64_t &s) { return float(double(lcg(s)) * 0x1.0p-53); }

void init_params(HostMlp &m, uint64_t seed) {
  for (size_t k = 0; k < m.n_params(); ++k) {
    m.P.p[k] = (unit(seed) - 0.5f) * 0.4f;
    m.Mo.p[k] = m.Vo.p[k] = 0.0f;
  }
}

int schedule(bool prefix) {
  const int D = 42, NP = 168, O = 4, H = 64, every = 5, gens = 20, steps = 3;
  CpuOps ops;
  HostMlp m;
  fsa::mlp_state_init(m, D, H);
  init_params(m, 1);
  std::vector<float> X(size_t(2048) * D), t(2048), cand(size_t(O) * NP * D), pred(size_t(O) * NP);
  uint64_t rs = 7;
  long n = NP; // samples stored: the initial population
  bool ready = false;
  for (int g = 1; g <= gens; ++g) {
    if ((g - 1) % every == 0 && n >= 2L * NP) {
      const int B = int(std::min<long>(n, 2048));
      for (int st = 0; st < steps; ++st) {
        for (size_t e = 0; e < size_t(B) * D; ++e) X[e] = 2.0f * unit(rs) - 1.0f;
        for (int b = 0; b < B; ++b) t[b] = unit(rs);
        if (prefix) backward_prefix(ops, m, B, X.data(), t.data());
        else fsa::mlp_backward_seq(ops, m, B, X.data(), t.data());
        fsa::mlp_adam_seq(ops, m, 1e-3f);
      }
      std::printf("generation %d: retrained on %d-row minibatches, loss %.6f\n", g, B, m.loss.p[0]);
      std::fflush(stdout);
      ready = true;
    }
    if (ready) {
      for (auto &v : cand) v = 2.0f * unit(rs) - 1.0f;
      fsa::mlp_forward_seq(ops, m, O * NP, cand.data(), pred.data());
    }
    n += NP;
  }
  std::printf("schedule done\n");
  return 0;
}

// backward against central differences of the loss (double-checked wiring)
int grad() {
  const int D = 5, H = 7, B = 9;
  CpuOps ops;
  HostMlp m;
  fsa::mlp_state_init(m, D, H);
  init_params(m, 3);
  for (size_t k = 0; k < m.n_params(); ++k) m.P.p[k] *= 5.0f;
  std::vector<float> X(size_t(B) * D), t(B);
  uint64_t rs = 11;
  for (auto &v : X) v = 2.0f * unit(rs) - 1.0f;
  for (auto &v : t) v = unit(rs);
  fsa::mlp_backward_seq(ops, m, B, X.data(), t.data());
  std::vector<float> g(m.G.p, m.G.p + m.n_params());
  double worst = 0.0;
  for (size_t k = 0; k < m.n_params(); ++k) {
    const float old = m.P.p[k], h = 1e-2f;
    m.P.p[k] = old + h;
    fsa::mlp_forward_seq(ops, m, B, X.data(), nullptr);
    ops.loss(B, m.y.p, t.data(), m.loss.p);
    const double lp = m.loss.p[0];
    m.P.p[k] = old - h;
    fsa::mlp_forward_seq(ops, m, B, X.data(), nullptr);
    ops.loss(B, m.y.p, t.data(), m.loss.p);
    const double lm = m.loss.p[0];
    m.P.p[k] = old;
    const double fd = (lp - lm) / (2.0 * h);
    worst = std::fmax(worst, std::fabs(fd - g[k]) / (1e-3 + std::fabs(fd)));
  }
  std::printf("gradient worst relative error %.3e\n", worst);
  return worst < 2e-2 ? 0 : 3;
}

} // namespace

int main(int argc, char **argv) {
  if (argc >= 3 && std::strcmp(argv[1], "schedule") == 0)
    return schedule(std::strcmp(argv[2], "prefix") == 0);
  if (argc >= 2 && std::strcmp(argv[1], "grad") == 0) return grad();
  std::fprintf(stderr, "usage: mlp_host_harness schedule fixed|prefix | grad\n");
  return 2;
}
