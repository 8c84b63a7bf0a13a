// Host-side step sequence of the MLP surrogate (D -> H -> H -> 1, GELU),
// written once over a backend. The GPU backend (mlp.hip) launches the
// kernels; a CPU backend (tests/mlp_host_harness.cpp) runs host loops over
// malloc'd buffers, so the order of workspace sizing, pointer reads and
// launches that the device runs can be checked under AddressSanitizer.
//
// Buf: a growable buffer with a data pointer `p` and `alloc(count)` that may
// release and reallocate (every pointer read before a growing alloc() is
// stale afterwards). Ops: gemm<TA, TB, EPI>, head, loss, head_back, colsum,
// gelu_back, adam (signatures at the call sites below). No HIP dependency.
#pragma once

#include <cmath>
#include <cstddef>
#include <cstdint>

namespace fsa {

enum { MLP_EPI_STORE = 0, MLP_EPI_BIAS_GELU = 1 };

// Parameters are one flat FP32 buffer W1[H x D], b1[H], W2[H x H], b2[H],
// W3[H], b3[1]; G holds the gradient in the same layout, Mo / Vo the Adam
// moments; Z1, H1, Z2, H2, y, gy, tmp and loss are the per-batch workspace.
template <class Buf> struct MlpState {
  int D = 0, Hn = 0, cap = 0;
  Buf P, G, Mo, Vo;
  Buf Z1, H1, Z2, H2, y, gy, tmp, loss;
  int step = 0;
  size_t n_params() const { return size_t(Hn) * D + Hn + size_t(Hn) * Hn + Hn + Hn + 1; }
  float *W1() { return P.p; }
  float *b1() { return P.p + size_t(Hn) * D; }
  float *W2() { return b1() + Hn; }
  float *b2() { return W2() + size_t(Hn) * Hn; }
  float *W3() { return b2() + Hn; }
  float *b3() { return W3() + Hn; }
  // sizes the workspace for a batch; a batch larger than every earlier one
  // reallocates every workspace buffer
  void ensure(int batch) {
    if (batch <= cap) return;
    cap = batch;
    const size_t bh = size_t(batch) * Hn;
    Z1.alloc(bh);
    H1.alloc(bh);
    Z2.alloc(bh);
    H2.alloc(bh);
    tmp.alloc(bh);
    y.alloc(batch);
    gy.alloc(batch);
    loss.alloc(1);
  }
};

template <class M> void mlp_state_init(M &m, int D, int Hn) {
  m.D = D;
  m.Hn = Hn;
  m.cap = 0;
  const size_t np = m.n_params();
  m.P.alloc(np);
  m.G.alloc(np);
  m.Mo.alloc(np);
  m.Vo.alloc(np);
  m.step = 0;
}

// y_out: caller-owned outputs, or null for the workspace's own m.y, which is
// read only after the workspace is sized for B
template <class Ops, class M>
void mlp_forward_seq(Ops &ops, M &m, int B, const float *X, float *y_out) {
  m.ensure(B);
  ops.template gemm<false, true, MLP_EPI_BIAS_GELU>(B, m.Hn, m.D, X, m.D, m.W1(), m.D, m.Z1.p, m.Hn,
                                                     m.b1(), m.H1.p);
  ops.template gemm<false, true, MLP_EPI_BIAS_GELU>(B, m.Hn, m.Hn, m.H1.p, m.Hn, m.W2(), m.Hn, m.Z2.p,
                                                     m.Hn, m.b2(), m.H2.p);
  ops.head(B, m.Hn, m.H2.p, m.W3(), m.b3(), y_out ? y_out : m.y.p);
}

// the backward pass proper, after a forward pass of the same batch wrote
// m.Z1, m.H1, m.Z2, m.H2 and m.y
template <class Ops, class M>
void mlp_backward_tail_seq(Ops &ops, M &m, int B, const float *X, const float *t) {
  const int Hn = m.Hn, D = m.D;
  ops.loss(B, m.y.p, t, m.loss.p);
  // dZ2 in tmp
  ops.head_back(B, Hn, m.y.p, t, m.W3(), m.Z2.p, m.gy.p, m.tmp.p);
  float *gW1 = m.G.p, *gb1 = gW1 + size_t(Hn) * D, *gW2 = gb1 + Hn, *gb2 = gW2 + size_t(Hn) * Hn,
        *gW3 = gb2 + Hn, *gb3 = gW3 + Hn;
  // gW3 = sum_b gy_b H2_b ; gb3 = sum_b gy_b
  ops.colsum(B, Hn, m.H2.p, m.gy.p, gW3);
  ops.colsum(B, 1, m.gy.p, static_cast<const float *>(nullptr), gb3);
  // gW2 = dZ2^T H1 ; gb2 = colsum dZ2
  ops.template gemm<true, false, MLP_EPI_STORE>(Hn, Hn, B, m.tmp.p, Hn, m.H1.p, Hn, gW2, Hn, nullptr,
                                               nullptr);
  ops.colsum(B, Hn, m.tmp.p, static_cast<const float *>(nullptr), gb2);
  // dH1 = dZ2 W2 (into H2, no longer needed); dZ1 = dH1 * gelu'(Z1)
  ops.template gemm<false, false, MLP_EPI_STORE>(B, Hn, Hn, m.tmp.p, Hn, m.W2(), Hn, m.H2.p, Hn,
                                                nullptr, nullptr);
  ops.gelu_back(int64_t(B) * Hn, m.Z1.p, m.H2.p);
  // gW1 = dZ1^T X ; gb1 = colsum dZ1
  ops.template gemm<true, false, MLP_EPI_STORE>(Hn, D, B, m.H2.p, Hn, X, D, gW1, D, nullptr, nullptr);
  ops.colsum(B, Hn, m.H2.p, static_cast<const float *>(nullptr), gb1);
}

// gradient of mean 0.5 (y - t)^2 over the batch into m.G, the loss into m.loss
template <class Ops, class M>
void mlp_backward_seq(Ops &ops, M &m, int B, const float *X, const float *t) {
  // outputs into m.y: passed as null so that the forward pass takes the
  // pointer after sizing the workspace
  mlp_forward_seq(ops, m, B, X, nullptr);
  mlp_backward_tail_seq(ops, m, B, X, t);
}

template <class Ops, class M> void mlp_adam_seq(Ops &ops, M &m, float lr) {
  ++m.step;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.f - std::pow(b1, float(m.step)), c2 = 1.f - std::pow(b2, float(m.step));
  ops.adam(int64_t(m.n_params()), m.P.p, m.G.p, m.Mo.p, m.Vo.p, lr, b1, b2, c1, c2, 1e-8f);
}

} // namespace fsa
