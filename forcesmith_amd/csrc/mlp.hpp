// MLP surrogate state shared by the DE driver, the C-ABI test hooks and
// the kernels in mlp.hip.
#pragma once

#include "common.hpp"

namespace fsa {

// ---------------------------------------------------------------------------
// Device MLP state and host-callable steps (used by the DE driver and the
// C-ABI test hooks). Parameters are one flat FP32 buffer:
//   W1[H x D], b1[H], W2[H x H], b2[H], W3[H], b3[1]
// ---------------------------------------------------------------------------
struct Mlp {
  int D = 0, Hn = 0, cap = 0;
  DBuf<float> P, G, Mo, Vo;          // params, grads, Adam moments
  DBuf<float> Z1, H1, Z2, H2, y, gy, tmp, loss;
  int step = 0;
  size_t n_params() const { return size_t(Hn) * D + Hn + size_t(Hn) * Hn + Hn + Hn + 1; }
  float *W1() { return P.p; }
  float *b1() { return P.p + size_t(Hn) * D; }
  float *W2() { return b1() + Hn; }
  float *b2() { return W2() + size_t(Hn) * Hn; }
  float *W3() { return b2() + Hn; }
  float *b3() { return W3() + Hn; }
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


void mlp_init(Mlp &m, int D, int Hn);
void mlp_forward(Mlp &m, hipStream_t s, int B, const float *X, float *y_out);
void mlp_backward(Mlp &m, hipStream_t s, int B, const float *X, const float *t);
void mlp_adam(Mlp &m, hipStream_t s, float lr);

} // namespace fsa
