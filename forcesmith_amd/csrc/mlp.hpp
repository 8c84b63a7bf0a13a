// MLP surrogate state shared by the DE driver, the C-ABI test hooks and
// the kernels in mlp.hip.
#pragma once

#include "common.hpp"
#include "mlp_seq.hpp"

namespace fsa {

// Device MLP state (the workspace layout and the step order live in
// mlp_seq.hpp, shared with the CPU tests).
struct Mlp : MlpState<DBuf<float>> {
  DBuf<float> ws; // split-K partial tiles of the weight-gradient products
};

void mlp_init(Mlp &m, int D, int Hn);
void mlp_forward(Mlp &m, hipStream_t s, int B, const float *X, float *y_out);
void mlp_backward(Mlp &m, hipStream_t s, int B, const float *X, const float *t);
void mlp_adam(Mlp &m, hipStream_t s, float lr);

} // namespace fsa
