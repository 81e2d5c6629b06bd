// MLP continuation-value model shared by the MLP kernels and the
// Longstaff-Schwartz driver.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tqf {

constexpr int kMlpIn = 6;        // inputs per sample
constexpr int kMlpStride = 8;    // floats per input row (6 used)
constexpr int kMlpParams = 4673;

int mlp_forward(const float* params, const float* x, int64_t n, float* out,
                hipStream_t s);
int mlp_grad_step(float* params, float* m, float* v, const float* x,
                  const float* tgt, const float* w, int64_t first, int64_t n,
                  float inv_norm, float lr, int step, float* partial,
                  int n_blocks, float* grad_out, float* loss_out,
                  hipStream_t s);

}  // namespace tqf
