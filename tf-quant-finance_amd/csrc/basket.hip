// Correlated multi-asset GBM basket call with pathwise adjoint Greeks.
//
// Paths follow MultivariateGeometricBrownianMotion._sample_paths
// (tff/models/geometric_brownian_motion/multivariate_geometric_brownian_motion.py:
// 229-282): ln S_i(T) = ln S0_i + (r - sigma_i^2/2) T + sigma_i W_i with
// W = sum_steps sqrt(dt) L z_step, L = chol(corr).  Because L is constant,
// W = L (sum_steps sqrt(dt) z_step): the kernel sums the per-step normals
// and applies L once per path (same value, 1/steps of the matvec work).
// Payoff e^{-rT} (A - K)^+ with A = sum_i w_i S_i(T).  The adjoint of this
// scheme does not depend on the path's intermediate states, so the reverse
// sweep is the closed form
//   dP/dS0_i   = e^{-rT} 1{A>K} w_i S_i / S0_i
//   dP/dsig_i  = e^{-rT} 1{A>K} w_i S_i (W_i - sigma_i T)
//   dP/dr      = -T P + e^{-rT} 1{A>K} T A
// Normals: Philox stream, draw j = step * d + i of path p.
#include "common.h"

using namespace tqf;

namespace {

constexpr int kMaxD = 16;
constexpr int kBlockK = 256;

template <int D>
__global__ void __launch_bounds__(kBlockK)
basket_kernel(const float* __restrict__ prm, int n_steps, float t, float r,
              float strike, int64_t n_paths, uint32_t k0, uint32_t k1,
              uint32_t stream, int greeks, double* __restrict__ partial) {
  // prm: s0[D], sigma[D], weight[D], chol[D*D]
  constexpr int NS = 2 + 2 * D;  // price, rho, delta[D], vega[D]
  __shared__ float sp[3 * D + D * D];
  __shared__ double red[kBlockK / kWave][2 * NS];
  for (int e = threadIdx.x; e < 3 * D + D * D; e += kBlockK) sp[e] = prm[e];
  __syncthreads();
  const float dt = t / (float)n_steps;
  const float sdt = sqrtf(dt);
  const float disc = __expf(-r * t);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int nq = greeks ? NS : 1;
  // per-wave sums in LDS (float64): each round of paths is summed over the
  // wave and added there, so no per-thread accumulators stay live across
  // the step loop (occupancy)
  for (int e = threadIdx.x; e < (kBlockK / kWave) * 2 * NS; e += kBlockK)
    (&red[0][0])[e] = 0.0;
  __syncthreads();
  // rounds are uniform over the block; lanes past n_paths compute a valid
  // path and contribute zero
  for (int64_t p0 = (int64_t)blockIdx.x * kBlockK; p0 < n_paths;
       p0 += (int64_t)gridDim.x * kBlockK) {
    const int64_t p = p0 + threadIdx.x;
    const float valid = p < n_paths ? 1.f : 0.f;
    float zs[D];
#pragma unroll
    for (int i = 0; i < D; ++i) zs[i] = 0.f;
#pragma unroll 1
    for (int st = 0; st < n_steps; ++st) {
      // D normals of this step: draw indices st*D .. st*D + D - 1
#pragma unroll
      for (int i = 0; i < D; i += 4) {
        const int j = st * D + i;  // D % 4 == 0 keeps j % 4 == 0
        const u32x4 w = philox4x32_10(
            u32x4{(uint32_t)p, (uint32_t)(j >> 2), stream, (uint32_t)(p >> 32)},
            k0, k1);
        float z0, z1, z2, z3;
        box_muller(w.x, w.y, z0, z1);
        box_muller(w.z, w.w, z2, z3);
        zs[i] += z0;
        if (i + 1 < D) zs[i + 1] += z1;
        if (i + 2 < D) zs[i + 2] += z2;
        if (i + 3 < D) zs[i + 3] += z3;
      }
    }
    // the parameters are read here, after the steps: an opaque offset keeps
    // their LDS loads from being hoisted out of the path loop, where the
    // D (D + 7) / 2 values would stay live in VGPRs through every step
    int o = 0;
    asm volatile("" : "+s"(o));
    const float* s0 = sp + o;
    const float* sig = s0 + D;
    const float* wt = s0 + 2 * D;
    const float* L = s0 + 3 * D;
    float wv[D], sv[D];
    float a = 0.f;
#pragma unroll
    for (int i = 0; i < D; ++i) {
      float acc_w = 0.f;
#pragma unroll
      for (int c = 0; c <= i; ++c) acc_w += L[i * D + c] * zs[c];
      wv[i] = acc_w * sdt;
      sv[i] = s0[i] * __expf((r - 0.5f * sig[i] * sig[i]) * t + sig[i] * wv[i]);
      a += wt[i] * sv[i];
    }
    const float itm = a > strike ? valid : 0.f;
    const float pay = disc * itm * (a - strike);
    float g[NS];
    g[0] = pay;
    g[1] = -t * pay + disc * itm * t * a;
#pragma unroll
    for (int i = 0; i < D; ++i) {
      g[2 + i] = disc * itm * wt[i] * sv[i] / s0[i];
      g[2 + D + i] = disc * itm * wt[i] * sv[i] * (wv[i] - sig[i] * t);
    }
#pragma unroll
    for (int q = 0; q < NS; ++q) {
      if (q < nq) {
        float v = g[q], v2 = g[q] * g[q];
        for (int off = 32; off > 0; off >>= 1) {
          v += __shfl_xor(v, off, 64);
          v2 += __shfl_xor(v2, off, 64);
        }
        if (lane == 0) {
          red[wave][q] += (double)v;
          red[wave][NS + q] += (double)v2;
        }
      }
    }
  }
  __syncthreads();
  for (int q = threadIdx.x; q < 2 * NS; q += kBlockK) {
    double v = 0;
    for (int w = 0; w < kBlockK / kWave; ++w) v += red[w][q];
    partial[(int64_t)blockIdx.x * 2 * NS + q] = v;
  }
}

// One workgroup per output: each thread sums a strided subset of the block
// partials in order, then a fixed-shape LDS tree (deterministic).
__global__ void __launch_bounds__(256)
basket_finalize_kernel(const double* __restrict__ partial, int n_blocks,
                       int ns, int64_t n_paths, double* __restrict__ out) {
  __shared__ double r1[256], r2[256];
  const int q = blockIdx.x;
  double s1 = 0, s2 = 0;
  for (int b = threadIdx.x; b < n_blocks; b += 256) {
    s1 += partial[(int64_t)b * 2 * ns + q];
    s2 += partial[(int64_t)b * 2 * ns + ns + q];
  }
  r1[threadIdx.x] = s1;
  r2[threadIdx.x] = s2;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      r1[threadIdx.x] += r1[threadIdx.x + w];
      r2[threadIdx.x] += r2[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  const double n = (double)n_paths;
  const double m = r1[0] / n;
  double var = n > 1 ? (r2[0] - n * m * m) / (n - 1) : 0.0;
  out[q] = m;
  out[ns + q] = sqrt((var > 0 ? var : 0) / n);
}

}  // namespace

extern "C" int64_t tqf_basket_workspace_size(int32_t dim, int32_t n_blocks) {
  const int d4 = (dim + 3) / 4 * 4;
  return (int64_t)n_blocks * 2 * (2 + 2 * d4) * 8;
}

extern "C" int tqf_basket_call(const float* params, int32_t dim,
                               int32_t n_steps, double maturity, double rate,
                               double strike, int64_t n_paths, uint64_t seed,
                               uint32_t stream_id, int32_t greeks,
                               int32_t n_blocks, void* workspace, double* out,
                               void* stream) {
  clear_error();
  TQF_REQUIRE(params && workspace && out, "null argument");
  TQF_REQUIRE(dim >= 1 && dim <= kMaxD, "basket supports 1..16 assets");
  TQF_REQUIRE(dim % 4 == 0, "dim must be a multiple of 4 (pad with zero "
              "weights)");
  TQF_REQUIRE(n_steps >= 1 && n_paths >= 2 && n_blocks >= 1, "bad sizes");
  const hipStream_t s = (hipStream_t)stream;
  const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
  double* partial = (double*)workspace;
  const int ns = 2 + 2 * dim;
#define TQF_BASKET(DD)                                                      \
  case DD:                                                                  \
    hipLaunchKernelGGL(basket_kernel<DD>, n_blocks, kBlockK, 0, s, params,  \
                       n_steps, (float)maturity, (float)rate, (float)strike, \
                       n_paths, k0, k1, stream_id, greeks, partial);        \
    break;
  switch (dim) {
    TQF_BASKET(4)
    TQF_BASKET(8)
    TQF_BASKET(12)
    TQF_BASKET(16)
  }
#undef TQF_BASKET
  hipLaunchKernelGGL(basket_finalize_kernel, ns, 256, 0, s, partial, n_blocks,
                     ns, n_paths, out);
  return check_launch("tqf_basket_call");
}
