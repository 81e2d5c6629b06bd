// Small MLP (6 -> 64 -> 64 -> 1, SiLU) for the Longstaff-Schwartz
// continuation value: forward, backward and Adam as hand-written kernels.
//
// The 6->64 and 64x64 layers run on MFMA (bf16 operands, float32
// accumulation: v_mfma_f32_16x16x32_bf16 wherever two k-steps of 16 pair up,
// i.e. layer 2, dH1, and dW1/dW2 over two groups of samples;
// v_mfma_f32_16x16x16_bf16 for layer 1's single step), the 64->1 layer on
// VALU.  Layer 1 is
// computed transposed, Z1^T = W1^T X^T (inputs padded to k = 16, the bias
// as the accumulator's initial value), so its output lands in the A-operand
// layout layer 2 needs; dW1 = X^T dZ1 is one more MFMA per 16-column tile.
// One wave handles 16 samples per MFMA group and keeps the layer-1 and
// layer-2 weights as MFMA operands in registers:
//   A operand (16 x 16, K chunk c): lane l holds row l % 16, k = c*16 +
//     4*(l/16) + {0..3};   B operand: column l % 16, same k;
//   C (16 x 16 float32): lane l holds column l % 16, rows 4*(l/16) + i.
// So the transposed layer-1 output of lane l (C layout of Z1^T: sample
// l % 16, features c*16 + 4*(l/16) + j) is already the A operand of
// layer 2, and the layer-2
// output in C layout is already the B operand of the weight-gradient
// product dW2 = H1^T dZ2.  Two small LDS transposes remain (H1^T and dZ2 as
// an A operand).
//
// Parameter vector layout (float32, kParams = 4673):
//   W1[6][64] | b1[64] | W2[64][64] (in, out) | b2[64] | w3[64] | b3
#include "common.h"
#include "mlp.h"

namespace tqf {
namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kIn = kMlpIn, kH = 64;
constexpr int kG1Stride = 68;
constexpr int oW1 = 0, oB1 = oW1 + kIn * kH, oW2 = oB1 + kH,
              oB2 = oW2 + kH * kH, oW3 = oB2 + kH, oB3 = oW3 + kH;
static_assert(oB3 + 1 == kMlpParams, "parameter layout");

// float -> bf16, round to nearest even: gfx950's v_cvt_pk_bf16_f32
// converts two values per instruction (in place of a shift-add-round
// sequence and a pack per value)
typedef __bf16 bf16x2h __attribute__((ext_vector_type(2)));

__device__ __forceinline__ bf16x4 pack_bf16(float a, float b, float c,
                                            float d) {
  const bf16x2h lo = {(__bf16)a, (__bf16)b};
  const bf16x2h hi = {(__bf16)c, (__bf16)d};
  const uint32_t ul = __builtin_bit_cast(uint32_t, lo);
  const uint32_t uh = __builtin_bit_cast(uint32_t, hi);
  return bf16x4{(short)(ul & 0xffffu), (short)(ul >> 16),
                (short)(uh & 0xffffu), (short)(uh >> 16)};
}


// logistic sigmoid through the hardware reciprocal (1 ulp; the network runs
// its products in bf16)
__device__ __forceinline__ float sigmoid(float z) {
  return __builtin_amdgcn_rcpf(1.0f + __expf(-z));
}
// SiLU'(z) from z and its sigmoid s
__device__ __forceinline__ float silu_grad_s(float z, float s) {
  return s * (1.0f + z * (1.0f - s));
}

__device__ __forceinline__ f32x4 mfma(bf16x4 a, bf16x4 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// Two k-steps of 16 in one gfx950 v_mfma_f32_16x16x32_bf16: the operands of
// step 0 and step 1 concatenated per lane (A lane l holds k 8(l/16) + j of
// the instruction, here step 0's 4(l/16) + j for j < 4 and step 1's for
// j >= 4; B the same), so the product sums both steps' k.
typedef short s16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
__device__ __forceinline__ f32x4 mfma2(bf16x4 a0, bf16x4 a1, bf16x4 b0,
                                       bf16x4 b1, f32x4 c) {
  const s16x8 a = {a0[0], a0[1], a0[2], a0[3], a1[0], a1[1], a1[2], a1[3]};
  const s16x8 b = {b0[0], b0[1], b0[2], b0[3], b1[0], b1[1], b1[2], b1[3]};
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(
      __builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c, 0, 0,
      0);
}

// Orders this wave's LDS writes and reads (the transpose buffers are private
// to a wave; the waves of a block run different numbers of groups, so no
// workgroup barrier may be used inside the group loop).  LDS instructions of
// one wave execute in order; this keeps the compiler from moving them.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Sum over the 16 lanes sharing l / 16 (one DPP row), by row-local DPP
// moves: quad_perm [1,0,3,2], quad_perm [2,3,0,1], row_half_mirror,
// row_mirror.  Every lane adds the same two partial sums at each step, so
// all 16 lanes end with the same value.
template <int kCtrl>
__device__ __forceinline__ float dpp(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(
      0, __float_as_int(v), kCtrl, 0xf, 0xf, false));
}
__device__ __forceinline__ float sum16(float v) {
  v += dpp<0xb1>(v);
  v += dpp<0x4e>(v);
  v += dpp<0x141>(v);
  v += dpp<0x140>(v);
  return v;
}

// Forward for one group.  Returns z1 (A layout, 16), h1 bf16 A operands,
// z2 (C layout, 16), y (4 samples of the lane's row block).
struct Fwd {
  bf16x4 h1a[4];      // A operands of layer 2: sample l%16, features
                      // c*16 + 4*(l/16) + j
  f32x4 z2[4];        // [n-tile t]: rows 4*(l/16)+i, column t*16 + l%16
  f32x4 s2[4];        // sigmoid(z2), same layout
  float y[4];         // outputs of rows 4*(l/16)+i
};

// g1 (optional, per-wave LDS [sample][kG1Stride]) receives SiLU'(z1) for
// the backward pass, written as soon as it is known so z1 and its sigmoid
// need no registers past layer 1.
__device__ __forceinline__ void forward_group(const float* __restrict__ P,
                                              const bf16x4* w1a,
                                              const bf16x4* w2b, bf16x4 xb,
                                              Fwd& f, float* g1 = nullptr) {
  const int lane = threadIdx.x & 63;
  const int q = lane >> 4, r = lane & 15;
  // layer 1 (MFMA, transposed): tile c of Z1^T = W1^T[c] X^T + b1, A = the
  // W1^T tile (rows features c*16 + r, k inputs 4q + j), B = X^T (column
  // sample r, k inputs 4q + j); the C layout (column sample r, rows
  // features c*16 + 4q + j) is the A layout of layer 2
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    f32x4 z;
#pragma unroll
    for (int j = 0; j < 4; ++j) z[j] = P[oB1 + c * 16 + 4 * q + j];
    z = mfma(w1a[c], xb, z);
    float h[4];
    f32x4 gv;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const float sg = sigmoid(z[j]);
      h[j] = z[j] * sg;
      gv[j] = silu_grad_s(z[j], sg);
    }
    f.h1a[c] = pack_bf16(h[0], h[1], h[2], h[3]);
    if (g1) *(f32x4*)(g1 + r * kG1Stride + c * 16 + 4 * q) = gv;
  }
  // layer 2 (MFMA): Z2[16 x 64] = H1[16 x 64] W2[64 x 64]
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int c = 0; c < 4; c += 2)
      acc = mfma2(f.h1a[c], f.h1a[c + 1], w2b[t * 4 + c], w2b[t * 4 + c + 1],
                  acc);
    const float bb = P[oB2 + t * 16 + r];
#pragma unroll
    for (int i = 0; i < 4; ++i) acc[i] += bb;
    f.z2[t] = acc;
  }
  // layer 3 (VALU + 16-lane sums): y for rows 4q + i
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    float s = 0.f;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const float sg = sigmoid(f.z2[t][i]);
      f.s2[t][i] = sg;
      s += f.z2[t][i] * sg * P[oW3 + t * 16 + r];
    }
    f.y[i] = sum16(s) + P[oB3];
  }
}

// W1^T as layer-1 A operands: tile c, row (feature) c*16 + r, k (input)
// 4q + j; inputs past kIn are zero.
__device__ __forceinline__ void load_w1(const float* __restrict__ P,
                                        bf16x4* w1a) {
  const int lane = threadIdx.x & 63, q = lane >> 4, r = lane & 15;
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    float v[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int k = 4 * q + j;
      v[j] = k < kIn ? P[oW1 + k * kH + c * 16 + r] : 0.f;
    }
    w1a[c] = pack_bf16(v[0], v[1], v[2], v[3]);
  }
}

// X^T as the layer-1 B operand: sample row s (column r of the tile),
// inputs 4q + j (zero past kIn or past the end)
__device__ __forceinline__ void load_x_op(const float* __restrict__ x,
                                          int64_t s, bool ok, float* v) {
  const int q = (threadIdx.x & 63) >> 4;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int k = 4 * q + j;
    v[j] = (ok && k < kIn) ? x[s * kMlpStride + k] : 0.f;
  }
}

// W2 as layer-2 B operands: tile (t, c): column t*16 + r, rows
// c*16 + 4q + {0..3}.  W2 is stored [in][out].
__device__ __forceinline__ void load_w2(const float* __restrict__ P,
                                        bf16x4* w2b) {
  const int lane = threadIdx.x & 63, q = lane >> 4, r = lane & 15;
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      float v[4];
#pragma unroll
      for (int j = 0; j < 4; ++j)
        v[j] = P[oW2 + (c * 16 + 4 * q + j) * kH + t * 16 + r];
      w2b[t * 4 + c] = pack_bf16(v[0], v[1], v[2], v[3]);
    }
}

// W2^T as B operands of dH1 = dZ2 W2^T: tile (t, c) over (features f =
// t*16 + r, k = n = c*16 + 4q + j): value W2[f][n].
__device__ __forceinline__ void load_w2t(const float* __restrict__ P,
                                         bf16x4* w2t) {
  const int lane = threadIdx.x & 63, q = lane >> 4, r = lane & 15;
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      float v[4];
#pragma unroll
      for (int j = 0; j < 4; ++j)
        v[j] = P[oW2 + (t * 16 + r) * kH + c * 16 + 4 * q + j];
      w2t[t * 4 + c] = pack_bf16(v[0], v[1], v[2], v[3]);
    }
}

// Copies the parameter vector into LDS (and clears the gradient row when
// G is given) with every global load of a thread issued before the first
// LDS store: a load-store loop would wait out the memory latency once per
// iteration (about 19 times per thread).
constexpr int kStageIters = (kMlpParams + 255) / 256;

__device__ __forceinline__ void stage_params(const float* __restrict__ params,
                                             float* P, float* G) {
  float v[kStageIters];
#pragma unroll
  for (int k = 0; k < kStageIters; ++k) {
    const int e = threadIdx.x + k * 256;
    v[k] = e < kMlpParams ? params[e] : 0.f;
  }
#pragma unroll
  for (int k = 0; k < kStageIters; ++k) {
    const int e = threadIdx.x + k * 256;
    if (e < kMlpParams) {
      P[e] = v[k];
      if (G) G[e] = 0.f;
    }
  }
}

__global__ void __launch_bounds__(256)
mlp_forward_kernel(const float* __restrict__ params,
                   const float* __restrict__ x, int64_t n,
                   float* __restrict__ out) {
  __shared__ float P[kMlpParams];
  stage_params(params, P, nullptr);
  __syncthreads();
  bf16x4 w1a[4], w2b[16];
  load_w1(P, w1a);
  load_w2(P, w2b);
  const int lane = threadIdx.x & 63, q = lane >> 4, r = lane & 15;
  const int64_t waves = (int64_t)gridDim.x * 4;
  const int64_t wid = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  // the next group's inputs are loaded while this one is processed
  float nx[4];
  auto fetch = [&](int64_t g) {
    const int64_t sa = g * 16 + r;
    load_x_op(x, sa, sa < n, nx);
  };
  fetch(wid);
  for (int64_t g = wid; g * 16 < n; g += waves) {
    const int64_t base = g * 16;
    const bf16x4 xb = pack_bf16(nx[0], nx[1], nx[2], nx[3]);
    fetch(g + waves);
    Fwd f;
    forward_group(P, w1a, w2b, xb, f);
    if (r == 0) {
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int64_t s = base + 4 * q + i;
        if (s < n) out[s] = f.y[i];
      }
    }
  }
}

// Gradient of L = sum_s w_s (y_s - t_s)^2 / norm over [0, n), per-block
// partial sums of the kMlpParams gradient entries (and the loss at the end).
template <int kF>
__global__ void __launch_bounds__(256, kF == 1 ? 2 : 1)
mlp_grad_kernel(const float* __restrict__ params, const float* __restrict__ x,
                const float* __restrict__ tgt, const float* __restrict__ wts,
                int64_t first, int64_t n, float inv_norm,
                float* __restrict__ partial) {
  // LDS: the parameters P (after the group loop: wave 0's gradient row),
  // then the loop buffers, which waves 1-3 reuse for their gradient rows
  // once every wave has left the loop
  //   tr   per-wave transpose buffers (bf16) of the kF groups in flight
  //   g1s  per-wave SiLU'(z1) of the kF groups, [sample][feature] with a
  //        row stride of 68 floats (16-byte rows, the four row groups of a
  //        read in distinct banks)
  //   w2t  the W2^T operands in operand order
  constexpr int kRow = kMlpParams + 1;
  constexpr int kTrBytes = 4 * kF * 16 * 64 * 2;
  constexpr int kG1Bytes = 4 * kF * 16 * kG1Stride * 4;
  constexpr int kW2tBytes = 16 * 64 * 8;
  constexpr int kLoopBytes = kTrBytes + kG1Bytes + kW2tBytes;
  constexpr int kRowsBytes = 3 * kRow * 4;
  constexpr int kPBytes = (kRow * 4 + 15) / 16 * 16;
  __shared__ __attribute__((aligned(16)))
      char smem[kPBytes + (kLoopBytes > kRowsBytes ? kLoopBytes : kRowsBytes)];
  float* P = (float*)smem;
  short (*tr)[kF][16 * 64] = (short (*)[kF][16 * 64])(smem + kPBytes);
  float (*g1s)[kF][16 * kG1Stride] =
      (float (*)[kF][16 * kG1Stride])(smem + kPBytes + kTrBytes);
  bf16x4* w2t_s = (bf16x4*)(smem + kPBytes + kTrBytes + kG1Bytes);
  float* rows = (float*)(smem + kPBytes);
  stage_params(params, P, nullptr);
  __syncthreads();
  const int lane = threadIdx.x & 63, q = lane >> 4, r = lane & 15;
  const int wave = threadIdx.x >> 6;
  bf16x4 w1a[4], w2b[16];
  load_w1(P, w1a);
  load_w2(P, w2b);
  // the W2^T operands stay in LDS in operand order (one 8-byte read per
  // lane and tile in the backward pass), written once by wave 0
  if (wave == 0) {
    bf16x4 w2t[16];
    load_w2t(P, w2t);
#pragma unroll
    for (int k = 0; k < 16; ++k) w2t_s[k * 64 + lane] = w2t[k];
  }
  __syncthreads();
  f32x4 dw2[16], dw1t[4];
#pragma unroll
  for (int e = 0; e < 16; ++e) dw2[e] = f32x4{0.f, 0.f, 0.f, 0.f};
  float db1[4], db2[4], dw3[4], db3 = 0.f, loss = 0.f;
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    db1[t] = db2[t] = dw3[t] = 0.f;
    dw1t[t] = f32x4{0.f, 0.f, 0.f, 0.f};
  }
  const int64_t waves = (int64_t)gridDim.x * 4;
  const int64_t wid = (int64_t)blockIdx.x * 4 + wave;
  const int64_t end = first + n;
  // Two groups in flight per wave, g and g + waves: every stage below runs
  // for both, so the two independent dependency chains hide each other's
  // latencies (one wave per SIMD leaves nothing else to switch to).  Each
  // accumulator takes group g's terms before group g + waves', the order
  // of one group at a time; a partner past the end has zero inputs and
  // weight and adds exact zeros.
  // inputs of a group, loaded one pair ahead: lane l holds, of sample
  // base + l % 16, inputs 4q .. 4q + 3 (the layer-1 B operand), its weight
  // and target (zero past the end), and, of samples base + 4q + j, input
  // l % 16 (the B operand of dW1 = X^T dZ1; zero past kIn or the end)
  float nx[kF][4], nxt[kF][4], nw[kF], nt[kF];
  auto fetch = [&](int u, int64_t g) {
    const int64_t sa = first + g * 16 + r;
    const bool ok = g * 16 < n && sa < end;
    load_x_op(x, sa, ok, nx[u]);
    nw[u] = ok ? wts[sa] : 0.f;
    nt[u] = ok ? tgt[sa] : 0.f;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int64_t sb = first + g * 16 + 4 * q + j;
      nxt[u][j] = (g * 16 < n && sb < end && r < kIn) ? x[sb * kMlpStride + r]
                                                       : 0.f;
    }
  };
#pragma unroll
  for (int u = 0; u < kF; ++u) fetch(u, wid + u * waves);
  for (int64_t g = wid; g * 16 < n; g += kF * waves) {
    bf16x4 xb[kF], xtb[kF];
    float wa[kF], ta[kF];
#pragma unroll
    for (int u = 0; u < kF; ++u) {
      xb[u] = pack_bf16(nx[u][0], nx[u][1], nx[u][2], nx[u][3]);
      xtb[u] = pack_bf16(nxt[u][0], nxt[u][1], nxt[u][2], nxt[u][3]);
      wa[u] = nw[u];
      ta[u] = nt[u];
    }
#pragma unroll
    for (int u = 0; u < kF; ++u) fetch(u, g + (kF + u) * waves);
    Fwd f[kF];
#pragma unroll
    for (int u = 0; u < kF; ++u)
      forward_group(P, w1a, w2b, xb[u], f[u], g1s[wave][u]);
    // dy for the lane's rows
    float dy[kF][4];
#pragma unroll
    for (int u = 0; u < kF; ++u)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int src = 4 * q + i;
        const float w = __shfl(wa[u], src, 64), t = __shfl(ta[u], src, 64);
        const float e = f[u].y[i] - t;
        dy[u][i] = 2.f * w * e * inv_norm;
        if (r == 0) loss += w * e * e * inv_norm;
        db3 += r == 0 ? dy[u][i] : 0.f;
      }
    // layer 3 and SiLU' of layer 2: dZ2 in C layout (rows 4q+i, col t*16+r)
    f32x4 dz2[kF][4];
#pragma unroll
    for (int u = 0; u < kF; ++u)
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const float w3 = P[oW3 + t * 16 + r];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const float z = f[u].z2[t][i];
          const float sg = f[u].s2[t][i];
          dw3[t] += z * sg * dy[u][i];
          const float d = dy[u][i] * w3 * silu_grad_s(z, sg);
          dz2[u][t][i] = d;
          db2[t] += d;
        }
      }
    // dW2 += H1^T dZ2: A = H1^T tile (features mt*16 + r, samples 4q + j)
    // from the transpose of the A-layout h1 via LDS.
#pragma unroll
    for (int u = 0; u < kF; ++u)
#pragma unroll
      for (int c = 0; c < 4; ++c)
#pragma unroll
        for (int j = 0; j < 4; ++j)  // [sample][feature]
          tr[wave][u][r * 64 + c * 16 + 4 * q + j] = f[u].h1a[c][j];
    wave_lds_sync();
    bf16x4 h1t[kF][4], dzb[kF][4];
#pragma unroll
    for (int u = 0; u < kF; ++u) {
#pragma unroll
      for (int mt = 0; mt < 4; ++mt) {
        short v[4];
#pragma unroll
        for (int j = 0; j < 4; ++j)
          v[j] = tr[wave][u][(4 * q + j) * 64 + mt * 16 + r];
        h1t[u][mt] = bf16x4{v[0], v[1], v[2], v[3]};
      }
      // dZ2 as B operand: col t*16+r, k (samples) 4q + i
#pragma unroll
      for (int t = 0; t < 4; ++t)
        dzb[u][t] = pack_bf16(dz2[u][t][0], dz2[u][t][1], dz2[u][t][2],
                              dz2[u][t][3]);
    }
    // the groups in flight in pairs: one k = 32 step over both groups'
    // samples
#pragma unroll
    for (int u = 0; u + 1 < kF; u += 2)
#pragma unroll
      for (int mt = 0; mt < 4; ++mt)
#pragma unroll
        for (int t = 0; t < 4; ++t)
          dw2[mt * 4 + t] = mfma2(h1t[u][mt], h1t[u + 1][mt], dzb[u][t],
                                  dzb[u + 1][t], dw2[mt * 4 + t]);
    if (kF % 2)
#pragma unroll
      for (int mt = 0; mt < 4; ++mt)
#pragma unroll
        for (int t = 0; t < 4; ++t)
          dw2[mt * 4 + t] = mfma(h1t[kF - 1][mt], dzb[kF - 1][t],
                                 dw2[mt * 4 + t]);
    // dH1 = dZ2 W2^T: A = dZ2 (rows = samples r, k = n c*16 + 4q + j) via LDS
    wave_lds_sync();
#pragma unroll
    for (int u = 0; u < kF; ++u)
#pragma unroll
      for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int i = 0; i < 4; ++i)  // [sample][n]
          tr[wave][u][(4 * q + i) * 64 + t * 16 + r] = dzb[u][t][i];
    wave_lds_sync();
    bf16x4 dza[kF][4];
#pragma unroll
    for (int u = 0; u < kF; ++u)
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        short v[4];
#pragma unroll
        for (int j = 0; j < 4; ++j)
          v[j] = tr[wave][u][r * 64 + c * 16 + 4 * q + j];
        dza[u][c] = bf16x4{v[0], v[1], v[2], v[3]};
      }
    wave_lds_sync();
    // dH1 tiles (C layout): rows (samples) 4q + i, features t*16 + r
    f32x4 dh[kF][4];
#pragma unroll
    for (int u = 0; u < kF; ++u)
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int c = 0; c < 4; c += 2)
          acc = mfma2(dza[u][c], dza[u][c + 1],
                      w2t_s[(t * 4 + c) * 64 + lane],
                      w2t_s[(t * 4 + c + 1) * 64 + lane], acc);
        dh[u][t] = acc;
      }
    // dZ1 = dH1 SiLU'(z1) (C layout: rows samples 4q + i, column feature
    // t*16 + r) is the A operand of dW1^T (row feature, k sample): tile t
    // of dW1^T += dZ1^T X
    bf16x4 dzp[kF][4];
#pragma unroll
    for (int u = 0; u < kF; ++u) {
      const float* G1 = g1s[wave][u];
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const int fo = t * 16 + r;
        float d[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          d[i] = dh[u][t][i] * G1[(4 * q + i) * kG1Stride + fo];
          db1[t] += d[i];
        }
        dzp[u][t] = pack_bf16(d[0], d[1], d[2], d[3]);
      }
    }
    // the groups in pairs, as dW2
#pragma unroll
    for (int u = 0; u + 1 < kF; u += 2)
#pragma unroll
      for (int t = 0; t < 4; ++t)
        dw1t[t] = mfma2(dzp[u][t], dzp[u + 1][t], xtb[u], xtb[u + 1],
                        dw1t[t]);
    if (kF % 2)
#pragma unroll
      for (int t = 0; t < 4; ++t)
        dw1t[t] = mfma(dzp[kF - 1][t], xtb[kF - 1], dw1t[t]);
  }
  // block reduction in a fixed order: each wave folds its four lane groups
  // (q = lane / 16, which hold the same features) with shuffles where
  // needed and writes its gradient row (every entry has exactly one owning
  // lane per wave): wave 0 into P, waves 1-3 into the loop buffers, all
  // free once every wave has left the group loop.  The rows are then added
  // in wave order, one entry per thread.
  auto qsum = [](float v) {
    v += __shfl_xor(v, 16, 64);
    v += __shfl_xor(v, 32, 64);
    return v;
  };
  float sb1[4], sb2[4], sw3[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    sb1[t] = qsum(db1[t]);
    sb2[t] = qsum(db2[t]);
    sw3[t] = qsum(dw3[t]);
  }
  // db3 and the loss live on the lanes with r == 0
  const float sb3 = qsum(db3), sl = qsum(loss);
  __syncthreads();
  float* row = wave == 0 ? P : rows + (wave - 1) * kRow;
#pragma unroll
  for (int mt = 0; mt < 4; ++mt)
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int i = 0; i < 4; ++i)
        row[oW2 + (mt * 16 + 4 * q + i) * kH + t * 16 + r] =
            dw2[mt * 4 + t][i];
  // dW1^T tile t (C layout): column input r, rows features t*16 + 4q + i
  if (r < kIn) {
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int i = 0; i < 4; ++i)
        row[oW1 + r * kH + t * 16 + 4 * q + i] = dw1t[t][i];
  }
  if (q == 0) {
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int fo = t * 16 + r;
      row[oB1 + fo] = sb1[t];
      row[oB2 + fo] = sb2[t];
      row[oW3 + fo] = sw3[t];
    }
  }
  if (lane == 0) {
    row[oB3] = sb3;
    row[kMlpParams] = sl;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < kRow; e += 256)
    partial[(int64_t)blockIdx.x * kRow + e] =
        ((P[e] + rows[e]) + rows[kRow + e]) + rows[2 * kRow + e];
}

// The block partials summed and the Adam update in one kernel: workgroup
// x owns parameters [32 x, 32 x + 32) (147 workgroups, so most CUs share
// the read); half-wave q = 2 w + (lane >> 5) of its kRedSlices adds the rows
// of blocks [q per, (q + 1) per) in block order (consecutive lanes read
// consecutive parameters), then the slice sums are added in slice order.
// groups of 16 samples each wave keeps in flight in the gradient kernel
constexpr int kGroupsInFlight = 2;

constexpr int kRedWaves = 16;
constexpr int kRedParams = 32;
constexpr int kRedSlices = kRedWaves * 64 / kRedParams;

__global__ void __launch_bounds__(kRedWaves * 64)
mlp_reduce_adam_kernel(float* __restrict__ params, float* __restrict__ m,
                       float* __restrict__ v,
                       const float* __restrict__ partial, int n_blocks,
                       float lr, float b1, float b2, float eps, int step,
                       float* __restrict__ grad_out,
                       float* __restrict__ loss_out) {
  __shared__ float red[kRedSlices][kRedParams];
  const int c = threadIdx.x % kRedParams, q = threadIdx.x / kRedParams;
  const int e = blockIdx.x * kRedParams + c;
  const int per = (n_blocks + kRedSlices - 1) / kRedSlices;
  const int b0 = q * per;
  const int b1e = b0 + per < n_blocks ? b0 + per : n_blocks;
  float g = 0.f;
  if (e <= kMlpParams) {
#pragma unroll 8
    for (int bb = b0; bb < b1e; ++bb)
      g += partial[(int64_t)bb * (kMlpParams + 1) + e];
  }
  red[q][c] = g;
  __syncthreads();
  if (q != 0 || e > kMlpParams) return;
  g = red[0][c];
#pragma unroll
  for (int k = 1; k < kRedSlices; ++k) g += red[k][c];
  if (e == kMlpParams) {
    if (loss_out) *loss_out = g;
    return;
  }
  if (grad_out) grad_out[e] = g;
  if (lr == 0.f) return;  // gradient only
  const float mm = b1 * m[e] + (1.f - b1) * g;
  const float vv = b2 * v[e] + (1.f - b2) * g * g;
  m[e] = mm;
  v[e] = vv;
  const float mh = mm / (1.f - __powf(b1, (float)step));
  const float vh = vv / (1.f - __powf(b2, (float)step));
  params[e] -= lr * mh / (sqrtf(vh) + eps);
}

}  // namespace

int mlp_forward(const float* params, const float* x, int64_t n, float* out,
                hipStream_t s) {
  const int64_t groups = (n + 15) / 16;
  int blocks = (int)((groups + 3) / 4);
  if (blocks > 4096) blocks = 4096;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(mlp_forward_kernel, blocks, 256, 0, s, params, x, n, out);
  return check_launch("mlp_forward");
}

int mlp_grad_step(float* params, float* m, float* v, const float* x,
                  const float* tgt, const float* w, int64_t first, int64_t n,
                  float inv_norm, float lr, int step, float* partial,
                  int n_blocks, float* grad_out, float* loss_out,
                  hipStream_t s) {
  hipLaunchKernelGGL(mlp_grad_kernel<kGroupsInFlight>, n_blocks, 256, 0, s,
                     params, x, tgt, w, first, n, inv_norm, partial);
  hipLaunchKernelGGL(mlp_reduce_adam_kernel,
                     (kMlpParams + 1 + kRedParams - 1) / kRedParams,
                     kRedWaves * 64, 0, s, params, m, v, partial, n_blocks, lr,
                     0.9f, 0.999f, 1e-8f, step, grad_out, loss_out);
  return check_launch("mlp_grad_step");
}

}  // namespace tqf

using namespace tqf;

extern "C" int tqf_mlp_forward(const float* params, const float* x, int64_t n,
                               float* out, void* stream) {
  clear_error();
  TQF_REQUIRE(params && x && out, "null argument");
  TQF_REQUIRE(n >= 0, "n must be non-negative");
  if (n == 0) return TQF_OK;
  return mlp_forward(params, x, n, out, (hipStream_t)stream);
}

extern "C" int64_t tqf_mlp_workspace_size(int32_t n_blocks) {
  // one row of gradient partial sums (and the loss) per workgroup
  return (int64_t)n_blocks * (kMlpParams + 1) * 4;
}

extern "C" int tqf_mlp_train_step(float* params, float* adam_m, float* adam_v,
                                  const float* x, const float* targets,
                                  const float* weights, int64_t first,
                                  int64_t n, float inv_norm, float lr,
                                  int32_t step, int32_t n_blocks,
                                  void* workspace, float* grad_out,
                                  float* loss_out, void* stream) {
  clear_error();
  TQF_REQUIRE(params && x && targets && weights && workspace, "null argument");
  TQF_REQUIRE(lr == 0.f || (adam_m && adam_v), "Adam state is null");
  TQF_REQUIRE(n >= 1 && first >= 0 && n_blocks >= 1 && step >= 1,
              "bad sizes");
  return mlp_grad_step(params, adam_m, adam_v, x, targets, weights, first, n,
                       inv_norm, lr, step, (float*)workspace, n_blocks,
                       grad_out, loss_out, (hipStream_t)stream);
}
