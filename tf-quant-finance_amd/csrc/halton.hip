// Halton sequences (radical inverses in prime bases), plain or randomised
// by digit permutations, as uniforms or as normal draws.
//
// Element k of the sequence is the radical inverse of the integer k + 1 in
// the prime base b_d of each axis, truncated to max_digits[d] base-b digits
// (tff/math/random_ops/halton/halton_impl.py:59-285, indices 1-based at
// :386-408).  Randomised (Owen 2017, :252-285, :288-330): digit j of axis d
// goes through perms[j * radix_sum + offset[d] + digit] -- the zeros past
// the integer's length included, up to max_digits[d] -- and
// zero_correction[d] is added.  The value is accumulated in float64 and
// rounded once to the output type; rows of up to 128 axes are built as
// 64-point tiles with one axis per wave (halton_tile_kernel), others one
// thread per (point, axis).  Digits come from float64 division by the base
// (exact quotients while the index fits in 32 bits; see BaseDiv).
#include "common.h"

using namespace tqf;

namespace {

struct HaltonArgs {
  const int32_t* bases;
  const int32_t* max_digits;
  const int32_t* offsets;
  int dim;
  int64_t n;
  int64_t first;
  const int64_t* seq;
  const int32_t* perms;
  int radix_sum;
  const double* zero_correction;
  int kind;  // 1 float32, 2 float64 uniforms; 3, 4 the same as normals
  void* out;
  int64_t stride;
};

// Division by the axis's base in float64, for k < 2^32 and 2 <= b < 2^13
// (the 1000th prime is 7919): with inv the double just above 1/b, the
// product k * inv lies in [k/b, k/b + 2^-20) (two roundings of at most
// 2^-52 relative on a value below 2^32 / b), and the fraction of k/b is 0
// or at most 1 - 1/b, so the product stays below the next integer and its
// truncation is the quotient; the remainder k - q b is exact.  Four
// instructions in place of the integer division's software sequence.
struct BaseDiv {
  double b, inv;
};

__device__ __forceinline__ BaseDiv base_div(uint32_t b) {
  const double bd = (double)b;
  return {bd, __longlong_as_double(__double_as_longlong(1.0 / bd) + 1)};
}

// One base-b digit of k: c = k mod b, k /= b (float64 while k fits in 32
// bits, the common case; the 64-bit division is a long software sequence).
__device__ __forceinline__ int next_digit(uint64_t& k, uint32_t b,
                                          const BaseDiv& dv) {
  if (k <= 0xffffffffull) {
    const double kd = (double)(uint32_t)k;
    const double q = trunc(kd * dv.inv);
    k = (uint64_t)(uint32_t)q;
    return (int)fma(-q, dv.b, kd);
  }
  const uint64_t q = k / b;
  const int c = (int)(k - q * (uint64_t)b);
  k = q;
  return c;
}

// Element (point i, axis d) as float64: the radical inverse of the 1-based
// index in base b_d (digit-permuted when randomised), plus the axis's zero
// correction.
__device__ __forceinline__ double halton_value(const HaltonArgs& a, int64_t i,
                                              int d, const BaseDiv& dv) {
  uint64_t k = (uint64_t)((a.seq ? a.seq[i] : a.first + i) + 1);
  const uint32_t b = (uint32_t)a.bases[d];
  const int nd = a.max_digits[d];
  // sum_j c_j b^-(j+1) as R / B with the digit-reversed integer
  // R = sum_j c_j b^(m-1-j) and B = b^m over the m digits taken (one
  // rounding instead of one per digit); exact in 64 bits while
  // m * bits(b) <= 63: always for plain points of an index below 2^32
  // (m = the index's digit count), not for every randomised axis, which
  // then adds digit by digit
  const int bits = 32 - __builtin_clz(b);
  double acc = 0.0;
  if (a.perms) {
    const int32_t* pd = a.perms + a.offsets[d];
    if (nd * bits <= 63) {
      uint64_t R = 0, B = 1;
      for (int j = 0; j < nd; ++j) {
        const int c = next_digit(k, b, dv);
        R = R * b + (uint64_t)pd[(int64_t)j * a.radix_sum + c];
        B *= b;
      }
      acc = (double)R / (double)B;
    } else {
      double w = 1.0;
      for (int j = 0; j < nd; ++j) {
        const int c = next_digit(k, b, dv);
        w *= (double)b;
        acc += (double)pd[(int64_t)j * a.radix_sum + c] / w;
      }
    }
  } else if (k <= 0xffffffffull) {
    // R < B <= b k < 2^45: exact in float64, and R / B is the same
    // correctly rounded quotient as from 64-bit integers
    double kd = (double)k, R = 0.0, B = 1.0;
    for (int j = 0; j < nd && kd != 0.0; ++j) {
      const double q = trunc(kd * dv.inv);
      R = fma(R, dv.b, fma(-q, dv.b, kd));
      B *= dv.b;
      kd = q;
    }
    acc = R / B;
  } else {
    double w = 1.0;
    for (int j = 0; j < nd && k; ++j) {
      const int c = next_digit(k, b, dv);
      w *= (double)b;
      acc += (double)c / w;
    }
  }
  if (a.zero_correction) acc += a.zero_correction[d];
  return acc;
}

// The output value of kind 1 / 3 (float32 uniform / normal) or 2 / 4
// (float64): the reference's transform for normals, erfinv((u - .5) 2) sqrt 2.
template <typename T>
__device__ __forceinline__ T halton_out(int kind, double acc) {
  if constexpr (sizeof(T) == sizeof(float)) {
    const float u = (float)acc;
    return kind == 1 ? u : erfinv_f32((u - 0.5f) * 2.0f) * 1.41421356f;
  } else {
    return kind == 2 ? acc : erfinv((acc - 0.5) * 2.0) * 1.4142135623730951;
  }
}

__global__ void __launch_bounds__(256) halton_kernel(HaltonArgs a) {
  // (point, axis) of this thread: 32-bit arithmetic when the element
  // count fits
  int64_t i;
  int d;
  if (a.n * (int64_t)a.dim <= 0xffffffffll - 256) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if ((int64_t)g >= a.n * (int64_t)a.dim) return;
    const uint32_t i32 = g / (uint32_t)a.dim;
    i = i32;
    d = (int)(g - i32 * (uint32_t)a.dim);
  } else {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= a.n * (int64_t)a.dim) return;
    i = g / a.dim;
    d = (int)(g - i * a.dim);
  }
  const double acc = halton_value(a, i, d, base_div((uint32_t)a.bases[d]));
  const int64_t o = i * a.stride + d;
  if (a.kind == 1 || a.kind == 3)
    ((float*)a.out)[o] = halton_out<float>(a.kind, acc);
  else
    ((double*)a.out)[o] = halton_out<double>(a.kind, acc);
}

// Tiles of 64 consecutive points x every axis (dim <= kHaltonTileDim, rows
// contiguous): lane p of a wave takes point p of one axis at a time, so a
// wave's 64 elements share the base and the digit loop runs the same count
// on every lane (one thread per element mixed bases in a wave and ran the
// smallest base's count, about 20 digits at 2^20 points); the tile goes
// through LDS and out as one contiguous stretch of 64 rows.  The element
// values are halton_value's.
constexpr int kHaltonTileDim = 128;

template <typename T>
__global__ void __launch_bounds__(256) halton_tile_kernel(HaltonArgs a) {
  extern __shared__ unsigned char smem[];
  T* tile = (T*)smem;  // [64][dim]
  // each axis's base and reciprocal, formed once per workgroup
  __shared__ double sb[kHaltonTileDim], sinv[kHaltonTileDim];
  if ((int)threadIdx.x < a.dim) {
    const BaseDiv dv = base_div((uint32_t)a.bases[threadIdx.x]);
    sb[threadIdx.x] = dv.b;
    sinv[threadIdx.x] = dv.inv;
  }
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int64_t i0 = (int64_t)blockIdx.x * 64;
  const int64_t i = i0 + lane;
  const bool valid = i < a.n;
  for (int d = wave; d < a.dim; d += 4)
    tile[lane * a.dim + d] =
        valid ? halton_out<T>(a.kind, halton_value(a, i, d, {sb[d], sinv[d]}))
              : (T)0;
  __syncthreads();
  int64_t np = a.n - i0;
  if (np > 64) np = 64;
  const int cnt = (int)np * a.dim;
  T* out = (T*)a.out + i0 * a.dim;  // rows contiguous (stride == dim)
  for (int e = threadIdx.x; e < cnt; e += 256) out[e] = tile[e];
}

}  // namespace

extern "C" int tqf_halton_points(const int32_t* bases,
                                 const int32_t* max_digits,
                                 const int32_t* offsets, int32_t dim,
                                 int64_t n, int64_t first_index,
                                 const int64_t* seq_indices,
                                 const int32_t* perms, int32_t radix_sum,
                                 const double* zero_correction,
                                 int32_t out_kind, void* out,
                                 int64_t out_stride, void* stream) {
  clear_error();
  TQF_REQUIRE(dim >= 1 && dim <= 1000,
              "`dim` should be in [1, 1000], got %d", dim);
  TQF_REQUIRE(n >= 0 && first_index >= 0, "bad sizes");
  TQF_REQUIRE(bases && max_digits && out, "null argument");
  TQF_REQUIRE(perms == nullptr || (offsets != nullptr && radix_sum >= 2),
              "randomised Halton needs offsets and radix_sum");
  TQF_REQUIRE(out_kind >= 1 && out_kind <= 4, "bad out_kind %d", out_kind);
  TQF_REQUIRE(out_stride >= dim, "out_stride < dim");
  if (n == 0) return TQF_OK;
  HaltonArgs a{bases, max_digits, offsets, dim, n, first_index, seq_indices,
               perms, radix_sum, zero_correction, out_kind, out, out_stride};
  const int64_t total = n * (int64_t)dim;
  if (out_stride == dim && dim <= kHaltonTileDim) {
    const unsigned blocks = (unsigned)((n + 63) / 64);
    if (out_kind == 1 || out_kind == 3)
      hipLaunchKernelGGL(halton_tile_kernel<float>, dim3(blocks), dim3(256),
                         64 * dim * sizeof(float), (hipStream_t)stream, a);
    else
      hipLaunchKernelGGL(halton_tile_kernel<double>, dim3(blocks), dim3(256),
                         64 * dim * sizeof(double), (hipStream_t)stream, a);
  } else {
    hipLaunchKernelGGL(halton_kernel, dim3((unsigned)((total + 255) / 256)),
                       dim3(256), 0, (hipStream_t)stream, a);
  }
  return check_launch("tqf_halton_points");
}
