// Shared device and host helpers of the tqf HIP library (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <string>

#include "../../include/tqf.h"

namespace tqf {

// ---------------------------------------------------------------- errors --

// Sets the calling thread's error message and returns `code`.
int set_error(int code, const char* fmt, ...);
void clear_error();

#define TQF_REQUIRE(cond, ...)                         \
  do {                                                 \
    if (!(cond)) return ::tqf::set_error(TQF_EINVAL, __VA_ARGS__); \
  } while (0)

// Returns 0 or the launch error (positive hipError_t) with a message.
int check_launch(const char* what);

constexpr int kWave = 64;

// Workgroup barrier that orders LDS only: the fences are limited to the
// local address space, so the wave's outstanding global stores are not
// waited for (__syncthreads waits for them).  For loops that stage data in
// LDS and stream it out: the next stage may overwrite the LDS buffer while
// the previous stage's stores are still in flight.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// ------------------------------------------------------------ Philox4x32 --
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11): 10 rounds of
// (hi, lo) = mulhilo(M, c) with a Weyl key schedule.
struct u32x4 {
  uint32_t x, y, z, w;
};

__device__ __forceinline__ u32x4 philox4x32_10(u32x4 c, uint32_t k0,
                                               uint32_t k1) {
  constexpr uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  constexpr uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    // one 32x32->64 multiply per word (v_mad_u64_u32) gives both halves
    const uint64_t p0 = (uint64_t)M0 * c.x;
    const uint64_t p1 = (uint64_t)M1 * c.z;
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    c = u32x4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
    k0 += W0;
    k1 += W1;
  }
  return c;
}

// Box-Muller on one word pair.  float: 24-bit uniforms, u in (0, 1];
// v_sin/v_cos take their argument in turns, so 2*pi*t needs no multiply.
__device__ __forceinline__ void box_muller(uint32_t a, uint32_t b, float& z0,
                                           float& z1) {
  const float u = (float)((a >> 8) + 1u) * 0x1p-24f;
  const float t = (float)(b >> 8) * 0x1p-24f;
  // v_log_f32 is log2: -2 ln u = -2 ln2 log2 u (u >= 2^-24, no denormals)
  const float r =
      __builtin_amdgcn_sqrtf(-1.3862943611198906f * __builtin_amdgcn_logf(u));
  z0 = r * __builtin_amdgcn_cosf(t);
  z1 = r * __builtin_amdgcn_sinf(t);
}

__device__ __forceinline__ void box_muller(uint32_t a, uint32_t b, double& z0,
                                           double& z1) {
  const double u = ((double)a + 1.0) * 0x1p-32;
  const double t = (double)b * 0x1p-32;
  const double r = sqrt(-2.0 * log(u));
  double s, c;
  sincospi(2.0 * t, &s, &c);
  z0 = r * c;
  z1 = r * s;
}

// ---------------------------------------------------------------- erfinv --
// float32: erfinv(x) = x P(w), w = -ln(1 - x^2); P is a degree-8 polynomial
// in w - 2.5 (w < 5) or sqrt(w) - 3 (w >= 5).  The form (the variable w, the
// split at w = 5 and the two shifted polynomial variables) is the one M.
// Giles published for single precision ("Approximating the erfinv
// function", GPU Computing Gems Jade Edition, 2011); the coefficients were
// refitted for this library by least squares against scipy
// (csrc/fit_erfinv.py); max relative error 2.9e-7 over float32 inputs with
// 1 - x^2 rounded once.
__device__ __forceinline__ float erfinv_f32(float x) {
  // w = -ln(1 - x^2) = -ln2 log2(1 - x^2); 1 - x^2 >= 2^-24 here, so the
  // raw v_log_f32 needs no denormal scaling
  float w = -0.69314718055994531f *
            __builtin_amdgcn_logf(__builtin_fmaf(-x, x, 1.0f));
  float p;
  if (w < 5.0f) {
    w = w - 2.5f;
    p = 2.798955610e-08f;
    p = __builtin_fmaf(p, w, 3.430734327e-07f);
    p = __builtin_fmaf(p, w, -3.522236486e-06f);
    p = __builtin_fmaf(p, w, -4.389634286e-06f);
    p = __builtin_fmaf(p, w, 2.185773789e-04f);
    p = __builtin_fmaf(p, w, -1.253729840e-03f);
    p = __builtin_fmaf(p, w, -4.177678720e-03f);
    p = __builtin_fmaf(p, w, 2.466407436e-01f);
    p = __builtin_fmaf(p, w, 1.501409417e+00f);
  } else {
    w = __builtin_amdgcn_sqrtf(w) - 3.0f;
    p = -1.435658875e-04f;
    p = __builtin_fmaf(p, w, 6.800158921e-05f);
    p = __builtin_fmaf(p, w, 1.270877357e-03f);
    p = __builtin_fmaf(p, w, -3.640709435e-03f);
    p = __builtin_fmaf(p, w, 5.775239294e-03f);
    p = __builtin_fmaf(p, w, -7.631052573e-03f);
    p = __builtin_fmaf(p, w, 9.433467592e-03f);
    p = __builtin_fmaf(p, w, 1.001674574e+00f);
    p = __builtin_fmaf(p, w, 2.832977048e+00f);
  }
  return p * x;
}

// float64 uses the device math library's erfinv.
__device__ __forceinline__ double erfinv_any(double x) { return erfinv(x); }
__device__ __forceinline__ float erfinv_any(float x) { return erfinv_f32(x); }

// Normal from a Sobol integer, reference transform (multivariate_normal.py:420):
// u = cast(int) * 2^-nd in the working type, z = erfinv((u - .5) * 2) sqrt 2.
template <typename R>
__device__ __forceinline__ R sobol_ref_normal(uint32_t v, R scale) {
  const R u = (R)v * scale;
  return erfinv_any((u - (R)0.5) * (R)2) * (R)1.4142135623730951;
}

// Engine transform of a 32-digit shifted point in float64: q = 2u - 1
// formed exactly from u = (v + 1/2) 2^-32.  (float32 uses the LDS table of
// normals.h.)
__device__ __forceinline__ double sobol_shifted_normal(uint32_t v, double) {
  const double q = ((double)v + 0.5) * 0x1p-31 - 1.0;
  return erfinv(q) * 1.4142135623730951;
}

// Uniform loads through the constant address space: per-step tables that
// every lane reads at the same index become scalar (s_load) loads.
template <typename T>
__device__ __forceinline__ T ld_uniform(const T* p, int64_t i) {
  return ((const __attribute__((address_space(4))) T*)p)[i];
}

// ------------------------------------------------------------ wave helpers --
__device__ __forceinline__ float read_lane(float v, int lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}
__device__ __forceinline__ double read_lane(double v, int lane) {
  const int64_t b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffff), lane);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
  return __longlong_as_double(((int64_t)hi << 32) | (uint32_t)lo);
}

__device__ __forceinline__ float fma_r(float a, float b, float c) {
  return __builtin_fmaf(a, b, c);
}
__device__ __forceinline__ double fma_r(double a, double b, double c) {
  return __builtin_fma(a, b, c);
}

// max(x, 0) as one integer max on the bit pattern: negative values (and
// -0) have the sign bit set, so they compare below 0; NaN with the sign bit
// clear passes through unchanged.
__device__ __forceinline__ float pos_part(float x) {
  return __int_as_float(max(__float_as_int(x), 0));
}
__device__ __forceinline__ double pos_part(double x) {
  return __longlong_as_double(max(__double_as_longlong(x), 0ll));
}

__device__ __forceinline__ float fast_exp(float x) { return __expf(x); }
__device__ __forceinline__ double fast_exp(double x) { return exp(x); }
__device__ __forceinline__ float fast_sqrt(float x) {
  return __builtin_amdgcn_sqrtf(x);
}
__device__ __forceinline__ double fast_sqrt(double x) { return sqrt(x); }
// a / b: float32 as a times the hardware reciprocal (about 1 ulp each, two
// instructions in place of the correctly rounded division's sequence);
// float64 exact
__device__ __forceinline__ float fast_div(float a, float b) {
  return a * __builtin_amdgcn_rcpf(b);
}
__device__ __forceinline__ double fast_div(double a, double b) {
  return a / b;
}

inline int dtype_size(int dtype) { return dtype == TQF_F64 ? 8 : 4; }

}  // namespace tqf
