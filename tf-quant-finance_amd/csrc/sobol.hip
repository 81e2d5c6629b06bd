// Sobol sequences: direction numbers (host), point generation and the
// per-lane low-byte table used by the fused steppers.
//
// A point of sequence index k in dimension d is the XOR of the generating
// columns v[d][b] over the set bits b of k (Joe & Kuo notes; natural, not
// Gray-code, order as tff/math/random_ops/sobol/sobol_impl.py:139-157 and
// tff/math/qmc/digital_net.py:379-399), XOR the digital shift.
#include "common.h"
#include "normals.h"

using namespace tqf;

extern "C" int tqf_sobol_direction_numbers(const int64_t* poly,
                                           const int64_t* minit,
                                           int32_t n_poly, int32_t dim,
                                           uint32_t* m_out) {
  TQF_REQUIRE(dim >= 1, "Dimension must be greater than zero. Supplied %d",
              dim);
  TQF_REQUIRE(dim - 1 <= n_poly,
              "Sobol dimension %d exceeds the %d dimensions of the direction "
              "number table", dim, n_poly + 1);
  TQF_REQUIRE(m_out != nullptr, "m_out is null");
  for (int j = 0; j < 32; ++j) m_out[j] = 1u;  // van der Corput
  for (int k = 0; k < dim - 1; ++k) {
    const uint64_t a = (uint64_t)poly[k];
    int deg = 0;
    while ((a >> (deg + 1)) != 0) ++deg;
    TQF_REQUIRE(deg >= 1 && deg <= 18, "bad polynomial %llu at row %d",
                (unsigned long long)a, k);
    uint64_t m[32];
    for (int j = 0; j < deg; ++j) m[j] = (uint64_t)minit[(int64_t)k * 18 + j];
    for (int j = deg; j < 32; ++j) {
      uint64_t v = m[j - deg];
      for (int i = 0; i < deg; ++i) {
        if ((a >> i) & 1) v ^= m[j - deg + i] << (deg - i);
      }
      m[j] = v;
    }
    for (int j = 0; j < 32; ++j) m_out[(int64_t)(k + 1) * 32 + j] = (uint32_t)m[j];
  }
  return TQF_OK;
}

namespace {

__device__ __forceinline__ uint32_t sobol_point(const uint32_t* __restrict__ v,
                                                uint64_t k, uint32_t x) {
  k &= 0xffffffffull;
  while (k) {
    const int b = __builtin_ctzll(k);
    x ^= v[b];
    k &= k - 1;
  }
  return x;
}

__global__ void __launch_bounds__(256)
sobol_points_kernel(const uint32_t* __restrict__ v, int dim, double scale,
                    int64_t n, int64_t first, const int64_t* __restrict__ seq,
                    const uint32_t* __restrict__ shift, int kind, int tent,
                    void* __restrict__ out, int64_t stride) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n * (int64_t)dim) return;
  const int64_t i = g / dim;
  const int d = (int)(g - i * dim);
  const uint64_t k = seq ? (uint64_t)seq[i] : (uint64_t)(first + i);
  const uint32_t x = sobol_point(v + (int64_t)d * 32, k, shift ? shift[d] : 0u);
  const int64_t o = i * stride + d;
  if (kind == 0) {
    ((uint32_t*)out)[o] = x;
  } else if (kind == 1) {
    float u = (float)x * (float)scale;
    if (tent) u = u < 0.5f ? 2.0f * u : 2.0f * (1.0f - u);
    ((float*)out)[o] = u;
  } else {
    double u = (double)x * scale;
    if (tent) u = u < 0.5 ? 2.0 * u : 2.0 * (1.0 - u);
    ((double*)out)[o] = u;
  }
}

// Entry (d, t) at ((d / group) * 256 + t) * group + d % group.
__global__ void __launch_bounds__(256)
sobol_lo_kernel(const uint32_t* __restrict__ v, int dim, int group,
                uint32_t* __restrict__ lo) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= (int64_t)dim * 256) return;
  const int64_t d = g >> 8;
  const uint32_t t = (uint32_t)(g & 255);
  lo[((d / group) * 256 + t) * group + d % group] =
      sobol_point(v + d * 32, t, 0u);
}

template <typename R>
__global__ void __launch_bounds__(256)
sobol_normals_kernel(const uint32_t* __restrict__ v, int dim, R scale,
                     int64_t n, int64_t first,
                     const uint32_t* __restrict__ shift, int mode,
                     R* __restrict__ out) {
  // float32 shifted transform: the same LDS table as the fused steppers
  constexpr bool kF32 = sizeof(R) == sizeof(float);
  unsigned char* icdf = nullptr;
  if constexpr (kF32) {
    icdf = icdf_lds();
    if (mode == 1) {
      icdf_stage(icdf);
      __syncthreads();
    }
  }
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n * (int64_t)dim) return;
  const int64_t i = g / dim;
  const int d = (int)(g - i * dim);
  const uint32_t x = sobol_point(v + (int64_t)d * 32, (uint64_t)(first + i),
                                 shift ? shift[d] : 0u);
  R z;
  if (mode == 0)
    z = sobol_ref_normal<R>(x, scale);
  else if constexpr (kF32)
    z = (R)icdf_shifted_f32(x, icdf);
  else
    z = sobol_shifted_normal(x, R(0));
  out[g] = z;
}

inline unsigned blocks_for(int64_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

extern "C" int tqf_sobol_points(const uint32_t* v, int32_t dim, int32_t nd,
                                int64_t n, int64_t first_index,
                                const int64_t* seq_indices,
                                const uint32_t* shift, int32_t out_kind,
                                int32_t tent, void* out, int64_t out_stride,
                                void* stream) {
  clear_error();
  TQF_REQUIRE(dim >= 1, "Dimension must be greater than zero. Supplied %d",
              dim);
  TQF_REQUIRE(n >= 0, "num_results must be non-negative, got %lld",
              (long long)n);
  TQF_REQUIRE(nd >= 1 && nd <= 32, "num_digits must be in [1, 32], got %d",
              nd);
  TQF_REQUIRE(first_index >= 0, "skip must be non-negative");
  TQF_REQUIRE(seq_indices != nullptr ||
                  first_index + n <= (int64_t)0xffffffffll + 1,
              "Sobol index range exceeds 32 bits");
  TQF_REQUIRE(out_kind >= 0 && out_kind <= 2, "bad out_kind %d", out_kind);
  TQF_REQUIRE(out_stride >= dim, "out_stride %lld < dim %d",
              (long long)out_stride, dim);
  TQF_REQUIRE(v != nullptr && out != nullptr, "null pointer argument");
  if (n == 0) return TQF_OK;
  const double scale = std::ldexp(1.0, -nd);
  hipLaunchKernelGGL(sobol_points_kernel, dim3(blocks_for(n * dim)), dim3(256),
                     0, (hipStream_t)stream, v, dim, scale, n, first_index,
                     seq_indices, shift, out_kind, tent, out, out_stride);
  return check_launch("tqf_sobol_points");
}

extern "C" int tqf_sobol_lo_table(const uint32_t* v, int32_t dim,
                                  int32_t group, uint32_t* lo, void* stream) {
  clear_error();
  TQF_REQUIRE(dim >= 1, "dim must be positive");
  TQF_REQUIRE(group >= 1 && dim % group == 0,
              "group must divide dim (dim %d, group %d)", dim, group);
  TQF_REQUIRE(v != nullptr && lo != nullptr, "null pointer argument");
  hipLaunchKernelGGL(sobol_lo_kernel, dim3(blocks_for((int64_t)dim * 256)),
                     dim3(256), 0, (hipStream_t)stream, v, dim, group, lo);
  return check_launch("tqf_sobol_lo_table");
}

extern "C" int tqf_sobol_normals(const uint32_t* v, int32_t dim, int32_t nd,
                                 int64_t n, int64_t first_index,
                                 const uint32_t* shift, int32_t mode,
                                 int32_t dtype, void* out, void* stream) {
  clear_error();
  TQF_REQUIRE(dim >= 1, "Dimension must be greater than zero. Supplied %d",
              dim);
  TQF_REQUIRE(n >= 0, "n must be non-negative");
  TQF_REQUIRE(nd >= 1 && nd <= 32, "num_digits must be in [1, 32], got %d",
              nd);
  TQF_REQUIRE(mode == 0 || mode == 1, "bad mode %d", mode);
  TQF_REQUIRE(mode == 0 || nd == 32, "the shifted transform needs nd = 32");
  TQF_REQUIRE(first_index >= 0 && first_index + n <= (int64_t)0xffffffffll + 1,
              "Sobol index range exceeds 32 bits");
  TQF_REQUIRE(dtype == TQF_F32 || dtype == TQF_F64, "bad dtype %d", dtype);
  TQF_REQUIRE(v != nullptr && out != nullptr, "null pointer argument");
  if (n == 0) return TQF_OK;
  const dim3 grid(blocks_for(n * dim));
  if (dtype == TQF_F32) {
    hipLaunchKernelGGL(sobol_normals_kernel<float>, grid, dim3(256), 0,
                       (hipStream_t)stream, v, dim, std::ldexp(1.0f, -nd), n,
                       first_index, shift, mode, (float*)out);
  } else {
    hipLaunchKernelGGL(sobol_normals_kernel<double>, grid, dim3(256), 0,
                       (hipStream_t)stream, v, dim, std::ldexp(1.0, -nd), n,
                       first_index, shift, mode, (double*)out);
  }
  return check_launch("tqf_sobol_normals");
}
