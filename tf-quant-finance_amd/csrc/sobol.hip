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

// Runs of consecutive indices.  Thread (d, r) produces the 256 indices of
// the aligned run [base + 256 r, base + 256 r + 256) (base = first rounded
// down to 256) in dimension d: the first point directly, each next one by
// one XOR, since (k - 1) ^ k = 2^(ctz(k) + 1) - 1 and so
// x(k) = x(k - 1) ^ p[ctz(k)] with p[t] = v[0] ^ ... ^ v[t].  Inside an
// aligned run, index 32 m + j has ctz(j) < 5 for j > 0 (fixed once the inner
// loop is unrolled) and ctz = 5 + ctz(m) for j = 0.  A wave covers 64
// consecutive dimensions (or 64 / dim runs of all dim dimensions when
// dim < 64), so the stores of one index are adjacent.
constexpr int kRun = 256;

template <class Emit>
__device__ __forceinline__ void sobol_run(const uint32_t* __restrict__ v,
                                          int dim, int64_t n, int64_t first,
                                          const uint32_t* __restrict__ shift,
                                          int dw, int rpw, Emit emit) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int rl = lane / dw, dl = lane - rl * dw;
  if (rl >= rpw) return;
  const int d = blockIdx.y * dw + dl;
  if (d >= dim) return;
  const int64_t base = first & ~(int64_t)(kRun - 1);
  const int64_t run = ((int64_t)blockIdx.x * 4 + wave) * rpw + rl;
  const int64_t k0 = base + run * kRun;
  const int64_t end = first + n;
  if (k0 >= end) return;
  const uint32_t* vd = v + (int64_t)d * 32;
  uint32_t p[8];
  p[0] = vd[0];
#pragma unroll
  for (int t = 1; t < 8; ++t) p[t] = p[t - 1] ^ vd[t];
  uint32_t x = sobol_point(vd, (uint64_t)k0, shift ? shift[d] : 0u);
  for (int m = 0; m < kRun / 32; ++m) {
    const int64_t km = k0 + 32 * m;
    if (km >= end) break;
    if (m > 0) {
      const int c = __builtin_ctz(m);  // 0, 1 or 2
      x ^= c == 0 ? p[5] : c == 1 ? p[6] : p[7];
    }
#pragma unroll
    for (int j = 0; j < 32; ++j) {
      if (j > 0) x ^= p[__builtin_ctz(j)];
      const int64_t k = km + j;
      if (k >= first && k < end) emit(k - first, d, x);
    }
  }
}

// Whole-row tiles for 4-byte outputs of stride dim: workgroup b builds
// rows [base + b R, base + b R + R) (base = first rounded down to 16) of
// every dimension in LDS, as 16-row runs per (dimension, run) item with the
// XOR recurrence of sobol_run, then writes the tile's valid rows as one
// contiguous stretch with W-word stores.  Rows of [N, dim] whose byte length
// is no multiple of the cache line then no longer leave lines half-written
// by different workgroups.
constexpr int kTileRun = 16;

// x of index k in one dimension with the 32 generating columns loaded
// together (no chain of dependent loads)
__device__ __forceinline__ uint32_t sobol_point_par(
    const uint32_t* __restrict__ vd, uint64_t k, uint32_t x) {
  uint32_t w[32];
#pragma unroll
  for (int b = 0; b < 32; ++b) w[b] = vd[b];
#pragma unroll
  for (int b = 0; b < 32; ++b) x ^= ((k >> b) & 1) ? w[b] : 0u;
  return x;
}

template <int W, class Val>
__device__ __forceinline__ void sobol_tile(const uint32_t* __restrict__ v,
                                           int dim, int64_t n, int64_t first,
                                           const uint32_t* __restrict__ shift,
                                           int rows, uint32_t* tile,
                                           uint32_t* __restrict__ out,
                                           Val val) {
  const int64_t base = first & ~(int64_t)(kTileRun - 1);
  const int64_t k_start = base + (int64_t)blockIdx.x * rows;
  const int64_t end = first + n;
  const int nrun = rows / kTileRun;
  for (int item = threadIdx.x; item < dim * nrun; item += blockDim.x) {
    const int r = item / dim, d = item - r * dim;
    const int64_t k0 = k_start + (int64_t)kTileRun * r;
    if (k0 >= end) continue;
    const uint32_t* vd = v + (int64_t)d * 32;
    uint32_t p[4];
    p[0] = vd[0];
#pragma unroll
    for (int t = 1; t < 4; ++t) p[t] = p[t - 1] ^ vd[t];
    uint32_t x = sobol_point_par(vd, (uint64_t)k0, shift ? shift[d] : 0u);
    uint32_t* t = tile + (int64_t)(kTileRun * r) * dim + d;
#pragma unroll
    for (int j = 0; j < kTileRun; ++j) {
      if (j > 0) x ^= p[__builtin_ctz(j)];
      t[j * dim] = val(x, d);
    }
  }
  __syncthreads();
  const int64_t lo = k_start > first ? k_start : first;
  const int64_t hi = k_start + rows < end ? k_start + rows : end;
  if (hi <= lo) return;
  const int64_t cnt = (hi - lo) * dim;  // a multiple of W
  typedef uint32_t vec __attribute__((ext_vector_type(W)));
  const vec* src = (const vec*)(tile + (lo - k_start) * dim);
  vec* dst = (vec*)(out + (lo - first) * dim);
  for (int64_t e = threadIdx.x; e < cnt / W; e += blockDim.x)
    __builtin_nontemporal_store(src[e], &dst[e]);  // streaming: written once
}

template <int W>
__global__ void __launch_bounds__(256)
sobol_points_tile_kernel(const uint32_t* __restrict__ v, int dim,
                         double scale, int64_t n, int64_t first,
                         const uint32_t* __restrict__ shift, int kind,
                         int tent, uint32_t* __restrict__ out, int rows) {
  extern __shared__ uint32_t tile[];
  const float fs = (float)scale;
  sobol_tile<W>(v, dim, n, first, shift, rows, tile, out,
                [&](uint32_t x, int) {
                  if (kind == 0) return x;
                  float u = (float)x * fs;
                  if (tent) u = u < 0.5f ? 2.0f * u : 2.0f * (1.0f - u);
                  return __float_as_uint(u);
                });
}

// Rows per tile: 16-row runs, the tile within 48 KB (three workgroups per
// CU), about 4096 words when the dimension is small.  0: no tile (the
// dimension alone exceeds the budget).
inline int tile_rows(int dim) {
  const int64_t budget = 48 * 1024 / 4;  // words
  if ((int64_t)kTileRun * dim > budget) return 0;
  int64_t runs = 4096 / ((int64_t)kTileRun * dim);
  if (runs < 1) runs = 1;
  if (runs * kTileRun * dim > budget) runs = budget / ((int64_t)kTileRun * dim);
  return (int)(runs * kTileRun);
}

// Lanes to (dimension, run): dw dimensions per wave row, rpw runs per wave.
struct RunGrid {
  int dw, rpw;
  dim3 grid;
};

inline RunGrid run_grid(int dim, int64_t n, int64_t first) {
  RunGrid g;
  g.dw = dim < 64 ? dim : 64;
  g.rpw = 64 / g.dw;
  const int64_t base = first & ~(int64_t)(kRun - 1);
  const int64_t runs = (first + n - base + kRun - 1) / kRun;
  const int64_t per_block = 4 * (int64_t)g.rpw;
  g.grid = dim3((unsigned)((runs + per_block - 1) / per_block),
                (unsigned)((dim + g.dw - 1) / g.dw));
  return g;
}

__global__ void __launch_bounds__(256)
sobol_points_run_kernel(const uint32_t* __restrict__ v, int dim, double scale,
                        int64_t n, int64_t first,
                        const uint32_t* __restrict__ shift, int kind, int tent,
                        void* __restrict__ out, int64_t stride, int dw,
                        int rpw) {
  sobol_run(v, dim, n, first, shift, dw, rpw,
            [&](int64_t i, int d, uint32_t x) {
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
            });
}

template <typename R>
__global__ void __launch_bounds__(256)
sobol_normals_run_kernel(const uint32_t* __restrict__ v, int dim, R scale,
                         int64_t n, int64_t first,
                         const uint32_t* __restrict__ shift, int mode,
                         R* __restrict__ out, int dw, int rpw) {
  constexpr bool kF32 = sizeof(R) == sizeof(float);
  unsigned char* icdf = nullptr;
  if constexpr (kF32) {
    icdf = icdf_lds();
    if (mode == 1) {
      icdf_stage(icdf);
      __syncthreads();
    }
  }
  sobol_run(v, dim, n, first, shift, dw, rpw,
            [&](int64_t i, int d, uint32_t x) {
              R z;
              if (mode == 0)
                z = sobol_ref_normal<R>(x, scale);
              else if constexpr (kF32)
                z = (R)icdf_shifted_f32(x, icdf);
              else
                z = sobol_shifted_normal(x, R(0));
              out[i * dim + d] = z;
            });
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
  if (seq_indices) {
    hipLaunchKernelGGL(sobol_points_kernel, dim3(blocks_for(n * dim)),
                       dim3(256), 0, (hipStream_t)stream, v, dim, scale, n,
                       first_index, seq_indices, shift, out_kind, tent, out,
                       out_stride);
  } else if (out_kind != 2 && out_stride == dim && tile_rows(dim) > 0) {
    // whole-row tiles with the widest store the row length and the output
    // address allow
    const int rows = tile_rows(dim);
    const int64_t base = first_index & ~(int64_t)(kTileRun - 1);
    const unsigned blocks =
        (unsigned)((first_index + n - base + rows - 1) / rows);
    const size_t lds = (size_t)rows * dim * 4;
    const uintptr_t a = (uintptr_t)out;
    if (dim % 4 == 0 && a % 16 == 0)
      hipLaunchKernelGGL(sobol_points_tile_kernel<4>, blocks, 256, lds,
                         (hipStream_t)stream, v, dim, scale, n, first_index,
                         shift, out_kind, tent, (uint32_t*)out, rows);
    else if (dim % 2 == 0 && a % 8 == 0)
      hipLaunchKernelGGL(sobol_points_tile_kernel<2>, blocks, 256, lds,
                         (hipStream_t)stream, v, dim, scale, n, first_index,
                         shift, out_kind, tent, (uint32_t*)out, rows);
    else
      hipLaunchKernelGGL(sobol_points_tile_kernel<1>, blocks, 256, lds,
                         (hipStream_t)stream, v, dim, scale, n, first_index,
                         shift, out_kind, tent, (uint32_t*)out, rows);
  } else {
    const RunGrid g = run_grid(dim, n, first_index);
    hipLaunchKernelGGL(sobol_points_run_kernel, g.grid, dim3(256), 0,
                       (hipStream_t)stream, v, dim, scale, n, first_index,
                       shift, out_kind, tent, out, out_stride, g.dw, g.rpw);
  }
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
  const RunGrid g = run_grid(dim, n, first_index);
  if (dtype == TQF_F32) {
    hipLaunchKernelGGL(sobol_normals_run_kernel<float>, g.grid, dim3(256), 0,
                       (hipStream_t)stream, v, dim, std::ldexp(1.0f, -nd), n,
                       first_index, shift, mode, (float*)out, g.dw, g.rpw);
  } else {
    hipLaunchKernelGGL(sobol_normals_run_kernel<double>, g.grid, dim3(256), 0,
                       (hipStream_t)stream, v, dim, std::ldexp(1.0, -nd), n,
                       first_index, shift, mode, (double*)out, g.dw, g.rpw);
  }
  return check_launch("tqf_sobol_normals");
}
