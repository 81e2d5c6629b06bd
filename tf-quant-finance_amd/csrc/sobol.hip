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

// Whole-row tiles for 4-byte outputs of stride dim: tile T holds rows
// [base + T R, base + T R + R) (base = first rounded down to 16) of every
// dimension, built in LDS as 16-row runs per (dimension, run) item with the
// XOR recurrence of sobol_run, then written as one contiguous stretch of
// the tile's valid rows with W-word stores.  Rows of [N, dim] whose byte
// length is no multiple of the cache line then no longer leave lines
// half-written by different workgroups.  The grid is persistent (a few
// workgroups per CU), workgroup b taking a contiguous range of tiles: an
// item's first point is formed from all 32 generating columns once, and
// each later tile's from the previous one by the columns of the index bits
// that change (x(k') = x(k) ^ XOR of v_i over the bits of k ^ k'), about
// two loads where a fresh start would read the whole 128-byte column set
// (per lane a separate line).
constexpr int kTileRun = 16;
constexpr int kTileItems = 3;  // items per thread: dim * runs <= 768
constexpr int kTileBits = 4;   // changed index bits loaded together

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
                                           int rows, int64_t tiles_per_block,
                                           uint32_t* tile,
                                           uint32_t* __restrict__ out,
                                           Val val) {
  const int64_t base = first & ~(int64_t)(kTileRun - 1);
  const int64_t end = first + n;
  const int64_t n_tiles = (end - base + rows - 1) / rows;
  const int64_t t_begin = (int64_t)blockIdx.x * tiles_per_block;
  int64_t t_end = t_begin + tiles_per_block;
  if (t_end > n_tiles) t_end = n_tiles;
  const int nrun = rows / kTileRun;
  const int n_items = dim * nrun;  // <= kTileItems * blockDim.x (host)
  // per item u of this thread: the run's start point and the prefix XORs
  // p[t] of the first four columns (the steps inside a 16-row run)
  uint32_t xs[kTileItems], p[kTileItems][4];
#pragma unroll
  for (int u = 0; u < kTileItems; ++u) {
    xs[u] = 0;
    const int item = threadIdx.x + u * blockDim.x;
    const int d = item < n_items ? item % dim : 0;
    const uint32_t* vd = v + (int64_t)d * 32;
    p[u][0] = vd[0];
#pragma unroll
    for (int t = 1; t < 4; ++t) p[u][t] = p[u][t - 1] ^ vd[t];
  }
  for (int64_t T = t_begin; T < t_end; ++T) {
    const int64_t k_tile = base + T * rows;
    if (T == t_begin) {
#pragma unroll
      for (int u = 0; u < kTileItems; ++u) {
        const int item = threadIdx.x + u * blockDim.x;
        const int r = item / dim, d = item - r * dim;
        const int64_t k0 = k_tile + (int64_t)kTileRun * r;
        if (item < n_items && k0 < end)
          xs[u] = sobol_point_par(v + (int64_t)d * 32, (uint64_t)k0,
                                  shift ? shift[d] : 0u);
      }
    } else {
      // each run start from the previous tile's by the columns of the
      // index bits that change (both starts below end <= 2^32: bits 0 ..
      // 31); the loads of the first kTileBits changed bits of every item
      // are issued together, further bits (rare) one by one.  A run past
      // the end stays past it in later tiles and is not advanced.
      uint32_t ld[kTileItems][kTileBits];
      uint64_t rest[kTileItems];
#pragma unroll
      for (int u = 0; u < kTileItems; ++u) {
        const int item = threadIdx.x + u * blockDim.x;
        const int r = item / dim, d = item - r * dim;
        const int64_t k0 = k_tile + (int64_t)kTileRun * r;
        const bool live = item < n_items && k0 < end;
        uint64_t m = live ? ((uint64_t)(k0 - rows) ^ (uint64_t)k0) : 0;
        const uint32_t* vd = v + (int64_t)(live ? d : 0) * 32;
#pragma unroll
        for (int i = 0; i < kTileBits; ++i) {
          ld[u][i] = m ? vd[__builtin_ctzll(m)] : 0u;
          m &= m - 1;
        }
        rest[u] = m;
      }
#pragma unroll
      for (int u = 0; u < kTileItems; ++u) {
        uint32_t x = xs[u];
#pragma unroll
        for (int i = 0; i < kTileBits; ++i) x ^= ld[u][i];
        if (rest[u]) {
          const int item = threadIdx.x + u * blockDim.x;
          const uint32_t* vd = v + (int64_t)(item - (item / dim) * dim) * 32;
          for (uint64_t m = rest[u]; m; m &= m - 1)
            x ^= vd[__builtin_ctzll(m)];
        }
        xs[u] = x;
      }
    }
#pragma unroll
    for (int u = 0; u < kTileItems; ++u) {
      const int item = threadIdx.x + u * blockDim.x;
      if (item >= n_items) continue;
      const int r = item / dim, d = item - r * dim;
      const int64_t k0 = k_tile + (int64_t)kTileRun * r;
      if (k0 >= end) continue;
      uint32_t x = xs[u];
      uint32_t* t = tile + (int64_t)(kTileRun * r) * dim + d;
#pragma unroll
      for (int j = 0; j < kTileRun; ++j) {
        if (j > 0) x ^= p[u][__builtin_ctz(j)];
        t[j * dim] = val(x, d);
      }
    }
    lds_barrier();
    const int64_t lo = k_tile > first ? k_tile : first;
    const int64_t hi = k_tile + rows < end ? k_tile + rows : end;
    if (hi > lo) {
      // (hi - lo) dim <= the tile's 12288 words: 32-bit indices; four LDS
      // reads in flight per group of stores.  Ordinary stores: streaming
      // (nontemporal) ones measured 0.88 ms against 0.65 ms for
      // sobol.sample(730, 2^20) here (the step-major path store is the
      // other way round)
      const int cnt = (int)((hi - lo) * dim) / W;  // dim (hi - lo) % W == 0
      typedef uint32_t vec __attribute__((ext_vector_type(W)));
      const vec* src = (const vec*)(tile + (lo - k_tile) * dim);
      vec* dst = (vec*)(out + (lo - first) * dim);
      const int bd = blockDim.x;
      int e = threadIdx.x;
      for (; e + 3 * bd < cnt; e += 4 * bd) {
        vec a[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) a[i] = src[e + i * bd];
#pragma unroll
        for (int i = 0; i < 4; ++i)
          dst[e + i * bd] = a[i];
      }
      for (; e < cnt; e += bd) dst[e] = src[e];
    }
    // the next tile overwrites the LDS tile; this tile's stores stay in
    // flight
    lds_barrier();
  }
}

// KIND 0: the uint32 points; 1: float32 x 2^-nd (TENT: tent-transformed).
template <int W, int KIND, bool TENT>
__global__ void __launch_bounds__(256)
sobol_points_tile_kernel(const uint32_t* __restrict__ v, int dim,
                         double scale, int64_t n, int64_t first,
                         const uint32_t* __restrict__ shift,
                         uint32_t* __restrict__ out, int rows,
                         int64_t tiles_per_block) {
  extern __shared__ uint32_t tile[];
  const float fs = (float)scale;
  sobol_tile<W>(v, dim, n, first, shift, rows, tiles_per_block, tile, out,
                [&](uint32_t x, int) {
                  if constexpr (KIND == 0) {
                    return x;
                  } else {
                    float u = (float)x * fs;
                    if constexpr (TENT)
                      u = u < 0.5f ? 2.0f * u : 2.0f * (1.0f - u);
                    return __float_as_uint(u);
                  }
                });
}

template <int W>
void launch_points_tile(int kind, int tent, unsigned blocks, size_t lds,
                        hipStream_t s, const uint32_t* v, int dim,
                        double scale, int64_t n, int64_t first,
                        const uint32_t* shift, uint32_t* out, int rows,
                        int64_t tiles_per_block) {
  if (kind == 0)
    hipLaunchKernelGGL((sobol_points_tile_kernel<W, 0, false>), blocks, 256,
                       lds, s, v, dim, scale, n, first, shift, out, rows,
                       tiles_per_block);
  else if (tent)
    hipLaunchKernelGGL((sobol_points_tile_kernel<W, 1, true>), blocks, 256,
                       lds, s, v, dim, scale, n, first, shift, out, rows,
                       tiles_per_block);
  else
    hipLaunchKernelGGL((sobol_points_tile_kernel<W, 1, false>), blocks, 256,
                       lds, s, v, dim, scale, n, first, shift, out, rows,
                       tiles_per_block);
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
    const int64_t n_tiles = (first_index + n - base + rows - 1) / rows;
    // persistent: three workgroups per CU (the LDS tile allows three)
    int cus = 0, dev = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount,
                              dev) != hipSuccess || cus < 1)
      cus = 256;
    int64_t grid = 3 * (int64_t)cus;
    if (grid > n_tiles) grid = n_tiles;
    const int64_t tiles_per_block = (n_tiles + grid - 1) / grid;
    const unsigned blocks =
        (unsigned)((n_tiles + tiles_per_block - 1) / tiles_per_block);
    const size_t lds = (size_t)rows * dim * 4;
    const uintptr_t a = (uintptr_t)out;
    const hipStream_t st = (hipStream_t)stream;
    if (dim % 4 == 0 && a % 16 == 0)
      launch_points_tile<4>(out_kind, tent, blocks, lds, st, v, dim, scale, n,
                            first_index, shift, (uint32_t*)out, rows,
                            tiles_per_block);
    else if (dim % 2 == 0 && a % 8 == 0)
      launch_points_tile<2>(out_kind, tent, blocks, lds, st, v, dim, scale, n,
                            first_index, shift, (uint32_t*)out, rows,
                            tiles_per_block);
    else
      launch_points_tile<1>(out_kind, tent, blocks, lds, st, v, dim, scale, n,
                            first_index, shift, (uint32_t*)out, rows,
                            tiles_per_block);
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
