// In-register normal sources shared by the fused path kernels.
//
// A kernel thread owns path `p`; it asks for normal k of step i, where the
// flat draw index j = i*D + k is uniform across the block (every thread
// walks the same steps).  Sources:
//  - injected draws:   draws[p*stride_path + i*stride_step + k]
//  - Philox4x32-10:    counter (p_lo, j/4, stream, p_hi), one call per four
//                      normals (Box-Muller pairs), antithetic rows p + N/2
//  - Sobol:            sequence index first + p, dimension j (the layout of
//                      tff/models/utils.py:114-127); low byte of the index
//                      from a [steps][256][D] table (the D dimensions of a
//                      step sit side by side, one wide load), high bits
//                      XOR-ed once per chunk of steps into LDS for the block.
#pragma once

#include "common.h"

namespace tqf {

constexpr int kPathBlock = 256;   // threads (= paths) per block
constexpr int kSobolChunk = 128;  // steps per LDS refill of the high-bit XOR

template <typename R>
struct RngArgs {
  uint32_t k0, k1, stream_id;
  const R* draws;
  int64_t stride_path, stride_step;
  const uint32_t* sv;      // [ndims][32] generating columns
  const uint32_t* slo;     // [ndims / D][256][D] low-byte XOR table
  const uint32_t* sshift;  // [ndims] or null
  int64_t sfirst;          // sequence index of path 0
  R sscale;                // 2^-nd
  int64_t n_paths;         // antithetic split
};

template <typename R>
inline void fill_rng_args(RngArgs<R>& a, const tqf_rng* r, int64_t n_paths) {
  a.k0 = (uint32_t)r->seed;
  a.k1 = (uint32_t)(r->seed >> 32);
  a.stream_id = r->stream_id;
  a.draws = (const R*)r->draws;
  a.stride_path = r->stride_path;
  a.stride_step = r->stride_step;
  a.sv = r->sobol_v;
  a.slo = r->sobol_lo;
  a.sshift = r->sobol_shift;
  a.sfirst = r->sobol_first_index;
  a.sscale = (R)std::ldexp(1.0, -r->sobol_nd);
  a.n_paths = n_paths;
}

// LDS words the Sobol high-bit table needs for D normals per step.
constexpr int sobol_lds_words(int D) { return 2 * kSobolChunk * D; }


template <typename R, int RNG>
struct Normals {
  static constexpr bool kSobol =
      RNG == TQF_RNG_SOBOL || RNG == TQF_RNG_SOBOL_SHIFTED;
  R c0, c1, c2, c3;  // Philox cache
  uint32_t ctr0, ctr3;
  R sign;
  int64_t p;
  uint32_t lo_idx, hi_sel;
  int chunk_base;
  int D;  // normals per step

  __device__ void init(const RngArgs<R>& a, int64_t path, int d) {
    D = d;
    p = path;
    sign = (R)1;
    chunk_base = 0;
    if (RNG == TQF_RNG_PHILOX_ANTITHETIC) {
      const int64_t half = a.n_paths / 2;
      if (path >= half) {
        p = path - half;
        sign = (R)-1;
      }
    }
    ctr0 = (uint32_t)p;
    ctr3 = (uint32_t)(p >> 32);
    if (kSobol) {
      const uint64_t idx = (uint64_t)(a.sfirst + path);
      const uint64_t hi0 =
          (uint64_t)(a.sfirst + (int64_t)blockIdx.x * kPathBlock) >> 8;
      lo_idx = (uint32_t)(idx & 255);
      hi_sel = (uint32_t)((idx >> 8) - hi0);
    }
  }

  // Call at the top of step i with every thread of the block (it may
  // synchronise).  n_dims bounds the Sobol dimensions that exist.
  __device__ __forceinline__ void begin_step(const RngArgs<R>& a, uint32_t* hx,
                                             int i, int n_dims) {
    if (kSobol && (i % kSobolChunk) == 0) {
      chunk_base = i;
      __syncthreads();
      const uint64_t hi0 =
          (uint64_t)(a.sfirst + (int64_t)blockIdx.x * kPathBlock) >> 8;
      const int n = kSobolChunk * D;
      for (int t = threadIdx.x; t < 2 * n; t += blockDim.x) {
        const int c = t / n;
        const int j = i * D + (t - c * n);
        uint32_t x = 0;
        if (j < n_dims) {
          x = a.sshift ? a.sshift[j] : 0u;
          uint64_t hb = (hi0 + c) & 0xffffffull;  // index bits 8..31
          const uint32_t* v = a.sv + (int64_t)j * 32 + 8;
          while (hb) {
            x ^= v[__builtin_ctzll(hb)];
            hb &= hb - 1;
          }
        }
        hx[t] = x;
      }
      __syncthreads();
    }
  }

  __device__ __forceinline__ R get(const RngArgs<R>& a, const uint32_t* hx,
                                   int i, int k) {
    const int j = i * D + k;
    if (RNG == TQF_RNG_DRAWS) {
      return a.draws[p * a.stride_path + (int64_t)i * a.stride_step + k];
    } else if (RNG == TQF_RNG_PHILOX || RNG == TQF_RNG_PHILOX_ANTITHETIC) {
      const int r = j & 3;
      if (r == 0) {
        const u32x4 w = philox4x32_10(
            u32x4{ctr0, (uint32_t)(j >> 2), a.stream_id, ctr3}, a.k0, a.k1);
        box_muller(w.x, w.y, c0, c1);
        box_muller(w.z, w.w, c2, c3);
      }
      const R z = r == 0 ? c0 : r == 1 ? c1 : r == 2 ? c2 : c3;
      return sign * z;
    } else {
      const int jj = j - chunk_base * D;
      // low-byte table interleaved by step: [step][256][D]
      const uint32_t v = a.slo[((int64_t)i * 256 + lo_idx) * D + k] ^
                         hx[hi_sel * (kSobolChunk * D) + jj];
      if (RNG == TQF_RNG_SOBOL) return sobol_ref_normal<R>(v, a.sscale);
      return sobol_shifted_normal(v, R(0));
    }
  }
};

}  // namespace tqf
