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
#include "icdf_table.h"

namespace tqf {

constexpr int kPathBlock = 256;   // threads (= paths) per block
// Word 3 of a raw (untyped) buffer descriptor on gfx9: data format 32.
constexpr int kRawBufferWord3 = 0x00020000;
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
constexpr int kSobolChunk = 128;  // steps per LDS refill of the high-bit XOR

template <typename R>
struct RngArgs {
  uint32_t k0, k1, stream_id;
  const R* draws;
  int64_t stride_path, stride_step;
  const uint32_t* sv;      // [ndims][32] generating columns
  const uint32_t* slo;     // [ndims / D][256][D] low-byte XOR table
  uint32_t slo_bytes;      // its size (bounds the buffer loads)
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
  const int64_t lo_bytes = (int64_t)r->sobol_ndims * 256 * 4;
  a.slo_bytes = (uint32_t)(lo_bytes < 0xffffffffll ? lo_bytes : 0xffffffffll);
  a.sshift = r->sobol_shift;
  a.sfirst = r->sobol_first_index;
  a.sscale = (R)std::ldexp(1.0, -r->sobol_nd);
  a.n_paths = n_paths;
}

// LDS words the Sobol high-bit table needs for D normals per step.
constexpr int sobol_lds_words(int D) { return 2 * kSobolChunk * D; }


// ---- float32 shifted-Sobol inverse CDF (piecewise cubic, table in LDS) ----
// z = Phi^-1((n + 1/2) 2^-25) of the top 25 bits n of a 32-bit point.  The
// point folds onto the lower half with its sign mask s (m = (v ^ s) >> 7 in
// [0, 2^24); the upper half, s all ones, reads the negated copy of the
// table); the float32 value of m * 128 (exact, formed by masking the low 7
// bits) indexes a segment by its exponent and top 3 mantissa bits, and its
// low 20 mantissa bits minus 2^20 are the cubic's variable (csrc/
// fit_icdf.py derives the table; exhaustive float32 error: 5.2e-7
// absolute, 3.3e-7 relative for |z| >= 1e-3).
//
// LDS image: segment k sits at byte ((k + 1072) & 255) * 16, which is bits
// 20..27 of the float's bit pattern times 16 (one AND of its upper half):
// segments 0..191 land on bytes [768, 3840), and m = 0 (the float 0) on
// byte 0, which holds the constant kIcdfZero.  The negated copy is 4096
// bytes above (bit 12, set from the sign mask with one AND-OR).
constexpr int kIcdfHalf = 4096;
constexpr int kIcdfSlot0 = (134 << kIcdfSegBits) & 255;  // 48
constexpr int kIcdfLdsBytes = 2 * kIcdfHalf;
static_assert(kIcdfSegBits == 3, "the LDS addressing assumes 8 sub-segments");
static_assert((kIcdfSlot0 + kIcdfSegs) * 16 <= kIcdfHalf,
              "a table half must fit below bit 12");

// The table's block-shared copy (static LDS of every kernel that uses it).
__device__ __forceinline__ unsigned char* icdf_lds() {
  __shared__ __align__(16) unsigned char t[kIcdfLdsBytes];
  return t;
}

// Cooperative copy of both halves into LDS (caller synchronises).
__device__ __forceinline__ void icdf_stage(unsigned char* lds) {
  float4* t = (float4*)lds;
  for (int e = threadIdx.x; e < kIcdfSegs; e += blockDim.x) {
    const float4 c = make_float4(kIcdfCoef[4 * e], kIcdfCoef[4 * e + 1],
                                 kIcdfCoef[4 * e + 2], kIcdfCoef[4 * e + 3]);
    t[kIcdfSlot0 + e] = c;
    t[kIcdfHalf / 16 + kIcdfSlot0 + e] = make_float4(-c.x, -c.y, -c.z, -c.w);
  }
  if (threadIdx.x == 0) {
    t[0] = make_float4(kIcdfZero, 0.f, 0.f, 0.f);
    t[kIcdfHalf / 16] = make_float4(-kIcdfZero, 0.f, 0.f, 0.f);
  }
}

__device__ __forceinline__ float icdf_shifted_f32(uint32_t v,
                                                  const unsigned char* lds) {
  const uint32_t s = (uint32_t)((int32_t)v >> 31);
  const uint32_t w = (v ^ s) & 0xFFFFFF80u;       // m * 128
  const uint32_t fi = __float_as_uint((float)w);  // exact
  // m * 128 is 0 or in [2^7, 2^31): exponent 134..157, so bits 20..27 of
  // fi are 48..239 (no wrap), or 0 for m = 0; the offset stays in its half
  uint32_t off;  // (fi >> 16) & 0xFF0: one AND reading the upper half-word
  asm("v_and_b32_sdwa %0, %1, %2 dst_sel:DWORD dst_unused:UNUSED_PAD "
      "src0_sel:WORD_1 src1_sel:DWORD"
      : "=v"(off)
      : "v"(fi), "v"(0xFF0u));
  const float g = (float)(int32_t)(fi | 0xFFF00000u);  // [-2^20, -1]
  const float4 c = *(const float4*)(lds + ((s & (uint32_t)kIcdfHalf) | off));
  return __builtin_fmaf(__builtin_fmaf(__builtin_fmaf(c.w, g, c.z), g, c.y),
                        g, c.x);
}

// Bytes of dynamic LDS the normal source of a kernel needs (the Sobol
// high-bit words; the ICDF table is static LDS).
__host__ __device__ constexpr int rng_lds_bytes(int rng_kind, int D) {
  return (rng_kind == TQF_RNG_SOBOL || rng_kind == TQF_RNG_SOBOL_SHIFTED)
             ? sobol_lds_words(D) * 4
             : 0;
}

template <typename R, int RNG>
struct Normals {
  static constexpr bool kSobol =
      RNG == TQF_RNG_SOBOL || RNG == TQF_RNG_SOBOL_SHIFTED;
  static constexpr bool kIcdf =
      RNG == TQF_RNG_SOBOL_SHIFTED && sizeof(R) == sizeof(float);
  R c0, c1, c2, c3;  // Philox cache
  uint32_t ctr0, ctr3;
  R sign;
  int64_t p;
  uint32_t lo_idx, hi_sel;
  int chunk_base;
  int D;  // normals per step
  unsigned char* icdf;
  uint32_t* hx;
  __amdgpu_buffer_rsrc_t lo_rsrc;  // low-byte table (raw buffer, bounds-checked)
  u32x2 lo_next;                   // get2: the next step's low-byte words
  int last_step;

  // Every thread of the block calls init (it stages LDS tables; the first
  // begin_step synchronises before any use).
  __device__ void init(const RngArgs<R>& a, int64_t path, int d,
                       unsigned char* lds, int n_steps) {
    D = d;
    last_step = n_steps > 0 ? n_steps - 1 : 0;
    p = path;
    sign = (R)1;
    chunk_base = 0;
    hx = (uint32_t*)lds;
    if constexpr (kIcdf) {
      icdf = icdf_lds();
      icdf_stage(icdf);
    }
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
      lo_rsrc = __builtin_amdgcn_make_buffer_rsrc((void*)a.slo, (short)0,
                                                  (int)a.slo_bytes, kRawBufferWord3);
      if (D == 2 && n_steps > 0) lo_next = load_lo2(0);
    }
  }

  // Step i's two low-byte words: voffset = the lane's entry, soffset = the
  // step's row (uniform).
  __device__ __forceinline__ u32x2 load_lo2(int i) const {
    return __builtin_amdgcn_raw_buffer_load_b64(lo_rsrc, lo_idx * 8u,
                                                (uint32_t)i * 2048u, 0);
  }

  // Call at the top of step i with every thread of the block (it may
  // synchronise).  n_dims bounds the Sobol dimensions that exist.
  __device__ __forceinline__ void begin_step(const RngArgs<R>& a, int i,
                                             int n_dims) {
    if (kSobol && (i % kSobolChunk) == 0) refill(a, i, n_dims);
  }

  // Builds the block's high-bit words for steps [i, i + kSobolChunk) (Sobol
  // only; every thread of the block calls it, it synchronises).
  __device__ __forceinline__ void refill(const RngArgs<R>& a, int i,
                                         int n_dims) {
    if (kSobol) {
      chunk_base = i;
      __syncthreads();
      const uint64_t hi0 =
          (uint64_t)(a.sfirst + (int64_t)blockIdx.x * kPathBlock) >> 8;
      const int n = kSobolChunk * D;
      // row c holds index bits 8..31 of hi0 + c; the second row is needed
      // only when the block's 256 paths straddle a multiple of 256
      const int rows =
          ((a.sfirst + (int64_t)blockIdx.x * kPathBlock) & 255) != 0 ? 2 : 1;
      for (int c = 0; c < rows; ++c) {
        // uniform per pass: the column masks live in scalar registers
        const uint32_t hb = (uint32_t)((hi0 + c) & 0xffffffull);
        for (int t = threadIdx.x; t < n; t += blockDim.x) {
          const int j = i * D + t;
          uint32_t x = 0;
          if (j < n_dims) {
            // columns 8..31 of dimension j in six 16-byte loads issued
            // together, then masked by the index bits 8..31
            const uint4* col = (const uint4*)(a.sv + (int64_t)j * 32 + 8);
            uint4 cv[6];
#pragma unroll
            for (int q = 0; q < 6; ++q) cv[q] = col[q];
            x = a.sshift ? a.sshift[j] : 0u;
#pragma unroll
            for (int q = 0; q < 6; ++q) {
              x ^= cv[q].x & (0u - ((hb >> (4 * q)) & 1u));
              x ^= cv[q].y & (0u - ((hb >> (4 * q + 1)) & 1u));
              x ^= cv[q].z & (0u - ((hb >> (4 * q + 2)) & 1u));
              x ^= cv[q].w & (0u - ((hb >> (4 * q + 3)) & 1u));
            }
          }
          hx[c * n + t] = x;
        }
      }
      __syncthreads();
    }
  }

  __device__ __forceinline__ R sobol_value(const RngArgs<R>& a, uint32_t v) {
    if constexpr (RNG == TQF_RNG_SOBOL) {
      return sobol_ref_normal<R>(v, a.sscale);
    } else if constexpr (kIcdf) {
      return (R)icdf_shifted_f32(v, icdf);
    } else {
      return sobol_shifted_normal(v, R(0));
    }
  }

  __device__ __forceinline__ R get(const RngArgs<R>& a, int i, int k) {
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
      return sobol_value(a, v);
    }
  }

  // Both normals of a step with two (D == 2): the Sobol words come in one
  // 8-byte global load and one 8-byte LDS read.
  __device__ __forceinline__ void get2(const RngArgs<R>& a, int i, R* z) {
    if (kSobol) {
      const int jj = (i - chunk_base) * 2;
      // software pipeline: this step's words were loaded one step ago
      const u32x2 lo = lo_next;
      lo_next = load_lo2(i < last_step ? i + 1 : last_step);
      const uint2 hi = *(const uint2*)(hx + hi_sel * (kSobolChunk * 2) + jj);
      z[0] = sobol_value(a, lo.x ^ hi.x);
      z[1] = sobol_value(a, lo.y ^ hi.y);
    } else {
      z[0] = get(a, i, 0);
      z[1] = get(a, i, 1);
    }
  }
};

}  // namespace tqf
