// The engine's counter-based normal stream as a standalone generator (the
// fused steppers inline the same counter layout).
#include "common.h"

using namespace tqf;

namespace {

template <typename R>
__global__ void __launch_bounds__(256)
philox_normals_kernel(uint32_t k0, uint32_t k1, uint32_t stream_id,
                      int64_t n_paths, int64_t n_draws, int64_t n_blk,
                      int antithetic, R* __restrict__ out) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n_paths * n_blk) return;
  const int64_t p = g / n_blk;
  const int64_t b = g - p * n_blk;
  int64_t src = p;
  R sign = (R)1;
  if (antithetic && p >= n_paths / 2) {
    src = p - n_paths / 2;
    sign = (R)-1;
  }
  const u32x4 w = philox4x32_10(
      u32x4{(uint32_t)src, (uint32_t)b, stream_id, (uint32_t)(src >> 32)}, k0,
      k1);
  R z[4];
  box_muller(w.x, w.y, z[0], z[1]);
  box_muller(w.z, w.w, z[2], z[3]);
  R* o = out + p * n_draws + b * 4;
  const int64_t rem = n_draws - b * 4;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    if (r < rem) o[r] = sign * z[r];
  }
}

}  // namespace

extern "C" int tqf_philox_normals(uint64_t seed, uint32_t stream_id,
                                  int64_t n_paths, int64_t n_draws,
                                  int32_t antithetic, int32_t dtype, void* out,
                                  void* stream) {
  clear_error();
  TQF_REQUIRE(n_paths >= 0 && n_draws >= 0, "sizes must be non-negative");
  TQF_REQUIRE(!antithetic || n_paths % 2 == 0,
              "First dimension of `sample_shape` should be even for "
              "PSEUDO_ANTITHETIC random type");
  TQF_REQUIRE(dtype == TQF_F32 || dtype == TQF_F64, "bad dtype %d", dtype);
  TQF_REQUIRE(n_draws < ((int64_t)1 << 34), "too many draws per path");
  if (n_paths == 0 || n_draws == 0) return TQF_OK;
  TQF_REQUIRE(out != nullptr, "null output");
  const int64_t n_blk = (n_draws + 3) / 4;
  const dim3 grid((unsigned)((n_paths * n_blk + 255) / 256));
  const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
  if (dtype == TQF_F32) {
    hipLaunchKernelGGL(philox_normals_kernel<float>, grid, dim3(256), 0,
                       (hipStream_t)stream, k0, k1, stream_id, n_paths,
                       n_draws, n_blk, antithetic, (float*)out);
  } else {
    hipLaunchKernelGGL(philox_normals_kernel<double>, grid, dim3(256), 0,
                       (hipStream_t)stream, k0, k1, stream_id, n_paths,
                       n_draws, n_blk, antithetic, (double*)out);
  }
  return check_launch("tqf_philox_normals");
}
