// Exact-scheme (log-normal) geometric Brownian motion paths, univariate
// and correlated multivariate, fused with the normal generation.
//
// Reference: tff/models/geometric_brownian_motion/
//   univariate_geometric_brownian_motion.py:261-317 and
//   multivariate_geometric_brownian_motion.py:229-282:
//     inc_i = (mu_i - sigma_i^2 / 2) dt + sqrt(dt) sigma_i (L z)_i
//     S_i(t_k) = S0_i exp(sum_{steps <= k} inc_i)
// The reference forms the cumulative sum as a [k, k] lower-triangular
// matvec over materialised draws; here each thread carries the running sum
// in registers over the grid, with the normals produced in place.
#include "common.h"
#include "normals.h"

using namespace tqf;

namespace {

template <typename R>
struct GbmArgs {
  int d;
  const R* drift;   // [d] mu_i - sigma_i^2 / 2
  const R* vol;     // [d] sigma_i
  const R* chol;    // [d][d] lower Cholesky factor of the correlation, or null
  const R* s0;      // [d] or [n_paths][d]
  int64_t s0_stride;
  const R* dt;
  const R* sqrt_dt;
  const int* slot;
  int n_steps, n_slots, init_slot;
  RngArgs<R> rng;
  int mode;
  R* out;
  int64_t n_paths;
  int stage_off;   // LDS byte offset of the path-major staging buffer
  int stage_recs;  // records per flush (stage_recs * d <= kStageWords)
};

// Path-major output ([N][k][d]) is staged: a block keeps up to kStageWords
// values per path in LDS ([path][kStageStride], odd stride: the per-path
// writes and the flush's reads fall in distinct banks) and flushes them as
// one contiguous segment per path, in place of one scattered store per
// path and record.
constexpr int kStageWords = 32;
constexpr int kStageStride = kStageWords + 1;

template <typename R, int RNG, int MAXD>
__global__ void __launch_bounds__(kPathBlock) gbm_exact_kernel(GbmArgs<R> a) {
  extern __shared__ __align__(16) unsigned char lds[];
  const int64_t path = (int64_t)blockIdx.x * kPathBlock + threadIdx.x;
  const bool active = path < a.n_paths;
  const int64_t pc = active ? path : a.n_paths - 1;
  const int d = a.d;

  R cum[MAXD], s0[MAXD];
#pragma unroll
  for (int c = 0; c < MAXD; ++c) {
    cum[c] = 0;
    s0[c] = c < d ? a.s0[pc * a.s0_stride + c] : (R)0;
  }
  Normals<R, RNG> nrm;
  nrm.init(a.rng, pc, d, lds, a.n_steps);

  const bool stage = a.mode == TQF_OUT_PATH_MAJOR;
  R* buf = (R*)(lds + a.stage_off);
  int nrec = 0, first_slot = 0;  // block-uniform
  // all threads of the block: write the staged records of every path as
  // one segment of slots [first_slot, first_slot + nrec)
  auto flush = [&]() {
    __syncthreads();
    const int seg = nrec * d;
    const int64_t p0 = (int64_t)blockIdx.x * kPathBlock;
    for (int e = threadIdx.x; e < kPathBlock * seg; e += kPathBlock) {
      const int p = e / seg, rem = e - p * seg;
      const int64_t pg = p0 + p;
      if (pg < a.n_paths)
        a.out[(pg * a.n_slots + first_slot) * d + rem] =
            buf[p * kStageStride + rem];
    }
    __syncthreads();
    nrec = 0;
  };
  auto record = [&](int slot, bool initial) {
    if (!stage) {
      if (!active) return;
#pragma unroll
      for (int c = 0; c < MAXD; ++c) {
        if (c < d) {
          const R v = initial ? s0[c] : s0[c] * fast_exp(cum[c]);
          a.out[((int64_t)slot * d + c) * a.n_paths + path] = v;
        }
      }
      return;
    }
    if (nrec > 0 && slot != first_slot + nrec) flush();
    if (nrec == 0) first_slot = slot;
    if (active) {
#pragma unroll
      for (int c = 0; c < MAXD; ++c) {
        if (c < d)
          buf[threadIdx.x * kStageStride + nrec * d + c] =
              initial ? s0[c] : s0[c] * fast_exp(cum[c]);
      }
    }
    if (++nrec == a.stage_recs) flush();
  };
  if (a.init_slot >= 0) record(a.init_slot, true);

  const int n_dims = a.n_steps * d;
  for (int i = 0; i < a.n_steps; ++i) {
    nrm.begin_step(a.rng, i, n_dims);
    const R dt = ld_uniform(a.dt, i), sdt = ld_uniform(a.sqrt_dt, i);
    R z[MAXD];
#pragma unroll
    for (int c = 0; c < MAXD; ++c) z[c] = c < d ? nrm.get(a.rng, i, c) : (R)0;
    if (a.chol) {
#pragma unroll
      for (int r = MAXD - 1; r >= 0; --r) {
        if (r < d) {
          R w = 0;
#pragma unroll
          for (int c = 0; c <= r; ++c) w += ld_uniform(a.chol, r * d + c) * z[c];
          z[r] = w;
        }
      }
    }
#pragma unroll
    for (int c = 0; c < MAXD; ++c)
      if (c < d)
        cum[c] += ld_uniform(a.drift, c) * dt + sdt * ld_uniform(a.vol, c) * z[c];
    const int s = ld_uniform(a.slot, i);
    if (s >= 0) record(s, false);
  }
  if (stage && nrec > 0) flush();
}

template <typename R, int MAXD>
int launch_d(const GbmArgs<R>& a, int rng_kind, unsigned blocks, size_t lds,
             hipStream_t s) {
  switch (rng_kind) {
    case TQF_RNG_DRAWS:
      hipLaunchKernelGGL((gbm_exact_kernel<R, TQF_RNG_DRAWS, MAXD>), blocks,
                         kPathBlock, lds, s, a);
      break;
    case TQF_RNG_PHILOX:
      hipLaunchKernelGGL((gbm_exact_kernel<R, TQF_RNG_PHILOX, MAXD>), blocks,
                         kPathBlock, lds, s, a);
      break;
    case TQF_RNG_PHILOX_ANTITHETIC:
      hipLaunchKernelGGL((gbm_exact_kernel<R, TQF_RNG_PHILOX_ANTITHETIC, MAXD>),
                         blocks, kPathBlock, lds, s, a);
      break;
    case TQF_RNG_SOBOL:
      hipLaunchKernelGGL((gbm_exact_kernel<R, TQF_RNG_SOBOL, MAXD>), blocks,
                         kPathBlock, lds, s, a);
      break;
    default:
      hipLaunchKernelGGL((gbm_exact_kernel<R, TQF_RNG_SOBOL_SHIFTED, MAXD>),
                         blocks, kPathBlock, lds, s, a);
  }
  return check_launch("tqf_gbm_exact");
}

template <typename R>
int launch(const tqf_gbm* g, const tqf_grid* grid, const tqf_rng* r,
           int mode, void* out, int64_t n_paths, hipStream_t s) {
  GbmArgs<R> a{};
  a.d = g->dim;
  a.drift = (const R*)g->drift;
  a.vol = (const R*)g->vol;
  a.chol = (const R*)g->chol;
  a.s0 = (const R*)g->s0;
  a.s0_stride = g->s0_stride_path;
  a.dt = (const R*)grid->dt;
  a.sqrt_dt = (const R*)grid->sqrt_dt;
  a.slot = grid->slot;
  a.n_steps = grid->n_steps;
  a.n_slots = grid->n_slots;
  a.init_slot = grid->init_slot;
  fill_rng_args(a.rng, r, n_paths);
  a.mode = mode;
  a.out = (R*)out;
  a.n_paths = n_paths;
  size_t lds = ((size_t)rng_lds_bytes(r->kind, g->dim) + 15) & ~(size_t)15;
  a.stage_off = (int)lds;
  a.stage_recs = kStageWords / g->dim;  // dim <= 16: at least 2
  if (mode == TQF_OUT_PATH_MAJOR)
    lds += (size_t)kPathBlock * kStageStride * sizeof(R);
  if (lds == 0) lds = 16;
  const unsigned blocks = (unsigned)((n_paths + kPathBlock - 1) / kPathBlock);
  if (g->dim <= 1) return launch_d<R, 1>(a, r->kind, blocks, lds, s);
  if (g->dim <= 4) return launch_d<R, 4>(a, r->kind, blocks, lds, s);
  if (g->dim <= 8) return launch_d<R, 8>(a, r->kind, blocks, lds, s);
  return launch_d<R, 16>(a, r->kind, blocks, lds, s);
}

}  // namespace

extern "C" int tqf_gbm_exact(const tqf_gbm* g, const tqf_grid* grid,
                             const tqf_rng* r, int32_t out_mode, void* out,
                             int64_t n_paths, int32_t dtype, void* stream) {
  clear_error();
  TQF_REQUIRE(g && grid && r, "null descriptor");
  TQF_REQUIRE(g->dim >= 1 && g->dim <= 16,
              "exact GBM supports 1 <= dim <= 16, got %d", g->dim);
  TQF_REQUIRE(g->drift && g->vol && g->s0, "null GBM parameters");
  TQF_REQUIRE(g->s0_stride_path == 0 || g->s0_stride_path == g->dim,
              "s0_stride_path must be 0 or dim");
  TQF_REQUIRE(n_paths >= 1, "num_samples must be positive");
  TQF_REQUIRE(dtype == TQF_F32 || dtype == TQF_F64, "bad dtype %d", dtype);
  TQF_REQUIRE(grid->n_steps >= 0 && grid->n_slots >= 1, "bad grid sizes");
  TQF_REQUIRE(grid->n_steps == 0 || (grid->dt && grid->sqrt_dt && grid->slot),
              "null grid arrays");
  TQF_REQUIRE(out_mode == TQF_OUT_PATH_MAJOR || out_mode == TQF_OUT_STEP_MAJOR,
              "exact GBM records paths (mode 0 or 1)");
  TQF_REQUIRE(out != nullptr, "null output");
  TQF_REQUIRE(r->kind >= TQF_RNG_DRAWS && r->kind <= TQF_RNG_SOBOL_SHIFTED,
              "unknown rng kind %d", r->kind);
  if (r->kind == TQF_RNG_DRAWS) TQF_REQUIRE(r->draws, "null normal_draws");
  if (r->kind == TQF_RNG_PHILOX_ANTITHETIC)
    TQF_REQUIRE(n_paths % 2 == 0,
                "First dimension of `sample_shape` should be even for "
                "PSEUDO_ANTITHETIC random type");
  if (r->kind == TQF_RNG_SOBOL || r->kind == TQF_RNG_SOBOL_SHIFTED) {
    TQF_REQUIRE(r->sobol_v && r->sobol_lo, "null Sobol tables");
    TQF_REQUIRE((int64_t)r->sobol_ndims >= (int64_t)grid->n_steps * g->dim,
                "Sobol tables hold too few dimensions");
    TQF_REQUIRE(r->sobol_first_index >= 0 &&
                    r->sobol_first_index + n_paths <= (int64_t)0xffffffffll + 1,
                "Sobol index range exceeds 32 bits");
  }
  const hipStream_t s = (hipStream_t)stream;
  if (dtype == TQF_F32) return launch<float>(g, grid, r, out_mode, out, n_paths, s);
  return launch<double>(g, grid, r, out_mode, out, n_paths, s);
}
