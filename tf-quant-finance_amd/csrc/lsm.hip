// Longstaff-Schwartz regression on the GPU (tff/models/longstaff_schwartz/
// lsm.py:128-436) for the reference's polynomial basis
// (make_polynomial_basis, lsm.py:50-125) and the built-in payoffs.
//
// Per exercise date (backward induction):
//   tqf_lsm_prepare   exercise value, continuation target y = df_ratio *
//                     (cashflow + values), per-asset sums for the centring
//   tqf_lsm_gram      G = X^T W X and X^T W y in float64 over the ITM
//                     calibration rows, with the basis built on the fly
//                     (X is never written to HBM)
//   tqf_lsm_solve     float64 Cholesky of the column-scaled G with pivot
//                     truncation (rank-deficient columns are dropped, which
//                     leaves the fitted values X beta of the pseudo-inverse
//                     solution unchanged) and the triangular solves
//   tqf_lsm_update    continuation = relu(X beta), exercise decision,
//                     cashflow / value update for every path
// and tqf_lsm_finalize for the discounted mean over the pricing rows.
//
// Basis: the centred variables c_j = (x_j - mean_j) * scale_j are scaled by
// a per-asset factor (the host uses 1 / std); column k is
// prod_j c_j^{e[k][j]} with the reference's exponent order.  Scaling a
// column by a constant does not change the fitted values.
#include "common.h"
#include "mlp.h"

using namespace tqf;

namespace {

constexpr int kMaxAssets = 8;
constexpr int kMaxPw = 64;      // (degree + 1) * dim power-table entries
constexpr int kTile = 64;       // Gram output tile (kTile x kTile)
constexpr int kSubPaths = 32;   // paths staged per LDS round

typedef double f64x4 __attribute__((ext_vector_type(4)));

// LDS written by one lane of a wave is visible to the others after this
// (no workgroup barrier).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ f64x4 mfma_f64(double a, double b, f64x4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// Lane `l`'s value of v in every lane (l wave-uniform).
__device__ __forceinline__ double read_lane_f64(double v, int l) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, l);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

struct LsmArgs {
  const double* paths;      // float64, element (p, t, j) at
                            // paths[p sp + t st + j sa]
  int64_t sp, st, sa;
  int64_t n_paths;
  int n_times, d;
  int t_idx;                // time index of the exercise date
  int payoff;               // TQF_PAYOFF_*
  double strike;
  const double* mean;       // [d]  (device) centring
  const double* scale;      // [d]
  const uint8_t* expo;      // [K][d] exponents
  int K;
  int deg;
  int basis_log;            // 0: basis on the stored values x; 1: on
                            // ln x (payoff on x); 2: the stored values
                            // are ln x (basis on them, payoff on exp)
  double log_strike;        // ln(strike) (-inf for strike <= 0)
  const double* table;      // TQF_PAYOFF_TABLE: [n_times][n_paths]
  // per-path discount rows of the current date (numerator e + 1 and
  // denominator e of the padded factors), or null for the scalar ratio
  const double* dnum;
  const double* dden;
};

// Discount ratio of path p: the per-path rows when given, else the date's
// scalar ratio.
__device__ __forceinline__ double disc_ratio(const LsmArgs& a, double ratio,
                                             int64_t p) {
  return a.dnum ? a.dnum[p] / a.dden[p] : ratio;
}

// Asset j of path p at the date's time index.  With the step-major layout
// (sp = 1) consecutive threads read consecutive doubles.
__device__ __forceinline__ double path_x(const LsmArgs& a, int64_t p, int j) {
  return a.paths[p * a.sp + (int64_t)a.t_idx * a.st + (int64_t)j * a.sa];
}

// The variable the basis is built on: x, or ln x when the caller regresses
// on log prices (the pricing platform's American options regress on
// ln S: tff/experimental/pricing_platform/framework/equity_instruments/
// american_option/utils.py:150-167).
__device__ __forceinline__ double basis_input(const LsmArgs& a, double x) {
  return a.basis_log == 1 ? log(x) : x;
}

// The payoff of the stored values x[0 .. d) of path p (exp of them when
// the paths hold log prices).
__device__ __forceinline__ double payoff_value(const LsmArgs& a,
                                               const double* x, int64_t p) {
  if (a.payoff == TQF_PAYOFF_TABLE)
    return a.table[(int64_t)a.t_idx * a.n_paths + p];
  auto value = [&](int j) { return a.basis_log == 2 ? exp(x[j]) : x[j]; };
  if (a.payoff == TQF_PAYOFF_BASKET_PUT) {
    double s = 0;
    for (int j = 0; j < a.d; ++j) s += value(j);
    const double v = a.strike - s / a.d;
    return v > 0 ? v : 0;
  } else {  // TQF_PAYOFF_MAX_CALL
    double m = value(0);
    for (int j = 1; j < a.d; ++j) {
      const double v = value(j);
      m = v > m ? v : m;
    }
    const double v = m - a.strike;
    return v > 0 ? v : 0;
  }
}

// payoff_value(a, x, p) > 0.  One asset held as a log price: the put is in
// the money iff ln S < ln K, the call iff ln S > ln K, with no exp.
__device__ __forceinline__ bool in_the_money(const LsmArgs& a,
                                             const double* x, int64_t p) {
  if (a.basis_log == 2 && a.d == 1 && a.payoff != TQF_PAYOFF_TABLE)
    return a.payoff == TQF_PAYOFF_BASKET_PUT ? x[0] < a.log_strike
                                             : x[0] > a.log_strike;
  return payoff_value(a, x, p) > 0;
}

// Exercise values, continuation targets and the per-block asset moments
// (sum and sum of squares, for the centring and the column scaling).
__global__ void __launch_bounds__(256)
lsm_prepare_kernel(LsmArgs a, double df_ratio, const double* __restrict__ cash,
                   const double* __restrict__ vals, double* __restrict__ exv,
                   double* __restrict__ y, double* __restrict__ block_sums) {
  // per-asset sums of the basis input and its square: wave butterflies,
  // then the four wave sums in wave order (deterministic)
  __shared__ double red[4][2 * kMaxAssets];
  const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  double x[kMaxAssets];
  for (int j = 0; j < kMaxAssets; ++j) x[j] = 0;
  if (p < a.n_paths) {
    for (int j = 0; j < a.d; ++j) x[j] = path_x(a, p, j);
    exv[p] = payoff_value(a, x, p);
    y[p] = disc_ratio(a, df_ratio, p) * (cash[p] + vals[p]);
  }
  for (int j = 0; j < a.d; ++j) {
    const double b = p < a.n_paths ? basis_input(a, x[j]) : 0.0;
    double s1 = b, s2 = b * b;
    for (int o = 32; o > 0; o >>= 1) {
      s1 += __shfl_xor(s1, o, 64);
      s2 += __shfl_xor(s2, o, 64);
    }
    if (lane == 0) {
      red[wave][j] = s1;
      red[wave][kMaxAssets + j] = s2;
    }
  }
  __syncthreads();
  if (threadIdx.x < a.d) {
    const int j = threadIdx.x;
    const double s1 = ((red[0][j] + red[1][j]) + red[2][j]) + red[3][j];
    const double s2 = ((red[0][kMaxAssets + j] + red[1][kMaxAssets + j]) +
                       red[2][kMaxAssets + j]) + red[3][kMaxAssets + j];
    block_sums[(int64_t)blockIdx.x * 2 * kMaxAssets + j] = s1;
    block_sums[(int64_t)blockIdx.x * 2 * kMaxAssets + kMaxAssets + j] = s2;
  }
}

// mean[d] and scale[d] = 1 / std (1 when the spread is zero).  One
// workgroup of kMeanThreads reads the block sums [n_blocks][2 kMaxAssets] as
// one flat array (consecutive threads, consecutive doubles); since
// kMeanThreads and 64 are multiples of 2 kMaxAssets, thread t only ever adds
// column t % (2 kMaxAssets), in block order.  Lanes l, l^16, l^32, l^48 hold
// the same column and are combined by two butterflies, then the wave sums
// are added in wave order (deterministic).  Columns past d are never mixed
// into the ones in use.
constexpr int kMeanThreads = 1024;
static_assert(64 % (2 * kMaxAssets) == 0 && 2 * kMaxAssets == 16,
              "the butterflies assume 16 columns");

__global__ void __launch_bounds__(kMeanThreads)
lsm_mean_kernel(const double* __restrict__ block_sums, int64_t n_blocks,
                int d, int64_t n_paths, double* __restrict__ mean,
                double* __restrict__ scale) {
  __shared__ double red[kMeanThreads / 64][2 * kMaxAssets];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int64_t total = n_blocks * 2 * kMaxAssets;
  double acc = 0;
#pragma unroll 8
  for (int64_t i = t; i < total; i += kMeanThreads) acc += block_sums[i];
  acc += __shfl_xor(acc, 16, 64);
  acc += __shfl_xor(acc, 32, 64);
  if (lane < 2 * kMaxAssets) red[wave][lane] = acc;
  __syncthreads();
  if (t >= d) return;
  double s1 = 0, s2 = 0;
  for (int w = 0; w < kMeanThreads / 64; ++w) {
    s1 += red[w][t];
    s2 += red[w][kMaxAssets + t];
  }
  const double n = (double)n_paths;
  const double m = s1 / n;
  const double var = s2 / n - m * m;
  mean[t] = m;
  scale[t] = var > 1e-300 && sqrt(var) > 1e-12 * fabs(m) ? 1.0 / sqrt(var) : 1.0;
}

// Basis column k of one path from its per-asset power table.
__device__ __forceinline__ double basis_col(const double* pw, int deg1,
                                            const uint8_t* e, int d) {
  double v = 1.0;
  for (int j = 0; j < d; ++j) v *= pw[j * deg1 + e[j]];
  return v;
}

// G tile (I, J) partial over one chunk of paths.  Column K of the extended
// design is y (so G[:, K] = X^T W y).  grid = (n_chunks, n_pairs).
__global__ void __launch_bounds__(256)
lsm_gram_kernel(LsmArgs a, const double* __restrict__ exv,
                const double* __restrict__ y, int64_t n_calib,
                int64_t chunk, const int2* __restrict__ pairs,
                double* __restrict__ partial) {
  __shared__ double xi[kSubPaths][kTile];
  __shared__ double xj[kSubPaths][kTile];
  __shared__ double pw[kSubPaths][kMaxPw];
  __shared__ double wrow[kSubPaths];
  __shared__ double yrow[kSubPaths];
  __shared__ uint8_t expo[256 * kMaxAssets];
  // MFMA operands: wave w owns rows w*16 .. w*16 + 15 of the 64 x 64 tile
  // and all four 16-column subtiles; per step of 4 paths, lane l holds
  // A = X^T[w*16 + l%16][path l/16] and B = X[path l/16][t*16 + l%16]
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  const int2 pr = pairs[blockIdx.y];
  const int i0 = pr.x * kTile, j0 = pr.y * kTile;
  const int deg1 = a.deg + 1;
  const int kext = a.K + 1;
  for (int e = threadIdx.x; e < a.K * a.d; e += 256) expo[e] = a.expo[e];

  f64x4 acc[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) acc[t] = f64x4{0.0, 0.0, 0.0, 0.0};

  const int64_t p_begin = (int64_t)blockIdx.x * chunk;
  int64_t p_end = p_begin + chunk;
  if (p_end > n_calib) p_end = n_calib;
  for (int64_t p0 = p_begin; p0 < p_end; p0 += kSubPaths) {
    __syncthreads();
    // per-path powers and weights
    if (threadIdx.x < kSubPaths) {
      const int s = threadIdx.x;
      const int64_t p = p0 + s;
      double w = 0;
      if (p < p_end) {
        for (int j = 0; j < a.d; ++j) {
          const double c =
              (basis_input(a, path_x(a, p, j)) - a.mean[j]) * a.scale[j];
          double v = 1.0;
          for (int e = 0; e <= a.deg; ++e) {
            pw[s][j * deg1 + e] = v;
            v *= c;
          }
        }
        w = exv[p] > 0 ? 1.0 : 0.0;  // in-the-money rows only
        yrow[s] = y[p];
      } else {
        yrow[s] = 0;
      }
      wrow[s] = w;
    }
    __syncthreads();
    // design tiles: row s, column i0 + c and j0 + c (column K is y)
    for (int e = threadIdx.x; e < kSubPaths * kTile; e += 256) {
      const int s = e / kTile, c = e - s * kTile;
      const int ci = i0 + c, cj = j0 + c;
      double vi = 0, vj = 0;
      if (wrow[s] != 0) {
        if (ci < a.K) vi = basis_col(pw[s], deg1, expo + ci * a.d, a.d);
        else if (ci == a.K) vi = yrow[s];
        if (cj < a.K) vj = basis_col(pw[s], deg1, expo + cj * a.d, a.d);
        else if (cj == a.K) vj = yrow[s];
      }
      xi[s][c] = vi;
      xj[s][c] = vj;
    }
    __syncthreads();
    // G tile += X_i^T X_j over the 32 staged paths, 4 paths per MFMA
    // (v_mfma_f64_16x16x4_f64, float64 products and sums)
#pragma unroll
    for (int k0 = 0; k0 < kSubPaths; k0 += 4) {
      const double av = xi[k0 + lk][wave * 16 + lr];
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const double bv = xj[k0 + lk][t * 16 + lr];
        acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc[t], 0, 0, 0);
      }
    }
  }
  // partial[(chunk, pair)][64][64]
  double* out = partial + ((int64_t)blockIdx.x * gridDim.y + blockIdx.y) *
                              (kTile * kTile);
  // accumulator layout of the float64 16 x 16 block: lane l, entry r holds
  // row l/16 + 4*r, column l%16
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int r = 0; r < 4; ++r)
      out[(wave * 16 + lk + 4 * r) * kTile + t * 16 + lr] = acc[t][r];
  (void)kext;
}

// Sums chunk partials into the dense extended Gram G[kext][kext] (symmetric
// fill).  One thread per tile entry per pair.
__global__ void __launch_bounds__(256)
lsm_gram_reduce_kernel(const double* __restrict__ partial, int n_chunks,
                       int n_pairs, const int2* __restrict__ pairs, int kext,
                       double* __restrict__ g) {
  const int e = blockIdx.x * 256 + threadIdx.x;
  const int pair = blockIdx.y;
  if (e >= kTile * kTile) return;
  double s = 0;
#pragma unroll 8
  for (int c = 0; c < n_chunks; ++c)
    s += partial[((int64_t)c * n_pairs + pair) * (kTile * kTile) + e];
  const int2 pr = pairs[pair];
  const int r = pr.x * kTile + e / kTile, col = pr.y * kTile + e % kTile;
  if (r < kext && col < kext) {
    g[(int64_t)r * kext + col] = s;
    g[(int64_t)col * kext + r] = s;
  }
}

// ---------------------------------------------------------------------------
// Moment Gram.  For a monomial basis the Gram entry
//   G[k][l] = sum_p w_p prod_j c_pj^(e_kj + e_lj)
// depends only on e_k + e_l, a point of the box [0, 2 deg]^d, and the rhs
// entry G[k][K] = sum_p w_p y_p prod_j c_pj^e_kj on a point of [0, deg]^d.
// The box moments are one float64 MFMA product A^T B over the paths, the
// dimensions split in a row group (j < dA) and a column group:
//   A[p][alpha] = prod_{j < dA} c_pj^alpha_j            (alpha in [0,2deg]^dA)
//   B[p][beta]  = w_p prod_{j >= dA} c_pj^beta_j         (moment columns)
//               = w_p y_p prod_{j >= dA} c_pj^beta'_j    (rhs columns,
//                                                          beta' in [0,deg])
// so M[alpha][beta] holds every moment once; G is gathered from M.  For the
// tensor basis of 5 assets at degree 2 this is a 32 x 160 product per path
// in place of the 244 x 244 Gram.
struct MomentPlan {
  int small;      // 1: every moment in registers (lsm_moment_small_kernel)
  int dA;         // dimensions in the row group
  int radix;      // 2 deg + 1
  int ra, rap;    // rows radix^dA, padded to a multiple of 16
  int cbm, cby;   // moment columns radix^(d - dA), rhs columns (deg+1)^(d-dA)
  int cbp;        // cbm + cby padded to a multiple of 16
  int nblk;       // 16 x 16 output blocks: (rap / 16) * (cbp / 16)
  int lds;        // dynamic LDS bytes of lsm_moment_kernel
};
constexpr int kMomMaxBlocks = 64;  // 16 accumulator blocks per wave

constexpr int kMomThreads = 512;
constexpr int kMomMaxLds = 160 * 1024;  // gfx950 LDS per workgroup
constexpr int kMomMaxHi = 128;          // high-digit combinations per group

// Columns of one dimension group for the staged paths: column
// sum_q digit_q rad^q (digit q of dimension j0 + q) is
// wgt_s prod_q c_s,j0+q ^ digit_q.  One item per (path, high digits): the
// high digits' product times the rad powers of dimension j0.
__device__ __forceinline__ void moment_group(
    const double* __restrict__ pwb, int npw, int R,
    const double* __restrict__ wgt, int j0, int nd, int rad,
    const uint8_t* __restrict__ hioff, double* __restrict__ dst, int ld,
    int tid) {
  if (nd == 0) {
    if (tid < kSubPaths) dst[tid * ld] = wgt ? wgt[tid] : 1.0;
    return;
  }
  int n_hi = 1;
  for (int q = 1; q < nd; ++q) n_hi *= rad;
  for (int it = tid; it < kSubPaths * n_hi; it += kMomThreads) {
    const int s = it & (kSubPaths - 1), hi = it / kSubPaths;
    const double* pr = pwb + s * npw;
    double base = wgt ? wgt[s] : 1.0;
    for (int q = 0; q < nd - 1; ++q)
      base *= pr[hioff[hi * kMaxAssets + q]];
    const double* p0 = pr + j0 * R;
    double* o = dst + s * ld + hi * rad;
    for (int e = 0; e < rad; ++e) o[e] = base * p0[e];
  }
}

// The product runs as a pipeline over 32-path sub-batches with one barrier
// per phase: in phase i the waves multiply the tiles of sub-batch i, build
// the tiles of sub-batch i + 1 from its powers, write the powers of
// sub-batch i + 2 from registers and load the inputs of sub-batch i + 3.
// Tiles, powers and weights are double-buffered in LDS; the tiles' leading
// dimensions are odd, so the 32 paths of a column fall in distinct banks.
__global__ void __launch_bounds__(kMomThreads)
lsm_moment_kernel(LsmArgs a, MomentPlan mp, const double* __restrict__ exv,
                  const double* __restrict__ y, int64_t n_calib,
                  int64_t chunk, double* __restrict__ partial) {
  extern __shared__ double mlds[];
  const int R = mp.radix, npw = a.d * R, dB = a.d - mp.dA;
  const int lda = mp.rap + 1, ldb = mp.cbp + 1;
  const int tid = threadIdx.x;
  double* const pw0 = mlds;                        // [2][kSubPaths][npw]
  double* const xa0 = pw0 + 2 * kSubPaths * npw;   // [2][kSubPaths][lda]
  double* const xb0 = xa0 + 2 * kSubPaths * lda;   // [2][kSubPaths][ldb]
  double* const wv0 = xb0 + 2 * kSubPaths * ldb;   // [2][2][kSubPaths]
  // high-digit offsets into a path's power row (j * radix + power) of the
  // groups A, B moments, B rhs: [3][kMomMaxHi][kMaxAssets]
  uint8_t* hioff = (uint8_t*)(wv0 + 4 * kSubPaths);
  {
    const int j0s[3] = {0, mp.dA, mp.dA};
    const int nds[3] = {mp.dA, dB, dB};
    const int rads[3] = {R, R, a.deg + 1};
    for (int gq = 0; gq < 3; ++gq) {
      int n_hi = 1;
      for (int q = 1; q < nds[gq]; ++q) n_hi *= rads[gq];
      for (int hi = tid; hi < n_hi; hi += kMomThreads) {
        int rem = hi;
        for (int q = 1; q < nds[gq]; ++q) {
          hioff[(gq * kMomMaxHi + hi) * kMaxAssets + q - 1] =
              (uint8_t)((j0s[gq] + q) * R + rem % rads[gq]);
          rem /= rads[gq];
        }
      }
    }
  }
  // padding columns stay zero (both buffers)
  for (int e = tid; e < 2 * kSubPaths * lda; e += kMomThreads)
    if (e % lda >= mp.ra) xa0[e] = 0.0;
  for (int e = tid; e < 2 * kSubPaths * ldb; e += kMomThreads)
    if (e % ldb >= mp.cbm + mp.cby) xb0[e] = 0.0;

  // the wave index is uniform: an SGPR keeps the block guards scalar
  const int lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int lr = lane & 15, lk = lane >> 4;
  const int ncb = mp.cbp >> 4;
  constexpr int kPerWave = kMomMaxBlocks / (kMomThreads / 64);
  const int nmine = (mp.nblk - wave + (kMomThreads / 64) - 1) /
                    (kMomThreads / 64);  // blocks of this wave
  // LDS offsets of this lane's A and B operands in block wave + 8 q
  int aoff[kPerWave], boff[kPerWave];
#pragma unroll
  for (int q = 0; q < kPerWave; ++q) {
    const int blk = wave + (kMomThreads / 64) * q;
    const int bi = blk / ncb, bj = blk - bi * ncb;
    aoff[q] = lk * lda + bi * 16 + lr;
    boff[q] = lk * ldb + bj * 16 + lr;
  }
  f64x4 acc[kPerWave];
#pragma unroll
  for (int q = 0; q < kPerWave; ++q) acc[q] = f64x4{0.0, 0.0, 0.0, 0.0};

  const int64_t p_begin = (int64_t)blockIdx.x * chunk;
  int64_t p_end = p_begin + chunk;
  if (p_end > n_calib) p_end = n_calib;
  const int nsub = p_end > p_begin
      ? (int)((p_end - p_begin + kSubPaths - 1) / kSubPaths) : 0;

  // power threads: one per (path s, dimension j) of a sub-batch
  const bool pthread = tid < kSubPaths * a.d;
  const int ps = tid & (kSubPaths - 1), pj = tid / kSubPaths;
  double in_e = 0.0, in_x = 0.0, in_y = 0.0;  // inputs of the next powers
  auto load = [&](int i) {
    const int64_t p = p_begin + (int64_t)i * kSubPaths + ps;
    in_e = 0.0;
    if (pthread && i < nsub && p < p_end) {
      in_e = exv[p];
      in_x = path_x(a, p, pj);
      if (pj == 0) in_y = y[p];
    }
  };
  // powers 0 .. 2 deg of the centred, scaled basis variable; rows of paths
  // past the chunk or out of the money are zero
  auto powers = [&](int buf) {
    if (!pthread) return;
    const double w = in_e > 0 ? 1.0 : 0.0;  // in-the-money rows only
    double* pr = pw0 + (buf * kSubPaths + ps) * npw + pj * R;
    if (w != 0.0) {
      const double cv = (basis_input(a, in_x) - a.mean[pj]) * a.scale[pj];
      double v = 1.0;
      for (int m = 0; m < R; ++m) {
        pr[m] = v;
        v *= cv;
      }
    } else {
      for (int m = 0; m < R; ++m) pr[m] = 0.0;
    }
    if (pj == 0) {
      double* wv = wv0 + buf * 2 * kSubPaths;
      wv[ps] = w;
      wv[kSubPaths + ps] = w != 0.0 ? in_y : 0.0;
    }
  };
  // the A and B tiles of one sub-batch
  auto tiles = [&](int buf) {
    const double* pwb = pw0 + buf * kSubPaths * npw;
    const double* wv = wv0 + buf * 2 * kSubPaths;
    double* xa = xa0 + buf * kSubPaths * lda;
    double* xb = xb0 + buf * kSubPaths * ldb;
    moment_group(pwb, npw, R, nullptr, 0, mp.dA, R, hioff, xa, lda, tid);
    moment_group(pwb, npw, R, wv, mp.dA, dB, R, hioff + kMomMaxHi * kMaxAssets,
                 xb, ldb, tid);
    moment_group(pwb, npw, R, wv + kSubPaths, mp.dA, dB, a.deg + 1,
                 hioff + 2 * kMomMaxHi * kMaxAssets, xb + mp.cbm, ldb, tid);
  };
  // M += A^T B over one sub-batch, 4 paths per MFMA; block wave + 8 q of
  // the (rap / 16) x (cbp / 16) grid of 16 x 16 blocks
  auto multiply = [&](int buf) {
    const double* xa = xa0 + buf * kSubPaths * lda;
    const double* xb = xb0 + buf * kSubPaths * ldb;
#pragma unroll
    for (int k0 = 0; k0 < kSubPaths; k0 += 4) {
      double av[kPerWave], bv[kPerWave];
#pragma unroll
      for (int q = 0; q < kPerWave; ++q) {
        if (q < nmine) {
          av[q] = xa[k0 * lda + aoff[q]];
          bv[q] = xb[k0 * ldb + boff[q]];
        }
      }
#pragma unroll
      for (int q = 0; q < kPerWave; ++q)
        if (q < nmine)
          acc[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[q], bv[q], acc[q],
                                                        0, 0, 0);
    }
  };

  load(0);
  __syncthreads();  // offset tables and padding
  powers(0);
  load(1);
  __syncthreads();
  if (nsub > 0) tiles(0);
  powers(1);
  load(2);
  __syncthreads();
  for (int i = 0; i < nsub; ++i) {
    multiply(i & 1);
    if (i + 1 < nsub) tiles((i + 1) & 1);
    if (i + 2 < nsub) {
      powers(i & 1);
      load(i + 3);
    }
    __syncthreads();
  }
  // partial[chunk][rap][cbp]; accumulator entry r of lane l is row
  // l/16 + 4 r, column l%16 of its block
  double* out = partial + (int64_t)blockIdx.x * (mp.rap * mp.cbp);
#pragma unroll
  for (int q = 0; q < kPerWave; ++q) {
    if (q < nmine) {
      const int blk = wave + (kMomThreads / 64) * q;
      const int bi = blk / ncb, bj = blk - bi * ncb;
#pragma unroll
      for (int r = 0; r < 4; ++r)
        out[(bi * 16 + lk + 4 * r) * mp.cbp + bj * 16 + lr] = acc[q][r];
    }
  }
}

// The moment product specialised on its plan (D assets, degree DEG, DA
// dimensions in the row group, DA <= 2 and at most 3 column dimensions).
// Each wave owns every 16 x 16 output block and a private run of 64-path
// sub-batches (sub-batches wave, wave + W, ... of the chunk for W waves),
// so the product runs without a workgroup barrier until the fixed-order sum
// of the waves' blocks at the end.  Per sub-batch each lane writes its
// path's record of factor tables to the wave's LDS slice:
//   TA[alpha]  prod_{j < DA} c_j^alpha_j       (alpha < R^DA, R = 2 DEG + 1)
//   TU[u]      prod of the column dimensions but the last, c^u_q
//              (u < R^NU, NU = D - DA - 1 of them)
//   TW[e]      w c_{D-1}^e       (w = 1 in the money, else 0; e < R)
//   TY[e]      w y c_{D-1}^e     (e <= DEG)
//   and a zero,
// so for each group of 4 paths a lane's A operand (row alpha) is one entry
// TA[alpha] and its B operand (column beta) the product TU[u] TW[e] of a
// moment column or TU[u'] TY[e'] of an rhs column (digits of beta in radix
// R, or in radix DEG + 1 re-encoded for TU); padding rows and columns read
// the zero.  The offsets are fixed by the lane and the block.  The output
// is the same [chunk][rap][cbp] partial as lsm_moment_kernel.
// Plans of more than 8 blocks run one wave per SIMD (4 per workgroup), so
// a wave has registers for its accumulators (8 per block), the operand
// offsets and the next group's reads in flight; smaller ones run 8 waves.
// The inputs of a wave's next sub-batch are loaded while it multiplies the
// current one.
constexpr int kFastSub = 64;    // paths per wave sub-batch
constexpr int kFastMaxBlocks = 20;

constexpr int cpow(int b, int e) { return e == 0 ? 1 : b * cpow(b, e - 1); }

template <int D, int DEG, int DA>
struct FastPlan {
  static constexpr int R = 2 * DEG + 1, NBD = D - DA, NU = NBD - 1;
  static constexpr int RA = cpow(R, DA), CBM = cpow(R, NBD);
  static constexpr int CBY = cpow(DEG + 1, NBD);
  static constexpr int RAP = (RA + 15) / 16 * 16;
  static constexpr int CBP = (CBM + CBY + 15) / 16 * 16;
  static constexpr int NBI = RAP / 16, NBJ = CBP / 16;
  // record layout (doubles)
  static constexpr int TA = 0, TU = RA, TW = TU + cpow(R, NU);
  static constexpr int TY = TW + R, ZERO = TY + DEG + 1;
  static constexpr int S = (ZERO + 1) | 1;  // odd record stride
  static constexpr int WAVES = NBI * NBJ > 8 ? 4 : 8;
  static constexpr int THREADS = 64 * WAVES;
  static constexpr int REC = WAVES * kFastSub * S;
  static constexpr int LDS = REC > RAP * CBP ? REC : RAP * CBP;
  // operand reads per group of 4 paths, spread over half of its MFMAs
  static constexpr int HALF = (NBI * NBJ + 1) / 2;
  static constexpr int READS_PER_MFMA =
      (NBI + NBJ * (NU > 0 ? 2 : 1) + HALF - 1) / HALF;
  static_assert(DA >= 1 && DA <= 2 && NBD >= 1 && NU <= 2,
                "factor tables of one or two dimensions");
  static_assert(NBI * NBJ <= kFastMaxBlocks, "accumulators of one wave");
  static_assert(LDS * 8 <= 160 * 1024, "LDS of one workgroup");
};

// The product of the powers of one or two dimensions: t[a + R b] =
// p0[a] p1[b] (N = 2) or t[a] = p0[a] (N = 1).
template <int N, int R>
__device__ __forceinline__ void factor_table(double* t, const double* p0,
                                             const double* p1) {
  if constexpr (N == 1) {
#pragma unroll
    for (int e = 0; e < R; ++e) t[e] = p0[e];
  } else if constexpr (N == 2) {
#pragma unroll
    for (int b = 0; b < R; ++b)
#pragma unroll
      for (int e = 0; e < R; ++e) t[e + R * b] = p0[e] * p1[b];
  }
}

// LOG: the basis variable is ln of the stored value (basis_log 1).
template <int D, int DEG, int DA, bool LOG>
__global__ void __launch_bounds__((FastPlan<D, DEG, DA>::THREADS))
lsm_moment_fast_kernel(LsmArgs a, const double* __restrict__ exv,
                       const double* __restrict__ y, int64_t n_calib,
                       int64_t chunk, double* __restrict__ partial) {
  using P = FastPlan<D, DEG, DA>;
  constexpr int R = P::R, NU = P::NU, NBI = P::NBI, NBJ = P::NBJ;
  constexpr int S = P::S;
  __shared__ double lds[P::LDS];
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int g = lane >> 4, r = lane & 15;

  // lane-constant record offsets of the operands
  int aoff[NBI], uoff[NBJ], woff[NBJ];
#pragma unroll
  for (int bi = 0; bi < NBI; ++bi) {
    const int alpha = bi * 16 + r;
    aoff[bi] = alpha < P::RA ? P::TA + alpha : P::ZERO;
  }
#pragma unroll
  for (int bj = 0; bj < NBJ; ++bj) {
    const int beta = bj * 16 + r;
    int u = 0, o = P::ZERO;
    if (beta < P::CBM) {
      u = beta % cpow(R, NU);
      o = P::TW + beta / cpow(R, NU);
    } else if (beta < P::CBM + P::CBY) {
      int rem = beta - P::CBM;
#pragma unroll
      for (int q = 0, m = 1; q < NU; ++q, m *= R) {
        u += (rem % (DEG + 1)) * m;
        rem /= DEG + 1;
      }
      o = P::TY + rem;
    }
    uoff[bj] = P::TU + u;
    woff[bj] = o;
  }

  f64x4 acc[NBI][NBJ];
#pragma unroll
  for (int bi = 0; bi < NBI; ++bi)
#pragma unroll
    for (int bj = 0; bj < NBJ; ++bj) acc[bi][bj] = f64x4{0.0, 0.0, 0.0, 0.0};
  // the loop-carried accumulators live in accumulator registers on entry
  // to and around the sub-batch loop (no copies to and from VGPRs at its
  // edges)
  auto pin_acc = [&]() {
#pragma unroll
    for (int bi = 0; bi < NBI; ++bi)
#pragma unroll
      for (int bj = 0; bj < NBJ; ++bj) asm volatile("" : "+a"(acc[bi][bj]));
  };
  pin_acc();

  const int64_t p_begin = (int64_t)blockIdx.x * chunk;
  int64_t p_end = p_begin + chunk;
  if (p_end > n_calib) p_end = n_calib;
  const int nsub = p_end > p_begin
      ? (int)((p_end - p_begin + kFastSub - 1) / kFastSub) : 0;
  double* const rec = lds + wave * (kFastSub * S);
  // inputs of sub-batch i: the exercise value, y and the D stored values.
  // The loop body has no branch (rows past the chunk read the chunk's first
  // row and get w = 0), so the accumulators stay in place across it.
  double in_e = 0.0, in_y = 0.0, in_x[D], mu[D], sc[D];
#pragma unroll
  for (int j = 0; j < D; ++j) {
    mu[j] = a.mean[j];
    sc[j] = a.scale[j];
  }
  auto load = [&](int i) {
    int64_t p = p_begin + (int64_t)i * kFastSub + lane;
    const bool ok = i < nsub && p < p_end;
    p = ok ? p : p_begin;
    in_e = exv[p];
    in_y = y[p];
#pragma unroll
    for (int j = 0; j < D; ++j) in_x[j] = path_x(a, p, j);
    in_e = ok ? in_e : 0.0;
  };
  if (nsub > 0) load(wave);
  for (int i = wave; i < nsub; i += P::WAVES) {
    // this lane's path record; rows past the chunk or out of the money
    // have w = 0 and unit powers
    const bool itm = in_e > 0;
    double pw[D][R];
#pragma unroll
    for (int j = 0; j < D; ++j) {
      const double b = LOG ? log(in_x[j]) : in_x[j];
      const double c = itm ? (b - mu[j]) * sc[j] : 0.0;
      pw[j][0] = 1.0;
#pragma unroll
      for (int e = 1; e < R; ++e) pw[j][e] = pw[j][e - 1] * c;
    }
    const double w = itm ? 1.0 : 0.0, wy = itm ? in_y : 0.0;
    load(i + P::WAVES);
    double* mine = rec + lane * S;
    factor_table<DA, R>(mine + P::TA, pw[0], pw[DA > 1 ? 1 : 0]);
    factor_table<NU, R>(mine + P::TU, pw[DA], pw[NU > 1 ? DA + 1 : DA]);
#pragma unroll
    for (int e = 0; e < R; ++e) mine[P::TW + e] = w * pw[D - 1][e];
#pragma unroll
    for (int e = 0; e <= DEG; ++e) mine[P::TY + e] = wy * pw[D - 1][e];
    mine[P::ZERO] = 0.0;
    wave_lds_sync();
    // operands of the group of paths 4 k0 .. 4 k0 + 3
    auto fetch = [&](int k0, double (&av)[NBI], double (&bv)[NBJ]) {
      const double* pr = rec + (k0 * 4 + g) * S;
#pragma unroll
      for (int bi = 0; bi < NBI; ++bi) av[bi] = pr[aoff[bi]];
#pragma unroll
      for (int bj = 0; bj < NBJ; ++bj)
        bv[bj] = NU > 0 ? pr[uoff[bj]] * pr[woff[bj]] : pr[woff[bj]];
    };
    // software pipeline: the reads and products of group k0 + 1 issue
    // between the MFMAs of group k0
    double av[NBI], bv[NBJ];
    fetch(0, av, bv);
#pragma unroll
    for (int k0 = 0; k0 < kFastSub / 4; ++k0) {
      double an[NBI], bn[NBJ];
      if (k0 + 1 < kFastSub / 4) fetch(k0 + 1, an, bn);
#pragma unroll
      for (int bi = 0; bi < NBI; ++bi)
#pragma unroll
        for (int bj = 0; bj < NBJ; ++bj)
          acc[bi][bj] = mfma_f64(av[bi], bv[bj], acc[bi][bj]);
      if (k0 + 1 < kFastSub / 4) {
        // first half of the MFMAs: the reads of group k0 + 1 in flight;
        // second half: their products
#pragma unroll
        for (int m = 0; m < P::HALF; ++m) {
          __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // MFMA
          __builtin_amdgcn_sched_group_barrier(0x100, P::READS_PER_MFMA, 0);
        }
#pragma unroll
        for (int m = P::HALF; m < NBI * NBJ; ++m) {
          __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // MFMA
          __builtin_amdgcn_sched_group_barrier(0x002, 1, 0);  // VALU
        }
#pragma unroll
        for (int bi = 0; bi < NBI; ++bi) av[bi] = an[bi];
#pragma unroll
        for (int bj = 0; bj < NBJ; ++bj) bv[bj] = bn[bj];
      }
      __builtin_amdgcn_sched_barrier(0);  // groups do not mix
    }
    wave_lds_sync();  // the next sub-batch overwrites the records
    pin_acc();
  }

  // fixed-order sum of the waves' blocks: accumulator entry q of lane l is
  // row l / 16 + 4 q, column l % 16 of its block
  __syncthreads();
  double* const buf = lds;  // [rap][cbp]
  for (int wv = 0; wv < P::WAVES; ++wv) {
    if (wave == wv) {
#pragma unroll
      for (int bi = 0; bi < NBI; ++bi)
#pragma unroll
        for (int bj = 0; bj < NBJ; ++bj)
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            double* o = buf + (bi * 16 + g + 4 * q) * P::CBP + bj * 16 + r;
            *o = wv == 0 ? acc[bi][bj][q] : *o + acc[bi][bj][q];
          }
    }
    __syncthreads();
  }
  double* out = partial + (int64_t)blockIdx.x * (P::RAP * P::CBP);
  for (int e = tid; e < P::RAP * P::CBP; e += P::THREADS) out[e] = buf[e];
}

// Launches the specialised moment product when one is built for the plan;
// false otherwise (the generic kernel then runs).
template <int D, int DEG, int DA>
bool try_moment_fast(const MomentPlan& mp, int d, int deg, const LsmArgs& a,
                     const double* exv, const double* y, int64_t n_calib,
                     int64_t chunk, int n_chunks, double* partial,
                     hipStream_t s) {
  using P = FastPlan<D, DEG, DA>;
  if (d != D || deg != DEG || mp.dA != DA || mp.rap != P::RAP ||
      mp.cbp != P::CBP || mp.cbm != P::CBM)
    return false;
  if (a.basis_log == 1)
    hipLaunchKernelGGL((lsm_moment_fast_kernel<D, DEG, DA, true>), n_chunks,
                       P::THREADS, 0, s, a, exv, y, n_calib, chunk, partial);
  else
    hipLaunchKernelGGL((lsm_moment_fast_kernel<D, DEG, DA, false>), n_chunks,
                       P::THREADS, 0, s, a, exv, y, n_calib, chunk, partial);
  return true;
}

bool launch_moment_fast(const MomentPlan& mp, int d, int deg,
                        const LsmArgs& a, const double* exv, const double* y,
                        int64_t n_calib, int64_t chunk, int n_chunks,
                        double* partial, hipStream_t s) {
  return try_moment_fast<2, 2, 1>(mp, d, deg, a, exv, y, n_calib, chunk,
                                  n_chunks, partial, s) ||
         try_moment_fast<3, 2, 2>(mp, d, deg, a, exv, y, n_calib, chunk,
                                  n_chunks, partial, s) ||
         try_moment_fast<4, 2, 2>(mp, d, deg, a, exv, y, n_calib, chunk,
                                  n_chunks, partial, s) ||
         try_moment_fast<5, 2, 2>(mp, d, deg, a, exv, y, n_calib, chunk,
                                  n_chunks, partial, s);
}

// One asset: the moments are c^0 .. c^(2 deg) and y c^0 .. y c^deg
// (5 + 3 at degree 2).  Each thread accumulates them for its paths in
// registers by running products; a fixed-order workgroup sum gives one row
// of partials [chunk][cbp].
constexpr int kSmallMoments = 40;

// NM: the accumulators a thread keeps (at least the plan's moment count),
// so degree 2 holds 8 float64 sums in registers rather than 40.
template <int NM>
__global__ void __launch_bounds__(256)
lsm_moment_small_kernel(LsmArgs a, MomentPlan mp,
                        const double* __restrict__ exv,
                        const double* __restrict__ y, int64_t n_calib,
                        int64_t chunk, double* __restrict__ partial) {
  __shared__ double red[4][NM];
  const int nm = mp.cbp, nx = mp.cbm;  // nx = 2 deg + 1 moments, then rhs
  double acc[NM];
#pragma unroll
  for (int m = 0; m < NM; ++m) acc[m] = 0.0;
  const int64_t p_begin = (int64_t)blockIdx.x * chunk;
  int64_t p_end = p_begin + chunk;
  if (p_end > n_calib) p_end = n_calib;
  for (int64_t p = p_begin + threadIdx.x; p < p_end; p += 256) {
    if (!(exv[p] > 0)) continue;  // in-the-money rows only
    const double x = path_x(a, p, 0);
    const double c = (basis_input(a, x) - a.mean[0]) * a.scale[0];
    double v = 1.0, u = y[p];
#pragma unroll
    for (int m = 0; m < NM; ++m) {
      if (m < nx) {
        acc[m] += v;
        v *= c;
      } else if (m < nm) {
        acc[m] += u;
        u *= c;
      }
    }
  }
  // fixed-order sums: wave butterflies, then the four waves in order
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int m = 0; m < NM; ++m) {
    if (m < nm) {
      double v = acc[m];
      for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
      if (lane == 0) red[wave][m] = v;
    }
  }
  __syncthreads();
  if (threadIdx.x < nm)
    partial[(int64_t)blockIdx.x * nm + threadIdx.x] =
        ((red[0][threadIdx.x] + red[1][threadIdx.x]) + red[2][threadIdx.x]) +
        red[3][threadIdx.x];
}

// Sums the chunk partials of the moment product (fixed order).
__global__ void __launch_bounds__(256)
lsm_moment_reduce_kernel(const double* __restrict__ partial, int n_chunks,
                         int size, double* __restrict__ m) {
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= size) return;
  double s = 0;
#pragma unroll 8
  for (int c = 0; c < n_chunks; ++c) s += partial[(int64_t)c * size + e];
  m[e] = s;
}

// Gathers the extended Gram G[kext][kext] from the moments: one thread per
// entry.  An exponent above the degree has no moment: its row becomes NaN.
__global__ void __launch_bounds__(256)
lsm_moment_gram_kernel(const double* __restrict__ m, MomentPlan mp,
                       const uint8_t* __restrict__ expo, int K, int d,
                       int deg, double* __restrict__ g) {
  const int kext = K + 1;
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= kext * kext) return;
  const int k = e / kext, l = e - k * kext;
  double v = 0.0;
  if (k < K || l < K) {
    const bool rhs = k == K || l == K;
    const int r0 = rhs ? (k == K ? l : k) : k;
    bool ok = true;
    int alpha = 0, beta = 0, mul = 1;
    for (int j = 0; j < mp.dA; ++j) {
      const int ek = expo[r0 * d + j];
      const int el = rhs ? 0 : expo[l * d + j];
      ok = ok && ek <= deg && el <= deg;
      alpha += (ek + el) * mul;
      mul *= mp.radix;
    }
    const int rad = rhs ? deg + 1 : mp.radix;
    mul = 1;
    for (int j = mp.dA; j < d; ++j) {
      const int ek = expo[r0 * d + j];
      const int el = rhs ? 0 : expo[l * d + j];
      ok = ok && ek <= deg && el <= deg;
      beta += (ek + el) * mul;
      mul *= rad;
    }
    if (rhs) beta += mp.cbm;
    const double mv = m[ok ? alpha * mp.cbp + beta : 0];
    v = ok ? mv : __longlong_as_double(0x7ff8000000000000LL);
  }
  g[e] = v;
}

// The moment plan for (d, deg): the split with the smallest padded product
// whose buffers fit the LDS of a workgroup and 64 output blocks, and only
// when the product is cheaper than the dense Gram tiles (whose partial
// buffer then holds the moment partials and their sum).  False when none
// qualifies.
bool moment_plan(int d, int deg, int n_pairs, MomentPlan* out) {
  const int R = 2 * deg + 1;
  if (d * R > 256) return false;
  if (d == 1 && R + deg + 1 <= kSmallMoments) {
    // one asset: every moment in registers
    MomentPlan p{};
    p.small = 1;
    p.dA = 0;
    p.radix = R;
    p.ra = p.rap = 1;
    p.cbm = R;
    p.cby = deg + 1;
    p.cbp = R + deg + 1;
    *out = p;
    return true;
  }
  bool found = false;
  MomentPlan best{};
  int64_t ra = 1;
  for (int dA = 0; dA <= d; ++dA, ra *= R) {
    if (ra > 128) break;
    int64_t cbm = 1, cby = 1;
    for (int j = dA; j < d; ++j) {
      cbm *= R;
      cby *= deg + 1;
    }
    if (cbm + cby > 256) continue;
    MomentPlan p{};
    p.dA = dA;
    p.radix = R;
    p.ra = (int)ra;
    p.rap = (int)((ra + 15) / 16 * 16);
    p.cbm = (int)cbm;
    p.cby = (int)cby;
    p.cbp = (int)((cbm + cby + 15) / 16 * 16);
    p.nblk = (p.rap / 16) * (p.cbp / 16);
    if (p.nblk > kMomMaxBlocks) continue;
    p.lds = (2 * kSubPaths * (d * R + p.rap + 1 + p.cbp + 1) +
             4 * kSubPaths) * 8 + 3 * kMomMaxHi * kMaxAssets;
    if (p.lds > kMomMaxLds) continue;
    // high-digit combinations of each group
    int64_t hi_a = 1, hi_b = 1;
    for (int j = 1; j < dA; ++j) hi_a *= R;
    for (int j = dA + 1; j < d; ++j) hi_b *= R;
    if (hi_a > kMomMaxHi || hi_b > kMomMaxHi) continue;
    if (!found || (int64_t)p.rap * p.cbp < (int64_t)best.rap * best.cbp) {
      best = p;
      found = true;
    }
  }
  if (!found) return false;
  const int64_t size = (int64_t)best.rap * best.cbp;
  if (size * 257 > (int64_t)256 * n_pairs * kTile * kTile) return false;
  *out = best;
  return true;
}

// Column-scaled Cholesky with pivot truncation and the two triangular
// solves, one workgroup of 512 threads, float64.  beta[K] out.
//
// The factorization runs on the extended matrix [[G, b], [b^T, .]] of order
// n = K + 1 <= 256 (b = X^T W y is column K of the Gram buffer): its last
// row of L is z = L^-1 b, so the forward solve comes out of the
// factorization.  Columns are scaled by 1 / sqrt(G_kk); a pivot at or below
// rtol * (largest scaled diagonal) drops its column (L column zero), which
// leaves the fitted values of the pseudo-inverse solution unchanged.
//
// Blocked right-looking form on 16 x 16 tiles (T = ceil(n / 16) tile rows):
// the lower tiles stay in registers for the whole solve, in the float64
// MFMA accumulator layout (lane l: column l % 16, rows l / 16 + 4 r); wave
// w of eight keeps tile rows 15 - w and w (17 tiles, two waves per SIMD so
// they fit its registers).  Per panel p:
//   1. the owner of tile (p, p) factors it in registers (16 sequential
//      columns on one wave, operands exchanged between lanes, no LDS round
//      and no workgroup barrier) and forms Y = L_pp^-1 in the same column
//      steps, with zero rows and columns for dropped pivots;
//   2. each panel tile (a, p) becomes X = A_ap Y^T (4 MFMAs), staged in LDS;
//   3. each trailing tile (a, b), a >= b > p, takes -= X_ap X_bp^T (4 MFMAs).
// Two workgroup barriers per panel in place of one per column.  The back
// solve L^T w = z runs by blocks from the last: w_b = Y_bb^T r_b, then every
// tile (b, c), c < b, subtracts L_bc^T w_b from r_c (one tile per c and
// step, so the order of the sums is fixed; the tiles' row sums meet in LDS
// and are added all at once).  beta = D w.
constexpr int kSolveWaves = 8;
constexpr int kSolveThreads = 64 * kSolveWaves;
constexpr int kBT = 16;         // tile edge
constexpr int kBLd = 17;        // LDS row stride of a staged tile (doubles)
constexpr int kBTiles = 17;     // two tile rows per wave: 136 / 8

__global__ void __launch_bounds__(kSolveThreads)
lsm_solve_kernel(const double* __restrict__ g, int K, double rtol,
                 double* __restrict__ beta, int* __restrict__ rank_out) {
  __shared__ double dsc[256];
  __shared__ double Ys[16][kBT * kBLd];      // Y_pp of every panel
  __shared__ double P[16][kBT * kBLd];       // panel tiles X_ap
  __shared__ double rv[256], wv[256];
  __shared__ double maxdiag_s;
  __shared__ int rank_s;
  const int n = K + 1, kext = K + 1;
  const int T = (n + kBT - 1) / kBT;
  const int t = threadIdx.x, lane = t & 63;
  const int wave = __builtin_amdgcn_readfirstlane(t >> 6);
  const int cl = lane & 15, rl = lane >> 4;
  if (t < 256) {
    double sc = 0.0;
    if (t < K) {
      const double d = g[t * kext + t];
      sc = d > 0 ? 1.0 / sqrt(d) : 0.0;
    } else if (t == K) {
      sc = 1.0;
    }
    dsc[t] = sc;
    rv[t] = 0.0;
    wv[t] = 0.0;
  }
  if (t == 0) rank_s = 0;
  __syncthreads();
  // the largest scaled diagonal entry (wave 0)
  if (t < 64) {
    double m = 0.0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int i = t + 64 * q;
      if (i < K) m = fmax(m, g[i * kext + i] * (dsc[i] * dsc[i]));
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
    if (t == 0) maxdiag_s = m;
  }
  // this wave's tiles q: (a, b) = (tile_a(q), tile_b(q)), a = -1 past the
  // last tile (wave-uniform scalars), scaled
  // wave w owns tile rows 15 - w (slots 0 .. 15 - w) and w (slots
  // 16 - w .. 16): 17 tiles each; rows at or past T hold nothing
  // (wo is the wave index passed through an empty asm statement inside the
  // loops, so the 34 tile coordinates and their LDS addresses are rebuilt
  // where used rather than hoisted out of the panel loop, which would
  // exhaust the registers)
  int wo = wave;
  auto tile_a = [&](int q) {
    const int a = q < 16 - wo ? 15 - wo : wo;
    return a < T ? a : -1;
  };
  auto tile_b = [&](int q) { return q < 16 - wo ? q : q - 16 + wo; };
  auto refresh = [&]() {
    wo = wave;
    asm volatile("" : "+s"(wo));
  };
  f64x4 tile[kBTiles];
#pragma unroll
  for (int q = 0; q < kBTiles; ++q) {
    const int a = tile_a(q), b = tile_b(q);
    f64x4 v = {0.0, 0.0, 0.0, 0.0};
    if (a >= 0) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = kBT * a + rl + 4 * r, col = kBT * b + cl;
        v[r] = (row < n && col < n)
                   ? g[row * kext + col] * (dsc[row] * dsc[col]) : 0.0;
      }
    }
    tile[q] = v;
    __builtin_amdgcn_sched_barrier(0);
  }
  __syncthreads();
  const double thr = rtol * maxdiag_s;
  // 1. factor diagonal tile (p, p) on its owner (wave p or 15 - p, slot 16
  // or p) in registers, in the tile's own layout (lane l: column l % 16,
  // rows l / 16 + 4 r), together with Y_pp = L_pp^-1 (zero rows and columns
  // where a pivot is dropped).  The tile is symmetric and both triangles are
  // updated alike, so row k of the trailing square equals its column k.
  // Column k: the pivot from its lane (v_readlane), its reciprocal root from
  // v_rsq_f64 and two Newton steps, L_ik from lane 16 (l / 16) + k (the same
  // row), L_jk = D_kj and the row k of R = I - (sum of the steps so far)
  // from lane 16 (k % 4) + l % 16 (the same column); then
  // D_ij -= L_ik L_jk (i, j > k) and R_ij -= L_ik Y_kj (i > k) with
  // Y_kj = R_kj / L_kk, the forward substitution L Y = I done column by
  // column alongside the factorization.
  auto factor_diag = [&](int pd) {
    const int qd = pd < 8 ? 16 : pd;
    f64x4 dv = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int q = 0; q < kBTiles; ++q)
      if (q == qd) dv = tile[q];
    f64x4 yv;
#pragma unroll
    for (int r = 0; r < 4; ++r) yv[r] = rl + 4 * r == cl ? 1.0 : 0.0;
    int kept = 0;
#pragma unroll
    for (int k = 0; k < kBT; ++k) {
      const int rk = k >> 2, gk = k & 3, src = 16 * gk + cl;
      const double dk = read_lane_f64(dv[rk], 16 * gk + k);
      double dik[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) dik[r] = __shfl(dv[r], (lane & 48) + k, 64);
      const double dkj = __shfl(dv[rk], src, 64);
      const double rkj = __shfl(yv[rk], src, 64);
      const bool keep = kBT * pd + k < K && dk > thr && dk > 0;
      double rinv = 0.0;
      if (keep) {
        double y = __builtin_amdgcn_rsq(dk);
        y = y * fma(-0.5 * dk * y, y, 1.5);
        y = y * fma(-0.5 * dk * y, y, 1.5);
        rinv = y;
      }
      const double root = dk * rinv;  // 0 for a dropped pivot
      kept += keep ? 1 : 0;
      // branch-free: rinv = 0 leaves the updates no-ops and zeroes row and
      // column k of L and Y
      const double ljk = dkj * rinv, ykj = rkj * rinv;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = rl + 4 * r;
        const double lik = dik[r] * rinv;
        double v = dv[r];
        v = (i > k && cl > k) ? v - lik * ljk : v;
        v = (cl == k && i > k) ? lik : v;
        v = (cl == k && i == k) ? root : v;
        dv[r] = v;
        double y = yv[r];
        y = i > k ? y - lik * ykj : y;
        y = i == k ? ykj : y;
        yv[r] = y;
      }
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) Ys[pd][(rl + 4 * r) * kBLd + cl] = yv[r];
    // L_pp (lower, zero above) back into the register tile
#pragma unroll
    for (int q = 0; q < kBTiles; ++q)
      if (q == qd)
#pragma unroll
        for (int r = 0; r < 4; ++r)
          tile[q][r] = rl + 4 * r >= cl ? dv[r] : 0.0;
    if (lane == 0) rank_s += kept;
  };
  // p = -1 only factors tile (0, 0)
  for (int p = -1; p < T; ++p) {
   if (p >= 0) {
    // Y_pp is ready (and the reads of the previous panel are done)
    __syncthreads();
    // 2. panel tiles X_ap = A_ap Y^T: A operand (row l % 16, k = 4 s + l / 16)
    // from the staged tile, B operand Y^T (k, column l % 16) = Y[l % 16][k]
    refresh();
#pragma unroll
    for (int q = 0; q < kBTiles; ++q) {
      const int a = tile_a(q), b = tile_b(q);
      if (a > p && b == p) {
        double* pa = P[a];
#pragma unroll
        for (int r = 0; r < 4; ++r) pa[(rl + 4 * r) * kBLd + cl] = tile[q][r];
        wave_lds_sync();
        f64x4 x = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int s4 = 0; s4 < kBT; s4 += 4)
          x = mfma_f64(pa[cl * kBLd + s4 + rl], Ys[p][cl * kBLd + s4 + rl], x);
        wave_lds_sync();
#pragma unroll
        for (int r = 0; r < 4; ++r) pa[(rl + 4 * r) * kBLd + cl] = x[r];
        tile[q] = x;
      }
      // one tile's loads at a time (hoisting all of them would exhaust
      // the registers)
      __builtin_amdgcn_sched_barrier(0);
    }
    __syncthreads();
   }
    // 3. trailing tiles (a, b), a >= b > p: C -= X_ap X_bp^T.  The owner
    // of tile (p + 1, p + 1) updates it first and factors it at once
    // (lookahead), then goes on with its other tiles.
    refresh();
    const int pn = p + 1;
    const bool next_owner = pn < T && wave == (pn < 8 ? pn : 15 - pn);
    auto trail = [&](int q) {
      const int a = tile_a(q), b = tile_b(q);
      const double* pa = P[a];
      const double* pb = P[b];
      f64x4 c = tile[q];
#pragma unroll
      for (int s4 = 0; s4 < kBT; s4 += 4)
        c = mfma_f64(-pa[cl * kBLd + s4 + rl], pb[cl * kBLd + s4 + rl], c);
      tile[q] = c;
    };
    if (next_owner) {
      const int qn = pn < 8 ? 16 : pn;
      if (p >= 0)
#pragma unroll
        for (int q = 0; q < kBTiles; ++q)
          if (q == qn) trail(q);
      factor_diag(pn);
      refresh();
    }
    if (p >= 0)
#pragma unroll
      for (int q = 0; q < kBTiles; ++q) {
        const int a = tile_a(q), b = tile_b(q);
        if (a >= 0 && b > p && !(next_owner && a == pn && b == pn)) trail(q);
        __builtin_amdgcn_sched_barrier(0);
      }
    // (the barrier at the top of the next panel orders these reads of P
    // before its writes)
  }
  __syncthreads();
  // r = z = row K of L
  const int bk = K / kBT, rk = K % kBT;
#pragma unroll
  for (int q = 0; q < kBTiles; ++q)
    if (tile_a(q) == bk)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int col = kBT * tile_b(q) + cl;
        if (rl + 4 * r == rk && col < K) rv[col] = tile[q][r];
      }
  __syncthreads();
  // one barrier per block: the owner of tile row b (wave b or 15 - b)
  // forms w_b = Y_bb^T r_b and subtracts its row's L_bc^T w_b from every
  // r_c, c < b, before the next block starts
  for (int b = T - 1; b >= 0; --b) {
    if (wave == (b < 8 ? b : 15 - b)) {
      refresh();
      if (lane < kBT) {
        double w = 0.0;
#pragma unroll
        for (int j = 0; j < kBT; ++j)
          w += Ys[b][j * kBLd + lane] * rv[kBT * b + j];
        wv[kBT * b + lane] = kBT * b + lane < K ? w : 0.0;
      }
      wave_lds_sync();
      double wb[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) wb[r] = wv[kBT * b + rl + 4 * r];
      // every tile (b, c), c < b: each lane's sum over its four rows into
      // P[c] (free after the factorization), then lane l adds the four row
      // groups of column l % 16 of blocks c = l / 16 + 4 m into r_c
#pragma unroll
      for (int q = 0; q < kBTiles; ++q) {
        const int c = tile_b(q);
        if (tile_a(q) == b && c < b) {
          double part = 0.0;
#pragma unroll
          for (int r = 0; r < 4; ++r) part += tile[q][r] * wb[r];
          P[c][rl * kBT + cl] = part;
        }
      }
      wave_lds_sync();
#pragma unroll
      for (int m = 0; m < 4; ++m) {
        const int c = rl + 4 * m;
        if (c < b) {
          const double* pc = P[c];
          rv[kBT * c + cl] -= (pc[cl] + pc[kBT + cl]) +
                              (pc[2 * kBT + cl] + pc[3 * kBT + cl]);
        }
      }
    }
    __syncthreads();
  }
  if (t < K) beta[t] = wv[t] * dsc[t];
  if (t == 0) *rank_out = rank_s;
}

// continuation = relu(X beta); exercise iff ex > continuation.  The
// per-path power table lives in dynamic LDS, [d * (deg + 1)][blockDim.x]
// (consecutive threads on consecutive words): as a private array indexed by
// the exponents it would sit in scratch memory.
__global__ void __launch_bounds__(256)
lsm_update_kernel(LsmArgs a, double df_ratio, const double* __restrict__ exv,
                  const double* __restrict__ beta, double* __restrict__ cash,
                  double* __restrict__ vals) {
  __shared__ double sb[256];
  __shared__ uint8_t expo[256 * kMaxAssets];
  extern __shared__ double pws[];
  const int nt = blockDim.x, tid = threadIdx.x;
  for (int k = tid; k < a.K; k += nt) sb[k] = beta[k];
  for (int e = tid; e < a.K * a.d; e += nt) expo[e] = a.expo[e];
  __syncthreads();
  const int64_t p = (int64_t)blockIdx.x * nt + tid;
  if (p >= a.n_paths) return;
  const int deg1 = a.deg + 1;
  for (int j = 0; j < a.d; ++j) {
    const double c =
        (basis_input(a, path_x(a, p, j)) - a.mean[j]) * a.scale[j];
    double v = 1.0;
    for (int e = 0; e <= a.deg; ++e) {
      pws[(j * deg1 + e) * nt + tid] = v;
      v *= c;
    }
  }
  double cont = 0;
  for (int k = 0; k < a.K; ++k) {
    const uint8_t* e = expo + k * a.d;
    double b = 1.0;
    for (int j = 0; j < a.d; ++j) b *= pws[(j * deg1 + e[j]) * nt + tid];
    cont += b * sb[k];
  }
  cont = cont > 0 ? cont : 0;
  const double ex = exv[p];
  if (ex > cont) {
    vals[p] = 0;
    cash[p] = ex;
  } else {
    vals[p] = disc_ratio(a, df_ratio, p) * (cash[p] + vals[p]);
    cash[p] = 0;
  }
}

// ---------------------------------------------------------------------------
// Dense update.  The continuation value of a monomial basis is the
// polynomial sum_k beta_k prod_j c_j^e_kj; with its coefficients scattered
// into the dense tensor T[m_0 + (deg+1) m_1 + ...] it is evaluated by nested
// Horner steps, about (deg+1)^d FMAs per path, the coefficients read through
// uniform (scalar) loads.

constexpr int kMaxDense = 1024;  // dense coefficients

// T from beta: entry i is the sum, in basis order, of the beta_k whose
// exponent row indexes i (rows with an exponent above the degree have no
// entry).  One workgroup.
__global__ void __launch_bounds__(kMaxDense)
lsm_dense_beta_kernel(const double* __restrict__ beta,
                      const uint8_t* __restrict__ expo, int K, int d, int deg,
                      int ndense, double* __restrict__ dense) {
  __shared__ int idx[256];
  const int t = threadIdx.x;
  if (t < K) {
    int i = 0, mul = 1;
    bool ok = true;
    for (int j = 0; j < d; ++j) {
      const int e = expo[t * d + j];
      ok = ok && e <= deg;
      i += e * mul;
      mul *= deg + 1;
    }
    idx[t] = ok ? i : -1;
  }
  __syncthreads();
  if (t >= ndense) return;
  double v = 0.0;
  for (int k = 0; k < K; ++k)
    if (idx[k] == t) v += beta[k];
  dense[t] = v;
}

// sum_m c_J^m (polynomial of the lower dimensions at t + m st_J), by Horner
template <int J>
struct DenseHorner {
  __device__ __forceinline__ static double eval(const double* __restrict__ t,
                                                const double* c, const int* st,
                                                int deg) {
    double h = DenseHorner<J - 1>::eval(t + deg * st[J], c, st, deg);
    for (int m = deg - 1; m >= 0; --m)
      h = fma(h, c[J], DenseHorner<J - 1>::eval(t + m * st[J], c, st, deg));
    return h;
  }
};
template <>
struct DenseHorner<0> {
  __device__ __forceinline__ static double eval(const double* __restrict__ t,
                                                const double* c, const int*,
                                                int deg) {
    double h = t[deg];
    for (int m = deg - 1; m >= 0; --m) h = fma(h, c[0], t[m]);
    return h;
  }
};

// lsm_update_kernel with the continuation value from the dense tensor.
template <int D>
__global__ void __launch_bounds__(256)
lsm_update_dense_kernel(LsmArgs a, double df_ratio,
                        const double* __restrict__ exv,
                        const double* __restrict__ dense,
                        double* __restrict__ cash, double* __restrict__ vals) {
  const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (p >= a.n_paths) return;
  double c[D];
  int st[D];
#pragma unroll
  for (int j = 0; j < D; ++j) {
    c[j] = (basis_input(a, path_x(a, p, j)) - a.mean[j]) * a.scale[j];
    st[j] = j == 0 ? 1 : st[j - 1] * (a.deg + 1);
  }
  double cont = DenseHorner<D - 1>::eval(dense, c, st, a.deg);
  cont = cont > 0 ? cont : 0;
  const double ex = exv[p];
  if (ex > cont) {
    vals[p] = 0;
    cash[p] = ex;
  } else {
    vals[p] = disc_ratio(a, df_ratio, p) * (cash[p] + vals[p]);
    cash[p] = 0;
  }
}

// The dense Horner with the degree fixed: the nest unrolls to constant
// coefficient offsets into the workgroup's LDS copy of the tensor (one
// broadcast read per coefficient, no scalar load and wait per coefficient)
// and a thread evaluates NP paths, so each read feeds NP FMAs.  The
// operations of each path are DenseHorner's, in the same order.
template <int J, int DEG, int NP>
struct DenseHornerFixed {
  __device__ __forceinline__ static void eval(const double* __restrict__ t,
                                              const double (*c)[NP],
                                              double (&h)[NP]) {
    constexpr int ST = cpow(DEG + 1, J);
    DenseHornerFixed<J - 1, DEG, NP>::eval(t + DEG * ST, c, h);
    if constexpr (J >= 3) {
      // dimensions above the third loop at run time: only a sub-nest of
      // (DEG + 1)^3 coefficients is unrolled, so its reads are the ones in
      // flight (not the whole tensor's)
#pragma unroll 1
      for (int m = DEG - 1; m >= 0; --m) {
        double g[NP];
        DenseHornerFixed<J - 1, DEG, NP>::eval(t + m * ST, c, g);
#pragma unroll
        for (int q = 0; q < NP; ++q) h[q] = fma(h[q], c[J][q], g[q]);
      }
    } else {
#pragma unroll
      for (int m = DEG - 1; m >= 0; --m) {
        double g[NP];
        DenseHornerFixed<J - 1, DEG, NP>::eval(t + m * ST, c, g);
#pragma unroll
        for (int q = 0; q < NP; ++q) h[q] = fma(h[q], c[J][q], g[q]);
      }
    }
  }
};
template <int DEG, int NP>
struct DenseHornerFixed<0, DEG, NP> {
  __device__ __forceinline__ static void eval(const double* __restrict__ t,
                                              const double (*c)[NP],
                                              double (&h)[NP]) {
#pragma unroll
    for (int q = 0; q < NP; ++q) h[q] = t[DEG];
#pragma unroll
    for (int m = DEG - 1; m >= 0; --m)
#pragma unroll
      for (int q = 0; q < NP; ++q) h[q] = fma(h[q], c[0][q], t[m]);
  }
};

constexpr int kDenseP = 4;  // paths per thread of the fixed-degree update

// lsm_update_dense_kernel at degree DEG: thread t of block b takes paths
// b 256 NP + t + 256 q, q < NP (coalesced per q).
template <int D, int DEG>
__global__ void __launch_bounds__(256)
lsm_update_dense_fixed_kernel(LsmArgs a, double df_ratio,
                              const double* __restrict__ exv,
                              const double* __restrict__ dense,
                              double* __restrict__ cash,
                              double* __restrict__ vals) {
  constexpr int NT = cpow(DEG + 1, D);
  __shared__ double tl[NT];
  for (int e = threadIdx.x; e < NT; e += 256) tl[e] = dense[e];
  const int64_t base = (int64_t)blockIdx.x * (256 * kDenseP) + threadIdx.x;
  const int64_t n = a.n_paths;
  double c[D][kDenseP];
#pragma unroll
  for (int q = 0; q < kDenseP; ++q) {
    int64_t p = base + 256 * q;
    p = p < n ? p : n - 1;  // rows past the end compute and are not stored
#pragma unroll
    for (int j = 0; j < D; ++j)
      c[j][q] = (basis_input(a, path_x(a, p, j)) - a.mean[j]) * a.scale[j];
  }
  __syncthreads();
  double h[kDenseP];
  DenseHornerFixed<D - 1, DEG, kDenseP>::eval(tl, c, h);
#pragma unroll
  for (int q = 0; q < kDenseP; ++q) {
    const int64_t p = base + 256 * q;
    if (p >= n) break;
    const double cont = h[q] > 0 ? h[q] : 0;
    const double ex = exv[p];
    if (ex > cont) {
      vals[p] = 0;
      cash[p] = ex;
    } else {
      vals[p] = disc_ratio(a, df_ratio, p) * (cash[p] + vals[p]);
      cash[p] = 0;
    }
  }
}

// Dense coefficients of (d, deg), or 0 when the dense update does not apply:
// more than kMaxDense of them, or more work than the basis-order sum.
int dense_size(int d, int deg, int K) {
  int64_t n = 1;
  for (int j = 0; j < d; ++j) {
    n *= deg + 1;
    if (n > kMaxDense) return 0;
  }
  return n <= (int64_t)K * (d + 1) ? (int)n : 0;
}

// Threads per workgroup of lsm_update_kernel: 256, or 64 when the power
// table would take more than 48 KB of LDS at 256.
inline int update_threads(const LsmArgs& a) {
  return a.d * (a.deg + 1) <= 24 ? 256 : 64;
}

// Terminal exercise values as the initial cashflow.
__global__ void __launch_bounds__(256)
lsm_terminal_kernel(LsmArgs a, double* __restrict__ cash,
                    double* __restrict__ vals) {
  const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (p >= a.n_paths) return;
  double x[kMaxAssets];
  for (int j = 0; j < a.d; ++j) x[j] = path_x(a, p, j);
  cash[p] = payoff_value(a, x, p);
  vals[p] = 0;
}

// Per-block sums of df0 * (cash + vals) over rows [first, n) (and squares);
// d1 / d0 (optional) are the per-path discount rows 1 and 0.
__global__ void __launch_bounds__(256)
lsm_final_kernel(const double* __restrict__ cash, const double* __restrict__ vals,
                 double df0, const double* __restrict__ d1,
                 const double* __restrict__ d0, int64_t first, int64_t n,
                 double* __restrict__ block_sums) {
  __shared__ double s1[256], s2[256];
  const int64_t p = first + (int64_t)blockIdx.x * 256 + threadIdx.x;
  double v = 0;
  if (p < n) v = (d1 ? d1[p] / d0[p] : df0) * (cash[p] + vals[p]);
  s1[threadIdx.x] = v;
  s2[threadIdx.x] = v * v;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      s1[threadIdx.x] += s1[threadIdx.x + s];
      s2[threadIdx.x] += s2[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    block_sums[2 * blockIdx.x] = s1[0];
    block_sums[2 * blockIdx.x + 1] = s2[0];
  }
}

__global__ void __launch_bounds__(256)
lsm_final_reduce_kernel(const double* __restrict__ block_sums,
                        int64_t n_blocks, int64_t count,
                        double* __restrict__ out) {
  // thread t adds blocks t, t + 256, ... in order; fixed LDS tree
  __shared__ double s1[256], s2[256];
  const int t = threadIdx.x;
  double a = 0, b = 0;
  for (int64_t i = t; i < n_blocks; i += 256) {
    a += block_sums[2 * i];
    b += block_sums[2 * i + 1];
  }
  s1[t] = a;
  s2[t] = b;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) {
      s1[t] += s1[t + w];
      s2[t] += s2[t + w];
    }
    __syncthreads();
  }
  if (t != 0) return;
  const double n = (double)count;
  const double mean = s1[0] / n;
  double var = n > 1 ? (s2[0] - n * mean * mean) / (n - 1) : 0.0;
  out[0] = mean;
  out[1] = sqrt((var > 0 ? var : 0.0) / n);
}

// MLP inputs of one date: standardised asset prices and exercise value / K
// (row stride kMlpStride), regression target y / K and ITM weight.
__global__ void __launch_bounds__(256)
lsm_features_kernel(LsmArgs a, const double* __restrict__ exv,
                    const double* __restrict__ y, float* __restrict__ x,
                    float* __restrict__ tgt, float* __restrict__ w) {
  const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (p >= a.n_paths) return;
  // the row of 8 floats in registers, stored as two 16-byte halves (a
  // wave's stores then cover whole lines)
  float v[kMlpStride];
#pragma unroll
  for (int j = 0; j < kMlpIn - 1; ++j)
    v[j] = j < a.d ? (float)((basis_input(a, path_x(a, p, j)) - a.mean[j]) *
                             a.scale[j])
                   : 0.f;
  const double e = exv[p];
  v[kMlpIn - 1] = (float)(e / a.strike);
  v[6] = v[7] = 0.f;
  typedef float f32x4v __attribute__((ext_vector_type(4)));
  f32x4v* xr = (f32x4v*)(x + p * kMlpStride);
  xr[0] = f32x4v{v[0], v[1], v[2], v[3]};
  xr[1] = f32x4v{v[4], v[5], v[6], v[7]};
  tgt[p] = (float)(y[p] / a.strike);
  w[p] = e > 0 ? 1.f : 0.f;
}

// Exercise update from a given continuation estimate (in units of K).
__global__ void __launch_bounds__(256)
lsm_update_cont_kernel(int64_t n, double strike, double df_ratio,
                       const double* __restrict__ exv,
                       const float* __restrict__ cont,
                       double* __restrict__ cash, double* __restrict__ vals) {
  const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (p >= n) return;
  double c = (double)cont[p] * strike;
  c = c > 0 ? c : 0;
  const double ex = exv[p];
  if (ex > c) {
    vals[p] = 0;
    cash[p] = ex;
  } else {
    vals[p] = df_ratio * (cash[p] + vals[p]);
    cash[p] = 0;
  }
}

// ---------------------------------------------------------------------------
// One-asset date chain, two launches per exercise date.  With one asset and
// the full polynomial basis the regression needs only the power sums
// c^0 .. c^(2 deg) and y c^0 .. y c^deg of the in-the-money calibration rows
// (the moment plan's small case), and the dense update evaluates the
// continuation by Horner in c.  So one pass over the paths (pass r)
//   - updates the cashflows of the date regressed before (date r + 1),
//   - accumulates the power sums of date r from the updated cashflows, and
//   - sums the basis input of date r - 1 and its square (its centring),
// and one workgroup (solve r) reduces the pass's per-block rows, builds the
// (K + 1)^2 extended Gram from the sums, solves it with the column-scaled,
// truncating Cholesky of lsm_solve_kernel (serially: K <= 13) and writes
// date r's dense coefficients and rank and date r - 1's centring.  The
// general chain takes eight launches per date for the same work.

constexpr int kSmallMaxDeg = 4;
constexpr int kSmallMaxK = 2 * kSmallMaxDeg + 1;  // exponent rows (K)
constexpr int kSmallMaxRows = 2048;  // workgroups (partial rows) per pass
constexpr int kSmallRows = 4;        // rows per thread and round of a pass

// One-asset payoff of the stored value x (exp of it for log prices) and
// the in-the-money test (ln S against ln K for log prices, no exp).
__device__ __forceinline__ double payoff1(const LsmArgs& a, double x,
                                          int64_t p) {
  if (a.payoff == TQF_PAYOFF_TABLE)
    return a.table[(int64_t)a.t_idx * a.n_paths + p];
  const double s = a.basis_log == 2 ? exp(x) : x;
  const double v = a.payoff == TQF_PAYOFF_BASKET_PUT ? a.strike - s
                                                     : s - a.strike;
  return v > 0 ? v : 0;
}
__device__ __forceinline__ bool itm1(const LsmArgs& a, double x, int64_t p) {
  if (a.basis_log == 2 && a.payoff != TQF_PAYOFF_TABLE)
    return a.payoff == TQF_PAYOFF_BASKET_PUT ? x < a.log_strike
                                             : x > a.log_strike;
  return payoff1(a, x, p) > 0;
}

// The per-date values one pass reads: the date's column of the paths
// (value of path p at col[p * sp]), its centring and discount rows.
struct SmallDate {
  const double* col;
  int t_idx;            // time index (the table payoff's row)
  double ratio;         // scalar discount ratio (when dnum is null)
  const double* dnum;
  const double* dden;
};

__device__ __forceinline__ double small_ratio(const SmallDate& d,
                                              int64_t p) {
  return d.dnum ? d.dnum[p] / d.dden[p] : d.ratio;
}

// Partial row of workgroup b: [0, NX) power sums c^m, [NX, NM) y c^m, then
// the sum and the sum of squares of date r - 1's basis input.  do_update
// = 0 skips the update (the first regressed date follows the terminal
// payoff); n_calib = 0 skips the power sums; do_stats = 0 the centring
// sums.  The centrings of dates r + 1 and r come from device slots
// (mean_u / mean_m, written by the solves).  cash holds each path's whole
// value (see small_date_chain).
template <int DEG>
__global__ void __launch_bounds__(256)
lsm_small_pass_kernel(LsmArgs a, SmallDate du, const double* __restrict__ cu,
                      const double* __restrict__ dense_u, int do_update,
                      SmallDate dm, const double* __restrict__ cm,
                      int64_t n_calib, SmallDate ds, int do_stats,
                      int64_t rows_per_block, double* __restrict__ cash,
                      double* __restrict__ partial) {
  constexpr int NX = 2 * DEG + 1, NM = NX + DEG + 1, W = NM + 2;
  __shared__ double red[4][W];
  double acc[W];
#pragma unroll
  for (int m = 0; m < W; ++m) acc[m] = 0.0;
  const int64_t n = a.n_paths, sp = a.sp;
  const double mu = do_update ? cu[0] : 0.0, su = do_update ? cu[kMaxAssets] : 0.0;
  const double mm = cm[0], sm = cm[kMaxAssets];
  double dn[DEG + 1];
#pragma unroll
  for (int m = 0; m <= DEG; ++m) dn[m] = do_update ? dense_u[m] : 0.0;
  LsmArgs au = a, am = a;
  au.t_idx = du.t_idx;
  am.t_idx = dm.t_idx;
  const int64_t p0 = (int64_t)blockIdx.x * rows_per_block;
  // kSmallRows rows per thread and round, every load of a round issued
  // before its arithmetic
  for (int64_t base = 0; base < rows_per_block; base += 256 * kSmallRows) {
    int64_t p[kSmallRows];
    double cs[kSmallRows], xu[kSmallRows], xm[kSmallRows], xs[kSmallRows];
#pragma unroll
    for (int i = 0; i < kSmallRows; ++i) {
      p[i] = p0 + base + 256 * i + threadIdx.x;
      const bool ok = base + 256 * i + threadIdx.x < rows_per_block &&
                      p[i] < n;
      if (!ok) p[i] = -1;
      cs[i] = ok ? cash[p[i]] : 0.0;
      xu[i] = ok && do_update ? du.col[p[i] * sp] : 1.0;
      xm[i] = ok && p[i] < n_calib ? dm.col[p[i] * sp] : 1.0;
      xs[i] = ok && do_stats ? ds.col[p[i] * sp] : 1.0;
    }
#pragma unroll
    for (int i = 0; i < kSmallRows; ++i) {
      if (p[i] < 0) continue;
      if (do_update) {
        const double ex = itm1(au, xu[i], p[i]) ? payoff1(au, xu[i], p[i])
                                                : 0.0;
        const double c = (basis_input(a, xu[i]) - mu) * su;
        double h = dn[DEG];
#pragma unroll
        for (int m = DEG - 1; m >= 0; --m) h = fma(h, c, dn[m]);
        const double cont = h > 0 ? h : 0;
        cs[i] = ex > cont ? ex : small_ratio(du, p[i]) * cs[i];
        cash[p[i]] = cs[i];
      }
      if (p[i] < n_calib && itm1(am, xm[i], p[i])) {
        // in-the-money calibration rows only
        const double c = (basis_input(a, xm[i]) - mm) * sm;
        double v = 1.0, u = small_ratio(dm, p[i]) * cs[i];
#pragma unroll
        for (int m = 0; m < NX; ++m) {
          acc[m] += v;
          v *= c;
        }
#pragma unroll
        for (int m = NX; m < NM; ++m) {
          acc[m] += u;
          u *= c;
        }
      }
      if (do_stats) {
        const double b = basis_input(a, xs[i]);
        acc[NM] += b;
        acc[NM + 1] += b * b;
      }
    }
  }
  // fixed-order workgroup sums: wave butterflies, then the waves in order
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int m = 0; m < W; ++m) {
    double v = acc[m];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if (lane == 0) red[wave][m] = v;
  }
  __syncthreads();
  if (threadIdx.x < W)
    partial[(int64_t)blockIdx.x * W + threadIdx.x] =
        ((red[0][threadIdx.x] + red[1][threadIdx.x]) + red[2][threadIdx.x]) +
        red[3][threadIdx.x];
}

// Solve r: column sums of the pass's n_rows partial rows (thread t adds
// rows t, t + 256, ...; then wave butterflies and the four wave sums in
// wave order), then, on one thread, the extended Gram
// G[k][l] = sum c^(e_k + e_l), G[k][K] = sum y c^e_k and the steps of
// lsm_solve_kernel: column scaling by 1 / sqrt(G_kk), pivots at or below
// rtol * max scaled diagonal dropped, L^T w = z, beta = D w.  dense[m] is
// the sum of the beta_k with exponent m, in basis order.  do_solve = 0 only
// reduces the centring sums (into mean_next[0] and mean_next[kMaxAssets]).
template <int DEG>
__global__ void __launch_bounds__(256)
lsm_small_solve_kernel(const double* __restrict__ partial, int n_rows,
                       int do_solve, int do_stats, int64_t n_paths,
                       const uint8_t* __restrict__ expo_g, int K, double rtol,
                       double* __restrict__ dense, int* __restrict__ rank_out,
                       double* __restrict__ mean_next) {
  constexpr int NX = 2 * DEG + 1, NM = NX + DEG + 1, W = NM + 2;
  __shared__ double red[4][W];
  __shared__ double A[(kSmallMaxK + 1) * (kSmallMaxK + 1)];
  __shared__ int expo[kSmallMaxK];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  if (t < K) expo[t] = expo_g[t];
  double acc[W];
#pragma unroll
  for (int c = 0; c < W; ++c) acc[c] = 0.0;
  // rows t + 256 i, eight per round with every load of a round issued
  // before its sums (the rows are added in increasing order)
  constexpr int kR = 8;
  for (int r0 = t; r0 < n_rows; r0 += 256 * kR) {
    double v[kR][W];
#pragma unroll
    for (int i = 0; i < kR; ++i) {
      const int r = r0 + 256 * i;
#pragma unroll
      for (int c = 0; c < W; ++c)
        v[i][c] = r < n_rows ? partial[(int64_t)r * W + c] : 0.0;
    }
#pragma unroll
    for (int i = 0; i < kR; ++i)
#pragma unroll
      for (int c = 0; c < W; ++c) acc[c] += v[i][c];
  }
#pragma unroll
  for (int c = 0; c < W; ++c) {
    double v = acc[c];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if (lane == 0) red[wave][c] = v;
  }
  __syncthreads();
  if (t != 0) return;
  double M[W];
#pragma unroll
  for (int c = 0; c < W; ++c)
    M[c] = ((red[0][c] + red[1][c]) + red[2][c]) + red[3][c];
  if (do_stats) {
    const double nn = (double)n_paths;
    const double m = M[NM] / nn;
    const double var = M[NM + 1] / nn - m * m;
    mean_next[0] = m;
    mean_next[kMaxAssets] =
        var > 1e-300 && sqrt(var) > 1e-12 * fabs(m) ? 1.0 / sqrt(var) : 1.0;
  }
  if (!do_solve) return;
  // the sums into LDS (indexed by exponent sums below)
  double* Ms = &red[0][0];
  for (int c = 0; c < W; ++c) Ms[c] = M[c];
  const int n1 = K + 1;
  double dsc[kSmallMaxK + 1];
  double maxdiag = 0.0;
  for (int k = 0; k < K; ++k) {
    const double gkk = Ms[2 * expo[k]];
    dsc[k] = gkk > 0 ? 1.0 / sqrt(gkk) : 0.0;
    maxdiag = fmax(maxdiag, gkk * (dsc[k] * dsc[k]));
  }
  dsc[K] = 1.0;
  // scaled lower triangle of the extended matrix (row K: the rhs sums)
  for (int i = 0; i < n1; ++i)
    for (int j = 0; j <= i; ++j) {
      double gij = 0.0;
      if (i < K) gij = Ms[expo[i] + expo[j]];
      else if (j < K) gij = Ms[NX + expo[j]];
      A[i * n1 + j] = gij * (dsc[i] * dsc[j]);
    }
  const double thr = rtol * maxdiag;
  int rank = 0;
  for (int k = 0; k < K; ++k) {
    const double d = A[k * n1 + k];
    const bool keep = d > thr && d > 0;
    const double root = keep ? sqrt(d) : 0.0;  // truncated: column dropped
    rank += keep ? 1 : 0;
    if (keep) {
      const double inv = 1.0 / d;
      for (int j = k + 1; j < n1; ++j) {
        const double cj = A[j * n1 + k];
        for (int i = j; i < n1; ++i) A[i * n1 + j] -= A[i * n1 + k] * cj * inv;
      }
    }
    A[k * n1 + k] = root;
    for (int i = k + 1; i < n1; ++i)
      A[i * n1 + k] = root > 0 ? A[i * n1 + k] / root : 0.0;
  }
  if (rank_out) *rank_out = rank;
  double r[kSmallMaxK], beta[kSmallMaxK];
  for (int j = 0; j < K; ++j) r[j] = A[K * n1 + j];
  for (int i = K - 1; i >= 0; --i) {
    const double li = A[i * n1 + i];
    const double w = li > 0 ? r[i] / li : 0.0;
    beta[i] = w * dsc[i];
    for (int j = 0; j < i; ++j) r[j] -= A[i * n1 + j] * w;
  }
  for (int m = 0; m <= DEG; ++m) {
    double v = 0.0;
    for (int k = 0; k < K; ++k)
      if (expo[k] == m) v += beta[k];
    dense[m] = v;
  }
}

inline unsigned blocks_for(int64_t n) { return (unsigned)((n + 255) / 256); }

LsmArgs make_args(const tqf_lsm* l, int t_idx) {
  LsmArgs a{};
  a.paths = l->paths;
  a.n_paths = l->n_paths;
  a.n_times = l->n_times;
  a.d = l->dim;
  // zero strides: the reference's [n_paths][n_times][dim] layout
  const bool given = l->stride_path || l->stride_time || l->stride_asset;
  a.sp = given ? l->stride_path : (int64_t)l->n_times * l->dim;
  a.st = given ? l->stride_time : l->dim;
  a.sa = given ? l->stride_asset : 1;
  a.t_idx = t_idx;
  a.payoff = l->payoff;
  a.strike = l->strike;
  a.mean = nullptr;
  a.scale = nullptr;
  a.expo = l->exponents;
  a.K = l->n_basis;
  a.deg = l->degree;
  a.basis_log = l->basis_log;
  a.log_strike = l->strike > 0 ? log(l->strike) : -HUGE_VAL;
  a.table = l->payoff_table;
  a.dnum = nullptr;
  a.dden = nullptr;
  return a;
}

int validate(const tqf_lsm* l) {
  TQF_REQUIRE(l != nullptr, "null descriptor");
  TQF_REQUIRE(l->dim >= 1 && l->dim <= kMaxAssets,
              "LSM supports 1..%d assets, got %d", kMaxAssets, l->dim);
  TQF_REQUIRE(l->degree >= 0 && (l->degree + 1) * l->dim <= kMaxPw,
              "(degree + 1) * dim must be at most %d", kMaxPw);
  TQF_REQUIRE(l->n_basis >= 1 && l->n_basis <= 255,
              "basis size must be in [1, 255], got %d", l->n_basis);
  TQF_REQUIRE(l->n_paths >= 1 && l->n_times >= 1, "empty sample paths");
  TQF_REQUIRE(l->payoff == TQF_PAYOFF_BASKET_PUT ||
                  l->payoff == TQF_PAYOFF_MAX_CALL ||
                  l->payoff == TQF_PAYOFF_TABLE,
              "unknown payoff %d", l->payoff);
  TQF_REQUIRE(l->payoff != TQF_PAYOFF_TABLE || l->payoff_table,
              "the table payoff needs payoff_table");
  TQF_REQUIRE(l->paths && l->exponents, "null LSM arrays");
  TQF_REQUIRE(l->basis_log >= 0 && l->basis_log <= 2,
              "basis_log must be 0, 1 or 2, got %d", l->basis_log);
  return TQF_OK;
}

}  // namespace

extern "C" int64_t tqf_lsm_workspace_size(int64_t n_paths, int32_t n_basis) {
  const int64_t nb = (n_paths + 255) / 256;
  const int tiles = (n_basis + 1 + kTile - 1) / kTile;
  const int pairs = tiles * (tiles + 1) / 2;
  const int64_t chunks = 256;
  const int64_t kext = n_basis + 1;
  int64_t bytes = 0;
  bytes += 3 * n_paths * 8;                     // exv, y, spare
  bytes += nb * 2 * kMaxAssets * 8 + 2 * nb * 8;  // block sums
  bytes += chunks * pairs * kTile * kTile * 8;  // Gram partials
  bytes += kext * kext * 8;                     // G
  bytes += 4 * kext * 8 + 64;                   // beta, rank (with slack)
  bytes += pairs * 8 + 4 * kMaxAssets * 8;      // pairs, mean, scale
  bytes += 1024;                                // one-asset chain state
  return bytes + 4096;
}

// The one-asset chain of lsm_small_pass_kernel / lsm_small_solve_kernel
// for the regressed dates r = n_ex - 2 .. 0 (exercise index ex_idx[r],
// discount ratio df[r + 2] / df[r + 1]), after the terminal payoff of date
// n_ex - 1 is in cash (vals zero).  The chain keeps each path's whole value
// in cash: the reference's pair (cashflow, discounted later values) always
// has one member zero after an update (exercise: (payoff, 0); hold:
// (0, ratio (cashflow + values))), so their sum is exact and is all the
// regression and the price read; vals is left at zero, and the final
// cash + vals is the same number.  slots holds three rotating centrings (mean
// at [0], scale at [kMaxAssets]; date r uses slot r % 3); dense the
// coefficients of the date solved last.
template <int DEG>
int small_date_chain(const tqf_lsm* l, const int32_t* ex_idx, int n_ex,
                     const double* df, int64_t n_calib, double* partial,
                     double* slots, double* dense, double* cash,
                     int32_t* ranks, hipStream_t s) {
  const int64_t n = l->n_paths;
  const LsmArgs a = make_args(l, 0);
  // rows per workgroup: a multiple of one round (256 kSmallRows rows)
  constexpr int64_t kRound = 256 * kSmallRows;
  int64_t rpb = (n + kSmallMaxRows - 1) / kSmallMaxRows;
  rpb = (rpb + kRound - 1) / kRound * kRound;
  const unsigned grid = (unsigned)((n + rpb - 1) / rpb);
  auto date = [&](int r) {
    SmallDate d{};
    d.col = a.paths + (int64_t)ex_idx[r] * a.st;
    d.t_idx = ex_idx[r];
    d.ratio = df[r + 2] / df[r + 1];
    if (l->path_discounts) {
      d.dnum = l->path_discounts + (int64_t)(r + 2) * n;
      d.dden = l->path_discounts + (int64_t)(r + 1) * n;
    }
    return d;
  };
  auto slot = [&](int r) { return slots + (r % 3) * 2 * kMaxAssets; };
  auto pass = [&](int r_upd, int r_mom, int r_stat) {
    const int r_any = r_upd >= 0 ? r_upd : (r_mom >= 0 ? r_mom : 0);
    const SmallDate du = date(r_upd >= 0 ? r_upd : r_any);
    const SmallDate dm = date(r_mom >= 0 ? r_mom : r_any);
    const SmallDate dst = date(r_stat >= 0 ? r_stat : r_any);
    hipLaunchKernelGGL(lsm_small_pass_kernel<DEG>, grid, 256, 0, s, a, du,
                       slot(r_upd >= 0 ? r_upd : r_any), dense,
                       r_upd >= 0 ? 1 : 0, dm,
                       slot(r_mom >= 0 ? r_mom : r_any),
                       r_mom >= 0 ? n_calib : (int64_t)0, dst,
                       r_stat >= 0 ? 1 : 0, rpb, cash, partial);
  };
  auto solve = [&](int r, int r_stat) {
    hipLaunchKernelGGL(lsm_small_solve_kernel<DEG>, 1, 256, 0, s, partial,
                       (int)grid, r >= 0 ? 1 : 0, r_stat >= 0 ? 1 : 0, n,
                       l->exponents, l->n_basis, l->rcond, dense,
                       (ranks && r >= 0) ? ranks + r : nullptr,
                       slot(r_stat >= 0 ? r_stat : 0));
  };
  // centring of the first regressed date
  const int r0 = n_ex - 2;
  pass(-1, -1, r0);
  solve(-1, r0);
  for (int r = r0; r >= 0; --r) {
    pass(r < r0 ? r + 1 : -1, r, r >= 1 ? r - 1 : -1);
    solve(r, r >= 1 ? r - 1 : -1);
  }
  pass(0, -1, -1);  // the exercise decision of the earliest date
  return check_launch("tqf_lsm_price (one-asset chain)");
}

extern "C" int tqf_lsm_price(const tqf_lsm* l, const int32_t* ex_idx,
                             int32_t n_ex, const double* df, void* workspace,
                             double* cash, double* vals, double* result,
                             int32_t* ranks, void* stream) {
  clear_error();
  int rc = validate(l);
  if (rc) return rc;
  TQF_REQUIRE(ex_idx && n_ex >= 1 && df && workspace && cash && vals && result,
              "null LSM argument");
  TQF_REQUIRE(l->num_calibration >= 0 && l->num_calibration <= l->n_paths,
              "num_calibration_samples out of range");
  for (int i = 0; i < n_ex; ++i)
    TQF_REQUIRE(ex_idx[i] >= 0 && ex_idx[i] < l->n_times,
                "exercise index %d out of range", ex_idx[i]);
  const hipStream_t s = (hipStream_t)stream;
  const int64_t n = l->n_paths;
  const int64_t nb = (n + 255) / 256;
  const int K = l->n_basis, kext = K + 1;
  // per-date centring and scaling live in the workspace
  const int tiles = (kext + kTile - 1) / kTile;
  const int n_pairs = tiles * (tiles + 1) / 2;
  const int64_t n_calib = l->num_calibration > 0 ? l->num_calibration : n;
  int64_t chunk = (n_calib + 255) / 256;
  chunk = ((chunk + kSubPaths - 1) / kSubPaths) * kSubPaths;
  if (chunk < kSubPaths) chunk = kSubPaths;
  const int n_chunks = (int)((n_calib + chunk - 1) / chunk);

  char* w = (char*)workspace;
  auto take = [&](int64_t bytes) {
    char* p = w;
    w += (bytes + 255) & ~(int64_t)255;
    return p;
  };
  double* exv = (double*)take(n * 8);
  double* y = (double*)take(n * 8);
  double* bsum = (double*)take(nb * 2 * kMaxAssets * 8);
  double* fsum = (double*)take(2 * nb * 8);
  double* partial = (double*)take((int64_t)256 * n_pairs * kTile * kTile * 8);
  double* g = (double*)take((int64_t)kext * kext * 8);
  double* beta = (double*)take(kext * 8);
  int* rank = (int*)take(64);
  int2* pairs = (int2*)take(n_pairs * 8);
  double* mean = (double*)take(kMaxAssets * 8);
  double* scale = (double*)take(kMaxAssets * 8);
  // one-asset chain: three (mean, scale) slots and the dense coefficients
  double* chain = (double*)take((6 * kMaxAssets + kSmallMaxDeg + 1) * 8);

  // pair table (host -> device, tiny)
  int2 hp[64];
  TQF_REQUIRE(n_pairs <= 64, "basis too large");
  // the pair table is a device copy of this host array; the copy must be
  // complete before `hp` leaves scope

  {
    int q = 0;
    for (int i = 0; i < tiles; ++i)
      for (int j = i; j < tiles; ++j) hp[q++] = int2{i, j};
  }
  // the exponent table (at most 255 x 8 bytes) comes back in the same
  // synchronisation: basis_col indexes the power table by exponent, so an
  // entry above the degree would read another asset's powers
  uint8_t he_expo[255 * kMaxAssets];
  hipError_t he = hipMemcpyAsync(pairs, hp, n_pairs * sizeof(int2),
                                 hipMemcpyHostToDevice, s);
  if (he == hipSuccess)
    he = hipMemcpyAsync(he_expo, l->exponents, (size_t)K * l->dim,
                        hipMemcpyDeviceToHost, s);
  if (he == hipSuccess) he = hipStreamSynchronize(s);
  if (he != hipSuccess) return set_error((int)he, "copy pairs / exponents");
  {
    double grid = 1;
    for (int j = 0; j < l->dim; ++j) grid *= l->degree + 1;
    TQF_REQUIRE(K <= grid, "%d basis functions exceed (degree + 1)^dim = %.0f",
                K, grid);
    for (int e = 0; e < K * l->dim; ++e)
      TQF_REQUIRE(he_expo[e] <= l->degree,
                  "exponent %d of basis function %d exceeds the degree %d",
                  (int)he_expo[e], e / l->dim, l->degree);
  }

  tqf_lsm l2 = *l;
  MomentPlan mp{};
  const bool use_moments = moment_plan(l->dim, l->degree, n_pairs, &mp);
  const int ndense = dense_size(l->dim, l->degree, K);
  LsmArgs a = make_args(&l2, ex_idx[n_ex - 1]);
  hipLaunchKernelGGL(lsm_terminal_kernel, blocks_for(n), 256, 0, s, a, cash,
                     vals);
  // one asset with the full polynomial basis: the two-launch date chain
  const bool small_chain = l->dim == 1 && use_moments && mp.small &&
                           ndense == l->degree + 1 &&
                           l->degree <= kSmallMaxDeg && K <= kSmallMaxK &&
                           n_ex >= 2;
  if (small_chain) {
    double* dense = chain + 6 * kMaxAssets;
    switch (l->degree) {
#define TQF_SMALL_CHAIN(DEG)                                               \
  case DEG:                                                                \
    rc = small_date_chain<DEG>(l, ex_idx, n_ex, df, n_calib, partial,      \
                               chain, dense, cash, ranks, s);              \
    break;
      TQF_SMALL_CHAIN(0)
      TQF_SMALL_CHAIN(1)
      TQF_SMALL_CHAIN(2)
      TQF_SMALL_CHAIN(3)
      TQF_SMALL_CHAIN(4)
#undef TQF_SMALL_CHAIN
    }
    if (rc) return rc;
  }
  for (int e = small_chain ? 0 : n_ex - 1; e > 0; --e) {
    a = make_args(&l2, ex_idx[e - 1]);
    a.mean = mean;
    a.scale = scale;
    if (l->path_discounts) {
      a.dnum = l->path_discounts + (int64_t)(e + 1) * n;
      a.dden = l->path_discounts + (int64_t)e * n;
    }
    const double ratio = df[e + 1] / df[e];
    hipLaunchKernelGGL(lsm_prepare_kernel, blocks_for(n), 256, 0, s, a, ratio,
                       cash, vals, exv, y, bsum);
    hipLaunchKernelGGL(lsm_mean_kernel, 1, kMeanThreads, 0, s, bsum, nb, a.d,
                       n, mean, scale);
    if (use_moments) {
      const int size = mp.rap * mp.cbp;
      double* msum = partial + (int64_t)n_chunks * size;
      if (mp.small && mp.cbp <= 8)
        hipLaunchKernelGGL(lsm_moment_small_kernel<8>, n_chunks, 256, 0, s,
                           a, mp, exv, y, n_calib, chunk, partial);
      else if (mp.small && mp.cbp <= 16)
        hipLaunchKernelGGL(lsm_moment_small_kernel<16>, n_chunks, 256, 0, s,
                           a, mp, exv, y, n_calib, chunk, partial);
      else if (mp.small)
        hipLaunchKernelGGL(lsm_moment_small_kernel<kSmallMoments>, n_chunks,
                           256, 0, s, a, mp, exv, y, n_calib, chunk, partial);
      else if (!launch_moment_fast(mp, a.d, a.deg, a, exv, y, n_calib, chunk,
                                   n_chunks, partial, s))
        hipLaunchKernelGGL(lsm_moment_kernel, n_chunks, kMomThreads, mp.lds,
                           s, a, mp, exv, y, n_calib, chunk, partial);
      hipLaunchKernelGGL(lsm_moment_reduce_kernel, (size + 255) / 256, 256, 0,
                         s, partial, n_chunks, size, msum);
      hipLaunchKernelGGL(lsm_moment_gram_kernel, (kext * kext + 255) / 256,
                         256, 0, s, msum, mp, a.expo, K, a.d, a.deg, g);
    } else {
      hipLaunchKernelGGL(lsm_gram_kernel, dim3(n_chunks, n_pairs), 256, 0, s,
                         a, exv, y, n_calib, chunk, pairs, partial);
      hipLaunchKernelGGL(lsm_gram_reduce_kernel,
                         dim3((kTile * kTile + 255) / 256, n_pairs), 256, 0, s,
                         partial, n_chunks, n_pairs, pairs, kext, g);
    }
    hipLaunchKernelGGL(lsm_solve_kernel, 1, kSolveThreads, 0, s, g, K,
                       l->rcond, beta, rank);
    if (ranks) {
      he = hipMemcpyAsync(ranks + (e - 1), rank, sizeof(int),
                          hipMemcpyDeviceToDevice, s);
      if (he != hipSuccess) return set_error((int)he, "copy rank");
    }
    if (ndense > 0) {
      // the Gram partials are spent: their buffer holds the dense tensor
      double* dense = partial;
      hipLaunchKernelGGL(lsm_dense_beta_kernel, 1, kMaxDense, 0, s, beta,
                         a.expo, K, a.d, a.deg, ndense, dense);
      // degree 2 (the tensor basis of up to 6 assets): the unrolled nest
      if (a.deg == 2 && a.d <= 6) {
        const unsigned fb =
            (unsigned)((n + 256 * kDenseP - 1) / (256 * kDenseP));
        switch (a.d) {
#define TQF_DENSE_UPDATE_DEG2(D)                                             \
  case D:                                                                    \
    hipLaunchKernelGGL((lsm_update_dense_fixed_kernel<D, 2>), fb, 256, 0, s, \
                       a, ratio, exv, dense, cash, vals);                    \
    break;
          TQF_DENSE_UPDATE_DEG2(1)
          TQF_DENSE_UPDATE_DEG2(2)
          TQF_DENSE_UPDATE_DEG2(3)
          TQF_DENSE_UPDATE_DEG2(4)
          TQF_DENSE_UPDATE_DEG2(5)
          TQF_DENSE_UPDATE_DEG2(6)
#undef TQF_DENSE_UPDATE_DEG2
        }
      } else {
        switch (a.d) {
#define TQF_DENSE_UPDATE(D)                                                  \
  case D:                                                                    \
    hipLaunchKernelGGL(lsm_update_dense_kernel<D>, blocks_for(n), 256, 0, s, \
                       a, ratio, exv, dense, cash, vals);                    \
    break;
          TQF_DENSE_UPDATE(1)
          TQF_DENSE_UPDATE(2)
          TQF_DENSE_UPDATE(3)
          TQF_DENSE_UPDATE(4)
          TQF_DENSE_UPDATE(5)
          TQF_DENSE_UPDATE(6)
          TQF_DENSE_UPDATE(7)
          TQF_DENSE_UPDATE(8)
#undef TQF_DENSE_UPDATE
        }
      }
    } else {
      const int ut = update_threads(a);
      const size_t lds = (size_t)a.d * (a.deg + 1) * ut * sizeof(double);
      hipLaunchKernelGGL(lsm_update_kernel, (unsigned)((n + ut - 1) / ut), ut,
                         lds, s, a, ratio, exv, beta, cash, vals);
    }
    rc = check_launch("tqf_lsm_price");
    if (rc) return rc;
  }
  const int64_t first = l->num_calibration > 0 ? l->num_calibration : 0;
  const int64_t count = n - first;
  TQF_REQUIRE(count >= 1, "no pricing rows after the calibration samples");
  const int64_t fb = (count + 255) / 256;
  const double* pd = l->path_discounts;
  hipLaunchKernelGGL(lsm_final_kernel, blocks_for(count), 256, 0, s, cash,
                     vals, df[1] / df[0], pd ? pd + n : nullptr, pd, first,
                     n, fsum);
  hipLaunchKernelGGL(lsm_final_reduce_kernel, 1, 256, 0, s, fsum, fb, count,
                     result);
  return check_launch("tqf_lsm_price finalize");
}

// ---- per-date building blocks (used by the MLP continuation model) --------

extern "C" int tqf_lsm_terminal(const tqf_lsm* l, int32_t t_idx, double* cash,
                                double* vals, void* stream) {
  clear_error();
  int rc = validate(l);
  if (rc) return rc;
  TQF_REQUIRE(t_idx >= 0 && t_idx < l->n_times, "time index out of range");
  TQF_REQUIRE(cash && vals, "null state");
  LsmArgs a = make_args(l, t_idx);
  hipLaunchKernelGGL(lsm_terminal_kernel, blocks_for(l->n_paths), 256, 0,
                     (hipStream_t)stream, a, cash, vals);
  return check_launch("tqf_lsm_terminal");
}

extern "C" int tqf_lsm_date_inputs(const tqf_lsm* l, int32_t t_idx,
                                   double df_ratio, const double* cash,
                                   const double* vals, double* exv, double* y,
                                   double* moments, float* x, float* tgt,
                                   float* w, void* stream) {
  clear_error();
  int rc = validate(l);
  if (rc) return rc;
  TQF_REQUIRE(t_idx >= 0 && t_idx < l->n_times, "time index out of range");
  TQF_REQUIRE(l->dim <= kMlpIn - 1, "the MLP takes at most %d assets",
              kMlpIn - 1);
  TQF_REQUIRE(cash && vals && exv && y && moments && x && tgt && w,
              "null argument");
  TQF_REQUIRE(((uintptr_t)x & 15) == 0, "x must be 16-byte aligned");
  const hipStream_t s = (hipStream_t)stream;
  const int64_t n = l->n_paths;
  const int64_t nb = (n + 255) / 256;
  // moments: [mean 8 | scale 8 | block sums nb * 16]
  double* mean = moments;
  double* scale = moments + kMaxAssets;
  double* bsum = moments + 2 * kMaxAssets;
  LsmArgs a = make_args(l, t_idx);
  hipLaunchKernelGGL(lsm_prepare_kernel, blocks_for(n), 256, 0, s, a, df_ratio,
                     cash, vals, exv, y, bsum);
  hipLaunchKernelGGL(lsm_mean_kernel, 1, kMeanThreads, 0, s, bsum, nb, a.d, n,
                     mean, scale);
  a.mean = mean;
  a.scale = scale;
  hipLaunchKernelGGL(lsm_features_kernel, blocks_for(n), 256, 0, s, a, exv, y,
                     x, tgt, w);
  return check_launch("tqf_lsm_date_inputs");
}

extern "C" int64_t tqf_lsm_moments_size(int64_t n_paths) {
  return (2 * kMaxAssets + ((n_paths + 255) / 256) * 2 * kMaxAssets) * 8;
}

extern "C" int tqf_lsm_update_cont(int64_t n_paths, double strike,
                                   double df_ratio, const double* exv,
                                   const float* cont, double* cash,
                                   double* vals, void* stream) {
  clear_error();
  TQF_REQUIRE(n_paths >= 1 && exv && cont && cash && vals, "bad argument");
  hipLaunchKernelGGL(lsm_update_cont_kernel, blocks_for(n_paths), 256, 0,
                     (hipStream_t)stream, n_paths, strike, df_ratio, exv, cont,
                     cash, vals);
  return check_launch("tqf_lsm_update_cont");
}

extern "C" int tqf_lsm_final(const double* cash, const double* vals,
                             double df0, int64_t first, int64_t n,
                             double* workspace, double* result, void* stream) {
  clear_error();
  TQF_REQUIRE(cash && vals && workspace && result && n > first && first >= 0,
              "bad argument");
  const int64_t count = n - first;
  const int64_t fb = (count + 255) / 256;
  const hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(lsm_final_kernel, blocks_for(count), 256, 0, s, cash,
                     vals, df0, (const double*)nullptr, (const double*)nullptr,
                     first, n, workspace);
  hipLaunchKernelGGL(lsm_final_reduce_kernel, 1, 256, 0, s, workspace, fb,
                     count, result);
  return check_launch("tqf_lsm_final");
}
