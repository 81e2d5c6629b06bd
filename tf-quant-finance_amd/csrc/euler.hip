// Fused Euler-Maruyama steppers.
//
// One thread owns one path for the whole time loop: the state lives in
// VGPRs, normals are produced in registers (Philox/Box-Muller, Sobol + an
// in-register inverse normal CDF, or read from injected draws), and the
// requested times are either recorded or priced on the spot.  This is the
// reference's `_while_loop` / `_euler_step` (tff/models/euler_sampling.py:
// 405-464, 513-537) with the same update order
//     t = times[i+1];  dw = z_i * sqrt_dt[i];
//     x' = (x + dt[i] * a(t, x)) + b(t, x) . dw
// and the same record rule (write at `written_count`, advance on
// keep_mask[i+1]; the host turns that rule into `slot[i]`).
//
// Sobol points: the sequence index of a path splits into a low byte
// (thread-dependent, looked up in a [dim][256] table that stays in L2) and
// high bits that at most two values share across a 256-thread block; the
// high-bit XOR for the block is built in LDS once per chunk of steps.
//
// Payoffs (TQF_OUT_PAYOFF): at a recorded time every lane of a wave holds
// one spot.  Lane s takes strike s and walks the 64 spots of the wave with
// v_readlane (scalar operands), so the wave produces the 64 strike sums of
// its 64 paths without any cross-lane reduction tree.  Per-wave sums go to
// LDS, the block combines its 4 waves in float64 and one finalize kernel
// sums the blocks in a fixed order (deterministic).
#include <mutex>

#include "common.h"
#include "normals.h"

using namespace tqf;

namespace {

constexpr int kBlock = kPathBlock;
constexpr int kWaves = kBlock / kWave;
// Path-major output ([N][k][D]) is staged in the payoff area of LDS, which
// that mode does not use: up to kStageWords values per path ([path]
// [kStageStride], odd stride) flushed as one contiguous segment per path,
// in place of one scattered store per path and record.  It is an
// instantiation of its own (kStage), so the other modes keep their code.
constexpr int kStageWords = 32;
constexpr int kStageStride = kStageWords + 1;
constexpr int kMaxLvLdsBytes = 32768;  // local-vol tables up to this size go to LDS

template <typename R>
struct EulerArgs {
  // model
  R p[8];
  const R* lv_vol;
  const R* lv_t;
  const R* coef;  // per-step constants: HESTON_QE [n_steps][12], HESTON_FT
                  // [n_steps][8] (StepRow)
  R lv_x0, lv_inv_dx;
  int lv_nt, lv_nx, lv_in_lds;
  int lv_edge;  // rows carry [below, interior..., above] values
  // grid
  const R* dt;
  const R* sqrt_dt;
  const R* t_next;
  const int* slot;
  int n_steps, n_slots, init_slot;
  RngArgs<R> rng;
  // output
  int mode, n_strikes;
  R* paths;
  const double* strikes;
  double* partials;
  int64_t n_paths;
  const R* x0;
  int64_t x0_stride;
  int lds_pay_off;   // byte offset of the payoff area in dynamic LDS
  int stage_recs;    // path-major: records staged per flush (D * it <= 32)
  int lds_spot_off;  // byte offset of the per-wave spot rows from it
};

// Kernel-internal model code: HESTON_FT with the variant flag p[5] != 0.
constexpr int kModelHestonOwn = 16;

template <int MODEL>
struct ModelDim {
  static constexpr int D = (MODEL == TQF_MODEL_HESTON_FT ||
                            MODEL == TQF_MODEL_HESTON_QE ||
                            MODEL == kModelHestonOwn)
                               ? 2
                               : 1;
};

// ---- local volatility lookup -----------------------------------------------
template <typename R>
__device__ __forceinline__ R lv_sigma(const EulerArgs<R>& a, const R* tab,
                                      int row, R wt, R x) {
  // interior columns [e, nx - e) with e = 1 in edge mode; column 0 and
  // nx - 1 then hold the values below the first and above the last node
  const int e = a.lv_edge;
  const int ni = a.lv_nx - 2 * e;
  const R fraw = (x - a.lv_x0) * a.lv_inv_dx;
  const R fmax = (R)(ni - 1);
  R fx = fraw < (R)0 ? (R)0 : fraw;
  fx = fx > fmax ? fmax : fx;
  int j = (int)fx;
  j = j > ni - 2 ? ni - 2 : j;
  const R wx = fx - (R)j;
  const R* r0 = tab + (int64_t)row * a.lv_nx;
  const R* r1 = row + 1 < a.lv_nt ? r0 + a.lv_nx : r0;
  int c0 = e + j, c1 = e + j + 1;
  R w = wx;
  if (e) {
    // outside the nodes: the edge column, no interpolation in S
    const bool lo = fraw < (R)0, hi = fraw > fmax;
    c0 = lo ? 0 : hi ? a.lv_nx - 1 : c0;
    c1 = lo ? 0 : hi ? a.lv_nx - 1 : c1;
    w = (lo || hi) ? (R)0 : w;
  }
  const R s0 = r0[c0] + w * (r0[c1] - r0[c0]);
  const R s1 = r1[c0] + w * (r1[c1] - r1[c0]);
  return s0 + wt * (s1 - s0);
}

// ---- per-step update -------------------------------------------------------
// Heston FT per-step row (the model's step_coef): {-dt/2, kappa theta dt,
// -kappa dt, r t_{i+1}, sqrt(dt), record slot (int bits), 0, 0}; one scalar
// load per step brings everything the step needs.
template <typename R>
struct StepRow {
  R c[8];
};
template <int MODEL>
constexpr bool kHasStepRow = MODEL == TQF_MODEL_HESTON_FT ||
                             MODEL == kModelHestonOwn ||
                             MODEL == TQF_MODEL_HULL_WHITE;

// Scalar load of row i (the compiler merges the eight into one s_load).
template <typename R>
__device__ __forceinline__ StepRow<R> ld_row(const R* coef, int64_t i) {
  StepRow<R> r;
#pragma unroll
  for (int k = 0; k < 8; ++k) r.c[k] = ld_uniform(coef, i * 8 + k);
  return r;
}

__device__ __forceinline__ int row_slot(float v) { return __float_as_int(v); }
__device__ __forceinline__ int row_slot(double v) {
  return (int)__double_as_longlong(v);
}

template <typename R, int MODEL>
struct Stepper {
  static constexpr int D = ModelDim<MODEL>::D;
  __device__ __forceinline__ static void step(const EulerArgs<R>& a, R* x,
                                              const R* z, R dt, R sdt,
                                              const R* tab, int lv_row,
                                              R lv_w, int i,
                                              const StepRow<R>& row) {
    if (MODEL == TQF_MODEL_GBM) {
      const R drift = a.p[0] * x[0];
      const R vol = a.p[1] * x[0];
      x[0] = (x[0] + dt * drift) + vol * (z[0] * sdt);
    } else if (MODEL == TQF_MODEL_CONST) {
      x[0] = (x[0] + dt * a.p[0]) + a.p[1] * (z[0] * sdt);
    } else if (MODEL == TQF_MODEL_HESTON_FT || MODEL == kModelHestonOwn) {
      // p: r, kappa, theta, xi*rho, xi*sqrt(1-rho^2).
      // HESTON_FT: full truncation, v+ = max(v, 0) in drift and vol.
      // kModelHestonOwn: HestonModel's own callables (heston_model.py:
      //   143-173): raw v in the drift, sqrt(|v|) in the vol.
      // x' = (x + dt a(x)) + b(x) (z sqrt_dt), with sqrt_dt folded into the
      // volatility factor and each "+ dt a" / "+ b z" a single FMA.
      // With the per-step row {-dt/2, kappa theta dt, -kappa dt, r t', ...}
      // the log-spot state carries x - r t (the deterministic drift is
      // added back where the state is recorded).
      const R vp = MODEL == TQF_MODEL_HESTON_FT ? pos_part(x[1]) : x[1];
      const R sq = fast_sqrt(MODEL == TQF_MODEL_HESTON_FT ? vp : fabs(x[1]));
      const R ssq = sq * sdt;
      x[0] = fma_r(vp, row.c[0], x[0]);
      x[1] = fma_r(vp, row.c[2], x[1] + row.c[1]);
      const R c1 = fma_r(a.p[4], z[1], a.p[3] * z[0]);
      x[0] = fma_r(ssq, z[0], x[0]);
      x[1] = fma_r(ssq, c1, x[1]);
    } else if (MODEL == TQF_MODEL_HULL_WHITE) {
      // exact Gaussian step of x = r - f(0, t) (vector_hull_white.py:
      // 738-760): x' = (e^{-a dt} x + m) + s z, row {e, m, s, f(0, t'), ...}
      x[0] = fma_r(row.c[2], z[0], fma_r(row.c[0], x[0], row.c[1]));
    } else if (MODEL == TQF_MODEL_HESTON_QE) {
      // Andersen QE (heston_model.py:522-572), psi_c = 1.5, gamma = 1/2.
      R c[10];
#pragma unroll
      for (int q = 0; q < 10; ++q) c[q] = ld_uniform(a.coef, (int64_t)i * 12 + q);
      if (c[9] != (R)0) {
        const R v = x[1];
        const R m = c[8] + (v - c[8]) * c[0];
        const R s2 = v * c[1] + c[2];
        // float32 divisions by the reciprocal (fast_div): the lanes of a
        // wave take both branches, and the correctly rounded sequence
        // dominated the step
        const R psi = fast_div(s2, m * m);
        R vn;
        if (psi < (R)1.5) {
          const R pinv = fast_div((R)2, psi);
          const R b2 = pinv - (R)1 + fast_sqrt(pinv * (pinv - (R)1));
          const R am = fast_div(m, (R)1 + b2);
          const R t = fast_sqrt(b2) + z[0];
          vn = am * t * t;
        } else {
          const R p = fast_div(psi - (R)1, psi + (R)1);
          const R beta = fast_div((R)1 - p, m);
          const R u = (R)0.5 * ((R)1 + erf(z[0] * (R)0.70710678118654752));
          vn = u > p ? fast_div(log((R)1 - p) - log((R)1 - u), beta) : (R)0;
        }
        x[0] = x[0] + c[3] + c[4] * v + c[5] * vn +
               fast_sqrt(c[6] * v + c[7] * vn) * z[1];
        x[1] = vn;
      }
    } else {  // LOCAL_VOL: drift (r - q) - sigma^2/2, vol sigma(t, e^x)
      const R s = lv_sigma(a, tab, lv_row, lv_w, x[0]);
      x[0] = (x[0] + dt * (a.p[0] - (R)0.5 * s * s)) + s * (z[0] * sdt);
    }
  }
};

// Payoff of one recorded slot for the calling wave (see file comment).  The
// wave's 64 spots go through a per-wave LDS row and come back as broadcast
// reads (every lane the same address), so a (path, strike) pair costs the
// four arithmetic instructions and no lane read.
template <typename R, int MODEL>
__device__ void wave_payoff(const EulerArgs<R>& a, R* lds_pay, int slot,
                            R spot, bool active) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  R* row = (R*)((unsigned char*)lds_pay + a.lds_spot_off) + wave * kWave;
  // lanes past the last path hold -inf: their payoff max(-inf - K, 0) is 0
  row[lane] = active ? spot : -(R)INFINITY;
  // one wave writes and reads its own row: LDS keeps a wave's accesses in
  // order, the barrier only stops the compiler from moving them
  __builtin_amdgcn_wave_barrier();
  for (int c = 0; c < a.n_strikes; c += kWave) {
    const int ks = c + lane;
    const R k = ks < a.n_strikes ? (R)a.strikes[ks] : (R)0;
    R s1 = 0, s2 = 0;
#pragma unroll 16
    for (int j = 0; j < kWave; ++j) {
      const R sj = row[j];
      R pay = sj - k;
      pay = pay > (R)0 ? pay : (R)0;
      s1 += pay;
      s2 += pay * pay;
    }
    if (ks < a.n_strikes) {
      R* o = lds_pay + ((int64_t)(wave * a.n_slots + slot) * a.n_strikes + ks) * 2;
      o[0] = s1;
      o[1] = s2;
    }
  }
  __builtin_amdgcn_wave_barrier();
}

template <typename R, int MODEL>
__device__ __forceinline__ void record(const EulerArgs<R>& a, R* lds_pay,
                                       int slot, const R* x, R off,
                                       int64_t path, bool active) {
  constexpr int D = ModelDim<MODEL>::D;
  R y[D];
#pragma unroll
  for (int c = 0; c < D; ++c) y[c] = x[c];
  y[0] += off;
  if (a.mode == TQF_OUT_PAYOFF) {
    const R spot = MODEL == TQF_MODEL_GBM ? y[0] : fast_exp(y[0]);
    wave_payoff<R, MODEL>(a, lds_pay, slot, spot, active);
  } else if (active) {
    if (a.mode == TQF_OUT_PATH_MAJOR) {
#pragma unroll
      for (int c = 0; c < D; ++c)
        a.paths[(path * a.n_slots + slot) * D + c] = y[c];
    } else {
      // coalesced rows, written once and not read back here: streaming
      // (nontemporal) stores
#pragma unroll
      for (int c = 0; c < D; ++c)
        __builtin_nontemporal_store(
            y[c], &a.paths[((int64_t)slot * D + c) * a.n_paths + path]);
    }
  }
}

template <typename R, int MODEL, int RNG, bool kStage>
__global__ void __launch_bounds__(kBlock) euler_kernel(EulerArgs<R> a) {
  constexpr int D = ModelDim<MODEL>::D;
  extern __shared__ __align__(16) unsigned char lds[];
  R* lds_pay = (R*)(lds + a.lds_pay_off);

  const int64_t path = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  const bool active = path < a.n_paths;
  const int64_t pc = active ? path : a.n_paths - 1;

  // Local-vol table: stage in LDS when small.
  const R* tab = a.lv_vol;
  if (MODEL == TQF_MODEL_LOCAL_VOL && a.lv_in_lds) {
    R* t = (R*)(lds + rng_lds_bytes(RNG, D));
    for (int e = threadIdx.x; e < a.lv_nt * a.lv_nx; e += kBlock) t[e] = a.lv_vol[e];
    tab = t;
    __syncthreads();
  }

  R x[D];
#pragma unroll
  for (int c = 0; c < D; ++c) x[c] = a.x0[pc * a.x0_stride + c];

  Normals<R, RNG> nrm;
  nrm.init(a.rng, pc, D, lds, a.n_steps);

  // path-major staging (block-uniform bookkeeping, flushed at the uniform
  // record steps).  Each wave flushes the rows of its own 64 paths, so no
  // workgroup barrier is needed: a wave's LDS accesses execute in order.
  // One store moves one whole record (D values); the lanes of a store cover
  // 64 / nrec consecutive paths, each as one contiguous run of nrec
  // records.  A full flush has nrec = stage_recs, a power of two.
  int nrec = 0, first_slot = 0;
  auto flush = [&]() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t p0 = (int64_t)blockIdx.x * kBlock + wave * kWave;
    const R* rows = lds_pay + wave * kWave * kStageStride;
    const bool pow2 = (nrec & (nrec - 1)) == 0;
    const int sh = 31 - __builtin_clz(nrec);
    for (int it = 0; it < nrec; ++it) {
      const int u = it * kWave + lane;
      const int pl = pow2 ? u >> sh : u / nrec;
      const int rec = u - pl * nrec;
      const int64_t pg = p0 + pl;
      const R* src = rows + pl * kStageStride + rec * D;
      R* dst = a.paths + (pg * a.n_slots + first_slot + rec) * D;
      if (pg < a.n_paths) {
        if constexpr (D == 2 && sizeof(R) == 4) {
          // one 8-byte store per record
          typedef float f32x2v __attribute__((ext_vector_type(2)));
          *(f32x2v*)dst = f32x2v{src[0], src[1]};
        } else {
#pragma unroll
          for (int c = 0; c < D; ++c) dst[c] = src[c];
        }
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    nrec = 0;
  };
  auto put = [&](int slot, R off) {
    if (nrec > 0 && slot != first_slot + nrec) flush();
    if (nrec == 0) first_slot = slot;
    if (active) {
      R* b = lds_pay + threadIdx.x * kStageStride + nrec * D;
#pragma unroll
      for (int c = 0; c < D; ++c) b[c] = c == 0 ? x[c] + off : x[c];
    }
    if (++nrec == a.stage_recs) flush();
  };

  if (a.init_slot >= 0) {
    const R off = MODEL == TQF_MODEL_HULL_WHITE ? a.p[0] : (R)0;
    if constexpr (kStage)
      put(a.init_slot, off);
    else
      record<R, MODEL>(a, lds_pay, a.init_slot, x, off, path, active);
  }

  int lv_row = 0;
  const int n_dims = a.n_steps * D;
  auto fetch = [&](int i, R* z) {
    if (D == 2) {
      nrm.get2(a.rng, i, z);
    } else {
#pragma unroll
      for (int k = 0; k < D; ++k) z[k] = nrm.get(a.rng, i, k);
    }
  };
  // Chunks of kSobolChunk steps (the Sobol high-bit words are rebuilt per
  // chunk).  Inside a chunk the normals of step i+1 are produced during
  // step i -- they do not depend on the state -- with no branch between
  // them and the update, so their LDS and memory latency overlaps it.
  for (int c0 = 0; c0 < a.n_steps; c0 += kSobolChunk) {
    nrm.refill(a.rng, c0, n_dims);
    const int c1 = c0 + kSobolChunk < a.n_steps ? c0 + kSobolChunk : a.n_steps;
    StepRow<R> row_next;
    if constexpr (kHasStepRow<MODEL>) row_next = ld_row(a.coef, c0);
    // one step: consumes zc (its normals, fetched a step earlier) and the
    // prefetched row, and fetches step i+1's normals into zn
    auto one_step = [&](int i, const R* zc, R* zn) {
      fetch(i + 1 < c1 ? i + 1 : i, zn);  // the chunk's last one is redundant
      // per-step scalars: the Heston FT row is loaded one step ahead, so the
      // update never waits on a scalar load issued behind the LDS reads
      StepRow<R> row;
      R dt, sdt;
      int s;
      if constexpr (kHasStepRow<MODEL>) {
        row = row_next;
        row_next = ld_row(a.coef, i + 1 < a.n_steps ? i + 1 : i);
        dt = (R)0;
        sdt = row.c[4];
        s = row_slot(row.c[5]);
      } else {
        dt = ld_uniform(a.dt, i);
        sdt = ld_uniform(a.sqrt_dt, i);
        s = ld_uniform(a.slot, i);
      }
      R lv_w = 0;
      if (MODEL == TQF_MODEL_LOCAL_VOL) {
        const R t = ld_uniform(a.t_next, i);
        while (lv_row + 1 < a.lv_nt && a.lv_t[lv_row + 1] <= t) ++lv_row;
        if (lv_row + 1 < a.lv_nt) {
          const R t0 = a.lv_t[lv_row], t1 = a.lv_t[lv_row + 1];
          lv_w = t <= t0 ? (R)0 : fast_div(t - t0, t1 - t0);
        }
      }
      Stepper<R, MODEL>::step(a, x, zc, dt, sdt, tab, lv_row, lv_w, i, row);
      if (s >= 0) {
        const R off = kHasStepRow<MODEL> ? row.c[3] : (R)0;
        if constexpr (kStage)
          put(s, off);
        else
          record<R, MODEL>(a, lds_pay, s, x, off, path, active);
      }
    };
    // two steps per iteration with the normals alternating between zA and
    // zB, so the pipelining needs no register copies
    R zA[D], zB[D];
    fetch(c0, zA);
    int i = c0;
    for (; i + 1 < c1; i += 2) {
      one_step(i, zA, zB);
      one_step(i + 1, zB, zA);
    }
    if (i < c1) one_step(i, zA, zB);
  }
  if constexpr (kStage) {
    if (nrec > 0) flush();
  }

  if (a.mode == TQF_OUT_PAYOFF) {
    __syncthreads();
    const int n_out = a.n_slots * a.n_strikes;
    for (int o = threadIdx.x; o < 2 * n_out; o += kBlock) {
      double acc = 0;
      for (int w = 0; w < kWaves; ++w) {
        // waves with no active path never wrote; they hold zeros
        const int64_t wb = (int64_t)blockIdx.x * kBlock + w * kWave;
        if (wb < a.n_paths) acc += (double)lds_pay[(int64_t)w * 2 * n_out + o];
      }
      a.partials[(int64_t)blockIdx.x * 2 * n_out + o] = acc;
    }
  }
}

// Level 1 of the block-partial reduction: thread t of workgroup (c, y) sums
// entry y*256 + t (of the 2 * n_out interleaved sums) over chunk c of
// kReduceChunk consecutive blocks; consecutive threads read consecutive
// doubles.
constexpr int kReduceChunk = 32;

__global__ void __launch_bounds__(256)
payoff_reduce_kernel(const double* __restrict__ partials, int64_t n_blocks,
                     int n_out2, double* __restrict__ chunks) {
  const int e = blockIdx.y * 256 + threadIdx.x;
  if (e >= n_out2) return;
  const int64_t b0 = (int64_t)blockIdx.x * kReduceChunk;
  int64_t b1 = b0 + kReduceChunk;
  if (b1 > n_blocks) b1 = n_blocks;
  double s = 0;
  // unrolled: several blocks' loads in flight, the sum keeps its order
#pragma unroll 8
  for (int64_t b = b0; b < b1; ++b) s += partials[b * n_out2 + e];
  chunks[(int64_t)blockIdx.x * n_out2 + e] = s;
}

// Level 2: one workgroup per output sums the chunk sums with a fixed-shape
// LDS tree (deterministic), then forms the mean and standard error.
__global__ void __launch_bounds__(256)
payoff_finalize_kernel(const double* __restrict__ partials, int64_t n_blocks,
                       int n_out, int n_strikes, int64_t n_paths,
                       const double* __restrict__ discount,
                       double* __restrict__ prices,
                       double* __restrict__ stderrs, double* __restrict__ sums) {
  // `partials` here are the level-1 chunk sums, [n_blocks][n_out][2]
  __shared__ double r1[256], r2[256];
  const int o = blockIdx.x;
  double s1 = 0, s2 = 0;
  for (int64_t b = threadIdx.x; b < n_blocks; b += 256) {
    s1 += partials[(b * n_out + o) * 2];
    s2 += partials[(b * n_out + o) * 2 + 1];
  }
  r1[threadIdx.x] = s1;
  r2[threadIdx.x] = s2;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      r1[threadIdx.x] += r1[threadIdx.x + w];
      r2[threadIdx.x] += r2[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  s1 = r1[0];
  s2 = r2[0];
  const double n = (double)n_paths;
  const double mean = s1 / n;
  double var = n > 1 ? (s2 - n * mean * mean) / (n - 1) : 0.0;
  var = var > 0 ? var : 0.0;
  const double df = discount ? discount[o / n_strikes] : 1.0;
  if (prices) prices[o] = df * mean;
  if (stderrs) stderrs[o] = df * sqrt(var / n);
  if (sums) {
    sums[2 * o] = s1;
    sums[2 * o + 1] = s2;
  }
}

// Reduces the per-block payoff sums in `ws` (layout [block][slot][strike]
// [2]) to prices and standard errors: level 1 into chunk sums stored after
// the block sums, level 2 one workgroup per output.
int launch_finalize(double* ws, int64_t n_paths, int n_slots, int n_strikes,
                    const double* discount, double* prices, double* stderrs,
                    double* sums, hipStream_t stream) {
  const int64_t n_blocks = (n_paths + kBlock - 1) / kBlock;
  const int n_out = n_slots * n_strikes;
  const int64_t n_chunks = (n_blocks + kReduceChunk - 1) / kReduceChunk;
  double* chunks = ws + n_blocks * n_out * 2;
  hipLaunchKernelGGL(payoff_reduce_kernel,
                     dim3((unsigned)n_chunks, (unsigned)((2 * n_out + 255) / 256)),
                     dim3(256), 0, stream, ws, n_blocks, 2 * n_out, chunks);
  int rc = check_launch("payoff reduce");
  if (rc) return rc;
  hipLaunchKernelGGL(payoff_finalize_kernel, dim3(n_out), dim3(256), 0, stream,
                     chunks, n_chunks, n_out, n_strikes, n_paths, discount,
                     prices, stderrs, sums);
  return check_launch("payoff finalize");
}

template <typename R>
int launch_typed(const tqf_model* m, const tqf_grid* g, const tqf_rng* r,
                 const tqf_output* out, int64_t n_paths, const void* x0,
                 int64_t x0_stride, hipStream_t stream) {
  EulerArgs<R> a{};
  for (int i = 0; i < 8; ++i) a.p[i] = (R)m->p[i];
  const int D =
      (m->kind == TQF_MODEL_HESTON_FT || m->kind == TQF_MODEL_HESTON_QE) ? 2 : 1;
  if (m->kind == TQF_MODEL_HESTON_FT) {
    const double xi = m->p[3], rho = m->p[4];
    a.p[3] = (R)(xi * rho);
    a.p[4] = (R)(xi * std::sqrt(1.0 - rho * rho));
  }
  a.lv_vol = (const R*)m->lv_vol;
  a.coef = (const R*)m->step_coef;
  a.lv_t = (const R*)m->lv_t;
  a.lv_x0 = (R)m->lv_x0;
  a.lv_inv_dx = m->lv_dx > 0 ? (R)(1.0 / m->lv_dx) : (R)0;
  a.lv_nt = m->lv_nt;
  a.lv_nx = m->lv_nx;
  a.lv_edge = (m->kind == TQF_MODEL_LOCAL_VOL && m->p[1] != 0.0) ? 1 : 0;
  a.dt = (const R*)g->dt;
  a.sqrt_dt = (const R*)g->sqrt_dt;
  a.t_next = (const R*)g->t_next;
  a.slot = g->slot;
  a.n_steps = g->n_steps;
  a.n_slots = g->n_slots;
  a.init_slot = g->init_slot;
  fill_rng_args(a.rng, r, n_paths);
  a.mode = out->mode == TQF_OUT_PAYOFF_PARTIALS ? TQF_OUT_PAYOFF : out->mode;
  a.n_strikes = out->n_strikes;
  a.paths = (R*)out->paths;
  a.strikes = out->strikes;
  a.partials = out->workspace;
  a.n_paths = n_paths;
  a.x0 = (const R*)x0;
  a.x0_stride = x0_stride;

  // dynamic LDS: Sobol high-bit words | local-vol table | payoff sums and
  // spot rows; the float32 inverse-CDF table is static LDS (kIcdfLdsBytes)
  size_t pay = 0;
  const bool stage = a.mode == TQF_OUT_PATH_MAJOR;
  a.stage_recs = kStageWords / D;
  if (stage) pay = (size_t)kBlock * kStageStride * sizeof(R);
  if (a.mode == TQF_OUT_PAYOFF) {
    const size_t sums = (size_t)kWaves * g->n_slots * out->n_strikes * 2 * sizeof(R);
    a.lds_spot_off = (int)((sums + 15) & ~(size_t)15);
    pay = (size_t)a.lds_spot_off + (size_t)kWaves * kWave * sizeof(R);
  }
  size_t lds = (size_t)rng_lds_bytes(r->kind, D);
  if (m->kind == TQF_MODEL_LOCAL_VOL) {
    const int64_t lv = (int64_t)m->lv_nt * m->lv_nx * (int64_t)sizeof(R);
    const int64_t total = (int64_t)((lds + lv + 15) & ~(size_t)15) +
                          (int64_t)pay + kIcdfLdsBytes;
    a.lv_in_lds = lv <= kMaxLvLdsBytes && total <= 64 * 1024;
    if (a.lv_in_lds) lds += (size_t)lv;
  }
  lds = (lds + 15) & ~(size_t)15;
  a.lds_pay_off = (int)lds;
  lds += pay;
  if (lds == 0) lds = 16;

  const unsigned n_blocks = (unsigned)((n_paths + kBlock - 1) / kBlock);
  void (*kern)(EulerArgs<R>) = nullptr;
  const int kind = (m->kind == TQF_MODEL_HESTON_FT && m->p[5] != 0.0)
                       ? kModelHestonOwn
                       : m->kind;
#define TQF_PICK(MODEL, RNG) \
  if (kind == MODEL && r->kind == RNG)                                   \
    kern = stage ? euler_kernel<R, MODEL, RNG, true>                     \
                 : euler_kernel<R, MODEL, RNG, false>;
#define TQF_PICK_RNG(MODEL)                       \
  TQF_PICK(MODEL, TQF_RNG_DRAWS)                  \
  TQF_PICK(MODEL, TQF_RNG_PHILOX)                 \
  TQF_PICK(MODEL, TQF_RNG_PHILOX_ANTITHETIC)      \
  TQF_PICK(MODEL, TQF_RNG_SOBOL)                  \
  TQF_PICK(MODEL, TQF_RNG_SOBOL_SHIFTED)
  TQF_PICK_RNG(TQF_MODEL_GBM)
  TQF_PICK_RNG(TQF_MODEL_CONST)
  TQF_PICK_RNG(TQF_MODEL_HESTON_FT)
  TQF_PICK_RNG(kModelHestonOwn)
  TQF_PICK_RNG(TQF_MODEL_LOCAL_VOL)
  TQF_PICK_RNG(TQF_MODEL_HESTON_QE)
  TQF_PICK_RNG(TQF_MODEL_HULL_WHITE)
#undef TQF_PICK_RNG
#undef TQF_PICK
  if (!kern) return set_error(TQF_ENOTIMPL, "unsupported model/rng combination");
  hipLaunchKernelGGL(kern, dim3(n_blocks), dim3(kBlock), lds, stream, a);
  int rc = check_launch("tqf_euler");
  if (rc) return rc;
  if (out->mode == TQF_OUT_PAYOFF)
    rc = launch_finalize(out->workspace, n_paths, g->n_slots, out->n_strikes,
                         out->discount, out->prices, out->stderrs, out->sums,
                         stream);
  return rc;
}

// ---- forward tangents (differentiable sampling) ----------------------------
// The fused paths as a function of (x0, model parameters): one thread per
// path re-simulates it with the forward kernel's own normals and Stepper
// (the same path bit for bit) and carries the tangents of the state with
// respect to every input, as custom_loops.for_loop carries the Jacobian of
// the loop state (tff/math/custom_loops.py:20-214, euler_sampling.py:
// 467-510).  Memory does not grow with the number of steps.
//   VJP: sum over paths, records and components of g * dX/dtheta (g the
//        incoming gradient of the [N][k][d] paths), per-block partials
//        reduced in a fixed order.
//   JVP: the tangent paths sum_q u_q dX/dtheta_q in the paths' layout.
// Inputs theta: x0[0..d) then the model's parameters
//   GBM (mu, sigma); CONST (a, b); HESTON_FT (r, kappa, theta, xi, rho).
constexpr int kMaxTan = 7;

template <typename R>
struct TangentArgs {
  int mode;
  int n_par;                 // model parameters (inputs after x0)
  R xi, rho, kappa, theta;   // Heston raw parameters
  const R* grad_paths;       // VJP [N][k][d]
  double* partials;          // VJP [blocks][kMaxTan]
  R* grad_x0;                // VJP, per-path x0: [N][d] or null
  const double* direction;   // JVP [d + n_par]
  const R* dir_x0;           // JVP, per-path x0 direction [N][d] or null
  R* tangent_paths;          // JVP [N][k][d]
};

template <typename R, int MODEL, int RNG>
__global__ void __launch_bounds__(kBlock)
euler_tangent_kernel(EulerArgs<R> a, TangentArgs<R> t) {
  constexpr int D = ModelDim<MODEL>::D;
  constexpr bool kHeston =
      MODEL == TQF_MODEL_HESTON_FT || MODEL == kModelHestonOwn;
  extern __shared__ __align__(16) unsigned char lds[];
  __shared__ double red[kWaves][kMaxTan];
  const int64_t path = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  const bool active = path < a.n_paths;
  const int64_t pc = active ? path : a.n_paths - 1;
  const int m = D + t.n_par;  // inputs
  R x[D];
#pragma unroll
  for (int c = 0; c < D; ++c) x[c] = a.x0[pc * a.x0_stride + c];
  // T[c][q] = dX_c / dtheta_q
  R T[D][kMaxTan];
#pragma unroll
  for (int c = 0; c < D; ++c)
#pragma unroll
    for (int q = 0; q < kMaxTan; ++q) T[c][q] = (c == q) ? (R)1 : (R)0;
  double acc[kMaxTan];
#pragma unroll
  for (int q = 0; q < kMaxTan; ++q) acc[q] = 0;
  double accx0[D];
#pragma unroll
  for (int c = 0; c < D; ++c) accx0[c] = 0;
  // JVP direction of this path
  R u[kMaxTan];
  if (t.mode == TQF_TANGENT_JVP) {
#pragma unroll
    for (int q = 0; q < kMaxTan; ++q)
      u[q] = q < m ? (R)ld_uniform(t.direction, q) : (R)0;
    if (t.dir_x0) {
#pragma unroll
      for (int c = 0; c < D; ++c) u[c] = t.dir_x0[pc * D + c];
    }
  }
  // the r-input of the Heston log-spot: the recorded value carries + r t
  const int q_r = D;  // HESTON_FT parameter 0 is r
  auto rec = [&](int slot, R time) {
    if (!active) return;
    const int64_t o = (path * a.n_slots + slot) * D;
    if (t.mode == TQF_TANGENT_VJP) {
#pragma unroll
      for (int c = 0; c < D; ++c) {
        const double g = (double)t.grad_paths[o + c];
#pragma unroll
        for (int q = 0; q < kMaxTan; ++q) acc[q] += g * (double)T[c][q];
        if (kHeston && c == 0) acc[q_r] += g * (double)time;
      }
    } else {
#pragma unroll
      for (int c = 0; c < D; ++c) {
        R v = 0;
#pragma unroll
        for (int q = 0; q < kMaxTan; ++q) v += u[q] * T[c][q];
        if (kHeston && c == 0) v += u[q_r] * time;
        t.tangent_paths[o + c] = v;
      }
    }
  };
  if (a.init_slot >= 0) rec(a.init_slot, (R)0);

  Normals<R, RNG> nrm;
  nrm.init(a.rng, pc, D, lds, a.n_steps);
  const int n_dims = a.n_steps * D;
  for (int c0 = 0; c0 < a.n_steps; c0 += kSobolChunk) {
    nrm.refill(a.rng, c0, n_dims);
    const int c1 = c0 + kSobolChunk < a.n_steps ? c0 + kSobolChunk : a.n_steps;
    for (int i = c0; i < c1; ++i) {
      R z[D];
#pragma unroll
      for (int k = 0; k < D; ++k) z[k] = nrm.get(a.rng, i, k);
      StepRow<R> row;
      R sdt;
      int s;
      const R dt = ld_uniform(a.dt, i);
      if constexpr (kHasStepRow<MODEL>) {
        row = ld_row(a.coef, i);
        sdt = row.c[4];
        s = row_slot(row.c[5]);
      } else {
        sdt = ld_uniform(a.sqrt_dt, i);
        s = ld_uniform(a.slot, i);
      }
      // tangents from the state before the step
      if (MODEL == TQF_MODEL_GBM) {
        const R dw = z[0] * sdt;
        const R f = ((R)1 + dt * a.p[0]) + a.p[1] * dw;
#pragma unroll
        for (int q = 0; q < kMaxTan; ++q) T[0][q] *= f;
        T[0][D + 0] += dt * x[0];   // mu
        T[0][D + 1] += x[0] * dw;   // sigma
      } else if (MODEL == TQF_MODEL_CONST) {
        T[0][D + 0] += dt;          // a
        T[0][D + 1] += z[0] * sdt;  // b
      } else if (kHeston) {
        const R v = x[1];
        const bool ft = MODEL == TQF_MODEL_HESTON_FT;
        const bool pos = v > (R)0;
        const R vp = ft ? (pos ? v : (R)0) : v;
        const R sq = sqrt(ft ? vp : fabs(v));
        // d sq / d v
        const R hs = sq > (R)0 ? (ft ? (R)0.5 / sq : (pos ? (R)0.5 : (R)-0.5) / sq)
                               : (R)0;
        const R dw0 = z[0] * sdt, dw1 = z[1] * sdt;
        const R s1r = sqrt((R)1 - t.rho * t.rho);
        const R cz = a.p[3] * dw0 + a.p[4] * dw1;  // xi rho dw0 + xi s dw1
#pragma unroll
        for (int q = 0; q < kMaxTan; ++q) {
          const R tvp = ft ? (pos ? T[1][q] : (R)0) : T[1][q];
          const R tsq = T[1][q] * hs;
          T[0][q] += (R)-0.5 * dt * tvp + tsq * dw0;
          T[1][q] += -t.kappa * dt * tvp + tsq * cz;
        }
        T[1][D + 1] += dt * (t.theta - vp);                    // kappa
        T[1][D + 2] += dt * t.kappa;                           // theta
        T[1][D + 3] += sq * (t.rho * dw0 + s1r * dw1);         // xi
        T[1][D + 4] += t.xi * sq * (dw0 - t.rho / s1r * dw1);  // rho
      }
      Stepper<R, MODEL>::step(a, x, z, dt, sdt, nullptr, 0, (R)0, i, row);
      if (s >= 0) rec(s, ld_uniform(a.t_next, i));
    }
  }
  if (t.mode != TQF_TANGENT_VJP) return;
  if (t.grad_x0) {
    // per-path initial states: their gradient is per path
    if (active) {
#pragma unroll
      for (int c = 0; c < D; ++c) t.grad_x0[path * D + c] = (R)acc[c];
    }
#pragma unroll
    for (int c = 0; c < D; ++c) acc[c] = 0;
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < kMaxTan; ++q) {
    double v = acc[q];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if (lane == 0) red[wave][q] = v;
  }
  __syncthreads();
  if (threadIdx.x < kMaxTan) {
    double v = 0;
    for (int w = 0; w < kWaves; ++w) v += red[w][threadIdx.x];
    t.partials[(int64_t)blockIdx.x * kMaxTan + threadIdx.x] = v;
  }
}

// Fixed-order sum of the per-block tangent partials: thread q of one
// workgroup per input adds the blocks in order through a 256-lane tree.
__global__ void __launch_bounds__(256)
tangent_reduce_kernel(const double* __restrict__ partials, int64_t n_blocks,
                      int m, double* __restrict__ out) {
  __shared__ double r[256];
  const int q = blockIdx.x;
  double s = 0;
  for (int64_t b = threadIdx.x; b < n_blocks; b += 256)
    s += partials[b * kMaxTan + q];
  r[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) r[threadIdx.x] += r[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0 && q < m) out[q] = r[0];
}

template <typename R>
int launch_tangent(const tqf_model* m, const tqf_grid* g, const tqf_rng* r,
                   const tqf_tangent* tg, int64_t n_paths, const void* x0,
                   int64_t x0_stride, hipStream_t stream) {
  EulerArgs<R> a{};
  for (int i = 0; i < 8; ++i) a.p[i] = (R)m->p[i];
  TangentArgs<R> t{};
  const int kind = (m->kind == TQF_MODEL_HESTON_FT && m->p[5] != 0.0)
                       ? kModelHestonOwn
                       : m->kind;
  const int D = m->kind == TQF_MODEL_HESTON_FT ? 2 : 1;
  if (m->kind == TQF_MODEL_HESTON_FT) {
    const double xi = m->p[3], rho = m->p[4];
    a.p[3] = (R)(xi * rho);
    a.p[4] = (R)(xi * std::sqrt(1.0 - rho * rho));
    t.kappa = (R)m->p[1];
    t.theta = (R)m->p[2];
    t.xi = (R)xi;
    t.rho = (R)rho;
    t.n_par = 5;
  } else {
    t.n_par = 2;
  }
  a.coef = (const R*)m->step_coef;
  a.dt = (const R*)g->dt;
  a.sqrt_dt = (const R*)g->sqrt_dt;
  a.t_next = (const R*)g->t_next;
  a.slot = g->slot;
  a.n_steps = g->n_steps;
  a.n_slots = g->n_slots;
  a.init_slot = g->init_slot;
  fill_rng_args(a.rng, r, n_paths);
  a.n_paths = n_paths;
  a.x0 = (const R*)x0;
  a.x0_stride = x0_stride;
  t.mode = tg->mode;
  t.grad_paths = (const R*)tg->grad_paths;
  t.partials = (double*)tg->workspace;
  t.grad_x0 = (R*)tg->grad_x0;
  t.direction = tg->direction;
  t.dir_x0 = (const R*)tg->dir_x0;
  t.tangent_paths = (R*)tg->tangent_paths;
  void (*kern)(EulerArgs<R>, TangentArgs<R>) = nullptr;
#define TQF_PICK_T(MODEL, RNG) \
  if (kind == MODEL && r->kind == RNG) kern = euler_tangent_kernel<R, MODEL, RNG>;
#define TQF_PICK_T_RNG(MODEL)                   \
  TQF_PICK_T(MODEL, TQF_RNG_DRAWS)              \
  TQF_PICK_T(MODEL, TQF_RNG_PHILOX)             \
  TQF_PICK_T(MODEL, TQF_RNG_PHILOX_ANTITHETIC)
  TQF_PICK_T_RNG(TQF_MODEL_GBM)
  TQF_PICK_T_RNG(TQF_MODEL_CONST)
  TQF_PICK_T_RNG(TQF_MODEL_HESTON_FT)
  TQF_PICK_T_RNG(kModelHestonOwn)
#undef TQF_PICK_T_RNG
#undef TQF_PICK_T
  if (!kern)
    return set_error(TQF_ENOTIMPL, "no tangent kernel for this model/rng");
  const unsigned n_blocks = (unsigned)((n_paths + kBlock - 1) / kBlock);
  hipLaunchKernelGGL(kern, dim3(n_blocks), dim3(kBlock), 16, stream, a, t);
  int rc = check_launch("tqf_euler_tangent");
  if (rc || tg->mode != TQF_TANGENT_VJP) return rc;
  hipLaunchKernelGGL(tangent_reduce_kernel, dim3(D + t.n_par), dim3(256), 0,
                     stream, (const double*)tg->workspace, (int64_t)n_blocks,
                     D + t.n_par, tg->grad_params);
  return check_launch("tqf_euler_tangent reduce");
}

}  // namespace

extern "C" int64_t tqf_euler_workspace_size(int64_t n_paths, int32_t n_slots,
                                            int32_t n_strikes) {
  const int64_t n_blocks = (n_paths + kBlock - 1) / kBlock;
  const int64_t n_chunks = (n_blocks + kReduceChunk - 1) / kReduceChunk;
  return (n_blocks + n_chunks) * (int64_t)n_slots * n_strikes * 2 *
         (int64_t)sizeof(double);
}

extern "C" int tqf_euler(const tqf_model* m, const tqf_grid* g,
                         const tqf_rng* r, const tqf_output* out,
                         int64_t n_paths, const void* x0,
                         int64_t x0_stride_path, int32_t dtype, void* stream) {
  clear_error();
  TQF_REQUIRE(m && g && r && out, "null descriptor");
  TQF_REQUIRE(dtype == TQF_F32 || dtype == TQF_F64, "bad dtype %d", dtype);
  TQF_REQUIRE(n_paths >= 1, "num_samples must be positive, got %lld",
              (long long)n_paths);
  TQF_REQUIRE(m->kind >= TQF_MODEL_GBM && m->kind <= TQF_MODEL_HULL_WHITE,
              "unknown model kind %d", m->kind);
  const int D =
      (m->kind == TQF_MODEL_HESTON_FT || m->kind == TQF_MODEL_HESTON_QE) ? 2 : 1;
  if (m->kind == TQF_MODEL_HESTON_QE || m->kind == TQF_MODEL_HESTON_FT) {
    TQF_REQUIRE(m->step_coef != nullptr || g->n_steps == 0,
                "Heston models need their per-step constants (step_coef)");
  }
  if (m->kind == TQF_MODEL_HULL_WHITE) {
    TQF_REQUIRE(m->step_coef != nullptr || g->n_steps == 0,
                "the Hull-White model needs its per-step constants "
                "(step_coef)");
    TQF_REQUIRE(out->mode == TQF_OUT_PATH_MAJOR ||
                    out->mode == TQF_OUT_STEP_MAJOR,
                "the Hull-White model records rates, not option payoffs");
  }
  TQF_REQUIRE(m->dim == D, "model dim %d does not match kind (%d)", m->dim, D);
  TQF_REQUIRE(g->n_steps >= 0 && g->n_slots >= 1, "bad grid sizes");
  TQF_REQUIRE(g->n_rng_steps >= g->n_steps, "n_rng_steps < n_steps");
  TQF_REQUIRE(g->init_slot >= -1 && g->init_slot < g->n_slots, "bad init_slot");
  TQF_REQUIRE(g->n_steps == 0 || (g->dt && g->sqrt_dt && g->slot),
              "null grid arrays");
  TQF_REQUIRE(x0 != nullptr, "null initial state");
  TQF_REQUIRE(x0_stride_path == 0 || x0_stride_path == D,
              "x0_stride_path must be 0 or dim");
  TQF_REQUIRE(r->kind >= TQF_RNG_DRAWS && r->kind <= TQF_RNG_SOBOL_SHIFTED,
              "unknown rng kind %d", r->kind);
  if (r->kind == TQF_RNG_DRAWS) {
    TQF_REQUIRE(r->draws != nullptr, "null normal_draws");
  }
  if (r->kind == TQF_RNG_PHILOX_ANTITHETIC) {
    TQF_REQUIRE(n_paths % 2 == 0,
                "First dimension of `sample_shape` should be even for "
                "PSEUDO_ANTITHETIC random type");
  }
  if (r->kind == TQF_RNG_SOBOL || r->kind == TQF_RNG_SOBOL_SHIFTED) {
    TQF_REQUIRE(r->sobol_v && r->sobol_lo, "null Sobol tables");
    TQF_REQUIRE((int64_t)r->sobol_ndims >= (int64_t)g->n_steps * D,
                "Sobol tables hold %d dimensions, the grid needs %lld",
                r->sobol_ndims, (long long)g->n_steps * D);
    TQF_REQUIRE(r->sobol_first_index >= 0 &&
                    r->sobol_first_index + n_paths <= (int64_t)0xffffffffll + 1,
                "Sobol index range exceeds 32 bits");
    TQF_REQUIRE(r->sobol_nd >= 1 && r->sobol_nd <= 32, "bad Sobol digits");
    TQF_REQUIRE(r->kind == TQF_RNG_SOBOL || r->sobol_nd == 32,
                "the shifted transform needs nd = 32");
  }
  if (m->kind == TQF_MODEL_LOCAL_VOL) {
    TQF_REQUIRE(m->lv_vol && m->lv_t && m->lv_nt >= 1 && m->lv_nx >= 2 &&
                    m->lv_dx > 0,
                "bad local volatility table");
    TQF_REQUIRE(m->p[1] == 0.0 || m->lv_nx >= 4,
                "an edge-mode local volatility table needs two interior "
                "columns and the two edge columns");
    TQF_REQUIRE(g->t_next != nullptr, "null t_next");
  }
  TQF_REQUIRE(out->mode >= TQF_OUT_PATH_MAJOR &&
                  out->mode <= TQF_OUT_PAYOFF_PARTIALS,
              "unknown output mode %d", out->mode);
  if (out->mode == TQF_OUT_PAYOFF || out->mode == TQF_OUT_PAYOFF_PARTIALS) {
    TQF_REQUIRE(out->n_strikes >= 1 && out->strikes && out->workspace,
                "payoff mode needs strikes and a workspace");
    TQF_REQUIRE(out->mode == TQF_OUT_PAYOFF_PARTIALS ||
                    (out->discount && out->prices && out->stderrs),
                "payoff mode needs discount, prices and stderrs");
    const int64_t lds = (int64_t)kWaves * g->n_slots * out->n_strikes * 2 *
                        dtype_size(dtype);
    TQF_REQUIRE(lds <= 48 * 1024,
                "payoff mode supports n_slots * n_strikes <= %d for this dtype",
                (int)(48 * 1024 / (kWaves * 2 * dtype_size(dtype))));
  } else {
    TQF_REQUIRE(out->paths != nullptr, "null output paths");
    TQF_REQUIRE(((uintptr_t)out->paths % (D * dtype_size(dtype))) == 0,
                "output paths must be aligned to one record (%d bytes)",
                (int)(D * dtype_size(dtype)));
  }
  const hipStream_t s = (hipStream_t)stream;
  if (dtype == TQF_F32)
    return launch_typed<float>(m, g, r, out, n_paths, x0, x0_stride_path, s);
  return launch_typed<double>(m, g, r, out, n_paths, x0, x0_stride_path, s);
}

extern "C" int tqf_payoff_finalize(double* workspace, int64_t n_paths,
                                   int32_t n_slots, int32_t n_strikes,
                                   const double* discount, double* prices,
                                   double* stderrs, double* sums,
                                   void* stream) {
  clear_error();
  TQF_REQUIRE(n_paths >= 2, "need at least two paths, got %lld",
              (long long)n_paths);
  TQF_REQUIRE(n_slots >= 1 && n_strikes >= 1, "bad payoff sizes");
  TQF_REQUIRE(workspace && discount && prices && stderrs,
              "null payoff buffers");
  return launch_finalize(workspace, n_paths, n_slots, n_strikes, discount,
                         prices, stderrs, sums, (hipStream_t)stream);
}

extern "C" int64_t tqf_euler_tangent_workspace_size(int64_t n_paths) {
  const int64_t n_blocks = (n_paths + kBlock - 1) / kBlock;
  return n_blocks * kMaxTan * (int64_t)sizeof(double);
}

extern "C" int tqf_euler_tangent(const tqf_model* m, const tqf_grid* g,
                                 const tqf_rng* r, const tqf_tangent* tg,
                                 int64_t n_paths, const void* x0,
                                 int64_t x0_stride_path, int32_t dtype,
                                 void* stream) {
  clear_error();
  TQF_REQUIRE(m && g && r && tg, "null descriptor");
  TQF_REQUIRE(dtype == TQF_F32 || dtype == TQF_F64, "bad dtype %d", dtype);
  TQF_REQUIRE(n_paths >= 1, "num_samples must be positive");
  TQF_REQUIRE(m->kind == TQF_MODEL_GBM || m->kind == TQF_MODEL_CONST ||
                  m->kind == TQF_MODEL_HESTON_FT,
              "tangents exist for the GBM, constant-coefficient and Heston "
              "Euler models, not kind %d", m->kind);
  TQF_REQUIRE(r->kind == TQF_RNG_DRAWS || r->kind == TQF_RNG_PHILOX ||
                  r->kind == TQF_RNG_PHILOX_ANTITHETIC,
              "tangents read injected draws or the Philox stream, not rng "
              "kind %d", r->kind);
  const int D = m->kind == TQF_MODEL_HESTON_FT ? 2 : 1;
  TQF_REQUIRE(m->dim == D, "model dim %d does not match kind", m->dim);
  TQF_REQUIRE(m->kind != TQF_MODEL_HESTON_FT || m->step_coef ||
                  g->n_steps == 0,
              "Heston models need their per-step constants (step_coef)");
  TQF_REQUIRE(g->n_steps >= 0 && g->n_slots >= 1, "bad grid sizes");
  TQF_REQUIRE(g->init_slot >= -1 && g->init_slot < g->n_slots, "bad init_slot");
  TQF_REQUIRE(g->n_steps == 0 || (g->dt && g->sqrt_dt && g->slot && g->t_next),
              "null grid arrays");
  TQF_REQUIRE(x0 != nullptr, "null initial state");
  TQF_REQUIRE(x0_stride_path == 0 || x0_stride_path == D,
              "x0_stride_path must be 0 or dim");
  TQF_REQUIRE(r->kind != TQF_RNG_DRAWS || r->draws, "null normal_draws");
  TQF_REQUIRE(r->kind != TQF_RNG_PHILOX_ANTITHETIC || n_paths % 2 == 0,
              "antithetic sampling needs an even number of paths");
  if (tg->mode == TQF_TANGENT_VJP) {
    TQF_REQUIRE(tg->grad_paths && tg->grad_params && tg->workspace,
                "VJP needs grad_paths, grad_params and a workspace");
    TQF_REQUIRE(x0_stride_path == 0 || tg->grad_x0,
                "per-path initial states need grad_x0");
  } else {
    TQF_REQUIRE(tg->mode == TQF_TANGENT_JVP, "unknown tangent mode %d",
                tg->mode);
    TQF_REQUIRE(tg->direction && tg->tangent_paths,
                "JVP needs a direction and tangent_paths");
  }
  const hipStream_t s = (hipStream_t)stream;
  if (dtype == TQF_F32)
    return launch_tangent<float>(m, g, r, tg, n_paths, x0, x0_stride_path, s);
  return launch_tangent<double>(m, g, r, tg, n_paths, x0, x0_stride_path, s);
}

// ---- allocation-free payoff pricing ----------------------------------------
// For callers that cannot allocate device memory (code under stream or graph
// capture, tight serving loops): tqf_euler in TQF_OUT_PAYOFF mode with its
// workspace (per-block payoff partials and chunk sums) in a fixed device
// array owned by the library.  Each GPU the library runs on holds its own
// copy for the life of the process; every call on one GPU shares it, so
// calls that overlap on one GPU (two streams, two threads) race.  Nothing
// has to start at zero: every block writes all of its partial sums before
// the reductions read them.
static __device__ __align__(256) double scratch[TQF_NOALLOC_SCRATCH_BYTES / 8];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on the current device, looked up once
// per device (the array's name is only a host-side stand-in).
int scratch_on_current_device(void** out) {
  int dev = -1;
  hipError_t he = hipGetDevice(&dev);
  if (he != hipSuccess) return set_error((int)he, "hipGetDevice");
  TQF_REQUIRE(dev >= 0 && dev < kMaxDevices, "device index %d out of range",
              dev);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[dev] == nullptr) {
    void* p = nullptr;
    he = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (he != hipSuccess || p == nullptr)
      return set_error(he != hipSuccess ? (int)he : (int)hipErrorUnknown,
                       "hipGetSymbolAddress(scratch)");
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return TQF_OK;
}
}  // namespace

extern "C" int tqf_euler_payoff_noalloc(const tqf_model* m, const tqf_grid* g,
                                        const tqf_rng* r,
                                        const tqf_output* out, int64_t n_paths,
                                        const void* x0, int64_t x0_stride_path,
                                        int32_t dtype, void* stream) {
  clear_error();
  TQF_REQUIRE(m && g && r && out, "null descriptor");
  TQF_REQUIRE(out->mode == TQF_OUT_PAYOFF,
              "the allocation-free entry point prices payoffs "
              "(TQF_OUT_PAYOFF)");
  TQF_REQUIRE(out->workspace == nullptr,
              "workspace must be NULL: the library's scratch array is used");
  TQF_REQUIRE(n_paths >= 1 && g->n_slots >= 1 && out->n_strikes >= 1,
              "bad sizes");
  const int64_t need =
      tqf_euler_workspace_size(n_paths, g->n_slots, out->n_strikes);
  TQF_REQUIRE(need <= (int64_t)TQF_NOALLOC_SCRATCH_BYTES,
              "%lld paths x %d dates x %d strikes need %lld bytes of scratch; "
              "the allocation-free entry point holds %lld",
              (long long)n_paths, g->n_slots, out->n_strikes, (long long)need,
              (long long)TQF_NOALLOC_SCRATCH_BYTES);
  void* ws = nullptr;
  int rc = scratch_on_current_device(&ws);
  if (rc) return rc;
  tqf_output o = *out;
  o.workspace = (double*)ws;
  return tqf_euler(m, g, r, &o, n_paths, x0, x0_stride_path, dtype, stream);
}
