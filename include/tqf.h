/*
 * tqf.h - C ABI of the MI355X Monte Carlo path-simulation engine.
 *
 * The reference (google/tf-quant-finance) is pure Python on TensorFlow and
 * has no FFI; its operator API is a handful of Python functions.  Each entry
 * point below replaces the TF ops behind one of those functions and names
 * the reference function (file:line) it serves.  The Python package `tqf`
 * binds these symbols with ctypes and keeps the reference's signatures.
 *
 * Conventions
 *  - Every pointer argument marked "device" is HBM memory owned by the
 *    caller (torch allocates it); the library allocates no device memory.
 *    Work buffers are passed in; *_workspace_size() gives their size.
 *  - `stream` is a hipStream_t (torch.cuda.current_stream().cuda_stream).
 *    Launches are asynchronous on it; nothing synchronises the host.
 *  - Return value: 0 ok; TQF_EINVAL (< 0) for an invalid argument (Python
 *    raises ValueError), TQF_ENOTIMPL for an unsupported option
 *    (NotImplementedError); > 0 is a hipError_t (RuntimeError).
 *    tqf_last_error() returns the calling thread's message.  Nothing aborts.
 *  - dtype: TQF_F32 or TQF_F64 selects the arithmetic of a call; all real
 *    device arrays of that call use it unless stated otherwise.
 */
#ifndef TQF_H_
#define TQF_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TQF_OK 0
#define TQF_EINVAL (-1)
#define TQF_ENOTIMPL (-2)

#define TQF_F32 0
#define TQF_F64 1

/* ---- library ----------------------------------------------------------- */
int tqf_version(void);
const char* tqf_last_error(void);

/* ---- Sobol ------------------------------------------------------------- */

/* Host.  Joe-Kuo direction numbers m[dim][32] (uint32, low 32 bits) from the
 * parsed data file: poly[k] = 2^s + 2a + 1 and minit[k*18 + i] = m_{i+1} of
 * line k.  Replaces tff/math/random_ops/sobol/sobol_impl.py:171-197
 * (_compute_direction_numbers). */
int tqf_sobol_direction_numbers(const int64_t* poly, const int64_t* minit,
                                int32_t n_poly, int32_t dim, uint32_t* m_out);

/* Device.  Integer or real Sobol points, row-major out[i*out_stride + d].
 *   v        device [dim][32] generating columns already aligned to `nd`
 *            digits (column j = m_j << (nd - 1 - j)).
 *   index_i  = first_index + i, or seq_indices[i] when seq_indices != NULL.
 *   shift    device [dim] digital shift (XOR) or NULL.
 *   out_kind 0: uint32 integers; 1: float32 cast(int) * 2^-nd;
 *            2: float64 cast(int) * 2^-nd.  tent != 0 applies the tent map.
 * Replaces sobol_impl.py:39-167 (sobol.sample: first_index = skip + 1,
 * nd = ceil(log2(skip + N + 1))) and tff/math/qmc/digital_net.py:205-419
 * (digital_net_sample / qmc.sobol_sample: first_index = 0). */
int tqf_sobol_points(const uint32_t* v, int32_t dim, int32_t nd, int64_t n,
                     int64_t first_index, const int64_t* seq_indices,
                     const uint32_t* shift, int32_t out_kind, int32_t tent,
                     void* out, int64_t out_stride, void* stream);

/* Device.  lo = XOR of v[d][b] over the set bits b < 8 of the table index
 * t: the per-lane part of a point, used by the fused steppers.  Entry
 * (d, t) is stored at ((d / group) * 256 + t) * group + d % group, so the
 * `group` dimensions a step consumes are adjacent (group divides dim). */
int tqf_sobol_lo_table(const uint32_t* v, int32_t dim, int32_t group,
                       uint32_t* lo, void* stream);

/* Device.  Normal draws from Sobol points, z = sqrt(2) erfinv(2u - 1),
 * out[i*dim + d] in `dtype`.  mode 0: the reference's transform of
 * sobol.sample (u = cast(int) 2^-nd, tff/math/random_ops/
 * multivariate_normal.py:356-424); mode 1: the engine's shifted transform
 * (u = (top25 + 1/2) 2^-25 for float32, evaluated with a piecewise-cubic
 * table within 5.2e-7 absolute / 3.3e-7 relative for |z| >= 1e-3, the same
 * code the fused steppers run; (int + 1/2) 2^-32 for float64). */
int tqf_sobol_normals(const uint32_t* v, int32_t dim, int32_t nd, int64_t n,
                      int64_t first_index, const uint32_t* shift,
                      int32_t mode, int32_t dtype, void* out, void* stream);

/* ---- Halton ------------------------------------------------------------ */

/* Device.  Halton points: element i is the radical inverse of
 * (seq_indices ? seq_indices[i] : first_index + i) + 1 in base bases[d],
 * truncated to max_digits[d] digits; out[i*out_stride + d].  perms (device
 * [max(max_digits)][radix_sum], or NULL) permutes digit j of axis d through
 * perms[j*radix_sum + offsets[d] + digit] and zero_correction (device
 * [dim] or NULL) is added (Owen's randomisation).  out_kind 1/2: float32/64
 * uniforms; 3/4: float32/64 normals sqrt(2) erfinv(2u - 1).  bases,
 * max_digits, offsets: device [dim] int32.  Replaces halton.sample
 * (tff/math/random_ops/halton/halton_impl.py:59-285) and the Halton branch
 * of the quasi normal draws (multivariate_normal.py:393-424). */
int tqf_halton_points(const int32_t* bases, const int32_t* max_digits,
                      const int32_t* offsets, int32_t dim, int64_t n,
                      int64_t first_index, const int64_t* seq_indices,
                      const int32_t* perms, int32_t radix_sum,
                      const double* zero_correction, int32_t out_kind,
                      void* out, int64_t out_stride, void* stream);

/* ---- Philox normals ---------------------------------------------------- */

/* Device.  out[p*n_draws + j] = normal j of path p of the engine's Philox
 * stream (Philox4x32-10 counter (p_lo, j/4, stream_id, p_hi), key = seed,
 * Box-Muller pairs).  antithetic != 0: rows p and p + n_paths/2 are (z, -z).
 * Replaces tf.random.normal / stateless_normal at
 * tff/math/random_ops/multivariate_normal.py:248-311. */
int tqf_philox_normals(uint64_t seed, uint32_t stream_id, int64_t n_paths,
                       int64_t n_draws, int32_t antithetic, int32_t dtype,
                       void* out, void* stream);

/* ---- Fused Euler-Maruyama steppers ------------------------------------- */

#define TQF_MODEL_GBM 0       /* dX = mu X dt + sigma X dW (Euler on S)     */
#define TQF_MODEL_CONST 1     /* dX = a dt + b dW, constant a, b (log GBM)  */
#define TQF_MODEL_HESTON_FT 2 /* (log S, v), full truncation, drift r      */
#define TQF_MODEL_LOCAL_VOL 3 /* log S, sigma(t, S) from a bilinear table   */
#define TQF_MODEL_HESTON_QE 4 /* (log S, v), Andersen QE (HestonModel)      */
#define TQF_MODEL_HULL_WHITE 5 /* short rate r = x + f(0, t), exact Gaussian
                                  step of x (HullWhiteModel1F)              */

#define TQF_RNG_DRAWS 0            /* injected normals                      */
#define TQF_RNG_PHILOX 1
#define TQF_RNG_PHILOX_ANTITHETIC 2
#define TQF_RNG_SOBOL 3            /* reference transform (mode 0 above)    */
#define TQF_RNG_SOBOL_SHIFTED 4    /* engine transform (mode 1 above)       */

#define TQF_OUT_PATH_MAJOR 0  /* out[p][slot][d]  (reference layout)        */
#define TQF_OUT_STEP_MAJOR 1  /* out[slot][d][p]  (coalesced path store)    */
#define TQF_OUT_PAYOFF 2      /* European calls reduced in the kernel       */
#define TQF_OUT_PAYOFF_PARTIALS 3 /* as PAYOFF, per-block sums only: the
                                     caller runs tqf_payoff_finalize */

typedef struct tqf_model {
  int32_t kind;
  int32_t dim;
  double p[8];  /* GBM: mu, sigma.  CONST: a, b.
                   HESTON_FT: r, kappa, theta, xi, rho, variant
                   (variant 0: full truncation; 1: HestonModel's own
                   coefficients, raw v drift and sqrt|v| vol).
                   LOCAL_VOL: p[0] drift rate (r - q); p[1] != 0 selects
                   edge mode: each table row is [value below column 1,
                   nx - 2 interior columns, value above], lv_x0 is the
                   log-spot of column 1, and outside the interior nodes
                   the edge value is used as it is.
                   HESTON_QE: theta (per-step values in step_coef). */
  const void* lv_vol;   /* LOCAL_VOL: device [nt][nx] sigma table (bilinear
                           in t and ln S, flat outside; see p[1])          */
  const void* lv_t;     /* device [nt] table times (increasing)            */
  double lv_x0;         /* log-spot of table column 0                      */
  double lv_dx;         /* uniform log-spot spacing                        */
  int32_t lv_nt;
  int32_t lv_nx;
  const void* step_coef;  /* HESTON_QE: device [n_steps][12] per-step
                             constants: 0 e = exp(-kappa dt), 1 c1, 2 c2
                             (s^2 = v c1 + c2), 3..7 k0..k4 of the log-spot
                             update, 8 theta, 9 update flag (0 when
                             dt <= tolerance), 10-11 unused.
                             HULL_WHITE: device [n_steps][8] = {e^{-a dt},
                             conditional mean, conditional standard
                             deviation, f(0, t_{i+1}), 1, slot[i] (bits as
                             for HESTON_FT), 0, 0}: x' = e x + m + s z and
                             the recorded rate is x + f(0, t); p[0] =
                             f(0, t_0) for a record at the first time.
                             HESTON_FT: device [n_steps][8] = {-dt/2,
                             kappa theta dt, -kappa dt, r t_{i+1},
                             sqrt(dt), slot[i] (the grid's record slot,
                             stored as the integer's bit pattern), 0, 0}
                             with r = p[0]; the kernel carries log S - r t
                             and adds r t back where it records or
                             prices. */
} tqf_model;

typedef struct tqf_grid {
  const void* dt;       /* device [n_steps] times[i+1] - times[i]          */
  const void* sqrt_dt;  /* device [n_steps]                                */
  const void* t_next;   /* device [n_steps] times[i+1]                     */
  const int32_t* slot;  /* device [n_steps] slot written after step i, -1  */
  int32_t n_steps;      /* steps the loop executes                         */
  int32_t n_slots;      /* requested times k                               */
  int32_t init_slot;    /* slot holding the initial state, or -1           */
  int32_t n_rng_steps;  /* steps the draws are laid out for (>= n_steps)   */
} tqf_grid;

typedef struct tqf_rng {
  int32_t kind;
  uint32_t stream_id;
  uint64_t seed;
  const void* draws;           /* DRAWS: element (p, i, k) at
                                  draws[p*stride_path + i*stride_step + k]  */
  int64_t stride_path;
  int64_t stride_step;
  const uint32_t* sobol_v;     /* SOBOL*: device [ndims][32]                */
  const uint32_t* sobol_lo;    /* SOBOL*: device [ndims/D][256][D] (D =
                                  normals per step, tqf_sobol_lo_table group) */
  const uint32_t* sobol_shift; /* SOBOL*: device [ndims] or NULL            */
  int64_t sobol_first_index;   /* sequence index of path 0                  */
  int32_t sobol_nd;
  int32_t sobol_ndims;
} tqf_rng;

typedef struct tqf_output {
  int32_t mode;
  int32_t n_strikes;         /* PAYOFF                                      */
  void* paths;               /* PATH_MAJOR / STEP_MAJOR, aligned to one
                                record (dim values of the dtype)           */
  const double* strikes;     /* PAYOFF: device [n_strikes] (calls)          */
  const double* discount;    /* PAYOFF: device [n_slots] discount factors   */
  double* workspace;         /* PAYOFF: tqf_euler_workspace_size() bytes    */
  double* prices;            /* PAYOFF: device [n_slots][n_strikes]         */
  double* stderrs;           /* PAYOFF: device [n_slots][n_strikes]         */
  double* sums;              /* PAYOFF: device [n_slots][n_strikes][2]
                                (sum, sum of squares) of raw payoffs, or NULL */
} tqf_output;

/* Workspace bytes for TQF_OUT_PAYOFF and TQF_OUT_PAYOFF_PARTIALS. */
int64_t tqf_euler_workspace_size(int64_t n_paths, int32_t n_slots,
                                 int32_t n_strikes);

/* Device.  Second stage of TQF_OUT_PAYOFF, run by tqf_euler itself in that
 * mode: reduces the per-block payoff sums a TQF_OUT_PAYOFF_PARTIALS launch
 * left in `workspace` to discounted prices, standard errors and (if `sums`
 * is not NULL) raw sums.  Replaces the reduce_mean over samples that follows
 * euler_sampling.sample in the reference's pricing notebooks (tff/models/
 * euler_sampling.py:27-195 returns the paths; the mean is the caller's). */
int tqf_payoff_finalize(double* workspace, int64_t n_paths, int32_t n_slots,
                        int32_t n_strikes, const double* discount,
                        double* prices, double* stderrs, double* sums,
                        void* stream);

/* Device.  Fused Euler-Maruyama: draws, steps, records and (optionally)
 * payoffs in one kernel with the state in registers.  x0: device
 * [n_paths][dim] (x0_stride_path = dim) or [dim] (x0_stride_path = 0).
 * Replaces tff/models/euler_sampling.py:335-537 (_sample, _while_loop,
 * _euler_step) for the known coefficient functions of
 * GeometricBrownianMotion (univariate...py:127-153), HestonModel
 * (heston_model.py:143-173, with truncation and drift r) and
 * LocalVolatilityModel (local_volatility_model.py:461-476). */
int tqf_euler(const tqf_model* model, const tqf_grid* grid,
              const tqf_rng* rng, const tqf_output* out, int64_t n_paths,
              const void* x0, int64_t x0_stride_path, int32_t dtype,
              void* stream);

/* Allocation-free variant of tqf_euler in TQF_OUT_PAYOFF mode.  No
 * counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under stream or graph capture, tight serving loops).
 * Same arguments and results as tqf_euler with out->mode == TQF_OUT_PAYOFF,
 * except that out->workspace must be NULL: the payoff partials live in a
 * fixed device array of TQF_NOALLOC_SCRATCH_BYTES that the library holds
 * on each GPU it runs on until the process ends.  Calls whose
 * tqf_euler_workspace_size(n_paths, n_slots, n_strikes) exceeds it return
 * TQF_EINVAL.  Every call on one GPU shares that array: calls that overlap
 * on one GPU (two streams, two threads) race.  Makes no hipMalloc, hipFree
 * or other allocating call; the stream must belong to the current device. */
#define TQF_NOALLOC_SCRATCH_BYTES (32 << 20)
int tqf_euler_payoff_noalloc(const tqf_model* model, const tqf_grid* grid,
                             const tqf_rng* rng, const tqf_output* out,
                             int64_t n_paths, const void* x0,
                             int64_t x0_stride_path, int32_t dtype,
                             void* stream);

/* ---- Forward tangents of the fused Euler paths -------------------------- */

#define TQF_TANGENT_VJP 0  /* gradient of sum(g * paths) w.r.t. the inputs */
#define TQF_TANGENT_JVP 1  /* tangent paths in a direction of the inputs   */

/* Inputs of a model, in order: the initial state x0[0..dim), then its
 * parameters -- GBM: mu, sigma; CONST: a, b; HESTON_FT (either variant):
 * r, kappa, theta, xi, rho.  The paths are those tqf_euler writes in
 * TQF_OUT_PATH_MAJOR layout [N][k][dim] for the same descriptors. */
typedef struct tqf_tangent {
  int32_t mode;               /* TQF_TANGENT_*                              */
  int32_t pad0;
  const void* grad_paths;     /* VJP: device [N][k][dim] incoming gradient  */
  double* grad_params;        /* VJP: device [dim + n_par] sums over paths;
                                 the x0 entries are 0 with per-path x0      */
  void* grad_x0;              /* VJP with x0_stride_path = dim: device
                                 [N][dim] per-path x0 gradient              */
  const double* direction;    /* JVP: device [dim + n_par]                  */
  const void* dir_x0;         /* JVP: device [N][dim] per-path x0 direction,
                                 or NULL to use direction[0..dim)           */
  void* tangent_paths;        /* JVP: device [N][k][dim] output             */
  void* workspace;            /* VJP: tqf_euler_tangent_workspace_size()    */
} tqf_tangent;

int64_t tqf_euler_tangent_workspace_size(int64_t n_paths);

/* Device.  Re-simulates the paths of tqf_euler (same normals, same update,
 * bit for bit) carrying the forward tangents of the state with respect to
 * every input, in registers: memory does not grow with the number of steps.
 * VJP: grad_params[q] = sum_{p, slot, c} g[p][slot][c] dX[p][slot][c] /
 * dinput_q (fixed-order reduction).  JVP: tangent_paths = sum_q
 * direction_q dX / dinput_q.  rng kinds: DRAWS, PHILOX, PHILOX_ANTITHETIC
 * (Sobol draws are passed as DRAWS).  Replaces the forward-Jacobian loop of
 * euler_sampling.sample with watch_params (tff/models/euler_sampling.py:
 * 467-510, tff/math/custom_loops.py:20-214) and the tape gradients taken
 * through it (tff/math/gradient.py:22-235). */
int tqf_euler_tangent(const tqf_model* model, const tqf_grid* grid,
                      const tqf_rng* rng, const tqf_tangent* tangent,
                      int64_t n_paths, const void* x0, int64_t x0_stride_path,
                      int32_t dtype, void* stream);

/* ---- Exact geometric Brownian motion ----------------------------------- */

typedef struct tqf_gbm {
  int32_t dim;              /* assets, 1..16                                 */
  const void* drift;        /* device [dim] mu_i - sigma_i^2 / 2             */
  const void* vol;          /* device [dim] sigma_i                          */
  const void* chol;         /* device [dim][dim] lower Cholesky factor of the
                               correlation matrix, or NULL (independent)     */
  const void* s0;           /* device [dim] or [n_paths][dim]                */
  int64_t s0_stride_path;   /* 0 or dim                                      */
} tqf_gbm;

/* Device.  S_i(t) = S0_i exp(sum (mu_i - sigma_i^2/2) dt + sqrt(dt) sigma_i
 * (L z)_i) over the grid, recorded at the grid's slots into `out`
 * (TQF_OUT_PATH_MAJOR [N][k][dim] or TQF_OUT_STEP_MAJOR [k][dim][N]).
 * Replaces GeometricBrownianMotion._sample_paths (tff/models/
 * geometric_brownian_motion/univariate_geometric_brownian_motion.py:261-317)
 * and MultivariateGeometricBrownianMotion._sample_paths
 * (multivariate_geometric_brownian_motion.py:229-282). */
int tqf_gbm_exact(const tqf_gbm* gbm, const tqf_grid* grid, const tqf_rng* rng,
                  int32_t out_mode, void* out, int64_t n_paths, int32_t dtype,
                  void* stream);

/* ---- Longstaff-Schwartz ------------------------------------------------- */

#define TQF_PAYOFF_BASKET_PUT 0  /* relu(K - mean_j S_j) (payoff_utils.py:27) */
#define TQF_PAYOFF_MAX_CALL 1    /* relu(max_j S_j - K)                        */
#define TQF_PAYOFF_TABLE 2       /* exercise values given per (time, path):
                                    payoff_table[t_idx][p]                    */

typedef struct tqf_lsm {
  const double* paths;      /* device float64, [n_paths][n_times][dim] or the
                               layout the strides below give               */
  int64_t n_paths;
  int32_t n_times;
  int32_t dim;              /* assets, 1..8                                  */
  int32_t payoff;           /* TQF_PAYOFF_*                                  */
  int32_t degree;           /* make_polynomial_basis(degree),
                               (degree + 1) * dim <= 64                       */
  int32_t n_basis;          /* (degree + 1)^dim, at most 255                 */
  int32_t basis_log;        /* 0: basis on the paths' values; 1: on their
                               logarithms (the payoff on the values); 2:
                               the paths hold log prices: basis on the
                               stored values, payoff on their exp          */
  const uint8_t* exponents; /* device [n_basis][dim] exponent table in the
                               reference's meshgrid order (lsm.py:116-118);
                               every entry <= degree and n_basis <=
                               (degree + 1)^dim (tqf_lsm_price checks both
                               and returns TQF_EINVAL)                       */
  double strike;
  int64_t num_calibration;  /* rows used for the regression (0 = all)       */
  double rcond;             /* pivot cutoff relative to the largest scaled
                               diagonal of X^T X                            */
  const double* payoff_table;   /* TQF_PAYOFF_TABLE: device
                                   [n_times][n_paths] exercise values (>= 0;
                                   in the money iff > 0)                     */
  const double* path_discounts; /* optional device [n_ex + 1][n_paths]: the
                                   padded discount factors of every path
                                   (row 0 = 1); when set they replace df in
                                   tqf_lsm_price                             */
  int64_t stride_path;  /* element strides of `paths`: value (p, t, j) at   */
  int64_t stride_time;  /* paths[p stride_path + t stride_time + j          */
  int64_t stride_asset; /* stride_asset]; all 0 = the reference layout
                           [n_paths][n_times][dim].  The step-major layout
                           [n_times][dim][n_paths] (stride_path = 1) makes
                           every per-date read coalesced.                   */
} tqf_lsm;

int64_t tqf_lsm_workspace_size(int64_t n_paths, int32_t n_basis);

/* Device.  Longstaff-Schwartz price of one payoff: backward induction over
 * the exercise indices ex_idx[n_ex] (time indices into the paths), with the
 * padded discount factors df[n_ex + 1] (df[0] = 1, df[i] = discount of
 * exercise date i-1; per path when lsm->path_discounts is set, the
 * reference's rank-3 discount factors), the ITM-masked float64 normal
 * equations built in the
 * kernel from the centred tensor-product basis, a column-scaled Cholesky
 * solve with pivot truncation, and the exercise update.  cash/vals are
 * device [n_paths] state buffers; result[2] = (price, standard error) over
 * the rows after num_calibration; ranks[n_ex - 1] (optional) receives the
 * rank kept at each regression.  Replaces least_square_mc and its loop body
 * (tff/models/longstaff_schwartz/lsm.py:128-436). */
int tqf_lsm_price(const tqf_lsm* lsm, const int32_t* ex_idx, int32_t n_ex,
                  const double* df, void* workspace, double* cash,
                  double* vals, double* result, int32_t* ranks, void* stream);

/* Per-date building blocks of the same backward induction, for callers that
 * supply their own continuation estimate (the MLP model below):
 *   tqf_lsm_terminal     cash = payoff at t_idx, vals = 0
 *   tqf_lsm_date_inputs  exv = payoff, y = df_ratio (cash + vals), centring
 *                        moments (tqf_lsm_moments_size bytes), and MLP rows
 *                        x[n][8] = (standardised S_1..S_dim, exv / K, 0...)
 *                        (x 16-byte aligned),
 *                        tgt = y / K, w = 1{exv > 0}   (dim <= 5)
 *   tqf_lsm_update_cont  exercise iff exv > K max(cont, 0) (lsm.py:381-400)
 *   tqf_lsm_final        (mean, stderr) of df0 (cash + vals) over [first, n)
 */
int tqf_lsm_terminal(const tqf_lsm* lsm, int32_t t_idx, double* cash,
                     double* vals, void* stream);
int64_t tqf_lsm_moments_size(int64_t n_paths);
int tqf_lsm_date_inputs(const tqf_lsm* lsm, int32_t t_idx, double df_ratio,
                        const double* cash, const double* vals, double* exv,
                        double* y, double* moments, float* x, float* tgt,
                        float* w, void* stream);
int tqf_lsm_update_cont(int64_t n_paths, double strike, double df_ratio,
                        const double* exv, const float* cont, double* cash,
                        double* vals, void* stream);
int tqf_lsm_final(const double* cash, const double* vals, double df0,
                  int64_t first, int64_t n, double* workspace, double* result,
                  void* stream);

/* ---- MLP continuation value (6 -> 64 -> 64 -> 1, SiLU) ------------------ */

/* Parameters: float32 [4673] = W1[6][64] | b1[64] | W2[64][64] | b2[64] |
 * w3[64] | b3.  Inputs x: [n][8] float32 rows (6 used).  The 64x64 layer
 * uses bf16 MFMA with float32 accumulation.  No counterpart in the
 * reference: this model stands in for the polynomial regression of
 * least_square_mc (lsm.py:328-383), which stays the fallback. */
int tqf_mlp_forward(const float* params, const float* x, int64_t n,
                    float* out, void* stream);
int64_t tqf_mlp_workspace_size(int32_t n_blocks);
/* One Adam step on L = sum_s w_s (f(x_s) - t_s)^2 * inv_norm over rows
 * [first, first + n).  lr = 0 computes the gradient only (grad_out
 * [4673], loss_out [1] optional, device). */
int tqf_mlp_train_step(float* params, float* adam_m, float* adam_v,
                       const float* x, const float* targets,
                       const float* weights, int64_t first, int64_t n,
                       float inv_norm, float lr, int32_t step,
                       int32_t n_blocks, void* workspace, float* grad_out,
                       float* loss_out, void* stream);

/* ---- Batched heterogeneous pricing with adjoint Greeks ----------------- */

#define TQF_CONTRACT_GBM 0     /* log-Euler GBM: Greeks delta, vega, rho     */
#define TQF_CONTRACT_HESTON 1  /* full-truncation Heston: delta, d/dv0, rho,
                                  d/dkappa, d/dtheta, d/dxi, d/drho          */

#define TQF_GREEKS_ADJOINT 0  /* Heston: reverse sweep over checkpoints     */
#define TQF_GREEKS_TANGENT 1  /* Heston: five tangents in the forward pass  */

/* tqf_batch_price flags */
#define TQF_BATCH_COUNTER 1       /* items from an atomic counter (else a
                                     static round robin)                    */
#define TQF_BATCH_TANGENT_ONLY 2  /* no reverse sweep in the launch: Heston
                                     contracts get forward-tangent Greeks
                                     whatever their greeks_mode (the same
                                     pathwise estimator); fewer registers,
                                     more waves per SIMD                    */

typedef struct tqf_contract {
  int32_t kind;
  int32_t n_steps;      /* uniform grid dt = maturity / n_steps             */
  uint32_t stream_id;   /* Philox sub-stream of the contract's paths        */
  int32_t greeks_mode;  /* TQF_GREEKS_* (Heston contracts; same estimator)  */
  int64_t n_paths;
  double s0, strike, maturity, rate;
  double sigma;         /* GBM                                              */
  double v0, kappa, theta, xi, rho;  /* Heston                              */
} tqf_contract;

typedef struct tqf_work_item {
  int32_t contract;
  int32_t pad0;
  int64_t first_path;   /* Philox path index of the item's first path       */
  int64_t n_paths;
} tqf_work_item;

int64_t tqf_batch_workspace_size(int32_t n_items, int32_t n_blocks,
                                 int32_t max_steps);

/* Device.  Prices European calls for n_contracts contracts in one
 * persistent launch over n_items work items (device arrays; the items of a
 * contract are contiguous rows; `order` (device [n_items], or NULL for row
 * order) is the processing order, longest item first), with
 * pathwise Greeks (Heston: checkpoints every 16 steps and a reverse sweep,
 * or forward tangents, per contract).  The workspace holds checkpoints for
 * max_steps steps (tqf_batch_workspace_size): a Heston contract with
 * adjoint Greeks and n_steps > max_steps gets the forward tangents (the
 * same estimator) instead of the reverse sweep.  Philox stream: key = seed, sub-stream = the contract's
 * stream_id, path index = first_path + offset.  out: device [n_contracts][16] doubles:
 * 8 means (price, Greeks) then their 8 standard errors.  The launch records
 * no memset, so it can be captured in a hipGraph.  Replaces sampling with
 * euler_sampling.sample per contract (tff/models/euler_sampling.py:27-537)
 * plus tf.GradientTape / fwd_gradient Greeks (tff/math/gradient.py:22-235,
 * tff/math/custom_loops.py:20-214). */
int tqf_batch_price(const tqf_contract* contracts, int32_t n_contracts,
                    const tqf_work_item* items, const int32_t* order,
                    int32_t n_items,
                    const int32_t* contract_first_item,
                    const int32_t* contract_n_items, uint64_t seed,
                    int32_t n_blocks, int32_t max_steps, int32_t flags,
                    void* workspace, double* out, void* stream);

/* ---- Basket call with adjoint Greeks ------------------------------------ */

int64_t tqf_basket_workspace_size(int32_t dim, int32_t n_blocks);

/* Device.  Arithmetic-average basket call e^{-rT} (sum_i w_i S_i(T) - K)^+
 * on correlated GBM (exact scheme over n_steps uniform steps, Philox
 * normals: key = seed, sub-stream stream_id, draw j = step * dim + asset),
 * with pathwise adjoint Greeks when greeks != 0.
 * params: device float32 [s0[dim], sigma[dim], weight[dim], chol[dim][dim]]
 * (dim a multiple of 4, <= 16; pad assets with zero weight).
 * out: device [2 * (2 + 2 dim)] doubles: means of (price, rho, delta[dim],
 * vega[dim]) then their standard errors.  Replaces
 * MultivariateGeometricBrownianMotion.sample_paths (tff/models/
 * geometric_brownian_motion/multivariate_geometric_brownian_motion.py:
 * 183-282) plus tf.GradientTape Greeks of the payoff mean. */
int tqf_basket_call(const float* params, int32_t dim, int32_t n_steps,
                    double maturity, double rate, double strike,
                    int64_t n_paths, uint64_t seed, uint32_t stream_id,
                    int32_t greeks, int32_t n_blocks, void* workspace,
                    double* out, void* stream);

#ifdef __cplusplus
}
#endif

#endif  /* TQF_H_ */
