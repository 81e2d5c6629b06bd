// Batched heterogeneous pricing with pathwise adjoint Greeks.
//
// A contract is a European call under GBM (log-Euler, exact in law) or
// Heston (full-truncation log-Euler), with its own parameters, path count,
// uniform step count and maturity.  The host splits every contract into
// work items of `tile` paths (one thread per path, several paths per thread
// when tile > block) and orders them longest first; a persistent grid pulls
// items from an atomic counter, so a 640x spread in work per contract does
// not leave CUs idle.  Each item writes its sums (price, Greeks and their
// squares) to its own row; a second kernel adds the rows of each contract
// in a fixed order.
//
// Greeks are pathwise.  GBM's adjoint does not depend on the path, so it is
// closed form at maturity.  Heston's does: in reverse mode (the default) the
// forward pass stores the state every kSeg steps (about sqrt(steps)), and
// the reverse sweep re-runs each segment from its checkpoint with the same
// Philox counters, keeps the segment's states and normals in registers, and
// walks it backwards; in tangent mode (a contract's greeks_mode) five
// tangents ride along the forward pass instead.  Truncation kinks use the
// subgradient 0 at v = 0.
//
// Forward step (euler_sampling.py:513-537 with the truncated Heston
// coefficients of heston_model.py:143-173, drift r):
//   x' = x + dt (r - v+/2) + sqrt(v+) dw0
//   v' = v + dt k (th - v+) + xi rho sqrt(v+) dw0 + xi sqrt(1-rho^2) sqrt(v+) dw1
#include "common.h"

using namespace tqf;

namespace {

constexpr int kBlockB = 256;
constexpr int kSeg = 16;  // steps per checkpoint segment

__device__ __forceinline__ void normals4(uint32_t p_lo, uint32_t p_hi,
                                         uint32_t blk, uint32_t stream,
                                         uint32_t k0, uint32_t k1, float* z) {
  const u32x4 w = philox4x32_10(u32x4{p_lo, blk, stream, p_hi}, k0, k1);
  box_muller(w.x, w.y, z[0], z[1]);
  box_muller(w.z, w.w, z[2], z[3]);
}

struct Contract {
  int kind, n_steps, greeks_mode;
  uint32_t stream;
  float s0, strike, t, r;
  float sigma;                         // GBM
  float v0, kappa, theta, xi, rho;     // Heston
  int64_t n_paths;
};

__device__ __forceinline__ void load_contract(const tqf_contract* __restrict__ c,
                                              Contract& k) {
  k.kind = c->kind;
  k.n_steps = c->n_steps;
  k.greeks_mode = c->greeks_mode;
  k.stream = c->stream_id;
  k.n_paths = c->n_paths;
  k.s0 = (float)c->s0;
  k.strike = (float)c->strike;
  k.t = (float)c->maturity;
  k.r = (float)c->rate;
  k.sigma = (float)c->sigma;
  k.v0 = (float)c->v0;
  k.kappa = (float)c->kappa;
  k.theta = (float)c->theta;
  k.xi = (float)c->xi;
  k.rho = (float)c->rho;
}

// One GBM path: price and (delta, vega, rho).
__device__ __forceinline__ void gbm_path(const Contract& c, int64_t p,
                                         uint32_t stream, uint32_t k0,
                                         uint32_t k1, float* g) {
  const float dt = c.t / (float)c.n_steps;
  const float sdt = sqrtf(dt);
  const float drift = (c.r - 0.5f * c.sigma * c.sigma) * dt;
  float x = 0.f, w = 0.f;
  float z[4];
  // four steps per Philox block (block i / 4 serves steps i .. i + 3)
  int i = 0;
  for (; i + 4 <= c.n_steps; i += 4) {
    normals4((uint32_t)p, (uint32_t)(p >> 32), (uint32_t)(i >> 2), stream, k0,
             k1, z);
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const float dw = z[u] * sdt;
      x = (x + drift) + c.sigma * dw;
      w += dw;
    }
  }
  if (i < c.n_steps) {
    normals4((uint32_t)p, (uint32_t)(p >> 32), (uint32_t)(i >> 2), stream, k0,
             k1, z);
#pragma unroll
    for (int u = 0; u < 3; ++u) {
      if (i + u < c.n_steps) {
        const float dw = z[u] * sdt;
        x = (x + drift) + c.sigma * dw;
        w += dw;
      }
    }
  }
  const float s = c.s0 * __expf(x);
  const float disc = __expf(-c.r * c.t);
  const float itm = s > c.strike ? 1.f : 0.f;
  const float pay = disc * itm * (s - c.strike);
  const float sb = disc * itm;  // dP/dS_T
  g[0] = pay;
  g[1] = sb * s / c.s0;                                   // delta
  g[2] = sb * s * (w - c.sigma * c.t);                    // vega
  g[3] = -c.t * pay + sb * s * c.t;                       // rho
  g[4] = g[5] = g[6] = g[7] = 0.f;
}

// Forward Heston step; returns the new state.
__device__ __forceinline__ void heston_step(const Contract& c, float xr,
                                            float xs, float dt, float sdt,
                                            float z0, float z1, float& x,
                                            float& v) {
  const float vp = v > 0.f ? v : 0.f;
  const float sq = __builtin_amdgcn_sqrtf(vp);
  const float dw0 = z0 * sdt, dw1 = z1 * sdt;
  const float xn = (x + dt * (c.r - 0.5f * vp)) + sq * dw0;
  const float vn = (v + dt * (c.kappa * (c.theta - vp))) +
                   ((xr * sq) * dw0 + (xs * sq) * dw1);
  x = xn;
  v = vn;
}

// One Heston path with the checkpointed reverse sweep.  ck: this thread's
// checkpoint column (stride `cstride` floats between segments).
__device__ void heston_path(const Contract& c, int64_t p, uint32_t stream,
                            uint32_t k0, uint32_t k1, float* __restrict__ ck,
                            int64_t cstride, float* g) {
  const int n = c.n_steps;
  const float dt = c.t / (float)n;
  const float sdt = sqrtf(dt);
  const float s1r = sqrtf(1.f - c.rho * c.rho);
  const float xr = c.xi * c.rho, xs = c.xi * s1r;
  const uint32_t plo = (uint32_t)p, phi = (uint32_t)(p >> 32);
  float x = __logf(c.s0), v = c.v0;
  float z[4];
  const int nseg = (n + kSeg - 1) / kSeg;
  for (int i = 0; i < n; ++i) {
    if (i % kSeg == 0) {
      ck[(int64_t)(i / kSeg) * 2 * cstride] = x;
      ck[(int64_t)(i / kSeg) * 2 * cstride + cstride] = v;
    }
    if ((i & 1) == 0)
      normals4(plo, phi, (uint32_t)(i >> 1), stream, k0, k1, z);
    const float z0 = (i & 1) ? z[2] : z[0];
    const float z1 = (i & 1) ? z[3] : z[1];
    heston_step(c, xr, xs, dt, sdt, z0, z1, x, v);
  }
  const float s = __expf(x);
  const float disc = __expf(-c.r * c.t);
  const float itm = s > c.strike ? 1.f : 0.f;
  const float pay = disc * itm * (s - c.strike);
  // adjoints
  float xb = disc * itm * s, vb = 0.f;
  float rb = -c.t * pay, kb = 0.f, tb = 0.f, xib = 0.f, rhob = 0.f;
  const float drho_s = -c.rho / s1r;  // d sqrt(1-rho^2) / d rho
  for (int sgi = nseg - 1; sgi >= 0; --sgi) {
    const int i0 = sgi * kSeg;
    float xs_[kSeg], vs_[kSeg], z0s[kSeg], z1s[kSeg];
    float cx = ck[(int64_t)sgi * 2 * cstride];
    float cv = ck[(int64_t)sgi * 2 * cstride + cstride];
#pragma unroll
    for (int u = 0; u < kSeg; ++u) {
      const int i = i0 + u;
      if (i < n) {
        if ((u & 1) == 0)
          normals4(plo, phi, (uint32_t)(i >> 1), stream, k0, k1, z);
        z0s[u] = (u & 1) ? z[2] : z[0];
        z1s[u] = (u & 1) ? z[3] : z[1];
        xs_[u] = cx;
        vs_[u] = cv;
        heston_step(c, xr, xs, dt, sdt, z0s[u], z1s[u], cx, cv);
      }
    }
#pragma unroll
    for (int u = kSeg - 1; u >= 0; --u) {
      const int i = i0 + u;
      if (i < n) {
        const float v0 = vs_[u];
        const float vp = v0 > 0.f ? v0 : 0.f;
        const float sq = __builtin_amdgcn_sqrtf(vp);
        const float dw0 = z0s[u] * sdt, dw1 = z1s[u] * sdt;
        const float sqb = xb * dw0 + vb * (xr * dw0 + xs * dw1);
        float vpb = -0.5f * dt * xb - dt * c.kappa * vb;
        // d sqrt(v+) / dv+ = 1 / (2 sqrt(v+)) by the hardware reciprocal
        vpb += sq > 0.f ? sqb * 0.5f * __builtin_amdgcn_rcpf(sq) : 0.f;
        rb += xb * dt;
        kb += vb * dt * (c.theta - vp);
        tb += vb * dt * c.kappa;
        xib += vb * (c.rho * sq * dw0 + s1r * sq * dw1);
        rhob += vb * (c.xi * sq * dw0 + c.xi * drho_s * sq * dw1);
        vb = vb + (v0 > 0.f ? vpb : 0.f);
        // xb unchanged: dx'/dx = 1
      }
    }
  }
  g[0] = pay;
  g[1] = xb / c.s0;  // delta
  g[2] = vb;         // d/dv0
  g[3] = rb;         // rho (rate)
  g[4] = kb;
  g[5] = tb;
  g[6] = xib;
  g[7] = rhob;
}

// One Heston path with forward-mode (tangent) Greeks: the tangents of
// (x, v) for v0, kappa, theta, xi and rho ride along the forward pass, so
// there is no checkpoint, no second pass and no second draw of the normals
// (which dominate a step's cost).  Same pathwise estimator as the reverse
// sweep above, with the same subgradient 0 at v <= 0: the price depends on
// x_T only, so each Greek is x-bar ("xb" below) times the x-tangent, and
// delta and rho are closed form as in the reverse sweep.
//   v+' = v > 0 ? v' : 0,  sqrt(v+)' = v+' / (2 sqrt(v+))
//   x'  <- x' - dt/2 v+' + sqrt(v+)' dw0
//   v'  <- v' - dt kappa v+' + sqrt(v+)' (xi rho dw0 + xi s dw1) + dv'/dp
__device__ void heston_path_tangent(const Contract& c, int64_t p,
                                    uint32_t stream, uint32_t k0,
                                    uint32_t k1, float* g) {
  const int n = c.n_steps;
  const float dt = c.t / (float)n;
  const float sdt = sqrtf(dt);
  const float s1r = sqrtf(1.f - c.rho * c.rho);
  const float xr = c.xi * c.rho, xs = c.xi * s1r;
  const float drho_s = -c.rho / s1r;  // d sqrt(1-rho^2) / d rho
  const float mdt2 = -0.5f * dt, mdtk = -dt * c.kappa, dtk = dt * c.kappa;
  const uint32_t plo = (uint32_t)p, phi = (uint32_t)(p >> 32);
  float x = __logf(c.s0), v = c.v0;
  // tangents of (x, v) for v0, kappa, theta, xi, rho
  float tx[5] = {0.f, 0.f, 0.f, 0.f, 0.f};
  float tv[5] = {1.f, 0.f, 0.f, 0.f, 0.f};
  float z[4];
  for (int i = 0; i < n; ++i) {
    if ((i & 1) == 0)
      normals4(plo, phi, (uint32_t)(i >> 1), stream, k0, k1, z);
    const float z0 = (i & 1) ? z[2] : z[0];
    const float z1 = (i & 1) ? z[3] : z[1];
    const bool pos = v > 0.f;
    const float vp = pos ? v : 0.f;
    const float sq = __builtin_amdgcn_sqrtf(vp);
    const float dw0 = z0 * sdt, dw1 = z1 * sdt;
    const float cc = xr * dw0 + xs * dw1;
    // d sqrt(v+) / dv+ by the hardware reciprocal (one instruction in
    // place of a correctly rounded division by sq)
    const float hs = sq > 0.f ? 0.5f * __builtin_amdgcn_rcpf(sq) : 0.f;
#pragma unroll
    for (int q = 0; q < 5; ++q) {
      const float tvp = pos ? tv[q] : 0.f;
      const float tsq = tvp * hs;
      tx[q] = (tx[q] + mdt2 * tvp) + tsq * dw0;
      tv[q] = (tv[q] + mdtk * tvp) + tsq * cc;
    }
    tv[1] += dt * (c.theta - vp);                    // kappa
    tv[2] += dtk;                                    // theta
    tv[3] += sq * (c.rho * dw0 + s1r * dw1);         // xi
    tv[4] += c.xi * sq * (dw0 + drho_s * dw1);       // rho
    // the state update of the reverse-mode path, so both modes price the
    // same paths bit for bit
    heston_step(c, xr, xs, dt, sdt, z0, z1, x, v);
  }
  const float s = __expf(x);
  const float disc = __expf(-c.r * c.t);
  const float itm = s > c.strike ? 1.f : 0.f;
  const float pay = disc * itm * (s - c.strike);
  const float xb = disc * itm * s;
  g[0] = pay;
  g[1] = xb / c.s0;             // delta
  g[2] = xb * tx[0];            // d/dv0
  g[3] = -c.t * pay + xb * c.t; // rho (rate): x_T carries r T
  g[4] = xb * tx[1];
  g[5] = xb * tx[2];
  g[6] = xb * tx[3];
  g[7] = xb * tx[4];
}

// Persistent kernel: items from `counter`, sums per item into `item_sums`
// [n_items][2 * kStats].
constexpr int kStats = 8;

// kAdjoint false: the kernel has no reverse sweep (a Heston contract that
// asks for adjoint Greeks gets the forward tangents, the same pathwise
// estimator), so it needs far fewer registers and runs more waves per SIMD.
template <bool kAdjoint>
__global__ void __launch_bounds__(kBlockB)
batch_kernel(const tqf_contract* __restrict__ contracts,
             const tqf_work_item* __restrict__ items,
             const int32_t* __restrict__ order, int n_items,
             uint32_t k0, uint32_t k1, int* __restrict__ counter,
             float* __restrict__ ckpt, int64_t ckpt_per_block,
             int max_seg, double* __restrict__ item_sums, int use_counter) {
  __shared__ int s_item;
  __shared__ double red[kBlockB / kWave][2 * kStats];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  float* ck = ckpt + (int64_t)blockIdx.x * ckpt_per_block + threadIdx.x;
  int next = blockIdx.x;
  for (;;) {
    if (use_counter) {
      __syncthreads();
      if (threadIdx.x == 0) s_item = atomicAdd(counter, 1);
      __syncthreads();
      next = s_item;
    }
    if (next >= n_items) break;
    const int row = order ? order[next] : next;
    const tqf_work_item it = items[row];
    Contract c;
    load_contract(contracts + it.contract, c);
    float acc[2 * kStats];
#pragma unroll
    for (int q = 0; q < 2 * kStats; ++q) acc[q] = 0.f;
    for (int64_t p = it.first_path + threadIdx.x; p < it.first_path + it.n_paths;
         p += kBlockB) {
      float g[kStats];
      if (c.kind == TQF_CONTRACT_GBM)
        gbm_path(c, p, c.stream, k0, k1, g);
      else if (!kAdjoint || c.greeks_mode == TQF_GREEKS_TANGENT ||
               (c.n_steps + kSeg - 1) / kSeg > max_seg)
        // a contract with more steps than the checkpoint workspace holds
        // (n_steps > max_steps) takes the forward tangents, the same
        // estimator, rather than write past its checkpoint column
        heston_path_tangent(c, p, c.stream, k0, k1, g);
      else
        heston_path(c, p, c.stream, k0, k1, ck, kBlockB, g);
#pragma unroll
      for (int q = 0; q < kStats; ++q) {
        acc[q] += g[q];
        acc[kStats + q] += g[q] * g[q];
      }
    }
    // wave reduction (float), then block (double)
#pragma unroll
    for (int q = 0; q < 2 * kStats; ++q) {
      float v = acc[q];
      for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
      if (lane == 0) red[wave][q] = (double)v;
    }
    __syncthreads();
    if (threadIdx.x < 2 * kStats) {
      double s = 0;
      for (int w = 0; w < kBlockB / kWave; ++w) s += red[w][threadIdx.x];
      item_sums[(int64_t)row * 2 * kStats + threadIdx.x] = s;
    }
    __syncthreads();
    if (!use_counter) next += gridDim.x;
  }
}

__global__ void zero_counter_kernel(int* c) { *c = 0; }

// Per contract: sums its items (contiguous rows, fixed order) -> mean, SE.
__global__ void __launch_bounds__(256)
batch_finalize_kernel(const double* __restrict__ item_sums,
                      const int* __restrict__ contract_first_item,
                      const int* __restrict__ contract_n_items,
                      const tqf_contract* __restrict__ contracts, int n_contracts,
                      double* __restrict__ out) {
  const int c = blockIdx.x * 256 + threadIdx.x;
  if (c >= n_contracts) return;
  double s[2 * kStats];
  for (int q = 0; q < 2 * kStats; ++q) s[q] = 0;
  const int f = contract_first_item[c], m = contract_n_items[c];
#pragma unroll 4
  for (int i = 0; i < m; ++i)
    for (int q = 0; q < 2 * kStats; ++q) s[q] += item_sums[(int64_t)(f + i) * 2 * kStats + q];
  const double n = (double)contracts[c].n_paths;
  for (int q = 0; q < kStats; ++q) {
    const double mean = s[q] / n;
    double var = n > 1 ? (s[kStats + q] - n * mean * mean) / (n - 1) : 0.0;
    out[(int64_t)c * 2 * kStats + q] = mean;
    out[(int64_t)c * 2 * kStats + kStats + q] = sqrt((var > 0 ? var : 0) / n);
  }
}

}  // namespace

extern "C" int64_t tqf_batch_workspace_size(int32_t n_items, int32_t n_blocks,
                                            int32_t max_steps) {
  const int64_t nseg = (max_steps + kSeg - 1) / kSeg;
  return (int64_t)n_items * 2 * kStats * 8 + 256 +
         (int64_t)n_blocks * kBlockB * nseg * 2 * 4;
}

extern "C" int tqf_batch_price(const tqf_contract* contracts,
                               int32_t n_contracts,
                               const tqf_work_item* items,
                               const int32_t* order, int32_t n_items,
                               const int32_t* contract_first_item,
                               const int32_t* contract_n_items, uint64_t seed,
                               int32_t n_blocks, int32_t max_steps,
                               int32_t flags, void* workspace,
                               double* out, void* stream) {
  clear_error();
  TQF_REQUIRE(contracts && items && contract_first_item && contract_n_items &&
                  workspace && out,
              "null argument");
  TQF_REQUIRE(n_contracts >= 1 && n_items >= 1 && n_blocks >= 1,
              "empty batch");
  TQF_REQUIRE(max_steps >= 1 && max_steps <= (1 << 20), "bad max_steps");
  const hipStream_t s = (hipStream_t)stream;
  char* w = (char*)workspace;
  double* item_sums = (double*)w;
  w += ((int64_t)n_items * 2 * kStats * 8 + 255) & ~(int64_t)255;
  int* counter = (int*)w;
  w += 256;
  float* ckpt = (float*)w;
  const int nseg = (max_steps + kSeg - 1) / kSeg;
  const int64_t per_block = (int64_t)kBlockB * nseg * 2;
  hipLaunchKernelGGL(zero_counter_kernel, 1, 1, 0, s, counter);
  const int use_counter = flags & TQF_BATCH_COUNTER;
  if (flags & TQF_BATCH_TANGENT_ONLY)
    hipLaunchKernelGGL(batch_kernel<false>, n_blocks, kBlockB, 0, s, contracts,
                       items, order, n_items, (uint32_t)seed,
                       (uint32_t)(seed >> 32), counter, ckpt, per_block,
                       nseg, item_sums, use_counter);
  else
    hipLaunchKernelGGL(batch_kernel<true>, n_blocks, kBlockB, 0, s, contracts,
                       items, order, n_items, (uint32_t)seed,
                       (uint32_t)(seed >> 32), counter, ckpt, per_block,
                       nseg, item_sums, use_counter);
  hipLaunchKernelGGL(batch_finalize_kernel, (n_contracts + 255) / 256, 256, 0,
                     s, item_sums, contract_first_item, contract_n_items,
                     contracts, n_contracts, out);
  return check_launch("tqf_batch_price");
}
