"""GPU parity tests: Sobol, Philox, fused Euler / exact GBM / Heston QE
kernels against the CPU oracle through the C ABI.

Tolerances: integers bit-exact; float64 paths 1e-12 relative (scale-aware),
float32 paths 2e-5 relative, with identical injected draws (the parity
bar of the engine against euler_sampling).  Streams generated in the kernels (Philox/Box-Muller, Sobol + inverse
CDF) are compared with the oracle's streams at the precision of the device
transcendentals, stated per test.
"""
import numpy as np
import pytest
import torch

from conftest import COEF_TIMING_GRID, load_golden, right_riemann_drift
from oracle import closed_form, euler as oeuler, gbm as ogbm
from oracle import heston as oheston, normal as onormal, rng as orng
from oracle import sobol as osobol

pytestmark = pytest.mark.gpu

if torch.cuda.is_available():
  import tqf
  from tqf.engine import euler as engine
  from tqf.engine import pricing
  from tqf.math import qmc
  from tqf.math.random import sobol as tsobol
  from tqf.math.random import multivariate_normal as mvn


def _np(t):
  return t.detach().cpu().numpy()


def _close(got, want, rtol, scale=None):
  got, want = np.asarray(got), np.asarray(want)
  scale = np.max(np.abs(want)) if scale is None else scale
  err = np.max(np.abs(got - want) / (np.abs(want) + 1e-3 * scale + 1e-300))
  assert err <= rtol, 'max scaled rel err %.3e > %.1e' % (err, rtol)


# ---------------------------------------------------------------- Sobol ----

# the dimension picks the kernel: whole-row tiles written with 4-, 8- or
# 16-byte stores (dim odd, even, a multiple of 4; dim <= 768), else one
# thread per (dimension, 256-index run)
@pytest.mark.parametrize('dim,n,skip', [
    (1, 1, 0), (2, 31, 0), (5, 1000, 17), (1, 3, 2**31 - 5), (2, 2**20, 0),
    (730, 2**13, 0), (730, 4096, 12345), (37, 777, 2**20 - 3),
    (4, 1000, 3), (8, 333, 100), (768, 100, 15), (769, 64, 0),
    (800, 500, 7)])
def test_sobol_integers_bit_exact(dim, n, skip):
  print('sobol ints dim=%d n=%d skip=%d' % (dim, n, skip), flush=True)
  got, nd = tsobol.sample_int(dim, n, skip)
  want, nd_o = osobol.sample_int(dim, n, skip)
  assert nd == nd_o
  np.testing.assert_array_equal(_np(got).view(np.uint32).astype(np.uint64),
                                want)


# whole-row tiles run on a persistent grid (three workgroups per CU), each
# workgroup advancing its item starts from tile to tile: sizes with more
# tiles than workgroups (16-row tiles of 730 dimensions, 4096-row tiles of
# one), every row compared
@pytest.mark.parametrize('dim,n,skip', [(730, 2**15, 7), (1, 2**23, 3),
                                        (37, 2**18, 1000)])
def test_sobol_tiles_advanced_across_a_workgroup_bit_exact(dim, n, skip):
  print('sobol tiles dim=%d n=%d skip=%d' % (dim, n, skip), flush=True)
  got, _ = tsobol.sample_int(dim, n, skip)
  want, _ = osobol.sample_int(dim, n, skip)
  np.testing.assert_array_equal(_np(got).view(np.uint32).astype(np.uint64),
                                want)


def test_sobol_sample_full_size_rows_match_the_oracle():
  # the C2 draw set, sobol.sample(730, 2^20) float32 (about 86 tiles per
  # workgroup): rows at tile boundaries and spread over the range against
  # the oracle's row of the same index
  got = _np(tqf.math.random.sobol.sample(730, 2**20))
  rows = [0, 1, 14, 15, 16, 17, 31, 32, 1023, 1024, 2**19 - 1, 2**19,
          2**20 - 17, 2**20 - 16, 2**20 - 1]
  rows += list(np.random.RandomState(4).randint(0, 2**20, 16))
  for r in rows:
    want = osobol.sample(730, 1, int(r))[0]
    np.testing.assert_array_equal(got[r], want, err_msg='row %d' % r)


def test_sobol_sample_goldens_and_dtypes():
  g = load_golden('sobol_golden.json')
  e = g['random_ops_dim2_first5']
  for dt in (np.float16, np.float32, np.float64):
    got = tqf.math.random.sobol.sample(2, 5, dtype=dt)
    np.testing.assert_allclose(_np(got).astype(np.float64), e['expected'],
                               rtol=1e-6)
  e = g['random_ops_large_skip']
  np.testing.assert_allclose(_np(tqf.math.random.sobol.sample(1, 3, e['skip'])),
                             e['expected'])
  s = _np(tqf.math.random.sobol.sample(5, 31))
  assert sorted(map(tuple, s.tolist())) == sorted(
      map(tuple, g['random_ops_dim5_31_as_set']['expected']))
  a = _np(tqf.math.random.sobol.sample(10, 67))
  b = _np(tqf.math.random.sobol.sample(10, 50, skip=17))
  np.testing.assert_array_equal(a[17:], b)
  # float32 rounding of cast(int) / 2^nd matches the oracle at nd = 31
  got = _np(tqf.math.random.sobol.sample(3, 1000, skip=2**31 - 1002))
  np.testing.assert_array_equal(got, osobol.sample(3, 1000, 2**31 - 1002))


@pytest.mark.parametrize('dim,n,skip', [(730, 300, 5), (800, 100, 0),
                                         (3, 1000, 17)])
def test_sobol_float64_matches_oracle(dim, n, skip):
  """float64 points (8-byte outputs take the per-(dimension, run) kernel)
  are cast(int) / 2^nd exactly."""
  got = _np(tqf.math.random.sobol.sample(dim, n, skip, dtype=np.float64))
  np.testing.assert_array_equal(got, osobol.sample(dim, n, skip, np.float64))


def test_sobol_validate_args():
  with pytest.raises(tqf.errors.InvalidArgumentError):
    tqf.math.random.sobol.sample(1, 4, skip=2**31 - 5, validate_args=True)


def test_qmc_goldens_shift_indices_tent():
  g = load_golden('sobol_golden.json')
  np.testing.assert_allclose(_np(qmc.sobol_sample(5, 29)),
                             g['qmc_sobol_29x5']['expected'])
  e = g['qmc_sequence_indices']
  np.testing.assert_allclose(
      _np(qmc.sobol_sample(5, 29, sequence_indices=e['indices'])),
      e['expected'])
  np.testing.assert_allclose(
      _np(qmc.sobol_sample(6, 8, apply_tent_transform=True)),
      g['qmc_tent_6x8']['expected'])
  np.testing.assert_allclose(_np(qmc.sobol_sample(6, 8, dtype=np.float64)),
                             g['qmc_dtype_6x8']['expected'])
  gm = qmc.sobol_generating_matrices(5, 31, 5)
  np.testing.assert_array_equal(gm, g['qmc_generating_matrices']['expected'])
  # digital shift: XOR with random 20-digit words
  shift = qmc.random_digital_shift(9, 20, seed=11)
  assert np.all(shift >= 2**19) and np.all(shift < 2**20)
  gm = qmc.sobol_generating_matrices(9, 5000, 20)
  got = _np(qmc.digital_net_sample(gm, 5000, 20, digital_shift=shift,
                                   dtype=np.float64))
  want = osobol.digital_net_sample(gm, 5000, 20, digital_shift=shift,
                                   dtype=np.float64)
  np.testing.assert_array_equal(got, want)


# ---------------------------------------------------------------- Philox ---

@pytest.mark.parametrize('dtype,tol', [(torch.float64, 1e-13),
                                       (torch.float32, 2e-5)])
def test_philox_normals_match_oracle(dtype, tol):
  npd = np.float64 if dtype == torch.float64 else np.float32
  got = _np(mvn.philox_normals(3000, 7, 1234, dtype, stream_id=3))
  want = orng.normals(3000, 7, 1234, stream=3, dtype=npd)
  assert np.max(np.abs(got - want)) <= tol * 10
  got = _np(mvn.philox_normals(64, 5, 99, dtype, antithetic=True))
  np.testing.assert_array_equal(got[:32], -got[32:])
  big = _np(mvn.philox_normals(2**18, 8, 7, dtype)).ravel()
  assert abs(big.mean()) < 5 / np.sqrt(big.size)
  assert abs(big.var() - 1) < 5 * np.sqrt(2 / big.size)


def test_mv_normal_sample_sobol_layout():
  # SOBOL through mv_normal_sample: erfinv((u - .5) * 2) sqrt 2 of sobol.sample
  z = _np(tqf.math.random.mv_normal_sample([1000], mean=np.zeros(6),
                                           random_type=mvn.RandomType.SOBOL,
                                           skip=5, dtype=np.float64))
  u = osobol.sample(6, 1000, skip=5, dtype=np.float64)
  np.testing.assert_allclose(z, onormal.quasi_normals(u), rtol=1e-13,
                             atol=1e-13)


# ---------------------------------------------------------------- Euler ----

def _models():
  return {
      'gbm': (('gbm', 0.05, 0.2), oeuler.gbm_fns(0.05, 0.2), [100.0]),
      'const': (('const', 0.05 - 0.02, 0.2), oeuler.gbm_log_fns(0.05, 0.2),
                [np.log(100.0)]),
      'heston_ft': (('heston', 0.03, 2.0, 0.04, 0.3, -0.7, 0.0),
                    oeuler.heston_ft_fns(2.0, 0.04, 0.3, -0.7, 0.03),
                    [np.log(100.), 0.04]),
  }


def _heston_ref_fns(kappa, theta, xi, rho):
  def drift(t, x):
    return oheston.drift_vol(x, kappa, theta, xi, rho)[0]

  def vol(t, x):
    return oheston.drift_vol(x, kappa, theta, xi, rho)[1]
  return drift, vol


@pytest.mark.parametrize('name', ['gbm', 'const', 'heston_ft', 'heston_ref'])
@pytest.mark.parametrize('dtype,tol', [(torch.float64, 1e-12),
                                       (torch.float32, 2e-5)])
def test_euler_injected_draws_parity(name, dtype, tol):
  print('euler parity', name, dtype, flush=True)
  npd = np.float64 if dtype == torch.float64 else np.float32
  if name == 'heston_ref':
    spec = ('heston', 0.0, 2.0, 0.04, 0.3, -0.7, 1.0)
    fns, x0 = _heston_ref_fns(2.0, 0.04, 0.3, -0.7), [np.log(100.), 0.04]
  else:
    spec, fns, x0 = _models()[name]
  d = engine.model_dim(spec)
  times = [0.25, 0.5, 1.0]
  rs = np.random.RandomState(7)
  from oracle import grid as ogrid
  t, mask, _ = ogrid.prepare_grid(times, npd, num_time_steps=50)
  z = rs.standard_normal((4000, t.shape[0] - 1, d)).astype(npd)
  want = oeuler.sample(d, fns[0], fns[1], np.asarray(times, npd),
                       num_samples=4000, initial_state=np.asarray(x0, npd),
                       normal_draws=z, num_time_steps=50, dtype=npd)
  drift = engine.tag(lambda t, x: None, spec)
  vol = engine.tag(lambda t, x: None, spec)
  got = tqf.models.euler_sampling.sample(
      d, drift, vol, np.asarray(times, npd), num_time_steps=50,
      initial_state=np.asarray(x0, npd), normal_draws=z, dtype=dtype)
  assert got.shape == want.shape
  for c in range(d):
    _close(_np(got)[..., c], want[..., c], tol)


@pytest.mark.parametrize('dtype', [torch.float64, torch.float32])
def test_euler_sobol_stream_parity(dtype):
  npd = np.float64 if dtype == torch.float64 else np.float32
  n, steps, skip = 2048, 40, 3
  spec, fns, x0 = _models()['heston_ft']
  u = osobol.sample(steps * 2, n, skip=skip, dtype=npd)
  z = onormal.quasi_normals(u.astype(np.float64)).reshape(n, steps, 2)
  want = oeuler.sample(2, fns[0], fns[1], np.asarray([1.0], npd),
                       num_samples=n, initial_state=np.asarray(x0, npd),
                       normal_draws=z.astype(npd), num_time_steps=steps,
                       dtype=npd)
  drift = engine.tag(lambda t, x: None, spec)
  got = tqf.models.euler_sampling.sample(
      2, drift, drift, np.asarray([1.0], npd), num_time_steps=steps,
      num_samples=n, initial_state=np.asarray(x0, npd),
      random_type=mvn.RandomType.SOBOL, skip=skip, dtype=dtype)
  # float32: the device erfinv is within 3e-7 relative of scipy's
  _close(_np(got), want, 1e-11 if dtype == torch.float64 else 5e-5)


@pytest.mark.parametrize('dtype', [torch.float64, torch.float32])
def test_euler_philox_stream_parity(dtype):
  npd = np.float64 if dtype == torch.float64 else np.float32
  n, steps = 2048, 30
  spec, fns, x0 = _models()['gbm']
  key = mvn.seed_to_key(42)
  z = orng.normals(n, steps, key, dtype=npd).reshape(n, steps, 1)
  want = oeuler.sample(1, fns[0], fns[1], np.asarray([1.0], npd),
                       num_samples=n, initial_state=np.asarray(x0, npd),
                       normal_draws=z.astype(npd), num_time_steps=steps,
                       dtype=npd)
  drift = engine.tag(lambda t, x: None, spec)
  got = tqf.models.euler_sampling.sample(
      1, drift, drift, np.asarray([1.0], npd), num_time_steps=steps,
      num_samples=n, initial_state=np.asarray(x0, npd),
      random_type=mvn.RandomType.PSEUDO, seed=42, dtype=dtype)
  _close(_np(got), want, 1e-11 if dtype == torch.float64 else 1e-4)


def test_euler_payoff_mode_matches_recorded_paths():
  n = 50000
  spec = ('heston', 0.03, 2.0, 0.04, 0.3, -0.7, 0.0)
  all_times, keep = pricing.heston_grid_for([0.5, 1.0], 100, 'reference',
                                            torch.float64)
  rng = engine.RngSpec('philox', seed=5)
  x0 = np.array([np.log(100.), 0.04])
  paths = engine.run(spec, all_times, keep, 2, x0, n, rng, torch.float64,
                     mode='paths')
  strikes = np.linspace(80, 120, 70)  # more than one 64-lane chunk
  res = engine.run(spec, all_times, keep, 2, x0, n, rng, torch.float64,
                   mode='payoff', strikes=strikes,
                   discount=np.exp(-0.03 * np.array([0.5, 1.0])))
  s = np.exp(_np(paths)[:, :, 0])
  pay = np.maximum(s[:, :, None] - strikes[None, None, :], 0)
  disc = np.exp(-0.03 * np.array([0.5, 1.0]))[:, None]
  np.testing.assert_allclose(_np(res['prices']), disc * pay.mean(0),
                             rtol=1e-12)
  np.testing.assert_allclose(
      _np(res['stderrs']), disc * pay.std(0, ddof=1) / np.sqrt(n), rtol=1e-9)
  # the two-call form (per-block sums, then tqf_payoff_finalize) gives the
  # same bits; 196 blocks = 6 full chunks of 32 and a ragged one
  lz = engine.prepare(spec, all_times, keep, 2, x0, n, rng, torch.float64,
                      mode='payoff', strikes=strikes,
                      discount=np.exp(-0.03 * np.array([0.5, 1.0])))
  paths_call, finalize_call = lz.launch_split()
  paths_call()
  split = finalize_call()
  for key in ('prices', 'stderrs', 'sums'):
    np.testing.assert_array_equal(_np(split[key]), _np(res[key]))
  # store mode is the transpose of path mode
  store = engine.run(spec, all_times, keep, 2, x0, n, rng, torch.float64,
                     mode='store')
  np.testing.assert_array_equal(_np(store).transpose(2, 0, 1), _np(paths))


def test_sampling_apis_return_the_store_as_a_path_major_view():
  # euler_sampling.sample, HestonModel.sample_paths and the GBM sample_paths
  # hand back [N, k, d] views of the coalesced [k, d, N] store: same shape
  # and, bit for bit, the values of the contiguous path-major mode
  from tqf.models import geometric_brownian_motion as gbm_lib
  n = 30000
  spec = ('heston', 0.03, 2.0, 0.04, 0.3, -0.7, 0.0)
  all_times, keep = pricing.heston_grid_for([0.25, 0.5, 1.0], 120,
                                            'reference', torch.float32)
  x0 = np.array([np.log(100.), 0.04])
  shift = pricing.sobol_digital_shift(240, 3)
  for rng in (engine.RngSpec('philox', seed=9),
              engine.RngSpec('sobol_shifted', shift=shift, first_index=1)):
    a = engine.run(spec, all_times, keep, 3, x0, n, rng, torch.float32,
                   mode='paths')
    b = engine.run(spec, all_times, keep, 3, x0, n, rng, torch.float32,
                   mode='view')
    assert b.shape == a.shape and b.stride(0) == 1 and a.is_contiguous()
    assert torch.equal(a, b)
  times = np.linspace(0.1, 1.0, 10)
  rng = engine.RngSpec('philox_antithetic', seed=4)
  for dt in (torch.float32, torch.float64):
    a = gbm_lib.run_exact(times, [0.03], [0.2], None, 100.0, n, rng, dt,
                          mode='paths')
    b = gbm_lib.run_exact(times, [0.03], [0.2], None, 100.0, n, rng, dt,
                          mode='view')
    assert b.shape == a.shape == (n, 10, 1) and b.stride(0) == 1
    assert torch.equal(a, b)
  p = tqf.models.GeometricBrownianMotion(0.05, 0.2).sample_paths(
      times, initial_state=100.0, num_samples=1000, seed=(1, 2),
      random_type=mvn.RandomType.STATELESS_ANTITHETIC)
  assert p.shape == (1000, 10, 1) and p.stride(0) == 1


def test_gbm_euler_prices_within_3se_of_black_scholes():
  # Plain method: 1-asset GBM Euler, float64, 1e5 paths, 100 steps
  strikes = np.linspace(80, 120, 8)
  all_times, keep = pricing.heston_grid_for([1.0], 100, 'reference',
                                            torch.float64)
  res = pricing.european_calls(('gbm', 0.05, 0.2), [100.0], strikes,
                               all_times, keep, 100000,
                               engine.RngSpec('philox', seed=42),
                               discount_rate=0.05, dtype=torch.float64)
  p, se = _np(res['prices'])[0], _np(res['stderrs'])[0]
  bs = closed_form.bs_price(100, strikes, 1.0, 0.05, 0.2)
  assert np.all(np.abs(p - bs) <= 3 * se), (p - bs) / se


def test_generic_loop_and_local_vol_kernel_read_coefficients_at_next_time():
  # zero draws: each value is the right Riemann sum of the drift, which
  # tells times[i + 1] (euler_sampling.py:519) from times[i]
  mu, x0 = 0.2, 0.1
  n_steps = COEF_TIMING_GRID.shape[0] - 1
  zeros = np.zeros((3, n_steps, 1))
  right, left = right_riemann_drift(mu, x0, COEF_TIMING_GRID)
  got = tqf.models.euler_sampling.sample(
      1, lambda t, x: mu * torch.sqrt(t) * torch.ones_like(x),
      lambda t, x: torch.zeros(x.shape + (1,), dtype=x.dtype, device=x.device),
      [0.3, 1.0], times_grid=COEF_TIMING_GRID, initial_state=[x0],
      normal_draws=zeros, dtype=torch.float64)
  np.testing.assert_allclose(_np(got)[:, :, 0], [[right[3], right[-1]]] * 3,
                             rtol=1e-14)
  # the fused local-volatility kernel on a time-only surface whose nodes are
  # the grid times: log S_T = log S_0 + sum dt_i (r - q - sigma(t_{i+1})^2/2)
  sig = 0.2 + 0.3 * COEF_TIMING_GRID
  xs = np.linspace(np.log(50.), np.log(200.), 5)
  surf = tqf.models.LocalVolSurface(COEF_TIMING_GRID, xs[0], xs[1] - xs[0],
                                    np.repeat(sig[:, None], 5, 1))
  model = tqf.models.LocalVolatilityModel(
      1, risk_free_rate=0.03, dividend_yield=0.01, local_volatility_fn=surf,
      dtype=torch.float64)
  assert engine.spec_of(model.drift_fn(), model.volatility_fn()) is not None
  s = model.sample_paths([0.3, 1.0], initial_state=[100.0],
                         times_grid=COEF_TIMING_GRID, normal_draws=zeros)
  dt = np.diff(COEF_TIMING_GRID)
  want = np.log(100.) + np.cumsum(dt * (0.02 - 0.5 * sig[1:]**2))
  want_left = np.log(100.) + np.cumsum(dt * (0.02 - 0.5 * sig[:-1]**2))
  np.testing.assert_allclose(np.log(_np(s))[:, :, 0],
                             [[want[3], want[-1]]] * 3, rtol=1e-13)
  assert abs(want[-1] - want_left[-1]) > 1e-3


def test_generic_loop_time_dependent_process_mean():
  # euler_sampling_test.py:206-294: dX = mu sqrt(t) dt + (a t + b) dW has
  # E[X_t] = x0 + 2/3 mu t^1.5
  mu, a, b, x0 = 0.2, 0.4, 0.33, 0.1
  times = np.array([0.0, 0.1, 0.21, 0.32, 0.43, 0.55])
  kw = dict(num_samples=10000, initial_state=[x0],
            random_type=mvn.RandomType.PSEUDO_ANTITHETIC, seed=42,
            time_step=0.01, dtype=torch.float64)
  paths = _np(tqf.models.euler_sampling.sample(
      1, lambda t, x: mu * torch.sqrt(t) * torch.ones_like(x),
      lambda t, x: (a * t + b) * torch.ones(x.shape + (1,), dtype=x.dtype,
                                            device=x.device), times, **kw))
  assert paths.shape == (10000, 6, 1)
  np.testing.assert_allclose(paths.mean(0)[:, 0],
                             x0 + 2.0 / 3.0 * mu * times**1.5, rtol=1e-2,
                             atol=1e-2)
  # antithetic pairs cancel the noise: the mean is the right Riemann sum
  from oracle import grid as ogrid
  t, _, idx = ogrid.prepare_grid(times, np.float64, time_step=0.01)
  right, _ = right_riemann_drift(mu, x0, t)
  np.testing.assert_allclose(paths.mean(0)[1:, 0], right[idx[1:] - 1],
                             rtol=1e-12)


def test_euler_float32_parity_on_the_headline_grid():
  # the headline workload's own grid (num_time_steps=365, maturities .25 ..
  # 1: 366 points) in float32 with injected draws, 2^14 paths
  print('euler float32 parity, headline grid', flush=True)
  spec, fns, x0 = _models()['heston_ft']
  times = np.array([0.25, 0.5, 0.75, 1.0], np.float32)
  from oracle import grid as ogrid
  t, _, _ = ogrid.prepare_grid(times, np.float32, num_time_steps=365)
  assert t.shape[0] == 366
  n = 2**14
  z = np.random.RandomState(29).standard_normal((n, 365, 2)).astype(
      np.float32)
  want = oeuler.sample(2, fns[0], fns[1], times, num_samples=n,
                       initial_state=np.asarray(x0, np.float32),
                       normal_draws=z, num_time_steps=365, dtype=np.float32)
  drift = engine.tag(lambda t, x: None, spec)
  vol = engine.tag(lambda t, x: None, spec)
  got = _np(tqf.models.euler_sampling.sample(
      2, drift, vol, times, num_time_steps=365,
      initial_state=np.asarray(x0, np.float32), normal_draws=z,
      dtype=torch.float32))
  assert got.shape == want.shape == (n, 4, 2)
  for c in range(2):
    _close(got[..., c], want[..., c], 2e-5)


def test_heston_qe_sample_paths_price_the_smile():
  # the reference's cross-check of HestonModel.sample_paths against the
  # semi-analytic price (heston_model_test.py:290-380), here at rho = -0.7
  # and xi = 0.6 where the QE spot update's correlation terms carry the
  # skew, 2^20 antithetic paths, every strike within 3 standard errors
  print('heston QE vs semi-analytic', flush=True)
  kappa, theta, xi, rho, v0 = 1.5, 0.04, 0.6, -0.7, 0.04
  model = tqf.models.HestonModel(kappa, theta, xi, rho, dtype=torch.float64)
  paths = model.sample_paths([0.5, 1.0], [np.log(100.), v0],
                             num_samples=2**20, num_time_steps=100,
                             random_type=mvn.RandomType.PSEUDO_ANTITHETIC,
                             seed=5)
  s = np.exp(_np(paths)[:, -1, 0])
  for k in (70.0, 85.0, 100.0, 115.0, 130.0):
    pay = np.maximum(s - k, 0)
    want = oheston.european_call(100.0, k, 1.0, 0.0, v0, kappa, theta, xi,
                                 rho)
    se = pay.std(ddof=1) / np.sqrt(s.size)
    assert abs(pay.mean() - want) < 3 * se, (k, pay.mean(), want, se)
  assert abs(s.mean() - 100.0) < 3 * s.std() / np.sqrt(s.size)


def test_batch_shape_generic_loop_reference_case():
  # euler_sampling_test.py:206-294 with use_batch: initial_state [2, 1, 1]
  # and a volatility with batch axes [2, 1, 1, 1]; paths [2, N, 6, 1]
  mu, a, b = 0.2, 0.4, 0.33
  times = np.array([0.0, 0.1, 0.21, 0.32, 0.43, 0.55])
  x0 = np.array([[[0.1]], [[0.1]]])
  paths = _np(tqf.models.euler_sampling.sample(
      1, lambda t, x: mu * torch.sqrt(t) * torch.ones_like(x),
      lambda t, x: (a * t + b) * torch.ones([2, 1, 1, 1], dtype=x.dtype,
                                            device=x.device),
      times, num_samples=10000, initial_state=x0,
      random_type=mvn.RandomType.PSEUDO_ANTITHETIC, seed=42, time_step=0.01,
      dtype=torch.float64))
  assert paths.shape == (2, 10000, 6, 1)
  means = paths.mean(1).reshape(2, 1, 6)
  np.testing.assert_allclose(means, x0 + 2.0 / 3.0 * mu * times**1.5,
                             rtol=1e-2, atol=1e-2)
  # the batch members have independent draws
  assert not np.allclose(paths[0], paths[1])


def test_batch_shape_fused_equals_one_run_per_member():
  # a batch of initial states and batched draws for one fused model: one
  # launch over the flattened paths, equal to a run per member
  rs = np.random.RandomState(12)
  n, steps = 500, 20
  z = rs.standard_normal((3, n, steps, 1))
  x0 = np.array([90.0, 100.0, 110.0]).reshape(3, 1, 1)
  gbm = tqf.models.GeometricBrownianMotion(0.05, 0.2, dtype=torch.float64)
  got = _np(tqf.models.euler_sampling.sample(
      1, gbm.drift_fn(), gbm.volatility_fn(), [0.5, 1.0],
      num_time_steps=steps, initial_state=x0, normal_draws=z,
      dtype=torch.float64))
  assert got.shape == (3, n, 2, 1)
  for b in range(3):
    one = _np(tqf.models.euler_sampling.sample(
        1, gbm.drift_fn(), gbm.volatility_fn(), [0.5, 1.0],
        num_time_steps=steps, initial_state=x0[b, 0], normal_draws=z[b],
        dtype=torch.float64))
    np.testing.assert_array_equal(got[b], one)
  # and the same batch through the generic loop (plain callables)
  loop = _np(tqf.models.euler_sampling.sample(
      1, lambda t, x: 0.05 * x, lambda t, x: 0.2 * x.unsqueeze(-1),
      [0.5, 1.0], num_time_steps=steps, initial_state=x0, normal_draws=z,
      dtype=torch.float64))
  np.testing.assert_allclose(loop, got, rtol=1e-12)


def test_generic_callables_match_fused_kernel():
  times = [0.5, 1.0]
  rs = np.random.RandomState(3)
  z = rs.standard_normal((1000, 20, 2))
  model = tqf.models.HestonModel(2.0, 0.04, 0.3, -0.7, dtype=np.float64)
  fd, fv = model.euler_ft_fns(0.03)
  fused = tqf.models.euler_sampling.sample(
      2, fd, fv, times, num_time_steps=20, initial_state=[np.log(100), .04],
      normal_draws=z, dtype=torch.float64)
  plain_d = lambda t, x: fd(t, x)
  plain_v = lambda t, x: fv(t, x)
  generic = tqf.models.euler_sampling.sample(
      2, plain_d, plain_v, times, num_time_steps=20,
      initial_state=[np.log(100), .04], normal_draws=z, dtype=torch.float64)
  _close(_np(generic), _np(fused), 1e-12)


# ---------------------------------------------------------------- GBM exact

@pytest.mark.parametrize('dtype,tol', [(torch.float64, 1e-12),
                                       (torch.float32, 2e-5)])
def test_gbm_exact_parity(dtype, tol):
  npd = np.float64 if dtype == torch.float64 else np.float32
  rs = np.random.RandomState(11)
  times = np.array([0.1, 0.4, 1.0, 2.0])
  z = rs.standard_normal((5000, 4, 1)).astype(npd)
  p = tqf.models.GeometricBrownianMotion(0.05, 0.3, dtype=dtype)
  got = p.sample_paths(times, initial_state=100.0, normal_draws=z)
  want = ogbm.sample_paths(times, 0.05, 0.3, 100.0, z, dtype=npd)
  _close(_np(got), want, tol)
  corr = np.array([[1, .3, .1], [.3, 1, .2], [.1, .2, 1]])
  z = rs.standard_normal((5000, 4, 3)).astype(npd)
  mp = tqf.models.MultivariateGeometricBrownianMotion(
      3, means=[0.01, 0.02, 0.03], volatilities=[0.1, 0.2, 0.3],
      corr_matrix=corr, dtype=dtype)
  got = mp.sample_paths(times, initial_state=[1., 2., 3.], normal_draws=z)
  want = ogbm.mv_sample_paths(times, [0.01, 0.02, 0.03], [0.1, 0.2, 0.3],
                              corr, [1., 2., 3.], z, dtype=npd)
  _close(_np(got), want, tol)


def test_gbm_exact_moments_sobol_and_philox():
  p = tqf.models.GeometricBrownianMotion(0.05, 0.2, dtype=torch.float64)
  for rt in (mvn.RandomType.SOBOL, mvn.RandomType.PSEUDO,
             mvn.RandomType.PSEUDO_ANTITHETIC):
    s = _np(p.sample_paths([1.0], initial_state=1.0, num_samples=100000,
                           random_type=rt, seed=3))[:, 0, 0]
    lg = np.log(s)
    assert abs(lg.mean() - 0.03) < 3 * lg.std() / np.sqrt(lg.size)
    assert abs(lg.var() - 0.04) < 3 * 0.04 * np.sqrt(2 / (lg.size - 1))


# ------------------------------------------- float32 shifted-Sobol normals

def test_shifted_normal_transform_exhaustive():
  # dimension 0 with 32 digits is the bit reversal of the index: indices
  # 0 .. 2^25 - 1 put every value of the top 25 bits through the kernel
  import ctypes
  from tqf import _native
  n = 2**25
  v = tsobol.device_columns(1, 32)
  want = None
  for shift in (None, 0x5A):  # the low 7 bits must not matter
    out = torch.empty(n, dtype=torch.float32, device='cuda')
    sh = None
    if shift is not None:
      sh = torch.tensor([shift], dtype=torch.int32, device='cuda')
    _native.call('tqf_sobol_normals', _native.ptr(v), ctypes.c_int32(1),
                 ctypes.c_int32(32), ctypes.c_int64(n), ctypes.c_int64(0),
                 _native.ptr(sh) if sh is not None else None,
                 ctypes.c_int32(1), ctypes.c_int32(_native.TQF_F32),
                 _native.ptr(out), _native.stream_ptr(out.device))
    got = _np(out).astype(np.float64)
    if want is None:
      k = np.arange(n, dtype=np.uint64)
      rev = np.zeros(n, np.uint64)
      for b in range(25):
        rev |= ((k >> np.uint64(b)) & np.uint64(1)) << np.uint64(31 - b)
      want = onormal.shifted_sobol_normals(rev.astype(np.uint32))
      ref_bits = got
    else:
      np.testing.assert_array_equal(got, ref_bits)
    err = np.abs(got - want)
    big = np.abs(want) >= 1e-3
    assert err.max() <= 5.5e-7, err.max()
    assert (err[big] / np.abs(want[big])).max() <= 3.5e-7
    assert err[~big].max() <= 2e-10


def test_euler_shifted_sobol_parity():
  # fused float32 Heston FT with the shifted-Sobol stream against the
  # oracle driven by the same points through the exact inverse CDF
  n, steps = 4096, 50
  spec, fns, x0 = _models()['heston_ft']
  shift = pricing.sobol_digital_shift(2 * steps, 7)
  ints = oheston.shifted_sobol_ints(2 * steps, n, shift, first=5)
  z = onormal.shifted_sobol_normals(ints.astype(np.uint32)).reshape(
      n, steps, 2)
  times = np.asarray([1.0], np.float32)
  want = oeuler.sample(2, fns[0], fns[1], times, num_samples=n,
                       initial_state=np.asarray(x0, np.float32),
                       normal_draws=z.astype(np.float32),
                       num_time_steps=steps, dtype=np.float32)
  all_times, keep = pricing.heston_grid_for([1.0], steps, 'reference',
                                            torch.float32)
  rng = engine.RngSpec('sobol_shifted', shift=shift, first_index=5)
  got = engine.run(spec, all_times, keep, 1, np.asarray(x0), n, rng,
                   torch.float32, mode='paths')
  # inverse CDF within 5e-7 of the exact one, float32 path rounding
  _close(_np(got), want, 5e-5)


# ---------------------------------------------------------- local vol

def _lv_surface(nt=4, nx=9):
  times = np.linspace(0.0, 1.2, nt)
  x = np.linspace(np.log(40.), np.log(250.), nx)
  vols = (0.25 - 0.06 * (x[None, :] - np.log(100.)) +
          0.04 * np.sin(3 * x[None, :]) + 0.05 * times[:, None])
  return times, x, vols


@pytest.mark.parametrize('nx', [9, 900])  # LDS-staged and global table
@pytest.mark.parametrize('dtype,tol', [(torch.float64, 1e-12),
                                       (torch.float32, 2e-5)])
def test_local_vol_parity(dtype, tol, nx):
  from oracle import local_vol as olv
  npd = np.float64 if dtype == torch.float64 else np.float32
  times, x, vols = _lv_surface(7 if nx > 100 else 4, nx)
  surf = tqf.models.LocalVolSurface(times, x[0], x[1] - x[0], vols)
  model = tqf.models.LocalVolatilityModel(1, risk_free_rate=0.03,
                                          dividend_yield=0.01,
                                          local_volatility_fn=surf,
                                          dtype=dtype)
  rs = np.random.RandomState(21)
  n, steps = 3000, 40
  z = rs.standard_normal((n, steps, 1)).astype(npd)
  got = _np(model.sample_paths([0.5, 1.0], initial_state=[100.0],
                               num_time_steps=steps, normal_draws=z))
  fns = olv.lv_fns(times, x[0], x[1] - x[0], vols, 0.03, 0.01)
  want = oeuler.sample(1, fns[0], fns[1], np.asarray([0.5, 1.0], npd),
                       num_samples=n,
                       initial_state=np.asarray([np.log(100.)], npd),
                       normal_draws=z, num_time_steps=steps, dtype=npd)
  _close(got, np.exp(want), tol)
  # the same surface as an arbitrary callable takes the generic loop
  generic = tqf.models.LocalVolatilityModel(
      1, risk_free_rate=0.03, dividend_yield=0.01,
      local_volatility_fn=lambda t, s: surf(t, s), dtype=dtype)
  got2 = _np(generic.sample_paths([0.5, 1.0], initial_state=[100.0],
                                  num_time_steps=steps, normal_draws=z))
  _close(got2, got, 10 * tol)


def test_local_vol_flat_surface_is_black_scholes():
  times = np.array([0.0, 1.0])
  x = np.linspace(np.log(20.), np.log(500.), 5)
  surf = tqf.models.LocalVolSurface(times, x[0], x[1] - x[0],
                                    np.full((2, 5), 0.25))
  model = tqf.models.LocalVolatilityModel(1, risk_free_rate=0.04,
                                          local_volatility_fn=surf,
                                          dtype=torch.float64)
  s = _np(model.sample_paths([1.0], initial_state=[100.0], num_time_steps=20,
                             num_samples=200000,
                             random_type=mvn.RandomType.PSEUDO, seed=5))
  st = s[:, 0, 0]
  for k in (80.0, 100.0, 120.0):
    pay = np.exp(-0.04) * np.maximum(st - k, 0)
    bs = closed_form.bs_price(100.0, k, 1.0, 0.04, 0.25)
    assert abs(pay.mean() - bs) <= 3 * pay.std() / np.sqrt(st.size), k


def test_local_vol_precompute_iv_steps_on_times_grid_and_rejects_grids():
  # a user callable is kept as it is; precompute_iv only makes the model
  # step on times_grid (local_volatility_model.py:425-436, 521-529)
  surf_fn = lambda t, s: 0.2 + 0.0 * s
  model = tqf.models.LocalVolatilityModel(
      1, risk_free_rate=0.02, local_volatility_fn=surf_fn,
      times_grid=np.linspace(0.0, 1.0, 11), spot_grid=[50.0, 200.0],
      precompute_iv=True, dtype=torch.float64)
  assert model.local_volatility_fn() is surf_fn
  with pytest.raises(ValueError):
    model.sample_paths([1.0], initial_state=[100.0], time_step=0.1)
  p = _np(model.sample_paths([1.0], initial_state=[100.0], num_samples=1000,
                             random_type=mvn.RandomType.PSEUDO, seed=1))
  assert p.shape == (1000, 1, 1) and np.all(p > 0)


# ---------------------------------------------------------------- Heston QE

@pytest.mark.parametrize('dtype,tol', [(torch.float64, 1e-12),
                                       (torch.float32, 2e-5)])
def test_heston_qe_parity(dtype, tol):
  npd = np.float64 if dtype == torch.float64 else np.float32
  rs = np.random.RandomState(13)
  times = np.array([0.3, 1.0])
  from oracle import grid as ogrid
  all_t, _ = ogrid.heston_prepare_grid(times, npd, time_step=0.05)
  z = rs.standard_normal((4000, all_t.shape[0] - 1, 2)).astype(npd)
  model = tqf.models.HestonModel(1.2, 0.05, 0.9, -0.6, dtype=dtype)
  got = _np(model.sample_paths(times, [np.log(100.), 0.03], time_step=0.05,
                               normal_draws=z))
  want = oheston.sample_paths_qe(times, np.log(100.), 0.03, 1.2, 0.05, 0.9,
                                 -0.6, z, time_step=0.05, dtype=npd)
  _close(got[..., 0], want[..., 0], tol)
  # the variance has exact zeros on the psi > 1.5 branch: absolute scale
  _close(got[..., 1], want[..., 1], tol, scale=1e3 * np.max(want[..., 1]))


# --------------------------------------------------------------- Halton

@pytest.mark.parametrize('dtype,rtol', [(torch.float32, 1e-6),
                                        (torch.float64, 1e-13)])
def test_halton_matches_oracle(dtype, rtol):
  from oracle import halton as ohalton
  from tqf.math.random import halton
  npd = np.float32 if dtype == torch.float32 else np.float64
  got, params = halton.sample(2, num_results=5, randomized=False, dtype=dtype)
  assert params is None
  np.testing.assert_allclose(
      _np(got), [[1. / 2, 1. / 3], [1. / 4, 2. / 3], [3. / 4, 1. / 9],
                 [1. / 8, 4. / 9], [5. / 8, 7. / 9]], rtol=1e-6)
  for dim, idx in ((17, np.arange(3000)),
                   (40, np.array([0, 5, 2**20 + 3, 12345678]))):
    got, _ = halton.sample(dim, sequence_indices=idx, randomized=False,
                           dtype=dtype)
    want = ohalton.sample(dim, sequence_indices=idx, dtype=npd)
    np.testing.assert_allclose(_np(got), want, rtol=rtol, atol=1e-30)
  # randomized: the oracle given the same permutations and correction
  got, params = halton.sample(7, num_results=2000, seed=11, dtype=dtype)
  want = ohalton.sample(7, num_results=2000,
                        perms=_np(params.perms),
                        zero_correction=_np(params.zero_correction),
                        dtype=npd)
  np.testing.assert_allclose(_np(got), want, rtol=rtol, atol=1e-30)
  # 1000 axes (bases up to 7919): in float64 the randomised axes whose
  # b^max_digits passes 2^63 add digit by digit, the others in one division
  g2, p2 = halton.sample(1000, num_results=16, seed=3, dtype=dtype)
  w2 = ohalton.sample(1000, num_results=16, perms=_np(p2.perms),
                      zero_correction=_np(p2.zero_correction), dtype=npd)
  np.testing.assert_allclose(_np(g2), w2, rtol=rtol, atol=1e-30)
  if dtype == torch.float64:
    # indices past 2^32 take the 64-bit digit division
    idx = np.array([2**32 - 2, 2**32 - 1, 2**33 + 5, 2**40 + 12345])
    g3, _ = halton.sample(40, sequence_indices=idx, randomized=False,
                          dtype=dtype)
    w3 = ohalton.sample(40, sequence_indices=idx, dtype=npd)
    np.testing.assert_allclose(_np(g3), w3, rtol=rtol, atol=1e-30)
  # the same seed gives the same points; params reproduce them by index
  again, _ = halton.sample(7, num_results=2000, seed=11, dtype=dtype)
  np.testing.assert_array_equal(_np(again), _np(got))
  part, _ = halton.sample(7, sequence_indices=np.arange(1000, 2000),
                          randomization_params=params, dtype=dtype)
  np.testing.assert_array_equal(_np(part), _np(got)[1000:])


def test_halton_full_size_rows_match_the_oracle():
  # halton.sample(100, 2^20) float32 (64-point tiles, one axis per wave):
  # rows at tile boundaries and spread over the range against the oracle
  from oracle import halton as ohalton
  from tqf.math.random import halton
  got, _ = halton.sample(100, num_results=2**20, randomized=False,
                         dtype=torch.float32)
  got = _np(got)
  rows = np.array([0, 1, 62, 63, 64, 65, 127, 128, 2**19 - 1, 2**19,
                   2**20 - 65, 2**20 - 64, 2**20 - 1] +
                  list(np.random.RandomState(8).randint(0, 2**20, 24)))
  want = ohalton.sample(100, sequence_indices=rows, dtype=np.float32)
  np.testing.assert_allclose(got[rows], want, rtol=1e-6, atol=1e-30)


def test_halton_randomized_integration_and_normal_draws():
  from oracle import halton as ohalton
  from tqf.math.random import halton
  # E[x1 x2^2 x3^3] over the unit cube = 1/24 (halton_test.py:139-154)
  est = []
  for seed in range(10):
    x, _ = halton.sample(3, num_results=1000, seed=seed,
                         dtype=torch.float64)
    x = _np(x)
    est.append(np.mean(x[:, 0] * x[:, 1]**2 * x[:, 2]**3))
  assert abs(np.mean(est) - 1 / 24) <= 2 * np.std(est) + 1e-4
  # Halton normal draws: sqrt(2) erfinv(2u - 1) of the points, quasi layout
  z = _np(mvn.multivariate_normal([500], mean=np.zeros(3),
                                  random_type=mvn.RandomType.HALTON,
                                  skip=10, dtype=torch.float64))
  u = ohalton.sample(3, sequence_indices=np.arange(10, 510),
                     dtype=np.float64)
  np.testing.assert_allclose(z, onormal.quasi_normals(u), rtol=1e-12,
                             atol=1e-12)


def test_euler_halton_draws_parity():
  from oracle import halton as ohalton
  n, steps = 1024, 16
  spec, fns, x0 = _models()['heston_ft']
  u = ohalton.sample(2 * steps, sequence_indices=np.arange(n),
                     dtype=np.float64)
  z = onormal.quasi_normals(u).reshape(n, steps, 2)
  want = oeuler.sample(2, fns[0], fns[1], np.asarray([1.0]), num_samples=n,
                       initial_state=np.asarray(x0), normal_draws=z,
                       num_time_steps=steps, dtype=np.float64)
  drift = engine.tag(lambda t, x: None, spec)
  got = tqf.models.euler_sampling.sample(
      2, drift, drift, np.asarray([1.0]), num_time_steps=steps,
      num_samples=n, initial_state=np.asarray(x0),
      random_type=mvn.RandomType.HALTON, dtype=torch.float64)
  _close(_np(got), want, 1e-11)


def test_qmc_scrambled_digital_net():
  nd, n = 12, 4096
  g = qmc.sobol_generating_matrices(5, n, nd)
  sm = qmc.random_scrambling_matrices(5, nd, seed=(2, 3))
  got = _np(qmc.digital_net_sample(g, n, nd, scrambling_matrices=sm,
                                   dtype=torch.float64))
  want = osobol.digital_net_sample_int(
      osobol.scramble_generating_matrices(g, sm, nd), n, nd) / 2.0**nd
  np.testing.assert_array_equal(got, want)


def test_heston_ft_fp64_engine_vs_semi_analytic():
  # the headline scheme in float64 (2^20 shifted-Sobol paths, the reference
  # grid of 365 steps) against the independent Lewis-integral price: the
  # difference is the Euler bias plus Monte Carlo error
  strikes = np.linspace(60, 140, 64)
  mats = (0.25, 0.5, 0.75, 1.0)
  res = pricing.heston_ft_calls(num_paths=2**20, num_time_steps=365,
                                maturities=mats, grid='reference',
                                random='sobol_shifted', seed=7,
                                dtype=torch.float64, s0=100.0, v0=0.04,
                                kappa=2.0, theta=0.04, xi=0.3, rho=-0.7,
                                r=0.03)
  p, se = _np(res['prices']), _np(res['stderrs'])
  sa = np.array([[oheston.european_call(100.0, k, t, 0.03, 0.04, 2.0, 0.04,
                                        0.3, -0.7) for k in strikes]
                 for t in mats])
  assert np.all(np.abs(p - sa) <= 3 * se + 4e-3), np.max(np.abs(p - sa))


def test_batch_float32_matches_float64_plain_kernel_on_the_same_streams():
  from tqf.engine import batch as batch_lib
  cs = batch_lib.random_contracts(64, seed=5)
  cs = [dict(c, n_paths=min(c['n_paths'], 8192)) for c in cs[:12]]
  plan = batch_lib.BatchPlan(cs, tile=1024, seed=11)
  out = _np(plan.launch())
  for i, c in enumerate(cs):
    n, t, r = c['n_steps'], c['maturity'], c['rate']
    grid = np.linspace(0.0, t, n + 1)
    keep = np.zeros(n + 1, bool)
    keep[-1] = True
    if c['kind'] == 0:
      spec = ('const', r - 0.5 * c['sigma']**2, c['sigma'])
      x0 = [np.log(c['s0'])]
    else:
      spec = ('heston', r, c['kappa'], c['theta'], c['xi'], c['rho'], 0.0)
      x0 = [np.log(c['s0']), c['v0']]
    ref = engine.run(spec, grid, keep, 1, x0, c['n_paths'],
                     engine.RngSpec('philox', seed=11, stream_id=i),
                     torch.float64, mode='payoff', strikes=[c['strike']],
                     discount=[np.exp(-r * t)])
    want = float(_np(ref['prices'])[0, 0])
    assert abs(out[i, 0] - want) <= 1e-3 * max(want, 1e-1), (i, out[i, 0],
                                                             want)
