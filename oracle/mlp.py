"""Oracle: the 6 -> 64 -> 64 -> 1 SiLU network of the learned
continuation value, forward and backward in numpy.

There is no counterpart in the reference (its continuation value is the
polynomial regression of tff/models/longstaff_schwartz/lsm.py:328-383).
The restatement is pinned by central finite differences of its own loss
(tests/test_oracle.py), which a wrong transpose, sign or chain-rule factor
fails.  `bf=True` rounds the operands of every matrix product the MFMA
kernels run in bfloat16 -- layer 1 (inputs and W1), layer 2 (H1 and W2),
dW1 (inputs and dZ1), dW2 and dH1 (dZ2) -- so GPU gradients can be compared
tightly; biases, activations and the 64 -> 1 layer stay float.

Parameter vector layout: W1[6][64] | b1[64] | W2[64][64] | b2[64] | w3[64] | b3.
"""
import numpy as np

N_IN, N_H = 6, 64
N_PARAMS = N_IN * N_H + N_H + N_H * N_H + N_H + N_H + 1


def unpack(p):
  p = np.asarray(p, np.float64)
  o = 0
  w1 = p[o:o + N_IN * N_H].reshape(N_IN, N_H); o += N_IN * N_H
  b1 = p[o:o + N_H]; o += N_H
  w2 = p[o:o + N_H * N_H].reshape(N_H, N_H); o += N_H * N_H
  b2 = p[o:o + N_H]; o += N_H
  w3 = p[o:o + N_H]; o += N_H
  b3 = p[o]
  return w1, b1, w2, b2, w3, b3


def pack(w1, b1, w2, b2, w3, b3):
  return np.concatenate([w1.ravel(), b1, w2.ravel(), b2, w3, [b3]])


def bf16(x):
  """Round to bfloat16 (nearest even) and back."""
  a = np.asarray(x, np.float32)
  u = a.view(np.uint32).astype(np.uint64)
  u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
  return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def silu(z):
  return z / (1 + np.exp(-z))


def silu_grad(z):
  s = 1 / (1 + np.exp(-z))
  return s * (1 + z * (1 - s))


def init_params(seed=0):
  rs = np.random.RandomState(seed)
  w1 = rs.standard_normal((N_IN, N_H)) * np.sqrt(2.0 / N_IN)
  w2 = rs.standard_normal((N_H, N_H)) * np.sqrt(2.0 / N_H)
  w3 = rs.standard_normal(N_H) * np.sqrt(1.0 / N_H)
  return pack(w1, np.zeros(N_H), w2, np.zeros(N_H), w3, 0.0)


def forward(p, x, bf=False):
  w1, b1, w2, b2, w3, b3 = unpack(p)
  xin = x[:, :N_IN]
  z1 = (bf16(xin) @ bf16(w1) if bf else xin @ w1) + b1
  h1 = silu(z1)
  w2e = bf16(w2) if bf else w2
  h1e = bf16(h1) if bf else h1
  z2 = h1e @ w2e + b2
  h2 = silu(z2)
  y = h2 @ w3 + b3
  return y, (z1, h1e, w2e, z2, h2)


def loss_and_grad(p, x, t, w, inv_norm, bf=False):
  """L = sum w (y - t)^2 * inv_norm and dL/dp."""
  w1, b1, w2, b2, w3, b3 = unpack(p)
  y, (z1, h1e, w2e, z2, h2) = forward(p, x, bf)
  e = y - t
  loss = np.sum(w * e * e) * inv_norm
  dy = 2 * w * e * inv_norm
  dw3 = h2.T @ dy
  db3 = dy.sum()
  dz2 = np.outer(dy, w3) * silu_grad(z2)
  dz2e = bf16(dz2) if bf else dz2
  dw2 = h1e.T @ dz2e
  db2 = dz2.sum(0)
  dh1 = dz2e @ w2e.T
  dz1 = dh1 * silu_grad(z1)
  xin = x[:, :N_IN]
  dw1 = bf16(xin).T @ bf16(dz1) if bf else xin.T @ dz1
  db1 = dz1.sum(0)
  return loss, pack(dw1, db1, dw2, db2, dw3, db3)


def adam(p, m, v, g, step, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
  m = b1 * m + (1 - b1) * g
  v = b2 * v + (1 - b2) * g * g
  mh = m / (1 - b1**step)
  vh = v / (1 - b2**step)
  return p - lr * mh / (np.sqrt(vh) + eps), m, v
