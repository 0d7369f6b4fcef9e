"""CPU oracle of the moment surrogate's network: forward, backward, loss and
Adam of the MLP ensemble in NumPy, with the GPU's BF16 operand rounding
restated exactly.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and
``bench.py``'s CPU baseline leg may import this module, and only as the
checker; the product package ``kwant_amd`` never imports it.

The surrogate has no counterpart in the reference (kwant/kpm.py has no
learned part), so this oracle restates textbook definitions:
  * BF16 = the upper 16 bits of an IEEE binary32 after round-to-nearest-even
    on the lower 16 (``bf16``); the three-term split x = hi + lo with
    hi = bf16(x), lo = bf16(x - hi) keeps hi*hi + hi*lo + lo*hi;
  * GELU(x) = x Phi(x) with the exact normal CDF (erf form) and
    GELU'(x) = Phi(x) + x phi(x);
  * the loss is the mean squared error per member, L = sum (y - t)^2 / (B O);
  * Adam is Kingma & Ba (2015), Algorithm 1, bias-corrected.

Pinning (so that the oracle is not its own reference), in
tests/test_mlp_oracle.py:
  * ``bf16``: torch's own float32 -> bfloat16 conversion;
  * ``gelu`` / ``gelu_grad``: torch.nn.functional.gelu and its autograd;
  * ``backward``: central finite differences of ``loss`` in float64 (without
    rounding), and torch autograd of the same network;
  * ``adam``: its first step equals lr * g / (|g| + eps) elementwise (closed
    form of a bias-corrected first step), and torch.optim.Adam over steps.
"""
import numpy as np
from scipy.special import erf


def bf16(x):
    """Round float32 values to bfloat16 (round to nearest, ties to even);
    returns float32 arrays holding the rounded values."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def split3(x):
    """(hi, lo) with hi = bf16(x), lo = bf16(x - hi), x float32."""
    x = np.asarray(x, dtype=np.float32)
    hi = bf16(x)
    lo = bf16(x - hi)
    return hi, lo


def gemm(a, b, split=False, rounding=True):
    """a (M, K) @ b (N, K).T with the GPU's operand rounding, in float64.
    rounding=False: plain float64 product (the unrounded network)."""
    if not rounding:
        return np.asarray(a, np.float64) @ np.asarray(b, np.float64).T
    if split:
        ah, al = split3(a)
        bh, bl = split3(b)
        f = np.float64
        return (ah.astype(f) @ bh.astype(f).T + ah.astype(f) @ bl.astype(f).T
                + al.astype(f) @ bh.astype(f).T)
    return bf16(a).astype(np.float64) @ bf16(b).astype(np.float64).T


def gelu(x):
    return 0.5 * x * (1.0 + erf(x / np.sqrt(2.0)))


def gelu_grad(x):
    return 0.5 * (1.0 + erf(x / np.sqrt(2.0))) + x * np.exp(-0.5 * x * x) / np.sqrt(2 * np.pi)


def forward(weights, biases, x, split=None, rounding=True):
    """One member: weights[l] (out, in), biases[l] (out,), x (B, in).
    Returns (pre-activations z_l, activations a_l, output y).  With rounding,
    activations are held in float32 between layers as on the GPU."""
    nl = len(weights)
    split = [False] * nl if split is None else split
    dt = np.float32 if rounding else np.float64
    act = np.asarray(x, dt)
    zs, acts = [], [act]
    for l in range(nl):
        z = gemm(act, np.asarray(weights[l], dt), split[l], rounding) + biases[l]
        z = z.astype(dt)
        zs.append(z)
        if l < nl - 1:
            act = gelu(z.astype(np.float64)).astype(dt)
            acts.append(act)
    return zs, acts, zs[-1]


def mse(y, t):
    """(loss, dL/dy) of the mean squared error over all entries."""
    e = np.asarray(y, np.float64) - np.asarray(t, np.float64)
    return float(np.mean(e * e)), 2.0 * e / e.size


def backward(weights, zs, acts, dy, rounding=True):
    """Gradients (dW_l, db_l) for dL/dy = dy, reverse mode through forward's
    cache.  With rounding the GEMM operands are BF16 as on the GPU (no split:
    the backward pass runs plain BF16 products)."""
    nl = len(weights)
    dt = np.float32 if rounding else np.float64
    g = np.asarray(dy, dt)
    dws, dbs = [None] * nl, [None] * nl
    for l in range(nl - 1, -1, -1):
        dws[l] = gemm(g.T, np.asarray(acts[l], dt).T, False, rounding).astype(dt)
        dbs[l] = g.astype(np.float64).sum(axis=0).astype(dt)
        if l > 0:
            gp = gemm(g, np.asarray(weights[l], dt).T, False, rounding)
            g = (gp * gelu_grad(np.asarray(zs[l - 1], np.float64))).astype(dt)
    return dws, dbs


def loss(weights, biases, x, t, rounding=False):
    return mse(forward(weights, biases, x, rounding=rounding)[2], t)[0]


def adam(p, g, m, v, lr, b1, b2, eps, t):
    """One bias-corrected Adam step at step count t >= 1; returns (p, m, v)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v


def dos_l1(d_a, d_b, nodes=None):
    """L1 distance of two KPM densities from their damped moments, by
    direct evaluation of rho(x) = (d_0 + 2 sum d_n T_n(x)) / (pi sqrt(1-x^2))
    on a fine uniform grid in theta (x = cos theta, dx = sin theta dtheta)."""
    d = np.atleast_2d(np.asarray(d_a, np.float64) - np.asarray(d_b, np.float64))
    k = 8 * d.shape[-1] if nodes is None else nodes
    theta = (np.arange(k) + 0.5) * np.pi / k
    n = np.arange(d.shape[-1])
    series = d[:, 0][None, :] + 2 * np.cos(np.outer(theta, n[1:])) @ d[:, 1:].T
    # int |rho| dx = int_0^pi |series(theta)| / pi dtheta (midpoint rule)
    return np.abs(series).mean(axis=0)
