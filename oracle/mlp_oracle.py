"""CPU oracle of the learned GNC initialisation (test infrastructure only).

A numpy restatement of the device MLP (kimera-rpgo_amd/csrc/learned.hip):
16 -> 128 -> 128 -> 2, ReLU, BF16 operands (round to nearest even) at every
matrix product with FP32 accumulation, the same loss (BCE-with-logits on the
weight target, mu_w/2 squared error on log mu, per-row sample weights, mean
over rows) and the same gradient flow (the ReLU masks read from the BF16
activations).  The learned part has no counterpart in the reference, so it is
pinned here by mathematics: tests compare the analytic gradient of this oracle
with central finite differences of its own loss in FP64 without BF16 rounding
(`exact=True`), and the device gradients with this oracle.

Only tests/ may import this module.
"""
from __future__ import annotations

import numpy as np

F, H, O = 16, 128, 2
NPARAMS = F * H + H + H * H + H + H * O + O


def unpack(p):
    p = np.asarray(p)
    if p.dtype not in (np.float32, np.float64):
        p = p.astype(np.float64)
    o = 0
    W1 = p[o:o + F * H].reshape(F, H); o += F * H
    b1 = p[o:o + H]; o += H
    W2 = p[o:o + H * H].reshape(H, H); o += H * H
    b2 = p[o:o + H]; o += H
    W3 = p[o:o + H * O].reshape(H, O); o += H * O
    b3 = p[o:o + O]
    return W1, b1, W2, b2, W3, b3


def pack(W1, b1, W2, b2, W3, b3):
    return np.concatenate([W1.ravel(), b1, W2.ravel(), b2, W3.ravel(), b3]).astype(np.float32)


def init_params(seed=0):
    """He-uniform weights, zero biases."""
    rng = np.random.default_rng(seed)
    W1 = rng.uniform(-1, 1, (F, H)) * np.sqrt(6.0 / F)
    W2 = rng.uniform(-1, 1, (H, H)) * np.sqrt(6.0 / H)
    W3 = rng.uniform(-1, 1, (H, O)) * np.sqrt(6.0 / H)
    return pack(W1, np.zeros(H), W2, np.zeros(H), W3, np.zeros(O))


def bf16(x):
    """Round float32 to bfloat16 (nearest even), returned as float32."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def _q(exact):
    """Rounding of a matrix-product operand: BF16 on the device path, none in
    the exact FP64 reference used by the finite-difference check."""
    if exact:
        return lambda a: np.asarray(a, np.float64)
    return lambda a: bf16(np.asarray(a, np.float32))


def forward(p, x, norm, exact=False):
    W1, b1, W2, b2, W3, b3 = unpack(p)
    q = _q(exact)
    dt = np.float64 if exact else np.float32
    mean, inv = np.asarray(norm[:F], dt), np.asarray(norm[F:], dt)
    xb = q((np.asarray(x, dt) - mean) * inv)
    h1 = q(np.maximum(xb @ q(W1) + b1.astype(dt), 0))
    h2 = q(np.maximum(h1 @ q(W2) + b2.astype(dt), 0))
    y = h2 @ q(W3) + b3.astype(dt)
    return y, (xb, h1, h2)


def loss_and_grad(p, x, target, norm, sample_w=None, mu_weight=1.0, exact=False):
    """Mean loss and its gradient with respect to the packed parameters."""
    W1, b1, W2, b2, W3, b3 = unpack(p)
    q = _q(exact)
    dt = np.float64 if exact else np.float32
    y, (xb, h1, h2) = forward(p, x, norm, exact)
    n = y.shape[0]
    t = np.asarray(target, dt)
    sw = np.ones(n, dt) if sample_w is None else np.asarray(sample_w, dt)
    z, m = y[:, 0], y[:, 1]
    sp = np.where(z > 0, z + np.log1p(np.exp(-np.abs(z))), np.log1p(np.exp(np.minimum(z, 0))))
    sig = 1.0 / (1.0 + np.exp(-z))
    loss = float(np.sum(sw * (sp - t[:, 0] * z) + sw * 0.5 * mu_weight * (m - t[:, 1]) ** 2) / n)
    dY = np.stack([sw * (sig - t[:, 0]) / n, sw * mu_weight * (m - t[:, 1]) / n], 1).astype(dt)
    dH2 = q((dY @ q(W3).T) * (h2 > 0))
    dH1 = q((dH2 @ q(W2).T) * (h1 > 0))
    g = np.concatenate([(xb.T @ dH1).ravel(), dH1.sum(0), (h1.T @ dH2).ravel(), dH2.sum(0),
                        (h2.T @ q(dY)).ravel(), dY.sum(0)])
    return loss, g
