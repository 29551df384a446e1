"""CPU checks of the MLP oracle (the learned initialisation has no reference
counterpart, so the oracle is pinned by mathematics): its analytic gradient
matches central finite differences of its loss in exact FP64, and its BF16
rounding is round-to-nearest-even on known values."""
import numpy as np

import mlp_oracle as M


def test_bf16_rounding_known_values():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -2.5, 0.0], np.float32)
    got = M.bf16(x)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: even mantissa (1.0) wins;
    # 1 + 3*2^-8 ties between 1 + 2^-7 and 1 + 2^-6: even (1 + 2^-6) wins
    want = np.array([1.0, 1.0, 1.0 + 2 ** -6, 1.0, -2.5, 0.0], np.float32)
    assert np.array_equal(got, want)


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(1)
    p = M.init_params(3).astype(np.float64)
    n = 37
    x = rng.normal(0, 1, (n, 16))
    t = np.stack([rng.uniform(0, 1, n), rng.normal(0, 1, n)], 1)
    norm = np.concatenate([rng.normal(0, 0.1, 16), rng.uniform(0.5, 2, 16)])
    sw = rng.uniform(0.5, 1.5, n)
    _, g = M.loss_and_grad(p, x, t, norm, sw, 0.7, exact=True)
    h = 1e-6
    for i in rng.choice(M.NPARAMS, 60, replace=False):
        e = np.zeros(M.NPARAMS)
        e[i] = h
        lp, _ = M.loss_and_grad(p + e, x, t, norm, sw, 0.7, exact=True)
        lm, _ = M.loss_and_grad(p - e, x, t, norm, sw, 0.7, exact=True)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-5 * (abs(fd) + abs(g[i])) + 1e-10, (i, fd, g[i])


def test_bf16_path_close_to_exact():
    rng = np.random.default_rng(2)
    p = M.init_params(4)
    x = rng.normal(0, 1, (200, 16))
    t = np.stack([rng.integers(0, 2, 200), rng.normal(0, 1, 200)], 1)
    norm = np.concatenate([np.zeros(16), np.ones(16)])
    _, ge = M.loss_and_grad(p.astype(np.float64), x, t, norm, exact=True)
    _, gb = M.loss_and_grad(p, x, t, norm)
    assert np.linalg.norm(gb - ge) <= 0.05 * np.linalg.norm(ge)
