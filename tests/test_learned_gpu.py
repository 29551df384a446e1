"""Device MLP (BF16 matrix-core forward / backward, Adam) against the CPU
oracle, and the loop-closure feature kernel against a numpy restatement."""
import numpy as np
import pytest

import mlp_oracle as M
from rpgo_amd import _capi, synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = _capi.Context(0)
    yield c
    c.close()


def _data(seed, n):
    rng = np.random.default_rng(seed)
    x = rng.normal(0, 1, (n, 16)).astype(np.float32)
    t = np.stack([rng.uniform(0, 1, n), rng.normal(0, 1, n)], 1).astype(np.float32)
    norm = np.concatenate([rng.normal(0, 0.2, 16), rng.uniform(0.5, 2, 16)]).astype(np.float32)
    return x, t, norm


@pytest.mark.parametrize("n", [1, 63, 64, 1000])
def test_forward_matches_oracle(ctx, n):
    x, _, norm = _data(n, n)
    p = M.init_params(n)
    ctx.mlp_set_params(p, norm)
    y = ctx.mlp_forward(x)
    yr, _ = M.forward(p, x, norm)
    err = np.abs(y - yr)
    print(f"[mlp] forward n={n} max|dy|={err.max():.3e} median={np.median(err):.3e}", flush=True)
    # FP32 accumulation order differs; a BF16 activation may round to the other side
    assert np.median(err) <= 1e-3 * (1 + np.median(np.abs(yr)))
    assert np.all(err <= 3e-2 * (1 + np.abs(yr)))


def test_gradients_match_oracle(ctx):
    x, t, norm = _data(7, 777)
    sw = np.random.default_rng(8).uniform(0.5, 1.5, 777).astype(np.float32)
    p = M.init_params(7)
    ctx.mlp_set_params(p, norm)
    g, loss = ctx.mlp_gradients(x, t, sw, 0.7)
    lr, gr = M.loss_and_grad(p, x, t, norm, sw, 0.7)
    le, ge = M.loss_and_grad(p.astype(np.float64), x, t, norm, sw, 0.7, exact=True)
    rel = np.linalg.norm(g - gr) / np.linalg.norm(gr)
    rel_exact = np.linalg.norm(g - ge) / np.linalg.norm(ge)
    print(f"[mlp] grad rel vs BF16 oracle {rel:.3e}, vs exact FP64 {rel_exact:.3e}; loss {loss:.6f} "
          f"oracle {lr:.6f}", flush=True)
    assert abs(loss - lr) <= 1e-3 * abs(lr)
    assert rel <= 1e-2
    assert rel_exact <= 5e-2
    # every parameter block carries gradient
    o = 0
    for size in (16 * 128, 128, 128 * 128, 128, 128 * 2, 2):
        blk = slice(o, o + size)
        assert np.linalg.norm(g[blk] - gr[blk]) <= 3e-2 * np.linalg.norm(gr[blk]) + 1e-7
        o += size


def test_training_reduces_loss(ctx):
    rng = np.random.default_rng(3)
    n = 8192
    x = rng.normal(0, 1, (n, 16)).astype(np.float32)
    logit = 2.0 * x[:, 0] - 1.5 * x[:, 3] + 0.5 * x[:, 5] * x[:, 6]
    t = np.stack([(logit > 0).astype(np.float32), 0.3 * x[:, 1] - 0.2]).T.astype(np.float32)
    norm = np.concatenate([np.zeros(16), np.ones(16)]).astype(np.float32)
    ctx.mlp_set_params(M.init_params(11), norm)
    hist = ctx.mlp_train(x, t, None, 1.0, batch=512, epochs=15, lr=2e-3, seed=5)
    y = ctx.mlp_forward(x)
    acc = np.mean((y[:, 0] > 0) == (t[:, 0] > 0.5))
    print(f"[mlp] loss {hist[0]:.4f} -> {hist[-1]:.4f}, accuracy {acc:.3f}", flush=True)
    assert hist[-1] < 0.5 * hist[0]
    assert acc > 0.95


def test_lc_features_match_numpy(ctx):
    import oracle as O
    import solver_oracle as S

    g = synth.se3_helix(n_poses=400, n_lc=60, outlier_frac=0.3, seed=9)
    ctx.pcm_upload(g.dim, g.pose0, g.odom, g.odom_cov, g.lc_from, g.lc_to, g.lc_pose, g.lc_cov)
    ctx.pcm_run(-1.0, 5.0, precision=_capi.PCM_FP64_REPLAY)
    got = ctx.pcm_download()
    k, idx, _ = ctx.max_clique(_capi.CLIQUE_HEU_REF)
    f = ctx.lc_features(idx[:k])
    acc = got["accepted"]
    m = len(acc)
    adj = _capi.unpack_bits(got["adj_bits"], m)
    traj = O.fold(g.dim, g.pose0, g.odom, g.odom_cov)
    member = np.zeros(m, bool)
    member[idx[:k]] = True
    deg = adj.sum(1)
    assert f.shape == (m, 16)
    assert np.allclose(f[:, 1], deg / max(m - 1, 1), atol=1e-6)
    assert np.allclose(f[:, 2], np.log1p(deg), atol=1e-5)
    assert np.allclose(f[:, 3], (adj & member[None, :]).sum(1) / max(k, 1), atol=1e-6)
    assert np.array_equal(f[:, 4] > 0.5, member)
    assert np.all(f[:, 6] == 0) and np.all(f[:, 7] == 0)  # no residuals given
    resid = np.random.default_rng(3).uniform(0, 50, m)
    fr = ctx.lc_features(idx[:k], resid, 2.0)
    assert np.allclose(fr[:, 6], np.log1p(resid / 2.0), rtol=1e-6)
    assert np.allclose(fr[:, 7], np.minimum(resid / 2.0, 10.0), rtol=1e-6)
    assert np.array_equal(np.delete(fr, [6, 7], 1), np.delete(f, [6, 7], 1))
    # chi^2 of each loop closure at the chain poses
    R = traj[:, :9].reshape(-1, 3, 3)
    t = traj[:, 9:12]
    for a in range(m):
        o = acc[a]
        i, j = g.lc_from[o], g.lc_to[o]
        Rij = R[i].T @ R[j]
        tij = R[i].T @ (t[j] - t[i])
        Rz = g.lc_pose[o, :9].reshape(3, 3)
        tz = g.lc_pose[o, 9:]
        E_R = Rz.T @ Rij
        E_t = Rz.T @ (tij - tz)
        e = S.se3_log(E_R, E_t)
        cov = g.lc_cov[o].reshape(6, 6)
        chi2 = e @ np.linalg.solve(cov, e)
        assert abs(f[a, 5] - np.log1p(chi2)) <= 1e-4 * (1 + np.log1p(chi2)), (a, f[a, 5], np.log1p(chi2))
        assert abs(f[a, 10] - abs(i - j) / g.n_poses) <= 1e-6
