"""GPU parity of landmark observation groups (rpgo_pcm_problem.landmark)
against the FP64 oracle (orc_pcm_adjacency_landmark).

Bar: accepted set (all observations), adjacency bits bit-exact; distances
within 1e-12 relative.  Cases: the scenario of tests/testLandmark.cpp:22-345
(PCM and PcmSimple), seeded landmarks seen from a helix with outlying
observations, and the upload checks.
"""
import numpy as np
import pytest

import oracle as O
from rpgo_amd import _capi, synth

pytestmark = pytest.mark.gpu

DIST_RTOL = 1e-12
I3 = np.eye(3)
R90 = np.array([[0, -1, 0], [1, 0, 0], [0, 0, 1.0]])


@pytest.fixture(scope="module")
def ctx():
    c = _capi.Context(0)
    yield c
    c.close()


def pose3(R=I3, t=(0, 0, 0)):
    return np.concatenate([np.asarray(R, float).reshape(9), np.asarray(t, float)])


def landmark_cov(n, var_t=0.04):
    """Diagonal::Precisions((0, 0, 0, p, p, p)) covariance: NaN rotation block."""
    c = np.full((6, 6), np.nan)
    c[3:, 3:] = var_t * np.eye(3)
    return np.tile(c.reshape(36), (n, 1))


def _check(ctx, pose0, odom, ocov, frm, zs, cov, thr, simple=None):
    print(f"[landmark] n_obs={len(frm)} thr={thr} simple={simple}", flush=True)
    ctx.pcm_upload(3, pose0, odom, ocov, frm, None, zs, cov, landmark=True)
    if simple is None:
        st = ctx.pcm_run(-1.0, thr, precision=_capi.PCM_FP32_RECHECK, want_dist=True)
    else:
        st = ctx.pcm_simple_run(-1.0, -1.0, *simple, want_dist=True)
    got = ctx.pcm_download(want_dist=True)
    traj = O.fold(3, pose0, odom, ocov)
    assert np.array_equal(ctx.pcm_traj(), traj)
    adj, dist = O.pcm_adjacency_landmark(3, traj, frm, zs, cov, thr, simple=simple)
    m = len(frm)
    assert np.array_equal(got["accepted"], np.arange(m))
    assert np.all(got["odom_dist"] == 0.0)
    A = _capi.unpack_bits(got["adj_bits"], m)
    assert np.array_equal(A, adj), np.argwhere(A != adj)[:10]
    np.testing.assert_allclose(got["dist"], dist, rtol=DIST_RTOL, atol=1e-300)
    assert st.n_consistent == int(adj.sum()) // 2
    return adj


def test_reference_landmark_scenario(ctx):
    odom = np.stack([pose3(I3, (1, 0, 0))] * 2 + [pose3(R90, (1, 0, 0))] * 3)
    ocov = np.tile((0.1 * np.eye(6)).reshape(36), (5, 1))
    l0 = ([1, 5, 4], np.stack([pose3(I3, (0, 1, 0)), pose3(I3, (0, -1, 0)), pose3(I3, (1, 0, 0))]))
    l1 = ([2, 5], np.stack([pose3(I3, (0, -1, 0)), pose3(I3, (2, 0, 0))]))
    for simple, thr in ((None, 2.5), ((0.3, 0.05), 0.0)):
        adj = _check(ctx, pose3(), odom, ocov, l0[0], l0[1], landmark_cov(3), thr, simple)
        k, idx, _ = ctx.max_clique(_capi.CLIQUE_HEU_REF)
        assert k == 2 and sorted(idx.tolist()) == [0, 1]
        adj = _check(ctx, pose3(), odom, ocov, l1[0], l1[1], landmark_cov(2), thr, simple)
        assert adj.sum() == 2


def _observations(g, n_obs, outlier_frac, seed):
    """One landmark seen from n_obs helix poses: true relative position with
    2 cm noise, or a random one for the outliers."""
    rng = np.random.default_rng(seed)
    Rg, tg = synth.se3_unpack(g.gt)
    lw = tg[g.n_poses // 2] + rng.normal(0, 3, 3)
    frm = np.sort(rng.choice(g.n_poses, n_obs, replace=False))
    rng.shuffle(frm)
    zs = []
    for k in frm:
        p = Rg[k].T @ (lw - tg[k])
        if rng.uniform() < outlier_frac:
            p = p + rng.normal(0, 5, 3)
        else:
            p = p + rng.normal(0, 0.02, 3)
        zs.append(pose3(I3, p))
    return frm, np.stack(zs)


@pytest.mark.parametrize("seed", [0, 1])
def test_random_landmarks(ctx, seed):
    g = synth.se3_helix(n_poses=600, n_lc=10, outlier_frac=0.0, seed=seed)
    frm, zs = _observations(g, 120, 0.25, seed)
    adj = _check(ctx, g.pose0, g.odom, g.odom_cov, frm, zs, landmark_cov(len(frm), 0.05 ** 2), 5.0)
    assert 0 < adj.sum() < len(frm) ** 2 - len(frm)
    k, idx, _ = ctx.max_clique(_capi.CLIQUE_HEU_REF)
    k_ref, idx_ref = O.max_clique_heu(adj)
    assert k == k_ref and np.array_equal(idx, idx_ref[:k])
    # full rotation information on half of the observations
    cov = landmark_cov(len(frm), 0.05 ** 2)
    cov[::2] = np.diag([1e-3] * 3 + [0.05 ** 2] * 3).reshape(36)
    _check(ctx, g.pose0, g.odom, g.odom_cov, frm, zs, cov, 5.0)
    _check(ctx, g.pose0, g.odom, g.odom_cov, frm, zs, cov, 0.0, simple=(0.05, 0.01))


def test_landmark_upload_checks(ctx):
    g = synth.se3_helix(n_poses=200, n_lc=10, outlier_frac=0.0, seed=3)
    frm, zs = _observations(g, 10, 0.0, 3)
    keep = [k for k in range(g.n_poses - 1) if k != 99]
    with pytest.raises(RuntimeError):  # one trajectory only
        ctx.pcm_upload(3, np.stack([g.gt[0], g.gt[100]]), g.odom[keep], g.odom_cov[keep], frm, None, zs,
                       landmark_cov(10), chain_off=[0, 100, 200], landmark=True)
    ctx.pcm_upload(3, g.pose0, g.odom, g.odom_cov, frm, None, zs, landmark_cov(10), landmark=True)
    ctx.pcm_run(-1.0, 5.0)
    with pytest.raises(RuntimeError):  # not streamed
        ctx.pcm_append(-1.0, 5.0, odom=g.odom[:1], odom_cov=g.odom_cov[:1])
