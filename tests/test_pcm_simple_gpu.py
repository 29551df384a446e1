"""GPU parity of the PcmSimple stages (rpgo_pcm_simple_run) against the FP64
oracle (orc_pcm_simple_adjacency).

Bar: accepted set and adjacency bits bit-exact; per-node translation norms
within 1e-12 relative (device and host transcendentals may differ in the last
ulp).  Thresholds are placed between observed norms so that no decision sits
within rounding of its threshold.  Runs through the C ABI (rpgo_amd._capi).
"""
import math

import numpy as np
import pytest

import oracle as O
from rpgo_amd import _capi, synth

pytestmark = pytest.mark.gpu

DIST_RTOL = 1e-12


@pytest.fixture(scope="module")
def ctx():
    c = _capi.Context(0)
    yield c
    c.close()


def _gap_threshold(values, q):
    """A threshold near quantile q, halfway between two distinct observed values."""
    v = np.unique(np.asarray(values, float))
    if len(v) < 2:
        return float(v[0]) * 2.0 + 1.0 if len(v) else 1.0
    k = min(max(int(q * (len(v) - 1)), 0), len(v) - 2)
    return float(0.5 * (v[k] + v[k + 1]))


def _oracle_norms(g, traj):
    pl = 12 if g.dim == 3 else 4
    d = 6 if g.dim == 3 else 3
    t = np.zeros(g.n_lc)
    r = np.zeros(g.n_lc)
    for k in range(g.n_lc):
        t[k], r[k] = O.simple_odom(g.dim, traj, g.lc_from[k], g.lc_to[k], g.lc_pose.reshape(-1, pl)[k],
                                   g.lc_cov.reshape(-1, d * d)[k])
    return t, r


def _pair_norms(g, traj, acc, limit=4000):
    pl = 12 if g.dim == 3 else 4
    d = 6 if g.dim == 3 else 3
    P = g.lc_pose.reshape(-1, pl)
    C = g.lc_cov.reshape(-1, d * d)
    rng = np.random.default_rng(1)
    ts, rs = [], []
    m = len(acc)
    for _ in range(min(limit, m * (m - 1) // 2)):
        i, j = sorted(rng.choice(m, 2, replace=False))
        a, c = acc[i], acc[j]
        t, r = O.simple_pair(g.dim, traj, g.lc_from[a], g.lc_to[a], P[a], C[a], g.lc_from[c],
                             g.lc_to[c], P[c], C[c])
        ts.append(t)
        rs.append(r)
    return np.array(ts), np.array(rs)


def _run_both(ctx, g, thr):
    print(f"[pcm_simple] kind={g.meta.get('kind')} N={g.n_poses} n_lc={g.n_lc} thr={thr}", flush=True)
    ctx.pcm_upload(g.dim, g.pose0, g.odom, g.odom_cov, g.lc_from, g.lc_to, g.lc_pose, g.lc_cov)
    st = ctx.pcm_simple_run(*thr, want_dist=True)
    got = ctx.pcm_download(want_dist=True)
    traj = O.fold(g.dim, g.pose0, g.odom, g.odom_cov)
    acc, od, adj, dist = O.pcm_simple_adjacency(g.dim, traj, g.lc_from, g.lc_to, g.lc_pose, g.lc_cov,
                                                *thr)
    return st, got, dict(accepted=acc, odom_dist=od, adj=adj, dist=dist)


def _assert_parity(st, got, ref, loop_check=True):
    assert np.array_equal(got["accepted"], ref["accepted"])
    m = len(ref["accepted"])
    assert st.n_accepted == m
    np.testing.assert_allclose(got["odom_dist"], ref["odom_dist"], rtol=DIST_RTOL, atol=0)
    adj = _capi.unpack_bits(got["adj_bits"], m) if m else np.zeros((0, 0), np.uint8)
    assert np.array_equal(adj, ref["adj"]), np.argwhere(adj != ref["adj"])[:10]
    if m and loop_check:
        np.testing.assert_allclose(got["dist"], ref["dist"], rtol=DIST_RTOL, atol=1e-300)
    assert st.n_consistent == int(ref["adj"].sum()) // 2


def _thresholds(g, q_odom=0.7, q_pair=0.6):
    traj = O.fold(g.dim, g.pose0, g.odom, g.odom_cov)
    t, r = _oracle_norms(g, traj)
    ot, orr = _gap_threshold(t, q_odom), _gap_threshold(r, q_odom)
    acc = np.nonzero((t < ot) & (r < orr))[0]
    pt, pr = _pair_norms(g, traj, acc)
    return (ot, orr, _gap_threshold(pt, q_pair), _gap_threshold(pr, q_pair))


def test_se2_random_walk(ctx):
    g, _, _ = synth.workload("se2_500", seed=0)
    thr = _thresholds(g)
    st, got, ref = _run_both(ctx, g, thr)
    assert 0 < len(ref["accepted"]) < g.n_lc
    assert 0 < ref["adj"].sum() < len(ref["accepted"]) ** 2
    _assert_parity(st, got, ref)


@pytest.mark.parametrize("seed", [0, 1])
def test_se3_helix(ctx, seed):
    g = synth.se3_helix(n_poses=2000, n_lc=400, outlier_frac=0.3, seed=seed)
    thr = _thresholds(g)
    st, got, ref = _run_both(ctx, g, thr)
    assert 0 < len(ref["accepted"]) < g.n_lc
    assert ref["adj"].sum() > 0
    _assert_parity(st, got, ref)


def test_se3_5k_full_size(ctx):
    g, _, _ = synth.workload("se3_5k", seed=0)
    thr = _thresholds(g, q_odom=0.9, q_pair=0.5)
    st, got, ref = _run_both(ctx, g, thr)
    assert len(ref["accepted"]) > 500
    _assert_parity(st, got, ref)
    k, idx, _ = ctx.max_clique(_capi.CLIQUE_HEU_REF)
    k_ref, idx_ref = O.max_clique_heu(ref["adj"])
    assert k == k_ref and np.array_equal(idx, idx_ref[:k])


def test_checks_disabled(ctx):
    g = synth.se3_helix(n_poses=600, n_lc=120, outlier_frac=0.3, seed=5)
    # odometry check off: every loop closure accepted; pair check on
    thr = (-1.0, 0.1, 0.05, 0.01)
    st, got, ref = _run_both(ctx, g, thr)
    assert len(ref["accepted"]) == g.n_lc
    _assert_parity(st, got, ref)
    # pair check off: empty adjacency
    thr = (0.5, 0.05, 0.05, -1.0)
    st, got, ref = _run_both(ctx, g, thr)
    assert ref["adj"].sum() == 0
    _assert_parity(st, got, ref, loop_check=False)


def test_nan_rotation_covariance(ctx):
    """Loop closures and one odometry factor with NaN rotation variances: the
    rotation norm drops to 0 on every path that touches them."""
    g = synth.se3_helix(n_poses=800, n_lc=160, outlier_frac=0.3, seed=7)
    lc_cov = g.lc_cov.reshape(-1, 36).copy()
    lc_cov[::3, 0] = np.nan
    odom_cov = g.odom_cov.reshape(-1, 36).copy()
    odom_cov[400, 7] = np.nan
    g2 = synth.PoseGraph(**{**g.__dict__, "lc_cov": lc_cov, "odom_cov": odom_cov})
    thr = _thresholds(g2)
    st, got, ref = _run_both(ctx, g2, thr)
    _assert_parity(st, got, ref)


def _ref_chain(odoms):
    I3 = np.eye(3)
    pose0 = np.concatenate([I3.reshape(9), np.zeros(3)])
    odom = np.array([np.concatenate([np.asarray(R, float).reshape(9), t]) for R, t in odoms])
    cov = np.tile((0.1 * np.eye(6)).reshape(36), (len(odoms), 1))
    return pose0, odom, cov


def _rz(a):
    c, s = math.cos(a), math.sin(a)
    return np.array([[c, -s, 0], [s, c, 0], [0, 0, 1.0]])


def test_reference_consistency_checks(ctx):
    """testPcmSimple.cpp:175-400 through the C ABI: LC1..LC3 form the clique, LC4
    stays out, for the translation and the rotation threshold."""
    R90 = np.array([[0, -1, 0], [1, 0, 0], [0, 0, 1.0]])
    R180 = np.diag([-1.0, -1.0, 1.0])
    pose0, odom, cov = _ref_chain([(R90, (1, 0, 0))] * 2 + [(np.eye(3), (1, 0, 0))] * 4)
    cases = [((-1.0, -1.0, 0.05, 100.0),
              [(R180, (0, 0.9, 0)), (_rz(3.1416), (-0.9, 0.9, 0)), (_rz(3.1416), (-1.9, 0.8, 0)),
               (_rz(3.1416), (-2.8, 0.75, 0))]),
             ((-1.0, -1.0, 100.0, 0.005),
              [(R180, (0, 1.0, 0)), (_rz(3.141), (-1.0, 1.0, 0)), (_rz(3.13), (-2.0, 1.0, 0)),
               (_rz(3.12), (-3.0, 1.0, 0))])]
    for thr, lcs in cases:
        for m, want in ((2, 2), (3, 3), (4, 3)):
            lp = np.array([np.concatenate([R.reshape(9), t]) for R, t in lcs[:m]])
            lcov = np.tile((0.1 * np.eye(6)).reshape(36), (m, 1))
            ctx.pcm_upload(3, pose0, odom, cov, list(range(3, 3 + m)), [0] * m, lp, lcov)
            st = ctx.pcm_simple_run(*thr)
            assert st.n_accepted == m
            k, idx, _ = ctx.max_clique(_capi.CLIQUE_HEU_REF)
            assert k == want, (thr, m, k)
            if m == 4:
                assert 3 not in set(idx.tolist())


def test_reference_odometry_checks(ctx):
    """testPcmSimple.cpp:18-172: odometry translation 0.05 / 0.052 against 0.051,
    rotation 0.0049 / 0.0052 against 0.005."""
    R90 = np.array([[0, -1, 0], [1, 0, 0], [0, 0, 1.0]])
    pose0, odom, cov = _ref_chain([(R90, (1, 0, 0))] * 3)
    for thr, lcs in [((0.051, 100.0, -1.0, -1.0), [(R90, (0.8, 0, 0)), (R90, (0.792, 0, 0))]),
                     ((100.0, 0.005, -1.0, -1.0), [(_rz(1.551), (1, 0, 0)), (_rz(1.55), (1, 0, 0))])]:
        lp = np.array([np.concatenate([R.reshape(9), t]) for R, t in lcs])
        lcov = np.tile((0.1 * np.eye(6)).reshape(36), (2, 1))
        ctx.pcm_upload(3, pose0, odom, cov, [3, 3], [0, 0], lp, lcov)
        st = ctx.pcm_simple_run(*thr)
        assert st.n_accepted == 1
        assert list(ctx.pcm_download()["accepted"]) == [0]
