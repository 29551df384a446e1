"""GPU parity of the PCM adjacency stage against the FP64 oracle.

Bar: the fold, odometry decisions, accepted set and adjacency bits are
bit-exact; distances of the FP64 replay agree with the oracle to 1e-12
relative (transcendentals may differ in the last ulp between the device
math library and glibc).  Runs through the C ABI (rpgo_amd._capi).
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


def _run_both(ctx, g, odom_thr, lc_thr, precision=_capi.PCM_FP64_REPLAY):
    print(f"[pcm] kind={g.meta.get('kind')} N={g.n_poses} n_lc={g.n_lc} thr=({odom_thr},{lc_thr}) "
          f"prec={precision}", flush=True)
    ctx.pcm_upload(g.dim, g.pose0, g.odom, g.odom_cov, g.lc_from, g.lc_to, g.lc_pose, g.lc_cov)
    st = ctx.pcm_run(odom_thr, lc_thr, precision=precision, want_dist=True)
    got = ctx.pcm_download(want_dist=True)
    traj_ref = O.fold(g.dim, g.pose0, g.odom, g.odom_cov)
    acc, od, adj, dist = O.pcm_adjacency(g.dim, traj_ref, g.lc_from, g.lc_to, g.lc_pose, g.lc_cov,
                                         odom_thr, lc_thr)
    return st, got, dict(traj=traj_ref, accepted=acc, odom_dist=od, adj=adj, dist=dist)


def _assert_parity(ctx, st, got, ref, check_traj=True):
    if check_traj:
        traj = ctx.pcm_traj()
        assert np.array_equal(traj, ref["traj"]), np.abs(traj - ref["traj"]).max()
    assert np.array_equal(got["accepted"], ref["accepted"])
    od_ref = ref["odom_dist"]
    np.testing.assert_allclose(got["odom_dist"], od_ref, rtol=DIST_RTOL, atol=0)
    m = len(ref["accepted"])
    assert st.n_accepted == m
    adj = _capi.unpack_bits(got["adj_bits"], m) if m else np.zeros((0, 0), np.uint8)
    assert np.array_equal(adj, ref["adj"]), np.argwhere(adj != ref["adj"])[:10]
    if m:
        np.testing.assert_allclose(got["dist"], ref["dist"], rtol=DIST_RTOL, atol=1e-300)
    assert st.n_consistent == int(ref["adj"].sum()) // 2


def test_se2_random_walk_parity(ctx):
    g, ot, lt = synth.workload("se2_500", seed=0)
    st, got, ref = _run_both(ctx, g, ot, lt)
    _assert_parity(ctx, st, got, ref)
    # odometry check active: the reference drops most inliers at threshold 1.0
    assert 0 < st.n_accepted < g.n_lc


def test_se2_no_odom_check_parity(ctx):
    g, _, _ = synth.workload("se2_500", seed=1)
    st, got, ref = _run_both(ctx, g, -1.0, 3.0)
    _assert_parity(ctx, st, got, ref)
    assert st.n_accepted == g.n_lc


def test_se3_5k_parity(ctx):
    g, ot, lt = synth.workload("se3_5k", seed=0)
    st, got, ref = _run_both(ctx, g, ot, lt)
    _assert_parity(ctx, st, got, ref)
    assert st.n_pairs == 1000 * 999 // 2


def test_se3_odom_check_and_small_thresholds(ctx):
    g = synth.se3_helix(n_poses=800, n_lc=150, outlier_frac=0.3, seed=4)
    for ot, lt in [(5.0, 3.0), (-1.0, 1.0), (2.0, -1.0), (0.0, 10.0)]:
        st, got, ref = _run_both(ctx, g, ot, lt)
        _assert_parity(ctx, st, got, ref, check_traj=(ot == 5.0))


def test_nan_rotation_covariance_and_edge_sizes(ctx):
    g = synth.se3_helix(n_poses=400, n_lc=70, outlier_frac=0.2, seed=5)
    cov = g.lc_cov.copy()
    cov[::7, 0] = np.nan  # rotation block unknown: translation-only distance
    ocov = g.odom_cov.copy()
    ocov[100, 7] = np.nan
    g2 = synth.PoseGraph(**{**g.__dict__, "lc_cov": cov, "odom_cov": ocov})
    st, got, ref = _run_both(ctx, g2, -1.0, 4.0)
    _assert_parity(ctx, st, got, ref)
    # zero and one loop closure
    for n in (0, 1):
        g3 = synth.PoseGraph(**{**g.__dict__, "lc_from": g.lc_from[:n], "lc_to": g.lc_to[:n],
                                "lc_pose": g.lc_pose[:n], "lc_cov": g.lc_cov[:n],
                                "is_outlier": g.is_outlier[:n]})
        st, got, ref = _run_both(ctx, g3, 5.0, 4.0)
        _assert_parity(ctx, st, got, ref)


def test_pcm_kat_scenario_on_gpu(ctx):
    """testPcm.cpp:90-192 scenario through the GPU path (threshold 0.5)."""
    def pose3(R, t):
        return np.concatenate([np.asarray(R, float).reshape(9), np.asarray(t, float)])

    def rz(a):
        c, s = math.cos(a), math.sin(a)
        return np.array([[c, -s, 0], [s, c, 0], [0, 0, 1.0]])

    R90 = np.array([[0, -1, 0], [1, 0, 0], [0, 0, 1.0]])
    odom = np.array([pose3(R90, (1, 0, 0))] * 2 + [pose3(np.eye(3), (1, 0, 0))] * 4)
    ocov = np.tile((0.1 * np.eye(6)).reshape(36), (6, 1))
    zs = np.array([pose3(rz(3.1416), (0, 0.9, 0)), pose3(rz(3.1416), (-1, 0.8, 0)),
                   pose3(rz(0.99 * 3.1416), (-1.8, 0.8, 0)), pose3(rz(0.98 * 3.1416), (-2.6, 0.6, 0))])
    g = synth.PoseGraph(dim=3, gt=np.zeros((7, 12)), values=np.tile(pose3(np.eye(3), (0, 0, 0)), (7, 1)),
                        odom=odom, odom_cov=ocov, lc_from=np.array([3, 4, 5, 6], np.int32),
                        lc_to=np.zeros(4, np.int32), lc_pose=zs,
                        lc_cov=np.tile((0.1 * np.eye(6)).reshape(36), (4, 1)),
                        is_outlier=np.zeros(4, bool), meta={"kind": "testPcm"})
    st, got, ref = _run_both(ctx, g, -1.0, 0.5)
    _assert_parity(ctx, st, got, ref)
    k, idx, _ = ctx.max_clique(_capi.CLIQUE_HEU_REF)
    assert k == 3 and 3 not in idx.tolist()


def test_clique_heu_ref_parity(ctx):
    g, ot, lt = synth.workload("se3_5k", seed=0)
    ctx.pcm_upload(g.dim, g.pose0, g.odom, g.odom_cov, g.lc_from, g.lc_to, g.lc_pose, g.lc_cov)
    ctx.pcm_run(ot, lt)
    got = ctx.pcm_download()
    m = len(got["accepted"])
    adj = _capi.unpack_bits(got["adj_bits"], m)
    k, idx, ms = ctx.max_clique(_capi.CLIQUE_HEU_REF)
    kr, idxr = O.max_clique_heu_dense(adj)
    print(f"[clique] m={m} k={k} ms={ms:.3f}", flush=True)
    assert k == kr and np.array_equal(idx, idxr)


def test_clique_heu_ref_random_graphs(ctx):
    rng = np.random.default_rng(21)
    for trial in range(25):
        n = int(rng.integers(1, 700))
        A = (rng.uniform(size=(n, n)) < rng.uniform(0.01, 0.6)).astype(np.uint8)
        blk = rng.choice(n, size=min(n, int(rng.integers(1, 200))), replace=False)
        A[np.ix_(blk, blk)] = rng.uniform(size=(len(blk), len(blk))) < 0.95
        A = np.triu(A, 1); A = A + A.T
        print(f"[clique] random trial={trial} n={n}", flush=True)
        k, idx = ctx.max_clique_bits(_capi.pack_bits(A), n)
        kr, idxr = O.max_clique_heu_dense(A)
        assert k == kr and np.array_equal(idx, idxr), (trial, n)
    # empty graph: FMC returns -1
    k, idx = ctx.max_clique_bits(np.zeros(1, np.uint64), 0)
    assert k == -1


def test_clique_above_register_bitset_limit_is_an_argument_error(ctx):
    """The heuristic keeps a wavefront's candidate set in registers: at most
    64 x 64 x 8 = 32,768 vertices.  One more is RPGO_ERR_ARG before any
    device work (the adjacency is not even copied)."""
    n = 32769
    W = (n + 63) // 64
    with pytest.raises(_capi.RpgoError, match="exceed the 32768 supported") as e:
        ctx.max_clique_bits(np.zeros(n * W, np.uint64), n)
    assert "status -1" in str(e.value)
    # the largest supported size still runs (an empty graph: every vertex alone)
    n = 32768
    k, idx = ctx.max_clique_bits(np.zeros(n * (n // 64), np.uint64), n)
    assert k == 1 and list(idx) == [0]
