"""Streaming PCM (rpgo_pcm_append / rpgo_pcm_simple_append): one reference
update() after another, as examples/RpgoReadG2oIncremental.cpp:52-76 drives
it.  After every increment the device state must equal a full run over the
concatenated problem bit for bit (trajectory records, accepted loop closures,
adjacency bits), and the incremental clique (findMaxCliqueHeuIncremental,
GraphUtils.cpp:30-44) must equal the oracle's on that adjacency."""
import numpy as np
import pytest

import oracle as O
from rpgo_amd import _capi, synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = _capi.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def ref():
    c = _capi.Context(0)
    yield c
    c.close()


def _stream_order(g):
    """Loop closures sorted (stably) by the later endpoint: each can be added as
    soon as both its poses exist."""
    return np.argsort(np.maximum(g.lc_from, g.lc_to), kind="stable")


def _chunks(g, order, step):
    """Increments: odometry rows [p0, p1) and the loop closures whose later
    endpoint is below the new chain length."""
    later = np.maximum(g.lc_from, g.lc_to)[order]
    n = g.n_poses
    out = []
    p0, t0 = 1, 0
    while p0 < n:
        p1 = min(n, p0 + step)
        t1 = int(np.searchsorted(later, p1, side="left"))
        out.append((p0, p1, order[t0:t1]))
        p0, t0 = p1, t1
    return out


def _start(c, g):
    c.pcm_upload(g.dim, g.pose0, np.zeros((0, g.odom.shape[1])), np.zeros((0, g.odom_cov.shape[1])),
                 np.zeros(0, np.int32), np.zeros(0, np.int32), np.zeros((0, g.lc_pose.shape[1])),
                 np.zeros((0, g.lc_cov.shape[1])))


def _full(c, g, n_poses, lcs, simple, thr):
    c.pcm_upload(g.dim, g.pose0, g.odom[: n_poses - 1], g.odom_cov[: n_poses - 1], g.lc_from[lcs],
                 g.lc_to[lcs], g.lc_pose[lcs], g.lc_cov[lcs])
    if simple:
        st = c.pcm_simple_run(*thr)
    else:
        st = c.pcm_run(*thr, precision=_capi.PCM_FP64_REPLAY)
    return st, c.pcm_download(), c.pcm_traj()


def _run_stream(ctx, ref, g, step, simple, thr, check_every=1, incremental=False):
    order = _stream_order(g)
    _start(ctx, g)
    if simple:
        ctx.pcm_simple_run(*thr)
    else:
        ctx.pcm_run(*thr, precision=_capi.PCM_FP64_REPLAY)
    seen = np.zeros(0, np.int64)
    prev_size = 0
    m_prev = 0
    pairs = 0
    for it, (p0, p1, lcs) in enumerate(_chunks(g, order, step)):
        inc = dict(odom=g.odom[p0 - 1:p1 - 1], odom_cov=g.odom_cov[p0 - 1:p1 - 1], lc_from=g.lc_from[lcs],
                   lc_to=g.lc_to[lcs], lc_pose=g.lc_pose[lcs], lc_cov=g.lc_cov[lcs])
        st = ctx.pcm_simple_append(*thr, **inc) if simple else ctx.pcm_append(*thr, **inc)
        seen = np.concatenate([seen, lcs])
        pairs += st.n_pairs
        m = st.n_accepted
        if incremental and m > 0:
            k, idx = ctx.max_clique_incremental(m - m_prev, prev_size)
        if it % check_every == 0 or p1 == g.n_poses:
            got = ctx.pcm_download()
            traj = ctx.pcm_traj()
            adj_bits = got["adj_bits"].copy()
            acc = got["accepted"].copy()
            st_f, full, traj_f = _full(ref, g, p1, seen, simple, thr)
            assert np.array_equal(traj, traj_f), it
            assert np.array_equal(acc, full["accepted"]), it
            assert np.array_equal(adj_bits, full["adj_bits"]), it
            assert m == st_f.n_accepted
            if incremental and m > 0:
                adj = _capi.unpack_bits(adj_bits, m)
                k_ref, idx_ref = O.max_clique_heu_incremental(adj, m - m_prev, prev_size)
                assert k == k_ref and np.array_equal(idx, idx_ref[: max(k, 0)]), it
        if incremental and m > 0 and k > 0:
            prev_size = k
        m_prev = m
    assert len(seen) == g.n_lc
    return pairs


@pytest.mark.parametrize("thr", [(-1.0, 5.0), (10.0, 3.0)])
def test_stream_se3_chunks(ctx, ref, thr):
    g = synth.se3_helix(n_poses=1500, n_lc=320, outlier_frac=0.3, seed=11)
    print(f"[stream] se3 chunks thr={thr}", flush=True)
    pairs = _run_stream(ctx, ref, g, step=100, simple=False, thr=thr)
    assert pairs > 0


def test_stream_se3_one_lc_per_call(ctx, ref):
    """One loop closure (and the odometry it needs) per call; the accepted count
    crosses the 64-row word boundaries of the adjacency."""
    g = synth.se3_helix(n_poses=700, n_lc=150, outlier_frac=0.25, seed=12)
    print("[stream] se3 single-LC increments", flush=True)
    _run_stream(ctx, ref, g, step=4, simple=False, thr=(-1.0, 5.0), check_every=7, incremental=True)


def test_stream_se2(ctx, ref):
    g, ot, lt = synth.workload("se2_500", seed=1)
    print("[stream] se2", flush=True)
    _run_stream(ctx, ref, g, step=37, simple=False, thr=(ot, lt), incremental=True)


def test_stream_simple(ctx, ref):
    g = synth.se3_helix(n_poses=900, n_lc=200, outlier_frac=0.3, seed=13)
    print("[stream] PcmSimple", flush=True)
    _run_stream(ctx, ref, g, step=60, simple=True, thr=(0.03, 0.003, 0.002, 0.0002))


def test_stream_mode_mismatch_is_an_error(ctx):
    g = synth.se3_helix(n_poses=50, n_lc=5, outlier_frac=1.0, seed=1)
    _start(ctx, g)
    ctx.pcm_run(-1.0, 5.0)
    with pytest.raises(RuntimeError):
        ctx.pcm_simple_append(0.1, 0.1, 0.1, 0.1, odom=g.odom[:3], odom_cov=g.odom_cov[:3])
    with pytest.raises(RuntimeError):
        # loop closure beyond the extended chain
        ctx.pcm_append(-1.0, 5.0, odom=g.odom[:2], odom_cov=g.odom_cov[:2], lc_from=[0], lc_to=[9],
                       lc_pose=g.lc_pose[:1], lc_cov=g.lc_cov[:1])
