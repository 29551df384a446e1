"""GPU parity of the FP32 PCM path (anchor-blocked segment covariances, FP64
replay of near-threshold pairs) against the FP64 oracle.

Bar: adjacency bit-exact against the oracle on every graph.  The FP32
distances of pairs that were not replayed agree with the oracle within the
recheck band (reported; asserted below 1e-3 relative).
"""
import numpy as np
import pytest

import oracle as O
from rpgo_amd import _capi, synth

pytestmark = pytest.mark.gpu

FP32 = _capi.PCM_FP32_RECHECK


@pytest.fixture(scope="module")
def ctx():
    c = _capi.Context(0)
    yield c
    c.close()


def _fp32_vs_oracle(ctx, g, odom_thr, lc_thr, check_dist=True):
    print(f"[fp32] kind={g.meta.get('kind')} N={g.n_poses} n_lc={g.n_lc} thr=({odom_thr},{lc_thr})",
          flush=True)
    ctx.pcm_upload(g.dim, g.pose0, g.odom, g.odom_cov, g.lc_from, g.lc_to, g.lc_pose, g.lc_cov)
    st = ctx.pcm_run(odom_thr, lc_thr, precision=FP32, want_dist=True)
    got = ctx.pcm_download(want_dist=True)
    traj = O.fold(g.dim, g.pose0, g.odom, g.odom_cov)
    acc, od, adj_ref, dist_ref = O.pcm_adjacency(g.dim, traj, g.lc_from, g.lc_to, g.lc_pose,
                                                 g.lc_cov, odom_thr, lc_thr)
    m = len(acc)
    assert np.array_equal(got["accepted"], acc)
    adj = _capi.unpack_bits(got["adj_bits"], m) if m else np.zeros((0, 0), np.uint8)
    bad = np.argwhere(adj != adj_ref)
    assert bad.size == 0, (bad[:10], [(dist_ref[i, j], got["dist"][i, j]) for i, j in bad[:5]])
    assert st.n_consistent == int(adj_ref.sum()) // 2
    rel = 0.0
    if m > 1 and check_dist:
        iu = np.triu_indices(m, 1)
        d32, d64 = got["dist"][iu], dist_ref[iu]
        fin = np.isfinite(d64) & (d64 > 0) & (d64 < 1e6)
        rel = float(np.max(np.abs(d32[fin] - d64[fin]) / d64[fin])) if fin.any() else 0.0
        print(f"[fp32] pairs={st.n_pairs} rechecked={st.n_rechecked} flipped={st.n_flipped} "
              f"max_rel_err={rel:.3e}", flush=True)
        assert rel < 1e-3
        # replayed pairs carry the FP64 replay's distance: the oracle's up to
        # the last ulp of the transcendentals (FP32 values are ~1e-6 off)
        assert int(np.sum(np.abs(d32 - d64) <= 1e-12 * np.abs(d64))) >= st.n_rechecked
    return st, rel


def test_fp32_se3_5k(ctx):
    g, ot, lt = synth.workload("se3_5k", seed=0)
    st, rel = _fp32_vs_oracle(ctx, g, ot, lt)
    assert st.n_rechecked < st.n_pairs // 20


@pytest.mark.parametrize("n_poses,n_lc", [(5000, 1000), (20000, 3000)])
def test_fp32_replay_is_the_fp64_path_bitwise(ctx, n_poses, n_lc):
    """Pairs replayed after the FP32 pass (one wavefront per pair for short
    queues, one lane per pair for long ones) carry the FP64 path's distance
    bit for bit."""
    g = synth.se3_helix(n_poses=n_poses, n_lc=n_lc, outlier_frac=0.3, seed=3)
    ctx.pcm_upload(g.dim, g.pose0, g.odom, g.odom_cov, g.lc_from, g.lc_to, g.lc_pose, g.lc_cov)
    ctx.pcm_run(-1.0, 5.0, precision=_capi.PCM_FP64_REPLAY, want_dist=True)
    ref = ctx.pcm_download(want_dist=True)
    st = ctx.pcm_pairs(5.0, precision=FP32, want_dist=True)
    got = ctx.pcm_download(want_dist=True)
    print(f"[fp32] replay vs FP64 path: N={n_poses} rechecked={st.n_rechecked}", flush=True)
    assert np.array_equal(got["adj_bits"], ref["adj_bits"])
    m = len(ref["accepted"])
    iu = np.triu_indices(m, 1)
    exact = int(np.sum(got["dist"][iu] == ref["dist"][iu]))
    assert st.n_rechecked > 0 and exact >= st.n_rechecked


def test_fp32_several_seeds_and_thresholds(ctx):
    for seed, (n, nl, of) in enumerate([(3000, 500, 0.3), (1500, 400, 0.5), (700, 300, 0.1)]):
        g = synth.se3_helix(n_poses=n, n_lc=nl, outlier_frac=of, seed=10 + seed)
        for lt in (1.0, 3.0, 5.0):
            _fp32_vs_oracle(ctx, g, -1.0, lt)


def test_fp32_odom_check_and_short_segments(ctx):
    # dense, short loop closures (many same-block and adjacent-block segments)
    g = synth.se3_helix(n_poses=900, n_lc=300, outlier_frac=0.2, seed=31, period=20, pitch=0.5,
                        lc_radius=1.5)
    _fp32_vs_oracle(ctx, g, 3.0, 4.0)
    _fp32_vs_oracle(ctx, g, -1.0, 4.0)


def test_fp32_shared_endpoints(ctx):
    # loop closures sharing their first key (a == c) and reversed directions
    g = synth.se3_helix(n_poses=600, n_lc=120, outlier_frac=0.25, seed=41)
    frm, to = g.lc_from.copy(), g.lc_to.copy()
    frm[1::3] = frm[0:-1:3][: len(frm[1::3])]
    g2 = synth.PoseGraph(**{**g.__dict__, "lc_from": frm, "lc_to": to})
    bad = np.abs(frm - to) < 2
    g2.lc_to[bad] = (g2.lc_to[bad] + 7) % g.n_poses
    _fp32_vs_oracle(ctx, g2, -1.0, 5.0)


def test_fp32_nan_rotation_falls_back_to_fp64(ctx):
    g = synth.se3_helix(n_poses=400, n_lc=80, outlier_frac=0.2, seed=5)
    cov = g.lc_cov.copy()
    cov[::5, 7] = np.nan
    g2 = synth.PoseGraph(**{**g.__dict__, "lc_cov": cov})
    st, _ = _fp32_vs_oracle(ctx, g2, -1.0, 4.0, check_dist=False)


def test_fp32_se2_falls_back_to_fp64(ctx):
    g, ot, lt = synth.workload("se2_500", seed=2)
    _fp32_vs_oracle(ctx, g, -1.0, 2.0)


def test_pairs_rerun_matches_fresh_run(ctx):
    g, ot, lt = synth.workload("se3_5k", seed=3)
    ctx.pcm_upload(g.dim, g.pose0, g.odom, g.odom_cov, g.lc_from, g.lc_to, g.lc_pose, g.lc_cov)
    ctx.pcm_run(ot, lt, precision=FP32)
    a1 = ctx.pcm_download()["adj_bits"]
    for prec in (FP32, _capi.PCM_FP64_REPLAY, FP32):
        st = ctx.pcm_pairs(lt, precision=prec)
        assert np.array_equal(ctx.pcm_download()["adj_bits"], a1)
    st = ctx.pcm_pairs(2.0, precision=FP32)
    ctx.pcm_run(ot, 2.0, precision=_capi.PCM_FP64_REPLAY)
    a2 = ctx.pcm_download()["adj_bits"]
    ctx.pcm_run(ot, lt, precision=FP32)
    st2 = ctx.pcm_pairs(2.0, precision=FP32)
    assert np.array_equal(ctx.pcm_download()["adj_bits"], a2)


def test_stage_timing_off_same_results(ctx):
    """rpgo_ctx_set_stage_timing(0): passes record no events, the ms_* fields of
    the pass read 0, and adjacency and counts equal the timed passes'."""
    g, ot, lt = synth.workload("se3_5k", seed=4)
    ctx.pcm_upload(g.dim, g.pose0, g.odom, g.odom_cov, g.lc_from, g.lc_to, g.lc_pose, g.lc_cov)
    ctx.pcm_run(ot, lt, precision=FP32)
    ref = ctx.pcm_download()["adj_bits"].copy()
    on = ctx.pcm_pairs(lt, precision=FP32)
    assert on.ms_pairs > 0 and on.ms_total > 0
    try:
        ctx.set_stage_timing(False)
        for prec in (FP32, _capi.PCM_FP64_REPLAY):
            off = ctx.pcm_pairs(lt, precision=prec)
            assert off.ms_pairs == 0 and off.ms_recheck == 0 and off.ms_total == 0
            assert off.n_consistent == on.n_consistent
            assert np.array_equal(ctx.pcm_download()["adj_bits"], ref)
        run = ctx.pcm_run(ot, lt, precision=FP32)
        assert run.ms_total == 0 and run.n_consistent == on.n_consistent
        assert np.array_equal(ctx.pcm_download()["adj_bits"], ref)
    finally:
        ctx.set_stage_timing(True)
    assert ctx.pcm_pairs(lt, precision=FP32).ms_pairs > 0


def test_fp32_large_graph_full_adjacency(ctx):
    """50,000 poses / 10,000 loop closures: all 49,995,000 pairs of the FP32
    path against the FP64 oracle (oracle row-parallel on the host cores)."""
    import os
    g, ot, lt = synth.workload("se3_50k", seed=1)
    print(f"[fp32] se3_50k full N={g.n_poses} n_lc={g.n_lc}", flush=True)
    ctx.pcm_upload(g.dim, g.pose0, g.odom, g.odom_cov, g.lc_from, g.lc_to, g.lc_pose, g.lc_cov)
    st = ctx.pcm_run(ot, lt, precision=FP32)
    got = ctx.pcm_download()
    m = st.n_accepted
    adj = _capi.unpack_bits(got["adj_bits"], m)
    traj = O.fold(g.dim, g.pose0, g.odom, g.odom_cov)
    O.set_threads(min(16, os.cpu_count() or 1))
    try:
        dist = O.pcm_dist_rows(3, traj, g.lc_from, g.lc_to, g.lc_pose, g.lc_cov)
    finally:
        O.set_threads(1)
    ref = (dist < lt).astype(np.uint8)
    np.fill_diagonal(ref, 0)
    bad = np.argwhere(adj != ref)
    print(f"[fp32] 50k full: pairs={st.n_pairs} rechecked={st.n_rechecked} flipped={st.n_flipped} "
          f"mismatches={len(bad)}", flush=True)
    assert len(bad) == 0, [(int(i), int(j), dist[i, j]) for i, j in bad[:8]]


def test_fp32_large_graph_sampled_rows(ctx):
    """50,000-pose graph: adjacency rows spot-checked against the oracle, plus
    size-independent properties (symmetry, zero diagonal, consistent count)."""
    g, ot, lt = synth.workload("se3_50k", seed=0)
    print(f"[fp32] se3_50k N={g.n_poses} n_lc={g.n_lc}", flush=True)
    ctx.pcm_upload(g.dim, g.pose0, g.odom, g.odom_cov, g.lc_from, g.lc_to, g.lc_pose, g.lc_cov)
    st = ctx.pcm_run(ot, lt, precision=FP32)
    got = ctx.pcm_download()
    m = st.n_accepted
    assert m == g.n_lc
    adj = _capi.unpack_bits(got["adj_bits"], m)
    assert np.array_equal(adj, adj.T) and not adj.diagonal().any()
    assert int(adj.sum()) // 2 == st.n_consistent
    traj = O.fold(g.dim, g.pose0, g.odom, g.odom_cov)
    rng = np.random.default_rng(0)
    rows = np.sort(rng.choice(np.arange(1, m), size=24, replace=False))
    pi, pj = [], []
    for j in rows:
        pi.extend(range(j)); pj.extend([j] * j)
    pi, pj = np.array(pi), np.array(pj)
    acc = got["accepted"]
    d = O.pair_dist_list(3, traj, g.lc_from[acc], g.lc_to[acc], g.lc_pose[acc], g.lc_cov[acc], pi, pj)
    ref = (d < lt).astype(np.uint8)
    mism = np.nonzero(adj[pj, pi] != ref)[0]
    print(f"[fp32] 50k: pairs={st.n_pairs} rechecked={st.n_rechecked} flipped={st.n_flipped} "
          f"checked={len(pi)} mismatches={len(mism)} ms_pairs={st.ms_pairs:.2f}", flush=True)
    assert len(mism) == 0
