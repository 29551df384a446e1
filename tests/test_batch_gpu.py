"""Batched PCM (rpgo_pcm_batch_run): every graph of a batch of uneven graphs
gets exactly what the single-graph FP64 path gives it alone -- accepted loop
closures, bit-packed adjacency and FMC::maxCliqueHeu clique -- for SE(3) and
SE(2), with the odometry check on and off and lc_threshold < 0."""
import numpy as np
import pytest

from rpgo_amd import _capi, synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = _capi.Context(0)
    yield c
    c.close()


def _graphs3(seed, n):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        # the helix generator needs more than one turn (100 poses) for inliers
        npz = int(np.exp(rng.uniform(np.log(150), np.log(1500))))
        out.append(synth.se3_helix(n_poses=npz, n_lc=max(1, int(npz * rng.uniform(0.1, 0.3))),
                                   outlier_frac=float(rng.uniform(0.05, 0.4)), seed=seed * 100 + k))
    return out


def _single(ctx, g, odom_thr, lc_thr):
    ctx.pcm_upload(g.dim, g.pose0, g.odom, g.odom_cov, g.lc_from, g.lc_to, g.lc_pose, g.lc_cov)
    ctx.pcm_run(odom_thr, lc_thr, precision=_capi.PCM_FP64_REPLAY)
    d = ctx.pcm_download()
    k, idx, _ = ctx.max_clique(_capi.CLIQUE_HEU_REF)
    return d["accepted"], np.asarray(d["adj_bits"]).reshape(len(d["accepted"]), -1), k, idx


@pytest.mark.parametrize("odom_thr,lc_thr", [(-1.0, 5.0), (10.0, 3.0), (-1.0, -1.0)])
def test_batch_se3_matches_single(ctx, odom_thr, lc_thr):
    graphs = _graphs3(7, 12)
    res = ctx.pcm_batch_run(graphs, odom_thr, lc_thr)
    adjs = [ctx.pcm_batch_adj(g).copy() for g in range(len(graphs))]
    for g, gr in enumerate(graphs):
        acc, adj, k, idx = _single(ctx, gr, odom_thr, lc_thr)
        assert np.array_equal(res["accepted"][g], acc), g
        assert np.array_equal(adjs[g], adj), g
        assert int(res["clique_size"][g]) == k
        assert np.array_equal(res["clique_idx"][g], idx[:max(k, 0)])
    st = res["stats"]
    print(f"[batch] SE(3) {len(graphs)} graphs odom {odom_thr} lc {lc_thr}: {st.n_pairs} pairs, "
          f"{st.ms_pairs:.2f} ms pairs, {st.ms_total:.2f} ms total", flush=True)


def test_batch_large_graphs_match_single(ctx):
    """Graphs of several thousand loop closures: the clique scan runs over
    several chip-wide rounds (1,024 candidates per graph each) and, for the
    largest, beyond them in the per-graph scan."""
    graphs = [synth.se3_helix(n_poses=n, n_lc=l, outlier_frac=f, seed=s)
              for n, l, f, s in ((9000, 2600, 0.3, 1), (400, 90, 0.2, 2), (40000, 12000, 0.4, 3), (3000, 700, 0.1, 4))]
    res = ctx.pcm_batch_run(graphs, -1.0, 5.0)
    for g, gr in enumerate(graphs):
        acc, adj, k, idx = _single(ctx, gr, -1.0, 5.0)
        print(f"[batch] large graph {g}: {len(acc)} loop closures, clique {k}", flush=True)
        assert np.array_equal(res["accepted"][g], acc), g
        assert int(res["clique_size"][g]) == k
        assert np.array_equal(res["clique_idx"][g], idx[:max(k, 0)])


def test_batch_se2_matches_single(ctx):
    rng = np.random.default_rng(3)
    graphs = [synth.se2_random_walk(n_poses=int(rng.integers(100, 600)), n_inlier=int(rng.integers(10, 60)),
                                    n_outlier=int(rng.integers(2, 20)), seed=k) for k in range(6)]
    res = ctx.pcm_batch_run(graphs, 1.0, 1.0)
    adjs = [ctx.pcm_batch_adj(g).copy() for g in range(len(graphs))]
    for g, gr in enumerate(graphs):
        acc, adj, k, idx = _single(ctx, gr, 1.0, 1.0)
        assert np.array_equal(res["accepted"][g], acc)
        assert np.array_equal(adjs[g], adj)
        assert int(res["clique_size"][g]) == k
        assert np.array_equal(res["clique_idx"][g], idx[:max(k, 0)])


def test_batch_single_pose_and_empty_graphs(ctx):
    graphs = _graphs3(9, 3)
    tiny = synth.se3_helix(n_poses=5, n_lc=1, outlier_frac=1.0, seed=1)  # one outlier loop closure
    res = ctx.pcm_batch_run([graphs[0], tiny, graphs[1]], -1.0, 5.0)
    for g, gr in enumerate([graphs[0], tiny, graphs[1]]):
        acc, adj, k, idx = _single(ctx, gr, -1.0, 5.0)
        assert np.array_equal(res["accepted"][g], acc)
        assert int(res["clique_size"][g]) == k


def _assert_same(res, ref):
    assert np.array_equal(res["n_accepted"], ref["n_accepted"])
    for g in range(len(res["accepted"])):
        assert np.array_equal(res["accepted"][g], ref["accepted"][g]), g
        assert np.array_equal(res["clique_idx"][g], ref["clique_idx"][g]), g
    assert np.array_equal(res["clique_size"], ref["clique_size"])


def test_batch_graph_replay_and_autotune(ctx):
    """RPGO_BATCH_GRAPH captures the stages once and replays them while the layout
    holds; RPGO_BATCH_AUTOTUNE picks launch shapes.  Neither changes a bit, and a
    replay on new values of the same layout gives the new graphs' results."""
    graphs = _graphs3(21, 10)
    print("[batch] graph/autotune: plain run", flush=True)
    plain = ctx.pcm_batch_run(graphs, 10.0, 5.0)
    adj_plain = [ctx.pcm_batch_adj(g).copy() for g in range(len(graphs))]
    print("[batch] graph/autotune: tuned + captured", flush=True)
    first = ctx.pcm_batch_run(graphs, 10.0, 5.0, flags=_capi.BATCH_GRAPH | _capi.BATCH_AUTOTUNE)
    assert first["stats"].graph_replayed == 0
    assert first["stats"].pair_waves in (1, 2, 4, 8) and first["stats"].clique_waves in (4, 8, 16)
    _assert_same(first, plain)
    again = ctx.pcm_batch_run(graphs, 10.0, 5.0, flags=_capi.BATCH_GRAPH | _capi.BATCH_AUTOTUNE)
    assert again["stats"].graph_replayed == 1 and again["stats"].ms_tune == 0.0
    _assert_same(again, plain)
    for g in range(len(graphs)):
        assert np.array_equal(ctx.pcm_batch_adj(g), adj_plain[g]), g
    # same sizes, other values: the replay must compute the new graphs
    other = [synth.se3_helix(n_poses=gr.n_poses, n_lc=gr.n_lc, outlier_frac=0.35, seed=500 + k)
             for k, gr in enumerate(graphs)]
    print("[batch] graph/autotune: replay on new values", flush=True)
    rep = ctx.pcm_batch_run(other, 10.0, 5.0, flags=_capi.BATCH_GRAPH | _capi.BATCH_AUTOTUNE)
    assert rep["stats"].graph_replayed == 1
    for g, gr in enumerate(other):
        acc, adj, k, idx = _single(ctx, gr, 10.0, 5.0)
        assert np.array_equal(rep["accepted"][g], acc), g
        assert int(rep["clique_size"][g]) == k
        assert np.array_equal(rep["clique_idx"][g], idx[:max(k, 0)]), g
    # a changed threshold is part of the layout: captured again
    thr = ctx.pcm_batch_run(other, 10.0, 3.0, flags=_capi.BATCH_GRAPH)
    assert thr["stats"].graph_replayed == 0
    ref = ctx.pcm_batch_run(other, 10.0, 3.0)
    _assert_same(thr, ref)


def _oracle_graph(gr, odom_thr, lc_thr):
    import oracle as O

    traj = O.fold(gr.dim, gr.pose0, gr.odom, gr.odom_cov)
    acc, _, adj, _ = O.pcm_adjacency(gr.dim, traj, gr.lc_from, gr.lc_to, gr.lc_pose, gr.lc_cov, odom_thr, lc_thr,
                                     want_dist=False)
    k, idx = O.max_clique_heu_dense(adj)
    return acc, adj, k, idx


@pytest.mark.parametrize("odom_thr,lc_thr", [(-1.0, 5.0), (10.0, 3.0)])
def test_batch_matches_fp64_oracle(ctx, odom_thr, lc_thr):
    """Every graph of an uneven batch against the FP64 CPU oracle directly:
    accepted loop closures, the whole adjacency and the reference heuristic
    clique."""
    import oracle as O

    O.set_threads(16)
    graphs = _graphs3(31, 10)
    res = ctx.pcm_batch_run(graphs, odom_thr, lc_thr)
    for g, gr in enumerate(graphs):
        acc, adj, k, idx = _oracle_graph(gr, odom_thr, lc_thr)
        got = _capi.unpack_bits(ctx.pcm_batch_adj(g), len(acc))
        assert np.array_equal(res["accepted"][g], acc), g
        assert np.array_equal(got, adj), g
        assert int(res["clique_size"][g]) == k and np.array_equal(res["clique_idx"][g], idx[:max(k, 0)]), g


def test_batch_c4_full_size():
    """The 256-graph batch of the benchmark (poses log-uniform in [200, 20000]):
    properties of every graph at full size, and oracle rows of eight graphs.
    Every loop closure accepted (odometry check off), adjacency symmetric with
    a zero diagonal, its consistent pairs summing to the batch count, the
    clique's k entries accepted loop closures; sampled rows of eight
    graphs (the largest among them) equal the FP64 oracle's distances against
    the threshold bit for bit, and the cliques of the graphs up to 1,200 loop
    closures equal the oracle's heuristic on the oracle's own adjacency."""
    import oracle as O

    O.set_threads(16)
    graphs = synth.c4_batch(0, 256)
    with _capi.Context(0) as c:
        print(f"[batch] C4: {len(graphs)} graphs, {sum(g.n_poses for g in graphs)} poses", flush=True)
        res = c.pcm_batch_run(graphs, -1.0, 5.0, flags=_capi.BATCH_GRAPH)
        total = 0
        for g, gr in enumerate(graphs):
            m = int(res["n_accepted"][g])
            assert m == gr.n_lc and np.array_equal(res["accepted"][g], np.arange(m)), g
            A = _capi.unpack_bits(c.pcm_batch_adj(g), m)
            assert np.array_equal(A, A.T) and not A.diagonal().any(), g
            total += int(A.sum()) // 2
            k = int(res["clique_size"][g])
            ids = res["clique_idx"][g]
            # FMC::maxCliqueHeu's stale buffer may repeat a vertex (findCliqueHeu.cpp:102,110-113)
            assert 1 <= k <= m and len(ids) == k and ids.min() >= 0 and ids.max() < m, g
        assert total == res["stats"].n_consistent
        order = np.argsort([gr.n_lc for gr in graphs])
        rng = np.random.default_rng(1)
        for g in list(order[-3:]) + list(rng.choice(order[:-3], 5, replace=False)):
            gr = graphs[g]
            m = gr.n_lc
            traj = O.fold(gr.dim, gr.pose0, gr.odom, gr.odom_cov)
            rows = np.sort(rng.choice(np.arange(1, m), min(24, m - 1), replace=False))
            pi = np.concatenate([np.arange(j) for j in rows]).astype(np.int32)
            pj = np.concatenate([np.full(j, j) for j in rows]).astype(np.int32)
            d = O.pair_dist_list(gr.dim, traj, gr.lc_from, gr.lc_to, gr.lc_pose, gr.lc_cov, pi, pj)
            A = _capi.unpack_bits(c.pcm_batch_adj(g), m)
            assert np.array_equal(A[pj, pi], (d < 5.0).astype(np.uint8)), g
            print(f"[batch] C4 graph {g}: {m} loop closures, {len(rows)} oracle rows", flush=True)
            if m <= 1200:
                acc, adj, k, idx = _oracle_graph(gr, -1.0, 5.0)
                assert np.array_equal(A, adj), g
                assert int(res["clique_size"][g]) == k and np.array_equal(res["clique_idx"][g], idx[:max(k, 0)]), g


def _adjs(ctx, n):
    return [ctx.pcm_batch_adj(g).copy() for g in range(n)]


def _assert_fp32_equals_fp64(ctx, graphs, odom_thr, lc_thr, flags=0):
    ref = ctx.pcm_batch_run(graphs, odom_thr, lc_thr, flags=flags)
    adj_ref = _adjs(ctx, len(graphs))
    got = ctx.pcm_batch_run(graphs, odom_thr, lc_thr, flags=flags, precision=_capi.PCM_FP32_RECHECK)
    adj_got = _adjs(ctx, len(graphs))
    _assert_same(got, ref)
    for g in range(len(graphs)):
        assert np.array_equal(adj_got[g], adj_ref[g]), g
    assert got["stats"].n_consistent == ref["stats"].n_consistent
    return got["stats"], ref["stats"]


@pytest.mark.parametrize("odom_thr,lc_thr", [(-1.0, 5.0), (10.0, 3.0)])
def test_batch_fp32_equals_fp64_batch(ctx, odom_thr, lc_thr):
    """The FP32 pair stage of an SE(3) batch (per-graph anchor tables, one pair
    launch over all graphs, the pairs near the threshold replayed in FP64)
    gives every graph the FP64 stage's adjacency bit for bit, hence the same
    cliques; a one-pose-pair graph and graphs of 15-450 loop closures mixed."""
    graphs = _graphs3(41, 14)
    graphs.insert(3, synth.se3_helix(n_poses=5, n_lc=1, outlier_frac=1.0, seed=1))
    st, st_ref = _assert_fp32_equals_fp64(ctx, graphs, odom_thr, lc_thr)
    print(f"[batch] FP32 stage odom {odom_thr} lc {lc_thr}: {st.n_pairs} pairs, {st.n_rechecked} replayed, "
          f"{st.n_flipped} overturned; pairs {st.ms_pairs:.2f} ms vs FP64 {st_ref.ms_pairs:.2f} ms", flush=True)
    assert st.fp32 == 1 and st_ref.fp32 == 0
    _assert_fp32_equals_fp64(ctx, graphs, odom_thr, lc_thr, flags=_capi.BATCH_GRAPH)


def test_batch_fp32_c4_equals_fp64_batch(ctx):
    """The 256-graph C4 batch (up to several thousand loop closures per graph)
    in the captured FP32 stage equals the FP64 stage graph by graph."""
    graphs = synth.c4_batch(0, 256)
    st, st_ref = _assert_fp32_equals_fp64(ctx, graphs, -1.0, 5.0, flags=_capi.BATCH_GRAPH)
    print(f"[batch] C4 FP32 stage: {st.n_pairs} pairs, {st.n_rechecked} replayed, {st.n_flipped} overturned; "
          f"{st.ms_total:.1f} ms vs FP64 {st_ref.ms_total:.1f} ms", flush=True)
    assert st.fp32 == 1


def test_batch_fp32_without_rotation_information_runs_fp64(ctx):
    """A loop-closure covariance with NaN rotation entries sends the batch's
    pairs to the FP64 stage (as the single-graph FP32 path does); the results
    still equal the FP64 batch."""
    graphs = _graphs3(43, 6)
    g = graphs[2]
    cov = g.lc_cov.copy()
    cov[::5, 7] = np.nan
    graphs[2] = synth.PoseGraph(**{**g.__dict__, "lc_cov": cov})
    st, _ = _assert_fp32_equals_fp64(ctx, graphs, -1.0, 4.0)
    assert st.fp32 == 0 and st.n_rechecked == 0
