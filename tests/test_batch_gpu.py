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
        npz = int(np.exp(rng.uniform(np.log(60), np.log(1500))))
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
    tiny = synth.se3_helix(n_poses=5, n_lc=1, outlier_frac=0.0, seed=1)
    res = ctx.pcm_batch_run([graphs[0], tiny, graphs[1]], -1.0, 5.0)
    for g, gr in enumerate([graphs[0], tiny, graphs[1]]):
        acc, adj, k, idx = _single(ctx, gr, -1.0, 5.0)
        assert np.array_equal(res["accepted"][g], acc)
        assert int(res["clique_size"][g]) == k
