"""Exact max-clique mode (RPGO_CLIQUE_EXACT) against the literal CPU
restatement of FMC::maxClique (findClique.cpp:31-145): bit-exact size and
vertex order on random graphs (with planted cliques), edge cases, and a PCM
consistency graph small enough for the sequential search."""
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


def _random_graph(rng, n, p, plant):
    a = np.triu(rng.random((n, n)) < p, 1)
    a = a | a.T
    if plant and n > 4:
        s = rng.choice(n, int(rng.integers(3, max(4, n // 2))), replace=False)
        a[np.ix_(s, s)] = True
        np.fill_diagonal(a, False)
    return a.astype(np.uint8)


def _gpu_exact(ctx, adj):
    n = adj.shape[0]
    k, idx = ctx.max_clique_bits(_capi.pack_bits(adj), n, _capi.CLIQUE_EXACT)
    return k, idx


@pytest.mark.parametrize("seed", range(6))
def test_random_graphs_bit_exact(ctx, seed):
    rng = np.random.default_rng(100 + seed)
    for t in range(40):
        n = int(rng.integers(1, 70))
        a = _random_graph(rng, n, float(rng.uniform(0.05, 0.9)), t % 3 == 0)
        kr, ir = O.max_clique_exact(a)
        k, idx = _gpu_exact(ctx, a)
        assert k == kr, (seed, t, n, k, kr)
        assert list(idx[:k]) == list(ir), (seed, t, n, list(idx[:k]), list(ir))
    print(f"[clique] exact seed {seed}: 40 graphs bit-exact", flush=True)


def test_edge_cases(ctx):
    # no edges: the highest vertex alone
    a = np.zeros((5, 5), np.uint8)
    k, idx = _gpu_exact(ctx, a)
    assert k == 1 and list(idx[:1]) == [4]
    assert (k, list(idx[:k])) == (O.max_clique_exact(a)[0], list(O.max_clique_exact(a)[1]))
    # complete graph: every vertex, deepest first
    a = 1 - np.eye(7, dtype=np.uint8)
    k, idx = _gpu_exact(ctx, a)
    kr, ir = O.max_clique_exact(a)
    assert k == kr == 7 and list(idx[:k]) == list(ir)
    # more than 64 vertices (several words per row)
    rng = np.random.default_rng(5)
    a = _random_graph(rng, 150, 0.3, True)
    kr, ir = O.max_clique_exact(a)
    k, idx = _gpu_exact(ctx, a)
    assert k == kr and list(idx[:k]) == list(ir)


def test_pcm_graph_bit_exact(ctx):
    g = synth.se3_helix(n_poses=2000, n_lc=300, outlier_frac=0.3, seed=1)
    ctx.pcm_upload(g.dim, g.pose0, g.odom, g.odom_cov, g.lc_from, g.lc_to, g.lc_pose, g.lc_cov)
    ctx.pcm_run(-1.0, 5.0, precision=_capi.PCM_FP64_REPLAY)
    adj = _capi.unpack_bits(ctx.pcm_download()["adj_bits"], len(ctx.pcm_download()["accepted"]))
    k, idx, _ = ctx.max_clique(_capi.CLIQUE_EXACT)
    kr, ir = O.max_clique_exact(adj)
    print(f"[clique] PCM m={adj.shape[0]} exact {k} (oracle {kr})", flush=True)
    assert k == kr and list(idx[:k]) == list(ir)


@pytest.mark.parametrize("seed", range(4))
def test_incremental_heuristic_bit_exact(ctx, seed):
    """findMaxCliqueHeuIncremental: size (0 unless it grew) and entries."""
    rng = np.random.default_rng(200 + seed)
    for t in range(40):
        n = int(rng.integers(1, 120))
        a = _random_graph(rng, n, float(rng.uniform(0.05, 0.9)), t % 3 == 0)
        num_new = int(rng.integers(0, n + 1))
        prev = int(rng.integers(0, 8))
        kr, ir = O.max_clique_heu_incremental(a, num_new, prev)
        kr = kr if kr > prev else 0
        k, idx = ctx.max_clique_incremental_bits(_capi.pack_bits(a), n, num_new, prev)
        assert k == kr, (seed, t, n, num_new, prev, k, kr)
        assert list(idx[:k]) == list(ir[:kr]), (seed, t)
    print(f"[clique] incremental seed {seed}: 40 graphs bit-exact", flush=True)


def _dense_block_graph(rng, n, frac, missing, p_noise):
    """A near-complete block (frac of the vertices, `missing` random non-edges
    inside it) in a sparse random graph: the PCM inlier-block shape, where
    degrees bound nothing and the greedy-colouring bound does the pruning."""
    a = np.triu(rng.random((n, n)) < p_noise, 1)
    a = a | a.T
    blk = np.sort(rng.choice(n, max(2, int(frac * n)), replace=False))
    a[np.ix_(blk, blk)] = True
    for _ in range(missing):
        u, v = rng.choice(blk, 2, replace=False)
        a[u, v] = a[v, u] = False
    np.fill_diagonal(a, False)
    return a.astype(np.uint8)


@pytest.mark.parametrize("seed", range(4))
def test_dense_blocks_bit_exact(ctx, seed):
    """Near-complete blocks with up to 12 missing edges (the sequential search
    still finishes): size and vertex order equal FMC::maxClique's."""
    rng = np.random.default_rng(300 + seed)
    for t in range(12):
        n = int(rng.integers(30, 130))
        a = _dense_block_graph(rng, n, float(rng.uniform(0.5, 0.95)), int(rng.integers(0, 13)),
                               float(rng.uniform(0.02, 0.3)))
        kr, ir = O.max_clique_exact(a)
        k, idx = _gpu_exact(ctx, a)
        assert k == kr, (seed, t, n, k, kr)
        assert list(idx[:k]) == list(ir), (seed, t, n)
    print(f"[clique] exact dense blocks seed {seed}: 12 graphs bit-exact", flush=True)


def test_exact_on_the_5k_pcm_graph(ctx):
    """The 5k workload's consistency graph (1,000 loop closures, a 768-member
    heuristic clique in a near-complete inlier block): the exact search
    finishes, returns a clique of the adjacency at least as large as the
    heuristic's, listed deepest vertex first and with the seed (its highest
    vertex) last."""
    import time

    g, ot, lt = synth.workload("se3_5k", seed=0)
    ctx.pcm_upload(g.dim, g.pose0, g.odom, g.odom_cov, g.lc_from, g.lc_to, g.lc_pose, g.lc_cov)
    ctx.pcm_run(ot, lt, precision=_capi.PCM_FP32_RECHECK)
    m = len(ctx.pcm_download()["accepted"])
    kh, ih, _ = ctx.max_clique(_capi.CLIQUE_HEU_REF)
    t = time.perf_counter()
    k, idx, _ = ctx.max_clique(_capi.CLIQUE_EXACT)
    dt = time.perf_counter() - t
    A = _capi.unpack_bits(ctx.pcm_download()["adj_bits"], m)
    print(f"[clique] 5k exact: size {k} (heuristic {kh}) in {dt:.2f} s", flush=True)
    assert k >= kh
    sel = np.asarray(idx[:k])
    assert len(set(sel.tolist())) == k
    sub = A[np.ix_(sel, sel)]
    assert sub.sum() == k * (k - 1), "not a clique"
    assert sel[-1] == sel.max()
