"""Pin the oracle functions that the reference's own tests leave open.

GNC-TLS (oracle/solver_oracle.py: tls_weights, tls_mu_init, gnc_converged,
gnc_optimize), restated from GTSAM's GncOptimizer (not vendored; called at
src/RobustSolver.cpp:142-143) and the GNC paper (Yang et al., RA-L 2020):
  * the weight is the minimiser over [0, 1] of the Black-Rangarajan dual
    objective w r + Phi_mu(w), Phi_mu(w) = mu (1 - w) barc / (mu + w), found
    here by a dense grid search rather than by the closed form;
  * the band edges follow from continuity (w = 1 at mu/(mu+1) barc, 0 at
    (mu+1)/mu barc), so a swapped or mis-scaled bound fails;
  * Remark 5's mu_0 puts the largest initial residual at half the upper band
    edge (ub(mu_0) = 2 r_max);
  * a three-factor trace whose inner least-squares problem has a closed form
    (the weighted mean of three translation priors), with mu_0 and the first
    weights checked against literal values computed by hand.

Heuristic and exact max clique (oracle/rpgo_oracle.c):
  * maxCliqueHeuIncremental (findCliqueHeu.cpp:119-209) through the wrapper
    findMaxCliqueHeuIncremental (src/utils/GraphUtils.cpp:30-44): hand-traced
    KATs of the start index, the seeding with the previous size and the
    "0 unless it grew" rule, including the stale-buffer output [3, 4, 4];
    with every vertex new and bound 0 it equals maxCliqueHeu;
  * maxClique (findClique.cpp:80-145): hand-traced KATs of the scan order
    (top-level vertices high to low, U consumed from the back) and of the
    output order (reverse discovery: deepest vertex first, seed last).
"""
import math

import numpy as np
import pytest
from scipy.stats import chi2

import oracle as O
import solver_oracle as S

BARC_09 = 0.5 * chi2.ppf(0.9, 6)  # 5.322320


# ------------------------------------------------------------------ GNC-TLS
def _surrogate_argmin(r, mu, barc, n=200001):
    w = np.linspace(0.0, 1.0, n)
    obj = w * r + mu * (1.0 - w) * barc / (mu + w)
    return w[np.argmin(obj)]


def test_tls_weight_minimises_the_gnc_surrogate():
    rng = np.random.default_rng(0)
    for _ in range(300):
        mu = float(np.exp(rng.uniform(np.log(1e-4), np.log(50.0))))
        barc = float(rng.uniform(0.2, 10.0))
        lb, ub = mu / (mu + 1) * barc, (mu + 1) / mu * barc
        r = float(np.exp(rng.uniform(np.log(0.2 * lb), np.log(5.0 * ub))))
        w = S.tls_weights(np.array([r]), mu, np.array([barc]), np.zeros(1, bool))[0]
        assert abs(w - _surrogate_argmin(r, mu, barc)) <= 2e-5, (r, mu, barc, w)


def test_tls_weight_band_edges_and_known_inliers():
    barc = np.array([BARC_09])
    for mu in (1e-3, 0.05, 0.7, 3.0, 40.0):
        lb, ub = mu / (mu + 1) * barc[0], (mu + 1) / mu * barc[0]
        w = lambda r: S.tls_weights(np.array([r]), mu, barc, np.zeros(1, bool))[0]
        assert w(lb) == 1.0 and w(0.5 * lb) == 1.0 and w(0.0) == 1.0
        assert w(ub) == 0.0 and w(2.0 * ub) == 0.0
        assert abs(w(lb * (1 + 1e-9)) - 1.0) < 1e-6 and abs(w(ub * (1 - 1e-9))) < 1e-6  # continuous at both edges
        rs = np.linspace(lb, ub, 101)
        ws = S.tls_weights(rs, mu, np.full(101, barc[0]), np.zeros(101, bool))
        assert np.all(np.diff(ws) <= 0.0)  # non-increasing in the residual
    known = np.array([True, False])
    w = S.tls_weights(np.array([1e6, 1e6]), 0.1, np.full(2, BARC_09), known)
    assert w[0] == 1.0 and w[1] == 0.0


def test_mu_init_remark5():
    rng = np.random.default_rng(1)
    for _ in range(50):
        r0 = rng.exponential(20.0, 30)
        barc = np.full(30, BARC_09)
        mu = S.tls_mu_init(r0, barc)
        ub = (mu + 1) / mu * BARC_09
        assert ub == pytest.approx(2.0 * r0.max(), rel=1e-12)
    # per-factor thresholds: the smallest candidate wins
    r0 = np.array([10.0, 10.0]); barc = np.array([1.0, 4.0])
    assert S.tls_mu_init(r0, barc) == pytest.approx(min(1.0 / 19.0, 4.0 / 16.0))
    # every residual at or below barc/2: no robust iteration (-1)
    assert S.tls_mu_init(np.array([0.5 * BARC_09, 0.1]), np.full(2, BARC_09)) == -1.0
    # very large residuals: floored at 1e-6
    assert S.tls_mu_init(np.array([1e12]), np.array([BARC_09])) == 1e-6


def test_gnc_convergence_rule():
    p = S.GncParams()
    w_frac = np.array([1.0, 0.5])
    assert S.gnc_converged(100.0, 100.0 * (1 + 0.9e-5), w_frac, p)      # cost change below 1e-5
    assert not S.gnc_converged(100.0, 100.0 * (1 + 2e-5), w_frac, p)
    assert S.gnc_converged(1.0, 5.0, np.array([1.0, 1e-4, 1 - 1e-4]), p)  # binary within weightsTol
    assert not S.gnc_converged(1.0, 5.0, np.array([1.0, 1.1e-4]), p)
    # the previous cost is floored at 1e-7 in the denominator
    assert S.gnc_converged(1e-13, 0.0, w_frac, p)
    assert not S.gnc_converged(1e-9, 0.0, w_frac, p)


def _three_priors(z, info_scale=1.0):
    """One pose, three translation priors along x (factor order = prior order)."""
    k = len(z)
    pt = np.zeros((k, 3)); pt[:, 0] = z
    return S.Graph(R=np.eye(3)[None].copy(), t=np.zeros((1, 3)), bi=np.zeros(0, int), bj=np.zeros(0, int),
                   zR=np.zeros((0, 3, 3)), zt=np.zeros((0, 3)), binfo=np.zeros((0, 6, 6)),
                   pk=np.zeros(k, int), pR=np.tile(np.eye(3), (k, 1, 1)), pt=pt,
                   pinfo=np.tile(np.eye(6) * info_scale, (k, 1, 1)))


def _scalar_gnc(z, barc, p: S.GncParams):
    """GNC-TLS on x in R with residuals r_k = (x - z_k)^2 / 2 and the exact
    weighted mean as the inner solve, started at x0 = 0."""
    z = np.asarray(z, float)
    x0 = 0.0
    x = z.mean()                                        # first solve, all weights 1
    r0 = 0.5 * (x0 - z) ** 2
    cands = [barc / (2 * r - barc) for r in r0 if 2 * r - barc > 0]
    mu = max(min(cands), 1e-6) if cands else -1.0
    prev = float(np.sum(0.5 * (x - z) ** 2))
    mus, ws = [mu], []
    for _ in range(p.max_iterations):
        r = 0.5 * (x - z) ** 2
        w = np.empty(len(z))
        for k, rk in enumerate(r):
            if rk <= mu / (mu + 1) * barc:
                w[k] = 1.0
            elif rk >= (mu + 1) / mu * barc:
                w[k] = 0.0
            else:
                w[k] = math.sqrt(barc * mu * (mu + 1) / rk) - mu
        ws.append(w)
        x = float(np.sum(w * z) / np.sum(w))
        cost = float(np.sum(w * 0.5 * (x - z) ** 2))
        if abs(cost - prev) / max(prev, 1e-7) < p.relative_cost_tol or \
                np.all(np.abs(w - np.round(w)) <= p.weights_tol):
            break
        mu *= p.mu_step
        mus.append(mu)
        prev = cost
    return x, mus, ws


def test_gnc_three_factor_trace():
    z = [0.0, 0.1, 5.0]
    g = _three_priors(z)
    # hand-computed start: r0 = (0, 0.005, 12.5) at x0 = 0, so mu0 = barc / (25 - barc);
    # first solve x = 1.7, r = (1.445, 1.28, 5.445)
    mu0 = BARC_09 / (25.0 - BARC_09)
    assert mu0 == pytest.approx(0.270475, abs=2e-6)
    one = S.gnc_optimize(g, S.GncParams(max_iterations=1))
    assert one["mu_trace"][0] == pytest.approx(mu0, rel=1e-12)
    np.testing.assert_allclose(one["weights"], [0.854553, 0.924867, 0.309085], atol=2e-6)
    # full trace against the closed-form inner solve
    p = S.GncParams()
    full = S.gnc_optimize(g, p)
    x, mus, ws = _scalar_gnc(z, BARC_09, p)
    assert full["iterations"] == len(ws)
    np.testing.assert_allclose(full["mu_trace"], mus, rtol=1e-12)
    np.testing.assert_allclose(full["weights"], ws[-1], atol=1e-6)
    assert full["t"][0, 0] == pytest.approx(x, abs=1e-6)
    # the gross outlier ends rejected, the two agreeing priors kept
    assert list(np.round(full["weights"])) == [1.0, 1.0, 0.0]


# ------------------------------------------------------------------ cliques
def _adj(n, edges):
    a = np.zeros((n, n), np.uint8)
    for u, v in edges:
        a[u, v] = a[v, u] = 1
    return a


# triangles {0,1,2}, {3,4,5}, {2,3,4}
G6 = _adj(6, [(0, 1), (0, 2), (1, 2), (2, 3), (3, 4), (3, 5), (4, 5), (2, 4)])


def test_incremental_heuristic_kats():
    # candidates 3, 4, 5 with bound 2: the seed 3 walk leaves the stale
    # buffer [3, 4, 4, 5]; its first 3 entries are returned (size grew 2 -> 3)
    k, idx = O.max_clique_heu_incremental(G6, 3, 2)
    assert k == 3 and list(idx) == [3, 4, 4]
    # bound 3: nothing grows, the wrapper returns 0 (keep the previous set)
    k, idx = O.max_clique_heu_incremental(G6, 3, 3)
    assert k == 0 and len(idx) == 0
    # only vertex 5 is new (degree 2, not pruned by bound 2): its walk
    # 5 -> 4 -> 3 leaves the buffer [5, 3, 4]
    k, idx = O.max_clique_heu_incremental(G6, 1, 2)
    assert k == 3 and list(idx) == [5, 3, 4]
    # the full heuristic on the same graph picks the first seed's triangle
    k, idx = O.max_clique_heu(G6)
    assert k == 3 and list(idx) == [0, 1, 2]
    k, idx = O.max_clique_heu_incremental(G6, 6, 0)
    assert k == 3 and list(idx) == [0, 1, 2]


def test_incremental_with_all_new_equals_full_heuristic():
    rng = np.random.default_rng(3)
    for t in range(60):
        n = int(rng.integers(2, 40))
        a = (rng.random((n, n)) < rng.uniform(0.2, 0.9)).astype(np.uint8)
        a = np.triu(a, 1); a = a + a.T
        kf, idxf = O.max_clique_heu(a)
        ki, idxi = O.max_clique_heu_incremental(a, n, 0)
        assert ki == kf and np.array_equal(idxi, idxf)
        # a previous size at or above the full result: "keep the old set"
        assert O.max_clique_heu_incremental(a, n, kf)[0] == 0


def test_exact_clique_kats():
    # two disjoint triangles: the scan runs from the highest vertex, so the
    # triangle {3,4,5} is met first and {0,1,2} is pruned (degree 2 < 3)
    k, idx = O.max_clique_exact(_adj(6, [(0, 1), (0, 2), (1, 2), (3, 4), (3, 5), (4, 5)]))
    assert k == 3 and list(idx) == [3, 4, 5]
    # hub 4 joined to 0..3 with edges 0-1 and 2-3: U = [0,1,2,3] is consumed
    # from the back, so {4,3,2} is found first, output deepest first: [2,3,4]
    k, idx = O.max_clique_exact(_adj(5, [(4, 0), (4, 1), (4, 2), (4, 3), (0, 1), (2, 3)]))
    assert k == 3 and list(idx) == [2, 3, 4]
    # K4 on 0..3 plus the pendant edge 4-0: seed 4 finds [0, 4] (size 2),
    # seed 3 then finds the K4 in the order 0, 1, 2 below it: [0, 1, 2, 3]
    k, idx = O.max_clique_exact(_adj(5, [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3), (4, 0)]))
    assert k == 4 and list(idx) == [0, 1, 2, 3]
