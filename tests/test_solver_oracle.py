"""Pin the FP64 GNC / LM oracle (oracle/solver_oracle.py).

Checks chosen so that a dropped term, a wrong sign or a transposed operand
fails one of them: Jacobians against central differences of the error,
exact recovery of a noiseless graph, first-order optimality of the LM result,
and the GNC weight outcomes of tests/testGnc.cpp:25-103 on robot_a.g2o.
"""
import os

import numpy as np
import pytest

import solver_oracle as S
from rpgo_amd import g2o, synth

DATA = os.path.join(os.path.dirname(__file__), "golden", "data")


def between_error(Ri, ti, Rj, tj, zR, zt):
    RiT = Ri.T
    Rr, tr = RiT @ Rj, RiT @ (tj - ti)
    return S.se3_log((zR.T @ Rr)[None], (zR.T @ (tr - zt))[None])[0]


def test_between_jacobians_match_finite_differences():
    rng = np.random.default_rng(0)
    for _ in range(20):
        Ri, Rj, zR = (S.rot_exp(rng.normal(0, 1.0, 3)) for _ in range(3))
        ti, tj, zt = (rng.normal(0, 3, 3) for _ in range(3))
        g = S.Graph(R=np.stack([Ri, Rj]), t=np.stack([ti, tj]), bi=np.array([0]), bj=np.array([1]),
                    zR=zR[None], zt=zt[None], binfo=np.eye(6)[None])
        # J from the oracle's Gauss-Newton blocks with identity information:
        # H_ii = Ji^T Ji ... recover Ji, Jj from a direct recomputation
        e0 = between_error(Ri, ti, Rj, tj, zR, zt)
        Jr = S.se3_logmap_derivative(e0[None])[0]
        Rr, tr = Ri.T @ Rj, Ri.T @ (tj - ti)
        Ji = -Jr @ S.adjoint(Rr.T, -(Rr.T @ tr))
        Jj = Jr
        h = 1e-6
        for k in range(6):
            d = np.zeros(6); d[k] = h
            Rp, tp = S.retract(Ri[None], ti[None], d)
            Rm, tm = S.retract(Ri[None], ti[None], -d)
            num = (between_error(Rp[0], tp[0], Rj, tj, zR, zt) - between_error(Rm[0], tm[0], Rj, tj, zR, zt)) / (2 * h)
            np.testing.assert_allclose(Ji[:, k], num, atol=2e-6)
            Rp, tp = S.retract(Rj[None], tj[None], d)
            Rm, tm = S.retract(Rj[None], tj[None], -d)
            num = (between_error(Ri, ti, Rp[0], tp[0], zR, zt) - between_error(Ri, ti, Rm[0], tm[0], zR, zt)) / (2 * h)
            np.testing.assert_allclose(Jj[:, k], num, atol=2e-6)
        # the assembled system agrees with J^T J and -J^T e
        H, b = S.linearize(g, g.R, g.t, np.ones(1))
        Hd = H.toarray()
        np.testing.assert_allclose(Hd[:6, :6], Ji.T @ Ji, atol=1e-9)
        np.testing.assert_allclose(Hd[:6, 6:], Ji.T @ Jj, atol=1e-9)
        np.testing.assert_allclose(b[6:], -Jj.T @ e0, atol=1e-9)


def test_exp_log_and_logmap_derivative():
    rng = np.random.default_rng(1)
    xi = rng.normal(0, 0.8, (50, 6))
    R, t = S.se3_exp(xi)
    np.testing.assert_allclose(S.se3_log(R, t), xi, atol=1e-10)
    # d/dd Log(T Exp(d)) at d=0 equals LogmapDerivative(Log T)
    h = 1e-6
    for k in range(5):
        J = S.se3_logmap_derivative(xi[k][None])[0]
        for c in range(6):
            d = np.zeros(6); d[c] = h
            Rp, tp = S.retract(R[k][None], t[k][None], d)
            Rm, tm = S.retract(R[k][None], t[k][None], -d)
            num = (S.se3_log(Rp, tp)[0] - S.se3_log(Rm, tm)[0]) / (2 * h)
            np.testing.assert_allclose(J[:, c], num, atol=1e-6)


def _graph_from(gr: synth.PoseGraph, use_gt=False, exact=False):
    Rv, tv = synth.se3_unpack(gr.gt if use_gt else gr.values)
    zR, zt = synth.se3_unpack(np.concatenate([gr.odom, gr.lc_pose]))
    bi = np.concatenate([np.arange(gr.n_poses - 1), gr.lc_from])
    bj = np.concatenate([np.arange(1, gr.n_poses), gr.lc_to])
    if exact:  # measurements replaced by the true relative poses
        Rg, tg = synth.se3_unpack(gr.gt)
        zR, zt = synth.se3_rel(Rg[bi], tg[bi], Rg[bj], tg[bj])
    cov = np.concatenate([gr.odom_cov, gr.lc_cov]).reshape(-1, 6, 6)
    info = np.linalg.inv(cov)
    R0, t0 = synth.se3_unpack(gr.gt[:1])
    return S.Graph(R=Rv.copy(), t=tv.copy(), bi=bi, bj=bj, zR=zR, zt=zt, binfo=info,
                   pk=np.array([0]), pR=R0, pt=t0, pinfo=(np.eye(6) / 0.01)[None])


def test_lm_recovers_noiseless_graph():
    gr = synth.se3_helix(n_poses=300, n_lc=40, outlier_frac=0.0, seed=3)
    g = _graph_from(gr, use_gt=True, exact=True)
    # perturb the initial values
    rng = np.random.default_rng(2)
    g.R = g.R @ S.rot_exp(rng.normal(0, 0.02, (g.N, 3)))
    g.t = g.t + rng.normal(0, 0.2, (g.N, 3))
    R, t, err, it = S.least_squares(g, np.ones(g.F))
    Rgt, tgt = synth.se3_unpack(gr.gt)
    assert err < 1e-12
    np.testing.assert_allclose(t, tgt, atol=1e-6)
    np.testing.assert_allclose(R, Rgt, atol=1e-7)


def test_lm_first_order_optimality():
    gr = synth.se3_helix(n_poses=400, n_lc=60, outlier_frac=0.0, seed=4)
    g = _graph_from(gr)
    R, t, err, it = S.least_squares(g, np.ones(g.F))
    H, b = S.linearize(g, R, t, np.ones(g.F))
    assert np.abs(b).max() < 1e-6 * max(1.0, np.abs(H.diagonal()).max()) ** 0.5
    # the optimum is below the error of the ground truth and of the initial values
    Rgt, tgt = synth.se3_unpack(gr.gt)
    assert err <= float(S.factor_errors(g, Rgt, tgt).sum()) + 1e-9
    assert err < float(S.factor_errors(g, g.R, g.t).sum())


def test_barc_sq_golden():
    """chi2 quantiles quoted for GncOptimizer's barcSq (0.5 * chi2inv(alpha, 6))."""
    g = S.Graph(R=np.eye(3)[None], t=np.zeros((1, 3)), bi=np.zeros(0, int), bj=np.zeros(0, int),
                zR=np.zeros((0, 3, 3)), zt=np.zeros((0, 3)), binfo=np.zeros((0, 6, 6)),
                pk=np.array([0]), pR=np.eye(3)[None], pt=np.zeros((1, 3)), pinfo=np.eye(6)[None])
    for a, v in [(0.01, 0.436045), (0.9, 5.322320), (0.99, 8.405947)]:
        assert abs(S.barc_sq(g, S.GncParams(inlier_threshold=a))[0] - v) < 1e-5


def robot_a_graph():
    gg = g2o.read(os.path.join(DATA, "robot_a.g2o"))
    keys = gg.keys
    idx = {k: n for n, k in enumerate(keys)}
    R = np.stack([gg.poses[k][:9].reshape(3, 3) for k in keys])
    t = np.stack([gg.poses[k][9:] for k in keys])
    odom = [e for e in gg.edges if e[1] == e[0] + 1]
    lcs = [e for e in gg.edges if e[1] != e[0] + 1]
    edges = odom + lcs  # Pcm.h:977-1005 order: odometry, special (prior), loop closures
    bi = np.array([idx[e[0]] for e in edges]); bj = np.array([idx[e[1]] for e in edges])
    zR = np.stack([e[2][:9].reshape(3, 3) for e in edges]); zt = np.stack([e[2][9:] for e in edges])
    info = np.stack([e[3] for e in edges])
    order = [("b", k) for k in range(len(odom))] + [("p", 0)] + \
            [("b", len(odom) + k) for k in range(len(lcs))]
    g = S.Graph(R=R, t=t, bi=bi, bj=bj, zR=zR, zt=zt, binfo=info, pk=np.array([0]),
                pR=R[:1].copy(), pt=t[:1].copy(), pinfo=(np.eye(6) / 0.01)[None], order=order)
    return g, len(odom) + 1


@pytest.mark.parametrize("alpha,expect", [(0.01, 0), (0.99, 3)])
def test_gnc_robot_a_weights(alpha, expect):
    """tests/testGnc.cpp:25-103: alpha 0.01 rejects the 3 loop closures of
    robot_a, alpha 0.99 accepts them (LM inner solver)."""
    g, n_known = robot_a_graph()
    assert g.F == 53 and n_known == 50
    res = S.gnc_optimize(g, S.GncParams(inlier_threshold=alpha, known_inliers=np.arange(n_known)))
    w = res["weights"]
    assert w[50:].sum() == expect
    assert w.sum() == 50 + expect


def test_gnc_rejects_planted_outliers():
    gr = synth.se3_helix(n_poses=400, n_lc=60, outlier_frac=0.3, seed=7)
    g = _graph_from(gr)
    n_known = gr.n_poses - 1
    order = [("b", k) for k in range(n_known)] + [("p", 0)] + [("b", n_known + k) for k in range(gr.n_lc)]
    g = S.Graph(**{k: getattr(g, k) for k in ("R", "t", "bi", "bj", "zR", "zt", "binfo", "pk", "pR", "pt", "pinfo")},
                order=order)
    res = S.gnc_optimize(g, S.GncParams(inlier_threshold=0.9, known_inliers=np.arange(n_known + 1)))
    w_lc = res["weights"][n_known + 1:] > 0.5
    # every gross outlier is rejected; TLS at alpha = 0.9 also cuts the true
    # loop closures whose residual lies above the 0.9 chi^2_6 quantile
    assert not np.any(w_lc & gr.is_outlier)
    assert (w_lc & ~gr.is_outlier).sum() >= 0.75 * (~gr.is_outlier).sum()
    assert np.all(np.isin(res["weights"], [0.0, 1.0]))
    assert res["iterations"] > 5
