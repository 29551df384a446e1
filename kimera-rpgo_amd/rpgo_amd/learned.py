"""Learned GNC initialisation: the plain PCM + GNC pipeline, its learned-start
variant and the training-set builder.

Pipeline (one SE(3) pose graph, all stages on the GPU through the C ABI):
  1. PCM: odometry check + pairwise consistency + max clique (rpgo_pcm_run,
     rpgo_max_clique) -> the loop closures that go to the solver (Pcm.h:851-1005);
  2. plain: GNC-TLS over LM on odometry + prior + those loop closures, GNC
     started as GTSAM starts it (RobustSolver.cpp:115-306);
  3. learned: per-loop-closure features (rpgo_lc_features) -> MLP
     (rpgo_mlp_forward) -> initial weights (predicted inliers 1, outliers 0)
     and a starting mu a few GNC steps before the predicted final mu; the
     solve starts there.  The result is accepted only through the cost check
     below; otherwise the plain solve is run and returned instead.

Cost check (plain GNC is the arbiter).  The learned result must
  * be a GNC fixed point: every final weight equals the TLS decision at the
    final poses (report.n_tls_disagree == 0) and the weights are binary;
  * not raise the TLS objective  sum_known e_k + sum_lc min(e_k, barc)
    (e_k = 1/2 e^T Lambda e) above its value at GNC's own first solve (all
    weights 1), the point plain GNC starts its descent from;
  * when plain GNC's result is at hand (evaluation: the bench and the tests),
    keep plain's inlier set with poses within 1e-5 m / 1e-5 rad of plain's.

Labels for training are the plain method's own outputs on the same graphs:
the final GNC weight of each loop closure and log of the final mu.
"""
from __future__ import annotations

import os
from dataclasses import dataclass

import numpy as np

from . import _capi, problem, synth

HERE = os.path.dirname(os.path.abspath(__file__))
MODEL_PATH = os.path.join(os.path.dirname(HERE), "weights", "learned_init.npz")


@dataclass
class PcmResult:
    accepted: np.ndarray   # accepted loop closures (original indices)
    clique: np.ndarray     # clique members (indices into accepted)
    inliers: np.ndarray    # original indices of the clique members, clique order
    ms: float


def pcm_stage(ctx, gr, odom_thr=-1.0, lc_thr=5.0, precision=_capi.PCM_FP32_RECHECK) -> PcmResult:
    ctx.pcm_upload(gr.dim, gr.pose0, gr.odom, gr.odom_cov, gr.lc_from, gr.lc_to, gr.lc_pose, gr.lc_cov)
    st = ctx.pcm_run(odom_thr, lc_thr, precision=precision)
    k, idx, ms = ctx.max_clique(_capi.CLIQUE_HEU_REF)
    acc = ctx.pcm_download()["accepted"]
    clique = np.asarray(idx[:max(k, 0)], np.int64)
    return PcmResult(acc, clique, acc[clique], st.ms_total + ms)


def solve_plain(ctx, gr, inliers, **solve_params):
    arrs = problem.solve_arrays(gr, inliers)
    poses, w, rep = ctx.solve(*arrs[:-1], use_gnc=1, known_inliers=arrs[-1], **solve_params)
    return poses, w, rep, arrs[-1]


@dataclass
class LearnedModel:
    params: np.ndarray  # RPGO_MLP_NPARAMS float32
    norm: np.ndarray    # mean[16], 1/std[16]
    back_steps: int = 2

    def save(self, path=MODEL_PATH):
        os.makedirs(os.path.dirname(path), exist_ok=True)
        np.savez(path, params=self.params, norm=self.norm, back_steps=self.back_steps)

    @staticmethod
    def load(path=MODEL_PATH):
        z = np.load(path, allow_pickle=False)
        return LearnedModel(z["params"].astype(np.float32), z["norm"].astype(np.float32), int(z["back_steps"]))


def predict(ctx, model: LearnedModel, feats):
    """(initial weights in {0,1}, starting mu) for the rows of feats."""
    ctx.mlp_set_params(model.params, model.norm)
    y = ctx.mlp_forward(feats)
    w0 = (y[:, 0] > 0.0).astype(np.float64)
    mu_final = float(np.exp(np.median(y[:, 1]))) if len(y) else 1.0
    return w0, mu_final


GNC_ALPHA = 0.9  # RobustSolverParams default (SolverParams.h:60-68)


def clique_features(ctx, gr, pcm: PcmResult, with_cost=False, **solve_params):
    """Features of the solver's loop closures (clique order), with the
    residuals after GNC's first solve (LM with every weight 1).  Returns
    (features, device ms of the first solve[, TLS objective there])."""
    arrs = problem.solve_arrays(gr, pcm.inliers)
    nk = arrs[-1]
    _, _, rep1 = ctx.solve(*arrs[:-1], **solve_params)
    err = ctx.solve_errors(len(arrs[1]))
    resid = np.zeros(len(pcm.accepted))
    resid[pcm.clique] = err[nk:]
    barc = _capi.lib().rpgo_gnc_threshold(GNC_ALPHA, 6)
    feats = ctx.lc_features(pcm.clique, resid, barc)[pcm.clique]
    if with_cost:
        return feats, rep1.ms_total, tls_cost(err, nk, barc)
    return feats, rep1.ms_total


def tls_cost(err, nk, barc):
    """GNC-TLS objective: known inliers at full cost, loop closures truncated at barc."""
    return float(np.sum(err[:nk]) + np.sum(np.minimum(err[nk:], barc)))


@dataclass
class LearnedOutcome:
    poses: np.ndarray
    weights: np.ndarray
    report: object
    ms: float          # device ms: first solve + learned-start solve (+ plain solve on fallback)
    accepted: bool     # the learned result passed the cost check
    check: dict        # the quantities the check looked at


POSE_TOL = 1e-5


def solve_learned(ctx, gr, pcm: PcmResult, model: LearnedModel, mu_step=1.4, plain=None, **solve_params):
    """Learned-start GNC, accepted through the cost check, plain GNC as the
    fallback.  plain: (poses, weights, report) of plain GNC on the same
    problem when already computed (evaluation); it is then also the judge of
    the inlier set and the poses, and the fallback reuses it (its device time
    is added to ms)."""
    feats, ms, tls_first = clique_features(ctx, gr, pcm, with_cost=True, **solve_params)
    w0, mu_final = predict(ctx, model, feats)
    arrs = problem.solve_arrays(gr, pcm.inliers)
    nk = arrs[-1]
    init_w = np.concatenate([np.ones(nk), w0])
    mu0 = max(mu_final / mu_step ** model.back_steps, 1e-6)
    poses, w, rep = ctx.solve(*arrs[:-1], use_gnc=1, known_inliers=nk, init_weights=init_w, init_mu=mu0,
                              **solve_params)
    ms += rep.ms_total
    barc = _capi.lib().rpgo_gnc_threshold(GNC_ALPHA, 6)
    tls_l = tls_cost(ctx.solve_errors(len(arrs[1])), nk, barc)
    binary = bool(np.all(np.abs(w - np.rint(w)) <= 1e-4))
    check = {"tls_fixed_point": rep.n_tls_disagree == 0, "binary_weights": binary,
             "tls_cost_learned": tls_l, "tls_cost_first_solve": tls_first,
             "tls_not_above_first_solve": tls_l <= tls_first}
    ok = check["tls_fixed_point"] and binary and check["tls_not_above_first_solve"]
    if plain is not None:
        pp, wp, rp = plain
        dt = float(np.abs(pp[:, 9:] - poses[:, 9:]).max())
        same = bool(np.array_equal(np.rint(wp), np.rint(w)))
        check.update({"same_inlier_set_as_plain": same, "max_translation_diff_vs_plain_m": dt,
                      "tls_cost_plain": tls_cost(_errors_at(ctx, arrs, pp), nk, barc)})
        ok = ok and same and dt <= POSE_TOL
    if ok:
        return LearnedOutcome(poses, w, rep, ms, True, check)
    if plain is None:
        poses, w, rep, _ = solve_plain(ctx, gr, pcm.inliers, **solve_params)
    else:
        poses, w, rep = plain
    return LearnedOutcome(poses, w, rep, ms + rep.ms_total, False, check)


def _errors_at(ctx, arrs, poses):
    """1/2 e^T Lambda e per factor of the last solve's problem at other poses."""
    return ctx.solve_errors(len(arrs[1]), poses=poses)


def graph_rows(ctx, gr, **solve_params):
    """Features and labels of one graph's solver loop closures (plain method)."""
    pcm = pcm_stage(ctx, gr)
    if len(pcm.clique) == 0:
        return None
    feats, _ = clique_features(ctx, gr, pcm, **solve_params)
    _, w, rep, nk = solve_plain(ctx, gr, pcm.inliers, **solve_params)
    wl = w[nk:]
    mu = rep.mu_final if rep.mu_final > 0 else 1.0
    target = np.stack([wl, np.full(len(wl), np.log(mu))], 1).astype(np.float32)
    return feats, target, rep


def training_graph(seed, rng=None):
    """Seeded training graph: 1k-4k poses, loop closures 10-30 % of poses,
    10-60 % outliers (the generator of the 50k benchmark graph)."""
    rng = np.random.default_rng(seed) if rng is None else rng
    n = int(np.exp(rng.uniform(np.log(1000), np.log(4000))))
    n_lc = int(n * rng.uniform(0.1, 0.3))
    of = float(rng.uniform(0.1, 0.6))
    return synth.se3_helix(n_poses=n, n_lc=n_lc, outlier_frac=of, seed=int(seed))


def init_params(seed=0):
    """He-uniform weights, zero biases, in the RPGO_MLP parameter layout."""
    rng = np.random.default_rng(seed)
    F, H, O = _capi.LC_FEATURES, _capi.MLP_HIDDEN, _capi.MLP_OUT
    W1 = rng.uniform(-1, 1, (F, H)) * np.sqrt(6.0 / F)
    W2 = rng.uniform(-1, 1, (H, H)) * np.sqrt(6.0 / H)
    W3 = rng.uniform(-1, 1, (H, O)) * np.sqrt(6.0 / H)
    return np.concatenate([W1.ravel(), np.zeros(H), W2.ravel(), np.zeros(H), W3.ravel(),
                           np.zeros(O)]).astype(np.float32)


def fit_norm(x):
    mean = x.mean(0)
    std = x.std(0)
    std[std < 1e-6] = 1.0
    return np.concatenate([mean, 1.0 / std]).astype(np.float32)
