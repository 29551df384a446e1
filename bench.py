"""PCM pair-check throughput on one MI355X (device-resident inputs).

Workload (default): SE(3) helix, 5,000 poses, 1,000 loop closures (20 %
outliers), odometry check off (odom_threshold = -1) so all 499,500 unique
pairs (i<j) are evaluated, lc_threshold = 5.0.  A step is one pass of the PCM
pairwise stage over the uploaded graph: FP32 consistency check of every pair,
FP64 replay of the pairs near the threshold, adjacency bit packing, consistent
pair count read back.  The odometry fold and check run once at ingest (the
reference folds the trajectory as odometry arrives) and are reported apart.
The timed steps run without the library's stage timers (rpgo_ctx_set_stage_
timing off); the kernel and stage times come from a second pass of the same
number of steps with the timers on.

Prints one JSON line (metric, value, roofline of the dominant kernel, CPU
baseline of the FP64 oracle on the host cores).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (os.path.join(ROOT, "kimera-rpgo_amd"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

import numpy as np  # noqa: E402

# Algorithmic FP32 work per SE(3) pair check (5 congruences, 6 group ops,
# 5 adjoints, sym add/sub, 2 PSD checks, Log, Cholesky Mahalanobis); SE(2) ~720.
FLOP_PER_PAIR = {3: 4540.0, 2: 720.0}
PEAK_TFLOPS = {"fp32": 157.3, "fp64": 78.6}
# HBM bytes per k_pairs32 launch on se3_5k, from the rocprofv3 FETCH_SIZE and
# WRITE_SIZE passes written up in this file (scaled by a plain copy's counts)
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_pairs32_se3_5k_traffic.md")


def measured_traffic():
    """bytes per launch recorded in TRAFFIC_FILE ('traffic_bytes_per_launch: N'), or None"""
    try:
        with open(TRAFFIC_FILE) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None

WORKLOADS = {
    "se3_5k": dict(kind="se3_helix", n_poses=5000, n_lc=1000, outlier_frac=0.2, odom_thr=-1.0, lc_thr=5.0),
    "se3_50k": dict(kind="se3_helix", n_poses=50000, n_lc=10000, outlier_frac=0.5, odom_thr=-1.0, lc_thr=5.0),
}


def cpu_baseline(g, odom_thr, lc_thr, budget_s=15.0):
    """FP64 oracle (single core) on a bounded row sample of the same graph."""
    import oracle as O

    O.set_threads(1)  # the reference is single-threaded
    t0 = time.perf_counter()
    traj = O.fold(g.dim, g.pose0, g.odom, g.odom_cov)
    t_fold = time.perf_counter() - t0
    m = g.n_lc
    # rows j in [0, j1): pairs j1*(j1-1)/2; grow j1 until the budget is spent
    pairs = 0
    t_pairs = 0.0
    j0 = 1
    step = 64
    while j0 < m and t_pairs < budget_s:
        j1 = min(m, j0 + step)
        t = time.perf_counter()
        O.pcm_dist_rows(g.dim, traj, g.lc_from[:j1], g.lc_to[:j1], g.lc_pose[:j1], g.lc_cov[:j1], j0, j1)
        t_pairs += time.perf_counter() - t
        pairs += (j1 * (j1 - 1) - j0 * (j0 - 1)) // 2
        j0 = j1
        step *= 2
    return dict(value=pairs / t_pairs, unit="pair-checks/s", cores=1, kind="port",
                sample=f"FP64 oracle, rows 1..{j0 - 1} of the workload ({pairs} pairs, "
                       f"{t_pairs:.1f} s; fold {t_fold * 1e3:.1f} ms)")


def robust_solve(ctx, gl, **solve_params):
    """Robust-solve ms on the 50,000-pose graph: PCM (FP32 + FP64 recheck, the
    reference's heuristic clique) then GNC-TLS over LM on odometry + prior +
    the clique's loop closures; plain GNC (started as GTSAM starts it) against
    the learned start (features + MLP + GNC from the predicted weights and mu,
    plain GNC as the fallback), on the same GPU.  Device times (HIP events)."""
    from rpgo_amd import _capi, learned

    t0 = time.perf_counter()
    pcm = learned.pcm_stage(ctx, gl)
    ms_pcm = pcm.ms
    poses_p, w_p, rep_p, nk = learned.solve_plain(ctx, gl, pcm.inliers, **solve_params)
    true_in = ~gl.is_outlier[pcm.inliers]

    def quality(poses, w):
        """kept true inliers, kept outliers, max translation error vs ground truth (m)"""
        keep = np.rint(w[nk:]) > 0.5
        return {"true_inliers_kept": int((keep & true_in).sum()), "outliers_kept": int((keep & ~true_in).sum()),
                "max_translation_error_vs_truth_m": float(np.abs(poses[:, 9:] - gl.gt[:, 9:]).max())}

    res = {"workload": f"SE(3) helix {gl.n_poses} poses, {gl.n_lc} loop closures, "
                       f"{int(round(100 * gl.is_outlier.mean()))}% outliers",
           "pcm_ms": ms_pcm, "solver_loop_closures": int(len(pcm.inliers)),
           "outliers_after_pcm": int(gl.is_outlier[pcm.inliers].sum()),
           "plain": {"ms": rep_p.ms_total, "gnc_iterations": rep_p.gnc_iterations,
                     "lm_iterations": rep_p.lm_iterations, "pcg_iterations": rep_p.pcg_iterations,
                     "inliers": rep_p.n_inliers, **quality(poses_p, w_p)},
           "end_to_end_plain_ms": ms_pcm + rep_p.ms_total}
    if os.path.exists(learned.MODEL_PATH):
        model = learned.LearnedModel.load()
        poses_l, w_l, rep_l, fb, ms_l = learned.solve_learned(ctx, gl, pcm, model, **solve_params)
        res["learned"] = {"ms": ms_l, "gnc_iterations": rep_l.gnc_iterations,
                          "lm_iterations": rep_l.lm_iterations, "fallback": bool(fb),
                          "tls_fixed_point": rep_l.n_tls_disagree == 0,
                          "same_inlier_set": bool(np.array_equal(np.rint(w_p), np.rint(w_l))),
                          "inlier_decisions_differing": int((np.rint(w_p) != np.rint(w_l)).sum()),
                          "max_pose_diff_vs_plain_m": float(np.abs(poses_p[:, 9:] - poses_l[:, 9:]).max()),
                          **quality(poses_l, w_l)}
        res["end_to_end_learned_ms"] = ms_pcm + ms_l
        res["speedup_learned_vs_plain"] = (ms_pcm + rep_p.ms_total) / (ms_pcm + ms_l)
    else:
        res["learned"] = None
    res["wall_s"] = time.perf_counter() - t0
    return res


def batch_c4(ctx, seed=0, n_graphs=256, steps=3):
    """C4 batch: 256 independent SE(3) graphs, poses log-uniform in [200, 20000],
    loop closures 0.1-0.3 of the poses, outliers 5-40 %, odometry check off and
    lc_threshold 5.0.  One launch per stage (fold, odometry check, compaction,
    pairs, symmetrize, clique of every graph), captured in a hipGraph, launch
    shapes picked by the autotuner on the first call.  Device times."""
    from rpgo_amd import _capi, synth

    rng = np.random.default_rng(seed)
    graphs = []
    for k in range(n_graphs):
        n = int(np.exp(rng.uniform(np.log(200), np.log(20000))))
        graphs.append(synth.se3_helix(n_poses=n, n_lc=max(1, int(n * rng.uniform(0.1, 0.3))),
                                      outlier_frac=float(rng.uniform(0.05, 0.4)), seed=seed * 1000 + k))
    flags = _capi.BATCH_GRAPH | _capi.BATCH_AUTOTUNE
    first = ctx.pcm_batch_run(graphs, -1.0, 5.0, flags=flags)
    staged = ctx.pcm_batch_run(graphs, -1.0, 5.0)
    t0 = time.perf_counter()
    reps = [ctx.pcm_batch_run(graphs, -1.0, 5.0, flags=flags) for _ in range(steps)]
    wall = (time.perf_counter() - t0) / steps
    ms = float(np.mean([r["stats"].ms_total for r in reps]))
    st = reps[-1]["stats"]
    same = all(np.array_equal(reps[-1]["clique_size"], r["clique_size"]) for r in (first, staged))
    ss = staged["stats"]
    adj_ms = ss.ms_fold + ss.ms_odom + ss.ms_pairs
    return {"graphs": n_graphs, "poses": int(sum(g.n_poses for g in graphs)),
            "loop_closures": int(sum(g.n_lc for g in graphs)), "pairs": int(st.n_pairs),
            "ms_per_batch_device": ms, "ms_per_batch_wall": wall * 1e3,
            "pair_checks_per_s": st.n_pairs / (ms * 1e-3),
            "pair_checks_per_s_adjacency_stage": st.n_pairs / (adj_ms * 1e-3),
            "stages_ms": {"fold": staged["stats"].ms_fold, "odom": staged["stats"].ms_odom,
                          "pairs": staged["stats"].ms_pairs, "clique": staged["stats"].ms_clique,
                          "total_staged": staged["stats"].ms_total},
            "tuned": {"pair_waves": st.pair_waves, "clique_waves": st.clique_waves,
                      "clique_prefetch": st.clique_prefetch, "ms_tune": first["stats"].ms_tune},
            "graph_replayed": bool(st.graph_replayed), "results_identical_across_modes": bool(same)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="se3_5k", choices=sorted(WORKLOADS))
    ap.add_argument("--precision", default="fp32", choices=["fp64", "fp32"])
    ap.add_argument("--no-large", dest="large", action="store_false",
                    help="skip the extra 50,000-pose pass reported beside the main line")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-solve", dest="solve", action="store_false",
                    help="skip the robust-solve measurement on the 50,000-pose graph")
    ap.add_argument("--no-batch", dest="batch", action="store_false",
                    help="skip the 256-graph batch measurement")
    ap.add_argument("--solve-poses", type=int, default=50000,
                    help="poses of the robust-solve graph (SE(3) helix, loop closures 1/5 of the poses, 50 %% outliers)")
    ap.add_argument("--pcg-graphs", type=int, default=1, choices=[0, 1],
                    help="replay PCG iterations as captured HIP graphs (1) or launch them directly (0)")
    ap.add_argument("--seed", type=int, default=0)
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {a.gpus} is not supported "
              "(the multi-GPU sharding is a design only)", file=sys.stderr)
        sys.exit(2)

    from rpgo_amd import _capi, synth

    w = WORKLOADS[a.workload]
    g = synth.se3_helix(n_poses=w["n_poses"], n_lc=w["n_lc"], outlier_frac=w["outlier_frac"], seed=a.seed)
    prec = _capi.PCM_FP64_REPLAY if a.precision == "fp64" else _capi.PCM_FP32_RECHECK
    ctx = _capi.Context(0)
    ctx.pcm_upload(g.dim, g.pose0, g.odom, g.odom_cov, g.lc_from, g.lc_to, g.lc_pose, g.lc_cov)
    # ingest: fold, odometry check, FP32 tables (outside the timed region, as in
    # the reference where the trajectory is folded when odometry arrives)
    ingest = ctx.pcm_run(w["odom_thr"], w["lc_thr"], precision=prec)
    # the timed steps run as a caller would, without stage timers (their HIP
    # events cost a 5k pass about 20 us); a second pass of the same steps with
    # the timers on gives the kernel and stage times of the roofline
    ctx.set_stage_timing(False)
    for _ in range(a.warmup):
        ctx.pcm_pairs(w["lc_thr"], precision=prec)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        last = ctx.pcm_pairs(w["lc_thr"], precision=prec)
    ctx.sync()
    wall = time.perf_counter() - t0
    ms_step = wall * 1e3 / a.steps
    ctx.set_stage_timing(True)
    stats = [ctx.pcm_pairs(w["lc_thr"], precision=prec) for _ in range(a.steps)]
    ctx.sync()
    assert stats[-1].n_consistent == last.n_consistent
    pairs = stats[-1].n_pairs
    ms_pairs = float(np.mean([s.ms_pairs for s in stats]))
    ms_total = float(np.mean([s.ms_total for s in stats]))
    flop = FLOP_PER_PAIR[g.dim] * pairs
    achieved = flop / (ms_pairs * 1e-3) / 1e12
    peak = PEAK_TFLOPS["fp64" if a.precision == "fp64" else "fp32"]
    out = {
        "metric": "PCM pair-checks/s",
        "value": pairs / (ms_step * 1e-3),
        "unit": "pair-checks/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if a.precision == "fp64" else "f32+f64",
        "data": "synthetic",
        "config": {"workload": f"{a.workload}: SE(3) helix {w['n_poses']} poses, {w['n_lc']} loop closures, "
                               f"{int(w['outlier_frac'] * 100)}% outliers, odom_thr {w['odom_thr']}, "
                               f"lc_thr {w['lc_thr']}", "pairs": pairs, "precision": a.precision,
                   "parallelism": "single GPU"},
        "roofline": {"bound": "valu", "kernel": "k_pairs64" if a.precision == "fp64" else "k_pairs32",
                     "achieved": achieved, "peak": peak, "unit": "TFLOP/s", "frac": achieved / peak,
                     "traffic": measured_traffic() if (a.workload == "se3_5k" and a.precision == "fp32") else None,
                     "flop_per_pair": FLOP_PER_PAIR[g.dim], "ms_kernel": ms_pairs},
        "stages_ms": {"pairs_kernel": ms_pairs, "recheck_fp64": float(np.mean([s.ms_recheck for s in stats])),
                      "pairwise_stage_device": ms_total,
                      "ingest_fold_fp64": ingest.ms_fold, "ingest_odom_check": ingest.ms_odom,
                      "ingest_total_device": ingest.ms_total},
        "n_consistent": stats[-1].n_consistent,
        "n_rechecked": stats[-1].n_rechecked,
        "n_flipped": stats[-1].n_flipped,
    }
    k, _, ms_clq = ctx.max_clique(_capi.CLIQUE_HEU_REF)
    out["clique_heu_ref"] = {"size": k, "ms": ms_clq}
    if a.large:
        wl = WORKLOADS["se3_50k"]
        gl = synth.se3_helix(n_poses=wl["n_poses"], n_lc=wl["n_lc"], outlier_frac=wl["outlier_frac"],
                             seed=a.seed)
        ctx.pcm_upload(gl.dim, gl.pose0, gl.odom, gl.odom_cov, gl.lc_from, gl.lc_to, gl.lc_pose, gl.lc_cov)
        ing = ctx.pcm_run(wl["odom_thr"], wl["lc_thr"], precision=prec)
        ctx.pcm_pairs(wl["lc_thr"], precision=prec)
        t1 = time.perf_counter()
        sl = [ctx.pcm_pairs(wl["lc_thr"], precision=prec) for _ in range(3)]
        wl_ms = (time.perf_counter() - t1) * 1e3 / 3
        mk = float(np.mean([s.ms_pairs for s in sl]))
        kl, _, msl = ctx.max_clique(_capi.CLIQUE_HEU_REF)
        out["se3_50k"] = {
            "pairs": sl[-1].n_pairs, "pair_checks_per_s": sl[-1].n_pairs / (wl_ms * 1e-3),
            "ms_per_pass": wl_ms, "ms_pairs_kernel": mk,
            "tflops_kernel": FLOP_PER_PAIR[3] * sl[-1].n_pairs / (mk * 1e-3) / 1e12,
            "ms_recheck": float(np.mean([s.ms_recheck for s in sl])), "n_rechecked": sl[-1].n_rechecked,
            "n_flipped": sl[-1].n_flipped, "ingest_ms": ing.ms_total, "ingest_fold_ms": ing.ms_fold,
            "clique_heu_ref": {"size": kl, "ms": msl}}
        if a.solve:
            gs = gl if a.solve_poses == wl["n_poses"] else synth.se3_helix(
                n_poses=a.solve_poses, n_lc=a.solve_poses // 5, outlier_frac=0.5, seed=a.seed)
            out["robust_solve_50k" if gs is gl else f"robust_solve_{a.solve_poses}"] = robust_solve(ctx, gs, pcg_graphs=a.pcg_graphs)
    if a.batch:
        out["batch_c4"] = batch_c4(ctx, seed=a.seed)
    ctx.close()
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(g, w["odom_thr"], w["lc_thr"])
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
