"""Writes the FP64 oracle's robust-solve fixtures (test data, run in the build
container; the GPU tests and bench.py read the .npz files only).

For each workload: the seeded SE(3) helix graph (rpgo_amd.synth), the FP64
oracle's PCM (fold, pairwise check, reference heuristic clique), then the
oracle's GNC-TLS over LM (oracle/solver_oracle.gnc_optimize: GTSAM
GncOptimizer semantics, RobustSolver.cpp:130-180) on odometry + prior + the
clique's loop closures in the order Pcm::buildGraphToOptimize gives them
(Pcm.h:977-1005; rpgo_amd.problem.solve_arrays).  Stored: the clique's
loop-closure indices, the final weights, poses (unit quaternion of R and t),
GNC iterations and cost, and the converged optimum of those weights (LM from
the GNC result to a relative cost decrease of 1e-12), stored as its FP32
offsets from the GNC result (dq_opt, dt_opt: about 1e-6 m, so FP32 keeps them
to 1e-13 m) and its cost (cost_opt).

  python tests/golden/make_solve_fixtures.py [se3_5k] [se3_50k]
  python tests/golden/make_solve_fixtures.py --optimum se3_50k   (adds the
      converged optimum to an existing file)

se3_5k takes about a minute, se3_50k about an hour on one core.
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (os.path.join(ROOT, "kimera-rpgo_amd"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

import oracle as O  # noqa: E402
import solver_oracle as S  # noqa: E402
from rpgo_amd import problem, synth  # noqa: E402

WORKLOADS = {
    # n_poses, n_lc, outlier fraction, seed; odometry check off, lc_threshold 5.0
    "se3_5k": (5000, 1000, 0.2, 0),
    "se3_50k": (50000, 10000, 0.5, 0),
}
LC_THR = 5.0
GNC_ALPHA = 0.9  # RobustSolverParams default (SolverParams.h:60-68)


def pcm_inliers(g, threads=8):
    O.set_threads(threads)
    traj = O.fold(g.dim, g.pose0, g.odom, g.odom_cov)
    acc, _, adj, _ = O.pcm_adjacency(g.dim, traj, g.lc_from, g.lc_to, g.lc_pose, g.lc_cov, -1.0, LC_THR,
                                     want_dist=False)
    k, idx = O.max_clique_heu_dense(adj)
    return acc[idx[:k]].astype(np.int64)


def make(name):
    n, nlc, of, seed = WORKLOADS[name]
    g = synth.se3_helix(n_poses=n, n_lc=nlc, outlier_frac=of, seed=seed)
    t0 = time.time()
    inl = pcm_inliers(g)
    print(f"[{name}] PCM clique {len(inl)} ({time.time() - t0:.0f} s)", flush=True)
    poses, kind, ki, kj, meas, info, nk = problem.solve_arrays(g, inl)
    G = S.graph_from_arrays(poses, kind, ki, kj, meas, info)
    t0 = time.time()
    ref = S.gnc_optimize(G, S.GncParams(inlier_threshold=GNC_ALPHA, known_inliers=np.arange(nk)))
    print(f"[{name}] GNC {ref['iterations']} iterations, {int((ref['weights'][nk:] > 0.5).sum())} of "
          f"{len(inl)} kept ({time.time() - t0:.0f} s)", flush=True)
    out = os.path.join(HERE, f"solve_{name}.npz")
    fields = dict(inliers=inl, weights=ref["weights"], q=synth.rot_to_quat(ref["R"]), t=ref["t"],
                  iterations=ref["iterations"], cost=ref["cost"], mu=np.asarray(ref["mu_trace"]))
    fields.update(optimum(G, ref["R"], ref["t"], ref["weights"]))
    np.savez_compressed(out, **fields)
    print(f"[{name}] wrote {out}", flush=True)


def optimum(G, R, t, w):
    """The converged optimum of the final weights, from the GNC result."""
    p = S.LMParams(relative_error_tol=1e-12, absolute_error_tol=0.0, max_iterations=50)
    Ro, to, cost, it = S.lm_optimize(G, R, t, w, p)
    print(f"  optimum: {it} more LM iterations, cost {cost:.12f}", flush=True)
    return dict(dq_opt=(synth.rot_to_quat(Ro) - synth.rot_to_quat(R)).astype(np.float32),
                dt_opt=(to - t).astype(np.float32), cost_opt=cost)


def add_optimum(name):
    n, nlc, of, seed = WORKLOADS[name]
    path = os.path.join(HERE, f"solve_{name}.npz")
    z = dict(np.load(path, allow_pickle=False))
    g = synth.se3_helix(n_poses=n, n_lc=nlc, outlier_frac=of, seed=seed)
    G = S.graph_from_arrays(*problem.solve_arrays(g, z["inliers"])[:-1])
    z.update(optimum(G, synth.quat_to_rot(z["q"]), z["t"], z["weights"]))
    np.savez_compressed(path, **z)
    print(f"[{name}] added the optimum to {path}", flush=True)


if __name__ == "__main__":
    args = sys.argv[1:]
    if args and args[0] == "--optimum":
        for name in args[1:]:
            add_optimum(name)
    else:
        for name in args or ["se3_5k"]:
            make(name)
