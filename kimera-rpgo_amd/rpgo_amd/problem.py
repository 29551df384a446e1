"""Solve-problem arrays (the rpgo_solve C-ABI layout) for a synthetic pose graph.

Factor order follows Pcm::buildGraphToOptimize (include/KimeraRPGO/outlier/
Pcm.h:977-1005): odometry factors, then the special factors (a prior on pose
0), then the selected loop closures; the odometry and the prior are GNC's
known inliers (src/RobustSolver.cpp:105-113).  Information matrices are the
inverses of the generator's covariances (noise models built from covariances,
as the reference's factors carry them).
"""
from __future__ import annotations

import numpy as np


def solve_arrays(gr, lc_idx=None, prior_var=0.01):
    """Returns (poses, kind, key_i, key_j, measured, information, n_known).

    gr: synth.PoseGraph (SE(3)); lc_idx: loop-closure indices to include (all
    when None), in the given order.  Poses are the generator's initial values."""
    if gr.dim != 3:
        raise ValueError("solve_arrays: SE(3) graphs only")
    n = gr.n_poses
    lc = np.arange(gr.n_lc) if lc_idx is None else np.asarray(lc_idx, dtype=np.int64)
    n_odom = n - 1
    F = n_odom + 1 + len(lc)
    kind = np.zeros(F, np.int32)
    ki = np.zeros(F, np.int32)
    kj = np.zeros(F, np.int32)
    meas = np.zeros((F, 12))
    cov = np.zeros((F, 36))
    ki[:n_odom] = np.arange(n_odom)
    kj[:n_odom] = np.arange(1, n)
    meas[:n_odom] = gr.odom
    cov[:n_odom] = gr.odom_cov
    kind[n_odom] = 1
    meas[n_odom] = gr.gt[0]
    cov[n_odom] = (np.eye(6) * prior_var).reshape(36)
    s = n_odom + 1
    ki[s:] = gr.lc_from[lc]
    kj[s:] = gr.lc_to[lc]
    meas[s:] = gr.lc_pose[lc]
    cov[s:] = gr.lc_cov[lc]
    info = np.linalg.inv(cov.reshape(-1, 6, 6)).reshape(F, 36)
    return np.ascontiguousarray(gr.values), kind, ki, kj, meas, info, n_odom + 1
