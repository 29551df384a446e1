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


# Pose2 (x, y, theta) inside the SE(3) tangent (omega, v): theta -> omega_z,
# x -> v_x, y -> v_y; the out-of-plane directions carry unit information
# (they decouple from the planar ones for planar poses and their residuals
# vanish), as the C++ facade embeds Pose2 graphs.
PLANE = (3, 4, 2)


def embed_pose2(p):
    """(..., 4) Pose2 rows (cos, sin, x, y) -> (..., 12) planar Pose3 rows."""
    p = np.asarray(p, dtype=np.float64)
    c, s, x, y = p[..., 0], p[..., 1], p[..., 2], p[..., 3]
    z, o = np.zeros_like(c), np.ones_like(c)
    return np.stack([c, -s, z, s, c, z, z, z, o, x, y, z], -1)


def embed_info2(I3):
    """(..., 3, 3) Pose2 information in (x, y, theta) -> (..., 36) SE(3) information."""
    I3 = np.asarray(I3, dtype=np.float64)
    out = np.zeros(I3.shape[:-2] + (6, 6))
    for k in (0, 1, 5):
        out[..., k, k] = 1.0
    for a in range(3):
        for b in range(3):
            out[..., PLANE[a], PLANE[b]] = I3[..., a, b]
    return out.reshape(I3.shape[:-2] + (36,))


def solve_arrays_se2(gr, lc_idx=None, prior_var=0.01):
    """solve_arrays for a Pose2 graph, embedded in the plane; solve it with
    gnc_factor_dim=3 and pose2_chart=1 (GTSAM's default Pose2 chart)."""
    if gr.dim != 2:
        raise ValueError("solve_arrays_se2: SE(2) graphs only")
    n = gr.n_poses
    lc = np.arange(gr.n_lc) if lc_idx is None else np.asarray(lc_idx, dtype=np.int64)
    n_odom = n - 1
    F = n_odom + 1 + len(lc)
    kind = np.zeros(F, np.int32)
    ki = np.zeros(F, np.int32)
    kj = np.zeros(F, np.int32)
    meas = np.zeros((F, 12))
    cov = np.zeros((F, 9))
    ki[:n_odom] = np.arange(n_odom)
    kj[:n_odom] = np.arange(1, n)
    meas[:n_odom] = embed_pose2(gr.odom)
    cov[:n_odom] = gr.odom_cov
    kind[n_odom] = 1
    meas[n_odom] = embed_pose2(gr.gt[0])
    cov[n_odom] = (np.eye(3) * prior_var).reshape(9)
    s = n_odom + 1
    ki[s:] = gr.lc_from[lc]
    kj[s:] = gr.lc_to[lc]
    meas[s:] = embed_pose2(gr.lc_pose[lc])
    cov[s:] = gr.lc_cov[lc]
    info = embed_info2(np.linalg.inv(cov.reshape(-1, 3, 3)))
    return np.ascontiguousarray(embed_pose2(gr.values)), kind, ki, kj, meas, info, n_odom + 1


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
